/* jac3d.cdv as converted by cdvmh (runtime calls + handler registration);
 * handlers are emitted from jac3d.loops.  Entry points: setup / iterate /
 * fetch / teardown (main() in jac3d_main.c runs them in sequence). */
#include <math.h>
#include <stdlib.h>

#include "dvmhlib2.h"
#include "dvmh_amd.h"
#include "dvmh_prog_common.h"

void loop_jac3d_21_host(DvmType *, DvmType[], DvmType[], long *, unsigned long long *);
void loop_jac3d_21_cuda(DvmType *, DvmType[], DvmType[], long *, unsigned long long *);
void loop_jac3d_41_host(DvmType *, DvmType[], DvmType[]);
void loop_jac3d_41_cuda(DvmType *, DvmType[], DvmType[]);
void loop_jac3d_54_host(DvmType *, DvmType[], DvmType[]);
void loop_jac3d_54_cuda(DvmType *, DvmType[], DvmType[]);

static DvmType A[64], B[64];
static long n;
static unsigned long long seed;
static int itDone;
static DvmType targets = DEVICE_TYPE_HOST | DEVICE_TYPE_CUDA;
static int regionFlags = 0;

void jac3d_set_targets(DvmType deviceTypes, int flags) {
    targets = deviceTypes;
    regionFlags = flags;
}

void jac3d_setup(long size, unsigned long long s) {
    DvmType cur_region, cur_loop;
    n = size;
    seed = s;
    itDone = 0;
    dvmh_init_lib_C(0);
    dvmh_line_C(14, "jac3d.cdv");
    dvmh_array_declare_C(A, 3, -rt_DOUBLE, (DvmType)0, (DvmType)1, (DvmType)1, (DvmType)n, (DvmType)1, (DvmType)1, (DvmType)n,
            (DvmType)1, (DvmType)1);
    dvmh_array_declare_C(B, 3, -rt_DOUBLE, (DvmType)0, (DvmType)1, (DvmType)1, (DvmType)n, (DvmType)1, (DvmType)1, (DvmType)n,
            (DvmType)1, (DvmType)1);
    dvmh_line_C(16, "jac3d.cdv");
    dvmh_array_alloc_C(A, n * n * n * sizeof(double));
    dvmh_distribute_C(A, 3, dvmh_distribution_block_C(1), dvmh_distribution_block_C(2), dvmh_distribution_block_C(3));
    dvmh_line_C(17, "jac3d.cdv");
    dvmh_array_alloc_C(B, n * n * n * sizeof(double));
    dvmh_align_C(B, A, 3, dvmh_alignment_linear_C(1, 1, 0), dvmh_alignment_linear_C(2, 1, 0), dvmh_alignment_linear_C(3, 1, 0));

    dvmh_line_C(19, "jac3d.cdv");
    cur_region = dvmh_region_create_C(regionFlags & ~REGION_AMD_GRAPH);
    dvmh_region_register_array_C(cur_region, INTENT_OUT, A, "A");
    dvmh_region_register_array_C(cur_region, INTENT_OUT, B, "B");
    dvmh_region_register_scalar_C(cur_region, INTENT_IN, &n, sizeof(n), "n");
    dvmh_region_register_scalar_C(cur_region, INTENT_IN, &seed, sizeof(seed), "seed");
    dvmh_region_execute_on_targets_C(cur_region, targets);
    dvmh_line_C(21, "jac3d.cdv");
    cur_loop = dvmh_loop_create_C(cur_region, 3, (DvmType)0, (DvmType)(n - 1), (DvmType)1, (DvmType)0, (DvmType)(n - 1), (DvmType)1,
            (DvmType)0, (DvmType)(n - 1), (DvmType)1);
    dvmh_loop_map_C(cur_loop, A, 3, dvmh_alignment_linear_C(1, 1, 0), dvmh_alignment_linear_C(2, 1, 0), dvmh_alignment_linear_C(3, 1, 0));
    dvmh_loop_register_handler_C(cur_loop, DEVICE_TYPE_HOST, HANDLER_TYPE_PARALLEL,
            dvmh_handler_func_C((DvmHandlerFunc)loop_jac3d_21_host, 4, A, B, &n, &seed));
    dvmh_loop_register_handler_C(cur_loop, DEVICE_TYPE_CUDA, 0, dvmh_handler_func_C((DvmHandlerFunc)loop_jac3d_21_cuda, 4, A, B, &n, &seed));
    dvmh_loop_perform_C(cur_loop);
    dvmh_region_end_C(cur_region);
}

static void sweep(int odd, double *eps, double *s) {
    DvmType cur_region, cur_loop;
    dvmh_line_C(odd ? 39 : 52, "jac3d.cdv");
    cur_region = dvmh_region_create_C(regionFlags);
    dvmh_region_register_array_C(cur_region, INTENT_INOUT, A, "A");
    dvmh_region_register_array_C(cur_region, INTENT_INOUT, B, "B");
    dvmh_region_execute_on_targets_C(cur_region, targets);
    dvmh_line_C(odd ? 41 : 54, "jac3d.cdv");
    dvmh_shadow_renew_C(odd ? A : B, 0, 0);
    cur_loop = dvmh_loop_create_C(cur_region, 3, (DvmType)1, (DvmType)(n - 2), (DvmType)1, (DvmType)1, (DvmType)(n - 2), (DvmType)1,
            (DvmType)1, (DvmType)(n - 2), (DvmType)1);
    dvmh_loop_map_C(cur_loop, odd ? B : A, 3, dvmh_alignment_linear_C(1, 1, 0), dvmh_alignment_linear_C(2, 1, 0),
            dvmh_alignment_linear_C(3, 1, 0));
    dvmh_loop_reduction_C(cur_loop, rf_MAX, eps, rt_DOUBLE, 1, 0, 0);
    dvmh_loop_reduction_C(cur_loop, rf_SUM, s, rt_DOUBLE, 1, 0, 0);
    if (odd) {
        dvmh_loop_register_handler_C(cur_loop, DEVICE_TYPE_HOST, HANDLER_TYPE_PARALLEL,
                dvmh_handler_func_C((DvmHandlerFunc)loop_jac3d_41_host, 2, A, B));
        dvmh_loop_register_handler_C(cur_loop, DEVICE_TYPE_CUDA, 0, dvmh_handler_func_C((DvmHandlerFunc)loop_jac3d_41_cuda, 2, A, B));
    } else {
        dvmh_loop_register_handler_C(cur_loop, DEVICE_TYPE_HOST, HANDLER_TYPE_PARALLEL,
                dvmh_handler_func_C((DvmHandlerFunc)loop_jac3d_54_host, 2, A, B));
        dvmh_loop_register_handler_C(cur_loop, DEVICE_TYPE_CUDA, 0, dvmh_handler_func_C((DvmHandlerFunc)loop_jac3d_54_cuda, 2, A, B));
    }
    dvmh_loop_perform_C(cur_loop);
    dvmh_region_end_C(cur_region);
}

/* Runs itmax more iterations (continuing the A/B alternation across calls).
 * Returns the wall time of the iteration phase (barrier to barrier). */
double jac3d_iterate(int itmax, double eps_hist[], double sum_hist[]) {
    double eps, s, startt, endt;
    dvmh_barrier();
    startt = dvmh_wtime();
    for (int it = 1; it <= itmax; it++) {
        eps = 0;
        s = 0;
        dvmh_line_C(37, "jac3d.cdv");
        dvmh_actual_variable2_(&eps);
        dvmh_actual_variable2_(&s);
        itDone++;
        sweep(itDone % 2 == 1, &eps, &s);
        dvmh_line_C(65, "jac3d.cdv");
        dvmh_get_actual_variable2_(&eps);
        dvmh_get_actual_variable2_(&s);
        if (eps_hist) eps_hist[it - 1] = eps;
        if (sum_hist) sum_hist[it - 1] = s;
    }
    dvmh_barrier();
    endt = dvmh_wtime();
    return endt - startt;
}

/* Copies the array written last (B after an odd count, A after an even one). */
void jac3d_fetch(double *outA, double *outB) {
    dvmh_get_actual_array2_(A);
    dvmh_get_actual_array2_(B);
    long total = n * n * n;
    if (outA) {
        const double *a = (const double *)dvmh_get_natural_base_C(0, A);
        for (long q = 0; q < total; q++) outA[q] = a[q];
    }
    if (outB) {
        const double *b = (const double *)dvmh_get_natural_base_C(0, B);
        for (long q = 0; q < total; q++) outB[q] = b[q];
    }
}

/* get_actual(X[lo0:hi0][lo1:hi1][lo2:hi2]) of A (which = 0) or B (1), then
 * copies that block, row-major, into out. */
void jac3d_fetch_block(int which, const long lo[3], const long hi[3], double *out) {
    DvmType *X = which ? B : A;
    dvmh_get_actual_subarray_C(X, 3, (DvmType)lo[0], (DvmType)hi[0], (DvmType)lo[1], (DvmType)hi[1], (DvmType)lo[2],
            (DvmType)hi[2]);
    const double *x = (const double *)dvmh_get_natural_base_C(0, X);
    for (long i = lo[0]; i <= hi[0]; i++)
        for (long j = lo[1]; j <= hi[1]; j++)
            for (long k = lo[2]; k <= hi[2]; k++) *out++ = x[(i * n + j) * n + k];
}

/* Writes the block from in into the host copy of A (which = 0) or B (1), then
 * actual(X[lo0:hi0][lo1:hi1][lo2:hi2]). */
void jac3d_store_block(int which, const long lo[3], const long hi[3], const double *in) {
    DvmType *X = which ? B : A;
    double *x = (double *)dvmh_get_natural_base_C(0, X);
    for (long i = lo[0]; i <= hi[0]; i++)
        for (long j = lo[1]; j <= hi[1]; j++)
            for (long k = lo[2]; k <= hi[2]; k++) x[(i * n + j) * n + k] = *in++;
    dvmh_actual_subarray_C(X, 3, (DvmType)lo[0], (DvmType)hi[0], (DvmType)lo[1], (DvmType)hi[1], (DvmType)lo[2],
            (DvmType)hi[2]);
}

void jac3d_teardown(void) {
    dvmh_delete_object_(B);
    dvmh_delete_object_(A);
}
