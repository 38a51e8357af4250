/* sor3d.cdv as converted by cdvmh; handlers are emitted from sor3d.loops. */
#include <math.h>
#include <stdlib.h>

#include "dvmhlib2.h"
#include "dvmh_amd.h"
#include "dvmh_prog_common.h"

void loop_sor3d_18_host(DvmType *, DvmType[], long *, unsigned long long *);
void loop_sor3d_18_cuda(DvmType *, DvmType[], long *, unsigned long long *);
void loop_sor3d_32_host(DvmType *, DvmType[], double *);
void loop_sor3d_32_cuda(DvmType *, DvmType[], double *);
void loop_sor3d_32_cuda_chain(DvmType *);

static DvmType A[64];
static long n;
static double w;
static unsigned long long seed;
static DvmType targets = DEVICE_TYPE_HOST | DEVICE_TYPE_CUDA;
static int regionFlags = 0;

void sor3d_set_targets(DvmType deviceTypes, int flags) {
    targets = deviceTypes;
    regionFlags = flags;
}

void sor3d_setup(long size, double omega, unsigned long long s) {
    DvmType cur_region, cur_loop;
    n = size;
    w = omega;
    seed = s;
    dvmh_init_lib_C(0);
    dvmh_line_C(13, "sor3d.cdv");
    dvmh_array_declare_C(A, 3, -rt_DOUBLE, (DvmType)0, (DvmType)1, (DvmType)1, (DvmType)n, (DvmType)1, (DvmType)1, (DvmType)n,
            (DvmType)1, (DvmType)1);
    dvmh_line_C(14, "sor3d.cdv");
    dvmh_array_alloc_C(A, n * n * n * sizeof(double));
    dvmh_distribute_C(A, 3, dvmh_distribution_block_C(1), dvmh_distribution_block_C(2), dvmh_distribution_block_C(3));
    dvmh_line_C(16, "sor3d.cdv");
    cur_region = dvmh_region_create_C(regionFlags & ~REGION_AMD_GRAPH);
    dvmh_region_register_array_C(cur_region, INTENT_OUT, A, "A");
    dvmh_region_register_scalar_C(cur_region, INTENT_IN, &n, sizeof(n), "n");
    dvmh_region_register_scalar_C(cur_region, INTENT_IN, &seed, sizeof(seed), "seed");
    dvmh_region_execute_on_targets_C(cur_region, targets);
    dvmh_line_C(18, "sor3d.cdv");
    cur_loop = dvmh_loop_create_C(cur_region, 3, (DvmType)0, (DvmType)(n - 1), (DvmType)1, (DvmType)0, (DvmType)(n - 1), (DvmType)1,
            (DvmType)0, (DvmType)(n - 1), (DvmType)1);
    dvmh_loop_map_C(cur_loop, A, 3, dvmh_alignment_linear_C(1, 1, 0), dvmh_alignment_linear_C(2, 1, 0), dvmh_alignment_linear_C(3, 1, 0));
    dvmh_loop_register_handler_C(cur_loop, DEVICE_TYPE_HOST, HANDLER_TYPE_PARALLEL,
            dvmh_handler_func_C((DvmHandlerFunc)loop_sor3d_18_host, 3, A, &n, &seed));
    dvmh_loop_register_handler_C(cur_loop, DEVICE_TYPE_CUDA, 0, dvmh_handler_func_C((DvmHandlerFunc)loop_sor3d_18_cuda, 3, A, &n, &seed));
    dvmh_loop_perform_C(cur_loop);
    dvmh_region_end_C(cur_region);
}

double sor3d_iterate(int itmax, double eps_hist[]) {
    DvmType cur_region, cur_loop;
    double eps, startt, endt;
    dvmh_barrier();
    startt = dvmh_wtime();
    for (int it = 1; it <= itmax; it++) {
        eps = 0;
        dvmh_line_C(29, "sor3d.cdv");
        dvmh_actual_variable2_(&eps);
        cur_region = dvmh_region_create_C(regionFlags);
        dvmh_region_register_array_C(cur_region, INTENT_INOUT, A, "A");
        dvmh_region_register_scalar_C(cur_region, INTENT_IN, &w, sizeof(w), "w");
        dvmh_region_execute_on_targets_C(cur_region, targets);
        dvmh_line_C(32, "sor3d.cdv");
        cur_loop = dvmh_loop_create_C(cur_region, 3, (DvmType)1, (DvmType)(n - 2), (DvmType)1, (DvmType)1, (DvmType)(n - 2), (DvmType)1,
                (DvmType)1, (DvmType)(n - 2), (DvmType)1);
        dvmh_loop_map_C(cur_loop, A, 3, dvmh_alignment_linear_C(1, 1, 0), dvmh_alignment_linear_C(2, 1, 0),
                dvmh_alignment_linear_C(3, 1, 0));
        dvmh_loop_across_C(cur_loop, 0, A, 3, (DvmType)1, (DvmType)1, (DvmType)1, (DvmType)1, (DvmType)1, (DvmType)1);
        dvmh_loop_reduction_C(cur_loop, rf_MAX, &eps, rt_DOUBLE, 1, 0, 0);
        dvmh_loop_register_handler_C(cur_loop, DEVICE_TYPE_HOST, HANDLER_TYPE_PARALLEL,
                dvmh_handler_func_C((DvmHandlerFunc)loop_sor3d_32_host, 2, A, &w));
        dvmh_loop_register_handler_C(cur_loop, DEVICE_TYPE_CUDA, 0, dvmh_handler_func_C((DvmHandlerFunc)loop_sor3d_32_cuda, 2, A, &w));
        dvmh_loop_perform_C(cur_loop);
        dvmh_region_end_C(cur_region);
        dvmh_line_C(44, "sor3d.cdv");
        dvmh_get_actual_variable2_(&eps);
        if (eps_hist) eps_hist[it - 1] = eps;
    }
    dvmh_barrier();
    endt = dvmh_wtime();
    return endt - startt;
}

/* The same sweeps, `group` at a time between dvmh_amd_batch_begin_C and
 * dvmh_amd_batch_end_C: each sweep keeps its own region, loop and
 * REDUCTION(MAX) variable (eps[q] of the group), and the loop also registers
 * the chain handler, so a group of sweeps runs as one pipelined launch
 * sequence; the residuals are read after the group.  eps_hist as in
 * sor3d_iterate. */
double sor3d_iterate_chain(int itmax, int group, double eps_hist[]) {
    DvmType cur_region, cur_loop;
    double startt, endt;
    double *eps = (double *)calloc((size_t)(group > 0 ? group : 1), sizeof(double));
    dvmh_barrier();
    startt = dvmh_wtime();
    for (int it0 = 1; it0 <= itmax; it0 += group) {
        int g = itmax - it0 + 1 < group ? itmax - it0 + 1 : group;
        dvmh_amd_batch_begin_C();
        for (int q = 0; q < g; q++) {
            eps[q] = 0;
            dvmh_line_C(29, "sor3d.cdv");
            cur_region = dvmh_region_create_C(regionFlags & ~REGION_AMD_GRAPH);
            dvmh_region_register_array_C(cur_region, INTENT_INOUT, A, "A");
            dvmh_region_register_scalar_C(cur_region, INTENT_IN, &w, sizeof(w), "w");
            dvmh_region_execute_on_targets_C(cur_region, targets);
            dvmh_line_C(32, "sor3d.cdv");
            cur_loop = dvmh_loop_create_C(cur_region, 3, (DvmType)1, (DvmType)(n - 2), (DvmType)1, (DvmType)1, (DvmType)(n - 2), (DvmType)1,
                    (DvmType)1, (DvmType)(n - 2), (DvmType)1);
            dvmh_loop_map_C(cur_loop, A, 3, dvmh_alignment_linear_C(1, 1, 0), dvmh_alignment_linear_C(2, 1, 0),
                    dvmh_alignment_linear_C(3, 1, 0));
            dvmh_loop_across_C(cur_loop, 0, A, 3, (DvmType)1, (DvmType)1, (DvmType)1, (DvmType)1, (DvmType)1, (DvmType)1);
            dvmh_loop_reduction_C(cur_loop, rf_MAX, &eps[q], rt_DOUBLE, 1, 0, 0);
            dvmh_loop_register_handler_C(cur_loop, DEVICE_TYPE_HOST, HANDLER_TYPE_PARALLEL,
                    dvmh_handler_func_C((DvmHandlerFunc)loop_sor3d_32_host, 2, A, &w));
            dvmh_loop_register_handler_C(cur_loop, DEVICE_TYPE_CUDA, 0, dvmh_handler_func_C((DvmHandlerFunc)loop_sor3d_32_cuda, 2, A, &w));
            dvmh_amd_loop_register_chain_handler_C(cur_loop, dvmh_handler_func_C((DvmHandlerFunc)loop_sor3d_32_cuda_chain, 0));
            dvmh_loop_perform_C(cur_loop);
            dvmh_region_end_C(cur_region);
        }
        dvmh_amd_batch_end_C();
        dvmh_line_C(44, "sor3d.cdv");
        for (int q = 0; q < g; q++) {
            dvmh_get_actual_variable2_(&eps[q]);
            if (eps_hist) eps_hist[it0 - 1 + q] = eps[q];
        }
    }
    dvmh_barrier();
    endt = dvmh_wtime();
    free(eps);
    return endt - startt;
}

void sor3d_fetch(double *outA) {
    dvmh_get_actual_array2_(A);
    long total = n * n * n;
    const double *a = (const double *)dvmh_get_natural_base_C(0, A);
    for (long q = 0; q < total; q++) outA[q] = a[q];
}

/* get_actual(A[lo0:hi0][lo1:hi1][lo2:hi2]), then copies that block, row-major,
 * into out. */
void sor3d_fetch_block(const long lo[3], const long hi[3], double *out) {
    dvmh_get_actual_subarray_C(A, 3, (DvmType)lo[0], (DvmType)hi[0], (DvmType)lo[1], (DvmType)hi[1], (DvmType)lo[2],
            (DvmType)hi[2]);
    const double *a = (const double *)dvmh_get_natural_base_C(0, A);
    for (long i = lo[0]; i <= hi[0]; i++)
        for (long j = lo[1]; j <= hi[1]; j++)
            for (long k = lo[2]; k <= hi[2]; k++) *out++ = a[(i * n + j) * n + k];
}

void sor3d_teardown(void) { dvmh_delete_object_(A); }
