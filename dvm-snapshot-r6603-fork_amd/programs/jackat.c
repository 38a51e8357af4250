/* The reference's FP32 Jacobi benchmarks (examples/Performance/jac2d_perf.c,
 * jac3d_perf.c) as cdvmh would convert them.  Known-answer tests: at L = 8000
 * (2-D) and L = 384 (3-D), 100 iterations, the programs check their final eps
 * against 58.37598 (+-1e-3) and 5.058044 (+-1e-4).  Handlers from jackat.loops. */
#include <math.h>
#include <stdlib.h>

#include "dvmhlib2.h"
#include "dvmh_amd.h"
#include "dvmh_prog_common.h"

#define DECL(N, ...) void N##_host(DvmType *, __VA_ARGS__); void N##_cuda(DvmType *, __VA_ARGS__);
DECL(loop_jac2kat_init, DvmType[], DvmType[], long *)
DECL(loop_jac2kat_max, DvmType[], DvmType[])
DECL(loop_jac2kat_st, DvmType[], DvmType[])
DECL(loop_jac3kat_init, DvmType[], DvmType[], long *)
DECL(loop_jac3kat_max, DvmType[], DvmType[])
DECL(loop_jac3kat_st, DvmType[], DvmType[])

static DvmType targets = DEVICE_TYPE_HOST | DEVICE_TYPE_CUDA;
static long L;

void jackat_set_targets(DvmType deviceTypes, int flags) {
    (void)flags;
    targets = deviceTypes;
}

static DvmType loopOn(DvmType region, int rank, DvmType base[], long lo, long hi) {
    DvmType l;
    if (rank == 2) {
        l = dvmh_loop_create_C(region, 2, (DvmType)lo, (DvmType)hi, (DvmType)1, (DvmType)lo, (DvmType)hi, (DvmType)1);
        dvmh_loop_map_C(l, base, 2, dvmh_alignment_linear_C(1, 1, 0), dvmh_alignment_linear_C(2, 1, 0));
    } else {
        l = dvmh_loop_create_C(region, 3, (DvmType)lo, (DvmType)hi, (DvmType)1, (DvmType)lo, (DvmType)hi, (DvmType)1, (DvmType)lo,
                (DvmType)hi, (DvmType)1);
        dvmh_loop_map_C(l, base, 3, dvmh_alignment_linear_C(1, 1, 0), dvmh_alignment_linear_C(2, 1, 0),
                dvmh_alignment_linear_C(3, 1, 0));
    }
    return l;
}

typedef void (*H2)(DvmType *, DvmType[], DvmType[]);
typedef void (*H3)(DvmType *, DvmType[], DvmType[], long *);

/* rank 2 or 3; returns the final eps, eps_hist per iteration. */
float jackat_run(int rank, long size, int itmax, float maxeps, float eps_hist[]) {
    DvmType A[64], B[64], region, l;
    float eps = 0;
    L = size;
    dvmh_init_lib_C(0);
    if (rank == 2) {
        dvmh_array_create_C(A, 2, -rt_FLOAT, (DvmType)L, (DvmType)1, (DvmType)1, (DvmType)L, (DvmType)1, (DvmType)1);
        dvmh_distribute_C(A, 2, dvmh_distribution_block_C(1), dvmh_distribution_block_C(2));
        dvmh_array_create_C(B, 2, -rt_FLOAT, (DvmType)L, (DvmType)1, (DvmType)1, (DvmType)L, (DvmType)1, (DvmType)1);
        dvmh_align_C(B, A, 2, dvmh_alignment_linear_C(1, 1, 0), dvmh_alignment_linear_C(2, 1, 0));
    } else {
        dvmh_array_create_C(A, 3, -rt_FLOAT, (DvmType)L, (DvmType)1, (DvmType)1, (DvmType)L, (DvmType)1, (DvmType)1, (DvmType)L,
                (DvmType)1, (DvmType)1);
        dvmh_distribute_C(A, 3, dvmh_distribution_block_C(1), dvmh_distribution_block_C(2), dvmh_distribution_block_C(3));
        dvmh_array_create_C(B, 3, -rt_FLOAT, (DvmType)L, (DvmType)1, (DvmType)1, (DvmType)L, (DvmType)1, (DvmType)1, (DvmType)L,
                (DvmType)1, (DvmType)1);
        dvmh_align_C(B, A, 3, dvmh_alignment_linear_C(1, 1, 0), dvmh_alignment_linear_C(2, 1, 0), dvmh_alignment_linear_C(3, 1, 0));
    }
    H3 initH = rank == 2 ? loop_jac2kat_init_host : loop_jac3kat_init_host;
    H3 initC = rank == 2 ? loop_jac2kat_init_cuda : loop_jac3kat_init_cuda;
    H2 maxH = rank == 2 ? loop_jac2kat_max_host : loop_jac3kat_max_host;
    H2 maxC = rank == 2 ? loop_jac2kat_max_cuda : loop_jac3kat_max_cuda;
    H2 stH = rank == 2 ? loop_jac2kat_st_host : loop_jac3kat_st_host;
    H2 stC = rank == 2 ? loop_jac2kat_st_cuda : loop_jac3kat_st_cuda;

    dvmh_line_C(rank == 2 ? 44 : 27, rank == 2 ? "jac2d_perf.c" : "jac3d_perf.c");
    region = dvmh_region_create_C(0);
    dvmh_region_register_array_C(region, INTENT_OUT, A, "A");
    dvmh_region_register_array_C(region, INTENT_OUT, B, "B");
    dvmh_region_register_scalar_C(region, INTENT_IN, &L, sizeof(L), "L");
    dvmh_region_execute_on_targets_C(region, targets);
    l = loopOn(region, rank, A, 0, L - 1);
    dvmh_loop_register_handler_C(l, DEVICE_TYPE_HOST, HANDLER_TYPE_PARALLEL, dvmh_handler_func_C((DvmHandlerFunc)initH, 3, A, B, &L));
    dvmh_loop_register_handler_C(l, DEVICE_TYPE_CUDA, 0, dvmh_handler_func_C((DvmHandlerFunc)initC, 3, A, B, &L));
    dvmh_loop_perform_C(l);
    dvmh_region_end_C(region);

    for (int it = 1; it <= itmax; it++) {
        eps = 0;
        dvmh_actual_variable2_(&eps);
        region = dvmh_region_create_C(0);
        dvmh_region_register_array_C(region, INTENT_INOUT, A, "A");
        dvmh_region_register_array_C(region, INTENT_INOUT, B, "B");
        dvmh_region_execute_on_targets_C(region, targets);
        dvmh_line_C(rank == 2 ? 57 : 60, rank == 2 ? "jac2d_perf.c" : "jac3d_perf.c");
        l = loopOn(region, rank, A, 1, L - 2);
        dvmh_loop_reduction_C(l, rf_MAX, &eps, rt_FLOAT, 1, 0, 0);
        dvmh_loop_register_handler_C(l, DEVICE_TYPE_HOST, HANDLER_TYPE_PARALLEL, dvmh_handler_func_C((DvmHandlerFunc)maxH, 2, A, B));
        dvmh_loop_register_handler_C(l, DEVICE_TYPE_CUDA, 0, dvmh_handler_func_C((DvmHandlerFunc)maxC, 2, A, B));
        dvmh_loop_perform_C(l);
        dvmh_line_C(rank == 2 ? 65 : 72, rank == 2 ? "jac2d_perf.c" : "jac3d_perf.c");
        dvmh_shadow_renew_C(A, 0, 0);
        l = loopOn(region, rank, B, 1, L - 2);
        dvmh_loop_register_handler_C(l, DEVICE_TYPE_HOST, HANDLER_TYPE_PARALLEL, dvmh_handler_func_C((DvmHandlerFunc)stH, 2, A, B));
        dvmh_loop_register_handler_C(l, DEVICE_TYPE_CUDA, 0, dvmh_handler_func_C((DvmHandlerFunc)stC, 2, A, B));
        dvmh_loop_perform_C(l);
        dvmh_region_end_C(region);
        dvmh_get_actual_variable2_(&eps);
        if (eps_hist) eps_hist[it - 1] = eps;
        if (eps < maxeps) break;
    }
    dvmh_delete_object_(B);
    dvmh_delete_object_(A);
    return eps;
}
