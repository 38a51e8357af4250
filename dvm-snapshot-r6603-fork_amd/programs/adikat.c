/* The reference's ADI benchmark (examples/Adi/adi.c) as cdvmh would convert
 * it: n^3 FP64 array, three ACROSS sweeps with one dependency per iteration,
 * MAX reduction in the third, stop when eps < maxeps.  Used as a known-answer
 * test: at n = 384 and 100 iterations the program checks |eps - 0.07249074| < 1e-6.
 * Handlers are emitted from adikat.loops. */
#include <math.h>
#include <stdlib.h>

#include "dvmhlib2.h"
#include "dvmh_amd.h"
#include "dvmh_prog_common.h"

void loop_adikat_init_host(DvmType *, DvmType[], long *);
void loop_adikat_init_cuda(DvmType *, DvmType[], long *);
void loop_adikat_x_host(DvmType *, DvmType[]);
void loop_adikat_x_cuda(DvmType *, DvmType[]);
void loop_adikat_y_host(DvmType *, DvmType[]);
void loop_adikat_y_cuda(DvmType *, DvmType[]);
void loop_adikat_z_host(DvmType *, DvmType[]);
void loop_adikat_z_cuda(DvmType *, DvmType[]);

static DvmType a[64];
static long n;
static DvmType targets = DEVICE_TYPE_HOST | DEVICE_TYPE_CUDA;

void adikat_set_targets(DvmType deviceTypes, int flags) {
    (void)flags;
    targets = deviceTypes;
}

static DvmType sweepLoop(DvmType region, int axis) {
    DvmType w[6] = {0, 0, 0, 0, 0, 0};
    w[2 * axis] = 1;
    w[2 * axis + 1] = 1;
    DvmType l = dvmh_loop_create_C(region, 3, (DvmType)1, (DvmType)(n - 2), (DvmType)1, (DvmType)1, (DvmType)(n - 2), (DvmType)1,
            (DvmType)1, (DvmType)(n - 2), (DvmType)1);
    dvmh_loop_map_C(l, a, 3, dvmh_alignment_linear_C(1, 1, 0), dvmh_alignment_linear_C(2, 1, 0), dvmh_alignment_linear_C(3, 1, 0));
    dvmh_loop_across_C(l, 0, a, 3, w[0], w[1], w[2], w[3], w[4], w[5]);
    return l;
}

/* Runs the whole program; eps_hist gets eps per iteration; returns the number
 * of iterations done (fewer than itmax when eps drops below maxeps). */
int adikat_run(long size, int itmax, double maxeps, double eps_hist[]) {
    DvmType region, l;
    double eps = 0;
    int it;
    n = size;
    dvmh_init_lib_C(0);
    dvmh_array_declare_C(a, 3, -rt_DOUBLE, (DvmType)0, (DvmType)1, (DvmType)1, (DvmType)n, (DvmType)1, (DvmType)1, (DvmType)n,
            (DvmType)1, (DvmType)1);
    dvmh_array_alloc_C(a, n * n * n * sizeof(double));
    dvmh_distribute_C(a, 3, dvmh_distribution_block_C(1), dvmh_distribution_block_C(2), dvmh_distribution_block_C(3));

    dvmh_line_C(86, "adi.c");
    region = dvmh_region_create_C(0);
    dvmh_region_register_array_C(region, INTENT_OUT, a, "a");
    dvmh_region_register_scalar_C(region, INTENT_IN, &n, sizeof(n), "n");
    dvmh_region_execute_on_targets_C(region, targets);
    l = dvmh_loop_create_C(region, 3, (DvmType)0, (DvmType)(n - 1), (DvmType)1, (DvmType)0, (DvmType)(n - 1), (DvmType)1, (DvmType)0,
            (DvmType)(n - 1), (DvmType)1);
    dvmh_loop_map_C(l, a, 3, dvmh_alignment_linear_C(1, 1, 0), dvmh_alignment_linear_C(2, 1, 0), dvmh_alignment_linear_C(3, 1, 0));
    dvmh_loop_register_handler_C(l, DEVICE_TYPE_HOST, HANDLER_TYPE_PARALLEL, dvmh_handler_func_C((DvmHandlerFunc)loop_adikat_init_host, 2, a, &n));
    dvmh_loop_register_handler_C(l, DEVICE_TYPE_CUDA, 0, dvmh_handler_func_C((DvmHandlerFunc)loop_adikat_init_cuda, 2, a, &n));
    dvmh_loop_perform_C(l);
    dvmh_region_end_C(region);

    for (it = 1; it <= itmax; it++) {
        eps = 0;
        dvmh_actual_variable2_(&eps);
        region = dvmh_region_create_C(0);
        dvmh_region_register_array_C(region, INTENT_INOUT, a, "a");
        dvmh_region_execute_on_targets_C(region, targets);
        dvmh_line_C(43, "adi.c");
        l = sweepLoop(region, 0);
        dvmh_loop_register_handler_C(l, DEVICE_TYPE_HOST, HANDLER_TYPE_PARALLEL, dvmh_handler_func_C((DvmHandlerFunc)loop_adikat_x_host, 1, a));
        dvmh_loop_register_handler_C(l, DEVICE_TYPE_CUDA, 0, dvmh_handler_func_C((DvmHandlerFunc)loop_adikat_x_cuda, 1, a));
        dvmh_loop_perform_C(l);
        dvmh_line_C(48, "adi.c");
        l = sweepLoop(region, 1);
        dvmh_loop_register_handler_C(l, DEVICE_TYPE_HOST, HANDLER_TYPE_PARALLEL, dvmh_handler_func_C((DvmHandlerFunc)loop_adikat_y_host, 1, a));
        dvmh_loop_register_handler_C(l, DEVICE_TYPE_CUDA, 0, dvmh_handler_func_C((DvmHandlerFunc)loop_adikat_y_cuda, 1, a));
        dvmh_loop_perform_C(l);
        dvmh_line_C(53, "adi.c");
        l = sweepLoop(region, 2);
        dvmh_loop_reduction_C(l, rf_MAX, &eps, rt_DOUBLE, 1, 0, 0);
        dvmh_loop_register_handler_C(l, DEVICE_TYPE_HOST, HANDLER_TYPE_PARALLEL, dvmh_handler_func_C((DvmHandlerFunc)loop_adikat_z_host, 1, a));
        dvmh_loop_register_handler_C(l, DEVICE_TYPE_CUDA, 0, dvmh_handler_func_C((DvmHandlerFunc)loop_adikat_z_cuda, 1, a));
        dvmh_loop_perform_C(l);
        dvmh_region_end_C(region);
        dvmh_get_actual_variable2_(&eps);
        if (eps_hist) eps_hist[it - 1] = eps;
        if (eps < maxeps) break;
    }
    dvmh_delete_object_(a);
    return it > itmax ? itmax : it;
}
