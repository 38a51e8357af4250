/* adi3d.cdv as converted by cdvmh; handlers are emitted from adi3d.loops. */
#include <math.h>
#include <stdlib.h>

#include "dvmhlib2.h"
#include "dvmh_amd.h"
#include "dvmh_prog_common.h"

void loop_adi3d_20_host(DvmType *, DvmType[], long *, unsigned long long *);
void loop_adi3d_20_cuda(DvmType *, DvmType[], long *, unsigned long long *);
#define SWEEP_DECL(L) void loop_adi3d_##L##_host(DvmType *, DvmType[], double *); void loop_adi3d_##L##_cuda(DvmType *, DvmType[], double *);
SWEEP_DECL(34) SWEEP_DECL(39) SWEEP_DECL(44) SWEEP_DECL(49) SWEEP_DECL(54) SWEEP_DECL(59)
void loop_adi3d_64_host(DvmType *, DvmType[], long *);
void loop_adi3d_64_cuda(DvmType *, DvmType[], long *);

static DvmType a[64];
static long n;
static double c;
static unsigned long long seed;
static DvmType targets = DEVICE_TYPE_HOST | DEVICE_TYPE_CUDA;
static int regionFlags = 0;

void adi3d_set_targets(DvmType deviceTypes, int flags) {
    targets = deviceTypes;
    regionFlags = flags;
}

void adi3d_setup(long size, double coef, unsigned long long s) {
    DvmType cur_region, cur_loop;
    n = size;
    c = coef;
    seed = s;
    dvmh_init_lib_C(0);
    dvmh_line_C(15, "adi3d.cdv");
    dvmh_array_declare_C(a, 3, -rt_DOUBLE, (DvmType)0, (DvmType)1, (DvmType)1, (DvmType)n, (DvmType)1, (DvmType)1, (DvmType)n,
            (DvmType)1, (DvmType)1);
    dvmh_line_C(16, "adi3d.cdv");
    dvmh_array_alloc_C(a, n * n * n * sizeof(double));
    dvmh_distribute_C(a, 3, dvmh_distribution_block_C(1), dvmh_distribution_block_C(2), dvmh_distribution_block_C(3));
    dvmh_line_C(18, "adi3d.cdv");
    cur_region = dvmh_region_create_C(regionFlags & ~REGION_AMD_GRAPH);
    dvmh_region_register_array_C(cur_region, INTENT_OUT, a, "a");
    dvmh_region_register_scalar_C(cur_region, INTENT_IN, &n, sizeof(n), "n");
    dvmh_region_register_scalar_C(cur_region, INTENT_IN, &seed, sizeof(seed), "seed");
    dvmh_region_execute_on_targets_C(cur_region, targets);
    dvmh_line_C(20, "adi3d.cdv");
    cur_loop = dvmh_loop_create_C(cur_region, 3, (DvmType)0, (DvmType)(n - 1), (DvmType)1, (DvmType)0, (DvmType)(n - 1), (DvmType)1,
            (DvmType)0, (DvmType)(n - 1), (DvmType)1);
    dvmh_loop_map_C(cur_loop, a, 3, dvmh_alignment_linear_C(1, 1, 0), dvmh_alignment_linear_C(2, 1, 0), dvmh_alignment_linear_C(3, 1, 0));
    dvmh_loop_register_handler_C(cur_loop, DEVICE_TYPE_HOST, HANDLER_TYPE_PARALLEL,
            dvmh_handler_func_C((DvmHandlerFunc)loop_adi3d_20_host, 3, a, &n, &seed));
    dvmh_loop_register_handler_C(cur_loop, DEVICE_TYPE_CUDA, 0, dvmh_handler_func_C((DvmHandlerFunc)loop_adi3d_20_cuda, 3, a, &n, &seed));
    dvmh_loop_perform_C(cur_loop);
    dvmh_region_end_C(cur_region);
}

typedef void (*Handler)(DvmType *, DvmType[], double *);

/* One ACROSS sweep: axis 0..2, dir +1 forward (x[i] += c*x[i-1]), -1 backward. */
static void sweep(DvmType cur_region, int line, int axis, int dir, Handler hh, Handler hc) {
    DvmType lo[3] = {0, 0, 0}, hi[3] = {n - 1, n - 1, n - 1}, st[3] = {1, 1, 1};
    DvmType w[6] = {0, 0, 0, 0, 0, 0};
    if (dir > 0) { lo[axis] = 1; hi[axis] = n - 1; w[2 * axis] = 1; }
    else { lo[axis] = n - 2; hi[axis] = 0; st[axis] = -1; w[2 * axis + 1] = 1; }
    dvmh_line_C(line, "adi3d.cdv");
    DvmType cur_loop = dvmh_loop_create_C(cur_region, 3, lo[0], hi[0], st[0], lo[1], hi[1], st[1], lo[2], hi[2], st[2]);
    dvmh_loop_map_C(cur_loop, a, 3, dvmh_alignment_linear_C(1, 1, 0), dvmh_alignment_linear_C(2, 1, 0), dvmh_alignment_linear_C(3, 1, 0));
    dvmh_loop_across_C(cur_loop, 0, a, 3, w[0], w[1], w[2], w[3], w[4], w[5]);
    dvmh_loop_register_handler_C(cur_loop, DEVICE_TYPE_HOST, HANDLER_TYPE_PARALLEL, dvmh_handler_func_C((DvmHandlerFunc)hh, 2, a, &c));
    dvmh_loop_register_handler_C(cur_loop, DEVICE_TYPE_CUDA, 0, dvmh_handler_func_C((DvmHandlerFunc)hc, 2, a, &c));
    dvmh_loop_perform_C(cur_loop);
}

double adi3d_iterate(int steps, double max_hist[], long loc_hist[]) {
    DvmType cur_region, cur_loop;
    double mx, startt, endt;
    long loc;
    dvmh_barrier();
    startt = dvmh_wtime();
    for (int stp = 1; stp <= steps; stp++) {
        mx = -1.0;
        loc = -1;
        dvmh_line_C(30, "adi3d.cdv");
        dvmh_actual_variable2_(&mx);
        dvmh_actual_variable2_(&loc);
        cur_region = dvmh_region_create_C(regionFlags);
        dvmh_region_register_array_C(cur_region, INTENT_INOUT, a, "a");
        dvmh_region_register_scalar_C(cur_region, INTENT_IN, &c, sizeof(c), "c");
        dvmh_region_register_scalar_C(cur_region, INTENT_IN, &n, sizeof(n), "n");
        dvmh_region_execute_on_targets_C(cur_region, targets);
        sweep(cur_region, 34, 0, 1, loop_adi3d_34_host, loop_adi3d_34_cuda);
        sweep(cur_region, 39, 0, -1, loop_adi3d_39_host, loop_adi3d_39_cuda);
        sweep(cur_region, 44, 1, 1, loop_adi3d_44_host, loop_adi3d_44_cuda);
        sweep(cur_region, 49, 1, -1, loop_adi3d_49_host, loop_adi3d_49_cuda);
        sweep(cur_region, 54, 2, 1, loop_adi3d_54_host, loop_adi3d_54_cuda);
        sweep(cur_region, 59, 2, -1, loop_adi3d_59_host, loop_adi3d_59_cuda);
        dvmh_line_C(64, "adi3d.cdv");
        cur_loop = dvmh_loop_create_C(cur_region, 3, (DvmType)0, (DvmType)(n - 1), (DvmType)1, (DvmType)0, (DvmType)(n - 1), (DvmType)1,
                (DvmType)0, (DvmType)(n - 1), (DvmType)1);
        dvmh_loop_map_C(cur_loop, a, 3, dvmh_alignment_linear_C(1, 1, 0), dvmh_alignment_linear_C(2, 1, 0),
                dvmh_alignment_linear_C(3, 1, 0));
        dvmh_loop_reduction_C(cur_loop, rf_MAXLOC, &mx, rt_DOUBLE, 1, &loc, sizeof(long) * (1));
        dvmh_loop_register_handler_C(cur_loop, DEVICE_TYPE_HOST, HANDLER_TYPE_PARALLEL,
                dvmh_handler_func_C((DvmHandlerFunc)loop_adi3d_64_host, 2, a, &n));
        dvmh_loop_register_handler_C(cur_loop, DEVICE_TYPE_CUDA, 0, dvmh_handler_func_C((DvmHandlerFunc)loop_adi3d_64_cuda, 2, a, &n));
        dvmh_loop_perform_C(cur_loop);
        dvmh_region_end_C(cur_region);
        dvmh_line_C(72, "adi3d.cdv");
        dvmh_get_actual_variable2_(&mx);
        dvmh_get_actual_variable2_(&loc);
        if (max_hist) max_hist[stp - 1] = mx;
        if (loc_hist) loc_hist[stp - 1] = loc;
    }
    dvmh_barrier();
    endt = dvmh_wtime();
    return endt - startt;
}

void adi3d_fetch(double *out) {
    dvmh_get_actual_array2_(a);
    long total = n * n * n;
    const double *p = (const double *)dvmh_get_natural_base_C(0, a);
    for (long q = 0; q < total; q++) out[q] = p[q];
}

void adi3d_teardown(void) { dvmh_delete_object_(a); }
