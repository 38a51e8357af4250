/*
 * CPU oracle for the DVMH loop path -- TEST INFRASTRUCTURE ONLY (see dvmh_oracle.h).
 * Compiled with -O2 -fopenmp -ffp-contract=off so that every floating-point
 * expression is evaluated exactly as written, in C's left-to-right order.
 */
#include "dvmh_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define Max(a, b) ((a) > (b) ? (a) : (b))
#define GOLDEN 0x9E3779B97F4A7C15ULL

uint64_t orc_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double orc_u01(uint64_t seed, uint64_t n) {
    return (double)(orc_mix64(seed + (n + 1) * GOLDEN) >> 11) * 0x1.0p-53;
}

void orc_init_grid2d(double *A, long n, uint64_t seed, double bnd) {
    long i, j;
#pragma omp parallel for private(j) schedule(static)
    for (i = 0; i < n; i++)
        for (j = 0; j < n; j++)
            A[i * n + j] = (i == 0 || j == 0 || i == n - 1 || j == n - 1) ? bnd : orc_u01(seed, (uint64_t)(i * n + j));
}

void orc_init_grid3d(double *A, long n, uint64_t seed, double bnd, int interior_only) {
    long i, j, k;
#pragma omp parallel for private(j, k) schedule(static)
    for (i = 0; i < n; i++)
        for (j = 0; j < n; j++)
            for (k = 0; k < n; k++) {
                long idx = (i * n + j) * n + k;
                int edge = i == 0 || j == 0 || k == 0 || i == n - 1 || j == n - 1 || k == n - 1;
                A[idx] = (interior_only && edge) ? bnd : orc_u01(seed, (uint64_t)idx);
            }
}

void orc_batched_sides(long *sides, int count, uint64_t seed) {
    for (int g = 0; g < count; g++)
        sides[g] = 64 + (long)(orc_u01(seed, (uint64_t)g) * 961.0);
}

uint64_t orc_batched_grid_seed(uint64_t seed, int g) {
    return orc_mix64(seed + (uint64_t)g + 1);
}

static int pick_threads(int threads) {
#ifdef _OPENMP
    return threads > 0 ? threads : omp_get_max_threads();
#else
    (void)threads;
    return 1;
#endif
}

/* ------------------------------------------------------------------ C1 */
/* Host handler shape (converter_main.cpp:1650-1822): omp parallel, per-thread
 * red_init, "omp for" over the first loop axis, red_post (MAX: order-free). */
void orc_jac2d(double *A, double *B, long n, int iters, double *eps_hist, int threads) {
    int nt = pick_threads(threads);
    for (int it = 0; it < iters; it++) {
        double eps = 0.0; /* eps = 0 before the region, as in the program */
        long i, j;
#pragma omp parallel for private(j) schedule(static) num_threads(nt)
        for (i = 1; i <= n - 2; i++)
            for (j = 1; j <= n - 2; j++)
                B[i * n + j] = 0.25 * (((A[(i - 1) * n + j] + A[(i + 1) * n + j]) + A[i * n + j - 1]) + A[i * n + j + 1]);
#pragma omp parallel num_threads(nt)
        {
            double leps = eps; /* red_init: MAX starts from the variable's value */
            long ii, jj;
#pragma omp for schedule(static) nowait
            for (ii = 1; ii <= n - 2; ii++)
                for (jj = 1; jj <= n - 2; jj++) {
                    double d = fabs(B[ii * n + jj] - A[ii * n + jj]);
                    leps = Max(d, leps);
                    A[ii * n + jj] = B[ii * n + jj];
                }
#pragma omp critical
            { if (eps < leps) eps = leps; } /* red_post: dest < src => replace */
        }
        if (eps_hist) eps_hist[it] = eps;
    }
}

/* ------------------------------------------------------------------ C2 */
static inline void neumaier_add(double *s, double *c, double v) {
    double t = *s + v;
    if (fabs(*s) >= fabs(v)) *c += (*s - t) + v;
    else *c += (v - t) + *s;
    *s = t;
}

static void jac3d_step(const double *src, double *dst, long n, double *eps_out, double *sum_out, int nt) {
    double eps = 0.0, sum = 0.0, comp = 0.0;
    long nn = n * n;
#pragma omp parallel num_threads(nt)
    {
        double leps = 0.0, ls = 0.0, lc = 0.0;
        long i, j, k;
#pragma omp for schedule(static) nowait
        for (i = 1; i <= n - 2; i++)
            for (j = 1; j <= n - 2; j++)
                for (k = 1; k <= n - 2; k++) {
                    long c = i * nn + j * n + k;
                    dst[c] = ((((((src[c - nn] + src[c + nn]) + src[c - n]) + src[c + n]) + src[c - 1]) + src[c + 1]) + src[c]) / 7.0;
                    leps = Max(leps, fabs(dst[c] - src[c]));
                    neumaier_add(&ls, &lc, dst[c]);
                }
#pragma omp for ordered schedule(static, 1)
        for (int t = 0; t < omp_get_num_threads(); t++) {
#pragma omp ordered
            {
                if (eps < leps) eps = leps;
                neumaier_add(&sum, &comp, ls);
                neumaier_add(&sum, &comp, lc);
            }
        }
    }
    *eps_out = eps;
    *sum_out = sum + comp;
}

void orc_jac3d(double *A, double *B, long n, int iters, double *eps_hist, double *sum_hist, int threads) {
    int nt = pick_threads(threads);
    for (int it = 1; it <= iters; it++) {
        double eps, s;
        if (it % 2 == 1) jac3d_step(A, B, n, &eps, &s, nt);
        else jac3d_step(B, A, n, &eps, &s, nt);
        if (eps_hist) eps_hist[it - 1] = eps;
        if (sum_hist) sum_hist[it - 1] = s;
    }
}

/* ------------------------------------------------------------------ C3 */
/* ACROSS with every loop axis dependent: the pipelined handshake of
 * converter_main.cpp:1710-1799.  Thread t owns a static chunk of j; for each i
 * it waits until thread t-1 finished row i, sweeps its chunk, then signals t+1. */
void orc_sor3d(double *A, long n, int iters, double w, double *eps_hist, int threads) {
    int nt = pick_threads(threads);
    long nn = n * n;
    int *slotReady = (int *)calloc((size_t)nt, sizeof(int));
    for (int it = 0; it < iters; it++) {
        double eps = 0.0;
#pragma omp parallel num_threads(nt)
        {
            double leps = eps;
            int cur = 0, working = nt;
            long i, j, k;
#ifdef _OPENMP
            cur = omp_get_thread_num();
            working = omp_get_num_threads();
#endif
            long outerIters = (n - 2 - 1) / 1 + 1;
            if (outerIters < working) working = (int)outerIters;
            __atomic_store_n(&slotReady[cur], 0, __ATOMIC_SEQ_CST);
#pragma omp barrier
            for (i = 1; i <= n - 2; i++) {
                if (cur > 0 && cur < working) {
                    while (!__atomic_load_n(&slotReady[cur - 1], __ATOMIC_ACQUIRE)) {}
                    __atomic_store_n(&slotReady[cur - 1], 0, __ATOMIC_RELEASE);
                }
#pragma omp for schedule(static) nowait
                for (j = 1; j <= n - 2; j++)
                    for (k = 1; k <= n - 2; k++) {
                        long c = i * nn + j * n + k;
                        double s = A[c];
                        A[c] = (w / 6.0) * (((((A[c - nn] + A[c + nn]) + A[c - n]) + A[c + n]) + A[c - 1]) + A[c + 1]) + (1.0 - w) * s;
                        leps = Max(leps, fabs(A[c] - s));
                    }
                if (cur < working - 1) {
                    while (__atomic_load_n(&slotReady[cur], __ATOMIC_ACQUIRE)) {}
                    __atomic_store_n(&slotReady[cur], 1, __ATOMIC_RELEASE);
                }
            }
#pragma omp critical
            { if (eps < leps) eps = leps; }
        }
        if (eps_hist) eps_hist[it] = eps;
    }
    free(slotReady);
}

/* ------------------------------------------------------------------ C4 */
/* One dependent axis: "omp for" goes on the first independent loop axis
 * (converter_main.cpp:1690-1708); the dependent axis runs sequentially. */
static void adi_sweep(double *a, long n, int axis, int forward, double c, int nt) {
    long nn = n * n;
    long i, j, k;
    if (axis == 0) {
        /* for (i = ...) for (j = 0..n-1) for (k = 0..n-1): omp for on j */
#pragma omp parallel num_threads(nt) private(i, j, k)
        {
            if (forward) {
                for (i = 1; i <= n - 1; i++) {
#pragma omp for schedule(static) nowait
                    for (j = 0; j <= n - 1; j++)
                        for (k = 0; k <= n - 1; k++)
                            a[i * nn + j * n + k] = a[i * nn + j * n + k] + c * a[(i - 1) * nn + j * n + k];
                }
            } else {
                for (i = n - 2; i >= 0; i--) {
#pragma omp for schedule(static) nowait
                    for (j = 0; j <= n - 1; j++)
                        for (k = 0; k <= n - 1; k++)
                            a[i * nn + j * n + k] = a[i * nn + j * n + k] + c * a[(i + 1) * nn + j * n + k];
                }
            }
        }
    } else if (axis == 1) {
#pragma omp parallel for num_threads(nt) private(j, k) schedule(static)
        for (i = 0; i <= n - 1; i++) {
            if (forward) {
                for (j = 1; j <= n - 1; j++)
                    for (k = 0; k <= n - 1; k++)
                        a[i * nn + j * n + k] = a[i * nn + j * n + k] + c * a[i * nn + (j - 1) * n + k];
            } else {
                for (j = n - 2; j >= 0; j--)
                    for (k = 0; k <= n - 1; k++)
                        a[i * nn + j * n + k] = a[i * nn + j * n + k] + c * a[i * nn + (j + 1) * n + k];
            }
        }
    } else {
#pragma omp parallel for num_threads(nt) private(j, k) schedule(static)
        for (i = 0; i <= n - 1; i++) {
            for (j = 0; j <= n - 1; j++) {
                if (forward) {
                    for (k = 1; k <= n - 1; k++)
                        a[i * nn + j * n + k] = a[i * nn + j * n + k] + c * a[i * nn + j * n + k - 1];
                } else {
                    for (k = n - 2; k >= 0; k--)
                        a[i * nn + j * n + k] = a[i * nn + j * n + k] + c * a[i * nn + j * n + k + 1];
                }
            }
        }
    }
}

void orc_adi3d(double *A, long n, int steps, double c, double *max_hist, long *loc_hist, int threads) {
    int nt = pick_threads(threads);
    long nn = n * n;
    for (int st = 0; st < steps; st++) {
        for (int axis = 0; axis < 3; axis++) {
            adi_sweep(A, n, axis, 1, c, nt);
            adi_sweep(A, n, axis, 0, c, nt);
        }
        double mx = -1.0;
        long loc = -1;
#pragma omp parallel num_threads(nt)
        {
            /* per-thread partials start from the reduction's initial value, never
             * from the shared result another thread may already have merged */
            double lmx = -1.0;
            long lloc = -1;
            long i, j, k;
#pragma omp for schedule(static) nowait
            for (i = 0; i <= n - 1; i++)
                for (j = 0; j <= n - 1; j++)
                    for (k = 0; k <= n - 1; k++)
                        if (fabs(A[i * nn + j * n + k]) > lmx) {
                            lmx = fabs(A[i * nn + j * n + k]);
                            lloc = i * nn + j * n + k;
                        }
#pragma omp critical
            {
                /* Deterministic tie rule: larger value, then lower linear index
                 * (what the sequential strict-'>' body yields). */
                if (mx < lmx || (mx == lmx && lloc >= 0 && (loc < 0 || lloc < loc))) {
                    mx = lmx;
                    loc = lloc;
                }
            }
        }
        if (max_hist) max_hist[st] = mx;
        if (loc_hist) loc_hist[st] = loc;
    }
}

/* ---------------------------------------------------- reference KATs */
/* examples/Adi/adi.c: init loop, then per iteration three ACROSS sweeps with
 * one dependency each; the third computes eps = Max(eps, |a - tmp1|). */
double orc_adi_kat(long n, int iters, int threads, double *eps_hist) {
    int nt = pick_threads(threads);
    long nn = n * n;
    double *a = (double *)malloc(sizeof(double) * (size_t)(n * nn));
    double eps = 0.0;
    long i, j, k;
#pragma omp parallel for num_threads(nt) private(j, k) schedule(static)
    for (i = 0; i < n; i++)
        for (j = 0; j < n; j++)
            for (k = 0; k < n; k++)
                if (k == 0 || k == n - 1 || j == 0 || j == n - 1 || i == 0 || i == n - 1)
                    a[i * nn + j * n + k] = 10.0 * i / (n - 1) + 10.0 * j / (n - 1) + 10.0 * k / (n - 1);
                else
                    a[i * nn + j * n + k] = 0;
    for (int it = 1; it <= iters; it++) {
        eps = 0;
#pragma omp parallel num_threads(nt) private(i, j, k)
        {
            for (i = 1; i < n - 1; i++) {
#pragma omp for schedule(static) nowait
                for (j = 1; j < n - 1; j++)
                    for (k = 1; k < n - 1; k++)
                        a[i * nn + j * n + k] = (a[(i - 1) * nn + j * n + k] + a[(i + 1) * nn + j * n + k]) / 2;
            }
        }
#pragma omp parallel for num_threads(nt) private(j, k) schedule(static)
        for (i = 1; i < n - 1; i++)
            for (j = 1; j < n - 1; j++)
                for (k = 1; k < n - 1; k++)
                    a[i * nn + j * n + k] = (a[i * nn + (j - 1) * n + k] + a[i * nn + (j + 1) * n + k]) / 2;
#pragma omp parallel num_threads(nt)
        {
            double leps = eps;
            long ii, jj, kk;
#pragma omp for schedule(static) nowait
            for (ii = 1; ii < n - 1; ii++)
                for (jj = 1; jj < n - 1; jj++)
                    for (kk = 1; kk < n - 1; kk++) {
                        double tmp1 = (a[ii * nn + jj * n + kk - 1] + a[ii * nn + jj * n + kk + 1]) / 2;
                        double tmp2 = fabs(a[ii * nn + jj * n + kk] - tmp1);
                        leps = Max(leps, tmp2);
                        a[ii * nn + jj * n + kk] = tmp1;
                    }
#pragma omp critical
            { if (eps < leps) eps = leps; }
        }
        if (eps_hist) eps_hist[it - 1] = eps;
    }
    free(a);
    return eps;
}

/* examples/Performance/jac3d_perf.c (float). */
float orc_jac3d_f32_kat(long n, int iters, int threads, float *eps_hist) {
    int nt = pick_threads(threads);
    long nn = n * n;
    float *A = (float *)malloc(sizeof(float) * (size_t)(n * nn));
    float *B = (float *)malloc(sizeof(float) * (size_t)(n * nn));
    float eps = 0;
    long i, j, k;
#pragma omp parallel for num_threads(nt) private(j, k) schedule(static)
    for (i = 0; i < n; i++)
        for (j = 0; j < n; j++)
            for (k = 0; k < n; k++) {
                A[i * nn + j * n + k] = 0;
                if (i == 0 || j == 0 || k == 0 || i == n - 1 || j == n - 1 || k == n - 1)
                    B[i * nn + j * n + k] = 0;
                else
                    B[i * nn + j * n + k] = 4 + i + j + k;
            }
    for (int it = 1; it <= iters; it++) {
        eps = 0;
#pragma omp parallel num_threads(nt)
        {
            float leps = eps;
            long ii, jj, kk;
#pragma omp for schedule(static) nowait
            for (ii = 1; ii < n - 1; ii++)
                for (jj = 1; jj < n - 1; jj++)
                    for (kk = 1; kk < n - 1; kk++) {
                        float tmp = fabs(B[ii * nn + jj * n + kk] - A[ii * nn + jj * n + kk]);
                        leps = Max(tmp, leps);
                        A[ii * nn + jj * n + kk] = B[ii * nn + jj * n + kk];
                    }
#pragma omp critical
            { if (eps < leps) eps = leps; }
        }
#pragma omp parallel for num_threads(nt) private(j, k) schedule(static)
        for (i = 1; i < n - 1; i++)
            for (j = 1; j < n - 1; j++)
                for (k = 1; k < n - 1; k++)
                    B[i * nn + j * n + k] = (A[(i - 1) * nn + j * n + k] + A[i * nn + (j - 1) * n + k] + A[i * nn + j * n + k - 1] +
                                             A[i * nn + j * n + k + 1] + A[i * nn + (j + 1) * n + k] + A[(i + 1) * nn + j * n + k]) / 6.0f;
        if (eps_hist) eps_hist[it - 1] = eps;
    }
    free(A);
    free(B);
    return eps;
}

/* examples/Performance/jac2d_perf.c (float). */
float orc_jac2d_f32_kat(long n, int iters, int threads, float *eps_hist) {
    int nt = pick_threads(threads);
    float *A = (float *)malloc(sizeof(float) * (size_t)(n * n));
    float *B = (float *)malloc(sizeof(float) * (size_t)(n * n));
    float eps = 0;
    long i, j;
#pragma omp parallel for num_threads(nt) private(j) schedule(static)
    for (i = 0; i < n; i++)
        for (j = 0; j < n; j++) {
            A[i * n + j] = 0;
            if (i == 0 || j == 0 || i == n - 1 || j == n - 1)
                B[i * n + j] = 0;
            else
                B[i * n + j] = 3 + i + j;
        }
    for (int it = 1; it <= iters; it++) {
        eps = 0;
#pragma omp parallel num_threads(nt)
        {
            float leps = eps;
            long ii, jj;
#pragma omp for schedule(static) nowait
            for (ii = 1; ii < n - 1; ii++)
                for (jj = 1; jj < n - 1; jj++) {
                    float tmp = fabs(B[ii * n + jj] - A[ii * n + jj]);
                    leps = Max(tmp, leps);
                    A[ii * n + jj] = B[ii * n + jj];
                }
#pragma omp critical
            { if (eps < leps) eps = leps; }
        }
#pragma omp parallel for num_threads(nt) private(j) schedule(static)
        for (i = 1; i < n - 1; i++)
            for (j = 1; j < n - 1; j++)
                B[i * n + j] = (A[(i - 1) * n + j] + A[i * n + j - 1] + A[i * n + j + 1] + A[(i + 1) * n + j]) / 4.0f;
        if (eps_hist) eps_hist[it - 1] = eps;
    }
    free(A);
    free(B);
    return eps;
}

/* ------------------------------------------------------ integer ABI */
long orc_convert_xy(long x, long y, long Rx, long Ry, int slash) {
    long idx;
    if (!slash) {
        long d = x + y;
        if (Rx == Ry) {
            if (d < Rx) idx = y + (1 + d) * d / 2;
            else idx = Rx * (Rx - 1) + x - (2 * Rx - d - 1) * (2 * Rx - d - 2) / 2;
        } else if (Rx < Ry) {
            if (d < Rx) idx = y + (1 + d) * d / 2;
            else if (d < Ry) idx = (1 + Rx) * Rx / 2 + Rx - x - 1 + Rx * (d - Rx);
            else idx = Rx * Ry - Ry + y - (Rx + Ry - d - 1) * (Rx + Ry - d - 2) / 2;
        } else {
            if (d < Ry) idx = x + (1 + d) * d / 2;
            else if (d < Rx) idx = (1 + Ry) * Ry / 2 + (Ry - y - 1) + Ry * (d - Ry);
            else idx = Rx * Ry - Rx + x - (Rx + Ry - d - 1) * (Rx + Ry - d - 2) / 2;
        }
    } else {
        /* anti-slash: diagonals run along x - y; e = x + (R - 1 - y) */
        if (Rx == Ry) {
            long e = x + Rx - 1 - y;
            if (e < Rx) idx = Rx - 1 - y + (e + 1) * e / 2;
            else idx = Rx * (Rx - 1) + x - (Rx - x + y) * (Rx - x + y - 1) / 2;
        } else if (Rx < Ry) {
            long e = x + Ry - 1 - y;
            if (e < Rx) idx = Ry - 1 - y + (e + 1) * e / 2;
            else if (e < Ry) idx = (1 + Rx) * Rx / 2 + Rx - x - 1 + Rx * (e - Rx);
            else idx = Rx * Ry - 1 - y - (Rx + y - x) * (Rx + y - x - 1) / 2;
        } else {
            long e = x + Ry - 1 - y;
            if (e < Ry) idx = x + (1 + e) * e / 2;
            else if (e < Rx) idx = (1 + Ry) * Ry / 2 + y + Ry * (x - y - 1);
            else idx = Rx * Ry - Rx + x - (Rx + y - x) * (Rx + y - x - 1) / 2;
        }
    }
    return idx;
}

void orc_change_filled_bounds2(long *lo, long *hi, long *st, long rank, long depMask, long *idxPerm) {
    int depCount = 0;
    for (long m = depMask; m; m &= m - 1) depCount++;
    int h = 0, hd = depCount;
    for (long i = 1; i <= rank; i++) {
        if ((depMask >> (rank - i)) & 1) idxPerm[h++] = i;
        else idxPerm[hd++] = i;
    }
    if (depCount > 0 && depCount < rank) {
        long t0[16], t1[16], t2[16];
        for (long i = 0; i < rank; i++) {
            t0[i] = lo[idxPerm[i] - 1];
            t1[i] = hi[idxPerm[i] - 1];
            t2[i] = st[idxPerm[i] - 1];
        }
        memcpy(lo, t0, sizeof(long) * (size_t)rank);
        memcpy(hi, t1, sizeof(long) * (size_t)rank);
        memcpy(st, t2, sizeof(long) * (size_t)rank);
    }
    for (long i = 0; i < rank; i++) idxPerm[i]--;
}
