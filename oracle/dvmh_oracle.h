/*
 * CPU oracle for the DVMH loop path.  TEST INFRASTRUCTURE ONLY: the product
 * (libdvmh_amd and the HIP handlers) never links, loads or calls this code.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * and only as the checker / the timed CPU baseline.
 *
 * It restates, in plain C + OpenMP compiled with -ffp-contract=off, what the
 * reference's cdvmh host handlers compute for the frozen program texts of the
 * five workloads and of the reference's own known-answer examples:
 *   - handler shape: cdvmh/src/converter_main.cpp:1503-1859 (genHostHandler):
 *     "#pragma omp for schedule(runtime)" on the first independent loop axis,
 *     loop body verbatim in loop-variable order, per-thread red_init/red_post;
 *     ACROSS loops with every axis dependent use the pipelined handshake of
 *     converter_main.cpp:1710-1799.
 *   - reduction semantics: rts/dvmh/loop.cpp:17-254 (identity init for SUM,
 *     initial-value init for MAX/MIN/MAXLOC, strict-compare replacement,
 *     final fold with the pre-loop value, rts/dvmh/dvmh2.cpp:3950-3953).
 *   - index math: rts/dvmh/include/dvmhlib_device.h:5-98 (diagonal layout),
 *     rts/dvmh/dvmh2.cpp:4556-4583 (change_filled_bounds2).
 *
 * Parity status of each function is stated next to its declaration.
 */
#ifndef DVMH_ORACLE_H
#define DVMH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- seeded portable generator (SplitMix64 finaliser, counter based) ----
 * value(seed, n) = (mix64(seed + (n + 1) * 0x9E3779B97F4A7C15) >> 11) * 2^-53
 * Pinned: tests/test_oracle.py checks the published SplitMix64 vectors. */
uint64_t orc_mix64(uint64_t z);
double orc_u01(uint64_t seed, uint64_t n);

/* ---- input generators (identical to the programs' host-side init) ---- */
/* 2-D grid n x n: boundary = bnd, interior u01(seed, i*n+j). */
void orc_init_grid2d(double *A, long n, uint64_t seed, double bnd);
/* 3-D grid n^3: boundary = bnd, interior u01(seed, (i*n+j)*n+k).
 * interior_only = 0 fills every point with u01 (ADI workload). */
void orc_init_grid3d(double *A, long n, uint64_t seed, double bnd, int interior_only);
/* Side lengths of the batched workload: 64 + floor(u01(seed, g) * 961). */
void orc_batched_sides(long *sides, int count, uint64_t seed);
/* Per-grid seed of the batched workload. */
uint64_t orc_batched_grid_seed(uint64_t seed, int g);

/* ---- workloads (frozen program texts, see dvm-snapshot-r6603-fork_amd/programs/<name>.cdv) ---- */

/* 2-D Jacobi, 2 loops per iteration:
 *   B[i][j] = 0.25 * (((A[i-1][j] + A[i+1][j]) + A[i][j-1]) + A[i][j+1]);
 *   { double d = fabs(B[i][j] - A[i][j]); eps = Max(d, eps); A[i][j] = B[i][j]; }
 * eps_hist[it] = eps after iteration it.  Pinned by harmonic-invariance and
 * separable-eigenmode tests (tests/test_oracle.py). */
void orc_jac2d(double *A, double *B, long n, int iters, double *eps_hist, int threads);

/* 3-D 7-point Jacobi, A->B on odd iterations, B->A on even ones:
 *   dst = ((((((src[i-1]+src[i+1])+src[j-1])+src[j+1])+src[k-1])+src[k+1])+src[c]) / 7.0;
 *   eps = Max(eps, fabs(dst - src));  s = s + dst;
 * SUM is accumulated with per-thread Neumaier compensation, merged in thread order.
 * Pinned by harmonic invariance, eigenmode decay and closed-form sums. */
void orc_jac3d(double *A, double *B, long n, int iters, double *eps_hist, double *sum_hist, int threads);

/* 3-D lexicographic SOR in place, ACROSS(A[1:1][1:1][1:1]):
 *   s = A[i][j][k];
 *   A[i][j][k] = (w / 6.0) * (((((A[i-1][j][k] + A[i+1][j][k]) + A[i][j-1][k]) + A[i][j+1][k])
 *                + A[i][j][k-1]) + A[i][j][k+1]) + (1.0 - w) * s;
 *   eps = Max(eps, fabs(A[i][j][k] - s));
 * Run with the pipelined OpenMP handshake (threads = 1: plain sequential order).
 * Pinned against the matrix form (D + wL) x' = ... solved by scipy in tests. */
void orc_sor3d(double *A, long n, int iters, double w, double *eps_hist, int threads);

/* ADI line recurrences, per step for axis = i, j, k:
 *   forward  (index 1..n-1 ascending):  a[..x..] = a[..x..] + c * a[..x-1..];
 *   backward (index n-2..0 descending): a[..x..] = a[..x..] + c * a[..x+1..];
 * then MAXLOC over all points: if (fabs(a) > mx) { mx = fabs(a); loc = linear index; }
 * with mx starting at -1.0 and loc at -1 (first occurrence wins on ties).
 * Pinned against closed-form bidiagonal solves and numpy argmax. */
void orc_adi3d(double *A, long n, int steps, double c, double *max_hist, long *loc_hist, int threads);

/* ---- the reference's own known-answer programs ----
 * examples/Adi/adi.c (FP64, golden eps 0.07249074 +- 1e-6 at n=384, 100 its). */
double orc_adi_kat(long n, int iters, int threads, double *eps_hist);
/* examples/Performance/jac3d_perf.c (FP32, golden eps 5.058044 +- 1e-4 at 384, 100 its). */
float orc_jac3d_f32_kat(long n, int iters, int threads, float *eps_hist);
/* examples/Performance/jac2d_perf.c (FP32, golden eps 58.37598 +- 1e-3 at 8000, 100 its). */
float orc_jac2d_f32_kat(long n, int iters, int threads, float *eps_hist);

/* ---- integer ABI functions ---- */
/* Diagonal position of (x, y) in an Rx x Ry plane, rts/dvmh/include/dvmhlib_device.h:56-98.
 * Pinned by brute-force bijection tests. */
long orc_convert_xy(long x, long y, long Rx, long Ry, int slash);
/* rts/dvmh/dvmh2.cpp:4556-4583: dependent axes first, 0-based idxPerm out. */
void orc_change_filled_bounds2(long *lo, long *hi, long *st, long rank, long depMask, long *idxPerm);

#ifdef __cplusplus
}
#endif
#endif
