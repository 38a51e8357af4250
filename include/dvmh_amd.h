/*
 * dvmh_amd.h -- entry points of libdvmh_amd that have no counterpart in the
 * reference interface.  Emitted HIP handlers use the tuning query; programs
 * and harnesses use batching, statistics and the tuner/cost-model controls.
 * Everything here is extern "C" with plain pointers and integers.
 */
#ifndef DVMH_AMD_H
#define DVMH_AMD_H

#include "dvmhlib_types.h"

#pragma GCC visibility push(default)
#ifdef __cplusplus
extern "C" {
#endif

/* ---- launch tuning beyond the thread block --------------------------------
 * An emitted handler first calls dvmh_loop_cuda_get_config_C (block, stream),
 * then this, naming the kind of its kernel family.  The runtime answers with
 * the configuration chosen by its tuner (default, exhaustive sweep, or the
 * learned cost model); the answer never changes arithmetic, only the mapping
 * of iterations to threads. */
#define DVMH_AMD_TUNE_N 8
#define DVMH_AMD_KIND_PARALLEL 0   /* independent iterations, one kernel */
#define DVMH_AMD_KIND_LINE 1       /* ACROSS with one dependent axis: line sweeps
                                      (VARIANT 1, dependent axis innermost: lines staged in
                                      LDS chunks of TILE_A steps; VARIANT 2, other axes: each
                                      thread's line in register chunks of TILE_A steps;
                                      VARIANT 3, dependent axis innermost and unit stride:
                                      a wave streams 64 lines in aligned 128-byte chunks,
                                      coalesced, transposed through LDS, a lane per line) */
#define DVMH_AMD_KIND_WAVEFRONT 2  /* ACROSS with >= 2 dependent axes: hyperplane wavefront */
#define DVMH_AMD_KIND_STENCIL 3    /* PARALLEL whose handler also has the LDS march variant
                                      (VARIANT 1: TILE_A = points per thread on the innermost
                                      axis, CHUNK = planes per workgroup on the outermost) */
#define DVMH_AMD_KIND_BATCH 4      /* PARALLEL loops of many instances run by one batch kernel */
#define DVMH_AMD_KIND_CHAIN 5      /* consecutive sweeps of a tiled-wavefront ACROSS loop run by its chain handler */
#define DVMH_AMD_KIND_PAIR 6       /* a stencil loop and the loop after it run fused by their pair handler
                                      (VARIANT 1, TILE_A and CHUNK as for STENCIL) */
/* tune[] slots */
#define DVMH_AMD_T_DEPTH 0     /* iterations of the slowest mapped axis per thread */
#define DVMH_AMD_T_XCD 1       /* 1: XCD-aware block-to-tile order */
#define DVMH_AMD_T_VARIANT 2   /* kernel variant (0 = direct global-memory body) */
#define DVMH_AMD_T_CHUNK 3     /* LINE: steps per LDS window; WAVEFRONT: tile depth */
#define DVMH_AMD_T_TILE_A 4    /* WAVEFRONT: tile extent on the first dependent axis; LINE: chunk steps */
#define DVMH_AMD_T_TILE_B 5    /* WAVEFRONT: tile extent on the second dependent axis */
#define DVMH_AMD_T_TRANSFORM 6 /* 1: arrays may be re-laid out for this loop */
#define DVMH_AMD_T_OWN 7       /* BATCH: 1 = a batch of stencil loops runs its handler's own LDS march shape;
                                  PAIR: 1 = the pair handler picks its march shape from the pairs' sizes */
void dvmh_amd_loop_get_tuning_C(DvmType curLoop, DvmType kind, DvmType tune[]);

/* ---- kernel timing -----------------------------------------------------------
 * Emitted handlers bracket their main kernel launches with phase 0 / phase 1
 * marks.  With timing enabled the runtime records HIP events on the loop's
 * stream at the marks and accumulates, per loop site (file, line), the number
 * of bracketed launches and their total duration.  enable = 1 times every
 * launch; enable = k > 1 times the first of every k launches of a site (the
 * events cost host time and a few microseconds of stream time each). */
void dvmh_amd_mark_kernel_C(DvmType curLoop, DvmType phase);
void dvmh_amd_set_kernel_timing_C(DvmType enable);
/* out[0] launches, out[1] total nanoseconds; returns 0 if the site is unknown. */
DvmType dvmh_amd_kernel_stats_C(const char *file, DvmType line, DvmType out[2]);
void dvmh_amd_reset_kernel_stats_C(void);

/* ---- execution options --------------------------------------------------------
 * sync mode 0: synchronise after every loop (the reference's behaviour);
 * 1: only when a result is needed (get_actual, region end, barrier). */
void dvmh_amd_set_sync_mode_C(DvmType mode);
/* 1: regions created with REGION_AMD_GRAPH are run lazily and replayed as graphs
 * (same as DVMH_AMD_GRAPH=1); 0: every region runs eagerly; 2: such regions
 * still run lazily (their loops at the region's end, consecutive pairs fused)
 * but nothing is captured, so every launch can carry timing marks. */
void dvmh_amd_set_graph_C(DvmType enable);
/* A lazy region is captured once; its graph is replayed only when it has at
 * least `nodes` nodes (kernels and copies; default 6, DVMH_AMD_GRAPH_MIN_NODES):
 * smaller steps run faster launched one by one.  Negative: the default. */
void dvmh_amd_set_graph_min_nodes_C(DvmType nodes);

/* ---- statistics ------------------------------------------------------------
 * out[0] loops run on the host, out[1] loops run on the GPU,
 * out[2] GPU loop time in nanoseconds (stream time of the executions the
 * tuner timed while exploring candidates), out[3] GPU present (0/1), out[4] compute units,
 * out[5] graph replays, out[6] batched groups, out[7] batched loops. */
void dvmh_amd_get_stats_C(DvmType out[8]);
void dvmh_amd_reset_stats_C(void);
/* Device layout changes (axis permutations and diagonal re-layouts of arrays
 * done for ACROSS loops, and their undoing) since the runtime started. */
DvmType dvmh_amd_layout_changes_C(void);
/* Host threads the host handlers of a loop run on (DVMH_NUM_THREADS, else the
 * OpenMP default). */
DvmType dvmh_amd_host_slots_C(void);
/* Bytes of array and variable data moved host -> device (out[0]) and device ->
 * host (out[1]) by actualisation since start-up or the last
 * dvmh_amd_reset_stats_C. */
void dvmh_amd_transfer_stats_C(DvmType out[2]);
/* GPU reduction results computed but not yet merged into the program's
 * variables (they merge at the next synchronisation point: the loop's end by
 * default, get_actual / actual in an async region or with DVMH_AMD_SYNC=1). */
DvmType dvmh_amd_pending_posts_C(void);
/* Honour REGION_ASYNC (as DVMH_ALLOW_ASYNC does at start-up, rts/dvmh/settings.cpp:361):
 * loops of regions created with it afterwards end without a synchronisation. */
void dvmh_amd_set_allow_async_C(DvmType allow);
/* The box-set algebra of the actual-state tracking, for inspection: applies
 * nops operations (records of 1 + 2 * rank values: 1 add / 2 remove, lo[rank],
 * hi[rank], inclusive) to an empty set and writes the parts of the box query
 * (lo[rank], hi[rank]) the set misses to outBoxes (at most maxBoxes records of
 * 2 * rank values); returns how many there are. */
DvmType dvmh_amd_pieces_eval_C(DvmType rank, const DvmType *ops, DvmType nops, const DvmType *query, DvmType *outBoxes,
        DvmType maxBoxes);
/* on != 0: every loop's handlers get rt_LONG from dvmh_loop_guess_index_type_C
 * and run their 64-bit index kernels whatever the array sizes (also
 * DVMH_AMD_INDEX64=1); 0 restores the choice by extent. */
void dvmh_amd_set_index64_C(DvmType on);
/* Seconds per launch of the runtime's streaming kernels on a device buffer of
 * `bytes` (allocated for the call): a copy that reads and writes every byte
 * once, and a read-only pass; timed with HIP events over `reps` launches.
 * They are the measured ceiling of a read-once/write-once kernel and the
 * known-traffic dispatches that scale HBM counters.  -1 without a GPU. */
double dvmh_amd_stream_copy_seconds_C(DvmType bytes, DvmType reps);
double dvmh_amd_stream_read_seconds_C(DvmType bytes, DvmType reps);
/* The layout-transform kernels timed on an FP64 n0 x n1 x n2 array (buffers
 * allocated for the measurement): which 0 swaps the two fastest axes (tiled
 * transpose), 1 converts the (n1, n2) planes to the diagonal layout, 2 rotates
 * the axes (general gather).  out[0] / out[1]: seconds of one forward /
 * backward transform, averaged over `reps` round trips; returns -1 without a GPU. */
DvmType dvmh_amd_layout_transform_seconds_C(DvmType which, DvmType n0, DvmType n1, DvmType n2, DvmType reps, double out[2]);
/* Folds n >= 1 partial values of one reduction, in device memory (vals; for
 * rf_MAXLOC / rf_MINLOC also locs, n records of locSize bytes), into one value
 * and its loc record written to outVal / outLoc (device or mapped memory), on
 * the runtime's stream; returns when the result is written.  func: rf_SUM ..
 * rf_MINLOC; rtType: rt_CHAR .. rt_DOUBLE element type.  The fold a loop's GPU
 * reduction finish uses (replacing rts/dvmh/cuda_reduction.cu:168-312's
 * per-warp results folded on the host): fixed partition and tree, so the
 * result is deterministic; equal MAX/MIN values keep the lower loc (integer
 * records of 1, 2, 4 or 8 bytes), else the lower partial index.  More than
 * 16384 partials fold in two passes through scratch taken from the runtime's
 * device cache.  Returns 0, -1 without a GPU. */
DvmType dvmh_amd_reduce_partials_C(DvmType func, DvmType rtType, DvmType n, const void *vals, const void *locs, DvmType locSize,
        void *outVal, void *outLoc);
/* No counterpart in the reference: the same fold without allocating anything,
 * for callers that cannot call hipMalloc (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation).  Its scratch is a fixed
 * 1 MiB array in device memory, one per GPU, held until the process ends, so
 * it folds at most 715,784,192 partials (returns -2 with a message above
 * that) and calls that overlap on one GPU race.  Same arguments and results
 * as dvmh_amd_reduce_partials_C; -1 without a GPU. */
DvmType dvmh_amd_reduce_partials_noalloc_C(DvmType func, DvmType rtType, DvmType n, const void *vals, const void *locs,
        DvmType locSize, void *outVal, void *outLoc);
/* Device name / architecture string of the GPU in use ("" without one). */
const char *dvmh_amd_device_name_C(void);
/* Position of (x, y) in an Rx x Ry plane stored diagonal by diagonal (slash 0:
 * diagonals of constant x + y; 1: of constant x - y), the index map of the
 * diagonal device layout (rts/dvmh/include/dvmhlib_device.h:56-98).  The same
 * function the layout kernels and the handlers' DiagView use. */
long dvmh_amd_convert_xy_C(long x, long y, long Rx, long Ry, long slash);
/* 1 if a block[0] x block[1] x block[2] workgroup of a kernel using
 * regsPerThread vector registers fits one compute unit's register file (its
 * waves spread over the 4 SIMDs, each with 512 registers per lane, allocated
 * in granules of 8), else 0.  dvmh_loop_cuda_get_config_C applies this rule
 * to the count a handler passes (rts/dvmh/loop.cpp:2177-2194 with ptxas
 * counts): a block that does not fit becomes (min(256, fitting threads), 1, 1),
 * or is an error if the program set it. */
DvmType dvmh_amd_block_fits_registers_C(const DvmType block[3], DvmType regsPerThread);

/* ---- batching of many small loops -------------------------------------------
 * Between begin and end, dvmh_loop_perform_C of loops that registered a batch
 * handler queues them; end groups queued loops by batch handler and calls each
 * once with all loop references, so one fused launch covers many grids. */
void dvmh_amd_batch_begin_C(void);
void dvmh_amd_batch_end_C(void);
void dvmh_amd_loop_register_batch_handler_C(DvmType curLoop, DvmType handlerHelper);
/* A chain handler instead: consecutive queued loops of the site join one group
 * in program order even when they share arrays (ACROSS sweeps of one array),
 * and the handler keeps their dependencies.  A group whose loop site the tuner
 * is still exploring runs loop by loop through the ordinary GPU handler. */
void dvmh_amd_loop_register_chain_handler_C(DvmType curLoop, DvmType handlerHelper);
/* Inside a batch handler (called as f(&batchRef)): number of loops, the i-th
 * loop reference (usable with every handler query), and the p-th parameter the
 * i-th loop registered for its ordinary GPU handler. */
DvmType dvmh_amd_batch_count_C(DvmType batchRef);
DvmType dvmh_amd_batch_loop_C(DvmType batchRef, DvmType i);
void *dvmh_amd_batch_param_C(DvmType batchRef, DvmType i, DvmType p);
/* Copies bytes from host memory to device memory dst on the runtime's stream. */
void dvmh_amd_batch_upload_C(DvmType batchRef, void *dst, const void *src, DvmType bytes);
/* A device copy of `bytes` host bytes that the runtime keeps for the caller's
 * (tag, slot) across executions of a batch handler: uploaded (asynchronously,
 * stream-ordered) only when the bytes differ from the last ones given for that
 * key.  For descriptor tables that repeat from one time step to the next. */
void *dvmh_amd_batch_persist_C(DvmType batchRef, const void *tag, DvmType slot, const void *src, DvmType bytes);
/* Inside a batch handler: finish reduction redIndex of every loop in the batch.
 * partials holds per-loop segments [offsets[i], offsets[i+1]) of values (and
 * locs) in device memory; one kernel folds each segment and the results are
 * posted to each loop's reduction variable. */
void dvmh_amd_batch_red_finish_C(DvmType batchRef, DvmType redIndex, const void *partials, const void *locPartials,
        const DvmType *offsets);
/* Device scratch for batch handlers: bytes of device memory kept by the runtime
 * until the batch handler returns. */
void *dvmh_amd_batch_scratch_C(DvmType batchRef, DvmType bytes);

/* ---- fused loop pairs -----------------------------------------------------------
 * An emitted HIP file registers, when its library is loaded, a pair handler
 * for two consecutive loops of its program that may run as one pass (a
 * stencil loop, then a loop touching every array only at the loop point).
 * When the runtime holds the two loops back together -- the loops of a lazy
 * (graph) region, or of an open batch -- and they follow each other in one
 * region with equal bounds, it calls the pair handler as f(&batchRef) instead
 * of the two loops' handlers; several such pairs of one batch go to one call.
 * Inside, dvmh_amd_batch_count_C is the number of pairs, member 0 is the first
 * loop of a pair and member 1 the second. */
void dvmh_amd_register_pair_C(DvmHandlerFunc first, DvmHandlerFunc second, DvmHandlerFunc pair);
DvmType dvmh_amd_pair_loop_C(DvmType batchRef, DvmType member, DvmType i);
void *dvmh_amd_pair_param_C(DvmType batchRef, DvmType member, DvmType i, DvmType p);
/* The pair handler runs nothing (the pairs do not meet its conditions): each
 * loop then runs through its own handler, in program order, when it returns. */
void dvmh_amd_pair_decline_C(DvmType batchRef);
/* As dvmh_amd_batch_red_finish_C, for reduction redIndex of the member loops. */
void dvmh_amd_pair_red_finish_C(DvmType batchRef, DvmType member, DvmType redIndex, const void *partials,
        const void *locPartials, const DvmType *offsets);
/* 1 (default, DVMH_AMD_FUSE): registered pairs run fused; 0: loop by loop. */
void dvmh_amd_set_fuse_C(DvmType enable);
/* out[0] pair-handler calls, out[1] loop pairs they ran. */
void dvmh_amd_pair_stats_C(DvmType out[2]);

/* ---- region graphs -----------------------------------------------------------
 * A region created with this flag bit is executed lazily: its loops are recorded
 * and launched at dvmh_region_end_C; the first repetition with identical
 * launches is captured into a hipGraph and later repetitions replay it. */
#define REGION_AMD_GRAPH 256

/* ---- tuner / learned cost model ---------------------------------------------
 * mode: 0 default configuration, 1 exhaustive sweep then best, 2 cost model
 * (falls back to the sweep for a loop the model has not seen), 3 fixed best
 * from the records.  Records: one line per (loop signature, candidate, seconds). */
void dvmh_amd_tuner_set_mode_C(DvmType mode);
/* Timed executions per candidate before the sweep moves on (the site's time
 * for a candidate is the minimum of its repeats); DVMH_AMD_TUNE_REPEATS. */
void dvmh_amd_tuner_set_repeats_C(DvmType repeats);
DvmType dvmh_amd_tuner_save_C(const char *path);
DvmType dvmh_amd_tuner_load_C(const char *path);
void dvmh_amd_tuner_clear_C(void);
/* Candidates still to be timed at the sites seen so far in the current mode
 * (every candidate for the exhaustive sweep; the default and the model's best
 * four for the cost model; none for modes 0 and 3). */
DvmType dvmh_amd_tuner_pending_C(void);
/* The same for the sites of one source loop (file, line). */
DvmType dvmh_amd_tuner_site_pending_C(const char *file, DvmType line);
/* Best candidate so far for the loop site (file, line): out[0] candidate index,
 * out[1] time in ns, out[2..4] block x,y,z, out[5..12] tune[0..7]; returns 0
 * if unknown. */
DvmType dvmh_amd_tuner_query_C(const char *file, DvmType line, DvmType out[13]);
/* The same for the chain site of that loop (DVMH_AMD_KIND_CHAIN). */
DvmType dvmh_amd_tuner_query_chain_C(const char *file, DvmType line, DvmType out[13]);
/* The same for the pair site of a fused pair whose first loop is at (file, line)
 * (DVMH_AMD_KIND_PAIR). */
DvmType dvmh_amd_tuner_query_pair_C(const char *file, DvmType line, DvmType out[13]);
/* How the layout-transform candidates of the loop's sites fared: out[0]
 * transform candidates, out[1] of them timed, out[2] best timed one in ns and
 * out[3] best timed natural-layout candidate in ns (-1 where none was timed);
 * returns the number of sites. */
DvmType dvmh_amd_tuner_transform_stats_C(const char *file, DvmType line, DvmType out[4]);
/* Forces the launch configuration (block[3], tune[DVMH_AMD_TUNE_N]) of every
 * loop of the given kind, over whatever the mode would choose; forced
 * executions are not timed.  block == NULL removes the forcing for the kind. */
void dvmh_amd_tuner_force_C(DvmType kind, const DvmType block[3], const DvmType tune[DVMH_AMD_TUNE_N]);
/* Keys of the known loop sites, one per line, into buf (truncated to bufLen);
 * returns the length needed including the terminating zero. */
DvmType dvmh_amd_tuner_sites_C(char *buf, DvmType bufLen);

/* ---- learned cost model -------------------------------------------------------
 * A perceptron with DVMH_AMD_MODEL_F inputs, two ReLU layers of
 * DVMH_AMD_MODEL_H units and one output predicts log(seconds per iteration
 * point) of a loop site under a candidate configuration.  Training and
 * inference run on the GPU.  Parameters are one FP32 vector of
 * dvmh_amd_model_param_count_C() values: W1[H][F] b1[H] W2[H][H] b2[H] W3[H] b3. */
#define DVMH_AMD_MODEL_F 24
#define DVMH_AMD_MODEL_H 64
DvmType dvmh_amd_model_param_count_C(void);
/* Trains a fresh network (Adam, `epochs` full-batch steps, cosine-decayed step
 * size from lr, weights seeded by seed) on every recorded timing except those
 * of sites whose key contains `holdout` (NULL or "": none).  Returns the
 * number of samples; *finalLoss (may be NULL) gets the standardized MSE. */
DvmType dvmh_amd_model_train_C(DvmType epochs, double lr, DvmType seed, const char *holdout, double *finalLoss);
DvmType dvmh_amd_model_save_C(const char *path);
DvmType dvmh_amd_model_load_C(const char *path);
DvmType dvmh_amd_model_ready_C(void);
/* For the first site whose key contains keyPart: predicted seconds (pred) and
 * recorded seconds (measured, -1 if none) of every candidate.  Returns the
 * candidate count, -1 if no site matches. */
DvmType dvmh_amd_model_site_predict_C(const char *keyPart, double *pred, double *measured, DvmType maxCands);
/* Raw (unstandardized) feature rows of that site's candidates. */
DvmType dvmh_amd_model_site_features_C(const char *keyPart, float *X, DvmType maxCands);
/* Direct access to the kernels (host pointers): gradient and loss of
 * mean((net(X) - y)^2), network outputs, one Adam step (step counts from 1). */
DvmType dvmh_amd_model_grad_C(const float *params, const float *X, const float *y, DvmType n, float *grad, float *loss);
DvmType dvmh_amd_model_forward_C(const float *params, const float *X, DvmType n, float *out);
DvmType dvmh_amd_model_adam_C(float *params, const float *grad, float *m, float *v, DvmType step, double lr, double beta1,
        double beta2, double eps);

#ifdef __cplusplus
}
#endif
#pragma GCC visibility pop

#endif
