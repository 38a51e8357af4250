/*
 * dvmhlib2.h -- C interface of the MI355X DVMH device runtime (libdvmh_amd.so).
 *
 * This is the drop-in boundary for programs converted by cdvmh: every entry
 * point keeps the name, argument list, argument meaning and calling convention
 * of the reference interface rts/dvmh/include/dvmhlib2.h (line numbers cited
 * per declaration), so converted host code and emitted handlers link against
 * this library unchanged.  The "_cuda_" spelling inside names is ABI spelling
 * only: the accelerator behind it is one gfx950 GPU driven through HIP, the
 * dim3 out-parameter is 3 x uint32 and the stream out-parameter is a
 * hipStream_t.
 *
 * Scope: one process, one GPU (plus the host).  Calls that only make sense with
 * several processes (indirect distributions, remote access, redistribution)
 * are exported so programs link, and abort with a clear message when called.
 *
 * Errors follow the reference: there are no return codes; a user error or an
 * internal error prints a message naming the call and aborts the process.
 */
#ifndef DVMHLIB2_H
#define DVMHLIB2_H

#include <stdlib.h>

#include "dvmhlib_types.h"
#include "dvmhlib_const.h"
#include "dvmhlib_stdio.h"
#include "dvmh_runtime_api.h"

#pragma GCC visibility push(default)
#ifdef __cplusplus
extern "C" {
#endif

/* ---- initialisation and exit (ref dvmhlib2.h:19-27) ---- */
void dvmh_line_C(DvmType lineNumber, const char fileName[]);                  /* :19 */
void dvmh_init_C(DvmType flags, int *pArgc, char ***pArgv);                     /* :22 */
void dvmh_init_lib_C(DvmType flags);                                             /* :24 */
void dvmh_exit_C(DvmType exitCode);                                              /* :26 */

/* ---- distributed arrays and templates (ref :29-79) ----
 * Variadic tails: per axis (axisSize, shadowLow, shadowHigh) for arrays,
 * (axisSize) for templates, one helper index per axis for distribute/align. */
void dvmh_array_declare_C(DvmType dvmDesc[], DvmType rank, DvmType typeSize, ...); /* :29 */
void dvmh_array_create_C(DvmType dvmDesc[], DvmType rank, DvmType typeSize, ...);  /* :30 */
void dvmh_template_create_C(DvmType dvmDesc[], DvmType rank, ...);                 /* :33 */
void dvmh_array_alloc_C(DvmType dvmDesc[], DvmType byteCount);                    /* :35 */
DvmType dvmh_handler_func_C(DvmHandlerFunc handlerFunc, DvmType customParamCount, ...); /* :37 */
DvmType dvmh_distribution_replicated_();                                          /* :46 */
DvmType dvmh_distribution_block_C(DvmType mpsAxis);                               /* :47 */
DvmType dvmh_distribution_wgtblock_C(DvmType mpsAxis, DvmType elemType, const void *arrayAddr, DvmType elemCount); /* :49 */
DvmType dvmh_distribution_genblock_C(DvmType mpsAxis, DvmType elemType, const void *arrayAddr);                  /* :51 */
DvmType dvmh_distribution_multblock_C(DvmType mpsAxis, DvmType multBlock);                                      /* :53 */
DvmType dvmh_distribution_indirect_C(DvmType mpsAxis, DvmType elemType, const void *arrayAddr);                 /* :55 */
DvmType dvmh_distribution_derived_C(DvmType mpsAxis, DvmType derivedRhsHelper, DvmType countingHandlerHelper,
        DvmType fillingHandlerHelper);                                            /* :57 */
void dvmh_distribute_C(DvmType dvmDesc[], DvmType rank, ...);                     /* :61 */
void dvmh_redistribute_C(DvmType dvmDesc[], DvmType rank, ...);                   /* :63 */
DvmType dvmh_alignment_linear_C(DvmType axis, DvmType multiplier, DvmType summand); /* :65 */
void dvmh_align_C(DvmType dvmDesc[], const DvmType templDesc[], DvmType templRank, ...); /* :67 */
void dvmh_realign_C(DvmType dvmDesc[], DvmType newValueFlag, const DvmType templDesc[], DvmType templRank, ...); /* :69 */
void dvmh_array_free_C(DvmType dvmDesc[]);                                        /* :77 */
void dvmh_delete_object_(DvmType dvmDesc[]);                                      /* :78 */
void dvmh_forget_header_(DvmType dvmDesc[]);                                      /* :79 */

/* ---- regular variables and memory (ref :81-95) ---- */
DvmType *dvmh_variable_gen_header_C(const void *addr, DvmType rank, DvmType typeSize, ...); /* :81 */
DvmType *dvmh_variable_get_header_C(const void *addr);                            /* :86 */
void dvmh_data_enter_C(const void *addr, DvmType size);                           /* :87 */
void dvmh_data_exit_C(const void *addr, DvmType saveFlag);                        /* :89 */
void *dvmh_malloc_C(size_t size);                                                  /* :91 */
void *dvmh_calloc_C(size_t nmemb, size_t size);                                    /* :92 */
void *dvmh_realloc_C(void *ptr, size_t size);                                      /* :93 */
char *dvmh_strdup_C(const char *s);                                                /* :94 */
void dvmh_free_C(void *ptr);                                                       /* :95 */

/* ---- device-side addressing (ref :103-107) ----
 * Header layout written by dvmh_fill_header_C (ref rts/dvmh/dvmh_buffer.cpp:149-182):
 *   [0] object handle, [1..rank] element stride of each ORIGINAL axis (0 for the
 *   two diagonalised axes), [rank+1] element offset from baseAddr, [rank+2] baseAddr.
 * extendedParams (7 entries, written only when the return value is 2):
 *   {x_axis, x_first, x_length, y_axis, y_first, y_length, slash}.
 * Return value: 0 natural layout, 1 permuted axes, 2 permuted + diagonalised. */
void *dvmh_get_natural_base_C(DvmType deviceNum, const DvmType dvmDesc[]);      /* :103 */
void *dvmh_get_device_addr_C(DvmType deviceNum, const void *addr);               /* :104 */
DvmType dvmh_fill_header_C(DvmType deviceNum, const void *baseAddr, const DvmType dvmDesc[], DvmType devHeader[],
        DvmType extendedParams[]);                                                /* :105 */

/* ---- actual-state management (ref :109-123) ---- */
void dvmh_get_actual_subvariable_C(void *addr, DvmType rank, ...);              /* :109 */
void dvmh_get_actual_variable2_(void *addr);                                     /* :111 */
void dvmh_get_actual_subarray_C(const DvmType dvmDesc[], DvmType rank, ...);    /* :112 */
void dvmh_get_actual_array2_(const DvmType dvmDesc[]);                           /* :114 */
void dvmh_get_actual_all2_();                                                    /* :115 */
void dvmh_actual_subvariable_C(const void *addr, DvmType rank, ...);            /* :117 */
void dvmh_actual_variable2_(const void *addr);                                   /* :119 */
void dvmh_actual_subarray_C(const DvmType dvmDesc[], DvmType rank, ...);        /* :120 */
void dvmh_actual_array2_(const DvmType dvmDesc[]);                               /* :122 */
void dvmh_actual_all2_();                                                        /* :123 */

/* ---- shadow edges and element access (ref :125-140) ---- */
void dvmh_shadow_renew_C(const DvmType dvmDesc[], DvmType cornerFlag, DvmType specifiedRank, ...); /* :125 */
void dvmh_indirect_shadow_renew_C(const DvmType dvmDesc[], DvmType axis, const char shadowName[]); /* :128 */
void dvmh_remote_access_C(DvmType rmaDesc[], const DvmType dvmDesc[], DvmType rank, ...);          /* :133 */
DvmType dvmh_has_element_C(const DvmType dvmDesc[], DvmType rank, ...);          /* :135 */
DvmType dvmh_calc_linear_C(const DvmType dvmDesc[], DvmType rank, ...);          /* :137 */
void *dvmh_get_own_element_addr_C(const DvmType dvmDesc[], DvmType rank, ...);   /* :139 */
void *dvmh_get_element_addr_C(const DvmType dvmDesc[], DvmType rank, ...);       /* :140 */

/* ---- regions (ref :142-155) ---- */
DvmType dvmh_region_create_C(DvmType regionFlags);                                /* :142 */
void dvmh_region_register_subarray_C(DvmType curRegion, DvmType intent, const DvmType dvmDesc[], const char varName[],
        DvmType rank, ...);                                                       /* :144 */
void dvmh_region_register_array_C(DvmType curRegion, DvmType intent, const DvmType dvmDesc[], const char varName[]); /* :148 */
void dvmh_region_register_scalar_C(DvmType curRegion, DvmType intent, const void *addr, DvmType typeSize,
        const char varName[]);                                                    /* :150 */
void dvmh_region_execute_on_targets_C(DvmType curRegion, DvmType deviceTypes);    /* :152 */
void dvmh_region_end_C(DvmType curRegion);                                        /* :154 */

/* ---- loop registration (ref :157-183) ----
 * dvmh_loop_create_C tail: per loop axis (start, end, step) with end inclusive. */
DvmType dvmh_loop_create_C(DvmType curRegion, DvmType rank, ...);                /* :157 */
void dvmh_loop_map_C(DvmType curLoop, const DvmType templDesc[], DvmType templRank, ...); /* :159 */
void dvmh_loop_set_cuda_block_C(DvmType curLoop, DvmType xSize, DvmType ySize, DvmType zSize); /* :161 */
void dvmh_loop_set_stage_C(DvmType curLoop, DvmType stage);                       /* :163 */
void dvmh_loop_reduction_C(DvmType curLoop, DvmType redType, void *arrayAddr, DvmType varType, DvmType arrayLength,
        void *locAddr, DvmType locSize);                                          /* :165 */
void dvmh_loop_across_C(DvmType curLoop, DvmType isOut, const DvmType dvmDesc[], DvmType rank, ...); /* :168 */
void dvmh_loop_shadow_compute_C(DvmType curLoop, const DvmType templDesc[], DvmType specifiedRank, ...); /* :170 */
void dvmh_loop_shadow_compute_array_C(DvmType curLoop, const DvmType dvmDesc[]); /* :173 */
void dvmh_loop_consistent_C(DvmType curLoop, const DvmType dvmDesc[], DvmType rank, ...); /* :175 */
void dvmh_loop_remote_access_C(DvmType curLoop, const DvmType dvmDesc[], DvmType rank, ...); /* :177 */
void dvmh_loop_array_correspondence_C(DvmType curLoop, const DvmType dvmDesc[], DvmType rank, ...); /* :179 */
void dvmh_loop_register_handler_C(DvmType curLoop, DvmType deviceType, DvmType handlerType, DvmType handlerHelper); /* :181 */
void dvmh_loop_perform_C(DvmType curLoop);                                        /* :183 */

/* ---- handler-side queries (ref :185-212) ----
 * curLoop inside a handler is the value *pLoopRef the handler was called with. */
DvmType dvmh_loop_get_dependency_mask_C(DvmType curLoop);                         /* :185 */
DvmType dvmh_loop_get_device_num_C(DvmType curLoop);                              /* :187 */
DvmType dvmh_loop_autotransform_C(DvmType curLoop, DvmType dvmDesc[]);            /* :189 */
void dvmh_loop_fill_bounds_C(DvmType curLoop, DvmType boundsLow[], DvmType boundsHigh[], DvmType loopSteps[]); /* :191 */
DvmType dvmh_loop_get_slot_count_C(DvmType curLoop);                              /* :193 */
void dvmh_loop_fill_local_part_C(DvmType curLoop, const DvmType dvmDesc[], DvmType part[]); /* :195 */
DvmType dvmh_loop_guess_index_type_C(DvmType curLoop);                            /* :197 */
const void *dvmh_loop_cuda_get_local_part_C(DvmType curLoop, const DvmType dvmDesc[], DvmType indexType); /* :199 */
void dvmh_loop_cuda_register_red_C(DvmType curLoop, DvmType redIndex, void **arrayAddrPtr, void **locAddrPtr); /* :202 */
void dvmh_loop_red_init_C(DvmType curLoop, DvmType redIndex, void *arrayAddr, void *locAddr); /* :203 */
void dvmh_loop_cuda_red_init_C(DvmType curLoop, DvmType redIndex, void **devArrayAddrPtr, void **devLocAddrPtr); /* :205 */
void dvmh_loop_cuda_get_config_C(DvmType curLoop, DvmType sharedPerThread, DvmType regsPerThread, void *inOutThreads,
        void *outStream, DvmType *outSharedPerBlock);                             /* :206 */
void dvmh_loop_cuda_red_prepare_C(DvmType curLoop, DvmType redIndex, DvmType count, DvmType fillFlag); /* :208 */
DvmType dvmh_loop_has_element_C(DvmType curLoop, const DvmType dvmDesc[], DvmType rank, ...); /* :209 */
void dvmh_loop_cuda_red_finish_C(DvmType curLoop, DvmType redIndex);              /* :211 */
void dvmh_loop_red_post_C(DvmType curLoop, DvmType redIndex, const void *arrayAddr, const void *locAddr); /* :212 */

/* ---- intervals (ref :219-223) ---- */
void dvmh_par_interval_start_C();                                                 /* :219 */
void dvmh_seq_interval_start_C();                                                 /* :220 */
void dvmh_sp_interval_end_();                                                     /* :221 */
void dvmh_usr_interval_start_C(DvmType userID);                                   /* :222 */
void dvmh_usr_interval_end_();                                                    /* :224 */

/* ---- loop-bound permutation for ACROSS handlers (ref :229-230) ---- */
void dvmh_change_filled_bounds_C(DvmType boundsLow[], DvmType boundsHigh[], DvmType loopSteps[], DvmType rank,
        DvmType depMask, DvmType idxPerm[]);                                      /* :229 */
void dvmh_change_filled_bounds2_C(DvmType boundsLow[], DvmType boundsHigh[], DvmType loopSteps[], DvmType rank,
        DvmType depMask, DvmType idxPerm[]);                                      /* :230 */

#ifdef __cplusplus
}
#endif
#pragma GCC visibility pop

#endif
