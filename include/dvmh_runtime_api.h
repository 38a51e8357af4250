/* User-callable runtime API (rts/dvmh/include/dvmh_runtime_api.h:14-35). */
#ifndef DVMH_RUNTIME_API_H
#define DVMH_RUNTIME_API_H

#include <stdio.h>

#pragma GCC visibility push(default)
#ifdef __cplusplus
extern "C" {
#endif

/* Barrier over the current multiprocessor system (one process here: a full
 * device synchronisation). */
void dvmh_barrier();
int dvmh_get_total_num_procs();
int dvmh_get_num_proc_axes();
int dvmh_get_num_procs(int axis);
/* Wall-clock seconds. */
double dvmh_wtime();

int dvmh_remove_local(const char *filename);
int dvmh_rename_local(const char *oldFN, const char *newFN);
FILE *dvmh_tmpfile_local(void);
char *dvmh_tmpnam_local(char *s);

#ifdef __cplusplus
}
#endif
#pragma GCC visibility pop

#endif
