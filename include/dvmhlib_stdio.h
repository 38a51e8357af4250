/* stdio replacements that cdvmh substitutes into converted programs by default
 * (cdvmh/src/project_ctx.cpp:36,432-448).  Same names and signatures as the
 * reference (rts/dvmh/include/dvmhlib_stdio.h, line numbers per declaration).
 * One process owns all data here, so each is a thin pass-through to libc. */
#ifndef DVMHLIB_STDIO_H
#define DVMHLIB_STDIO_H

#include <stdarg.h>
#include <stdio.h>

#pragma GCC visibility push(default)
#ifdef __cplusplus
extern "C" {
#endif

int dvmh_remove(const char *filename);                                          /* :16 */
int dvmh_rename(const char *oldFN, const char *newFN);                          /* :17 */
int dvmh_fclose(FILE *stream);                                                  /* :21 */
int dvmh_fflush(FILE *stream);                                                  /* :22 */
FILE *dvmh_fopen(const char *filename, const char *mode);                       /* :23 */
int dvmh_fprintf(FILE *stream, const char *format, ...);                        /* :28 */
void dvmh_void_fprintf(FILE *stream, const char *format, ...);                  /* :29 */
int dvmh_printf(const char *format, ...);                                       /* :32 */
void dvmh_void_printf(const char *format, ...);                                 /* :33 */
int dvmh_vprintf(const char *format, va_list arg);                              /* :40 */
void dvmh_void_vprintf(const char *format, va_list arg);                        /* :41 */
int dvmh_fputs(const char *s, FILE *stream);                                    /* :50 */
void dvmh_void_fputs(const char *s, FILE *stream);                              /* :51 */
int dvmh_putchar(int c);                                                        /* :58 */
void dvmh_void_putchar(int c);                                                  /* :59 */
int dvmh_puts(const char *s);                                                   /* :60 */
void dvmh_void_puts(const char *s);                                             /* :61 */
size_t dvmh_fread(void *ptr, size_t size, size_t nmemb, FILE *stream);          /* :65 */
void dvmh_void_fread(void *ptr, size_t size, size_t nmemb, FILE *stream);       /* :67 */
size_t dvmh_fwrite(const void *ptr, size_t size, size_t nmemb, FILE *stream);   /* :69 */
void dvmh_void_fwrite(const void *ptr, size_t size, size_t nmemb, FILE *stream); /* :71 */

#ifdef __cplusplus
}
#endif
#pragma GCC visibility pop

#endif
