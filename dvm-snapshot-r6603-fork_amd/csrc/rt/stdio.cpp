// dvmh_* stdio replacements (rts/dvmh/include/dvmhlib_stdio.h).  One process
// owns every file, so each is a pass-through to the C library.
#include <cstdarg>
#include <cstdio>

#include "dvmhlib_stdio.h"

extern "C" {
int dvmh_remove(const char *filename) { return remove(filename); }
int dvmh_rename(const char *oldFN, const char *newFN) { return rename(oldFN, newFN); }
int dvmh_fclose(FILE *stream) { return fclose(stream); }
int dvmh_fflush(FILE *stream) { return fflush(stream); }
FILE *dvmh_fopen(const char *filename, const char *mode) { return fopen(filename, mode); }
int dvmh_fprintf(FILE *stream, const char *format, ...) {
    va_list ap;
    va_start(ap, format);
    int r = vfprintf(stream, format, ap);
    va_end(ap);
    return r;
}
void dvmh_void_fprintf(FILE *stream, const char *format, ...) {
    va_list ap;
    va_start(ap, format);
    vfprintf(stream, format, ap);
    va_end(ap);
}
int dvmh_printf(const char *format, ...) {
    va_list ap;
    va_start(ap, format);
    int r = vprintf(format, ap);
    va_end(ap);
    return r;
}
void dvmh_void_printf(const char *format, ...) {
    va_list ap;
    va_start(ap, format);
    vprintf(format, ap);
    va_end(ap);
}
int dvmh_vprintf(const char *format, va_list arg) { return vprintf(format, arg); }
void dvmh_void_vprintf(const char *format, va_list arg) { vprintf(format, arg); }
int dvmh_fputs(const char *s, FILE *stream) { return fputs(s, stream); }
void dvmh_void_fputs(const char *s, FILE *stream) { fputs(s, stream); }
int dvmh_putchar(int c) { return putchar(c); }
void dvmh_void_putchar(int c) { putchar(c); }
int dvmh_puts(const char *s) { return puts(s); }
void dvmh_void_puts(const char *s) { puts(s); }
size_t dvmh_fread(void *ptr, size_t size, size_t nmemb, FILE *stream) { return fread(ptr, size, nmemb, stream); }
void dvmh_void_fread(void *ptr, size_t size, size_t nmemb, FILE *stream) {
    size_t got = fread(ptr, size, nmemb, stream);
    (void)got;
}
size_t dvmh_fwrite(const void *ptr, size_t size, size_t nmemb, FILE *stream) { return fwrite(ptr, size, nmemb, stream); }
void dvmh_void_fwrite(const void *ptr, size_t size, size_t nmemb, FILE *stream) { (void)fwrite(ptr, size, nmemb, stream); }
}
