/* Basic ABI types of the DVMH runtime C interface.
 * Same encodings as the reference (rts/dvmh/include/dvmhlib_types.h:8-17):
 * DvmType is `long` on LP64, handlers are called as int f(...). */
#ifndef DVMHLIB_TYPES_H
#define DVMHLIB_TYPES_H

typedef long DvmType;
typedef unsigned long UDvmType;

#ifdef __cplusplus
typedef int (*DvmHandlerFunc)(...);
#else
typedef int (*DvmHandlerFunc)();
#endif

#endif
