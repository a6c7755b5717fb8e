/*
 * topaz_index_compat.h -- drop-in for the suffix-array step of `topaz index`.
 *
 * libtopaz_amd.so exports divsufsort64 with libdivsufsort64's signature
 * (divsufsort64.h:79), the one call index.c:66 makes to build the .SA file,
 * so TOPAZ's objects link against it in place of divsufsort.o, sssort.o and
 * trsort.o.  The suffix array is built on the GPU (tpz_suffix_array); being
 * unique, it is the reference's byte for byte.
 */
#ifndef TOPAZ_INDEX_COMPAT_H
#define TOPAZ_INDEX_COMPAT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* divsufsort64.h:79.  Returns 0 when done, -1 for a null pointer or
 * n < 0 (divsufsort.c:338), -2 when the GPU build fails (the reference's
 * code for a failed allocation), with the reason on stderr. */
int32_t divsufsort64(const uint8_t* T, int64_t* SA, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
