/*
 * spike_memory.h -- the reference's host allocator ABI (OpenMP/spike_memory.h:25-37), backed
 * by pinned host memory as the reference's GPU build does (spike_memory_cuda.c:8-48 uses
 * cudaMallocHost): buffers from spike_malloc are page-locked and device-mapped, so a caller's
 * panels staged through them move to and from HBM by DMA without an extra copy.
 *
 * Kept from the reference: the names, the spike_malloc(alignment, nmemb, size) macro that
 * records the calling function and line, the ALIGN_* constants, and the cnt_alloc / cnt_free
 * counters (one increment per _spike_malloc / spike_free call).
 * Changed: a failed allocation prints one line to stderr and returns NULL instead of aborting;
 * nmemb * size overflow is rejected the same way; the memory is pinned when a HIP device is
 * visible and HSPIKE_PINNED is not 0, and ordinary aligned host memory otherwise (spike_free
 * releases either kind); the counters are updated atomically.
 */
#ifndef HSPIKE_SPIKE_MEMORY_H
#define HSPIKE_SPIKE_MEMORY_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

extern unsigned int cnt_alloc;   /* spike_memory.h:26 */
extern unsigned int cnt_free;    /* spike_memory.h:27 */

#ifndef ALIGN_INT
#define ALIGN_INT     (const int) 16   /* spike_memory.h:29-31 */
#define ALIGN_REAL    (const int) 32
#define ALIGN_COMPLEX (const int) 32
#endif

#ifndef spike_malloc
#define spike_malloc(alignment, nmemb, size) _spike_malloc(alignment, nmemb, size, __FUNCTION__, __LINE__)
#endif

/* spike_memory.h:35: nmemb * size bytes aligned to `alignment` (a power of two; at least 64) */
void *_spike_malloc(const int alignment, const size_t nmemb, const size_t size, const char *function, const int line);
/* spike_memory.h:36: releases a buffer from _spike_malloc (NULL is ignored) */
void spike_free(void *ptr);
/* spike_memory.h:37: same as spike_free (the reference's assignment to its by-value copy of
 * the pointer has no effect on the caller) */
void spike_nullify(void *ptr);
/* 1 when ptr (from _spike_malloc) is pinned, device-mapped host memory, 0 otherwise */
int spike_memory_is_pinned(const void *ptr);

#ifdef __cplusplus
}
#endif
#endif
