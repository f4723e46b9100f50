/*
 * spike_memory.c -- host allocator of the reference's memory ABI (OpenMP/spike_memory.h:33-37;
 * pinned like spike_memory_cuda.c:22, whose cudaMallocHost is hipHostMalloc here).
 *
 * Every buffer carries a small header just below the address handed out, recording which
 * allocator produced it, so spike_free releases pinned and ordinary memory alike.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "spike_memory.h"

unsigned int cnt_alloc = 0;
unsigned int cnt_free = 0;

enum { KIND_PINNED = 1, KIND_HOST = 2 };
#define MEM_MAGIC 0x4853504b4d454d31ULL   /* "HSPKMEM1" */

typedef struct {
    uint64_t magic;
    uint32_t kind;
    uint32_t offset;    /* bytes from the allocation's base to the address handed out */
    uint64_t bytes;
    uint64_t pad_;
} mem_hdr_t;

/* pinned memory is used when a HIP device is visible and HSPIKE_PINNED is not 0 (decided once) */
static int pinned_enabled(void)
{
    static int state = -1;
    int s = __atomic_load_n(&state, __ATOMIC_ACQUIRE);
    if (s >= 0) return s;
    const char *v = getenv("HSPIKE_PINNED");
    int ndev = 0;
    s = !(v && atoi(v) == 0) && hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0;
    if (!s) (void)hipGetLastError();
    __atomic_store_n(&state, s, __ATOMIC_RELEASE);
    return s;
}

static mem_hdr_t *header_of(const void *ptr) { return (mem_hdr_t *)((char *)ptr - sizeof(mem_hdr_t)); }

void *_spike_malloc(const int alignment, const size_t nmemb, const size_t size, const char *function, const int line)
{
    size_t align = 64;
    while ((int64_t)align < (int64_t)alignment) align <<= 1;
    if (size && nmemb > (SIZE_MAX - 2 * align) / size) {
        fprintf(stderr, "%s (from %s:%d): %zu elements of %zu bytes overflow size_t\n", __func__,
                function ? function : "?", line, nmemb, size);
        return NULL;
    }
    const size_t bytes = nmemb * size, total = bytes + align;   /* header fits in the first `align` bytes */
    void *base = NULL;
    uint32_t kind = 0;
    if (pinned_enabled()) {
        if (hipHostMalloc(&base, total, hipHostMallocDefault) == hipSuccess && base) kind = KIND_PINNED;
        else { (void)hipGetLastError(); base = NULL; }
    }
    if (!base) {
        if (posix_memalign(&base, align, total) != 0) base = NULL;
        kind = KIND_HOST;
    }
    if (!base) {
        fprintf(stderr, "%s (from %s:%d): cannot allocate %zu elements of %zu bytes\n", __func__,
                function ? function : "?", line, nmemb, size);
        return NULL;
    }
    char *user = (char *)base + align;
    mem_hdr_t *h = header_of(user);
    h->magic = MEM_MAGIC;
    h->kind = kind;
    h->offset = (uint32_t)align;
    h->bytes = bytes;
    __atomic_add_fetch(&cnt_alloc, 1u, __ATOMIC_RELAXED);
    return user;
}

void spike_free(void *ptr)
{
    __atomic_add_fetch(&cnt_free, 1u, __ATOMIC_RELAXED);
    if (!ptr) return;
    mem_hdr_t *h = header_of(ptr);
    if (h->magic != MEM_MAGIC) {
        fprintf(stderr, "%s: %p was not allocated by spike_malloc; not released\n", __func__, ptr);
        return;
    }
    void *base = (char *)ptr - h->offset;
    const uint32_t kind = h->kind;
    h->magic = 0;
    if (kind == KIND_PINNED) {
        if (hipHostFree(base) != hipSuccess) (void)hipGetLastError();
    } else {
        free(base);
    }
}

void spike_nullify(void *ptr) { spike_free(ptr); }

int spike_memory_is_pinned(const void *ptr)
{
    if (!ptr) return 0;
    const mem_hdr_t *h = header_of(ptr);
    return h->magic == MEM_MAGIC && h->kind == KIND_PINNED;
}
