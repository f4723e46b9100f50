// spike_core.hip -- host orchestration of the MI355X SPIKE solver and the spike_ext.h C ABI.
//
// Pipeline per batch of systems (a single-system handle is a batch of one):
//   validate + bandwidth (one launch over every system's rows)
//   group systems by block size b = 16 * NT, NT in {1, 2, 4, 8, 16}
//   per group: CSR -> blocks, block LU with the top block rows R_j (k_factor; k_rtop where the
//              R panels do not fit in LDS), V/W spike tips keeping their forward sweeps
//              (k_sweep), reduced-system setup (k_red_setup_mfma)                 -- spike_factor
//   (b = 256 and complex FP64 at b = 128: the same stages with the factor state in global
//   memory -- k_factor_gm, k_sweep_gm, k_recover_gm, k_red_setup_gm)
//   per RHS chunk: y tips keeping the forward sweep (k_sweep), block-Jacobi reduced solve in
//              FP64 (k_red_jacobi_q / k_red_jacobi), recovery as one backward chain into the FP64
//              solution (k_recover), FP64 residual (k_residual); mixed precision repeats on the
//              residual until every column meets tol.                                -- solve
// The reference's flow is OpenMP/spike_interfaces.c:43-419; partition rows follow
// OpenMP/spike_analysis.c:38-76.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <vector>
#include <algorithm>
#include <chrono>
#include <mutex>
#include <type_traits>
#include <memory>
#include <thread>
#include <condition_variable>
#include <deque>
#include <map>
#include <numeric>
#include <tuple>
#include <initializer_list>

#include "spike_dev.h"
#include "spike_kernels.inc"
#include "spike_kernels_wide.inc"
#include "spike_kernels_aux.inc"
#include "spike_band.inc"
#include "spike_ext.h"

// ============================================================================ utilities
namespace {

thread_local char g_errmsg[512];

__attribute__((format(printf, 1, 2))) void set_err(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_errmsg, sizeof g_errmsg, fmt, ap);
    va_end(ap);
}

#define HIPCK(call)                                                                  \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            set_err("HIP error %s (line %ld)", hipGetErrorString(e_), (long)__LINE__); \
            return SPIKE_ERR_DEVICE;                                                 \
        }                                                                            \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    int ensure(size_t need) {
        if (need == 0) need = 256;
        if (bytes >= need) return 0;
        release();
        if (hipMalloc(&p, need) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            set_err("device allocation of %s%ld bytes failed", "", (long)need);
            return SPIKE_ERR_OOM;
        }
        bytes = need;
        if (poison_new()) (void)hipMemset(p, 0xff, need);   // NaN in every float type
        return 0;
    }
    // HSPIKE_POISON=1: fill new device buffers with NaN bytes, so any read of memory no kernel
    // wrote shows up in the results (debug aid)
    static bool poison_new() {
        static const bool on = [] { const char *v = getenv("HSPIKE_POISON"); return v && *v == '1'; }();
        return on;
    }
    template <typename X> X *as() const { return (X *)p; }
};

// Pinned host memory mapped into the device address space: the staging area of small transfers.
struct PinBuf {
    char *h = nullptr;
    char *d = nullptr;
    size_t bytes = 0;
    PinBuf() = default;
    PinBuf(const PinBuf &) = delete;
    PinBuf &operator=(const PinBuf &) = delete;
    ~PinBuf() { if (h) (void)hipHostFree(h); }
    int ensure(size_t need) {
        if (bytes >= need) return 0;
        if (h) (void)hipHostFree(h);
        h = d = nullptr;
        bytes = 0;
        void *p = nullptr, *dp = nullptr;
        if (hipHostMalloc(&p, need, hipHostMallocDefault) != hipSuccess ||
            hipHostGetDevicePointer(&dp, p, 0) != hipSuccess) {
            (void)hipGetLastError();
            if (p) (void)hipHostFree(p);
            set_err("pinned host allocation of %s%ld bytes failed", "", (long)need);
            return SPIKE_ERR_OOM;
        }
        h = (char *)p;
        d = (char *)dp;
        bytes = need;
        return 0;
    }
};

// Host -> device copy of a small descriptor array; completes before returning because the host
// vectors are transient or rewritten right after the call.
int upload(DevBuf &d, const void *h, size_t bytes, hipStream_t s) {
    int e = d.ensure(bytes);
    if (e) return e;
    if (bytes) {
        HIPCK(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, s));
        HIPCK(hipStreamSynchronize(s));
    }
    return 0;
}

inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

// Copies between pageable host arrays (what callers hand over) and device memory through two
// pinned buffers: the runtime would move pageable data through small staging copies run on the
// compute units.  An upload copies chunk k into one buffer while the DMA of chunk k - 1 drains
// from the other; a download starts the DMA of chunk k + 1 before copying chunk k out.  Chunk
// copies on the host are split over `threads` threads (1 where several stagers copy at once, as
// the server's workers do).  The _cols forms move ncols columns of `width`
// bytes between a strided host panel (pitch >= width) and a contiguous device one.  Without
// pinned buffers (allocation refused) the copies go to the runtime directly.  One stager serves one
// thread at a time.
struct HostStager {
    static constexpr size_t CHUNK = (size_t)32 << 20;
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool pend[2] = {false, false};          // a DMA reading buf[k] may still run
    bool ok = false;
    int threads = 4;
    explicit HostStager(int copy_threads = 4) : threads(copy_threads) {
        ok = true;
        for (int k = 0; k < 2 && ok; ++k)
            ok = hipHostMalloc(&buf[k], CHUNK, hipHostMallocDefault) == hipSuccess &&
                 hipEventCreateWithFlags(&ev[k], hipEventDisableTiming) == hipSuccess;
        if (!ok) (void)hipGetLastError();
    }
    ~HostStager() {
        for (int k = 0; k < 2; ++k) {
            if (ev[k]) { (void)hipEventSynchronize(ev[k]); (void)hipEventDestroy(ev[k]); }
            if (buf[k]) (void)hipHostFree(buf[k]);
        }
    }
    HostStager(const HostStager &) = delete;
    HostStager &operator=(const HostStager &) = delete;
    void par_copy(void *dst, const void *src, size_t len) const {
        const int nth = std::max(1, std::min(threads, 8));
        if (nth == 1 || len < ((size_t)4 << 20)) { memcpy(dst, src, len); return; }
        const size_t part = (len / nth + 63) & ~(size_t)63;
        std::thread th[7];
        for (int t = 1; t < nth; ++t) {
            const size_t o = std::min(len, part * t), l = std::min(len, part * (t + 1)) - o;
            th[t - 1] = std::thread([=] { memcpy((char *)dst + o, (const char *)src + o, l); });
        }
        memcpy(dst, src, std::min(len, part));
        for (int t = 1; t < nth; ++t) th[t - 1].join();
    }
    // bytes [off, off + len) of the column image <-> the strided host panel
    void pack(void *chunk, const void *src, size_t spitch, size_t width, size_t off, size_t len) const {
        if (spitch == width) { par_copy(chunk, (const char *)src + off, len); return; }
        for (size_t o = off; o < off + len;) {
            const size_t c = o / width, in = o % width, piece = std::min(width - in, off + len - o);
            memcpy((char *)chunk + (o - off), (const char *)src + c * spitch + in, piece);
            o += piece;
        }
    }
    void unpack(void *dst, size_t dpitch, const void *chunk, size_t width, size_t off, size_t len) const {
        if (dpitch == width) { par_copy((char *)dst + off, chunk, len); return; }
        for (size_t o = off; o < off + len;) {
            const size_t c = o / width, in = o % width, piece = std::min(width - in, off + len - o);
            memcpy((char *)dst + c * dpitch + in, (const char *)chunk + (o - off), piece);
            o += piece;
        }
    }
    int h2d_cols(void *dst, const void *src, size_t spitch, size_t width, size_t ncols, hipStream_t st) {
        if (!ok) {
            HIPCK(hipMemcpy2DAsync(dst, width, src, spitch, width, ncols, hipMemcpyHostToDevice, st));
            return 0;
        }
        const size_t total = width * ncols;
        int k = 0;
        for (size_t off = 0; off < total; off += CHUNK, k ^= 1) {
            const size_t len = std::min(CHUNK, total - off);
            if (pend[k]) HIPCK(hipEventSynchronize(ev[k]));
            pack(buf[k], src, spitch, width, off, len);
            HIPCK(hipMemcpyAsync((char *)dst + off, buf[k], len, hipMemcpyHostToDevice, st));
            HIPCK(hipEventRecord(ev[k], st));
            pend[k] = true;
        }
        return 0;
    }
    int h2d(void *dst, const void *src, size_t bytes, hipStream_t st) { return h2d_cols(dst, src, bytes, bytes, 1, st); }
    // returns once the data is in dst
    int d2h_cols(void *dst, size_t dpitch, const void *src, size_t width, size_t ncols, hipStream_t st) {
        if (!ok) {
            HIPCK(hipMemcpy2DAsync(dst, dpitch, src, width, width, ncols, hipMemcpyDeviceToHost, st));
            HIPCK(hipStreamSynchronize(st));
            return 0;
        }
        const size_t total = width * ncols;
        int k = 0, wk = -1;
        size_t woff = 0, wlen = 0;
        for (size_t off = 0; off < total || wk >= 0; off += CHUNK, k ^= 1) {
            int nk = -1;
            size_t noff = 0, nlen = 0;
            if (off < total) {              // buf[k] is free: its previous contents were copied out
                nlen = std::min(CHUNK, total - off);
                noff = off;
                HIPCK(hipMemcpyAsync(buf[k], (const char *)src + off, nlen, hipMemcpyDeviceToHost, st));
                HIPCK(hipEventRecord(ev[k], st));
                nk = k;
            }
            if (wk >= 0) {
                HIPCK(hipEventSynchronize(ev[wk]));
                unpack(dst, dpitch, buf[wk], width, woff, wlen);
            }
            wk = nk; woff = noff; wlen = nlen;
        }
        pend[0] = pend[1] = false;          // every DMA of the stream so far has completed
        return 0;
    }
    int d2h(void *dst, const void *src, size_t bytes, hipStream_t st) { return d2h_cols(dst, bytes, src, bytes, 1, st); }
    int drain() {                           // uploads issued so far have completed
        for (int k = 0; k < 2; ++k)
            if (pend[k]) { HIPCK(hipEventSynchronize(ev[k])); pend[k] = false; }
        return 0;
    }
};

// Stagers lent to the host entry points (kept for the life of the process; one per concurrent call)
struct StagerLease {
    static std::mutex &mu() { static std::mutex m; return m; }
    static std::vector<HostStager *> &pool() { static std::vector<HostStager *> p; return p; }
    HostStager *s;
    StagerLease() {
        {
            std::lock_guard<std::mutex> lk(mu());
            if (!pool().empty()) { s = pool().back(); pool().pop_back(); return; }
        }
        s = new HostStager();
    }
    ~StagerLease() {
        std::lock_guard<std::mutex> lk(mu());
        pool().push_back(s);
    }
    HostStager *operator->() { return s; }
};

template <typename T> T s_zero_host() {
    T z;
    memset(&z, 0, sizeof z);
    return z;
}

int block_nt(int ku, int kl) {
    const int k = std::max(ku, kl);
    if (k <= 16) return 1;
    if (k <= 32) return 2;
    if (k <= 64) return 4;
    if (k <= 128) return 8;
    if (k <= 256) return 16;
    return 0;
}

// Groups whose factor state does not fit in registers run the global-memory kernels of
// spike_kernels_wide.inc: b = 256, and complex FP64 at b = 128.
bool wide_group(int tcode, int nt) { return nt == 16 || (tcode == 3 && nt == 8); }
// block rows per conversion workgroup: 64 for the narrow bands (several whole blocks, so that a
// workgroup has work enough to amortize its launch), otherwise a 16-row slab of one block (8 for
// complex FP64 at b = 256, where three 16-row slabs exceed the LDS)
constexpr int conv_rows(int tbytes, int nt) { return nt <= 2 ? 64 : ((tbytes == 16 && nt == 16) ? 8 : 16); }

// tcode: 0 float, 1 double, 2 complex float, 3 complex double
int type_code(int dtype, int precision) { return (dtype == SPIKE_DTYPE_Z ? 2 : 0) + (precision == SPIKE_PREC_FP64 ? 1 : 0); }
size_t type_bytes(int tcode) { return tcode == 0 ? 4 : (tcode == 3 ? 16 : 8); }
size_t h_bytes(int dtype) { return dtype == SPIKE_DTYPE_Z ? 16 : 8; }

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

#define NT_SWITCH(NTV, BODY)                                 \
    switch (NTV) {                                           \
    case 1: { constexpr int NT = 1; BODY; } break;           \
    case 2: { constexpr int NT = 2; BODY; } break;           \
    case 4: { constexpr int NT = 4; BODY; } break;           \
    case 8: { constexpr int NT = 8; BODY; } break;           \
    default: break;                                          \
    }

// every block size; the launch helpers route the wide groups to the global-memory kernels
#define NTA_SWITCH(NTV, BODY)                                \
    switch (NTV) {                                           \
    case 1: { constexpr int NT = 1; BODY; } break;           \
    case 2: { constexpr int NT = 2; BODY; } break;           \
    case 4: { constexpr int NT = 4; BODY; } break;           \
    case 8: { constexpr int NT = 8; BODY; } break;           \
    case 16: { constexpr int NT = 16; BODY; } break;         \
    default: break;                                          \
    }

// the wide groups: (T, NT) pairs of wide_group()
#define WIDE_SWITCH(TC, NTV, BODY)                                              \
    {                                                                           \
    if ((NTV) == 16) {                                                          \
        constexpr int NT = 16;                                                  \
        switch (TC) {                                                           \
        case 0: { typedef float T; BODY; } break;                               \
        case 1: { typedef double T; BODY; } break;                              \
        case 2: { typedef zf T; BODY; } break;                                  \
        case 3: { typedef zd T; BODY; } break;                                  \
        default: break;                                                         \
        }                                                                       \
    } else if ((TC) == 3 && (NTV) == 8) {                                       \
        constexpr int NT = 8;                                                   \
        typedef zd T;                                                           \
        BODY;                                                                   \
    }                                                                           \
    }

#define T_SWITCH(TC, BODY)                                   \
    switch (TC) {                                            \
    case 0: { typedef float T; BODY; } break;                \
    case 1: { typedef double T; BODY; } break;               \
    case 2: { typedef zf T; BODY; } break;                   \
    case 3: { typedef zd T; BODY; } break;                   \
    default: break;                                          \
    }

// ============================================================================ autotuner
namespace {

// Factor workgroups the device runs at once for block size 16*NT and factor type tcode: CUs x
// workgroups per CU of k_factor (HIP occupancy query; 0 when no device is visible).  Cached.
int factor_slots(int tcode, int nt) {
    static std::mutex mu;
    static std::map<int, int> cache;
    std::lock_guard<std::mutex> lk(mu);
    const int key = tcode * 16 + nt;
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int dev = 0, cus = 0, per = 0;
    const void *fn = nullptr;
    if (wide_group(tcode, nt)) WIDE_SWITCH(tcode, nt, { fn = (const void *)(k_factor_gm<T, NT>); })
    else T_SWITCH(tcode, NT_SWITCH(nt, { fn = (const void *)(k_factor<T, NT>); }));
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        !fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, nt * 64, 0) != hipSuccess)
        cus = per = 0;
    (void)hipGetLastError();
    const int slots = cus * per;
    if (getenv("HSPIKE_VERBOSE"))
        fprintf(stderr, "[hspike] factor slots tcode=%d NT=%d: %d CUs x %d workgroups\n", tcode, nt, cus, per);
    cache[key] = slots;
    return slots;
}

}  // namespace

// Partition count from (n, k, nrhs, dtype, precision) and the device only, so a system's
// schedule never depends on the batch it is solved in.  Each partition is one workgroup chain of
// nb blocks.  About 12 blocks per partition keep the spike couplings small enough for the reduced
// block-Jacobi iteration to converge in a few sweeps; a large system is capped at one round of
// the factor kernel (the workgroups the device runs at once): past that, more partitions only
// add rounds of the same chains plus a larger reduced system.
// Throughput mode (a batch of many systems: far more partitions than the device runs at once, so
// the per-partition work -- V spikes, tips, interfaces of the reduced system, its sweeps -- is
// what the partition count buys) uses partitions of 96 blocks; measured on the 4096-system batch
// (ms per step): 12 blocks 253.5, 24 227.3, 48 214.2, 96 208.8.
extern "C" int spike_autotune_partitions_tuned(int n, int ku, int kl, int nrhs, int dtype, int precision, int tune) {
    (void)nrhs;
    const int NT = block_nt(ku, kl);
    if (NT == 0 || n <= 0) return 1;
    const int b = 16 * NT;
    static const long bpp_latency = [] {
        const char *v = getenv("HSPIKE_BLOCKS_PER_PART");
        const long x = v ? atol(v) : 0;
        return x > 0 ? x : 12L;
    }();
    static const long bpp_throughput = [] {
        const char *v = getenv("HSPIKE_BLOCKS_PER_PART_BATCH");
        const long x = v ? atol(v) : 0;
        return x > 0 ? x : 96L;
    }();
    const long blocks_per_part = tune == SPIKE_TUNE_THROUGHPUT ? bpp_throughput : bpp_latency;
    const long rows_target = blocks_per_part * b;
    long P = n / rows_target;
    const long slots = factor_slots(type_code(dtype, precision), NT);
    if (slots > 0) P = std::min(P, slots);
    const long min_rows = std::max(2 * b, ku + kl);
    P = std::min(P, (long)n / std::max(1L, min_rows));
    P = std::max(1L, std::min(P, 8192L));
    return (int)P;
}

extern "C" int spike_autotune_partitions(int n, int ku, int kl, int nrhs, int dtype, int precision) {
    return spike_autotune_partitions_tuned(n, ku, kl, nrhs, dtype, precision, SPIKE_TUNE_LATENCY);
}

// ============================================================================ engine
namespace {

struct SysHost {
    int n = 0;
    int64_t nnz = 0;
    const int *d_rowptr = nullptr, *d_colind = nullptr;
    const void *d_aij = nullptr;
    int ku = 0, kl = 0, P = 1, NT = 1;
    int group = -1, gidx = 0;     // group and index inside the group
    int nrhs = 0;
    const void *d_b = nullptr;
    int ldb = 0;
    void *d_x = nullptr;
    int ldx = 0;
    double maxrel = 0.0;
    double maxrel_before = 0.0;             // largest final residual of the earlier RHS chunks
    int refine_iters = 0, red_iters = 0;
    double red_update = 0.0;
    bool red_direct = false;                // reduced system solved by the band LU (until the next factor)
};

struct Group {
    int NT = 1, b = 16, tcode = 1;
    std::vector<int> sys;                   // batch indices
    std::vector<PartDesc> parts;
    std::vector<IfaceDesc> ifs;
    std::vector<int> part_iface;
    std::vector<RowChunk> chunks;
    std::vector<int> chunk0;                // per system in the group: first chunk, +1 end
    std::vector<SysDev> sysdev;
    std::vector<int> sys_part0;             // global partition of each system
    int64_t nblk = 0;
    int64_t rows_pad = 0;                   // nblk * b
    DevBuf d_parts, d_ifs, d_part_iface, d_chunks, d_chunk0, d_sysdev, d_active, d_red_active;
    DevBuf blkmap;                          // partition of each block
    DevBuf D, U, C, R, Bt, Ct;              // factor type
    DevBuf Wf, Vf;                          // forward-swept W spike (all blocks) and V spike (last two blocks)
    DevBuf wstop;                           // per partition: blocks of Wf the W sweep wrote (decayed spike)
    DevBuf rstop;                           // per partition: blocks of R the factor formed (decayed R_j)
    DevBuf vt, vb, wt, wb, T;               // H type
    DevBuf yt, yb, red0, red1, scratch, rwork, xrm, partials, norms, conv, pivinfo, status;
    DevBuf items_vw, items_yb, items_yr, items_rb, items_rr;
    int n_items_vw = 0, n_items_y = 0, n_items_r = 0;
    int ct_vw = 4, ct_y = 1;                // column tiles per sweep workgroup (V/W items, RHS items)
    int ymax = 0;
    std::vector<uint64_t> solve_key;        // descriptors of the last solve chunk (reused when equal)
    std::vector<uint64_t> plan_key;         // members, sizes, bandwidths, P and CSR arrays of the plan
    bool plan_uploaded = false;             // d_parts .. d_sysdev hold this plan
    bool vw_uploaded = false;               // items_vw holds this plan's V/W items
    int red_last = 0;                       // buffer (0: red0, 1: red1) holding the last iterate
    std::vector<double> piv_host;
    // direct solve of the reduced system: R of every system in rd_sys assembled in band storage
    // and factored (banded LU with partial pivoting); the factors stay valid until the next factor
    std::vector<int> rd_sys;                // group-local systems, in the order of rd_bs
    bool rd_factored = false;
    std::vector<BandSys> rd_bs;
    std::vector<RedDirect> rd_desc_h;
    DevBuf rd_desc, rd_pass, rd_bsd, rd_ab, rd_piv, rd_x, rd_info, rd_items;
    // device-side solve control
    DevBuf d_sysstat, d_summary, d_direct;
    std::vector<int> direct_h;              // 1: the system's reduced system is solved directly
    bool direct_dirty = true;
};

struct Engine {
    spike_opts_t o;
    int dtype = SPIKE_DTYPE_D;
    hipStream_t stream = nullptr;
    std::vector<SysHost> sys;
    std::vector<std::unique_ptr<Group>> groups;
    int ngroups = 0;
    bool factored = false;
    DevBuf d_bw, d_bwchunks, d_bwsys;
    // small transfers staged through pinned host memory and moved by a copy kernel on the
    // engine's stream: uploads take consecutive slices (no wait), a readback waits for the
    // stream and then recycles the whole area
    PinBuf pin;
    size_t pin_off = 0;
    static constexpr size_t PIN_BYTES = (size_t)1 << 20, PIN_MAX = (size_t)256 << 10;
    int copy_words(const void *src, void *dst, size_t bytes) {
        const int64_t nw = (int64_t)(bytes / 4);
        if (nw <= 0) return 0;
        const int grid = (int)std::min<int64_t>((nw + 255) / 256, 256);
        hipLaunchKernelGGL(k_copy_words, dim3(grid), dim3(256), 0, stream, (const uint32_t *)src, (uint32_t *)dst, nw);
        HIPCK(hipGetLastError());
        return 0;
    }
    // device <- host, asynchronous; falls back to a DMA copy (synchronous) for large arrays
    int put_small(DevBuf &d, const void *host, size_t bytes) {
        int e;
        if ((e = d.ensure(bytes))) return e;
        if (!bytes) return 0;
        if (bytes > PIN_MAX || bytes % 4) return upload(d, host, bytes, stream);
        if ((e = pin.ensure(PIN_BYTES))) return e;
        if (pin_off + bytes > pin.bytes) { HIPCK(hipStreamSynchronize(stream)); pin_off = 0; }
        memcpy(pin.h + pin_off, host, bytes);
        if ((e = copy_words(pin.d + pin_off, d.p, bytes))) return e;
        pin_off += (bytes + 255) & ~(size_t)255;
        return 0;
    }
    // host <- device for several arrays with one wait: list of (host dst, device src, bytes)
    int get_small(std::initializer_list<std::tuple<void *, const void *, size_t>> parts) {
        int e;
        if ((e = pin.ensure(PIN_BYTES))) return e;
        size_t need = 0;
        bool dma = false;
        for (auto &p : parts) {
            need += (std::get<2>(p) + 255) & ~(size_t)255;
            dma = dma || std::get<2>(p) % 4;
        }
        if (dma || need > PIN_MAX) {
            for (auto &p : parts)
                if (std::get<2>(p)) HIPCK(hipMemcpyAsync(std::get<0>(p), std::get<1>(p), std::get<2>(p), hipMemcpyDeviceToHost, stream));
            HIPCK(hipStreamSynchronize(stream));
            pin_off = 0;
            return 0;
        }
        if (pin_off + need > pin.bytes) { HIPCK(hipStreamSynchronize(stream)); pin_off = 0; }
        size_t off = pin_off;
        for (auto &p : parts) {
            if ((e = copy_words(std::get<1>(p), pin.d + off, std::get<2>(p)))) return e;
            off += (std::get<2>(p) + 255) & ~(size_t)255;
        }
        HIPCK(hipStreamSynchronize(stream));
        off = pin_off;
        for (auto &p : parts) {
            memcpy(std::get<0>(p), pin.h + off, std::get<2>(p));
            off += (std::get<2>(p) + 255) & ~(size_t)255;
        }
        pin_off = 0;
        return 0;
    }
    int zero_ints(DevBuf &d, size_t count) {
        if (!count) return 0;
        hipLaunchKernelGGL((k_fill<int>), dim3(1), dim3(256), 0, stream, d.as<int>(), (int64_t)count, 0);
        HIPCK(hipGetLastError());
        return 0;
    }
    std::vector<uint64_t> bw_key;           // systems described by d_bwchunks / d_bwsys
    size_t bw_nchunks = 0;                  // row chunks in d_bwchunks
    DevBuf own_rowptr, own_colind, own_aij;   // host-matrix copies (spike_set_matrix_host)
    DevBuf stage_b, stage_x;                   // host solve staging
    spike_stats_t st;
    hipEvent_t ev[16];
    bool events = false;
    Engine() {
        spike_default_opts(&o);
        memset(&st, 0, sizeof st);
    }
    ~Engine() {
        if (events)
            for (auto &e : ev) (void)hipEventDestroy(e);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (own_stream) (void)hipStreamDestroy(own_stream);
    }
    int init_events() {
        if (events) return 0;
        for (auto &e : ev) HIPCK(hipEventCreate(&e));
        events = true;
        return 0;
    }
    float ev_ms(int a, int b) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev[a], ev[b]) != hipSuccess) { (void)hipGetLastError(); return 0.f; }
        return ms;
    }
    size_t device_bytes() const {
        size_t s = 0;
        for (int gi = 0; gi < ngroups; ++gi) {
            const Group &g = *groups[gi];
            const DevBuf *bs[] = {&g.D, &g.U, &g.C, &g.R, &g.Wf, &g.Vf, &g.wstop, &g.rstop, &g.Bt, &g.Ct, &g.vt, &g.vb,
                                  &g.wt, &g.wb, &g.T, &g.yt, &g.yb, &g.red0, &g.red1, &g.scratch, &g.rwork, &g.xrm,
                                  &g.rd_ab, &g.rd_x};
            for (auto b : bs) s += b->bytes;
        }
        return s;
    }
    // Graph mode captures work on a non-blocking stream of the engine's own when the caller gave
    // none (the legacy default stream cannot be captured); the caller's default-stream work is
    // waited for first, as the default stream itself would have ordered it.
    int prepare_stream() {
        if (o.graph && !stream) {
            if (!own_stream) HIPCK(hipStreamCreateWithFlags(&own_stream, hipStreamNonBlocking));
            stream = own_stream;
        }
        if (own_stream && stream == own_stream) HIPCK(hipStreamSynchronize(nullptr));
        return 0;
    }
    int factor();
    int reduced_direct(Group &g, const std::vector<int> &which, DevBuf &xo, int ymax);
    int direct_pass(Group &g, int ymax);
    int launch_jacobi(Group &g, int ncc, int ymax, const DevBuf &xo, DevBuf &xn);
    int run_passes(int ymax, bool capture, double rtol, int max_refine, double &t_y, double &t_red, double &t_rec,
                   double &t_res, double &t_norm, int &passes);
    int read_stats(bool *incomplete = nullptr);
    int solve_chunk(int c0, int m);
    int solve_setup(int c0, int ymax, bool &cached);
    int solve_graph(int ymax);
    std::vector<uint64_t> graph_key(int ymax) const;
    // graph of one single-chunk solve (replayed while graph_key() is unchanged); the budget of
    // passes and sweep pairs comes from the last eager solve
    hipGraphExec_t gexec = nullptr;
    std::vector<uint64_t> gkey;
    int graph_passes = 0, graph_pairs = 0;
    int eager_pairs = 0;                     // largest sweep-pair count of a pass in the last eager solve
    hipStream_t own_stream = nullptr;        // non-blocking stream for graph mode when the caller gave none
    bool graph_replayed = false;
    int solve(int nrhs_hint);
};

// ---------------------------------------------------------------------------- dispatch
// Column tiles per sweep workgroup: CT * NT * sizeof(T) <= 64 keeps the two LDS exchange panels
// (2 * 16*CT * b elements) within 32 KB and the per-lane registers (CT accumulator tiles per
// product, NT-deep operand slices) low enough for two or more waves per SIMD.
constexpr bool sweep_fits(int tbytes, int nt, int ct) { return ct * nt * tbytes <= 64; }
int ct_max(int tcode, int NT) {
    int ct = 4;
    while (ct > 1 && !sweep_fits((int)type_bytes(tcode), NT, ct)) ct >>= 1;
    return ct;
}
int ct_for(int ncols, int ctmax) { return std::min(ctmax, ncols <= 16 ? 1 : (ncols <= 32 ? 2 : 4)); }

// wide groups (wide_group(): b = 256, complex FP64 at b = 128) sweep one column tile per item

template <typename T, int NT, int CT, bool VW>
void launch_sweep(int nitems, hipStream_t s, const SweepItem *items, const PartDesc *parts, const T *D, const T *C,
                  const T *U, const T *R, const T *Bt, const T *Ct, typename tr<T>::H *vt, typename tr<T>::H *vb,
                  typename tr<T>::H *wt, typename tr<T>::H *wb, typename tr<T>::H *yt, typename tr<T>::H *yb, T *Wf, T *Vf,
                  T *wsave, int ymax, const int *active, int *wstop, const int *rstop) {
    if constexpr (wide_kernels<T, NT>()) {
        if constexpr (CT == 1)
            hipLaunchKernelGGL((k_sweep_gm<T, NT, VW>), dim3(nitems), dim3(NT * 64), 0, s, items, parts, D, C, U, R, Bt, Ct,
                               vt, vb, wt, wb, yt, yb, Wf, Vf, (Q4<T> *)wsave, ymax, active, wstop, rstop);
    } else if constexpr (sweep_fits((int)sizeof(T), NT, CT))
        hipLaunchKernelGGL((k_sweep<T, NT, CT, VW>), dim3(nitems), dim3(NT * 64), 0, s, items, parts, D, C, U, R, Bt, Ct, vt,
                           vb, wt, wb, yt, yb, Wf, Vf, (Q4<T> *)wsave, ymax, active, wstop, rstop);
}

template <typename T, int NT, int CT>
void launch_recover(int nitems, hipStream_t s, const SweepItem *items, const PartDesc *parts, const T *U, const T *Wf,
                    const T *Vf, const typename tr<T>::H *red, const int *part_iface, int ymax, const T *wsave,
                    int accumulate, const int *active, const int *wstop) {
    if constexpr (wide_kernels<T, NT>()) {
        if constexpr (CT == 1)
            hipLaunchKernelGGL((k_recover_gm<T, NT>), dim3(nitems), dim3(NT * 64), 0, s, items, parts, U, Wf, Vf, red,
                               part_iface, ymax, (const Q4<T> *)wsave, accumulate, active, wstop);
    } else if constexpr (sweep_fits((int)sizeof(T), NT, CT))
        hipLaunchKernelGGL((k_recover<T, NT, CT>), dim3(nitems), dim3(NT * 64), 0, s, items, parts, U, Wf, Vf, red,
                           part_iface, ymax, (const Q4<T> *)wsave, accumulate, active, wstop);
}

// FP64 residual with TW column lanes per row: the narrowest of 4, 8, 16 covering the chunk's
// columns (all chunks share the partials stride `ymax`)
// (TW lanes per row, CPL columns per lane): pairs of columns per lane when the solver keeps the
// row-major iterate (row stride even), and the fewest lanes per row that cover the columns
int launch_residual(int dtype, int mcols, int ymax, bool rowmajor, unsigned nchunks, hipStream_t s,
                    const RowChunk *chunks, const SysDev *sys, double *partials, const int *active) {
    const int cpl = (rowmajor && dtype == SPIKE_DTYPE_D) ? 2 : 1;     // complex pairs cost too many registers
    const int need = (mcols + cpl - 1) / cpl;
    const int tw = need <= 4 ? 4 : (need <= 8 ? 8 : 16);
#define RES_LAUNCH(H, TW, CPL) hipLaunchKernelGGL((k_residual<H, TW, CPL>), dim3(nchunks), dim3(256), 0, s, chunks, sys, partials, ymax, active)
#define RES_TW(H, CPL) do { if (tw == 4) RES_LAUNCH(H, 4, CPL); else if (tw == 8) RES_LAUNCH(H, 8, CPL); else RES_LAUNCH(H, 16, CPL); } while (0)
    if (dtype == SPIKE_DTYPE_Z) {
        if (cpl == 2) RES_TW(zd, 2); else RES_TW(zd, 1);
    } else {
        if (cpl == 2) RES_TW(double, 2); else RES_TW(double, 1);
    }
#undef RES_TW
#undef RES_LAUNCH
    HIPCK(hipGetLastError());
    return 0;
}

#define CT_SWITCH(CTV, BODY)                                 \
    switch (CTV) {                                           \
    case 1: { constexpr int CT = 1; BODY; } break;           \
    case 2: { constexpr int CT = 2; BODY; } break;           \
    case 4: { constexpr int CT = 4; BODY; } break;           \
    default: break;                                          \
    }

template <typename T>
int fill_zero(DevBuf &d, size_t count, hipStream_t s) {
    if (!count) return 0;
    const int grid = (int)std::min<size_t>((count + 255) / 256, 65536);
    hipLaunchKernelGGL((k_fill<T>), dim3(grid), dim3(256), 0, s, d.as<T>(), (int64_t)count, s_zero_host<T>());
    HIPCK(hipGetLastError());
    return 0;
}

}  // namespace

// ============================================================================ factor
namespace {

int Engine::factor() {
    if (sys.empty()) { set_err("no matrix bound%s", ""); return SPIKE_ERR_INVALID_ARG; }
    int e = init_events();
    if (e || (e = prepare_stream())) return e;
    const double t0 = now_ms();
    hipStream_t s = stream;
    const int nsys = (int)sys.size();
    // a new factorization invalidates the direct solves of the reduced systems
    for (auto &S : sys) S.red_direct = false;
    for (auto &gp : groups) { gp->rd_factored = false; gp->rd_sys.clear(); gp->direct_h.clear(); gp->direct_dirty = true; }
    HIPCK(hipEventRecord(ev[0], s));

    // ---- validation and bandwidth of every system (one launch)
    {
        // the chunk and system descriptors of the scan are rebuilt only when the systems' arrays
        // or sizes change (a batch factored again keeps them)
        std::vector<uint64_t> key;
        key.reserve(3 * (size_t)nsys);
        for (int i = 0; i < nsys; ++i)
            key.insert(key.end(), {(uint64_t)(uintptr_t)sys[i].d_rowptr, (uint64_t)(uintptr_t)sys[i].d_colind,
                                   (uint64_t)sys[i].n});
        if (key != bw_key) {
            std::vector<RowChunk> ch;
            std::vector<SysDev> sd(nsys);
            for (int i = 0; i < nsys; ++i) {
                memset(&sd[i], 0, sizeof(SysDev));
                sd[i].rowptr = sys[i].d_rowptr;
                sd[i].colind = sys[i].d_colind;
                sd[i].n = sys[i].n;
                for (int r = 0; r < sys[i].n; r += BW_ROWS) ch.push_back({i, r});
            }
            if ((e = upload(d_bwchunks, ch.data(), ch.size() * sizeof(RowChunk), s))) return e;
            if ((e = upload(d_bwsys, sd.data(), sd.size() * sizeof(SysDev), s))) return e;
            bw_key = key;
            bw_nchunks = ch.size();
        }
        if ((e = d_bw.ensure(3 * sizeof(int) * nsys)) || (e = zero_ints(d_bw, 3 * (size_t)nsys))) return e;
        if (bw_nchunks)
            hipLaunchKernelGGL(k_bandwidth, dim3((unsigned)bw_nchunks), dim3(256), 0, s, d_bwchunks.as<RowChunk>(),
                               d_bwsys.as<SysDev>(), d_bw.as<int>());
        HIPCK(hipGetLastError());
        std::vector<int> bw(3 * nsys);
        if ((e = get_small({{bw.data(), d_bw.p, bw.size() * sizeof(int)}}))) return e;
        for (int i = 0; i < nsys; ++i) {
            if (bw[3 * i + 2]) { set_err("malformed CSR in system %s%ld", "", (long)i); return SPIKE_ERR_MALFORMED_CSR; }
            SysHost &S = sys[i];
            S.ku = bw[3 * i];
            S.kl = bw[3 * i + 1];
            S.NT = block_nt(S.ku, S.kl);
            if (S.NT == 0) {
                set_err("bandwidth %ld exceeds this build's limit", (long)std::max(S.ku, S.kl));
                return SPIKE_ERR_UNSUPPORTED;
            }
            S.P = o.partitions > 0 ? o.partitions
                                   : spike_autotune_partitions_tuned(S.n, S.ku, S.kl, S.nrhs, dtype, o.precision, o.tune);
            if (S.P > S.n) S.P = std::max(1, S.n);
            // a partition needs at least max(ku, kl) rows, so that its coupling blocks B_i, C_i
            // fit inside it; its top and bottom tips may overlap (short partitions of the
            // reference's fixtures at P = 4), which leaves R nonsingular
            const int nreg = S.n / S.P;
            if (S.P > 1 && (nreg < std::max(S.ku, S.kl) || nreg < 1)) {
                set_err("partitions too small for the bandwidth (rows per partition %ld)", (long)nreg);
                return SPIKE_ERR_INVALID_ARG;
            }
        }
    }
    HIPCK(hipEventRecord(ev[1], s));

    // ---- group systems by block size
    const int tcode = type_code(dtype, o.precision);
    const size_t ts = type_bytes(tcode), hs = h_bytes(dtype);
    int nts[5] = {1, 2, 4, 8, 16};
    ngroups = 0;
    for (int q = 0; q < 5; ++q) {
        std::vector<int> members;
        for (int i = 0; i < nsys; ++i)
            if (sys[i].NT == nts[q]) members.push_back(i);
        if (members.empty()) continue;
        if ((int)groups.size() <= ngroups) groups.emplace_back(new Group());
        Group &g = *groups[ngroups];
        // a plan already uploaded for the same members (sizes, bandwidths, partition counts and
        // CSR arrays) is kept as it is: a batch factored again skips the rebuild of its
        // partition, interface and row-chunk lists
        std::vector<uint64_t> pk;
        pk.reserve(8 * members.size() + 2);
        pk.push_back((uint64_t)nts[q]);
        pk.push_back((uint64_t)tcode);
        for (int i : members) {
            const SysHost &S = sys[i];
            pk.insert(pk.end(), {(uint64_t)i, (uint64_t)S.n, (uint64_t)S.P, (uint64_t)S.ku, (uint64_t)S.kl,
                                 (uint64_t)(uintptr_t)S.d_rowptr, (uint64_t)(uintptr_t)S.d_colind,
                                 (uint64_t)(uintptr_t)S.d_aij});
        }
        if (g.plan_uploaded && pk == g.plan_key) {
            for (int gs = 0; gs < (int)members.size(); ++gs) {
                sys[members[gs]].group = ngroups;
                sys[members[gs]].gidx = gs;
            }
            ++ngroups;
            continue;
        }
        g.plan_key = pk;
        std::vector<PartDesc> old_parts;
        old_parts.swap(g.parts);
        const bool same_type = g.NT == nts[q] && g.tcode == tcode && g.sys == members;
        std::vector<SysDev> old_sysdev;
        old_sysdev.swap(g.sysdev);
        g.NT = nts[q];
        g.b = 16 * g.NT;
        g.tcode = tcode;
        g.sys = members;
        g.ifs.clear(); g.part_iface.clear(); g.chunks.clear(); g.chunk0.clear();
        g.sys_part0.clear();
        int64_t blk = 0;
        for (int gs = 0; gs < (int)members.size(); ++gs) {
            SysHost &S = sys[members[gs]];
            S.group = ngroups;
            S.gidx = gs;
            const int P = S.P, nreg = S.n / P;
            const int part0 = (int)g.parts.size(), if0 = (int)g.ifs.size();
            g.sys_part0.push_back(part0);
            for (int p = 0; p < P; ++p) {
                PartDesc d;
                memset(&d, 0, sizeof d);
                d.sys = gs; d.p = p; d.P = P;
                d.r0 = p * nreg;
                d.ni = (p == P - 1) ? S.n - d.r0 : nreg;
                d.nb = ceil_div(d.ni, g.b);
                d.ku = S.ku; d.kl = S.kl;
                d.blk = blk;
                blk += d.nb;
                g.parts.push_back(d);
                g.part_iface.push_back(if0 + p);
            }
            for (int i = 0; i + 1 < P; ++i) {
                IfaceDesc f;
                memset(&f, 0, sizeof f);
                f.sys = gs; f.i = i; f.P = P; f.ku = S.ku; f.kl = S.kl; f.part = part0 + i;
                g.ifs.push_back(f);
            }
            g.chunk0.push_back((int)g.chunks.size());
            for (int r = 0; r < S.n; r += RES_ROWS) g.chunks.push_back({gs, r});
            SysDev sd;
            memset(&sd, 0, sizeof sd);
            sd.rowptr = S.d_rowptr; sd.colind = S.d_colind; sd.aij = S.d_aij;
            sd.n = S.n; sd.ku = S.ku; sd.kl = S.kl; sd.P = P; sd.part0 = part0; sd.nreg = nreg;
            g.sysdev.push_back(sd);
        }
        g.chunk0.push_back((int)g.chunks.size());
        g.nblk = blk;
        g.rows_pad = blk * g.b;
        bool same = same_type && old_parts.size() == g.parts.size() && old_sysdev.size() == g.sysdev.size() &&
                    !memcmp(old_parts.data(), g.parts.data(), g.parts.size() * sizeof(PartDesc));
        for (size_t gs = 0; same && gs < g.sysdev.size(); ++gs)
            same = old_sysdev[gs].rowptr == g.sysdev[gs].rowptr && old_sysdev[gs].colind == g.sysdev[gs].colind &&
                   old_sysdev[gs].aij == g.sysdev[gs].aij;
        if (same) g.sysdev.swap(old_sysdev);    // keep the solve fields of the uploaded copy
        g.plan_uploaded = same && g.plan_uploaded;
        if (!g.plan_uploaded) g.solve_key.clear();
        ++ngroups;
    }

    for (int gi = 0; gi < ngroups; ++gi) {
        Group &g = *groups[gi];
        const size_t bb = (size_t)g.b * g.b;
        const size_t np = g.parts.size(), nif = g.ifs.size();
        if (!g.plan_uploaded) {
            if ((e = upload(g.d_parts, g.parts.data(), np * sizeof(PartDesc), s))) return e;
            if ((e = upload(g.d_ifs, g.ifs.data(), nif * sizeof(IfaceDesc), s))) return e;
            if ((e = upload(g.d_part_iface, g.part_iface.data(), g.part_iface.size() * sizeof(int), s))) return e;
            if ((e = upload(g.d_chunks, g.chunks.data(), g.chunks.size() * sizeof(RowChunk), s))) return e;
            if ((e = upload(g.d_chunk0, g.chunk0.data(), g.chunk0.size() * sizeof(int), s))) return e;
            if ((e = upload(g.d_sysdev, g.sysdev.data(), g.sysdev.size() * sizeof(SysDev), s))) return e;
            if ((e = g.blkmap.ensure(g.nblk * sizeof(int)))) return e;
            hipLaunchKernelGGL(k_blkmap, dim3((unsigned)np), dim3(256), 0, s, g.d_parts.as<PartDesc>(), g.blkmap.as<int>());
            HIPCK(hipGetLastError());
            g.vw_uploaded = false;
        }
        DevBuf *tb[] = {&g.D, &g.U, &g.C, &g.R, &g.Wf};
        for (auto b : tb)
            if ((e = b->ensure(g.nblk * bb * ts))) return e;
        if ((e = g.Vf.ensure(np * 2 * bb * ts))) return e;
        if ((e = g.wstop.ensure(np * sizeof(int))) || (e = g.rstop.ensure(np * sizeof(int)))) return e;
        if ((e = g.Bt.ensure(np * bb * ts)) || (e = g.Ct.ensure(np * bb * ts))) return e;
        DevBuf *hb[] = {&g.vt, &g.vb, &g.wt, &g.wb};
        for (auto b : hb)
            if ((e = b->ensure(np * bb * hs))) return e;
        if ((e = g.T.ensure(std::max<size_t>(nif, 1) * bb * hs))) return e;
        if ((e = g.pivinfo.ensure(np * 2 * sizeof(double)))) return e;
        if ((e = g.status.ensure(4 * sizeof(int))) || (e = zero_ints(g.status, 4))) return e;
        // zero the coupling tiles, which the conversion scatters into (fill kernels; no memset
        // nodes in code a caller may capture); the conversion writes D whole, and U and C whole
        // except, with FP32 factors, the column groups no entry can reach (never read)
        T_SWITCH(tcode, {
            if ((e = fill_zero<T>(g.Bt, np * bb, s)) || (e = fill_zero<T>(g.Ct, np * bb, s))) return e;
        });
        g.plan_uploaded = true;
    }
    HIPCK(hipEventRecord(ev[2], s));
    for (int gi = 0; gi < ngroups; ++gi) {
        Group &g = *groups[gi];
        const int np = (int)g.parts.size();
        // the wide groups' factor reads B_j from the Wf array and writes -Uq_j to its own block
        const bool wide = wide_group(g.tcode, g.NT);
#define CONV_LAUNCH                                                                                            \
    {                                                                                                          \
        constexpr int RS = conv_rows((int)sizeof(T), NT);                                                      \
        hipLaunchKernelGGL((k_csr_to_blocks<T, NT, RS>), dim3((unsigned)((g.nblk * 16 * NT + RS - 1) / RS)), dim3(256), 0, s, \
                           g.blkmap.as<int>(), g.d_sysdev.as<SysDev>(), g.d_parts.as<PartDesc>(), (int64_t)g.nblk, g.D.as<T>(), \
                           wide ? g.Wf.as<T>() : g.U.as<T>(), g.C.as<T>(), g.Bt.as<T>(), g.Ct.as<T>(),         \
                           g.status.as<int>());                                                                \
    }
        if (g.NT == 16) WIDE_SWITCH(g.tcode, 16, CONV_LAUNCH)
        else T_SWITCH(g.tcode, NT_SWITCH(g.NT, CONV_LAUNCH));
#undef CONV_LAUNCH
        HIPCK(hipGetLastError());
        (void)np;
    }
    HIPCK(hipEventRecord(ev[3], s));
    for (int gi = 0; gi < ngroups; ++gi) {
        Group &g = *groups[gi];
        const int np = (int)g.parts.size();
        if (wide_group(g.tcode, g.NT)) {
            WIDE_SWITCH(g.tcode, g.NT, {
                hipLaunchKernelGGL((k_factor_gm<T, NT>), dim3(np), dim3(NT * 64), 0, s, g.d_parts.as<PartDesc>(),
                                   g.D.as<T>(), g.U.as<T>(), (const T *)g.C.as<T>(), g.R.as<T>(), (const T *)g.Wf.as<T>(),
                                   g.pivinfo.as<double>(), g.rstop.as<int>());
            });
        } else {
            // one copy of R's rows where it adds workgroups and the launch has more partitions than
            // the device runs at once with two
            T_SWITCH(g.tcode, NT_SWITCH(g.NT, {
                bool one = false;
                if constexpr (factor_one_copy<T, NT>()) one = np > factor_slots(g.tcode, NT);
                if (one)
                    hipLaunchKernelGGL((k_factor<T, NT, factor_one_copy<T, NT>() ? 1 : 2>), dim3(np), dim3(NT * 64), 0, s,
                                       g.d_parts.as<PartDesc>(), g.D.as<T>(), g.U.as<T>(), (const T *)g.C.as<T>(),
                                       g.R.as<T>(), g.pivinfo.as<double>(), (const int *)nullptr, g.rstop.as<int>());
                else
                    hipLaunchKernelGGL((k_factor<T, NT, 2>), dim3(np), dim3(NT * 64), 0, s, g.d_parts.as<PartDesc>(),
                                       g.D.as<T>(), g.U.as<T>(), (const T *)g.C.as<T>(), g.R.as<T>(),
                                       g.pivinfo.as<double>(), (const int *)nullptr, g.rstop.as<int>());
            }));
        }
        HIPCK(hipGetLastError());
    }
    HIPCK(hipEventRecord(ev[4], s));
    for (int gi = 0; gi < ngroups; ++gi) {
        Group &g = *groups[gi];
        const int np = (int)g.parts.size();
        if (wide_group(g.tcode, g.NT)) continue;          // k_factor_gm forms R
        T_SWITCH(g.tcode, NT_SWITCH(g.NT, {
            if constexpr (!factor_forms_r<T, NT>())        // else the factor kernel formed R
                hipLaunchKernelGGL((k_rtop<T, NT>), dim3(np), dim3(NT * 64), 0, s, g.d_parts.as<PartDesc>(),
                                   (const T *)g.U.as<T>(), g.R.as<T>(), (const int *)nullptr, g.rstop.as<int>());
        }));
        HIPCK(hipGetLastError());
    }
    HIPCK(hipEventRecord(ev[5], s));
    // ---- V / W spike tips (partitions of systems with P > 1); wstop starts at zero (fill kernel)
    for (int gi = 0; gi < ngroups; ++gi) {
        Group &g = *groups[gi];
        if ((e = fill_zero<int>(g.wstop, g.parts.size(), s))) return e;
        if (g.vw_uploaded) {
            if (g.n_items_vw) {
                T_SWITCH(g.tcode, NTA_SWITCH(g.NT, CT_SWITCH(g.ct_vw, {
                    typedef typename tr<T>::H H;
                    (launch_sweep<T, NT, CT, true>)(g.n_items_vw, s, g.items_vw.as<SweepItem>(), g.d_parts.as<PartDesc>(),
                                            g.D.as<T>(), g.C.as<T>(), g.U.as<T>(), g.R.as<T>(), g.Bt.as<T>(),
                                            g.Ct.as<T>(), g.vt.as<H>(), g.vb.as<H>(), g.wt.as<H>(), g.wb.as<H>(),
                                            (H *)nullptr, (H *)nullptr, g.Wf.as<T>(), g.Vf.as<T>(), (T *)nullptr, 0,
                                            (const int *)nullptr, g.wstop.as<int>(), (const int *)g.rstop.as<int>());
                })));
                HIPCK(hipGetLastError());
            }
            continue;
        }
        std::vector<SweepItem> items;
        // a spike has at most b = 16 NT columns: NT column tiles cover it (more would multiply zeros)
        g.ct_vw = wide_group(g.tcode, g.NT) ? 1 : std::min(ct_max(g.tcode, g.NT), g.NT);
        const int cpi = 16 * g.ct_vw;
        for (int q = 0; q < (int)g.parts.size(); ++q) {
            const PartDesc &d = g.parts[q];
            if (d.P == 1) continue;
            SweepItem it;
            memset(&it, 0, sizeof it);
            it.part = q;
            if (d.p < d.P - 1)
                for (int c = 0; c < d.ku; c += cpi) { it.kind = ITEM_V; it.col0 = c; it.ncols = std::min(cpi, d.ku - c); items.push_back(it); }
            if (d.p > 0)
                for (int c = 0; c < d.kl; c += cpi) { it.kind = ITEM_W; it.col0 = c; it.ncols = std::min(cpi, d.kl - c); items.push_back(it); }
        }
        g.n_items_vw = (int)items.size();
        g.vw_uploaded = true;
        if (items.empty()) continue;
        if ((e = upload(g.items_vw, items.data(), items.size() * sizeof(SweepItem), s))) return e;
        T_SWITCH(g.tcode, NTA_SWITCH(g.NT, CT_SWITCH(g.ct_vw, {
            typedef typename tr<T>::H H;
            (launch_sweep<T, NT, CT, true>)(g.n_items_vw, s, g.items_vw.as<SweepItem>(), g.d_parts.as<PartDesc>(),
                                    g.D.as<T>(), g.C.as<T>(), g.U.as<T>(), g.R.as<T>(), g.Bt.as<T>(),
                                    g.Ct.as<T>(), g.vt.as<H>(), g.vb.as<H>(), g.wt.as<H>(), g.wb.as<H>(),
                                    (H *)nullptr, (H *)nullptr, g.Wf.as<T>(), g.Vf.as<T>(), (T *)nullptr, 0,
                                    (const int *)nullptr, g.wstop.as<int>(), (const int *)g.rstop.as<int>());
        })));
        HIPCK(hipGetLastError());
    }
    HIPCK(hipEventRecord(ev[6], s));
    // ---- reduced-system setup: T_i = (I - Wt_{i+1} Vb_i)^{-1}
    for (int gi = 0; gi < ngroups; ++gi) {
        Group &g = *groups[gi];
        if (g.ifs.empty()) continue;
        if (g.NT == 16) {                                   // global-memory inversion (16-wide panels)
            if (dtype == SPIKE_DTYPE_Z)
                hipLaunchKernelGGL((k_red_setup_gm<zd, 16>), dim3((unsigned)g.ifs.size()), dim3(16 * 64), 0, s,
                                   g.d_ifs.as<IfaceDesc>(), (const zd *)g.vb.p, (const zd *)g.wt.p, g.T.as<zd>(),
                                   g.status.as<int>() + 1, (const int *)nullptr);
            else
                hipLaunchKernelGGL((k_red_setup_gm<double, 16>), dim3((unsigned)g.ifs.size()), dim3(16 * 64), 0, s,
                                   g.d_ifs.as<IfaceDesc>(), (const double *)g.vb.p, (const double *)g.wt.p,
                                   g.T.as<double>(), g.status.as<int>() + 1, (const int *)nullptr);
        } else if (dtype == SPIKE_DTYPE_Z) {
            NT_SWITCH(g.NT, {
                hipLaunchKernelGGL((k_red_setup_mfma<zd, NT>), dim3((unsigned)g.ifs.size()), dim3(NT * 64), 0, s,
                                   g.d_ifs.as<IfaceDesc>(), (const zd *)g.vb.p, (const zd *)g.wt.p, g.T.as<zd>(),
                                   g.status.as<int>() + 1, (const int *)nullptr);
            });
        } else {
            NT_SWITCH(g.NT, {
                hipLaunchKernelGGL((k_red_setup_mfma<double, NT>), dim3((unsigned)g.ifs.size()), dim3(NT * 64), 0, s,
                                   g.d_ifs.as<IfaceDesc>(), (const double *)g.vb.p, (const double *)g.wt.p,
                                   g.T.as<double>(), g.status.as<int>() + 1, (const int *)nullptr);
            });
        }
        HIPCK(hipGetLastError());
    }
    HIPCK(hipEventRecord(ev[7], s));
    // ---- status: malformed entries, singular pivots
    for (int gi = 0; gi < ngroups; ++gi) {
        Group &g = *groups[gi];
        int stv[4];
        g.piv_host.resize(g.parts.size() * 2);
        if ((e = get_small({{stv, g.status.p, sizeof stv}, {g.piv_host.data(), g.pivinfo.p, g.piv_host.size() * sizeof(double)}})))
            return e;
        if (stv[0]) { set_err("matrix entry outside the detected band%s", ""); return SPIKE_ERR_MALFORMED_CSR; }
        for (size_t q = 0; q < g.parts.size(); ++q)
            if (g.piv_host[2 * q + 1] != 0.0) {
                set_err("zero or non-finite pivot in partition %s%ld", "", (long)q);
                return SPIKE_ERR_SINGULAR;
            }
        // a 4x4 panel pivot of the MFMA inversion was zero: redo T_i with full partial pivoting
        // (the whole block in LDS; b = 256 pivots inside 16 x 16 panels already)
        const size_t smem = ((size_t)g.b * g.b + 2 * g.b) * hs;
        if (stv[1] && !g.ifs.empty() && smem <= 160 * 1024) {
            if ((e = zero_ints(g.status, 4))) return e;
            if (dtype == SPIKE_DTYPE_Z) {
                HIPCK(hipFuncSetAttribute((const void *)k_red_setup<zd>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
                hipLaunchKernelGGL(k_red_setup<zd>, dim3((unsigned)g.ifs.size()), dim3(256), smem, s, g.d_ifs.as<IfaceDesc>(),
                                   (const zd *)g.vb.p, (const zd *)g.wt.p, g.T.as<zd>(), g.b, g.status.as<int>() + 1,
                                   (const int *)nullptr);
            } else {
                HIPCK(hipFuncSetAttribute((const void *)k_red_setup<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
                hipLaunchKernelGGL(k_red_setup<double>, dim3((unsigned)g.ifs.size()), dim3(256), smem, s,
                                   g.d_ifs.as<IfaceDesc>(), (const double *)g.vb.p, (const double *)g.wt.p, g.T.as<double>(),
                                   g.b, g.status.as<int>() + 1, (const int *)nullptr);
            }
            HIPCK(hipGetLastError());
            if ((e = get_small({{stv, g.status.p, sizeof stv}}))) return e;
        }
        if (stv[1]) { set_err("singular reduced-system block%s", ""); return SPIKE_ERR_SINGULAR; }
    }
    factored = true;
    st.ms_convert = ev_ms(2, 3);
    st.ms_factor = ev_ms(3, 4);
    st.ms_rtop = ev_ms(4, 5);
    st.ms_spikes = ev_ms(5, 6);
    st.ms_red_setup = ev_ms(6, 7);
    st.ms_factor_total = now_ms() - t0;
    return SPIKE_SUCCESS;
}

// ============================================================================ direct reduced solve
// R of the group-local systems `which` (assembled and factored once per factorization; the set
// only grows between factorizations), its right-hand side from the RHS tips of this pass, the
// band solve, and the solution scattered into the interface iterate `xo` the recovery reads.
int Engine::reduced_direct(Group &g, const std::vector<int> &which, DevBuf &xo, int ymax) {
    hipStream_t s = stream;
    int e;
    const size_t hs = h_bytes(dtype);
    bool grow = false;
    for (int gs : which)
        if (std::find(g.rd_sys.begin(), g.rd_sys.end(), gs) == g.rd_sys.end()) { g.rd_sys.push_back(gs); grow = true; }
    const int nd = (int)g.rd_sys.size();
    if (grow || !g.rd_factored) {
        g.rd_bs.assign(nd, BandSys{});
        g.rd_desc_h.assign(nd, RedDirect{});
        int64_t ab = 0, pv = 0;
        int maxP = 1;
        size_t lds = 0;
        for (int q = 0; q < nd; ++q) {
            const SysHost &S = sys[g.sys[g.rd_sys[q]]];
            int klr, kur;
            red_band_shape(S.ku, S.kl, klr, kur);
            BandSys &B = g.rd_bs[q];
            B.n = S.P * (S.ku + S.kl);
            B.kl = klr; B.ku = kur; B.ldab = 2 * klr + kur + 1;
            B.ab = ab; B.piv = pv;
            ab += (int64_t)B.n * B.ldab;
            pv += B.n;
            RedDirect &D = g.rd_desc_h[q];
            D.part0 = g.sys_part0[g.rd_sys[q]];
            D.iface0 = g.part_iface[D.part0];
            D.P = S.P; D.ku = S.ku; D.kl = S.kl; D.band = q;
            maxP = std::max(maxP, S.P);
            lds = std::max(lds, (size_t)(std::max(klr, 1) + klr + kur + 1) * hs);
        }
        if (lds > 64 * 1024) {
            set_err("reduced system bandwidth %ld too wide for the direct solver", (long)(lds / hs));
            return SPIKE_ERR_UNSUPPORTED;
        }
        if ((e = g.rd_ab.ensure((size_t)ab * hs)) || (e = g.rd_piv.ensure((size_t)pv * sizeof(int))) ||
            (e = g.rd_info.ensure(nd * sizeof(int))))
            return e;
        if (dtype == SPIKE_DTYPE_Z) { if ((e = fill_zero<zd>(g.rd_ab, (size_t)ab, s))) return e; }
        else if ((e = fill_zero<double>(g.rd_ab, (size_t)ab, s))) return e;
        if ((e = upload(g.rd_bsd, g.rd_bs.data(), nd * sizeof(BandSys), s)) ||
            (e = upload(g.rd_desc, g.rd_desc_h.data(), nd * sizeof(RedDirect), s)))
            return e;
        if (dtype == SPIKE_DTYPE_Z) {
            hipLaunchKernelGGL(k_red_band<zd>, dim3(maxP, nd), dim3(256), 0, s, g.rd_desc.as<RedDirect>(), g.rd_bsd.as<BandSys>(),
                               (const zd *)g.vt.p, (const zd *)g.vb.p, (const zd *)g.wt.p, (const zd *)g.wb.p, g.b,
                               g.rd_ab.as<zd>());
            hipLaunchKernelGGL(k_gbtrf<zd>, dim3(nd), dim3(GB_THREADS), lds, s, g.rd_bsd.as<BandSys>(), g.rd_ab.as<zd>(),
                               g.rd_piv.as<int>(), g.rd_info.as<int>());
        } else {
            hipLaunchKernelGGL(k_red_band<double>, dim3(maxP, nd), dim3(256), 0, s, g.rd_desc.as<RedDirect>(),
                               g.rd_bsd.as<BandSys>(), (const double *)g.vt.p, (const double *)g.vb.p,
                               (const double *)g.wt.p, (const double *)g.wb.p, g.b, g.rd_ab.as<double>());
            hipLaunchKernelGGL(k_gbtrf<double>, dim3(nd), dim3(GB_THREADS), lds, s, g.rd_bsd.as<BandSys>(),
                               g.rd_ab.as<double>(), g.rd_piv.as<int>(), g.rd_info.as<int>());
        }
        HIPCK(hipGetLastError());
        std::vector<int> info(nd);
        if ((e = get_small({{info.data(), g.rd_info.p, nd * sizeof(int)}}))) return e;
        for (int q = 0; q < nd; ++q)
            if (info[q]) {
                set_err("singular reduced system (zero pivot in column %ld of R)", (long)(info[q] - 1));
                return SPIKE_ERR_SINGULAR;
            }
        g.rd_factored = true;
    }
    // right-hand sides of this pass for the systems in `which`, band solve, scatter into the iterate
    int64_t xe = 0;
    int maxnr = 1, nrow = 0;
    std::vector<int> items;
    std::vector<int> items_g;
    std::vector<RedDirect> pass;
    const size_t lds_cap = 48 * 1024;
    size_t lds = 0;
    for (int q = 0; q < nd; ++q) {
        if (std::find(which.begin(), which.end(), g.rd_sys[q]) == which.end()) continue;
        pass.push_back(g.rd_desc_h[q]);
        BandSys &B = g.rd_bs[q];
        B.m = ymax; B.ldx = B.n; B.x = xe;
        xe += (int64_t)B.n * ymax;
        maxnr = std::max(maxnr, B.n);
        nrow = std::max(nrow, (g.rd_desc_h[q].P - 1) * (g.rd_desc_h[q].ku + g.rd_desc_h[q].kl) * ymax);
        const size_t per_col = (size_t)B.n * hs;
        if (per_col <= lds_cap) {
            const int mc = (int)std::min<size_t>({64, (size_t)ymax, lds_cap / per_col});
            for (int c = 0; c < ymax; c += mc) items.insert(items.end(), {q, c, std::min(mc, ymax - c)});
            lds = std::max(lds, per_col * mc);
        } else {
            for (int c = 0; c < ymax; c += 64) items_g.insert(items_g.end(), {q, c, std::min(64, ymax - c)});
        }
    }
    const int npass = (int)pass.size();
    if (!npass) return SPIKE_SUCCESS;
    if ((e = g.rd_x.ensure((size_t)xe * hs))) return e;
    if ((e = upload(g.rd_bsd, g.rd_bs.data(), nd * sizeof(BandSys), s)) ||
        (e = upload(g.rd_pass, pass.data(), npass * sizeof(RedDirect), s)))
        return e;
    std::vector<int> all(items);
    all.insert(all.end(), items_g.begin(), items_g.end());
    if ((e = upload(g.rd_items, all.data(), all.size() * sizeof(int), s))) return e;
    const unsigned rgrid = (unsigned)std::min<int64_t>(((int64_t)maxnr * ymax + 255) / 256, 4096);
    const unsigned sgrid = (unsigned)std::min<int64_t>(((int64_t)nrow + 255) / 256 + 1, 4096);
    const int ni = (int)items.size() / 3, ng = (int)items_g.size() / 3;
    if (dtype == SPIKE_DTYPE_Z) {
        hipLaunchKernelGGL(k_red_rhs<zd>, dim3(rgrid, npass), dim3(256), 0, s, g.rd_pass.as<RedDirect>(), g.rd_bsd.as<BandSys>(),
                           (const zd *)g.yt.p, (const zd *)g.yb.p, g.b, ymax, g.rd_x.as<zd>());
        if (ni) hipLaunchKernelGGL((k_gbtrs<zd, true>), dim3(ni), dim3(256), lds, s, g.rd_bsd.as<BandSys>(), g.rd_items.as<int>(),
                                   (const zd *)g.rd_ab.p, (const int *)g.rd_piv.p, g.rd_x.as<zd>());
        if (ng) hipLaunchKernelGGL((k_gbtrs<zd, false>), dim3(ng), dim3(256), 0, s, g.rd_bsd.as<BandSys>(),
                                   g.rd_items.as<int>() + 3 * ni, (const zd *)g.rd_ab.p, (const int *)g.rd_piv.p,
                                   g.rd_x.as<zd>());
        hipLaunchKernelGGL(k_red_scatter<zd>, dim3(sgrid, npass), dim3(256), 0, s, g.rd_pass.as<RedDirect>(),
                           g.rd_bsd.as<BandSys>(), (const zd *)g.rd_x.p, g.b, ymax, xo.as<zd>());
    } else {
        hipLaunchKernelGGL(k_red_rhs<double>, dim3(rgrid, npass), dim3(256), 0, s, g.rd_pass.as<RedDirect>(),
                           g.rd_bsd.as<BandSys>(), (const double *)g.yt.p, (const double *)g.yb.p, g.b, ymax,
                           g.rd_x.as<double>());
        if (ni) hipLaunchKernelGGL((k_gbtrs<double, true>), dim3(ni), dim3(256), lds, s, g.rd_bsd.as<BandSys>(),
                                   g.rd_items.as<int>(), (const double *)g.rd_ab.p, (const int *)g.rd_piv.p,
                                   g.rd_x.as<double>());
        if (ng) hipLaunchKernelGGL((k_gbtrs<double, false>), dim3(ng), dim3(256), 0, s, g.rd_bsd.as<BandSys>(),
                                   g.rd_items.as<int>() + 3 * ni, (const double *)g.rd_ab.p, (const int *)g.rd_piv.p,
                                   g.rd_x.as<double>());
        hipLaunchKernelGGL(k_red_scatter<double>, dim3(sgrid, npass), dim3(256), 0, s, g.rd_pass.as<RedDirect>(),
                           g.rd_bsd.as<BandSys>(), (const double *)g.rd_x.p, g.b, ymax, xo.as<double>());
    }
    HIPCK(hipGetLastError());
    return SPIKE_SUCCESS;
}

// ============================================================================ solve
// Workspaces, per-system views and sweep items of an RHS chunk (uploaded only when they change).
int Engine::solve_setup(int c0, int ymax, bool &cached_all) {
    hipStream_t s = stream;
    int e;
    const size_t hs = h_bytes(dtype);
    cached_all = true;
    // ---- per-group workspaces, descriptors and items
    for (int gi = 0; gi < ngroups; ++gi) {
        Group &g = *groups[gi];
        const size_t np = g.parts.size(), nif = g.ifs.size(), ngs = g.sys.size();
        const size_t ts = type_bytes(g.tcode);
        g.ymax = ymax;
        if ((e = g.yt.ensure(np * g.b * ymax * hs)) || (e = g.yb.ensure(np * g.b * ymax * hs)) ||
            (e = g.red0.ensure(std::max<size_t>(nif, 1) * 2 * g.b * ymax * hs)) ||
            (e = g.red1.ensure(std::max<size_t>(nif, 1) * 2 * g.b * ymax * hs)) ||
            (e = g.scratch.ensure((size_t)g.rows_pad * 16 * ceil_div(ymax, 16) * ts)) ||
            (e = g.partials.ensure(g.chunks.size() * ymax * 2 * sizeof(double))) ||
            (e = g.norms.ensure(ngs * ymax * 2 * sizeof(double))) ||
            (e = g.conv.ensure(std::max<size_t>(nif, 1) * ceil_div(ymax, 16) * 2 * sizeof(double))))
            return e;
        // row-major copy of each system's iterate (the residual's x gathers): a row stride of the
        // system's own columns in this chunk, rounded up to even (a lane's column pair is one
        // aligned load), so a system with few columns gathers from densely packed rows
        size_t rrows = 0, xrows = 0;
        auto sys_ldrm = [&](const SysHost &S) { return (std::max(1, std::min(ymax, S.nrhs - c0)) + 1) & ~1; };
        for (size_t gs = 0; gs < ngs; ++gs) {
            const SysHost &S = sys[g.sys[gs]];
            rrows += S.n;
            xrows += (size_t)S.n * sys_ldrm(S);
        }
        if ((e = g.rwork.ensure(rrows * ymax * hs)) || (e = g.xrm.ensure(xrows * hs))) return e;
        std::vector<uint64_t> key = {(uint64_t)c0, (uint64_t)ymax, (uint64_t)(uintptr_t)g.rwork.p,
                                     (uint64_t)(uintptr_t)g.xrm.p};
        for (size_t gs = 0; gs < ngs; ++gs) {
            const SysHost &S = sys[g.sys[gs]];
            key.insert(key.end(), {(uint64_t)(uintptr_t)S.d_b, (uint64_t)S.ldb, (uint64_t)(uintptr_t)S.d_x,
                                   (uint64_t)S.ldx, (uint64_t)S.nrhs});
        }
        const bool cached = key == g.solve_key;
        size_t roff = 0, xoff = 0;
        for (size_t gs = 0; gs < ngs; ++gs) {
            SysHost &S = sys[g.sys[gs]];
            SysDev &sd = g.sysdev[gs];
            const int ldrm = sys_ldrm(S);
            const int m = std::max(0, std::min(ymax, S.nrhs - c0));
            sd.m = m;
            sd.ldb = S.ldb; sd.ldx = S.ldx; sd.ldr = S.n;
            sd.bptr = (const char *)S.d_b + (size_t)c0 * S.ldb * hs;
            sd.xptr = (char *)S.d_x + (size_t)c0 * S.ldx * hs;
            sd.rptr = (char *)g.rwork.p + roff * ymax * hs;
            sd.xrm = (char *)g.xrm.p + xoff * hs;
            sd.ldrm = ldrm;
            roff += S.n;
            xoff += (size_t)S.n * ldrm;
            S.refine_iters = 0;
        }
        if ((e = g.d_active.ensure(ngs * sizeof(int))) || (e = g.d_red_active.ensure(ngs * sizeof(int))) ||
            (e = g.d_sysstat.ensure(ngs * sizeof(SysStat))) || (e = g.d_summary.ensure(4 * sizeof(int))))
            return e;
        if (g.direct_h.size() != ngs) {              // systems on the direct reduced solve (set anew after a factor)
            g.direct_h.assign(ngs, 0);
            if (o.reduced_solver == SPIKE_RED_DIRECT)
                for (size_t gs = 0; gs < ngs; ++gs) {
                    SysHost &S = sys[g.sys[gs]];
                    if (S.P > 1 && S.ku + S.kl > 0) { g.direct_h[gs] = 1; S.red_direct = true; }
                }
            g.direct_dirty = true;
        }
        if (g.direct_dirty) {
            if ((e = put_small(g.d_direct, g.direct_h.data(), ngs * sizeof(int)))) return e;
            g.direct_dirty = false;
        }
        if (cached) continue;
        cached_all = false;
        if ((e = upload(g.d_sysdev, g.sysdev.data(), ngs * sizeof(SysDev), s))) return e;
        // items: Y sweeps and recoveries, from b (first pass) and from r (refinement passes)
        g.ct_y = wide_group(g.tcode, g.NT) ? 1 : ct_for(ymax, ct_max(g.tcode, g.NT));
        const int cpi = 16 * g.ct_y;
        std::vector<SweepItem> yb_, yr_, rb_, rr_;
        for (int q = 0; q < (int)np; ++q) {
            const PartDesc &d = g.parts[q];
            const SysDev &sd = g.sysdev[d.sys];
            const SysHost &S = sys[g.sys[d.sys]];
            for (int c = 0; c < sd.m; c += cpi) {
                SweepItem it;
                memset(&it, 0, sizeof it);
                it.part = q; it.kind = ITEM_Y; it.col0 = c; it.ncols = std::min(cpi, sd.m - c);
                it.dst = (char *)sd.xptr + (size_t)c * S.ldx * hs; it.lddst = S.ldx;
                it.dst_rm = (void *)sd.xrm; it.ldrm = sd.ldrm;
                it.src = (const char *)sd.bptr + (size_t)c * S.ldb * hs; it.ldsrc = S.ldb;
                yb_.push_back(it);        // every partition: the recovery reads its forward sweep
                rb_.push_back(it);
                it.src = (const char *)sd.rptr + (size_t)c * S.n * hs; it.ldsrc = S.n;
                yr_.push_back(it);
                rr_.push_back(it);
            }
        }
        g.n_items_y = (int)yb_.size();
        g.n_items_r = (int)rb_.size();
        if ((e = upload(g.items_yb, yb_.data(), yb_.size() * sizeof(SweepItem), s)) ||
            (e = upload(g.items_yr, yr_.data(), yr_.size() * sizeof(SweepItem), s)) ||
            (e = upload(g.items_rb, rb_.data(), rb_.size() * sizeof(SweepItem), s)) ||
            (e = upload(g.items_rr, rr_.data(), rr_.size() * sizeof(SweepItem), s)))
            return e;
        g.solve_key = key;
    }
    return SPIKE_SUCCESS;
}

int Engine::solve_chunk(int c0, int ymax) {
    hipStream_t s = stream;
    int e;
    const double rtol = (o.precision == SPIKE_PREC_FP64) ? o.reduced_tol : std::max(o.reduced_tol, 1e-9);
    const int max_refine = std::max(0, o.max_refine);
    double t_y = 0, t_red = 0, t_rec = 0, t_res = 0, t_norm = 0;
    int passes = 0;
    bool cached;
    if ((e = solve_setup(c0, ymax, cached))) return e;
    for (int gi = 0; gi < ngroups; ++gi) {
        Group &g = *groups[gi];
        const int ngs = (int)g.sys.size();
        hipLaunchKernelGGL(k_stat_init, dim3((ngs + 255) / 256), dim3(256), 0, s, g.d_sysstat.as<SysStat>(), ngs,
                           (const SysDev *)g.d_sysdev.as<SysDev>(), g.d_active.as<int>(), c0 == 0 ? 1 : 0);
        HIPCK(hipGetLastError());
    }
    e = run_passes(ymax, false, rtol, max_refine, t_y, t_red, t_rec, t_res, t_norm, passes);
    if (e) return e;
    st.ms_ysweep = t_y;
    st.ms_reduced = t_red;
    st.ms_recover = t_rec;
    st.ms_residual = t_res;
    st.ms_norms = t_norm;
    st.solve_passes = passes;
    st.launch_groups = ngroups;
    return read_stats();
}

// Per-system results of the last solve chunk into the host view.  *incomplete (optional) is set
// when a system would have needed more than the launches made: still active after the last pass,
// or a reduced iteration that ran out of sweeps or diverged.
int Engine::read_stats(bool *incomplete) {
    if (incomplete) *incomplete = false;
    for (int gi = 0; gi < ngroups; ++gi) {
        Group &g = *groups[gi];
        const size_t ngs = g.sys.size();
        std::vector<SysStat> h(ngs);
        std::vector<int> act(ngs);
        int e;
        if ((e = get_small({{h.data(), g.d_sysstat.p, ngs * sizeof(SysStat)}, {act.data(), g.d_active.p, ngs * sizeof(int)}})))
            return e;
        for (size_t gs = 0; gs < ngs; ++gs) {
            SysHost &S = sys[g.sys[gs]];
            S.maxrel = h[gs].maxrel;
            S.refine_iters = h[gs].refine_iters;
            S.red_iters = h[gs].red_iters;
            S.red_update = h[gs].red_update;
            if (incomplete && (act[gs] || h[gs].unfinished || h[gs].red_active || h[gs].diverged)) *incomplete = true;
        }
    }
    return SPIKE_SUCCESS;
}

std::vector<uint64_t> Engine::graph_key(int ymax) const {
    auto bits = [](double v) { uint64_t u; memcpy(&u, &v, sizeof u); return u; };
    std::vector<uint64_t> k = {(uint64_t)ymax, (uint64_t)graph_passes, (uint64_t)graph_pairs, (uint64_t)dtype,
                               (uint64_t)o.max_refine, bits(o.tol), bits(o.reduced_tol), (uint64_t)o.precision,
                               (uint64_t)o.reduced_maxit, (uint64_t)(uintptr_t)stream, (uint64_t)ngroups};
    for (int gi = 0; gi < ngroups; ++gi) {
        const Group &g = *groups[gi];
        k.insert(k.end(), {(uint64_t)g.NT, (uint64_t)g.tcode, (uint64_t)g.ct_y, (uint64_t)g.n_items_y, (uint64_t)g.n_items_r,
                           (uint64_t)g.ifs.size(), (uint64_t)g.chunks.size(), (uint64_t)g.sys.size(), (uint64_t)g.ymax});
        const DevBuf *bufs[] = {&g.D, &g.U, &g.C, &g.R, &g.Wf, &g.Vf, &g.wstop, &g.rstop, &g.Bt, &g.Ct, &g.vt, &g.vb, &g.wt, &g.wb, &g.T,
                                &g.yt, &g.yb, &g.red0, &g.red1, &g.scratch, &g.partials, &g.norms, &g.conv, &g.rwork,
                                &g.xrm, &g.d_active, &g.d_red_active, &g.d_sysstat, &g.d_direct, &g.d_summary,
                                &g.items_yb, &g.items_yr, &g.items_rb, &g.items_rr, &g.d_parts, &g.d_ifs, &g.d_part_iface,
                                &g.d_chunks, &g.d_chunk0, &g.d_sysdev};
        for (auto b : bufs) k.push_back((uint64_t)(uintptr_t)b->p);
        k.insert(k.end(), g.solve_key.begin(), g.solve_key.end());
    }
    return k;
}

// One single-chunk solve as a graph replay: the launches of the eager solve (stat init, then
// graph_passes passes of graph_pairs sweep pairs) captured once and replayed while graph_key()
// holds.  SPIKE_ERR_UNSUPPORTED sends the caller to the eager solve: no budget yet, descriptors
// just changed, a system on the direct reduced solve, or a replay that needed more launches than
// the budget (then the graph is dropped and the eager solve sets a new budget).
int Engine::solve_graph(int ymax) {
    if (graph_passes <= 0 || graph_pairs <= 0) return SPIKE_ERR_UNSUPPORTED;
    for (int gi = 0; gi < ngroups; ++gi)
        if (std::any_of(groups[gi]->direct_h.begin(), groups[gi]->direct_h.end(), [](int d) { return d != 0; }))
            return SPIKE_ERR_UNSUPPORTED;
    bool cached;
    int e;
    if ((e = solve_setup(0, ymax, cached))) return e;
    if (!cached) return SPIKE_ERR_UNSUPPORTED;
    const double rtol = (o.precision == SPIKE_PREC_FP64) ? o.reduced_tol : std::max(o.reduced_tol, 1e-9);
    const int max_refine = std::max(0, o.max_refine);
    const std::vector<uint64_t> key = graph_key(ymax);
    if (!gexec || key != gkey) {
        if (gexec) { (void)hipGraphExecDestroy(gexec); gexec = nullptr; }
        hipGraph_t graph = nullptr;
        HIPCK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int rc = SPIKE_SUCCESS;
        for (int gi = 0; gi < ngroups && !rc; ++gi) {
            Group &g = *groups[gi];
            const int ngs = (int)g.sys.size();
            hipLaunchKernelGGL(k_stat_init, dim3((ngs + 255) / 256), dim3(256), 0, stream, g.d_sysstat.as<SysStat>(), ngs,
                               (const SysDev *)g.d_sysdev.p, g.d_active.as<int>(), 1);
            if (hipGetLastError() != hipSuccess) rc = SPIKE_ERR_DEVICE;
        }
        double t_y = 0, t_red = 0, t_rec = 0, t_res = 0, t_norm = 0;
        int passes = 0;
        if (!rc) rc = run_passes(ymax, true, rtol, max_refine, t_y, t_red, t_rec, t_res, t_norm, passes);
        const hipError_t ce = hipStreamEndCapture(stream, &graph);
        if (rc || ce != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipGetLastError();
            if (!rc) set_err("graph capture failed (%s)", hipGetErrorString(ce));
            return rc ? rc : SPIKE_ERR_DEVICE;
        }
        const hipError_t ie = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) {
            (void)hipGetLastError();
            gexec = nullptr;
            set_err("graph instantiation failed (%s)", hipGetErrorString(ie));
            return SPIKE_ERR_DEVICE;
        }
        gkey = key;
    }
    HIPCK(hipGraphLaunch(gexec, stream));
    bool incomplete = false;
    if ((e = read_stats(&incomplete))) return e;
    if (incomplete) {
        (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        return SPIKE_ERR_UNSUPPORTED;
    }
    st.ms_ysweep = st.ms_reduced = st.ms_recover = st.ms_residual = st.ms_norms = 0.0;   // not timed per stage
    st.solve_passes = graph_passes;
    st.launch_groups = ngroups;
    graph_replayed = true;
    return SPIKE_SUCCESS;
}

// The solve and refinement passes of one RHS chunk.  Every decision is taken on the device
// (k_red_check, k_refine_check); kernels skip finished systems.  Eager (capture = false): the
// host reads a 16-byte summary after each pair of sweeps and each pass to stop early and to move
// diverging reduced systems to the direct solve.  Capture: a fixed budget of graph_passes passes
// and graph_pairs pairs of sweeps, no host call at all.
int Engine::run_passes(int ymax, bool capture, double rtol, int max_refine, double &t_y, double &t_red, double &t_rec,
                       double &t_res, double &t_norm, int &passes) {
    hipStream_t s = stream;
    int e;
    const int npass = capture ? graph_passes : max_refine + 1;
    auto rec = [&](int i) -> int { if (!capture) HIPCK(hipEventRecord(ev[i], s)); return 0; };
    for (int it = 0; it < npass; ++it) {
        if ((e = rec(8))) return e;
        // ---- y tips
        for (int gi = 0; gi < ngroups; ++gi) {
            Group &g = *groups[gi];
            if (!g.n_items_y) continue;
            const DevBuf &items = it == 0 ? g.items_yb : g.items_yr;
            T_SWITCH(g.tcode, NTA_SWITCH(g.NT, CT_SWITCH(g.ct_y, {
                typedef typename tr<T>::H H;
                (launch_sweep<T, NT, CT, false>)(g.n_items_y, s, items.as<SweepItem>(), g.d_parts.as<PartDesc>(), g.D.as<T>(),
                                        g.C.as<T>(), g.U.as<T>(), g.R.as<T>(), g.Bt.as<T>(), g.Ct.as<T>(), (H *)nullptr,
                                        (H *)nullptr, (H *)nullptr, (H *)nullptr, g.yt.as<H>(), g.yb.as<H>(),
                                        (T *)nullptr, (T *)nullptr, g.scratch.as<T>(), g.ymax, g.d_active.as<int>(),
                                        (int *)nullptr, (const int *)g.rstop.as<int>());
            })));
            HIPCK(hipGetLastError());
        }
        if ((e = rec(9))) return e;
        // ---- reduced system: block-Jacobi sweeps in FP64, two per convergence check; the last
        // iterate of every system is in red0 (an even number of sweeps)
        for (int gi = 0; gi < ngroups; ++gi) {
            Group &g = *groups[gi];
            if (g.ifs.empty()) continue;
            const int nif = (int)g.ifs.size(), ngs = (int)g.sys.size();
            const int ncc = ceil_div(ymax, 16);
            const int slot = 2 * g.b * ymax;
            const dim3 zg((unsigned)std::min(nif, 65536));
            if (dtype == SPIKE_DTYPE_Z)
                hipLaunchKernelGGL(k_red_zero<zd>, zg, dim3(256), 0, s, (const IfaceDesc *)g.d_ifs.p, nif,
                                   (const int *)g.d_active.p, g.red0.as<zd>(), slot);
            else
                hipLaunchKernelGGL(k_red_zero<double>, zg, dim3(256), 0, s, (const IfaceDesc *)g.d_ifs.p, nif,
                                   (const int *)g.d_active.p, g.red0.as<double>(), slot);
            hipLaunchKernelGGL(k_red_begin, dim3((ngs + 255) / 256), dim3(256), 0, s, g.d_sysstat.as<SysStat>(), ngs,
                               (const SysDev *)g.d_sysdev.p, (const int *)g.d_active.p, (const int *)g.d_direct.p,
                               g.d_red_active.as<int>());
            HIPCK(hipGetLastError());
            const int maxpairs = capture ? graph_pairs : std::max(1, o.reduced_maxit / 2);
            bool sweeping = std::find(g.direct_h.begin(), g.direct_h.end(), 0) != g.direct_h.end();
            int summary[4] = {0, 0, 0, 0};
            int pairs_done = 0;
            for (int pr = 0; sweeping && pr < maxpairs; ++pr, ++pairs_done) {
                for (int rep = 0; rep < 2; ++rep) {
                    const DevBuf &xo = rep == 0 ? g.red0 : g.red1;
                    DevBuf &xn = rep == 0 ? g.red1 : g.red0;
                    if ((e = launch_jacobi(g, ncc, ymax, xo, xn))) return e;
                }
                hipLaunchKernelGGL(k_red_check, dim3((unsigned)ngs), dim3(64), 0, s, g.d_sysstat.as<SysStat>(), ngs,
                                   (const SysDev *)g.d_sysdev.p, (const int *)g.d_part_iface.p, (const double *)g.conv.p,
                                   ncc, rtol, 2 * (pr + 1), g.d_red_active.as<int>());
                HIPCK(hipGetLastError());
                if (!capture) {
                    hipLaunchKernelGGL(k_stat_summary, dim3(1), dim3(256), 0, s, (const SysStat *)g.d_sysstat.p, ngs,
                                       (const int *)g.d_active.p, g.d_summary.as<int>());
                    HIPCK(hipGetLastError());
                    if ((e = get_small({{summary, g.d_summary.p, sizeof summary}}))) return e;
                    sweeping = summary[0] > 0;
                }
            }
            g.red_last = 0;
            if (!capture) eager_pairs = std::max(eager_pairs, pairs_done);
            if (capture || o.reduced_solver != SPIKE_RED_AUTO) {
                if (!capture && std::any_of(g.direct_h.begin(), g.direct_h.end(), [](int d) { return d != 0; }))
                    if ((e = direct_pass(g, ymax))) return e;
                continue;
            }
            // diverged, or still sweeping after reduced_maxit sweeps: the direct solve from now on
            if (summary[0] > 0 || summary[1] > 0) {
                std::vector<SysStat> h(ngs);
                std::vector<int> act(ngs);
                if ((e = get_small({{h.data(), g.d_sysstat.p, ngs * sizeof(SysStat)}, {act.data(), g.d_active.p, ngs * sizeof(int)}})))
                    return e;
                for (int gs = 0; gs < ngs; ++gs)
                    if (act[gs] && (h[gs].diverged || h[gs].red_active)) {
                        g.direct_h[gs] = 1;
                        sys[g.sys[gs]].red_direct = true;
                        g.direct_dirty = true;
                    }
                if (g.direct_dirty) {
                    if ((e = put_small(g.d_direct, g.direct_h.data(), ngs * sizeof(int)))) return e;
                    g.direct_dirty = false;
                }
            }
            if (std::any_of(g.direct_h.begin(), g.direct_h.end(), [](int d) { return d != 0; }))
                if ((e = direct_pass(g, ymax))) return e;
        }
        if ((e = rec(10))) return e;
        // ---- recovery into the FP64 solution
        for (int gi = 0; gi < ngroups; ++gi) {
            Group &g = *groups[gi];
            if (!g.n_items_r) continue;
            const DevBuf &items = it == 0 ? g.items_rb : g.items_rr;
            T_SWITCH(g.tcode, NTA_SWITCH(g.NT, CT_SWITCH(g.ct_y, {
                typedef typename tr<T>::H H;
                (launch_recover<T, NT, CT>)(g.n_items_r, s, items.as<SweepItem>(), g.d_parts.as<PartDesc>(), g.U.as<T>(),
                                          g.Wf.as<T>(), g.Vf.as<T>(), (const H *)g.red0.p, g.d_part_iface.as<int>(),
                                          g.ymax, g.scratch.as<T>(), it > 0 ? 1 : 0, g.d_active.as<int>(),
                                          (const int *)g.wstop.as<int>());
            })));
            HIPCK(hipGetLastError());
        }
        if ((e = rec(11))) return e;
        // ---- FP64 residual of every column (residual launches timed apart from the norm sums)
        for (int gi = 0; gi < ngroups; ++gi) {
            Group &g = *groups[gi];
            if ((e = launch_residual(dtype, ymax, ymax, true, (unsigned)g.chunks.size(), s, g.d_chunks.as<RowChunk>(),
                                     g.d_sysdev.as<SysDev>(), g.partials.as<double>(), g.d_active.as<int>())))
                return e;
        }
        if ((e = rec(13))) return e;
        for (int gi = 0; gi < ngroups; ++gi) {
            Group &g = *groups[gi];
            const int ngs = (int)g.sys.size();
            hipLaunchKernelGGL(k_norm_reduce, dim3((unsigned)(ngs * ymax)), dim3(256), 0, s, g.d_chunk0.as<int>(),
                               g.d_sysdev.as<SysDev>(), g.partials.as<double>(), g.norms.as<double>(), ymax);
            hipLaunchKernelGGL(k_refine_check, dim3((ngs + 255) / 256), dim3(256), 0, s, g.d_sysstat.as<SysStat>(), ngs,
                               (const SysDev *)g.d_sysdev.p, (const double *)g.norms.p, ymax, o.tol, it, max_refine,
                               g.d_active.as<int>());
            HIPCK(hipGetLastError());
        }
        if ((e = rec(12))) return e;
        ++passes;
        if (!capture) {
            bool any = false;
            for (int gi = 0; gi < ngroups && !any; ++gi) {
                Group &g = *groups[gi];
                int summary[4];
                hipLaunchKernelGGL(k_stat_summary, dim3(1), dim3(256), 0, s, (const SysStat *)g.d_sysstat.p,
                                   (int)g.sys.size(), (const int *)g.d_active.p, g.d_summary.as<int>());
                HIPCK(hipGetLastError());
                if ((e = get_small({{summary, g.d_summary.p, sizeof summary}}))) return e;
                any = summary[2] > 0;
            }
            HIPCK(hipEventSynchronize(ev[12]));            // the pass's events are complete before they are read
            t_y += ev_ms(8, 9); t_red += ev_ms(9, 10); t_rec += ev_ms(10, 11); t_res += ev_ms(11, 13); t_norm += ev_ms(13, 12);
            if (!any) break;
        }
    }
    return SPIKE_SUCCESS;
}

// One pass of the direct reduced solve for the group's active systems on it.
int Engine::direct_pass(Group &g, int ymax) {
    std::vector<int> act(g.sys.size());
    int e;
    if ((e = get_small({{act.data(), g.d_active.p, act.size() * sizeof(int)}}))) return e;
    std::vector<int> which;
    for (size_t gs = 0; gs < g.sys.size(); ++gs)
        if (act[gs] && g.direct_h[gs]) which.push_back((int)gs);
    if (which.empty()) return SPIKE_SUCCESS;
    return reduced_direct(g, which, g.red0, ymax);
}

int Engine::launch_jacobi(Group &g, int ncc, int ymax, const DevBuf &xo, DevBuf &xn) {
    hipStream_t s = stream;
    const unsigned nb = (unsigned)(g.ifs.size() * ncc);
    if (g.NT == 16) {               // column sub-chunks keep three b x CK vectors within the LDS
        if (dtype == SPIKE_DTYPE_Z)
            hipLaunchKernelGGL((k_red_jacobi<zd, 16, 8>), dim3(nb), dim3(256), 0, s, g.d_ifs.as<IfaceDesc>(), ncc,
                               (const zd *)g.vt.p, (const zd *)g.vb.p, (const zd *)g.wt.p, (const zd *)g.wb.p,
                               (const zd *)g.yt.p, (const zd *)g.yb.p, (const zd *)g.T.p, (const zd *)xo.p, xn.as<zd>(),
                               g.conv.as<double>(), ymax, ymax, g.d_red_active.as<int>());
        else
            hipLaunchKernelGGL((k_red_jacobi<double, 16, 16>), dim3(nb), dim3(256), 0, s, g.d_ifs.as<IfaceDesc>(), ncc,
                               (const double *)g.vt.p, (const double *)g.vb.p, (const double *)g.wt.p,
                               (const double *)g.wb.p, (const double *)g.yt.p, (const double *)g.yb.p,
                               (const double *)g.T.p, (const double *)xo.p, xn.as<double>(), g.conv.as<double>(), ymax,
                               ymax, g.d_red_active.as<int>());
    } else if (dtype == SPIKE_DTYPE_Z) {
        NT_SWITCH(g.NT, {
            hipLaunchKernelGGL((k_red_jacobi<zd, NT, 16>), dim3(nb), dim3(256), 0, s, g.d_ifs.as<IfaceDesc>(), ncc, (const zd *)g.vt.p,
                               (const zd *)g.vb.p, (const zd *)g.wt.p, (const zd *)g.wb.p, (const zd *)g.yt.p, (const zd *)g.yb.p,
                               (const zd *)g.T.p, (const zd *)xo.p, xn.as<zd>(), g.conv.as<double>(), ymax, ymax,
                               g.d_red_active.as<int>());
        });
    } else {
        NT_SWITCH(g.NT, {
            if constexpr (NT <= 4)
                hipLaunchKernelGGL((k_red_jacobi_q<double, NT>), dim3(nb), dim3(NT * 64), 0, s, g.d_ifs.as<IfaceDesc>(), ncc,
                                   (const double *)g.vt.p, (const double *)g.vb.p, (const double *)g.wt.p,
                                   (const double *)g.wb.p, (const double *)g.yt.p, (const double *)g.yb.p,
                                   (const double *)g.T.p, (const double *)xo.p, xn.as<double>(), g.conv.as<double>(), ymax,
                                   ymax, g.d_red_active.as<int>());
            else
                hipLaunchKernelGGL((k_red_jacobi<double, NT, 16>), dim3(nb), dim3(256), 0, s, g.d_ifs.as<IfaceDesc>(), ncc,
                                   (const double *)g.vt.p, (const double *)g.vb.p, (const double *)g.wt.p,
                                   (const double *)g.wb.p, (const double *)g.yt.p, (const double *)g.yb.p,
                                   (const double *)g.T.p, (const double *)xo.p, xn.as<double>(), g.conv.as<double>(), ymax,
                                   ymax, g.d_red_active.as<int>());
        });
    }
    HIPCK(hipGetLastError());
    return SPIKE_SUCCESS;
}

int Engine::solve(int) {
    if (!factored) {
        int e = factor();
        if (e) return e;
    } else if (int e = prepare_stream()) {
        return e;
    }
    const double t0 = now_ms();
    int maxn = 0;
    for (auto &S : sys) { maxn = std::max(maxn, S.nrhs); S.maxrel = 0.0; }
    const int chunk = 64;
    graph_replayed = false;
    bool done = false;
    if (o.graph && maxn > 0 && maxn <= chunk) {
        const int e = solve_graph(maxn);
        if (e == SPIKE_SUCCESS) done = true;
        else if (e != SPIKE_ERR_UNSUPPORTED) return e;
    }
    if (!done) {
        eager_pairs = 0;
        for (int c0 = 0; c0 < maxn; c0 += chunk) {
            const int e = solve_chunk(c0, std::min(chunk, maxn - c0));
            if (e) return e;
        }
        // budget of the graph of the next single-chunk solve: this solve's passes and sweep pairs
        // (a replay that needs more is detected and solved eagerly)
        bool direct = false;
        for (auto &S : sys) direct = direct || S.red_direct;
        if (o.graph && maxn <= chunk && !direct) {
            graph_passes = st.solve_passes;
            graph_pairs = std::max(1, eager_pairs);
        }
    }
    st.ms_solve_total = now_ms() - t0;
    for (auto &S : sys)
        if (!(S.maxrel <= o.tol)) {
            set_err("relative residual %s%ld e-12 above tolerance", "", (long)(S.maxrel * 1e12));
            return SPIKE_ERR_NOT_CONVERGED;
        }
    return SPIKE_SUCCESS;
}

// per-stage times of `src` added to (or copied into) `dst`
void stage_times(spike_stats_t &dst, const spike_stats_t &src, bool add) {
    double spike_stats_t::*f[] = {&spike_stats_t::ms_convert, &spike_stats_t::ms_factor, &spike_stats_t::ms_rtop,
                                  &spike_stats_t::ms_spikes, &spike_stats_t::ms_red_setup, &spike_stats_t::ms_ysweep,
                                  &spike_stats_t::ms_reduced, &spike_stats_t::ms_recover, &spike_stats_t::ms_residual,
                                  &spike_stats_t::ms_factor_total, &spike_stats_t::ms_solve_total, &spike_stats_t::ms_norms};
    for (auto m : f) dst.*m = add ? dst.*m + src.*m : src.*m;
}

// Rows the V / W chains of the last factorization swept (the W chains' stop blocks read back):
// a W chain covers blocks [0, wstop), a V chain blocks [j0, nb), j0 the block holding row ni - ku.
void spike_work(const Engine &E, double *w) {
    w[0] = w[1] = w[2] = 0.0;
    if (!E.factored) return;
    for (int gi = 0; gi < E.ngroups; ++gi) {
        const Group &g = *E.groups[gi];
        const size_t np = g.parts.size();
        if (!np || !g.wstop.p) continue;
        std::vector<int> ws(np);
        if (hipStreamSynchronize(E.stream) != hipSuccess ||
            hipMemcpy(ws.data(), g.wstop.p, np * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
            return;
        for (size_t q = 0; q < np; ++q) {
            const PartDesc &d = g.parts[q];
            if (d.P == 1) continue;
            const bool hw = d.p > 0, hv = d.p < d.P - 1;
            const int64_t wr = hw ? std::min<int64_t>((int64_t)ws[q] * g.b, d.ni) : 0;
            const int64_t v0 = hv ? (int64_t)((d.ni - d.ku) / g.b) * g.b : d.ni;
            const int64_t rows = wr + (d.ni - std::max(v0, wr));
            const double band = d.kl + d.ku + 1.0, k2 = 2.0 * (d.kl + d.ku);
            w[0] += (double)rows;
            w[1] += (double)rows * band;
            w[2] += k2 * ((double)wr * d.kl + (double)(d.ni - v0) * d.ku);
        }
    }
}

void fill_stats(const Engine &E, spike_stats_t *s) {
    *s = E.st;
    if (!E.sys.empty()) {
        const SysHost &S = E.sys[0];
        s->n = S.n; s->nnz = S.nnz; s->ku = S.ku; s->kl = S.kl; s->b = 16 * S.NT; s->P = S.P; s->nrhs = S.nrhs;
        s->refine_iters = S.refine_iters; s->reduced_iters = S.red_iters; s->reduced_update = S.red_update;
        double mr = 0;
        for (auto &X : E.sys) mr = std::max(mr, X.maxrel);
        s->max_rel_residual = mr;
    }
    s->precision = E.o.precision;
    s->dtype = E.dtype;
    s->device_bytes = E.device_bytes();
    s->reduced_direct = 0;
    for (auto &X : E.sys) s->reduced_direct += X.red_direct ? 1 : 0;
    s->graph_replayed = E.graph_replayed ? 1 : 0;
}

}  // namespace

// ============================================================================ C ABI
struct spike_handle {
    Engine E;
    DevBuf xr_host_dummy;
};

extern "C" {

void spike_default_opts(spike_opts_t *o) {
    memset(o, 0, sizeof *o);
    o->precision = SPIKE_PREC_FP64;
    o->partitions = 0;
    o->max_refine = 8;
    o->tol = 1e-10;
    o->reduced_maxit = 64;
    o->reduced_tol = 1e-15;
    o->sweep_waves = 0;
    o->verbose = 0;
    o->reduced_solver = SPIKE_RED_AUTO;
    o->graph = 0;
    o->tune = SPIKE_TUNE_LATENCY;
}

const char *spike_error_string(int code) {
    switch (code) {
    case SPIKE_SUCCESS: return "success";
    case SPIKE_ERR_INVALID_ARG: return "invalid argument";
    case SPIKE_ERR_MALFORMED_CSR: return "malformed CSR";
    case SPIKE_ERR_OOM: return "out of memory";
    case SPIKE_ERR_SINGULAR: return "singular pivot";
    case SPIKE_ERR_NOT_CONVERGED: return "not converged";
    case SPIKE_ERR_UNSUPPORTED: return "unsupported";
    case SPIKE_ERR_DEVICE: return "device error";
    case SPIKE_ERR_IO: return "io error";
    case SPIKE_ERR_CANCELLED: return "cancelled";
    default: return "unknown";
    }
}

const char *spike_last_error_message(void) { return g_errmsg; }

int spike_create(spike_handle_t **h, const spike_opts_t *o) {
    if (!h) return SPIKE_ERR_INVALID_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        set_err("no HIP device%s", "");
        *h = nullptr;
        return SPIKE_ERR_DEVICE;
    }
    spike_handle *x = new (std::nothrow) spike_handle();
    if (!x) return SPIKE_ERR_OOM;
    if (o) x->E.o = *o;
    *h = x;
    return SPIKE_SUCCESS;
}

int spike_destroy(spike_handle_t *h) {
    if (!h) return SPIKE_ERR_INVALID_ARG;
    (void)hipStreamSynchronize(h->E.stream);
    delete h;
    return SPIKE_SUCCESS;
}

int spike_set_opts(spike_handle_t *h, const spike_opts_t *o) {
    if (!h || !o) return SPIKE_ERR_INVALID_ARG;
    const bool refactor = o->precision != h->E.o.precision || o->partitions != h->E.o.partitions || o->tune != h->E.o.tune;
    h->E.o = *o;
    if (refactor) h->E.factored = false;
    return SPIKE_SUCCESS;
}

int spike_set_stream(spike_handle_t *h, void *stream) {
    if (!h) return SPIKE_ERR_INVALID_ARG;
    h->E.stream = (hipStream_t)stream;
    return SPIKE_SUCCESS;
}

int spike_set_matrix_dev(spike_handle_t *h, int dtype, int n, int64_t nnz, const int *d_rowptr, const int *d_colind,
                         const void *d_aij) {
    if (!h || n <= 0 || nnz < 0 || nnz > INT32_MAX || !d_rowptr || (nnz > 0 && (!d_colind || !d_aij)) ||
        (dtype != SPIKE_DTYPE_D && dtype != SPIKE_DTYPE_Z))
        return SPIKE_ERR_INVALID_ARG;
    Engine &E = h->E;
    E.dtype = dtype;
    E.sys.assign(1, SysHost());
    SysHost &S = E.sys[0];
    S.n = n; S.nnz = nnz; S.d_rowptr = d_rowptr; S.d_colind = d_colind; S.d_aij = d_aij;
    E.factored = false;
    return SPIKE_SUCCESS;
}

int spike_set_matrix_host(spike_handle_t *h, int dtype, int n, int64_t nnz, const int *rowptr, const int *colind,
                          const void *aij) {
    if (!h || n <= 0 || nnz < 0 || nnz > INT32_MAX || !rowptr || (nnz > 0 && (!colind || !aij)) ||
        (dtype != SPIKE_DTYPE_D && dtype != SPIKE_DTYPE_Z))
        return SPIKE_ERR_INVALID_ARG;
    if (matrix_CheckCSR(n, (uLong_t)nnz, colind, rowptr) != SPIKE_SUCCESS) return SPIKE_ERR_MALFORMED_CSR;
    Engine &E = h->E;
    const size_t hs = h_bytes(dtype);
    int e;
    if ((e = E.own_rowptr.ensure((size_t)(n + 1) * 4)) || (e = E.own_colind.ensure((size_t)nnz * 4)) ||
        (e = E.own_aij.ensure((size_t)nnz * hs)))
        return e;
    {
        StagerLease sg;
        if ((e = sg->h2d(E.own_rowptr.p, rowptr, (size_t)(n + 1) * 4, nullptr))) return e;
        if (nnz && ((e = sg->h2d(E.own_colind.p, colind, (size_t)nnz * 4, nullptr)) ||
                    (e = sg->h2d(E.own_aij.p, aij, (size_t)nnz * hs, nullptr))))
            return e;
        if ((e = sg->drain())) return e;
        HIPCK(hipStreamSynchronize(nullptr));
    }
    return spike_set_matrix_dev(h, dtype, n, nnz, E.own_rowptr.as<int>(), E.own_colind.as<int>(), E.own_aij.p);
}

int spike_factor(spike_handle_t *h) {
    if (!h) return SPIKE_ERR_INVALID_ARG;
    return h->E.factor();
}

int spike_solve_dev(spike_handle_t *h, int nrhs, const void *d_b, int ldb, void *d_x, int ldx) {
    if (!h || h->E.sys.empty() || nrhs < 0 || (nrhs > 0 && (!d_b || !d_x))) return SPIKE_ERR_INVALID_ARG;
    SysHost &S = h->E.sys[0];
    if (ldb < S.n || ldx < S.n) return SPIKE_ERR_INVALID_ARG;
    if (nrhs == 0) return SPIKE_SUCCESS;
    S.nrhs = nrhs; S.d_b = d_b; S.ldb = ldb; S.d_x = d_x; S.ldx = ldx;
    const int e = h->E.solve(nrhs);
    if (h->E.o.verbose) {
        spike_stats_t st;
        fill_stats(h->E, &st);
        fprintf(stderr, "[spike] n=%d ku=%d kl=%d b=%d P=%d nrhs=%d prec=%d refine=%d red_iters=%d maxrel=%.3e "
                        "factor=%.3fms (conv %.3f lu %.3f spikes %.3f) solve=%.3fms\n",
                st.n, st.ku, st.kl, st.b, st.P, st.nrhs, st.precision, st.refine_iters, st.reduced_iters,
                st.max_rel_residual, st.ms_factor_total, st.ms_convert, st.ms_factor, st.ms_spikes, st.ms_solve_total);
    }
    return e;
}

int spike_solve_host(spike_handle_t *h, int nrhs, const void *b, int ldb, void *x, int ldx) {
    if (!h || h->E.sys.empty() || nrhs < 0 || (nrhs > 0 && (!b || !x))) return SPIKE_ERR_INVALID_ARG;
    Engine &E = h->E;
    const int n = E.sys[0].n;
    if (ldb < n || ldx < n) return SPIKE_ERR_INVALID_ARG;
    if (nrhs == 0) return SPIKE_SUCCESS;
    const size_t hs = h_bytes(E.dtype), bytes = (size_t)n * nrhs * hs;
    int e;
    if ((e = E.stage_b.ensure(bytes)) || (e = E.stage_x.ensure(bytes))) return e;
    StagerLease sg;
    if ((e = sg->h2d_cols(E.stage_b.p, b, (size_t)ldb * hs, (size_t)n * hs, nrhs, nullptr)) || (e = sg->drain())) return e;
    HIPCK(hipStreamSynchronize(nullptr));
    const int rc = spike_solve_dev(h, nrhs, E.stage_b.p, n, E.stage_x.p, n);
    if (rc == SPIKE_SUCCESS || rc == SPIKE_ERR_NOT_CONVERGED) {
        HIPCK(hipStreamSynchronize(E.stream));      // the solve may run on the handle's own stream
        if ((e = sg->d2h_cols(x, (size_t)ldx * hs, E.stage_x.p, (size_t)n * hs, nrhs, nullptr))) return e;
    }
    return rc;
}

int spike_get_spike_work(const spike_handle_t *h, double w[3]) {
    if (!h || !w) return SPIKE_ERR_INVALID_ARG;
    spike_work(h->E, w);
    return SPIKE_SUCCESS;
}

int spike_get_stats(const spike_handle_t *h, spike_stats_t *s) {
    if (!h || !s) return SPIKE_ERR_INVALID_ARG;
    fill_stats(h->E, s);
    return SPIKE_SUCCESS;
}

int spike_get_schedule(const spike_handle_t *h, int *P, int *rows, int cap) {
    if (!h || h->E.sys.empty() || !P) return SPIKE_ERR_INVALID_ARG;
    const SysHost &S = h->E.sys[0];
    *P = S.P;
    if (rows) {
        if (cap < S.P + 1) return SPIKE_ERR_INVALID_ARG;
        const int nreg = S.n / S.P;
        for (int p = 0; p < S.P; ++p) rows[p] = p * nreg;
        rows[S.P] = S.n;
    }
    return SPIKE_SUCCESS;
}

// Compact per-partition tips (reference layout) from the b-strided device slots.
int spike_get_tips(const spike_handle_t *h, double *Vt, double *Vb, double *Wt, double *Wb, double *yt, double *yb) {
    if (!h || h->E.sys.empty() || !h->E.factored || h->E.dtype != SPIKE_DTYPE_D) return SPIKE_ERR_INVALID_ARG;
    const Engine &E = h->E;
    const SysHost &S = E.sys[0];
    const Group &g = *E.groups[S.group];
    const int b = g.b, P = S.P, ku = S.ku, kl = S.kl, m = std::min(S.nrhs, g.ymax);
    const size_t bb = (size_t)b * b;
    const int p0 = g.sys_part0[S.gidx];
    auto fetch = [&](const DevBuf &d, size_t slot, int rows, int cols, double *out) -> int {
        if (!out || rows == 0 || cols == 0) return 0;
        std::vector<double> tmp(slot * P);
        HIPCK(hipMemcpy(tmp.data(), (const double *)d.p + (size_t)p0 * slot, tmp.size() * sizeof(double),
                        hipMemcpyDeviceToHost));
        for (int p = 0; p < P; ++p)
            for (int c = 0; c < cols; ++c)
                for (int r = 0; r < rows; ++r) out[(size_t)p * rows * cols + r + (size_t)c * rows] = tmp[p * slot + r + (size_t)c * b];
        return 0;
    };
    int e;
    if ((e = fetch(g.vt, bb, ku, ku, Vt)) || (e = fetch(g.vb, bb, kl, ku, Vb)) || (e = fetch(g.wt, bb, ku, kl, Wt)) ||
        (e = fetch(g.wb, bb, kl, kl, Wb)))
        return e;
    if (g.ymax > 0) {
        if ((e = fetch(g.yt, (size_t)b * g.ymax, ku, m, yt)) || (e = fetch(g.yb, (size_t)b * g.ymax, kl, m, yb))) return e;
    }
    // the reference leaves V of the last partition and W of the first unset: report zeros
    if (Vt && P > 0) memset(Vt + (size_t)(P - 1) * ku * ku, 0, sizeof(double) * ku * ku);
    if (Vb && P > 0) memset(Vb + (size_t)(P - 1) * kl * ku, 0, sizeof(double) * kl * ku);
    if (Wt) memset(Wt, 0, sizeof(double) * ku * kl);
    if (Wb) memset(Wb, 0, sizeof(double) * kl * kl);
    return SPIKE_SUCCESS;
}

// Reference-order reduced solution [x^t_p ; x^b_p] of every partition for the last inner solve
// (the solve of A x = b itself when no refinement step followed), from the last interface
// iterate and the tips; the end unknowns x^t_0 and x^b_{P-1} are formed on the host.
int spike_get_reduced_solution(const spike_handle_t *h, double *xr) {
    if (!h || !xr || h->E.sys.empty() || !h->E.factored || h->E.dtype != SPIKE_DTYPE_D) return SPIKE_ERR_INVALID_ARG;
    const Engine &E = h->E;
    const SysHost &S = E.sys[0];
    const Group &g = *E.groups[S.group];
    const int P = S.P, ku = S.ku, kl = S.kl, k2 = ku + kl, m = std::min(S.nrhs, g.ymax), b = g.b;
    if (P < 2 || m < 1) return SPIKE_ERR_INVALID_ARG;
    std::vector<double> Vt((size_t)P * ku * ku), Vb((size_t)P * kl * ku), Wt((size_t)P * ku * kl), Wb((size_t)P * kl * kl),
        yt((size_t)P * ku * m), yb((size_t)P * kl * m);
    int e = spike_get_tips(h, Vt.data(), Vb.data(), Wt.data(), Wb.data(), yt.data(), yb.data());
    if (e) return e;
    const size_t slot = (size_t)2 * b * g.ymax;
    std::vector<double> X((size_t)(P - 1) * slot);
    const int f0 = g.part_iface[g.sys_part0[S.gidx]];
    const DevBuf &last = g.red_last ? g.red1 : g.red0;
    HIPCK(hipMemcpy(X.data(), (const double *)last.p + (size_t)f0 * slot, X.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int c = 0; c < m; ++c) {
        double *col = xr + (size_t)c * P * k2;
        for (int p = 0; p < P; ++p) {
            for (int r = 0; r < ku; ++r) {
                double v;
                if (p > 0) v = X[(size_t)(p - 1) * slot + b + r + (size_t)c * 2 * b];
                else {
                    v = yt[r + (size_t)c * ku];
                    for (int k = 0; k < ku; ++k) v -= Vt[r + (size_t)k * ku] * X[b + k + (size_t)c * 2 * b];
                }
                col[p * k2 + r] = v;
            }
            for (int r = 0; r < kl; ++r) {
                double v;
                if (p < P - 1) v = X[(size_t)p * slot + r + (size_t)c * 2 * b];
                else {
                    v = yb[(size_t)p * kl * m + r + (size_t)c * kl];
                    for (int k = 0; k < kl; ++k)
                        v -= Wb[(size_t)p * kl * kl + r + (size_t)k * kl] * X[(size_t)(p - 1) * slot + k + (size_t)c * 2 * b];
                }
                col[p * k2 + ku + r] = v;
            }
        }
    }
    return SPIKE_SUCCESS;
}

}  // extern "C"

// ============================================================================ batched systems
struct spike_batch {
    Engine E;
    // A batch whose arrays do not fit in HBM is solved in windows of at most `cap` systems, one
    // after another (each system's result is the one it gets alone, so windows change no bits);
    // `cap` halves on every failed allocation and is kept for later calls.
    int cap = INT32_MAX;
    int windows = 0;
    std::vector<double> rel;        // final residual of every system of the last call
    spike_stats_t acc{};            // stage times summed over the windows of the last call
};

extern "C" {

int spike_batch_create(spike_batch_t **bt, const spike_opts_t *o) {
    if (!bt) return SPIKE_ERR_INVALID_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { (void)hipGetLastError(); *bt = nullptr; return SPIKE_ERR_DEVICE; }
    spike_batch *x = new (std::nothrow) spike_batch();
    if (!x) return SPIKE_ERR_OOM;
    if (o) x->E.o = *o;
    if (const char *v = getenv("HSPIKE_BATCH_WINDOW"))      // systems per window from the start (tests, tuning)
        if (atoi(v) > 0) x->cap = atoi(v);
    *bt = x;
    return SPIKE_SUCCESS;
}

int spike_batch_destroy(spike_batch_t *bt) {
    if (!bt) return SPIKE_ERR_INVALID_ARG;
    (void)hipStreamSynchronize(bt->E.stream);
    delete bt;
    return SPIKE_SUCCESS;
}

int spike_batch_solve_dev(spike_batch_t *bt, int dtype, int nsys, const spike_system_t *sys) {
    if (!bt || nsys <= 0 || !sys || (dtype != SPIKE_DTYPE_D && dtype != SPIKE_DTYPE_Z)) return SPIKE_ERR_INVALID_ARG;
    for (int i = 0; i < nsys; ++i) {
        const spike_system_t &q = sys[i];
        if (q.n <= 0 || q.nnz < 0 || q.nnz > INT32_MAX || !q.d_rowptr || q.nrhs < 0 || (q.nrhs > 0 && (!q.d_b || !q.d_x)))
            return SPIKE_ERR_INVALID_ARG;
    }
    Engine &E = bt->E;
    bt->rel.assign(nsys, 0.0);
    bt->acc = spike_stats_t{};
    bt->windows = 0;
    const double t0 = now_ms();
    int e = SPIKE_SUCCESS;
    const int graph = E.o.graph;      // windows rebind every system: replay only a batch solved whole
    for (int lo = 0; lo < nsys;) {
        const int hi = (int)std::min<int64_t>(nsys, (int64_t)lo + bt->cap);
        E.dtype = dtype;
        E.sys.assign(hi - lo, SysHost());
        for (int i = lo; i < hi; ++i) {
            const spike_system_t &q = sys[i];
            SysHost &S = E.sys[i - lo];
            S.n = q.n; S.nnz = q.nnz; S.d_rowptr = q.d_rowptr; S.d_colind = q.d_colind; S.d_aij = q.d_aij;
            S.nrhs = q.nrhs; S.d_b = q.d_b; S.ldb = q.n; S.d_x = q.d_x; S.ldx = q.n;
        }
        E.factored = false;
        E.o.graph = (lo == 0 && hi == nsys) ? graph : 0;
        e = E.factor();
        if (!e) e = E.solve(0);
        E.o.graph = graph;
        if (e == SPIKE_ERR_OOM && hi - lo > 1) {      // free this window's arrays, retry it halved
            (void)hipStreamSynchronize(E.stream);
            E.groups.clear();
            E.ngroups = 0;
            bt->cap = (hi - lo) / 2;
            continue;
        }
        for (int i = lo; i < hi; ++i) bt->rel[i] = E.sys[i - lo].maxrel;
        stage_times(bt->acc, E.st, true);
        ++bt->windows;
        if (e) break;
        lo = hi;
    }
    if (E.o.verbose) {
        double mr = 0;
        for (double r : bt->rel) mr = std::max(mr, r);
        fprintf(stderr, "[spike batch] %d systems in %d window(s), factor %.3f ms, solve %.3f ms, total %.3f ms, max rel %.3e\n",
                nsys, bt->windows, bt->acc.ms_factor_total, bt->acc.ms_solve_total, now_ms() - t0, mr);
    }
    return e;
}

int spike_batch_get_spike_work(const spike_batch_t *bt, double w[3]) {
    if (!bt || !w) return SPIKE_ERR_INVALID_ARG;
    spike_work(bt->E, w);
    return SPIKE_SUCCESS;
}

int spike_batch_get_stats(const spike_batch_t *bt, spike_stats_t *s, double *rel) {
    if (!bt) return SPIKE_ERR_INVALID_ARG;
    if (s) {
        fill_stats(bt->E, s);
        stage_times(*s, bt->acc, false);
        double mr = 0;
        for (double r : bt->rel) mr = std::max(mr, r);
        s->max_rel_residual = mr;
    }
    if (rel)
        for (size_t i = 0; i < bt->rel.size(); ++i) rel[i] = bt->rel[i];
    return SPIKE_SUCCESS;
}

// ============================================================================ residual helper
}  // extern "C"

namespace {

// Per-column ||b - A x|| / ||b|| of device arrays (column-major, leading dimension n) with the
// residual r = b - A x left in d_r; small staging buffers come from `ws` (kept by the caller).
struct ResidualWs {
    DevBuf dch, dsys, dpart, dnorm, dc0;
};

int residual_dev(int dtype, int n, const int *d_rowptr, const int *d_colind, const void *d_aij, int nrhs, const void *d_x,
                 const void *d_b, void *d_r, double *rel, ResidualWs &ws, hipStream_t st) {
    const size_t hs = h_bytes(dtype);
    std::vector<RowChunk> ch;
    for (int r = 0; r < n; r += RES_ROWS) ch.push_back({0, r});
    const int nch = (int)ch.size();
    int e;
    if ((e = upload(ws.dch, ch.data(), ch.size() * sizeof(RowChunk), st)) || (e = ws.dpart.ensure((size_t)nch * 64 * 2 * 8)) ||
        (e = ws.dnorm.ensure(64 * 2 * 8)))
        return e;
    for (int c0 = 0; c0 < nrhs; c0 += 64) {
        const int m = std::min(64, nrhs - c0);
        SysDev sd;
        memset(&sd, 0, sizeof sd);
        sd.rowptr = d_rowptr; sd.colind = d_colind; sd.aij = d_aij; sd.n = n; sd.m = m;
        sd.ldb = n; sd.ldx = n; sd.ldr = n;
        sd.bptr = (const char *)d_b + (size_t)c0 * n * hs;
        sd.xptr = (char *)d_x + (size_t)c0 * n * hs;
        sd.rptr = (char *)d_r + (size_t)c0 * n * hs;
        int c0s[2] = {0, nch};
        if ((e = upload(ws.dsys, &sd, sizeof sd, st)) || (e = upload(ws.dc0, c0s, sizeof c0s, st))) return e;
        if ((e = launch_residual(dtype, m, 64, false, (unsigned)nch, st, ws.dch.as<RowChunk>(), ws.dsys.as<SysDev>(),
                                 ws.dpart.as<double>(), (const int *)nullptr)))
            return e;
        hipLaunchKernelGGL(k_norm_reduce, dim3(64), dim3(256), 0, st, ws.dc0.as<int>(), ws.dsys.as<SysDev>(),
                           ws.dpart.as<double>(), ws.dnorm.as<double>(), 64);
        HIPCK(hipGetLastError());
        double nr[128];
        HIPCK(hipMemcpyAsync(nr, ws.dnorm.p, sizeof nr, hipMemcpyDeviceToHost, st));
        HIPCK(hipStreamSynchronize(st));
        for (int c = 0; c < m; ++c) rel[c0 + c] = nr[2 * c + 1] > 0 ? sqrt(nr[2 * c] / nr[2 * c + 1]) : sqrt(nr[2 * c]);
    }
    return SPIKE_SUCCESS;
}

// ---------------------------------------------------------------- band solve on a workspace
// The whole device workspace of one band solve is carved from one region in a fixed order, so
// the allocating entry point (one device buffer per call) and the allocation-free one (a static
// device array) run the same kernels on the same layout and produce the same bits.
struct Carve {
    char *base;
    size_t used = 0, cap;
    Carve(void *p, size_t c) : base((char *)p), cap(c) {}
    template <typename X> X *take(size_t count) {
        const size_t at = (used + 255) & ~(size_t)255;
        used = at + count * sizeof(X);
        return (base && used <= cap) ? (X *)(base + at) : nullptr;
    }
};

struct BandWork {
    int *flags;            // [0..2] ku, kl, malformed (k_bandwidth); [3] LU info; [4] CSR ends
    RowChunk *chunks;      // ceil(n / RES_ROWS) row chunks (>= the ceil(n / BW_ROWS) of k_bandwidth)
    SysDev *sysdev;
    int *c0s;              // {0, chunks} for k_norm_reduce
    double *partials, *norms;
    BandSys *sys;
    int *items;
    int *piv;
    void *ab, *r;
};

// Carves the workspace of an n x n band solve with nrhs columns; null members when it does not fit.
BandWork band_carve(Carve &c, int dtype, int n, int ku, int kl, int nrhs) {
    const size_t hs = h_bytes(dtype);
    const int nch = (n + RES_ROWS - 1) / RES_ROWS, ldab = 2 * kl + ku + 1;
    BandWork w;
    w.flags = c.take<int>(8);
    w.chunks = c.take<RowChunk>(nch);
    w.sysdev = c.take<SysDev>(1);
    w.c0s = c.take<int>(2);
    w.partials = c.take<double>((size_t)nch * 64 * 2);
    w.norms = c.take<double>(128);
    w.sys = c.take<BandSys>(1);
    w.items = c.take<int>(3 * (size_t)(nrhs + 1));
    w.piv = c.take<int>(n);
    w.ab = c.take<char>((size_t)n * ldab * hs);
    w.r = c.take<char>((size_t)n * std::max(nrhs, 1) * hs);
    return w;
}

size_t band_ws_bytes(int dtype, int n, int ku, int kl, int nrhs) {
    Carve c(nullptr, SIZE_MAX);
    band_carve(c, dtype, n, ku, kl, nrhs);
    return c.used;
}

// Column-major panels of H per column chunk of 64: r = b - A x and each column's relative
// residual into rel (caller-sized, nrhs entries); w.chunks already hold the RES_ROWS chunks.
int band_residual(int dtype, int n, const int *rowptr, const int *colind, const void *aij, int nrhs, const void *b, int ldb,
                  const void *x, int ldx, void *r, const BandWork &w, double *rel, hipStream_t st) {
    const size_t hs = h_bytes(dtype);
    const int nch = (n + RES_ROWS - 1) / RES_ROWS;
    for (int c0 = 0; c0 < nrhs; c0 += 64) {
        const int m = std::min(64, nrhs - c0);
        SysDev sd;
        memset(&sd, 0, sizeof sd);
        sd.rowptr = rowptr; sd.colind = colind; sd.aij = aij; sd.n = n; sd.m = m;
        sd.ldb = ldb; sd.ldx = ldx; sd.ldr = n;
        sd.bptr = (const char *)b + (size_t)c0 * ldb * hs;
        sd.xptr = (char *)const_cast<void *>(x) + (size_t)c0 * ldx * hs;
        sd.rptr = (char *)r + (size_t)c0 * n * hs;
        hipLaunchKernelGGL(k_put<SysDev>, dim3(1), dim3(64), 0, st, w.sysdev, sd);
        int e;
        if ((e = launch_residual(dtype, m, 64, false, (unsigned)nch, st, w.chunks, w.sysdev, w.partials, (const int *)nullptr)))
            return e;
        hipLaunchKernelGGL(k_norm_reduce, dim3(64), dim3(256), 0, st, (const int *)w.c0s, (const SysDev *)w.sysdev,
                           (const double *)w.partials, w.norms, 64);
        HIPCK(hipGetLastError());
        double nr[128];
        HIPCK(hipMemcpyAsync(nr, w.norms, sizeof nr, hipMemcpyDeviceToHost, st));
        HIPCK(hipStreamSynchronize(st));
        for (int c = 0; c < m; ++c) rel[c0 + c] = nr[2 * c + 1] > 0 ? sqrt(nr[2 * c] / nr[2 * c + 1]) : sqrt(nr[2 * c]);
    }
    return SPIKE_SUCCESS;
}

// gbtrs on a panel p (n x nrhs, leading dimension ldp) in place
int band_trs(int dtype, int n, int ku, int kl, int nrhs, void *p, int ldp, const BandWork &w, hipStream_t st) {
    const size_t hs = h_bytes(dtype), per_col = (size_t)n * hs, cap = 48 * 1024;
    const bool lx = per_col <= cap;
    const int mc = lx ? (int)std::min<size_t>({64, (size_t)nrhs, cap / per_col}) : std::min(64, nrhs);
    BandSys B{};
    B.n = n; B.kl = kl; B.ku = ku; B.ldab = 2 * kl + ku + 1; B.m = nrhs; B.ldx = ldp; B.ab = 0; B.piv = 0; B.x = 0;
    const int ni = (nrhs + mc - 1) / mc;
    hipLaunchKernelGGL(k_put<BandSys>, dim3(1), dim3(64), 0, st, w.sys, B);
    hipLaunchKernelGGL(k_band_items, dim3((ni + 255) / 256), dim3(256), 0, st, w.items, nrhs, mc);
    const size_t lds = lx ? per_col * mc : 0;
    if (dtype == SPIKE_DTYPE_Z) {
        if (lx) hipLaunchKernelGGL((k_gbtrs<zd, true>), dim3(ni), dim3(256), lds, st, (const BandSys *)w.sys, (const int *)w.items,
                                   (const zd *)w.ab, (const int *)w.piv, (zd *)p);
        else hipLaunchKernelGGL((k_gbtrs<zd, false>), dim3(ni), dim3(256), 0, st, (const BandSys *)w.sys, (const int *)w.items,
                                (const zd *)w.ab, (const int *)w.piv, (zd *)p);
    } else {
        if (lx) hipLaunchKernelGGL((k_gbtrs<double, true>), dim3(ni), dim3(256), lds, st, (const BandSys *)w.sys,
                                   (const int *)w.items, (const double *)w.ab, (const int *)w.piv, (double *)p);
        else hipLaunchKernelGGL((k_gbtrs<double, false>), dim3(ni), dim3(256), 0, st, (const BandSys *)w.sys,
                                (const int *)w.items, (const double *)w.ab, (const int *)w.piv, (double *)p);
    }
    HIPCK(hipGetLastError());
    return SPIKE_SUCCESS;
}

// Band LU with partial pivoting of the device CSR (bandwidths ku, kl), x = A^{-1} b, then FP64
// refinement x += A^{-1}(b - A x) while the largest relative residual halves (at most three
// corrections, the reference's PARDISO setting for R, spike_algebra.c:299).  rel receives every
// column's final relative residual (nrhs entries, caller memory).  All device memory is `w`.
int band_run(int dtype, int n, int ku, int kl, const int *rowptr, const int *colind, const void *aij, int nrhs,
             const void *b, int ldb, void *x, int ldx, double *rel, const BandWork &w, hipStream_t st) {
    const size_t hs = h_bytes(dtype);
    const int ldab = 2 * kl + ku + 1, nch = (n + RES_ROWS - 1) / RES_ROWS;
    const size_t lds_f = (size_t)(std::max(kl, 1) + kl + ku + 1) * hs;
    if (lds_f > 64 * 1024) {
        set_err("band too wide for the device band LU (kl %ld, ku %ld)", (long)kl, (long)ku);
        return SPIKE_ERR_UNSUPPORTED;
    }
    BandSys B{};
    B.n = n; B.kl = kl; B.ku = ku; B.ldab = ldab;
    const size_t elems = (size_t)n * ldab;
    const unsigned rg = (unsigned)((n + 255) / 256), fg = (unsigned)std::min<size_t>((elems + 255) / 256, 65536);
    hipLaunchKernelGGL(k_put<BandSys>, dim3(1), dim3(64), 0, st, w.sys, B);
    hipLaunchKernelGGL(k_make_chunks, dim3((nch + 255) / 256), dim3(256), 0, st, w.chunks, nch, RES_ROWS);
    hipLaunchKernelGGL(k_put<int>, dim3(1), dim3(64), 0, st, w.c0s, 0);
    hipLaunchKernelGGL(k_put<int>, dim3(1), dim3(64), 0, st, w.c0s + 1, nch);
    if (dtype == SPIKE_DTYPE_Z) {
        hipLaunchKernelGGL(k_fill<zd>, dim3(fg), dim3(256), 0, st, (zd *)w.ab, (int64_t)elems, s_zero_host<zd>());
        hipLaunchKernelGGL(k_csr_to_band<zd>, dim3(rg), dim3(256), 0, st, rowptr, colind, (const zd *)aij, n, kl, ku, ldab,
                           (zd *)w.ab);
        hipLaunchKernelGGL(k_gbtrf<zd>, dim3(1), dim3(GB_THREADS), lds_f, st, (const BandSys *)w.sys, (zd *)w.ab, w.piv,
                           w.flags + 3);
    } else {
        hipLaunchKernelGGL(k_fill<double>, dim3(fg), dim3(256), 0, st, (double *)w.ab, (int64_t)elems, 0.0);
        hipLaunchKernelGGL(k_csr_to_band<double>, dim3(rg), dim3(256), 0, st, rowptr, colind, (const double *)aij, n, kl, ku,
                           ldab, (double *)w.ab);
        hipLaunchKernelGGL(k_gbtrf<double>, dim3(1), dim3(GB_THREADS), lds_f, st, (const BandSys *)w.sys, (double *)w.ab,
                           w.piv, w.flags + 3);
    }
    HIPCK(hipGetLastError());
    int info = 0;
    HIPCK(hipMemcpyAsync(&info, w.flags + 3, sizeof info, hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    if (info) {
        set_err("singular matrix (no nonzero pivot in column %ld)", (long)(info - 1));
        return SPIKE_ERR_SINGULAR;
    }
    if (nrhs == 0) return SPIKE_SUCCESS;
    const unsigned pg = (unsigned)std::min<int64_t>(((int64_t)n * nrhs + 255) / 256, 65536);
    if (dtype == SPIKE_DTYPE_Z) hipLaunchKernelGGL(k_copy2d<zd>, dim3(pg), dim3(256), 0, st, (zd *)x, ldx, (const zd *)b, ldb, n, nrhs);
    else hipLaunchKernelGGL(k_copy2d<double>, dim3(pg), dim3(256), 0, st, (double *)x, ldx, (const double *)b, ldb, n, nrhs);
    HIPCK(hipGetLastError());
    int e;
    if ((e = band_trs(dtype, n, ku, kl, nrhs, x, ldx, w, st))) return e;
    if ((e = band_residual(dtype, n, rowptr, colind, aij, nrhs, b, ldb, x, ldx, w.r, w, rel, st))) return e;
    double worst = 0.0;
    for (int c = 0; c < nrhs; ++c) worst = std::max(worst, rel[c]);
    for (int it = 0; it < 3 && worst > 0.0 && std::isfinite(worst); ++it) {
        if ((e = band_trs(dtype, n, ku, kl, nrhs, w.r, n, w, st))) return e;          // delta = A^{-1} r, in place
        if (dtype == SPIKE_DTYPE_Z) hipLaunchKernelGGL(k_add2d<zd>, dim3(pg), dim3(256), 0, st, (zd *)x, ldx, (const zd *)w.r, n, n, nrhs);
        else hipLaunchKernelGGL(k_add2d<double>, dim3(pg), dim3(256), 0, st, (double *)x, ldx, (const double *)w.r, n, n, nrhs);
        HIPCK(hipGetLastError());
        if ((e = band_residual(dtype, n, rowptr, colind, aij, nrhs, b, ldb, x, ldx, w.r, w, rel, st))) return e;
        double wn = 0.0;
        for (int c = 0; c < nrhs; ++c) wn = std::max(wn, rel[c]);
        if (!(wn < 0.5 * worst)) break;               // stagnating: the correction is at rounding level
        worst = wn;
    }
    return SPIKE_SUCCESS;
}

}  // namespace

extern "C" {

int spike_residual_host(int dtype, int n, int64_t nnz, const int *rowptr, const int *colind, const void *aij, int nrhs,
                        const void *x, const void *b, double *rel) {
    if (n <= 0 || nnz < 0 || nrhs < 0 || !rowptr || !rel || (dtype != SPIKE_DTYPE_D && dtype != SPIKE_DTYPE_Z))
        return SPIKE_ERR_INVALID_ARG;
    if (nrhs == 0) return SPIKE_SUCCESS;
    const size_t hs = h_bytes(dtype);
    DevBuf drp, dci, da, dx, db, dr;
    ResidualWs ws;
    int e;
    if ((e = drp.ensure((size_t)(n + 1) * 4)) || (e = dci.ensure((size_t)nnz * 4)) || (e = da.ensure((size_t)nnz * hs)) ||
        (e = dx.ensure((size_t)n * nrhs * hs)) || (e = db.ensure((size_t)n * nrhs * hs)) || (e = dr.ensure((size_t)n * nrhs * hs)))
        return e;
    HIPCK(hipMemcpy(drp.p, rowptr, (size_t)(n + 1) * 4, hipMemcpyHostToDevice));
    if (nnz) {
        HIPCK(hipMemcpy(dci.p, colind, (size_t)nnz * 4, hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(da.p, aij, (size_t)nnz * hs, hipMemcpyHostToDevice));
    }
    HIPCK(hipMemcpy(dx.p, x, (size_t)n * nrhs * hs, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(db.p, b, (size_t)n * nrhs * hs, hipMemcpyHostToDevice));
    return residual_dev(dtype, n, drp.as<int>(), dci.as<int>(), da.p, nrhs, dx.p, db.p, dr.p, rel, ws, 0);
}

// Direct solve of a banded CSR matrix (host arrays) by the device band LU with partial pivoting
// and FP64 refinement (band_run); the workspace is one device buffer of this call.
int spike_band_solve_host(int dtype, int n, int64_t nnz, const int *rowptr, const int *colind, const void *aij, int nrhs,
                          const void *b, int ldb, void *x, int ldx, double *rel) {
    if (n <= 0 || nnz < 0 || nnz > INT32_MAX || nrhs < 0 || !rowptr || (nnz > 0 && (!colind || !aij)) ||
        (nrhs > 0 && (!b || !x)) || ldb < n || ldx < n || (dtype != SPIKE_DTYPE_D && dtype != SPIKE_DTYPE_Z))
        return SPIKE_ERR_INVALID_ARG;
    if (matrix_CheckCSR(n, (uLong_t)nnz, colind, rowptr) != SPIKE_SUCCESS) return SPIKE_ERR_MALFORMED_CSR;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { (void)hipGetLastError(); set_err("no HIP device%s", ""); return SPIKE_ERR_DEVICE; }
    integer_t ku = 0, kl = 0;
    matrix_ComputeBandwidth(n, colind, rowptr, nullptr, &ku, &kl);
    const size_t hs = h_bytes(dtype), panel = (size_t)n * nrhs * hs;
    DevBuf drp, dci, da, dx, db, ws;
    int e;
    if ((e = drp.ensure((size_t)(n + 1) * 4)) || (e = dci.ensure((size_t)std::max<int64_t>(nnz, 1) * 4)) ||
        (e = da.ensure((size_t)std::max<int64_t>(nnz, 1) * hs)) || (e = dx.ensure(panel)) || (e = db.ensure(panel)) ||
        (e = ws.ensure(band_ws_bytes(dtype, n, ku, kl, nrhs))))
        return e;
    StagerLease sg;
    if ((e = sg->h2d(drp.p, rowptr, (size_t)(n + 1) * 4, nullptr))) return e;
    if (nnz && ((e = sg->h2d(dci.p, colind, (size_t)nnz * 4, nullptr)) || (e = sg->h2d(da.p, aij, (size_t)nnz * hs, nullptr))))
        return e;
    if (nrhs && (e = sg->h2d_cols(db.p, b, (size_t)ldb * hs, (size_t)n * hs, nrhs, nullptr))) return e;
    if ((e = sg->drain())) return e;
    HIPCK(hipStreamSynchronize(nullptr));
    Carve c(ws.p, ws.bytes);
    const BandWork w = band_carve(c, dtype, n, ku, kl, nrhs);
    std::vector<double> rl(std::max(nrhs, 1));
    if ((e = band_run(dtype, n, ku, kl, drp.as<int>(), dci.as<int>(), da.p, nrhs, db.p, n, dx.p, n, rl.data(), w, 0))) return e;
    HIPCK(hipStreamSynchronize(nullptr));
    if (nrhs && (e = sg->d2h_cols(x, (size_t)ldx * hs, dx.p, (size_t)n * hs, nrhs, nullptr))) return e;
    if (rel) for (int q = 0; q < nrhs; ++q) rel[q] = rl[q];
    return SPIKE_SUCCESS;
}

}  // extern "C"

// ============================================================================ allocation-free band solve
// Workspace of spike_band_solve_dev_noalloc: a static device array, so the entry point calls no
// allocator.  Each GPU the library runs on holds its own copy until the process ends.  In host
// code the array's name is not a device address: the address on the current device comes from
// hipGetSymbolAddress, resolved once per device and kept in scratch_dev[].
static __device__ __align__(256) unsigned char scratch[SPIKE_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex scratch_mu;
void *scratch_dev[SCRATCH_MAX_DEVICES];

int scratch_on_current_device(void **out) {
    int dev = 0;
    HIPCK(hipGetDevice(&dev));
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) {
        set_err("device ordinal %ld beyond the scratch table", (long)dev);
        return SPIKE_ERR_DEVICE;
    }
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            set_err("no device address for the scratch area (%s)", hipGetErrorString(e));
            return SPIKE_ERR_DEVICE;
        }
        scratch_dev[dev] = p;
    }
    *out = scratch_dev[dev];
    return SPIKE_SUCCESS;
}

}  // namespace

extern "C" int spike_band_solve_dev_noalloc(int dtype, int n, int64_t nnz, const int *d_rowptr, const int *d_colind,
                                            const void *d_aij, int nrhs, const void *d_b, int ldb, void *d_x, int ldx,
                                            double *rel, void *stream) {
    if (n <= 0 || nnz < 0 || nnz > INT32_MAX || nrhs < 0 || !d_rowptr || (nnz > 0 && (!d_colind || !d_aij)) ||
        (nrhs > 0 && (!d_b || !d_x || !rel)) || ldb < n || ldx < n || (dtype != SPIKE_DTYPE_D && dtype != SPIKE_DTYPE_Z))
        return SPIKE_ERR_INVALID_ARG;
    void *base = nullptr;
    int e = scratch_on_current_device(&base);
    if (e) return e;
    hipStream_t st = (hipStream_t)stream;
    // CSR checks and bandwidths on the device, from the head of the area: flags, row chunks, view
    const int nbw = (n + BW_ROWS - 1) / BW_ROWS;
    Carve head(base, SPIKE_NOALLOC_SCRATCH_BYTES);
    int *flags = head.take<int>(8);
    RowChunk *ch = head.take<RowChunk>(nbw);
    SysDev *sd = head.take<SysDev>(1);
    if (!sd) {
        set_err("n = %ld needs more than the %ld-byte scratch area", (long)n, (long)SPIKE_NOALLOC_SCRATCH_BYTES);
        return SPIKE_ERR_OOM;
    }
    SysDev sv;
    memset(&sv, 0, sizeof sv);
    sv.rowptr = d_rowptr; sv.colind = d_colind; sv.n = n;
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(256), 0, st, flags, (int64_t)8, 0);   // the area keeps the last call's bytes
    hipLaunchKernelGGL(k_make_chunks, dim3((nbw + 255) / 256), dim3(256), 0, st, ch, nbw, BW_ROWS);
    hipLaunchKernelGGL(k_put<SysDev>, dim3(1), dim3(64), 0, st, sd, sv);
    hipLaunchKernelGGL(k_bandwidth, dim3(nbw), dim3(256), 0, st, (const RowChunk *)ch, (const SysDev *)sd, flags);
    hipLaunchKernelGGL(k_csr_ends, dim3(1), dim3(64), 0, st, d_rowptr, n, nnz, flags + 4);
    HIPCK(hipGetLastError());
    int fl[8];
    HIPCK(hipMemcpyAsync(fl, flags, sizeof fl, hipMemcpyDeviceToHost, st));
    HIPCK(hipStreamSynchronize(st));
    if (fl[2] || fl[4]) {
        set_err("malformed CSR%s", "");
        return SPIKE_ERR_MALFORMED_CSR;
    }
    const int ku = fl[0], kl = fl[1];
    const size_t need = band_ws_bytes(dtype, n, ku, kl, nrhs);
    if (need > SPIKE_NOALLOC_SCRATCH_BYTES) {
        set_err("this solve needs %ld bytes of workspace; the allocation-free entry point has %ld", (long)need,
                (long)SPIKE_NOALLOC_SCRATCH_BYTES);
        return SPIKE_ERR_OOM;
    }
    Carve c(base, SPIKE_NOALLOC_SCRATCH_BYTES);
    const BandWork w = band_carve(c, dtype, n, ku, kl, nrhs);
    return band_run(dtype, n, ku, kl, d_rowptr, d_colind, d_aij, nrhs, d_b, ldb, d_x, ldx, rel, w, st);
}

#include "spike_server.inc"
