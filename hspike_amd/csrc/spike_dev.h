// spike_dev.h -- device-side building blocks of the gfx950 SPIKE kernels.
//
// Scalar types: float, double, zf (complex float), zd (complex double).  Every banded block
// operation is a chain of 16x16 MFMA tiles:
//   v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32, lane l of a wave64 supplies
//     A[l % 16][l / 16]   (16 x 4 slice of the left operand)
//     B[l / 16][l % 16]   (4 x 16 slice of the right operand)
//   and holds D[trow(l / 16, e)][l % 16], e = 0..3, of the 16 x 16 result (trow below).
// A K = 16*NT product is issued as NT*4 MFMAs.  At MFMA step (t, e) the hardware k-slot
// g = l / 16 is mapped to the logical index k = 16 t + trow(g, e): the right operand a lane must
// supply is then exactly element e of its own accumulator tile t, so the result of one GEMM is
// the right operand of the next with no data movement ("K-permuted chaining"); the left operand
// is read at column 16 t + trow(g, e).  Complex products use four real MFMAs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

template <typename R>
struct cx {
    R re, im;
};
typedef cx<float> zf;
typedef cx<double> zd;

// ---------------------------------------------------------------- memory through descriptors
// Most array pointers reach the kernels inside descriptors in memory (systems, work items); the
// compiler treats them as generic and accesses them with flat instructions.  Flat operations are
// counted with the LDS operations as well, so every wait for an LDS access (the exchanges and
// staging of each chain) also waits for the flat loads issued just before it -- the prefetches
// of the next block.  These pointers address device or host-mapped memory, never LDS or
// scratch, so gld / gst access them through global-address-space pointers.
#define HS_GLOBAL __attribute__((address_space(1)))
template <typename T> __device__ __forceinline__ T gld(const T *p) { return *(const HS_GLOBAL T *)p; }
template <typename R> __device__ __forceinline__ cx<R> gld(const cx<R> *p) {
    const HS_GLOBAL R *q = (const HS_GLOBAL R *)p;
    return cx<R>{q[0], q[1]};
}
template <typename T> __device__ __forceinline__ void gst(T *p, T v) { *(HS_GLOBAL T *)p = v; }
template <typename R> __device__ __forceinline__ void gst(cx<R> *p, cx<R> v) {
    HS_GLOBAL R *q = (HS_GLOBAL R *)p;
    q[0] = v.re;
    q[1] = v.im;
}

template <typename T> struct tr;
template <> struct tr<float>  { typedef float  R; typedef f4v V; static constexpr bool C = false; typedef double H; };
template <> struct tr<double> { typedef double R; typedef d4v V; static constexpr bool C = false; typedef double H; };
template <> struct tr<zf>     { typedef float  R; typedef f4v V; static constexpr bool C = true;  typedef zd H; };
template <> struct tr<zd>     { typedef double R; typedef d4v V; static constexpr bool C = true;  typedef zd H; };

// ---------------------------------------------------------------- scalar arithmetic
__device__ __forceinline__ float  s_zero(float)  { return 0.f; }
__device__ __forceinline__ double s_zero(double) { return 0.0; }
template <typename R> __device__ __forceinline__ cx<R> s_zero(cx<R>) { return {R(0), R(0)}; }
__device__ __forceinline__ float  s_one(float)  { return 1.f; }
__device__ __forceinline__ double s_one(double) { return 1.0; }
template <typename R> __device__ __forceinline__ cx<R> s_one(cx<R>) { return {R(1), R(0)}; }

template <typename R> __device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R> __device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R> __device__ __forceinline__ cx<R> operator-(cx<R> a) { return {-a.re, -a.im}; }
template <typename R> __device__ __forceinline__ cx<R> operator*(cx<R> a, cx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R> __device__ __forceinline__ cx<R> s_recip(cx<R> a) {
    // 1 / a with scaling by max(|re|, |im|) to avoid overflow (Smith's method, one division)
    R ar = a.re < 0 ? -a.re : a.re, ai = a.im < 0 ? -a.im : a.im;
    if (ar >= ai) {
        R r = a.im / a.re, d = a.re + a.im * r;
        R inv = R(1) / d;
        return {inv, -r * inv};
    } else {
        R r = a.re / a.im, d = a.re * r + a.im;
        R inv = R(1) / d;
        return {r * inv, -inv};
    }
}
__device__ __forceinline__ float  s_recip(float a)  { return 1.f / a; }
__device__ __forceinline__ double s_recip(double a) { return 1.0 / a; }
__device__ __forceinline__ double s_abs2(float a)  { return (double)a * a; }
__device__ __forceinline__ double s_abs2(double a) { return a * a; }
template <typename R> __device__ __forceinline__ double s_abs2(cx<R> a) { return (double)a.re * a.re + (double)a.im * a.im; }
// cheap modulus for pivot selection: |re| + |im|
__device__ __forceinline__ double s_mag(float a)  { return a < 0 ? -(double)a : (double)a; }
__device__ __forceinline__ double s_mag(double a) { return a < 0 ? -a : a; }
template <typename R> __device__ __forceinline__ double s_mag(cx<R> a) {
    return (a.re < 0 ? -(double)a.re : (double)a.re) + (a.im < 0 ? -(double)a.im : (double)a.im);
}
// the same modulus in the component type (pivot search without conversions)
__device__ __forceinline__ float  s_magr(float a)  { return __builtin_fabsf(a); }
__device__ __forceinline__ double s_magr(double a) { return __builtin_fabs(a); }
template <typename R> __device__ __forceinline__ R s_magr(cx<R> a) { return s_magr(a.re) + s_magr(a.im); }
// reciprocal of a pivot: FP32 takes the hardware reciprocal (1 ulp; FP32 factors only seed the
// refinement), FP64 and complex the correctly rounded division
__device__ __forceinline__ float  s_recip_piv(float a)  { return __builtin_amdgcn_rcpf(a); }
__device__ __forceinline__ double s_recip_piv(double a) { return 1.0 / a; }
template <typename R> __device__ __forceinline__ cx<R> s_recip_piv(cx<R> a) { return s_recip(a); }
__device__ __forceinline__ bool s_finite(float a)  { return __builtin_isfinite(a); }
__device__ __forceinline__ bool s_finite(double a) { return __builtin_isfinite(a); }
template <typename R> __device__ __forceinline__ bool s_finite(cx<R> a) { return __builtin_isfinite(a.re) && __builtin_isfinite(a.im); }

// precision conversions (to the factor type T from the FP64 / complex128 type H and back)
template <typename T, typename H> __device__ __forceinline__ T s_cvt(H x);
template <> __device__ __forceinline__ float  s_cvt<float, double>(double x)  { return (float)x; }
template <> __device__ __forceinline__ double s_cvt<double, double>(double x) { return x; }
template <> __device__ __forceinline__ double s_cvt<double, float>(float x)   { return (double)x; }
template <> __device__ __forceinline__ zf s_cvt<zf, zd>(zd x) { return {(float)x.re, (float)x.im}; }
template <> __device__ __forceinline__ zd s_cvt<zd, zd>(zd x) { return x; }
template <> __device__ __forceinline__ zd s_cvt<zd, zf>(zf x) { return {(double)x.re, (double)x.im}; }
template <> __device__ __forceinline__ float s_cvt<float, float>(float x) { return x; }
template <> __device__ __forceinline__ zf s_cvt<zf, zf>(zf x) { return x; }

// ---------------------------------------------------------------- MFMA tiles
__device__ __forceinline__ d4v mfma_step(double a, double b, d4v c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ __forceinline__ f4v mfma_step(float a, float b, f4v c)   { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// Row of the 16x16 result held by lane group g = l / 16 in element e.  Measured on gfx950:
// v_mfma_f32_16x16x4f32 holds row 4g + e, v_mfma_f64_16x16x4f64 holds row g + 4e (column l % 16
// for both).  Every accumulator-layout index and the K-permuted chaining go through trow().
template <typename T> __device__ __forceinline__ constexpr bool row_major_groups() {
    return sizeof(typename tr<T>::R) == 4;
}
template <typename T> __device__ __forceinline__ int trow(int g, int e) {
    return row_major_groups<T>() ? (4 * g + e) : (g + 4 * e);
}

// ---------------------------------------------------------------- block storage
// Element (r, c) of a b x b factor block (D/Dinv, B/Uq, C, R) is stored at r * b + acol(c):
// rows are contiguous and the columns inside each 16-column group are permuted so that the
// four left-operand values a lane feeds to MFMA steps e = 0..3 of one K tile (columns
// 16 t2 + trow(g, e)) are adjacent -- one Q4 (16-64 byte) load per lane and K tile.
template <typename T> __device__ __forceinline__ int acol(int c) {
    return row_major_groups<T>() ? c : ((c & ~15) | ((c & 3) << 2) | ((c >> 2) & 3));
}
template <typename T> struct alignas(4 * sizeof(T)) Q4 { T v[4]; };

// 16 x 16 accumulator tile distributed over a wave64 (4 elements per lane)
template <typename T, bool C = tr<T>::C> struct Tile;

template <typename T> struct Tile<T, false> {
    typename tr<T>::V v;
    __device__ __forceinline__ void zero() { v = typename tr<T>::V{0, 0, 0, 0}; }
    __device__ __forceinline__ T get(int e) const { return v[e]; }
    __device__ __forceinline__ void set(int e, T x) { v[e] = x; }
    // this += a * b  (one K=4 slice; a, b per the MFMA lane layout)
    __device__ __forceinline__ void mma(T a, T b) { v = mfma_step(a, b, v); }
};

template <typename T> struct Tile<T, true> {
    typedef typename tr<T>::R R;
    typename tr<T>::V re, im;
    __device__ __forceinline__ void zero() { re = typename tr<T>::V{0, 0, 0, 0}; im = re; }
    __device__ __forceinline__ T get(int e) const { return {re[e], im[e]}; }
    __device__ __forceinline__ void set(int e, T x) { re[e] = x.re; im[e] = x.im; }
    __device__ __forceinline__ void mma(T a, T b) {
        re = mfma_step(a.re, b.re, re);
        re = mfma_step(-a.im, b.im, re);
        im = mfma_step(a.re, b.im, im);
        im = mfma_step(a.im, b.re, im);
    }
};

// Extra wait states between the last MFMA of an accumulation chain and the first read of its
// result.  On gfx950 the wait states the compiler inserts before reading a 16x16x4 MFMA result
// were measured to be too few when the wave issues alone on its SIMD (the tail of a launch, or a
// small system on an idle GPU): the read returned partly stale accumulator values -- wrong tips
// in the last workgroup of an FP64 sweep, and rounding-level differences in FP32 between a
// system solved alone and in a batch.  Sixteen more wait states remove both; the asm takes the
// accumulator as an operand, so it is placed after the chain.
template <typename T> __device__ __forceinline__ void tile_settle(Tile<T> &x) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (tr<T>::C) {
        asm volatile("s_nop 7\n\ts_nop 7" : "+a"(x.re), "+a"(x.im));
    } else {
        asm volatile("s_nop 7\n\ts_nop 7" : "+a"(x.v));
    }
#else
    (void)x;
#endif
}

// The same wait states once after the chains of N independent accumulator tiles: one asm that
// takes every tile keeps all of their MFMAs ahead of it (settling tile by tile lets the
// scheduler put sixteen wait states between consecutive independent MFMAs).
template <typename T, int N> __device__ __forceinline__ void tiles_settle(Tile<T> (&x)[N]) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (!tr<T>::C && N == 2) {
        asm volatile("s_nop 7\n\ts_nop 7" : "+a"(x[0].v), "+a"(x[1].v));
    } else if constexpr (!tr<T>::C && N == 3) {
        asm volatile("s_nop 7\n\ts_nop 7" : "+a"(x[0].v), "+a"(x[1].v), "+a"(x[2].v));
    } else if constexpr (!tr<T>::C && N == 4) {
        asm volatile("s_nop 7\n\ts_nop 7" : "+a"(x[0].v), "+a"(x[1].v), "+a"(x[2].v), "+a"(x[3].v));
    } else if constexpr (tr<T>::C && N == 2) {
        asm volatile("s_nop 7\n\ts_nop 7" : "+a"(x[0].re), "+a"(x[0].im), "+a"(x[1].re), "+a"(x[1].im));
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) tile_settle(x[i]);
    }
#else
    (void)x;
#endif
}

template <typename T> __device__ __forceinline__ Q4<T> tile_q(const Tile<T> &x) {
    Q4<T> r;
#pragma unroll
    for (int e = 0; e < 4; ++e) r.v[e] = x.get(e);
    return r;
}

// ---------------------------------------------------------------- block descriptors
// One SPIKE partition of one system.  Blocks of b = 16*NT rows; the partition's last block is
// padded with identity rows.  All block arrays are column-major b x b.
struct PartDesc {
    int sys;        // system index in the batch
    int p, P;       // partition index inside its system, partitions of the system
    int r0, ni;     // first row (in the system) and row count
    int nb;         // number of b-row blocks
    int ku, kl;     // bandwidths of the system
    int pad_;
    int64_t blk;    // index of the partition's first block in the block arrays
};

// One interface (coupling between partitions i and i+1 of a system) of the reduced system.
struct IfaceDesc {
    int sys, i, P, ku, kl;
    int part;       // global index of partition i (partition i+1 is part+1)
    int active_;    // unused
    int pad_;
};

// One sweep work item: a group of up to 16*NWS columns of one partition.
enum { ITEM_Y = 0, ITEM_V = 1, ITEM_W = 2 };
struct SweepItem {
    int part;       // global partition index
    int kind;       // ITEM_Y / ITEM_V / ITEM_W
    int col0;       // first column (RHS column for Y, spike column for V/W)
    int ncols;      // number of columns in this item
    const void *src;   // Y / recovery: system RHS (H type) offset to column col0, row 0 of the system
    int ldsrc;         // leading dimension of src
    int pad0_;
    void *dst;         // recovery: solution (H type) offset to column col0, row 0 of the system
    int lddst;
    int pad1_;
    void *dst_rm;      // recovery: row-major copy of the solution (H type, row stride ldrm), row 0, column 0
    int ldrm;
    int pad2_;
};

// Per-system view used by the row-parallel kernels (validation, conversion, residual).
struct SysDev {
    const int *rowptr;
    const int *colind;
    const void *aij;    // H values
    int n, ku, kl, P;
    int part0;          // first global partition of the system
    int nreg;           // rows per partition except the last: floor(n / P)
    int m;              // RHS columns of the current chunk
    int ldb, ldx, ldr;
    const void *bptr;   // RHS (H), offset to the chunk's first column
    void *xptr;         // solution (H), column-major with leading dimension ldx
    void *rptr;         // residual work (H)
    const void *xrm;    // row-major copy of the solution (row stride ldrm), or null to read xptr
    int ldrm;
    int pad_;
};

struct RowChunk {
    int sys;
    int row0;
};

// Per-system state of a solve kept on the device: the decisions of the refinement passes and of
// the reduced iteration are made by small kernels from it, so a solve's launches need no host
// round trip (eager runs read a summary only to stop early) and can be captured in a graph.
struct SysStat {
    double maxrel;          // largest column residual after the latest pass (with earlier chunks)
    double maxrel_before;   // of the earlier RHS chunks
    double red_update;      // last relative update of the reduced iteration
    int red_iters;          // sweeps of the latest reduced iteration
    int refine_iters;       // refinement passes applied
    int red_active;         // still sweeping in this pass
    int diverged;           // the sweeps stopped contracting in this pass
    int unfinished;         // an earlier pass ended with the sweeps still running or diverged
    int pad_;
};
