// bem_linalg.hip -- dense LU with partial pivoting, triangular solves,
// FP64 residuals and mixed-precision iterative refinement.
//
// Replaces torch.linalg.solve in BEMManager.solve
// (src/managers/BEM_manager.py:1032-1073).  Matrices are row-major (the
// layout of the reference's Taichi field _mat_A and of torch), complex
// interleaved.  Right-looking blocked LU, panel width NB:
//   1. panel: per column, pivot search (|re|+|im|, first maximum, as LAPACK
//      izamax) and row swap inside the panel, then scale + rank-1 update
//   2. the panel's row interchanges applied to the other columns
//   3. U12 = L11^{-1} A12 (unit lower triangular solve)
//   4. A22 -= L21 U12 on MFMA (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32),
//      complex products as four real products
#include <math.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "bem_device.h"
#include "bem_internal.h"

namespace bem {

size_t dtype_size(int dtype) {
    switch (dtype) {
        case BEM_F64: return 8;
        case BEM_C128: return 16;
        case BEM_F32: return 4;
        default: return 8;
    }
}
bool dtype_complex(int dtype) { return dtype == BEM_C128 || dtype == BEM_C64; }

namespace {

constexpr int NB = 64;   // panel width
// Outer block width: NB-column panels inside, one rank-W trailing update per
// block.  512 columns for FP32 / complex64 factors, 256 for FP64 / complex128
// (measured at n = 20,480: the c64 factorisation with split-fp16 updates 160
// -> 146 ms at 512, whose rank-512 updates read and write C half as often;
// the c128 one 471 -> 474 ms, and at n = 5120 slower, so it keeps 256).
template <typename T>
__host__ __device__ constexpr int outer_w() { return sizeof(T) == 4 ? 512 : 256; }

template <typename T>
struct cx {
    T re, im;
};

template <typename T, bool C>
struct Acc {
    // element access of a row-major matrix of T (C: interleaved complex)
    static __device__ __forceinline__ cx<T> ld(const T* a, int64_t idx) {
        if constexpr (C) return {a[2 * idx], a[2 * idx + 1]};
        else return {a[idx], T(0)};
    }
    static __device__ __forceinline__ void st(T* a, int64_t idx, cx<T> v) {
        if constexpr (C) { a[2 * idx] = v.re; a[2 * idx + 1] = v.im; }
        else a[idx] = v.re;
    }
};

template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cx<T> cdiv(cx<T> a, cx<T> b) {
    // Smith's algorithm
    if (fabs((double)b.re) >= fabs((double)b.im)) {
        T r = b.im / b.re, d = b.re + b.im * r;
        return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    T r = b.re / b.im, d = b.im + b.re * r;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}

// 1 / b for b != 0: conj(b) / |b|^2 with one reciprocal (correctly rounded
// division), or Smith's quotient where |b|^2 would leave the normal range.
template <typename T>
__device__ __forceinline__ cx<T> crecip(cx<T> b) {
    const double m = fmax(fabs((double)b.re), fabs((double)b.im));
    const double lo = sizeof(T) == 4 ? 1e-18 : 1e-150, hi = sizeof(T) == 4 ? 1e18 : 1e150;
    if (m < lo || m > hi) return cdiv(cx<T>{T(1), T(0)}, b);
    const T d = T(1) / (b.re * b.re + b.im * b.im);
    return {b.re * d, -b.im * d};
}

// One factorisation of a (possibly batched) LU: the matrix and its
// per-problem scratch.  Kernels take it by value (one problem) or from a
// device table indexed by blockIdx.z (B = true: a batch of uneven sizes
// stepping through the same panel index k together).
struct LuProb {
    void* A;
    int64_t ld;
    int n;
    int32_t* ipiv;
    int* info;
    unsigned* counter;
    double* pval;
    int* pidx;
    void* SL;
    int* tids;   // tournament candidates: two levels of ceil(n / TRW) groups x NB row ids
    void* SU;    // -inv(U11) of the current panel (NB x NB)
    void* pan;   // exchange area of the persistent panel kernel (PanelX<T>)
    long long wait;  // bound of every wait on other workgroups (wall-clock ticks, 100 MHz)
};

template <bool B>
__device__ __forceinline__ LuProb get_prob(const LuProb& inl, const LuProb* tbl) {
    if constexpr (B) return tbl[blockIdx.z];
    else return inl;
}

// Row interchanges of pivots [p_lo, p_hi) applied to the columns in
// [c_lo, c_hi) outside [x_lo, x_hi) (all bounds clamped to the problem's n).
template <typename T, bool C, bool B>
__global__ void laswp_kernel(LuProb inl, const LuProb* tbl, int p_lo, int p_hi, int c_lo, int c_hi, int x_lo,
                             int x_hi) {
    const LuProb P = get_prob<B>(inl, tbl);
    T* A = (T*)P.A;
    const int64_t ld = P.ld;
    const int n = P.n;
    const int32_t* ipiv = P.ipiv;
    p_hi = min(n, p_hi);
    c_hi = min(n, c_hi);
    x_hi = min(n, x_hi);
    const int col = c_lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (p_lo >= p_hi || col >= c_hi || (col >= x_lo && col < x_hi)) return;
    for (int c = p_lo; c < p_hi; ++c) {
        const int p = ipiv[c];
        if (p != c && p >= 0 && p < n) {  // rows outside the matrix (a failed panel) are skipped
            cx<T> a = Acc<T, C>::ld(A, (int64_t)c * ld + col);
            cx<T> b = Acc<T, C>::ld(A, (int64_t)p * ld + col);
            Acc<T, C>::st(A, (int64_t)c * ld + col, b);
            Acc<T, C>::st(A, (int64_t)p * ld + col, a);
        }
    }
}

// The same for block-local pivoting, where every pivot of [p_lo, p_hi)
// (at most NB of them) lies inside those rows: each wave composes the
// interchanges into one permutation of the w rows (lane i ends holding the
// row that lands at p_lo + i), then every column moves with w independent
// loads followed by w stores instead of w dependent swaps, and rows that keep
// their place are not touched.  Pivots outside the rows (a failed panel's)
// are skipped.  FP32 / complex64 (the column's w entries stay in registers).
template <bool C, bool B>
__global__ __launch_bounds__(256) void laswp_local_kernel(LuProb inl, const LuProb* tbl, int p_lo, int p_hi, int c_lo,
                                                          int c_hi, int x_lo, int x_hi) {
    using T = float;
    const LuProb P = get_prob<B>(inl, tbl);
    T* A = (T*)P.A;
    const int64_t ld = P.ld;
    const int n = P.n;
    p_hi = min(n, min(p_hi, p_lo + NB));
    c_hi = min(n, c_hi);
    x_hi = min(n, x_hi);
    const int w = p_hi - p_lo;
    if (w <= 0) return;
    const int lane = threadIdx.x & 63;
    const int pv = lane < w ? P.ipiv[p_lo + lane] - p_lo : lane;
    int src = lane;
    for (int c = 0; c < w; ++c) {
        const int p = __builtin_amdgcn_readlane(pv, c);
        if (p != c && p >= 0 && p < w) {  // uniform
            const int a = __builtin_amdgcn_readlane(src, c), b = __builtin_amdgcn_readlane(src, p);
            src = lane == c ? b : (lane == p ? a : src);
        }
    }
    const int col = c_lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= c_hi || (col >= x_lo && col < x_hi)) return;
    // rows that keep their place (most of them, for diagonally dominant
    // blocks) are neither read nor written
    cx<T> v[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int si = __builtin_amdgcn_readlane(src, i);
        if (i < w && si != i) v[i] = Acc<T, C>::ld(A, (int64_t)(p_lo + si) * ld + col);
    }
#pragma unroll
    for (int i = 0; i < NB; ++i)
        if (i < w && __builtin_amdgcn_readlane(src, i) != i) Acc<T, C>::st(A, (int64_t)(p_lo + i) * ld + col, v[i]);
}

// ---------------------------------------------------------------------------
// Panel column step c (one launch per column of the NB-wide panel [k, kend)):
//   - rows r > c-1 get the rank-1 update of column c-1 (multiplier
//     l_r = A[r][c-1] / A[c-1][c-1], the pivot row c-1 being in place);
//   - the updated column c is searched for the pivot (largest |re|+|im|,
//     first row on ties: LAPACK's rule) with a per-block partial result;
//   - the last block to finish (threadfence + counter) reduces the partials,
//     writes ipiv[c] and swaps rows c and p inside the panel columns.
// Wave lanes map to panel columns so row segments are read coalesced; no
// block ever waits for another.  With c == kend only the update runs.
// ---------------------------------------------------------------------------
constexpr int RPB = 64;  // rows per block (16 per wave)

// Lane l's value of v (l wave-uniform) through scalar registers.
__device__ __forceinline__ double readlane_f64(double v, int l) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Coherent (agent-scope) load of a value another workgroup published.
template <typename T>
__device__ __forceinline__ T ld_agent(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T, bool C, bool B>
__global__ __launch_bounds__(256) void col_step_kernel(LuProb inl, const LuProb* tbl, int k, int c) {
    const LuProb P = get_prob<B>(inl, tbl);
    T* A = (T*)P.A;
    const int64_t ld = P.ld;
    const int n = P.n;
    const int kend = min(n, k + NB);
    if (c > kend) return;  // this problem's panel (or matrix) has ended
    const int nrows = n - (c < kend ? c : kend);
    if (nrows <= 0) return;
    const unsigned nblk = (unsigned)((nrows + RPB - 1) / RPB);  // blocks working on this problem
    if (blockIdx.x >= nblk) return;
    double* pval = P.pval;
    int* pidx = P.pidx;
    unsigned* counter = P.counter;
    int32_t* ipiv = P.ipiv;
    int* info = P.info;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = k + lane;  // panel column of this lane
    const bool search = c < kend;
    const int rbeg = search ? c : kend;  // first row touched (rows >= c; row c-1 is the pivot row)
    __shared__ double wv[4];
    __shared__ int wi[4];
    __shared__ bool s_last;
    cx<T> u = {T(0), T(0)}, diag = {T(1), T(0)};
    const bool upd = c > k;
    if (upd) {
        if (col >= c - 1 && col < kend) u = Acc<T, C>::ld(A, (int64_t)(c - 1) * ld + col);
        diag.re = __shfl(u.re, c - 1 - k, 64);
        diag.im = __shfl(u.im, c - 1 - k, 64);
    }
    const bool dzero = diag.re == T(0) && diag.im == T(0);
    // LAPACK's scaling of the column: multiply by the reciprocal of the pivot
    const cx<T> rdiag = dzero ? cx<T>{T(1), T(0)} : crecip(diag);
    double best = -1.0;
    int brow = 0x7fffffff;
    constexpr int RW = RPB / 4;  // rows per wave
    const int r0 = rbeg + blockIdx.x * RPB + wave * RW;
    const bool mine = col >= (upd ? c - 1 : c) && col < kend;
    // all loads first (independent, in flight together), then update + store
    cx<T> x[RW];
#pragma unroll
    for (int q = 0; q < RW; ++q) {
        x[q] = {T(0), T(0)};
        if (mine && r0 + q < n) x[q] = Acc<T, C>::ld(A, (int64_t)(r0 + q) * ld + col);
    }
#pragma unroll
    for (int q = 0; q < RW; ++q) {
        const int r = r0 + q;
        if (r >= n) break;
        if (upd) {
            cx<T> l = {__shfl(x[q].re, c - 1 - k, 64), __shfl(x[q].im, c - 1 - k, 64)};
            if (!dzero) l = cmul(l, rdiag);
            if (col == c - 1) x[q] = l;
            else if (col >= c && col < kend) {
                cx<T> lu = cmul(l, u);
                x[q].re -= lu.re; x[q].im -= lu.im;
            }
            if (mine) Acc<T, C>::st(A, (int64_t)r * ld + col, x[q]);
        }
        if (search && col == c) {
            double m = fabs((double)x[q].re) + fabs((double)x[q].im);
            if (m > best) { best = m; brow = r; }  // rows visited in increasing order
        }
    }
    if (!search) return;
    if (col == c) { wv[wave] = best; wi[wave] = brow; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = wv[0];
        int bi = wi[0];
        for (int q = 1; q < 4; ++q)
            if (wv[q] > b || (wv[q] == b && wi[q] < bi)) { b = wv[q]; bi = wi[q]; }
        pval[blockIdx.x] = b;
        pidx[blockIdx.x] = bi;
        __threadfence();
        const unsigned done = atomicAdd(counter, 1u);
        s_last = done == nblk - 1;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    __shared__ double rv[256];
    __shared__ int ri[256];
    double b = -1.0;
    int bi = 0x7fffffff;
    for (int g = threadIdx.x; g < (int)nblk; g += 256) {
        double v = ld_agent(&pval[g]);
        int i2 = ld_agent(&pidx[g]);
        if (v > b || (v == b && i2 < bi)) { b = v; bi = i2; }
    }
    rv[threadIdx.x] = b;
    ri[threadIdx.x] = bi;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            double ov = rv[threadIdx.x + off];
            int oi = ri[threadIdx.x + off];
            if (ov > rv[threadIdx.x] || (ov == rv[threadIdx.x] && oi < ri[threadIdx.x])) {
                rv[threadIdx.x] = ov; ri[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    const int p = ri[0];
    if (threadIdx.x == 0) {
        ipiv[c] = p;
        if (rv[0] == 0.0 && *info == 0) *info = c + 1;
        *counter = 0;
    }
    if (p != c && threadIdx.x < kend - k) {
        const int cc = k + threadIdx.x;
        cx<T> a = Acc<T, C>::ld(A, (int64_t)c * ld + cc);
        cx<T> bb = Acc<T, C>::ld(A, (int64_t)p * ld + cc);
        Acc<T, C>::st(A, (int64_t)c * ld + cc, bb);
        Acc<T, C>::st(A, (int64_t)p * ld + cc, a);
    }
}

// ---------------------------------------------------------------------------
// Persistent panel factorisation with partial pivoting: one launch per
// NB-column panel of a single problem.  G workgroups, each owning the R
// consecutive rows [k + gR, k + (g+1)R) of the panel, stay resident and
// synchronise through a grid barrier.  The panel is processed in sub-panels
// of SB columns whose entries (for the workgroup's rows) stay in LDS:
//   per column c: every workgroup publishes its best pivot candidate (row,
//     |re|+|im|, its SB sub-panel values) and, if it owns row c, row c's
//     values; barrier; every workgroup picks the same pivot row p (largest
//     |re|+|im|, first row on ties: LAPACK's rule); workgroup 0 records
//     ipiv[c] and swaps rows c and p in the panel columns outside the
//     sub-panel; the owners of rows c and p swap them in LDS; every
//     workgroup eliminates column c from its rows below c (the same
//     arithmetic as col_step_kernel).
//   per sub-panel: the factored sub-panel is written back; barrier; every
//     workgroup forms U_s = inv(L_ss) A[sub-panel rows, right of it]
//     (SB x < NB) and updates its rows below the sub-panel right of it;
//     workgroup 0 stores U_s after the next barrier, when nobody reads
//     those entries any more.
// A panel costs NB + NB/SB - 1 barriers instead of NB + 1 dependent
// launches, and its right part is read and written once per sub-panel
// instead of once per column.  Every barrier ends in an agent-scope acquire,
// so the data other workgroups stored before it are read with plain loads.
// Every wait is bounded (two seconds of wall clock by default, the context's
// wait limit): a workgroup that gives up raises the shared fail word, every
// other one then leaves at its next wait, and info becomes -2.
// ---------------------------------------------------------------------------
constexpr int SB = 8;        // sub-panel width
constexpr int PGMAX = 64;    // largest workgroup count of a persistent panel (one wave reduces them)
constexpr int PNT = 512;     // threads per workgroup: eight waves (measured: 1024 and 256 were slower)
template <typename T>
constexpr int panel_rmax() { return (128 * 1024) / (SB * (int)sizeof(cx<T>)); }

template <typename T>
struct PanelX {
    unsigned count;       // grid barrier arrivals since the factorisation began
    unsigned pad0;
    int fail, pad1;
    double cval[2][PGMAX];        // candidates per workgroup, double-buffered by column parity
    int crow[2][PGMAX];
    cx<T> cvals[2][PGMAX][SB];
    cx<T> rowc[2][SB];            // row c's sub-panel values, from its owner
};


// Barrier over the workgroups of a persistent launch on a counter that only
// grows during a factorisation: each arrival adds one (release: this
// workgroup's stores before it), then waits with relaxed loads until the
// count reaches target and takes one acquire fence.  (Acquire loads in the
// wait, or a release per poll, would write back or invalidate the L2 of
// the workgroup's XCD on every iteration, beside the trailing-update
// kernels running there.)  Returns false for every workgroup once any wait
// has timed out.
__device__ bool grid_barrier(unsigned* count, int* fail, unsigned target, long long ticks) {
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        const long long t0 = wall_clock64();
        // one load round trip per poll; the fail word and the clock every 16th
        unsigned it = 0;
        while ((int)(__hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) < 0) {
            if ((++it & 15u) == 0 &&
                (__hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || wall_clock64() - t0 > ticks)) {
                ok = 0;
                break;
            }
        }
        if (!ok) __hip_atomic_store(fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// (b, bi) <- (v, r) if v is the better pivot candidate: larger value, then
// smaller row (an order-free rule, so any reduction order picks the same
// candidate).  Bitwise logic and selects, no branches.
__device__ __forceinline__ void take_better(double v, int r, double& b, int& bi) {
    const bool take = (v > b) | ((v == b) & (r < bi));
    b = take ? v : b;
    bi = take ? r : bi;
}

// One step of a wave reduction through a DPP lane pattern (a VALU move, no
// LDS round trip); the patterns used keep every source lane inside the row.
template <int CTRL>
__device__ __forceinline__ void dpp_best(double& b, int& bi) {
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(b), __double2hiint(b), CTRL, 0xF, 0xF, false);
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(b), __double2loint(b), CTRL, 0xF, 0xF, false);
    const int oi = __builtin_amdgcn_update_dpp(bi, bi, CTRL, 0xF, 0xF, false);
    take_better(__hiloint2double(hi, lo), oi, b, bi);
}

// Orders this wave's LDS accesses for the compiler (the hardware runs one
// wave's LDS instructions in order): a store before it is seen by a load of
// another lane of the wave after it.
__device__ __forceinline__ void wave_lds_order() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// x + (x of the lane the DPP pattern selects), for FP64 and FP32 values.
template <int CTRL>
__device__ __forceinline__ double dpp_add(double v) {
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), CTRL, 0xF, 0xF, false);
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), CTRL, 0xF, 0xF, false);
    return v + __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_add(float v) {
    return v + __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over the four lanes of a quad, in every lane of it (whole wave active).
template <typename T>
__device__ __forceinline__ T quad_sum(T v) {
    v = dpp_add<0xB1>(v);     // quad_perm [1, 0, 3, 2]
    return dpp_add<0x4E>(v);  // quad_perm [2, 3, 0, 1]
}

// Best candidate of the wave, returned in every lane (the whole wave must be
// active): within each row of 16 lanes by xor-1 and xor-2 quad permutations,
// the half-row mirror and the row mirror, after which every lane of a row
// holds the row's best; then the four rows through lane reads.
__device__ __forceinline__ void wave_best(double& b, int& bi) {
    dpp_best<0xB1>(b, bi);   // quad_perm [1, 0, 3, 2]
    dpp_best<0x4E>(b, bi);   // quad_perm [2, 3, 0, 1]
    dpp_best<0x141>(b, bi);  // row_half_mirror
    dpp_best<0x140>(b, bi);  // row_mirror
    double rb[4];
    int rr[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        rb[q] = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(b), 16 * q),
                                 __builtin_amdgcn_readlane(__double2loint(b), 16 * q));
        rr[q] = __builtin_amdgcn_readlane(bi, 16 * q);
    }
    take_better(rb[1], rr[1], rb[0], rr[0]);
    take_better(rb[3], rr[3], rb[2], rr[2]);
    take_better(rb[2], rr[2], rb[0], rr[0]);
    b = rb[0];
    bi = rr[0];
}

// Barriers one persistent panel launch of w columns takes (one per column,
// one after every sub-panel but the last).
__host__ __device__ inline int panel_barriers(int w) { return w + (w + SB - 1) / SB - 1; }

template <typename T, bool C>
__global__ __launch_bounds__(PNT) void panel_persist_kernel(LuProb P, int k, int G, int R, unsigned base) {
    constexpr int RMAX = panel_rmax<T>();
    __shared__ cx<T> sub[RMAX][SB];   // this workgroup's rows, sub-panel columns
    __shared__ cx<T> us[SB][NB];      // U_s of the sub-panel
    __shared__ cx<T> lss[SB][SB];     // L_ss (unit lower) of the sub-panel
    __shared__ cx<T> urow[SB], orow[SB];
    __shared__ cx<T> allc[PGMAX][SB];   // every workgroup's candidate row values
    __shared__ double cv_s[PGMAX];
    __shared__ int cr_s[PGMAX];
    __shared__ double wv[PNT / 64];   // the waves' candidates
    __shared__ int wi[PNT / 64];
    T* A = (T*)P.A;
    const int64_t ld = P.ld;
    const int n = P.n, kend = min(n, k + NB);
    PanelX<T>* X = (PanelX<T>*)P.pan;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, g = blockIdx.x;
    const int rb = k + g * R, re = min(n, rb + R), nr = max(0, re - rb);
    int par = 0;
    unsigned target = base;  // barrier targets: base + G, base + 2G, ...
    bool pending_u = false;
    int pu_cs = 0, pu_sw = 0, pu_ce = 0;
    auto fail_out = [&]() {
        if (t == 0) __hip_atomic_store(P.info, -2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    for (int cs = k; cs < kend; cs += SB) {
        const int ce = min(kend, cs + SB), sw = ce - cs;
        const int i0 = max(0, cs - rb);
        // (plain loads: every wait above ended in an agent-scope acquire)
        for (int e = i0 * SB + t; e < nr * SB; e += PNT) {
            const int i = e / SB, j = e % SB;
            if (j < sw) sub[i][j] = Acc<T, C>::ld(A, (int64_t)(rb + i) * ld + cs + j);
        }
        __syncthreads();
        // this thread's best row of the column being searched (rows ascending
        // per thread, ties to the first row); later columns of the sub-panel
        // are searched during the previous column's elimination
        double b = -1.0;
        int bi = 0x7fffffff;
        for (int i = i0 + t; i < nr; i += PNT) {
            const cx<T> v = sub[i][0];
            const double m = fabs((double)v.re) + fabs((double)v.im);
            const bool tk = m > b;
            b = tk ? m : b;
            bi = tk ? rb + i : bi;
        }
        for (int c = cs; c < ce; ++c) {
            const int jc = c - cs;
            // the workgroup's candidate: a DPP reduction per wave, then every
            // thread folds the PNT / 64 wave results as a tree
            wave_best(b, bi);
            if (lane == 0) { wv[wave] = b; wi[wave] = bi; }
            __syncthreads();
            {
                double x[PNT / 64];
                int y[PNT / 64];
#pragma unroll
                for (int q = 0; q < PNT / 64; ++q) { x[q] = wv[q]; y[q] = wi[q]; }
#pragma unroll
                for (int w = PNT / 128; w > 0; w >>= 1)
#pragma unroll
                    for (int q = 0; q < w; ++q) take_better(x[q + w], y[q + w], x[q], y[q]);
                b = x[0];
                bi = y[0];
            }
            if (t == 0) { X->cval[par][g] = b; X->crow[par][g] = bi; }
            if (t < sw) {
                if (bi != 0x7fffffff) X->cvals[par][g][t] = sub[bi - rb][t];
                if (c >= rb && c < re) X->rowc[par][t] = sub[c - rb][t];
            }
            target += (unsigned)G;
            if (!grid_barrier(&X->count, &X->fail, target, P.wait)) { fail_out(); return; }
            if (c == cs && pending_u) {
                // every workgroup has read the previous sub-panel's rows: store U_s
                if (g == 0) {
                    const int W = kend - pu_ce;
                    for (int e = t; e < pu_sw * W; e += PNT)
                        Acc<T, C>::st(A, (int64_t)(pu_cs + e / W) * ld + pu_ce + e % W, us[e / W][e % W]);
                }
                pending_u = false;
            }
            // every candidate (value, row, sub-panel values) and row c in one
            // round of loads, then the same reduction in every workgroup
            for (int e = t; e < G * (SB + 1) + SB; e += PNT) {
                if (e >= G * (SB + 1)) {
                    const int jj = e - G * (SB + 1);
                    if (jj < sw) orow[jj] = X->rowc[par][jj];
                    continue;
                }
                const int q = e / (SB + 1), jj = e % (SB + 1);
                if (jj == SB) {
                    cv_s[q] = X->cval[par][q];
                    cr_s[q] = X->crow[par][q];
                } else if (jj < sw) {
                    allc[q][jj] = X->cvals[par][q][jj];
                }
            }
            __syncthreads();
            // the pivot: every wave reduces the G workgroup candidates (lane q
            // holds workgroup q's) and finds the owner of the winner by ballot
            int p, gw;
            {
                const double v0 = lane < G ? cv_s[lane] : -1.0;
                const int r0 = lane < G ? cr_s[lane] : 0x7fffffff;
                b = v0;
                bi = r0;
                wave_best(b, bi);
                const unsigned long long win = __ballot(lane < G && v0 == b && r0 == bi);
                p = bi;
                gw = win ? __ffsll((unsigned long long)win) - 1 : 0;
            }
            const cx<T> piv = allc[gw][jc];
            const bool dz = piv.re == T(0) && piv.im == T(0);
            // LAPACK's scaling of the column: multiply by the reciprocal of the pivot
            const cx<T> rpiv = dz ? cx<T>{T(1), T(0)} : crecip(piv);
            // workgroup 0 interchanges rows c and p in the panel columns outside
            // the sub-panel: loads now, stores after the elimination below
            const int scol = k + t;
            const bool swp = g == 0 && p != c && t < kend - k && (scol < cs || scol >= ce);
            cx<T> sa = {T(0), T(0)}, sb2 = {T(0), T(0)};
            if (swp) {
                sa = Acc<T, C>::ld(A, (int64_t)c * ld + scol);
                sb2 = Acc<T, C>::ld(A, (int64_t)p * ld + scol);
            }
            if (g == 0 && t == 0) {
                P.ipiv[c] = p;
                if (dz && *P.info == 0) *P.info = c + 1;
            }
            if (t < sw) urow[t] = allc[gw][t];
            if (p != c && t < sw) {
                if (p >= rb && p < re) sub[p - rb][t] = orow[t];
                if (c >= rb && c < re) sub[c - rb][t] = allc[gw][t];
            }
            __syncthreads();
            b = -1.0;
            bi = 0x7fffffff;
            for (int i = max(0, c + 1 - rb) + t; i < nr; i += PNT) {
                cx<T> l = sub[i][jc];
                if (!dz) l = cmul(l, rpiv);
                sub[i][jc] = l;
#pragma unroll
                for (int j = 1; j < SB; ++j) {
                    if (j > jc && j < sw) {
                        const cx<T> lu = cmul(l, urow[j]);
                        sub[i][j].re -= lu.re;
                        sub[i][j].im -= lu.im;
                    }
                }
                if (jc + 1 < sw) {
                    const double m = fabs((double)sub[i][jc + 1].re) + fabs((double)sub[i][jc + 1].im);
                    const bool tk = m > b;
                    b = tk ? m : b;
                    bi = tk ? rb + i : bi;
                }
            }
            if (swp) {
                Acc<T, C>::st(A, (int64_t)c * ld + scol, sb2);
                Acc<T, C>::st(A, (int64_t)p * ld + scol, sa);
            }
            __syncthreads();
            par ^= 1;
        }
        for (int e = i0 * SB + t; e < nr * SB; e += PNT) {
            const int i = e / SB, j = e % SB;
            if (j < sw) Acc<T, C>::st(A, (int64_t)(rb + i) * ld + cs + j, sub[i][j]);
        }
        if (ce >= kend) break;
        target += (unsigned)G;
        if (!grid_barrier(&X->count, &X->fail, target, P.wait)) { fail_out(); return; }
        // U_s = inv(L_ss) A[cs:ce, ce:kend], formed redundantly by every workgroup
        const int W = kend - ce;
        if (t < SB * SB) {
            const int m = t / SB, q = t % SB;
            lss[m][q] = (m < sw && q < m) ? Acc<T, C>::ld(A, (int64_t)(cs + m) * ld + cs + q) : cx<T>{T(0), T(0)};
        }
        __syncthreads();
        if (t < W) {
            for (int m = 0; m < sw; ++m) {
                cx<T> v = Acc<T, C>::ld(A, (int64_t)(cs + m) * ld + ce + t);
                for (int q = 0; q < m; ++q) {
                    const cx<T> lu = cmul(lss[m][q], us[q][t]);
                    v.re -= lu.re;
                    v.im -= lu.im;
                }
                us[m][t] = v;
            }
        }
        __syncthreads();
        // rows below the sub-panel: A[r, ce:kend] -= L[r, cs:ce] U_s
        // in chunks of U elements per thread whose loads are all issued first
        const int i1 = max(0, ce - rb), tot = max(0, nr - i1) * W;
        constexpr int U = 8;
        for (int base = 0; base < tot; base += PNT * U) {
            cx<T> v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = base + u * PNT + t;
                if (e < tot) v[u] = Acc<T, C>::ld(A, (int64_t)(rb + i1 + e / W) * ld + ce + e % W);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = base + u * PNT + t;
                if (e < tot) {
                    const int i = i1 + e / W, j = e % W;
                    for (int m = 0; m < sw; ++m) {
                        const cx<T> lu = cmul(sub[i][m], us[m][j]);
                        v[u].re -= lu.re;
                        v[u].im -= lu.im;
                    }
                    Acc<T, C>::st(A, (int64_t)(rb + i) * ld + ce + j, v[u]);
                }
            }
        }
        __syncthreads();  // the next sub-panel's load reads these rows through other threads
        pending_u = true;
        pu_cs = cs;
        pu_sw = sw;
        pu_ce = ce;
    }
}

// ---------------------------------------------------------------------------
// Tournament pivoting of a panel (columns [k, kend), rows [k, n)), the
// communication-avoiding alternative to one global pivot search per column:
//   level 0: every block of TRW rows runs GEPP in LDS and keeps its w pivot
//            rows as candidates;
//   level L: a block stacks the candidates of four groups of level L-1
//            (their original rows, gathered from the matrix) and runs GEPP
//            again, until one group is left;
//   the surviving rows, in their GEPP order, are the panel's pivot rows.
// Selection runs in FP32 (|re| + |im|, first row on ties).  The panel is
// then permuted, its top w x w block factored without pivoting and
// L21 = A21 inv(U11) formed on MFMA.  This replaces 64 dependent launches per
// panel by about ten.
// ---------------------------------------------------------------------------
constexpr int TRW = 256;     // candidate rows per tournament block
constexpr int TLD = NB + 1;  // padded LDS row

__device__ __forceinline__ int tourn_stride(int n) { return ((n + TRW - 1) / TRW) * NB; }

// GEPP of the TRW rows held in sm (w columns); rows with id < 0 are empty.
// Leaves the chosen rows at the top of sm/id, in pivot order, and returns how
// many were chosen.  256 threads, thread t owns row t.  The row update runs
// in aligned chunks of 8 columns (vector LDS access); columns at or left of
// the pivot column get harmless values, they are never read again.
__device__ int tourn_gepp(float2 (*sm)[TLD], int* id, int w, float* rv, int* ri) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    int chosen = 0;
    for (int j = 0; j < w; ++j) {
        float v = -1.f;
        if (t >= j && id[t] >= 0) {
            const float2 a = sm[t][j];
            v = fabsf(a.x) + fabsf(a.y);
        }
        int bi = t;
        for (int off = 32; off > 0; off >>= 1) {
            const float ov = __shfl_xor(v, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            if (ov > v || (ov == v && oi < bi)) { v = ov; bi = oi; }
        }
        if (lane == 0) { rv[wave] = v; ri[wave] = bi; }
        __syncthreads();
        float bv = rv[0];
        int p = ri[0];
        for (int q = 1; q < 4; ++q)
            if (rv[q] > bv || (rv[q] == bv && ri[q] < p)) { bv = rv[q]; p = ri[q]; }
        if (bv < 0.f) break;  // no candidate rows left (uniform)
        if (p != j) {
            if (t < w) { const float2 x = sm[j][t]; sm[j][t] = sm[p][t]; sm[p][t] = x; }
            if (t == 0) { const int x = id[j]; id[j] = id[p]; id[p] = x; }
        }
        __syncthreads();
        if (t > j && id[t] >= 0 && bv > 0.f) {
            const cx<float> d = {sm[j][j].x, sm[j][j].y};
            const cx<float> l = cdiv(cx<float>{sm[t][j].x, sm[t][j].y}, d);
            for (int c0 = (j + 1) & ~7; c0 < w; c0 += 8) {
                float2 u[8], x[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) { u[q] = sm[j][c0 + q]; x[q] = sm[t][c0 + q]; }
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    x[q].x -= l.re * u[q].x - l.im * u[q].y;
                    x[q].y -= l.re * u[q].y + l.im * u[q].x;
                }
#pragma unroll
                for (int q = 0; q < 8; ++q) sm[t][c0 + q] = x[q];
            }
        }
        __syncthreads();
        ++chosen;
    }
    return chosen;
}

template <typename T, bool C, bool B>
__global__ __launch_bounds__(256) void tourn_kernel(LuProb inl, const LuProb* tbl, int k, int level) {
    const LuProb P = get_prob<B>(inl, tbl);
    const int n = P.n;
    if (k >= n) return;
    const int kend = min(n, k + NB), w = kend - k, m = n - k;
    int G = (m + TRW - 1) / TRW, Gprev = 0;
    for (int L = 0; L < level; ++L) { Gprev = G; G = (G + 3) / 4; }
    const int g = blockIdx.x;
    if (g >= G) return;
    __shared__ float2 sm[TRW][TLD];
    __shared__ int id[TRW];
    __shared__ float rv[4];
    __shared__ int ri[4];
    const T* A = (const T*)P.A;
    const int stride = tourn_stride(n);
    const int t = threadIdx.x;
    if (level == 0) {
        const int r = k + g * TRW + t;
        id[t] = r < n ? r : -1;
    } else {
        const int child = 4 * g + (t >> 6), slot = t & 63;
        id[t] = child < Gprev ? P.tids[((level - 1) & 1) * stride + child * NB + slot] : -1;
    }
    __syncthreads();
    // thread t gathers column t % 64 of rows t / 64 + 4 q, 16 loads in flight
    {
        const int c = t & 63, r0 = t >> 6;
        for (int q0 = 0; q0 < TRW / 4; q0 += 16) {
            cx<T> a[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int r = r0 + 4 * (q0 + q);
                const int rid = id[r];
                a[q] = (c < w && rid >= 0) ? Acc<T, C>::ld(A, (int64_t)rid * P.ld + k + c) : cx<T>{T(0), T(0)};
            }
#pragma unroll
            for (int q = 0; q < 16; ++q) sm[r0 + 4 * (q0 + q)][c] = make_float2((float)a[q].re, (float)a[q].im);
        }
    }
    __syncthreads();
    const int chosen = tourn_gepp(sm, id, w, rv, ri);
    if (t < NB) P.tids[(level & 1) * stride + g * NB + t] = t < chosen ? id[t] : -1;
}

// LAPACK-style interchanges from the chosen rows sel[0..w): the row chosen
// for position k+c sits at its own index unless a previous interchange moved
// it, and a moved row is found by following ipiv from its index.
template <bool B>
__global__ void tourn_ipiv_kernel(LuProb inl, const LuProb* tbl, int k, int level) {
    const LuProb P = get_prob<B>(inl, tbl);
    const int n = P.n;
    if (k >= n || threadIdx.x != 0) return;
    const int kend = min(n, k + NB), w = kend - k;
    const int* sel = P.tids + (level & 1) * tourn_stride(n);
    for (int c = 0; c < w; ++c) {
        int pos = sel[c];
        if (pos < 0) pos = k + c;  // cannot happen: the panel has at least w rows
        while (pos < k + c) pos = P.ipiv[pos];
        P.ipiv[k + c] = pos;
    }
}

// LU without pivoting of the top w x w block of the panel in LDS (tournament
// pivoting has already moved the pivot rows up), then S_U = -inv(U11) for
// L21 = A21 inv(U11) (GEMM mode 5 stores -(A21 S_U)).
template <typename T, bool C, bool B>
__global__ __launch_bounds__(256) void topblock_kernel(LuProb inl, const LuProb* tbl, int k) {
    const LuProb P = get_prob<B>(inl, tbl);
    const int n = P.n;
    if (k >= n) return;
    const int kend = min(n, k + NB), w = kend - k;
    T* A = (T*)P.A;
    const int64_t ld = P.ld;
    __shared__ cx<T> M[NB][NB + 1];
    __shared__ cx<T> X[NB][NB + 1];
    const int t = threadIdx.x;
    {
        cx<T> a[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int e = t + 256 * q, r = e / NB, c = e % NB;
            a[q] = (r < w && c < w) ? Acc<T, C>::ld(A, (int64_t)(k + r) * ld + k + c) : cx<T>{T(r == c ? 1 : 0), T(0)};
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int e = t + 256 * q;
            M[e / NB][e % NB] = a[q];
        }
    }
    __syncthreads();
    for (int j = 0; j < w; ++j) {
        const cx<T> d = M[j][j];
        const bool zero = d.re == T(0) && d.im == T(0);
        if (zero && t == 0 && *P.info == 0) *P.info = k + j + 1;
        if (!zero)
            for (int i = j + 1 + t; i < w; i += 256) M[i][j] = cdiv(M[i][j], d);
        __syncthreads();
        const int rem = w - j - 1;
        for (int e = t; e < rem * rem; e += 256) {
            const int i = j + 1 + e / rem, c = j + 1 + e % rem;
            const cx<T> lu = cmul(M[i][j], M[j][c]);
            M[i][c].re -= lu.re;
            M[i][c].im -= lu.im;
        }
        __syncthreads();
    }
    for (int e = t; e < w * w; e += 256) {
        const int r = e / w, c = e % w;
        Acc<T, C>::st(A, (int64_t)(k + r) * ld + k + c, M[r][c]);
    }
    // X = inv(U11) by anti-diagonals d = c - i: X[i][i+d] needs only entries
    // of smaller d.  Four threads per entry split the dot product.
    for (int e = t; e < NB * NB; e += 256) X[e / NB][e % NB] = cx<T>{T(0), T(0)};
    __syncthreads();
    {
        const int g = t >> 2, q = t & 3;
        for (int d = 0; d < NB; ++d) {
            cx<T> sum = {T(0), T(0)};
            const int i = g, c = g + d;
            if (c < NB)
                for (int j = i + 1 + q; j <= c; j += 4) {
                    const cx<T> ux = cmul(M[i][j], X[j][c]);
                    sum.re += ux.re;
                    sum.im += ux.im;
                }
            sum.re += __shfl_xor(sum.re, 1, 64);
            sum.im += __shfl_xor(sum.im, 1, 64);
            sum.re += __shfl_xor(sum.re, 2, 64);
            sum.im += __shfl_xor(sum.im, 2, 64);
            if (c < NB && q == 0) {
                const cx<T> num = {T(d == 0 ? 1 : 0) - sum.re, -sum.im};
                const cx<T> dii = M[i][i];
                const bool z = dii.re == T(0) && dii.im == T(0);
                X[i][c] = z ? num : cdiv(num, dii);
            }
            __syncthreads();
        }
    }
    T* SU = (T*)P.SU;
    for (int e = t; e < NB * NB; e += 256) {
        const int i = e / NB, c = e % NB;
        Acc<T, C>::st(SU, e, cx<T>{-X[i][c].re, -X[i][c].im});
    }
}

// Block-local pivoting panel top in one workgroup of eight waves, the block
// held in registers (wave q: columns 8q .. 8q + 7 of every row, lane r:
// row r), rows never moved (each lane tracks its row's position):
//  * column j: its owner wave finds the pivot (largest |re|+|im| among the
//    rows not yet pivoted, first position on ties; ipiv as LAPACK's),
//    forms the multipliers and publishes them with the pivot lane in LDS;
//    after one workgroup barrier every wave updates its columns right of j
//    with the pivot row's values read from the pivot lane (readlane);
//  * the factored block is written in pivoted order (LDS, then the matrix);
//  * waves 0-3 form SU = -inv(U11) and waves 4-7 S_L = I - inv(L11), 16
//    columns each, by Gauss-Jordan elimination on the identity held in
//    registers (row i of the inverse read from lane i), the factor's
//    entries read from LDS: no barrier per row;
//  * S_L goes to slot (k mod W) / 64 of the outer block, so the separate
//    linv_kernel launch is not needed for this pivoting.
// The block past w (the matrix's last partial block) is the identity.
template <typename T>
__device__ __forceinline__ cx<T> readlane_cx(cx<T> v, int l) {
    if constexpr (sizeof(T) == 8) {
        const double re = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v.re), l),
                                           __builtin_amdgcn_readlane(__double2loint(v.re), l));
        const double im = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v.im), l),
                                           __builtin_amdgcn_readlane(__double2loint(v.im), l));
        return {re, im};
    } else {
        return {__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.re), l)),
                __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.im), l))};
    }
}

template <typename T, bool C, bool B>
__global__ __launch_bounds__(512) void topblock_local_kernel(LuProb inl, const LuProb* tbl, int k) {
    const LuProb P = get_prob<B>(inl, tbl);
    const int n = P.n;
    if (k >= n) return;
    const int kend = min(n, k + NB), w = kend - k;
    T* A = (T*)P.A;
    const int64_t ld = P.ld;
    // the block in LDS, entry (r, c) at column (c + r) mod 64 of row r: rows
    // and columns both read without bank conflicts, no padding, so that with
    // the multipliers in the same space the kernel needs 32 KB (complex64)
    // and fits beside two workgroups of the trailing-update GEMM
    __shared__ cx<T> M[NB][NB];
    cx<T>* const Lc = &M[0][0];           // [2][NB] multipliers of column j, by parity of j (factorisation only)
    int* const Pv = (int*)&M[2][0];       // [2] pivot lane of column j (factorisation only)
    auto at = [&](int r, int c) -> cx<T>& { return M[r][(c + r) & (NB - 1)]; };
    // the wave index as a scalar, so that every per-wave choice below is a scalar branch
    const int t = threadIdx.x, lane = t & 63, wave = __builtin_amdgcn_readfirstlane(t >> 6);
    for (int r = wave; r < NB; r += 8)
        at(r, lane) = (r < w && lane < w) ? Acc<T, C>::ld(A, (int64_t)(k + r) * ld + k + lane)
                                          : cx<T>{T(r == lane ? 1 : 0), T(0)};
    __syncthreads();
    cx<T> v[8];
#pragma unroll
    for (int cc = 0; cc < 8; ++cc) v[cc] = at(lane, 8 * wave + cc);
    __syncthreads();  // the block is in registers: its LDS space now holds the multipliers
    int pos = lane;
    bool done = false;
    for (int jq = 0; jq < 8; ++jq) {
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
            const int j = 8 * jq + jj, par = j & 1;
            if (wave == jq) {
                double b = done ? -1.0 : fabs((double)v[jj].re) + fabs((double)v[jj].im);
                int bi = pos;
                wave_best(b, bi);
                const int p = __builtin_ffsll(__ballot(pos == bi)) - 1;
                const cx<T> d = readlane_cx(v[jj], p);
                const bool zero = d.re == T(0) && d.im == T(0);
                const bool upd = !done && lane != p;
                // the column times the pivot's reciprocal (as LAPACK scales)
                const cx<T> rd = zero ? cx<T>{T(1), T(0)} : crecip(d);
                cx<T> l = {T(0), T(0)};
                if (upd) {
                    l = cmul(v[jj], rd);
                    v[jj] = l;
                }
                Lc[par * NB + lane] = l;
                if (lane == 0) {
                    Pv[par] = p;
                    if (j < w) P.ipiv[k + j] = k + bi;
                    if (zero && *P.info == 0) *P.info = k + j + 1;
                }
            }
            __syncthreads();
            const int p = Pv[par];
            const cx<T> l = Lc[par * NB + lane];
#pragma unroll
            for (int cc = 0; cc < 8; ++cc)
                if (8 * wave + cc > j) {
                    const cx<T> u = readlane_cx(v[cc], p);
                    v[cc].re -= l.re * u.re - l.im * u.im;
                    v[cc].im -= l.re * u.im + l.im * u.re;
                }
            const int ppos = __builtin_amdgcn_readlane(pos, p);
            if (pos == j) pos = ppos;
            if (lane == p) {
                pos = j;
                done = true;
            }
        }
    }
    __syncthreads();  // every wave has read the last multipliers
#pragma unroll
    for (int cc = 0; cc < 8; ++cc) at(pos, 8 * wave + cc) = v[cc];
    __syncthreads();
    for (int r = wave; r < w; r += 8)
        if (lane < w) Acc<T, C>::st(A, (int64_t)(k + r) * ld + k + lane, at(r, lane));
    // inverses: lane = row (in pivoted order), 16 columns per wave
    const int c0 = 16 * (wave & 3);
    cx<T> x[16];
#pragma unroll
    for (int cc = 0; cc < 16; ++cc) x[cc] = cx<T>{T(lane == c0 + cc ? 1 : 0), T(0)};
    if (wave < 4) {
        // X = inv(U11) = inv(D^-1 U11) D^-1, D = diag(U11): the unit upper
        // factor's inverse by back substitution (for i = 63 .. 0, rows above
        // subtract U[r][i] / U[r][r] times row i, which is final; only
        // columns c >= i are nonzero), then column c times 1 / U[c][c].  The
        // same work per step as inv(L11) on the other four waves.
        const cx<T> dl = at(lane, lane);
        const cx<T> rdl = dl.re == T(0) && dl.im == T(0) ? cx<T>{T(1), T(0)} : crecip(dl);
        for (int i = NB - 1; i >= 0; --i) {
            // rows above i subtract; row i and the rows below keep theirs
            // (a zero multiplier, no branch)
            const cx<T> m = cmul(at(lane, i), rdl);
            const cx<T> mm = lane < i ? m : cx<T>{T(0), T(0)};
#pragma unroll
            for (int cc = 0; cc < 16; ++cc)
                if (c0 + cc >= i) {
                    const cx<T> xi = readlane_cx(x[cc], i);
                    x[cc].re -= mm.re * xi.re - mm.im * xi.im;
                    x[cc].im -= mm.re * xi.im + mm.im * xi.re;
                }
        }
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) {
            const cx<T> d = at(c0 + cc, c0 + cc);
            const cx<T> rd = d.re == T(0) && d.im == T(0) ? cx<T>{T(1), T(0)} : crecip(d);
            x[cc] = cmul(x[cc], rd);
        }
    } else {
        // X = inv(L11), unit lower: for i = 0 .. 63, rows below subtract
        // L[r][i] times row i (only columns c <= i are nonzero)
        for (int i = 0; i < NB; ++i) {
            const cx<T> m0 = at(lane, i);
            const cx<T> m = lane > i ? m0 : cx<T>{T(0), T(0)};
#pragma unroll
            for (int cc = 0; cc < 16; ++cc)
                if (c0 + cc <= i) {
                    const cx<T> xi = readlane_cx(x[cc], i);
                    x[cc].re -= m.re * xi.re - m.im * xi.im;
                    x[cc].im -= m.re * xi.im + m.im * xi.re;
                }
        }
    }
    // SU = -inv(U11) (zero below the diagonal), S_L = I - inv(L11) (zero on and above it)
    T* S = wave < 4 ? (T*)P.SU : (T*)P.SL + (C ? 2 : 1) * (int64_t)((k % outer_w<T>()) / NB) * NB * NB;
#pragma unroll
    for (int cc = 0; cc < 16; ++cc) {
        const int c = c0 + cc;
        const bool keep = wave < 4 ? lane <= c : lane > c;
        Acc<T, C>::st(S, lane * NB + c, keep ? cx<T>{-x[cc].re, -x[cc].im} : cx<T>{T(0), T(0)});
    }
}

// S_L = I - inv(L11) for the unit lower w x w block at (k, k), so that
// U12 = inv(L11) A12 = A12 - S_L A12 runs on the MFMA GEMM (|l| <= 1 keeps
// S_L moderate, so the subtraction does not cancel).
template <typename T, bool C, bool B>
__global__ __launch_bounds__(256) void linv_kernel(LuProb inl, const LuProb* tbl, int k, int slot) {
    const LuProb P = get_prob<B>(inl, tbl);
    const T* A = (const T*)P.A;
    const int64_t ld = P.ld;
    const int kend = min(P.n, k + NB), w = kend - k;
    if (kend >= P.n) return;  // no columns right of the block
    T* SL = (T*)P.SL + (C ? 2 : 1) * (int64_t)slot * NB * NB;
    __shared__ cx<T> M[NB][NB + 1];
    __shared__ cx<T> X[NB][NB + 1];
    const int t = threadIdx.x;
    {
        cx<T> a[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int e = t + 256 * q, r = e / NB, c = e % NB;
            a[q] = (r < w && c < w) ? Acc<T, C>::ld(A, (int64_t)(k + r) * ld + k + c) : cx<T>{T(r == c ? 1 : 0), T(0)};
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int e = t + 256 * q;
            M[e / NB][e % NB] = a[q];
            X[e / NB][e % NB] = cx<T>{T(0), T(0)};
        }
    }
    __syncthreads();
    // X = inv(L11) (unit lower) by forward substitution, one column per quad
    // of lanes: X[i][c] = -sum_{c <= j < i} M[i][j] X[j][c] for i > c.  A
    // column only reads its own entries, so each wave runs its 16 columns on
    // its own, ordered through its LDS accesses (one wave's LDS instructions
    // execute in order), with no barrier between rows; the four lanes of a
    // quad split the sum and add it by DPP.
    {
        const int lane = t & 63, c = 16 * (t >> 6) + (lane >> 2), q = lane & 3;
        if (q == 0) X[c][c] = cx<T>{T(1), T(0)};
        for (int i = 1; i < NB; ++i) {
            wave_lds_order();
            cx<T> sum = {T(0), T(0)};
            if (i > c)
                for (int j = c + q; j < i; j += 4) {
                    const cx<T> lx = cmul(M[i][j], X[j][c]);
                    sum.re += lx.re;
                    sum.im += lx.im;
                }
            sum.re = quad_sum(sum.re);
            sum.im = quad_sum(sum.im);
            if (i > c && q == 0) X[i][c] = cx<T>{-sum.re, -sum.im};
        }
    }
    __syncthreads();
    for (int e = t; e < NB * NB; e += 256) {
        const int i = e / NB, c = e % NB;
        Acc<T, C>::st(SL, e, cx<T>{T(i == c ? 1 : 0) - X[i][c].re, -X[i][c].im});
    }
}

// ---------------------------------------------------------------------------
// C[M x N] -= A[M x K] B[K x N] on MFMA.  Block tile 64 x 64, four waves in
// a 2 x 2 arrangement, each wave 32 x 32 = 2 x 2 MFMA 16x16 blocks, K staged
// through LDS in chunks of KC (32 for FP32, 16 for FP64).  MFMA 16x16x4 operand layout: lane l supplies
// A[l%16][l/16] and B[l/16][l%16].  Result register r of lane l holds column
// l%16 of row 4*(l/16)+r for v_mfma_f32_16x16x4f32 but of row l/16+4*r for
// v_mfma_f64_16x16x4f64 (both measured on gfx950).  Complex products use four
// real MFMAs.  C may alias B when M <= 64 and each column tile is owned by
// one block (all of a block's B reads precede its C writes).
// ---------------------------------------------------------------------------
template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 op(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 op(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int GT = 64;   // block tile
constexpr int WT = 32;   // wave tile

// Operands of one C -= A B.
template <typename T>
struct GemmArgs {
    T* Cm;
    int64_t ldc;
    const T* Am;
    int64_t lda;
    const T* Bm;
    int64_t ldb;
    int M, N, K;
};

// The LU products of one step for problem P (bounds clamped to its n).  The
// factorisation is blocked twice: inner panels of NB columns inside outer
// blocks [K0, K1) of outer_w<T>() columns.
//   mode 0  inner U12:      A[k:kend, kend:K1] -= S_L A[k:kend, kend:K1]   (in place)
//   mode 1  inner trailing: A[kend:n, kend:K1] -= A[kend:n, k:kend] A[k:kend, kend:K1]
//   mode 2  outer U12, block b, elimination of the solved blocks above it:
//           A[r0:r0+wb, K1:n] -= A[r0:r0+wb, K0:r0] A[K0:r0, K1:n]
//   mode 3  outer U12, block b: A[r0:r0+wb, K1:n] -= S_L,b A[r0:r0+wb, K1:n]   (in place)
//   mode 4  outer trailing: A[K1:n, c0:c1] -= A[K1:n, K0:K1] A[K0:K1, c0:c1]
//   mode 5  tournament panel: A[kend:n, k:kend] = A[kend:n, k:kend] inv(U11)
//           (the kernel stores -(A B) with B = -inv(U11) instead of
//           subtracting; in place: N <= 64, so every block owns whole rows)
//           (columns c0 < 0 -> K1, c1 <= 0 -> n: the whole trailing matrix)
// with r0 = K0 + b NB, S_L = I - inv(L_bb) of the block's unit lower factor.

struct LuStep {
    int mode, K0, k, b;
    int c0 = -1, c1 = 0;
};

template <typename T, bool C>
__host__ __device__ inline GemmArgs<T> lu_gemm_args(const LuProb& P, const LuStep& st) {
    constexpr int E = C ? 2 : 1;
    T* A = (T*)P.A;
    const int64_t ld = P.ld;
    const int n = P.n;
    const int K1 = n < st.K0 + outer_w<T>() ? n : st.K0 + outer_w<T>();
    auto at = [&](int r, int c) { return A + E * ((int64_t)r * ld + c); };
    GemmArgs<T> g{A, ld, A, ld, A, ld, 0, 0, 0};
    if (st.K0 >= n) return g;
    if (st.mode == 5) {
        const int k = st.k, kend = n < k + NB ? n : k + NB, w = kend - k;
        if (k >= n) return g;
        return {at(kend, k), ld, at(kend, k), ld, (const T*)P.SU, (int64_t)NB, n - kend, w, w};
    }
    if (st.mode <= 1) {
        const int k = st.k, kend = n < k + NB ? n : k + NB, w = kend - k;
        if (k >= n) return g;
        if (st.mode == 0)
            return {at(k, kend), ld, (const T*)P.SL + E * (int64_t)st.b * NB * NB, (int64_t)NB, at(k, kend), ld, w,
                    K1 - kend, w};
        return {at(kend, kend), ld, at(kend, k), ld, at(k, kend), ld, n - kend, K1 - kend, w};
    }
    if (st.mode == 4) {
        const int c0 = st.c0 < 0 ? K1 : (st.c0 < n ? st.c0 : n);
        const int c1 = st.c1 <= 0 ? n : (st.c1 < n ? st.c1 : n);
        return {at(K1, c0), ld, at(K1, st.K0), ld, at(st.K0, c0), ld, n - K1, c1 - c0, K1 - st.K0};
    }
    const int r0 = st.K0 + st.b * NB;
    if (r0 >= K1) return g;
    const int wb = K1 - r0 < NB ? K1 - r0 : NB;
    if (st.mode == 2) return {at(r0, K1), ld, at(r0, st.K0), ld, at(st.K0, K1), ld, wb, n - K1, r0 - st.K0};
    return {at(r0, K1), ld, (const T*)P.SL + E * (int64_t)st.b * NB * NB, (int64_t)NB, at(r0, K1), ld, wb, n - K1,
            wb};
}

template <typename T, bool C, bool B>
__global__ __launch_bounds__(256) void gemm_sub_kernel(GemmArgs<T> inl, const LuProb* tbl, LuStep st) {
    GemmArgs<T> ga = inl;
    if constexpr (B) ga = lu_gemm_args<T, C>(tbl[blockIdx.z], st);
    T* Cm = ga.Cm;
    const T* Am = ga.Am;
    const T* Bm = ga.Bm;
    const int64_t ldc = ga.ldc, lda = ga.lda, ldb = ga.ldb;
    const int M = ga.M, N = ga.N, K = ga.K;
    if (M <= 0 || N <= 0 || (int)blockIdx.y * GT >= M || (int)blockIdx.x * GT >= N) return;
    using V4 = typename Mfma<T>::v4;
    constexpr int NBLK = WT / 16;
    // K chunks of 32 (FP32) or 16 (FP64): these products are short (K = 64 ..
    // 448 over one or a few tile rows) and latency-bound on operands fresh
    // from HBM, so fewer, larger chunks mean fewer dependent load rounds
    constexpr int KC = sizeof(T) == 4 ? 32 : 16;
    __shared__ T sAr[GT][KC + 1], sAi[C ? GT : 1][KC + 1];
    __shared__ T sBr[KC][GT + 1], sBi[C ? KC : 1][GT + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int bm = blockIdx.y * GT, bn = blockIdx.x * GT;
    const int wm = (wave >> 1) * WT, wn = (wave & 1) * WT;
    V4 accR[NBLK][NBLK], accI[C ? NBLK : 1][C ? NBLK : 1];
#pragma unroll
    for (int a = 0; a < NBLK; ++a)
#pragma unroll
        for (int b = 0; b < NBLK; ++b) {
            accR[a][b] = V4{0, 0, 0, 0};
            if constexpr (C) accI[a][b] = V4{0, 0, 0, 0};
        }
    // C's entries of this thread are loaded before the K loop, so their
    // latency is hidden behind it
    cx<T> cpre[NBLK][NBLK][4];
#pragma unroll
    for (int a = 0; a < NBLK; ++a)
#pragma unroll
        for (int b = 0; b < NBLK; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gr = bm + wm + 16 * a + Mfma<T>::row(lane, r);
                const int gc = bn + wn + 16 * b + (lane & 15);
                cpre[a][b][r] = (st.mode != 5 && gr < M && gc < N) ? Acc<T, C>::ld(Cm, (int64_t)gr * ldc + gc)
                                                                    : cx<T>{T(0), T(0)};
            }
    // register prefetch: the next K chunk is loaded while the current one is
    // multiplied out of LDS
    constexpr int PE = GT * KC / 256;
    cx<T> pa[PE], pb[PE];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < PE; ++q) {
            const int e = tid + 256 * q;
            const int ra = e / KC, ca = e % KC, rb = e / GT, cb = e % GT;
            const int gra = bm + ra, gca = k0 + ca, grb = k0 + rb, gcb = bn + cb;
            pa[q] = (gra < M && gca < K) ? Acc<T, C>::ld(Am, (int64_t)gra * lda + gca) : cx<T>{T(0), T(0)};
            pb[q] = (grb < K && gcb < N) ? Acc<T, C>::ld(Bm, (int64_t)grb * ldb + gcb) : cx<T>{T(0), T(0)};
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < K; k0 += KC) {
#pragma unroll
        for (int q = 0; q < PE; ++q) {
            const int e = tid + 256 * q;
            const int ra = e / KC, ca = e % KC, rb = e / GT, cb = e % GT;
            sAr[ra][ca] = pa[q].re;
            if constexpr (C) sAi[ra][ca] = pa[q].im;
            sBr[rb][cb] = pb[q].re;
            if constexpr (C) sBi[rb][cb] = pb[q].im;
        }
        __syncthreads();
        if (k0 + KC < K) fetch(k0 + KC);
#pragma unroll
        for (int kk = 0; kk < KC; kk += 4) {
            const int ka = kk + (lane >> 4);
            T ar[NBLK], ai[NBLK], br[NBLK], bi[NBLK];
#pragma unroll
            for (int a = 0; a < NBLK; ++a) {
                ar[a] = sAr[wm + 16 * a + (lane & 15)][ka];
                if constexpr (C) ai[a] = sAi[wm + 16 * a + (lane & 15)][ka];
            }
#pragma unroll
            for (int b = 0; b < NBLK; ++b) {
                br[b] = sBr[ka][wn + 16 * b + (lane & 15)];
                if constexpr (C) bi[b] = sBi[ka][wn + 16 * b + (lane & 15)];
            }
#pragma unroll
            for (int a = 0; a < NBLK; ++a)
#pragma unroll
                for (int b = 0; b < NBLK; ++b) {
                    accR[a][b] = Mfma<T>::op(ar[a], br[b], accR[a][b]);
                    if constexpr (C) {
                        accR[a][b] = Mfma<T>::op(-ai[a], bi[b], accR[a][b]);
                        accI[a][b] = Mfma<T>::op(ar[a], bi[b], accI[a][b]);
                        accI[a][b] = Mfma<T>::op(ai[a], br[b], accI[a][b]);
                    }
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < NBLK; ++a)
#pragma unroll
        for (int b = 0; b < NBLK; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gr = bm + wm + 16 * a + Mfma<T>::row(lane, r);
                const int gc = bn + wn + 16 * b + (lane & 15);
                if (gr < M && gc < N) {
                    cx<T> v = cpre[a][b][r];
                    v.re -= accR[a][b][r];
                    if constexpr (C) v.im -= accI[a][b][r];
                    Acc<T, C>::st(Cm, (int64_t)gr * ldc + gc, v);
                }
            }
}

// ---------------------------------------------------------------------------
// The outer U12 rows of one panel under block-local pivoting (FP32 factors),
// one launch per panel instead of an interchange, a product with the solved
// rows above and a product with inv(L_bb):
//   U_b = inv(L_bb) (P_b A[r0:r0+64, K1:n] - L[b, 0:b] U[0:b]),  r0 = K0 + 64 b,
// with P_b the panel's interchanges composed into one permutation (all of its
// pivots lie in its 64 rows) and inv(L_bb) = I - S_L,b.  Each workgroup owns
// all 64 rows of a 64-column tile: it reads the tile's rows through P_b
// before any store (so the tile is updated in place), forms the first
// product over K = 64 b on MFMA, stages the difference in LDS as the second
// product's operand (K = 64, S_L from the panel's slot) and stores the rows
// in order.  The columns left of the outer block are interchanged by a
// launch of their own.  Only full panels of blocks with columns right of
// them exist here (K1 < n).
// ---------------------------------------------------------------------------
template <bool C>
__global__ __launch_bounds__(256) void u12_panel_kernel(LuProb P, int K0, int b) {
    using T = float;
    using V4 = typename Mfma<T>::v4;
    constexpr int NBLK = WT / 16, KC = 32, PE = GT * KC / 256;
    const int n = P.n, K1 = min(n, K0 + outer_w<T>()), r0 = K0 + b * NB, N = n - K1;
    const int bn = blockIdx.x * GT;
    if (N <= 0 || bn >= N || r0 + NB > K1) return;  // uniform
    T* A = (T*)P.A;
    const int64_t ld = P.ld;
    __shared__ T sAr[GT][KC + 1], sAi[C ? GT : 1][KC + 1];
    // the first product's B chunks (rows 0 .. KC), then the tile
    __shared__ T sTr[GT][GT + 1], sTi[C ? GT : 1][GT + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * WT, wn = (wave & 1) * WT;
    // the composed permutation: lane i ends holding the row that lands at r0 + i
    const int pv = P.ipiv[r0 + lane] - r0;
    int src = lane;
    for (int c = 0; c < NB; ++c) {
        const int p = __builtin_amdgcn_readlane(pv, c);
        if (p != c && p >= 0 && p < NB) {  // uniform
            const int x = __builtin_amdgcn_readlane(src, c), y = __builtin_amdgcn_readlane(src, p);
            src = lane == c ? y : (lane == p ? x : src);
        }
    }
    // this thread's tile entries, read through the permutation
    cx<T> tv[NBLK][NBLK][4];
#pragma unroll
    for (int a = 0; a < NBLK; ++a)
#pragma unroll
        for (int bb = 0; bb < NBLK; ++bb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = wm + 16 * a + Mfma<T>::row(lane, r);
                const int sr = __shfl(src, row, 64);
                const int gc = bn + wn + 16 * bb + (lane & 15);
                tv[a][bb][r] = gc < N ? Acc<T, C>::ld(A, (int64_t)(r0 + sr) * ld + K1 + gc) : cx<T>{T(0), T(0)};
            }
    V4 accR[NBLK][NBLK], accI[C ? NBLK : 1][C ? NBLK : 1];
    auto zero_acc = [&]() {
#pragma unroll
        for (int a = 0; a < NBLK; ++a)
#pragma unroll
            for (int bb = 0; bb < NBLK; ++bb) {
                accR[a][bb] = V4{0, 0, 0, 0};
                if constexpr (C) accI[a][bb] = V4{0, 0, 0, 0};
            }
    };
    // accR/accI += (sA chunk) (B chunk in sT rows 0 .. KC, or the tile rows k0 .. k0 + KC)
    auto mma = [&](int brow0) {
#pragma unroll
        for (int kk = 0; kk < KC; kk += 4) {
            const int ka = kk + (lane >> 4);
            T ar[NBLK], ai[NBLK], br[NBLK], bi[NBLK];
#pragma unroll
            for (int a = 0; a < NBLK; ++a) {
                ar[a] = sAr[wm + 16 * a + (lane & 15)][ka];
                if constexpr (C) ai[a] = sAi[wm + 16 * a + (lane & 15)][ka];
            }
#pragma unroll
            for (int bb = 0; bb < NBLK; ++bb) {
                br[bb] = sTr[brow0 + ka][wn + 16 * bb + (lane & 15)];
                if constexpr (C) bi[bb] = sTi[brow0 + ka][wn + 16 * bb + (lane & 15)];
            }
#pragma unroll
            for (int a = 0; a < NBLK; ++a)
#pragma unroll
                for (int bb = 0; bb < NBLK; ++bb) {
                    accR[a][bb] = Mfma<T>::op(ar[a], br[bb], accR[a][bb]);
                    if constexpr (C) {
                        accR[a][bb] = Mfma<T>::op(-ai[a], bi[bb], accR[a][bb]);
                        accI[a][bb] = Mfma<T>::op(ar[a], bi[bb], accI[a][bb]);
                        accI[a][bb] = Mfma<T>::op(ai[a], br[bb], accI[a][bb]);
                    }
                }
        }
    };
    zero_acc();
    const int K = b * NB;  // the solved rows above
    if (K > 0) {
        cx<T> pa[PE], pb[PE];
        auto fetch = [&](int k0) {
#pragma unroll
            for (int q = 0; q < PE; ++q) {
                const int e = tid + 256 * q;
                const int ra = e / KC, ca = e % KC, rb = e / GT, cb = e % GT;
                pa[q] = Acc<T, C>::ld(A, (int64_t)(r0 + ra) * ld + K0 + k0 + ca);
                pb[q] = bn + cb < N ? Acc<T, C>::ld(A, (int64_t)(K0 + k0 + rb) * ld + K1 + bn + cb) : cx<T>{T(0), T(0)};
            }
        };
        fetch(0);
        for (int k0 = 0; k0 < K; k0 += KC) {
#pragma unroll
            for (int q = 0; q < PE; ++q) {
                const int e = tid + 256 * q;
                const int ra = e / KC, ca = e % KC, rb = e / GT, cb = e % GT;
                sAr[ra][ca] = pa[q].re;
                if constexpr (C) sAi[ra][ca] = pa[q].im;
                sTr[rb][cb] = pb[q].re;
                if constexpr (C) sTi[rb][cb] = pb[q].im;
            }
            __syncthreads();
            if (k0 + KC < K) fetch(k0 + KC);
            mma(0);
            __syncthreads();
        }
    }
    // the difference, kept in registers and staged in LDS as the tile
#pragma unroll
    for (int a = 0; a < NBLK; ++a)
#pragma unroll
        for (int bb = 0; bb < NBLK; ++bb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                tv[a][bb][r].re -= accR[a][bb][r];
                if constexpr (C) tv[a][bb][r].im -= accI[a][bb][r];
                const int row = wm + 16 * a + Mfma<T>::row(lane, r), col = wn + 16 * bb + (lane & 15);
                sTr[row][col] = tv[a][bb][r].re;
                if constexpr (C) sTi[row][col] = tv[a][bb][r].im;
            }
    // second product: S_L,b (tile), K = 64 in chunks of KC
    const T* SL = (const T*)P.SL + (C ? 2 : 1) * (int64_t)b * NB * NB;
    zero_acc();
    for (int k0 = 0; k0 < NB; k0 += KC) {
#pragma unroll
        for (int q = 0; q < PE; ++q) {
            const int e = tid + 256 * q;
            const int ra = e / KC, ca = e % KC;
            const cx<T> v = Acc<T, C>::ld(SL, (int64_t)ra * NB + k0 + ca);
            sAr[ra][ca] = v.re;
            if constexpr (C) sAi[ra][ca] = v.im;
        }
        __syncthreads();  // (the first pass: also the tile's stores)
        mma(k0);
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < NBLK; ++a)
#pragma unroll
        for (int bb = 0; bb < NBLK; ++bb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = wm + 16 * a + Mfma<T>::row(lane, r);
                const int gc = bn + wn + 16 * bb + (lane & 15);
                if (gc < N) {
                    cx<T> v = tv[a][bb][r];
                    v.re -= accR[a][bb][r];
                    if constexpr (C) v.im -= accI[a][bb][r];
                    Acc<T, C>::st(A, (int64_t)(r0 + row) * ld + K1 + gc, v);
                }
            }
}

// ---------------------------------------------------------------------------
// Trailing updates of large factorisations (modes 1 and 4): C -= A B with a
// 128 x BN block tile (BN = 128 or 64), eight waves of 64 x BN/4 (4 x BN/64
// MFMA 16x16 blocks each), so every A and B element staged in LDS feeds
// twice as many MFMAs as in the 64 x 64 kernel.  K runs in chunks of one
// 128-byte row segment (KB = 8 complex doubles, 16 complex floats or
// doubles, 32 floats) through two LDS buffers: the next chunk is loaded into
// registers before the MFMAs of the current one and stored to the other
// buffer after them, one barrier per chunk.  A is staged k-major (transposed)
// with one element of padding so both the transposing stores and the MFMA
// operand reads are conflict-free.  The complex products are issued as four
// sweeps over the accumulator grid (re*re, re*im, -im*im, im*re), so
// dependent MFMAs on one accumulator are a whole sweep apart.
// ---------------------------------------------------------------------------
constexpr int BT = 128;

template <typename T, bool C>
struct Vec2 {
    // one element of a row-major matrix (C: interleaved complex) as one load
    typedef T v2 __attribute__((ext_vector_type(2)));
    static __device__ __forceinline__ cx<T> ld(const T* a, int64_t idx) {
        if constexpr (C) {
            const v2 v = *(const v2*)(a + 2 * idx);
            return {v.x, v.y};
        } else {
            return {a[idx], T(0)};
        }
    }
};

template <typename T, bool C, int BN, bool B>
__global__ __launch_bounds__(512) void gemm_big_kernel(GemmArgs<T> inl, const LuProb* tbl, LuStep st) {
    GemmArgs<T> ga = inl;
    if constexpr (B) ga = lu_gemm_args<T, C>(tbl[blockIdx.z], st);
    T* Cm = ga.Cm;
    const T* Am = ga.Am;
    const T* Bm = ga.Bm;
    const int64_t ldc = ga.ldc, lda = ga.lda, ldb = ga.ldb;
    const int M = ga.M, N = ga.N, K = ga.K;
    if (M <= 0 || N <= 0 || K <= 0 || (int)blockIdx.y * BT >= M || (int)blockIdx.x * BN >= N) return;
    using V4 = typename Mfma<T>::v4;
    constexpr int E = C ? 2 : 1;
    constexpr int KB = 128 / (int)(sizeof(T) * E);
    constexpr int NT = 512;                      // eight waves, 2 x 4
    constexpr int MA = 4, NBB = BN / 64;         // MFMA blocks per wave (rows, columns)
    constexpr int PA = BT * KB / NT, PB = KB * BN / NT;
    __shared__ T sA[2][E][KB][BT + 1];
    __shared__ T sB[2][E][KB][BN];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int bm = blockIdx.y * BT, bn = blockIdx.x * BN;
    const int wm = (wave >> 2) * 64, wn = (wave & 3) * (BN / 4);
    V4 accR[MA][NBB], accI[C ? MA : 1][C ? NBB : 1];
#pragma unroll
    for (int a = 0; a < MA; ++a)
#pragma unroll
        for (int b = 0; b < NBB; ++b) {
            accR[a][b] = V4{0, 0, 0, 0};
            if constexpr (C) accI[a][b] = V4{0, 0, 0, 0};
        }
    cx<T> pa[PA], pb[PB];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < PA; ++q) {
            const int e = tid + NT * q, r = e / KB, c = e % KB;
            const int gr = bm + r, gc = k0 + c;
            pa[q] = (gr < M && gc < K) ? Vec2<T, C>::ld(Am, (int64_t)gr * lda + gc) : cx<T>{T(0), T(0)};
        }
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            const int e = tid + NT * q, r = e / BN, c = e % BN;
            const int gr = k0 + r, gc = bn + c;
            pb[q] = (gr < K && gc < N) ? Vec2<T, C>::ld(Bm, (int64_t)gr * ldb + gc) : cx<T>{T(0), T(0)};
        }
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int q = 0; q < PA; ++q) {
            const int e = tid + NT * q, r = e / KB, c = e % KB;
            sA[buf][0][c][r] = pa[q].re;
            if constexpr (C) sA[buf][1][c][r] = pa[q].im;
        }
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            const int e = tid + NT * q, r = e / BN, c = e % BN;
            sB[buf][0][r][c] = pb[q].re;
            if constexpr (C) sB[buf][1][r][c] = pb[q].im;
        }
    };
    fetch(0);
    stash(0);
    __syncthreads();
    const int nk = (K + KB - 1) / KB;
    for (int i = 0; i < nk; ++i) {
        const int cur = i & 1;
        if (i + 1 < nk) fetch((i + 1) * KB);
#pragma unroll
        for (int kk = 0; kk < KB; kk += 4) {
            const int ka = kk + (lane >> 4);
            T ar[MA], ai[MA], br[NBB], bi[NBB];
#pragma unroll
            for (int a = 0; a < MA; ++a) {
                ar[a] = sA[cur][0][ka][wm + 16 * a + (lane & 15)];
                if constexpr (C) ai[a] = sA[cur][1][ka][wm + 16 * a + (lane & 15)];
            }
#pragma unroll
            for (int b = 0; b < NBB; ++b) {
                br[b] = sB[cur][0][ka][wn + 16 * b + (lane & 15)];
                if constexpr (C) bi[b] = sB[cur][1][ka][wn + 16 * b + (lane & 15)];
            }
#pragma unroll
            for (int a = 0; a < MA; ++a)
#pragma unroll
                for (int b = 0; b < NBB; ++b) accR[a][b] = Mfma<T>::op(ar[a], br[b], accR[a][b]);
            if constexpr (C) {
#pragma unroll
                for (int a = 0; a < MA; ++a)
#pragma unroll
                    for (int b = 0; b < NBB; ++b) accI[a][b] = Mfma<T>::op(ar[a], bi[b], accI[a][b]);
#pragma unroll
                for (int a = 0; a < MA; ++a)
#pragma unroll
                    for (int b = 0; b < NBB; ++b) accR[a][b] = Mfma<T>::op(-ai[a], bi[b], accR[a][b]);
#pragma unroll
                for (int a = 0; a < MA; ++a)
#pragma unroll
                    for (int b = 0; b < NBB; ++b) accI[a][b] = Mfma<T>::op(ai[a], br[b], accI[a][b]);
            }
        }
        if (i + 1 < nk) stash(cur ^ 1);
        __syncthreads();
    }
    // read-modify-write of C, one row of MFMA blocks at a time (the compiler
    // memory barrier keeps it from hoisting every C load at once, which
    // would need as many registers as the accumulators)
#pragma unroll
    for (int a = 0; a < MA; ++a) {
#pragma unroll
        for (int b = 0; b < NBB; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gr = bm + wm + 16 * a + Mfma<T>::row(lane, r);
                const int gc = bn + wn + 16 * b + (lane & 15);
                if (gr < M && gc < N) {
                    cx<T> v = Vec2<T, C>::ld(Cm, (int64_t)gr * ldc + gc);
                    v.re -= accR[a][b][r];
                    if constexpr (C) v.im -= accI[a][b][r];
                    Acc<T, C>::st(Cm, (int64_t)gr * ldc + gc, v);
                }
            }
        asm volatile("" ::: "memory");
    }
}

// The same trailing update for complex operands with three real products
// per complex multiply-add (Gauss): P1 = Re a Re b, P2 = Im a Im b,
// P3 = (Re a + Im a)(Re b + Im b), Re = P1 - P2, Im = P3 - P1 - P2.  The
// operand sums are formed once when a chunk is staged (a third LDS plane per
// operand), so the MFMA count per chunk drops from 4 to 3 per complex
// product; the imaginary part carries the usual 3M rounding (normwise, a
// few ulps of |a||b|).
template <typename T, int BN, bool B>
__global__ __launch_bounds__(512) void gemm_big3m_kernel(GemmArgs<T> inl, const LuProb* tbl, LuStep st) {
    GemmArgs<T> ga = inl;
    if constexpr (B) ga = lu_gemm_args<T, true>(tbl[blockIdx.z], st);
    T* Cm = ga.Cm;
    const T* Am = ga.Am;
    const T* Bm = ga.Bm;
    const int64_t ldc = ga.ldc, lda = ga.lda, ldb = ga.ldb;
    const int M = ga.M, N = ga.N, K = ga.K;
    if (M <= 0 || N <= 0 || K <= 0 || (int)blockIdx.y * BT >= M || (int)blockIdx.x * BN >= N) return;
    using V4 = typename Mfma<T>::v4;
    constexpr int KB = 128 / (int)(sizeof(T) * 2);
    constexpr int NT = 512;
    constexpr int MA = 4, NBB = BN / 64;
    constexpr int PA = BT * KB / NT, PB = KB * BN / NT;
    __shared__ T sA[2][3][KB][BT + 1];
    __shared__ T sB[2][3][KB][BN];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int bm = blockIdx.y * BT, bn = blockIdx.x * BN;
    const int wm = (wave >> 2) * 64, wn = (wave & 3) * (BN / 4);
    V4 acc1[MA][NBB], acc2[MA][NBB], acc3[MA][NBB];
#pragma unroll
    for (int a = 0; a < MA; ++a)
#pragma unroll
        for (int b = 0; b < NBB; ++b) {
            acc1[a][b] = V4{0, 0, 0, 0};
            acc2[a][b] = V4{0, 0, 0, 0};
            acc3[a][b] = V4{0, 0, 0, 0};
        }
    cx<T> pa[PA], pb[PB];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < PA; ++q) {
            const int e = tid + NT * q, r = e / KB, c = e % KB;
            const int gr = bm + r, gc = k0 + c;
            pa[q] = (gr < M && gc < K) ? Vec2<T, true>::ld(Am, (int64_t)gr * lda + gc) : cx<T>{T(0), T(0)};
        }
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            const int e = tid + NT * q, r = e / BN, c = e % BN;
            const int gr = k0 + r, gc = bn + c;
            pb[q] = (gr < K && gc < N) ? Vec2<T, true>::ld(Bm, (int64_t)gr * ldb + gc) : cx<T>{T(0), T(0)};
        }
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int q = 0; q < PA; ++q) {
            const int e = tid + NT * q, r = e / KB, c = e % KB;
            sA[buf][0][c][r] = pa[q].re;
            sA[buf][1][c][r] = pa[q].im;
            sA[buf][2][c][r] = pa[q].re + pa[q].im;
        }
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            const int e = tid + NT * q, r = e / BN, c = e % BN;
            sB[buf][0][r][c] = pb[q].re;
            sB[buf][1][r][c] = pb[q].im;
            sB[buf][2][r][c] = pb[q].re + pb[q].im;
        }
    };
    fetch(0);
    stash(0);
    __syncthreads();
    const int nk = (K + KB - 1) / KB;
    for (int i = 0; i < nk; ++i) {
        const int cur = i & 1;
        if (i + 1 < nk) fetch((i + 1) * KB);
#pragma unroll
        for (int kk = 0; kk < KB; kk += 4) {
            const int ka = kk + (lane >> 4);
            T a0[MA], a1[MA], a2[MA], b0[NBB], b1[NBB], b2[NBB];
#pragma unroll
            for (int a = 0; a < MA; ++a) {
                a0[a] = sA[cur][0][ka][wm + 16 * a + (lane & 15)];
                a1[a] = sA[cur][1][ka][wm + 16 * a + (lane & 15)];
                a2[a] = sA[cur][2][ka][wm + 16 * a + (lane & 15)];
            }
#pragma unroll
            for (int b = 0; b < NBB; ++b) {
                b0[b] = sB[cur][0][ka][wn + 16 * b + (lane & 15)];
                b1[b] = sB[cur][1][ka][wn + 16 * b + (lane & 15)];
                b2[b] = sB[cur][2][ka][wn + 16 * b + (lane & 15)];
            }
#pragma unroll
            for (int a = 0; a < MA; ++a)
#pragma unroll
                for (int b = 0; b < NBB; ++b) acc1[a][b] = Mfma<T>::op(a0[a], b0[b], acc1[a][b]);
#pragma unroll
            for (int a = 0; a < MA; ++a)
#pragma unroll
                for (int b = 0; b < NBB; ++b) acc2[a][b] = Mfma<T>::op(a1[a], b1[b], acc2[a][b]);
#pragma unroll
            for (int a = 0; a < MA; ++a)
#pragma unroll
                for (int b = 0; b < NBB; ++b) acc3[a][b] = Mfma<T>::op(a2[a], b2[b], acc3[a][b]);
        }
        if (i + 1 < nk) stash(cur ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < MA; ++a) {
#pragma unroll
        for (int b = 0; b < NBB; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gr = bm + wm + 16 * a + Mfma<T>::row(lane, r);
                const int gc = bn + wn + 16 * b + (lane & 15);
                if (gr < M && gc < N) {
                    cx<T> v = Vec2<T, true>::ld(Cm, (int64_t)gr * ldc + gc);
                    const T p1 = acc1[a][b][r], p2 = acc2[a][b][r], p3 = acc3[a][b][r];
                    v.re -= p1 - p2;
                    v.im -= p3 - p1 - p2;
                    Acc<T, true>::st(Cm, (int64_t)gr * ldc + gc, v);
                }
            }
        asm volatile("" ::: "memory");
    }
}

// ---------------------------------------------------------------------------
// Triangular solves with one right-hand side, blocked by TB rows.
// ---------------------------------------------------------------------------
constexpr int TB = 64;

// One triangular solve of a (possibly batched) sweep: the factors, the
// complex128 work vector and the sweep's synchronisation words, by value or
// from a table indexed by blockIdx.z.  sync holds, for the lower and then
// the upper sweep, a ticket counter and one ready flag per TB-row block
// (2 (nblk + 1) words, zeroed before the forward sweep).
struct SolveProb {
    const void* LU;
    int64_t ld;
    int n;
    double* w;
    unsigned* sync;
    long long wait;  // bound of every wait on another block's flag (wall-clock ticks, 100 MHz)
    int* fail;       // raised when a wait gives up (the context's CTR_TRI_FAIL word)
};
template <bool B>
__device__ __forceinline__ SolveProb get_sp(const SolveProb& inl, const SolveProb* tbl) {
    if constexpr (B) return tbl[blockIdx.z];
    else return inl;
}

__host__ __device__ inline int tri_blocks(int n) { return (n + TB - 1) / TB; }

__global__ void tri_sync_reset_kernel(SolveProb inl, const SolveProb* tbl, int batched) {
    const SolveProb P = batched ? tbl[blockIdx.z] : inl;
    const int words = 2 * (tri_blocks(P.n) + 1);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += gridDim.x * blockDim.x) P.sync[i] = 0u;
}

// Chained blocked triangular sweep, one workgroup per TB-row block and one
// launch per sweep.  Workgroups take tickets in dispatch order and solve the
// row blocks in dependency order (forward: 0, 1, ...; backward: the last
// first), so a workgroup only ever waits for workgroups that are already
// running.  Each one
//   - stages its diagonal block in LDS and reads its right-hand side,
//   - subtracts M(b, j) x_j for every solved block j as soon as its ready
//     flag is up (x_j read with coherent loads),
//   - solves with the diagonal block (one wave, column by column),
//   - stores x_b and raises its flag (release).
// M is the lower (LOWER) or upper factor: without TRANS L (unit) and U, read
// by rows; with TRANS U^T (lower, non-unit) and L^T (upper, unit), read from
// the same storage by columns.  Arithmetic in FP64 whatever T is.  A wait
// that does not end within the wait limit (two seconds by default; never
// expected: every predecessor is resident) leaves NaN in the block instead of
// hanging, and raises the fail word the host reads after the solve.
template <typename T, bool C, bool LOWER, bool TRANS, bool B>
__global__ __launch_bounds__(256) void trsv_chain_kernel(SolveProb inl, const SolveProb* tbl) {
    const SolveProb P = get_sp<B>(inl, tbl);
    const T* A = (const T*)P.LU;
    const int64_t ld = P.ld;
    const int n = P.n;
    const int nblk = tri_blocks(n);
    double* x = P.w;
    unsigned* ticket = P.sync + (LOWER ? 0 : nblk + 1);
    unsigned* ready = ticket + 1;
    constexpr bool UNIT = LOWER != TRANS;
    __shared__ double dr[TB][TB + 1], di[C ? TB : 1][TB + 1];
    __shared__ double xr[TB], xi[TB];
    __shared__ double pr[4][TB], pi[4][TB];
    __shared__ unsigned s_tk;
    __shared__ int s_fail;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if ((int)blockIdx.x >= nblk) return;  // grid sized for the batch's largest n
    if (t == 0) {
        s_tk = atomicAdd(ticket, 1u);
        s_fail = 0;
    }
    __syncthreads();
    if (s_tk >= (unsigned)nblk) return;  // sync words not reset: nothing to own
    const int b = LOWER ? (int)s_tk : nblk - 1 - (int)s_tk;
    const int r0 = b * TB, nb = min(TB, n - r0);
    // element (i, j) of the operator M, rows/columns relative to r0/c0
    auto mel = [&](int i, int c0, int j) -> cx<T> {
        return TRANS ? Acc<T, C>::ld(A, (int64_t)(c0 + j) * ld + r0 + i) : Acc<T, C>::ld(A, (int64_t)(r0 + i) * ld + c0 + j);
    };
    for (int e = t; e < TB * TB; e += 256) {
        const int i = e / TB, j = e % TB;
        cx<T> v = {T(0), T(0)};
        if (i < nb && j < nb) v = mel(i, r0, j);
        dr[i][j] = (double)v.re;
        if constexpr (C) di[i][j] = (double)v.im;
    }
    __syncthreads();  // the diagonal block is complete before any wave reads it
    // right-hand side: only this workgroup touches rows [r0, r0 + nb)
    double acc_r = 0.0, acc_i = 0.0;
    if (t < nb) { acc_r = x[2 * (r0 + t)]; acc_i = x[2 * (r0 + t) + 1]; }
    const int jfirst = LOWER ? 0 : nblk - 1, jstep = LOWER ? 1 : -1;
    // partial products: without TRANS thread (row i = t / 4, quarter q) sums
    // columns q, q + 4, ...; with TRANS wave q sums columns q, q + 4, ... for
    // row i = lane (rows of the factor read contiguously)
    const int pi_row = TRANS ? lane : t >> 2, pq = TRANS ? wave : t & 3;
    for (int j = jfirst; j != b; j += jstep) {
        const int c0 = j * TB, nc = min(TB, n - c0);
        // this thread's entries of M(b, j): the factor does not depend on x_j,
        // so the loads are issued before the wait for block j
        cx<T> mv[TB / 4];
#pragma unroll
        for (int u = 0; u < TB / 4; ++u) {
            const int c = pq + 4 * u;
            mv[u] = (pi_row < nb && c < nc) ? mel(pi_row, c0, c) : cx<T>{T(0), T(0)};
        }
        if (t == 0) {
            const long long t0 = wall_clock64();
            while (__hip_atomic_load(&ready[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
                __builtin_amdgcn_s_sleep(2);
                if (wall_clock64() - t0 > P.wait) { s_fail = 1; break; }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
        if (s_fail) break;
        if (t < nc) { xr[t] = ld_agent(&x[2 * (c0 + t)]); xi[t] = ld_agent(&x[2 * (c0 + t) + 1]); }
        __syncthreads();
        double sr = 0.0, si = 0.0;
        const int i = pi_row, q = pq;
#pragma unroll
        for (int u = 0; u < TB / 4; ++u) {
            const int c = q + 4 * u;
            if (c < nc) {  // entries outside the block are zero
                sr += (double)mv[u].re * xr[c] - (double)mv[u].im * xi[c];
                si += (double)mv[u].re * xi[c] + (double)mv[u].im * xr[c];
            }
        }
        if (!TRANS) {
            sr += __shfl_xor(sr, 1, 64); si += __shfl_xor(si, 1, 64);
            sr += __shfl_xor(sr, 2, 64); si += __shfl_xor(si, 2, 64);
            if (q == 0) { pr[0][i] = sr; pi[0][i] = si; }
        } else {
            pr[q][i] = sr; pi[q][i] = si;
        }
        __syncthreads();
        if (t < nb) {
            if (TRANS) {
                acc_r -= pr[0][t] + pr[1][t] + pr[2][t] + pr[3][t];
                acc_i -= pi[0][t] + pi[1][t] + pi[2][t] + pi[3][t];
            } else {
                acc_r -= pr[0][t];
                acc_i -= pi[0][t];
            }
        }
        __syncthreads();  // xr/pr are rewritten for the next block
    }
    if (s_fail) {
        if (t < nb) { x[2 * (r0 + t)] = __builtin_nan(""); x[2 * (r0 + t) + 1] = __builtin_nan(""); }
        if (t == 0) __hip_atomic_store(P.fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;  // the flag stays down: workgroups after this one time out too
    }
    if (wave == 0) {
        // column-oriented solve with the diagonal block: lane i owns row i.
        // Lane i's reciprocal diagonal entry is formed before the chain, and
        // each solved value reaches the other lanes through scalar registers.
        cx<double> rd = {1.0, 0.0};
        if (!UNIT && lane < nb) rd = cdiv(cx<double>{1.0, 0.0}, cx<double>{dr[lane][lane], C ? di[lane][lane] : 0.0});
        // branch-free steps (selects instead of lane branches), so that the
        // LDS reads of the diagonal block do not wait for the solved values
        auto step = [&](int c) {
            if (!UNIT) {
                const cx<double> q = cmul(cx<double>{acc_r, acc_i}, rd);
                acc_r = lane == c ? q.re : acc_r;
                acc_i = lane == c ? q.im : acc_i;
            }
            const double vr = readlane_f64(acc_r, c), vi = readlane_f64(acc_i, c);
            const bool act = lane < nb && (LOWER ? lane > c : lane < c);
            const double mr0 = dr[lane][c], mi0 = C ? di[lane][c] : 0.0;
            const double mr = act ? mr0 : 0.0, mi = act ? mi0 : 0.0;
            acc_r -= mr * vr - mi * vi;
            acc_i -= mr * vi + mi * vr;
        };
        if (nb == TB) {
            // a full block: unrolled, so that the diagonal block's LDS reads
            // (which do not depend on the solved values) are issued ahead
#pragma unroll
            for (int s = 0; s < TB; ++s) step(LOWER ? s : TB - 1 - s);
        } else {
            for (int s = 0; s < nb; ++s) step(LOWER ? s : nb - 1 - s);
        }
        if (lane < nb) { x[2 * (r0 + lane)] = acc_r; x[2 * (r0 + lane) + 1] = acc_i; }
        __threadfence();
        if (lane == 0) __hip_atomic_store(&ready[b], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// r = b - A x in FP64 (A of T/C, x/b/r as complex128 work vectors): one wave
// per row.
template <typename T, bool C>
__global__ __launch_bounds__(256) void residual_kernel(const T* A, int64_t ld, int n, const double* x,
                                                       const double* b, double* r) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    double sr = 0.0, si = 0.0;
    for (int c = lane; c < n; c += 64) {
        cx<T> a = Acc<T, C>::ld(A, (int64_t)row * ld + c);
        double xr = x[2 * c], xi = x[2 * c + 1];
        sr += (double)a.re * xr - (double)a.im * xi;
        si += (double)a.re * xi + (double)a.im * xr;
    }
    sr = wave_sum(sr); si = wave_sum(si);
    if (lane == 0) { r[2 * row] = b[2 * row] - sr; r[2 * row + 1] = b[2 * row + 1] - si; }
}

// gather y[i] = x[perm[i]] (complex128 work vectors)
__global__ void permute_kernel(const double* x, const int32_t* perm, int n, double* y) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[2 * i] = x[2 * perm[i]];
    y[2 * i + 1] = x[2 * perm[i] + 1];
}
// scatter y[perm[i]] = x[i] (the inverse permutation)
__global__ void unpermute_kernel(const double* x, const int32_t* perm, int n, double* y) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[2 * perm[i]] = x[2 * i];
    y[2 * perm[i] + 1] = x[2 * i + 1];
}

// complex128 work vector <-> caller vector of dtype
__global__ void vec_to_work(const void* v, int dtype, int n, double* w) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double re, im;
    switch (dtype) {
        case BEM_F64: re = ((const double*)v)[i]; im = 0.0; break;
        case BEM_C128: re = ((const double*)v)[2 * i]; im = ((const double*)v)[2 * i + 1]; break;
        case BEM_F32: re = ((const float*)v)[i]; im = 0.0; break;
        default: re = ((const float*)v)[2 * i]; im = ((const float*)v)[2 * i + 1]; break;
    }
    w[2 * i] = re; w[2 * i + 1] = im;
}
__global__ void work_to_vec(const double* w, int n, void* v, int dtype, int accumulate) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double re = w[2 * i], im = w[2 * i + 1];
    switch (dtype) {
        case BEM_F64: { double* o = (double*)v + i; *o = accumulate ? *o + re : re; } break;
        case BEM_C128: {
            double* o = (double*)v + 2 * i;
            if (accumulate) { o[0] += re; o[1] += im; } else { o[0] = re; o[1] = im; }
        } break;
        case BEM_F32: { float* o = (float*)v + i; *o = (float)(accumulate ? *o + re : re); } break;
        default: {
            float* o = (float*)v + 2 * i;
            if (accumulate) { o[0] += (float)re; o[1] += (float)im; }
            else { o[0] = (float)re; o[1] = (float)im; }
        } break;
    }
}

// block sums of |w_i|^2 (complex128 work vector)
__global__ void norm2_kernel(const double* w, int n, double* out) {
    __shared__ double red[4];
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        s += w[2 * i] * w[2 * i] + w[2 * i + 1] * w[2 * i + 1];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, red[0] + red[1] + red[2] + red[3]);
}

template <typename TS, bool CS, typename TD, bool CD>
__global__ void convert_kernel(const TS* src, int64_t lds, TD* dst, int64_t ldd, int rows, int cols) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)rows * cols) return;
    int r = (int)(e / cols), c = (int)(e % cols);
    cx<TS> v = Acc<TS, CS>::ld(src, (int64_t)r * lds + c);
    Acc<TD, CD>::st(dst, (int64_t)r * ldd + c, cx<TD>{(TD)v.re, (TD)v.im});
}

// dst += alpha * op(src), op = identity or transpose; transposed reads go
// through a 32 x 32 LDS tile so both sides stay coalesced.
template <typename T, bool C>
__global__ __launch_bounds__(256) void axpy_kernel(const T* src, int64_t lds, T* dst, int64_t ldd, int rows,
                                                   int cols, double ar, double ai, int transpose) {
    __shared__ cx<T> tile[32][33];
    const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    if (transpose) {
        // dst[r][c] += alpha src[c][r]; load src rows c0.., cols r0..
        for (int yy = ty; yy < 32; yy += 8) {
            int sr = c0 + yy, sc = r0 + tx;
            if (sr < cols && sc < rows) tile[yy][tx] = Acc<T, C>::ld(src, (int64_t)sr * lds + sc);
        }
        __syncthreads();
        for (int yy = ty; yy < 32; yy += 8) {
            int r = r0 + yy, c = c0 + tx;
            if (r < rows && c < cols) {
                cx<T> v = tile[tx][yy];
                cx<T> d = Acc<T, C>::ld(dst, (int64_t)r * ldd + c);
                d.re += (T)(ar * v.re - ai * v.im);
                d.im += (T)(ar * v.im + ai * v.re);
                Acc<T, C>::st(dst, (int64_t)r * ldd + c, d);
            }
        }
    } else {
        for (int yy = ty; yy < 32; yy += 8) {
            int r = r0 + yy, c = c0 + tx;
            if (r < rows && c < cols) {
                cx<T> v = Acc<T, C>::ld(src, (int64_t)r * lds + c);
                cx<T> d = Acc<T, C>::ld(dst, (int64_t)r * ldd + c);
                d.re += (T)(ar * v.re - ai * v.im);
                d.im += (T)(ar * v.im + ai * v.re);
                Acc<T, C>::st(dst, (int64_t)r * ldd + c, d);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// c64 rank-512 trailing updates on the fp16 MFMA (v_mfma_f32_32x32x16_f16;
// measured on gfx950: 1.54 PFLOP/s sustained against 143 TFLOP/s for the
// FP32 MFMA).  The operands are split once per outer block into fp16 planes:
//  * each row of the left factor A (each column of the right factor B) is
//    scaled by the power of two 2^e that puts its largest |re| or |im| into
//    [2^14, 2^15), well inside fp16's range;
//  * a scaled value v becomes hi = fp16(v) and lo = fp16(v - hi), so
//    |v - hi - lo| <= 2^-22 |v| where both are normal, and below 2^-24 (an
//    fp16 subnormal step, 2^-38 of the row's largest value) where not;
//  * a real product keeps hi*hi + hi*lo + lo*hi (lo*lo is below 2^-22 |a b|)
//    in one FP32 accumulator, and a complex product takes four real ones
//    (re re - im im, re im + im re; the minus sign is a sign-bit flip of B's
//    imaginary parts), so one complex multiply-add is twelve fp16 products
//    against four FP32 ones; measured error of the update about 3e-7 of
//    sum |a||b|;
//  * C -= 2^-(e_i + f_j) (re, im), the exponents applied with ldexp.
// Plane layout: plane p (0 re hi, 1 re lo, 2 im hi, 3 im lo), stage s (16
// consecutive k), row r at element ((p HNS + s) cap + r) 16 + k % 16, so
// that one stage of consecutive rows is contiguous; stages past K are
// zero-padded up to a multiple of 16.
// ---------------------------------------------------------------------------
typedef _Float16 h16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 h16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
constexpr int HS = 16;          // K depth of one stage (the MFMA's)
constexpr int HNS = outer_w<float>() / HS;  // stages of a full outer block
constexpr int HT = 128;         // C tile (rows and columns)

// Split planes of one factorisation (device): the outer block's left-factor
// rows and right-factor columns, with their exponents; cap = n.
struct H16Ws {
    _Float16* a;
    int* ea;
    _Float16* b;
    int* eb;
    int cap;
};

__host__ __device__ inline int64_t hoff(int p, int s, int r, int cap) {
    return ((int64_t)(p * HNS + s) * cap + r) * HS;
}

size_t h16_ws_bytes(int n) { return 2 * (4 * (size_t)n * outer_w<float>() * sizeof(_Float16) + (size_t)n * sizeof(int) + 512); }

H16Ws h16_carve(char* base, int n) {
    H16Ws w;
    w.cap = n;
    const size_t planes = 4 * (size_t)n * outer_w<float>() * sizeof(_Float16);
    w.a = (_Float16*)base;
    w.ea = (int*)(base + planes);
    char* nb = base + ((planes + (size_t)n * sizeof(int) + 255) & ~(size_t)255) + 256;
    w.b = (_Float16*)nb;
    w.eb = (int*)(nb + planes);
    return w;
}

// e with m 2^e in [2^14, 2^15); 0 for m = 0 or not finite
__device__ __forceinline__ int h16_exponent(float m) {
    if (!(m > 0.f) || !(m < INFINITY)) return 0;
    int ex;
    frexpf(m, &ex);  // m in [2^(ex - 1), 2^ex)
    return max(-120, min(120, 15 - ex));
}

__device__ __forceinline__ void h16_split2(float x, int e, _Float16& h, _Float16& l) {
    const float v = ldexpf(x, e);
    h = (_Float16)v;
    l = (_Float16)(v - (float)h);
}

// Rows of the left factor (M x K, K <= 512, of a row-major c64 matrix) into
// the planes, one wave per row (lane l: k = 8 l .. 8 l + 7).
__global__ __launch_bounds__(256) void h16_split_rows_kernel(const float* A, int64_t lda, int M, int K, H16Ws w) {
    constexpr int KPL = outer_w<float>() / 64;  // consecutive k per lane
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= M) return;
    const float* a = A + 2 * (int64_t)row * lda;
    const int ns = (K + HS - 1) / HS;
    float re[KPL], im[KPL], m = 0.f;
#pragma unroll
    for (int i = 0; i < KPL; ++i) {
        const int k = KPL * lane + i;
        re[i] = 0.f;
        im[i] = 0.f;
        if (k < K) {
            const float2 v = *(const float2*)(a + 2 * k);
            re[i] = v.x;
            im[i] = v.y;
        }
        m = fmaxf(m, fmaxf(fabsf(re[i]), fabsf(im[i])));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    const int e = h16_exponent(m);
    if (lane == 0) w.ea[row] = e;
    const int s = (KPL * lane) / HS;
    if (s >= ns) return;
#pragma unroll
    for (int i = 0; i < KPL; ++i) {
        _Float16 h, l;
        const int k = KPL * lane + i;
        h16_split2(re[i], e, h, l);
        w.a[hoff(0, s, row, w.cap) + k % HS] = h;
        w.a[hoff(1, s, row, w.cap) + k % HS] = l;
        h16_split2(im[i], e, h, l);
        w.a[hoff(2, s, row, w.cap) + k % HS] = h;
        w.a[hoff(3, s, row, w.cap) + k % HS] = l;
    }
}

// Columns of the right factor (K x N, K <= 512, row-major c64) into the
// planes, 64 columns per workgroup: the column maxima, then one stage of a
// column per thread.
__global__ __launch_bounds__(256) void h16_split_cols_kernel(const float* Bm, int64_t ldb, int K, int N, H16Ws w) {
    __shared__ float sm[4][64];
    __shared__ int se[64];
    const int t = threadIdx.x, c = t & 63, g = t >> 6;
    const int col = blockIdx.x * 64 + c;
    const int ns = (K + HS - 1) / HS;
    float m = 0.f;
    if (col < N)
        for (int k = g; k < K; k += 4) {
            const float2 v = *(const float2*)(Bm + 2 * ((int64_t)k * ldb + col));
            m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
        }
    sm[g][c] = m;
    __syncthreads();
    if (t < 64) {
        const int e = h16_exponent(fmaxf(fmaxf(sm[0][t], sm[1][t]), fmaxf(sm[2][t], sm[3][t])));
        se[t] = e;
        if (blockIdx.x * 64 + t < N) w.eb[blockIdx.x * 64 + t] = e;
    }
    __syncthreads();
    if (col >= N) return;
    const int e = se[c];
    for (int s = g; s < ns; s += 4) {
        h16x8_t v[4][2];
#pragma unroll
        for (int i = 0; i < HS; ++i) {
            const int k = HS * s + i;
            float2 x = {0.f, 0.f};
            if (k < K) x = *(const float2*)(Bm + 2 * ((int64_t)k * ldb + col));
            _Float16 h, l;
            h16_split2(x.x, e, h, l);
            v[0][i >> 3][i & 7] = h;
            v[1][i >> 3][i & 7] = l;
            h16_split2(x.y, e, h, l);
            v[2][i >> 3][i & 7] = h;
            v[3][i >> 3][i & 7] = l;
        }
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            h16x8_t* o = (h16x8_t*)(w.b + hoff(p, s, col, w.cap));
            o[0] = v[p][0];
            o[1] = v[p][1];
        }
    }
}

__device__ __forceinline__ h16x8_t h16_neg(h16x8_t x) {
    typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));
    u32x4_t u = __builtin_bit_cast(u32x4_t, x);
    u ^= 0x80008000u;
    return __builtin_bit_cast(h16x8_t, u);
}

// C[M x N] -= A B from the planes (A rows from row 0, B columns from column
// boff).  128 x 128 tile per workgroup of four waves (2 x 2, each 64 x 64 =
// 2 x 2 MFMA blocks of 32 x 32), two workgroups per compute unit; stages of
// 16 k double-buffered in LDS (2 x 32 KB), the next stage's loads in flight
// in registers during the current one.  LDS rows hold 16 k (32 bytes); the
// two 16-byte halves of a row are swapped when bit 2 xor bit 3 of the row
// is set, so 16 consecutive rows read at one k offset hit distinct banks.
// Tiles are dealt so that each XCD (workgroup id mod 8) takes a contiguous
// run of the grouped order (4 tile rows, column-major inside a group): the
// tiles one XCD runs at a time share their A and B planes in its L2.
__global__ __launch_bounds__(256, 2) void gemm_h16_kernel(float* Cm, int64_t ldc, int M, int N, int K, H16Ws w,
                                                          int boff, int tiles_m, int tiles_n) {
    constexpr int G = 4;
    const int T = tiles_m * tiles_n, per = (T + 7) / 8;
    const int t = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (t >= T) return;
    const int grp = t / (G * tiles_n), rem = t - grp * G * tiles_n;
    const int rows_in = min(G, tiles_m - grp * G);
    const int bm = (grp * G + rem % rows_in) * HT, bn = (rem / rows_in) * HT;
    __shared__ _Float16 sm[2][2][4][HT][HS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    f32x16_t accR[2][2], accI[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                accR[a][b][j] = 0.f;
                accI[a][b][j] = 0.f;
            }
    // this thread's load slots: q = 0..3 plane q of A, q = 4..7 plane q - 4
    // of B, tile row r, half c (rows past M or N are clamped: they reach
    // only C rows or columns that are not written)
    h16x8_t pf[8];
    const int r = tid >> 1, c = tid & 1;
    const unsigned offA = (unsigned)min(bm + r, M - 1) * HS + 8 * c;
    const unsigned offB = (unsigned)(boff + min(bn + r, N - 1)) * HS + 8 * c;
    const int64_t sst = (int64_t)w.cap * HS, pst = HNS * sst;
    auto load = [&](int s) {
        const _Float16* a0 = w.a + s * sst;
        const _Float16* b0 = w.b + s * sst;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            pf[q] = *(const h16x8_t*)(a0 + q * pst + offA);
            pf[4 + q] = *(const h16x8_t*)(b0 + q * pst + offB);
        }
    };
    const int cs = 8 * (c ^ (((r >> 2) ^ (r >> 3)) & 1));
    auto store = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 8; ++q) *(h16x8_t*)&sm[buf][q >> 2][q & 3][r][cs] = pf[q];
    };
    const int ns = (K + HS - 1) / HS;
    load(0);
    store(0);
    __syncthreads();
    // MFMA 32x32x16 operands (measured on gfx950): lane l supplies
    // A[l % 32][8 (l / 32) + i] and B[8 (l / 32) + i][l % 32], i = 0..7
    const int l32 = lane & 31, kc = 8 * ((lane >> 5) ^ (((l32 >> 2) ^ (l32 >> 3)) & 1));
    for (int s = 0; s < ns; ++s) {
        const int buf = s & 1;
        if (s + 1 < ns) load(s + 1);
        h16x8_t af[2][4];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int p = 0; p < 4; ++p) af[a][p] = *(const h16x8_t*)&sm[buf][0][p][wm + 32 * a + l32][kc];
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            h16x8_t bf[4];
#pragma unroll
            for (int p = 0; p < 4; ++p) bf[p] = *(const h16x8_t*)&sm[buf][1][p][wn + 32 * b + l32][kc];
            const h16x8_t bn2 = h16_neg(bf[2]), bn3 = h16_neg(bf[3]);
            // (hi, lo) x (hi, lo) without lo lo; four accumulators in turn
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                accR[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(af[a][0], bf[0], accR[a][b], 0, 0, 0);
                accI[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(af[a][0], bf[2], accI[a][b], 0, 0, 0);
            }
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                accR[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(af[a][0], bf[1], accR[a][b], 0, 0, 0);
                accI[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(af[a][0], bf[3], accI[a][b], 0, 0, 0);
            }
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                accR[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(af[a][1], bf[0], accR[a][b], 0, 0, 0);
                accI[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(af[a][1], bf[2], accI[a][b], 0, 0, 0);
            }
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                accR[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(af[a][2], bn2, accR[a][b], 0, 0, 0);
                accI[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(af[a][2], bf[0], accI[a][b], 0, 0, 0);
            }
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                accR[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(af[a][2], bn3, accR[a][b], 0, 0, 0);
                accI[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(af[a][2], bf[1], accI[a][b], 0, 0, 0);
            }
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                accR[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(af[a][3], bn2, accR[a][b], 0, 0, 0);
                accI[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(af[a][3], bf[0], accI[a][b], 0, 0, 0);
            }
        }
        if (s + 1 < ns) store(buf ^ 1);
        __syncthreads();
    }
    // result register j of lane l: row 8 (j / 4) + 4 (l / 32) + j % 4,
    // column l % 32 of its 32 x 32 block (measured on gfx950)
    int ecol[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        const int gc = bn + wn + 32 * b + l32;
        ecol[b] = gc < N ? w.eb[boff + gc] : 0;
    }
#pragma unroll
    for (int a = 0; a < 2; ++a) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int gr = bm + wm + 32 * a + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
            if (gr >= M) continue;
            const int erow = w.ea[gr];
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int gc = bn + wn + 32 * b + l32;
                if (gc < N) {
                    float2* cp = (float2*)(Cm + 2 * ((int64_t)gr * ldc + gc));
                    float2 v = *cp;
                    const int e = -(erow + ecol[b]);
                    v.x -= ldexpf(accR[a][b][j], e);
                    v.y -= ldexpf(accI[a][b][j], e);
                    *cp = v;
                }
            }
        }
        asm volatile("" ::: "memory");
    }
}

// The outer block's left factor A[K1:n, K0:K1] and right factor
// A[K0:K1, K1:n] into the planes (once per outer block, after its U12).
void h16_split(const LuProb& P, int K0, const H16Ws& w, hipStream_t s) {
    const int n = P.n, K1 = std::min(n, K0 + outer_w<float>()), M = n - K1;
    if (M <= 0) return;
    float* A = (float*)P.A;
    hipLaunchKernelGGL(h16_split_rows_kernel, dim3((M + 3) / 4), dim3(256), 0, s,
                       (const float*)(A + 2 * ((int64_t)K1 * P.ld + K0)), P.ld, M, K1 - K0, w);
    hipLaunchKernelGGL(h16_split_cols_kernel, dim3((M + 63) / 64), dim3(256), 0, s,
                       (const float*)(A + 2 * ((int64_t)K0 * P.ld + K1)), P.ld, K1 - K0, M, w);
}

// A[K1:n, c0:c1] -= A[K1:n, K0:K1] A[K0:K1, c0:c1] from the planes of h16_split.
void h16_update(const LuProb& P, int K0, int c0, int c1, const H16Ws& w, hipStream_t s) {
    const int n = P.n, K1 = std::min(n, K0 + outer_w<float>()), M = n - K1, N = c1 - c0;
    if (M <= 0 || N <= 0) return;
    float* C = (float*)P.A + 2 * ((int64_t)K1 * P.ld + c0);
    const int tm = (M + HT - 1) / HT, tn = (N + HT - 1) / HT;
    hipLaunchKernelGGL(gemm_h16_kernel, dim3(8 * ((tm * tn + 7) / 8)), dim3(256), 0, s, C, P.ld, M, N, K1 - K0, w,
                       c0 - K1, tm, tn);
}

// gopts, the GEMM options of a factorisation: bit 4 (GEMM_SMALL_TILES) = 64 x
// 64 tiles for every update; bit 5 (GEMM_WIDE4M) = complex trailing updates on
// 128 x 128 tiles with four real products per complex product.  (The
// split-fp16 rank-512 updates are chosen by the caller of lu_launch.)
constexpr int GEMM_SMALL_TILES = 16;
constexpr int GEMM_WIDE4M = 32;

int gemm_opts(const Context* ctx) {
    return (ctx->lu_gemm_tile ? 0 : GEMM_SMALL_TILES) |
           (ctx->lu_gemm_tile == 2 ? GEMM_WIDE4M : 0);
}

template <typename T, bool C, bool B>
void lu_gemm(const LuProb& one, const LuProb* d_tbl, const LuStep& st, int M, int N, int active, int gopts,
             hipStream_t s) {
    if (M <= 0 || N <= 0) return;
    GemmArgs<T> g{};
    if constexpr (!B) g = lu_gemm_args<T, C>(one, st);
    const dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT, active);
    // trailing updates with room for 128-row tiles take the large-tile kernel
    // (64-wide column tiles when the update is narrower than 128 or would
    // leave a half-empty last 128-column tile)
    if ((st.mode == 1 || st.mode == 4) && M >= 2 * BT && !(gopts & GEMM_SMALL_TILES)) {
        // tile widths measured on a 20,224 x 20,224 x 256 update: complex128
        // 128 x 64 (two workgroups per CU; 128 x 128 holds one), complex64
        // 128 x 64 with three products per complex product, real 128 x 128
        const int rows = (M + BT - 1) / BT;
        const bool wide = (N >= BT && N % BT == 0) || N > 2 * BT;
        const dim3 g64((N + 63) / 64, rows, active), g128((N + BT - 1) / BT, rows, active);
        if constexpr (C) {
            if (gopts & GEMM_WIDE4M) {
                if (wide) hipLaunchKernelGGL((gemm_big_kernel<T, C, BT, B>), g128, dim3(512), 0, s, g, d_tbl, st);
                else hipLaunchKernelGGL((gemm_big_kernel<T, C, 64, B>), g64, dim3(512), 0, s, g, d_tbl, st);
            } else if constexpr (std::is_same<T, float>::value) {
                hipLaunchKernelGGL((gemm_big3m_kernel<T, 64, B>), g64, dim3(512), 0, s, g, d_tbl, st);
            } else {
                hipLaunchKernelGGL((gemm_big_kernel<T, C, 64, B>), g64, dim3(512), 0, s, g, d_tbl, st);
            }
        } else {
            if (wide) hipLaunchKernelGGL((gemm_big_kernel<T, C, BT, B>), g128, dim3(512), 0, s, g, d_tbl, st);
            else hipLaunchKernelGGL((gemm_big_kernel<T, C, 64, B>), g64, dim3(512), 0, s, g, d_tbl, st);
        }
        return;
    }
    hipLaunchKernelGGL((gemm_sub_kernel<T, C, B>), grid, dim3(256), 0, s, g, d_tbl, st);
}

// Factorisation of the NB-column panel at k (rows k..n of every active
// problem): partial pivoting with one fused update/search launch per column
// (pivoting 0, LAPACK's pivot rows), tournament pivoting (pivoting 1), or
// block-local pivoting (pivoting 2: rows interchanged only inside the
// panel's top NB x NB block; for factors that precondition an iterative
// refinement against the original matrix).
// pgrid > 0 (single problems): partial pivoting runs as one persistent
// launch of at most pgrid workgroups (one per compute unit of the stream)
// when the panel's rows fit their LDS.
template <typename T, bool C, bool B>
void panel_factor(const LuProb& one, const LuProb* d_tbl, const std::vector<int>& ns, int k, int active,
                  int pivoting, int gopts, int pgrid, unsigned* pbase, hipStream_t s) {
    auto max_over = [&](auto f) {
        int m = 0;
        for (int q = 0; q < active; ++q) m = std::max(m, f(ns[q]));
        return m;
    };
    if (pivoting == 0 && !B && pgrid > 0) {
        const int nrem = ns[0] - k;
        const int G = std::max(1, std::min({pgrid, PGMAX, (nrem + 63) / 64}));
        const int R = (nrem + G - 1) / G;
        if (R <= panel_rmax<T>()) {
            hipLaunchKernelGGL((panel_persist_kernel<T, C>), dim3(G), dim3(PNT), 0, s, one, k, G, R, *pbase);
            *pbase += (unsigned)(panel_barriers(std::min(NB, nrem)) * G);
            return;
        }
    }
    if (pivoting == 2) {
        // block-local pivoting: the top block in one workgroup, L21 on MFMA
        hipLaunchKernelGGL((topblock_local_kernel<T, C, B>), dim3(1, 1, active), dim3(512), 0, s, one, d_tbl, k);
        const int M = max_over([&](int n) { return n - std::min(n, k + NB); });
        lu_gemm<T, C, B>(one, d_tbl, LuStep{5, k - k % outer_w<T>(), k, 0}, M, NB, active, gopts, s);
        return;
    }
    if (pivoting == 0) {
        const int kendmax = std::min(ns[0], k + NB);
        for (int c = k; c <= kendmax; ++c) {
            const int G = max_over([&](int n) {
                const int kend = std::min(n, k + NB);
                if (c > kend) return 0;
                const int rows = n - (c < kend ? c : kend);
                return rows > 0 ? (rows + RPB - 1) / RPB : 0;
            });
            if (G <= 0) break;
            hipLaunchKernelGGL((col_step_kernel<T, C, B>), dim3(G, 1, active), dim3(256), 0, s, one, d_tbl, k, c);
        }
        return;
    }
    const int mmax = max_over([&](int n) { return n - k; });
    int G = (mmax + TRW - 1) / TRW, level = 0;
    hipLaunchKernelGGL((tourn_kernel<T, C, B>), dim3(G, 1, active), dim3(256), 0, s, one, d_tbl, k, 0);
    while (G > 1) {
        G = (G + 3) / 4;
        ++level;
        hipLaunchKernelGGL((tourn_kernel<T, C, B>), dim3(G, 1, active), dim3(256), 0, s, one, d_tbl, k, level);
    }
    hipLaunchKernelGGL((tourn_ipiv_kernel<B>), dim3(1, 1, active), dim3(64), 0, s, one, d_tbl, k, level);
    hipLaunchKernelGGL((laswp_kernel<T, C, B>), dim3(1, 1, active), dim3(256), 0, s, one, d_tbl, k, k + NB, k,
                       k + NB, 0, 0);
    hipLaunchKernelGGL((topblock_kernel<T, C, B>), dim3(1, 1, active), dim3(256), 0, s, one, d_tbl, k);
    const int M = max_over([&](int n) { return n - std::min(n, k + NB); });
    lu_gemm<T, C, B>(one, d_tbl, LuStep{5, k - k % outer_w<T>(), k, 0}, M, NB, active, gopts, s);
}

// Launch sequence of the two-level right-looking LU for one problem (B =
// false, the problem by value) or a batch (B = true, table on the device,
// problems sorted by decreasing n so that the first `active` entries are
// the ones with columns left at a given step).
// hw (single c64 problems, BEM_OPT_C64_LU_GEMM = 2): the rank-512 outer updates run on
// the split-fp16 planes.
template <typename T, bool C, bool B>
void lu_launch(const LuProb& one, const LuProb* d_tbl, const std::vector<int>& ns, int pivoting, int gopts,
               int pgrid, hipStream_t s, int K_begin = 0, int K_end = 1 << 30, const H16Ws* hw = nullptr) {
    const int nmax = ns[0];
    unsigned pbase = 0;
    auto count_active = [&](int from) {
        int a = 0;
        while (a < (int)ns.size() && ns[a] > from) ++a;
        return a;
    };
    // outer blocks [K_begin, min(nmax, K_end)), K_begin a multiple of outer_w<T>()
    for (int K0 = K_begin; K0 < std::min(nmax, K_end); K0 += outer_w<T>()) {
        const int K1max = std::min(nmax, K0 + outer_w<T>());
        const int act_outer = count_active(K0);
        for (int k = K0; k < K1max; k += NB) {
            const int active = count_active(k);
            auto max_over = [&](auto f) {
                int m = 0;
                for (int q = 0; q < active; ++q) m = std::max(m, f(ns[q]));
                return m;
            };
            panel_factor<T, C, B>(one, d_tbl, ns, k, active, pivoting, gopts, pgrid, &pbase, s);
            // this panel's interchanges inside the outer block
            hipLaunchKernelGGL((laswp_kernel<T, C, B>), dim3((outer_w<T>() + 255) / 256, 1, active), dim3(256), 0, s, one,
                               d_tbl, k, k + NB, K0, K0 + outer_w<T>(), k, k + NB);
            const int slot = (k - K0) / NB;
            if (pivoting != 2)  // block-local pivoting: the panel top wrote S_L
                hipLaunchKernelGGL((linv_kernel<T, C, B>), dim3(1, 1, active), dim3(256), 0, s, one, d_tbl, k, slot);
            const int Nin = max_over([&](int n) { return std::min(n, K0 + outer_w<T>()) - std::min(n, k + NB); });
            const int Min = max_over([&](int n) { return n - std::min(n, k + NB); });
            lu_gemm<T, C, B>(one, d_tbl, LuStep{0, K0, k, slot}, NB, Nin, active, gopts, s);
            lu_gemm<T, C, B>(one, d_tbl, LuStep{1, K0, k, slot}, Min, Nin, active, gopts, s);
        }
        // the outer block's interchanges on the columns left and right of it
        hipLaunchKernelGGL((laswp_kernel<T, C, B>), dim3((nmax + 255) / 256, 1, act_outer), dim3(256), 0, s, one,
                           d_tbl, K0, K0 + outer_w<T>(), 0, nmax, K0, K0 + outer_w<T>());
        int Nout = 0;
        for (int q = 0; q < act_outer; ++q) Nout = std::max(Nout, ns[q] - std::min(ns[q], K0 + outer_w<T>()));
        if (Nout > 0) {
            const int nb_blocks = (K1max - K0 + NB - 1) / NB;
            for (int b = 0; b < nb_blocks; ++b) {
                if (b > 0) lu_gemm<T, C, B>(one, d_tbl, LuStep{2, K0, K0, b}, NB, Nout, act_outer, gopts, s);
                lu_gemm<T, C, B>(one, d_tbl, LuStep{3, K0, K0, b}, NB, Nout, act_outer, gopts, s);
            }
            if (!B && hw) {
                h16_split(one, K0, *hw, s);
                h16_update(one, K0, std::min(nmax, K0 + outer_w<T>()), nmax, *hw, s);
            } else {
                lu_gemm<T, C, B>(one, d_tbl, LuStep{4, K0, K0, 0}, Nout, Nout, act_outer, gopts, s);
            }
        }
    }
}

// One problem with look-ahead over two internal streams: the inner panels of
// outer block j+1 (stream sp) run while the trailing update of block j
// beyond block j+1's columns (stream sg) is still in flight.  The streams
// may be restricted to disjoint compute units (context option), so the
// latency-bound panel kernels do not wait behind GEMM blocks.
//   sp: panels of block j  -> event ep
//   sg: wait ep; interchanges outside block j; U12; update of block j+1's
//       columns -> event eg; update of the remaining columns
//   sp: wait eg; panels of block j+1 ...
// Interchanges of pivots [p_lo, p_hi) over columns [c_lo, c_hi) outside
// [x_lo, x_hi) of one problem: the composed permutation for block-local
// pivoting of FP32 factors, the swap sequence otherwise.
template <typename T, bool C>
void laswp_launch(const LuProb& P, int pivoting, int p_lo, int p_hi, int c_lo, int c_hi, int x_lo, int x_hi,
                  hipStream_t s) {
    const dim3 grid((std::max(0, c_hi - c_lo) + 255) / 256);
    if (grid.x == 0) return;
    if constexpr (std::is_same<T, float>::value) {
        if (pivoting == 2) {
            hipLaunchKernelGGL((laswp_local_kernel<C, false>), grid, dim3(256), 0, s, P, nullptr, p_lo, p_hi, c_lo,
                               c_hi, x_lo, x_hi);
            return;
        }
    }
    hipLaunchKernelGGL((laswp_kernel<T, C, false>), grid, dim3(256), 0, s, P, nullptr, p_lo, p_hi, c_lo, c_hi, x_lo,
                       x_hi);
}

template <typename T, bool C>
void lu_launch_lookahead(const LuProb& P, int pivoting, int gopts, int pgrid, hipStream_t s, hipStream_t sp,
                         hipStream_t sg, hipEvent_t ep, hipEvent_t eg, hipEvent_t e0, const hipEvent_t* pev,
                         const H16Ws* hw = nullptr) {
    const int n = P.n;
    const std::vector<int> ns1{n};
    unsigned pbase = 0;
    hipEventRecord(e0, s);
    hipStreamWaitEvent(sp, e0, 0);
    hipStreamWaitEvent(sg, e0, 0);
    for (int K0 = 0; K0 < n; K0 += outer_w<T>()) {
        const int K1 = std::min(n, K0 + outer_w<T>());
        const int Nout = n - K1;
        for (int k = K0; k < K1; k += NB) {
            const int kend = std::min(n, k + NB);
            panel_factor<T, C, false>(P, nullptr, ns1, k, 1, pivoting, gopts, pgrid, &pbase, sp);
            // inside the block (a few hundred columns, most rows in place for
            // these matrices) the plain swap sequence is the shorter launch
            hipLaunchKernelGGL((laswp_kernel<T, C, false>), dim3((outer_w<T>() + 255) / 256), dim3(256), 0, sp, P, nullptr, k,
                               k + NB, K0, K0 + outer_w<T>(), k, k + NB);
            const int slot = (k - K0) / NB;
            if (pivoting != 2)  // block-local pivoting: the panel top wrote S_L
                hipLaunchKernelGGL((linv_kernel<T, C, false>), dim3(1), dim3(256), 0, sp, P, nullptr, k, slot);
            // The panel's rows of the outer U12 (columns K1..n) need only this
            // panel (its pivots, S_L and its rows of L) and the U12 rows above
            // it: the update stream takes them as each panel completes, beside
            // the rest of the block's panels (a later panel's interchanges never
            // reach the rows of an earlier one).
            hipEventRecord(pev[slot], sp);
            hipStreamWaitEvent(sg, pev[slot], 0);
            if (std::is_same<T, float>::value && pivoting == 2) {
                // FP32 block-local: the panel's U12 rows in one launch (reading
                // its rows of the right columns through its interchanges), then
                // the interchanges of the columns left of the block
                if (Nout > 0)
                    hipLaunchKernelGGL((u12_panel_kernel<C>), dim3((Nout + GT - 1) / GT), dim3(256), 0, sg, P, K0, slot);
                laswp_launch<T, C>(P, pivoting, k, k + NB, 0, K0, 0, 0, sg);
            } else {
                laswp_launch<T, C>(P, pivoting, k, k + NB, 0, n, K0, K0 + outer_w<T>(), sg);
                if (Nout > 0) {
                    if (slot > 0) lu_gemm<T, C, false>(P, nullptr, LuStep{2, K0, K0, slot}, NB, Nout, 1, gopts, sg);
                    lu_gemm<T, C, false>(P, nullptr, LuStep{3, K0, K0, slot}, NB, Nout, 1, gopts, sg);
                }
            }
            lu_gemm<T, C, false>(P, nullptr, LuStep{0, K0, k, slot}, NB, K1 - kend, 1, gopts, sp);
            lu_gemm<T, C, false>(P, nullptr, LuStep{1, K0, k, slot}, n - kend, K1 - kend, 1, gopts, sp);
        }
        if (Nout > 0) {
            const int c_next = std::min(n, K1 + outer_w<T>());
            LuStep st4{4, K0, K0, 0};
            st4.c0 = K1;
            st4.c1 = c_next;
            if (hw) {
                h16_split(P, K0, *hw, sg);
                h16_update(P, K0, K1, c_next, *hw, sg);
            } else {
                lu_gemm<T, C, false>(P, nullptr, st4, Nout, c_next - K1, 1, gopts, sg);
            }
            hipEventRecord(eg, sg);
            hipStreamWaitEvent(sp, eg, 0);
            if (c_next < n) {
                st4.c0 = c_next;
                st4.c1 = n;
                if (hw) h16_update(P, K0, c_next, n, *hw, sg);
                else lu_gemm<T, C, false>(P, nullptr, st4, Nout, n - c_next, 1, gopts, sg);
            }
        }
    }
    hipEventRecord(ep, sp);
    hipEventRecord(eg, sg);
    hipStreamWaitEvent(s, ep, 0);
    hipStreamWaitEvent(s, eg, 0);
}

// Per-problem scratch: counter and info words, the per-block pivot
// partials and S_L of each NB block of an outer block (outer_w / NB of them).
template <typename T, bool C>
size_t lu_scratch_bytes(int n) {
    const int gmax = (n + RPB - 1) / RPB;
    const size_t tids = 2 * (size_t)((n + TRW - 1) / TRW) * NB * sizeof(int);
    return 512 + (((size_t)gmax * (sizeof(double) + sizeof(int)) + 255) & ~(size_t)255) +
           (outer_w<T>() / NB) * NB * NB * sizeof(cx<T>) + 256 + ((tids + 255) & ~(size_t)255) + NB * NB * sizeof(cx<T>) + 256 +
           sizeof(PanelX<T>) + 256;
}

template <typename T, bool C>
LuProb lu_carve(char* base, void* A, int64_t ld, int n, int32_t* ipiv, int* info, long long wait) {
    const int gmax = (n + RPB - 1) / RPB;
    LuProb P;
    P.A = A; P.ld = ld; P.n = n; P.ipiv = ipiv; P.wait = wait;
    P.counter = (unsigned*)base;
    P.info = info ? info : (int*)(base + 64);
    P.pval = (double*)(base + 512);
    P.pidx = (int*)(P.pval + gmax);
    P.SL = (void*)(((uintptr_t)(P.pidx + gmax) + 255) & ~(uintptr_t)255);
    P.tids = (int*)(((uintptr_t)((cx<T>*)P.SL + (outer_w<T>() / NB) * NB * NB) + 255) & ~(uintptr_t)255);
    P.SU = (void*)(((uintptr_t)(P.tids + 2 * ((n + TRW - 1) / TRW) * NB) + 255) & ~(uintptr_t)255);
    P.pan = (void*)(((uintptr_t)((cx<T>*)P.SU + NB * NB) + 255) & ~(uintptr_t)255);
    return P;
}

int ensure_lu_scratch(Context* ctx, size_t need) {
    if (ctx->lu_scratch_bytes >= need) return BEM_OK;
    if (ctx->lu_scratch) hipFree(ctx->lu_scratch);
    ctx->lu_scratch = nullptr;
    ctx->lu_scratch_bytes = 0;
    ctx->lu_table.clear();
    int rc = check_hip(hipMalloc(&ctx->lu_scratch, need), "LU scratch");
    if (rc) return rc;
    ctx->lu_scratch_bytes = need;
    return BEM_OK;
}

// The two look-ahead streams (default priority).  With a positive
// lu_lookahead_cus the panel stream gets the first that many compute units
// (bits of the CU mask) and the update stream the others; -1 leaves both
// unrestricted.
// Only one look-ahead pair exists at a time: HIP maps a process's streams
// onto a few hardware queues (GPU_MAX_HW_QUEUES, 4 by default), so with
// both pairs and the null stream alive the pair created second shares
// queues, and its panel and update launches no longer overlap (measured on
// L5: the c64 block-local LU 120 -> 133 ms after a c128 LU had created the
// masked pair, or the c128 LU 480 -> 516 ms the other way round).  Switching
// between partial and block-local pivoting therefore drops the other pair
// (its work is finished first).
void drop_lu_pair(hipStream_t (&st)[2]) {
    for (hipStream_t& q : st)
        if (q) {
            (void)hipStreamSynchronize(q);
            (void)hipStreamDestroy(q);
            q = nullptr;
        }
}

int ensure_lu_streams(Context* ctx) {
    if (ctx->lu_stream[0] && ctx->lu_streams_cus == ctx->lu_lookahead_cus) return BEM_OK;
    drop_lu_pair(ctx->lu_stream_free);
    drop_lu_pair(ctx->lu_stream);
    int rc;
    if (ctx->lu_lookahead_cus > 0) {
        if (ctx->num_cu <= 0) {
            hipDeviceProp_t prop;
            rc = check_hip(hipGetDeviceProperties(&prop, ctx->device), "device properties");
            if (rc) return rc;
            ctx->num_cu = prop.multiProcessorCount;
        }
        const int ncu = ctx->num_cu, words = (ncu + 31) / 32;
        const int np = std::min(ctx->lu_lookahead_cus, ncu - 1);
        std::vector<uint32_t> mp(words, 0u), mg(words, 0u);
        for (int c = 0; c < ncu; ++c) (c < np ? mp : mg)[c / 32] |= 1u << (c % 32);
        rc = check_hip(hipExtStreamCreateWithCUMask(&ctx->lu_stream[0], words, mp.data()), "panel stream");
        if (rc) return rc;
        rc = check_hip(hipExtStreamCreateWithCUMask(&ctx->lu_stream[1], words, mg.data()), "update stream");
        if (rc) return rc;
    } else {
        for (int q = 0; q < 2; ++q) {
            rc = check_hip(hipStreamCreateWithFlags(&ctx->lu_stream[q], hipStreamNonBlocking), "LU stream");
            if (rc) return rc;
        }
    }
    for (int q = 0; q < 3; ++q)
        if (!ctx->lu_event[q]) {
            rc = check_hip(hipEventCreateWithFlags(&ctx->lu_event[q], hipEventDisableTiming), "LU event");
            if (rc) return rc;
        }
    for (hipEvent_t& e : ctx->lu_pevent)
        if (!e) {
            rc = check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "LU event");
            if (rc) return rc;
        }
    ctx->lu_streams_cus = ctx->lu_lookahead_cus;
    return BEM_OK;
}

// The unmasked look-ahead streams of block-local pivoting (default priority):
// its panels are two short launches, so reserving compute units for them
// only takes them from the trailing update (measured on the L5 c64
// factorisation: 216 ms unmasked, 226 ms with 32 masked units).
int ensure_lu_streams_free(Context* ctx) {
    drop_lu_pair(ctx->lu_stream);
    for (int q = 0; q < 2; ++q)
        if (!ctx->lu_stream_free[q]) {
            int rc = check_hip(hipStreamCreateWithFlags(&ctx->lu_stream_free[q], hipStreamNonBlocking), "LU stream");
            if (rc) return rc;
        }
    for (int q = 0; q < 3; ++q)
        if (!ctx->lu_event[q]) {
            int rc = check_hip(hipEventCreateWithFlags(&ctx->lu_event[q], hipEventDisableTiming), "LU event");
            if (rc) return rc;
        }
    for (hipEvent_t& e : ctx->lu_pevent)
        if (!e) {
            int rc = check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "LU event");
            if (rc) return rc;
        }
    return BEM_OK;
}

// counter = 0 and info = 0 of every problem (a kernel, not a memset, so the
// sequence can be captured in a graph)
// The counters of a factorisation (the persistent panel's barrier words sit
// at the start of its exchange area, the same layout for every T).
__device__ void lu_reset_words(const LuProb& P) {
    *P.counter = 0u;
    *P.info = 0;
    unsigned* w = (unsigned*)P.pan;
    w[0] = 0u;  // barrier arrivals
    w[2] = 0u;  // fail
}
__global__ void lu_reset_kernel(const LuProb* tbl, int count) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    lu_reset_words(tbl[q]);
}
__global__ void lu_reset_one_kernel(LuProb P) { lu_reset_words(P); }
// ipiv = identity before a factorisation, so the interchanges of a panel that
// gave up (a wait ran out) are no-ops instead of whatever the array held
template <bool B>
__global__ void ipiv_identity_kernel(LuProb inl, const LuProb* tbl) {
    const LuProb P = get_prob<B>(inl, tbl);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P.n) P.ipiv[i] = i;
}

// Largest grid of the persistent panel whose workgroups are all resident at
// once on `units` otherwise idle compute units (its waits need every one of
// them running).  Two facts fix it:
//  * per unit, the number of workgroups the kernel's registers and LDS allow
//    (hipOccupancyMaxActiveBlocksPerMultiprocessor; 0 means the kernel cannot
//    run at all, and the panels then take the per-column launches);
//  * the dispatcher deals workgroups round-robin over the XCDs the stream may
//    use (num_cu / 8 units each; a CU mask takes the first units, which fill
//    whole XCDs in order), so each XCD must hold ceil(G / XCDs) of them, and
//    the XCD with the fewest units bounds the grid.
// Example: a 48-unit mask spans 32 + 16 units, so at one workgroup per unit
// the grid is at most 2 x 16 = 32 (a grid of 48 puts 24 on the 16-unit XCD:
// measured never to become resident, the bounded waits reported it).
template <typename T, bool C>
int panel_grid_resident(const Context* ctx, int units) {
    static int per_unit = -1;  // one query per kernel instance and process
    if (per_unit < 0) {
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(panel_persist_kernel<T, C>),
                                                         PNT, 0) != hipSuccess)
            occ = 0;
        per_unit = occ;
    }
    if (per_unit <= 0 || units <= 0) return 0;
    const int per_xcd = std::max(1, ctx->num_cu / 8);
    const int full = units / per_xcd, rem = units % per_xcd;
    const int xcds = full + (rem ? 1 : 0);
    return xcds * (rem ? rem : per_xcd) * per_unit;
}

template <typename T, bool C>
int lu_factor_t(Context* ctx, T* A, int64_t ld, int n, int32_t* ipiv, hipStream_t s) {
    int rc = ensure_lu_scratch(ctx, lu_scratch_bytes<T, C>(n));
    if (rc) return rc;
    ctx->lu_table.clear();  // this call reuses the scratch a batch table may occupy
    const LuProb P = lu_carve<T, C>((char*)ctx->lu_scratch, A, ld, n, ipiv, nullptr, ctx->wait_ticks);
    hipLaunchKernelGGL(lu_reset_one_kernel, dim3(1), dim3(1), 0, s, P);
    hipLaunchKernelGGL(ipiv_identity_kernel<false>, dim3((n + 255) / 256), dim3(256), 0, s, P, (const LuProb*)nullptr);
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cap) != hipSuccess) cap = hipStreamCaptureStatusNone;
    if (ctx->num_cu <= 0) {
        hipDeviceProp_t prop;
        rc = check_hip(hipGetDeviceProperties(&prop, ctx->device), "device properties");
        if (rc) return rc;
        ctx->num_cu = prop.multiProcessorCount;
    }
    // workgroups of the persistent panel: as many as stay resident on the
    // compute units its stream may use
    int pgrid = ctx->lu_panel ? panel_grid_resident<T, C>(ctx, ctx->num_cu) : 0;
    // split-fp16 rank-512 updates (c64, an option)
    H16Ws hws{};
    const H16Ws* hw = nullptr;
    if (std::is_same<T, float>::value && C && ctx->c64_gemm_h16 && n > outer_w<T>()) {
        const size_t need = h16_ws_bytes(n);
        if (ctx->lu_h16_bytes < need) {
            if (ctx->lu_h16) hipFree(ctx->lu_h16);
            ctx->lu_h16 = nullptr;
            ctx->lu_h16_bytes = 0;
            rc = check_hip(hipMalloc(&ctx->lu_h16, need), "LU fp16 planes");
            if (rc) return rc;
            ctx->lu_h16_bytes = need;
        }
        hws = h16_carve((char*)ctx->lu_h16, n);
        hw = &hws;
    }
    if (ctx->lu_lookahead_cus != 0 && cap == hipStreamCaptureStatusNone && n > outer_w<T>() && ctx->lu_pivoting == 2) {
        rc = ensure_lu_streams_free(ctx);
        if (rc) return rc;
        lu_launch_lookahead<T, C>(P, ctx->lu_pivoting, gemm_opts(ctx), 0, s, ctx->lu_stream_free[0],
                                  ctx->lu_stream_free[1], ctx->lu_event[0], ctx->lu_event[1], ctx->lu_event[2],
                                  ctx->lu_pevent, hw);
    } else if (ctx->lu_lookahead_cus != 0 && cap == hipStreamCaptureStatusNone && n > outer_w<T>()) {
        rc = ensure_lu_streams(ctx);
        if (rc) return rc;
        if (pgrid > 0 && ctx->lu_lookahead_cus > 0)
            pgrid = panel_grid_resident<T, C>(ctx, std::min(ctx->lu_lookahead_cus, ctx->num_cu - 1));
        if (ctx->lu_panel > 1) pgrid = std::min(pgrid, ctx->lu_panel);
        lu_launch_lookahead<T, C>(P, ctx->lu_pivoting, gemm_opts(ctx), pgrid, s, ctx->lu_stream[0],
                                  ctx->lu_stream[1], ctx->lu_event[0], ctx->lu_event[1], ctx->lu_event[2],
                                  ctx->lu_pevent, hw);
    } else {
        if (ctx->lu_panel > 1) pgrid = std::min(pgrid, ctx->lu_panel);
        lu_launch<T, C, false>(P, nullptr, std::vector<int>{n}, ctx->lu_pivoting, gemm_opts(ctx), pgrid, s, 0,
                               1 << 30, hw);
    }
    rc = check_hip(hipGetLastError(), "LU launch");
    if (rc) return rc;
    int h_info = 0;
    rc = check_hip(hipMemcpyAsync(&h_info, P.info, sizeof(int), hipMemcpyDeviceToHost, s), "LU info");
    if (rc) return rc;
    rc = check_hip(hipStreamSynchronize(s), "LU sync");
    if (rc) return rc;
    if (h_info < 0) return set_error(BEM_ERR_HIP, "LU panel: a workgroup barrier timed out");
    if (h_info) return set_error(BEM_ERR_SINGULAR, "exactly singular matrix: zero pivot in column " + std::to_string(h_info - 1));
    return BEM_OK;
}

// Batched factorisation: one launch sequence for all problems.  The problem
// table is uploaded only when it differs from the previous call's, so a
// second call with the same arguments issues kernels only (capturable in a
// graph).  info[q] (device) gets 0 or 1 + the first zero-pivot column.
template <typename T, bool C>
int lu_factor_batched_t(Context* ctx, int count, void* const* A, const int64_t* ld, const int32_t* n,
                        int32_t* const* ipiv, int32_t* info, hipStream_t s) {
    std::vector<int> order(count);
    for (int q = 0; q < count; ++q) order[q] = q;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return n[a] > n[b]; });
    size_t off = 0;
    std::vector<size_t> offs(count);
    for (int r = 0; r < count; ++r) {
        offs[r] = off;
        off += (lu_scratch_bytes<T, C>(n[order[r]]) + 255) & ~(size_t)255;
    }
    const size_t tbl_bytes = (sizeof(LuProb) * count + 255) & ~(size_t)255;
    int rc = ensure_lu_scratch(ctx, off + tbl_bytes);
    if (rc) return rc;
    char* base = (char*)ctx->lu_scratch;
    std::vector<LuProb> tbl(count);
    std::vector<int> ns(count);
    for (int r = 0; r < count; ++r) {
        const int q = order[r];
        tbl[r] = lu_carve<T, C>(base + offs[r], A[q], ld[q], n[q], ipiv[q], info + q, ctx->wait_ticks);
        ns[r] = n[q];
    }
    LuProb* d_tbl = (LuProb*)(base + off);
    const size_t tb = sizeof(LuProb) * count;
    if (ctx->lu_table.size() != tb || memcmp(ctx->lu_table.data(), tbl.data(), tb) != 0 ||
        ctx->lu_table_dev != (void*)d_tbl) {
        rc = check_hip(hipMemcpyAsync(d_tbl, tbl.data(), tb, hipMemcpyHostToDevice, s), "LU table");
        if (rc) return rc;
        rc = check_hip(hipStreamSynchronize(s), "LU table sync");
        if (rc) return rc;
        ctx->lu_table.assign((const char*)tbl.data(), (const char*)tbl.data() + tb);
        ctx->lu_table_dev = (void*)d_tbl;
    }
    hipLaunchKernelGGL(lu_reset_kernel, dim3((count + 63) / 64), dim3(64), 0, s, (const LuProb*)d_tbl, count);
    hipLaunchKernelGGL(ipiv_identity_kernel<true>, dim3((ns[0] + 255) / 256, 1, count), dim3(256), 0, s, LuProb{},
                       (const LuProb*)d_tbl);
    // Once only the largest problem is left (from the first outer block past
    // the second-largest n), its remaining columns take the single-problem
    // path: persistent partial-pivoting panels instead of the batch's panel
    // launches, which for one problem are a dozen short launches per panel.
    // Any pivot rule per panel gives a valid factorisation; partial pivoting
    // is the stronger one.  (With persistent panels switched off the batch
    // runs its own steps to the end.)
    int K_switch = 1 << 30;
    if (ctx->lu_panel != 0) {
        const int second = count > 1 ? ns[1] : 0;
        const int ks = (second + outer_w<T>() - 1) / outer_w<T>() * outer_w<T>();
        if (ns[0] > ks) K_switch = ks;
    }
    if (K_switch > 0) lu_launch<T, C, true>(LuProb{}, d_tbl, ns, ctx->lu_pivoting, gemm_opts(ctx), 0, s, 0, K_switch);
    if (K_switch < ns[0]) {
        if (ctx->num_cu <= 0) {
            hipDeviceProp_t prop;
            rc = check_hip(hipGetDeviceProperties(&prop, ctx->device), "device properties");
            if (rc) return rc;
            ctx->num_cu = prop.multiProcessorCount;
        }
        int pgrid = panel_grid_resident<T, C>(ctx, ctx->num_cu);
        if (ctx->lu_panel > 1) pgrid = std::min(pgrid, ctx->lu_panel);
        // block-local pivoting keeps its own rule for the tail
        const int tail_piv = ctx->lu_pivoting == 2 ? 2 : 0;
        lu_launch<T, C, false>(tbl[0], nullptr, std::vector<int>{ns[0]}, tail_piv, gemm_opts(ctx), pgrid, s,
                               K_switch);
    }
    return check_hip(hipGetLastError(), "batched LU launch");
}

// Forward then backward sweep on the complex128 work vectors (already
// permuted): L U w = w, or with TRANS U^T L^T w = w.  nmax = largest n of
// the batch (B) or the one problem's n.  One launch zeroes both sweeps'
// synchronisation words, then one chained launch per sweep.
template <typename T, bool C, bool TRANS, bool B>
void tri_launch(const SolveProb& one, const SolveProb* d_tbl, int nmax, int count, hipStream_t s) {
    const int nblk = tri_blocks(nmax);
    hipLaunchKernelGGL(tri_sync_reset_kernel, dim3((2 * (nblk + 1) + 255) / 256, 1, count), dim3(256), 0, s, one,
                       d_tbl, (int)B);
    hipLaunchKernelGGL((trsv_chain_kernel<T, C, true, TRANS, B>), dim3(nblk, 1, count), dim3(256), 0, s, one, d_tbl);
    hipLaunchKernelGGL((trsv_chain_kernel<T, C, false, TRANS, B>), dim3(nblk, 1, count), dim3(256), 0, s, one, d_tbl);
}

template <typename T, bool C>
void tri_solve_t(const T* LU, int64_t ld, int n, double* w, unsigned* sync, bool trans, long long wait, int* fail,
                 hipStream_t s) {
    const SolveProb P{LU, ld, n, w, sync, wait, fail};
    if (trans) tri_launch<T, C, true, false>(P, nullptr, n, 1, s);
    else tri_launch<T, C, false, false>(P, nullptr, n, 1, s);
}

// The synchronisation words of single-problem triangular sweeps.
int ensure_tri_sync(Context* ctx, int n) {
    const size_t words = 2 * (size_t)(tri_blocks(n) + 1);
    if (ctx->tri_sync_words >= words) return BEM_OK;
    if (ctx->tri_sync) hipFree(ctx->tri_sync);
    ctx->tri_sync = nullptr;
    ctx->tri_sync_words = 0;
    int rc = check_hip(hipMalloc(&ctx->tri_sync, words * sizeof(unsigned)), "triangular solve sync words");
    if (rc) return rc;
    ctx->tri_sync_words = words;
    return BEM_OK;
}

int tri_solve_any(Context* ctx, const void* LU, int64_t ld, int n, int dtype, double* w, bool trans, hipStream_t s) {
    int rc = ensure_tri_sync(ctx, n);
    if (rc) return rc;
    unsigned* sync = (unsigned*)ctx->tri_sync;
    int* fail = ctx->counters + CTR_TRI_FAIL;
    const long long wt = ctx->wait_ticks;
    switch (dtype) {
        case BEM_F64: tri_solve_t<double, false>((const double*)LU, ld, n, w, sync, trans, wt, fail, s); break;
        case BEM_C128: tri_solve_t<double, true>((const double*)LU, ld, n, w, sync, trans, wt, fail, s); break;
        case BEM_F32: tri_solve_t<float, false>((const float*)LU, ld, n, w, sync, trans, wt, fail, s); break;
        default: tri_solve_t<float, true>((const float*)LU, ld, n, w, sync, trans, wt, fail, s); break;
    }
    return BEM_OK;
}

__global__ void fail_reset_kernel(int* f) {
    if (threadIdx.x == 0) *f = 0;
}

// The fail word of the triangular sweeps is cleared once per solve call (a
// kernel, so the sequence can be captured) and stays raised through the
// call's later sweeps.
void tri_fail_reset(Context* ctx, hipStream_t s) {
    hipLaunchKernelGGL(fail_reset_kernel, dim3(1), dim3(64), 0, s, ctx->counters + CTR_TRI_FAIL);
}

// Reports a sweep that gave up a wait as an error.  Under stream capture the
// word cannot be read back; the NaN the sweep left in its block is then the
// only sign.
int tri_fail_check(Context* ctx, hipStream_t s, const char* what) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    int rc = check_hip(hipStreamIsCapturing(s, &cs), "capture status");
    if (rc) return rc;
    if (cs != hipStreamCaptureStatusNone) return BEM_OK;
    int h = 0;
    rc = check_hip(hipMemcpyAsync(&h, ctx->counters + CTR_TRI_FAIL, sizeof(int), hipMemcpyDeviceToHost, s),
                   "triangular solve status");
    if (rc) return rc;
    rc = check_hip(hipStreamSynchronize(s), "triangular solve sync");
    if (rc) return rc;
    if (h) return set_error(BEM_ERR_HIP, std::string(what) + ": a triangular sweep's wait timed out");
    return BEM_OK;
}

int ensure_work(Context* ctx, size_t bytes) {
    if (ctx->work_bytes >= bytes) return BEM_OK;
    if (ctx->work) hipFree(ctx->work);
    ctx->work = nullptr;
    ctx->work_bytes = 0;
    int rc = check_hip(hipMalloc(&ctx->work, bytes), "solver workspace");
    if (rc) return rc;
    ctx->work_bytes = bytes;
    return BEM_OK;
}

// LAPACK-style interchanges -> gather permutation (host; ipiv is n ints).
int make_perm(const int32_t* d_ipiv, int n, int32_t* d_perm, hipStream_t s) {
    std::vector<int32_t> ip(n), perm(n);
    int rc = check_hip(hipMemcpyAsync(ip.data(), d_ipiv, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s), "ipiv");
    if (rc) return rc;
    rc = check_hip(hipStreamSynchronize(s), "ipiv sync");
    if (rc) return rc;
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int c = 0; c < n; ++c) {
        int p = ip[c];
        if (p < 0 || p >= n) return set_error(BEM_ERR_INVALID, "invalid pivot index");
        int t = perm[c]; perm[c] = perm[p]; perm[p] = t;
    }
    rc = check_hip(hipMemcpyAsync(d_perm, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s), "perm");
    if (rc) return rc;
    return check_hip(hipStreamSynchronize(s), "perm sync");
}

bool valid_dtype(int d) { return d >= BEM_F64 && d <= BEM_C64; }

void residual_any(const void* A, int64_t ld, int n, int dtype, const double* x, const double* b, double* r,
                  hipStream_t s) {
    dim3 g((n + 3) / 4);
    switch (dtype) {
        case BEM_F64: hipLaunchKernelGGL((residual_kernel<double, false>), g, dim3(256), 0, s, (const double*)A, ld, n, x, b, r); break;
        case BEM_C128: hipLaunchKernelGGL((residual_kernel<double, true>), g, dim3(256), 0, s, (const double*)A, ld, n, x, b, r); break;
        case BEM_F32: hipLaunchKernelGGL((residual_kernel<float, false>), g, dim3(256), 0, s, (const float*)A, ld, n, x, b, r); break;
        default: hipLaunchKernelGGL((residual_kernel<float, true>), g, dim3(256), 0, s, (const float*)A, ld, n, x, b, r); break;
    }
}

}  // namespace

int lu_factor(Context* ctx, void* A, int64_t ld, int n, int dtype, int32_t* ipiv, hipStream_t s) {
    if (!A || !ipiv) return set_error(BEM_ERR_INVALID, "null matrix or pivot array");
    if (n <= 0 || ld < n) return set_error(BEM_ERR_INVALID, "invalid matrix size");
    if (!valid_dtype(dtype)) return set_error(BEM_ERR_INVALID, "unknown dtype");
    switch (dtype) {
        case BEM_F64: return lu_factor_t<double, false>(ctx, (double*)A, ld, n, ipiv, s);
        case BEM_C128: return lu_factor_t<double, true>(ctx, (double*)A, ld, n, ipiv, s);
        case BEM_F32: return lu_factor_t<float, false>(ctx, (float*)A, ld, n, ipiv, s);
        default: return lu_factor_t<float, true>(ctx, (float*)A, ld, n, ipiv, s);
    }
}

int lu_factor_batched(Context* ctx, int count, void* const* A, const int64_t* ld, const int32_t* n, int dtype,
                      int32_t* const* ipiv, int32_t* info, hipStream_t s) {
    if (count <= 0) return BEM_OK;
    if (!A || !ld || !n || !ipiv || !info) return set_error(BEM_ERR_INVALID, "null argument");
    if (!valid_dtype(dtype)) return set_error(BEM_ERR_INVALID, "unknown dtype");
    for (int q = 0; q < count; ++q) {
        if (!A[q] || !ipiv[q]) return set_error(BEM_ERR_INVALID, "null matrix or pivot array in the batch");
        if (n[q] <= 0 || ld[q] < n[q]) return set_error(BEM_ERR_INVALID, "invalid matrix size in the batch");
    }
    switch (dtype) {
        case BEM_F64: return lu_factor_batched_t<double, false>(ctx, count, A, ld, n, ipiv, info, s);
        case BEM_C128: return lu_factor_batched_t<double, true>(ctx, count, A, ld, n, ipiv, info, s);
        case BEM_F32: return lu_factor_batched_t<float, false>(ctx, count, A, ld, n, ipiv, info, s);
        default: return lu_factor_batched_t<float, true>(ctx, count, A, ld, n, ipiv, info, s);
    }
}

int lu_solve(Context* ctx, const void* LU, int64_t ld, int n, int dtype, const int32_t* ipiv, void* b,
             hipStream_t s) {
    if (!LU || !ipiv || !b) return set_error(BEM_ERR_INVALID, "null argument");
    if (n <= 0 || ld < n) return set_error(BEM_ERR_INVALID, "invalid matrix size");
    if (!valid_dtype(dtype)) return set_error(BEM_ERR_INVALID, "unknown dtype");
    int rc = ensure_work(ctx, (size_t)n * 16 * 2 + (size_t)n * 4);
    if (rc) return rc;
    double* w0 = (double*)ctx->work;
    double* w1 = w0 + 2 * (size_t)n;
    int32_t* perm = (int32_t*)(w1 + 2 * (size_t)n);
    rc = make_perm(ipiv, n, perm, s);
    if (rc) return rc;
    dim3 g((n + 255) / 256);
    tri_fail_reset(ctx, s);
    hipLaunchKernelGGL(vec_to_work, g, dim3(256), 0, s, (const void*)b, dtype, n, w0);
    hipLaunchKernelGGL(permute_kernel, g, dim3(256), 0, s, (const double*)w0, (const int32_t*)perm, n, w1);
    if ((rc = tri_solve_any(ctx, LU, ld, n, dtype, w1, false, s))) return rc;
    hipLaunchKernelGGL(work_to_vec, g, dim3(256), 0, s, (const double*)w1, n, b, dtype, 0);
    if ((rc = check_hip(hipGetLastError(), "LU solve"))) return rc;
    return tri_fail_check(ctx, s, "LU solve");
}

// ---------------------------------------------------------------------------
// Allocation-free solve (bem_lu_solve_noalloc).  The scratch of its kernels
// (two complex128 work vectors, the pivot permutation, the sweeps' ticket and
// ready words and the fail word) is the file-scope array `scratch`, of
// NOALLOC_BYTES per device, which every call on that device shares: each call
// rewrites what it reads (the synchronisation and fail words are cleared by
// kernels of the call), so consecutive calls are independent, but calls that
// overlap on one device race.  Its device address is taken once per device by
// noalloc_setup (bem_create); host code uses that address, never the array's
// name, which on the host is only a stand-in.
// ---------------------------------------------------------------------------
namespace {

constexpr int NOALLOC_NMAX = 65536;
constexpr size_t noalloc_align(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t NOALLOC_W = noalloc_align((size_t)NOALLOC_NMAX * 16);  // one complex128 work vector
constexpr size_t NOALLOC_PERM = noalloc_align((size_t)NOALLOC_NMAX * 4);
constexpr size_t NOALLOC_SYNC = noalloc_align((size_t)2 * ((NOALLOC_NMAX + TB - 1) / TB + 1) * 4);
constexpr size_t NOALLOC_BYTES = 2 * NOALLOC_W + NOALLOC_PERM + NOALLOC_SYNC + 256;  // + the fail word

__device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mutex;
void* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// LAPACK interchanges -> gather permutation on the device: the swaps are
// applied in order by one thread (n steps), so no host round trip is needed.
__global__ void perm_from_ipiv_kernel(const int32_t* ipiv, int n, int32_t* perm) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) perm[i] = i;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int c = 0; c < n; ++c) {
            const int p = ipiv[c];
            if (p != c && p >= 0 && p < n) { const int t = perm[c]; perm[c] = perm[p]; perm[p] = t; }
        }
}

}  // namespace

// Called once per context with its device current: records this device's
// address of `scratch` (each device holds its own copy).
int noalloc_setup(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        return set_error(BEM_ERR_INVALID, "device index beyond the no-allocation scratch table");
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (g_noalloc_ptr[device]) return BEM_OK;
    void* ptr = nullptr;
    const int rc = check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "no-allocation scratch address");
    if (rc) return rc;
    g_noalloc_ptr[device] = ptr;
    return BEM_OK;
}

int lu_solve_noalloc(Context* ctx, const void* LU, int64_t ld, int n, int dtype, const int32_t* ipiv, void* b,
                     hipStream_t s) {
    if (!LU || !ipiv || !b) return set_error(BEM_ERR_INVALID, "null argument");
    if (n <= 0 || ld < n) return set_error(BEM_ERR_INVALID, "invalid matrix size");
    if (!valid_dtype(dtype)) return set_error(BEM_ERR_INVALID, "unknown dtype");
    if (n > NOALLOC_NMAX)
        return set_error(BEM_ERR_INVALID, "bem_lu_solve_noalloc: n = " + std::to_string(n) +
                                              " exceeds the scratch limit of " + std::to_string(NOALLOC_NMAX));
    char* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mutex);
        if (ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES) base = (char*)g_noalloc_ptr[ctx->device];
    }
    if (!base) return set_error(BEM_ERR_INVALID, "no-allocation scratch not set up on this device");
    double* w0 = (double*)base;
    double* w1 = (double*)(base + NOALLOC_W);
    int32_t* perm = (int32_t*)(base + 2 * NOALLOC_W);
    unsigned* sync = (unsigned*)(base + 2 * NOALLOC_W + NOALLOC_PERM);
    int* fail = (int*)(base + 2 * NOALLOC_W + NOALLOC_PERM + NOALLOC_SYNC);
    const dim3 g((n + 255) / 256);
    hipLaunchKernelGGL(fail_reset_kernel, dim3(1), dim3(64), 0, s, fail);
    hipLaunchKernelGGL(perm_from_ipiv_kernel, dim3(1), dim3(256), 0, s, ipiv, n, perm);
    hipLaunchKernelGGL(vec_to_work, g, dim3(256), 0, s, (const void*)b, dtype, n, w0);
    hipLaunchKernelGGL(permute_kernel, g, dim3(256), 0, s, (const double*)w0, (const int32_t*)perm, n, w1);
    const long long wt = ctx->wait_ticks;
    switch (dtype) {
        case BEM_F64: tri_solve_t<double, false>((const double*)LU, ld, n, w1, sync, false, wt, fail, s); break;
        case BEM_C128: tri_solve_t<double, true>((const double*)LU, ld, n, w1, sync, false, wt, fail, s); break;
        case BEM_F32: tri_solve_t<float, false>((const float*)LU, ld, n, w1, sync, false, wt, fail, s); break;
        default: tri_solve_t<float, true>((const float*)LU, ld, n, w1, sync, false, wt, fail, s); break;
    }
    hipLaunchKernelGGL(work_to_vec, g, dim3(256), 0, s, (const double*)w1, n, b, dtype, 0);
    int rc = check_hip(hipGetLastError(), "allocation-free LU solve");
    if (rc) return rc;
    // the time-out report, as bem_lu_solve's (skipped under stream capture)
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    rc = check_hip(hipStreamIsCapturing(s, &cs), "capture status");
    if (rc || cs != hipStreamCaptureStatusNone) return rc;
    int h = 0;
    rc = check_hip(hipMemcpyAsync(&h, fail, sizeof(int), hipMemcpyDeviceToHost, s), "triangular solve status");
    if (rc) return rc;
    rc = check_hip(hipStreamSynchronize(s), "triangular solve sync");
    if (rc) return rc;
    if (h) return set_error(BEM_ERR_HIP, "allocation-free LU solve: a triangular sweep's wait timed out");
    return BEM_OK;
}

// A^T x = b (plain transpose, no conjugation): with P A = L U,
// A^T = U^T L^T P, so z = L^-T U^-T b and x = P^T z.
int lu_solve_transpose(Context* ctx, const void* LU, int64_t ld, int n, int dtype, const int32_t* ipiv, void* b,
                       hipStream_t s) {
    if (!LU || !ipiv || !b) return set_error(BEM_ERR_INVALID, "null argument");
    if (n <= 0 || ld < n) return set_error(BEM_ERR_INVALID, "invalid matrix size");
    if (!valid_dtype(dtype)) return set_error(BEM_ERR_INVALID, "unknown dtype");
    int rc = ensure_work(ctx, (size_t)n * 16 * 2 + (size_t)n * 4);
    if (rc) return rc;
    double* w0 = (double*)ctx->work;
    double* w1 = w0 + 2 * (size_t)n;
    int32_t* perm = (int32_t*)(w1 + 2 * (size_t)n);
    rc = make_perm(ipiv, n, perm, s);
    if (rc) return rc;
    dim3 g((n + 255) / 256);
    tri_fail_reset(ctx, s);
    hipLaunchKernelGGL(vec_to_work, g, dim3(256), 0, s, (const void*)b, dtype, n, w0);
    if ((rc = tri_solve_any(ctx, LU, ld, n, dtype, w0, true, s))) return rc;
    hipLaunchKernelGGL(unpermute_kernel, g, dim3(256), 0, s, (const double*)w0, (const int32_t*)perm, n, w1);
    hipLaunchKernelGGL(work_to_vec, g, dim3(256), 0, s, (const double*)w1, n, b, dtype, 0);
    if ((rc = check_hip(hipGetLastError(), "transposed LU solve"))) return rc;
    return tri_fail_check(ctx, s, "transposed LU solve");
}

int gemv_residual(Context* ctx, const void* A, int64_t ld, int n, int dtype_a, const void* x,
                  const void* b, void* r, hipStream_t s) {
    if (!A || !x || !b || !r) return set_error(BEM_ERR_INVALID, "null argument");
    if (dtype_a != BEM_F64 && dtype_a != BEM_C128)
        return set_error(BEM_ERR_INVALID, "residual needs an FP64 matrix");
    int rc = ensure_work(ctx, (size_t)n * 16 * 3 + (size_t)n * 4);
    if (rc) return rc;
    double* wx = (double*)ctx->work;
    double* wb = wx + 2 * (size_t)n;
    double* wr = wb + 2 * (size_t)n;
    dim3 g((n + 255) / 256);
    hipLaunchKernelGGL(vec_to_work, g, dim3(256), 0, s, x, dtype_a, n, wx);
    hipLaunchKernelGGL(vec_to_work, g, dim3(256), 0, s, b, dtype_a, n, wb);
    residual_any(A, ld, n, dtype_a, wx, wb, wr, s);
    hipLaunchKernelGGL(work_to_vec, g, dim3(256), 0, s, (const double*)wr, n, r, dtype_a, 0);
    return check_hip(hipGetLastError(), "residual");
}

int refine_solve(Context* ctx, const void* A, int64_t lda, int dtype_a, const void* LU, int64_t ldlu, int n,
                 int dtype_lu, const int32_t* ipiv, const void* b, void* x, int max_iter, double tol,
                 bem_refine_stats* st, hipStream_t s) {
    if (!A || !LU || !ipiv || !b || !x) return set_error(BEM_ERR_INVALID, "null argument");
    if (!valid_dtype(dtype_lu) || !valid_dtype(dtype_a)) return set_error(BEM_ERR_INVALID, "unknown dtype");
    if (dtype_complex(dtype_lu) != dtype_complex(dtype_a))
        return set_error(BEM_ERR_INVALID, "matrix and LU must both be real or both complex");
    const int dtype64 = dtype_complex(dtype_lu) ? BEM_C128 : BEM_F64;
    // work: x, b, r, d (complex128 each), perm, then a 64-byte aligned norm
    // accumulator (up to 63 bytes of padding + 8 bytes)
    int rc = ensure_work(ctx, (size_t)n * 16 * 4 + (size_t)n * 4 + 128);
    if (rc) return rc;
    double* wx = (double*)ctx->work;
    double* wb = wx + 2 * (size_t)n;
    double* wr = wb + 2 * (size_t)n;
    double* wd = wr + 2 * (size_t)n;
    int32_t* perm = (int32_t*)(wd + 2 * (size_t)n);
    double* nrm = (double*)(((uintptr_t)(perm + n) + 63) & ~(uintptr_t)63);
    rc = make_perm(ipiv, n, perm, s);
    if (rc) return rc;
    dim3 g((n + 255) / 256);
    auto norm = [&](const double* w, double* out) -> int {
        hipMemsetAsync(nrm, 0, sizeof(double), s);
        hipLaunchKernelGGL(norm2_kernel, dim3(min(1024, (n + 255) / 256)), dim3(256), 0, s, w, n, nrm);
        int r2 = check_hip(hipMemcpyAsync(out, nrm, sizeof(double), hipMemcpyDeviceToHost, s), "norm");
        if (r2) return r2;
        r2 = check_hip(hipStreamSynchronize(s), "norm sync");
        *out = sqrt(*out);
        return r2;
    };
    tri_fail_reset(ctx, s);
    hipLaunchKernelGGL(vec_to_work, g, dim3(256), 0, s, b, dtype64, n, wb);
    double bn = 0.0;
    if ((rc = norm(wb, &bn))) return rc;
    // x0 = LU^{-1} b
    hipLaunchKernelGGL(permute_kernel, g, dim3(256), 0, s, (const double*)wb, (const int32_t*)perm, n, wx);
    if ((rc = tri_solve_any(ctx, LU, ldlu, n, dtype_lu, wx, false, s))) return rc;
    double r0n = -1.0, rn = 0.0;
    int it = 0, conv = 0;
    for (;;) {
        residual_any(A, lda, n, dtype_a, wx, wb, wr, s);
        if ((rc = norm(wr, &rn))) return rc;
        if (r0n < 0.0) r0n = rn;
        if (rn <= tol * bn) { conv = 1; break; }
        if (it >= max_iter) break;
        hipLaunchKernelGGL(permute_kernel, g, dim3(256), 0, s, (const double*)wr, (const int32_t*)perm, n, wd);
        if ((rc = tri_solve_any(ctx, LU, ldlu, n, dtype_lu, wd, false, s))) return rc;
        hipLaunchKernelGGL(work_to_vec, g, dim3(256), 0, s, (const double*)wd, n, (void*)wx, BEM_C128, 1);
        ++it;
    }
    hipLaunchKernelGGL(work_to_vec, g, dim3(256), 0, s, (const double*)wx, n, x, dtype64, 0);
    rc = check_hip(hipGetLastError(), "refinement");
    if (rc) return rc;
    rc = check_hip(hipStreamSynchronize(s), "refinement sync");
    if (st) {
        st->iterations = it; st->converged = conv; st->residual_norm0 = r0n;
        st->residual_norm = rn; st->b_norm = bn;
    }
    if (rc) return rc;
    return tri_fail_check(ctx, s, "refined solve");
}

int matrix_axpy(const void* src, int64_t lds, void* dst, int64_t ldd, int dtype, int rows, int cols,
                double ar, double ai, int transpose, hipStream_t s) {
    if (!src || !dst) return set_error(BEM_ERR_INVALID, "null argument");
    if (!valid_dtype(dtype)) return set_error(BEM_ERR_INVALID, "unknown dtype");
    if (rows <= 0 || cols <= 0) return BEM_OK;
    dim3 g((cols + 31) / 32, (rows + 31) / 32);
    switch (dtype) {
        case BEM_F64: hipLaunchKernelGGL((axpy_kernel<double, false>), g, dim3(256), 0, s, (const double*)src, lds, (double*)dst, ldd, rows, cols, ar, ai, transpose); break;
        case BEM_C128: hipLaunchKernelGGL((axpy_kernel<double, true>), g, dim3(256), 0, s, (const double*)src, lds, (double*)dst, ldd, rows, cols, ar, ai, transpose); break;
        case BEM_F32: hipLaunchKernelGGL((axpy_kernel<float, false>), g, dim3(256), 0, s, (const float*)src, lds, (float*)dst, ldd, rows, cols, ar, ai, transpose); break;
        default: hipLaunchKernelGGL((axpy_kernel<float, true>), g, dim3(256), 0, s, (const float*)src, lds, (float*)dst, ldd, rows, cols, ar, ai, transpose); break;
    }
    return check_hip(hipGetLastError(), "matrix axpy");
}

int convert_matrix(const void* src, int ds, int64_t lds, void* dst, int dd, int64_t ldd, int rows, int cols,
                   hipStream_t s) {
    if (!src || !dst) return set_error(BEM_ERR_INVALID, "null argument");
    if (!valid_dtype(ds) || !valid_dtype(dd)) return set_error(BEM_ERR_INVALID, "unknown dtype");
    if (dtype_complex(ds) != dtype_complex(dd))
        return set_error(BEM_ERR_INVALID, "conversion between real and complex types");
    int64_t tot = (int64_t)rows * cols;
    if (tot <= 0) return BEM_OK;
    dim3 g((unsigned)((tot + 255) / 256));
    if (ds == BEM_C128 && dd == BEM_C64)
        hipLaunchKernelGGL((convert_kernel<double, true, float, true>), g, dim3(256), 0, s, (const double*)src, lds, (float*)dst, ldd, rows, cols);
    else if (ds == BEM_C64 && dd == BEM_C128)
        hipLaunchKernelGGL((convert_kernel<float, true, double, true>), g, dim3(256), 0, s, (const float*)src, lds, (double*)dst, ldd, rows, cols);
    else if (ds == BEM_F64 && dd == BEM_F32)
        hipLaunchKernelGGL((convert_kernel<double, false, float, false>), g, dim3(256), 0, s, (const double*)src, lds, (float*)dst, ldd, rows, cols);
    else if (ds == BEM_F32 && dd == BEM_F64)
        hipLaunchKernelGGL((convert_kernel<float, false, double, false>), g, dim3(256), 0, s, (const float*)src, lds, (double*)dst, ldd, rows, cols);
    else if (ds == BEM_C128)
        hipLaunchKernelGGL((convert_kernel<double, true, double, true>), g, dim3(256), 0, s, (const double*)src, lds, (double*)dst, ldd, rows, cols);
    else if (ds == BEM_C64)
        hipLaunchKernelGGL((convert_kernel<float, true, float, true>), g, dim3(256), 0, s, (const float*)src, lds, (float*)dst, ldd, rows, cols);
    else if (ds == BEM_F64)
        hipLaunchKernelGGL((convert_kernel<double, false, double, false>), g, dim3(256), 0, s, (const double*)src, lds, (double*)dst, ldd, rows, cols);
    else
        hipLaunchKernelGGL((convert_kernel<float, false, float, false>), g, dim3(256), 0, s, (const float*)src, lds, (float*)dst, ldd, rows, cols);
    return check_hip(hipGetLastError(), "convert");
}

// ---------------------------------------------------------------------------
// Batched solves, FP64 refinement and 1-norm condition estimates over the
// factors of bem_lu_factor_batched.  Everything runs on the device (the
// pivot permutation included), so a repeated call with the same arguments
// issues kernels only and can be captured in a graph.
// ---------------------------------------------------------------------------
namespace {

struct VecProb {
    const void* A;      // matrix of the residuals (LU dtype), may be null
    int64_t lda;
    int n;
    const int32_t* ipiv;
    int32_t* perm;
    const double* b;    // complex128 right-hand side (caller)
    double* x;          // complex128 solution (caller)
    double* w;          // work vector (the triangular-solve vector)
    double* r;          // work vector
    double* acc;        // 8 doubles: reductions and the condition estimate state
    double* out;        // caller outputs (2 doubles) or null
};

__device__ __forceinline__ const VecProb& vp(const VecProb* t) { return t[blockIdx.z]; }

// perm from LAPACK interchanges, one problem per block (thread 0 replays
// the swaps in order; the identity is written by the whole block first)
__global__ void vb_perm_kernel(const VecProb* tbl) {
    const VecProb& P = tbl[blockIdx.x];
    for (int i = threadIdx.x; i < P.n; i += blockDim.x) P.perm[i] = i;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t* pm = P.perm;
        for (int c = 0; c < P.n; ++c) {
            const int p = P.ipiv[c];
            if (p != c && p >= 0 && p < P.n) { const int t = pm[c]; pm[c] = pm[p]; pm[p] = t; }
        }
    }
}

// w[i] = src[perm[i]], src = b (which == 0), r (1) or w-copy in r (2 same as 1)
__global__ void vb_gather_kernel(const VecProb* tbl, int which) {
    const VecProb& P = vp(tbl);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const double* src = which == 0 ? P.b : P.r;
    const int j = P.perm[i];
    P.w[2 * i] = src[2 * j];
    P.w[2 * i + 1] = src[2 * j + 1];
}

// r[perm[i]] = w[i] (the inverse permutation, transposed solves)
__global__ void vb_scatter_kernel(const VecProb* tbl) {
    const VecProb& P = vp(tbl);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const int j = P.perm[i];
    P.r[2 * j] = P.w[2 * i];
    P.r[2 * j + 1] = P.w[2 * i + 1];
}

// x = w (accumulate = 0) or x += w
__global__ void vb_update_x_kernel(const VecProb* tbl, int accumulate) {
    const VecProb& P = vp(tbl);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    if (accumulate) { P.x[2 * i] += P.w[2 * i]; P.x[2 * i + 1] += P.w[2 * i + 1]; }
    else { P.x[2 * i] = P.w[2 * i]; P.x[2 * i + 1] = P.w[2 * i + 1]; }
}

// r = b - A x in FP64 (one wave per row)
template <typename T, bool C>
__global__ __launch_bounds__(256) void vb_residual_kernel(const VecProb* tbl) {
    const VecProb& P = vp(tbl);
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= P.n) return;
    const T* A = (const T*)P.A;
    double sr = 0.0, si = 0.0;
    for (int c = lane; c < P.n; c += 64) {
        cx<T> a = Acc<T, C>::ld(A, (int64_t)row * P.lda + c);
        const double xr = P.x[2 * c], xi = P.x[2 * c + 1];
        sr += (double)a.re * xr - (double)a.im * xi;
        si += (double)a.re * xi + (double)a.im * xr;
    }
    sr = wave_sum(sr); si = wave_sum(si);
    if (lane == 0) { P.r[2 * row] = P.b[2 * row] - sr; P.r[2 * row + 1] = P.b[2 * row + 1] - si; }
}

__global__ void vb_zero_acc_kernel(const VecProb* tbl, int count, int first, int num) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    for (int e = first; e < first + num; ++e) tbl[q].acc[e] = 0.0;
}

// acc[slot] += sum |v_i|^2 (kind 0) or sum |v_i| (kind 1) of v = b (which 0),
// r (1) or w (2)
__global__ void vb_norm_kernel(const VecProb* tbl, int which, int kind, int slot) {
    const VecProb& P = vp(tbl);
    const double* v = which == 0 ? P.b : (which == 1 ? P.r : P.w);
    double sum = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.n; i += gridDim.x * blockDim.x) {
        const double a = v[2 * i], b = v[2 * i + 1];
        sum += kind ? sqrt(a * a + b * b) : a * a + b * b;
    }
    __shared__ double red[4];
    sum = wave_sum(sum);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(P.acc + slot, red[0] + red[1] + red[2] + red[3]);
}

// out[0] = sqrt(acc[1] / acc[0]): relative residual ||b - A x|| / ||b||
__global__ void vb_resid_out_kernel(const VecProb* tbl, int count) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const VecProb& P = tbl[q];
    P.out[0] = P.acc[0] > 0.0 ? sqrt(P.acc[1] / P.acc[0]) : sqrt(P.acc[1]);
}

// Hager-Higham steps.  acc[2] = current ||A^-1 x||_1 sum, acc[3] = estimate,
// acc[4] = argmax bits, acc[5] = max |z| bits.
__global__ void ce_init_kernel(const VecProb* tbl) {
    const VecProb& P = vp(tbl);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    P.r[2 * i] = 1.0 / P.n;  // x = (1/n) 1, kept in r for the gather
    P.r[2 * i + 1] = 0.0;
}
// est = max(est, scale * acc[2]); w (= y) -> conj(sign(y)) in w for the
// transposed solve
__global__ void ce_sign_kernel(const VecProb* tbl) {
    const VecProb& P = vp(tbl);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const double a = P.w[2 * i], b = P.w[2 * i + 1];
    const double m = sqrt(a * a + b * b);
    P.w[2 * i] = m > 0.0 ? a / m : 1.0;
    P.w[2 * i + 1] = m > 0.0 ? -b / m : 0.0;
}
__global__ void ce_est_kernel(const VecProb* tbl, int count, double scale_num, int scale_by_n) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const VecProb& P = tbl[q];
    const double v = scale_by_n ? P.acc[2] * scale_num / P.n : P.acc[2];
    if (v > P.acc[3]) P.acc[3] = v;
    P.acc[2] = 0.0;
}
// j = argmax |r_i| (r = A^-T conj(xi); |conj| is the same), first on ties;
// one block per problem
__global__ void ce_argmax_kernel(const VecProb* tbl) {
    const VecProb& P = tbl[blockIdx.x];
    double best = -1.0;
    int bi = 0;
    for (int i = threadIdx.x; i < P.n; i += blockDim.x) {
        const double a = P.r[2 * i], b = P.r[2 * i + 1];
        const double m = a * a + b * b;
        if (m > best) { best = m; bi = i; }
    }
    __shared__ double sv[256];
    __shared__ int si[256];
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            const double ov = sv[threadIdx.x + off];
            const int oi = si[threadIdx.x + off];
            if (ov > sv[threadIdx.x] || (ov == sv[threadIdx.x] && oi < si[threadIdx.x])) {
                sv[threadIdx.x] = ov; si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) P.acc[4] = (double)si[0];
}
// r = e_j (next x), or the alternating test vector (alt = 1):
// x_i = (-1)^i (1 + i / (n - 1))
__global__ void ce_next_kernel(const VecProb* tbl, int alt) {
    const VecProb& P = vp(tbl);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    double v;
    if (alt) v = (i & 1 ? -1.0 : 1.0) * (1.0 + (P.n > 1 ? (double)i / (P.n - 1) : 0.0));
    else v = i == (int)P.acc[4] ? 1.0 : 0.0;
    P.r[2 * i] = v;
    P.r[2 * i + 1] = 0.0;
}
// ||A||_1 = max_j sum_i |a_ij| (columns coalesced over threads); the
// maximum goes through the ordered bits of the non-negative doubles
template <typename T, bool C>
__global__ void ce_norm1_kernel(const VecProb* tbl) {
    const VecProb& P = vp(tbl);
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= P.n) return;
    const T* A = (const T*)P.A;
    double sum = 0.0;
    for (int r = 0; r < P.n; ++r) {
        cx<T> a = Acc<T, C>::ld(A, (int64_t)r * P.lda + col);
        sum += sqrt((double)a.re * a.re + (double)a.im * a.im);
    }
    atomicMax((unsigned long long*)(P.acc + 6), (unsigned long long)__double_as_longlong(sum));
}
__global__ void ce_out_kernel(const VecProb* tbl, int count) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const VecProb& P = tbl[q];
    P.out[0] = P.acc[6];
    P.out[1] = P.acc[3];
}

struct BatchPlan {
    std::vector<int> ns;
    int nmax = 0, count = 0;
    VecProb* d_vec = nullptr;
    SolveProb* d_sp = nullptr;
};

// Lays out the per-problem work vectors and both tables in the batch
// workspace of the context (reused while it is large enough) and uploads the
// tables when they changed.
int batch_plan(Context* ctx, int slot, int count, const void* const* LU, const int64_t* ld, const int32_t* n,
               const int32_t* const* ipiv, const void* const* A, const int64_t* lda, const void* const* b,
               void* const* x, double* out, int out_stride, BatchPlan& plan, hipStream_t s) {
    // slot: 0 = solves, 1 = condition estimates (separate workspaces, so an
    // alternating sequence of both keeps its cached tables)
    void*& ws = ctx->batch_ws[slot];
    size_t& ws_bytes = ctx->batch_bytes[slot];
    std::vector<char>& cached = ctx->batch_tables[slot];
    size_t bytes = 0;
    std::vector<size_t> off(count);
    for (int q = 0; q < count; ++q) {
        off[q] = bytes;
        bytes += ((size_t)n[q] * 4 + 255) / 256 * 256;          // perm
        bytes += ((size_t)n[q] * 16 + 255) / 256 * 256 * 2;     // w, r
        bytes += 256;                                           // acc
        bytes += (2 * ((size_t)tri_blocks(n[q]) + 1) * 4 + 255) / 256 * 256;  // sweep sync words
    }
    const size_t vt = (sizeof(VecProb) * count + 255) / 256 * 256;
    const size_t st = (sizeof(SolveProb) * count + 255) / 256 * 256;
    const size_t need = bytes + vt + st;
    if (ws_bytes < need) {
        if (ws) hipFree(ws);
        ws = nullptr;
        ws_bytes = 0;
        cached.clear();
        int rc = check_hip(hipMalloc(&ws, need), "batch workspace");
        if (rc) return rc;
        ws_bytes = need;
    }
    char* base = (char*)ws;
    std::vector<VecProb> vt_h(count);
    std::vector<SolveProb> st_h(count);
    plan.ns.assign(n, n + count);
    plan.nmax = 0;
    plan.count = count;
    for (int q = 0; q < count; ++q) {
        char* p = base + off[q];
        VecProb v;
        v.A = A ? A[q] : nullptr;
        v.lda = lda ? lda[q] : 0;
        v.n = n[q];
        v.ipiv = ipiv[q];
        v.perm = (int32_t*)p;
        p += ((size_t)n[q] * 4 + 255) / 256 * 256;
        v.b = b ? (const double*)b[q] : nullptr;
        v.x = x ? (double*)x[q] : nullptr;
        v.w = (double*)p;
        p += ((size_t)n[q] * 16 + 255) / 256 * 256;
        v.r = (double*)p;
        p += ((size_t)n[q] * 16 + 255) / 256 * 256;
        v.acc = (double*)p;
        p += 256;
        v.out = out ? out + (size_t)out_stride * q : nullptr;
        vt_h[q] = v;
        st_h[q] = SolveProb{LU[q], ld[q], n[q], v.w, (unsigned*)p, ctx->wait_ticks, ctx->counters + CTR_TRI_FAIL};
        plan.nmax = std::max(plan.nmax, n[q]);
    }
    // The sweep kernels take the problems in table order (blockIdx.z), and
    // workgroups are dispatched z by z: the largest problems go first, so
    // their long chains start at once and the short ones fill in beside them.
    // (Only the sweeps read this table; each entry carries its own vectors.)
    std::stable_sort(st_h.begin(), st_h.end(), [](const SolveProb& x, const SolveProb& y) { return x.n > y.n; });
    plan.d_vec = (VecProb*)(base + bytes);
    plan.d_sp = (SolveProb*)(base + bytes + vt);
    std::vector<char> blob(vt + st, 0);
    memcpy(blob.data(), vt_h.data(), sizeof(VecProb) * count);
    memcpy(blob.data() + vt, st_h.data(), sizeof(SolveProb) * count);
    if (cached != blob) {
        int rc = check_hip(hipMemcpyAsync(plan.d_vec, blob.data(), blob.size(), hipMemcpyHostToDevice, s),
                           "batch tables");
        if (rc) return rc;
        rc = check_hip(hipStreamSynchronize(s), "batch tables sync");
        if (rc) return rc;
        cached.swap(blob);
    }
    return BEM_OK;
}

template <typename T, bool C>
void batch_tri(const BatchPlan& pl, bool trans, hipStream_t s) {
    if (trans) tri_launch<T, C, true, true>(SolveProb{}, pl.d_sp, pl.nmax, pl.count, s);
    else tri_launch<T, C, false, true>(SolveProb{}, pl.d_sp, pl.nmax, pl.count, s);
}

template <typename T, bool C>
void batch_solve_t(const BatchPlan& pl, int iters, bool resid, hipStream_t s) {
    const dim3 gv((pl.nmax + 255) / 256, 1, pl.count), b256(256);
    const dim3 gr((pl.nmax + 3) / 4, 1, pl.count);
    hipLaunchKernelGGL(vb_perm_kernel, dim3(pl.count), dim3(256), 0, s, (const VecProb*)pl.d_vec);
    hipLaunchKernelGGL(vb_gather_kernel, gv, b256, 0, s, (const VecProb*)pl.d_vec, 0);
    batch_tri<T, C>(pl, false, s);
    hipLaunchKernelGGL(vb_update_x_kernel, gv, b256, 0, s, (const VecProb*)pl.d_vec, 0);
    for (int it = 0; it < iters; ++it) {
        hipLaunchKernelGGL((vb_residual_kernel<T, C>), gr, b256, 0, s, (const VecProb*)pl.d_vec);
        hipLaunchKernelGGL(vb_gather_kernel, gv, b256, 0, s, (const VecProb*)pl.d_vec, 1);
        batch_tri<T, C>(pl, false, s);
        hipLaunchKernelGGL(vb_update_x_kernel, gv, b256, 0, s, (const VecProb*)pl.d_vec, 1);
    }
    if (resid) {
        hipLaunchKernelGGL(vb_zero_acc_kernel, dim3((pl.count + 63) / 64), dim3(64), 0, s,
                           (const VecProb*)pl.d_vec, pl.count, 0, 2);
        hipLaunchKernelGGL((vb_residual_kernel<T, C>), gr, b256, 0, s, (const VecProb*)pl.d_vec);
        const dim3 gn(std::min(64, (pl.nmax + 255) / 256), 1, pl.count);
        hipLaunchKernelGGL(vb_norm_kernel, gn, b256, 0, s, (const VecProb*)pl.d_vec, 0, 0, 0);
        hipLaunchKernelGGL(vb_norm_kernel, gn, b256, 0, s, (const VecProb*)pl.d_vec, 1, 0, 1);
        hipLaunchKernelGGL(vb_resid_out_kernel, dim3((pl.count + 63) / 64), dim3(64), 0, s,
                           (const VecProb*)pl.d_vec, pl.count);
    }
}

template <typename T, bool C>
void batch_condest_t(const BatchPlan& pl, int iters, hipStream_t s) {
    const VecProb* tb = pl.d_vec;
    const dim3 gv((pl.nmax + 255) / 256, 1, pl.count), b256(256);
    const dim3 gn(std::min(64, (pl.nmax + 255) / 256), 1, pl.count);
    const dim3 gq((pl.count + 63) / 64), b64(64);
    hipLaunchKernelGGL(vb_perm_kernel, dim3(pl.count), dim3(256), 0, s, tb);
    hipLaunchKernelGGL(vb_zero_acc_kernel, gq, b64, 0, s, tb, pl.count, 0, 8);
    hipLaunchKernelGGL((ce_norm1_kernel<T, C>), gv, b256, 0, s, tb);
    hipLaunchKernelGGL(ce_init_kernel, gv, b256, 0, s, tb);
    for (int it = 0; it < iters; ++it) {
        // y = A^-1 x: w = P x, then the two sweeps
        hipLaunchKernelGGL(vb_gather_kernel, gv, b256, 0, s, tb, 1);
        batch_tri<T, C>(pl, false, s);
        hipLaunchKernelGGL(vb_norm_kernel, gn, b256, 0, s, tb, 2, 1, 2);
        hipLaunchKernelGGL(ce_est_kernel, gq, b64, 0, s, tb, pl.count, 1.0, 0);
        // z = A^-H sign(y) = conj(A^-T conj(sign(y))): the conjugate is taken
        // in ce_sign, the outer one does not change |z|
        hipLaunchKernelGGL(ce_sign_kernel, gv, b256, 0, s, tb);
        batch_tri<T, C>(pl, true, s);
        hipLaunchKernelGGL(vb_scatter_kernel, gv, b256, 0, s, tb);
        hipLaunchKernelGGL(ce_argmax_kernel, dim3(pl.count), b256, 0, s, tb);
        hipLaunchKernelGGL(ce_next_kernel, gv, b256, 0, s, tb, 0);
    }
    // the alternating-sign test vector: est >= 2 ||A^-1 x||_1 / (3 n)
    hipLaunchKernelGGL(ce_next_kernel, gv, b256, 0, s, tb, 1);
    hipLaunchKernelGGL(vb_gather_kernel, gv, b256, 0, s, tb, 1);
    batch_tri<T, C>(pl, false, s);
    hipLaunchKernelGGL(vb_norm_kernel, gn, b256, 0, s, tb, 2, 1, 2);
    hipLaunchKernelGGL(ce_est_kernel, gq, b64, 0, s, tb, pl.count, 2.0 / 3.0, 1);
    hipLaunchKernelGGL(ce_out_kernel, gq, b64, 0, s, tb, pl.count);
}

int batch_check(int count, const void* const* LU, const int64_t* ld, const int32_t* n, int dtype,
                const int32_t* const* ipiv) {
    if (!LU || !ld || !n || !ipiv) return set_error(BEM_ERR_INVALID, "null argument");
    if (!valid_dtype(dtype)) return set_error(BEM_ERR_INVALID, "unknown dtype");
    for (int q = 0; q < count; ++q) {
        if (!LU[q] || !ipiv[q]) return set_error(BEM_ERR_INVALID, "null factor or pivot array in the batch");
        if (n[q] <= 0 || ld[q] < n[q]) return set_error(BEM_ERR_INVALID, "invalid matrix size in the batch");
    }
    return BEM_OK;
}

}  // namespace

int batch_solve(Context* ctx, int count, const void* const* LU, const int64_t* ld, const int32_t* n, int dtype,
                const int32_t* const* ipiv, const void* const* A, const int64_t* lda, const void* const* b,
                void* const* x, int refine_iters, double* resid, hipStream_t s) {
    if (count <= 0) return BEM_OK;
    int rc = batch_check(count, LU, ld, n, dtype, ipiv);
    if (rc) return rc;
    if (!b || !x) return set_error(BEM_ERR_INVALID, "null right-hand side or solution array");
    if ((refine_iters > 0 || resid) && (!A || !lda))
        return set_error(BEM_ERR_INVALID, "refinement and residuals need the matrices");
    if (refine_iters < 0) return set_error(BEM_ERR_INVALID, "negative refinement count");
    BatchPlan pl;
    rc = batch_plan(ctx, 0, count, LU, ld, n, ipiv, A, lda, b, x, resid, 1, pl, s);
    if (rc) return rc;
    tri_fail_reset(ctx, s);
    switch (dtype) {
        case BEM_F64: batch_solve_t<double, false>(pl, refine_iters, resid != nullptr, s); break;
        case BEM_C128: batch_solve_t<double, true>(pl, refine_iters, resid != nullptr, s); break;
        case BEM_F32: batch_solve_t<float, false>(pl, refine_iters, resid != nullptr, s); break;
        default: batch_solve_t<float, true>(pl, refine_iters, resid != nullptr, s); break;
    }
    if ((rc = check_hip(hipGetLastError(), "batched solve"))) return rc;
    return tri_fail_check(ctx, s, "batched solve");
}

int batch_condest(Context* ctx, int count, const void* const* LU, const int64_t* ld, const int32_t* n, int dtype,
                  const int32_t* const* ipiv, const void* const* A, const int64_t* lda, int iters, double* out,
                  hipStream_t s) {
    if (count <= 0) return BEM_OK;
    int rc = batch_check(count, LU, ld, n, dtype, ipiv);
    if (rc) return rc;
    if (!A || !lda || !out) return set_error(BEM_ERR_INVALID, "null matrix or output array");
    if (iters < 1 || iters > 20) return set_error(BEM_ERR_INVALID, "iteration count must be 1..20");
    BatchPlan pl;
    rc = batch_plan(ctx, 1, count, LU, ld, n, ipiv, A, lda, nullptr, nullptr, out, 2, pl, s);
    if (rc) return rc;
    tri_fail_reset(ctx, s);
    switch (dtype) {
        case BEM_F64: batch_condest_t<double, false>(pl, iters, s); break;
        case BEM_C128: batch_condest_t<double, true>(pl, iters, s); break;
        case BEM_F32: batch_condest_t<float, false>(pl, iters, s); break;
        default: batch_condest_t<float, true>(pl, iters, s); break;
    }
    if ((rc = check_hip(hipGetLastError(), "batched condition estimate"))) return rc;
    return tri_fail_check(ctx, s, "batched condition estimate");
}

}  // namespace bem
