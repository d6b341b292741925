// spread_impl.h -- exponential-of-semicircle spreading fused with the x FFT
// (the spread / FFT / deconvolve steps of a type-1 NUFFT, first dimension):
// the kernels and their launch per kernel width.  spread_kernel.hip holds the
// host entry points; spread_w*.hip instantiate the launches of a few widths
// each, so the widths compile in parallel.
//
// One workgroup owns a tile of the periodic fine grid that spans the whole x
// extent and (by x bz) rows in (y, z).  It
//   1. lists the sorted-point ranges that can reach the tile: points are
//      sorted by bin key (z bin major, y bin minor), so the bins a tile
//      visits in one z-bin row form at most two contiguous runs (two when
//      the y range wraps); wave 0 reads their starts and scans their sizes;
//   2. lets every wave walk the candidate list on its own, one candidate per
//      lane, with the next candidate's record loaded while the current one
//      is spread -- there is no block-wide synchronisation in this loop.  A
//      lane computes the rows of the tile its stencil covers and the kernel
//      values phi(d + m) (registers) and adds its w cells per covered row
//      with LDS atomics into accumulator rows (fixed point, see below) that
//      carry w ghost cells past x = n1, so a stencil never wraps inside the
//      loop and every cell address is base + constant;
//   3. runs the x FFT of its rows in LDS (the first Stockham stage reads the
//      accumulators and folds the ghost cells back onto x < w, rows padded by
//      one element per 16 against bank conflicts, twiddles in LDS), keeps the
//      N1 central frequencies, divides by phi_hat and writes them: the only
//      HBM write of the pass.
//
// Accumulation.  On gfx950 an FP32 LDS atomic add costs ~190 cycles per wave
// instruction (lanes serialised), an FP64 one ~21-26, a 64-bit integer one
// ~8-13 and a 32-bit integer one ~7-9 (measured).  The FP32 form therefore
// accumulates in fixed point, and the result does not depend on the order
// of the atomics:
//  * narrow (k_spread_narrow): 32-bit fixed point for each of re and im,
//    packed into one 8-byte cell as the integer im * 2^32 + re (re
//    sign-extended), so a tile holds twice the rows of a 64-bit one and a
//    contribution is one 64-bit LDS add (~13 cycles) instead of two 32-bit
//    ones; the low word of the sum is re's sum, the high word minus re's
//    sign is im's, both exact while below 2^31.  A tile's scale 2^q is set
//    from its covering bound C -- the sum over the sort bins it visits of
//    each bin's bound on how many of its points cover any one cell -- so
//    that C * max|c| * 2^q < 2^30: every cell sum stays below 2^31 in
//    magnitude (the 32-bit adds wrap, the final sums are exact) and the
//    resolution 2^-q is ~ C * max|c| * 2^-30.  Tiles with C > NARROW_MAX
//    (dense inputs, where 32 bits would lose precision) are queued for the
//    wide kernel.
//  * wide (k_spread_wide): 64-bit cells, G rows per pass.  A contribution v
//    becomes round(v * 2^q) through one FP64 fma with 1.5 * 2^52, whose bit
//    pattern is 0x4338000000000000 + round(v * 2^q) for |v * 2^q| < 2^51;
//    the low 48 bits of the integer sum of such patterns are the exact two's
//    complement sum (the constant's low 48 bits are zero); q keeps the tile's
//    candidate count times max|c| below 2^45.  The FP64 form runs this
//    kernel for every tile with FP64 atomics.
#pragma once
#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>

#include "mpsq_kernels.h"
#include "spread_tiles.h"

namespace mpsq {

// wide kernel: 512 threads, x FFT of up to E * NT cells (G rows) per pass
template <typename R> struct SpreadCfg;
template <> struct SpreadCfg<float> { static constexpr int E = 8, NT = 512; };
template <> struct SpreadCfg<double> { static constexpr int E = 8, NT = 512; };
// narrow kernel: 512 threads whose tile (up to 16 rows of a 512-cell grid,
// ~74 KB of LDS, two workgroups per CU) is transformed at once, 2E cells per
// thread.  (One 1024-thread workgroup per CU with 32-row tiles reads each
// point 32% less often but ran 13% slower: with a single workgroup per CU
// nothing overlaps its barrier-separated phases.)
struct NarrowCfg {
    static constexpr int E = 8, NT = 512;
};
constexpr int MAX_WIDE_PASSES = 8;  // narrow tile rows / wide pass rows (32-row real tiles: 8)

constexpr int MAX_TILE_B = 16;   // by, bz <= 16
constexpr int NARROW_MAX = 256;  // largest covering bound of a 32-bit (narrow) FP32 tile
// counters after the 8 per-XCD tile counters: fallback items queued, taken
constexpr int CTR_FB_COUNT = 8 * XCD_CTR_STRIDE, CTR_FB_TAKE = 9 * XCD_CTR_STRIDE;

// phi(t) = exp(beta (sqrt(1 - u^2) - 1)) with u = 2t/w, |u| <= 1.  FP32:
// the exponent is written without the cancellation of sqrt(1 - u^2) - 1,
//   beta (sqrt(1 - u^2) - 1) = -beta u^2 / (1 + sqrt(1 - u^2)),
// so its relative error is a few ulp everywhere (the direct form loses
// ~1e-6 of phi near the kernel's centre, which shows as ~1e-5 errors on the
// weak modes of near-lattice inputs); then the hardware exp2 with
// nbl = -beta log2(e).  A stencil's cells lie on the support (|u| <= 1), so
// a rounding of 1 - u^2 below zero is clamped to the support's edge.
__device__ __forceinline__ float es_kernel_uu(float uu, float nbl)
{
    const float s = __builtin_amdgcn_sqrtf(fmaxf(1.0f - uu, 0.0f));
    return __builtin_amdgcn_exp2f((nbl * uu) * __builtin_amdgcn_rcpf(1.0f + s));
}
__device__ __forceinline__ float es_kernel(float t, float nbl, float ic) { return es_kernel_uu(t * t * ic, nbl); }
// phi at u = t * 2 / w
__device__ __forceinline__ float es_kernel_u(float u, float nbl) { return es_kernel_uu(u * u, nbl); }
// the kernel's shape argument: -beta * log2(e) for FP32, beta for FP64
template <typename R> __host__ __device__ inline R es_shape(double beta);
template <> __host__ __device__ inline float es_shape<float>(double beta) { return (float)(-beta * 1.4426950408889634); }
template <> __host__ __device__ inline double es_shape<double>(double beta) { return beta; }
__device__ __forceinline__ double es_kernel(double t, double beta, double ic)
{
    const double a = 1.0 - t * t * ic;
    return (a > 0.0) ? exp(beta * (sqrt(a) - 1.0)) : 0.0;
}

// nearest integer of x, ties up (floor(x + 0.5)): one v_cvt_rpi_i32_f32
__device__ __forceinline__ int rint_up(float x)
{
    int r;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// Rows [f, f + cnt) of a tile of b rows starting at grid row r0 covered by a
// stencil of w cells starting at grid row l (periodic, period n; b + w <= n
// so the covered rows are one contiguous range); m0 = first stencil index.
__device__ __forceinline__ void overlap(int l, int r0, int b, int w, int n, int &f, int &cnt, int &m0)
{
    int o = l - r0;
    if (o < 0) o += n;
    if (o < b) {
        f = o;
        m0 = 0;
        cnt = min(b - o, w);
    } else {
        m0 = n - o;
        f = 0;
        cnt = (m0 < w) ? min(w - m0, b) : 0;
    }
}

// LDS layout of a tile of `rows` rows (RS = n1 + w ghost cells rounded to
// even, RSF = padded FFT row, RSX = max(RS, RSF)), G = rows of one x-FFT
// group:
//   FFT tile      [rows][RSX] complex R, row r at r * RSX (a wide pass uses
//                 the first G rows);
//   narrow cells  packed 8-byte cells of row r at [r * RSX, r * RSX + RS),
//                 counted from byte narrow_bias(w) (so that a lane's base
//                 address minus 8w is never negative, see k_spread_narrow);
//   wide planes   one pass of G rows: [2][G][RS] 8-byte cells;
// then the n1 twiddles.
__host__ __device__ inline int spread_row_stride(int n1, int w) { return (n1 + w + 1) & ~1; }
__host__ __device__ inline int spread_row_stride_x(int n1, int w)
{
    const int a = spread_row_stride(n1, w), b = padded_len(n1);
    return a > b ? a : b;
}
template <typename R> int spread_max_rows(int n1)
{
    const int te = (sizeof(R) == 4) ? 2 * NarrowCfg::E * NarrowCfg::NT : SpreadCfg<R>::E * SpreadCfg<R>::NT;
    int g = 1;
    while (2 * g * n1 <= te) g *= 2;
    return g;
}
template <typename R> int spread_group_rows(int rows, int n1)
{
    const int te = SpreadCfg<R>::E * SpreadCfg<R>::NT;
    int g = 1;
    while (2 * g <= rows && 2 * g * n1 <= te) g *= 2;
    return g;
}
// bytes of the tile region (accumulators / FFT rows); the twiddles follow it
// (real: the narrow kernel's real form, 4-byte cells)
__host__ __device__ constexpr int narrow_bias(int w) { return (8 * w + 15) / 16 * 16; }
template <typename R> size_t spread_tile_bytes(int rows, int n1, int w, bool real = false)
{
    const size_t G = (size_t)spread_group_rows<R>(rows, n1);
    const size_t RSX = (size_t)spread_row_stride_x(n1, w);
    const size_t narrowb = sizeof(R) == 4 ? (size_t)rows * RSX * (real ? 4 : 8) + (size_t)narrow_bias(w) : 0;
    const size_t wideb = std::max(2 * G * spread_row_stride(n1, w) * 8, G * RSX * sizeof(typename cplx<R>::T));
    return (std::max(narrowb, wideb) + 15) / 16 * 16;
}
template <typename R> size_t spread_lds_bytes(int rows, int n1, int w, bool real = false)
{
    // + n1 twiddles; FP32: + the tile's scaled deconvolution factors (N1 <= n1 / 2
    // floats) and its rows' output offsets (rows int64)
    const size_t extra = sizeof(R) == 4 ? (size_t)(n1 / 2) * 4 + (size_t)rows * 8 : 0;
    return spread_tile_bytes<R>(rows, n1, w, real) + (size_t)n1 * sizeof(typename cplx<R>::T) + extra;
}
// byte offset of the narrow kernel's per-tile tables after the twiddles
__host__ __device__ inline int spread_tab_offset(int lds_tw, int n1) { return lds_tw + n1 * 8; }

// Real strengths (the narrow kernel's REAL form): one 32-bit cell per row
// cell, rows 2p and 2p + 1 of the accumulators inside complex FFT row p's
// slot (RSX complex elements), so the tile region is rows * RSX * 4 bytes
// past the bias and the twiddles follow it.
__host__ __device__ inline int spread_real_tw_offset(int rows, int n1, int w)
{
    return (narrow_bias(w) + rows * spread_row_stride_x(n1, w) * 4 + 15) / 16 * 16;
}
// rows pair up
__host__ __device__ inline bool spread_real_fits(int rows) { return rows >= 2 && rows % 2 == 0; }
// each wave can transform its own pairs: they fit its lanes' registers
// (2E elements per lane) and a pair's accumulators end inside its FFT slot
// (else the block-wide FFT)
__host__ __device__ inline bool spread_real_wave_fft(int rows, int n1, int w)
{
    constexpr int NW = NarrowCfg::NT / 64;
    const int pairs = rows / 2;
    return ((pairs + NW - 1) / NW) * n1 <= 2 * NarrowCfg::E * 64 &&
           narrow_bias(w) + spread_row_stride(n1, w) * 4 <= spread_row_stride_x(n1, w) * 4;
}

constexpr double FIX_MAGIC = 6755399441055744.0;  // 1.5 * 2^52
// Narrow cells from two FP64 fmas.  fma(v, k, FIX_MAGIC) is FIX_MAGIC +
// round(v k) exactly for |v k| < 2^51, and as a 64-bit pattern that is
// 0x4338000000000000 + round(v k): its low word is re, its high word
// 0x43380000 plus re's sign extension.  With FIX_MAGIC_IM = FIX_MAGIC +
// 0xBCC80000 the low word of the imaginary fma is im + 0xBCC80000 (mod 2^32),
// and 0x43380000 + 0xBCC80000 = 2^32, so
//   packed = {lo: lo(dr), hi: hi(dr) + lo(di)} = im * 2^32 + re (re sign-extended)
// -- two fmas and one 32-bit add per cell.
constexpr double FIX_MAGIC_IM = 6755402608279552.0;  // 1.5 * 2^52 + 0xBCC80000
// {lo: lo(br), hi: hi(br) + lo(bi)}: the 32-bit add is written as one
// instruction (left to itself the compiler forms a 64-bit add plus a move)
__device__ __forceinline__ unsigned long long pack_cell(unsigned long long br, unsigned long long bi)
{
    unsigned hi;
    asm("v_add_u32 %0, %1, %2" : "=v"(hi) : "v"((unsigned)(br >> 32)), "v"((unsigned)bi));
    return ((unsigned long long)hi << 32) | (unsigned)br;
}

// One Stockham stage of `cnt` row transforms of length n (row-major) whose
// input element i of row t comes from load(t, i) and whose output element f
// goes to store(t, f, v): the first stage reads the accumulator planes, the
// middle ones the padded LDS tile, the last one writes the kept frequencies
// to HBM.  sync_mid separates every read of the stage from every write
// (needed whenever the stage writes the memory it reads from).
template <typename C, int E, int NT, int R, class Load, class Store>
__device__ __forceinline__ void stage_io(int n, int cnt, FDiv dnb, FDiv dns, int Ns, int tws,
                                         const C *__restrict__ tw, Load load, Store store, bool sync_mid)
{
    constexpr int TPT = (E + R - 1) / R;
    const int nb = n / R;
    const int total = nb * cnt;
    C v[TPT * R];
    int tt[TPT], bb[TPT];
#pragma unroll
    for (int i = 0; i < TPT; ++i) {
        const int task = threadIdx.x + i * NT;
        bb[i] = -1;
        if (task < total) {
            const int t = (int)fdiv((uint32_t)task, dnb);
            const int j = task - t * nb;
            const int k = j - (int)fdiv((uint32_t)j, dns) * Ns;
            const int tk = k * tws;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                C a = load(t, j + r * nb);
                if (r > 0 && Ns > 1) a = cmul(a, tw[r * tk]);
                v[i * R + r] = a;
            }
            Bfly<R, C>::run(&v[i * R]);
            tt[i] = t;
            bb[i] = (j - k) * R + k;
        }
    }
    if (sync_mid) __syncthreads();
#pragma unroll
    for (int i = 0; i < TPT; ++i)
        if (bb[i] >= 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) store(tt[i], bb[i] + r * Ns, v[i * R + r]);
        }
}

template <typename C, int E, int NT, class Load, class Store>
__device__ __forceinline__ void stage_any(int R, int n, int cnt, FDiv dnb, FDiv dns, int Ns, int tws,
                                          const C *__restrict__ tw, Load load, Store store, bool sync_mid)
{
    switch (R) {
    case 8: stage_io<C, E, NT, 8>(n, cnt, dnb, dns, Ns, tws, tw, load, store, sync_mid); break;
    case 4: stage_io<C, E, NT, 4>(n, cnt, dnb, dns, Ns, tws, tw, load, store, sync_mid); break;
    case 2: stage_io<C, E, NT, 2>(n, cnt, dnb, dns, Ns, tws, tw, load, store, sync_mid); break;
    case 3: stage_io<C, E, NT, 3>(n, cnt, dnb, dns, Ns, tws, tw, load, store, sync_mid); break;
    case 5: stage_io<C, E, NT, 5>(n, cnt, dnb, dns, Ns, tws, tw, load, store, sync_mid); break;
    default: break;
    }
}

// x FFT of `cnt` rows: accumulators -> padded tile -> ... -> kept frequencies
// (LSYNC: the last stage writes back to the tile, so it also separates its
// reads from its writes)
template <typename C, int E, int NT, bool LSYNC = false, class LoadAcc, class LoadTile, class StoreTile,
          class StoreOut>
__device__ __forceinline__ void x_fft(const RadixPlan &rp1, int n1, int cnt, const C *__restrict__ tw, LoadAcc la,
                                      LoadTile lt, StoreTile st, StoreOut so)
{
    int Ns = 1;
    for (int s = 0; s < rp1.nst; ++s) {
        const int Rr = rp1.r[s];
        const bool first = (s == 0), last = (s == rp1.nst - 1);
        if (first && last)
            stage_any<C, E, NT>(Rr, n1, cnt, rp1.dnb[s], rp1.dns[s], Ns, rp1.tws[s], tw, la, so, LSYNC);
        else if (first)
            stage_any<C, E, NT>(Rr, n1, cnt, rp1.dnb[s], rp1.dns[s], Ns, rp1.tws[s], tw, la, st, true);
        else if (last)
            stage_any<C, E, NT>(Rr, n1, cnt, rp1.dnb[s], rp1.dns[s], Ns, rp1.tws[s], tw, lt, so, LSYNC);
        else
            stage_any<C, E, NT>(Rr, n1, cnt, rp1.dnb[s], rp1.dns[s], Ns, rp1.tws[s], tw, lt, st, true);
        if (!last) __syncthreads();
        Ns *= Rr;
    }
}

// ---------------------------------------------------------------------------
// The same x FFT with each wave owning whole rows (rows wv, wv + nwv, ...):
// a row's stages read and write only its own LDS slot, so the stages of a
// wave are separated by a wave-level wait instead of a workgroup barrier and
// the waves run their rows independently.  Usable when every wave's rows fit
// its lanes' registers (ceil(cnt / nwv) * n <= E * 64).
// ---------------------------------------------------------------------------
// LDS writes of this wave complete before its next LDS reads (a wave's LDS
// operations are ordered; the "memory" clobber keeps the compiler from moving
// loads or stores across)
__device__ __forceinline__ void wave_lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

template <typename C, int E, int R, class Load, class Store>
__device__ __forceinline__ void stage_io_w(int n, int cntw, int wv, int nwv, FDiv dnb, FDiv dns, int Ns, int tws,
                                           const C *__restrict__ tw, Load load, Store store, bool sync_mid)
{
    constexpr int TPT = (E + R - 1) / R;
    const int lane = threadIdx.x & 63;
    const int nb = n / R;
    const int total = nb * cntw;
    C v[TPT * R];
    int tt[TPT], bb[TPT];
#pragma unroll
    for (int i = 0; i < TPT; ++i) {
        const int task = lane + i * 64;
        bb[i] = -1;
        if (task < total) {
            const int tl = (int)fdiv((uint32_t)task, dnb);
            const int j = task - tl * nb;
            const int k = j - (int)fdiv((uint32_t)j, dns) * Ns;
            const int tk = k * tws;
            const int t = wv + tl * nwv;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                C a = load(t, j + r * nb);
                if (r > 0 && Ns > 1) a = cmul(a, tw[r * tk]);
                v[i * R + r] = a;
            }
            Bfly<R, C>::run(&v[i * R]);
            tt[i] = t;
            bb[i] = (j - k) * R + k;
        }
    }
    if (sync_mid) wave_lds_sync();
#pragma unroll
    for (int i = 0; i < TPT; ++i)
        if (bb[i] >= 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) store(tt[i], bb[i] + r * Ns, v[i * R + r]);
        }
}

template <typename C, int E, class Load, class Store>
__device__ __forceinline__ void stage_any_w(int R, int n, int cntw, int wv, int nwv, FDiv dnb, FDiv dns, int Ns,
                                            int tws, const C *__restrict__ tw, Load load, Store store, bool sync_mid)
{
    switch (R) {
    case 8: stage_io_w<C, E, 8>(n, cntw, wv, nwv, dnb, dns, Ns, tws, tw, load, store, sync_mid); break;
    case 4: stage_io_w<C, E, 4>(n, cntw, wv, nwv, dnb, dns, Ns, tws, tw, load, store, sync_mid); break;
    case 2: stage_io_w<C, E, 2>(n, cntw, wv, nwv, dnb, dns, Ns, tws, tw, load, store, sync_mid); break;
    case 3: stage_io_w<C, E, 3>(n, cntw, wv, nwv, dnb, dns, Ns, tws, tw, load, store, sync_mid); break;
    case 5: stage_io_w<C, E, 5>(n, cntw, wv, nwv, dnb, dns, Ns, tws, tw, load, store, sync_mid); break;
    default: break;
    }
}

// rows wv, wv + nwv, ... of cnt; FDiv of n / r per stage as in the plan (the
// task index counts this wave's rows only).  LSYNC: the last stage writes the
// memory it reads (its output goes back to the tile), so its reads complete
// before its writes as in the other stages.
template <typename C, int E, bool LSYNC = false, class LoadAcc, class LoadTile, class StoreTile, class StoreOut>
__device__ __forceinline__ void x_fft_w(const RadixPlan &rp1, int n1, int cnt, int wv, int nwv,
                                        const C *__restrict__ tw, LoadAcc la, LoadTile lt, StoreTile st, StoreOut so)
{
    const int cntw = (cnt - wv + nwv - 1) / nwv;
    if (cntw <= 0) return;
    int Ns = 1;
    for (int s = 0; s < rp1.nst; ++s) {
        const int Rr = rp1.r[s];
        const bool first = (s == 0), last = (s == rp1.nst - 1);
        if (first && last)
            stage_any_w<C, E>(Rr, n1, cntw, wv, nwv, rp1.dnb[s], rp1.dns[s], Ns, rp1.tws[s], tw, la, so, LSYNC);
        else if (first)
            stage_any_w<C, E>(Rr, n1, cntw, wv, nwv, rp1.dnb[s], rp1.dns[s], Ns, rp1.tws[s], tw, la, st, true);
        else if (last)
            stage_any_w<C, E>(Rr, n1, cntw, wv, nwv, rp1.dnb[s], rp1.dns[s], Ns, rp1.tws[s], tw, lt, so, LSYNC);
        else
            stage_any_w<C, E>(Rr, n1, cntw, wv, nwv, rp1.dnb[s], rp1.dns[s], Ns, rp1.tws[s], tw, lt, st, true);
        if (!last) wave_lds_sync();
        Ns *= Rr;
    }
}

// binary search of the job owning tile id (tile0s ascending, tile0s[0] = 0)
__device__ __forceinline__ int job_of(const int *__restrict__ tile0s, int njobs, int id)
{
    int lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (tile0s[mid] <= id) lo = mid;
        else hi = mid;
    }
    return __builtin_amdgcn_readfirstlane(lo);  // uniform: id is the workgroup's tile
}

// Tile ids: job-major.  Inside a job whose transforms differ in cost (one
// per species), id - tile0 = (tile position) * ntr + transform, so the XCDs'
// ranges get the same mix of species (3x different point counts); when they
// cost the same (one per frame, tmajor), id - tile0 = transform * ntiles +
// position, so the tiles an XCD runs at once are neighbours in one grid and
// share their candidates in its L2.
template <typename R>
__device__ __forceinline__ void tile_decode(const SpreadJob<R> &J, int id, int &t, int &y0, int &z0)
{
    const int loc = id - J.tile0;
    int ti;
    if (J.tmajor) {
        t = (int)fdiv((uint32_t)loc, J.dnt);
        ti = loc - t * J.ntiles;
    } else {
        ti = (int)fdiv((uint32_t)loc, J.dntr);
        t = loc - ti * J.ntr;
    }
    const int by = (J.g.dim >= 2) ? J.g.by : 1, bz = (J.g.dim >= 3) ? J.g.bz : 1;
    const int tz = (int)fdiv((uint32_t)ti, J.dnty);
    y0 = (ti - tz * ((J.g.dim >= 2) ? J.g.n2 / by : 1)) * by;
    z0 = tz * bz;
}

// candidate runs, strength bound and covering bound of the current and the
// next tile (double-buffered), and the next tile's id, in LDS
struct RunTable {
    int64_t rbeg[2][MAX_RUNS];
    int rpre[2][MAX_RUNS + 1];
    float amax[2];
    int cb[2];
    int next;
};

// Wave 0: the runs of a tile (start and size of run `lane`, inclusive scan of
// the sizes), its strength bound and, with `bound`, its covering bound (the
// sum of the visited bins' bounds), published in buffer pb.
template <typename R>
__device__ __forceinline__ void runs_publish(RunTable &S, int pb, const SpreadJob<R> &J, int t, int y0, int z0,
                                             int lane, bool bound)
{
    const Geo &g = J.g;
    int ylo, nyb, zlo, nzb;
    tile_bins(g, y0, z0, ylo, nyb, zlo, nzb);
    const int nruns = 2 * nzb;  // <= MAX_RUNS (checked by the host)
    int64_t b0 = 0;
    int cnt = 0;
    if (lane < nruns)
        tile_run(g, J.binstart + (int64_t)t * J.bin_tstride * (g.nbins + 1), ylo, nyb, zlo, lane, b0, cnt,
                 [](const int64_t *q) { return gload(q); });
    int cb = 0x7fffffff;
    if (bound && J.binmax) {
        const int *bm = J.binmax + (int64_t)t * J.bin_tstride * (g.nbins + 1);
        const int nv = nyb * nzb;
        int sum = 0;
        for (int i = lane; i < nv; i += 64) {
            const int k = i / nyb, jy = i - k * nyb;
            sum += gload(bm + posmod(zlo + k, g.nbZ) * g.nbY + posmod(ylo + jy, g.nbY));
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
        cb = sum;
    }
    int x = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane < nruns) {
        S.rbeg[pb][lane] = b0;
        S.rpre[pb][lane + 1] = x;
    }
    if (lane == 0) {
        S.rpre[pb][0] = 0;
        S.amax[pb] = J.amax ? J.amax[t] : J.amax_s;
        S.cb[pb] = cb;
    }
}

// one lane's candidate (sorted record and strength): candidate v of the
// tile whose runs are in buffer b; e is the lane's run cursor (v ascending
// between resets)
// The run table is published before it is read (barriers), so rpre is the
// nondecreasing prefix of the tile's run sizes with rpre[nruns] = total and
// the walk stops at e < nruns; the bounds below only guard against a table
// that breaks this (the walk stops at the table's end, the load is skipped
// and the launch's error flag is raised instead of reading out of range).
template <typename R> struct Cand {
    PtRec<R> pr;
    typename cplx<R>::T cv;
    bool have;
    __device__ __forceinline__ void fetch(const SpreadJob<R> &J, const RunTable &S, int b, int v, int tt, int total,
                                          int &e)
    {
        have = v < total;
        if (have) {
            while (e < MAX_RUNS - 1 && S.rpre[b][e + 1] <= v) ++e;
            const int64_t p = S.rbeg[b][e] + (v - S.rpre[b][e]);
            if ((uint64_t)p >= (uint64_t)J.nrec) {
                have = false;
                atomicOr(J.err, 2);
                return;
            }
            pr = gload(J.srec + p);
            cv = gload(J.cs + (int64_t)tt * J.c_tstride + p);
        }
    }
};

// kept frequency f of tile row r, times 2^e2 / phi_hat (conjugated if asked), to G1
template <typename R>
__device__ __forceinline__ void store_kept(const SpreadJob<R> &J, int t, int y0, int z0, int by, int r, int f,
                                           typename cplx<R>::T v, int e2)
{
    const Geo &g = J.g;
    const int N1 = g.N1, h1 = N1 / 2, n1 = g.n1;
    int kk;
    if (f < N1 - h1) kk = f + h1;
    else if (f >= n1 - h1) kk = f - (n1 - h1);
    else return;
    if (kk >= J.g1_keep) return;  // half spectrum: k1 <= 0 only
    v = cscale(v, J.phinv1[kk]);
    if (e2) v = cmk<typename cplx<R>::T>(ldexp(v.x, e2), ldexp(v.y, e2));
    if (J.final_conj) v.y = -v.y;
    const int rz = (int)fdiv((uint32_t)r, J.dby), ry = r - rz * by;
    J.G1[(((int64_t)t * g.n3 + z0 + rz) * g.n2 + y0 + ry) * J.g1_len + kk] = v;
}

template <typename R, int NT>
__device__ __forceinline__ void load_twiddles(const SpreadJob<R> &J, unsigned char *smem, int off)
{
    using C = typename cplx<R>::T;
    C *tw = reinterpret_cast<C *>(smem + off);
    for (int i = threadIdx.x; i < J.g.n1; i += NT) tw[i] = gload(J.tw1 + i);
}

// ---------------------------------------------------------------------------
// Narrow (32-bit) FP32 spread + x FFT, persistent over the tiles of a job
// table.  Workgroup b runs on XCD x = b % 8 (round-robin dispatch of the
// resident grid) and takes the next tile of its XCD's range [xr[x], xr[x+1])
// from the counter wc[x * XCD_CTR_STRIDE]; the host cuts the ranges at equal
// estimated cost.  While a tile is spread, wave 0 already holds the next
// tile's id and loads its runs and bounds; while it is transformed, every
// lane already loads its first candidate of the next tile.  A tile whose
// covering bound exceeds NARROW_MAX (or whose strengths are not finite) is
// appended to fbl as its wide passes (tile id * MAX_WIDE_PASSES + pass)
// instead.
//
// REAL: the strengths are real (the host guarantees it; a point with a
// nonzero imaginary part is skipped and raises err).  A cell is then one
// 32-bit fixed-point word (the low word of the same FP64 fma pattern, so the
// integer sums equal the real parts of the complex form's), one FP64 fma and
// one 32-bit LDS add per contribution; rows 2p and 2p + 1 enter one complex
// x FFT as re and im, and the last stage's output is split by Hermitian
// symmetry, X(f) = (Z(f) + conj Z(-f)) / 2, Y(f) = (Z(f) - conj Z(-f)) / 2i,
// with the 1/2 in the output factor.
// ---------------------------------------------------------------------------
template <int W, bool REAL>
__global__ __launch_bounds__(NarrowCfg::NT, 4) void k_spread_narrow(const SpreadJob<float> *__restrict__ jobs,
                                                          const int *__restrict__ tile0s, int njobs,
                                                          const int *__restrict__ xr, int *__restrict__ wc,
                                                          int *__restrict__ fbl)
{
    using R = float;
    using C = float2;
    constexpr int E = NarrowCfg::E, NT = NarrowCfg::NT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ RunTable S;

    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // first candidate of this wave: the last wave takes the first 64, so that
    // wave 0, which also loads the next tile's runs, gets the ragged tail
    const int vw = (NT / 64 - 1 - wv) * 64;
    const int xcd = blockIdx.x & 7;
    const int xbeg = xr[xcd], xend = xr[xcd + 1];
    // packed cells: im * 2^32 + re (two's complement, mod 2^64), row r at r * RSX;
    // REAL: re alone in 32 bits
    using Cell = typename std::conditional<REAL, unsigned, unsigned long long>::type;
    constexpr int CB = (int)sizeof(Cell);
    Cell *acc = reinterpret_cast<Cell *>(smem_raw + narrow_bias(W));
    // Lane order of a stencil's cells.  Near-lattice inputs put many lanes of
    // a wave on cells of one LDS bank at once (atoms of a lattice row lie a
    // fixed number of cells apart, and the rows are bank-aligned), so lane
    // group g = lane mod NG adds its W cells in the rotated order
    // lx + (m + rot) mod W, rot = g * (W / NG): lanes that start on the same
    // cell use NG different banks at every atomic instruction.  The sums are
    // integers, so the order does not change them.  In the unrolled cell loop
    // the wrap (m + rot >= W) is a compile-time function of m per group, so a
    // row needs NG base addresses and every atomic keeps an immediate offset:
    // cell m of the lane is at base[seg(m)] + CB m with the lane constant
    // rofs[s] = CB rot - (g >= NG - s ? CB W : 0) (>= -CB W: hence the bias;
    // CB = bytes per cell).
    constexpr int NG = (W >= 4) ? 4 : (W >= 2 ? 2 : 1), RSTEP = W / NG;
    const int rg = lane & (NG - 1), rot = rg * RSTEP;
    int rofs[NG];
#pragma unroll
    for (int s = 0; s < NG; ++s) rofs[s] = CB * rot - ((rg >= NG - s) ? CB * W : 0);
    C *tile = reinterpret_cast<C *>(smem_raw);

    if (threadIdx.x == 0) S.next = xbeg + atomicAdd(wc + xcd * XCD_CTR_STRIDE, 1);
    __syncthreads();
    int cur = S.next;
    if (cur >= xend) return;
    // twiddles after the tile region of the job's geometry
    auto tw_off = [](const SpreadJob<R> &Jq) -> int {
        if constexpr (REAL) {
            const int rq = ((Jq.g.dim >= 2) ? Jq.g.by : 1) * ((Jq.g.dim >= 3) ? Jq.g.bz : 1);
            return spread_real_tw_offset(rq, Jq.g.n1, W);
        } else {
            return Jq.lds_tw;
        }
    };
    int j = job_of(tile0s, njobs, cur);
    int t, y0, z0;
    tile_decode(jobs[j], cur, t, y0, z0);
    load_twiddles<R, NT>(jobs[j], smem_raw, tw_off(jobs[j]));
    int pb = 0;
    if (wv == 0) runs_publish(S, pb, jobs[j], t, y0, z0, lane, true);
    __syncthreads();
    // two candidates per lane in flight: this one and the one NT further
    Cand<R> c0, c1;
    int e = 0;
    int total = S.rpre[pb][nruns_of(jobs[j].g, y0, z0)];
    c0.fetch(jobs[j], S, pb, vw + lane, t, total, e);
    c1.fetch(jobs[j], S, pb, vw + NT + lane, t, total, e);

    for (;;) {
        const SpreadJob<R> &J = jobs[j];
        const Geo &g = J.g;
        const int n1 = g.n1, w = g.w;
        const int RSX = spread_row_stride_x(n1, W);  // >= n1 + W: a stencil never leaves its row
        const int by = (g.dim >= 2) ? g.by : 1;
        const int bz = (g.dim >= 3) ? g.bz : 1;
        const int rows = by * bz;
        const float bl = es_shape<R>(g.beta);      // -beta * log2(e)
        const float sw = (float)(2.0 / (double)w);  // u = t * 2/w
        const int nyb = min(W, by);                 // kernel values along y a tile can use
        const int RSXB = RSX * CB;                  // bytes per accumulator row
        const int ltw = tw_off(J);
        const C *tw = reinterpret_cast<const C *>(smem_raw + ltw);
        const float am = S.amax[pb];
        const int cb = S.cb[pb];
        const bool ok = (am < 3.0e38f) && cb <= NARROW_MAX;
        int q = 0;
        if (ok && am > 0.0f) q = 29 - ilogb((double)max(cb, 1) * (double)am);

        if (ok) {
            uint4 *a4 = reinterpret_cast<uint4 *>(acc);
            const int n16 = (rows * RSX * CB + 15) / 16;  // inside the tile region (16-byte multiple)
            for (int i = threadIdx.x; i < n16; i += NT) a4[i] = make_uint4(0u, 0u, 0u, 0u);
            // per tile: the kept frequencies' factors 2^-q / phi_hat and the
            // G1 offset of every tile row, read by the FFT's last stage (the
            // barrier after the spread makes them visible; the previous
            // tile's FFT finished reading them at the loop's last barrier)
            const int N1 = g.N1;
            float *phs = reinterpret_cast<float *>(smem_raw + spread_tab_offset(ltw, n1));
            int64_t *rowoff = reinterpret_cast<int64_t *>(phs + n1 / 2);
            for (int i = threadIdx.x; i < N1; i += NT) phs[i] = ldexpf(gload(J.phinv1 + i), -q - (REAL ? 1 : 0));
            for (int r = threadIdx.x; r < rows; r += NT) {
                const int rz = (int)fdiv((uint32_t)r, J.dby), ry = r - rz * by;
                rowoff[r] = (((int64_t)t * g.n3 + z0 + rz) * g.n2 + y0 + ry) * J.g1_len;
            }
        } else if (threadIdx.x == 0) {
            const int np = rows >> J.grp_sh;  // wide passes of G rows
            const int k = atomicAdd(wc + CTR_FB_COUNT, np);
            for (int p = 0; p < np; ++p) fbl[k + p] = cur * MAX_WIDE_PASSES + p;
        }
        __syncthreads();
        if (wv == 0) {
            // the next tile's id, runs and bounds (other buffer) while this one is spread
            int nxt = 0;
            if (lane == 0) nxt = xbeg + atomicAdd(wc + xcd * XCD_CTR_STRIDE, 1);
            nxt = __shfl(nxt, 0, 64);
            if (lane == 0) S.next = nxt;
            if (nxt < xend) {
                int jq, tq, y0q, z0q;
                jq = job_of(tile0s, njobs, nxt);
                tile_decode(jobs[jq], nxt, tq, y0q, z0q);
                runs_publish(S, pb ^ 1, jobs[jq], tq, y0q, z0q, lane, true);
            }
        }

        // ---- spread: every wave walks candidates v = vw + lane + k*NT
        if (ok) {
            for (int v0 = vw; v0 < total; v0 += NT) {
                const PtRec<R> p = c0.pr;
                const C cc = c0.cv;
                const bool hv = c0.have;
                c0 = c1;
                c1.fetch(J, S, pb, v0 + 2 * NT + lane, t, total, e);
                if (!hv) continue;
                if constexpr (REAL) {
                    if (cc.y != 0.0f) {  // 1 for a NaN (non-finite input), else 2
                        atomicOr(J.err, cc.y == cc.y ? 2 : 1);
                        continue;
                    }
                }
                int fy = 0, ny = 1, my = 0, fz = 0, nz = 1, mz = 0;
                if (g.dim >= 2) overlap(p.ly, y0, by, w, g.n2, fy, ny, my);
                if (g.dim >= 3) overlap(p.lz, z0, bz, w, g.n3, fz, nz, mz);
                if (ny <= 0 || nz <= 0) continue;
                // kernel values at u = (d + m) * 2/w: x for all W cells (FP64
                // for the cell fmas), y only for the tile's rows (nyb <= by,
                // uniform), z one per z row inside the row loop
                // kxd[m]: the kernel at the lane's m-th cell lx + (m + rot) mod W
                double kxd[W];
                const float ux = p.dx * sw;
#pragma unroll
                for (int m = 0; m < W; ++m) {
                    const int mm = (m + rot < W) ? m + rot : m + rot - W;
                    kxd[m] = (double)es_kernel_u(fmaf((float)mm, sw, ux), bl);
                }
                float ky[W];
                const float uy = (p.dy + (float)my) * sw;
#pragma unroll
                for (int iy = 0; iy < W; ++iy) {
                    ky[iy] = 0.0f;
                    if (iy < nyb) ky[iy] = (g.dim < 2) ? 1.0f : (iy < ny) ? es_kernel_u(fmaf((float)iy, sw, uy), bl) : 0.0f;
                }
                const float uz = (p.dz + (float)mz) * sw;
                const float csx = ldexpf(cc.x, q), csy = ldexpf(cc.y, q);  // strength * 2^q
                const int abase = (int)(p.lx * (unsigned)CB);              // byte offset of the first cell
                for (int iz = 0; iz < nz; ++iz) {
                    const float wz = (g.dim >= 3) ? es_kernel_u(fmaf((float)iz, sw, uz), bl) : 1.0f;
                    const int rr0 = __mul24(fz + iz, by) + fy;  // tile row of iy = 0
#pragma unroll
                    for (int iy = 0; iy < W; ++iy) {
                        if (iy < nyb && iy < ny) {  // iy < nyb is uniform
                            const float wyz = ky[iy] * wz;
                            const double vr = (double)(csx * wyz), vi = (double)(csy * wyz);
                            unsigned char *ab = reinterpret_cast<unsigned char *>(acc) +
                                                (__mul24(rr0 + iy, RSXB) + abase);
                            unsigned char *base[NG];
#pragma unroll
                            for (int s = 0; s < NG; ++s) base[s] = ab + rofs[s];
#pragma unroll
                            for (int m = 0; m < W; ++m) {
                                // segment of m: groups g >= NG - seg have wrapped
                                int seg = 0;
#pragma unroll
                                for (int g = 1; g < NG; ++g) seg += (m >= W - g * RSTEP) ? 1 : 0;
                                if constexpr (REAL) {
                                    const unsigned br = (unsigned)(unsigned long long)__double_as_longlong(
                                        __builtin_fma(vr, kxd[m], FIX_MAGIC));
                                    __hip_atomic_fetch_add(reinterpret_cast<unsigned *>(base[seg] + 4 * m), br,
                                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                                    continue;
                                }
                                const unsigned long long br =
                                    (unsigned long long)__double_as_longlong(__builtin_fma(vr, kxd[m], FIX_MAGIC));
                                const unsigned long long bi =
                                    (unsigned long long)__double_as_longlong(__builtin_fma(vi, kxd[m], FIX_MAGIC_IM));
                                __hip_atomic_fetch_add(reinterpret_cast<unsigned long long *>(base[seg] + 8 * m),
                                                       pack_cell(br, bi), __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();
        // the next tile's first candidates are in flight during this FFT
        const int nxt = S.next;
        const bool has_next = nxt < xend;
        int jn = j, tn = 0, y0n = 0, z0n = 0, total_n = 0;
        if (has_next) {
            jn = job_of(tile0s, njobs, nxt);
            tile_decode(jobs[jn], nxt, tn, y0n, z0n);
            total_n = S.rpre[pb ^ 1][nruns_of(jobs[jn].g, y0n, z0n)];
            e = 0;
            c0.fetch(jobs[jn], S, pb ^ 1, vw + lane, tn, total_n, e);
            c1.fetch(jobs[jn], S, pb ^ 1, vw + NT + lane, tn, total_n, e);
        }

        // ---- x FFT of the tile's rows (2E elements per thread)
        if constexpr (REAL) {
            if (ok) {
                // rows 2p and 2p + 1 as one complex row p; each wave its own pairs
                const int N1 = g.N1, h1 = N1 / 2;
                const float *phs = reinterpret_cast<const float *>(smem_raw + spread_tab_offset(ltw, n1));
                const int64_t *rowoff = reinterpret_cast<const int64_t *>(phs + n1 / 2);
                const unsigned *acc32 = reinterpret_cast<const unsigned *>(acc);
                auto load_pair = [&](int pr, int i) -> C {
                    const int a = __mul24(2 * pr, RSX) + i;
                    unsigned x0 = acc32[a], x1 = acc32[a + RSX];
                    if (i < W) {
                        x0 += acc32[a + n1];
                        x1 += acc32[a + RSX + n1];
                    }
                    return cmk<C>((float)(int)x0, (float)(int)x1);  // |sum| < 2^31: exact
                };
                auto load_tile = [&](int row, int i) -> C { return tile[__mul24(row, RSX) + padx<true>(i)]; };
                auto store_tile = [&](int row, int i, C v) { tile[__mul24(row, RSX) + padx<true>(i)] = v; };
                constexpr int NW = NT / 64;
                const int pairs = rows >> 1;
                const int KO = J.g1_keep;  // kept outputs per row (N1, or the half spectrum k1 <= 0)
                const bool cj = J.final_conj != 0;
                C *G1 = J.G1;
                // kept frequency kk of rows 2 pr and 2 pr + 1 from Z = FFT(row 2 pr + i row 2 pr + 1)
                auto split = [&](int pr, int kk) {
                    const C *zr = tile + __mul24(pr, RSX);
                    const int f = (kk >= h1) ? kk - h1 : kk + n1 - h1;
                    const int fm = (f == 0) ? 0 : n1 - f;
                    const C z = zr[padx<true>(f)], y = zr[padx<true>(fm)];
                    const float sc = phs[kk];  // 2^-q / phi_hat / 2
                    const float xr = (z.x + y.x) * sc, xi = (z.y - y.y) * sc;
                    const float yr = (z.y + y.y) * sc, yi = (y.x - z.x) * sc;
                    gstore(G1 + rowoff[2 * pr] + kk, cmk<C>(xr, cj ? -xi : xi));
                    gstore(G1 + rowoff[2 * pr + 1] + kk, cmk<C>(yr, cj ? -yi : yi));
                };
                if (spread_real_wave_fft(rows, n1, W)) {
                    x_fft_w<C, 2 * E, true>(J.rp1, n1, pairs, wv, NW, tw, load_pair, load_tile, store_tile, store_tile);
                    wave_lds_sync();
                    for (int pr = wv; pr < pairs; pr += NW)
                        for (int kk = lane; kk < KO; kk += 64) split(pr, kk);
                } else {
                    x_fft<C, 2 * E, NT, true>(J.rp1, n1, pairs, tw, load_pair, load_tile, store_tile, store_tile);
                    __syncthreads();
                    for (int i = threadIdx.x; i < pairs * KO; i += NT) {
                        const int pr = i / KO;
                        split(pr, i - pr * KO);
                    }
                }
            }
        } else if (ok) {
            const int N1 = g.N1, h1 = N1 / 2;
            const float *phs = reinterpret_cast<const float *>(smem_raw + spread_tab_offset(ltw, n1));
            const int64_t *rowoff = reinterpret_cast<const int64_t *>(phs + n1 / 2);
            // the ghost cells x = n1 + i fold onto i < w here (32-bit adds
            // wrap; the final sums are exact); 2^-q goes into the output factor
            auto load_acc = [&](int row, int i) -> C {
                const int a = __mul24(row, RSX) + i;
                unsigned long long T = acc[a];
                if (i < W) T += acc[a + n1];
                const int re = (int)(unsigned)T;                // |sum| < 2^31: exact
                const int im = (int)(unsigned)(T >> 32) - (re >> 31);  // undo re's sign extension
                return cmk<C>((float)re, (float)im);
            };
            auto load_tile = [&](int row, int i) -> C { return tile[__mul24(row, RSX) + padx<true>(i)]; };
            auto store_tile = [&](int row, int i, C v) { tile[__mul24(row, RSX) + padx<true>(i)] = v; };
            const bool cj = J.final_conj != 0;
            C *G1 = J.G1;
            auto store_out = [&](int row, int f, C v) {
                int kk;
                if (f < N1 - h1) kk = f + h1;
                else if (f >= n1 - h1) kk = f - (n1 - h1);
                else return;
                if (kk >= J.g1_keep) return;
                const float sc = phs[kk];
                gstore(G1 + rowoff[row] + kk, cmk<C>(v.x * sc, cj ? -(v.y * sc) : v.y * sc));
            };
            // each wave transforms whole rows when its share fits its lanes'
            // registers (rows * n1 <= 2E * NT always holds, so a multiple of
            // 8 rows always does) and when every row's accumulators lie in
            // the row's own FFT slot (the accumulators start narrow_bias(W)
            // bytes into the tile, so a row needs that much slack past its
            // n1 + W cells: else the first stage of one row's wave could
            // overwrite the next row's last cells before its wave read
            // them); else all rows at once with workgroup barriers between
            // the stages, whose first stage reads every row before writing
            constexpr int NW = NT / 64;
            if (((rows + NW - 1) / NW) * n1 <= 2 * E * 64 &&
                narrow_bias(W) + spread_row_stride(n1, W) * 8 <= RSX * 8)
                x_fft_w<C, 2 * E>(J.rp1, n1, rows, wv, NW, tw, load_acc, load_tile, store_tile, store_out);
            else
                x_fft<C, 2 * E, NT>(J.rp1, n1, rows, tw, load_acc, load_tile, store_tile, store_out);
        }
        if (!has_next) break;
        __syncthreads();  // every read of the tile (and the twiddles) done
        if (jn != j) load_twiddles<R, NT>(jobs[jn], smem_raw, tw_off(jobs[jn]));
        cur = nxt;
        pb ^= 1;
        j = jn;
        t = tn;
        y0 = y0n;
        z0 = z0n;
        total = total_n;
    }
}

// ---------------------------------------------------------------------------
// Wide spread + x FFT: one pass of G rows of a tile per item.  Items are
// either every tile of the table (fbl == nullptr; per-XCD ranges as in the
// narrow kernel) or the list fbl[0 .. wc[CTR_FB_COUNT]) the narrow kernel
// queued (item = tile id * MAX_WIDE_PASSES + pass), taken through
// wc[CTR_FB_TAKE].
// ---------------------------------------------------------------------------
template <typename R, int W>
__global__ __launch_bounds__(512, 4) void k_spread_wide(const SpreadJob<R> *__restrict__ jobs,
                                                        const int *__restrict__ tile0s, int njobs,
                                                        const int *__restrict__ xr, int *__restrict__ wc,
                                                        const int *__restrict__ fbl)
{
    using C = typename cplx<R>::T;
    constexpr bool F32 = (sizeof(R) == 4);
    constexpr int E = SpreadCfg<R>::E, NT = SpreadCfg<R>::NT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ RunTable S;

    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int vw = (NT / 64 - 1 - wv) * 64;
    const int xcd = blockIdx.x & 7;
    const int xbeg = fbl ? 0 : xr[xcd], xend = fbl ? wc[CTR_FB_COUNT] : xr[xcd + 1];
    unsigned long long *acc64 = reinterpret_cast<unsigned long long *>(smem_raw);
    double *accd = reinterpret_cast<double *>(smem_raw);
    C *tile = reinterpret_cast<C *>(smem_raw);
    // next item: its position k in [xbeg, xend) (k >= xend: none) -> tile id, pass
    auto take = [&]() -> int {
        return fbl ? atomicAdd(wc + CTR_FB_TAKE, 1) : xbeg + atomicAdd(wc + xcd * XCD_CTR_STRIDE, 1);
    };
    auto item_of = [&](int k, int &id, int &pass) {
        int it = fbl ? fbl[k] : MAX_WIDE_PASSES * k;
        it = __builtin_amdgcn_readfirstlane(it);
        id = it / MAX_WIDE_PASSES;
        pass = it % MAX_WIDE_PASSES;
    };

    if (threadIdx.x == 0) S.next = take();
    __syncthreads();
    const int kfirst = S.next;
    if (kfirst >= xend) return;
    int cur, pass;
    item_of(kfirst, cur, pass);
    int j = job_of(tile0s, njobs, cur);
    int t, y0, z0;
    tile_decode(jobs[j], cur, t, y0, z0);
    load_twiddles<R, NT>(jobs[j], smem_raw, jobs[j].lds_tw);
    int pb = 0;
    if (wv == 0) runs_publish(S, pb, jobs[j], t, y0, z0, lane, false);
    __syncthreads();
    Cand<R> cd;
    int e = 0;
    int total = S.rpre[pb][nruns_of(jobs[j].g, y0, z0)];
    cd.fetch(jobs[j], S, pb, vw + lane, t, total, e);

    for (;;) {
        const SpreadJob<R> &J = jobs[j];
        const Geo &g = J.g;
        const int n1 = g.n1, w = g.w;
        const int RS = spread_row_stride(n1, W);
        const int RSX = spread_row_stride_x(n1, W);
        const int by = (g.dim >= 2) ? g.by : 1;
        const int bz = (g.dim >= 3) ? g.bz : 1;
        const int G = J.grp;
        const int r0 = pass * G;    // tile rows [r0, r0 + G) of this item
        const int wplane = G * RS;  // cells of one plane
        const R beta = es_shape<R>(g.beta), ic = (R)g.ic;
        const C *tw = reinterpret_cast<const C *>(smem_raw + J.lds_tw);
        const float am = S.amax[pb];

        {
            uint4 *a4 = reinterpret_cast<uint4 *>(smem_raw);
            for (int i = threadIdx.x; i < wplane; i += NT) a4[i] = make_uint4(0u, 0u, 0u, 0u);
        }
        __syncthreads();
        if (wv == 0) {
            int kn = 0;
            if (lane == 0) kn = take();
            kn = __shfl(kn, 0, 64);
            if (lane == 0) S.next = kn;
            if (kn < xend) {
                int idq, pq, jq, tq, y0q, z0q;
                item_of(kn, idq, pq);
                jq = job_of(tile0s, njobs, idq);
                tile_decode(jobs[jq], idq, tq, y0q, z0q);
                runs_publish(S, pb ^ 1, jobs[jq], tq, y0q, z0q, lane, false);
            }
        }

        // fixed-point scale 2^q (FP32): total * amax * 2^q < 2^45
        double fscale = 1.0, finv = 1.0;
        bool fbad = false;
        if (F32) {
            fbad = !(am < 3.0e38f);
            const double bound = (double)max(total, 1) * (double)(am > 0.0f && !fbad ? am : 1.0f);
            const int qq = 44 - ilogb(bound);
            fscale = ldexp(1.0, qq);
            finv = ldexp(1.0, -qq);
        }

        for (int v0 = vw; v0 < total; v0 += NT) {
            const PtRec<R> p = cd.pr;
            const C cc = cd.cv;
            const bool hv = cd.have;
            cd.fetch(J, S, pb, v0 + NT + lane, t, total, e);
            if (!hv) continue;
            int fy = 0, ny = 1, my = 0, fz = 0, nz = 1, mz = 0;
            if (g.dim >= 2) overlap(p.ly, y0, by, w, g.n2, fy, ny, my);
            if (g.dim >= 3) overlap(p.lz, z0, bz, w, g.n3, fz, nz, mz);
            if (ny <= 0 || nz <= 0) continue;
            double kxd[W];
            R kx[W];
#pragma unroll
            for (int m = 0; m < W; ++m) {
                kx[m] = es_kernel(p.dx + (R)m, beta, ic);
                kxd[m] = (double)kx[m] * fscale;
            }
            R ky[W];
#pragma unroll
            for (int iy = 0; iy < W; ++iy)
                ky[iy] = (g.dim >= 2) ? ((iy < ny) ? es_kernel(p.dy + (R)(my + iy), beta, ic) : (R)0) : (R)1;
            for (int iz = 0; iz < nz; ++iz) {
                const R wz = (g.dim >= 3) ? es_kernel(p.dz + (R)(mz + iz), beta, ic) : (R)1;
                const int rr0 = (fz + iz) * by + fy - r0;  // pass row of iy = 0
#pragma unroll
                for (int iy = 0; iy < W; ++iy) {
                    const int r = rr0 + iy;
                    if (iy < ny && r >= 0 && r < G) {
                        const R wyz = ky[iy] * wz;
                        const int c0 = r * RS + p.lx;
                        if (F32) {
                            const double vr = (double)(cc.x * wyz), vi = (double)(cc.y * wyz);
                            unsigned long long *ar = acc64 + c0, *ai = acc64 + wplane + c0;
#pragma unroll
                            for (int m = 0; m < W; ++m) {
                                const double dr = __builtin_fma(vr, kxd[m], FIX_MAGIC);
                                const double di = __builtin_fma(vi, kxd[m], FIX_MAGIC);
                                __hip_atomic_fetch_add(ar + m, (unsigned long long)__double_as_longlong(dr),
                                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                                __hip_atomic_fetch_add(ai + m, (unsigned long long)__double_as_longlong(di),
                                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                            }
                        } else {
                            const R vr = cc.x * wyz, vi = cc.y * wyz;
                            double *ar = accd + c0, *ai = accd + wplane + c0;
#pragma unroll
                            for (int m = 0; m < W; ++m) {
                                __hip_atomic_fetch_add(ar + m, (double)(vr * kx[m]), __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                                __hip_atomic_fetch_add(ai + m, (double)(vi * kx[m]), __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();
        const int kn = S.next;
        const bool has_next = kn < xend;
        int nid = 0, npass = 0, jn = j, tn = 0, y0n = 0, z0n = 0, total_n = 0;
        if (has_next) {
            item_of(kn, nid, npass);
            jn = job_of(tile0s, njobs, nid);
            tile_decode(jobs[jn], nid, tn, y0n, z0n);
            total_n = S.rpre[pb ^ 1][nruns_of(jobs[jn].g, y0n, z0n)];
            e = 0;
            cd.fetch(jobs[jn], S, pb ^ 1, vw + lane, tn, total_n, e);
        }

        // ---- x FFT of the item's G rows
        {
            // the ghost cells x = n1 + i fold onto i < w here
            auto load_acc = [&](int row, int i) -> C {
                const int a = row * RS + i;
                if (F32) {
                    unsigned long long ur = acc64[a], ui = acc64[wplane + a];
                    if (i < W) {
                        ur += acc64[a + n1];
                        ui += acc64[wplane + a + n1];
                    }
                    const long long br = (long long)(ur << 16) >> 16;
                    const long long bi = (long long)(ui << 16) >> 16;
                    return cmk<C>(fbad ? (R)NAN : (R)((double)br * finv), fbad ? (R)NAN : (R)((double)bi * finv));
                } else {
                    double dr = accd[a], di = accd[wplane + a];
                    if (i < W) {
                        dr += accd[a + n1];
                        di += accd[wplane + a + n1];
                    }
                    return cmk<C>((R)dr, (R)di);
                }
            };
            auto load_tile = [&](int row, int i) -> C { return tile[row * RSX + padx<true>(i)]; };
            auto store_tile = [&](int row, int i, C v) { tile[row * RSX + padx<true>(i)] = v; };
            auto store_out = [&](int row, int f, C v) { store_kept(J, t, y0, z0, by, r0 + row, f, v, 0); };
            x_fft<C, E, NT>(J.rp1, n1, G, tw, load_acc, load_tile, store_tile, store_out);
        }
        if (!has_next) break;
        __syncthreads();  // every read of the tile (and the twiddles) done
        if (jn != j) load_twiddles<R, NT>(jobs[jn], smem_raw, jobs[jn].lds_tw);
        pb ^= 1;
        j = jn;
        t = tn;
        y0 = y0n;
        z0 = z0n;
        pass = npass;
        total = total_n;
    }
}

// workgroups of a persistent launch: CUs x resident workgroups per CU at the
// launch's LDS size (cached per kernel and size), rounded to the 8 XCDs
inline int persistent_grid(const void *fn, int nt, size_t lds)
{
    static std::mutex mu;
    static std::map<std::pair<const void *, size_t>, int> cache;
    std::lock_guard<std::mutex> lk(mu);
    const auto key = std::make_pair(fn, lds);
    auto it = cache.find(key);
    if (it == cache.end()) {
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 156 * 1024);
        int dev = 0, nb = 0, ncu = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, nt, lds) != hipSuccess || nb < 1) nb = 1;
        const int grid = std::max(1, ncu) * nb;
        it = cache.emplace(key, (grid + 7) / 8 * 8).first;
    }
    return it->second;
}

// the narrow (FP32; real or complex strengths) and wide launches of width W
template <typename R, int W>
void launch_spread_w(const SpreadJob<R> *jobs, const int *tile0s, int njobs, const int *xr, int *wc, int *fbl,
                     size_t lds, hipStream_t st, size_t lds_real)
{
    constexpr int NT = SpreadCfg<R>::NT;
    launch_zero_i32(wc, XCD_CTR_INTS, st);
    const void *fw = reinterpret_cast<const void *>(&k_spread_wide<R, W>);
    if constexpr (sizeof(R) == 4) {
        if (lds_real > 0) {
            const void *fn = reinterpret_cast<const void *>(&k_spread_narrow<W, true>);
            hipLaunchKernelGGL((k_spread_narrow<W, true>), dim3((unsigned)persistent_grid(fn, NarrowCfg::NT, lds_real)),
                               dim3(NarrowCfg::NT), lds_real, st, jobs, tile0s, njobs, xr, wc, fbl);
        } else {
            const void *fn = reinterpret_cast<const void *>(&k_spread_narrow<W, false>);
            hipLaunchKernelGGL((k_spread_narrow<W, false>), dim3((unsigned)persistent_grid(fn, NarrowCfg::NT, lds)),
                               dim3(NarrowCfg::NT), lds, st, jobs, tile0s, njobs, xr, wc, fbl);
        }
        // the tiles the narrow kernel queued (usually none: its workgroups then
        // only read the empty count)
        hipLaunchKernelGGL((k_spread_wide<R, W>), dim3((unsigned)persistent_grid(fw, NT, lds)), dim3(NT), lds, st,
                           jobs, tile0s, njobs, xr, wc, (const int *)fbl);
    } else {
        hipLaunchKernelGGL((k_spread_wide<R, W>), dim3((unsigned)persistent_grid(fw, NT, lds)), dim3(NT), lds, st,
                           jobs, tile0s, njobs, xr, wc, (const int *)nullptr);
    }
}

#define MPSQ_SPREAD_W_ARGS(R)                                                                                         \
    const SpreadJob<R> *, const int *, int, const int *, int *, int *, size_t, hipStream_t, size_t
// declared here, instantiated in one spread_w*.hip each
#define MPSQ_SPREAD_W_EXTERN(W)                                                                                       \
    extern template void launch_spread_w<float, W>(MPSQ_SPREAD_W_ARGS(float));                                        \
    extern template void launch_spread_w<double, W>(MPSQ_SPREAD_W_ARGS(double));
#define MPSQ_SPREAD_W_INST(W)                                                                                         \
    template void launch_spread_w<float, W>(MPSQ_SPREAD_W_ARGS(float));                                               \
    template void launch_spread_w<double, W>(MPSQ_SPREAD_W_ARGS(double));

}  // namespace mpsq
