// spread_kernel.hip -- exponential-of-semicircle spreading fused with the x FFT
// (the spread / FFT / deconvolve steps of a type-1 NUFFT, first dimension).
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
//      with LDS atomics into planar accumulator rows (real plane, imaginary
//      plane) that carry w ghost cells past x = n1, so a stencil never wraps
//      inside the loop and every cell address is base + constant;
//   3. folds the ghost cells back onto x < w, converts the tile to complex R
//      in place, runs the x FFT of its rows in LDS (rows padded by one
//      element per 16 against bank conflicts; twiddles in LDS), keeps the N1
//      central frequencies, divides by phi_hat and writes them: the only HBM
//      write of the pass.
//
// Accumulation.  On gfx950 an FP32 LDS atomic add costs ~190 cycles per wave
// instruction (lanes serialised), an FP64 one ~21-26 and a 64-bit integer
// one ~8-13 (measured).  The FP32 form therefore accumulates in 64-bit
// fixed point: a contribution v becomes round(v * 2^q) through one FP64
// fma with the constant 1.5 * 2^52, whose bit pattern is 0x4338000000000000
// + round(v * 2^q) for |v * 2^q| < 2^51; the low 48 bits of the integer sum of
// those patterns are the exact two's complement sum of the rounded values
// (the constant's low 48 bits are zero).  q is chosen per tile so that the
// tile's candidate count times the largest strength component stays below
// 2^45, which bounds every cell sum below 2^46: resolution 2^-q relative to
// that bound (about 1e-11 of the largest cell for the trajectory workloads),
// and the result does not depend on the order of the atomics.  The plain
// FP64 form accumulates with FP64 atomics.
//
// Tiles are numbered so that each XCD (workgroups are dealt to the 8 XCDs
// round robin) runs a contiguous range of tiles: neighbouring tiles share
// most of their candidates, which then come from that XCD's L2.
#include <algorithm>

#include "mpsq_kernels.h"

namespace mpsq {

template <typename R> struct SpreadCfg;
template <> struct SpreadCfg<float> { static constexpr int E = 8, NT = 512; };
template <> struct SpreadCfg<double> { static constexpr int E = 8, NT = 512; };

constexpr int MAX_RUNS = 64;    // candidate runs of one tile (two per visited z-bin row)
constexpr int MAX_TILE_B = 8;   // by, bz <= 8

__device__ __forceinline__ int floordiv(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }
__device__ __forceinline__ int posmod(int a, int n)
{
    const int r = a % n;
    return r < 0 ? r + n : r;
}

// phi(t) = exp(beta (sqrt(1 - ic t^2) - 1)), |t| <= w/2.  FP32: hardware
// exp2 / sqrt (relative error ~1e-6 only where phi < 1e-6, i.e. absolute
// error < 1e-12); FP64: the accurate library functions.
__device__ __forceinline__ float es_kernel(float t, float beta, float ic)
{
    const float a = 1.0f - t * t * ic;
    return (a > 0.0f) ? __expf(beta * (__builtin_amdgcn_sqrtf(a) - 1.0f)) : 0.0f;
}
__device__ __forceinline__ double es_kernel(double t, double beta, double ic)
{
    const double a = 1.0 - t * t * ic;
    return (a > 0.0) ? exp(beta * (sqrt(a) - 1.0)) : 0.0;
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

// LDS layout: accumulator planes [2][rows][RS] of 8-byte cells (RS = n1 + w
// ghost cells, rounded to even), reused in place for the FFT tile
// [rows][padded_len(n1)] of complex R; then n1 twiddles.
__host__ __device__ inline int spread_row_stride(int n1, int w) { return (n1 + w + 1) & ~1; }
template <typename R> size_t spread_lds_bytes(int rows, int n1, int w)
{
    const size_t accb = (size_t)2 * rows * spread_row_stride(n1, w) * 8;
    const size_t fftb = (size_t)rows * padded_len(n1) * sizeof(typename cplx<R>::T);
    return ((accb > fftb ? accb : fftb) + 15) / 16 * 16 + (size_t)n1 * sizeof(typename cplx<R>::T);
}

constexpr double FIX_MAGIC = 6755399441055744.0;  // 1.5 * 2^52

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

// Tile id sequence of a persistent workgroup: workgroup b runs on XCD b % 8
// (round-robin dispatch of the resident grid) and takes the tiles
// x*per + b/8 + k*(grid/8) of its XCD's contiguous range [x*per, (x+1)*per),
// so the 2-per-CU workgroups of one XCD always work on neighbouring tiles.
// Tile id = (tile position) * ntr + transform: every XCD gets the same mix of
// transforms (species differ 3x in point count; a transform-major order left
// XCDs unevenly loaded), and the tiles in flight on one XCD still cover a few
// neighbouring positions, whose candidates they share through its L2.
struct TileSeq {
    int cur, step, end;
    __device__ TileSeq(int ntot)
    {
        const int x = blockIdx.x & 7, li = blockIdx.x >> 3, nl = gridDim.x >> 3;
        const int per = (ntot + 7) >> 3;
        cur = x * per + li;
        step = nl;
        end = min(ntot, (x + 1) * per);
    }
};

template <typename R, int E, int NT, int W>
__global__ __launch_bounds__(NT, 4) void k_spread_xfft(Geo g, RadixPlan rp1, const PtRec<R> *__restrict__ srec,
                                                    const typename cplx<R>::T *__restrict__ cs, int64_t c_tstride,
                                                    const int64_t *__restrict__ binstart, int bin_tstride,
                                                    const typename cplx<R>::T *__restrict__ tw1,
                                                    const R *__restrict__ phinv1, typename cplx<R>::T *__restrict__ G1,
                                                    int final_conj, int ntr, int ntiles_all, FDiv dntr,
                                                    FDiv dnty, FDiv dby, const float *__restrict__ amax)
{
    using C = typename cplx<R>::T;
    constexpr bool FIX = (sizeof(R) == 4);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int64_t rbeg[MAX_RUNS];
    __shared__ int rpre[MAX_RUNS + 1];
    __shared__ float s_amax;

    const int n1 = g.n1, w = g.w;
    const int RS = spread_row_stride(n1, W);
    const int RSF = padded_len(n1);
    const int by = (g.dim >= 2) ? g.by : 1;
    const int bz = (g.dim >= 3) ? g.bz : 1;
    const int rows = by * bz;
    const R beta = (R)g.beta, ic = (R)g.ic;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int N1 = g.N1, h1 = N1 / 2;

    unsigned long long *accu = reinterpret_cast<unsigned long long *>(smem_raw);
    double *accd = reinterpret_cast<double *>(smem_raw);
    C *tile = reinterpret_cast<C *>(smem_raw);
    const int plane = rows * RS;
    const size_t accb = (size_t)2 * plane * 8, fftb = (size_t)rows * RSF * sizeof(C);
    C *tw = reinterpret_cast<C *>(smem_raw + ((accb > fftb ? accb : fftb) + 15) / 16 * 16);
    for (int i = threadIdx.x; i < n1; i += NT) tw[i] = tw1[i];

    // decode a tile id: transform, first row y0, first plane z0
    auto decode = [&](int id, int &t, int &y0, int &z0) {
        const int ti = (int)fdiv((uint32_t)id, dntr);
        t = id - ti * ntr;
        const int tz = (int)fdiv((uint32_t)ti, dnty);
        y0 = (ti - tz * (g.dim >= 2 ? g.n2 / by : 1)) * by;
        z0 = tz * bz;
    };
    // wave 0: the candidate runs of a tile (lane e = 2 * z-bin row + part):
    // start and size, loaded into registers
    auto runs_load = [&](int t, int y0, int z0, int &nruns, int64_t &b0, int &cnt) {
        int ylo = 0, nyb = 1, zlo = 0, nzb = 1;
        if (g.dim >= 2) {
            ylo = floordiv(y0 - (w - 1), g.BY);
            nyb = floordiv(y0 + by - 1, g.BY) - ylo + 1;
            if (nyb >= g.nbY) {
                ylo = 0;
                nyb = g.nbY;
            }
        }
        if (g.dim >= 3) {
            zlo = floordiv(z0 - (w - 1), g.BZ);
            nzb = floordiv(z0 + bz - 1, g.BZ) - zlo + 1;
            if (nzb >= g.nbZ) {
                zlo = 0;
                nzb = g.nbZ;
            }
        }
        nruns = 2 * nzb;  // <= MAX_RUNS (checked by the host)
        b0 = 0;
        cnt = 0;
        if (lane < nruns) {
            const int64_t *bs = binstart + (int64_t)t * bin_tstride * (g.nbins + 1);
            const int k = lane >> 1, part = lane & 1;
            const int zb = posmod(zlo + k, g.nbZ);
            int ya = 0, yb = 0;  // bin range [ya, yb) of this part
            const int a = ylo, b = ylo + nyb;
            if (nyb >= g.nbY) {
                if (part == 0) yb = g.nbY;
            } else if (a < 0) {
                if (part == 0) {
                    ya = a + g.nbY;
                    yb = g.nbY;
                } else {
                    yb = b;
                }
            } else if (b > g.nbY) {
                if (part == 0) {
                    ya = a;
                    yb = g.nbY;
                } else {
                    yb = b - g.nbY;
                }
            } else if (part == 0) {
                ya = a;
                yb = b;
            }
            if (yb > ya) {
                b0 = bs[zb * g.nbY + ya];
                cnt = (int)(bs[zb * g.nbY + yb] - b0);
            }
        }
    };
    // wave 0: publish the runs (inclusive scan of the sizes)
    auto runs_store = [&](int nruns, int64_t b0, int cnt, float am) {
        int x = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane < nruns) {
            rbeg[lane] = b0;
            rpre[lane + 1] = x;
        }
        if (lane == 0) {
            rpre[0] = 0;
            s_amax = am;
        }
    };

    TileSeq seq(ntiles_all);
    if (seq.cur >= seq.end) return;
    int t, y0, z0;
    decode(seq.cur, t, y0, z0);
    int nruns = 0;
    if (wv == 0) {
        int64_t b0;
        int cnt;
        runs_load(t, y0, z0, nruns, b0, cnt);
        runs_store(nruns, b0, cnt, amax ? amax[t] : 1.0f);
    }
    __syncthreads();

    // candidate fetch (runs of the tile currently published)
    PtRec<R> pr;
    C cv;
    bool have = false;
    int e = 0, total = 0;
    auto fetch = [&](int v, int tt) {
        have = v < total;
        if (have) {
            while (rpre[e + 1] <= v) ++e;
            const int64_t p = rbeg[e] + (v - rpre[e]);
            pr = srec[p];
            cv = cs[(int64_t)tt * c_tstride + p];
        }
    };
    // number of runs of a tile (wave 0 knows it from runs_load; every
    // thread needs it to read the tile's candidate total rpre[nruns])
    auto nruns_of = [&](int yy0, int zz0) {
        int nzb = 1;
        if (g.dim >= 3) {
            const int zlo = floordiv(zz0 - (w - 1), g.BZ);
            nzb = floordiv(zz0 + bz - 1, g.BZ) - zlo + 1;
            if (nzb >= g.nbZ) nzb = g.nbZ;
        }
        (void)yy0;
        return 2 * nzb;
    };
    total = rpre[nruns_of(y0, z0)];
    e = 0;
    fetch(wv * 64 + lane, t);

    for (;;) {
        // ---- tile (t, y0, z0): zero the accumulator planes
        {
            uint4 *a4 = reinterpret_cast<uint4 *>(smem_raw);
            for (int i = threadIdx.x; i < plane; i += NT) a4[i] = make_uint4(0u, 0u, 0u, 0u);
        }
        const float am = s_amax;
        __syncthreads();
        const int next = seq.cur + seq.step;
        const bool has_next = next < seq.end;
        int tn = 0, y0n = 0, z0n = 0;
        if (has_next) decode(next, tn, y0n, z0n);
        // wave 0: the next tile's runs are in flight during this tile's spread
        int nruns_n = 0, cnt_n = 0;
        int64_t b0_n = 0;
        float am_n = 1.0f;
        if (wv == 0 && has_next) {
            runs_load(tn, y0n, z0n, nruns_n, b0_n, cnt_n);
            if (amax) am_n = amax[tn];
        }

        // fixed-point scale 2^q: total * amax * 2^q < 2^45
        double fscale = 1.0, finv = 1.0;
        bool fbad = false;
        if (FIX) {
            fbad = !(am < 3.0e38f);
            const double bound = (double)max(total, 1) * (double)(am > 0.0f && !fbad ? am : 1.0f);
            const int q = 44 - ilogb(bound);
            fscale = ldexp(1.0, q);
            finv = ldexp(1.0, -q);
        }

        // ---- spread: every wave walks candidates v = wv*64 + lane + k*NT
        for (int v0 = wv * 64; v0 < total; v0 += NT) {
            const PtRec<R> cur = pr;
            const C cc = cv;
            const bool ok = have;
            fetch(v0 + NT + lane, t);
            if (!ok) continue;
            int fy = 0, ny = 1, my = 0, fz = 0, nz = 1, mz = 0;
            if (g.dim >= 2) overlap(cur.ly, y0, by, w, g.n2, fy, ny, my);
            if (g.dim >= 3) overlap(cur.lz, z0, bz, w, g.n3, fz, nz, mz);
            if (ny <= 0 || nz <= 0) continue;
            double kxd[W];
            R kx[W];
#pragma unroll
            for (int m = 0; m < W; ++m) {
                kx[m] = es_kernel(cur.dx + (R)m, beta, ic);
                kxd[m] = (double)kx[m] * fscale;
            }
            R ky[MAX_TILE_B];
#pragma unroll
            for (int iy = 0; iy < MAX_TILE_B; ++iy) {
                ky[iy] = (R)0;
                if (iy < by)
                    ky[iy] = (g.dim >= 2) ? ((iy < ny) ? es_kernel(cur.dy + (R)(my + iy), beta, ic) : (R)0) : (R)1;
            }
            for (int iz = 0; iz < nz; ++iz) {
                const R wz = (g.dim >= 3) ? es_kernel(cur.dz + (R)(mz + iz), beta, ic) : (R)1;
                const int rbase = ((fz + iz) * by + fy) * RS + cur.lx;
#pragma unroll
                for (int iy = 0; iy < MAX_TILE_B; ++iy) {
                    if (iy < by && iy < ny) {
                        const R wyz = ky[iy] * wz;
                        const int cell0 = rbase + iy * RS;
                        if (FIX) {
                            const double vr = (double)(cc.x * wyz), vi = (double)(cc.y * wyz);
                            unsigned long long *pr_ = accu + cell0, *pi_ = accu + plane + cell0;
#pragma unroll
                            for (int m = 0; m < W; ++m) {
                                const double dr = __builtin_fma(vr, kxd[m], FIX_MAGIC);
                                const double di = __builtin_fma(vi, kxd[m], FIX_MAGIC);
                                __hip_atomic_fetch_add(pr_ + m, (unsigned long long)__double_as_longlong(dr),
                                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                                __hip_atomic_fetch_add(pi_ + m, (unsigned long long)__double_as_longlong(di),
                                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                            }
                        } else {
                            const R vr = cc.x * wyz, vi = cc.y * wyz;
                            double *pr_ = accd + cell0, *pi_ = accd + plane + cell0;
#pragma unroll
                            for (int m = 0; m < W; ++m) {
                                __hip_atomic_fetch_add(pr_ + m, (double)(vr * kx[m]), __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                                __hip_atomic_fetch_add(pi_ + m, (double)(vi * kx[m]), __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP);
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();

        // ---- publish the next tile's runs (wave 0), fold the ghost cells
        if (wv == 0 && has_next) runs_store(nruns_n, b0_n, cnt_n, am_n);
        for (int i = threadIdx.x; i < 2 * rows * W; i += NT) {
            const int pl = i / (rows * W), rr = (i / W) % rows, x = i % W;
            const int a = pl * plane + rr * RS;
            if (FIX) accu[a + x] += accu[a + n1 + x];
            else accd[a + x] += accd[a + n1 + x];
        }
        __syncthreads();
        // the next tile's first candidates are in flight during this FFT
        const int total_n = has_next ? rpre[nruns_of(y0n, z0n)] : 0;
        {
            const int tot_cur = total;
            total = total_n;
            e = 0;
            fetch(wv * 64 + lane, tn);
            total = tot_cur;
        }

        // ---- x FFT: accumulator -> padded tile -> ... -> kept frequencies
        auto load_acc = [&](int row, int i) -> C {
            const int a = row * RS + i;
            if (FIX) {
                const long long br = (long long)(accu[a] << 16) >> 16;
                const long long bi = (long long)(accu[plane + a] << 16) >> 16;
                return cmk<C>(fbad ? (R)NAN : (R)((double)br * finv), fbad ? (R)NAN : (R)((double)bi * finv));
            } else {
                return cmk<C>((R)accd[a], (R)accd[plane + a]);
            }
        };
        auto load_tile = [&](int row, int i) -> C { return tile[row * RSF + padx<true>(i)]; };
        auto store_tile = [&](int row, int i, C v) { tile[row * RSF + padx<true>(i)] = v; };
        auto store_out = [&](int row, int f, C v) {
            int kk;
            if (f < N1 - h1) kk = f + h1;
            else if (f >= n1 - h1) kk = f - (n1 - h1);
            else return;
            v = cscale(v, phinv1[kk]);
            if (final_conj) v.y = -v.y;
            const int rz = (int)fdiv((uint32_t)row, dby), ry = row - rz * by;
            G1[(((int64_t)t * g.n3 + z0 + rz) * g.n2 + y0 + ry) * N1 + kk] = v;
        };
        {
            int Ns = 1;
            for (int s = 0; s < rp1.nst; ++s) {
                const int Rr = rp1.r[s];
                const bool first = (s == 0), last = (s == rp1.nst - 1);
                if (first && last)
                    stage_any<C, E, NT>(Rr, n1, rows, rp1.dnb[s], rp1.dns[s], Ns, rp1.tws[s], tw, load_acc, store_out,
                                        false);
                else if (first)
                    stage_any<C, E, NT>(Rr, n1, rows, rp1.dnb[s], rp1.dns[s], Ns, rp1.tws[s], tw, load_acc,
                                        store_tile, true);
                else if (last)
                    stage_any<C, E, NT>(Rr, n1, rows, rp1.dnb[s], rp1.dns[s], Ns, rp1.tws[s], tw, load_tile,
                                        store_out, false);
                else
                    stage_any<C, E, NT>(Rr, n1, rows, rp1.dnb[s], rp1.dns[s], Ns, rp1.tws[s], tw, load_tile,
                                        store_tile, true);
                if (!last) __syncthreads();
                Ns *= Rr;
            }
        }
        if (!has_next) break;
        __syncthreads();  // every read of the tile done before it is zeroed
        seq.cur = next;
        t = tn;
        y0 = y0n;
        z0 = z0n;
        total = total_n;
    }
}

template <typename R> int spread_tile_elems() { return SpreadCfg<R>::E * SpreadCfg<R>::NT; }
int spread_max_runs() { return MAX_RUNS; }
int spread_max_tile_b() { return MAX_TILE_B; }
template <typename R> size_t spread_lds(int rows, int n1, int w) { return spread_lds_bytes<R>(rows, n1, w); }

template <typename R, int W>
static void launch_w(const Geo &g, const RadixPlan &rp1, int ntr, const PtRec<R> *srec,
                     const typename cplx<R>::T *cs, int64_t c_tstride, const int64_t *binstart, int bin_tstride,
                     const typename cplx<R>::T *tw1, const R *phinv1, typename cplx<R>::T *G1, int final_conj,
                     const float *amax, hipStream_t st)
{
    constexpr int E = SpreadCfg<R>::E, NT = SpreadCfg<R>::NT;
    const int by = (g.dim >= 2) ? g.by : 1, bz = (g.dim >= 3) ? g.bz : 1;
    const int nty = (g.dim >= 2) ? g.n2 / by : 1;
    const int ntiles = nty * ((g.dim >= 3) ? g.n3 / bz : 1);
    const size_t lds = spread_lds_bytes<R>(by * bz, g.n1, W);
    const void *fn = reinterpret_cast<const void *>(&k_spread_xfft<R, E, NT, W>);
    static int resident = 0;  // workgroups per CU at this kernel's LDS size
    static size_t resident_lds = 0;
    static int ncu = 0;
    if (resident == 0 || resident_lds != lds) {
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 156 * 1024);
        int dev = 0, nb = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, NT, lds) != hipSuccess || nb < 1) nb = 1;
        resident = nb;
        resident_lds = lds;
    }
    const int ntot = ntiles * ntr;
    int grid = std::max(1, ncu) * resident;
    grid = ((grid + 7) / 8) * 8;
    grid = std::min(grid, ((ntot + 7) / 8) * 8);
    hipLaunchKernelGGL((k_spread_xfft<R, E, NT, W>), dim3(grid), dim3(NT), lds, st, g, rp1, srec, cs, c_tstride,
                       binstart, bin_tstride, tw1, phinv1, G1, final_conj, ntr, ntot, make_fdiv((uint32_t)ntr),
                       make_fdiv((uint32_t)nty), make_fdiv((uint32_t)by), amax);
}

template <typename R>
void launch_spread_xfft(const Geo &g, const RadixPlan &rp1, int ntr, const PtRec<R> *srec,
                        const typename cplx<R>::T *cs, int64_t c_tstride, const int64_t *binstart, int bin_tstride,
                        const typename cplx<R>::T *tw1, const R *phinv1, typename cplx<R>::T *G1, int final_conj,
                        const float *amax, hipStream_t st)
{
    if (ntr <= 0) return;
#define MPSQ_SPREAD_W(WV)                                                                                             \
    case WV:                                                                                                          \
        launch_w<R, WV>(g, rp1, ntr, srec, cs, c_tstride, binstart, bin_tstride, tw1, phinv1, G1, final_conj, amax,  \
                        st);                                                                                          \
        break;
    switch (g.w) {
        MPSQ_SPREAD_W(2)
        MPSQ_SPREAD_W(3)
        MPSQ_SPREAD_W(4)
        MPSQ_SPREAD_W(5)
        MPSQ_SPREAD_W(6)
        MPSQ_SPREAD_W(7)
        MPSQ_SPREAD_W(8)
        MPSQ_SPREAD_W(9)
        MPSQ_SPREAD_W(10)
        MPSQ_SPREAD_W(11)
        MPSQ_SPREAD_W(12)
        MPSQ_SPREAD_W(13)
        MPSQ_SPREAD_W(14)
        MPSQ_SPREAD_W(15)
        MPSQ_SPREAD_W(16)
    default: break;
    }
#undef MPSQ_SPREAD_W
}

#define MPSQ_SPREAD_INST(R)                                                                                           \
    template int spread_tile_elems<R>();                                                                              \
    template size_t spread_lds<R>(int, int, int);                                                                     \
    template void launch_spread_xfft<R>(const Geo &, const RadixPlan &, int, const PtRec<R> *, const cplx<R>::T *,   \
                                        int64_t, const int64_t *, int, const cplx<R>::T *, const R *, cplx<R>::T *,   \
                                        int, const float *, hipStream_t);

MPSQ_SPREAD_INST(float)
MPSQ_SPREAD_INST(double)

}  // namespace mpsq
