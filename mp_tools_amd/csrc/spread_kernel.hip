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
#include <map>
#include <mutex>

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

// Persistent spread + x FFT over the tiles of a job table.
//
// Tile ids: job-major; inside a job, id - tile0 = (tile position) * ntr +
// transform, so the tiles taken one after another cover every transform of
// a few neighbouring positions (their candidates overlap and come from the
// XCD's L2), and the species (3x different point counts) are spread evenly.
// Workgroup b runs on XCD x = b % 8 (round-robin dispatch of the resident
// grid) and takes the next tile of its XCD's range [xr[x], xr[x+1]) from the
// counter wc[x]; the host cuts the ranges at equal estimated cost.  While a
// tile is spread, wave 0 already holds the next tile's id and loads its
// candidate runs; while it is transformed, every lane already loads its first
// candidate of the next tile.
template <typename R, int E, int NT, int W>
__global__ __launch_bounds__(NT, 4) void k_spread_xfft(const SpreadJob<R> *__restrict__ jobs,
                                                       const int *__restrict__ tile0s, int njobs,
                                                       const int *__restrict__ xr, int *__restrict__ wc)
{
    using C = typename cplx<R>::T;
    constexpr bool FIX = (sizeof(R) == 4);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int64_t rbeg[2][MAX_RUNS];  // candidate runs, double-buffered by tile parity
    __shared__ int rpre[2][MAX_RUNS + 1];
    __shared__ float s_amax[2];
    __shared__ int s_next;

    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int xcd = blockIdx.x & 7;
    const int xbeg = xr[xcd], xend = xr[xcd + 1];
    unsigned long long *accu = reinterpret_cast<unsigned long long *>(smem_raw);
    double *accd = reinterpret_cast<double *>(smem_raw);
    C *tile = reinterpret_cast<C *>(smem_raw);

    // geometry of a job's tile: transform t, first row y0, first plane z0
    auto decode = [&](const SpreadJob<R> &J, int id, int &t, int &y0, int &z0) {
        const int loc = id - J.tile0;
        const int ti = (int)fdiv((uint32_t)loc, J.dntr);
        t = loc - ti * J.ntr;
        const int by = (J.g.dim >= 2) ? J.g.by : 1, bz = (J.g.dim >= 3) ? J.g.bz : 1;
        const int tz = (int)fdiv((uint32_t)ti, J.dnty);
        y0 = (ti - tz * ((J.g.dim >= 2) ? J.g.n2 / by : 1)) * by;
        z0 = tz * bz;
    };
    // number of candidate runs of a tile: two per visited z-bin row
    auto nruns_of = [&](const SpreadJob<R> &J, int z0) {
        const Geo &g = J.g;
        int nzb = 1;
        if (g.dim >= 3) {
            const int zlo = floordiv(z0 - (g.w - 1), g.BZ);
            nzb = floordiv(z0 + g.bz - 1, g.BZ) - zlo + 1;
            if (nzb >= g.nbZ) nzb = g.nbZ;
        }
        return 2 * nzb;
    };
    // wave 0: start and size of run `lane` of a tile (registers)
    auto runs_load = [&](const SpreadJob<R> &J, int t, int y0, int z0, int &nruns, int64_t &b0, int &cnt) {
        const Geo &g = J.g;
        const int w = g.w, by = (g.dim >= 2) ? g.by : 1, bz = (g.dim >= 3) ? g.bz : 1;
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
            const int64_t *bs = J.binstart + (int64_t)t * J.bin_tstride * (g.nbins + 1);
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
    // wave 0: publish the runs (inclusive scan of the sizes) in buffer pb
    auto runs_store = [&](int pb, int nruns, int64_t b0, int cnt, float am) {
        int x = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane < nruns) {
            rbeg[pb][lane] = b0;
            rpre[pb][lane + 1] = x;
        }
        if (lane == 0) {
            rpre[pb][0] = 0;
            s_amax[pb] = am;
        }
    };
    auto load_tw = [&](const SpreadJob<R> &J) {
        C *tw = reinterpret_cast<C *>(smem_raw + J.lds_tw);
        for (int i = threadIdx.x; i < J.g.n1; i += NT) tw[i] = J.tw1[i];
    };

    // ---- first tile
    if (threadIdx.x == 0) s_next = xbeg + atomicAdd(wc + xcd, 1);
    __syncthreads();
    int cur = s_next;
    if (cur >= xend) return;
    int j = job_of(tile0s, njobs, cur);
    int t, y0, z0;
    decode(jobs[j], cur, t, y0, z0);
    load_tw(jobs[j]);
    int pb = 0;  // run buffer of the current tile
    if (wv == 0) {
        int nr, cnt;
        int64_t b0;
        runs_load(jobs[j], t, y0, z0, nr, b0, cnt);
        runs_store(pb, nr, b0, cnt, jobs[j].amax ? jobs[j].amax[t] : 1.0f);
    }
    __syncthreads();

    // candidate fetch from the runs currently published
    PtRec<R> pr;
    C cv;
    bool have = false;
    int e = 0, total = 0;
    auto fetch = [&](const SpreadJob<R> &J, int b, int v, int tt) {
        have = v < total;
        if (have) {
            while (rpre[b][e + 1] <= v) ++e;
            const int64_t p = rbeg[b][e] + (v - rpre[b][e]);
            pr = J.srec[p];
            cv = J.cs[(int64_t)tt * J.c_tstride + p];
        }
    };
    total = rpre[pb][nruns_of(jobs[j], z0)];
    fetch(jobs[j], pb, wv * 64 + lane, t);

    for (;;) {
        const SpreadJob<R> &J = jobs[j];
        const Geo &g = J.g;
        const int n1 = g.n1, w = g.w;
        const int RS = spread_row_stride(n1, W);
        const int RSF = padded_len(n1);
        const int by = (g.dim >= 2) ? g.by : 1;
        const int bz = (g.dim >= 3) ? g.bz : 1;
        const int rows = by * bz;
        const int plane = rows * RS;
        const R beta = (R)g.beta, ic = (R)g.ic;
        const C *tw = reinterpret_cast<const C *>(smem_raw + J.lds_tw);

        // ---- zero the accumulator planes; take the next tile's id
        {
            uint4 *a4 = reinterpret_cast<uint4 *>(smem_raw);
            for (int i = threadIdx.x; i < plane; i += NT) a4[i] = make_uint4(0u, 0u, 0u, 0u);
        }
        const float am = s_amax[pb];
        __syncthreads();
        if (wv == 0) {
            // the next tile's id and runs (other buffer) while this one is spread
            int nxt = 0;
            if (lane == 0) nxt = xbeg + atomicAdd(wc + xcd, 1);
            nxt = __shfl(nxt, 0, 64);
            if (lane == 0) s_next = nxt;
            if (nxt < xend) {
                int jn, tn, y0n, z0n, nr, cnt;
                int64_t b0;
                jn = job_of(tile0s, njobs, nxt);
                decode(jobs[jn], nxt, tn, y0n, z0n);
                runs_load(jobs[jn], tn, y0n, z0n, nr, b0, cnt);
                runs_store(pb ^ 1, nr, b0, cnt, jobs[jn].amax ? jobs[jn].amax[tn] : 1.0f);
            }
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
            const PtRec<R> cur_p = pr;
            const C cc = cv;
            const bool ok = have;
            fetch(J, pb, v0 + NT + lane, t);
            if (!ok) continue;
            int fy = 0, ny = 1, my = 0, fz = 0, nz = 1, mz = 0;
            if (g.dim >= 2) overlap(cur_p.ly, y0, by, w, g.n2, fy, ny, my);
            if (g.dim >= 3) overlap(cur_p.lz, z0, bz, w, g.n3, fz, nz, mz);
            if (ny <= 0 || nz <= 0) continue;
            double kxd[W];
            R kx[W];
#pragma unroll
            for (int m = 0; m < W; ++m) {
                kx[m] = es_kernel(cur_p.dx + (R)m, beta, ic);
                kxd[m] = (double)kx[m] * fscale;
            }
            R ky[MAX_TILE_B];
#pragma unroll
            for (int iy = 0; iy < MAX_TILE_B; ++iy) {
                ky[iy] = (R)0;
                if (iy < by)
                    ky[iy] = (g.dim >= 2) ? ((iy < ny) ? es_kernel(cur_p.dy + (R)(my + iy), beta, ic) : (R)0) : (R)1;
            }
            for (int iz = 0; iz < nz; ++iz) {
                const R wz = (g.dim >= 3) ? es_kernel(cur_p.dz + (R)(mz + iz), beta, ic) : (R)1;
                const int rbase = ((fz + iz) * by + fy) * RS + cur_p.lx;
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

        // ---- fold the ghost cells
        for (int i = threadIdx.x; i < 2 * rows * W; i += NT) {
            const int pl = i / (rows * W), rr = (i / W) % rows, x = i % W;
            const int a = pl * plane + rr * RS;
            if (FIX) accu[a + x] += accu[a + n1 + x];
            else accd[a + x] += accd[a + n1 + x];
        }
        __syncthreads();
        // the next tile's first candidates are in flight during this FFT
        const int nxt = s_next;
        const bool has_next = nxt < xend;
        int jn = j, tn = 0, y0n = 0, z0n = 0, total_n = 0;
        if (has_next) {
            jn = job_of(tile0s, njobs, nxt);
            decode(jobs[jn], nxt, tn, y0n, z0n);
            total_n = rpre[pb ^ 1][nruns_of(jobs[jn], z0n)];
            const int tot_cur = total;
            total = total_n;
            e = 0;
            fetch(jobs[jn], pb ^ 1, wv * 64 + lane, tn);
            total = tot_cur;
        }

        // ---- x FFT: accumulator -> padded tile -> ... -> kept frequencies
        const int N1 = g.N1, h1 = N1 / 2;
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
            v = cscale(v, J.phinv1[kk]);
            if (J.final_conj) v.y = -v.y;
            const int rz = (int)fdiv((uint32_t)row, J.dby), ry = row - rz * by;
            J.G1[(((int64_t)t * g.n3 + z0 + rz) * g.n2 + y0 + ry) * N1 + kk] = v;
        };
        {
            const RadixPlan &rp1 = J.rp1;
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
        __syncthreads();  // every read of the tile (and the twiddles) done
        if (jn != j) load_tw(jobs[jn]);
        cur = nxt;
        pb ^= 1;
        j = jn;
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

__global__ void k_zero_i32(int *__restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

void launch_zero_i32(int *p, int n, hipStream_t st)
{
    if (n > 0) hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
}

template <typename R, int W>
static void launch_w(const SpreadJob<R> *jobs, const int *tile0s, int njobs, const int *xr, int *wc, size_t lds,
                     hipStream_t st)
{
    constexpr int E = SpreadCfg<R>::E, NT = SpreadCfg<R>::NT;
    const void *fn = reinterpret_cast<const void *>(&k_spread_xfft<R, E, NT, W>);
    static std::mutex mu;
    static std::map<size_t, int> resident_at;  // workgroups per CU at an LDS size
    static int ncu = 0;
    int resident;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = resident_at.find(lds);
        if (it == resident_at.end()) {
            (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 156 * 1024);
            int dev = 0, nb = 0;
            (void)hipGetDevice(&dev);
            (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, NT, lds) != hipSuccess || nb < 1) nb = 1;
            it = resident_at.emplace(lds, nb).first;
        }
        resident = it->second;
    }
    int grid = std::max(1, ncu) * resident;
    grid = ((grid + 7) / 8) * 8;
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, wc, 8);
    hipLaunchKernelGGL((k_spread_xfft<R, E, NT, W>), dim3(grid), dim3(NT), lds, st, jobs, tile0s, njobs, xr, wc);
}

template <typename R>
void launch_spread_jobs(const SpreadJob<R> *jobs, const int *tile0s, int njobs, int w, const int *xr, int *wc,
                        size_t lds, hipStream_t st)
{
    if (njobs <= 0) return;
#define MPSQ_SPREAD_W(WV)                                                                                             \
    case WV: launch_w<R, WV>(jobs, tile0s, njobs, xr, wc, lds, st); break;
    switch (w) {
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
    template void launch_spread_jobs<R>(const SpreadJob<R> *, const int *, int, int, const int *, int *, size_t,     \
                                        hipStream_t);

MPSQ_SPREAD_INST(float)
MPSQ_SPREAD_INST(double)

}  // namespace mpsq
