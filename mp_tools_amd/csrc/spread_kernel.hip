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
//      lane computes the rows of the tile its stencil covers, the w x-kernel
//      values (registers), and for each covered row the y/z kernel product,
//      and adds the w cells with FP64 LDS atomics.  gfx950 executes FP32 LDS
//      atomic adds lane by lane (~190 cycles per wave instruction, measured)
//      but FP64 ones at close to the plain LDS rate (~10-20 cycles), so the
//      accumulator tile is FP64 in both precision forms;
//   3. rounds the tile to R in place, runs the x FFT of its rows in LDS
//      (twiddles copied to LDS), keeps the N1 central frequencies, divides by
//      phi_hat and writes them: the only HBM write of the pass.
//
// Tiles are numbered so that each XCD (workgroups are dealt to the 8 XCDs
// round robin) runs a contiguous range of tiles: neighbouring tiles share
// most of their candidates, which then come from that XCD's L2.
#include "mpsq_kernels.h"

namespace mpsq {

template <typename R> struct SpreadCfg;
template <> struct SpreadCfg<float> { static constexpr int E = 8, NT = 512; };
template <> struct SpreadCfg<double> { static constexpr int E = 8, NT = 512; };

constexpr int MAX_RUNS = 64;  // candidate runs of one tile (two per visited z-bin row)

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

template <typename R> size_t spread_lds_bytes(int rows, int n1)
{
    return (size_t)rows * (n1 + 1) * 16 + (size_t)n1 * sizeof(typename cplx<R>::T);
}

template <typename R, int E, int NT, int W>
__global__ __launch_bounds__(NT) void k_spread_xfft(Geo g, RadixPlan rp1, const PtRec<R> *__restrict__ srec,
                                                    const typename cplx<R>::T *__restrict__ cs, int64_t c_tstride,
                                                    const int64_t *__restrict__ binstart, int bin_tstride,
                                                    const typename cplx<R>::T *__restrict__ tw1,
                                                    const R *__restrict__ phinv1, typename cplx<R>::T *__restrict__ G1,
                                                    int final_conj, int ntiles, int nblocks)
{
    using C = typename cplx<R>::T;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int64_t rbeg[MAX_RUNS];
    __shared__ int rpre[MAX_RUNS + 1];

    // XCD-aware numbering: block L runs on XCD L % 8; give XCD k the
    // contiguous id range [k * per, (k + 1) * per)
    const int L = blockIdx.x;
    const int per = nblocks >> 3;
    const int id = (L < per * 8) ? (L & 7) * per + (L >> 3) : L;
    const int t = id / ntiles, tile_i = id % ntiles;

    const int n1 = g.n1, w = g.w, RS = g.n1 + 1;
    const int by = (g.dim >= 2) ? g.by : 1;
    const int bz = (g.dim >= 3) ? g.bz : 1;
    const int rows = by * bz;
    const int nty = (g.dim >= 2) ? g.n2 / by : 1;
    const int ty = tile_i % nty, tz = tile_i / nty;
    const int y0 = ty * by, z0 = tz * bz;
    const R beta = (R)g.beta, ic = (R)g.ic;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;

    double2 *acc = reinterpret_cast<double2 *>(smem_raw);
    C *tw = reinterpret_cast<C *>(smem_raw + (size_t)rows * RS * 16);

    for (int i = threadIdx.x; i < rows * RS; i += NT) acc[i] = make_double2(0.0, 0.0);
    for (int i = threadIdx.x; i < n1; i += NT) tw[i] = tw1[i];

    // ---- 1. candidate runs (wave 0): entry e = 2 * (z-bin row k) + part
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
    const int nruns = 2 * nzb;  // <= MAX_RUNS (checked by the host)
    if (wv == 0) {
        const int64_t *bs = binstart + (int64_t)t * bin_tstride * (g.nbins + 1);
        int64_t b0 = 0;
        int cnt = 0;
        if (lane < nruns) {
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
        if (lane == 0) rpre[0] = 0;
    }
    __syncthreads();
    const int total = rpre[nruns];
    const C *cst = cs + (int64_t)t * c_tstride;

    // ---- 2. spread: every wave walks candidates v = wv*64 + lane + k*NT
    PtRec<R> pr;
    C cv;
    bool have = false;
    auto fetch = [&](int v) {
        have = v < total;
        if (have) {
            int e = 0;
            while (rpre[e + 1] <= v) ++e;
            const int64_t p = rbeg[e] + (v - rpre[e]);
            pr = srec[p];
            cv = cst[p];
        }
    };
    fetch(wv * 64 + lane);
    for (int v0 = wv * 64; v0 < total; v0 += NT) {
        const PtRec<R> cur = pr;
        const C cc = cv;
        const bool ok = have;
        fetch(v0 + NT + lane);
        if (!ok) continue;
        int fy = 0, ny = 1, my = 0, fz = 0, nz = 1, mz = 0;
        if (g.dim >= 2) overlap(cur.ly, y0, by, w, g.n2, fy, ny, my);
        if (g.dim >= 3) overlap(cur.lz, z0, bz, w, g.n3, fz, nz, mz);
        if (ny <= 0 || nz <= 0) continue;
        R kx[W];
#pragma unroll
        for (int m = 0; m < W; ++m) kx[m] = (m < w) ? es_kernel(cur.dx + (R)m, beta, ic) : (R)0;
        for (int iz = 0; iz < nz; ++iz) {
            const R wz = (g.dim >= 3) ? es_kernel(cur.dz + (R)(mz + iz), beta, ic) : (R)1;
            for (int iy = 0; iy < ny; ++iy) {
                const R wyz = ((g.dim >= 2) ? es_kernel(cur.dy + (R)(my + iy), beta, ic) : (R)1) * wz;
                const R vr = cc.x * wyz, vi = cc.y * wyz;
                double *row = reinterpret_cast<double *>(acc + ((fz + iz) * by + fy + iy) * RS);
                int x = cur.lx;
#pragma unroll
                for (int m = 0; m < W; ++m) {
                    if (m < w) {
                        double *cell = row + 2 * x;
                        __hip_atomic_fetch_add(cell, (double)(vr * kx[m]), __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                        __hip_atomic_fetch_add(cell + 1, (double)(vi * kx[m]), __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                        x = (x + 1 == n1) ? 0 : x + 1;
                    }
                }
            }
        }
    }
    __syncthreads();

    // ---- 3. round to R in place, x FFT, truncate + deconvolve
    C *tile;
    if (sizeof(R) == 8) {
        tile = reinterpret_cast<C *>(acc);
    } else {
        constexpr int EP = E + 1;  // padded rows: rows*(n1+1) <= E*NT + NT
        double2 tmp[EP];
#pragma unroll
        for (int k = 0; k < EP; ++k) {
            const int i = threadIdx.x + k * NT;
            if (i < rows * RS) tmp[k] = acc[i];
        }
        __syncthreads();
        tile = reinterpret_cast<C *>(smem_raw);
#pragma unroll
        for (int k = 0; k < EP; ++k) {
            const int i = threadIdx.x + k * NT;
            if (i < rows * RS) tile[i] = cmk<C>((R)tmp[k].x, (R)tmp[k].y);
        }
    }
    __syncthreads();
    lds_fft<C, E>(tile, rp1, rows, 1, RS, tw, false);

    const int N1 = g.N1;
    const int h1 = N1 / 2;
    for (int i = threadIdx.x; i < rows * N1; i += NT) {
        const int row = i / N1, kk = i % N1;
        int f = kk - h1;
        if (f < 0) f += n1;
        C v = cscale(tile[row * RS + f], phinv1[kk]);
        if (final_conj) v.y = -v.y;
        const int yy = y0 + row % by, zz = z0 + row / by;
        G1[(((int64_t)t * g.n3 + zz) * g.n2 + yy) * N1 + kk] = v;
    }
}

template <typename R> int spread_tile_elems() { return SpreadCfg<R>::E * SpreadCfg<R>::NT; }
int spread_max_runs() { return MAX_RUNS; }

template <typename R, int W>
static void launch_w(const Geo &g, const RadixPlan &rp1, int ntr, const PtRec<R> *srec,
                     const typename cplx<R>::T *cs, int64_t c_tstride, const int64_t *binstart, int bin_tstride,
                     const typename cplx<R>::T *tw1, const R *phinv1, typename cplx<R>::T *G1, int final_conj,
                     hipStream_t st)
{
    constexpr int E = SpreadCfg<R>::E, NT = SpreadCfg<R>::NT;
    const int by = (g.dim >= 2) ? g.by : 1, bz = (g.dim >= 3) ? g.bz : 1;
    const int ntiles = ((g.dim >= 2) ? g.n2 / by : 1) * ((g.dim >= 3) ? g.n3 / bz : 1);
    const size_t lds = spread_lds_bytes<R>(by * bz, g.n1);
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_spread_xfft<R, E, NT, W>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024);
        attr = true;
    }
    const int nblocks = ntiles * ntr;
    hipLaunchKernelGGL((k_spread_xfft<R, E, NT, W>), dim3(nblocks), dim3(NT), lds, st, g, rp1, srec, cs, c_tstride,
                       binstart, bin_tstride, tw1, phinv1, G1, final_conj, ntiles, nblocks);
}

template <typename R>
void launch_spread_xfft(const Geo &g, const RadixPlan &rp1, int ntr, const PtRec<R> *srec,
                        const typename cplx<R>::T *cs, int64_t c_tstride, const int64_t *binstart, int bin_tstride,
                        const typename cplx<R>::T *tw1, const R *phinv1, typename cplx<R>::T *G1, int final_conj,
                        hipStream_t st)
{
    if (ntr <= 0) return;
    if (g.w <= 8)
        launch_w<R, 8>(g, rp1, ntr, srec, cs, c_tstride, binstart, bin_tstride, tw1, phinv1, G1, final_conj, st);
    else
        launch_w<R, MAXW>(g, rp1, ntr, srec, cs, c_tstride, binstart, bin_tstride, tw1, phinv1, G1, final_conj, st);
}

#define MPSQ_SPREAD_INST(R)                                                                                           \
    template int spread_tile_elems<R>();                                                                              \
    template size_t spread_lds_bytes<R>(int, int);                                                                    \
    template void launch_spread_xfft<R>(const Geo &, const RadixPlan &, int, const PtRec<R> *, const cplx<R>::T *,   \
                                        int64_t, const int64_t *, int, const cplx<R>::T *, const R *, cplx<R>::T *,   \
                                        int, hipStream_t);

MPSQ_SPREAD_INST(float)
MPSQ_SPREAD_INST(double)

}  // namespace mpsq
