// spread_kernel.hip -- exponential-of-semicircle spreading fused with the x FFT
// (the spread / FFT / deconvolve steps of a type-1 NUFFT, first dimension).
//
// One workgroup owns a tile of the periodic fine grid that spans the whole x
// extent and (by x bz) rows in (y, z).  It
//   1. finds the sort bins whose points can reach the tile (one thread per
//      bin, block scan of the bin sizes);
//   2. stages CH candidate points at a time into LDS: first kernel cell in
//      x, the rows of the tile its stencil covers, the kernel values
//      phi(d + m) of those rows and of the w x cells, and its strength; the
//      next chunk's records are loaded while the current one is spread;
//   3. flattens the staged work into (point, row) items -- every item is the
//      same w-cell update, so the lanes of a wave stay converged; each
//      thread takes a contiguous run of items -- and adds the cells with
//      FP64 LDS atomics.  gfx950 executes FP32 LDS atomic adds lane by lane
//      (~190 cycles per wave instruction, measured) but FP64 ones at close
//      to the plain LDS rate (~10-20 cycles), so the accumulator tile is FP64
//      in both precision forms.  Rows are padded by one cell so that items
//      with the same x in different rows fall in different banks;
//   4. rounds the tile to R in place, runs the x FFT of its rows in LDS
//      (twiddles copied to LDS), keeps the N1 central frequencies, divides by
//      phi_hat and writes them: the only HBM write of the pass.
#include "mpsq_kernels.h"

namespace mpsq {

template <typename R> struct SpreadCfg;
template <> struct SpreadCfg<float> { static constexpr int E = 8, NT = 512, CH = 128; };
template <> struct SpreadCfg<double> { static constexpr int E = 8, NT = 512, CH = 64; };

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

// exclusive scan of one int per thread over an NT-thread block; *total gets
// the block sum on every thread
template <int NT> __device__ __forceinline__ int block_scan(int v, int *wsum, int *total)
{
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        if (k < wv) base += wsum[k];
        tot += wsum[k];
    }
    *total = tot;
    __syncthreads();
    return base + x - v;
}

template <typename R> size_t spread_stage_bytes(int n1, int w, int by, int bz, int CH)
{
    size_t b = (size_t)CH * sizeof(typename cplx<R>::T);   // strengths
    b += (size_t)CH * (w + by + bz) * sizeof(R);           // kernel values
    b += (size_t)CH * 8 + (size_t)(CH + 1) * 4;            // lx, meta, item prefix
    const size_t tw = (size_t)n1 * sizeof(typename cplx<R>::T);  // reused for the FFT twiddles
    return ((b > tw ? b : tw) + 15) / 16 * 16;
}

template <typename R> size_t spread_lds_bytes(int rows, int n1, int w, int by, int bz, int CH)
{
    return (size_t)rows * (n1 + 1) * 16 + spread_stage_bytes<R>(n1, w, by, bz, CH);
}

template <typename R, int E, int NT, int CH>
__global__ __launch_bounds__(NT) void k_spread_xfft(Geo g, RadixPlan rp1, const PtRec<R> *__restrict__ srec,
                                                    const typename cplx<R>::T *__restrict__ cs, int64_t c_tstride,
                                                    const int64_t *__restrict__ binstart, int bin_tstride,
                                                    const typename cplx<R>::T *__restrict__ tw1,
                                                    const R *__restrict__ phinv1, typename cplx<R>::T *__restrict__ G1,
                                                    int final_conj)
{
    static_assert(MAX_NB_BINS <= NT, "one thread per visited bin");
    using C = typename cplx<R>::T;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int64_t rstart[MAX_NB_BINS];
    __shared__ int rpref[MAX_NB_BINS + 1];
    __shared__ int wsum[NT / 64];

    const int t = blockIdx.y;
    const int n1 = g.n1, w = g.w, RS = g.n1 + 1;
    const int by = (g.dim >= 2) ? g.by : 1;
    const int bz = (g.dim >= 3) ? g.bz : 1;
    const int rows = by * bz;
    const int nty = (g.dim >= 2) ? g.n2 / by : 1;
    const int ty = blockIdx.x % nty, tz = blockIdx.x / nty;
    const int y0 = ty * by, z0 = tz * bz;
    const R beta = (R)g.beta, ic = (R)g.ic;

    double2 *acc = reinterpret_cast<double2 *>(smem_raw);
    unsigned char *stage = smem_raw + (size_t)rows * RS * 16;
    unsigned char *sp = stage;
    C *st_c = reinterpret_cast<C *>(sp);
    sp += CH * sizeof(C);
    R *st_kx = reinterpret_cast<R *>(sp);
    sp += CH * w * sizeof(R);
    R *st_ky = reinterpret_cast<R *>(sp);
    sp += CH * by * sizeof(R);
    R *st_kz = reinterpret_cast<R *>(sp);
    sp += CH * bz * sizeof(R);
    int *st_lx = reinterpret_cast<int *>(sp);
    sp += CH * 4;
    int *st_meta = reinterpret_cast<int *>(sp);
    sp += CH * 4;
    int *st_pref = reinterpret_cast<int *>(sp);

    for (int i = threadIdx.x; i < rows * RS; i += NT) acc[i] = make_double2(0.0, 0.0);

    // ---- 1. sort bins that can reach the tile (one thread per bin)
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
    const int nb = min(nyb * nzb, MAX_NB_BINS);
    {
        const int64_t *bs = binstart + (int64_t)t * bin_tstride * (g.nbins + 1);
        int cntb = 0;
        int64_t s0 = 0;
        if ((int)threadIdx.x < nb) {
            const int iz = threadIdx.x / nyb, iy = threadIdx.x % nyb;
            const int b = (g.dim >= 3 ? posmod(zlo + iz, g.nbZ) * g.nbY : 0) + (g.dim >= 2 ? posmod(ylo + iy, g.nbY) : 0);
            s0 = bs[b];
            cntb = (int)(bs[b + 1] - s0);
        }
        int totb;
        const int pre = block_scan<NT>(cntb, wsum, &totb);
        if ((int)threadIdx.x < nb) {
            rstart[threadIdx.x] = s0;
            rpref[threadIdx.x] = pre;
        }
        if (threadIdx.x == 0) rpref[nb] = totb;
        __syncthreads();
    }
    const int total = rpref[nb];
    const C *cst = cs + (int64_t)t * c_tstride;

    // candidate v -> its sorted point (largest bin k with rpref[k] <= v)
    PtRec<R> pr;
    C cv;
    bool have = false;
    auto fetch = [&](int base) {
        const int v = base + (int)threadIdx.x;
        have = ((int)threadIdx.x < CH) && (v < total);
        if (have) {
            int lo = 0, hi = nb;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (rpref[mid] <= v) lo = mid;
                else hi = mid;
            }
            const int64_t p = rstart[lo] + (v - rpref[lo]);
            pr = srec[p];
            cv = cst[p];
        }
    };
    fetch(0);

    for (int base = 0; base < total; base += CH) {
        // ---- 2. stage this chunk from registers
        int cnt = 0;
        if (have) {
            int fy = 0, ny = 1, my = 0, fz = 0, nz = 1, mz = 0;
            if (g.dim >= 2) {
                int oy = pr.ly - y0;
                if (oy < 0) oy += g.n2;
                fy = -1;
                ny = 0;
                for (int m = 0; m < w; ++m) {
                    int q = oy + m;
                    if (q >= g.n2) q -= g.n2;
                    if (q < by) {
                        if (fy < 0) {
                            fy = q;
                            my = m;
                        }
                        ++ny;
                    }
                }
            }
            if (g.dim >= 3) {
                int oz = pr.lz - z0;
                if (oz < 0) oz += g.n3;
                fz = -1;
                nz = 0;
                for (int m = 0; m < w; ++m) {
                    int q = oz + m;
                    if (q >= g.n3) q -= g.n3;
                    if (q < bz) {
                        if (fz < 0) {
                            fz = q;
                            mz = m;
                        }
                        ++nz;
                    }
                }
            }
            if (ny > 0 && nz > 0) {
                cnt = ny * nz;
                const int s = threadIdx.x;
                st_lx[s] = pr.lx;
                st_meta[s] = fy | (ny << 8) | (fz << 16) | (nz << 24);
                st_c[s] = cv;
                for (int m = 0; m < w; ++m) st_kx[s * w + m] = es_kernel(pr.dx + (R)m, beta, ic);
                for (int i = 0; i < ny; ++i)
                    st_ky[s * by + i] = (g.dim >= 2) ? es_kernel(pr.dy + (R)(my + i), beta, ic) : (R)1;
                for (int i = 0; i < nz; ++i)
                    st_kz[s * bz + i] = (g.dim >= 3) ? es_kernel(pr.dz + (R)(mz + i), beta, ic) : (R)1;
            }
        }
        // the next chunk's records are in flight while this one is spread
        fetch(base + CH);
        int nitems;
        const int pre = block_scan<NT>(cnt, wsum, &nitems);
        if ((int)threadIdx.x < CH) st_pref[threadIdx.x] = pre;
        if (threadIdx.x == 0) st_pref[CH] = nitems;
        __syncthreads();
        // ---- 3. (point, row) items: thread k owns items [k*q, (k+1)*q)
        const int q = (nitems + NT - 1) / NT;
        int it = threadIdx.x * q;
        const int it_end = min(it + q, nitems);
        if (it < it_end) {
            int lo = 0, hi = CH;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (st_pref[mid] <= it) lo = mid;
                else hi = mid;
            }
            int s = lo;
            for (; it < it_end; ++it) {
                while (st_pref[s + 1] <= it) ++s;
                const int r = it - st_pref[s];
                const int meta = st_meta[s];
                const int fy = meta & 255, ny = (meta >> 8) & 255, fz = (meta >> 16) & 255;
                const int iy = r % ny, iz = r / ny;
                const R wyz = st_ky[s * by + iy] * st_kz[s * bz + iz];
                const C c = st_c[s];
                const double vr = (double)(c.x * wyz), vi = (double)(c.y * wyz);
                double2 *row = acc + ((fz + iz) * by + fy + iy) * RS;
                int x = st_lx[s];
                const R *kx = st_kx + s * w;
                for (int m = 0; m < w; ++m) {
                    const double km = (double)kx[m];
                    double *cell = reinterpret_cast<double *>(row + x);
                    __hip_atomic_fetch_add(cell, vr * km, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    __hip_atomic_fetch_add(cell + 1, vi * km, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (++x == n1) x = 0;
                }
            }
        }
        __syncthreads();
    }

    // ---- 4. twiddles to LDS, round to R in place, x FFT, truncate + deconvolve
    C *tw = reinterpret_cast<C *>(stage);
    for (int i = threadIdx.x; i < n1; i += NT) tw[i] = tw1[i];
    C *tile;
    if (sizeof(R) == 8) {
        tile = reinterpret_cast<C *>(acc);
    } else {
        constexpr int EP = E + 1;  // padded rows: rows*(n1+1) <= E*NT + 128
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

template <typename R>
void launch_spread_xfft(const Geo &g, const RadixPlan &rp1, int ntr, const PtRec<R> *srec,
                        const typename cplx<R>::T *cs, int64_t c_tstride, const int64_t *binstart, int bin_tstride,
                        const typename cplx<R>::T *tw1, const R *phinv1, typename cplx<R>::T *G1, int final_conj,
                        hipStream_t st)
{
    constexpr int E = SpreadCfg<R>::E, NT = SpreadCfg<R>::NT, CH = SpreadCfg<R>::CH;
    const int by = (g.dim >= 2) ? g.by : 1, bz = (g.dim >= 3) ? g.bz : 1;
    const int ntiles = ((g.dim >= 2) ? g.n2 / by : 1) * ((g.dim >= 3) ? g.n3 / bz : 1);
    const size_t lds = spread_lds_bytes<R>(by * bz, g.n1, g.w, by, bz, CH);
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_spread_xfft<R, E, NT, CH>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024);
        attr = true;
    }
    hipLaunchKernelGGL((k_spread_xfft<R, E, NT, CH>), dim3(ntiles, ntr), dim3(NT), lds, st, g, rp1, srec, cs,
                       c_tstride, binstart, bin_tstride, tw1, phinv1, G1, final_conj);
}

#define MPSQ_SPREAD_INST(R)                                                                                           \
    template int spread_tile_elems<R>();                                                                              \
    template void launch_spread_xfft<R>(const Geo &, const RadixPlan &, int, const PtRec<R> *, const cplx<R>::T *,   \
                                        int64_t, const int64_t *, int, const cplx<R>::T *, const R *, cplx<R>::T *,   \
                                        int, hipStream_t);

MPSQ_SPREAD_INST(float)
MPSQ_SPREAD_INST(double)

}  // namespace mpsq
