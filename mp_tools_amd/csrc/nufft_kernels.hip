// nufft_kernels.hip -- type-1 NUFFT kernels for gfx950.
//
// Pipeline for one batch of transforms (t = 0..ntr-1):
//   k_fold / k_stage_traj   FP64 fold of coordinates into the periodic fine
//                           grid: first kernel cell l (int) + offset d
//                           (FP32 in the FP32 form), bin key, strength.
//   k_sort_count            per-chunk bin histogram (LDS)
//   k_scan_*                exclusive scan of the (transform, bin, chunk)
//                           count matrix -> absolute destinations
//   k_sort_scatter          stable scatter: ranks inside a chunk are taken
//                           in point order, wave by wave, so the permutation
//                           equals a stable counting sort by key
//   k_spread_xfft           one workgroup per (y,z) pencil tile spanning the
//                           whole x extent: exponential-of-semicircle
//                           spreading of every point whose kernel touches
//                           the tile into LDS (ds_add atomics), then the x
//                           FFT of the tile's rows in LDS, truncation to the
//                           kept k1 and deconvolution; only N1 of n1 x
//                           frequencies ever reach HBM.
//   k_colfft                the y and z passes: column FFTs of LDS tiles,
//                           truncation, deconvolution; the last pass also
//                           applies species weights and sums the species
//                           of a frame (S(Q,w) combine, mp_sqom56.f90:1338-1356).
#include <algorithm>

#include "mpsq_device.h"
#include "mpsq_kernels.h"

namespace mpsq {

// ---------------------------------------------------------------------------
// FP64 fold (contraction off: matches the CPU restatement bit for bit)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void fold1(double xv, double s, double hw, int n, int &l, double &d, int &bad)
{
#pragma clang fp contract(off)
    const double u = xv * s;
    const double t = u - hw;
    const double lf = ceil(t);
    d = lf - u;
    long long lw = 0;
    if (isfinite(lf) && fabs(lf) < 9.0e18) lw = (long long)lf;
    else bad = 1;
    lw %= (long long)n;
    if (lw < 0) lw += n;
    l = (int)lw;
}

template <typename R>
__device__ __forceinline__ int bin_key(const Geo &g, int ly, int lz)
{
    if (g.dim == 1) return 0;
    if (g.dim == 2) return ly / g.BY;
    return (lz / g.BZ) * g.nbY + ly / g.BY;
}

template <typename R, typename Tin>
__global__ void k_fold(Geo g, int64_t M, const Tin *__restrict__ x, const Tin *__restrict__ y,
                       const Tin *__restrict__ z, PtRec<R> *__restrict__ rec, int *__restrict__ key,
                       int *__restrict__ bad_flag)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    int bad = 0;
    int l[3] = {0, 0, 0};
    double d[3] = {0.0, 0.0, 0.0};
    const Tin *co[3] = {x, y, z};
    const double sc[3] = {g.s1, g.s2, g.s3};
    const int nn[3] = {g.n1, g.n2, g.n3};
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (k < g.dim) {
            const double xv = (double)co[k][j];
            if (!isfinite(xv)) bad = 1;
            fold1(xv, sc[k], g.hw, nn[k], l[k], d[k], bad);
        }
    PtRec<R> r;
    r.lx = l[0];
    r.ly = l[1];
    r.lz = l[2];
    r.idx = (int)j;
    r.dx = (R)d[0];
    r.dy = (R)d[1];
    r.dz = (R)d[2];
    r.pad = (R)0;
    rec[j] = r;
    key[j] = bin_key<R>(g, l[1], l[2]);
    if (bad) atomicOr(bad_flag, 1);
}

// ---------------------------------------------------------------------------
// a2 + a3 staging of trajectory frames (mp_sqom56.f90:1104-1137):
//   point p of batch = frame b (of the batch) * n_tot + atom a, a species-major
//   transform t = b * nsp + species(a); transforms are contiguous in p.
//   keep = not a vacancy (0,0,0) and inside the selection window
//   x_d  = ((p0*e_d0 + p1*e_d1) + p2*e_d2) * tpe_d,  p = r - centre     (FP64)
//   c    = keep ? exp(-i 2pi q0.r) * [prod_d (1+cos x_d)/2] : 0
// Dropped atoms keep their slot with strength 0, which leaves every
// amplitude unchanged and keeps the per-transform point counts fixed.
// ---------------------------------------------------------------------------
// job j's blocks are [jobs[j].blk0, jobs[j+1].blk0): point p of the job is
// active atom a = p % mact of frame p / mact; active species k holds atoms
// [act_off[k], act_off[k+1]) of that numbering, taken from frame atoms
// act_src[k] + (a - act_off[k]) (species-major storage)
__device__ __forceinline__ int find_job(const int *__restrict__ blk0s, int njobs, int b)
{
    int lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (blk0s[mid] <= b) lo = mid;
        else hi = mid;
    }
    // uniform by construction (b is the block index); saying so keeps the
    // job's fields in scalar registers
    return __builtin_amdgcn_readfirstlane(lo);
}

template <typename R>
__global__ __launch_bounds__(256) void k_stage_traj(const StageJob *__restrict__ jobs, const int *__restrict__ blk0s,
                                                    int njobs, PtRec<R> *__restrict__ rec_all,
                                                    int *__restrict__ key_all,
                                                    typename cplx<R>::T *__restrict__ cstr_all,
                                                    int *__restrict__ bad_flag)
{
#pragma clang fp contract(off)
    using C = typename cplx<R>::T;
    const int jb = find_job(blk0s, njobs, blockIdx.x);
    const StageJob &J = jobs[jb];
    const Geo &g = J.g;
    const StageParams &sp = J.sp;
    const int64_t p = (int64_t)(blockIdx.x - J.blk0) * blockDim.x + threadIdx.x;
    if (p >= J.npts) return;
    const int64_t fb = p / J.mact, aa = p - fb * J.mact;
    int k = 0;
    while (k + 1 < J.nact && J.act_off[k + 1] <= aa) ++k;
    const int64_t a = J.act_src[k] + (aa - J.act_off[k]);
    const double wk = J.wst[k];
    const float4 r = J.pos[fb * J.nspan + a];
    const int4 *at_ind = J.at_ind;
    PtRec<R> *rec = rec_all + J.out0;
    int *key = key_all + J.out0;
    C *cstr = cstr_all + J.out0;
    bool keep = true;
    if (sp.vacancy_check) keep = (r.x != 0.0f) || (r.y != 0.0f) || (r.z != 0.0f);
    if (sp.select_mode == 1 && at_ind) {  // CELL window on cell indices
        const int4 ii = at_ind[a];
        const int iv[3] = {ii.x, ii.y, ii.z};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float v = (float)(iv[k] - sp.n_row[k] / 2 - 1);
            keep = keep && (v >= sp.pmin[k]) && (v <= sp.pmax[k]);
        }
    } else if (sp.select_mode == 2) {  // BULK box with margin, FP32 bounds
        const float rv[3] = {r.x, r.y, r.z};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float lo = sp.pmin[k] - sp.margin, hi = sp.pmax[k] + sp.margin;
            keep = keep && (rv[k] >= lo) && (rv[k] < hi);
        }
    }
    const double r0 = (double)r.x, r1 = (double)r.y, r2 = (double)r.z;
    const double p0 = r0 - sp.centre[0], p1 = r1 - sp.centre[1], p2 = r2 - sp.centre[2];
    int bad = 0;
    int l[3] = {0, 0, 0};
    double d[3] = {0.0, 0.0, 0.0};
    double xv[3] = {0.0, 0.0, 0.0};
    const double sc[3] = {g.s1, g.s2, g.s3};
    const int nn[3] = {g.n1, g.n2, g.n3};
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (k < g.dim) {
            const double e0 = p0 * sp.axes[3 * k];
            const double e1 = p1 * sp.axes[3 * k + 1];
            const double e2 = p2 * sp.axes[3 * k + 2];
            const double dot = (e0 + e1) + e2;
            xv[k] = dot * sp.tpe[k];
            fold1(xv[k], sc[k], g.hw, nn[k], l[k], d[k], bad);
        }
    double cr = 0.0, ci = 0.0;
    if (keep) {
        double sn = 0.0, cs = 1.0;  // q0 = 0 (uniform): the phase is 0 for every finite position
        if (sp.q0[0] != 0.0 || sp.q0[1] != 0.0 || sp.q0[2] != 0.0) {
            const double ph =
                6.283185307179586476925286766559 * ((sp.q0[0] * r0 + sp.q0[1] * r1) + sp.q0[2] * r2);
            sincos(ph, &sn, &cs);
        }
        cr = cs;
        ci = -sn;
        if (sp.window) {
            double wf = 1.0;
            for (int k = 0; k < g.dim; ++k) wf *= 0.5 * (1.0 + cos(xv[k]));
            cr *= wf;
            ci *= wf;
        }
        if (sp.conj) ci = -ci;
        cr *= wk;
        ci *= wk;
    }
    PtRec<R> rr;
    rr.lx = l[0];
    rr.ly = l[1];
    rr.lz = l[2];
    rr.idx = (int)a;
    rr.dx = (R)d[0];
    rr.dy = (R)d[1];
    rr.dz = (R)d[2];
    rr.pad = keep ? (R)1 : (R)0;  // selection flag (read by the engine's debug view only)
    rec[p] = rr;
    key[p] = bin_key<R>(g, l[1], l[2]);
    C cv;
    cv.x = (R)cr;
    cv.y = (R)ci;
    cstr[p] = cv;
    if (bad) atomicOr(bad_flag, 1);
}

// ---------------------------------------------------------------------------
// stable bin sort
// ---------------------------------------------------------------------------
__global__ void k_sort_count(const int *__restrict__ key, const SortChunk *__restrict__ chunks, int *__restrict__ H)
{
    extern __shared__ int hist[];
    const SortChunk ch = chunks[blockIdx.x];
    const int nbins = ch.nbins;
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) hist[b] = 0;
    __syncthreads();
    // a wave's lanes hold consecutive points, and points in storage order
    // (lattice rows) run in the same bin: one atomic per run of equal keys,
    // adding the run's length (integer counts: the result is the same)
    const int lane = threadIdx.x & 63;
    for (int64_t p0 = ch.p0; p0 < ch.p1; p0 += blockDim.x) {
        const int64_t p = p0 + threadIdx.x;
        const int k = p < ch.p1 ? key[p] : -1;
        const int prev = __shfl_up(k, 1, 64);
        const bool start = lane == 0 || k != prev;
        const unsigned long long starts = __ballot(start);
        const unsigned long long above = lane == 63 ? 0ull : starts & (~0ull << (lane + 1));
        const int len = (above ? __builtin_ctzll(above) : 64) - lane;
        if (start && k >= 0) __hip_atomic_fetch_add(&hist[k], len, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) H[ch.hbase + (int64_t)b * ch.nch + ch.lc] = hist[b];
}

// exclusive scan, three phases over blocks of SCAN_B elements
__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t *sh, int64_t &total)
{
    // blockDim.x == 256, wave64
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int64_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wv] = x;
    __syncthreads();
    int64_t wbase = 0;
    for (int k = 0; k < wv; ++k) wbase += sh[k];
    total = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return wbase + x - v;
}

// Block b of the scans owns elements [b * SCAN_B, (b + 1) * SCAN_B).
constexpr int SCAN_PER = 16, SCAN_B = 256 * SCAN_PER;

// total of each block (any order: coalesced loads)
__global__ __launch_bounds__(256) void k_scan_partial(const int *__restrict__ H, int64_t n,
                                                      int64_t *__restrict__ bsum)
{
    __shared__ int64_t sh[4];
    const int64_t b0 = (int64_t)blockIdx.x * SCAN_B;
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
        const int64_t i = b0 + k * 256 + threadIdx.x;
        if (i < n) s += H[i];
    }
    int64_t tot;
    block_excl_scan(s, sh, tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_scan_top(int64_t *__restrict__ bsum, int nb)
{
    __shared__ int64_t sh[4];
    int64_t carry = 0;
    for (int b0 = 0; b0 < nb; b0 += 256) {
        const int i = b0 + threadIdx.x;
        const int64_t v = (i < nb) ? bsum[i] : 0;
        int64_t tot;
        const int64_t ex = block_excl_scan(v, sh, tot);
        if (i < nb) bsum[i] = carry + ex;
        carry += tot;
    }
}

// exclusive scan of a block's SCAN_B counts: coalesced loads into LDS,
// thread t scans elements [16 t, 16 t + 16), coalesced stores from LDS
// (LDS index i + i / 16 keeps the per-thread runs off one bank)
__global__ __launch_bounds__(256) void k_scan_final(const int *__restrict__ H, int64_t n,
                                                    const int64_t *__restrict__ bsum, int64_t *__restrict__ Hs)
{
    __shared__ int64_t sh[4];
    __shared__ int cnt[SCAN_B + SCAN_B / 16];
    __shared__ int64_t out[SCAN_B + SCAN_B / 16];
    const int64_t b0 = (int64_t)blockIdx.x * SCAN_B;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
        const int i = k * 256 + threadIdx.x;
        cnt[i + (i >> 4)] = (b0 + i < n) ? H[b0 + i] : 0;
    }
    __syncthreads();
    const int j0 = threadIdx.x * SCAN_PER, l0 = j0 + (j0 >> 4);  // 16 elements: one padded run
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) s += cnt[l0 + k];
    int64_t tot;
    int64_t run = bsum[blockIdx.x] + block_excl_scan(s, sh, tot);
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
        out[l0 + k] = run;
        run += cnt[l0 + k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
        const int i = k * 256 + threadIdx.x;
        if (b0 + i < n) Hs[b0 + i] = out[i + (i >> 4)];
    }
}

// Stable scatter.  Each block owns one chunk and walks it in rounds of
// blockDim points; inside a round the waves take turns in order, and inside
// a wave lanes with equal keys are ranked by lane id (bit-sliced ballots), so
// the slot of a point is (#earlier points of the chunk with its key) +
// (destination of its (bin, chunk) cell).
template <typename R>
__global__ __launch_bounds__(256) void k_sort_scatter(const int *__restrict__ key, const SortChunk *__restrict__ chunks,
                                                      const int64_t *__restrict__ Hs,
                                                      const PtRec<R> *__restrict__ rec,
                                                      const typename cplx<R>::T *__restrict__ cin,
                                                      PtRec<R> *__restrict__ srec,
                                                      typename cplx<R>::T *__restrict__ cout,
                                                      int *__restrict__ perm_out)
{
    extern __shared__ int cnt[];
    const SortChunk ch = chunks[blockIdx.x];
    const int nbins = ch.nbins, keybits = ch.keybits;
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) cnt[b] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint64_t below = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    for (int64_t r0 = ch.p0; r0 < ch.p1; r0 += blockDim.x) {
        const int64_t p = r0 + threadIdx.x;
        const bool valid = p < ch.p1;
        const int k = valid ? key[p] : 0;
        uint64_t m = __ballot(valid);
        for (int b = 0; b < keybits; ++b) {
            const int bit = (k >> b) & 1;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const int rin = __popcll(m & below);
        const int same = __popcll(m);
        int rank = 0;
        for (int w = 0; w < nw; ++w) {
            if (wv == w && valid) {
                const int base = cnt[k];
                rank = base + rin;
                if (rin == 0) cnt[k] = base + same;
            }
            __syncthreads();
        }
        if (valid) {
            const int64_t dst = Hs[ch.hbase + (int64_t)k * ch.nch + ch.lc] + rank;
            srec[dst] = rec[p];
            if (cin) cout[dst] = cin[p];
            if (perm_out) perm_out[dst] = (int)(p - ch.pbase);
        }
    }
}

// binstart[bs0 + b] = first sorted slot of bin b of transform t;
// [bs0 + nbins] = end (grid: x over bins, y over transforms)
__global__ void k_binstart(const int64_t *__restrict__ Hs, const TransformInfo *__restrict__ tinfo,
                           int64_t *__restrict__ binstart)
{
    const TransformInfo ti = tinfo[blockIdx.y];
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > ti.nbins) return;
    binstart[ti.bs0 + b] = (b < ti.nbins && ti.nch > 0) ? Hs[ti.hbase + (int64_t)b * ti.nch] : ti.p1;
}

// Covering bound of the FP32 spread (one wave per bin): a cell x of a row is
// covered by the points whose first kernel cell lies in [x - w + 1, x]
// (cyclically), a window of w <= 2^sxsh cells that meets at most 3
// cyclically consecutive segments of 2^sxsh cells (the last segment may be
// short).  The largest count over such segment triples bounds the number of
// the bin's points covering any cell.
template <typename R>
__global__ void k_binmax(const PtRec<R> *__restrict__ srec, const int64_t *__restrict__ binstart,
                         const TransformInfo *__restrict__ tinfo, int *__restrict__ binmax)
{
    __shared__ int hist[4][64];
    const TransformInfo ti = tinfo[blockIdx.y];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = blockIdx.x * 4 + wv;
    hist[wv][lane] = 0;
    __syncthreads();
    if (b < ti.nbins) {
        const int64_t p0 = binstart[ti.bs0 + b], p1 = binstart[ti.bs0 + b + 1];
        for (int64_t p = p0 + lane; p < p1; p += 64) {
            const int sg = min(max(srec[p].lx, 0) >> ti.sxsh, ti.nseg - 1);
            atomicAdd(&hist[wv][sg], 1);
        }
    }
    __syncthreads();
    if (b < ti.nbins) {
        const int ns = ti.nseg;
        int v = 0;
        if (lane < ns) v = hist[wv][lane] + hist[wv][(lane + ns - 1) % ns] + hist[wv][(lane + ns - 2) % ns];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
        if (lane == 0) binmax[ti.bs0 + b] = v;
    }
}

// strengths of the plan API: cs[tr][k] = c[tr][perm[k]] (conjugated for iflag > 0)
template <typename R>
__global__ void k_gather_strengths(int64_t M, int ntr, const int *__restrict__ perm,
                                   const typename cplx<R>::T *__restrict__ c, typename cplx<R>::T *__restrict__ cs,
                                   int conj)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M * ntr) return;
    const int64_t tr = i / M, k = i % M;
    auto v = c[tr * M + perm[k]];
    if (conj) v.y = -v.y;
    cs[i] = v;
}

// amax[t] = max_k max(|Re c[t][k]|, |Im c[t][k]|): non-negative floats order
// like their bit patterns, so an integer max is exact and order-independent
// (a NaN strength yields a NaN bound)
__device__ __forceinline__ float nanmax(float a, float b) { return (a != a || b != b) ? __int_as_float(0x7fc00000) : fmaxf(a, b); }
__global__ void k_zero_u32(unsigned *__restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void k_absmax(int64_t M, const float2 *__restrict__ c, unsigned *__restrict__ amax)
{
    const int tr = blockIdx.y;
    float m = 0.0f;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < M; k += (int64_t)gridDim.x * blockDim.x) {
        const float2 v = c[(int64_t)tr * M + k];
        m = nanmax(m, nanmax(fabsf(v.x), fabsf(v.y)));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = nanmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(amax + tr, __float_as_uint(m));
}

// ---------------------------------------------------------------------------
// column FFT pass (y or z): transform t holds [outer][n][inner]; the block
// transforms `cols` adjacent columns of one outer slab along n in LDS,
// keeps NO (N, or N + 1 in a half-spectrum pass) of the n frequencies,
// deconvolves.  FINAL: additionally sums the nsp transforms of group
// blockIdx.y with weights and writes [g][N][inner] (conjugated for
// iflag > 0), or, herm, each kept mode and its mirror into the full map.
// ---------------------------------------------------------------------------
// RC is the arithmetic (and LDS tile) precision; the data in HBM stays
// cplx<R> (RC = double on FP32 data runs the pass in FP64 with FP64
// twiddles).
template <typename R, typename RC, int E, int NT, bool FINAL>
__global__ __launch_bounds__(NT) void k_colfft(const ColJob<R> *__restrict__ jobs, const int *__restrict__ blk0s,
                                               int njobs)
{
    using C = typename cplx<RC>::T;
    using CS = typename cplx<R>::T;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const ColJob<R> &J = jobs[find_job(blk0s, njobs, blockIdx.x)];
    const ColPass &cp = J.cp;
    C *tile = reinterpret_cast<C *>(smem_raw);
    const int n = cp.n, N = cp.N, cols = cp.cols, inner = cp.inner;
    C *tw_l = tile + n * cols;  // twiddles in LDS
    C *acc = tw_l + n;          // FINAL: species sum of the kept modes, [NO][cols]
    const C *twg;
    if constexpr (sizeof(RC) == sizeof(R)) twg = reinterpret_cast<const C *>(J.tw);
    else twg = reinterpret_cast<const C *>(J.tw64);
    for (int i = threadIdx.x; i < n; i += NT) tw_l[i] = twg[i];
    const int ncb = J.ncb;
    const int local = blockIdx.x - J.blk0;
    const int grp = local / (ncb * cp.outer);
    const int rem = local - grp * ncb * cp.outer;
    const int o = rem / ncb, cb = rem - o * ncb;
    const int col0 = cb * cols;
    const int ncol = min(cols, inner - col0);
    const int h = N / 2, NO = cp.NO;
    const int nsp = FINAL ? cp.nsp : 1;
    // deconvolution factor of kept output kk: phinv is even in k, so the
    // extra half-spectrum output k = N - h takes that of -k (entry 0 for
    // even N; for odd N that output is never a mirror source)
    auto phi_ix = [&](int kk) { return kk < N ? kk : 0; };
    for (int s = 0; s < nsp; ++s) {
        const int t = FINAL ? grp * cp.nsp + s : grp;
        {
            // all E loads of a thread are issued before the first LDS store;
            // tile element i = threadIdx.x + k * NT is row i / cols, column i % cols.
            // A non-final pass with nsp > 1 sums the group's nsp transforms
            // with the weights wsc before its FFT (the FFT is linear).
            const int nspl = FINAL ? 1 : cp.nsp;
            C v[E];
            for (int u = 0; u < nspl; ++u) {
                const int tu = FINAL ? t : grp * nspl + u;
                const CS *src = J.in + (int64_t)tu * cp.in_tstride + (int64_t)o * n * inner + col0;
                const RC wu = (!FINAL && J.wsc) ? (RC)J.wsc[u] : (RC)1;
#pragma unroll
                for (int k = 0; k < E; ++k) {
                    const int i = threadIdx.x + k * NT;
                    const int r = (int)fdiv((uint32_t)i, cp.dcols), c = i - r * cols;
                    const CS xs = (i < n * cols && c < ncol) ? src[r * inner + c] : cmk<CS>((R)0, (R)0);
                    const C x = cmk<C>((RC)xs.x, (RC)xs.y);
                    v[k] = (u == 0) ? cscale(x, wu) : cadd(v[k], cscale(x, wu));
                }
            }
#pragma unroll
            for (int k = 0; k < E; ++k) {
                const int i = threadIdx.x + k * NT;
                if (i < n * cols) tile[i] = v[k];
            }
        }
        __syncthreads();
        lds_fft<C, E>(tile, J.rp, cols, cp.dcols, cols, 1, tw_l, true);
        if constexpr (!FINAL) {
            CS *dst = J.out + (int64_t)t * cp.out_tstride + (int64_t)o * NO * inner + col0;
            for (int i = threadIdx.x; i < NO * cols; i += NT) {
                const int kk = (int)fdiv((uint32_t)i, cp.dcols), c = i - kk * cols;
                if (c >= ncol) continue;
                int f = kk - h;
                if (f < 0) f += n;
                const C x = cscale(tile[f * cols + c], (RC)J.phinv[phi_ix(kk)]);
                dst[(int64_t)kk * inner + c] = cmk<CS>((R)x.x, (R)x.y);
            }
        } else {
            for (int i = threadIdx.x; i < NO * cols; i += NT) {
                const int kk = (int)fdiv((uint32_t)i, cp.dcols), c = i - kk * cols;
                int f = kk - h;
                if (f < 0) f += n;
                R wv = J.phinv[phi_ix(kk)];
                // (a half-spectrum pass has no per-Q table: the host keeps those jobs full)
                if (J.wtab) wv *= (c < ncol) ? (R)J.wtab[(int64_t)s * cp.NQ + (int64_t)kk * inner + col0 + c] : (R)0;
                else if (J.wsc) wv *= J.wsc[s];
                const C x = cscale(tile[f * cols + c], wv);
                acc[i] = (s == 0) ? x : cadd(acc[i], x);
            }
        }
        __syncthreads();
    }
    if constexpr (FINAL) {
        if (!cp.herm) {
            C *dst = J.out + (int64_t)grp * cp.out_tstride + (int64_t)o * N * inner + col0;
            for (int i = threadIdx.x; i < N * cols; i += NT) {
                const int kk = (int)fdiv((uint32_t)i, cp.dcols), c = i - kk * cols;
                if (c >= ncol) continue;
                C x = acc[i];
                if (cp.conj) x.y = -x.y;
                dst[(int64_t)kk * inner + c] = x;
            }
        } else {
            // column col = km * hH1 + kh (3D; 2D: col = kh); the transformed
            // axis' index kk.  Mode (kk, km, kh) goes to its place and, when
            // k1 < 0, its mirror (2 h - kk, 2 h2 - km, 2 h1 - kh), where in
            // range, as the conjugate: every mode of the map is written once.
            C *dst = J.out + (int64_t)grp * cp.out_tstride;
            const int N1 = cp.hN1, h1 = N1 / 2, N2 = cp.hN2, h2 = N2 / 2;
            for (int i = threadIdx.x; i < NO * cols; i += NT) {
                const int kk = (int)fdiv((uint32_t)i, cp.dcols), c = i - kk * cols;
                if (c >= ncol) continue;
                const int col = col0 + c;
                int km = 0, kh = col;
                if (cp.hmid) {
                    km = (int)fdiv((uint32_t)col, cp.dH1);
                    kh = col - km * cp.hH1;
                }
                C x = acc[i];
                if (cp.conj) x.y = -x.y;
                if (kh >= cp.hH1v) continue;  // pad column
                if (kk < N && km < N2) dst[(int64_t)kk * cp.hsa + (int64_t)km * N1 + kh] = x;
                const int ka = 2 * h - kk, kmm = cp.hmid ? 2 * h2 - km : 0, kx = 2 * h1 - kh;
                if (kh < h1 && kx < N1 && ka >= 0 && ka < N && kmm >= 0 && kmm < N2)
                    dst[(int64_t)ka * cp.hsa + (int64_t)kmm * N1 + kx] = cmk<C>(x.x, -x.y);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// explicit instantiations + launch helpers
// ---------------------------------------------------------------------------
template <typename R> struct Tiles;
// y pass: 1024 threads of EY elements (FP32: 16 columns of 512 rows,
// 128-byte row segments; two workgroups, 32 waves, per CU); final pass:
// 256 threads of EF elements when the columns fit (FP32 on C2: 32 columns of
// 64 rows, 256-byte row segments), small tiles so that many workgroups
// share a CU while each walks its species one after another (C2, 24
// launches of 16 frames: 4.63 ms with 256 threads, 5.01 with 512, 4.96 with
// 128; 4 or 16 elements per thread 5.9-6.2), else 512 threads (the size
// limit of the column passes)
// RY: arithmetic of the y pass.  FP32 data keeps FP32 arithmetic: an FP64
// y pass (8 columns of 512 rows of double2, two workgroups per CU) measured
// 0.68 against 0.37 ms per 16-frame C2 launch, for errors of the FP32 form
// that stay of the same order (eps32 x the Bragg-to-diffuse amplitude ratio
// of near-lattice inputs, whichever pass order)
template <> struct Tiles<float> { static constexpr int EY = 8, EF = 8; using RY = float; };
template <> struct Tiles<double> { static constexpr int EY = 4, EF = 4; using RY = double; };
constexpr int colfft_nt(bool final) { return final ? 512 : 1024; }
// the final pass in 256-thread blocks when every column of the launch fits
constexpr int COLFFT_NT_FINAL_SMALL = 256;


template <typename R>
void launch_fold(const Geo &g, int64_t M, const void *x, const void *y, const void *z, bool in_f32, PtRec<R> *rec,
                 int *key, int *bad, hipStream_t st)
{
    const int NT = 256;
    const unsigned nb = (unsigned)((M + NT - 1) / NT);
    if (M == 0) return;
    if (in_f32)
        hipLaunchKernelGGL((k_fold<R, float>), dim3(nb), dim3(NT), 0, st, g, M, (const float *)x, (const float *)y,
                           (const float *)z, rec, key, bad);
    else
        hipLaunchKernelGGL((k_fold<R, double>), dim3(nb), dim3(NT), 0, st, g, M, (const double *)x,
                           (const double *)y, (const double *)z, rec, key, bad);
}

template <typename R>
void launch_stage_jobs(const StageJob *jobs, const int *blk0s, int njobs, int nblocks, PtRec<R> *rec, int *key,
                       typename cplx<R>::T *cstr, int *bad, hipStream_t st)
{
    if (nblocks <= 0) return;
    hipLaunchKernelGGL((k_stage_traj<R>), dim3((unsigned)nblocks), dim3(256), 0, st, jobs, blk0s, njobs, rec, key,
                       cstr, bad);
}

void launch_sort_count(const int *key, const SortChunk *chunks, int nchunks, int max_nbins, int *H, hipStream_t st)
{
    if (nchunks == 0) return;
    hipLaunchKernelGGL(k_sort_count, dim3(nchunks), dim3(256), max_nbins * sizeof(int), st, key, chunks, H);
}

void launch_scan(const int *H, int64_t n, int64_t *bsum, int64_t *Hs, hipStream_t st)
{
    if (n == 0) return;
    const int nb = (int)((n + SCAN_B - 1) / SCAN_B);
    hipLaunchKernelGGL(k_scan_partial, dim3(nb), dim3(256), 0, st, H, n, bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, st, bsum, nb);
    hipLaunchKernelGGL(k_scan_final, dim3(nb), dim3(256), 0, st, H, n, bsum, Hs);
}

int64_t scan_bsum_size(int64_t n) { return (n + SCAN_B - 1) / SCAN_B + 1; }

template <typename R>
void launch_sort_scatter(const int *key, const SortChunk *chunks, int nchunks, int max_nbins, const int64_t *Hs,
                         const PtRec<R> *rec, const typename cplx<R>::T *cin, PtRec<R> *srec,
                         typename cplx<R>::T *cout, int *perm, hipStream_t st)
{
    if (nchunks == 0) return;
    hipLaunchKernelGGL((k_sort_scatter<R>), dim3(nchunks), dim3(256), max_nbins * sizeof(int), st, key, chunks, Hs,
                       rec, cin, srec, cout, perm);
}

void launch_binstart(const int64_t *Hs, const TransformInfo *tinfo, int ntr, int max_nbins, int64_t *binstart,
                     hipStream_t st)
{
    if (ntr <= 0) return;
    hipLaunchKernelGGL(k_binstart, dim3((unsigned)((max_nbins + 1 + 255) / 256), (unsigned)ntr), dim3(256), 0, st, Hs,
                       tinfo, binstart);
}

template <typename R>
void launch_binmax(const PtRec<R> *srec, const int64_t *binstart, const TransformInfo *tinfo, int ntr, int max_nbins,
                   int *binmax, hipStream_t st)
{
    if (ntr <= 0 || max_nbins <= 0) return;
    hipLaunchKernelGGL((k_binmax<R>), dim3((unsigned)((max_nbins + 3) / 4), (unsigned)ntr), dim3(256), 0, st, srec,
                       binstart, tinfo, binmax);
}

template <typename R>
void launch_gather(int64_t M, int ntr, const int *perm, const typename cplx<R>::T *c, typename cplx<R>::T *cs,
                   int conj, hipStream_t st)
{
    const int64_t n = M * ntr;
    if (n == 0) return;
    hipLaunchKernelGGL((k_gather_strengths<R>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, M, ntr, perm, c,
                       cs, conj);
}

void launch_absmax(int64_t M, int ntr, const float2 *c, float *amax, hipStream_t st)
{
    if (ntr <= 0) return;
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((ntr + 255) / 256)), dim3(256), 0, st,
                       reinterpret_cast<unsigned *>(amax), ntr);
    if (M <= 0) return;
    const int64_t nb = std::min<int64_t>((M + 255) / 256, 1024);
    hipLaunchKernelGGL(k_absmax, dim3((unsigned)nb, (unsigned)ntr), dim3(256), 0, st, M, c,
                       reinterpret_cast<unsigned *>(amax));
}

template <typename R> int colfft_tile_elems(bool final)
{
    return (final ? Tiles<R>::EF : Tiles<R>::EY) * colfft_nt(final);
}

template <typename R> int colfft_final_small_elems() { return Tiles<R>::EF * COLFFT_NT_FINAL_SMALL; }

// LDS of one column pass of a job: tile + twiddles (+ species sum for the final pass)
template <typename R> size_t colfft_lds(const ColPass &cp, bool final)
{
    const size_t es = final ? sizeof(typename cplx<R>::T) : sizeof(typename cplx<typename Tiles<R>::RY>::T);
    return ((size_t)cp.n * (cp.cols + 1) + (final ? (size_t)cp.NO * cp.cols : 0)) * es;
}

template <typename R>
void launch_colfft_jobs(const ColJob<R> *jobs, const int *blk0s, int njobs, int nblocks, bool final, size_t lds,
                        hipStream_t st, int max_n)
{
    if (nblocks <= 0) return;
    constexpr int EF = Tiles<R>::EF, EY = Tiles<R>::EY, NTF = colfft_nt(true), NTY = colfft_nt(false);
    constexpr int NTS = COLFFT_NT_FINAL_SMALL;
    using RY = typename Tiles<R>::RY;
    // set once per process (a function-local static initialises thread-safely)
    static const bool attr = [] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_colfft<R, R, EF, NTF, true>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 156 * 1024);
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_colfft<R, R, EF, NTS, true>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 156 * 1024);
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_colfft<R, RY, EY, NTY, false>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 156 * 1024);
        return true;
    }();
    (void)attr;
    // the jobs' column counts were chosen for the small tile when their
    // columns fit it (colfft_final_small_elems), which the large tile holds too
    if (final && max_n <= colfft_final_small_elems<R>())
        hipLaunchKernelGGL((k_colfft<R, R, EF, NTS, true>), dim3((unsigned)nblocks), dim3(NTS), lds, st, jobs, blk0s,
                           njobs);
    else if (final)
        hipLaunchKernelGGL((k_colfft<R, R, EF, NTF, true>), dim3((unsigned)nblocks), dim3(NTF), lds, st, jobs, blk0s,
                           njobs);
    else
        hipLaunchKernelGGL((k_colfft<R, RY, EY, NTY, false>), dim3((unsigned)nblocks), dim3(NTY), lds, st, jobs,
                           blk0s, njobs);
}

#define MPSQ_INST(R)                                                                                                  \
    template void launch_fold<R>(const Geo &, int64_t, const void *, const void *, const void *, bool, PtRec<R> *,    \
                                 int *, int *, hipStream_t);                                                          \
    template void launch_stage_jobs<R>(const StageJob *, const int *, int, int, PtRec<R> *, int *, cplx<R>::T *,     \
                                       int *, hipStream_t);                                                           \
    template void launch_sort_scatter<R>(const int *, const SortChunk *, int, int, const int64_t *,                  \
                                         const PtRec<R> *, const cplx<R>::T *, PtRec<R> *, cplx<R>::T *, int *,       \
                                         hipStream_t);                                                                \
    template void launch_gather<R>(int64_t, int, const int *, const cplx<R>::T *, cplx<R>::T *, int, hipStream_t);    \
    template void launch_binmax<R>(const PtRec<R> *, const int64_t *, const TransformInfo *, int, int, int *,         \
                                   hipStream_t);                                                                      \
    template int colfft_tile_elems<R>(bool);                                                                          \
    template int colfft_final_small_elems<R>();                                                                       \
    template size_t colfft_lds<R>(const ColPass &, bool);                                                             \
    template void launch_colfft_jobs<R>(const ColJob<R> *, const int *, int, int, bool, size_t, hipStream_t, int);

MPSQ_INST(float)
MPSQ_INST(double)

}  // namespace mpsq
