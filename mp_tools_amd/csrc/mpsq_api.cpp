// mpsq_api.cpp -- host side of libmpsq.so: plan construction, the
// FINUFFT-shaped C ABI, the Fortran-ABI shims and the S(Q,w) engine.
//
// Everything here is orchestration: parameter choice, device buffers and
// kernel launches on one HIP stream per plan.  All arithmetic of the path
// runs in the kernels of nufft_kernels.hip and sqw_kernels.hip; there is no
// CPU fallback -- without a usable HIP device every call returns
// MPSQ_ERR_DEVICE with zeroed outputs.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <vector>

#include "mpsq.h"
#include "mpsq_kernels.h"

namespace mpsq {

struct Err {
    int code;
};

#define HIPCHK(x)                                                                                                     \
    do {                                                                                                              \
        hipError_t e_ = (x);                                                                                          \
        if (e_ != hipSuccess) throw Err{MPSQ_ERR_DEVICE};                                                             \
    } while (0)

// Error flags the kernels raise in a device int: bit 0 a non-finite point
// coordinate, bit 1 an internal bound check of the spread (a candidate run
// position outside the sorted records; never expected).
static void throw_bad(int hbad)
{
    if (hbad & 2) throw Err{MPSQ_ERR_DEVICE};
    if (hbad & 1) throw Err{MPSQ_ERR_NONFINITE_POINT};
}

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void ensure(size_t b)
    {
        if (b <= bytes && p) return;
        release();
        if (b == 0) return;
        if (hipMalloc(&p, b) != hipSuccess) {
            p = nullptr;
            throw Err{MPSQ_ERR_DEVICE};
        }
        bytes = b;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
    ~DevBuf() { release(); }
};

// ---------------------------------------------------------------------------
// kernel parameters and fine grid
// ---------------------------------------------------------------------------
// Width and shape of the exponential-of-semicircle kernel for upsampling
// factor 2, after the published FINUFFT parameter rule (Barnett, Magland &
// af Klinteberg, SIAM J. Sci. Comput. 41 (2019)): w = ceil(-log10(eps/10)),
// beta = gamma * w with gamma = 2.20, 2.26, 2.38 for w = 2, 3, 4 and 2.30
// otherwise.  FP32 plans floor eps at 1e-7 (w <= 8), FP64 plans at 1e-15.
static void es_params(double eps, bool single, int &w, double &beta, int &warn)
{
    warn = 0;
    if (!(eps > 0.0)) throw Err{MPSQ_ERR_BAD_ARG};
    const double floor_eps = single ? 1e-7 : 1e-15;
    if (eps < floor_eps) {
        warn = MPSQ_WARN_EPS_TOO_SMALL;
        eps = floor_eps;
    }
    w = (int)std::ceil(-std::log10(eps / 10.0));
    w = std::max(2, std::min(MAXW, w));
    const double gamma = (w == 2) ? 2.20 : (w == 3) ? 2.26 : (w == 4) ? 2.38 : 2.30;
    beta = gamma * w;
}

static int64_t next_smooth_even(int64_t n)
{
    if (n < 2) n = 2;
    if (n & 1) ++n;
    for (;; n += 2) {
        int64_t m = n;
        for (int p : {2, 3, 5})
            while (m % p == 0) m /= p;
        if (m == 1) return n;
    }
}

static bool radix_plan(int n, RadixPlan &rp)
{
    rp.n = n;
    rp.nst = 0;
    int m = n, e = 0;
    while (m % 2 == 0) {
        m /= 2;
        ++e;
    }
    std::vector<int> rs;
    int n8 = e / 3, rem = e % 3;
    if (rem == 1 && n8 >= 1) {
        --n8;
        rs.push_back(4);
        rs.push_back(4);
    } else if (rem == 1) {
        rs.push_back(2);
    } else if (rem == 2) {
        rs.push_back(4);
    }
    for (int i = 0; i < n8; ++i) rs.insert(rs.begin(), 8);
    while (m % 3 == 0) {
        m /= 3;
        rs.push_back(3);
    }
    while (m % 5 == 0) {
        m /= 5;
        rs.push_back(5);
    }
    if (m != 1 || rs.size() > 24) return false;
    int Ns = 1;
    for (size_t i = 0; i < rs.size(); ++i) {
        rp.r[rp.nst] = rs[i];
        rp.dnb[rp.nst] = make_fdiv((uint32_t)(n / rs[i]));
        rp.dns[rp.nst] = make_fdiv((uint32_t)Ns);
        rp.tws[rp.nst] = n / (Ns * rs[i]);
        Ns *= rs[i];
        ++rp.nst;
    }
    return true;
}

// Gauss-Legendre nodes/weights on [-1, 1] (Newton on P_q)
static void gauss_legendre(int q, std::vector<double> &x, std::vector<double> &wt)
{
    x.resize(q);
    wt.resize(q);
    for (int i = 0; i < q; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (q + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= q; ++k) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            const double dp = q * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= q; ++k) {
            const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        const double dp = q * (z * p1 - p0) / (z * z - 1.0);
        x[i] = z;
        wt[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

// 1 / phi_hat(k/n) for k = -N/2 .. (N-1)/2, where
// phi_hat(nu) = 2 int_0^{w/2} phi(t) cos(2 pi nu t) dt  (FP64 quadrature)
static std::vector<double> deconv_factors(int64_t N, int64_t n, int w, double beta)
{
    std::vector<double> gx, gw;
    gauss_legendre(128, gx, gw);
    const double hw = 0.5 * w;
    std::vector<double> tt(gx.size()), ff(gx.size());
    for (size_t i = 0; i < gx.size(); ++i) {
        tt[i] = 0.5 * hw * (gx[i] + 1.0);
        const double a = 1.0 - (tt[i] / hw) * (tt[i] / hw);
        ff[i] = 0.5 * hw * gw[i] * std::exp(beta * (std::sqrt(std::max(0.0, a)) - 1.0));
    }
    std::vector<double> out((size_t)N);
    for (int64_t kk = 0; kk < N; ++kk) {
        const double k = (double)(kk - N / 2);
        double s = 0.0;
        for (size_t i = 0; i < gx.size(); ++i) s += ff[i] * std::cos(2.0 * M_PI * k * tt[i] / (double)n);
        out[(size_t)kk] = 1.0 / (2.0 * s);
    }
    return out;
}

template <typename R> using Cx = typename cplx<R>::T;

// I(Q,t) by FFT correlation from this many lags on (below it the direct
// sums read less), with blocks sized to keep two per CU's LDS
constexpr int IQT_FFT_MIN_LAGS = 16;
constexpr size_t IQT_FFT_LDS_TARGET = 76 * 1024;

template <typename R> static void upload_twiddles(DevBuf &buf, int n, hipStream_t st, std::vector<Cx<R>> &host)
{
    host.resize((size_t)n);
    for (int m = 0; m < n; ++m) {
        const double a = -2.0 * M_PI * (double)m / (double)n;
        host[(size_t)m].x = (R)std::cos(a);
        host[(size_t)m].y = (R)std::sin(a);
    }
    buf.ensure(sizeof(Cx<R>) * (size_t)n);
    HIPCHK(hipMemcpyAsync(buf.p, host.data(), sizeof(Cx<R>) * (size_t)n, hipMemcpyHostToDevice, st));
}

static int ilog2_ceil(int64_t v)
{
    int b = 0;
    while ((int64_t(1) << b) < v) ++b;
    return b;
}

// HIP-event timing of launch groups on one stream (stage categories of
// mpsq_sqw_get_timing).  Events are pooled; elapsed times are collected
// after the stream has been synchronised.
struct StageTimer {
    static constexpr int K = 6;
    bool on = false;
    std::vector<hipEvent_t> pool;
    std::vector<int> cat;
    size_t used = 0;
    double ms[K] = {};
    int64_t n[K] = {};
    void reset()
    {
        used = 0;
        cat.clear();
        for (int k = 0; k < K; ++k) {
            ms[k] = 0.0;
            n[k] = 0;
        }
    }
    int begin(int k, hipStream_t st)
    {
        if (!on) return -1;
        if (2 * used >= pool.size()) {
            hipEvent_t a, b;
            HIPCHK(hipEventCreate(&a));
            HIPCHK(hipEventCreate(&b));
            pool.push_back(a);
            pool.push_back(b);
        }
        cat.push_back(k);
        HIPCHK(hipEventRecord(pool[2 * used], st));
        return (int)used++;
    }
    void end(int idx, hipStream_t st)
    {
        if (idx >= 0) HIPCHK(hipEventRecord(pool[2 * (size_t)idx + 1], st));
    }
    void collect()
    {
        for (size_t i = 0; i < used; ++i) {
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, pool[2 * i], pool[2 * i + 1]));
            ms[cat[i]] += t;
            n[cat[i]] += 1;
        }
        used = 0;
        cat.clear();
    }
    ~StageTimer()
    {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

enum { ST_STAGE = 0, ST_SORT = 1, ST_SPREAD = 2, ST_COLPASS = 3, ST_FINAL = 4, ST_REDUCE = 5 };

// ---------------------------------------------------------------------------
// Job tables: built in pinned host memory and uploaded with one copy per
// run, ahead of the kernels that read them (stream order).  Before the host
// rewrites the tables it waits for the previous upload, which keeps a
// pending copy from reading half-written tables.
// ---------------------------------------------------------------------------
struct Arena {
    unsigned char *h = nullptr;
    size_t cap = 0, used = 0;
    DevBuf d;
    hipEvent_t ev = nullptr;
    bool pending = false;
    void begin()
    {
        if (pending) HIPCHK(hipEventSynchronize(ev));
        pending = false;
        used = 0;
    }
    size_t put_raw(const void *src, size_t bytes)
    {
        const size_t off = (used + 255) & ~(size_t)255;
        if (off + bytes > cap) {
            size_t nc = std::max<size_t>(cap * 2, 1 << 16);
            while (nc < off + bytes) nc *= 2;
            unsigned char *nh = nullptr;
            HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&nh), nc, hipHostMallocDefault));
            if (h) {
                std::memcpy(nh, h, used);
                (void)hipHostFree(h);
            }
            h = nh;
            cap = nc;
        }
        if (bytes) std::memcpy(h + off, src, bytes);
        used = off + bytes;
        return off;
    }
    template <typename T> size_t put(const std::vector<T> &v) { return put_raw(v.data(), sizeof(T) * v.size()); }
    void upload(hipStream_t st)
    {
        if (used == 0) return;
        d.ensure(used);
        if (!ev) HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        HIPCHK(hipMemcpyAsync(d.p, h, used, hipMemcpyHostToDevice, st));
        HIPCHK(hipEventRecord(ev, st));
        pending = true;
    }
    template <typename T> const T *dev(size_t off) const
    {
        return reinterpret_cast<const T *>(static_cast<const unsigned char *>(d.p) + off);
    }
    ~Arena()
    {
        if (pending && ev) (void)hipEventSynchronize(ev);
        if (ev) (void)hipEventDestroy(ev);
        if (h) (void)hipHostFree(h);
    }
};

// Cut the tile-id space [0, cost.size()) into 8 contiguous ranges of about
// equal total cost, one per XCD.
static std::vector<int> xcd_ranges(const std::vector<double> &job_cost, const std::vector<int> &job_tiles)
{
    double tot = 0.0;
    int ntile = 0;
    for (size_t j = 0; j < job_cost.size(); ++j) {
        tot += job_cost[j] * job_tiles[j];
        ntile += job_tiles[j];
    }
    std::vector<int> xr(9, 0);
    xr[8] = ntile;
    int tile = 0;
    double acc = 0.0;
    size_t jj = 0;
    int in_job = 0;
    for (int x = 1; x < 8; ++x) {
        const double target = tot * x / 8.0;
        while (jj < job_cost.size() && acc + job_cost[jj] * (job_tiles[jj] - in_job) <= target) {
            acc += job_cost[jj] * (job_tiles[jj] - in_job);
            tile += job_tiles[jj] - in_job;
            ++jj;
            in_job = 0;
        }
        if (jj < job_cost.size() && job_cost[jj] > 0) {
            const int k = std::min(job_tiles[jj] - in_job, (int)((target - acc) / job_cost[jj]));
            acc += job_cost[jj] * k;
            tile += k;
            in_job += k;
        }
        xr[x] = tile;
    }
    return xr;
}

// Spread tile shapes chosen by the autotuner, per (dim, n1, n2, n3, w, fp64).
struct TuneKey {
    int v[6];
    bool operator<(const TuneKey &o) const
    {
        for (int i = 0; i < 6; ++i)
            if (v[i] != o.v[i]) return v[i] < o.v[i];
        return false;
    }
};
static std::mutex g_tune_mu;
static std::map<TuneKey, std::pair<int, int>> g_tuned;

// ---------------------------------------------------------------------------
// geometry shared by plans and the S(Q,w) engine
// ---------------------------------------------------------------------------
template <typename R> struct Core {
    int dim = 0, iflag = -1;
    double eps = 0;
    int64_t N[3] = {1, 1, 1}, n[3] = {1, 1, 1};
    Geo g{};
    RadixPlan rp[3]{};
    DevBuf tw[3], tw64[3], phinv[3];
    std::vector<Cx<R>> tw_host[3];
    std::vector<double2> tw64_host[3];
    std::vector<R> phinv_host[3];
    int device = 0;
    hipStream_t st = nullptr;
    bool own_stream = false;
    int warn = 0;
    int sort_chunk = 0;
    // FP32 with real strengths: the narrow spread kernel's real form, whose
    // 32-bit cells and paired-row x FFT allow tiles of twice the rows
    bool real = false;
    // real strengths (either precision): the half spectrum through the
    // column passes
    bool real_str = false;
    int rows_max = 1;

    void setup(int dim_, const int64_t *nm, int iflag_, double eps_, const mpsq_opts *o, bool real_ = false)
    {
        if (dim_ < 1 || dim_ > 3 || !nm) throw Err{MPSQ_ERR_BAD_ARG};
        dim = dim_;
        iflag = iflag_;
        eps = eps_;
        for (int d = 0; d < 3; ++d) N[d] = (d < dim) ? nm[d] : 1;
        for (int d = 0; d < dim; ++d)
            if (N[d] < 1 || N[d] > (1 << 20)) throw Err{MPSQ_ERR_BAD_ARG};
        mpsq_opts def;
        mpsq_default_opts(&def);
        if (!o) o = &def;
        device = o->device;
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1 || device < 0 || device >= ndev)
            throw Err{MPSQ_ERR_DEVICE};
        HIPCHK(hipSetDevice(device));
        if (o->stream) {
            st = (hipStream_t)o->stream;
            own_stream = false;
        } else {
            HIPCHK(hipStreamCreate(&st));  // blocking: ordered with the legacy default stream
            own_stream = true;
        }
        int w;
        double beta;
        es_params(eps, sizeof(R) == 4, w, beta, warn);
        for (int d = 0; d < 3; ++d) n[d] = (d < dim) ? next_smooth_even(std::max(2 * N[d], (int64_t)2 * w)) : 1;
        const int TE = spread_tile_elems<R>();
        const int TC = colfft_tile_elems<R>(true);
        if (n[0] > TE || n[1] > TC || n[2] > TC) throw Err{MPSQ_ERR_SIZE};
        g.dim = dim;
        g.n1 = (int)n[0];
        g.n2 = (int)n[1];
        g.n3 = (int)n[2];
        g.N1 = (int)N[0];
        g.N2 = (int)N[1];
        g.N3 = (int)N[2];
        g.w = w;
        g.beta = beta;
        g.ic = 4.0 / ((double)w * (double)w);
        g.s1 = (double)n[0] / (2.0 * M_PI);
        g.s2 = (double)n[1] / (2.0 * M_PI);
        g.s3 = (double)n[2] / (2.0 * M_PI);
        g.hw = 0.5 * (double)w;
        // spread tile: whole x extent x (by x bz) rows within the LDS tile;
        // by + w <= n2 and bz + w <= n3 keep the rows a kernel stencil has in
        // one tile contiguous
        real_str = real_;
        real = real_ && sizeof(R) == 4;
        // real strengths: up to 32 rows where the complex kernel's registers
        // stop short of that (n1 = 512: 16 -> 32); grids that already allow
        // 32 rows or more keep their cap, since taller tiles of dense inputs
        // overflow the narrow kernel's covering bound and every wide pass
        // re-walks the whole tile's candidates
        rows_max = spread_max_rows<R>((int)n[0]);
        if (real && rows_max < 32) rows_max = std::min(2 * rows_max, 32);
        const int bmax = spread_max_tile_b();
        int by = 1, bz = 1;
        if (dim == 2) {
            for (int b = 1; b <= rows_max && b <= bmax && n[1] % b == 0 && b + w <= n[1]; b *= 2) by = b;
        } else if (dim == 3) {
            double best = 1e300;
            for (int b1 = 1; b1 <= rows_max && b1 <= bmax && n[1] % b1 == 0 && b1 + w <= n[1]; b1 *= 2)
                for (int b2 = 1; b1 * b2 <= rows_max && b2 <= bmax && n[2] % b2 == 0 && b2 + w <= n[2]; b2 *= 2) {
                    // (ties: more rows, then the taller y extent, whose
                    // covering bounds, summed over fewer z-bin rows, are
                    // smaller: C2's real 8 x 4 tiles give 1.51e-5 at random
                    // Q against 1.58e-5 for 4 x 8, at the same speed)
                    const double cost = (double)(b1 + w - 1) * (b2 + w - 1) / (b1 * b2);
                    const bool tie = std::fabs(cost - best) < 1e-12;
                    if (cost < best - 1e-12 || (tie && b1 * b2 > by * bz) || (tie && b1 * b2 == by * bz && b1 > by)) {
                        best = cost;
                        by = b1;
                        bz = b2;
                    }
                }
        }
        if (o->tile_by <= 0 && o->tile_bz <= 0) {
            const TuneKey key{{dim, (int)n[0], (int)n[1], (int)n[2], w, sizeof(R) == 8 ? 1 : (real ? 2 : 0)}};
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tuned.find(key);
            if (it != g_tuned.end()) {
                by = it->second.first;
                bz = it->second.second;
            }
        }
        if (o->tile_by > 0 && dim >= 2 && n[1] % o->tile_by == 0) by = o->tile_by;
        if (o->tile_bz > 0 && dim >= 3 && n[2] % o->tile_bz == 0) bz = o->tile_bz;
        if (by * bz > rows_max || by > bmax || bz > bmax) throw Err{MPSQ_ERR_BAD_ARG};
        if (spread_lds<R>(by * bz, (int)n[0], w, real) > 156 * 1024) throw Err{MPSQ_ERR_SIZE};
        if ((dim >= 2 && by + w > n[1]) || (dim >= 3 && bz + w > n[2])) throw Err{MPSQ_ERR_BAD_ARG};
        g.by = by;
        g.bz = bz;
        // sort bins: start at the tile size, coarsen until the bin count fits
        // the LDS counters, a tile visits at most MAX_NB_BINS bins and has at
        // most spread_max_runs() runs
        int BY = (dim >= 2) ? by : 1, BZ = (dim >= 3) ? bz : 1;
        auto nvis = [&](int B, int b, int64_t nb) -> int64_t {
            int64_t v = (b + w - 2 + B - 1) / B + 1;
            return std::min(v, nb);
        };
        for (;;) {
            const int64_t nbY = (dim >= 2) ? n[1] / BY : 1, nbZ = (dim >= 3) ? n[2] / BZ : 1;
            const int64_t vis = (dim >= 2 ? nvis(BY, by, nbY) : 1) * (dim >= 3 ? nvis(BZ, bz, nbZ) : 1);
            const int64_t runs = 2 * (dim >= 3 ? nvis(BZ, bz, nbZ) : 1);
            if (nbY * nbZ <= 16384 && vis <= MAX_NB_BINS && runs <= spread_max_runs()) break;
            const bool cy = (dim >= 2) && (n[1] % (2 * BY) == 0) && (BY < n[1]);
            const bool cz = (dim >= 3) && (n[2] % (2 * BZ) == 0) && (BZ < n[2]);
            if (cy && (!cz || nbY >= nbZ)) BY *= 2;
            else if (cz) BZ *= 2;
            else throw Err{MPSQ_ERR_SIZE};
        }
        g.BY = BY;
        g.BZ = BZ;
        g.nbY = (dim >= 2) ? (int)(n[1] / BY) : 1;
        g.nbZ = (dim >= 3) ? (int)(n[2] / BZ) : 1;
        g.nbins = g.nbY * g.nbZ;
        sort_chunk = o->sort_chunk > 0 ? o->sort_chunk : std::max(1024, 2 * g.nbins);
        sort_chunk = ((sort_chunk + 255) / 256) * 256;
        for (int d = 0; d < dim; ++d) {
            if (!radix_plan((int)n[d], rp[d])) throw Err{MPSQ_ERR_SIZE};
            upload_twiddles<R>(tw[d], (int)n[d], st, tw_host[d]);
            upload_twiddles<double>(tw64[d], (int)n[d], st, tw64_host[d]);
            const std::vector<double> f = deconv_factors(N[d], n[d], w, beta);
            phinv_host[d].assign(f.begin(), f.end());
            phinv[d].ensure(sizeof(R) * f.size());
            HIPCHK(hipMemcpyAsync(phinv[d].p, phinv_host[d].data(), sizeof(R) * f.size(), hipMemcpyHostToDevice,
                                  st));
        }
        HIPCHK(hipStreamSynchronize(st));
    }

    int keybits() const { return g.nbins > 1 ? ilog2_ceil(g.nbins) : 0; }

    int colfft_cols(int64_t nlen, int64_t inner, bool final) const
    {
        const int TE = (final && nlen <= colfft_final_small_elems<R>()) ? colfft_final_small_elems<R>()
                                                                         : colfft_tile_elems<R>(final);
        int c = 1;
        while ((int64_t)(2 * c) * nlen <= TE) c *= 2;
        while (c > 1 && c / 2 >= inner) c /= 2;
        return c;
    }

    // Column passes after spreading.  2D: one final pass along y.  3D: a
    // y pass (every transform) and a final z pass that groups nsp
    // consecutive transforms (species) with their weights.  herm (real
    // strengths): the half spectrum, x frequencies k1 <= 0 (h1x of them)
    // and y, z frequencies -N/2 .. N - N/2 (one more than kept, so that
    // every mode's mirror is among them); the final pass writes the mirrors.
    // (rows of the half spectrum padded to 16 columns: 128-byte FP32 row
    // segments of the y pass start on cache lines)
    int64_t h1v(bool herm) const { return herm ? N[0] / 2 + 1 : N[0]; }
    int64_t h1x(bool herm) const { return herm ? (h1v(true) + 15) / 16 * 16 : N[0]; }
    ColPass colpass(int axis, bool final, int nsp, bool herm = false) const
    {
        ColPass c{};
        const int64_t NQ = N[0] * N[1] * N[2];
        const int64_t H1 = h1x(herm), N2o = herm ? N[1] + 1 : N[1];
        c.n = (int)n[axis];
        c.N = (int)N[axis];
        c.NO = herm ? c.N + 1 : c.N;
        if (axis == 1) {
            c.inner = (int)H1;
            c.outer = (dim == 3) ? (int)n[2] : 1;
            c.in_tstride = n[2] * n[1] * H1;
            c.out_tstride = (dim == 3) ? n[2] * N2o * H1 : NQ;
        } else {
            c.inner = (int)(N2o * H1);
            c.outer = 1;
            c.in_tstride = n[2] * N2o * H1;
            c.out_tstride = NQ;
        }
        if (herm && final) {
            c.herm = 1;
            c.hmid = (dim == 3) ? 1 : 0;
            c.hN1 = (int)N[0];
            c.hH1 = (int)H1;
            c.hH1v = (int)h1v(true);
            c.hN2 = (dim == 3) ? (int)N[1] : 1;
            c.hsa = (dim == 3) ? N[0] * N[1] : N[0];
            c.dH1 = make_fdiv((uint32_t)H1);
        }
        c.cols = colfft_cols(c.n, c.inner, final);
        c.dcols = make_fdiv((uint32_t)c.cols);
        c.nsp = nsp;
        c.NQ = NQ;
        c.conj = (final && iflag > 0) ? 1 : 0;
        return c;
    }
    ColJob<R> coljob(int axis, bool final, int nsp, int ngroups, const Cx<R> *in, Cx<R> *out, const R *wsc,
                     const float *wtab, bool herm = false) const
    {
        ColJob<R> J{};
        J.cp = colpass(axis, final, nsp, herm);
        J.rp = rp[axis];
        J.in = in;
        J.out = out;
        J.tw = tw[axis].template as<Cx<R>>();
        J.tw64 = tw64[axis].template as<double2>();
        J.phinv = phinv[axis].template as<R>();
        J.wsc = wsc;
        J.wtab = final ? wtab : nullptr;
        J.ncb = (J.cp.inner + J.cp.cols - 1) / J.cp.cols;
        J.ngroups = ngroups;
        J.blk0 = 0;
        return J;
    }
    static int64_t col_blocks(const ColJob<R> &J) { return (int64_t)J.ncb * J.cp.outer * J.ngroups; }

    int ntiles() const
    {
        const int by = (dim >= 2) ? g.by : 1, bz = (dim >= 3) ? g.bz : 1;
        return ((dim >= 2) ? g.n2 / by : 1) * ((dim >= 3) ? g.n3 / bz : 1);
    }
    // x segments of the covering bound: 2^sxsh >= w cells, at most 64 of them
    void segments(int &sxsh, int &nseg) const
    {
        sxsh = 0;
        while ((1 << sxsh) < g.w || ((g.n1 + (1 << sxsh) - 1) >> sxsh) > 64) ++sxsh;
        nseg = (g.n1 + (1 << sxsh) - 1) >> sxsh;
    }
    size_t spread_lds_size() const
    {
        const int by = (dim >= 2) ? g.by : 1, bz = (dim >= 3) ? g.bz : 1;
        return spread_lds<R>(by * bz, g.n1, g.w, real);
    }
    SpreadJob<R> spreadjob(int ntr, const PtRec<R> *srec, int64_t nrec, int *err, const Cx<R> *cs, int64_t c_tstride,
                           const int64_t *bstart, const int *bmax, int bin_tstride, Cx<R> *G1, const float *amax,
                           int final_conj) const
    {
        SpreadJob<R> J{};
        J.g = g;
        J.rp1 = rp[0];
        J.srec = srec;
        J.nrec = nrec;
        J.err = err;
        J.cs = cs;
        J.c_tstride = c_tstride;
        J.binstart = bstart;
        J.bin_tstride = bin_tstride;
        J.tw1 = tw[0].template as<Cx<R>>();
        J.phinv1 = phinv[0].template as<R>();
        J.G1 = G1;
        J.amax = amax;
        J.amax_s = 1.0f;
        J.binmax = bmax;
        J.final_conj = final_conj;
        J.ntr = ntr;
        J.ntiles = ntiles();
        J.tile0 = 0;
        const int by = (dim >= 2) ? g.by : 1, bz = (dim >= 3) ? g.bz : 1;
        J.grp = spread_group_rows<R>(by * bz, g.n1);
        J.grp_sh = ilog2_ceil(J.grp);
        J.dntr = make_fdiv((uint32_t)ntr);
        J.dnt = make_fdiv((uint32_t)J.ntiles);
        J.tmajor = 0;
        J.dnty = make_fdiv((uint32_t)((dim >= 2) ? g.n2 / by : 1));
        J.dby = make_fdiv((uint32_t)by);
        J.lds_tw = (int)spread_lds_twiddles<R>(by * bz, g.n1, g.w, real);
        J.g1_len = g.N1;
        J.g1_keep = g.N1;
        return J;
    }
    // estimated cost of one tile of this geometry with M points per transform
    // (candidates it visits plus a fixed part for zeroing / FFT / output)
    double tile_cost(int64_t M) const
    {
        const int by = (dim >= 2) ? g.by : 1, bz = (dim >= 3) ? g.bz : 1;
        const double halo = (double)(by + (dim >= 2 ? g.w - 1 : 0)) * (bz + (dim >= 3 ? g.w - 1 : 0));
        const double cand = (double)M * halo / ((double)g.n2 * g.n3);
        return cand * g.w + (double)by * bz * g.n1 * 1.5;
    }

    ~Core()
    {
        if (st) (void)hipStreamSynchronize(st);
        if (own_stream && st) (void)hipStreamDestroy(st);
    }
};

// ---------------------------------------------------------------------------
// FINUFFT-shaped plan
// ---------------------------------------------------------------------------
template <typename R> struct Plan {
    Core<R> core;
    int ntrans = 1;
    int64_t M = -1;
    DevBuf xyz[3], rec, key, srec, perm, H, Hs, bsum, binstart, binmax, bad, wc, fbl;
    DevBuf cin, cs, G1, G2, fout, amax;
    Arena sort_tab, run_tab;
    int nch = 0;
    size_t off_chunks = 0, off_tinfo = 0;

    void setpts(int64_t M_, const R *x, const R *y, const R *z, bool on_device)
    {
        Core<R> &c = core;
        if (M_ < 0 || M_ > INT32_MAX) throw Err{MPSQ_ERR_BAD_ARG};
        HIPCHK(hipSetDevice(c.device));
        M = -1;
        const R *in[3] = {x, y, z};
        const void *dptr[3] = {nullptr, nullptr, nullptr};
        for (int d = 0; d < c.dim; ++d) {
            if (!in[d] && M_ > 0) throw Err{MPSQ_ERR_BAD_ARG};
            if (on_device) {
                dptr[d] = in[d];
            } else {
                xyz[d].ensure(sizeof(R) * (size_t)std::max<int64_t>(M_, 1));
                if (M_ > 0)
                    HIPCHK(hipMemcpyAsync(xyz[d].p, in[d], sizeof(R) * (size_t)M_, hipMemcpyHostToDevice, c.st));
                dptr[d] = xyz[d].p;
            }
        }
        const size_t Mb = (size_t)std::max<int64_t>(M_, 1);
        rec.ensure(sizeof(PtRec<R>) * Mb);
        srec.ensure(sizeof(PtRec<R>) * Mb);
        key.ensure(sizeof(int) * Mb);
        perm.ensure(sizeof(int) * Mb);
        bad.ensure(sizeof(int));
        HIPCHK(hipMemsetAsync(bad.p, 0, sizeof(int), c.st));
        launch_fold<R>(c.g, M_, dptr[0], dptr[1], dptr[2], sizeof(R) == 4, rec.as<PtRec<R>>(), key.as<int>(),
                       bad.as<int>(), c.st);
        const int64_t CH = c.sort_chunk;
        nch = (int)((M_ + CH - 1) / CH);
        std::vector<SortChunk> chunks((size_t)std::max(nch, 1));
        for (int k = 0; k < nch; ++k) {
            SortChunk &ch = chunks[k];
            ch.p0 = (int64_t)k * CH;
            ch.p1 = std::min<int64_t>(M_, (int64_t)(k + 1) * CH);
            ch.pbase = 0;
            ch.hbase = 0;
            ch.nch = nch;
            ch.lc = k;
            ch.nbins = c.g.nbins;
            ch.keybits = c.keybits();
        }
        std::vector<TransformInfo> tinfo(1);
        tinfo[0].p0 = 0;
        tinfo[0].p1 = M_;
        tinfo[0].hbase = 0;
        tinfo[0].bs0 = 0;
        tinfo[0].nch = nch;
        tinfo[0].nbins = c.g.nbins;
        c.segments(tinfo[0].sxsh, tinfo[0].nseg);
        sort_tab.begin();
        off_chunks = sort_tab.put(chunks);
        off_tinfo = sort_tab.put(tinfo);
        sort_tab.upload(c.st);
        const int64_t Htot = (int64_t)c.g.nbins * nch;
        H.ensure(sizeof(int) * (size_t)std::max<int64_t>(Htot, 1));
        Hs.ensure(sizeof(int64_t) * (size_t)std::max<int64_t>(Htot, 1));
        bsum.ensure(sizeof(int64_t) * (size_t)scan_bsum_size(Htot));
        binstart.ensure(sizeof(int64_t) * (size_t)(c.g.nbins + 1));
        const SortChunk *chd = sort_tab.dev<SortChunk>(off_chunks);
        launch_sort_count(key.as<int>(), chd, nch, c.g.nbins, H.as<int>(), c.st);
        launch_scan(H.as<int>(), Htot, bsum.as<int64_t>(), Hs.as<int64_t>(), c.st);
        launch_sort_scatter<R>(key.as<int>(), chd, nch, c.g.nbins, Hs.as<int64_t>(), rec.as<PtRec<R>>(), nullptr,
                               srec.as<PtRec<R>>(), nullptr, perm.as<int>(), c.st);
        launch_binstart(Hs.as<int64_t>(), sort_tab.dev<TransformInfo>(off_tinfo), 1, c.g.nbins,
                        binstart.as<int64_t>(), c.st);
        if (sizeof(R) == 4) {
            binmax.ensure(sizeof(int) * (size_t)(c.g.nbins + 1));
            launch_binmax<R>(srec.as<PtRec<R>>(), binstart.as<int64_t>(), sort_tab.dev<TransformInfo>(off_tinfo), 1,
                             c.g.nbins, binmax.as<int>(), c.st);
        }
        HIPCHK(hipGetLastError());
        int hbad = 0;
        HIPCHK(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, c.st));
        HIPCHK(hipStreamSynchronize(c.st));
        throw_bad(hbad);
        M = M_;
    }

    void execute(const R *c_in, R *f_out, bool on_device)
    {
        Core<R> &c = core;
        if (M < 0) throw Err{MPSQ_ERR_NOT_SET};
        HIPCHK(hipSetDevice(c.device));
        const int64_t NQ = c.N[0] * c.N[1] * c.N[2];
        const size_t Mb = (size_t)std::max<int64_t>(M, 1);
        const Cx<R> *cd;
        if (on_device) {
            cd = reinterpret_cast<const Cx<R> *>(c_in);
        } else {
            cin.ensure(sizeof(Cx<R>) * Mb * ntrans);
            if (M > 0)
                HIPCHK(hipMemcpyAsync(cin.p, c_in, sizeof(Cx<R>) * (size_t)M * ntrans, hipMemcpyHostToDevice, c.st));
            cd = cin.as<Cx<R>>();
        }
        Cx<R> *fd;
        if (on_device) {
            fd = reinterpret_cast<Cx<R> *>(f_out);
        } else {
            fout.ensure(sizeof(Cx<R>) * (size_t)NQ * ntrans);
            fd = fout.as<Cx<R>>();
        }
        cs.ensure(sizeof(Cx<R>) * Mb * ntrans);
        launch_gather<R>(M, ntrans, perm.as<int>(), cd, cs.as<Cx<R>>(), c.iflag > 0 ? 1 : 0, c.st);
        Cx<R> *g1 = fd;
        if (c.dim >= 2) {
            G1.ensure(sizeof(Cx<R>) * (size_t)(c.n[2] * c.n[1] * c.N[0]) * ntrans);
            g1 = G1.as<Cx<R>>();
        }
        if (c.dim == 3) G2.ensure(sizeof(Cx<R>) * (size_t)(c.n[2] * c.N[1] * c.N[0]) * ntrans);
        const float *am = nullptr;
        if (sizeof(R) == 4) {
            amax.ensure(sizeof(float) * (size_t)ntrans);
            launch_absmax(M, ntrans, reinterpret_cast<const float2 *>(cs.p), amax.as<float>(), c.st);
            am = amax.as<float>();
        }
        // one-entry job tables: spread, then the column passes
        launch_zero_i32(bad.as<int>(), 1, c.st);
        std::vector<SpreadJob<R>> sj(1, c.spreadjob(ntrans, srec.as<PtRec<R>>(), M, bad.as<int>(), cs.as<Cx<R>>(), M,
                                                     binstart.as<int64_t>(),
                                                     sizeof(R) == 4 ? binmax.as<int>() : nullptr, 0, g1, am,
                                                     (c.dim == 1 && c.iflag > 0) ? 1 : 0));
        std::vector<int> zero1(1, 0);
        const int ntile = sj[0].ntiles * ntrans;
        std::vector<int> xr = xcd_ranges(std::vector<double>(1, 1.0), std::vector<int>(1, ntile));
        std::vector<ColJob<R>> cy, cz;
        if (c.dim == 2) {
            cz.push_back(c.coljob(1, true, 1, ntrans, g1, fd, nullptr, nullptr));
        } else if (c.dim == 3) {
            cy.push_back(c.coljob(1, false, 1, ntrans, g1, G2.as<Cx<R>>(), nullptr, nullptr));
            cz.push_back(c.coljob(2, true, 1, ntrans, G2.as<Cx<R>>(), fd, nullptr, nullptr));
        }
        run_tab.begin();
        const size_t o_sj = run_tab.put(sj), o_z = run_tab.put(zero1), o_xr = run_tab.put(xr);
        const size_t o_cy = run_tab.put(cy), o_cz = run_tab.put(cz);
        run_tab.upload(c.st);
        wc.ensure(sizeof(int) * XCD_CTR_INTS);
        fbl.ensure(sizeof(int) * (size_t)spread_max_wide_passes() * (size_t)ntile);
        launch_spread_jobs<R>(run_tab.dev<SpreadJob<R>>(o_sj), run_tab.dev<int>(o_z), 1, c.g.w,
                              run_tab.dev<int>(o_xr), wc.as<int>(), fbl.as<int>(), c.spread_lds_size(), c.st);
        if (!cy.empty())
            launch_colfft_jobs<R>(run_tab.dev<ColJob<R>>(o_cy), run_tab.dev<int>(o_z), 1,
                                  (int)Core<R>::col_blocks(cy[0]), false, colfft_lds<R>(cy[0].cp, false), c.st,
                                  cy[0].cp.n);
        if (!cz.empty())
            launch_colfft_jobs<R>(run_tab.dev<ColJob<R>>(o_cz), run_tab.dev<int>(o_z), 1,
                                  (int)Core<R>::col_blocks(cz[0]), true, colfft_lds<R>(cz[0].cp, true), c.st,
                                  cz[0].cp.n);
        HIPCHK(hipGetLastError());
        if (!on_device) {
            int hbad = 0;
            HIPCHK(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, c.st));
            HIPCHK(hipMemcpyAsync(f_out, fd, sizeof(Cx<R>) * (size_t)NQ * ntrans, hipMemcpyDeviceToHost, c.st));
            HIPCHK(hipStreamSynchronize(c.st));
            throw_bad(hbad);
        }
    }

    void zero_output(R *f_out, bool on_device)
    {
        if (!f_out) return;
        const int64_t NQ = core.N[0] * core.N[1] * core.N[2];
        const size_t b = sizeof(Cx<R>) * (size_t)NQ * ntrans;
        if (on_device) {
            (void)hipMemsetAsync(f_out, 0, b, core.st);
            (void)hipStreamSynchronize(core.st);
        } else {
            std::memset(f_out, 0, b);
        }
    }
};

template <typename R> static int makeplan_impl(int type, int dim, const int64_t *nm, int iflag, int ntrans,
                                               double eps, Plan<R> **out, const mpsq_opts *o)
{
    if (!out) return MPSQ_ERR_BAD_ARG;
    *out = nullptr;
    if (type != 1 || ntrans < 1) return MPSQ_ERR_BAD_ARG;
    Plan<R> *p = new (std::nothrow) Plan<R>();
    if (!p) return MPSQ_ERR_DEVICE;
    try {
        p->ntrans = ntrans;
        p->core.setup(dim, nm, iflag, eps, o);
    } catch (const Err &e) {
        delete p;
        return e.code;
    } catch (...) {
        delete p;
        return MPSQ_ERR_DEVICE;
    }
    *out = p;
    return p->core.warn;
}

template <typename R> static int setpts_impl(Plan<R> *p, int64_t M, const R *x, const R *y, const R *z, bool dev)
{
    if (!p) return MPSQ_ERR_BAD_ARG;
    try {
        p->setpts(M, x, y, z, dev);
    } catch (const Err &e) {
        p->M = -1;
        return e.code;
    } catch (...) {
        p->M = -1;
        return MPSQ_ERR_DEVICE;
    }
    return 0;
}

template <typename R> static int execute_impl(Plan<R> *p, const R *c, R *f, bool dev)
{
    if (!p) return MPSQ_ERR_BAD_ARG;
    try {
        p->execute(c, f, dev);
    } catch (const Err &e) {
        p->zero_output(f, dev);
        return e.code;
    } catch (...) {
        p->zero_output(f, dev);
        return MPSQ_ERR_DEVICE;
    }
    return p->core.warn;
}

template <typename R> static int perm_impl(Plan<R> *p, int32_t *perm)
{
    if (!p || !perm) return MPSQ_ERR_BAD_ARG;
    if (p->M < 0) return MPSQ_ERR_NOT_SET;
    if (p->M == 0) return 0;
    if (hipMemcpy(perm, p->perm.p, sizeof(int) * (size_t)p->M, hipMemcpyDeviceToHost) != hipSuccess)
        return MPSQ_ERR_DEVICE;
    return 0;
}

template <typename R> static int info_impl(Plan<R> *p, int64_t *out)
{
    if (!p || !out) return MPSQ_ERR_BAD_ARG;
    const Geo &g = p->core.g;
    out[0] = g.n1;
    out[1] = g.n2;
    out[2] = g.n3;
    out[3] = g.w;
    out[4] = g.by;
    out[5] = g.bz;
    out[6] = g.BY;
    out[7] = g.BZ;
    return 0;
}

// ---------------------------------------------------------------------------
// S(Q,w) engine
// ---------------------------------------------------------------------------
constexpr int MAXSP = 64;

template <typename R> struct Engine;

// Device buffers of one run (any number of engine parts): points (staged,
// sorted), the sort's count matrix, bin starts, the two intermediate grids,
// and the job tables.
template <typename R> struct Workspace {
    DevBuf rec, key, cstr, srec, cs, H, Hs, bsum, binstart, binmax, G1, G2, bad, wc, fbl;
    Arena tab;
};

// A part of a run: frames [f0, f0 + b) of one engine's trajectory.
template <typename R> struct Part {
    Engine<R> *e;
    int64_t f0;
    int b;
    const float *pos;  // frame f0 of the trajectory [T][n_tot][4]
    void *amp;         // output of frame f0 ([b][NQ] combined, [b][nact][NQ] per species)
    const R *wsc = nullptr;      // species weights for this part (nullptr: the engine's)
    const double *wsh = nullptr; // host copy of wsc (nullptr: the engine's)
    const float *wtab = nullptr; // per-Q weight table [nact][NQ] (nullptr: the engine's, if any)
};

template <typename R>
static void run_parts(const std::vector<Part<R>> &parts, Workspace<R> &ws, Arena &A, hipStream_t st, StageTimer *tm);

template <typename R> struct Engine {
    Core<R> core;
    int precision = MPSQ_FP32;
    int64_t n_tot = 0;
    int nsp = 0, nact = 0;
    std::vector<int64_t> sp_count, sp_off;
    std::vector<int> act;           // active species ids
    std::vector<int64_t> act_off;   // prefix over active species (size nact+1)
    int64_t Mact = 0;
    int combine = 1;
    StageParams sp{};
    int B = 1;                      // frames per batch when run alone
    DevBuf wsc, wtab, at_ind;
    std::vector<double> wsh;        // active species weights (host copy of wsc)
    Workspace<R> ws;                // buffers when run alone
    // reduction caches
    DevBuf win_d, twT;
    StageTimer tm;
    std::vector<double> win_h;
    std::vector<Cx<R>> twT_h;
    int win_n = -1, win_kind = -1, twT_n = -1;
    RadixPlan rpT{};
    DevBuf twI;                     // FP64 twiddles of the I(Q,t) correlation
    std::vector<Cx<double>> twI_h;
    int twI_n = -1;
    RadixPlan rpI{};

    int64_t nq() const { return core.N[0] * core.N[1] * core.N[2]; }
    int64_t amp_frame_elems() const { return combine ? nq() : nq() * nact; }
    // device bytes of intermediates per frame (points, sort keys, grids)
    // (a frame is one transform when its species are merged, else nact; the
    // default counts nact, which bounds both)
    size_t bytes_per_frame(int transforms = 0) const
    {
        return (size_t)Mact * (2 * sizeof(PtRec<R>) + 2 * sizeof(Cx<R>) + 2 * sizeof(int)) +
               (size_t)(transforms > 0 ? transforms : nact) * sizeof(Cx<R>) *
                   (size_t)(core.n[2] * core.n[1] * core.N[0] +
                            (core.dim == 3 ? core.n[2] * core.N[1] * core.N[0] : 0));
    }

    void create(const mpsq_sqw_spec *s)
    {
        if (!s || (s->dim != 2 && s->dim != 3) || s->nsp < 1 || s->nsp > MAXSP || !s->sp_count || s->n_tot < 1)
            throw Err{MPSQ_ERR_BAD_ARG};
        // q_center = 0: every strength (window and weights times
        // exp(-i 2 pi q0 . r)) is real
        core.setup(s->dim, s->n_modes, s->iflag, s->eps, &s->opts,
                   s->q0[0] == 0.0 && s->q0[1] == 0.0 && s->q0[2] == 0.0);
        precision = s->precision;
        n_tot = s->n_tot;
        nsp = s->nsp;
        combine = s->combine ? 1 : 0;
        sp_count.assign(s->sp_count, s->sp_count + nsp);
        sp_off.assign((size_t)nsp + 1, 0);
        for (int j = 0; j < nsp; ++j) {
            if (sp_count[j] < 0) throw Err{MPSQ_ERR_BAD_ARG};
            sp_off[j + 1] = sp_off[j] + sp_count[j];
        }
        if (sp_off[nsp] != n_tot) throw Err{MPSQ_ERR_BAD_ARG};
        act.clear();
        act_off.assign(1, 0);
        for (int j = 0; j < nsp; ++j)
            if (!s->mask || s->mask[j]) {
                act.push_back(j);
                act_off.push_back(act_off.back() + sp_count[j]);
            }
        nact = (int)act.size();
        if (nact == 0) throw Err{MPSQ_ERR_BAD_ARG};
        Mact = act_off.back();
        if (Mact > INT32_MAX) throw Err{MPSQ_ERR_SIZE};
        const int64_t NQ = nq();
        // weights of the active species
        std::vector<R> wh((size_t)nact);
        for (int k = 0; k < nact; ++k) wh[k] = s->w_scalar ? (R)s->w_scalar[act[k]] : (R)1;
        wsh.assign(wh.begin(), wh.end());
        wsc.ensure(sizeof(R) * (size_t)nact);
        HIPCHK(hipMemcpy(wsc.p, wh.data(), sizeof(R) * (size_t)nact, hipMemcpyHostToDevice));
        if (s->w_table) {
            wtab.ensure(sizeof(float) * (size_t)nact * NQ);
            for (int k = 0; k < nact; ++k)
                HIPCHK(hipMemcpy(wtab.as<float>() + (size_t)k * NQ, s->w_table + (size_t)act[k] * NQ,
                                 sizeof(float) * (size_t)NQ, hipMemcpyHostToDevice));
        }
        // staging parameters
        sp.n_tot = n_tot;
        sp.vacancy_check = s->vacancy_check;
        sp.select_mode = s->select_mode;
        for (int d = 0; d < 3; ++d) {
            sp.n_row[d] = s->n_row[d];
            sp.pmin[d] = s->pmin[d];
            sp.pmax[d] = s->pmax[d];
            sp.centre[d] = s->centre[d];
            sp.tpe[d] = s->tpe[d];
            sp.q0[d] = s->q0[d];
        }
        for (int k = 0; k < 9; ++k) sp.axes[k] = s->axes[k];
        sp.margin = s->margin;
        sp.window = s->window;
        sp.conj = core.iflag > 0 ? 1 : 0;
        if (s->select_mode == 1) {
            if (!s->at_ind) throw Err{MPSQ_ERR_BAD_ARG};
            at_ind.ensure(sizeof(int) * 4 * (size_t)n_tot);
            HIPCHK(hipMemcpy(at_ind.p, s->at_ind, sizeof(int) * 4 * (size_t)n_tot, hipMemcpyHostToDevice));
        }
        // batch size: the caller's, else intermediate buffers of up to 16 GiB
        // (at most 1/16 of the device's memory), <= 128 frames: fewer, longer
        // launches shorten the tails of the persistent passes (C2 on MI355X:
        // 8060 frames/s at 16 frames per launch, 8450 at 64, 8520 at 128)
        const int tpf = (combine && !wtab.p) ? 1 : nact;  // transforms per frame when run alone
        size_t tot = 0;
        if (hipDeviceTotalMem(&tot, core.device) != hipSuccess || tot == 0) tot = (size_t)32 << 30;
        const size_t budget = std::min((size_t)16 << 30, tot / 16);
        int b = s->batch_frames > 0 ? std::min(s->batch_frames, 1024)
                                    : (int)std::min<size_t>(std::max<size_t>(1, budget / bytes_per_frame(tpf)), 128);
        B = std::max(1, b);
    }

    void amplitudes(const float *d_pos, int64_t T, void *d_amp)
    {
        Core<R> &c = core;
        if (!d_pos || !d_amp || T < 1) throw Err{MPSQ_ERR_BAD_ARG};
        HIPCHK(hipSetDevice(c.device));
        ws.bad.ensure(sizeof(int));
        HIPCHK(hipMemsetAsync(ws.bad.p, 0, sizeof(int), c.st));
        for (int64_t f0 = 0; f0 < T; f0 += B) {
            const int b = (int)std::min<int64_t>(B, T - f0);
            std::vector<Part<R>> parts(1);
            parts[0].e = this;
            parts[0].f0 = f0;
            parts[0].b = b;
            parts[0].pos = d_pos + (size_t)f0 * n_tot * 4;
            parts[0].amp = static_cast<Cx<R> *>(d_amp) + (size_t)f0 * amp_frame_elems();
            run_parts<R>(parts, ws, ws.tab, c.st, &tm);
        }
        HIPCHK(hipGetLastError());
        int hbad = 0;
        HIPCHK(hipMemcpyAsync(&hbad, ws.bad.p, sizeof(int), hipMemcpyDeviceToHost, c.st));
        HIPCHK(hipStreamSynchronize(c.st));
        tm.collect();
        throw_bad(hbad);
    }

    // Stage and bin-sort frame f alone (the spread and column passes run too,
    // into a scratch amplitude buffer) and copy the sorted records back:
    // atom index, first kernel cell and selection flag per sorted position.
    void debug_sorted(const float *d_pos, int64_t f, int32_t *atom, int32_t *cell, int32_t *keep, int64_t n)
    {
        Core<R> &c = core;
        if (!d_pos || f < 0 || n != Mact) throw Err{MPSQ_ERR_BAD_ARG};
        HIPCHK(hipSetDevice(c.device));
        DevBuf amp;
        amp.ensure(sizeof(Cx<R>) * (size_t)amp_frame_elems());
        ws.bad.ensure(sizeof(int));
        HIPCHK(hipMemsetAsync(ws.bad.p, 0, sizeof(int), c.st));
        std::vector<Part<R>> parts(1);
        parts[0].e = this;
        parts[0].f0 = f;
        parts[0].b = 1;
        parts[0].pos = d_pos + (size_t)f * n_tot * 4;
        parts[0].amp = amp.p;
        run_parts<R>(parts, ws, ws.tab, c.st, nullptr);
        std::vector<PtRec<R>> h((size_t)std::max<int64_t>(n, 1));
        if (n > 0)
            HIPCHK(hipMemcpyAsync(h.data(), ws.srec.p, sizeof(PtRec<R>) * (size_t)n, hipMemcpyDeviceToHost, c.st));
        int hbad = 0;
        HIPCHK(hipMemcpyAsync(&hbad, ws.bad.p, sizeof(int), hipMemcpyDeviceToHost, c.st));
        HIPCHK(hipStreamSynchronize(c.st));
        throw_bad(hbad);
        for (int64_t k = 0; k < n; ++k) {
            if (atom) atom[k] = h[k].idx;
            if (cell) {
                cell[3 * k] = h[k].lx;
                cell[3 * k + 1] = h[k].ly;
                cell[3 * k + 2] = h[k].lz;
            }
            if (keep) keep[k] = h[k].pad != (R)0 ? 1 : 0;
        }
    }

    void reduce(const void *d_amp, int64_t T, const mpsq_reduce_spec *r, double *d_S, double *d_Stot)
    {
        reduce_on(d_amp, T, r, d_S, d_Stot, core.st, true);
    }

    // Enqueue the reduction on stream st (sync: wait and collect timings).
    // The window and time-FFT twiddles are uploaded when T or the window
    // kind change, so a repeated call enqueues kernels only.
    void reduce_on(const void *d_amp, int64_t T, const mpsq_reduce_spec *r, double *d_S, double *d_Stot,
                   hipStream_t st, bool sync)
    {
        Core<R> &c = core;
        if (!combine || !d_amp || !r || (!d_S && r->n_f > 0) || T < 1) throw Err{MPSQ_ERR_BAD_ARG};
        if (r->kind < 0 || r->kind > 2 || (r->d_qidx && r->n_qidx < 1)) throw Err{MPSQ_ERR_BAD_ARG};
        // the reference's window length rule (mp_sqom56.f90:683-686)
        int64_t n_int = r->n_int;
        if (r->odd_n_int) {
            if (n_int == 0) n_int = T / 2;
            if (r->kind == 0 && n_int % 2 == 0) n_int += 1;
        }
        if (n_int < 1 || n_int > T || r->n_f < 0 || !(r->nat > 0)) throw Err{MPSQ_ERR_BAD_ARG};
        if (r->kind != 0 && (r->m_lo < 0 || r->m_lo + r->n_f - 1 > T - n_int)) throw Err{MPSQ_ERR_BAD_ARG};
        HIPCHK(hipSetDevice(c.device));
        const int64_t NQ = c.N[0] * c.N[1] * c.N[2];
        SqwParams p{};
        p.T = T;
        p.NQs = NQ;
        p.NQ = r->d_qidx ? r->n_qidx : NQ;
        p.qidx = r->d_qidx;
        p.kind = r->kind;
        p.n_int = (int)n_int;
        p.n_frame = T - n_int + 1;
        p.m_lo = r->m_lo;
        p.n_f = (int)r->n_f;
        p.win_kind = r->win_kind;
        p.subtract_mean = r->subtract_mean;
        p.norm_s = r->dt / (r->nat * r->nat * 0.5 * (double)p.n_frame * (double)p.n_int);
        if (p.kind != 0) p.norm_s = 1.0 / (r->nat * r->nat * (p.kind == 1 ? (double)p.n_int : 1.0));
        p.norm_tot = 1.0 / ((double)T * r->nat * r->nat);
        const Cx<R> *amp = static_cast<const Cx<R> *>(d_amp);
        if (p.kind != 0) {
            // many lags: FFT correlation (three FP64 LDS transforms per Q);
            // a few lags, or a T without a radix plan / LDS room: direct sums
            RadixPlan rp{};
            const bool fft = p.n_f >= IQT_FFT_MIN_LAGS && T <= iqt_fft_tile_elems() &&
                             iqt_fft_lds(T, 1) <= SQW_MAX_LDS && radix_plan((int)T, rp);
            const int ev = tm.begin(ST_REDUCE, st);
            if (fft) {
                if (twI_n != (int)T) {
                    upload_twiddles<double>(twI, (int)T, st, twI_h);
                    twI_n = (int)T;
                    rpI = rp;
                }
                // columns per block: the most that keep two blocks per CU
                // (or one, for long series)
                int cols = 1;
                while (2 * cols * T <= iqt_fft_tile_elems() && 2 * cols <= SQW_MAX_COLS &&
                       iqt_fft_lds(T, 2 * cols) <= IQT_FFT_LDS_TARGET)
                    cols *= 2;
                launch_iqt_fft<R>(p, rpI, cols, amp, twI.as<Cx<double>>(), r->d_qfac, d_S, d_Stot, st);
            } else {
                launch_iqt<R>(p, amp, r->d_qfac, d_S, d_Stot, st);
            }
            tm.end(ev, st);
            HIPCHK(hipGetLastError());
            if (!sync) return;
            HIPCHK(hipStreamSynchronize(st));
            tm.collect();
            return;
        }
        RadixPlan rp{};
        const bool single = (p.n_frame == 1) && (T <= sqw_tile_elems<R>()) &&
                            sqw_single_lds<R>(T, 1) <= SQW_MAX_LDS && radix_plan((int)T, rp);
        if (p.n_frame == 1) {
            if (win_n != (int)T || win_kind != r->win_kind) {
                win_h.resize((size_t)T);
                for (int64_t i = 0; i < T; ++i) {
                    if (T <= 1) win_h[i] = 1.0;
                    else if (r->win_kind == 0) win_h[i] = 0.5 * (1.0 - std::cos(2.0 * M_PI * (double)i / (double)(T - 1)));
                    else win_h[i] = 0.5 * (1.0 - std::cos(2.0 * M_PI * (double)(i + 1) / (double)(T + 1)));
                }
                win_d.ensure(sizeof(double) * (size_t)T);
                HIPCHK(hipMemcpyAsync(win_d.p, win_h.data(), sizeof(double) * (size_t)T, hipMemcpyHostToDevice, st));
                win_n = (int)T;
                win_kind = r->win_kind;
            }
        }
        if (single) {
            if (twT_n != (int)T) {
                upload_twiddles<R>(twT, (int)T, st, twT_h);
                twT_n = (int)T;
                rpT = rp;
            }
            int cols = 1;
            while (2 * cols * T <= sqw_tile_elems<R>() && 2 * cols <= SQW_MAX_COLS &&
                   sqw_single_lds<R>(T, 2 * cols) <= SQW_MAX_LDS)
                cols *= 2;
            const int ev = tm.begin(ST_REDUCE, st);
            launch_sqw_single<R>(p, rpT, cols, amp, twT.as<Cx<R>>(), win_d.as<double>(), r->d_qfac, d_S, d_Stot, st);
            tm.end(ev, st);
        } else if (p.n_frame == 1) {
            // one window, T without a radix plan: one DFT recurrence per frequency
            const int ev = tm.begin(ST_REDUCE, st);
            launch_sqw_window1<R>(p, amp, r->d_qfac, win_d.as<double>(), d_S, d_Stot, st);
            tm.end(ev, st);
        } else {
            const int ev = tm.begin(ST_REDUCE, st);
            launch_sqw_sliding<R>(p, amp, r->d_qfac, d_S, d_Stot, st);
            tm.end(ev, st);
        }
        HIPCHK(hipGetLastError());
        if (!sync) return;
        HIPCHK(hipStreamSynchronize(st));
        tm.collect();
    }
};

// ---------------------------------------------------------------------------
// One run of engine parts: every phase is one launch over all parts.
//   stage    points of all parts (active species, species-major per frame)
//   sort     count / scan / stable scatter / bin starts over all transforms
//   spread   one persistent launch over the tiles of all transforms
//            (parts are grouped by kernel width; one launch per width)
//   columns  y pass over all transforms, final pass (species combine)
// ---------------------------------------------------------------------------
template <typename R>
static void run_parts(const std::vector<Part<R>> &parts, Workspace<R> &ws, Arena &A, hipStream_t st, StageTimer *tm)
{
    StageTimer dummy;
    if (!tm) tm = &dummy;
    const int np = (int)parts.size();
    if (np == 0) return;
    // ---- layout of the parts in the shared buffers
    std::vector<int64_t> p0(np + 1, 0), t0(np + 1, 0), g1o(np + 1, 0), g2o(np + 1, 0), bso(np + 1, 0);
    int64_t hsz = 0;
    std::vector<StageJob> stj(np);
    std::vector<int> stb(np);
    std::vector<SortChunk> chunks;
    std::vector<TransformInfo> tinfo;
    int max_nbins = 1;
    int sblocks = 0;
    // A combined map with scalar species weights is one transform per frame
    // over all active atoms with strengths w_s c_j (the NUFFT is linear in
    // the strengths); per-species amplitudes or per-Q weight tables keep one
    // transform per species.
    std::vector<char> merged(np);
    std::vector<float> wbound(np, 1.0f);
    auto part_wtab = [&](int q) -> const float * {
        const Engine<R> &e = *parts[q].e;
        return !e.combine ? nullptr : parts[q].wtab ? parts[q].wtab : (e.wtab.p ? e.wtab.template as<float>() : nullptr);
    };
    for (int q = 0; q < np; ++q) {
        const Engine<R> &e = *parts[q].e;
        merged[q] = e.combine && !part_wtab(q);
    }
    // Real strengths: the real form of the narrow spread (rows in pairs),
    // and, without a per-Q weight table, the half spectrum through the
    // column passes (F(-k) = conj F(k); the final pass writes the mirrors)
    std::vector<char> real_ok(np), herm(np);
    for (int q = 0; q < np; ++q) {
        const Core<R> &c = parts[q].e->core;
        const int rows = ((c.dim >= 2) ? c.g.by : 1) * ((c.dim >= 3) ? c.g.bz : 1);
        real_ok[q] = c.real && spread_real_ok(rows, c.g.n1, c.g.w);
        // FP32: only jobs of the real narrow form (the complex narrow kernel
        // writes whole rows); FP64: the wide kernel keeps the half itself
        herm[q] = (sizeof(R) == 8 ? c.real_str : real_ok[q]) && !part_wtab(q);
    }
    auto tpf = [&](int q) { return merged[q] ? 1 : parts[q].e->nact; };  // transforms per frame
    for (int q = 0; q < np; ++q) {
        const Engine<R> &e = *parts[q].e;
        const Core<R> &c = e.core;
        const int ntr = parts[q].b * tpf(q);
        p0[q + 1] = p0[q] + (int64_t)parts[q].b * e.Mact;
        t0[q + 1] = t0[q] + ntr;
        g1o[q + 1] = g1o[q] + (int64_t)ntr * c.n[2] * c.n[1] * c.h1x(herm[q]);
        g2o[q + 1] = g2o[q] + (c.dim == 3 ? (int64_t)ntr * c.n[2] * (herm[q] ? (c.N[1] + 1) * c.h1x(true)
                                                                           : c.N[1] * c.N[0])
                                          : 0);
        bso[q + 1] = bso[q] + (int64_t)ntr * (c.g.nbins + 1);
        max_nbins = std::max(max_nbins, c.g.nbins);
        StageJob &J = stj[q];
        std::memset(&J, 0, sizeof(J));
        J.g = c.g;
        J.sp = e.sp;
        J.pos = reinterpret_cast<const float4 *>(parts[q].pos);
        J.at_ind = e.at_ind.p ? e.at_ind.template as<int4>() : nullptr;
        J.npts = (int64_t)parts[q].b * e.Mact;
        J.out0 = p0[q];
        J.nspan = e.n_tot;
        J.mact = e.Mact;
        J.blk0 = sblocks;
        J.nact = e.nact;
        const double *wq = parts[q].wsh ? parts[q].wsh : e.wsh.data();
        double wmax = 0.0;
        for (int a = 0; a < e.nact; ++a) {
            J.act_src[a] = e.sp_off[e.act[a]];
            J.act_off[a] = e.act_off[a];
            J.wst[a] = merged[q] ? wq[a] : 1.0;
            wmax = std::max(wmax, std::fabs(wq[a]));
        }
        J.act_off[e.nact] = e.act_off[e.nact];
        // bound of the staged strengths' components (|w_s c_j| <= |w_s|),
        // rounded up past the staging's FP32 rounding
        if (merged[q]) wbound[q] = (float)(wmax * (1.0 + 1.0 / (1 << 20)));
        stb[q] = sblocks;
        sblocks += (int)((J.npts + 255) / 256);
        // sort chunks of every transform (frame fb, active species a, or
        // frame fb when merged)
        const int64_t CH = c.sort_chunk;
        for (int fb = 0; fb < parts[q].b; ++fb)
            for (int a = 0; a < tpf(q); ++a) {
                const int64_t pa = p0[q] + (int64_t)fb * e.Mact + (merged[q] ? 0 : e.act_off[a]);
                const int64_t pb = p0[q] + (int64_t)fb * e.Mact + (merged[q] ? e.Mact : e.act_off[a + 1]);
                const int nch = (int)((pb - pa + CH - 1) / CH);
                for (int k = 0; k < nch; ++k) {
                    SortChunk ch;
                    ch.p0 = pa + (int64_t)k * CH;
                    ch.p1 = std::min(pb, pa + (int64_t)(k + 1) * CH);
                    ch.pbase = pa;
                    ch.hbase = hsz;
                    ch.nch = nch;
                    ch.lc = k;
                    ch.nbins = c.g.nbins;
                    ch.keybits = c.keybits();
                    chunks.push_back(ch);
                }
                TransformInfo ti;
                ti.p0 = pa;
                ti.p1 = pb;
                ti.hbase = hsz;
                ti.bs0 = bso[q] + (int64_t)(fb * tpf(q) + a) * (c.g.nbins + 1);
                ti.nch = nch;
                ti.nbins = c.g.nbins;
                c.segments(ti.sxsh, ti.nseg);
                tinfo.push_back(ti);
                hsz += (int64_t)c.g.nbins * nch;
            }
    }
    const int64_t npts = p0[np], ntr_all = t0[np];
    // ---- buffers
    const size_t nb = (size_t)std::max<int64_t>(npts, 1);
    ws.rec.ensure(sizeof(PtRec<R>) * nb);
    ws.srec.ensure(sizeof(PtRec<R>) * nb);
    ws.key.ensure(sizeof(int) * nb);
    ws.cstr.ensure(sizeof(Cx<R>) * nb);
    ws.cs.ensure(sizeof(Cx<R>) * nb);
    ws.H.ensure(sizeof(int) * (size_t)std::max<int64_t>(hsz, 1));
    ws.Hs.ensure(sizeof(int64_t) * (size_t)std::max<int64_t>(hsz, 1));
    ws.bsum.ensure(sizeof(int64_t) * (size_t)scan_bsum_size(hsz));
    ws.binstart.ensure(sizeof(int64_t) * (size_t)std::max<int64_t>(bso[np], 1));
    if (sizeof(R) == 4) ws.binmax.ensure(sizeof(int) * (size_t)std::max<int64_t>(bso[np], 1));
    ws.G1.ensure(sizeof(Cx<R>) * (size_t)std::max<int64_t>(g1o[np], 1));
    ws.G2.ensure(sizeof(Cx<R>) * (size_t)std::max<int64_t>(g2o[np], 1));
    ws.bad.ensure(sizeof(int));
    ws.wc.ensure(sizeof(int) * XCD_CTR_INTS);
    // ---- spread and column job tables.  Parts are grouped by kernel width
    // and by how many workgroups of their LDS size fit a CU (a launch's LDS
    // is the largest of its jobs, so a small-grid job launched with a large
    // one would run at the large one's occupancy); one launch per group.
    // (real and complex strengths never share a launch: a real geometry may
    // have more rows than the complex kernel takes)
    auto sclass = [&](int q) {
        const Core<R> &c = parts[q].e->core;
        const int res = (int)std::min<size_t>(8, (size_t)(156 * 1024) / std::max<size_t>(c.spread_lds_size(), 1));
        return (real_ok[q] ? 4096 : 0) + c.g.w * 16 + res;
    };
    std::vector<int> widths;  // group keys
    for (int q = 0; q < np; ++q)
        if (std::find(widths.begin(), widths.end(), sclass(q)) == widths.end()) widths.push_back(sclass(q));
    struct SpreadLaunch {
        int w;
        size_t o_jobs, o_t0, o_xr, lds;
        int njobs;
        size_t lds_real;  // > 0: the real narrow kernel (FP32, every job's strengths real)
    };
    std::vector<SpreadLaunch> sl;
    int max_tiles = 0;  // tiles of the largest spread launch
    std::vector<std::vector<SpreadJob<R>>> sjobs(widths.size());
    std::vector<std::vector<int>> st0(widths.size()), sxr(widths.size());
    for (size_t wi = 0; wi < widths.size(); ++wi) {
        std::vector<double> cost;
        std::vector<int> ntl;
        int tile = 0;
        size_t lds = 0, lds_real = 0;
        bool real = sizeof(R) == 4;
        for (int q = 0; q < np; ++q) {
            const Engine<R> &e = *parts[q].e;
            const Core<R> &c = e.core;
            if (sclass(q) != widths[wi]) continue;
            const int rows = ((c.dim >= 2) ? c.g.by : 1) * ((c.dim >= 3) ? c.g.bz : 1);
            real = real && real_ok[q];
            if (real) lds_real = std::max(lds_real, spread_lds_real(rows, c.g.n1, c.g.w));
            const int ntr = parts[q].b * tpf(q);
            SpreadJob<R> J = c.spreadjob(ntr, ws.srec.template as<PtRec<R>>(), npts, ws.bad.template as<int>(),
                                         ws.cs.template as<Cx<R>>(), 0,
                                         ws.binstart.template as<int64_t>() + bso[q],
                                         sizeof(R) == 4 ? ws.binmax.template as<int>() + bso[q] : nullptr, 1,
                                         ws.G1.template as<Cx<R>>() + g1o[q], nullptr, 0);
            J.amax_s = wbound[q];
            J.tmajor = merged[q] ? 1 : 0;
            J.g1_len = (int)c.h1x(herm[q]);
            J.g1_keep = (int)c.h1v(herm[q]);
            J.tile0 = tile;
            tile += J.ntiles * ntr;
            sjobs[wi].push_back(J);
            st0[wi].push_back(J.tile0);
            cost.push_back(c.tile_cost(e.Mact / tpf(q)));
            ntl.push_back(J.ntiles * ntr);
            lds = std::max(lds, c.spread_lds_size());
        }
        sxr[wi] = xcd_ranges(cost, ntl);
        sl.push_back(SpreadLaunch{(widths[wi] % 4096) / 16, 0, 0, 0, lds, (int)sjobs[wi].size(), real ? lds_real : 0});
        max_tiles = std::max(max_tiles, tile);
    }
    ws.fbl.ensure(sizeof(int) * (size_t)spread_max_wide_passes() * (size_t)std::max(max_tiles, 1));
    std::vector<ColJob<R>> cyj, czj;
    std::vector<int> cyb, czb;
    int cyn = 0, czn = 0;
    size_t cyl = 0, czl = 0;
    for (int q = 0; q < np; ++q) {
        const Engine<R> &e = *parts[q].e;
        const Core<R> &c = e.core;
        const int ntr = parts[q].b * tpf(q);
        const int ngrp = e.combine ? parts[q].b : ntr;
        const int nspg = merged[q] ? 1 : e.combine ? e.nact : 1;
        Cx<R> *g1 = ws.G1.template as<Cx<R>>() + g1o[q];
        Cx<R> *g2 = ws.G2.template as<Cx<R>>() + g2o[q];
        Cx<R> *out = static_cast<Cx<R> *>(parts[q].amp);
        // merged: weights are in the strengths; per-Q weights (X-ray form
        // factors) are applied per species after the last FFT
        const R *wsc = (e.combine && !merged[q]) ? (parts[q].wsc ? parts[q].wsc : e.wsc.template as<R>()) : nullptr;
        const float *wtab = part_wtab(q);
        if (c.dim == 3) {
            ColJob<R> y = c.coljob(1, false, 1, ntr, g1, g2, nullptr, nullptr, herm[q]);
            y.blk0 = cyn;
            cyn += (int)Core<R>::col_blocks(y);
            cyl = std::max(cyl, colfft_lds<R>(y.cp, false));
            cyj.push_back(y);
            cyb.push_back(y.blk0);
            ColJob<R> z = c.coljob(2, true, nspg, ngrp, g2, out, wsc, wtab, herm[q]);
            z.blk0 = czn;
            czn += (int)Core<R>::col_blocks(z);
            czl = std::max(czl, colfft_lds<R>(z.cp, true));
            czj.push_back(z);
            czb.push_back(z.blk0);
        } else {
            ColJob<R> z = c.coljob(1, true, nspg, ngrp, g1, out, wsc, wtab, herm[q]);
            z.blk0 = czn;
            czn += (int)Core<R>::col_blocks(z);
            czl = std::max(czl, colfft_lds<R>(z.cp, true));
            czj.push_back(z);
            czb.push_back(z.blk0);
        }
    }
    // ---- tables
    A.begin();
    const size_t o_stj = A.put(stj), o_stb = A.put(stb), o_ch = A.put(chunks), o_ti = A.put(tinfo);
    for (size_t wi = 0; wi < sl.size(); ++wi) {
        sl[wi].o_jobs = A.put(sjobs[wi]);
        sl[wi].o_t0 = A.put(st0[wi]);
        sl[wi].o_xr = A.put(sxr[wi]);
    }
    const size_t o_cy = A.put(cyj), o_cyb = A.put(cyb), o_cz = A.put(czj), o_czb = A.put(czb);
    A.upload(st);
    // ---- launches
    int ev = tm->begin(ST_STAGE, st);
    launch_stage_jobs<R>(A.dev<StageJob>(o_stj), A.dev<int>(o_stb), np, sblocks, ws.rec.template as<PtRec<R>>(),
                         ws.key.template as<int>(), ws.cstr.template as<Cx<R>>(), ws.bad.template as<int>(), st);
    tm->end(ev, st);
    ev = tm->begin(ST_SORT, st);
    const int nch = (int)chunks.size();
    launch_sort_count(ws.key.template as<int>(), A.dev<SortChunk>(o_ch), nch, max_nbins, ws.H.template as<int>(), st);
    launch_scan(ws.H.template as<int>(), hsz, ws.bsum.template as<int64_t>(), ws.Hs.template as<int64_t>(), st);
    launch_sort_scatter<R>(ws.key.template as<int>(), A.dev<SortChunk>(o_ch), nch, max_nbins,
                           ws.Hs.template as<int64_t>(), ws.rec.template as<PtRec<R>>(),
                           ws.cstr.template as<Cx<R>>(), ws.srec.template as<PtRec<R>>(),
                           ws.cs.template as<Cx<R>>(), nullptr, st);
    launch_binstart(ws.Hs.template as<int64_t>(), A.dev<TransformInfo>(o_ti), (int)ntr_all, max_nbins,
                    ws.binstart.template as<int64_t>(), st);
    if (sizeof(R) == 4)
        launch_binmax<R>(ws.srec.template as<PtRec<R>>(), ws.binstart.template as<int64_t>(),
                         A.dev<TransformInfo>(o_ti), (int)ntr_all, max_nbins, ws.binmax.template as<int>(), st);
    tm->end(ev, st);
    ev = tm->begin(ST_SPREAD, st);
    for (const SpreadLaunch &L : sl)
        launch_spread_jobs<R>(A.dev<SpreadJob<R>>(L.o_jobs), A.dev<int>(L.o_t0), L.njobs, L.w, A.dev<int>(L.o_xr),
                              ws.wc.template as<int>(), ws.fbl.template as<int>(), L.lds, st, L.lds_real);
    tm->end(ev, st);
    if (cyn > 0) {
        ev = tm->begin(ST_COLPASS, st);
        launch_colfft_jobs<R>(A.dev<ColJob<R>>(o_cy), A.dev<int>(o_cyb), (int)cyj.size(), cyn, false, cyl, st, 0);
        tm->end(ev, st);
    }
    ev = tm->begin(ST_FINAL, st);
    int czmax = 0;
    for (const ColJob<R> &cj : czj) czmax = std::max(czmax, cj.cp.n);
    launch_colfft_jobs<R>(A.dev<ColJob<R>>(o_cz), A.dev<int>(o_czb), (int)czj.size(), czn, true, czl, st, czmax);
    tm->end(ev, st);
}

// ---------------------------------------------------------------------------
// Batch of S(Q,w) jobs (uneven geometries, frame counts, species): all jobs'
// frames go through one launch per phase (memory permitting: rounds cut at a
// device-memory budget), the reductions follow on the same stream, and the
// whole sequence can be captured once in a hipGraph and replayed.
// ---------------------------------------------------------------------------
template <typename R> struct Batch {
    std::vector<Engine<R> *> jobs;
    hipStream_t st = nullptr;
    bool own_stream = false;
    int device = 0;
    size_t budget = (size_t)16 << 30;  // intermediates per round
    Workspace<R> ws;
    std::vector<Arena *> arenas;  // job tables of each round (all alive in a captured graph)
    StageTimer tm;
    // captured sequence
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // pointers, sizes and every reduction argument the graph was captured
    // with (the reductions' parameters are baked into the captured launches)
    std::vector<uint64_t> gkey;

    ~Batch()
    {
        if (st) (void)hipStreamSynchronize(st);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (Engine<R> *e : jobs) delete e;
        for (Arena *a : arenas) delete a;
        if (own_stream && st) (void)hipStreamDestroy(st);
    }

    void create(int njobs, const mpsq_sqw_spec *specs)
    {
        if (njobs < 1 || njobs > MAX_JOBS || !specs) throw Err{MPSQ_ERR_BAD_ARG};
        device = specs[0].opts.device;
        HIPCHK(hipSetDevice(device));
        if (specs[0].opts.stream) {
            st = (hipStream_t)specs[0].opts.stream;
        } else {
            HIPCHK(hipStreamCreate(&st));  // blocking: ordered with the legacy default stream
            own_stream = true;
        }
        for (int j = 0; j < njobs; ++j) {
            if (specs[j].precision != (sizeof(R) == 4 ? MPSQ_FP32 : MPSQ_FP64) || !specs[j].combine)
                throw Err{MPSQ_ERR_BAD_ARG};
            mpsq_sqw_spec sj = specs[j];
            sj.opts.device = device;
            sj.opts.stream = st;
            Engine<R> *e = new Engine<R>();
            jobs.push_back(e);
            e->create(&sj);
        }
    }

    // frames of every job in rounds of at most `budget` bytes of intermediates
    std::vector<std::vector<Part<R>>> rounds(const float *const *d_pos, const int64_t *T, void *const *d_amp) const
    {
        std::vector<std::vector<Part<R>>> out(1);
        size_t used = 0;
        for (size_t j = 0; j < jobs.size(); ++j) {
            const Engine<R> &e = *jobs[j];
            const size_t per = std::max<size_t>(e.bytes_per_frame(), 1);
            int64_t f = 0;
            while (f < T[j]) {
                int64_t room = (int64_t)((budget > used ? budget - used : 0) / per);
                if (room < 1 && !out.back().empty()) {
                    out.emplace_back();
                    used = 0;
                    continue;
                }
                const int b = (int)std::max<int64_t>(1, std::min<int64_t>(room, T[j] - f));
                Part<R> p;
                p.e = jobs[j];
                p.f0 = f;
                p.b = b;
                p.pos = d_pos[j] + (size_t)f * e.n_tot * 4;
                p.amp = static_cast<Cx<R> *>(d_amp[j]) + (size_t)f * e.amp_frame_elems();
                out.back().push_back(p);
                used += per * (size_t)b;
                f += b;
            }
        }
        return out;
    }

    void enqueue(const float *const *d_pos, const int64_t *T, void *const *d_amp, const mpsq_reduce_spec *r,
                 double *const *d_S, double *const *d_Stot, StageTimer *timer)
    {
        ws.bad.ensure(sizeof(int));
        launch_zero_i32(ws.bad.template as<int>(), 1, st);
        const std::vector<std::vector<Part<R>>> rs = rounds(d_pos, T, d_amp);
        while (arenas.size() < rs.size()) arenas.push_back(new Arena());
        for (size_t k = 0; k < rs.size(); ++k) run_parts<R>(rs[k], ws, *arenas[k], st, timer);
        if (r)
            for (size_t j = 0; j < jobs.size(); ++j)
                jobs[j]->reduce_on(d_amp[j], T[j], &r[j], d_S ? d_S[j] : nullptr, d_Stot ? d_Stot[j] : nullptr, st,
                                   false);
    }

    void check_bad()
    {
        int hbad = 0;
        HIPCHK(hipMemcpyAsync(&hbad, ws.bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        throw_bad(hbad);
    }

    // use_graph: the first call with a given set of pointers runs eagerly
    // (buffers and tables are allocated), the second captures the sequence,
    // later calls replay it.
    void run(const float *const *d_pos, const int64_t *T, void *const *d_amp, const mpsq_reduce_spec *r,
             double *const *d_S, double *const *d_Stot, bool use_graph)
    {
        HIPCHK(hipSetDevice(device));
        std::vector<uint64_t> key;
        auto kp = [&](const void *p) { key.push_back((uint64_t)(uintptr_t)p); };
        auto kd = [&](double v) {
            uint64_t b;
            std::memcpy(&b, &v, sizeof b);
            key.push_back(b);
        };
        key.push_back(r ? 1u : 0u);
        for (size_t j = 0; j < jobs.size(); ++j) {
            kp(d_pos[j]);
            key.push_back((uint64_t)T[j]);
            kp(d_amp[j]);
            kp(d_S ? d_S[j] : nullptr);
            kp(d_Stot ? d_Stot[j] : nullptr);
            if (r) {
                key.push_back((uint64_t)r[j].n_int);
                key.push_back((uint64_t)r[j].win_kind);
                kd(r[j].dt);
                kd(r[j].nat);
                key.push_back((uint64_t)r[j].m_lo);
                key.push_back((uint64_t)r[j].n_f);
                key.push_back((uint64_t)r[j].subtract_mean);
                kp(r[j].d_qfac);
                key.push_back((uint64_t)r[j].kind);
                key.push_back((uint64_t)r[j].odd_n_int);
                kp(r[j].d_qidx);
                key.push_back((uint64_t)r[j].n_qidx);
            }
        }
        if (use_graph && gexec && key == gkey) {
            HIPCHK(hipGraphLaunch(gexec, st));
            check_bad();
            return;
        }
        if (use_graph && key == gkey && !gexec) {
            // capture: everything is allocated and the tables are current
            HIPCHK(hipStreamSynchronize(st));
            for (Arena *a : arenas) a->begin();  // nothing pending may be waited on inside the capture
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            enqueue(d_pos, T, d_amp, r, d_S, d_Stot, nullptr);
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamEndCapture(st, &g));
            graph = g;
            HIPCHK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
            HIPCHK(hipGraphLaunch(gexec, st));
            check_bad();
            return;
        }
        if (gexec) {
            (void)hipGraphExecDestroy(gexec);
            gexec = nullptr;
        }
        if (graph) {
            (void)hipGraphDestroy(graph);
            graph = nullptr;
        }
        gkey = key;
        enqueue(d_pos, T, d_amp, r, d_S, d_Stot, &tm);
        HIPCHK(hipGetLastError());
        check_bad();
        tm.collect();
    }
};

// ---- spread tile autotuner ----
template <typename R>
static int autotune_impl(const mpsq_sqw_spec *spec, const float *d_pos, int64_t T, int *tby, int *tbz, double *ms)
{
    const int64_t Tt = std::min<int64_t>(T, 4);
    // the geometry (fine grid, width) of the heuristic choice
    Engine<R> probe;
    probe.create(spec);
    const Core<R> &c = probe.core;
    const int dim = c.dim, w = c.g.w, n1 = c.g.n1, n2 = c.g.n2, n3 = c.g.n3;
    const int bmax = spread_max_tile_b();
    DevBuf amp;
    amp.ensure(sizeof(Cx<R>) * (size_t)Tt * (size_t)probe.amp_frame_elems());
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    double best = 1e300, t_heur = 1e300;
    int bb[2] = {c.g.by, c.g.bz};
    const int hb[2] = {c.g.by, c.g.bz};  // the shape the engine takes untuned (heuristic or earlier tuning)
    for (int by = 1; by <= (dim >= 2 ? bmax : 1); by *= 2)
        for (int bz = 1; bz <= (dim >= 3 ? bmax : 1); bz *= 2) {
            if (by * bz > c.rows_max) continue;
            if (dim >= 2 && (n2 % by || by + w > n2)) continue;
            if (dim >= 3 && (n3 % bz || bz + w > n3)) continue;
            if (spread_lds<R>(by * bz, n1, w, c.real) > 156 * 1024) continue;
            mpsq_sqw_spec sj = *spec;
            sj.opts.tile_by = by;
            sj.opts.tile_bz = bz;
            Engine<R> e;
            e.create(&sj);
            e.amplitudes(d_pos, Tt, amp.p);  // warm-up: buffers, tables
            HIPCHK(hipEventRecord(e0, e.core.st));
            e.amplitudes(d_pos, Tt, amp.p);
            HIPCHK(hipEventRecord(e1, e.core.st));
            HIPCHK(hipEventSynchronize(e1));
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, e0, e1));
            if (by == hb[0] && bz == hb[1]) t_heur = t;
            if (t < best) {
                best = t;
                bb[0] = by;
                bb[1] = bz;
            }
        }
    // within 2% of the fastest that shape stays (timing noise; the
    // heuristic's tie-breaks favour smaller covering bounds, i.e. a finer
    // FP32 fixed-point scale)
    if (t_heur <= 1.02 * best) {
        bb[0] = hb[0];
        bb[1] = hb[1];
        best = t_heur;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    {
        const TuneKey key{{dim, n1, n2, n3, w, sizeof(R) == 8 ? 1 : (c.real ? 2 : 0)}};
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[key] = std::make_pair(bb[0], bb[1]);
    }
    if (tby) *tby = bb[0];
    if (tbz) *tbz = bb[1];
    if (ms) *ms = best / (double)Tt;
    return 0;
}

// ---------------------------------------------------------------------------
// Map server: resident trajectories, HBM-budget admission, continuous batching
// ---------------------------------------------------------------------------
struct SrvTraj {
    const float *pos = nullptr;
    int64_t T = 0, n_tot = 0;
    int nsp = 0;
    std::vector<int64_t> counts;
    double tpe[3] = {1, 1, 1}, qs[3] = {1, 1, 1};
};

struct SrvReq {
    int64_t ticket = 0;
    mpsq_request r{};
    std::vector<int32_t> mask;
    std::vector<double> w, xff;
    int state = 0;  // 0 queued, 1 admitted, 2 done, < 0 error
    std::chrono::steady_clock::time_point t_sub, t_admit, t_done;
    Engine<float> *e = nullptr;
    int64_t nq = 0;
    size_t bytes = 0;
    void *amp = nullptr;
    double *S = nullptr, *St = nullptr, *coef = nullptr;
    float *wsc = nullptr, *wtab = nullptr;
    std::vector<double> wact;  // host copy of wsc
    int64_t next = 0;  // frames already given to a step
    bool reducing = false;
    hipEvent_t done_ev = nullptr;
    mpsq_reduce_spec red{};
};

struct Server {
    int device = 0;
    size_t budget = (size_t)64 << 30, used = 0, step_budget = (size_t)8 << 30;
    int max_active = 64;
    hipStream_t st = nullptr;
    std::vector<SrvTraj> trajs;
    std::mutex mu;
    std::condition_variable cv, idle_cv;
    std::deque<SrvReq *> queue;
    std::map<int64_t, SrvReq *> reqs;
    std::vector<SrvReq *> active;  // admitted and not finished (worker thread only)
    int64_t next_ticket = 1, pending = 0, steps = 0, ndone = 0;
    double busy_ms = 0.0;
    bool stop = false;
    std::thread worker;
    Workspace<float> ws;
    Arena ring[3];
    int ring_k = 0;
    std::map<std::vector<int64_t>, Engine<float> *> engines;  // by (traj, modes, eps, mask)

    void start(const mpsq_server_opts *o)
    {
        device = o ? o->device : 0;
        if (o && o->hbm_budget > 0) budget = (size_t)o->hbm_budget;
        if (o && o->step_budget > 0) step_budget = (size_t)o->step_budget;
        if (o && o->max_active > 0) max_active = o->max_active;
        HIPCHK(hipSetDevice(device));
        // non-blocking: the synchronous copies of request set-up (legacy
        // stream) must not wait for the steps in flight
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        worker = std::thread([this] { loop(); });
    }

    ~Server()
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv.notify_all();
        if (worker.joinable()) worker.join();
        if (st) (void)hipStreamSynchronize(st);
        for (auto &kv : reqs) free_req(kv.second);
        for (auto &kv : engines) delete kv.second;
        if (st) (void)hipStreamDestroy(st);
    }

    void free_req(SrvReq *q)
    {
        void *ptrs[6] = {q->amp, q->S, q->St, q->coef, q->wsc, q->wtab};
        for (void *p : ptrs)
            if (p) (void)hipFreeAsync(p, st);
        if (q->done_ev) (void)hipEventDestroy(q->done_ev);
        delete q;
    }

    static size_t req_bytes(const SrvReq &q, int nact, int64_t nq)
    {
        size_t b = sizeof(float2) * (size_t)q.r.T * (size_t)nq;               // amplitudes
        b += sizeof(double) * (size_t)(std::max<int64_t>(q.r.n_f, 1) + 1) * (size_t)nq;  // S, S_tot
        if (q.r.weight_kind == 1) b += sizeof(float) * (size_t)nact * (size_t)nq;
        return b + 4096;
    }

    Engine<float> *engine_for(const SrvReq &q)
    {
        const SrvTraj &tj = trajs[q.r.traj];
        std::vector<int64_t> key{q.r.traj, q.r.n_modes[0], q.r.n_modes[1], q.r.n_modes[2]};
        int64_t eb;
        std::memcpy(&eb, &q.r.eps, sizeof(eb));
        key.push_back(eb);
        for (int j = 0; j < tj.nsp; ++j) key.push_back(q.mask.empty() ? 1 : q.mask[j]);
        auto it = engines.find(key);
        if (it != engines.end()) return it->second;
        mpsq_sqw_spec sp;
        mpsq_sqw_default_spec(&sp);
        sp.dim = 3;
        for (int d = 0; d < 3; ++d) sp.n_modes[d] = q.r.n_modes[d];
        sp.eps = q.r.eps;
        sp.precision = MPSQ_FP32;
        sp.n_tot = tj.n_tot;
        sp.nsp = tj.nsp;
        sp.sp_count = tj.counts.data();
        sp.mask = q.mask.empty() ? nullptr : q.mask.data();
        for (int d = 0; d < 3; ++d) sp.tpe[d] = tj.tpe[d];
        sp.opts.device = device;
        Engine<float> *e = new Engine<float>();
        try {
            e->create(&sp);
        } catch (...) {
            delete e;
            throw;
        }
        engines[key] = e;
        return e;
    }

    // set-up of an admitted request (worker thread)
    void prepare(SrvReq *q)
    {
        Engine<float> *e = engine_for(*q);
        q->e = e;
        q->nq = e->nq();
        const int nact = e->nact;
        HIPCHK(hipMallocAsync(&q->amp, sizeof(float2) * (size_t)q->r.T * (size_t)q->nq, st));
        HIPCHK(hipMallocAsync(reinterpret_cast<void **>(&q->S),
                              sizeof(double) * (size_t)std::max<int64_t>(q->r.n_f, 1) * (size_t)q->nq, st));
        HIPCHK(hipMallocAsync(reinterpret_cast<void **>(&q->St), sizeof(double) * (size_t)q->nq, st));
        // weights of the active species, in the engine's active order
        std::vector<float> wa;
        std::vector<double> ca;
        const SrvTraj &tj = trajs[q->r.traj];
        for (int j = 0; j < tj.nsp; ++j) {
            if (!q->mask.empty() && !q->mask[j]) continue;
            wa.push_back(q->r.weight_kind == 0 ? (float)q->w[j] : 1.0f);
            q->wact.push_back((double)wa.back());
            if (q->r.weight_kind == 1) ca.insert(ca.end(), q->xff.begin() + 9 * j, q->xff.begin() + 9 * j + 9);
        }
        HIPCHK(hipMallocAsync(reinterpret_cast<void **>(&q->wsc), sizeof(float) * wa.size(), st));
        HIPCHK(hipMemcpyAsync(q->wsc, wa.data(), sizeof(float) * wa.size(), hipMemcpyHostToDevice, st));
        if (q->r.weight_kind == 1) {
            HIPCHK(hipMallocAsync(reinterpret_cast<void **>(&q->coef), sizeof(double) * ca.size(), st));
            HIPCHK(hipMemcpyAsync(q->coef, ca.data(), sizeof(double) * ca.size(), hipMemcpyHostToDevice, st));
            HIPCHK(hipMallocAsync(reinterpret_cast<void **>(&q->wtab), sizeof(float) * (size_t)nact * q->nq, st));
            launch_xray_table(q->r.n_modes, tj.qs, nact, q->coef, q->wtab, st);
        }
        HIPCHK(hipEventCreateWithFlags(&q->done_ev, hipEventDisableTiming));
        double nat = 0.0;
        for (int64_t c : tj.counts) nat += (double)c;
        q->red.n_int = q->r.T;
        q->red.win_kind = 0;
        q->red.dt = q->r.dt;
        q->red.nat = nat;
        q->red.m_lo = 0;
        q->red.n_f = q->r.n_f;
        q->red.subtract_mean = 1;
        q->red.d_qfac = nullptr;
    }

    void loop()
    {
        (void)hipSetDevice(device);
        for (;;) {
            std::vector<SrvReq *> admit;
            {
                std::unique_lock<std::mutex> lk(mu);
                if (stop) return;
                while (!queue.empty() && (int)active.size() + (int)admit.size() < max_active) {
                    SrvReq *q = queue.front();
                    if (used + q->bytes > budget) break;  // arrival order: the head waits for memory
                    queue.pop_front();
                    used += q->bytes;
                    admit.push_back(q);
                }
                if (admit.empty() && active.empty()) {
                    cv.wait_for(lk, std::chrono::milliseconds(2));
                    continue;
                }
            }
            for (SrvReq *q : admit) {
                int err = 0;
                try {
                    prepare(q);
                } catch (const Err &er) {
                    err = er.code;
                } catch (...) {
                    err = MPSQ_ERR_DEVICE;
                }
                std::lock_guard<std::mutex> lk(mu);
                q->t_admit = std::chrono::steady_clock::now();
                if (err) {
                    q->state = -err;
                    used -= q->bytes;
                    --pending;
                    idle_cv.notify_all();
                } else {
                    q->state = 1;
                    active.push_back(q);
                }
            }
            // one continuous-batching step: the next frames of every admitted
            // request that still has frames, in one launch per phase
            const auto t0 = std::chrono::steady_clock::now();
            std::vector<Part<float>> parts;
            std::vector<SrvReq *> in_step, finishing;
            int nrun = 0;
            for (SrvReq *q : active)
                if (!q->reducing) ++nrun;
            for (SrvReq *q : active) {
                if (q->reducing) continue;
                const SrvTraj &tj = trajs[q->r.traj];
                const size_t share = step_budget / (size_t)std::max(nrun, 1);
                const int64_t fit = (int64_t)(share / std::max<size_t>(q->e->bytes_per_frame(), 1));
                const int b = (int)std::max<int64_t>(1, std::min<int64_t>(fit, q->r.T - q->next));
                Part<float> p;
                p.e = q->e;
                p.f0 = q->r.f0 + q->next;
                p.b = b;
                p.pos = tj.pos + (size_t)p.f0 * tj.n_tot * 4;
                p.amp = static_cast<float2 *>(q->amp) + (size_t)q->next * q->nq;
                p.wsc = q->wsc;
                p.wsh = q->wact.data();
                p.wtab = q->wtab;
                parts.push_back(p);
                in_step.push_back(q);
                q->next += b;
                if (q->next == q->r.T) finishing.push_back(q);
            }
            try {
                if (!parts.empty()) run_parts<float>(parts, ws, ring[ring_k++ % 3], st, nullptr);
                for (SrvReq *q : finishing) {
                    q->e->reduce_on(q->amp, q->r.T, &q->red, q->S, q->St, st, false);
                    HIPCHK(hipEventRecord(q->done_ev, st));
                    q->reducing = true;
                }
            } catch (...) {
                std::lock_guard<std::mutex> lk(mu);
                for (SrvReq *q : in_step) q->state = -MPSQ_ERR_DEVICE;
            }
            if (!parts.empty()) {
                ++steps;
                busy_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            }
            // completions; the next step is already queued behind them
            bool any_running = false;
            for (SrvReq *q : active) any_running = any_running || (!q->reducing && q->state > 0);
            std::vector<SrvReq *> still;
            for (SrvReq *q : active) {
                bool fin = q->state < 0;
                if (!fin && q->reducing) {
                    if (!any_running) (void)hipEventSynchronize(q->done_ev);
                    fin = hipEventQuery(q->done_ev) == hipSuccess;
                }
                if (fin) {
                    std::lock_guard<std::mutex> lk(mu);
                    if (q->state > 0) q->state = 2;
                    else used -= q->bytes;
                    q->t_done = std::chrono::steady_clock::now();
                    ++ndone;
                    --pending;
                    idle_cv.notify_all();
                } else {
                    still.push_back(q);
                }
            }
            active.swap(still);
        }
    }
};

}  // namespace mpsq

using namespace mpsq;

struct mpsq_batch_s {
    Batch<float> *bf = nullptr;
    Batch<double> *bd = nullptr;
    ~mpsq_batch_s()
    {
        delete bf;
        delete bd;
    }
};

struct mpsq_sqw_s {
    int precision = MPSQ_FP32;
    Engine<float> *ef = nullptr;
    Engine<double> *ed = nullptr;
    ~mpsq_sqw_s()
    {
        delete ef;
        delete ed;
    }
};

// mp_dos spectra: the velocity columns' windowed time FFT (one window over
// all frames, no mean subtraction, Hann window of the (i + 1) / (n + 1) form,
// norm 1 / T) and the species sums.  With a radix plan for T, one fused
// kernel per group of blocks (FFT in LDS, per-block species sums) and a
// fixed-order reduction of the blocks' sums; otherwise one DFT recurrence
// per frequency (k_sqw_window1) into |FFT|^2 rows (chunks of columns within
// 512 MB) and k_dos_species.
template <typename R>
static void dos_run(const float *d_vel, int64_t T, int64_t n_tot, const int32_t *d_at_ind, int nsp, int64_t n_freq,
                    double *d_dos, hipStream_t st)
{
    const int64_t ncs = 2 * n_tot;  // complex columns per frame: (v_x, v_y), (v_z, pad)
    const int64_t n_f = 2 * n_freq - 1;
    SqwParams p{};
    p.T = T;
    p.kind = 0;
    p.n_int = (int)T;
    p.n_frame = 1;
    p.m_lo = -(n_freq - 1);
    p.n_f = (int)n_f;
    p.win_kind = 1;
    p.subtract_mean = 0;
    p.norm_s = 1.0 / (double)T;
    p.norm_tot = 0.0;
    RadixPlan rp{};
    const bool single = (T <= sqw_tile_elems<R>()) && dos_fused_lds<R>(T, 1) <= SQW_MAX_LDS && radix_plan((int)T, rp);
    DevBuf tw, win, S, buf;
    int cols = 1;
    std::vector<double> wh((size_t)T);
    for (int64_t t = 0; t < T; ++t)
        wh[t] = (T <= 1) ? 1.0 : 0.5 * (1.0 - std::cos(2.0 * M_PI * (double)(t + 1) / (double)(T + 1)));
    win.ensure(sizeof(double) * (size_t)T);
    HIPCHK(hipMemcpyAsync(win.p, wh.data(), sizeof(double) * (size_t)T, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    if (single) {
        std::vector<Cx<R>> twh;
        upload_twiddles<R>(tw, (int)T, st, twh);
        while (2 * cols * T <= sqw_tile_elems<R>() && 2 * cols <= SQW_MAX_COLS &&
               dos_fused_lds<R>(T, 2 * cols) <= SQW_MAX_LDS)
            cols *= 2;
    }
    launch_zero_i32(reinterpret_cast<int *>(d_dos), (int)(2 * n_freq * nsp), st);
    if (single) {
        // one pass: FFT, |X|^2 and per-block species sums (no |FFT|^2 rows
        // in HBM); blocks per launch keep their sums within 256 MB
        const int64_t nout = (int64_t)nsp * n_freq;
        const int64_t nblk = (ncs + cols - 1) / cols;
        const int64_t bpl = std::max<int64_t>(1, std::min<int64_t>(nblk, ((int64_t)256 << 20) / (8 * nout)));
        DevBuf part;
        part.ensure(sizeof(double) * (size_t)((bpl + dos_reduce_groups()) * nout));
        DosParams dp{};
        dp.T = T;
        dp.ncs = ncs;
        dp.n_freq = n_freq;
        dp.nsp = nsp;
        dp.norm = 1.0 / (double)T;
        for (int64_t b0 = 0; b0 < nblk; b0 += bpl) {
            const int64_t nb = std::min(bpl, nblk - b0);
            dp.c0 = b0 * cols;
            dp.nc = std::min(nb * cols, ncs - dp.c0);
            launch_dos_fused<R>(dp, rp, cols, d_vel, d_at_ind, tw.as<Cx<R>>(), win.as<double>(), part.as<double>(),
                                st);
            dos_reduce_parts(part.as<double>(), nb, nsp, (int)n_freq, part.as<double>() + bpl * nout, d_dos, st);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipStreamSynchronize(st));
        return;
    }
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(ncs, ((int64_t)512 << 20) / (8 * n_f)));
    S.ensure(sizeof(double) * (size_t)(n_f * chunk));
    if (sizeof(R) == 8) buf.ensure(sizeof(double2) * (size_t)(T * chunk));
    for (int64_t c0 = 0; c0 < ncs; c0 += chunk) {
        const int64_t nc = std::min(chunk, ncs - c0);
        p.NQ = nc;
        const Cx<R> *amp;
        if (sizeof(R) == 8) {
            launch_cols_to_f64(d_vel, T, ncs, c0, nc, buf.as<double>(), st);
            amp = buf.as<Cx<R>>();
            p.NQs = nc;
        } else {
            amp = reinterpret_cast<const Cx<R> *>(d_vel) + c0;
            p.NQs = ncs;
        }
        if (single)
            launch_sqw_single<R>(p, rp, cols, amp, tw.as<Cx<R>>(), win.as<double>(), nullptr, S.as<double>(), nullptr,
                                 st);
        else
            launch_sqw_window1<R>(p, amp, nullptr, win.as<double>(), S.as<double>(), nullptr, st);
        launch_dos_species(S.as<double>(), nc, c0, n_freq, d_at_ind, nsp, d_dos, st);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(st));
}

extern "C" {

void mpsq_default_opts(mpsq_opts *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
}

const char *mpsq_version(void) { return "mpsq 0.1 (gfx950)"; }

static int sql_geom(const void *d_amp, int nsp, int precision, const int64_t *n_modes, const float *dq,
                    const float *a_cell_inv, float q_range, float q_step, int n_hist, const double *d_hist, SqlGeom &g)
{
    if (!d_amp || !n_modes || !dq || !a_cell_inv || !d_hist || nsp < 1 || nsp > 8 || n_hist < 1 ||
        !(q_step > 0.0f) || (precision != MPSQ_FP32 && precision != MPSQ_FP64))
        return MPSQ_ERR_BAD_ARG;
    g = SqlGeom{};
    int64_t nm = 1;
    for (int d = 0; d < 3; ++d) {
        if (n_modes[d] < 1 || n_modes[d] > (int64_t)1 << 20) return MPSQ_ERR_BAD_ARG;
        g.N[d] = (int)n_modes[d];
        g.dq[d] = dq[d];
        nm *= n_modes[d];
    }
    for (int i = 0; i < 9; ++i) g.ainv[i] = a_cell_inv[i];
    g.q_range = q_range;
    g.q_step = q_step;
    g.n_hist = n_hist;
    g.nsp = nsp;
    g.npairs = nsp * (nsp + 1) / 2;
    g.nmodes = nm;
    return 0;
}

int mpsq_sql_hist_noalloc(const void *d_amp, int nsp, int precision, const int64_t *n_modes, const float *dq,
                          const float *a_cell_inv, float q_range, float q_step, int n_hist, double *d_hist,
                          void *stream)
{
    SqlGeom g;
    const int bad = sql_geom(d_amp, nsp, precision, n_modes, dq, a_cell_inv, q_range, q_step, n_hist, d_hist, g);
    if (bad) return bad;
    const int rc = launch_sql_hist_static(d_amp, precision == MPSQ_FP64, g, d_hist, static_cast<hipStream_t>(stream));
    if (rc == -1 || rc == -3) return MPSQ_ERR_SIZE;
    if (rc != 0 || hipGetLastError() != hipSuccess) return MPSQ_ERR_DEVICE;
    return 0;
}

int mpsq_sql_hist(const void *d_amp, int nsp, int precision, const int64_t *n_modes, const float *dq,
                  const float *a_cell_inv, float q_range, float q_step, int n_hist, double *d_hist, void *stream)
{
    SqlGeom g;
    const int bad = sql_geom(d_amp, nsp, precision, n_modes, dq, a_cell_inv, q_range, q_step, n_hist, d_hist, g);
    if (bad) return bad;
    const int rc = launch_sql_hist(d_amp, precision == MPSQ_FP64, g, d_hist, static_cast<hipStream_t>(stream));
    if (rc == -1) return MPSQ_ERR_SIZE;
    if (rc != 0 || hipGetLastError() != hipSuccess) return MPSQ_ERR_DEVICE;
    return 0;
}

int mpsq_sql_hist_mc(const void *d_amp, int nsp, int precision, const int64_t *n_modes, const float *dq,
                     const float *a_cell_inv, float q_range, float q_step, int n_hist, int64_t n_h, int64_t n_mc,
                     uint64_t seed, uint32_t tag, double *d_hist, int64_t *d_skip, void *stream)
{
    SqlGeom g;
    const int bad = sql_geom(d_amp, nsp, precision, n_modes, dq, a_cell_inv, q_range, q_step, n_hist, d_hist, g);
    if (bad) return bad;
    // chain and sample indices are 32-bit counter words; j + (stride mod n_mc)
    // must not wrap
    if (!d_skip || n_h < 1 || n_mc < 1 || n_h > (int64_t)UINT32_MAX || n_mc > (int64_t)INT32_MAX)
        return MPSQ_ERR_BAD_ARG;
    const SqlMc mc{seed, tag, n_h, n_mc};
    const int rc = launch_sql_hist_mc(d_amp, precision == MPSQ_FP64, g, mc, d_hist,
                                      reinterpret_cast<unsigned long long *>(d_skip), static_cast<hipStream_t>(stream));
    if (rc == -1) return MPSQ_ERR_SIZE;
    if (rc != 0 || hipGetLastError() != hipSuccess) return MPSQ_ERR_DEVICE;
    return 0;
}

// ---- FP64 plan ----
int mpsq_makeplan(int type, int dim, const int64_t *n_modes, int iflag, int ntrans, double eps, mpsq_plan *plan,
                  const mpsq_opts *opts)
{
    if (!plan) return MPSQ_ERR_BAD_ARG;
    *plan = nullptr;
    Plan<double> *p = nullptr;
    const int rc = makeplan_impl<double>(type, dim, n_modes, iflag, ntrans, eps, &p, opts);
    if (p) *plan = reinterpret_cast<mpsq_plan>(p);
    return rc;
}
int mpsq_setpts(mpsq_plan p, int64_t M, const double *x, const double *y, const double *z)
{
    return setpts_impl<double>(reinterpret_cast<Plan<double> *>(p), M, x, y, z, false);
}
int mpsq_setpts_dev(mpsq_plan p, int64_t M, const double *x, const double *y, const double *z)
{
    return setpts_impl<double>(reinterpret_cast<Plan<double> *>(p), M, x, y, z, true);
}
int mpsq_execute(mpsq_plan p, const double *c, double *f)
{
    return execute_impl<double>(reinterpret_cast<Plan<double> *>(p), c, f, false);
}
int mpsq_execute_dev(mpsq_plan p, const double *c, double *f)
{
    return execute_impl<double>(reinterpret_cast<Plan<double> *>(p), c, f, true);
}
int mpsq_destroy(mpsq_plan p)
{
    delete reinterpret_cast<Plan<double> *>(p);
    return 0;
}
int mpsq_get_sort_perm(mpsq_plan p, int32_t *perm) { return perm_impl<double>(reinterpret_cast<Plan<double> *>(p), perm); }
int mpsq_plan_info(mpsq_plan p, int64_t *out) { return info_impl<double>(reinterpret_cast<Plan<double> *>(p), out); }

// ---- FP32 plan ----
int mpsqf_makeplan(int type, int dim, const int64_t *n_modes, int iflag, int ntrans, double eps, mpsq_planf *plan,
                   const mpsq_opts *opts)
{
    if (!plan) return MPSQ_ERR_BAD_ARG;
    *plan = nullptr;
    Plan<float> *p = nullptr;
    const int rc = makeplan_impl<float>(type, dim, n_modes, iflag, ntrans, eps, &p, opts);
    if (p) *plan = reinterpret_cast<mpsq_planf>(p);
    return rc;
}
int mpsqf_setpts(mpsq_planf p, int64_t M, const float *x, const float *y, const float *z)
{
    return setpts_impl<float>(reinterpret_cast<Plan<float> *>(p), M, x, y, z, false);
}
int mpsqf_setpts_dev(mpsq_planf p, int64_t M, const float *x, const float *y, const float *z)
{
    return setpts_impl<float>(reinterpret_cast<Plan<float> *>(p), M, x, y, z, true);
}
int mpsqf_execute(mpsq_planf p, const float *c, float *f)
{
    return execute_impl<float>(reinterpret_cast<Plan<float> *>(p), c, f, false);
}
int mpsqf_execute_dev(mpsq_planf p, const float *c, float *f)
{
    return execute_impl<float>(reinterpret_cast<Plan<float> *>(p), c, f, true);
}
int mpsqf_destroy(mpsq_planf p)
{
    delete reinterpret_cast<Plan<float> *>(p);
    return 0;
}
int mpsqf_get_sort_perm(mpsq_planf p, int32_t *perm) { return perm_impl<float>(reinterpret_cast<Plan<float> *>(p), perm); }
int mpsqf_plan_info(mpsq_planf p, int64_t *out) { return info_impl<float>(reinterpret_cast<Plan<float> *>(p), out); }

// ---- simple interface with a one-entry plan cache ----
// The reference calls the simple interface once per frame and species from
// serial code (mp_sqom56.f90:1101-1156); reusing the device plan across
// calls with the same shape avoids re-creating twiddles and buffers.
static std::mutex g_simple_mu;
static Plan<double> *g_simple = nullptr;
static int g_simple_key[2] = {0, 0};
static int64_t g_simple_nm[3] = {0, 0, 0};
static double g_simple_eps = 0.0;
static int g_simple_dev = -1;

static int simple_nufft(int dim, int64_t nj, const double *x, const double *y, const double *z, const double *c,
                        int iflag, double eps, const int64_t *nm, double *fk, const mpsq_opts *opts)
{
    int64_t NQ = 1;
    for (int d = 0; d < dim; ++d) NQ *= (nm[d] > 0 ? nm[d] : 0);
    std::lock_guard<std::mutex> lk(g_simple_mu);
    const int dev = opts ? opts->device : 0;
    bool reuse = g_simple && g_simple_key[0] == dim && g_simple_key[1] == iflag && g_simple_eps == eps &&
                 g_simple_dev == dev && !(opts && opts->stream);
    for (int d = 0; reuse && d < dim; ++d) reuse = (g_simple_nm[d] == nm[d]);
    int rc = 0;
    if (!reuse) {
        delete g_simple;
        g_simple = nullptr;
        rc = makeplan_impl<double>(1, dim, nm, iflag, 1, eps, &g_simple, opts);
        if (rc >= 2) {
            if (fk && NQ > 0) std::memset(fk, 0, sizeof(double) * 2 * (size_t)NQ);
            return rc;
        }
        g_simple_key[0] = dim;
        g_simple_key[1] = iflag;
        for (int d = 0; d < 3; ++d) g_simple_nm[d] = d < dim ? nm[d] : 1;
        g_simple_eps = eps;
        g_simple_dev = dev;
    }
    const int warn = g_simple->core.warn;
    rc = setpts_impl<double>(g_simple, nj, x, y, z, false);
    if (rc >= 2) {
        if (fk && NQ > 0) std::memset(fk, 0, sizeof(double) * 2 * (size_t)NQ);
        return rc;
    }
    rc = execute_impl<double>(g_simple, c, fk, false);
    return rc ? rc : warn;
}

int mpsq_nufft1d1(int64_t nj, const double *xj, const double *cj, int iflag, double eps, int64_t ms, double *fk,
                  const mpsq_opts *opts)
{
    const int64_t nm[3] = {ms, 1, 1};
    return simple_nufft(1, nj, xj, nullptr, nullptr, cj, iflag, eps, nm, fk, opts);
}
int mpsq_nufft2d1(int64_t nj, const double *xj, const double *yj, const double *cj, int iflag, double eps, int64_t ms,
                  int64_t mt, double *fk, const mpsq_opts *opts)
{
    const int64_t nm[3] = {ms, mt, 1};
    return simple_nufft(2, nj, xj, yj, nullptr, cj, iflag, eps, nm, fk, opts);
}
int mpsq_nufft3d1(int64_t nj, const double *xj, const double *yj, const double *zj, const double *cj, int iflag,
                  double eps, int64_t ms, int64_t mt, int64_t mu, double *fk, const mpsq_opts *opts)
{
    const int64_t nm[3] = {ms, mt, mu};
    return simple_nufft(3, nj, xj, yj, zj, cj, iflag, eps, nm, fk, opts);
}

// ---- Fortran ABI ----
void finufft1d1_(const int64_t *nj, const double *xj, const double *cj, const int *iflag, const double *eps,
                 const int64_t *ms, double *fk, const void *opts, int *ier)
{
    const int rc = mpsq_nufft1d1(*nj, xj, cj, *iflag, *eps, *ms, fk, static_cast<const mpsq_opts *>(opts));
    if (ier) *ier = rc;
}
void finufft2d1_(const int64_t *nj, const double *xj, const double *yj, const double *cj, const int *iflag,
                 const double *eps, const int64_t *ms, const int64_t *mt, double *fk, const void *opts, int *ier)
{
    const int rc = mpsq_nufft2d1(*nj, xj, yj, cj, *iflag, *eps, *ms, *mt, fk, static_cast<const mpsq_opts *>(opts));
    if (ier) *ier = rc;
}
void finufft3d1_(const int64_t *nj, const double *xj, const double *yj, const double *zj, const double *cj,
                 const int *iflag, const double *eps, const int64_t *ms, const int64_t *mt, const int64_t *mu,
                 double *fk, const void *opts, int *ier)
{
    const int rc =
        mpsq_nufft3d1(*nj, xj, yj, zj, cj, *iflag, *eps, *ms, *mt, *mu, fk, static_cast<const mpsq_opts *>(opts));
    if (ier) *ier = rc;
}
void mpsq_makeplan_(const int *type, const int *dim, const int64_t *n_modes, const int *iflag, const int *ntrans,
                    const double *eps, int64_t *handle, const void *opts, int *ier)
{
    mpsq_plan p = nullptr;
    const int rc = mpsq_makeplan(*type, *dim, n_modes, *iflag, *ntrans, *eps, &p, static_cast<const mpsq_opts *>(opts));
    *handle = (int64_t)(intptr_t)p;
    if (ier) *ier = rc;
}
void mpsq_setpts_(const int64_t *handle, const int64_t *M, const double *x, const double *y, const double *z, int *ier)
{
    const int rc = mpsq_setpts((mpsq_plan)(intptr_t)*handle, *M, x, y, z);
    if (ier) *ier = rc;
}
void mpsq_execute_(const int64_t *handle, const double *c, double *f, int *ier)
{
    const int rc = mpsq_execute((mpsq_plan)(intptr_t)*handle, c, f);
    if (ier) *ier = rc;
}
void mpsq_destroy_(const int64_t *handle, int *ier)
{
    const int rc = mpsq_destroy((mpsq_plan)(intptr_t)*handle);
    if (ier) *ier = rc;
}

// ---- S(Q,w) engine ----
void mpsq_sqw_default_spec(mpsq_sqw_spec *s)
{
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    s->dim = 3;
    s->n_modes[0] = s->n_modes[1] = s->n_modes[2] = 1;
    s->eps = 1e-6;
    s->precision = MPSQ_FP32;
    s->iflag = -1;
    s->combine = 1;
    s->axes[0] = s->axes[4] = s->axes[8] = 1.0;
    s->tpe[0] = s->tpe[1] = s->tpe[2] = 1.0;
    s->vacancy_check = 1;
    s->margin = 0.1f;
    mpsq_default_opts(&s->opts);
}

int mpsq_sqw_create(const mpsq_sqw_spec *spec, mpsq_sqw *eng)
{
    if (!eng) return MPSQ_ERR_BAD_ARG;
    *eng = nullptr;
    if (!spec) return MPSQ_ERR_BAD_ARG;
    mpsq_sqw_s *e = new (std::nothrow) mpsq_sqw_s();
    if (!e) return MPSQ_ERR_DEVICE;
    e->precision = spec->precision;
    try {
        if (spec->precision == MPSQ_FP64) {
            e->ed = new Engine<double>();
            e->ed->create(spec);
        } else if (spec->precision == MPSQ_FP32) {
            e->ef = new Engine<float>();
            e->ef->create(spec);
        } else {
            throw Err{MPSQ_ERR_BAD_ARG};
        }
    } catch (const Err &er) {
        delete e;
        return er.code;
    } catch (...) {
        delete e;
        return MPSQ_ERR_DEVICE;
    }
    *eng = e;
    return e->ed ? e->ed->core.warn : e->ef->core.warn;
}

int64_t mpsq_sqw_nq(mpsq_sqw e)
{
    if (!e) return 0;
    if (e->ed) return e->ed->core.N[0] * e->ed->core.N[1] * e->ed->core.N[2];
    return e->ef->core.N[0] * e->ef->core.N[1] * e->ef->core.N[2];
}

int mpsq_sqw_nact(mpsq_sqw e)
{
    if (!e) return 0;
    return e->ed ? e->ed->nact : e->ef->nact;
}

int mpsq_sqw_amplitudes(mpsq_sqw e, const float *d_pos, int64_t T, void *d_amp)
{
    if (!e) return MPSQ_ERR_BAD_ARG;
    try {
        if (e->ed) e->ed->amplitudes(d_pos, T, d_amp);
        else e->ef->amplitudes(d_pos, T, d_amp);
    } catch (const Err &er) {
        return er.code;
    } catch (...) {
        return MPSQ_ERR_DEVICE;
    }
    return e->ed ? e->ed->core.warn : e->ef->core.warn;
}

int mpsq_sqw_reduce(mpsq_sqw e, const void *d_amp, int64_t T, const mpsq_reduce_spec *r, double *d_S, double *d_Stot)
{
    if (!e) return MPSQ_ERR_BAD_ARG;
    try {
        if (e->ed) e->ed->reduce(d_amp, T, r, d_S, d_Stot);
        else e->ef->reduce(d_amp, T, r, d_S, d_Stot);
    } catch (const Err &er) {
        return er.code;
    } catch (...) {
        return MPSQ_ERR_DEVICE;
    }
    return 0;
}

int mpsq_sqw_debug_sorted(mpsq_sqw e, const float *d_pos, int64_t frame, int32_t *atom, int32_t *cell,
                          int32_t *keep, int64_t n)
{
    if (!e) return MPSQ_ERR_BAD_ARG;
    try {
        if (e->ed) e->ed->debug_sorted(d_pos, frame, atom, cell, keep, n);
        else e->ef->debug_sorted(d_pos, frame, atom, cell, keep, n);
    } catch (const Err &er) {
        return er.code;
    } catch (...) {
        return MPSQ_ERR_DEVICE;
    }
    return 0;
}

int mpsq_sqw_info(mpsq_sqw e, int64_t *out)
{
    if (!e || !out) return MPSQ_ERR_BAD_ARG;
    const Geo &g = e->ed ? e->ed->core.g : e->ef->core.g;
    const int64_t v[8] = {g.n1, g.n2, g.n3, g.w, g.by, g.bz, g.BY, g.BZ};
    for (int k = 0; k < 8; ++k) out[k] = v[k];
    return 0;
}

int mpsq_dos(const float *d_vel, int64_t T, int64_t n_tot, const int32_t *d_at_ind, int nsp, int64_t n_freq,
             int precision, double *d_dos, void *stream)
{
    if (!d_vel || !d_at_ind || !d_dos || T < 1 || n_tot < 1 || nsp < 1 || n_freq < 1 || n_freq > T / 2 + 1 ||
        (precision != MPSQ_FP32 && precision != MPSQ_FP64) || 2 * n_freq * nsp > ((int64_t)1 << 30))
        return MPSQ_ERR_BAD_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MPSQ_ERR_DEVICE;
    try {
        hipStream_t st = static_cast<hipStream_t>(stream);
        if (precision == MPSQ_FP64) dos_run<double>(d_vel, T, n_tot, d_at_ind, nsp, n_freq, d_dos, st);
        else dos_run<float>(d_vel, T, n_tot, d_at_ind, nsp, n_freq, d_dos, st);
    } catch (const Err &er) {
        return er.code;
    } catch (...) {
        return MPSQ_ERR_DEVICE;
    }
    return 0;
}

int mpsq_oneph_amplitudes(const mpsq_oneph_spec *sp, const float *d_pos, const int32_t *d_at_ind, int64_t T,
                          const double *wsc, const float *d_wtab, void *d_amp, void *stream)
{
    if (!sp || !d_pos || !d_at_ind || !d_amp || T < 1 || !sp->at_base) return MPSQ_ERR_BAD_ARG;
    const int n1 = sp->n_row[0], n2 = sp->n_row[1], n3 = sp->n_row[2];
    if (n1 < 1 || n2 < 1 || n3 < 1 || n1 > 1024 || n2 > 1024 || n3 > 1024 || sp->n_bz < 1 || sp->nsp < 1 ||
        sp->nsp > ONEPH_MAX_SP || (sp->j_plane != 1 && sp->j_plane != 2) ||
        (sp->precision != MPSQ_FP32 && sp->precision != MPSQ_FP64))
        return MPSQ_ERR_BAD_ARG;
    // the reference's loops read the (n_row(1), n_row(2)) columns with the
    // zone pixels of either axis: square for (hk0), cubic for (hhl)
    if (n1 != n2 || (sp->j_plane == 2 && n3 != n1)) return MPSQ_ERR_BAD_ARG;
    const int64_t NQ = (int64_t)sp->n_bz * n1 * sp->n_bz * n2;
    if (NQ > ((int64_t)1 << 31)) return MPSQ_ERR_SIZE;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MPSQ_ERR_DEVICE;
    try {
        OnephParams p{};
        for (int d = 0; d < 3; ++d) p.n_row[d] = sp->n_row[d];
        p.n_bz = sp->n_bz;
        p.j_plane = sp->j_plane;
        p.nsp = sp->nsp;
        p.combine = (wsc || d_wtab) ? 1 : 0;
        p.nsuper = (int64_t)n1 * n2 * n3;
        p.n_tot = p.nsuper * sp->nsp;
        p.qx_min = sp->qx_min;
        p.qy_min = sp->qy_min;
        p.twopi_s = (float)(2.0 * M_PI);
        p.sqrt2 = std::sqrt(2.0f);
        // zone offset: n_row * (q_min - floor(q_min)) in REAL(4), truncated (:1039-1040)
        p.nqx_min = (int)((float)n1 * (sp->qx_min - std::floor(sp->qx_min)));
        p.nqy_min = (int)((float)n2 * (sp->qy_min - std::floor(sp->qy_min)));
        for (int j = 0; j < sp->nsp; ++j)
            for (int c = 0; c < 3; ++c) p.base[j][c] = sp->at_base[3 * j + c];
        hipStream_t st = static_cast<hipStream_t>(stream);
        // frames per chunk: the three work grids within ~256 MB
        const int64_t per_frame = (int64_t)sp->nsp * 2 * n1 * n2 * (8 + 16 + 16);
        const int nfc = (int)std::max<int64_t>(1, std::min<int64_t>(T, ((int64_t)256 << 20) / per_frame));
        DevBuf u, ua, U, tw1, tw2, wd, bad;
        u.ensure(sizeof(double) * (size_t)nfc * sp->nsp * 2 * n1 * n2);
        ua.ensure(sizeof(double2) * (size_t)nfc * sp->nsp * 2 * n1 * n2);
        U.ensure(sizeof(double2) * (size_t)nfc * sp->nsp * 2 * n1 * n2);
        auto table = [&](DevBuf &b, int n) {
            std::vector<double2> h((size_t)n);
            for (int m = 0; m < n; ++m) h[m] = make_double2(std::cos(2.0 * M_PI * m / n), -std::sin(2.0 * M_PI * m / n));
            b.ensure(sizeof(double2) * (size_t)n);
            HIPCHK(hipMemcpyAsync(b.p, h.data(), sizeof(double2) * (size_t)n, hipMemcpyHostToDevice, st));
            HIPCHK(hipStreamSynchronize(st));
        };
        table(tw1, n1);
        table(tw2, n2);
        if (wsc) {
            wd.ensure(sizeof(double) * (size_t)sp->nsp);
            HIPCHK(hipMemcpyAsync(wd.p, wsc, sizeof(double) * (size_t)sp->nsp, hipMemcpyHostToDevice, st));
        }
        bad.ensure(sizeof(int));
        launch_zero_i32(bad.as<int>(), 1, st);
        for (int64_t f0 = 0; f0 < T; f0 += nfc) {
            const int nf = (int)std::min<int64_t>(nfc, T - f0);
            if (sp->precision == MPSQ_FP64)
                launch_oneph<double>(p, d_pos, d_at_ind, f0, nf, u.as<double>(), ua.as<double2>(), U.as<double2>(),
                                     tw1.as<double2>(), tw2.as<double2>(), wd.as<double>(), d_wtab, d_amp,
                                     bad.as<int>(), st);
            else
                launch_oneph<float>(p, d_pos, d_at_ind, f0, nf, u.as<double>(), ua.as<double2>(), U.as<double2>(),
                                    tw1.as<double2>(), tw2.as<double2>(), wd.as<double>(), d_wtab, d_amp,
                                    bad.as<int>(), st);
            HIPCHK(hipGetLastError());
        }
        int hbad = 0;
        HIPCHK(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (hbad & 4) return MPSQ_ERR_BAD_ARG;
    } catch (const Err &er) {
        return er.code;
    } catch (...) {
        return MPSQ_ERR_DEVICE;
    }
    return 0;
}

int mpsq_qtables(const mpsq_qplane *qp, int nsp, const float *x_ffpar, float *d_q2_norm, float *d_x_ffq,
                 void *stream)
{
    if (!qp || qp->n_qx < 1 || qp->n_qy < 1 || (d_x_ffq && (nsp < 1 || nsp > 64 || !x_ffpar)))
        return MPSQ_ERR_BAD_ARG;
    QTableParams p{};
    p.n_qx = qp->n_qx;
    p.n_qy = qp->n_qy;
    p.nsp = d_x_ffq ? nsp : 0;
    p.qx_min = qp->qx_min;
    p.qy_min = qp->qy_min;
    p.dqx = qp->bz_nx / (float)qp->n_qx;
    p.dqy = qp->bz_ny / (float)qp->n_qy;
    for (int d = 0; d < 3; ++d) {
        p.e1[d] = qp->e1[d];
        p.e2[d] = qp->e2[d];
        p.q_v[d] = qp->q_v[d];
        p.e1p[d] = qp->e1p[d];
        p.e2p[d] = qp->e2p[d];
        p.q_vp[d] = qp->q_vp[d];
    }
    for (int s = 0; s < p.nsp; ++s)
        for (int k = 0; k < 9; ++k) p.ff[s][k] = x_ffpar[9 * s + k];
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MPSQ_ERR_DEVICE;
    launch_qtables(p, d_q2_norm, d_x_ffq, (hipStream_t)stream);
    return hipGetLastError() == hipSuccess ? 0 : MPSQ_ERR_DEVICE;
}

// The reference's speckle filter (mp_sqom56.f90:1530-1570), host memory:
// one sequential pass over the n_rows x n_cols row-major array in the walk
// order (row-major, or column-major when column_walk), cyclic neighbours,
// each value compared with its neighbours as the walk has left them.
int mpsq_speckle_filter(double *map, int64_t n_rows, int64_t n_cols, int column_walk, double s_trig)
{
    if (!map || n_rows < 1 || n_cols < 1) return MPSQ_ERR_BAD_ARG;
    if (!(s_trig > 0.0)) return 0;  // s_trig = 0 turns the filter off (:1531)
    const int64_t n = n_rows * n_cols;
    auto at = [&](int64_t k) -> double & { return column_walk ? map[(k % n_rows) * n_cols + k / n_rows] : map[k]; };
    for (int64_t k = 0; k < n; ++k) {
        const double a = at(k == 0 ? n - 1 : k - 1), b = at(k + 1 == n ? 0 : k + 1);
        double &v = at(k);
        if (v < std::min(a, b) / s_trig) v = std::min(a, b);
        if (v > s_trig * std::max(a, b)) v = std::max(a, b);
    }
    return 0;
}

int mpsq_sqw_sync(mpsq_sqw e)
{
    if (!e) return MPSQ_ERR_BAD_ARG;
    hipStream_t st = e->ed ? e->ed->core.st : e->ef->core.st;
    return hipStreamSynchronize(st) == hipSuccess ? 0 : MPSQ_ERR_DEVICE;
}

int mpsq_sqw_set_timing(mpsq_sqw e, int on)
{
    if (!e) return MPSQ_ERR_BAD_ARG;
    StageTimer &tm = e->ed ? e->ed->tm : e->ef->tm;
    tm.on = on != 0;
    tm.reset();
    return 0;
}

int mpsq_sqw_get_timing(mpsq_sqw e, double *ms, int64_t *launches, int n)
{
    if (!e || n < 0) return MPSQ_ERR_BAD_ARG;
    const StageTimer &tm = e->ed ? e->ed->tm : e->ef->tm;
    for (int k = 0; k < n && k < StageTimer::K; ++k) {
        if (ms) ms[k] = tm.ms[k];
        if (launches) launches[k] = tm.n[k];
    }
    return 0;
}

int mpsq_sqw_destroy(mpsq_sqw e)
{
    delete e;
    return 0;
}

// ---- batches ----
int mpsq_batch_create(int njobs, const mpsq_sqw_spec *specs, mpsq_batch *batch)
{
    if (!batch) return MPSQ_ERR_BAD_ARG;
    *batch = nullptr;
    if (njobs < 1 || !specs) return MPSQ_ERR_BAD_ARG;
    mpsq_batch_s *b = new (std::nothrow) mpsq_batch_s();
    if (!b) return MPSQ_ERR_DEVICE;
    try {
        if (specs[0].precision == MPSQ_FP64) {
            b->bd = new Batch<double>();
            b->bd->create(njobs, specs);
        } else {
            b->bf = new Batch<float>();
            b->bf->create(njobs, specs);
        }
    } catch (const Err &er) {
        delete b;
        return er.code;
    } catch (...) {
        delete b;
        return MPSQ_ERR_DEVICE;
    }
    *batch = b;
    int warn = 0;
    if (b->bd)
        for (auto *e : b->bd->jobs) warn = std::max(warn, e->core.warn);
    if (b->bf)
        for (auto *e : b->bf->jobs) warn = std::max(warn, e->core.warn);
    return warn;
}

int mpsq_batch_run(mpsq_batch b, const float *const *d_pos, const int64_t *T, void *const *d_amp,
                   const mpsq_reduce_spec *r, double *const *d_S, double *const *d_Stot, int use_graph)
{
    if (!b || !d_pos || !T || !d_amp) return MPSQ_ERR_BAD_ARG;
    try {
        const size_t nj = b->bf ? b->bf->jobs.size() : b->bd->jobs.size();
        for (size_t j = 0; j < nj; ++j) {
            if (!d_pos[j] || !d_amp[j] || T[j] < 1) throw Err{MPSQ_ERR_BAD_ARG};
            if (r && (r[j].n_int < (r[j].odd_n_int ? 0 : 1) || r[j].n_int > T[j] || r[j].n_f < 0 || !(r[j].nat > 0) ||
                      (r[j].n_f > 0 && (!d_S || !d_S[j]))))
                throw Err{MPSQ_ERR_BAD_ARG};
        }
        if (b->bf) b->bf->run(d_pos, T, d_amp, r, d_S, d_Stot, use_graph != 0);
        else b->bd->run(d_pos, T, d_amp, r, d_S, d_Stot, use_graph != 0);
    } catch (const Err &er) {
        return er.code;
    } catch (...) {
        return MPSQ_ERR_DEVICE;
    }
    return 0;
}

int mpsq_batch_set_budget(mpsq_batch b, int64_t bytes)
{
    if (!b || bytes < 1) return MPSQ_ERR_BAD_ARG;
    if (b->bf) b->bf->budget = (size_t)bytes;
    else b->bd->budget = (size_t)bytes;
    return 0;
}

int mpsq_batch_set_timing(mpsq_batch b, int on)
{
    if (!b) return MPSQ_ERR_BAD_ARG;
    StageTimer &tm = b->bf ? b->bf->tm : b->bd->tm;
    tm.on = on != 0;
    tm.reset();
    return 0;
}

int mpsq_batch_get_timing(mpsq_batch b, double *ms, int64_t *launches, int n)
{
    if (!b || n < 0) return MPSQ_ERR_BAD_ARG;
    const StageTimer &tm = b->bf ? b->bf->tm : b->bd->tm;
    for (int k = 0; k < n && k < StageTimer::K; ++k) {
        if (ms) ms[k] = tm.ms[k];
        if (launches) launches[k] = tm.n[k];
    }
    return 0;
}

int mpsq_batch_destroy(mpsq_batch b)
{
    delete b;
    return 0;
}

// ---- map server ----
struct mpsq_server_s {
    Server *s = nullptr;
    ~mpsq_server_s() { delete s; }
};

int mpsq_server_create(const mpsq_server_opts *o, mpsq_server *srv)
{
    if (!srv) return MPSQ_ERR_BAD_ARG;
    *srv = nullptr;
    mpsq_server_s *h = new (std::nothrow) mpsq_server_s();
    if (!h) return MPSQ_ERR_DEVICE;
    try {
        h->s = new Server();
        h->s->start(o);
    } catch (const Err &er) {
        delete h;
        return er.code;
    } catch (...) {
        delete h;
        return MPSQ_ERR_DEVICE;
    }
    *srv = h;
    return 0;
}

int mpsq_server_add_trajectory(mpsq_server srv, const float *d_pos, int64_t T, int nsp, const int64_t *sp_count,
                               const double *tpe, const double *qscale, int *traj_id)
{
    if (!srv || !d_pos || T < 1 || nsp < 1 || nsp > MAXSP || !sp_count || !tpe) return MPSQ_ERR_BAD_ARG;
    SrvTraj t;
    t.pos = d_pos;
    t.T = T;
    t.nsp = nsp;
    t.counts.assign(sp_count, sp_count + nsp);
    for (int64_t c : t.counts) {
        if (c < 0) return MPSQ_ERR_BAD_ARG;
        t.n_tot += c;
    }
    for (int d = 0; d < 3; ++d) {
        t.tpe[d] = tpe[d];
        t.qs[d] = qscale ? qscale[d] : 1.0;
    }
    std::lock_guard<std::mutex> lk(srv->s->mu);
    srv->s->trajs.push_back(t);
    if (traj_id) *traj_id = (int)srv->s->trajs.size() - 1;
    return 0;
}

int mpsq_server_submit(mpsq_server srv, const mpsq_request *r, int64_t *ticket)
{
    if (!srv || !r || !ticket) return MPSQ_ERR_BAD_ARG;
    Server &S = *srv->s;
    std::lock_guard<std::mutex> lk(S.mu);
    if (r->traj < 0 || r->traj >= (int)S.trajs.size()) return MPSQ_ERR_BAD_ARG;
    const SrvTraj &tj = S.trajs[r->traj];
    if (r->T < 1 || r->f0 < 0 || r->f0 + r->T > tj.T || r->n_f < 0 || !(r->eps > 0)) return MPSQ_ERR_BAD_ARG;
    for (int d = 0; d < 3; ++d)
        if (r->n_modes[d] < 1) return MPSQ_ERR_BAD_ARG;
    if (r->weight_kind == 0 ? !r->w_scalar : (r->weight_kind != 1 || !r->xff)) return MPSQ_ERR_BAD_ARG;
    SrvReq *q = new SrvReq();
    q->r = *r;
    if (r->mask) q->mask.assign(r->mask, r->mask + tj.nsp);
    int nact = 0;
    for (int j = 0; j < tj.nsp; ++j) nact += q->mask.empty() ? 1 : (q->mask[j] != 0);
    if (nact == 0) {
        delete q;
        return MPSQ_ERR_BAD_ARG;
    }
    if (r->weight_kind == 0) q->w.assign(r->w_scalar, r->w_scalar + tj.nsp);
    else q->xff.assign(r->xff, r->xff + 9 * tj.nsp);
    q->r.mask = nullptr;
    q->r.w_scalar = nullptr;
    q->r.xff = nullptr;
    q->bytes = Server::req_bytes(*q, nact, r->n_modes[0] * r->n_modes[1] * r->n_modes[2]);
    if (q->bytes > S.budget) {
        delete q;
        return MPSQ_ERR_SIZE;
    }
    q->ticket = S.next_ticket++;
    q->t_sub = std::chrono::steady_clock::now();
    S.reqs[q->ticket] = q;
    S.queue.push_back(q);
    ++S.pending;
    *ticket = q->ticket;
    S.cv.notify_all();
    return 0;
}

int mpsq_server_poll(mpsq_server srv, int64_t ticket, int *state, double *queue_ms, double *latency_ms)
{
    if (!srv) return MPSQ_ERR_BAD_ARG;
    Server &S = *srv->s;
    std::lock_guard<std::mutex> lk(S.mu);
    auto it = S.reqs.find(ticket);
    if (it == S.reqs.end()) return MPSQ_ERR_BAD_ARG;
    const SrvReq *q = it->second;
    if (state) *state = q->state;
    using ms = std::chrono::duration<double, std::milli>;
    if (queue_ms) *queue_ms = q->state == 0 ? -1.0 : ms(q->t_admit - q->t_sub).count();
    if (latency_ms) *latency_ms = q->state == 2 ? ms(q->t_done - q->t_sub).count() : -1.0;
    return 0;
}

int mpsq_server_result(mpsq_server srv, int64_t ticket, const double **d_S, const double **d_Stot, int64_t *nq)
{
    if (!srv) return MPSQ_ERR_BAD_ARG;
    Server &S = *srv->s;
    std::lock_guard<std::mutex> lk(S.mu);
    auto it = S.reqs.find(ticket);
    if (it == S.reqs.end() || it->second->state != 2) return MPSQ_ERR_NOT_SET;
    if (d_S) *d_S = it->second->S;
    if (d_Stot) *d_Stot = it->second->St;
    if (nq) *nq = it->second->nq;
    return 0;
}

int mpsq_server_copy_result(mpsq_server srv, int64_t ticket, double *d_S, double *d_Stot)
{
    if (!srv) return MPSQ_ERR_BAD_ARG;
    Server &S = *srv->s;
    const double *src_S = nullptr, *src_St = nullptr;
    int64_t nq = 0, nf = 0;
    {
        std::lock_guard<std::mutex> lk(S.mu);
        auto it = S.reqs.find(ticket);
        if (it == S.reqs.end() || it->second->state != 2) return MPSQ_ERR_NOT_SET;
        src_S = it->second->S;
        src_St = it->second->St;
        nq = it->second->nq;
        nf = it->second->r.n_f;
    }
    if (hipSetDevice(S.device) != hipSuccess) return MPSQ_ERR_DEVICE;
    if (d_S && nf > 0 && hipMemcpy(d_S, src_S, sizeof(double) * (size_t)(nf * nq), hipMemcpyDeviceToDevice))
        return MPSQ_ERR_DEVICE;
    if (d_Stot && hipMemcpy(d_Stot, src_St, sizeof(double) * (size_t)nq, hipMemcpyDeviceToDevice))
        return MPSQ_ERR_DEVICE;
    return 0;
}

int mpsq_server_release(mpsq_server srv, int64_t ticket)
{
    if (!srv) return MPSQ_ERR_BAD_ARG;
    Server &S = *srv->s;
    std::lock_guard<std::mutex> lk(S.mu);
    auto it = S.reqs.find(ticket);
    if (it == S.reqs.end()) return MPSQ_ERR_BAD_ARG;
    SrvReq *q = it->second;
    if (q->state == 0 || q->state == 1) return MPSQ_ERR_NOT_SET;  // queued or running
    if (q->state == 2) S.used -= q->bytes;
    (void)hipSetDevice(S.device);
    S.reqs.erase(it);
    S.free_req(q);
    S.cv.notify_all();
    return 0;
}

int mpsq_server_wait_idle(mpsq_server srv)
{
    if (!srv) return MPSQ_ERR_BAD_ARG;
    Server &S = *srv->s;
    std::unique_lock<std::mutex> lk(S.mu);
    S.idle_cv.wait(lk, [&] { return S.pending == 0; });
    return 0;
}

int mpsq_server_stats(mpsq_server srv, int64_t *steps, int64_t *done, double *busy_ms)
{
    if (!srv) return MPSQ_ERR_BAD_ARG;
    Server &S = *srv->s;
    std::lock_guard<std::mutex> lk(S.mu);
    if (steps) *steps = S.steps;
    if (done) *done = S.ndone;
    if (busy_ms) *busy_ms = S.busy_ms;
    return 0;
}

int mpsq_server_destroy(mpsq_server srv)
{
    delete srv;
    return 0;
}

int mpsq_autotune_tiles(const mpsq_sqw_spec *spec, const float *d_pos, int64_t T, int *tile_by, int *tile_bz,
                        double *ms_per_frame)
{
    if (!spec || !d_pos || T < 1) return MPSQ_ERR_BAD_ARG;
    try {
        if (spec->precision == MPSQ_FP64) return autotune_impl<double>(spec, d_pos, T, tile_by, tile_bz, ms_per_frame);
        return autotune_impl<float>(spec, d_pos, T, tile_by, tile_bz, ms_per_frame);
    } catch (const Err &er) {
        return er.code;
    } catch (...) {
        return MPSQ_ERR_DEVICE;
    }
}

}  // extern "C"
