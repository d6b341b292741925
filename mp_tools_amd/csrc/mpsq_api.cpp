// mpsq_api.cpp -- host side of libmpsq.so: plan construction, the
// FINUFFT-shaped C ABI, the Fortran-ABI shims and the S(Q,w) engine.
//
// Everything here is orchestration: parameter choice, device buffers and
// kernel launches on one HIP stream per plan.  All arithmetic of the path
// runs in the kernels of nufft_kernels.hip and sqw_kernels.hip; there is no
// CPU fallback -- without a usable HIP device every call returns
// MPSQ_ERR_DEVICE with zeroed outputs.
#include <algorithm>
#include <cmath>
#include <cstring>
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
// geometry shared by plans and the S(Q,w) engine
// ---------------------------------------------------------------------------
template <typename R> struct Core {
    int dim = 0, iflag = -1;
    double eps = 0;
    int64_t N[3] = {1, 1, 1}, n[3] = {1, 1, 1};
    Geo g{};
    RadixPlan rp[3]{};
    DevBuf tw[3], phinv[3];
    std::vector<Cx<R>> tw_host[3];
    std::vector<R> phinv_host[3];
    int device = 0;
    hipStream_t st = nullptr;
    bool own_stream = false;
    int warn = 0;
    int sort_chunk = 0;

    void setup(int dim_, const int64_t *nm, int iflag_, double eps_, const mpsq_opts *o)
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
            HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
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
        const int rows_max = (int)std::min<int64_t>(128, TE / n[0]);
        const int bmax = spread_max_tile_b();
        int by = 1, bz = 1;
        if (dim == 2) {
            for (int b = 1; b <= rows_max && b <= bmax && n[1] % b == 0 && b + w <= n[1]; b *= 2) by = b;
        } else if (dim == 3) {
            double best = 1e300;
            for (int b1 = 1; b1 <= rows_max && b1 <= bmax && n[1] % b1 == 0 && b1 + w <= n[1]; b1 *= 2)
                for (int b2 = 1; b1 * b2 <= rows_max && b2 <= bmax && n[2] % b2 == 0 && b2 + w <= n[2]; b2 *= 2) {
                    const double cost = (double)(b1 + w - 1) * (b2 + w - 1) / (b1 * b2);
                    if (cost < best - 1e-12 || (std::fabs(cost - best) < 1e-12 && b1 * b2 > by * bz)) {
                        best = cost;
                        by = b1;
                        bz = b2;
                    }
                }
        }
        if (o->tile_by > 0 && dim >= 2 && n[1] % o->tile_by == 0) by = o->tile_by;
        if (o->tile_bz > 0 && dim >= 3 && n[2] % o->tile_bz == 0) bz = o->tile_bz;
        if ((int64_t)by * bz * n[0] > TE || by > bmax || bz > bmax) throw Err{MPSQ_ERR_BAD_ARG};
        if (spread_lds<R>(by * bz, (int)n[0], w) > 156 * 1024) throw Err{MPSQ_ERR_SIZE};
        if ((dim >= 2 && by + w > n[1]) || (dim >= 3 && bz + w > n[2])) throw Err{MPSQ_ERR_BAD_ARG};
        g.by = by;
        g.bz = bz;
        // sort bins: start at the tile size, coarsen until the bin count fits
        // the LDS counters and a tile visits at most MAX_NB_BINS bins
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
        const int TE = colfft_tile_elems<R>(final);
        int c = 1;
        while ((int64_t)(2 * c) * nlen <= TE) c *= 2;
        while (c > 1 && c / 2 >= inner) c /= 2;
        return c;
    }

    // Launch the transform passes after spreading.  ntr transforms are laid
    // out back to back; FINAL groups nsp consecutive transforms.
    void passes(int ntr, int ngroups, int nsp, const Cx<R> *G1, Cx<R> *G2, Cx<R> *out, const R *wsc,
                const float *wtab, StageTimer *tm = nullptr)
    {
        StageTimer dummy;
        if (!tm) tm = &dummy;
        const int conj = iflag > 0 ? 1 : 0;
        const int64_t NQ = N[0] * N[1] * N[2];
        if (dim == 2) {
            ColPass cp{};
            cp.n = (int)n[1];
            cp.N = (int)N[1];
            cp.inner = (int)N[0];
            cp.outer = 1;
            cp.cols = colfft_cols(n[1], N[0], true);
            cp.dcols = make_fdiv((uint32_t)cp.cols);
            cp.in_tstride = n[1] * N[0];
            cp.out_tstride = N[1] * N[0];
            cp.nsp = nsp;
            cp.NQ = NQ;
            cp.conj = conj;
            const int ev = tm->begin(ST_FINAL, st);
            launch_colfft<R>(cp, rp[1], ngroups, G1, out, tw[1].as<Cx<R>>(), phinv[1].as<R>(), wsc, wtab, true, st);
            tm->end(ev, st);
        } else if (dim == 3) {
            ColPass cy{};
            cy.n = (int)n[1];
            cy.N = (int)N[1];
            cy.inner = (int)N[0];
            cy.outer = (int)n[2];
            cy.cols = colfft_cols(n[1], N[0], false);
            cy.dcols = make_fdiv((uint32_t)cy.cols);
            cy.in_tstride = n[2] * n[1] * N[0];
            cy.out_tstride = n[2] * N[1] * N[0];
            cy.nsp = 1;
            cy.NQ = NQ;
            cy.conj = 0;
            const int ev1 = tm->begin(ST_COLPASS, st);
            launch_colfft<R>(cy, rp[1], ntr, G1, G2, tw[1].as<Cx<R>>(), phinv[1].as<R>(), nullptr, nullptr, false,
                             st);
            tm->end(ev1, st);
            ColPass cz{};
            cz.n = (int)n[2];
            cz.N = (int)N[2];
            cz.inner = (int)(N[0] * N[1]);
            cz.outer = 1;
            cz.cols = colfft_cols(n[2], N[0] * N[1], true);
            cz.dcols = make_fdiv((uint32_t)cz.cols);
            cz.in_tstride = n[2] * N[1] * N[0];
            cz.out_tstride = N[2] * N[1] * N[0];
            cz.nsp = nsp;
            cz.NQ = NQ;
            cz.conj = conj;
            const int ev2 = tm->begin(ST_FINAL, st);
            launch_colfft<R>(cz, rp[2], ngroups, G2, out, tw[2].as<Cx<R>>(), phinv[2].as<R>(), wsc, wtab, true, st);
            tm->end(ev2, st);
        }
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
    DevBuf xyz[3], rec, key, srec, perm, H, Hs, bsum, chunks_d, tinfo_d, binstart, bad;
    DevBuf cin, cs, G1, G2, fout, amax;
    std::vector<SortChunk> chunks;
    std::vector<TransformInfo> tinfo;

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
        const int nch = (int)((M_ + CH - 1) / CH);
        chunks.resize((size_t)std::max(nch, 1));
        for (int k = 0; k < nch; ++k) {
            chunks[k].p0 = (int64_t)k * CH;
            chunks[k].p1 = std::min<int64_t>(M_, (int64_t)(k + 1) * CH);
            chunks[k].pbase = 0;
            chunks[k].hbase = 0;
            chunks[k].nch = nch;
            chunks[k].lc = k;
        }
        tinfo.assign(1, TransformInfo{0, M_, 0, nch, 0});
        const int64_t Htot = (int64_t)c.g.nbins * nch;
        H.ensure(sizeof(int) * (size_t)std::max<int64_t>(Htot, 1));
        Hs.ensure(sizeof(int64_t) * (size_t)std::max<int64_t>(Htot, 1));
        bsum.ensure(sizeof(int64_t) * (size_t)scan_bsum_size(Htot));
        chunks_d.ensure(sizeof(SortChunk) * chunks.size());
        tinfo_d.ensure(sizeof(TransformInfo));
        binstart.ensure(sizeof(int64_t) * (size_t)(c.g.nbins + 1));
        HIPCHK(hipMemcpyAsync(chunks_d.p, chunks.data(), sizeof(SortChunk) * chunks.size(), hipMemcpyHostToDevice,
                              c.st));
        HIPCHK(hipMemcpyAsync(tinfo_d.p, tinfo.data(), sizeof(TransformInfo), hipMemcpyHostToDevice, c.st));
        launch_sort_count(key.as<int>(), chunks_d.as<SortChunk>(), nch, c.g.nbins, H.as<int>(), c.st);
        launch_scan(H.as<int>(), Htot, bsum.as<int64_t>(), Hs.as<int64_t>(), c.st);
        launch_sort_scatter<R>(key.as<int>(), chunks_d.as<SortChunk>(), nch, c.g.nbins, c.keybits(),
                               Hs.as<int64_t>(), rec.as<PtRec<R>>(), nullptr, srec.as<PtRec<R>>(), nullptr,
                               perm.as<int>(), c.st);
        launch_binstart(Hs.as<int64_t>(), tinfo_d.as<TransformInfo>(), 1, c.g.nbins, binstart.as<int64_t>(), c.st);
        HIPCHK(hipGetLastError());
        int hbad = 0;
        HIPCHK(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, c.st));
        HIPCHK(hipStreamSynchronize(c.st));
        if (hbad) throw Err{MPSQ_ERR_NONFINITE_POINT};
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
        launch_spread_xfft<R>(c.g, c.rp[0], ntrans, srec.as<PtRec<R>>(), cs.as<Cx<R>>(), M,
                              binstart.as<int64_t>(),
                              0, c.tw[0].template as<Cx<R>>(), c.phinv[0].template as<R>(), g1,
                              (c.dim == 1 && c.iflag > 0) ? 1 : 0, am, c.st);
        c.passes(ntrans, ntrans, 1, g1, G2.as<Cx<R>>(), fd, nullptr, nullptr);
        HIPCHK(hipGetLastError());
        if (!on_device) {
            HIPCHK(hipMemcpyAsync(f_out, fd, sizeof(Cx<R>) * (size_t)NQ * ntrans, hipMemcpyDeviceToHost, c.st));
            HIPCHK(hipStreamSynchronize(c.st));
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
    int B = 1;                      // frames per batch
    DevBuf wsc, wtab, at_ind, rec, key, cstr, srec, cs, H, Hs, bsum, chunks_d, tinfo_d, binstart, bad, G1, G2;
    std::vector<SortChunk> chunks;      // for B frames
    std::vector<int> nchunks_upto;      // chunks of the first b frames
    std::vector<int64_t> H_upto;        // count-matrix size of the first b frames
    std::vector<TransformInfo> tinfo;
    // reduction caches
    DevBuf win_d, twT;
    StageTimer tm;
    std::vector<double> win_h;
    std::vector<Cx<R>> twT_h;
    int win_n = -1, win_kind = -1, twT_n = -1;
    RadixPlan rpT{};

    void create(const mpsq_sqw_spec *s)
    {
        if (!s || (s->dim != 2 && s->dim != 3) || s->nsp < 1 || s->nsp > MAXSP || !s->sp_count || s->n_tot < 1)
            throw Err{MPSQ_ERR_BAD_ARG};
        core.setup(s->dim, s->n_modes, s->iflag, s->eps, &s->opts);
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
        const int64_t NQ = core.N[0] * core.N[1] * core.N[2];
        // weights of the active species
        std::vector<R> wh((size_t)nact);
        for (int k = 0; k < nact; ++k) wh[k] = s->w_scalar ? (R)s->w_scalar[act[k]] : (R)1;
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
        // batch size: intermediate buffers of about 2 GiB per batch, <= 16 frames
        const size_t per_frame = (size_t)Mact * (2 * sizeof(PtRec<R>) + 2 * sizeof(Cx<R>) + 2 * sizeof(int)) +
                                 (size_t)nact * sizeof(Cx<R>) *
                                     (size_t)(core.n[2] * core.n[1] * core.N[0] +
                                              (core.dim == 3 ? core.n[2] * core.N[1] * core.N[0] : 0));
        int b = s->batch_frames > 0 ? s->batch_frames : (int)std::max<size_t>(1, ((size_t)2 << 30) / per_frame);
        B = std::max(1, std::min(b, 16));
        build_chunks();
        alloc_batch();
    }

    // sort chunk table for B frames; transform t = b*nact + k, points of
    // t are [b*Mact + act_off[k], b*Mact + act_off[k+1])
    void build_chunks()
    {
        const int64_t CH = core.sort_chunk;
        chunks.clear();
        tinfo.clear();
        nchunks_upto.assign((size_t)B + 1, 0);
        H_upto.assign((size_t)B + 1, 0);
        int64_t hbase = 0;
        for (int b = 0; b < B; ++b) {
            for (int k = 0; k < nact; ++k) {
                const int64_t p0 = (int64_t)b * Mact + act_off[k], p1 = (int64_t)b * Mact + act_off[k + 1];
                const int nch = (int)((p1 - p0 + CH - 1) / CH);
                for (int c = 0; c < nch; ++c) {
                    SortChunk ch;
                    ch.p0 = p0 + (int64_t)c * CH;
                    ch.p1 = std::min(p1, p0 + (int64_t)(c + 1) * CH);
                    ch.pbase = p0;
                    ch.hbase = hbase;
                    ch.nch = nch;
                    ch.lc = c;
                    chunks.push_back(ch);
                }
                tinfo.push_back(TransformInfo{p0, p1, hbase, nch, 0});
                hbase += (int64_t)core.g.nbins * nch;
            }
            nchunks_upto[b + 1] = (int)chunks.size();
            H_upto[b + 1] = hbase;
        }
    }

    void alloc_batch()
    {
        const size_t np = (size_t)B * (size_t)std::max<int64_t>(Mact, 1);
        rec.ensure(sizeof(PtRec<R>) * np);
        srec.ensure(sizeof(PtRec<R>) * np);
        key.ensure(sizeof(int) * np);
        cstr.ensure(sizeof(Cx<R>) * np);
        cs.ensure(sizeof(Cx<R>) * np);
        const int64_t Htot = std::max<int64_t>(H_upto[B], 1);
        H.ensure(sizeof(int) * (size_t)Htot);
        Hs.ensure(sizeof(int64_t) * (size_t)Htot);
        bsum.ensure(sizeof(int64_t) * (size_t)scan_bsum_size(Htot));
        chunks_d.ensure(sizeof(SortChunk) * std::max<size_t>(chunks.size(), 1));
        tinfo_d.ensure(sizeof(TransformInfo) * std::max<size_t>(tinfo.size(), 1));
        binstart.ensure(sizeof(int64_t) * (size_t)(core.g.nbins + 1) * tinfo.size());
        bad.ensure(sizeof(int));
        if (!chunks.empty())
            HIPCHK(hipMemcpy(chunks_d.p, chunks.data(), sizeof(SortChunk) * chunks.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(tinfo_d.p, tinfo.data(), sizeof(TransformInfo) * tinfo.size(), hipMemcpyHostToDevice));
        const size_t ntr = (size_t)B * nact;
        G1.ensure(sizeof(Cx<R>) * ntr * (size_t)(core.n[2] * core.n[1] * core.N[0]));
        if (core.dim == 3) G2.ensure(sizeof(Cx<R>) * ntr * (size_t)(core.n[2] * core.N[1] * core.N[0]));
    }

    void amplitudes(const float *d_pos, int64_t T, void *d_amp)
    {
        Core<R> &c = core;
        if (!d_pos || !d_amp || T < 1) throw Err{MPSQ_ERR_BAD_ARG};
        HIPCHK(hipSetDevice(c.device));
        const int64_t NQ = c.N[0] * c.N[1] * c.N[2];
        HIPCHK(hipMemsetAsync(bad.p, 0, sizeof(int), c.st));
        Cx<R> *amp = static_cast<Cx<R> *>(d_amp);
        for (int64_t f0 = 0; f0 < T; f0 += B) {
            const int b = (int)std::min<int64_t>(B, T - f0);
            const int ntr = b * nact;
            // stage all active atoms of b frames.  The kernel walks atoms of
            // the frame in the active-species order via the position offset.
            int ev = tm.begin(ST_STAGE, c.st);
            stage_batch(d_pos + (size_t)f0 * n_tot * 4, b);
            tm.end(ev, c.st);
            ev = tm.begin(ST_SORT, c.st);
            const int nch = nchunks_upto[b];
            launch_sort_count(key.as<int>(), chunks_d.as<SortChunk>(), nch, c.g.nbins, H.as<int>(), c.st);
            launch_scan(H.as<int>(), H_upto[b], bsum.as<int64_t>(), Hs.as<int64_t>(), c.st);
            launch_sort_scatter<R>(key.as<int>(), chunks_d.as<SortChunk>(), nch, c.g.nbins, c.keybits(),
                                   Hs.as<int64_t>(), rec.as<PtRec<R>>(), cstr.as<Cx<R>>(), srec.as<PtRec<R>>(),
                                   cs.as<Cx<R>>(), nullptr, c.st);
            launch_binstart(Hs.as<int64_t>(), tinfo_d.as<TransformInfo>(), ntr, c.g.nbins, binstart.as<int64_t>(),
                            c.st);
            tm.end(ev, c.st);
            ev = tm.begin(ST_SPREAD, c.st);
            launch_spread_xfft<R>(c.g, c.rp[0], ntr, srec.as<PtRec<R>>(), cs.as<Cx<R>>(), 0,
                                  binstart.as<int64_t>(), 1,
                                  c.tw[0].template as<Cx<R>>(), c.phinv[0].template as<R>(), G1.as<Cx<R>>(), 0,
                                  nullptr, c.st);
            tm.end(ev, c.st);
            if (combine)
                c.passes(ntr, b, nact, G1.as<Cx<R>>(), G2.as<Cx<R>>(), amp + (size_t)f0 * NQ, wsc.as<R>(),
                         wtab.p ? wtab.as<float>() : nullptr, &tm);
            else
                c.passes(ntr, ntr, 1, G1.as<Cx<R>>(), G2.as<Cx<R>>(), amp + (size_t)f0 * nact * NQ, nullptr, nullptr,
                         &tm);
        }
        HIPCHK(hipGetLastError());
        int hbad = 0;
        HIPCHK(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, c.st));
        HIPCHK(hipStreamSynchronize(c.st));
        tm.collect();
        if (hbad) throw Err{MPSQ_ERR_NONFINITE_POINT};
    }

    // Stage b frames.  When every species is active the batch's points are
    // the frames' atoms in storage order; otherwise one launch per active
    // species block maps its atoms onto the active point range.
    void stage_batch(const float *pos, int b)
    {
        Core<R> &c = core;
        const float4 *p4 = reinterpret_cast<const float4 *>(pos);
        const int4 *ai = at_ind.p ? at_ind.as<int4>() : nullptr;
        if (nact == nsp) {
            launch_stage<R>(c.g, sp, (int64_t)b * n_tot, p4, ai, rec.as<PtRec<R>>(), key.as<int>(),
                            cstr.as<Cx<R>>(), bad.as<int>(), c.st);
            return;
        }
        for (int fb = 0; fb < b; ++fb)
            for (int k = 0; k < nact; ++k) {
                const int j = act[k];
                const int64_t cnt = sp_count[j];
                if (cnt == 0) continue;
                StageParams s2 = sp;
                s2.n_tot = n_tot;  // atom index = p % n_tot stays valid with the offsets below
                const int64_t dst = (int64_t)fb * Mact + act_off[k];
                // the kernel indexes pos[p] and at_ind[p % n_tot]; pass views
                // shifted so that p = 0 is atom sp_off[j] of frame fb
                launch_stage_view(s2, p4 + (size_t)fb * n_tot + sp_off[j], ai ? ai + sp_off[j] : nullptr, cnt, dst);
            }
    }

    void launch_stage_view(StageParams s2, const float4 *p4, const int4 *ai, int64_t cnt, int64_t dst)
    {
        Core<R> &c = core;
        s2.n_tot = cnt;  // at_ind view indexed by p % cnt == p
        launch_stage<R>(c.g, s2, cnt, p4, ai, rec.as<PtRec<R>>() + dst, key.as<int>() + dst, cstr.as<Cx<R>>() + dst,
                        bad.as<int>(), c.st);
    }

    void reduce(const void *d_amp, int64_t T, const mpsq_reduce_spec *r, double *d_S, double *d_Stot)
    {
        Core<R> &c = core;
        if (!combine || !d_amp || !r || (!d_S && r->n_f > 0) || T < 1) throw Err{MPSQ_ERR_BAD_ARG};
        if (r->n_int < 1 || r->n_int > T || r->n_f < 0 || !(r->nat > 0)) throw Err{MPSQ_ERR_BAD_ARG};
        HIPCHK(hipSetDevice(c.device));
        const int64_t NQ = c.N[0] * c.N[1] * c.N[2];
        SqwParams p{};
        p.T = T;
        p.NQ = NQ;
        p.n_int = (int)r->n_int;
        p.n_frame = T - r->n_int + 1;
        p.m_lo = r->m_lo;
        p.n_f = (int)r->n_f;
        p.win_kind = r->win_kind;
        p.subtract_mean = r->subtract_mean;
        p.norm_s = r->dt / (r->nat * r->nat * 0.5 * (double)p.n_frame * (double)p.n_int);
        p.norm_tot = 1.0 / ((double)T * r->nat * r->nat);
        const Cx<R> *amp = static_cast<const Cx<R> *>(d_amp);
        RadixPlan rp{};
        const bool single = (p.n_frame == 1) && (T <= sqw_tile_elems<R>()) && radix_plan((int)T, rp);
        if (single) {
            if (win_n != (int)T || win_kind != r->win_kind) {
                win_h.resize((size_t)T);
                for (int64_t i = 0; i < T; ++i) {
                    if (T <= 1) win_h[i] = 1.0;
                    else if (r->win_kind == 0) win_h[i] = 0.5 * (1.0 - std::cos(2.0 * M_PI * (double)i / (double)(T - 1)));
                    else win_h[i] = 0.5 * (1.0 - std::cos(2.0 * M_PI * (double)(i + 1) / (double)(T + 1)));
                }
                win_d.ensure(sizeof(double) * (size_t)T);
                HIPCHK(hipMemcpyAsync(win_d.p, win_h.data(), sizeof(double) * (size_t)T, hipMemcpyHostToDevice, c.st));
                win_n = (int)T;
                win_kind = r->win_kind;
            }
            if (twT_n != (int)T) {
                upload_twiddles<R>(twT, (int)T, c.st, twT_h);
                twT_n = (int)T;
                rpT = rp;
            }
            int cols = 1;
            while (2 * cols * T <= sqw_tile_elems<R>() && 2 * cols <= 64) cols *= 2;
            const int ev = tm.begin(ST_REDUCE, c.st);
            launch_sqw_single<R>(p, rpT, cols, amp, twT.as<Cx<R>>(), win_d.as<double>(), r->d_qfac, d_S, d_Stot, c.st);
            tm.end(ev, c.st);
        } else {
            const int ev = tm.begin(ST_REDUCE, c.st);
            launch_sqw_sliding<R>(p, amp, r->d_qfac, d_S, d_Stot, c.st);
            tm.end(ev, c.st);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c.st));
        tm.collect();
    }
};

}  // namespace mpsq

using namespace mpsq;

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

extern "C" {

void mpsq_default_opts(mpsq_opts *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
}

const char *mpsq_version(void) { return "mpsq 0.1 (gfx950)"; }

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

}  // extern "C"
