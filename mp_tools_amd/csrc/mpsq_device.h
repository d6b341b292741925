// mpsq_device.h -- device-side building blocks shared by the NUFFT and the
// S(Q,w) reduction kernels (gfx950, wave64).
//
//  * complex helpers for float2 / double2
//  * in-register DFT butterflies of radix 2, 3, 4, 5, 8 (forward sign -)
//  * a block-cooperative, in-place Stockham FFT over an LDS tile holding
//    `cnt` independent length-n transforms at buf[t*ts + i*es]
//
// The Stockham stage follows the autosort formulation of Govindaraju et al.
// (SC'08): butterfly j of a radix-R stage reads x[j + r*n/R] (twiddled by
// W_n^{r*(j mod Ns)*n/(Ns*R)}) and writes y[(j/Ns)*Ns*R + (j mod Ns) + r*Ns];
// after all stages the data is in natural order.  The tile is transformed in
// place: every thread first loads all its butterflies into registers, the
// block synchronises, and then every thread stores, so no LDS ping-pong
// buffer is needed.  Twiddles come from a table tw[m] = exp(-2 pi i m/n)
// computed in FP64 on the host.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpsq {

template <typename R> struct cplx;
template <> struct cplx<float> { using T = float2; };
template <> struct cplx<double> { using T = double2; };

template <typename C> __device__ __forceinline__ C cmk(decltype(C::x) a, decltype(C::x) b)
{
    C r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return cmk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return cmk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b)
{
    return cmk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C> __device__ __forceinline__ C cscale(C a, decltype(C::x) s) { return cmk<C>(a.x * s, a.y * s); }
// multiply by -i
template <typename C> __device__ __forceinline__ C cmul_mi(C a) { return cmk<C>(a.y, -a.x); }
template <typename C> __device__ __forceinline__ C cconj(C a) { return cmk<C>(a.x, -a.y); }

// ---------------------------------------------------------------------------
// Loads and stores through global-address-space pointers.  A pointer read
// from a job table is generic, and generic accesses compile to flat
// instructions, which count against the LDS counter too: a wait for an LDS
// result then also waits for every flat load in flight (a candidate
// prefetch, say).  These helpers issue global_load / global_store.
// ---------------------------------------------------------------------------
#define MPSQ_GLOBAL __attribute__((opencl_global))
typedef unsigned int mpsq_u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int mpsq_u32x4 __attribute__((ext_vector_type(4)));
template <int N> struct U32x4s {
    mpsq_u32x4 v[N];
};
template <typename T> __device__ __forceinline__ T gload(const T *p)
{
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, *(const MPSQ_GLOBAL unsigned int *)p);
    } else if constexpr (sizeof(T) == 8) {
        return __builtin_bit_cast(T, *(const MPSQ_GLOBAL mpsq_u32x2 *)p);
    } else {
        static_assert(sizeof(T) % 16 == 0, "gload: 4, 8 or 16n bytes");
        U32x4s<sizeof(T) / 16> b;
#pragma unroll
        for (int k = 0; k < (int)(sizeof(T) / 16); ++k) b.v[k] = ((const MPSQ_GLOBAL mpsq_u32x4 *)p)[k];
        return __builtin_bit_cast(T, b);
    }
}
template <typename T> __device__ __forceinline__ void gstore(T *p, const T &v)
{
    if constexpr (sizeof(T) == 4) {
        *(MPSQ_GLOBAL unsigned int *)p = __builtin_bit_cast(unsigned int, v);
    } else if constexpr (sizeof(T) == 8) {
        *(MPSQ_GLOBAL mpsq_u32x2 *)p = __builtin_bit_cast(mpsq_u32x2, v);
    } else {
        static_assert(sizeof(T) % 16 == 0, "gstore: 4, 8 or 16n bytes");
        const U32x4s<sizeof(T) / 16> b = __builtin_bit_cast(U32x4s<sizeof(T) / 16>, v);
#pragma unroll
        for (int k = 0; k < (int)(sizeof(T) / 16); ++k) ((MPSQ_GLOBAL mpsq_u32x4 *)p)[k] = b.v[k];
    }
}

// ---------------------------------------------------------------------------
// forward DFT butterflies, X_k = sum_n a_n exp(-2 pi i n k / R)
// ---------------------------------------------------------------------------
template <int R, typename C> struct Bfly;

template <typename C> struct Bfly<2, C> {
    __device__ __forceinline__ static void run(C *v)
    {
        C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <typename C> struct Bfly<4, C> {
    __device__ __forceinline__ static void run(C *v)
    {
        C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        C t2 = cadd(v[1], v[3]), t3 = cmul_mi(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

template <typename C> struct Bfly<8, C> {
    __device__ __forceinline__ static void run(C *v)
    {
        using R = decltype(C::x);
        const R h = (R)0.70710678118654752440084436210485;
        C e[4] = {v[0], v[2], v[4], v[6]};
        C o[4] = {v[1], v[3], v[5], v[7]};
        Bfly<4, C>::run(e);
        Bfly<4, C>::run(o);
        // W8^1 = (1 - i)/sqrt2, W8^2 = -i, W8^3 = (-1 - i)/sqrt2
        C o1 = cmk<C>((o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h);
        C o2 = cmul_mi(o[2]);
        C o3 = cmk<C>((o[3].y - o[3].x) * h, -(o[3].x + o[3].y) * h);
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};

template <typename C> struct Bfly<3, C> {
    __device__ __forceinline__ static void run(C *v)
    {
        using R = decltype(C::x);
        const R s3 = (R)0.86602540378443864676372317075294;  // sin(2 pi/3)
        C t1 = cadd(v[1], v[2]);
        C t2 = csub(v[0], cscale(t1, (R)0.5));
        C d = csub(v[1], v[2]);
        C t3 = cmk<C>(d.y * s3, -d.x * s3);  // (-i sqrt3/2) * d
        v[0] = cadd(v[0], t1);
        v[1] = cadd(t2, t3);
        v[2] = csub(t2, t3);
    }
};

template <typename C> struct Bfly<5, C> {
    __device__ __forceinline__ static void run(C *v)
    {
        using R = decltype(C::x);
        const R c1 = (R)0.30901699437494742410229341718282;   // cos(2 pi/5)
        const R c2 = (R)-0.80901699437494742410229341718282;  // cos(4 pi/5)
        const R s1 = (R)0.95105651629515357211643933337938;   // sin(2 pi/5)
        const R s2 = (R)0.58778525229247312916870595463907;   // sin(4 pi/5)
        C b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]);
        C d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
        C a0 = v[0];
        C p1 = cadd(a0, cadd(cscale(b1, c1), cscale(b2, c2)));
        C p2 = cadd(a0, cadd(cscale(b1, c2), cscale(b2, c1)));
        C q1 = cadd(cscale(d1, s1), cscale(d2, s2));   // multiplied by -i below
        C q2 = csub(cscale(d1, s2), cscale(d2, s1));
        C mq1 = cmul_mi(q1), mq2 = cmul_mi(q2);
        v[0] = cadd(a0, cadd(b1, b2));
        v[1] = cadd(p1, mq1);
        v[4] = csub(p1, mq1);
        v[2] = cadd(p2, mq2);
        v[3] = csub(p2, mq2);
    }
};

// Division by a runtime constant d (1 <= d < 2^31) of x < 2^31 without a
// divide: q = (umulhi(x, m) + x) >> s with s = ceil(log2 d) and
// m = floor(2^(32+s) / d) + 1 - 2^32 (the round-up multiplier of Granlund &
// Montgomery, PLDI 1994; the 33-bit multiplier m + 2^32 is split so that
// every product fits 32-bit hardware).  Exact for all x < 2^31.
struct FDiv {
    uint32_t m;
    uint32_t s;
};
__host__ __device__ inline FDiv make_fdiv(uint32_t d)
{
    uint32_t s = 0;
    while (s < 31 && (1u << s) < d) ++s;
    const uint64_t two32s = (uint64_t)1 << (32 + s);
    FDiv f;
    f.m = (uint32_t)(two32s / d + 1 - ((uint64_t)1 << 32));
    f.s = s;
    return f;
}
__host__ __device__ __forceinline__ uint32_t fdiv(uint32_t x, FDiv f)
{
#ifdef __HIP_DEVICE_COMPILE__
    const uint32_t hi = __umulhi(x, f.m);
#else
    const uint32_t hi = (uint32_t)(((uint64_t)x * f.m) >> 32);
#endif
    return (hi + x) >> f.s;
}

// Radix plan of one transform length: stages applied in order, with the
// divisors every stage needs (n / R and the sub-transform length Ns).
struct RadixPlan {
    int n;
    int nst;
    int r[24];
    FDiv dnb[24];  // n / r[s]
    FDiv dns[24];  // Ns = r[0] * ... * r[s-1]
    int tws[24];   // twiddle stride n / (Ns * r[s])
};

// Position of element i of a row-major transform when rows are padded with
// one spare element per 16 (keeps the stride-R writes of the early Stockham
// stages off a single LDS bank pair).
template <bool PAD> __device__ __forceinline__ int padx(int i) { return PAD ? i + (i >> 4) : i; }
__host__ __device__ inline int padded_len(int n) { return n + (n >> 4) + 1; }

// One Stockham stage over an LDS tile.  NT threads, each holding at most
// ceil(E/R) butterflies (E = elements per thread the tile was sized for).
// col_major: consecutive threads take consecutive transforms (use when the
// transforms are the contiguous index, es > 1); otherwise consecutive threads
// take consecutive butterflies of one transform (es == 1, rows may be padded).
template <typename C, int E, int R, bool PAD>
__device__ __forceinline__ void stockham_stage(C *buf, int n, int cnt, FDiv dcnt, FDiv dnb, FDiv dns, int es, int ts,
                                               int Ns, const C *__restrict__ tw, bool col_major)
{
    constexpr int TPT = (E + R - 1) / R;
    const int nb = n / R;
    const int total = nb * cnt;
    const int tid = threadIdx.x, NT = blockDim.x;
    const int tstride = n / (Ns * R);
    C v[TPT * R];
    int tt[TPT], jj[TPT];
#pragma unroll
    for (int i = 0; i < TPT; ++i) {
        const int task = tid + i * NT;
        if (task < total) {
            int t, j;
            if (col_major) {
                j = (int)fdiv((uint32_t)task, dcnt);
                t = task - j * cnt;
            } else {
                t = (int)fdiv((uint32_t)task, dnb);
                j = task - t * nb;
            }
            tt[i] = t;
            jj[i] = j;
            const C *src = buf + t * ts;
            const int k = j - (int)fdiv((uint32_t)j, dns) * Ns;
            const int tk = k * tstride;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                C a = src[padx<PAD>(j + r * nb) * es];
                if (r > 0) a = cmul(a, tw[r * tk]);
                v[i * R + r] = a;
            }
            Bfly<R, C>::run(&v[i * R]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < TPT; ++i) {
        const int task = tid + i * NT;
        if (task < total) {
            const int t = tt[i], j = jj[i];
            C *dst = buf + t * ts;
            const int k = j - (int)fdiv((uint32_t)j, dns) * Ns;
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[padx<PAD>(base + r * Ns) * es] = v[i * R + r];
        }
    }
    __syncthreads();
}

// Full forward FFT of the tile (natural order in, natural order out).
// dcnt = make_fdiv(cnt) (used in col_major mode).
template <typename C, int E, bool PAD = false>
__device__ __forceinline__ void lds_fft(C *buf, const RadixPlan &rp, int cnt, FDiv dcnt, int es, int ts, const C *__restrict__ tw,
                        bool col_major)
{
    int Ns = 1;
    for (int s = 0; s < rp.nst; ++s) {
        const int R = rp.r[s];
        const FDiv dnb = rp.dnb[s], dns = rp.dns[s];
        switch (R) {
        case 8: stockham_stage<C, E, 8, PAD>(buf, rp.n, cnt, dcnt, dnb, dns, es, ts, Ns, tw, col_major); break;
        case 4: stockham_stage<C, E, 4, PAD>(buf, rp.n, cnt, dcnt, dnb, dns, es, ts, Ns, tw, col_major); break;
        case 2: stockham_stage<C, E, 2, PAD>(buf, rp.n, cnt, dcnt, dnb, dns, es, ts, Ns, tw, col_major); break;
        case 3: stockham_stage<C, E, 3, PAD>(buf, rp.n, cnt, dcnt, dnb, dns, es, ts, Ns, tw, col_major); break;
        case 5: stockham_stage<C, E, 5, PAD>(buf, rp.n, cnt, dcnt, dnb, dns, es, ts, Ns, tw, col_major); break;
        default: break;
        }
        Ns *= R;
    }
}

}  // namespace mpsq
