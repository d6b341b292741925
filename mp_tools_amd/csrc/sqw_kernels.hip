// sqw_kernels.hip -- S(Q) / S(Q,w) reduction of combined amplitudes C(t,Q)
// (layout [T][NQ], Q fastest) on gfx950.
//
// k_sqw_single   one Hann window over all T frames (n_int == T): a block
//                stages T x cols amplitudes in LDS, forms the per-Q mean and
//                sum |C|^2 in FP64, subtracts the mean, applies the window
//                (and the optional per-Q factor), runs the time FFT down the
//                LDS columns, and writes |X(m)|^2 in FP64
//                (mp_sqom56.f90:1177-1185, 1381-1387, 1480-1517, 1527, 1576).
// k_sqw_sliding  sliding windows with hop 1 (n_frame = T - n_int + 1 > 1),
//                any n_int: the Hann window is a sum of three exponentials,
//                so each windowed DFT X_s(m) is a combination of three
//                rectangular sliding DFTs that advance by an O(1) recurrence
//                per frame (FP64); sum_s |X_s(m)|^2 accumulates in FP64
//                (mp_sqom56.f90:1437-1478 computes the same sum per energy).
// k_sqw_window1  one window over all frames when T has no 2,3,5 radix plan:
//                one DFT recurrence per (Q, frequency), window from a table.
// k_iqt          I(Q,t) (mp_sqom56.f90:1467-1471, 1504-1510) by direct sums,
//                a thread per Q and group of lags;
// k_iqt_fft      the same for many lags by FFT correlation in LDS (FP64).
// k_dos_fused    mp_dos spectra: time FFT of packed velocity columns and
//                per-block species sums (mp_dos56.f90:495-525), added in a
//                fixed order by k_dos_groups / k_dos_final.
// k_qtables      the Q plane's q2_norm and X-ray form factor tables
//                (mp_sqom56.f90:901-936); k_xray_table the per-Q X-ray weights
//                of served requests on the mode grid.
#include "mpsq_kernels.h"

namespace mpsq {

// LDS tiles of the time-FFT kernels: E elements per thread, NT threads
template <typename R> struct SqwTiles;
template <> struct SqwTiles<float> { static constexpr int E = 16, NT = 512; };
template <> struct SqwTiles<double> { static constexpr int E = 8, NT = 512; };

// amplitude column of output q; -1 for an index list entry outside
// [0, NQs), whose outputs are written as NaN (nothing is read for it)
__device__ __forceinline__ int64_t src_q(const SqwParams &p, int64_t q)
{
    if (!p.qidx) return q;
    const int64_t s = p.qidx[q];
    return (s >= 0 && s < p.NQs) ? s : -1;
}
__device__ __forceinline__ double qnan() { return __longlong_as_double(0x7ff8000000000000LL); }

template <typename R, int E, int NT>
__global__ __launch_bounds__(NT) void k_sqw_single(SqwParams p, RadixPlan rp, int cols, FDiv dcols, FDiv dT,
                                                   const typename cplx<R>::T *__restrict__ amp,
                                                   const typename cplx<R>::T *__restrict__ tw,
                                                   const double *__restrict__ win, const float *__restrict__ qfac,
                                                   double *__restrict__ S, double *__restrict__ Stot)
{
    using C = typename cplx<R>::T;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C *tile = reinterpret_cast<C *>(smem_raw);
    C *tw_l = tile + (int)p.T * cols;  // twiddles in LDS
    double *red = reinterpret_cast<double *>(tw_l + (int)p.T);  // [NT / 64][cols][3]: per-wave column sums
    for (int i = threadIdx.x; i < (int)p.T; i += NT) tw_l[i] = tw[i];
    __shared__ double mean_r[SQW_MAX_COLS], mean_i[SQW_MAX_COLS];
    const int T = (int)p.T;
    const int64_t q0 = (int64_t)blockIdx.x * cols;
    const int ncol = (int)min((int64_t)cols, p.NQ - q0);
    // element i = threadIdx.x + k * NT of the tile is frame i / cols, column
    // i % cols; NT is a multiple of cols, so a thread's elements all belong
    // to column c = threadIdx.x % cols (frames seg, seg + nseg, ...)
    const int c = threadIdx.x % cols, seg = threadIdx.x / cols, nseg = NT / cols;
    const int64_t sc = c < ncol ? src_q(p, q0 + c) : -1;
    {
        // all loads of a thread in flight before any use; the column sums
        // from the registers, frames in increasing order, as the values go
        // to the tile
        C v[E];
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int t = seg + k * nseg;
            v[k] = (t < T && sc >= 0) ? amp[(int64_t)t * p.NQs + sc] : cmk<C>((R)0, (R)0);
        }
        double sr = 0.0, si = 0.0, ss = 0.0;
#pragma unroll
        for (int k = 0; k < E; ++k)
            if (seg + k * nseg < T) {
                tile[(seg + k * nseg) * cols + c] = v[k];
                sr += (double)v[k].x;
                si += (double)v[k].y;
                ss += (double)v[k].x * (double)v[k].x + (double)v[k].y * (double)v[k].y;
            }
        // the lanes of a wave that share c (64 / cols of them) are combined
        // by shuffles, then the waves' partials in LDS
        for (int o = cols; o < 64; o <<= 1) {
            sr += __shfl_xor(sr, o, 64);
            si += __shfl_xor(si, o, 64);
            ss += __shfl_xor(ss, o, 64);
        }
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        if (lane < cols) {
            red[(wv * cols + lane) * 3 + 0] = sr;
            red[(wv * cols + lane) * 3 + 1] = si;
            red[(wv * cols + lane) * 3 + 2] = ss;
        }
        __syncthreads();
        if (threadIdx.x < cols) {
            double a = 0.0, b = 0.0, q = 0.0;
            for (int k = 0; k < NT / 64; ++k) {
                a += red[(k * cols + threadIdx.x) * 3 + 0];
                b += red[(k * cols + threadIdx.x) * 3 + 1];
                q += red[(k * cols + threadIdx.x) * 3 + 2];
            }
            mean_r[threadIdx.x] = p.subtract_mean ? a / (double)T : 0.0;
            mean_i[threadIdx.x] = p.subtract_mean ? b / (double)T : 0.0;
            if (Stot && threadIdx.x < ncol)
                Stot[q0 + threadIdx.x] = src_q(p, q0 + threadIdx.x) >= 0 ? q * p.norm_tot : qnan();
        }
        __syncthreads();
    }
    {
        // mean subtracted, windowed (and the per-Q factor): each thread its
        // own elements of the tile
        const double mr = mean_r[c], mi = mean_i[c];
        const double qf = (qfac && sc >= 0) ? (double)qfac[sc] : 1.0;
        for (int t = seg; t < T; t += nseg) {
            const C x = tile[t * cols + c];
            double f = win[t];
            if (qfac && sc >= 0) f *= qf;
            tile[t * cols + c] = cmk<C>((R)(((double)x.x - mr) * f), (R)(((double)x.y - mi) * f));
        }
    }
    __syncthreads();
    lds_fft<C, E>(tile, rp, cols, dcols, cols, 1, tw_l, true);
    int m0 = (int)(p.m_lo % T);
    if (m0 < 0) m0 += T;
    for (int i = threadIdx.x; i < p.n_f * cols; i += NT) {
        const int mi = (int)fdiv((uint32_t)i, dcols), c = i - mi * cols;
        if (c >= ncol) continue;
        const int mm = m0 + mi;
        const int m = mm - (int)fdiv((uint32_t)mm, dT) * T;
        const C v = tile[m * cols + c];
        S[(int64_t)mi * p.NQ + q0 + c] = src_q(p, q0 + c) >= 0
                                             ? ((double)v.x * (double)v.x + (double)v.y * (double)v.y) * p.norm_s
                                             : qnan();
    }
}

// One window over all T frames (n_frame == 1) when T has no 2,3,5 radix
// plan: thread = (Q, frequency), X(m) = sum_i w_i C'(i) e^{-2 pi i m i / n}
// by a running twiddle with the window w from the table (one complex
// product per frame, against three for the sliding kernel's decomposition of
// the window into exponentials); |X|^2 and all sums in FP64.
template <typename R>
__global__ __launch_bounds__(256, 8) void k_sqw_window1(SqwParams p, const typename cplx<R>::T *__restrict__ amp,
                                                     const float *__restrict__ qfac, const double *__restrict__ win,
                                                     double *__restrict__ S, double *__restrict__ Stot)
{
    const int64_t q = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    const int mi_ = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (q >= p.NQ || mi_ >= p.n_f) return;
    const int n = p.n_int;
    const int64_t qs = src_q(p, q);
    if (qs < 0) {
        if (mi_ == 0 && Stot) Stot[q] = qnan();
        S[(int64_t)mi_ * p.NQ + q] = qnan();
        return;
    }
    amp += qs;
    double mr = 0.0, mim = 0.0, ss = 0.0;
    for (int i = 0; i < n; ++i) {
        const auto v = amp[(int64_t)i * p.NQs];
        mr += (double)v.x;
        mim += (double)v.y;
        ss += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
    }
    if (mi_ == 0 && Stot) Stot[q] = ss * p.norm_tot;
    if (p.subtract_mean) {
        mr /= (double)n;
        mim /= (double)n;
    } else {
        mr = mim = 0.0;
    }
    const double qf = qfac ? (double)qfac[qs] : 1.0;
    int64_t m = (p.m_lo + (int64_t)mi_) % n;
    if (m < 0) m += n;
    const double om = 6.283185307179586476925286766559 * (double)m / (double)n;
    const double rr = cos(om), ri = sin(om);
    double sr = 1.0, si = 0.0, dr = 0.0, di = 0.0;
    for (int i = 0; i < n; ++i) {
        const auto v = amp[(int64_t)i * p.NQs];
        const double f = win[i] * qf;
        const double cr = ((double)v.x - mr) * f, ci = ((double)v.y - mim) * f;
        dr += sr * cr - si * ci;
        di += sr * ci + si * cr;
        const double nr = sr * rr + si * ri;  // times e^{-i om}
        si = si * rr - sr * ri;
        sr = nr;
    }
    S[(int64_t)mi_ * p.NQ + q] = (dr * dr + di * di) * p.norm_s;
}

template <typename R>
void launch_sqw_window1(const SqwParams &p, const typename cplx<R>::T *amp, const float *qfac, const double *win,
                        double *S, double *Stot, hipStream_t st)
{
    const unsigned gx = (unsigned)((p.NQ + 63) / 64);
    const unsigned gy = (unsigned)((p.n_f + 3) / 4);
    hipLaunchKernelGGL((k_sqw_window1<R>), dim3(gx, gy > 0 ? gy : 1), dim3(256), 0, st, p, amp, qfac, win, S, Stot);
}

// sliding windows: thread = (Q, group of MG frequencies; the launch uses 1)
template <typename R, int MG>
__global__ __launch_bounds__(256, 8) void k_sqw_sliding(SqwParams p, const typename cplx<R>::T *__restrict__ amp,
                                                     const float *__restrict__ qfac, double *__restrict__ S,
                                                     double *__restrict__ Stot)
{
    const int64_t q = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    const int grp = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (q >= p.NQ || grp * MG >= p.n_f) return;
    const int64_t T = p.T;
    const int n = p.n_int;
    const int64_t qs = src_q(p, q);
    if (qs < 0) {
        if (grp == 0 && Stot) Stot[q] = qnan();
        for (int j = 0; j < MG; ++j)
            if (grp * MG + j < p.n_f) S[(int64_t)(grp * MG + j) * p.NQ + q] = qnan();
        return;
    }
    amp += qs;
    double mr = 0.0, mi = 0.0, ss = 0.0;
    for (int64_t t = 0; t < T; ++t) {
        const auto v = amp[t * p.NQs];
        mr += (double)v.x;
        mi += (double)v.y;
        ss += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
    }
    if (grp == 0 && Stot) Stot[q] = ss * p.norm_tot;
    if (p.subtract_mean) {
        mr /= (double)T;
        mi /= (double)T;
    } else {
        mr = mi = 0.0;
    }
    const double qf = qfac ? (double)qfac[qs] : 1.0;
    const double twopi = 6.283185307179586476925286766559;
    // window = a0 + a1 e^{i th i} + a2 e^{-i th i}
    double th = 0.0, a0 = 1.0, a1r = 0.0, a1i = 0.0;
    if (n > 1) {
        const int off = p.win_kind == 1 ? 1 : 0;
        th = twopi / (double)(p.win_kind == 1 ? n + 1 : n - 1);
        a0 = 0.5;
        a1r = -0.25 * cos(th * off);
        a1i = -0.25 * sin(th * off);
    }
    // a2 = conj(a1)
    double Dr[MG][3], Di[MG][3], rr[MG][3], ri[MG][3], trr[MG][3], tri[MG][3], sr[MG][3], si[MG][3];
    double acc[MG];
#pragma unroll
    for (int j = 0; j < MG; ++j) {
        int64_t m = (p.m_lo + (int64_t)(grp * MG + j)) % n;
        if (m < 0) m += n;
        const double om = twopi * (double)m / (double)n;
        const double oms[3] = {om, om - th, om + th};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            Dr[j][k] = Di[j][k] = 0.0;
            rr[j][k] = cos(oms[k]);          // e^{+i w}
            ri[j][k] = sin(oms[k]);
            trr[j][k] = cos(oms[k] * (double)(n - 1));   // e^{-i w (n-1)}
            tri[j][k] = -sin(oms[k] * (double)(n - 1));
            sr[j][k] = 1.0;                  // running e^{-i w i}
            si[j][k] = 0.0;
        }
        acc[j] = 0.0;
    }
    for (int i = 0; i < n; ++i) {
        const auto v = amp[(int64_t)i * p.NQs];
        const double cr = ((double)v.x - mr) * qf, ci = ((double)v.y - mi) * qf;
#pragma unroll
        for (int j = 0; j < MG; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                Dr[j][k] += sr[j][k] * cr - si[j][k] * ci;
                Di[j][k] += sr[j][k] * ci + si[j][k] * cr;
                // multiply running twiddle by e^{-i w}
                const double nr = sr[j][k] * rr[j][k] + si[j][k] * ri[j][k];
                const double ni = si[j][k] * rr[j][k] - sr[j][k] * ri[j][k];
                sr[j][k] = nr;
                si[j][k] = ni;
            }
    }
    for (int64_t s = 0;; ++s) {
#pragma unroll
        for (int j = 0; j < MG; ++j) {
            // X = a0 D0 + a1 D1 + conj(a1) D2
            const double xr = a0 * Dr[j][0] + (a1r * Dr[j][1] - a1i * Di[j][1]) + (a1r * Dr[j][2] + a1i * Di[j][2]);
            const double xi = a0 * Di[j][0] + (a1r * Di[j][1] + a1i * Dr[j][1]) + (a1r * Di[j][2] - a1i * Dr[j][2]);
            acc[j] += xr * xr + xi * xi;
        }
        if (s + 1 >= p.n_frame) break;
        const auto vo = amp[s * p.NQs];
        const auto vn = amp[(s + n) * p.NQs];
        const double or_ = ((double)vo.x - mr) * qf, oi = ((double)vo.y - mi) * qf;
        const double nr_ = ((double)vn.x - mr) * qf, ni = ((double)vn.y - mi) * qf;
#pragma unroll
        for (int j = 0; j < MG; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double ar = Dr[j][k] - or_, ai = Di[j][k] - oi;
                Dr[j][k] = (ar * rr[j][k] - ai * ri[j][k]) + (trr[j][k] * nr_ - tri[j][k] * ni);
                Di[j][k] = (ar * ri[j][k] + ai * rr[j][k]) + (trr[j][k] * ni + tri[j][k] * nr_);
            }
    }
#pragma unroll
    for (int j = 0; j < MG; ++j) {
        const int mi_ = grp * MG + j;
        if (mi_ < p.n_f) S[(int64_t)mi_ * p.NQ + q] = acc[j] * p.norm_s;
    }
}

// Per-Q X-ray weights of nsp species on the mode grid (CMCL order, k1
// fastest, k_d = i_d - N_d/2): w[s][Q] = c_s + sum_{k=1..4} a_sk exp(-b_sk |Q|^2 / (16 pi^2)),
// |Q|^2 = sum_d (k_d qs_d)^2 in inverse Angstrom (qs_d = 2 pi / box length
// along d); coef[s] = {a1..a4, b1..b4, c}
__global__ void k_xray_table(int64_t N1, int64_t N2, int64_t N3, double qs1, double qs2, double qs3, int nsp,
                             const double *__restrict__ coef, float *__restrict__ w)
{
    const int64_t NQ = N1 * N2 * N3;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= NQ) return;
    const int64_t i1 = q % N1, i2 = (q / N1) % N2, i3 = q / (N1 * N2);
    const double k1 = (double)(i1 - N1 / 2) * qs1, k2 = (double)(i2 - N2 / 2) * qs2, k3 = (double)(i3 - N3 / 2) * qs3;
    const double s2 = (k1 * k1 + k2 * k2 + k3 * k3) / (16.0 * M_PI * M_PI);
    for (int s = 0; s < nsp; ++s) {
        const double *c = coef + 9 * s;
        double f = c[8];
        for (int k = 0; k < 4; ++k) f += c[k] * exp(-c[4 + k] * s2);
        w[(int64_t)s * NQ + q] = (float)f;
    }
}

void launch_xray_table(const int64_t *N, const double *qs, int nsp, const double *coef, float *w, hipStream_t st)
{
    const int64_t NQ = N[0] * N[1] * N[2];
    if (NQ <= 0 || nsp <= 0) return;
    hipLaunchKernelGGL(k_xray_table, dim3((unsigned)((NQ + 255) / 256)), dim3(256), 0, st, N[0], N[1], N[2], qs[0],
                       qs[1], qs[2], nsp, coef, w);
}

// mp_sqom's per-Q tables of a Q plane, in the reference's REAL(4)
// arithmetic without contraction (mp_sqom56.f90:909-936): for pixel (i, j)
// (i fastest, i = k1 + n_qx/2 of the amplitudes' CMCL order)
//   q1 = (qx_min + i * dqx) e1,  q2 = (qy_min + j * dqy) e2   (r.l.u.)
//   q2_norm = 1 / sqrt(|q1 + q2|^2 + |q_v|^2)   (1 where that is below 1e-8)
//   x_ffq[s] = c_s + sum_k a_sk exp(-0.25 b_sk q_sq'),
//   q_sq' = 0.25 (|q1' + q2'|^2 + |q_vp|^2) with the 1/A vectors e1p, e2p, q_vp
__device__ __forceinline__ float dot3_nc(const float *a, const float *b)
{
#pragma clang fp contract(off)
    const float x = a[0] * b[0];
    const float y = a[1] * b[1];
    const float z = a[2] * b[2];
    return (x + y) + z;
}

__global__ void k_qtables(QTableParams p, float *__restrict__ q2_norm, float *__restrict__ x_ffq)
{
#pragma clang fp contract(off)
    const int64_t np = (int64_t)p.n_qx * p.n_qy;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const int i = (int)(q % p.n_qx), j = (int)(q / p.n_qx);
    const float ax = p.qx_min + (float)i * p.dqx;
    const float ay = p.qy_min + (float)j * p.dqy;
    if (q2_norm) {
        float v[3];
        for (int d = 0; d < 3; ++d) {
            const float q1 = ax * p.e1[d];
            const float q2 = ay * p.e2[d];
            v[d] = q1 + q2;
        }
        const float q_sq = dot3_nc(v, v) + dot3_nc(p.q_v, p.q_v);
        q2_norm[q] = (q_sq >= 1.0e-8f) ? 1.0f / sqrtf(q_sq) : 1.0f;
    }
    if (x_ffq) {
        float v[3];
        for (int d = 0; d < 3; ++d) {
            const float q1 = ax * p.e1p[d];
            const float q2 = ay * p.e2p[d];
            v[d] = q1 + q2;
        }
        const float q_sq = 0.25f * (dot3_nc(v, v) + dot3_nc(p.q_vp, p.q_vp));
        for (int s = 0; s < p.nsp; ++s) {
            float f = p.ff[s][8];
            for (int k = 0; k < 4; ++k) {
                const float e = expf(-0.25f * p.ff[s][2 * k + 1] * q_sq);
                const float t = p.ff[s][2 * k] * e;
                f = f + t;
            }
            x_ffq[(int64_t)s * np + q] = f;
        }
    }
}

void launch_qtables(const QTableParams &p, float *q2_norm, float *x_ffq, hipStream_t st)
{
    const int64_t np = (int64_t)p.n_qx * p.n_qy;
    if (np <= 0) return;
    hipLaunchKernelGGL(k_qtables, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, p, q2_norm, x_ffq);
}

// I(Q,t) (mp_sqom56.f90:1467-1471 map, :1504-1510 cut): thread = (output q,
// group of MG lags); C' = (C - <C>_t) * qfac as in the S(Q,w) kernels; the
// products and sums in FP64, the real part written times norm_s (1/nat^2,
// and 1/n_int for the map form)
template <typename R, int MG>
__global__ __launch_bounds__(256) void k_iqt(SqwParams p, const typename cplx<R>::T *__restrict__ amp,
                                             const float *__restrict__ qfac, double *__restrict__ S,
                                             double *__restrict__ Stot)
{
    const int64_t q = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    const int grp = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (q >= p.NQ || grp * MG >= p.n_f) return;
    const int64_t qs = src_q(p, q);
    if (qs < 0) {
        if (grp == 0 && Stot) Stot[q] = qnan();
        for (int j = 0; j < MG; ++j)
            if (grp * MG + j < p.n_f) S[(int64_t)(grp * MG + j) * p.NQ + q] = qnan();
        return;
    }
    amp += qs;
    double mr = 0.0, mi = 0.0, ss = 0.0;
    for (int64_t t = 0; t < p.T; ++t) {
        const auto v = amp[t * p.NQs];
        mr += (double)v.x;
        mi += (double)v.y;
        ss += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
    }
    if (grp == 0 && Stot) Stot[q] = ss * p.norm_tot;
    if (p.subtract_mean) {
        mr /= (double)p.T;
        mi /= (double)p.T;
    } else {
        mr = mi = 0.0;
    }
    const double qf = qfac ? (double)qfac[qs] : 1.0;
    double acc[MG];
    int64_t lag[MG];
#pragma unroll
    for (int j = 0; j < MG; ++j) {
        acc[j] = 0.0;
        lag[j] = p.m_lo + grp * MG + j;
    }
    for (int i = 0; i < p.n_int; ++i) {
        const auto a = amp[(int64_t)i * p.NQs];
        const double ar = ((double)a.x - mr) * qf, ai = ((double)a.y - mi) * qf;
#pragma unroll
        for (int j = 0; j < MG; ++j) {
            const int64_t t = i + lag[j];
            if (grp * MG + j < p.n_f && t >= 0 && t < p.T) {
                const auto b = amp[t * p.NQs];
                const double br = ((double)b.x - mr) * qf, bi = ((double)b.y - mi) * qf;
                acc[j] += br * ar + bi * ai;  // Re(b conj(a))
            }
        }
    }
#pragma unroll
    for (int j = 0; j < MG; ++j) {
        const int l = grp * MG + j;
        if (l < p.n_f) S[(int64_t)l * p.NQ + q] = acc[j] * p.norm_s;
    }
}

// I(Q,t) at many lags by FFT correlation (the sums of k_iqt, computed in
// O(T log T) per Q instead of O(n_f n_int)).  With x(t) = C'(t), t < T, and
// y(t) = C'(t) for t < n_int, 0 beyond, the circular correlation over T
// points is
//   sum_{s < n_int} x((s + l) mod T) conj(y(s)) = IFFT(X conj(Y))[l],
// which equals k_iqt's sum for 0 <= l <= T - n_int (s + l never wraps).
// The inverse transform is the forward one of the conjugate,
// IFFT(Z) = conj(FFT(conj(Z))) / T, and conj(X conj(Y)) = conj(X) Y, so
//   I(l) = Re FFT(conj(X) Y)[l] / T.
// A block holds `cols` Q columns of x and y in LDS (FP64 whatever the
// amplitudes' precision, like k_iqt's sums); mean and sum |C|^2 per column as
// in k_sqw_single.
template <typename R, int E, int NT>
__global__ __launch_bounds__(NT) void k_iqt_fft(SqwParams p, RadixPlan rp, int cols, FDiv dcols,
                                                const typename cplx<R>::T *__restrict__ amp,
                                                const double2 *__restrict__ tw, const float *__restrict__ qfac,
                                                double *__restrict__ S, double *__restrict__ Stot)
{
    using A = typename cplx<R>::T;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int T = (int)p.T;
    double2 *xt = reinterpret_cast<double2 *>(smem_raw);  // [T][cols]
    double2 *yt = xt + T * cols;                           // [T][cols]
    double2 *tw_l = yt + T * cols;                         // [T]
    double *red = reinterpret_cast<double *>(tw_l + T);    // [NT / 64][cols][3]
    __shared__ double mean_r[SQW_MAX_COLS], mean_i[SQW_MAX_COLS], qf_s[SQW_MAX_COLS];
    for (int i = threadIdx.x; i < T; i += NT) tw_l[i] = tw[i];
    const int64_t q0 = (int64_t)blockIdx.x * cols;
    const int ncol = (int)min((int64_t)cols, p.NQ - q0);
    // a thread's elements are column c = threadIdx.x % cols, frames seg,
    // seg + nseg, ... (NT is a multiple of cols)
    const int c = threadIdx.x % cols, seg = threadIdx.x / cols, nseg = NT / cols;
    const int64_t sc = c < ncol ? src_q(p, q0 + c) : -1;
    {
        A v[E];
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int t = seg + k * nseg;
            v[k] = (t < T && sc >= 0) ? amp[(int64_t)t * p.NQs + sc] : cmk<A>((R)0, (R)0);
        }
        double sr = 0.0, si = 0.0, ss = 0.0;
#pragma unroll
        for (int k = 0; k < E; ++k)
            if (seg + k * nseg < T) {
                const double2 x = make_double2((double)v[k].x, (double)v[k].y);
                xt[(seg + k * nseg) * cols + c] = x;
                sr += x.x;
                si += x.y;
                ss += x.x * x.x + x.y * x.y;
            }
        for (int o = cols; o < 64; o <<= 1) {
            sr += __shfl_xor(sr, o, 64);
            si += __shfl_xor(si, o, 64);
            ss += __shfl_xor(ss, o, 64);
        }
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        if (lane < cols) {
            red[(wv * cols + lane) * 3 + 0] = sr;
            red[(wv * cols + lane) * 3 + 1] = si;
            red[(wv * cols + lane) * 3 + 2] = ss;
        }
        __syncthreads();
        if (threadIdx.x < cols) {
            double a = 0.0, b = 0.0, q = 0.0;
            for (int k = 0; k < NT / 64; ++k) {
                a += red[(k * cols + threadIdx.x) * 3 + 0];
                b += red[(k * cols + threadIdx.x) * 3 + 1];
                q += red[(k * cols + threadIdx.x) * 3 + 2];
            }
            mean_r[threadIdx.x] = p.subtract_mean ? a / (double)T : 0.0;
            mean_i[threadIdx.x] = p.subtract_mean ? b / (double)T : 0.0;
            const int64_t s = threadIdx.x < ncol ? src_q(p, q0 + threadIdx.x) : -1;
            qf_s[threadIdx.x] = (qfac && s >= 0) ? (double)qfac[s] : 1.0;
            if (Stot && threadIdx.x < ncol) Stot[q0 + threadIdx.x] = s >= 0 ? q * p.norm_tot : qnan();
        }
        __syncthreads();
    }
    {
        // x = (C - mean) qf over all frames, y = x over the first n_int
        const double mr = mean_r[c], mi = mean_i[c], qf = qf_s[c];
        for (int t = seg; t < T; t += nseg) {
            const double2 v = xt[t * cols + c];
            const double2 x = make_double2((v.x - mr) * qf, (v.y - mi) * qf);
            xt[t * cols + c] = x;
            yt[t * cols + c] = t < p.n_int ? x : make_double2(0.0, 0.0);
        }
    }
    __syncthreads();
    lds_fft<double2, E>(xt, rp, cols, dcols, cols, 1, tw_l, true);
    lds_fft<double2, E>(yt, rp, cols, dcols, cols, 1, tw_l, true);
    for (int i = threadIdx.x; i < T * cols; i += NT) {
        const double2 a = xt[i], b = yt[i];
        xt[i] = make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);  // conj(a) b
    }
    __syncthreads();
    lds_fft<double2, E>(xt, rp, cols, dcols, cols, 1, tw_l, true);
    const double scale = p.norm_s / (double)T;
    for (int i = threadIdx.x; i < p.n_f * cols; i += NT) {
        const int l = (int)fdiv((uint32_t)i, dcols), c = i - l * cols;
        if (c >= ncol) continue;
        S[(int64_t)l * p.NQ + q0 + c] = src_q(p, q0 + c) >= 0 ? xt[(int)(p.m_lo + l) * cols + c].x * scale : qnan();
    }
}

int iqt_fft_tile_elems() { return SqwTiles<double>::E * SqwTiles<double>::NT; }

size_t iqt_fft_lds(int64_t T, int cols)
{
    return (size_t)T * (size_t)(2 * cols + 1) * sizeof(double2) +
           sizeof(double) * 3 * (SqwTiles<double>::NT / 64) * (size_t)cols;
}

template <typename R>
void launch_iqt_fft(const SqwParams &p, const RadixPlan &rp, int cols, const typename cplx<R>::T *amp,
                    const double2 *tw, const float *qfac, double *S, double *Stot, hipStream_t st)
{
    constexpr int E = SqwTiles<double>::E, NT = SqwTiles<double>::NT;
    const unsigned nb = (unsigned)((p.NQ + cols - 1) / cols);
    // set once per process (a function-local static initialises thread-safely)
    static const bool attr = [] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_iqt_fft<R, E, NT>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)SQW_MAX_LDS);
        return true;
    }();
    (void)attr;
    hipLaunchKernelGGL((k_iqt_fft<R, E, NT>), dim3(nb), dim3(NT), iqt_fft_lds(p.T, cols), st, p, rp, cols,
                       make_fdiv((uint32_t)cols), amp, tw, qfac, S, Stot);
}

template <typename R>
void launch_iqt(const SqwParams &p, const typename cplx<R>::T *amp, const float *qfac, double *S, double *Stot,
                hipStream_t st)
{
    constexpr int MG = 4;
    const unsigned gx = (unsigned)((p.NQ + 63) / 64);
    const unsigned gy = (unsigned)((p.n_f + 4 * MG - 1) / (4 * MG));
    hipLaunchKernelGGL((k_iqt<R, MG>), dim3(gx, gy > 0 ? gy : 1), dim3(256), 0, st, p, amp, qfac, S, Stot);
}

// Vibrational DOS species sums (mp_dos56.f90:519-525) from |FFT|^2 of
// packed velocity columns: column q of the chunk is complex column c0 + q of
// [n_tot][2] per frame, i.e. atom (c0 + q) / 2 with (v_x + i v_y) or
// (v_z + i 0).  For a pair of real signals packed as z = x + i y,
// |X(f)|^2 + |Y(f)|^2 = (|Z(f)|^2 + |Z(-f)|^2) / 2, which also holds for
// (v_z + i 0).  S rows hold m = -(n_freq - 1) .. n_freq - 1.
//   dos[f][s] += sum over the chunk's columns of atoms of type s (at_ind .w,
//                1-based) of (S(f) + S(-f)) / 2
// Block (f, s): threads sum the columns in a fixed stride order, then a
// fixed tree; the chunks are added in order, so the sums are deterministic.
__global__ __launch_bounds__(256) void k_dos_species(const double *__restrict__ S, int64_t ncols, int64_t c0,
                                                     int64_t n_freq, const int4 *__restrict__ at_ind, int nsp,
                                                     double *__restrict__ dos)
{
    __shared__ double red[256];
    const int64_t f = blockIdx.x;
    const int s = blockIdx.y;
    const double *Sp = S + (n_freq - 1 + f) * ncols, *Sm = S + (n_freq - 1 - f) * ncols;
    double acc = 0.0;
    for (int64_t q = threadIdx.x; q < ncols; q += 256)
        if (at_ind[(c0 + q) >> 1].w - 1 == s) acc += 0.5 * (Sp[q] + Sm[q]);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) dos[f * nsp + s] += red[0];
}

void launch_dos_species(const double *S, int64_t ncols, int64_t c0, int64_t n_freq, const int *at_ind, int nsp,
                        double *dos, hipStream_t st)
{
    hipLaunchKernelGGL(k_dos_species, dim3((unsigned)n_freq, (unsigned)nsp), dim3(256), 0, st, S, ncols, c0, n_freq,
                       reinterpret_cast<const int4 *>(at_ind), nsp, dos);
}

// The same spectra without the |FFT|^2 rows in HBM: a block loads `cols`
// packed velocity columns (first column c0 + block * cols), applies the
// window, runs the time FFT in LDS like k_sqw_single, and for every atom type
// s and frequency f writes
//   part[block][s][f] = sum over its columns of type s, in column order, of
//                       (|X(f)|^2 + |X(-f)|^2) / 2 * norm
// (FP64); dos_reduce_parts adds the blocks' rows in a fixed order.
template <typename R, int E, int NT>
__global__ __launch_bounds__(NT) void k_dos_fused(DosParams p, RadixPlan rp, int cols, FDiv dcols,
                                                  const float2 *__restrict__ vel, const int4 *__restrict__ at_ind,
                                                  const typename cplx<R>::T *__restrict__ tw,
                                                  const double *__restrict__ win, double *__restrict__ part)
{
    using C = typename cplx<R>::T;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int T = (int)p.T;
    C *tile = reinterpret_cast<C *>(smem_raw);  // [T][cols]
    C *tw_l = tile + T * cols;
    __shared__ int spc[SQW_MAX_COLS];
    for (int i = threadIdx.x; i < T; i += NT) tw_l[i] = tw[i];
    const int64_t q0 = p.c0 + (int64_t)blockIdx.x * cols;
    const int ncol = (int)min((int64_t)cols, p.c0 + p.nc - q0);
    if ((int)threadIdx.x < cols) spc[threadIdx.x] = (int)threadIdx.x < ncol ? at_ind[(q0 + threadIdx.x) >> 1].w - 1 : -1;
    {
        C v[E];
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int i = threadIdx.x + k * NT;
            const int t = (int)fdiv((uint32_t)i, dcols), c = i - t * cols;
            v[k] = cmk<C>((R)0, (R)0);
            if (i < T * cols && c < ncol) {
                const float2 a = vel[(int64_t)t * p.ncs + q0 + c];
                const double w = win[t];
                v[k] = cmk<C>((R)((double)a.x * w), (R)((double)a.y * w));
            }
        }
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int i = threadIdx.x + k * NT;
            if (i < T * cols) tile[i] = v[k];
        }
    }
    __syncthreads();
    lds_fft<C, E>(tile, rp, cols, dcols, cols, 1, tw_l, true);
    const int nf = (int)p.n_freq, nout = p.nsp * nf;
    double *dst = part + (int64_t)blockIdx.x * nout;
    for (int i = threadIdx.x; i < nout; i += NT) {
        const int s = i / nf, f = i - s * nf;
        const int fm = f == 0 ? 0 : T - f;
        double acc = 0.0;
        for (int c = 0; c < ncol; ++c)
            if (spc[c] == s) {
                const C a = tile[f * cols + c], b = tile[fm * cols + c];
                acc += 0.5 * (((double)a.x * (double)a.x + (double)a.y * (double)a.y) * p.norm +
                              ((double)b.x * (double)b.x + (double)b.y * (double)b.y) * p.norm);
            }
        dst[i] = acc;
    }
}

template <typename R> size_t dos_fused_lds(int64_t T, int cols)
{
    return (size_t)T * (size_t)(cols + 1) * sizeof(typename cplx<R>::T);
}

template <typename R>
void launch_dos_fused(const DosParams &p, const RadixPlan &rp, int cols, const float *vel, const int *at_ind,
                      const typename cplx<R>::T *tw, const double *win, double *part, hipStream_t st)
{
    constexpr int E = SqwTiles<R>::E, NT = SqwTiles<R>::NT;
    const unsigned nb = (unsigned)((p.nc + cols - 1) / cols);
    // set once per process (a function-local static initialises thread-safely)
    static const bool attr = [] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_dos_fused<R, E, NT>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)SQW_MAX_LDS);
        return true;
    }();
    (void)attr;
    hipLaunchKernelGGL((k_dos_fused<R, E, NT>), dim3(nb), dim3(NT), dos_fused_lds<R>(p.T, cols), st, p, rp, cols,
                       make_fdiv((uint32_t)cols), reinterpret_cast<const float2 *>(vel),
                       reinterpret_cast<const int4 *>(at_ind), tw, win, part);
}

// part2[g][i] = sum over the rows r = g * per .. of part[r][i], in order
__global__ void k_dos_groups(const double *__restrict__ part, int64_t nrows, int64_t per, int nout,
                             double *__restrict__ part2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nout) return;
    const int64_t r0 = (int64_t)blockIdx.y * per, r1 = min(r0 + per, nrows);
    double s = 0.0;
    for (int64_t r = r0; r < r1; ++r) s += part[r * nout + i];
    part2[(int64_t)blockIdx.y * nout + i] = s;
}

// dos[f][s] += sum over g of part2[g][s * n_freq + f], in order
__global__ void k_dos_final(const double *__restrict__ part2, int ng, int nsp, int n_freq, double *__restrict__ dos)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsp * n_freq) return;
    double s = 0.0;
    for (int g = 0; g < ng; ++g) s += part2[(int64_t)g * nsp * n_freq + i];
    const int sp = i / n_freq, f = i - sp * n_freq;
    dos[(int64_t)f * nsp + sp] += s;
}

int dos_reduce_groups() { return 64; }

void dos_reduce_parts(const double *part, int64_t nrows, int nsp, int n_freq, double *part2, double *dos,
                      hipStream_t st)
{
    const int nout = nsp * n_freq, G = dos_reduce_groups();
    const int64_t per = (nrows + G - 1) / G;
    const int ng = (int)((nrows + per - 1) / per);
    hipLaunchKernelGGL(k_dos_groups, dim3((unsigned)((nout + 255) / 256), (unsigned)ng), dim3(256), 0, st, part, nrows,
                       per, nout, part2);
    hipLaunchKernelGGL(k_dos_final, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, st, part2, ng, nsp, n_freq,
                       dos);
}

// complex128 copy of complex columns [c0, c0 + nc) of float [T][ncs] pairs
__global__ void k_cols_to_f64(const float2 *__restrict__ in, int64_t T, int64_t ncs, int64_t c0, int64_t nc,
                              double2 *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T * nc) return;
    const int64_t t = i / nc, q = i - t * nc;
    const float2 v = in[t * ncs + c0 + q];
    out[i] = make_double2((double)v.x, (double)v.y);
}

void launch_cols_to_f64(const float *in, int64_t T, int64_t ncs, int64_t c0, int64_t nc, double *out, hipStream_t st)
{
    if (T * nc == 0) return;
    hipLaunchKernelGGL(k_cols_to_f64, dim3((unsigned)((T * nc + 255) / 256)), dim3(256), 0, st,
                       reinterpret_cast<const float2 *>(in), T, ncs, c0, nc, reinterpret_cast<double2 *>(out));
}


template <typename R> int sqw_tile_elems() { return SqwTiles<R>::E * SqwTiles<R>::NT; }

template <typename R> size_t sqw_single_lds(int64_t T, int cols)
{
    return (size_t)T * (size_t)(cols + 1) * sizeof(typename cplx<R>::T) +
           sizeof(double) * 3 * (SqwTiles<R>::NT / 64) * (size_t)cols;
}

template <typename R>
void launch_sqw_single(const SqwParams &p, const RadixPlan &rp, int cols, const typename cplx<R>::T *amp,
                       const typename cplx<R>::T *tw, const double *win, const float *qfac, double *S, double *Stot,
                       hipStream_t st)
{
    constexpr int E = SqwTiles<R>::E, NT = SqwTiles<R>::NT;
    const unsigned nb = (unsigned)((p.NQ + cols - 1) / cols);
    const size_t lds = sqw_single_lds<R>(p.T, cols);  // <= SQW_MAX_LDS (the caller's choice of cols)
    // set once per process (a function-local static initialises thread-safely)
    static const bool attr = [] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_sqw_single<R, E, NT>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)SQW_MAX_LDS);
        return true;
    }();
    (void)attr;
    hipLaunchKernelGGL((k_sqw_single<R, E, NT>), dim3(nb), dim3(NT), lds, st, p, rp, cols, make_fdiv((uint32_t)cols),
                       make_fdiv((uint32_t)p.T), amp, tw, win, qfac, S, Stot);
}

template <typename R>
void launch_sqw_sliding(const SqwParams &p, const typename cplx<R>::T *amp, const float *qfac, double *S,
                        double *Stot, hipStream_t st)
{
    // one frequency per thread, at most 64 registers (8 waves per SIMD): the
    // recurrences of a frequency hold 36 FP64 values, so more frequencies per
    // thread cost occupancy (a 2.1 M mode map at 17 frequencies, 256
    // windows: 137.7 ms with 4 per thread, 87.1 with 2, 77.0 with 1 at 116
    // registers, 50.5 with 1 at 64)
    constexpr int MG = 1;
    const unsigned gx = (unsigned)((p.NQ + 63) / 64);
    const unsigned gy = (unsigned)((p.n_f + 4 * MG - 1) / (4 * MG));
    hipLaunchKernelGGL((k_sqw_sliding<R, MG>), dim3(gx, gy > 0 ? gy : 1), dim3(256), 0, st, p, amp, qfac, S, Stot);
}

#define MPSQ_SQW_INST(R)                                                                                              \
    template int sqw_tile_elems<R>();                                                                                 \
    template size_t sqw_single_lds<R>(int64_t, int);                                                                  \
    template void launch_sqw_single<R>(const SqwParams &, const RadixPlan &, int, const cplx<R>::T *,                \
                                       const cplx<R>::T *, const double *, const float *, double *, double *,        \
                                       hipStream_t);                                                                  \
    template void launch_sqw_sliding<R>(const SqwParams &, const cplx<R>::T *, const float *, double *, double *,    \
                                        hipStream_t);                                                                 \
    template void launch_sqw_window1<R>(const SqwParams &, const cplx<R>::T *, const float *, const double *,        \
                                        double *, double *, hipStream_t);                                             \
    template void launch_iqt<R>(const SqwParams &, const cplx<R>::T *, const float *, double *, double *, hipStream_t); \
    template void launch_iqt_fft<R>(const SqwParams &, const RadixPlan &, int, const cplx<R>::T *, const double2 *,  \
                                    const float *, double *, double *, hipStream_t);                                 \
    template size_t dos_fused_lds<R>(int64_t, int);                                                                   \
    template void launch_dos_fused<R>(const DosParams &, const RadixPlan &, int, const float *, const int *,          \
                                      const cplx<R>::T *, const double *, double *, hipStream_t);

MPSQ_SQW_INST(float)
MPSQ_SQW_INST(double)

}  // namespace mpsq
