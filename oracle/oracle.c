/*
 * oracle.c -- CPU FP64 restatement of MP_tools' S(Q,w) hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * mp_tools_amd/: it may be built and called from tests/, from
 * __graft_entry__.smoke() and from the cpu_baseline leg of bench.py, and from
 * nowhere else.  The product path never links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the jkulda/MP_tools tree).  The non-uniform FFT itself
 * lives in FINUFFT, which is not part of the reference tree; the oracle
 * therefore restates the *definition* that the reference's own brute-force
 * transform `dirft2d1` computes (code/src/mp_sqom56.f90:2530-2602) and the
 * FINUFFT method as published (code/contrib/barnett_FINUFFT.txt:38-40).
 *
 * How each function is pinned (tests/test_oracle.py):
 *   orc_dirft_type1         uniform points reduce to numpy.fft; a single point
 *                           at the origin; the shift theorem; a perfect
 *                           lattice gives the closed-form Dirichlet pattern;
 *                           linearity; 3D == stack of 2D slices with phase.
 *   orc_dirft_type1_sampled equals orc_dirft_type1 at the sampled modes and
 *                           a numpy brute force on tiny inputs.
 *   orc_dirft_type1_prod    equals orc_dirft_type1_sampled on every mode of
 *                           the product set, and orc_dirft_type1 where the
 *                           set is the whole mode box.
 *   orc_qtables             closed forms on an axis-aligned plane (1/|Q|,
 *                           the Q = 0 pixel, the vertical component), the
 *                           form factors at Q = 0 against the atomic numbers
 *                           of the reference's table, single-Gaussian decay.
 *   orc_fold / orc_counting_sort  closed-form fold checked in numpy (floor /
 *                           ceil / modulo) and numpy's stable argsort.
 *   orc_stage               numpy restatement of the dot products + the
 *                           single-point phase closed form.
 *   orc_select_cell/bulk    known answers on the Al/U frame fixture and on
 *                           hand-built vacancy/window cases.
 *   orc_sqw_direct          T=1 closed form, constant amplitudes -> S==0,
 *                           pure-tone closed form (peak at m0), numpy FFT of
 *                           the windowed series, Parseval.
 *   orc_iqt                 constant amplitudes -> 0 (mean subtraction); a
 *                           pure tone a e^{-i w t} -> n_int |a|^2 cos(w lag)
 *                           (scaled: |a|^2 cos(w lag)); lag 0 -> sum of |C'|^2;
 *                           equals numpy's correlation of each window.
 *   orc_oneph               numpy's fft2 for the column DFT, a single
 *                           displaced atom (closed form), a plane-wave
 *                           displacement pattern (peaks at its wave
 *                           vector), vacancies and the (hhl) column map.
 *   orc_dos                 numpy's FFT of the windowed series, a constant
 *                           velocity (f = 0 closed form), Parseval over the
 *                           half spectrum of real signals.
 *   orc_philox4x32_10       the published known-answer vectors of the
 *                           Random123 suite (zero, all-ones and pi-digit
 *                           counter/key).
 *   orc_sql_hist_mc         constant amplitudes: every kept sample adds
 *                           exactly Re(c_i c_j*), so bins hold integer
 *                           counts that with the skips add up to n_h n_mc;
 *                           the radius law (counts flat in |Q|); isotropy
 *                           (direction moments <cos^2> = 1/3, <cos^4> = 1/5
 *                           of the unit sphere, which cube corners would
 *                           break); the normalised estimate of a smooth
 *                           |A(Q)|^2 against its value and against the
 *                           Gauss sum over the same grid.
 *   orc_speckle_pass        known answers: a lone spike / dip takes its
 *                           neighbours' max / min, s_trig 1 flattens, and a
 *                           hand-worked vector whose result depends on the
 *                           sequential order of the pass.
 *   orc_es_spread           end-to-end NUFFT against orc_dirft_type1 at eps.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_TWO_PI (2.0 * M_PI)

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* a4 / a5: type-1 direct DFT.                                               */
/* F(k1,k2,k3) = sum_j c_j exp(s*i*(k1 x_j + k2 y_j + k3 z_j)), s = +1 if     */
/* iflag >= 0 else -1, k_d in [-N_d/2, (N_d-1)/2], output in CMCL order      */
/* (k1 fastest).  Same definition as dirft2d1 (mp_sqom56.f90:2539-2546,       */
/* sign rule :2549-2550, index range :2535); the 3D form is the stack of 2D   */
/* slices with strengths c*exp(-i k3 z) that mp_sql builds                   */
/* (mp_sql56.f90:1110-1132).  Like dirft2d1 it walks k1 with a per-point      */
/* twiddle recurrence (mp_sqom56.f90:2564-2571); the (k2,k3) row factor is    */
/* evaluated directly for accuracy.                                          */
/* c and f are interleaved complex (re, im).                                 */
/* ------------------------------------------------------------------------ */
int orc_dirft_type1(int dim, int64_t M, const double *x, const double *y,
                    const double *z, const double *c, int iflag,
                    int64_t N1, int64_t N2, int64_t N3, double *f)
{
    if (dim < 1 || dim > 3 || M < 0 || N1 < 1 || N2 < 1 || N3 < 1) return 2;
    if (dim < 2) N2 = 1;
    if (dim < 3) N3 = 1;
    const double s = (iflag >= 0) ? 1.0 : -1.0;
    const int64_t nrow = N2 * N3;
    const int64_t k1min = -(N1 / 2);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t r = 0; r < nrow; ++r) {
        const int64_t k2 = (r % N2) - N2 / 2;
        const int64_t k3 = (r / N2) - N3 / 2;
        double *row = f + 2 * r * N1;
        for (int64_t i = 0; i < 2 * N1; ++i) row[i] = 0.0;
        for (int64_t j = 0; j < M; ++j) {
            double ph = 0.0;
            if (dim >= 2) ph += (double)k2 * y[j];
            if (dim >= 3) ph += (double)k3 * z[j];
            ph += (double)k1min * x[j];
            const double er = cos(s * ph), ei = sin(s * ph);
            double cr = c[2 * j] * er - c[2 * j + 1] * ei;
            double ci = c[2 * j] * ei + c[2 * j + 1] * er;
            const double zr = cos(s * x[j]), zi = sin(s * x[j]);
            for (int64_t k = 0; k < N1; ++k) {
                row[2 * k] += cr;
                row[2 * k + 1] += ci;
                const double tr = cr * zr - ci * zi;
                ci = cr * zi + ci * zr;
                cr = tr;
            }
        }
    }
    return 0;
}

/* Direct sum at K explicit integer modes (k1[m], k2[m], k3[m]).  Used at the  */
/* full sizes where the dense direct DFT would take too long.  Points are     */
/* processed in blocks; per block the 1D exponentials exp(s i k x) are        */
/* tabulated with a recurrence re-anchored by an exact sincos every 16 steps, */
/* so each (point, mode) term costs two complex products.                     */
int orc_dirft_type1_sampled(int dim, int64_t M, const double *x,
                            const double *y, const double *z, const double *c,
                            int iflag, int64_t K, const int64_t *k1,
                            const int64_t *k2, const int64_t *k3, double *f)
{
    if (dim < 1 || dim > 3 || M < 0 || K < 0) return 2;
    const double s = (iflag >= 0) ? 1.0 : -1.0;
    int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int64_t m = 0; m < K; ++m) {
        const int64_t kk[3] = {k1[m], dim >= 2 ? k2[m] : 0, dim >= 3 ? k3[m] : 0};
        for (int d = 0; d < 3; ++d) {
            if (m == 0 || kk[d] < lo[d]) lo[d] = kk[d];
            if (m == 0 || kk[d] > hi[d]) hi[d] = kk[d];
        }
    }
    for (int64_t i = 0; i < 2 * K; ++i) f[i] = 0.0;
    if (K == 0 || M == 0) return 0;
    const int64_t span[3] = {hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
    if (4 * K < span[0] + span[1] + span[2]) {
        /* few modes: one exact sincos per (point, mode) beats the tables */
#pragma omp parallel
        {
            double *acc = (double *)calloc((size_t)(2 * K), sizeof(double));
#pragma omp for schedule(static)
            for (int64_t j = 0; j < M; ++j) {
                const double xv = x[j], yv = dim >= 2 ? y[j] : 0.0, zv = dim >= 3 ? z[j] : 0.0;
                const double cr = c[2 * j], ci = c[2 * j + 1];
                for (int64_t m = 0; m < K; ++m) {
                    double ph = (double)k1[m] * xv;
                    if (dim >= 2) ph += (double)k2[m] * yv;
                    if (dim >= 3) ph += (double)k3[m] * zv;
                    const double er = cos(s * ph), ei = sin(s * ph);
                    acc[2 * m] += cr * er - ci * ei;
                    acc[2 * m + 1] += cr * ei + ci * er;
                }
            }
#pragma omp critical
            for (int64_t i = 0; i < 2 * K; ++i) f[i] += acc[i];
            free(acc);
        }
        return 0;
    }
    const int64_t B = 256;
    const int64_t nblk = (M + B - 1) / B;
#pragma omp parallel
    {
        double *acc = (double *)calloc((size_t)(2 * K), sizeof(double));
        double *tab[3];
        for (int d = 0; d < 3; ++d)
            tab[d] = (double *)malloc((size_t)(2 * span[d] * B) * sizeof(double));
#pragma omp for schedule(dynamic, 1)
        for (int64_t b = 0; b < nblk; ++b) {
            const int64_t j0 = b * B;
            const int64_t nb = (j0 + B <= M) ? B : (M - j0);
            const double *co[3] = {x, y, z};
            for (int d = 0; d < 3; ++d) {
                for (int64_t jj = 0; jj < nb; ++jj) {
                    const double xv = (d < dim) ? co[d][j0 + jj] : 0.0;
                    const double zr = cos(s * xv), zi = sin(s * xv);
                    double er = 0.0, ei = 0.0;
                    for (int64_t q = 0; q < span[d]; ++q) {
                        if ((q & 15) == 0) {
                            const double ph = s * (double)(lo[d] + q) * xv;
                            er = cos(ph);
                            ei = sin(ph);
                        }
                        tab[d][2 * (q * B + jj)] = er;
                        tab[d][2 * (q * B + jj) + 1] = ei;
                        const double tr = er * zr - ei * zi;
                        ei = er * zi + ei * zr;
                        er = tr;
                    }
                }
            }
            for (int64_t m = 0; m < K; ++m) {
                const double *t1 = tab[0] + 2 * (k1[m] - lo[0]) * B;
                const double *t2 = tab[1] + 2 * ((dim >= 2 ? k2[m] : 0) - lo[1]) * B;
                const double *t3 = tab[2] + 2 * ((dim >= 3 ? k3[m] : 0) - lo[2]) * B;
                double sr = 0.0, si = 0.0;
                for (int64_t jj = 0; jj < nb; ++jj) {
                    const double ar = t1[2 * jj] * t2[2 * jj] - t1[2 * jj + 1] * t2[2 * jj + 1];
                    const double ai = t1[2 * jj] * t2[2 * jj + 1] + t1[2 * jj + 1] * t2[2 * jj];
                    const double br = ar * t3[2 * jj] - ai * t3[2 * jj + 1];
                    const double bi = ar * t3[2 * jj + 1] + ai * t3[2 * jj];
                    const double cr = c[2 * (j0 + jj)], ci = c[2 * (j0 + jj) + 1];
                    sr += cr * br - ci * bi;
                    si += cr * bi + ci * br;
                }
                acc[2 * m] += sr;
                acc[2 * m + 1] += si;
            }
        }
#pragma omp critical
        for (int64_t i = 0; i < 2 * K; ++i) f[i] += acc[i];
        free(acc);
        for (int d = 0; d < 3; ++d) free(tab[d]);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* a8: mp_sqom's per-Q tables of a Q plane (mp_sqom56.f90:909-936), REAL(4)  */
/* arithmetic without contraction, pixel (i, j) with i fastest:              */
/*   q1 = (qx_min + i*(bz_nx/n_qx))*e1, q2 = (qy_min + j*(bz_ny/n_qy))*e2    */
/*   q_sq = dot(q1+q2, q1+q2) + dot(q_v, q_v)                                */
/*   q2_norm = 1/sqrt(q_sq) if q_sq >= 1e-8 else 1                    (:914) */
/*   q_sq' = .25*(dot(q1'+q2', q1'+q2') + dot(q_vp, q_vp)) on e1p, e2p (:925)*/
/*   x_ffq(s) = ff(9) + sum_k ff(2k-1)*exp(-.25*ff(2k)*q_sq')     (:929-931) */
/* ff: [nsp][9] as read from the .par file (a1,b1,a2,b2,a3,b3,a4,b4,c).     */
/* ------------------------------------------------------------------------ */
static float orc_dot3f(const float *a, const float *b)
{
    const float x = a[0] * b[0];
    const float y = a[1] * b[1];
    const float z = a[2] * b[2];
    return (x + y) + z;
}

void orc_qtables(int n_qx, int n_qy, float qx_min, float qy_min, float bz_nx, float bz_ny,
                 const float *e1, const float *e2, const float *q_v, const float *e1p,
                 const float *e2p, const float *q_vp, int nsp, const float *ff,
                 float *q2_norm, float *x_ffq)
{
    const float dqx = bz_nx / (float)n_qx, dqy = bz_ny / (float)n_qy;
    const int64_t np = (int64_t)n_qx * n_qy;
    for (int j = 0; j < n_qy; ++j)
        for (int i = 0; i < n_qx; ++i) {
            const int64_t q = (int64_t)j * n_qx + i;
            const float ax = qx_min + (float)i * dqx;
            const float ay = qy_min + (float)j * dqy;
            float v[3];
            for (int d = 0; d < 3; ++d) {
                const float a = ax * e1[d];
                const float b = ay * e2[d];
                v[d] = a + b;
            }
            const float q_sq = orc_dot3f(v, v) + orc_dot3f(q_v, q_v);
            if (q2_norm) q2_norm[q] = (q_sq >= 1.0e-8f) ? 1.0f / sqrtf(q_sq) : 1.0f;
            if (!x_ffq) continue;
            for (int d = 0; d < 3; ++d) {
                const float a = ax * e1p[d];
                const float b = ay * e2p[d];
                v[d] = a + b;
            }
            const float qs = 0.25f * (orc_dot3f(v, v) + orc_dot3f(q_vp, q_vp));
            for (int s = 0; s < nsp; ++s) {
                const float *c = ff + 9 * s;
                float f = c[8];
                for (int k = 0; k < 4; ++k) {
                    const float e = expf(-0.25f * c[2 * k + 1] * qs);
                    const float t = c[2 * k] * e;
                    f = f + t;
                }
                x_ffq[(int64_t)s * np + q] = f;
            }
        }
}

/* Direct sum on a product set of modes: F[a3][a2][a1] at (k1s[a1], k2s[a2], */
/* k3s[a3]) (a1 fastest), the same definition as orc_dirft_type1.  Per point  */
/* the n1 + n2 + n3 one-dimensional exponentials are exact sincos values and  */
/* each mode costs one complex multiply-add, so a few hundred to ten thousand */
/* seeded random modes over millions of points stay cheap.  Threads sum       */
/* private accumulators over static point ranges, added in thread order.     */
int orc_dirft_type1_prod(int dim, int64_t M, const double *x, const double *y,
                         const double *z, const double *c, int iflag,
                         int64_t n1, const int64_t *k1s, int64_t n2,
                         const int64_t *k2s, int64_t n3, const int64_t *k3s,
                         double *f)
{
    if (dim < 1 || dim > 3 || M < 0 || n1 < 1 || n2 < 1 || n3 < 1) return 2;
    if ((dim < 2 && n2 != 1) || (dim < 3 && n3 != 1)) return 2;
    const double s = (iflag >= 0) ? 1.0 : -1.0;
    const int64_t K = n1 * n2 * n3;
    for (int64_t i = 0; i < 2 * K; ++i) f[i] = 0.0;
    const int nth = orc_num_threads();
    double *part = (double *)calloc((size_t)(2 * K) * (size_t)nth, sizeof(double));
    if (!part) return 2;
#pragma omp parallel num_threads(nth)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *ar = part + (size_t)(2 * K) * (size_t)tid, *ai = ar + K;
        double *e1r = (double *)malloc(sizeof(double) * (size_t)(2 * n1));
        double *e1i = e1r + n1;
        double e2r[4096], e2i[4096], e3r[4096], e3i[4096];
        const int64_t lo = M * tid / nth, hi = M * (tid + 1) / nth;
        for (int64_t j = lo; j < hi; ++j) {
            for (int64_t a = 0; a < n1; ++a) {
                const double ph = s * (double)k1s[a] * x[j];
                e1r[a] = cos(ph);
                e1i[a] = sin(ph);
            }
            for (int64_t a = 0; a < n2 && a < 4096; ++a) {
                const double ph = dim >= 2 ? s * (double)k2s[a] * y[j] : 0.0;
                e2r[a] = cos(ph);
                e2i[a] = sin(ph);
            }
            for (int64_t a = 0; a < n3 && a < 4096; ++a) {
                const double ph = dim >= 3 ? s * (double)k3s[a] * z[j] : 0.0;
                e3r[a] = cos(ph);
                e3i[a] = sin(ph);
            }
            const double cr = c[2 * j], ci = c[2 * j + 1];
            for (int64_t a3 = 0; a3 < n3; ++a3) {
                const double ur = cr * e3r[a3] - ci * e3i[a3], ui = cr * e3i[a3] + ci * e3r[a3];
                for (int64_t a2 = 0; a2 < n2; ++a2) {
                    const double tr = ur * e2r[a2] - ui * e2i[a2], ti = ur * e2i[a2] + ui * e2r[a2];
                    double *rr = ar + (a3 * n2 + a2) * n1, *ri = ai + (a3 * n2 + a2) * n1;
                    for (int64_t a = 0; a < n1; ++a) {
                        rr[a] += tr * e1r[a] - ti * e1i[a];
                        ri[a] += tr * e1i[a] + ti * e1r[a];
                    }
                }
            }
        }
        free(e1r);
    }
    for (int t = 0; t < nth; ++t) {
        const double *ar = part + (size_t)(2 * K) * (size_t)t, *ai = ar + K;
        for (int64_t m = 0; m < K; ++m) {
            f[2 * m] += ar[m];
            f[2 * m + 1] += ai[m];
        }
    }
    free(part);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* a4.1: FP64 periodic fold and bin key (the build's own definition; the     */
/* FINUFFT internals it replaces are not in the reference tree).             */
/* For each used dimension d: u = x * (n_d / 2pi); the kernel's first cell is */
/* l = ceil(u - w/2); its offset d = l - u lies in [-w/2, -w/2+1); l is then  */
/* wrapped into [0, n_d).  Points outside [-pi, pi) fold periodically.        */
/* Bin of a point: (l_y / by, l_z / bz), key = (l_z/bz) * (n2/by) + l_y/by;   */
/* in 1D every point is in bin 0.  The operation order below is fixed (no     */
/* fused multiply-add; built with -ffp-contract=off) so that the GPU fold     */
/* reproduces l, d and key bit for bit.                                       */
/* ------------------------------------------------------------------------ */
int orc_fold(int dim, int64_t M, const double *x, const double *y,
             const double *z, const int64_t *n, int w, int by, int bz,
             int32_t *l_out, double *d_out, int32_t *key_out)
{
    if (dim < 1 || dim > 3 || w < 2 || by < 1 || bz < 1) return 2;
    const double *co[3] = {x, y, z};
    const double hw = 0.5 * (double)w;
    double sc[3];
    for (int d = 0; d < 3; ++d) sc[d] = (double)n[d] / ORC_TWO_PI;
    const int64_t nby = (dim >= 2) ? n[1] / by : 1;
    int bad = 0;
#pragma omp parallel for reduction(| : bad)
    for (int64_t j = 0; j < M; ++j) {
        int32_t li[3] = {0, 0, 0};
        for (int d = 0; d < 3; ++d) {
            double dd = 0.0;
            if (d < dim) {
                const double xv = co[d][j];
                if (!isfinite(xv)) { bad = 1; }
                const double u = xv * sc[d];
                const double t = u - hw;
                const double l = ceil(t);
                dd = l - u;
                int64_t lw = (isfinite(l) && fabs(l) < 9.0e18) ? (int64_t)l : 0;
                lw %= n[d];
                if (lw < 0) lw += n[d];
                li[d] = (int32_t)lw;
            }
            l_out[3 * j + d] = li[d];
            d_out[3 * j + d] = dd;
        }
        int32_t key = 0;
        if (dim == 2) key = li[1] / by;
        if (dim == 3) key = (int32_t)((li[2] / bz) * nby + li[1] / by);
        key_out[j] = key;
    }
    return bad ? 3 : 0;
}

/* Stable counting sort: perm[k] = index of the k-th point in (key, index)   */
/* order.  offsets[b] = first sorted slot of key b, offsets[nkeys] = M.       */
/* Restates the stable species counting sort of mp_lbin56.f90:1336-1347 and   */
/* is the definition of the bin-sort permutation (a4.1).                     */
int orc_counting_sort(int64_t M, const int32_t *key, int32_t nkeys,
                      int32_t *perm, int64_t *offsets)
{
    if (nkeys < 1) return 2;
    int64_t *cnt = (int64_t *)calloc((size_t)nkeys + 1, sizeof(int64_t));
    if (!cnt) return 2;
    for (int64_t j = 0; j < M; ++j) {
        if (key[j] < 0 || key[j] >= nkeys) { free(cnt); return 2; }
        cnt[key[j] + 1]++;
    }
    for (int32_t b = 0; b < nkeys; ++b) cnt[b + 1] += cnt[b];
    if (offsets)
        for (int32_t b = 0; b <= nkeys; ++b) offsets[b] = cnt[b];
    for (int64_t j = 0; j < M; ++j) perm[cnt[key[j]]++] = (int32_t)j;
    free(cnt);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* a2: point selection + compaction (order preserved).                       */
/* CELL (mp_sqom56.f90:1104-1113): keep site i iff its position is not the   */
/* vacancy marker (0,0,0) and every cell index lies in the real-space window */
/* (ind[d] - n_row[d]/2 - 1) in [pmin[d], pmax[d]].                          */
/* ------------------------------------------------------------------------ */
int64_t orc_select_cell(int64_t nsuper, const float *pos, const int32_t *ind,
                        const int32_t *n_row, const float *pmin,
                        const float *pmax, int32_t *sel)
{
    int64_t k = 0;
    for (int64_t i = 0; i < nsuper; ++i) {
        const float *p = pos + 4 * i;
        const int nz = (p[0] != 0.0f) || (p[1] != 0.0f) || (p[2] != 0.0f);
        int inside = 1;
        for (int d = 0; d < 3; ++d) {
            const float v = (float)(ind[4 * i + d] - n_row[d] / 2 - 1);
            inside = inside && (v >= pmin[d]) && (v <= pmax[d]);
        }
        if (nz && inside) sel[k++] = (int32_t)i;
    }
    return k;
}

/* BULK (mp_sqom56.f90:1121-1132): keep atom i iff pmin-m <= r < pmax+m in    */
/* every direction, with margin m = eps_x (0.1, :214) for a single box        */
/* (nsuper==1) else 1.  The bounds are formed in FP32 as in the reference.    */
int64_t orc_select_bulk(int64_t n, const float *pos, const float *pmin,
                        const float *pmax, float margin, int32_t *sel)
{
    float lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = pmin[d] - margin;
        hi[d] = pmax[d] + margin;
    }
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i) {
        const float *p = pos + 4 * i;
        int inside = 1;
        for (int d = 0; d < 3; ++d) inside = inside && (p[d] >= lo[d]) && (p[d] < hi[d]);
        if (inside) sel[k++] = (int32_t)i;
    }
    return k;
}

/* ------------------------------------------------------------------------ */
/* a3: staging of NUFFT coordinates and strengths (mp_sqom56.f90:1115-1118,  */
/* 1134-1137, 809-811, 1167).  For every selected atom with FP32 position r: */
/*   x_d = ((p0*e_d0 + p1*e_d1) + p2*e_d2) * tpe_d,  p = r - centre (FP64)    */
/*   c   = weight * exp(-i * 2pi * (q0 . r))                                 */
/*   c  *= prod_d (1 + cos x_d) / 2   if window (the reference's             */
/*         (1+cos x)(1+cos y) with its 0.25 norm, :1118,:1167)               */
/* axes e is 3x3 (row d = plane vector d, already normalised).               */
/* ------------------------------------------------------------------------ */
void orc_stage(int dim, int64_t M, const float *pos, const int32_t *sel,
               const double *centre, const double *e, const double *tpe,
               const double *q0, int window, double weight, double *x,
               double *y, double *z, double *c)
{
    double *out[3] = {x, y, z};
#pragma omp parallel for
    for (int64_t j = 0; j < M; ++j) {
        const int64_t i = sel ? sel[j] : j;
        const double r0 = (double)pos[4 * i], r1 = (double)pos[4 * i + 1], r2 = (double)pos[4 * i + 2];
        const double p0 = r0 - centre[0], p1 = r1 - centre[1], p2 = r2 - centre[2];
        double xv[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) {
            const double a = p0 * e[3 * d];
            const double b = p1 * e[3 * d + 1];
            const double cc = p2 * e[3 * d + 2];
            const double dot = (a + b) + cc;
            xv[d] = dot * tpe[d];
            out[d][j] = xv[d];
        }
        const double ph = ORC_TWO_PI * ((q0[0] * r0 + q0[1] * r1) + q0[2] * r2);
        double cr = weight * cos(ph), ci = -weight * sin(ph);
        if (window) {
            double wf = 1.0;
            for (int d = 0; d < dim; ++d) wf *= 0.5 * (1.0 + cos(xv[d]));
            cr *= wf;
            ci *= wf;
        }
        c[2 * j] = cr;
        c[2 * j + 1] = ci;
    }
}

/* ------------------------------------------------------------------------ */
/* a7-a14: the S(Q) / S(Q,w) reduction, FP64, by direct windowed DFT.        */
/* F   : per-species amplitudes [T][nsp][NQ] interleaved complex.            */
/* wsc : per-species scalar weights [nsp] (used when wtab == NULL)          */
/* wtab: per-species per-Q weights [nsp][NQ] or NULL                        */
/* mask: [nsp] 0/1                                                          */
/* qfac: per-Q real factor applied to the E-resolved amplitude or NULL      */
/*       (q2_norm when j_qsq != 1, mp_sqom56.f90:1376, 901-917)              */
/* Outputs:                                                                 */
/*   stot[Q] = sum_t |C(t,Q)|^2 / (T * nat^2)          (:1381-1387, :1574)   */
/*      with C = sum_j m_j w_j F_j                      (:1338-1356)         */
/*   s[mi][Q], m = m_lo + mi, mi < n_f:                                     */
/*      C'(t,Q) = [sum_j m_j w_j (F_j - M_j)] * qfac    (:1358-1377)         */
/*      M_j(Q) = (1/T) sum_t F_j(t,Q)                   (:1177-1185)         */
/*      X_s(m) = sum_i win_i exp(-2 pi i m i / n_int) C'(s+i)  (:1437-1466;  */
/*               :1480-1501 computes the same |X| by FFT)                    */
/*      s = dt / (nat^2 * 0.5 * n_frame * n_int) * sum_s |X_s|^2  (:1527,    */
/*          :1576), n_frame = T - n_int + 1                                  */
/*   window: kind 0 = 0.5(1 - cos(2 pi i/(n_int-1))), i = 0..n_int-1 (:941-948) */
/*           kind 1 = 0.5(1 - cos(2 pi (i+1)/(n_int+1)))  (:2429-2431)       */
/*           n_int <= 1 -> window 1                                          */
/* subtract_mean = 0 skips the M_j term (the pair mode raw sums, :1460-1463) */
/* ------------------------------------------------------------------------ */
/* C(t,Q) and C'(t,Q) of one Q (mp_sqom56.f90:1338-1377): the weighted
 * species sum of F, and the same of F - <F>_t (or of F) times qfac.
 * cc, cp: interleaved complex [T]. */
static void orc_combine_q(int64_t T, int64_t NQ, int nsp, const double *F, const double *wsc,
                          const double *wtab, const int32_t *mask, const double *qfac,
                          int subtract_mean, int64_t q, double *cc, double *cp)
{
    double mean[2 * 64];
    double *mj = (nsp <= 64) ? mean : (double *)malloc((size_t)(2 * nsp) * sizeof(double));
    for (int j = 0; j < nsp; ++j) {
        double ar = 0.0, ai = 0.0;
        for (int64_t t = 0; t < T; ++t) {
            ar += F[2 * ((t * nsp + j) * NQ + q)];
            ai += F[2 * ((t * nsp + j) * NQ + q) + 1];
        }
        mj[2 * j] = ar / (double)T;
        mj[2 * j + 1] = ai / (double)T;
    }
    for (int64_t t = 0; t < T; ++t) {
        double sr = 0.0, si = 0.0, pr = 0.0, pi = 0.0;
        for (int j = 0; j < nsp; ++j) {
            if (!mask[j]) continue;
            const double wv = wtab ? wtab[(int64_t)j * NQ + q] : wsc[j];
            const double fr = F[2 * ((t * nsp + j) * NQ + q)];
            const double fi = F[2 * ((t * nsp + j) * NQ + q) + 1];
            sr += wv * fr;
            si += wv * fi;
            if (subtract_mean) {
                pr += wv * (fr - mj[2 * j]);
                pi += wv * (fi - mj[2 * j + 1]);
            } else {
                pr += wv * fr;
                pi += wv * fi;
            }
        }
        const double qf = qfac ? qfac[q] : 1.0;
        cc[2 * t] = sr;
        cc[2 * t + 1] = si;
        cp[2 * t] = pr * qf;
        cp[2 * t + 1] = pi * qf;
    }
    if (mj != mean) free(mj);
}

/* ------------------------------------------------------------------------ */
/* One-phonon amplitudes (mp_sqom56.f90:1037-1098), per frame and atom type, */
/* in the reference's loop order with 1-based indices: displacement columns  */
/* u_x, u_y(k, l) (REAL(4) displacements, FP64 sums), their 2-D forward DFT  */
/* (direct double sum) / n_row(1), then every pixel of the n_bz x n_bz       */
/* zones.  pos float [T][n_tot][4], at_ind int32 [n_tot][4], base float      */
/* [nsp][3]; out complex128 [T][nsp][NQ], NQ = n_bz n1 x n_bz n2, x fastest.  */
/* Returns 2 for a cell index outside the columns.                           */
/* ------------------------------------------------------------------------ */
int orc_oneph(int64_t T, int nsp, const int32_t *n_row, int n_bz, int j_plane, float qx_min, float qy_min,
              const float *pos, const int32_t *at_ind, const float *base, double *out)
{
    const int n1 = n_row[0], n2 = n_row[1], n3 = n_row[2];
    const int64_t nsuper = (int64_t)n1 * n2 * n3, n_tot = nsuper * nsp;
    const int nqx = n_bz * n1, nqy = n_bz * n2;
    const int64_t NQ = (int64_t)nqx * nqy;
    const float twopi_s = (float)(2.0 * M_PI), sqrt2 = sqrtf(2.0f);
    const int nqx_min = (int)((float)n1 * (qx_min - floorf(qx_min)));
    const int nqy_min = (int)((float)n2 * (qy_min - floorf(qy_min)));
    double *ux = (double *)malloc(sizeof(double) * (size_t)n1 * n2);
    double *uy = (double *)malloc(sizeof(double) * (size_t)n1 * n2);
    double *Ux = (double *)malloc(sizeof(double) * 2 * (size_t)n1 * n2);
    double *Uy = (double *)malloc(sizeof(double) * 2 * (size_t)n1 * n2);
    int rc = 0;
    for (int64_t f = 0; f < T && !rc; ++f)
        for (int j = 0; j < nsp && !rc; ++j) {
            memset(ux, 0, sizeof(double) * (size_t)n1 * n2);
            memset(uy, 0, sizeof(double) * (size_t)n1 * n2);
            for (int64_t i = 0; i < nsuper; ++i) {
                const int64_t a = (int64_t)j * nsuper + i;
                const float *r = pos + 4 * (f * n_tot + a);
                const int32_t *ind = at_ind + 4 * a;
                if (!(r[0] != 0.0f || r[1] != 0.0f || r[2] != 0.0f)) continue;
                int k, l;
                float dx, dy;
                if (j_plane == 1) {
                    k = ind[0];
                    l = ind[1];
                    dx = r[0] - base[3 * j] - (float)ind[0] + (float)(n1 / 2);
                    dy = r[1] - base[3 * j + 1] - (float)ind[1] + (float)(n2 / 2);
                } else {
                    const int m = (n1 + n2) / 2;
                    k = ((ind[0] - 1 + ind[1] - 1) % m + m) % m + 1;
                    l = ind[2];
                    dx = r[0] - base[3 * j] - (float)ind[0] + (float)(n1 / 2) + 1.0f;
                    dx = dx + (r[1] - base[3 * j + 1] - (float)ind[1] + (float)(n2 / 2) + 1.0f);
                    dx = dx / sqrt2;
                    dy = r[2] - base[3 * j + 2] - (float)ind[2] + (float)(n3 / 2) + 1.0f;
                }
                if (k < 1 || k > n1 || l < 1 || l > n2) {
                    rc = 2;
                    break;
                }
                ux[(k - 1) + (int64_t)n1 * (l - 1)] += (double)dx;
                uy[(k - 1) + (int64_t)n1 * (l - 1)] += (double)dy;
            }
            if (rc) break;
            /* U(kk, ll) = sum_{k,l} u(k, l) exp(-2 pi i ((k-1)(kk-1)/n1 + (l-1)(ll-1)/n2)) / n1 */
            for (int kk = 0; kk < n1; ++kk)
                for (int ll = 0; ll < n2; ++ll) {
                    double xr = 0, xi = 0, yr = 0, yi = 0;
                    for (int k = 0; k < n1; ++k)
                        for (int l = 0; l < n2; ++l) {
                            const double ph = -ORC_TWO_PI * ((double)((int64_t)k * kk % n1) / n1 +
                                                            (double)((int64_t)l * ll % n2) / n2);
                            const double c = cos(ph), s = sin(ph);
                            xr += ux[k + (int64_t)n1 * l] * c;
                            xi += ux[k + (int64_t)n1 * l] * s;
                            yr += uy[k + (int64_t)n1 * l] * c;
                            yi += uy[k + (int64_t)n1 * l] * s;
                        }
                    Ux[2 * (kk + (int64_t)n1 * ll)] = xr / n1;
                    Ux[2 * (kk + (int64_t)n1 * ll) + 1] = xi / n1;
                    Uy[2 * (kk + (int64_t)n1 * ll)] = yr / n1;
                    Uy[2 * (kk + (int64_t)n1 * ll) + 1] = yi / n1;
                }
            for (int ii = 1; ii <= n_bz; ++ii)
                for (int jj = 1; jj <= n_bz; ++jj)
                    for (int l = 1; l <= n2; ++l)
                        for (int k = 1; k <= n1; ++k) {
                            const float q_x = qx_min + (float)(ii - 1) + (float)(k - 1) / (float)n1;
                            const float q_y = qy_min + (float)(jj - 1) + (float)(l - 1) / (float)n2;
                            int kk = (nqx_min + k) % n1, ll = (nqy_min + l) % n2;
                            if (kk == 0) kk = n1;
                            if (ll == 0) ll = n2;
                            float arg;
                            if (j_plane == 1) arg = q_x * base[3 * j] + q_y * base[3 * j + 1];
                            else arg = q_x * (base[3 * j] + base[3 * j + 1]) + q_y * base[3 * j + 2];
                            arg = twopi_s * arg;
                            const double c = cos((double)arg), s = -sin((double)arg);
                            const int64_t u0 = 2 * ((kk - 1) + (int64_t)n1 * (ll - 1));
                            const double ar = q_x * Ux[u0] + q_y * Uy[u0], ai = q_x * Ux[u0 + 1] + q_y * Uy[u0 + 1];
                            const int64_t X = (int64_t)(ii - 1) * n1 + (k - 1), Y = (int64_t)(jj - 1) * n2 + (l - 1);
                            double *o = out + 2 * ((f * nsp + j) * NQ + X + (int64_t)nqx * Y);
                            o[0] = c * ar - s * ai;
                            o[1] = c * ai + s * ar;
                        }
        }
    free(ux);
    free(uy);
    free(Ux);
    free(Uy);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* mp_dos spectra (mp_dos56.f90:495-525): per atom and component, the direct */
/* DFT of v(t) w(t) with w = 0.5 (1 - cos(2 pi i / (T + 1))), i = 1 .. T,    */
/* and out[f][s] = sum over atoms of type s (at_ind[4a + 3], 1-based) of     */
/* sum_c |X(f)|^2 / T for f < n_freq.  vel float [T][n_tot][4].              */
/* ------------------------------------------------------------------------ */
int orc_dos(int64_t T, int64_t n_tot, const float *vel, const int32_t *at_ind, int nsp, int64_t n_freq, double *out)
{
    if (T < 1 || n_freq < 1 || n_freq > T) return 2;
    memset(out, 0, sizeof(double) * (size_t)(n_freq * nsp));
    double *w = (double *)malloc(sizeof(double) * (size_t)T);
    for (int64_t t = 0; t < T; ++t) w[t] = 0.5 * (1.0 - cos(ORC_TWO_PI * (double)(t + 1) / (double)(T + 1)));
    for (int64_t a = 0; a < n_tot; ++a) {
        const int s = at_ind[4 * a + 3] - 1;
        if (s < 0 || s >= nsp) continue;
        for (int c = 0; c < 3; ++c)
            for (int64_t f = 0; f < n_freq; ++f) {
                double xr = 0.0, xi = 0.0;
                for (int64_t t = 0; t < T; ++t) {
                    const double v = (double)vel[4 * (t * n_tot + a) + c] * w[t];
                    const double ph = -ORC_TWO_PI * (double)((f * t) % T) / (double)T;
                    xr += v * cos(ph);
                    xi += v * sin(ph);
                }
                out[f * nsp + s] += (xr * xr + xi * xi) / (double)T;
            }
    }
    free(w);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as  */
/* easy as 1, 2, 3", SC'11): ten rounds of                                   */
/*   (c0, c1, c2, c3) <- (hi(M1 c2) ^ c1 ^ k0, lo(M1 c2),                    */
/*                        hi(M0 c0) ^ c3 ^ k1, lo(M0 c0))                    */
/* with M0 = 0xD2511F53, M1 = 0xCD9E8D57 and the key bumped by the Weyl      */
/* constants (0x9E3779B9, 0xBB67AE85) after each round.                      */
/* ------------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out)
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static float orc_unit(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

/* |Q| of the mode offset (k1, k2, k3) in FP32 (mp_sql56.f90:1036-1044):      */
/* norm2(a_cell_inv . (k * d_q)), every operation rounded on its own.        */
static float orc_sql_qnorm(const float *dq, const float *ainv, int k1, int k2, int k3)
{
    const float u[3] = {(float)k1 * dq[0], (float)k2 * dq[1], (float)k3 * dq[2]};
    float v[3];
    for (int a = 0; a < 3; ++a) v[a] = (ainv[3 * a] * u[0] + ainv[3 * a + 1] * u[1]) + ainv[3 * a + 2] * u[2];
    return sqrtf((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
}

/* ------------------------------------------------------------------------ */
/* mp_sql's Monte Carlo integration (j_acc = 2, mp_sql56.f90:1177-1210) for  */
/* one frame: chains k < n_h of n_mc samples j.  Sample (j, k), attempt a:   */
/* x = philox(counter (j, k, a, tag), key (seed lo, seed hi)); the point     */
/* p_i = 2 u(x_i) - 1 (i < 3, u(x) = (x >> 8) 2^-24) is kept when            */
/* 0 < norm2(p) <= 1 (:1178-1185); then r = u(x_3) and                       */
/* d = nint(h (p / |p|) r), h_d = N_d - (N_d / 2 + 1) (:1187-1190), in FP32.  */
/* d = 0 -> skip[0]++ (:1192-1195); |Q(d)| > q_range -> skip[1]++ (:1208-1210)*/
/* else hist[nint(|Q| / q_step)][p] += Re(A_i A_j*) at d for pairs i >= j     */
/* (:1201-1207; bins >= n_hist are dropped).  64 misses -> skip[0]++.        */
/* amp [nsp][N3][N2][N1] complex (re, im), mode i_d - N_d / 2.  hist and     */
/* skip are accumulated.                                                     */
/* ------------------------------------------------------------------------ */
int orc_sql_hist_mc(int nsp, const int64_t *N, const float *dq, const float *ainv, float q_range, float q_step,
                    int n_hist, int64_t n_h, int64_t n_mc, uint64_t seed, uint32_t tag, const double *amp,
                    double *hist, int64_t *skip)
{
    if (nsp < 1 || nsp > 8 || n_hist < 1 || n_h < 1 || n_mc < 1) return 2;
    const int npairs = nsp * (nsp + 1) / 2;
    const int64_t nm = N[0] * N[1] * N[2];
    const int64_t total = n_h * n_mc;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    int c[3];
    float h[3];
    for (int d = 0; d < 3; ++d) {
        c[d] = (int)(N[d] / 2);
        h[d] = (float)(N[d] - (c[d] + 1));
    }
    const int nthr = orc_num_threads();
    double *th = (double *)calloc((size_t)nthr * n_hist * npairs, sizeof(double));
    int64_t *ts = (int64_t *)calloc((size_t)nthr * 2, sizeof(int64_t));
#pragma omp parallel num_threads(nthr)
    {
#ifdef _OPENMP
        const int tid = omp_get_thread_num();
#else
        const int tid = 0;
#endif
        double *hh = th + (size_t)tid * n_hist * npairs;
        int64_t *ss = ts + 2 * tid;
#pragma omp for schedule(static)
        for (int64_t idx = 0; idx < total; ++idx) {
            const uint32_t j = (uint32_t)(idx % n_mc), k = (uint32_t)(idx / n_mc);
            uint32_t x[4];
            float p[3], pn = 0.0f;
            int ok = 0;
            for (uint32_t a = 0; a < 64 && !ok; ++a) {
                const uint32_t ctr[4] = {j, k, a, tag};
                orc_philox4x32_10(ctr, key, x);
                for (int i = 0; i < 3; ++i) p[i] = 2.0f * orc_unit(x[i]) - 1.0f;
                pn = sqrtf((p[0] * p[0] + p[1] * p[1]) + p[2] * p[2]);
                ok = pn > 0.0f && pn <= 1.0f;
            }
            if (!ok) {
                ss[0]++;
                continue;
            }
            const float r = orc_unit(x[3]);
            int d[3];
            for (int i = 0; i < 3; ++i) d[i] = (int)roundf((h[i] * (p[i] / pn)) * r);
            if (d[0] == 0 && d[1] == 0 && d[2] == 0) {
                ss[0]++;
                continue;
            }
            int inside = 1;
            for (int i = 0; i < 3; ++i) inside &= (c[i] + d[i] >= 0 && c[i] + d[i] < N[i]);
            const float q = inside ? orc_sql_qnorm(dq, ainv, d[0], d[1], d[2]) : 0.0f;
            if (!inside || !(q <= q_range)) {
                ss[1]++;
                continue;
            }
            const int b = (int)roundf(q / q_step);
            if (b >= n_hist) continue;
            const int64_t m = (c[0] + d[0]) + N[0] * ((c[1] + d[1]) + N[1] * (int64_t)(c[2] + d[2]));
            int pp = 0;
            for (int s = 0; s < nsp; ++s)
                for (int t = 0; t <= s; ++t, ++pp) {
                    const double *as = amp + 2 * (s * nm + m), *at = amp + 2 * (t * nm + m);
                    hh[b * npairs + pp] += as[0] * at[0] + as[1] * at[1];
                }
        }
    }
    for (int t = 0; t < nthr; ++t) {
        for (int64_t i = 0; i < (int64_t)n_hist * npairs; ++i) hist[i] += th[(size_t)t * n_hist * npairs + i];
        skip[0] += ts[2 * t];
        skip[1] += ts[2 * t + 1];
    }
    free(th);
    free(ts);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Speckle filter (mp_sqom56.f90:1539-1548, one pass): csp(i2) for i2 = 1..n */
/* in order, i1 = i2 - 1 and i3 = i2 + 1 wrapped into 1..n, with the values */
/* already replaced earlier in the pass.  v is the pass's flat walk order.   */
/* ------------------------------------------------------------------------ */
void orc_speckle_pass(int64_t n, double *v, double s_trig)
{
    for (int64_t i2 = 1; i2 <= n; ++i2) {
        int64_t i1 = i2 - 1, i3 = i2 + 1;
        if (i1 < 1) i1 += n;
        if (i3 > n) i3 -= n;
        const double c1 = v[i1 - 1], c3 = v[i3 - 1];
        const double mn = c1 < c3 ? c1 : c3, mx = c1 > c3 ? c1 : c3;
        if (v[i2 - 1] < mn / s_trig) v[i2 - 1] = mn;
        if (v[i2 - 1] > s_trig * mx) v[i2 - 1] = mx;
    }
}

/* ------------------------------------------------------------------------ */
/* The intermediate function I(Q,t) (j_sq != 1): for lags lag = lag_lo + l,  */
/*   out[l][Q] = Re sum_{i=0}^{n_int-1} C'(i + lag, Q) conj(C'(i, Q))        */
/* times 1/n_int when scaled (the E = const map form, mp_sqom56.f90:1467-1471*/
/* -- cs_plot is REAL, so the real part), unscaled otherwise (the E(Q) cut,  */
/* :1504-1510), over nat^2 (:1574-1575, no time step for I(Q,t)).  C' as in  */
/* orc_sqw_direct; 0 <= lag <= T - n_int.                                    */
/* ------------------------------------------------------------------------ */
int orc_iqt(int64_t T, int64_t NQ, int nsp, const double *F, const double *wsc, const double *wtab,
            const int32_t *mask, const double *qfac, int64_t n_int, int64_t lag_lo, int64_t n_lag,
            int subtract_mean, int scaled, double nat, double *out)
{
    if (T < 1 || NQ < 0 || nsp < 1 || n_int < 1 || n_int > T || lag_lo < 0 || n_lag < 0 || !(nat > 0) ||
        lag_lo + n_lag - 1 > T - n_int)
        return 2;
#pragma omp parallel
    {
        double *cc = (double *)malloc((size_t)(2 * T) * sizeof(double));
        double *cp = (double *)malloc((size_t)(2 * T) * sizeof(double));
#pragma omp for schedule(static)
        for (int64_t q = 0; q < NQ; ++q) {
            orc_combine_q(T, NQ, nsp, F, wsc, wtab, mask, qfac, subtract_mean, q, cc, cp);
            for (int64_t l = 0; l < n_lag; ++l) {
                const int64_t lag = lag_lo + l;
                double acc = 0.0;
                for (int64_t i = 0; i < n_int; ++i)
                    acc += cp[2 * (i + lag)] * cp[2 * i] + cp[2 * (i + lag) + 1] * cp[2 * i + 1];
                out[l * NQ + q] = (scaled ? acc / (double)n_int : acc) / (nat * nat);
            }
        }
        free(cc);
        free(cp);
    }
    return 0;
}

int orc_sqw_direct(int64_t T, int64_t NQ, int nsp, const double *F,
                   const double *wsc, const double *wtab, const int32_t *mask,
                   const double *qfac, int64_t n_int, int win_kind, double dt,
                   double nat, int64_t m_lo, int64_t n_f, int subtract_mean,
                   double *s, double *stot)
{
    if (T < 1 || NQ < 0 || nsp < 1 || n_int < 1 || n_int > T || n_f < 0) return 2;
    const int64_t n_frame = T - n_int + 1;
    double *win = (double *)malloc((size_t)n_int * sizeof(double));
    double *twr = (double *)malloc((size_t)(n_f * n_int) * sizeof(double));
    double *twi = (double *)malloc((size_t)(n_f * n_int) * sizeof(double));
    for (int64_t i = 0; i < n_int; ++i) {
        if (n_int <= 1) win[i] = 1.0;
        else if (win_kind == 0) win[i] = 0.5 * (1.0 - cos(ORC_TWO_PI * (double)i / (double)(n_int - 1)));
        else win[i] = 0.5 * (1.0 - cos(ORC_TWO_PI * (double)(i + 1) / (double)(n_int + 1)));
    }
    for (int64_t mi = 0; mi < n_f; ++mi) {
        const int64_t m = m_lo + mi;
        for (int64_t i = 0; i < n_int; ++i) {
            int64_t p = (m * i) % n_int;
            if (p < 0) p += n_int;
            const double ph = -ORC_TWO_PI * (double)p / (double)n_int;
            twr[mi * n_int + i] = win[i] * cos(ph);
            twi[mi * n_int + i] = win[i] * sin(ph);
        }
    }
    const double inv_n2 = 1.0 / (nat * nat);
    const double norm_w = dt * inv_n2 / (0.5 * (double)n_frame * (double)n_int);
#pragma omp parallel
    {
        double *cc = (double *)malloc((size_t)(2 * T) * sizeof(double));
        double *cp = (double *)malloc((size_t)(2 * T) * sizeof(double));
#pragma omp for schedule(static)
        for (int64_t q = 0; q < NQ; ++q) {
            orc_combine_q(T, NQ, nsp, F, wsc, wtab, mask, qfac, subtract_mean, q, cc, cp);
            double tot = 0.0;
            for (int64_t t = 0; t < T; ++t) tot += cc[2 * t] * cc[2 * t] + cc[2 * t + 1] * cc[2 * t + 1];
            if (stot) stot[q] = tot / (double)T * inv_n2;
            for (int64_t mi = 0; mi < n_f; ++mi) {
                double acc = 0.0;
                for (int64_t sidx = 0; sidx < n_frame; ++sidx) {
                    double xr = 0.0, xi = 0.0;
                    for (int64_t i = 0; i < n_int; ++i) {
                        const double ar = cp[2 * (sidx + i)], ai = cp[2 * (sidx + i) + 1];
                        const double br = twr[mi * n_int + i], bi = twi[mi * n_int + i];
                        xr += ar * br - ai * bi;
                        xi += ar * bi + ai * br;
                    }
                    acc += xr * xr + xi * xi;
                }
                s[mi * NQ + q] = acc * norm_w;
            }
        }
        free(cc);
        free(cp);
    }
    free(win);
    free(twr);
    free(twi);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* a4.2: exponential-of-semicircle spreading onto the periodic fine grid,    */
/* FP64, OpenMP (the FINUFFT method, barnett_FINUFFT.txt:38-40).  Used by    */
/* the CPU FINUFFT-method restatement that bench.py times beside the GPU.    */
/*   phi(t) = exp(beta * (sqrt(1 - (2t/w)^2) - 1)),  |t| <= w/2              */
/*   grid[l] += c_j * prod_d phi(l_d - u_d)  for the w^dim cells of point j  */
/* The grid is partitioned into slabs along the slowest used dimension; each */
/* thread owns a slab and walks the points whose kernel rows touch it, so no */
/* two threads write one cell.  grid is interleaved complex [n3][n2][n1].    */
/* ------------------------------------------------------------------------ */
int orc_es_spread(int dim, int64_t M, const double *x, const double *y,
                  const double *z, const double *c, const int64_t *n, int w,
                  double beta, double *grid)
{
    if (dim < 1 || dim > 3 || w < 2 || w > 16) return 2;
    const int64_t n1 = n[0], n2 = dim >= 2 ? n[1] : 1, n3 = dim >= 3 ? n[2] : 1;
    const int64_t tot = n1 * n2 * n3;
    memset(grid, 0, (size_t)(2 * tot) * sizeof(double));
    int32_t *l = (int32_t *)malloc((size_t)(3 * M) * sizeof(int32_t));
    double *d = (double *)malloc((size_t)(3 * M) * sizeof(double));
    int32_t *key = (int32_t *)malloc((size_t)M * sizeof(int32_t));
    /* kernel values of every point, computed once: kv[(3 j + dd) w + m] */
    double *kv = (double *)malloc((size_t)M * 3 * (size_t)w * sizeof(double));
    const int64_t nn[3] = {n1, n2, n3};
    orc_fold(dim, M, x, y, z, nn, w, 1, 1, l, d, key);
    const double hw = 0.5 * w, ic = 1.0 / (hw * hw);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < M; ++j) {
        double *kj = kv + j * 3 * w;
        for (int dd = 0; dd < 3; ++dd)
            for (int m = 0; m < w; ++m) {
                if (dd < dim) {
                    const double t = d[3 * j + dd] + (double)m;
                    const double a = 1.0 - t * t * ic;
                    kj[w * dd + m] = a > 0.0 ? exp(beta * (sqrt(a) - 1.0)) : 0.0;
                } else {
                    kj[w * dd + m] = (m == 0) ? 1.0 : 0.0;
                }
            }
    }
    /* threads own contiguous slabs of the slowest grid dimension (z in 3D)
     * and spread the part of every stencil that falls in their slab, so
     * every grid cell is written by one thread in a fixed point order */
    const int sd = dim - 1;               /* slab dimension */
    const int64_t ns = nn[sd];
    /* bucket points by first kernel cell along the slab dimension */
    int64_t *start = (int64_t *)calloc((size_t)ns + 1, sizeof(int64_t));
    int32_t *ord = (int32_t *)malloc((size_t)M * sizeof(int32_t));
    for (int64_t j = 0; j < M; ++j) start[l[3 * j + sd] + 1]++;
    for (int64_t i = 0; i < ns; ++i) start[i + 1] += start[i];
    {
        int64_t *fill = (int64_t *)malloc((size_t)ns * sizeof(int64_t));
        memcpy(fill, start, (size_t)ns * sizeof(int64_t));
        for (int64_t j = 0; j < M; ++j) ord[fill[l[3 * j + sd]]++] = (int32_t)j;
        free(fill);
    }
#pragma omp parallel
    {
        int nt = 1, tid = 0;
#ifdef _OPENMP
        nt = omp_get_num_threads();
        tid = omp_get_thread_num();
#endif
        const int64_t s0 = ns * tid / nt, s1 = ns * (tid + 1) / nt;
        if (s1 > s0) {
            /* first-cell values whose rows reach [s0, s1): l in [s0-w+1, s1-1];
             * when that range wraps onto itself every bucket is visited once */
            const int64_t len = (s1 - s0) + w - 1;
            const int64_t lv0 = (len >= ns) ? 0 : s0 - w + 1;
            const int64_t lv1 = (len >= ns) ? ns : s1;
            for (int64_t lv = lv0; lv < lv1; ++lv) {
                const int64_t lb = ((lv % ns) + ns) % ns;
                for (int64_t q = start[lb]; q < start[lb + 1]; ++q) {
                    const int64_t j = ord[q];
                    const double *ker = kv + j * 3 * w;
                    const double cr = c[2 * j], ci = c[2 * j + 1];
                    const int w2 = dim >= 2 ? w : 1, w3 = dim >= 3 ? w : 1;
                    for (int m3 = 0; m3 < w3; ++m3) {
                        const int64_t i3 = (l[3 * j + 2] + m3) % n3;
                        if (sd == 2 && (i3 < s0 || i3 >= s1)) continue;
                        for (int m2 = 0; m2 < w2; ++m2) {
                            const int64_t i2 = (l[3 * j + 1] + m2) % n2;
                            if (sd == 1 && (i2 < s0 || i2 >= s1)) continue;
                            const double k23 = ker[w + m2] * ker[2 * w + m3];
                            const double sr = cr * k23, si = ci * k23;
                            double *row = grid + 2 * ((i3 * n2 + i2) * n1);
                            const int64_t l1 = l[3 * j];
                            if (sd != 0 && l1 + w <= n1) {
                                double *rr = row + 2 * l1;
                                for (int m1 = 0; m1 < w; ++m1) {
                                    rr[2 * m1] += sr * ker[m1];
                                    rr[2 * m1 + 1] += si * ker[m1];
                                }
                                continue;
                            }
                            for (int m1 = 0; m1 < w; ++m1) {
                                const int64_t i1 = (l1 + m1) % n1;
                                if (sd == 0 && (i1 < s0 || i1 >= s1)) continue;
                                row[2 * i1] += sr * ker[m1];
                                row[2 * i1 + 1] += si * ker[m1];
                            }
                        }
                    }
                }
            }
        }
    }
    free(kv);
    free(l); free(d); free(key); free(start); free(ord);
    return 0;
}
