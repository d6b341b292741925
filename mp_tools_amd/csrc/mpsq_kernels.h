// mpsq_kernels.h -- parameter blocks shared by host code and kernels, and
// the host-side launch helpers implemented in the .hip files.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpsq_device.h"

namespace mpsq {

constexpr int MAXW = 16;          // widest kernel (eps ~ 1e-15)
constexpr int MAX_NB_BINS = 256;  // sort bins a spread tile may visit

// Fine-grid / kernel geometry of one NUFFT plan.
struct Geo {
    int dim;
    int n1, n2, n3;   // fine grid (unused dims = 1)
    int N1, N2, N3;   // modes (unused dims = 1)
    int w;            // kernel width
    double beta, ic;  // phi(t) = exp(beta (sqrt(1 - ic t^2) - 1)), ic = (2/w)^2
    double s1, s2, s3;  // n_d / (2 pi)
    double hw;        // w / 2
    int by, bz;       // spread tile rows along y, z (tile spans all of x)
    int BY, BZ;       // sort-bin extent along y, z
    int nbY, nbZ;     // number of sort bins along y, z
    int nbins;
};

// Sorted point record: first kernel cell (wrapped) and offset of the point
// from it, in fine-grid units, plus the point's index in the caller's order.
template <typename R> struct alignas(16) PtRec {
    int lx, ly, lz, idx;
    R dx, dy, dz, pad;
};

// Trajectory staging (a2 selection + a3 coordinates/strengths).
struct StageParams {
    int64_t n_tot;       // atoms per frame
    int vacancy_check;   // drop atoms at exactly (0,0,0)
    int select_mode;     // 0 none, 1 CELL window on at_ind, 2 BULK box
    int n_row[3];
    float pmin[3], pmax[3];
    float margin;
    double centre[3];
    double axes[9];      // rows: unit plane vectors e1, e2, e3
    double tpe[3];
    double q0[3];        // q_center: strength phase exp(-i 2 pi q0 . r)
    int window;          // spatial Hann window
    int conj;            // conjugate strengths (iflag > 0)
};

struct SortChunk {
    int64_t p0, p1;   // point range of the chunk
    int64_t pbase;    // first point of the chunk's transform
    int64_t hbase;    // start of the transform's count matrix
    int nch, lc;      // chunks of the transform, index of this chunk
};

struct TransformInfo {
    int64_t p0, p1, hbase;
    int nch, pad;
};

struct ColPass {
    int n, N, cols, inner, outer;
    FDiv dcols;  // make_fdiv(cols)
    int64_t in_tstride, out_tstride;
    int nsp;
    int64_t NQ;
    int conj;
};

// S(Q,w) reduction parameters
struct SqwParams {
    int64_t T, NQ;
    int n_int;
    int64_t n_frame;
    int64_t m_lo;
    int n_f;
    int win_kind;       // 0: 0.5(1-cos(2pi i/(n-1))), 1: 0.5(1-cos(2pi(i+1)/(n+1)))
    int subtract_mean;
    double norm_s;      // dt / (nat^2 * 0.5 * n_frame * n_int)
    double norm_tot;    // 1 / (T * nat^2)
};

// --- launch helpers (nufft_kernels.hip) ---
template <typename R>
void launch_fold(const Geo &g, int64_t M, const void *x, const void *y, const void *z, bool in_f32, PtRec<R> *rec,
                 int *key, int *bad, hipStream_t st);
template <typename R>
void launch_stage(const Geo &g, const StageParams &sp, int64_t npts, const float4 *pos, const int4 *at_ind,
                  PtRec<R> *rec, int *key, typename cplx<R>::T *cstr, int *bad, hipStream_t st);
void launch_sort_count(const int *key, const SortChunk *chunks, int nchunks, int nbins, int *H, hipStream_t st);
void launch_scan(const int *H, int64_t n, int64_t *bsum, int64_t *Hs, hipStream_t st);
int64_t scan_bsum_size(int64_t n);
template <typename R>
void launch_sort_scatter(const int *key, const SortChunk *chunks, int nchunks, int nbins, int keybits,
                         const int64_t *Hs, const PtRec<R> *rec, const typename cplx<R>::T *cin, PtRec<R> *srec,
                         typename cplx<R>::T *cout, int *perm, hipStream_t st);
void launch_binstart(const int64_t *Hs, const TransformInfo *tinfo, int ntr, int nbins, int64_t *binstart,
                     hipStream_t st);
template <typename R>
void launch_gather(int64_t M, int ntr, const int *perm, const typename cplx<R>::T *c, typename cplx<R>::T *cs,
                   int conj, hipStream_t st);
template <typename R> int spread_tile_elems();
int spread_max_runs();
int spread_max_tile_b();
template <typename R> size_t spread_lds(int rows, int n1, int w);
template <typename R> int colfft_tile_elems(bool final);
// amax[t] bounds |Re c|, |Im c| of transform t's strengths (FP32 form only;
// nullptr: strengths are known to be bounded by 1)
template <typename R>
void launch_spread_xfft(const Geo &g, const RadixPlan &rp1, int ntr, const PtRec<R> *srec,
                        const typename cplx<R>::T *cs, int64_t c_tstride, const int64_t *binstart, int bin_tstride,
                        const typename cplx<R>::T *tw1, const R *phinv1, typename cplx<R>::T *G1, int final_conj,
                        const float *amax, hipStream_t st);
// amax[t] = max over the M strengths of transform t of max(|Re|, |Im|)
void launch_absmax(int64_t M, int ntr, const float2 *c, float *amax, hipStream_t st);
template <typename R>
void launch_colfft(const ColPass &cp, const RadixPlan &rp, int ngroups, const typename cplx<R>::T *in,
                   typename cplx<R>::T *out, const typename cplx<R>::T *tw, const R *phinv, const R *wsc,
                   const float *wtab, bool final, hipStream_t st);

// --- launch helpers (sqw_kernels.hip) ---
template <typename R> int sqw_tile_elems();
template <typename R>
void launch_sqw_single(const SqwParams &p, const RadixPlan &rp, int cols, const typename cplx<R>::T *amp,
                       const typename cplx<R>::T *tw, const double *win, const float *qfac, double *S, double *Stot,
                       hipStream_t st);
template <typename R>
void launch_sqw_sliding(const SqwParams &p, const typename cplx<R>::T *amp, const float *qfac, double *S,
                        double *Stot, hipStream_t st);

}  // namespace mpsq
