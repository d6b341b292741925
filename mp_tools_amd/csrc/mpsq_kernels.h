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
    int nbins, keybits;  // sort bins of the chunk's transform, bits of its keys
};

struct TransformInfo {
    int64_t p0, p1, hbase;
    int64_t bs0;      // first entry of the transform's binstart block (nbins + 1 entries)
    int nch, nbins;
    int sxsh, nseg;   // x segments of the covering bound: 2^sxsh cells (>= w), nseg of them
};

struct ColPass {
    int n, N, cols, inner, outer;
    FDiv dcols;  // make_fdiv(cols)
    int64_t in_tstride, out_tstride;
    int nsp;
    int64_t NQ;
    int conj;
    // Outputs kept per column: N, or N + 1 for a half-spectrum (Hermitian)
    // pass, which also keeps k = N - N/2 so that every mode's mirror -k is
    // among the kept ones.
    int NO;
    // herm (final pass): the columns are the half spectrum of a transform of
    // real strengths, x frequencies kh in [0, hH1v) (k1 = kh - N1/2 <= 0;
    // rows padded to hH1 columns, whose pad columns are not written) and,
    // with hmid, middle (y) indices in [0, hN2 + 1); the pass writes each
    // kept mode and its mirror F(-k) = conj F(k) into the full [N][hN2][hN1]
    // map (hsa: the map stride of the transformed axis)
    int herm, hmid, hN1, hH1, hH1v, hN2;
    int64_t hsa;
    FDiv dH1;
};

// ---------------------------------------------------------------------------
// Job tables.  One launch of a phase serves any number of jobs (different
// geometries, frame counts, species); a single-job run is a one-entry table.
// ---------------------------------------------------------------------------
constexpr int MAX_JOBS = 256;   // jobs of one launch (tile-id prefix cached in LDS)
// counters of the persistent spread, one 256-byte line each (a line shared
// by the 8 XCDs' L2s serialises every atomic on it): 8 per-XCD tile
// counters, then the count of queued wide passes and their take counter
constexpr int XCD_CTR_STRIDE = 64;
constexpr int XCD_CTR_INTS = 10 * XCD_CTR_STRIDE;

// trajectory staging of one job's batch of frames: points [out0, out0+npts)
// of the shared point buffers, blocks [blk0, blk0 + ceil(npts/256))
struct StageJob {
    Geo g;
    StageParams sp;
    const float4 *pos;   // first frame of the batch
    const int4 *at_ind;
    int64_t npts;        // frames * atoms staged (active species only)
    int64_t out0;
    int64_t nspan;       // atoms per frame in pos (n_tot) and active atoms per frame
    int64_t mact;
    int blk0;
    int nact;
    int64_t act_src[64]; // first atom of active species k in a frame
    int64_t act_off[65]; // prefix of active species counts
    double wst[64];      // strength factor of active species k (its weight when the
                         // species of a frame form one transform, else 1)
};

// spread + x FFT of ntr transforms sharing one geometry; tiles
// [tile0, tile0 + ntiles * ntr) of the launch's tile-id space
template <typename R> struct SpreadJob {
    Geo g;
    RadixPlan rp1;
    const PtRec<R> *srec;                // sorted records (absolute positions)
    int64_t nrec;                        // records in srec (bound of every run position)
    int *err;                            // error flags of the run (bit 1: a run position out of range)
    const typename cplx<R>::T *cs;       // sorted strengths, transform t at + t*c_tstride
    int64_t c_tstride;
    const int64_t *binstart;             // transform t's block at + t*bin_tstride*(nbins+1)
    int bin_tstride;
    const typename cplx<R>::T *tw1;
    const R *phinv1;
    typename cplx<R>::T *G1;             // [ntr][n3][n2][g1_len]
    const float *amax;                   // per transform, or nullptr (strengths bounded by amax_s)
    float amax_s;
    const int *binmax;                   // FP32: per-bin covering bound, indexed like binstart (or nullptr)
    int grp, grp_sh;                     // rows per x-FFT group (a power of two) and its log2
    int final_conj;
    int ntr, ntiles, tile0;
    int tmajor;                          // tile ids transform-major (else position-major)
    FDiv dntr, dnty, dby, dnt;           // ntr, tiles along y, by, ntiles
    int lds_tw;                          // byte offset of the twiddles in dynamic LDS
    int g1_len;                          // G1 row length: N1, or the half spectrum's padded to 16
    int g1_keep;                         // kept frequencies per row: N1, or N1 / 2 + 1 (k1 <= 0)
};

// one column-FFT pass of one job: blocks [blk0, blk0 + ncb*outer*ngroups)
template <typename R> struct ColJob {
    ColPass cp;
    RadixPlan rp;
    const typename cplx<R>::T *in;
    typename cplx<R>::T *out;
    const typename cplx<R>::T *tw;
    const double2 *tw64;                 // FP64 twiddles (passes that compute in FP64 on FP32 data)
    const R *phinv, *wsc;
    const float *wtab;
    int ncb, ngroups, blk0;
};

// S(Q,w) reduction parameters
// Q columns of one k_sqw_single workgroup: a power of two <= this
constexpr int SQW_MAX_COLS = 64;

struct SqwParams {
    int64_t T, NQ;      // frames; modes reduced (outputs [n_f][NQ], [NQ])
    int64_t NQs;        // modes per frame of the amplitudes (row stride)
    const int *qidx;    // mode of output q (a cut), or nullptr: q itself
    int kind;           // 0 S(Q,w); 1, 2 I(Q,t) (1: times 1/n_int)
    int n_int;
    int64_t n_frame;
    int64_t m_lo;
    int n_f;
    int win_kind;       // 0: 0.5(1-cos(2pi i/(n-1))), 1: 0.5(1-cos(2pi(i+1)/(n+1)))
    int subtract_mean;
    double norm_s;      // S(Q,w): dt / (nat^2 * 0.5 * n_frame * n_int); I(Q,t): 1 / (nat^2 [* n_int])
    double norm_tot;    // 1 / (T * nat^2)
};

// --- launch helpers (nufft_kernels.hip) ---
template <typename R>
void launch_fold(const Geo &g, int64_t M, const void *x, const void *y, const void *z, bool in_f32, PtRec<R> *rec,
                 int *key, int *bad, hipStream_t st);
template <typename R>
void launch_stage_jobs(const StageJob *jobs, const int *blk0s, int njobs, int nblocks, PtRec<R> *rec, int *key,
                       typename cplx<R>::T *cstr, int *bad, hipStream_t st);
void launch_sort_count(const int *key, const SortChunk *chunks, int nchunks, int max_nbins, int *H, hipStream_t st);
void launch_scan(const int *H, int64_t n, int64_t *bsum, int64_t *Hs, hipStream_t st);
int64_t scan_bsum_size(int64_t n);
template <typename R>
void launch_sort_scatter(const int *key, const SortChunk *chunks, int nchunks, int max_nbins, const int64_t *Hs,
                         const PtRec<R> *rec, const typename cplx<R>::T *cin, PtRec<R> *srec,
                         typename cplx<R>::T *cout, int *perm, hipStream_t st);
void launch_binstart(const int64_t *Hs, const TransformInfo *tinfo, int ntr, int max_nbins, int64_t *binstart,
                     hipStream_t st);
template <typename R>
void launch_gather(int64_t M, int ntr, const int *perm, const typename cplx<R>::T *c, typename cplx<R>::T *cs,
                   int conj, hipStream_t st);
template <typename R> int spread_tile_elems();
int spread_max_runs();
int spread_max_tile_b();
// wide passes a queued FP32 tile can take (entries of fbl per tile)
int spread_max_wide_passes();
// largest tile (rows of the whole x extent) of a fine grid of n1 cells in x,
// and the rows one x-FFT group of such a tile holds
template <typename R> int spread_max_rows(int n1);
template <typename R> int spread_group_rows(int rows, int n1);
// (real: the geometry of a real-strength plan, whose narrow tiles have 4-byte cells)
template <typename R> size_t spread_lds(int rows, int n1, int w, bool real = false);
// byte offset of the twiddles in a spread workgroup's dynamic LDS
template <typename R> size_t spread_lds_twiddles(int rows, int n1, int w, bool real = false);
// FP32 tiles of real strengths: whether a geometry can take the real narrow
// kernel (rows in pairs, one 32-bit cell per row cell, two rows per complex
// x FFT) and that kernel's dynamic LDS
bool spread_real_ok(int rows, int n1, int w);
size_t spread_lds_real(int rows, int n1, int w);
// binmax[bs0 + b] = largest number of bin b's points whose first kernel cell
// lies in 3 cyclically consecutive x segments (a bound on how many of them
// cover any one cell of a row)
template <typename R>
void launch_binmax(const PtRec<R> *srec, const int64_t *binstart, const TransformInfo *tinfo, int ntr, int max_nbins,
                   int *binmax, hipStream_t st);
template <typename R> int colfft_tile_elems(bool final);
// columns of length n <= colfft_final_small_elems() take the final pass's
// smaller tile (and its blocks of fewer threads)
template <typename R> int colfft_final_small_elems();
// Spread + x FFT of every tile of a job table in one persistent launch.
// Workgroups run on XCD blockIdx.x % 8 and take tiles of their XCD's range
// [xr[x], xr[x+1]) of the tile-id space from the counter
// wc[x * XCD_CTR_STRIDE] (wc: XCD_CTR_INTS ints, zeroed by this call).  FP32:
// tiles too dense for 32-bit accumulation are queued in fbl
// (spread_max_wide_passes() ints per tile of the table) and run by a second,
// 64-bit launch.  All jobs share the
// kernel width w; lds = max over jobs of spread_lds.  lds_real > 0 (FP32
// only): every job's strengths are real (imaginary parts exactly zero; a
// point with a nonzero one raises err) and every job's geometry passes
// spread_real_ok, and the narrow launch takes the real form with lds_real =
// max over jobs of spread_lds_real.
template <typename R>
void launch_spread_jobs(const SpreadJob<R> *jobs, const int *tile0s, int njobs, int w, const int *xr, int *wc,
                        int *fbl, size_t lds, hipStream_t st, size_t lds_real = 0);
// |Q| histogram of partial intensities (mp_sql's Gauss integration)
struct SqlGeom {
    int N[3];
    float dq[3];
    float ainv[9];  // row-major: v_i = sum_j ainv[3 i + j] u_j
    float q_range, q_step;
    int n_hist, nsp, npairs;
    int64_t nmodes;
};
// out[n_hist][npairs] += the histogram of amp ([nsp][nmodes] complex);
// returns 0, -1 if the histogram does not fit one workgroup's LDS (64 KB)
// or nsp > 8, -2 if the partial buffer could not be allocated
int launch_sql_hist(const void *amp, bool fp64, const SqlGeom &g, double *out, hipStream_t st);
// Monte Carlo integration: n_h chains of n_mc samples drawn by Philox4x32-10
// under key seed and counter (j, k, attempt, tag)
struct SqlMc {
    uint64_t seed;
    uint32_t tag;
    int64_t n_h, n_mc;
};
// out[n_hist][npairs] += the sampled histogram, skip[0..1] += the counts of
// samples at d = 0 and beyond q_range; returns 0, -1 (LDS or nsp), -2
// (partial buffer)
int launch_sql_hist_mc(const void *amp, bool fp64, const SqlGeom &g, const SqlMc &mc, double *out,
                       unsigned long long *skip, hipStream_t st);
// the same with its partial slices in a static device array (no allocation);
// returns 0, -1 (LDS), -2 (no scratch address), -3 (scratch too small)
int launch_sql_hist_static(const void *amp, bool fp64, const SqlGeom &g, double *out, hipStream_t st);
int64_t sql_scratch_doubles();
// p[0..n) = 0 by a kernel (graph-capturable, unlike a memset node)
void launch_zero_i32(int *p, int n, hipStream_t st);
// amax[t] = max over the M strengths of transform t of max(|Re|, |Im|)
void launch_absmax(int64_t M, int ntr, const float2 *c, float *amax, hipStream_t st);
template <typename R> size_t colfft_lds(const ColPass &cp, bool final);
template <typename R>
void launch_colfft_jobs(const ColJob<R> *jobs, const int *blk0s, int njobs, int nblocks, bool final, size_t lds,
                        hipStream_t st, int max_n);

// --- launch helpers (sqw_kernels.hip) ---
// X-ray form-factor weights w[s][Q] on the mode grid N (see sqw_kernels.hip)
void launch_xray_table(const int64_t *N, const double *qs, int nsp, const double *coef, float *w, hipStream_t st);
// mp_sqom's Q tables of a Q plane (q2_norm, x_ffq; see sqw_kernels.hip)
struct QTableParams {
    int n_qx, n_qy, nsp;
    float qx_min, qy_min, dqx, dqy;  // dqx = bz_nx / n_qx (REAL(4) division, as the reference)
    float e1[3], e2[3], q_v[3];
    float e1p[3], e2p[3], q_vp[3];
    float ff[64][9];                 // a1, b1, a2, b2, a3, b3, a4, b4, c per species
};
void launch_qtables(const QTableParams &p, float *q2_norm, float *x_ffq, hipStream_t st);
template <typename R> int sqw_tile_elems();
// dynamic LDS of a k_sqw_single workgroup over T frames and cols Q columns;
// launches keep it <= SQW_MAX_LDS (the kernel's own static LDS is 1 KB)
template <typename R> size_t sqw_single_lds(int64_t T, int cols);
constexpr size_t SQW_MAX_LDS = 150 * 1024;
template <typename R>
void launch_sqw_single(const SqwParams &p, const RadixPlan &rp, int cols, const typename cplx<R>::T *amp,
                       const typename cplx<R>::T *tw, const double *win, const float *qfac, double *S, double *Stot,
                       hipStream_t st);
template <typename R>
void launch_sqw_sliding(const SqwParams &p, const typename cplx<R>::T *amp, const float *qfac, double *S,
                        double *Stot, hipStream_t st);
// one window over all frames (n_frame 1) for a T without a radix plan: the
// window from the table win[n_int] (FP64)
template <typename R>
void launch_sqw_window1(const SqwParams &p, const typename cplx<R>::T *amp, const float *qfac, const double *win,
                        double *S, double *Stot, hipStream_t st);
// DOS: dos[f][s] += species sums of (S(f) + S(-f)) / 2 over columns [c0, c0 + ncols)
void launch_dos_species(const double *S, int64_t ncols, int64_t c0, int64_t n_freq, const int *at_ind, int nsp,
                        double *dos, hipStream_t st);
void launch_cols_to_f64(const float *in, int64_t T, int64_t ncs, int64_t c0, int64_t nc, double *out, hipStream_t st);
// mp_dos spectra in one pass: time FFT of packed velocity columns
// [c0, c0 + nc) (ncs per frame) and per-block species sums part[block][s][f];
// dos_reduce_parts adds nrows such rows into dos[f][s] in a fixed order
// (part2: dos_reduce_groups() * nsp * n_freq doubles)
struct DosParams {
    int64_t T, ncs, c0, nc, n_freq;
    int nsp;
    double norm;
};
template <typename R> size_t dos_fused_lds(int64_t T, int cols);
template <typename R>
void launch_dos_fused(const DosParams &p, const RadixPlan &rp, int cols, const float *vel, const int *at_ind,
                      const typename cplx<R>::T *tw, const double *win, double *part, hipStream_t st);
int dos_reduce_groups();
void dos_reduce_parts(const double *part, int64_t nrows, int nsp, int n_freq, double *part2, double *dos,
                      hipStream_t st);
// I(Q,t): S[l][q] = norm Re sum_{i<n_int} C'(i + m_lo + l) conj C'(i), and S_tot
template <typename R>
void launch_iqt(const SqwParams &p, const typename cplx<R>::T *amp, const float *qfac, double *S, double *Stot,
                hipStream_t st);
// the same sums by FFT correlation in LDS (FP64): cols Q columns per block,
// T * cols <= iqt_fft_tile_elems(), iqt_fft_lds(T, cols) <= SQW_MAX_LDS,
// rp the radix plan of T, tw the FP64 twiddles e^{-2 pi i m / T}
int iqt_fft_tile_elems();
size_t iqt_fft_lds(int64_t T, int cols);
template <typename R>
void launch_iqt_fft(const SqwParams &p, const RadixPlan &rp, int cols, const typename cplx<R>::T *amp,
                    const double2 *tw, const float *qfac, double *S, double *Stot, hipStream_t st);

// One-phonon amplitudes (oneph_kernels.hip)
constexpr int ONEPH_MAX_SP = 64;
struct OnephParams {
    int n_row[3];
    int n_bz, j_plane, nsp, combine;
    int64_t nsuper, n_tot;
    float qx_min, qy_min;
    float twopi_s, sqrt2;   // the reference's REAL(4) 2 pi and sqrt(2)
    int nqx_min, nqy_min;   // zone offset of the first pixel (truncated)
    float base[ONEPH_MAX_SP][3];
};
// frames [f0, f0 + nf): u (nf * nsp * 2 * n1 * n2 doubles), ua and U (as
// many double2) are work arrays; tw1 / tw2: exp(-2 pi i m / n) for n_row(1),
// n_row(2); out per species or combined (p.combine: wsc or wtab); bad |= 4
// when a cell index lies outside the columns
template <typename R>
void launch_oneph(const OnephParams &p, const float *pos, const int *at_ind, int64_t f0, int nf, double *u,
                  double2 *ua, double2 *U, const double2 *tw1, const double2 *tw2, const double *wsc,
                  const float *wtab, void *out, int *bad, hipStream_t st);

}  // namespace mpsq
