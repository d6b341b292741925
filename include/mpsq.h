/*
 * mpsq.h -- C ABI of the MI355X S(Q,w) engine (libmpsq.so).
 *
 * Two layers:
 *   1. A FINUFFT-shaped type-1 NUFFT (plan / setpts / execute / destroy, FP64
 *      and FP32 "f" variants, 1-3 dimensions, ntrans strength vectors), the
 *      drop-in for the reference's calls into FINUFFT, plus Fortran-ABI
 *      shims with the reference's by-reference argument list.
 *   2. A device-resident S(Q,w) engine: per-frame, per-species type-1 NUFFT
 *      of staged trajectory positions, species combine, and the Hann-windowed
 *      time-FFT reduction to S(Q) and S(Q,w).
 *
 * Return codes follow FINUFFT: 0 = ok, 1 = warning (eps below the precision
 * floor of the plan; the result is still computed at the floor), >= 2 =
 * error, in which case every output array is zero-filled.  No entry point
 * prints, exits or aborts.
 *
 * Complex arrays are interleaved (re, im) pairs: double[2] per value for the
 * FP64 API, float[2] for the FP32 API.  Mode order is CMCL: k1 fastest,
 * k_d = -N_d/2 .. (N_d-1)/2.
 *
 * Citations are to the jkulda/MP_tools tree.
 */
#ifndef MPSQ_H
#define MPSQ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MPSQ_OK = 0,
    MPSQ_WARN_EPS_TOO_SMALL = 1,
    MPSQ_ERR_BAD_ARG = 2,
    MPSQ_ERR_NONFINITE_POINT = 3,
    MPSQ_ERR_SIZE = 4,         /* a fine-grid dimension exceeds the LDS tile */
    MPSQ_ERR_DEVICE = 5,       /* HIP runtime / allocation failure */
    MPSQ_ERR_NOT_SET = 6       /* execute before setpts */
};

enum { MPSQ_FP32 = 0, MPSQ_FP64 = 1 };

typedef struct mpsq_opts {
    int device;        /* HIP device ordinal (default 0) */
    void *stream;      /* hipStream_t used by every launch of the plan, NULL = own stream */
    int tile_by;       /* spread tile rows along y (0 = tuned/heuristic) */
    int tile_bz;       /* spread tile rows along z (0 = tuned/heuristic) */
    int sort_chunk;    /* points per sort chunk (0 = heuristic) */
    int debug;         /* reserved, 0 */
} mpsq_opts;

/* Defaults: device 0, own stream, heuristic tiles. */
void mpsq_default_opts(mpsq_opts *o);

/* ---------------- FINUFFT-shaped plan API, FP64 ---------------- */
typedef struct mpsq_plan_s *mpsq_plan;

/* Replaces FINUFFT's finufft_makeplan (type 1 only); the reference calls the
 * simple interface finufft2d1 at mp_sqom56.f90:1148 and mp_sql56.f90:1129,
 * which is this plan's makeplan + setpts + execute + destroy. */
int mpsq_makeplan(int type, int dim, const int64_t *n_modes, int iflag, int ntrans, double eps,
                  mpsq_plan *plan, const mpsq_opts *opts);
/* Copies M points (host arrays; y/z ignored below dim 2/3) to the device,
 * folds them in FP64 and bin-sorts them.  Any finite coordinate is accepted
 * and folded periodically with period 2 pi. */
int mpsq_setpts(mpsq_plan plan, int64_t M, const double *x, const double *y, const double *z);
/* c: ntrans*M complex strengths (host), f: ntrans*N1*N2*N3 complex modes (host). */
int mpsq_execute(mpsq_plan plan, const double *c, double *f);
/* Device-resident variants: every pointer is device memory, work is queued on
 * the plan's stream and the call returns without synchronising. */
int mpsq_setpts_dev(mpsq_plan plan, int64_t M, const double *d_x, const double *d_y, const double *d_z);
int mpsq_execute_dev(mpsq_plan plan, const double *d_c, double *d_f);
int mpsq_destroy(mpsq_plan plan);
/* Bin-sort permutation of the last setpts (perm[k] = caller index of the
 * k-th point in bin order), copied to host memory of M int32. */
int mpsq_get_sort_perm(mpsq_plan plan, int32_t *perm);
/* Fine grid size, kernel width and tile geometry chosen by makeplan:
 * out[0..2] = n1,n2,n3, out[3] = w, out[4] = by, out[5] = bz, out[6] = BY,
 * out[7] = BZ. */
int mpsq_plan_info(mpsq_plan plan, int64_t *out);

/* ---------------- FINUFFT-shaped plan API, FP32 ---------------- */
typedef struct mpsq_planf_s *mpsq_planf;
int mpsqf_makeplan(int type, int dim, const int64_t *n_modes, int iflag, int ntrans, double eps,
                   mpsq_planf *plan, const mpsq_opts *opts);
int mpsqf_setpts(mpsq_planf plan, int64_t M, const float *x, const float *y, const float *z);
int mpsqf_execute(mpsq_planf plan, const float *c, float *f);
int mpsqf_setpts_dev(mpsq_planf plan, int64_t M, const float *d_x, const float *d_y, const float *d_z);
int mpsqf_execute_dev(mpsq_planf plan, const float *d_c, float *d_f);
int mpsqf_destroy(mpsq_planf plan);
int mpsqf_get_sort_perm(mpsq_planf plan, int32_t *perm);
int mpsqf_plan_info(mpsq_planf plan, int64_t *out);

/* ---------------- simple interface (FINUFFT's finufft{1,2,3}d1) ---------------- */
int mpsq_nufft1d1(int64_t nj, const double *xj, const double *cj, int iflag, double eps, int64_t ms,
                  double *fk, const mpsq_opts *opts);
int mpsq_nufft2d1(int64_t nj, const double *xj, const double *yj, const double *cj, int iflag, double eps,
                  int64_t ms, int64_t mt, double *fk, const mpsq_opts *opts);
int mpsq_nufft3d1(int64_t nj, const double *xj, const double *yj, const double *zj, const double *cj, int iflag,
                  double eps, int64_t ms, int64_t mt, int64_t mu, double *fk, const mpsq_opts *opts);

/* ---------------- Fortran ABI (gfortran/flang name mangling, by reference) ----------------
 * Exactly the argument list of the reference's implicit-interface call
 *   call finufft2d1(n_fft,xf,yf,cf,iflag,eps_fft,n_qx8,n_qy8,ampl_tot,nul_opt,ier)
 * (mp_sqom56.f90:1148; declarations :94-98,106,195-196; mp_sql56.f90:1129).
 * opts may be a null pointer (the reference passes an unallocated allocatable);
 * when non-null it points to an mpsq_opts.  fk is always fully overwritten. */
void finufft1d1_(const int64_t *nj, const double *xj, const double *cj, const int *iflag, const double *eps,
                 const int64_t *ms, double *fk, const void *opts, int *ier);
void finufft2d1_(const int64_t *nj, const double *xj, const double *yj, const double *cj, const int *iflag,
                 const double *eps, const int64_t *ms, const int64_t *mt, double *fk, const void *opts, int *ier);
void finufft3d1_(const int64_t *nj, const double *xj, const double *yj, const double *zj, const double *cj,
                 const int *iflag, const double *eps, const int64_t *ms, const int64_t *mt, const int64_t *mu,
                 double *fk, const void *opts, int *ier);
/* Fortran-callable plan handle wrappers (handle = int64). */
void mpsq_makeplan_(const int *type, const int *dim, const int64_t *n_modes, const int *iflag, const int *ntrans,
                    const double *eps, int64_t *handle, const void *opts, int *ier);
void mpsq_setpts_(const int64_t *handle, const int64_t *M, const double *x, const double *y, const double *z,
                  int *ier);
void mpsq_execute_(const int64_t *handle, const double *c, double *f, int *ier);
void mpsq_destroy_(const int64_t *handle, int *ier);

/* ---------------- S(Q,w) engine ----------------
 * Trajectory frames are device-resident FP32 records (x, y, z, charge) in the
 * MP_TOOLS binary layout (at_pos(4,n_tot), code/docs/1_mp_data_doc_154.txt:
 * 100-104), frame-major: pos[f][atom][4], atoms species-major (CELL record
 * order mp_lbin56.f90:1204; BULK species blocks mp_lbin56.f90:1336-1347).
 * Per frame and active species the engine stages the NUFFT coordinates and
 * strengths (mp_sqom56.f90:1104-1137), runs a type-1 NUFFT
 * (mp_sqom56.f90:1148), and combines the species with their weights
 * (mp_sqom56.f90:1338-1356) into C(t,Q), layout [T][NQ]. */
typedef struct mpsq_sqw_s *mpsq_sqw;

typedef struct mpsq_sqw_spec {
    int dim;                  /* 2 (mp_sqom Q plane) or 3 */
    int64_t n_modes[3];       /* N1, N2, N3 (unused = 1) */
    double eps;               /* NUFFT tolerance (reference: 1e-6, mp_sqom56.f90:195) */
    int precision;            /* MPSQ_FP32 or MPSQ_FP64 */
    int iflag;                /* sign of the exponent (reference: -1, :196) */
    int64_t n_tot;            /* atoms per frame */
    int nsp;                  /* species */
    const int64_t *sp_count;  /* [nsp] atoms of each species (species-major) */
    const int32_t *mask;      /* [nsp] 0/1 or NULL = all (at_mask, :1345) */
    const double *w_scalar;   /* [nsp] weights (unit / 0.1*b_coh, :1347-1350) or NULL = 1 */
    const float *w_table;     /* [nsp][NQ] per-Q weights (x_ffq, :1351-1352) host, or NULL */
    int combine;              /* 1: C(t,Q) = sum_j w_j F_j; 0: per-species F_j, layout [T][nact][NQ] */
    double centre[3];         /* at_pos_centre */
    double axes[9];           /* unit plane vectors e1, e2, e3 (rows) (:799-807) */
    double tpe[3];            /* 2 pi / (e_d . (pmax - pmin)) (:809-811) */
    double q0[3];             /* q_center: strength phase exp(-i 2 pi q0 . r) (:1117) */
    int window;               /* spatial Hann window (j_wind, :1118, :1167) */
    int vacancy_check;        /* drop atoms at (0,0,0) (:1106) */
    int select_mode;          /* 0 none, 1 CELL window on at_ind (:1107-1113), 2 BULK box (:1123-1132) */
    int32_t n_row[3];
    float pmin[3], pmax[3];   /* at_pos_min / at_pos_max */
    float margin;             /* BULK margin (eps_x = 0.1 for one box, else 1) */
    const int32_t *at_ind;    /* host [n_tot][4] cell indices for select_mode 1, or NULL */
    int batch_frames;         /* frames per launch batch (0 = auto: <= 128 frames, <= 16 GiB of intermediates) */
    mpsq_opts opts;
} mpsq_sqw_spec;

void mpsq_sqw_default_spec(mpsq_sqw_spec *s);
int mpsq_sqw_create(const mpsq_sqw_spec *spec, mpsq_sqw *eng);
/* d_pos: device float[T][n_tot][4]; d_amp: device complex [T][NQ] (combine)
 * or [T][nact][NQ], FP32 or FP64 per the spec's precision. */
int mpsq_sqw_amplitudes(mpsq_sqw eng, const float *d_pos, int64_t T, void *d_amp);

typedef struct mpsq_reduce_spec {
    int64_t n_int;            /* window length (frames); n_frame = T - n_int + 1 windows, hop 1 */
    int win_kind;             /* 0: 0.5(1-cos(2 pi i/(n_int-1))) (:941-948); 1: 0.5(1-cos(2 pi (i+1)/(n_int+1))) (:2429-2431) */
    double dt;                /* frame step t_step (ps) */
    double nat;               /* N_at = sum of nsuper_r (:1574-1576) */
    int64_t m_lo;             /* first frequency index (may be negative) */
    int64_t n_f;              /* number of frequency bins written */
    int subtract_mean;        /* 1: inelastic C' = C - <C>_t (:1358-1377) */
    const float *d_qfac;      /* device [NQ] factor on C' (q2_norm when j_qsq != 1), or NULL */
    int kind;                 /* 0: S(Q,w), frequencies m_lo .. m_lo+n_f-1 (E = const maps :1437-1478 and
                                 E(Q) cuts :1480-1517); 1, 2: the intermediate function I(Q,t) (j_sq != 1),
                                 d_S[l][Q] = Re sum_{i<n_int} C'(i+lag, Q) conj C'(i, Q) for lag = m_lo + l,
                                 0 <= lag <= T - n_int, times 1/n_int for kind 1 (the map, :1467-1471) and
                                 unscaled for kind 2 (the cut, :1504-1510), over nat^2 (:1574-1575) */
    int odd_n_int;            /* 1: the reference's rule (:683-686): n_int = 0 means T/2, and for S(Q,w) an
                                 even n_int grows by one */
    const int32_t *d_qidx;    /* device [n_qidx] mode indices to reduce (the pixels of an E(Q) cut), or NULL
                                 for every mode; d_S is then [n_f][n_qidx], d_Stot [n_qidx], d_qfac per mode;
                                 an index outside [0, NQ) gives NaN outputs in its column */
    int64_t n_qidx;
} mpsq_reduce_spec;

/* The one-phonon approximation of the amplitudes (j_oneph = 1,
 * mp_sqom56.f90:1037-1098) for lattice (CELL) data: per frame and atom type
 * j, the displacements d = pos - base_j - at_ind + n_row / 2 summed over the
 * vertical cell index into the columns u_x, u_y(k, l) of the Q plane ((hk0):
 * (k, l) = the cell's (a, b) index, d its (x, y); (hhl): k = (a + b - 2) mod
 * (n_row(1) + n_row(2)) / 2 + 1, l = c, d_x = (d_a + d_b + 2) / sqrt 2, d_y =
 * d_c + 1), U = 2-D forward FFT(u) / n_row(1), and for the pixels of n_bz x
 * n_bz zones (x fastest: n_qx = n_bz n_row(1) pixels of 1 / n_row(1)
 * r.l.u. from qx_min, n_qy likewise)
 *   A_j(Q) = exp(-i 2 pi Q . base_j) (q_x U_x(kk, ll) + q_y U_y(kk, ll)),
 * (kk, ll) = the pixel's position inside its zone shifted by the first
 * pixel's (:1083-1086); (hhl) phases use (base_a + base_b, base_c).  Atoms at
 * (0, 0, 0) are vacancies.  The reference's REAL(4) steps (displacements, Q,
 * phase argument, zone offset) are kept; sums, the DFT and the products are
 * FP64.  n_row(1) = n_row(2) (and = n_row(3) for (hhl)), as the reference's
 * loops assume; a cell index outside the columns is MPSQ_ERR_BAD_ARG. */
typedef struct mpsq_oneph_spec {
    int n_row[3];             /* cells along a, b, c */
    int n_bz;                 /* Brillouin zones per axis */
    int j_plane;              /* 1: (hk0), 2: (hhl) */
    int nsp;                  /* atom types: d_pos holds nsp blocks of n_row(1) n_row(2) n_row(3) atoms */
    float qx_min, qy_min;     /* first pixel of the map (r.l.u.) */
    const float *at_base;     /* host [nsp][3] fractional basis positions */
    int precision;            /* MPSQ_FP32: complex float output, MPSQ_FP64: complex double */
} mpsq_oneph_spec;
/* d_pos: device float [T][n_tot][4] (CELL positions in lattice units, as the
 * engine takes them); d_at_ind: device int32 [n_tot][4] (1-based cell indices
 * and atom type); d_amp: device [T][nsp][NQ] (wsc and d_wtab NULL) or the
 * species combined [T][NQ] with host weights wsc[nsp] or device per-Q weights
 * d_wtab[nsp][NQ] (mp_sqom56.f90:1338-1356).  Synchronises. */
int mpsq_oneph_amplitudes(const mpsq_oneph_spec *spec, const float *d_pos, const int32_t *d_at_ind, int64_t T,
                          const double *wsc, const float *d_wtab, void *d_amp, void *stream);

/* Vibrational spectra of mp_dos (mp_dos56.f90:495-527): for every atom and
 * velocity component the Hann-windowed forward FFT over the T frames
 * (window 0.5 (1 - cos(2 pi (t + 1) / (T + 1))), t = 0 .. T-1, :495-498;
 * no mean subtraction), and per atom type s and frequency f < n_freq
 *   d_dos[f][s] = sum over atoms of type s of sum_c |FFT(v_c w)(f)|^2 / T
 * (cs_plot of :519-525, type = at_ind .w of the atom, 1-based).  d_vel:
 * device float [T][n_tot][4] (v_x, v_y, v_z, unused), d_at_ind: device int32
 * [n_tot][4]; n_freq <= T / 2 + 1 (the reference's n_freq_plot, int(0.5 (T -
 * 1) t_step / t_step) + 1 in REAL(4)).  MPSQ_FP32 computes the FFT in FP32
 * (the reference's own precision), MPSQ_FP64 in FP64; the sums are FP64 in a
 * fixed order.  Synchronises. */
int mpsq_dos(const float *d_vel, int64_t T, int64_t n_tot, const int32_t *d_at_ind, int nsp, int64_t n_freq,
             int precision, double *d_dos, void *stream);

/* The reference's per-Q tables of a Q plane (mp_sqom56.f90:909-936),
 * computed on the device in its REAL(4) arithmetic: for pixel (i, j),
 * i = 0..n_qx-1 fastest (the CMCL order of the plane's amplitudes, k1 =
 * i - n_qx/2), q1 = (qx_min + i * (bz_nx / n_qx)) e1, q2 = (qy_min + j *
 * (bz_ny / n_qy)) e2 and
 *   q2_norm = 1 / sqrt(|q1 + q2|^2 + |q_v|^2)  (1 where that is below 1e-8),
 *             the factor on C' when S(Q)/Q^2 scaling is off (j_qsq != 1);
 *   x_ffq[s] = c + sum_k a_k exp(-0.25 b_k q_sq'),  q_sq' = 0.25 (|q1' + q2'|^2
 *             + |q_vp|^2) with the same pixel coefficients on e1p, e2p (1/A),
 *             the X-ray weights of species s (j_weight = 3).
 * x_ffpar: host [nsp][9] coefficients as the reference reads them
 * (a1, b1, a2, b2, a3, b3, a4, b4, c; mp_sqom56.f90:495); nsp <= 64.
 * d_q2_norm: device float[n_qx * n_qy] or NULL; d_x_ffq: device
 * float[nsp][n_qx * n_qy] or NULL.  Enqueued on stream (NULL: the default
 * stream). */
typedef struct mpsq_qplane {
    int32_t n_qx, n_qy;
    float qx_min, qy_min;      /* dot(e1, q_center)/|e1| - bz_nx/2 (and y), :895-906 */
    float bz_nx, bz_ny;        /* Brillouin-zone extents of the plane */
    float e1[3], e2[3];        /* normalised plane vectors (r.l.u.) */
    float q_v[3];              /* (dot(ev, q_center)/|ev|) ev (:911) */
    float e1p[3], e2p[3];      /* plane vectors in 1/A (:812-820) */
    float q_vp[3];             /* (dot(evp, q_center)/|evp|) evp (:920) */
} mpsq_qplane;
int mpsq_qtables(const mpsq_qplane *qp, int nsp, const float *x_ffpar, float *d_q2_norm, float *d_x_ffq,
                 void *stream);

/* The speckle filter (mp_sqom56.f90:1530-1570) on a map in host memory,
 * in place: one sequential pass over the n_rows x n_cols row-major array,
 * walking it row-major (column_walk 0) or column-major (1), the neighbours
 * of element k being k - 1 and k + 1 (cyclic) as the walk has left them: a
 * value below min(neighbours) / s_trig becomes that minimum, a value above
 * s_trig * max(neighbours) that maximum.  s_trig <= 0 leaves the map as it
 * is.  The reference filters an E = const map twice (x fastest, then y
 * fastest: column_walk 0 then 1 on a [n_qy][n_qx] map) and an E(Q) cut of
 * all n_freq_max frequencies once, frequency fastest (column_walk 1 on
 * d_S[n_f][n_q]); it does so before the 1/nat^2 norm, which does not change
 * the filter's decisions. */
int mpsq_speckle_filter(double *map, int64_t n_rows, int64_t n_cols, int column_walk, double s_trig);

/* d_S: device double[n_f][NQ]; d_Stot: device double[NQ] or NULL. */
int mpsq_sqw_reduce(mpsq_sqw eng, const void *d_amp, int64_t T, const mpsq_reduce_spec *r, double *d_S,
                    double *d_Stot);
int mpsq_sqw_destroy(mpsq_sqw eng);
/* Inspection of the engine's staged, bin-sorted points (a1-a3, a4.1), for
 * bit-exact parity checks of the selection, fold and sort: stages and sorts
 * frame `frame` of d_pos alone and writes, per sorted position k < n (n must
 * equal the active atoms per frame; transforms in order, one per frame when
 * the species are merged into one transform, else one per active species):
 * atom[k] = the atom's index in the frame (mp_sqom56.f90:1104-1137 keeps the
 * species-major order; the sort is stable), cell[3k..3k+2] = its first
 * kernel cell on the fine grid, keep[k] = 1 if it passed the vacancy and
 * window selection (mp_sqom56.f90:1106-1113,1123-1132), else 0 (its strength
 * is then 0).  Host arrays; any may be NULL.  Synchronises. */
int mpsq_sqw_debug_sorted(mpsq_sqw eng, const float *d_pos, int64_t frame, int32_t *atom, int32_t *cell,
                          int32_t *keep, int64_t n);
/* Fine grid, kernel width, spread tile and sort-bin shape of an engine, in
 * the layout of mpsq_plan_info. */
int mpsq_sqw_info(mpsq_sqw eng, int64_t *out);
/* Synchronise the engine's stream. */
int mpsq_sqw_sync(mpsq_sqw eng);
/* Per-stage timing with HIP events recorded on the engine's stream around
 * every launch group.  Stages: 0 staging, 1 bin sort, 2 spread + x FFT,
 * 3 non-final column FFT pass, 4 final column pass (deconvolve + combine),
 * 5 S(Q,w) reduction.  set_timing(1) clears the accumulators;
 * get_timing fills ms[k] (summed milliseconds) and launches[k] for k < n. */
int mpsq_sqw_set_timing(mpsq_sqw eng, int on);
int mpsq_sqw_get_timing(mpsq_sqw eng, double *ms, int64_t *launches, int n);
/* Number of modes N1*N2*N3 and active species of an engine. */
int64_t mpsq_sqw_nq(mpsq_sqw eng);
int mpsq_sqw_nact(mpsq_sqw eng);

/* ---------------- batches of S(Q,w) jobs ----------------
 * No counterpart in the reference, which computes one map per dialogue
 * session.  A batch runs many uneven jobs (different supercells, species,
 * frame counts and mode boxes) through one launch per phase: staging, bin
 * sort, one persistent spread + x-FFT launch over every tile of every job
 * (one per kernel width present), the column passes; the per-job reductions
 * follow on the same stream.  Every spec must have combine = 1 and the same
 * precision; specs[0].opts gives the device and the stream. */
typedef struct mpsq_batch_s *mpsq_batch;
int mpsq_batch_create(int njobs, const mpsq_sqw_spec *specs, mpsq_batch *batch);
/* d_pos[j]: device float[T[j]][n_tot_j][4]; d_amp[j]: device complex [T[j]][NQ_j];
 * r: njobs reduction specs or NULL (amplitudes only); d_S[j], d_Stot[j] as in
 * mpsq_sqw_reduce (d_S / d_Stot may be NULL).  use_graph = 1: the second call
 * with the same arguments captures the whole sequence in a hipGraph, later
 * calls replay it.  Returns after the stream has drained. */
int mpsq_batch_run(mpsq_batch b, const float *const *d_pos, const int64_t *T, void *const *d_amp,
                   const mpsq_reduce_spec *r, double *const *d_S, double *const *d_Stot, int use_graph);
/* Device bytes of intermediates per launch round (default 16 GiB): a batch
 * larger than this runs in several rounds. */
int mpsq_batch_set_budget(mpsq_batch b, int64_t bytes);
int mpsq_batch_set_timing(mpsq_batch b, int on);
int mpsq_batch_get_timing(mpsq_batch b, double *ms, int64_t *launches, int n);
int mpsq_batch_destroy(mpsq_batch b);

/* ---------------- spread tile autotuner ----------------
 * No counterpart in the reference.  Times every spread tile shape (by, bz)
 * that fits the spec's geometry on the first frames of d_pos (at most 4) and
 * returns the fastest with its time per frame.  The choice is cached for the
 * process per (dim, fine grid, kernel width, precision); engines and batches
 * created afterwards with opts.tile_by = tile_bz = 0 use it. */
int mpsq_autotune_tiles(const mpsq_sqw_spec *spec, const float *d_pos, int64_t T, int *tile_by, int *tile_bz,
                        double *ms_per_frame);

/* ---------------- S(Q,w) map server ----------------
 * No counterpart in the reference (an interactive, one-map-at-a-time tool).
 * Trajectories registered with the server stay resident on the device; map
 * requests (species set, mode box, frame range, neutron or X-ray weights)
 * are admitted in arrival order while the admitted requests' amplitudes,
 * maps and weight tables fit the HBM budget, and a worker thread runs
 * continuous batching: every step processes the next frames of every
 * admitted request in one launch per phase, a request joins at the next
 * step after admission and leaves once its S(Q,w) reduction has run. */
typedef struct mpsq_server_s *mpsq_server;
typedef struct mpsq_server_opts {
    int device;
    int64_t hbm_budget;    /* bytes for admitted requests' device data (0 = 64 GiB) */
    int64_t step_budget;   /* bytes of intermediates per step (0 = 8 GiB) */
    int max_active;        /* requests admitted at once (0 = 64) */
} mpsq_server_opts;
int mpsq_server_create(const mpsq_server_opts *o, mpsq_server *srv);
/* d_pos: device float[T][n_tot][4], owned by the caller while the server
 * lives; tpe: staging phase scale per axis (2 pi / box length in position
 * units); qscale: inverse Angstrom per mode index along each axis (2 pi /
 * box length in Angstrom), used by X-ray form factors. */
int mpsq_server_add_trajectory(mpsq_server srv, const float *d_pos, int64_t T, int nsp, const int64_t *sp_count,
                               const double *tpe, const double *qscale, int *traj_id);
typedef struct mpsq_request {
    int traj;
    const int32_t *mask;      /* [nsp] 0/1 species set, or NULL = all */
    int64_t n_modes[3];
    int64_t f0, T;            /* frames [f0, f0 + T); one Hann window over the T frames */
    double eps;
    int weight_kind;          /* 0: scalar weights w_scalar (neutron 0.1 b); 1: X-ray form factors xff */
    const double *w_scalar;   /* [nsp] */
    const double *xff;        /* [nsp][9]: a1..a4, b1..b4, c of c + sum a_k exp(-b_k |Q|^2 / 16 pi^2) */
    double dt;                /* frame step (ps) */
    int64_t n_f;              /* frequencies m = 0 .. n_f-1 of S(Q,w) */
} mpsq_request;
/* Copies the request; never blocks on the device. */
int mpsq_server_submit(mpsq_server srv, const mpsq_request *r, int64_t *ticket);
/* state: 0 queued, 1 admitted, 2 done, or a negative error code; queue_ms:
 * submit to admission; latency_ms: submit to completion (done only). */
int mpsq_server_poll(mpsq_server srv, int64_t ticket, int *state, double *queue_ms, double *latency_ms);
/* Device S [n_f][NQ] and S_tot [NQ] of a done request, valid until release. */
int mpsq_server_result(mpsq_server srv, int64_t ticket, const double **d_S, const double **d_Stot, int64_t *nq);
/* Copy a done request's S [n_f][NQ] and S_tot [NQ] into caller device buffers (either may be NULL). */
int mpsq_server_copy_result(mpsq_server srv, int64_t ticket, double *d_S, double *d_Stot);
int mpsq_server_release(mpsq_server srv, int64_t ticket);
/* Block until every submitted request is done or failed. */
int mpsq_server_wait_idle(mpsq_server srv);
/* steps run, requests done, summed step launch time on the host (ms) */
int mpsq_server_stats(mpsq_server srv, int64_t *steps, int64_t *done, double *busy_ms);
int mpsq_server_destroy(mpsq_server srv);

/* ---------------- mp_sql: spherically averaged partial intensities ----------------
 * Replaces the Gauss integration loop of mp_sql (code/src/mp_sql56.f90:1224-1254;
 * q_norm table :1036-1044): for every mode k of the nsp amplitude grids
 * d_amp ([nsp][N3][N2][N1] complex, FP32 or FP64 by precision, CMCL order,
 * k_d = i_d - N_d/2), q = |a_cell_inv . (k * dq)| in FP32 and, for q <= q_range,
 * d_hist[nint(q / q_step)][p] += Re(A_i A_j*) for the pairs p = (i >= j) in the
 * order (0,0), (1,0), (1,1), (2,0), ... (nsp (nsp + 1) / 2 of them).  d_hist is
 * device double [n_hist][npairs], accumulated (not zeroed); a_cell_inv row-major
 * (v_i = sum_j a_cell_inv[3 i + j] u_j).  Enqueued on stream (NULL: the default
 * stream); n_hist * npairs * 8 bytes must fit 64 KB (else MPSQ_ERR_SIZE). */
int mpsq_sql_hist(const void *d_amp, int nsp, int precision, const int64_t n_modes[3], const float dq[3],
                  const float a_cell_inv[9], float q_range, float q_step, int n_hist, double *d_hist, void *stream);

/* Replaces the Monte Carlo integration of mp_sql (j_acc = 2,
 * code/src/mp_sql56.f90:1146-1215): n_h chains of n_mc samples (the reference's
 * n_mc is 10^6, :40).  Sample (j, k) draws a point uniformly in [-1, 1)^3 until
 * it lies in the unit ball (0 < |p| <= 1), takes a radius fraction r in [0, 1)
 * and the mode d = nint(h * (p / |p|) * r) with h_d = N_d - (N_d / 2 + 1), all in
 * FP32; d = 0 adds 1 to d_skip[0], |Q(d)| > q_range adds 1 to d_skip[1], else
 * Re(A_i A_j*) at d is added to d_hist[nint(|Q| / q_step)][p] as in
 * mpsq_sql_hist.  The random numbers come from Philox4x32-10 with key
 * (seed & 0xffffffff, seed >> 32) and counter (j, k, attempt, tag), words 0..2
 * for the point and word 3 for r, each as (word >> 8) * 2^-24; a sample is a
 * pure function of (seed, tag, j, k), whatever the launch shape.  (The reference
 * draws from Fortran's random_number and retries without bound; a sample whose
 * first 64 attempts miss the ball counts as d = 0 here.)  d_hist and d_skip
 * (device int64[2]) are accumulated, not zeroed.  MPSQ_ERR_BAD_ARG for n_h < 1,
 * n_mc < 1, n_h > 2^32 - 1, n_mc > 2^31 - 1 or a null d_skip; otherwise the
 * arguments and limits of mpsq_sql_hist. */
int mpsq_sql_hist_mc(const void *d_amp, int nsp, int precision, const int64_t n_modes[3], const float dq[3],
                     const float a_cell_inv[9], float q_range, float q_step, int n_hist, int64_t n_h, int64_t n_mc,
                     uint64_t seed, uint32_t tag, double *d_hist, int64_t *d_skip, void *stream);

/* mpsq_sql_hist without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, result and error convention as mpsq_sql_hist; stream must
 * belong to the current device.  The kernels' partial histograms live in a
 * static device array of 2^20 doubles (8 MiB) that each GPU the library runs
 * on holds until the process ends; a call uses up to 2 x (compute units)
 * slices of n_hist * npairs doubles, and n_hist * npairs above 8192 (the
 * workgroup's LDS histogram) is rejected with MPSQ_ERR_SIZE.  Every call on
 * one GPU shares that array: calls that overlap on one GPU race. */
int mpsq_sql_hist_noalloc(const void *d_amp, int nsp, int precision, const int64_t n_modes[3], const float dq[3],
                          const float a_cell_inv[9], float q_range, float q_step, int n_hist, double *d_hist,
                          void *stream);

/* Library version string. */
const char *mpsq_version(void);

#ifdef __cplusplus
}
#endif

#endif /* MPSQ_H */
