// oneph_kernels.hip -- the one-phonon approximation of the S(Q,w) amplitudes
// (j_oneph = 1, mp_sqom56.f90:1037-1098), frame by frame on the device.
//
// For lattice (CELL) data the amplitude of atom type j at Q is taken to first
// order in the displacements, projected on the Q plane (hk0) or (hhl):
//   u_c(k, l)   = sum over the cells of column (k, l) of the displacement
//                 component c (in-plane x, y of the plane), the vertical
//                 index summed out (:1046-1072);
//   U_c         = FFT_2(u_c) / n_row(1)   (forward, unnormalised, :1074-1075);
//   A_j(Q)      = exp(-i 2pi Q . base_j) (q_x U_x(kk, ll) + q_y U_y(kk, ll))
//                 for the pixels of n_bz x n_bz Brillouin zones of n_row
//                 pixels each, (kk, ll) the pixel's position inside the zone
//                 (:1077-1096).
// Quantities the reference keeps in REAL(4) (displacements, Q, the phase
// argument, the zone offset) are formed in FP32 in the same order; the sums,
// the DFT and the products run in FP64.
//
// Kernels:
//   k_oneph_scatter   thread per (frame, atom): u_c(k, l) += d_c (FP64 atomics;
//                     the order of a column's additions is free, so the sums
//                     agree with a sequential sum to rounding)
//   k_oneph_dft       thread per output of one separable DFT pass (exact
//                     twiddle tables, n_row <= 1024 per axis)
//   k_oneph_assemble  thread per (frame, pixel): every species' amplitude,
//                     written per species or combined with scalar or per-Q
//                     weights
#include "mpsq_kernels.h"

namespace mpsq {

// u: [nf][nsp][2][n1 * n2] FP64, column (k, l) at k + n1 * l
__global__ void k_oneph_scatter(OnephParams p, const float4 *__restrict__ pos, const int4 *__restrict__ at_ind,
                                int64_t f0, int nf, double *__restrict__ u, int *__restrict__ bad)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t natoms = (int64_t)p.nsp * p.nsuper;
    if (gid >= (int64_t)nf * natoms) return;
    const int64_t fl = gid / natoms, a = gid - fl * natoms;
    const int j = (int)(a / p.nsuper);
    const float4 r = pos[(f0 + fl) * p.n_tot + a];
    if (r.x == 0.0f && r.y == 0.0f && r.z == 0.0f) return;  // vacancy (:1050, :1062)
    const int4 ind = at_ind[a];
    const int n1 = p.n_row[0], n2 = p.n_row[1], n3 = p.n_row[2];
    int k, l;
    float dx, dy;
    if (p.j_plane == 1) {
        // (hk0): d = pos - base - ind + n_row / 2 (:1051-1055)
        k = ind.x;
        l = ind.y;
        dx = __fadd_rn(__fsub_rn(__fsub_rn(r.x, p.base[j][0]), (float)ind.x), (float)(n1 / 2));
        dy = __fadd_rn(__fsub_rn(__fsub_rn(r.y, p.base[j][1]), (float)ind.y), (float)(n2 / 2));
    } else {
        // (hhl): column along [1 1 0] (:1061-1069)
        const int m = (n1 + n2) / 2;
        int kk = (ind.x - 1 + ind.y - 1) % m;
        if (kk < 0) kk += m;
        k = kk + 1;
        l = ind.z;
        float d1 = __fadd_rn(__fadd_rn(__fsub_rn(__fsub_rn(r.x, p.base[j][0]), (float)ind.x), (float)(n1 / 2)), 1.0f);
        d1 = __fadd_rn(d1, __fadd_rn(__fadd_rn(__fsub_rn(__fsub_rn(r.y, p.base[j][1]), (float)ind.y), (float)(n2 / 2)),
                                     1.0f));
        dx = __fdiv_rn(d1, p.sqrt2);
        dy = __fadd_rn(__fadd_rn(__fsub_rn(__fsub_rn(r.z, p.base[j][2]), (float)ind.z), (float)(n3 / 2)), 1.0f);
    }
    if (k < 1 || k > n1 || l < 1 || l > n2) {
        atomicOr(bad, 4);  // a cell index outside the (n_row(1), n_row(2)) columns
        return;
    }
    const int64_t npl = (int64_t)n1 * n2;
    double *uj = u + ((int64_t)fl * p.nsp + j) * 2 * npl + (k - 1) + (int64_t)n1 * (l - 1);
    atomicAdd(uj, (double)dx);
    atomicAdd(uj + npl, (double)dy);
}

// One separable DFT pass over nb batches of an n_a x n_b grid (a fastest):
//   axis 0: out[b][kk][l] = sum_k in[b][k][l] tw[(k kk) mod n_a]   (in real or complex)
//   axis 1: out[b][kk][ll] = scale sum_l in[b][kk][l] tw[(l ll) mod n_b]
// tw: exp(-2 pi i m / n) for the pass's length n.
template <bool REAL_IN>
__global__ void k_oneph_dft(int64_t nb, int n1, int n2, int axis, const void *__restrict__ in,
                            const double2 *__restrict__ tw, double scale, double2 *__restrict__ out)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t npl = (int64_t)n1 * n2;
    if (gid >= nb * npl) return;
    const int64_t b = gid / npl;
    const int rem = (int)(gid - b * npl);
    const int x = rem % n1, y = rem / n1;  // output (x, y)
    const int n = axis == 0 ? n1 : n2;
    const int o = axis == 0 ? x : y;  // output frequency along the axis
    double sr = 0.0, si = 0.0;
    int m = 0;  // (i * o) mod n
    for (int i = 0; i < n; ++i) {
        const int64_t src = b * npl + (axis == 0 ? i + (int64_t)n1 * y : x + (int64_t)n1 * i);
        const double2 t = tw[m];
        if (REAL_IN) {
            const double v = static_cast<const double *>(in)[src];
            sr += v * t.x;
            si += v * t.y;
        } else {
            const double2 v = static_cast<const double2 *>(in)[src];
            sr += v.x * t.x - v.y * t.y;
            si += v.x * t.y + v.y * t.x;
        }
        m += o;
        if (m >= n) m -= n;
    }
    out[gid] = make_double2(sr * scale, si * scale);
}

// U: [nf][nsp][2][n1 * n2] complex FP64; out per species [T][nsp][NQ] or
// combined [T][NQ] (weights wsc[j] or wtab[j][q]); NQ = (n_bz n1) x (n_bz n2)
template <typename R>
__global__ void k_oneph_assemble(OnephParams p, const double2 *__restrict__ U, int64_t f0, int nf,
                                 const double *__restrict__ wsc, const float *__restrict__ wtab,
                                 typename cplx<R>::T *__restrict__ out)
{
    const int n1 = p.n_row[0], n2 = p.n_row[1];
    const int nqx = p.n_bz * n1, nqy = p.n_bz * n2;
    const int64_t NQ = (int64_t)nqx * nqy;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t)nf * NQ) return;
    const int64_t fl = gid / NQ, q = gid - fl * NQ;
    const int X = (int)(q % nqx), Y = (int)(q / nqx);
    const int ii = X / n1, k = X - ii * n1;  // zone, pixel in zone (0-based)
    const int jj = Y / n2, l = Y - jj * n2;
    // q = q_min + zone + (pixel - 1) / n_row (REAL(4), :1081-1082)
    const float qx = __fadd_rn(__fadd_rn(p.qx_min, (float)ii), __fdiv_rn((float)k, (float)n1));
    const float qy = __fadd_rn(__fadd_rn(p.qy_min, (float)jj), __fdiv_rn((float)l, (float)n2));
    int kk = (p.nqx_min + k) % n1, ll = (p.nqy_min + l) % n2;  // (:1083-1086)
    if (kk < 0) kk += n1;
    if (ll < 0) ll += n2;
    const int64_t npl = (int64_t)n1 * n2;
    double sr = 0.0, si = 0.0;
    for (int j = 0; j < p.nsp; ++j) {
        float arg;
        if (p.j_plane == 1) arg = __fadd_rn(__fmul_rn(qx, p.base[j][0]), __fmul_rn(qy, p.base[j][1]));
        else arg = __fadd_rn(__fmul_rn(qx, __fadd_rn(p.base[j][0], p.base[j][1])), __fmul_rn(qy, p.base[j][2]));
        arg = __fmul_rn(p.twopi_s, arg);
        double sn, cs;
        sincos((double)arg, &sn, &cs);  // exp(-i arg) = cs - i sn
        const double2 *Uj = U + ((int64_t)fl * p.nsp + j) * 2 * npl;
        const double2 ux = Uj[kk + (int64_t)n1 * ll], uy = Uj[npl + kk + (int64_t)n1 * ll];
        const double ar = (double)qx * ux.x + (double)qy * uy.x, ai = (double)qx * ux.y + (double)qy * uy.y;
        const double vr = cs * ar + sn * ai, vi = cs * ai - sn * ar;
        if (!p.combine) {
            typename cplx<R>::T v;
            v.x = (R)vr;
            v.y = (R)vi;
            out[((f0 + fl) * p.nsp + j) * NQ + q] = v;
        } else {
            const double w = wtab ? (double)wtab[(int64_t)j * NQ + q] : wsc[j];
            sr += w * vr;
            si += w * vi;
        }
    }
    if (p.combine) {
        typename cplx<R>::T v;
        v.x = (R)sr;
        v.y = (R)si;
        out[(f0 + fl) * NQ + q] = v;
    }
}

__global__ void k_oneph_zero(double *__restrict__ u, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] = 0.0;
}

template <typename R>
void launch_oneph(const OnephParams &p, const float *pos, const int *at_ind, int64_t f0, int nf, double *u,
                  double2 *ua, double2 *U, const double2 *tw1, const double2 *tw2, const double *wsc,
                  const float *wtab, void *out, int *bad, hipStream_t st)
{
    const int64_t npl = (int64_t)p.n_row[0] * p.n_row[1];
    const int64_t nb = (int64_t)nf * p.nsp * 2;
    hipLaunchKernelGGL(k_oneph_zero, dim3((unsigned)((nb * npl + 255) / 256)), dim3(256), 0, st, u, nb * npl);
    const int64_t na = (int64_t)nf * p.nsp * p.nsuper;
    if (na > 0)
        hipLaunchKernelGGL(k_oneph_scatter, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, st, p,
                           reinterpret_cast<const float4 *>(pos), reinterpret_cast<const int4 *>(at_ind), f0, nf, u,
                           bad);
    const unsigned gb = (unsigned)((nb * npl + 255) / 256);
    hipLaunchKernelGGL((k_oneph_dft<true>), dim3(gb), dim3(256), 0, st, nb, p.n_row[0], p.n_row[1], 0,
                       static_cast<const void *>(u), tw1, 1.0, ua);
    hipLaunchKernelGGL((k_oneph_dft<false>), dim3(gb), dim3(256), 0, st, nb, p.n_row[0], p.n_row[1], 1,
                       static_cast<const void *>(ua), tw2, 1.0 / (double)p.n_row[0], U);
    const int64_t NQ = (int64_t)p.n_bz * p.n_row[0] * p.n_bz * p.n_row[1];
    hipLaunchKernelGGL((k_oneph_assemble<R>), dim3((unsigned)(((int64_t)nf * NQ + 255) / 256)), dim3(256), 0, st, p,
                       U, f0, nf, wsc, wtab, static_cast<typename cplx<R>::T *>(out));
}

template void launch_oneph<float>(const OnephParams &, const float *, const int *, int64_t, int, double *, double2 *,
                                  double2 *, const double2 *, const double2 *, const double *, const float *, void *,
                                  int *, hipStream_t);
template void launch_oneph<double>(const OnephParams &, const float *, const int *, int64_t, int, double *, double2 *,
                                   double2 *, const double2 *, const double2 *, const double *, const float *, void *,
                                   int *, hipStream_t);

}  // namespace mpsq
