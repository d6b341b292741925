// sql_kernels.hip -- spherical |Q| histogram of partial intensities (the
// "Gauss integration" of mp_sql, code/src/mp_sql56.f90:1224-1254).
//
// For every mode k of nsp amplitude grids A_s[k] (CMCL order, k_d from
// -(N_d/2)), the reference computes |Q| in FP32 as
//     q = norm2( a_cell_inv . ((k - centre) * d_q) )       (:1036-1044)
// and, for q <= q_range, adds Re(A_i A_j*) to bin nint(q / q_step) of the
// (i, j) partial (:1237-1244; the real part because sq_hist is real).  Here
// |Q| and the bin follow the same FP32 arithmetic (elementwise product,
// matrix-vector product, square root of the sum of squares, round half away
// from zero), the products are taken and summed in FP64, and only the pairs
// i >= j are kept (Re(A_i A_j*) is symmetric).
//
// Each workgroup sums its share of the modes into an LDS histogram (FP64
// LDS atomics; consecutive modes of a row fall in different bins except
// near the axes) and writes it to its own slice of a partial buffer; a
// second kernel adds the slices in a fixed order.
#include <algorithm>
#include <mutex>

#include "mpsq_kernels.h"

namespace mpsq {

template <typename C>
__global__ __launch_bounds__(256) void k_sql_hist(const C *__restrict__ amp, SqlGeom g, double *__restrict__ part)
{
    extern __shared__ double hs[];  // [n_hist][npairs]
    const int nb = g.n_hist * g.npairs;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) hs[i] = 0.0;
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < g.nmodes; m += stride) {
        const int i1 = (int)(m % g.N[0]);
        const int64_t r = m / g.N[0];
        const int i2 = (int)(r % g.N[1]);
        const int i3 = (int)(r / g.N[1]);
        const float u0 = (float)(i1 - g.N[0] / 2) * g.dq[0];
        const float u1 = (float)(i2 - g.N[1] / 2) * g.dq[1];
        const float u2 = (float)(i3 - g.N[2] / 2) * g.dq[2];
        float v[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = __fadd_rn(__fadd_rn(__fmul_rn(g.ainv[3 * a], u0), __fmul_rn(g.ainv[3 * a + 1], u1)),
                                                  __fmul_rn(g.ainv[3 * a + 2], u2));
        const float q = __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(v[0], v[0]), __fmul_rn(v[1], v[1])), __fmul_rn(v[2], v[2])));
        if (!(q <= g.q_range)) continue;
        const int b = (int)roundf(__fdiv_rn(q, g.q_step));  // nint: half away from zero
        if (b < 0 || b >= g.n_hist) continue;
        double ar[8], ai[8];
        for (int s = 0; s < g.nsp; ++s) {
            const C a = amp[(int64_t)s * g.nmodes + m];
            ar[s] = (double)a.x;
            ai[s] = (double)a.y;
        }
        int p = 0;
        for (int s = 0; s < g.nsp; ++s)
            for (int t = 0; t <= s; ++t, ++p)
                atomicAdd(&hs[b * g.npairs + p], ar[s] * ar[t] + ai[s] * ai[t]);
    }
    __syncthreads();
    double *dst = part + (int64_t)blockIdx.x * nb;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) dst[i] = hs[i];
}

// out[i] += sum over the nparts slices of part[slice][i], in slice order
__global__ void k_sql_reduce(const double *__restrict__ part, int nparts, int nb, double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    double s = 0.0;
    for (int k = 0; k < nparts; ++k) s += part[(int64_t)k * nb + i];
    out[i] += s;
}

// Returns 0, or -1 when the histogram does not fit one workgroup's LDS.
int launch_sql_hist(const void *amp, bool fp64, const SqlGeom &g, double *out, hipStream_t st)
{
    const int nb = g.n_hist * g.npairs;
    const size_t lds = sizeof(double) * (size_t)nb;
    if (lds > 64 * 1024 || g.nsp > 8) return -1;
    int dev = 0, ncu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const int nparts = std::max(1, ncu) * 2;
    double *part = nullptr;
    if (hipMallocAsync(reinterpret_cast<void **>(&part), sizeof(double) * (size_t)nparts * nb, st) != hipSuccess)
        return -2;
    if (fp64)
        hipLaunchKernelGGL((k_sql_hist<double2>), dim3(nparts), dim3(256), lds, st,
                           static_cast<const double2 *>(amp), g, part);
    else
        hipLaunchKernelGGL((k_sql_hist<float2>), dim3(nparts), dim3(256), lds, st,
                           static_cast<const float2 *>(amp), g, part);
    hipLaunchKernelGGL(k_sql_reduce, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, part, nparts, nb, out);
    (void)hipFreeAsync(part, st);
    return 0;
}

// ---------------------------------------------------------------------------
// The same histogram without any allocation: the partial slices live in
// `scratch`, a file-scope device array of SQL_SCRATCH_DOUBLES doubles
// (8 MiB) that every GPU the library runs on holds until the process ends.
// Host code never uses the array's name as a device address: once per
// device, on the first call with that device current, hipGetSymbolAddress
// gives that device's copy, kept in scratch_ptr[device].  Every call on one
// GPU shares the array (calls that overlap on one GPU race); each slice is
// written whole by its workgroup before the reduction reads it, so nothing
// needs clearing between calls.
// ---------------------------------------------------------------------------
constexpr int64_t SQL_SCRATCH_DOUBLES = (int64_t)1 << 20;
static __device__ __align__(256) double scratch[SQL_SCRATCH_DOUBLES];

constexpr int SQL_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[SQL_MAX_DEVICES] = {};

// the device's copy of `scratch` (nullptr if the lookup fails)
static double *scratch_on(int dev)
{
    if (dev < 0 || dev >= SQL_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return nullptr;
        g_scratch_ptr[dev] = static_cast<double *>(ptr);
    }
    return g_scratch_ptr[dev];
}

int64_t sql_scratch_doubles() { return SQL_SCRATCH_DOUBLES; }

// Returns 0; -1 when the histogram does not fit one workgroup's LDS (or nsp
// > 8); -3 when the scratch array would hold fewer than 64 slices; -2 when
// the device's scratch address cannot be had.
int launch_sql_hist_static(const void *amp, bool fp64, const SqlGeom &g, double *out, hipStream_t st)
{
    const int nb = g.n_hist * g.npairs;
    const size_t lds = sizeof(double) * (size_t)nb;
    if (lds > 64 * 1024 || g.nsp > 8) return -1;
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -2;
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const int64_t fit = SQL_SCRATCH_DOUBLES / std::max(nb, 1);
    if (fit < 64) return -3;
    const int nparts = (int)std::min<int64_t>((int64_t)std::max(1, ncu) * 2, fit);
    double *part = scratch_on(dev);
    if (!part) return -2;
    if (fp64)
        hipLaunchKernelGGL((k_sql_hist<double2>), dim3(nparts), dim3(256), lds, st,
                           static_cast<const double2 *>(amp), g, part);
    else
        hipLaunchKernelGGL((k_sql_hist<float2>), dim3(nparts), dim3(256), lds, st,
                           static_cast<const float2 *>(amp), g, part);
    hipLaunchKernelGGL(k_sql_reduce, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, part, nparts, nb, out);
    return 0;
}

}  // namespace mpsq
