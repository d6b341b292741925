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
// near the axes) and writes it to its own slice of a partial buffer; the
// slices are added in a fixed order, by groups over many blocks and then the
// group sums.
#include <algorithm>
#include <mutex>

#include "mpsq_kernels.h"

namespace mpsq {

// |Q| of the mode k = (k1, k2, k3) (offsets from the centre) in FP32:
// norm2(a_cell_inv . (k * d_q)) with every operation rounded on its own
__device__ inline float sql_qnorm(const SqlGeom &g, int k1, int k2, int k3)
{
    const float u0 = __fmul_rn((float)k1, g.dq[0]);
    const float u1 = __fmul_rn((float)k2, g.dq[1]);
    const float u2 = __fmul_rn((float)k3, g.dq[2]);
    float v[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] = __fadd_rn(__fadd_rn(__fmul_rn(g.ainv[3 * a], u0), __fmul_rn(g.ainv[3 * a + 1], u1)),
                                              __fmul_rn(g.ainv[3 * a + 2], u2));
    return __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(v[0], v[0]), __fmul_rn(v[1], v[1])), __fmul_rn(v[2], v[2])));
}

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
        const float q = sql_qnorm(g, i1 - g.N[0] / 2, i2 - g.N[1] / 2, i3 - g.N[2] / 2);
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

// part2[grp][i] = sum of part[k][i] over the slices k of group grp
// (k = grp * per .. min((grp + 1) * per, nparts) - 1, in order): the first of
// two levels, so that the slices are added by many blocks at once
__global__ void k_sql_reduce_groups(const double *__restrict__ part, int nparts, int per, int nb,
                                    double *__restrict__ part2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const int k0 = blockIdx.y * per, k1 = min(k0 + per, nparts);
    double s = 0.0;
    for (int k = k0; k < k1; ++k) s += part[(int64_t)k * nb + i];
    part2[(int64_t)blockIdx.y * nb + i] = s;
}

constexpr int SQL_REDUCE_GROUPS = 32;

// out[i] += the nparts slices of part, through SQL_REDUCE_GROUPS group sums
// in part2 (fixed order: the result does not depend on timing)
static void reduce_slices(const double *part, int nparts, int nb, double *part2, double *out, hipStream_t st)
{
    const int per = (nparts + SQL_REDUCE_GROUPS - 1) / SQL_REDUCE_GROUPS;
    const int ng = (nparts + per - 1) / per;
    hipLaunchKernelGGL(k_sql_reduce_groups, dim3((unsigned)((nb + 255) / 256), (unsigned)ng), dim3(256), 0, st, part,
                       nparts, per, nb, part2);
    hipLaunchKernelGGL(k_sql_reduce, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, part2, ng, nb, out);
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
    if (hipMallocAsync(reinterpret_cast<void **>(&part), sizeof(double) * (size_t)(nparts + SQL_REDUCE_GROUPS) * nb,
                       st) != hipSuccess)
        return -2;
    if (fp64)
        hipLaunchKernelGGL((k_sql_hist<double2>), dim3(nparts), dim3(256), lds, st,
                           static_cast<const double2 *>(amp), g, part);
    else
        hipLaunchKernelGGL((k_sql_hist<float2>), dim3(nparts), dim3(256), lds, st,
                           static_cast<const float2 *>(amp), g, part);
    reduce_slices(part, nparts, nb, part + (size_t)nparts * nb, out, st);
    (void)hipFreeAsync(part, st);
    return 0;
}

// ---------------------------------------------------------------------------
// Monte Carlo integration (mp_sql's j_acc = 2, code/src/mp_sql56.f90:1146-1215).
// Sample (j, k) -- j < n_mc within chain k < n_h -- draws a point uniformly
// in the cube [-1, 1)^3 until one lands in the unit ball (0 < |p| <= 1),
// projects it onto the unit sphere, scales it by a radius fraction r in
// [0, 1) and by the half-widths h_d = N_d - (N_d / 2 + 1) of the mode box, and
// rounds to the mode d = nint(h * p / |p| * r) (:1177-1190).  d = 0 counts as
// skip 1 (:1192-1195); |Q(d)| > q_range as skip 2 (:1208-1210); otherwise
// Re(A_i A_j*) at d is added to bin nint(|Q| / q_step) as in the Gauss
// kernel (:1201-1207).
//
// The reference draws from Fortran's random_number; here every number comes
// from the counter-based generator Philox4x32-10 (Salmon, Moraes, Dror and
// Shaw, SC'11), so a sample is a pure function of (seed, tag, k, j): attempt
// a of sample (j, k) is philox(counter (j, k, a, tag), key (seed_lo,
// seed_hi)); its words 0..2 give the cube point and word 3 the radius
// fraction of the accepted attempt, each as (word >> 8) * 2^-24.  The FP32
// arithmetic of the point, its norm, the projection and the rounding is the
// reference's, each operation rounded on its own.  A sample whose first
// MC_MAX_ATTEMPTS attempts all miss the ball (probability 0.477^64 < 1e-20)
// counts as skip 1 rather than looping on.
//
// Each lane walks its own samples (stride = grid size) one attempt per
// iteration, so lanes whose draw missed the ball do not hold up the others;
// the histogram is summed in FP64 LDS atomics per workgroup and written to
// the workgroup's slice like the Gauss kernel's; skip counts are exact
// integers added to d_skip.
// ---------------------------------------------------------------------------
constexpr uint32_t MC_MAX_ATTEMPTS = 64;

struct Ph4 {
    uint32_t x, y, z, w;
};

__device__ inline Ph4 philox4x32_10(Ph4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = Ph4{(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// [0, 1) with 24 random bits: exact in FP32
__device__ inline float mc_unit(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

template <typename C>
__global__ __launch_bounds__(256) void k_sql_hist_mc(const C *__restrict__ amp, SqlGeom g, SqlMc mc,
                                                     double *__restrict__ part, unsigned long long *__restrict__ skip)
{
    extern __shared__ double hs[];  // [n_hist][npairs]
    __shared__ unsigned long long sks[2];
    const int nb = g.n_hist * g.npairs;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) hs[i] = 0.0;
    if (threadIdx.x < 2) sks[threadIdx.x] = 0;
    __syncthreads();

    const uint64_t n_mc = (uint64_t)mc.n_mc;
    const uint64_t total = (uint64_t)mc.n_h * n_mc;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // (j, k) of idx = k n_mc + j, advanced by the stride's own (j, k) split
    uint32_t j = (uint32_t)(idx % n_mc), k = (uint32_t)(idx / n_mc);
    const uint32_t sj = (uint32_t)(stride % n_mc), sk = (uint32_t)(stride / n_mc);
    const uint32_t key0 = (uint32_t)mc.seed, key1 = (uint32_t)(mc.seed >> 32);
    float h[3];
    int c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        c[d] = g.N[d] / 2;
        h[d] = (float)(g.N[d] - (c[d] + 1));
    }
    uint32_t attempt = 0;
    unsigned s1 = 0, s2 = 0;
    while (idx < total) {
        const Ph4 x = philox4x32_10(Ph4{j, k, attempt, mc.tag}, key0, key1);
        const float p0 = 2.0f * mc_unit(x.x) - 1.0f;  // exact
        const float p1 = 2.0f * mc_unit(x.y) - 1.0f;
        const float p2 = 2.0f * mc_unit(x.z) - 1.0f;
        const float pn = __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(p0, p0), __fmul_rn(p1, p1)), __fmul_rn(p2, p2)));
        bool done = true;
        if (pn > 0.0f && pn <= 1.0f) {
            const float r = mc_unit(x.w);
            const int d0 = (int)roundf(__fmul_rn(__fmul_rn(h[0], __fdiv_rn(p0, pn)), r));
            const int d1 = (int)roundf(__fmul_rn(__fmul_rn(h[1], __fdiv_rn(p1, pn)), r));
            const int d2 = (int)roundf(__fmul_rn(__fmul_rn(h[2], __fdiv_rn(p2, pn)), r));
            const int i1 = c[0] + d0, i2 = c[1] + d1, i3 = c[2] + d2;
            if ((d0 | d1 | d2) == 0) {
                ++s1;
            } else if ((unsigned)i1 >= (unsigned)g.N[0] || (unsigned)i2 >= (unsigned)g.N[1] ||
                       (unsigned)i3 >= (unsigned)g.N[2]) {
                ++s2;  // outside the mode box: |p / |p|| <= 1 and r < 1 keep every sample inside
            } else {
                const float q = sql_qnorm(g, d0, d1, d2);
                if (!(q <= g.q_range)) {
                    ++s2;
                } else {
                    const int b = (int)roundf(__fdiv_rn(q, g.q_step));
                    if (b < g.n_hist) {
                        const int64_t m = i1 + (int64_t)g.N[0] * (i2 + (int64_t)g.N[1] * i3);
                        double ar[8], ai[8];
                        for (int s = 0; s < g.nsp; ++s) {
                            const C a = amp[(int64_t)s * g.nmodes + m];
                            ar[s] = (double)a.x;
                            ai[s] = (double)a.y;
                        }
                        int pp = 0;
                        for (int s = 0; s < g.nsp; ++s)
                            for (int t = 0; t <= s; ++t, ++pp)
                                atomicAdd(&hs[b * g.npairs + pp], ar[s] * ar[t] + ai[s] * ai[t]);
                    }
                }
            }
        } else if (++attempt == MC_MAX_ATTEMPTS) {
            ++s1;
        } else {
            done = false;
        }
        if (done) {
            attempt = 0;
            idx += stride;
            j += sj;
            k += sk;
            if (j >= (uint32_t)n_mc) {
                j -= (uint32_t)n_mc;
                ++k;
            }
        }
    }
    if (s1) atomicAdd(&sks[0], (unsigned long long)s1);
    if (s2) atomicAdd(&sks[1], (unsigned long long)s2);
    __syncthreads();
    if (threadIdx.x < 2 && sks[threadIdx.x]) atomicAdd(&skip[threadIdx.x], sks[threadIdx.x]);
    double *dst = part + (int64_t)blockIdx.x * nb;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) dst[i] = hs[i];
}

// Returns 0, -1 when the histogram does not fit one workgroup's LDS (or nsp
// > 8), -2 when the partial buffer cannot be allocated.
int launch_sql_hist_mc(const void *amp, bool fp64, const SqlGeom &g, const SqlMc &mc, double *out,
                       unsigned long long *skip, hipStream_t st)
{
    const int nb = g.n_hist * g.npairs;
    const size_t lds = sizeof(double) * (size_t)nb;
    if (lds > 64 * 1024 || g.nsp > 8) return -1;
    int dev = 0, ncu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    // enough lanes to fill the chip, no more workgroups than samples need
    const uint64_t total = (uint64_t)mc.n_h * (uint64_t)mc.n_mc;
    const int nparts = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)std::max(1, ncu) * 4, (total + 255) / 256));
    double *part = nullptr;
    if (hipMallocAsync(reinterpret_cast<void **>(&part), sizeof(double) * (size_t)(nparts + SQL_REDUCE_GROUPS) * nb,
                       st) != hipSuccess)
        return -2;
    if (fp64)
        hipLaunchKernelGGL((k_sql_hist_mc<double2>), dim3(nparts), dim3(256), lds, st,
                           static_cast<const double2 *>(amp), g, mc, part, skip);
    else
        hipLaunchKernelGGL((k_sql_hist_mc<float2>), dim3(nparts), dim3(256), lds, st,
                           static_cast<const float2 *>(amp), g, mc, part, skip);
    reduce_slices(part, nparts, nb, part + (size_t)nparts * nb, out, st);
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
