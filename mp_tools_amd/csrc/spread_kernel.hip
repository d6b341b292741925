// spread_kernel.hip -- host entry points of the spread + x FFT pass: LDS
// sizes of a tile geometry, the counter-zeroing kernel, and the dispatch of a
// job table to the launch of its kernel width (the kernels and launches are
// in spread_impl.h, instantiated per width in spread_w*.hip).
#include "spread_impl.h"

namespace mpsq {

MPSQ_SPREAD_W_EXTERN(2)
MPSQ_SPREAD_W_EXTERN(3)
MPSQ_SPREAD_W_EXTERN(4)
MPSQ_SPREAD_W_EXTERN(5)
MPSQ_SPREAD_W_EXTERN(6)
MPSQ_SPREAD_W_EXTERN(7)
MPSQ_SPREAD_W_EXTERN(8)
MPSQ_SPREAD_W_EXTERN(9)
MPSQ_SPREAD_W_EXTERN(10)
MPSQ_SPREAD_W_EXTERN(11)
MPSQ_SPREAD_W_EXTERN(12)
MPSQ_SPREAD_W_EXTERN(13)
MPSQ_SPREAD_W_EXTERN(14)
MPSQ_SPREAD_W_EXTERN(15)
MPSQ_SPREAD_W_EXTERN(16)

size_t spread_lds_real(int rows, int n1, int w)
{
    return (size_t)spread_real_tw_offset(rows, n1, w) + (size_t)n1 * 8 + (size_t)(n1 / 2) * 4 + (size_t)rows * 8;
}
bool spread_real_ok(int rows, int, int) { return spread_real_fits(rows); }
template <typename R> int spread_tile_elems() { return SpreadCfg<R>::E * SpreadCfg<R>::NT; }
int spread_max_runs() { return MAX_RUNS; }
int spread_max_tile_b() { return MAX_TILE_B; }
int spread_max_wide_passes() { return MAX_WIDE_PASSES; }
template <typename R> size_t spread_lds(int rows, int n1, int w, bool real)
{
    return spread_lds_bytes<R>(rows, n1, w, real);
}
template <typename R> size_t spread_lds_twiddles(int rows, int n1, int w, bool real)
{
    return spread_tile_bytes<R>(rows, n1, w, real);
}

__global__ void k_zero_i32(int *__restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

void launch_zero_i32(int *p, int n, hipStream_t st)
{
    if (n > 0) hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
}

template <typename R>
void launch_spread_jobs(const SpreadJob<R> *jobs, const int *tile0s, int njobs, int w, const int *xr, int *wc,
                        int *fbl, size_t lds, hipStream_t st, size_t lds_real)
{
    if (njobs <= 0) return;
#define MPSQ_SPREAD_W(WV)                                                                                             \
    case WV: launch_spread_w<R, WV>(jobs, tile0s, njobs, xr, wc, fbl, lds, st, lds_real); break;
    switch (w) {
        MPSQ_SPREAD_W(2)
        MPSQ_SPREAD_W(3)
        MPSQ_SPREAD_W(4)
        MPSQ_SPREAD_W(5)
        MPSQ_SPREAD_W(6)
        MPSQ_SPREAD_W(7)
        MPSQ_SPREAD_W(8)
        MPSQ_SPREAD_W(9)
        MPSQ_SPREAD_W(10)
        MPSQ_SPREAD_W(11)
        MPSQ_SPREAD_W(12)
        MPSQ_SPREAD_W(13)
        MPSQ_SPREAD_W(14)
        MPSQ_SPREAD_W(15)
        MPSQ_SPREAD_W(16)
    default: break;
    }
#undef MPSQ_SPREAD_W
}

#define MPSQ_SPREAD_INST(R)                                                                                           \
    template int spread_tile_elems<R>();                                                                              \
    template int spread_max_rows<R>(int);                                                                             \
    template int spread_group_rows<R>(int, int);                                                                      \
    template size_t spread_lds<R>(int, int, int, bool);                                                               \
    template size_t spread_lds_twiddles<R>(int, int, int, bool);                                                      \
    template void launch_spread_jobs<R>(const SpreadJob<R> *, const int *, int, int, const int *, int *, int *,      \
                                        size_t, hipStream_t, size_t);

MPSQ_SPREAD_INST(float)
MPSQ_SPREAD_INST(double)

}  // namespace mpsq
