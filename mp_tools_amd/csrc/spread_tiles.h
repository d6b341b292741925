// spread_tiles.h -- which sorted points a spread tile visits.
//
// Points are bin-sorted by key zb * nbY + yb (z bin major, y bin minor; bins
// of BY x BZ fine-grid rows, the whole x extent).  A tile of by x bz rows
// starting at (y0, z0) can receive contributions from points whose first
// kernel cell lies in rows [y0 - (w - 1), y0 + by) x [z0 - (w - 1), z0 + bz)
// (cyclically), i.e. from the y bins [ylo, ylo + nyb) of each z bin of
// [zlo, zlo + nzb).  In one z-bin row those y bins are at most two contiguous
// key ranges (two when the y range wraps past the end of the grid), so a
// tile's candidates are 2 * nzb runs of the sorted arrays.
//
// These functions are shared by the spread kernels and a host check that
// walks every tile of several geometries (tests/native/runs_check.cpp).
#pragma once
#include <stdint.h>

#include "mpsq_kernels.h"

namespace mpsq {

constexpr int MAX_RUNS = 64;  // candidate runs of one tile (two per visited z-bin row)

__host__ __device__ inline int floordiv(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }
__host__ __device__ inline int posmod(int a, int n)
{
    const int r = a % n;
    return r < 0 ? r + n : r;
}

// sort bins a tile visits: y bins [ylo, ylo + nyb), z bins [zlo, zlo + nzb) (cyclic)
__host__ __device__ inline void tile_bins(const Geo &g, int y0, int z0, int &ylo, int &nyb, int &zlo, int &nzb)
{
    const int w = g.w, by = (g.dim >= 2) ? g.by : 1, bz = (g.dim >= 3) ? g.bz : 1;
    ylo = 0;
    nyb = 1;
    zlo = 0;
    nzb = 1;
    if (g.dim >= 2) {
        ylo = floordiv(y0 - (w - 1), g.BY);
        nyb = floordiv(y0 + by - 1, g.BY) - ylo + 1;
        if (nyb >= g.nbY) {
            ylo = 0;
            nyb = g.nbY;
        }
    }
    if (g.dim >= 3) {
        zlo = floordiv(z0 - (w - 1), g.BZ);
        nzb = floordiv(z0 + bz - 1, g.BZ) - zlo + 1;
        if (nzb >= g.nbZ) {
            zlo = 0;
            nzb = g.nbZ;
        }
    }
}

// runs of a tile: 2 per visited z-bin row (the second one is empty unless
// the y range wraps)
__host__ __device__ inline int nruns_of(const Geo &g, int y0, int z0)
{
    int ylo, nyb, zlo, nzb;
    tile_bins(g, y0, z0, ylo, nyb, zlo, nzb);
    return 2 * nzb;
}

// Run `r` (< 2 * nzb) of a tile: the sorted positions [b0, b0 + cnt) of its
// bins, from the transform's bin starts bs (nbins + 1 entries).
struct PlainLoad {
    __host__ __device__ int64_t operator()(const int64_t *q) const { return *q; }
};
// ld(q) reads bs entries (the kernels pass a global-address-space load)
template <class Ld = PlainLoad>
__host__ __device__ inline void tile_run(const Geo &g, const int64_t *bs, int ylo, int nyb, int zlo, int r,
                                         int64_t &b0, int &cnt, Ld ld = Ld())
{
    b0 = 0;
    cnt = 0;
    const int k = r >> 1, part = r & 1;
    const int zb = posmod(zlo + k, g.nbZ);
    int ya = 0, yb = 0;  // bin range [ya, yb) of this part
    const int a = ylo, b = ylo + nyb;
    if (nyb >= g.nbY) {
        if (part == 0) yb = g.nbY;
    } else if (a < 0) {
        if (part == 0) {
            ya = a + g.nbY;
            yb = g.nbY;
        } else {
            yb = b;
        }
    } else if (b > g.nbY) {
        if (part == 0) {
            ya = a;
            yb = g.nbY;
        } else {
            yb = b - g.nbY;
        }
    } else if (part == 0) {
        ya = a;
        yb = b;
    }
    if (yb > ya) {
        b0 = ld(bs + zb * g.nbY + ya);
        cnt = (int)(ld(bs + zb * g.nbY + yb) - b0);
    }
}

}  // namespace mpsq
