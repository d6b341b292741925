// Host check of the spread tiles' candidate runs (spread_tiles.h, the code
// the spread kernels run).  For several fine grids, kernel widths, tile and
// sort-bin shapes (including the cyclic wrap of a tile's bin range and bins
// coarser than tiles), random points are bin-sorted with a stable counting
// sort; then, for every tile, the run table is built as the kernels build it
// (tile_bins, tile_run, prefix of the run sizes) and the candidates are
// walked as the kernels walk them (one cursor per lane, candidates ascending).
// Checked: the run count fits MAX_RUNS, the cursor stays below the run count,
// every run position lies inside the transform's points, no point is visited
// twice by a tile, and every point whose kernel stencil reaches a row of the
// tile is visited.  Prints "ok" or the first failure.
#include <cstdio>
#include <cstdint>
#include <random>
#include <vector>

#include "spread_tiles.h"

using namespace mpsq;

static int fail(const char *what, int dim, int n2, int n3, int w, int by, int bz, int BY, int BZ, int y0, int z0)
{
    std::printf("FAIL %s: dim %d n2 %d n3 %d w %d tile %dx%d bins %dx%d at (%d,%d)\n", what, dim, n2, n3, w, by, bz,
                BY, BZ, y0, z0);
    return 1;
}

static bool covers(int l, int r0, int b, int w, int n)
{
    // stencil rows l .. l+w-1 (mod n) meet tile rows r0 .. r0+b-1
    for (int m = 0; m < w; ++m) {
        const int r = ((l + m) % n + n) % n;
        if (r >= r0 && r < r0 + b) return true;
    }
    return false;
}

static int check(int dim, int n2, int n3, int w, int by, int bz, int BY, int BZ, int M, uint32_t seed)
{
    Geo g{};
    g.dim = dim;
    g.n1 = 64;
    g.n2 = dim >= 2 ? n2 : 1;
    g.n3 = dim >= 3 ? n3 : 1;
    g.w = w;
    g.by = by;
    g.bz = bz;
    g.BY = dim >= 2 ? BY : 1;
    g.BZ = dim >= 3 ? BZ : 1;
    g.nbY = dim >= 2 ? n2 / BY : 1;
    g.nbZ = dim >= 3 ? n3 / BZ : 1;
    g.nbins = g.nbY * g.nbZ;
    std::mt19937 rng(seed);
    std::vector<int> ly(M), lz(M), key(M);
    for (int i = 0; i < M; ++i) {
        ly[i] = dim >= 2 ? (int)(rng() % (uint32_t)g.n2) : 0;
        lz[i] = dim >= 3 ? (int)(rng() % (uint32_t)g.n3) : 0;
        key[i] = (lz[i] / g.BZ) * g.nbY + ly[i] / g.BY;
    }
    // stable counting sort
    std::vector<int64_t> bs(g.nbins + 1, 0);
    for (int i = 0; i < M; ++i) ++bs[key[i] + 1];
    for (int b = 0; b < g.nbins; ++b) bs[b + 1] += bs[b];
    std::vector<int64_t> pos(bs.begin(), bs.end() - 1);
    std::vector<int> sorted(M);
    for (int i = 0; i < M; ++i) sorted[pos[key[i]]++] = i;
    const int nty = g.n2 / (dim >= 2 ? by : 1), ntz = g.n3 / (dim >= 3 ? bz : 1);
    std::vector<int> seen(M, -1);
    int tile = 0;
    for (int tz = 0; tz < ntz; ++tz)
        for (int ty = 0; ty < nty; ++ty, ++tile) {
            const int y0 = ty * (dim >= 2 ? by : 1), z0 = tz * (dim >= 3 ? bz : 1);
            int ylo, nyb, zlo, nzb;
            tile_bins(g, y0, z0, ylo, nyb, zlo, nzb);
            const int nruns = nruns_of(g, y0, z0);
            if (nruns != 2 * nzb || nruns > MAX_RUNS) return fail("run count", dim, n2, n3, w, by, bz, BY, BZ, y0, z0);
            int64_t rbeg[MAX_RUNS];
            int rpre[MAX_RUNS + 1];
            rpre[0] = 0;
            for (int r = 0; r < nruns; ++r) {
                int cnt;
                tile_run(g, bs.data(), ylo, nyb, zlo, r, rbeg[r], cnt);
                if (cnt < 0) return fail("negative run", dim, n2, n3, w, by, bz, BY, BZ, y0, z0);
                rpre[r + 1] = rpre[r] + cnt;
            }
            const int total = rpre[nruns];
            // the kernels' walk: lanes take v = lane, lane + 64, ... with a
            // private cursor; one lane's sequence is enough (the walk of every
            // lane is the same code on an ascending subsequence)
            for (int lane = 0; lane < 64; ++lane) {
                int e = 0;
                for (int v = lane; v < total; v += 64) {
                    while (e < MAX_RUNS - 1 && rpre[e + 1] <= v) ++e;
                    if (e >= nruns) return fail("cursor past the runs", dim, n2, n3, w, by, bz, BY, BZ, y0, z0);
                    const int64_t p = rbeg[e] + (v - rpre[e]);
                    if (p < 0 || p >= M) return fail("run position out of range", dim, n2, n3, w, by, bz, BY, BZ, y0, z0);
                    const int i = sorted[p];
                    if (seen[i] == tile) return fail("point visited twice", dim, n2, n3, w, by, bz, BY, BZ, y0, z0);
                    seen[i] = tile;
                }
            }
            for (int i = 0; i < M; ++i) {
                const bool reach = (dim < 2 || covers(ly[i], y0, by, w, g.n2)) && (dim < 3 || covers(lz[i], z0, bz, w, g.n3));
                if (reach && seen[i] != tile) return fail("reaching point missed", dim, n2, n3, w, by, bz, BY, BZ, y0, z0);
            }
        }
    return 0;
}

int main()
{
    struct Case {
        int dim, n2, n3, w, by, bz, BY, BZ;
    };
    const Case cases[] = {
        {1, 1, 1, 7, 1, 1, 1, 1},
        {2, 64, 1, 7, 8, 1, 8, 1},     {2, 64, 1, 8, 16, 1, 16, 1},   {2, 16, 1, 8, 8, 1, 16, 1},
        {2, 30, 1, 4, 2, 1, 6, 1},     {2, 20, 1, 12, 4, 1, 4, 1},
        {3, 64, 32, 8, 4, 4, 4, 4},    {3, 64, 32, 8, 8, 2, 8, 2},    {3, 64, 32, 8, 4, 4, 8, 8},
        {3, 32, 16, 7, 4, 4, 16, 16},  {3, 48, 24, 6, 2, 8, 4, 8},    {3, 16, 16, 7, 8, 8, 8, 8},
        {3, 40, 20, 5, 1, 1, 2, 4},    {3, 128, 18, 3, 16, 2, 32, 6}, {3, 30, 30, 13, 2, 2, 10, 10},
        {3, 64, 64, 8, 16, 1, 16, 4},
    };
    uint32_t seed = 1;
    for (const Case &c : cases)
        for (int M : {0, 1, 37, 3000})
            if (check(c.dim, c.n2, c.n3, c.w, c.by, c.bz, c.BY, c.BZ, M, seed++)) return 1;
    std::printf("ok\n");
    return 0;
}
