// Order in which the workgroups of one accumulation launch take the tiles of a tiles_x x tiles_y
// grid (host and device code; plain C++ without HIP).
#pragma once

#ifdef __HIPCC__
#define SARBP_HD __host__ __device__
#else
#define SARBP_HD
#endif

namespace sarbp {

constexpr int kXcds = 8;

// The dispatcher deals a launch's workgroups round-robin over the device's 8 XCDs, each with its
// own L2: workgroup `id` runs on XCD id % 8. xcd_rank is the position in a list of n work items
// that workgroup `id` takes so that XCD k gets the k-th eighth of the list, in order (XCDs below
// n % 8 get one item more). A bijection of [0, n).
SARBP_HD inline int xcd_rank(int id, int n) {
    const int q = n / kXcds, r = n % kXcds;
    const int k = id % kXcds, j = id / kXcds;
    return k * q + (k < r ? k : r) + j;
}

// Tile (tx, ty) of workgroup `id`: XCD k takes the k-th eighth of the tiles in column-major order
// (`cols`: a band of whole columns) or row-major order (a band of whole rows). Tiles that lie at
// about the same range from the platform read the same stretch of each profile row, and they do so
// at about the same time when they start together on one XCD; a band across the range direction
// keeps those tiles on one L2. On configuration C2 (range along image x) column bands cut the
// accumulation's HBM reads from 9.1 GB to 1.06 GB per launch against the dispatch order, and row
// bands raise them to 17.5 GB. A bijection of [0, tiles_x * tiles_y) onto the grid.
SARBP_HD inline void xcd_band_tile(int id, int tiles_x, int tiles_y, bool cols, int &tx, int &ty) {
    const int t = xcd_rank(id, tiles_x * tiles_y);
    if (cols) {
        tx = t / tiles_y;
        ty = t % tiles_y;
    } else {
        tx = t % tiles_x;
        ty = t / tiles_x;
    }
}

}  // namespace sarbp
