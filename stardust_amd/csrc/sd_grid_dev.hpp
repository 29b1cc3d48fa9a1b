// Device-side grid helpers shared by the grid and contact kernels.
#pragma once

#include "sd_internal.hpp"

namespace sd {

// clamp(floor((p - lo) / cell), 0, nc - 1); NaN maps to 0.
__device__ __forceinline__ int axis_cell(float p, float lo, float cell, int nc) {
    float f = floorf((p - lo) / cell);
    f = fmaxf(f, 0.0f);
    f = fminf(f, (float)(nc - 1));
    return (int)f;
}

__device__ __forceinline__ void cell_coords(const GridDev& g, float x, float y, float z, int& cx, int& cy, int& cz) {
    cx = axis_cell(x, g.lox, g.cell, g.nx);
    cy = axis_cell(y, g.loy, g.cell, g.ny);
    cz = axis_cell(z, g.loz, g.cell, g.nz);
}

__device__ __forceinline__ uint32_t cell_id(const GridDev& g, float x, float y, float z) {
    int cx, cy, cz;
    cell_coords(g, x, y, z, cx, cy, cz);
    return (uint32_t)((cz * g.ny + cy) * g.nx + cx);
}

// Sorted-position range covering cells (cx-1 .. cx+1, y, z). The cells of one
// row are consecutive cell ids, so their sorted runs are contiguous; empty
// cells (start == end == 0) are skipped. Empty row: lo == hi == 0. The three
// range entries are loaded together (a clamped x repeats a cell, which does not
// change the union).
__device__ __forceinline__ void row_range(const GridDev& g, const int* __restrict__ ranges, int cx, int y, int z,
                                          int& lo, int& hi) {
    const int base = (z * g.ny + y) * g.nx;
    const int xa = cx > 0 ? cx - 1 : cx;
    const int xb = cx < g.nx - 1 ? cx + 1 : cx;
    const int2 ra = *reinterpret_cast<const int2*>(ranges + 2 * (size_t)(base + xa));
    const int2 rb = *reinterpret_cast<const int2*>(ranges + 2 * (size_t)(base + cx));
    const int2 rc = *reinterpret_cast<const int2*>(ranges + 2 * (size_t)(base + xb));
    lo = 0x7fffffff;
    hi = 0;
    if (ra.y > ra.x) {
        lo = ra.x < lo ? ra.x : lo;
        hi = ra.y > hi ? ra.y : hi;
    }
    if (rb.y > rb.x) {
        lo = rb.x < lo ? rb.x : lo;
        hi = rb.y > hi ? rb.y : hi;
    }
    if (rc.y > rc.x) {
        lo = rc.x < lo ? rc.x : lo;
        hi = rc.y > hi ? rc.y : hi;
    }
    if (hi == 0) lo = 0;
}

// Sorted-position window of cells (cx-1 .. cx+1, y, z) from the cell starts
// (every cell has one, empty or not): [start of the first cell, start of the
// cell after the last). Empty window: lo == hi.
__device__ __forceinline__ void row_window(const GridDev& g, const int* __restrict__ start, int cx, int y, int z,
                                           int& lo, int& hi) {
    const int base = (z * g.ny + y) * g.nx;
    const int xa = cx > 0 ? cx - 1 : cx;
    const int xe = cx < g.nx - 1 ? cx + 2 : cx + 1;
    lo = start[(size_t)(base + xa)];
    hi = start[(size_t)(base + xe)];
}

}  // namespace sd
