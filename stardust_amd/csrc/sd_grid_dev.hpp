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
// cells (start == end == 0) are skipped. Empty row: lo == hi == 0.
__device__ __forceinline__ void row_range(const GridDev& g, const int* __restrict__ ranges, int cx, int y, int z,
                                          int& lo, int& hi) {
    const int x0 = cx > 0 ? cx - 1 : 0;
    const int x1 = cx < g.nx - 1 ? cx + 1 : g.nx - 1;
    const int base = (z * g.ny + y) * g.nx;
    lo = 0x7fffffff;
    hi = 0;
    for (int x = x0; x <= x1; ++x) {
        const int2 r = *reinterpret_cast<const int2*>(ranges + 2 * (size_t)(base + x));
        if (r.y > r.x) {
            lo = r.x < lo ? r.x : lo;
            hi = r.y > hi ? r.y : hi;
        }
    }
    if (hi == 0) lo = 0;
}

}  // namespace sd
