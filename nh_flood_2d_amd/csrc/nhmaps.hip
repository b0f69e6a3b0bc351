// libnhflood.so, flood maps: cell values splatted onto a north-up raster tile.
//
// A mesh cell is an axis-aligned rectangle (centre x, y; half sizes hw, hh). The tile is `rows` x
// `cols` pixels of hr x vr metres whose top-left corner is (tile_min_x, tile_max_y). A cell owns
// the pixels whose centres lie in its half-open bounds [min, max) along x and (min, max] along the
// row axis, which gives (output/flood_map.py:196-222):
//   first column ceil((min_x - tile_min_x) / hr - 0.5), last ceil((max_x - tile_min_x) / hr - 0.5) - 1,
//   first row ceil((tile_max_y - max_y) / vr - 0.5), last ceil((tile_max_y - min_y) / vr - 0.5) - 1,
// clamped to the tile, evaluated in F32 in that order (no contraction), after an F32 bounding-box
// overlap test against the tile. Cells do not overlap, so every pixel has at most one writer.
//
// One thread per cell: a cell covers (2 hw / hr) x (2 hh / vr) pixels, 1 to 16 on a 3-level
// quadtree at the finest size. The pass is bound by its scattered tile stores; the tile stays
// resident in HBM between tiles and frames.
#include "nhf_device.h"

__global__ __launch_bounds__(256) void k_rasterize(const float *__restrict__ x, const float *__restrict__ y,
                                                    const float *__restrict__ hw, const float *__restrict__ hh,
                                                    const float *__restrict__ value, int n, float tile_min_x,
                                                    float tile_max_y, float hr, float vr, int rows, int cols,
                                                    float *__restrict__ tile, long long ld) {
    const float tile_max_x = tile_min_x + (float)cols * hr;
    const float tile_min_y = tile_max_y - (float)rows * vr;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float px = x[i], py = y[i], phw = hw[i], phh = hh[i];
        const float e_min_x = px - phw, e_max_x = px + phw;
        const float e_min_y = py - phh, e_max_y = py + phh;
        if (!(e_max_x >= tile_min_x && e_min_x < tile_max_x && e_max_y >= tile_min_y && e_min_y < tile_max_y))
            continue;
        int c0 = (int)ceilf((e_min_x - tile_min_x) / hr - 0.5f);
        int c1 = (int)(ceilf((e_max_x - tile_min_x) / hr - 0.5f) - 1.0f);
        int r0 = (int)ceilf((tile_max_y - e_max_y) / vr - 0.5f);
        int r1 = (int)(ceilf((tile_max_y - e_min_y) / vr - 0.5f) - 1.0f);
        c0 = c0 > 0 ? c0 : 0;
        c1 = c1 < cols - 1 ? c1 : cols - 1;
        r0 = r0 > 0 ? r0 : 0;
        r1 = r1 < rows - 1 ? r1 : rows - 1;
        const float v = value[i];
        for (int r = r0; r <= r1; ++r)
            for (int c = c0; c <= c1; ++c) tile[(long long)r * ld + c] = v;
    }
}

extern "C" {

int32_t nhf_rasterize(const float *x, const float *y, const float *half_w, const float *half_h, const float *value,
                      int32_t n, float tile_min_x, float tile_max_y, float hr, float vr, int32_t rows, int32_t cols,
                      float *tile, int64_t ld, void *stream) {
    if (n < 0 || rows < 0 || cols < 0 || ld < cols) return set_error_text("nhf_rasterize: negative size or ld < cols");
    if (!(hr > 0.0f) || !(vr > 0.0f)) return set_error_text("nhf_rasterize: pixel sizes must be positive");
    if (n == 0 || rows == 0 || cols == 0) return 0;
    if (!x || !y || !half_w || !half_h || !value || !tile) return set_error_text("nhf_rasterize: null array");
    const int blocks = (int)((n + 255) / 256 < 16384 ? (n + 255) / 256 : 16384);
    hipLaunchKernelGGL(k_rasterize, dim3(blocks), dim3(256), 0, (hipStream_t)stream, x, y, half_w, half_h, value, n,
                       tile_min_x, tile_max_y, hr, vr, rows, cols, tile, (long long)ld);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // extern "C"
