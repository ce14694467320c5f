// a17: WarpImg2D -- bilinear warp of an image by a per-pixel offset field
// (WarpImg2D.cpp:33-75), the building block of coarse-to-fine / warped HS.
//
// dst(x, y) = bilinear sample of src at (x + offx, y + offy) when all four
// taps lie inside the image, else src(x, y).  The offset splits into an
// integer part by truncation, moved down by one for negative offsets (so the
// fraction lies in [0, 1]).
//
// The generated code mixes float data with double literals (1.0); this
// kernel evaluates each statement in the type C gives it: the weights
// 1.0 - frac in double, float products where both operands are float, and
// each "out += term" as one fma in the statement's type, rounded to float on
// assignment -- the same convention as the other accumulations here.
#include "hmx_common.h"

#include <algorithm>

#pragma clang fp contract(off)

namespace hmx {

__device__ __forceinline__ float warp_pixel(const float *__restrict__ src, float ox, float oy,
                                            int r, int c, int H, int W) {
    int mx = (int)ox, my = (int)oy;
    float xf = ox - (float)mx, yf = oy - (float)my;
    if (ox < 0.0f) {
        --mx;
        xf = (float)(1.0 + (double)xf);
    }
    if (oy < 0.0f) {
        --my;
        yf = (float)(1.0 + (double)yf);
    }
    const int64_t sx = (int64_t)c + mx, sy = (int64_t)r + my;
    if (sx >= 0 && sx + 1 < W && sy >= 0 && sy + 1 < H) {
        const int64_t q = sy * W + sx;
        const float i00 = src[q], i10 = src[q + 1], i01 = src[q + W], i11 = src[q + W + 1];
        const double ax = 1.0 - (double)xf, ay = 1.0 - (double)yf;
        float out = 0.0f;
        out = (float)fma((double)i00 * ax, ay, (double)out);        // IN00*(1-xf)*(1-yf)
        const float p10 = i10 * xf;
        out = (float)fma((double)p10, ay, (double)out);             // IN10*xf*(1-yf)
        out = (float)fma((double)i01 * ax, (double)yf, (double)out);  // IN01*(1-xf)*yf
        const float p11 = i11 * xf;
        out = fmaf(p11, yf, out);                                   // IN11*xf*yf
        return out;
    }
    return src[(int64_t)r * W + c];
}

// 16 B/px: read offx, offy and (once, through the cache) src; write dst.
__global__ __launch_bounds__(256) void warp_img2d_kernel(const float *__restrict__ src,
                                                         const float *__restrict__ offx,
                                                         const float *__restrict__ offy,
                                                         float *__restrict__ dst, int H, int W) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= W) return;
    for (int r = blockIdx.y; r < H; r += gridDim.y) {
        const int64_t p = (int64_t)r * W + c;
        st_out(dst + p, warp_pixel(src, offx[p], offy[p], r, c, H, W), rows_whole_lines(W));
    }
}

int launch_warp_img2d(const float *src, const float *offx, const float *offy, float *dst, int H,
                      int W, hipStream_t s) {
    dim3 grid((W + 255) / 256, std::min(H, 65535));
    hipLaunchKernelGGL(warp_img2d_kernel, grid, dim3(256), 0, s, src, offx, offy, dst, H, W);
    HMX_LAUNCH_CHECK("warp_img2d");
    return HMX_OK;
}

}  // namespace hmx

using namespace hmx;

extern "C" int hmx_warp_img2d(const float *src, const float *offx, const float *offy, float *dst,
                              int H, int W, void *stream) {
    HMX_REQUIRE(src && offx && offy && dst, "hmx_warp_img2d: null argument");
    HMX_REQUIRE(H > 0 && W > 0, "hmx_warp_img2d: H and W must be positive");
    HMX_REQUIRE(dst != src && dst != offx && dst != offy,
                "hmx_warp_img2d: dst must be distinct from the inputs");
    return launch_warp_img2d(src, offx, offy, dst, H, W, as_stream(stream));
}
