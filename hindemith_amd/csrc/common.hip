// Error state, small fill kernels and tap packing shared by all entry points.
#include "hmx_common.h"

#include <cstring>

namespace hmx {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
const std::string &get_error() { return g_last_error; }

int fail(int code, const std::string &msg) {
    set_error(msg);
    return code;
}

int make_taps(Taps &t, int nd, const float *w, const int *ox, const int *oy) {
    std::memset(&t, 0, sizeof(t));
    if (nd < 1 || nd > HMX_MAX_TAPS)
        return fail(HMX_ERR_INVALID, "stencil must have 1.." + std::to_string(HMX_MAX_TAPS) +
                                         " taps, got " + std::to_string(nd));
    if (!w || !ox || !oy) return fail(HMX_ERR_INVALID, "stencil tap arrays must be non-null");
    t.nd = nd;
    for (int d = 0; d < nd; ++d) {
        t.w[d] = w[d];
        t.ox[d] = ox[d];
        t.oy[d] = oy[d];
        int ax = ox[d] < 0 ? -ox[d] : ox[d];
        int ay = oy[d] < 0 ? -oy[d] : oy[d];
        if (ax > t.rx) t.rx = ax;
        if (ay > t.ry) t.ry = ay;
    }
    return HMX_OK;
}

__global__ void fill_f32_kernel(float *p, float v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void fill_u32_kernel(unsigned *p, unsigned v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

static unsigned fill_grid(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    return (unsigned)b;
}

int launch_fill_f32(float *p, float v, int64_t n, hipStream_t s) {
    if (n <= 0) return HMX_OK;
    hipLaunchKernelGGL(fill_f32_kernel, dim3(fill_grid(n)), dim3(256), 0, s, p, v, n);
    HMX_LAUNCH_CHECK("fill_f32");
    return HMX_OK;
}

int launch_fill_u32(unsigned *p, unsigned v, int64_t n, hipStream_t s) {
    if (n <= 0) return HMX_OK;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(fill_grid(n)), dim3(256), 0, s, p, v, n);
    HMX_LAUNCH_CHECK("fill_u32");
    return HMX_OK;
}

}  // namespace hmx

extern "C" const char *hmx_last_error(void) { return hmx::get_error().c_str(); }
extern "C" const char *hmx_version(void) { return "hmx 0.1 (gfx950)"; }
