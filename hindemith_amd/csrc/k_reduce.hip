// a7 sum2d (Array2DSum) for gfx950.
//
// The reference tree-reduces through global temporaries fenced only by a
// local barrier (Array2DSum.cpp:81-101).  Here: pass 1 runs a fixed grid of
// HMX_SUM_PARTIALS work-groups; each thread accumulates a fixed strided
// subset in fp64, the wavefront folds with shuffles and the work-group folds
// through LDS; pass 2 (one work-group) folds the partials in a fixed order.
// The grid never depends on timing, so the result is bit-reproducible.
#include "hmx_common.h"

#include <mutex>
#include <vector>

namespace hmx {

constexpr int RED_NT = 256;

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ inline double block_sum(double v) {
    __shared__ double part[RED_NT / 64];
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) part[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < RED_NT / 64; ++i) r += part[i];
    return r;  // valid in thread 0
}

__global__ __launch_bounds__(RED_NT) void sum_pass1_kernel(const float *__restrict__ x, int64_t n,
                                                           double *__restrict__ partials) {
    double acc = 0.0;
    const int64_t stride = (int64_t)gridDim.x * RED_NT;
    for (int64_t i = (int64_t)blockIdx.x * RED_NT + threadIdx.x; i < n; i += stride)
        acc += (double)x[i];
    const double r = block_sum(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

__global__ __launch_bounds__(RED_NT) void sum_pass2_kernel(const double *__restrict__ partials,
                                                           int np, float *__restrict__ out,
                                                           double *__restrict__ out64) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < np; i += RED_NT) acc += partials[i];
    const double r = block_sum(acc);
    if (threadIdx.x == 0) {
        if (out) out[0] = (float)r;
        if (out64) out64[0] = r;
    }
}

int launch_sum(const float *x, int64_t n, double *partials, float *out, double *out64,
               hipStream_t s) {
    hipLaunchKernelGGL(sum_pass1_kernel, dim3(HMX_SUM_PARTIALS), dim3(RED_NT), 0, s, x, n,
                       partials);
    HMX_LAUNCH_CHECK("sum_pass1");
    hipLaunchKernelGGL(sum_pass2_kernel, dim3(1), dim3(RED_NT), 0, s, partials,
                       HMX_SUM_PARTIALS, out, out64);
    HMX_LAUNCH_CHECK("sum_pass2");
    return HMX_OK;
}

}  // namespace hmx

extern "C" int hmx_sum2d(const float *x, int64_t n, double *partials, float *out, void *stream) {
    HMX_REQUIRE(partials && out, "hmx_sum2d: null partials/out");
    HMX_REQUIRE(n >= 0 && (n == 0 || x), "hmx_sum2d: bad input");
    return hmx::launch_sum(x, n, partials, out, nullptr, hmx::as_stream(stream));
}

// ---------------------------------------------------------------------------
// Allocation-free form of hmx_sum2d.  The pass-1 partials live in a
// file-scope device array instead of a caller workspace; every GPU the
// library runs on holds its own copy until the process ends.  In host code
// the array's name is only a stand-in: the device address of this GPU's copy
// comes from hipGetSymbolAddress, taken once per device on the first call
// with that device current, and is the only form host code passes on.
static __device__ __align__(256) double scratch[HMX_SUM_PARTIALS];

namespace {
std::mutex g_scratch_mu;
std::vector<double *> g_scratch_ptr;  // by device ordinal; null until set up
}  // namespace

static int sum_scratch(double **out) {
    int dev = 0;
    HMX_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HMX_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return hmx::fail(HMX_ERR_HIP, "hmx_sum2d_noalloc: no device address for scratch");
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return HMX_OK;
}

extern "C" int hmx_sum2d_noalloc(const float *x, int64_t n, float *out, void *stream) {
    HMX_REQUIRE(out, "hmx_sum2d_noalloc: null out");
    HMX_REQUIRE(n >= 0 && (n == 0 || x), "hmx_sum2d_noalloc: bad input");
    // the partials are one per pass-1 work-group of the fixed grid, whatever n is
    static_assert(sizeof(scratch) == sizeof(double) * HMX_SUM_PARTIALS, "scratch size");
    double *partials = nullptr;
    int rc = sum_scratch(&partials);
    if (rc) return rc;
    return hmx::launch_sum(x, n, partials, out, nullptr, hmx::as_stream(stream));
}
