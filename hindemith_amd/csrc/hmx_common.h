// Shared definitions for libhmx (MI355X / gfx950 Horn-Schunck hot path).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/hmx.h"

namespace hmx {

// Thread-local error text returned by hmx_last_error().
void set_error(const std::string &msg);
const std::string &get_error();

// Status helpers: every entry point returns HMX_OK or one of the codes below
// after recording a message.
int fail(int code, const std::string &msg);

#define HMX_HIP_TRY(expr)                                                       \
    do {                                                                        \
        hipError_t _e = (expr);                                                 \
        if (_e != hipSuccess)                                                   \
            return ::hmx::fail(HMX_ERR_HIP, std::string(#expr) + ": " +         \
                                                hipGetErrorString(_e));         \
    } while (0)

#define HMX_LAUNCH_CHECK(what)                                                  \
    do {                                                                        \
        hipError_t _e = hipGetLastError();                                      \
        if (_e != hipSuccess)                                                   \
            return ::hmx::fail(HMX_ERR_HIP, std::string("launch ") + (what) +   \
                                                ": " + hipGetErrorString(_e));  \
    } while (0)

#define HMX_REQUIRE(cond, msg)                                                  \
    do {                                                                        \
        if (!(cond)) return ::hmx::fail(HMX_ERR_INVALID, (msg));                \
    } while (0)

// Stencil taps passed by value as a kernel argument (uniform -> SGPRs).
struct Taps {
    int nd;
    int rx, ry;  // max |offx|, max |offy|
    float w[HMX_MAX_TAPS];
    int ox[HMX_MAX_TAPS];
    int oy[HMX_MAX_TAPS];
};

// Fill a Taps struct from host arrays; validates nd.
int make_taps(Taps &t, int nd, const float *w, const int *ox, const int *oy);

// A 4-tap stencil inside the derivative cube: every offset in {0, 1}^2 (the
// canonical Dx, Dy, Dt), which reads only the 2 x 2 block at (c, r).
__host__ __device__ inline bool taps_in_cube(const Taps &t) {
    if (t.nd != 4) return false;
    for (int d = 0; d < 4; ++d)
        if (t.ox[d] < 0 || t.ox[d] > 1 || t.oy[d] < 0 || t.oy[d] > 1) return false;
    return true;
}

// Such a stencil applied to the block x00 = X(c, r), x10 = X(c+1, r),
// x01 = X(c, r+1), x11 = X(c+1, r+1) (clamped): each tap a select on its
// uniform offset, accumulated with fma in tap order from 0
__device__ __forceinline__ float cube_apply(const Taps &t, float x00, float x10, float x01,
                                            float x11) {
    float acc = 0.0f;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        const float lo = t.ox[d] ? x10 : x00, hi = t.ox[d] ? x11 : x01;
        acc = fmaf(t.w[d], t.oy[d] ? hi : lo, acc);
    }
    return acc;
}

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

__host__ __device__ inline int clampi(int v, int lo, int hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

// Every image plane lives in device (global) memory.  A pointer the compiler
// cannot trace to a kernel argument (read from a work descriptor in memory,
// the batched launch's per-tile pair descriptors) is generic to it, and it
// accesses it with FLAT instructions, which count against the LDS counter as
// well: every wait for such a read then also waits for the LDS traffic in
// flight, and a tile's prologue no longer keeps all its reads in flight
// (the batched launch's reciprocal-quotient build ran 1.5x the cycles of the
// same body launched with argument pointers).  These casts name the global
// address space, so the accesses are global ones.
using gfloat = __attribute__((address_space(1))) float;
__device__ __forceinline__ float ld_g(const float *p) { return *(const gfloat *)p; }

// A kernel's output store, written through the L2 (a system-scope store,
// sc0 sc1) when wt.  The next kernel reads the output from memory anyway (a
// kernel boundary invalidates the L2); written through, it leaves no dirty
// lines for the end of the launch to write back, which otherwise follows the
// last work-group's work.  Only whole cache lines gain: a partly written line
// becomes a partial write to memory that the L2 would have merged, so kernels
// pass wt = rows_whole_lines(W), true when every image row starts on a
// 128-byte line (a wave's 64 consecutive floats then fill whole lines).
__device__ __forceinline__ void st_out(float *p, float x, bool wt) {
    gfloat *g = (gfloat *)p;
    if (wt)
        __hip_atomic_store(g, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    else
        *g = x;
}
__host__ __device__ inline bool rows_whole_lines(int W) { return (W & 31) == 0; }

// blockIdx -> logical tile index so that the tiles one XCD receives (blocks
// are dealt round-robin over the 8 XCDs) are spatially contiguous, which lets
// neighbouring tiles share halo lines in that XCD's L2.
// rev: each XCD walks its range from the far end instead.  Launches that
// alternate rev begin with the tiles the previous launch ended with, whose
// data that XCD's L2 still holds.
__device__ inline int xcd_remap(int b, int n, bool rev = false) {
    constexpr int NX = 8;
    int per = n / NX, rem = n % NX;
    int x = b % NX, slot = b / NX;
    const int cnt = per + (x < rem ? 1 : 0);
    const int base = (x < rem) ? x * (per + 1) : rem * (per + 1) + (x - rem) * per;
    return base + (rev ? cnt - 1 - slot : slot);
}

// Device-side helpers used by several kernels.
__global__ void fill_f32_kernel(float *p, float v, int64_t n);
__global__ void fill_u32_kernel(unsigned *p, unsigned v, int64_t n);
int launch_fill_f32(float *p, float v, int64_t n, hipStream_t s);
int launch_fill_u32(unsigned *p, unsigned v, int64_t n, hipStream_t s);

}  // namespace hmx
