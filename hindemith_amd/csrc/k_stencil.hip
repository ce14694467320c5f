// a1 Stencil2DMVM, a2 SpmvDIA2D and a10 HS setup kernels for gfx950.
//
// Arithmetic follows the reference's generated OpenCL: the accumulation
// "y += D * x" of Stencil2DMVM.cpp:47 / SpmvDIA2D.cpp:43 is one fused
// multiply-add per tap, in tap order, starting from y = 0.0; every other
// binary op is rounded on its own (this file is compiled with contraction off).
#include "hmx_common.h"

#pragma clang fp contract(off)

namespace hmx {

// ---------------------------------------------------------------------------
// a1: LDS-staged clamped stencil.  A 256-thread work-group (64 x 4) owns a
// 64 x 16 output tile; it stages the (16+2R) x (64+2R) input window with
// clamped coordinates, so every tap of every output pixel is an LDS read and
// the border rule of Stencil2DMVM.cpp:40-46 (clamp each coordinate) holds by
// construction: window cell (rx, ry) holds x[clamp(x0-R+rx), clamp(y0-R+ry)].
constexpr int ST_TW = 64, ST_TH = 16;

template <int R, int NTAP>  // NTAP > 0: exactly NTAP taps, unrolled; 0: any count
__global__ __launch_bounds__(256) void stencil2d_lds_kernel(const float *__restrict__ x,
                                                            float *__restrict__ y, int H, int W,
                                                            Taps taps) {
    constexpr int LW = ST_TW + 2 * R, LH = ST_TH + 2 * R;
    __shared__ float tile[LH * LW];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int x0 = blockIdx.x * ST_TW, y0 = blockIdx.y * ST_TH;
    for (int i = threadIdx.x; i < LH * LW; i += 256) {
        int ry = i / LW, rx = i - ry * LW;
        int gx = clampi(x0 - R + rx, 0, W - 1), gy = clampi(y0 - R + ry, 0, H - 1);
        tile[i] = x[(int64_t)gy * W + gx];
    }
    // taps hoisted into registers once: LDS delta and weight
    constexpr int NT = NTAP ? NTAP : 1;
    int dl[NT];
    float wt[NT];
#pragma unroll
    for (int d = 0; d < NT; ++d) {
        dl[d] = taps.oy[d] * LW + taps.ox[d];
        wt[d] = taps.w[d];
    }
    __syncthreads();
    const int gx = x0 + tx;
    if (gx >= W) return;
    for (int r = ty; r < ST_TH; r += 4) {
        const int gy = y0 + r;
        if (gy >= H) break;
        const float *base = tile + (r + R) * LW + tx + R;
        float acc = 0.0f;
        if (NTAP) {
#pragma unroll
            for (int d = 0; d < NT; ++d) acc = fmaf(wt[d], base[dl[d]], acc);
        } else {
            for (int d = 0; d < taps.nd; ++d)
                acc = fmaf(taps.w[d], base[taps.oy[d] * LW + taps.ox[d]], acc);
        }
        y[(int64_t)gy * W + gx] = acc;
    }
}

// Radius-1 stencils: a thread takes one column and ST_V consecutive rows and
// reads every tap straight from global memory (the neighbours' lines are L1
// hits), all ST_V x NTAP loads independent, so a wave keeps many reads in
// flight; no LDS staging and no barrier.  Faster than the LDS tile at this
// radius (2048^2 5-point: 7.7 vs 10.4 us per launch).
constexpr int ST_V = 8;

template <int NTAP>  // NTAP > 0: exactly NTAP taps, unrolled; 0: any count
__global__ __launch_bounds__(256) void stencil2d_cols_kernel(const float *__restrict__ x,
                                                             float *__restrict__ y, int H,
                                                             int W, Taps taps) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    const int r0 = blockIdx.y * ST_V;
    if (c >= W) return;
    float acc[ST_V];
#pragma unroll
    for (int v = 0; v < ST_V; ++v) acc[v] = 0.0f;
    auto tap = [&](int d) {
        const int xx = clampi(c + taps.ox[d], 0, W - 1);
        const float w = taps.w[d];
#pragma unroll
        for (int v = 0; v < ST_V; ++v) {
            const int yy = clampi(r0 + v + taps.oy[d], 0, H - 1);
            acc[v] = fmaf(w, x[(int64_t)yy * W + xx], acc[v]);
        }
    };
    if (NTAP > 0) {
#pragma unroll
        for (int d = 0; d < (NTAP > 0 ? NTAP : 1); ++d) tap(d);
    } else {
        for (int d = 0; d < taps.nd; ++d) tap(d);
    }
#pragma unroll
    for (int v = 0; v < ST_V; ++v)
        if (r0 + v < H) st_out(y + (int64_t)(r0 + v) * W + c, acc[v], rows_whole_lines(W));
}

// Wide stencils (radius > 4): taps straight from global memory (L1/L2 serve
// the reuse), still clamped per coordinate.
__global__ __launch_bounds__(256) void stencil2d_direct_kernel(const float *__restrict__ x,
                                                               float *__restrict__ y, int H,
                                                               int W, Taps taps) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= W) return;
    for (int r = blockIdx.y; r < H; r += gridDim.y) {
        float acc = 0.0f;
        for (int d = 0; d < taps.nd; ++d) {
            int xx = clampi(c + taps.ox[d], 0, W - 1), yy = clampi(r + taps.oy[d], 0, H - 1);
            acc = fmaf(taps.w[d], x[(int64_t)yy * W + xx], acc);
        }
        y[(int64_t)r * W + c] = acc;
    }
}

int launch_stencil2d(const float *x, float *y, int H, int W, const Taps &t, hipStream_t s) {
    const int R = t.rx > t.ry ? t.rx : t.ry;
    if (R <= 1) {
        using KFn = void (*)(const float *, float *, int, int, Taps);
        static const KFn table[5] = {stencil2d_cols_kernel<4>, stencil2d_cols_kernel<5>,
                                     stencil2d_cols_kernel<8>, stencil2d_cols_kernel<9>,
                                     stencil2d_cols_kernel<0>};
        const int ci = t.nd == 4 ? 0 : t.nd == 5 ? 1 : t.nd == 8 ? 2 : t.nd == 9 ? 3 : 4;
        dim3 grid((W + 255) / 256, (H + ST_V - 1) / ST_V);
        hipLaunchKernelGGL(table[ci], grid, dim3(256), 0, s, x, y, H, W, t);
    } else if (R <= 4) {
        dim3 grid((W + ST_TW - 1) / ST_TW, (H + ST_TH - 1) / ST_TH);
        using KFn = void (*)(const float *, float *, int, int, Taps);
#define ST_ROW(RR) {stencil2d_lds_kernel<RR, 4>, stencil2d_lds_kernel<RR, 5>,          \
                    stencil2d_lds_kernel<RR, 8>, stencil2d_lds_kernel<RR, 9>,          \
                    stencil2d_lds_kernel<RR, 0>}
        static const KFn table[3][5] = {ST_ROW(1), ST_ROW(2), ST_ROW(4)};
#undef ST_ROW
        const int ri = R <= 1 ? 0 : (R <= 2 ? 1 : 2);
        const int ci = t.nd == 4 ? 0 : t.nd == 5 ? 1 : t.nd == 8 ? 2 : t.nd == 9 ? 3 : 4;
        hipLaunchKernelGGL(table[ri][ci], grid, dim3(256), 0, s, x, y, H, W, t);
    } else {
        dim3 grid((W + 255) / 256, H < 65535 ? H : 65535);
        hipLaunchKernelGGL(stencil2d_direct_kernel, grid, dim3(256), 0, s, x, y, H, W, t);
    }
    HMX_LAUNCH_CHECK("stencil2d");
    return HMX_OK;
}

// ---------------------------------------------------------------------------
// a2: 2D DIA SpMV.  y[p] = sum_d A[d][p+o_d] * x[p+o_d] over in-range p+o_d
// (SpmvDIA2D.cpp:41-43; coefficient index SpMat2D.cpp:214).  Each coefficient
// plane element is read exactly once over the whole launch, so the planes are
// streamed straight from HBM; the nd shifted reads of x hit in L1/L2.
// ND > 0: exactly ND diagonals, all loads of a pixel issued before its fma
// chain; ND == 0: any count (runtime loop).
template <int ND>
__global__ __launch_bounds__(256) void dia2d_kernel(const float *__restrict__ A,
                                                    const float *__restrict__ x,
                                                    float *__restrict__ y, int H, int W,
                                                    Taps taps) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= W) return;
    const int64_t HW = (int64_t)H * W;
    for (int r = blockIdx.y; r < H; r += gridDim.y) {
        const int64_t p = (int64_t)r * W + c;
        float acc = 0.0f;
        if (ND) {
            float av[ND ? ND : 1], xv[ND ? ND : 1];
            bool in[ND ? ND : 1];
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                const int xx = c + taps.ox[d], yy = r + taps.oy[d];
                in[d] = xx >= 0 && xx < W && yy >= 0 && yy < H;
                const int64_t q = in[d] ? (int64_t)yy * W + xx : p;
                av[d] = A[d * HW + q];
                xv[d] = x[q];
            }
#pragma unroll
            for (int d = 0; d < ND; ++d)
                if (in[d]) acc = fmaf(av[d], xv[d], acc);
        } else {
            for (int d = 0; d < taps.nd; ++d) {
                const int xx = c + taps.ox[d], yy = r + taps.oy[d];
                if (xx >= 0 && xx < W && yy >= 0 && yy < H) {
                    const int64_t q = (int64_t)yy * W + xx;
                    acc = fmaf(A[d * HW + q], x[q], acc);
                }
            }
        }
        st_out(y + p, acc, rows_whole_lines(W));
    }
}

// ---------------------------------------------------------------------------
// a10: the canonical derivative cube in one pass (20 B/px: read I1, I2;
// write Ix, Iy, It).  Same per-op rounding as six Stencil2DMVM launches and
// three Array2DOps.
__device__ inline float tap_apply(const float *__restrict__ x, int H, int W, int r, int c,
                                  const Taps &t) {
    float acc = 0.0f;
    for (int d = 0; d < t.nd; ++d) {
        int xx = clampi(c + t.ox[d], 0, W - 1), yy = clampi(r + t.oy[d], 0, H - 1);
        acc = fmaf(t.w[d], x[(int64_t)yy * W + xx], acc);
    }
    return acc;
}

__global__ __launch_bounds__(256) void hs_setup_kernel(const float *__restrict__ I1,
                                                       const float *__restrict__ I2, int H,
                                                       int W, Taps dx, Taps dy, Taps dt,
                                                       float *__restrict__ Ix,
                                                       float *__restrict__ Iy,
                                                       float *__restrict__ It) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= W) return;
    for (int r = blockIdx.y; r < H; r += gridDim.y) {
        const int64_t p = (int64_t)r * W + c;
        float a1 = tap_apply(I1, H, W, r, c, dx);
        float a2 = tap_apply(I2, H, W, r, c, dx);
        st_out(Ix + p, a1 + a2, rows_whole_lines(W));
        float b1 = tap_apply(I1, H, W, r, c, dy);
        float b2 = tap_apply(I2, H, W, r, c, dy);
        st_out(Iy + p, b1 + b2, rows_whole_lines(W));
        float t2 = tap_apply(I2, H, W, r, c, dt);
        float t1 = tap_apply(I1, H, W, r, c, dt);
        st_out(It + p, t2 - t1, rows_whole_lines(W));
    }
}

// The same cube for three 4-tap stencils (the canonical Dx, Dy, Dt): a thread
// takes one column and SETUP_V rows, its taps unrolled, so all of a row's 24
// reads (mostly L1 hits) are issued before the fma chains instead of one
// read-and-wait per tap of a runtime loop.
constexpr int SETUP_V = 4;

__device__ __forceinline__ float tap_apply4(const float *__restrict__ x, int H, int W, int r,
                                            int c, const Taps &t) {
    float v[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        const int xx = clampi(c + t.ox[d], 0, W - 1), yy = clampi(r + t.oy[d], 0, H - 1);
        v[d] = x[(int64_t)yy * W + xx];
    }
    float acc = 0.0f;
#pragma unroll
    for (int d = 0; d < 4; ++d) acc = fmaf(t.w[d], v[d], acc);
    return acc;
}

// COPY: the same pass also copies the solve's start state (u, v) into the
// work planes (wu, wv), where the first Jacobi launch reads it when the
// launch count is odd.
template <bool COPY>
__global__ __launch_bounds__(256) void hs_setup4_kernel(const float *__restrict__ I1,
                                                        const float *__restrict__ I2, int H,
                                                        int W, Taps dx, Taps dy, Taps dt,
                                                        float *__restrict__ Ix,
                                                        float *__restrict__ Iy,
                                                        float *__restrict__ It,
                                                        const float *__restrict__ u,
                                                        const float *__restrict__ v,
                                                        float *__restrict__ wu,
                                                        float *__restrict__ wv) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= W) return;
#pragma unroll
    for (int q = 0; q < SETUP_V; ++q) {
        const int r = blockIdx.y * SETUP_V + q;
        if (r >= H) break;
        const int64_t p = (int64_t)r * W + c;
        const float a1 = tap_apply4(I1, H, W, r, c, dx);
        const float a2 = tap_apply4(I2, H, W, r, c, dx);
        const float b1 = tap_apply4(I1, H, W, r, c, dy);
        const float b2 = tap_apply4(I2, H, W, r, c, dy);
        const float t2 = tap_apply4(I2, H, W, r, c, dt);
        const float t1 = tap_apply4(I1, H, W, r, c, dt);
        st_out(Ix + p, a1 + a2, rows_whole_lines(W));
        st_out(Iy + p, b1 + b2, rows_whole_lines(W));
        st_out(It + p, t2 - t1, rows_whole_lines(W));
        if (COPY) {
            st_out(wu + p, u[p], rows_whole_lines(W));
            st_out(wv + p, v[p], rows_whole_lines(W));
        }
    }
}

// The canonical cube's stencils all read the 2 x 2 block (c, r), (c+1, r),
// (c, r+1), (c+1, r+1) (offsets in {0, 1}^2).  For that case a thread takes
// one column and CUBE_V rows and reads the (CUBE_V + 1)-row, 2-column
// neighbourhood of each frame once into registers (row r+1 of one output row
// is row r of the next); each tap is then a register select on its
// work-group-uniform offset, accumulated with fma in tap order from 0 as in
// Stencil2DMVM.cpp:37-47.  Reads per pixel fall from 24 (L1-served) to
// 4 (CUBE_V + 1) / CUBE_V.
constexpr int CUBE_V = 8;

// (taps_in_cube and cube_apply: hmx_common.h)

template <bool COPY>
__global__ __launch_bounds__(256) void hs_setup_cube_kernel(
    const float *__restrict__ I1, const float *__restrict__ I2, int H, int W, Taps dx, Taps dy,
    Taps dt, float *__restrict__ Ix, float *__restrict__ Iy, float *__restrict__ It,
    const float *__restrict__ u, const float *__restrict__ v, float *__restrict__ wu,
    float *__restrict__ wv) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= W) return;
    const int c1 = min(c + 1, W - 1);
    const int r0 = blockIdx.y * CUBE_V;
    float a0[CUBE_V + 1], a1[CUBE_V + 1], b0[CUBE_V + 1], b1[CUBE_V + 1];
#pragma unroll
    for (int j = 0; j <= CUBE_V; ++j) {
        const int64_t row = (int64_t)min(r0 + j, H - 1) * W;
        a0[j] = I1[row + c];
        a1[j] = I1[row + c1];
        b0[j] = I2[row + c];
        b1[j] = I2[row + c1];
    }
#pragma unroll
    for (int q = 0; q < CUBE_V; ++q) {
        const int r = r0 + q;
        if (r >= H) break;
        const int64_t p = (int64_t)r * W + c;
        const float e1 = cube_apply(dx, a0[q], a1[q], a0[q + 1], a1[q + 1]);
        const float e2 = cube_apply(dx, b0[q], b1[q], b0[q + 1], b1[q + 1]);
        const float f1 = cube_apply(dy, a0[q], a1[q], a0[q + 1], a1[q + 1]);
        const float f2 = cube_apply(dy, b0[q], b1[q], b0[q + 1], b1[q + 1]);
        const float g2 = cube_apply(dt, b0[q], b1[q], b0[q + 1], b1[q + 1]);
        const float g1 = cube_apply(dt, a0[q], a1[q], a0[q + 1], a1[q + 1]);
        st_out(Ix + p, e1 + e2, rows_whole_lines(W));
        st_out(Iy + p, f1 + f2, rows_whole_lines(W));
        st_out(It + p, g2 - g1, rows_whole_lines(W));
        if (COPY) {
            st_out(wu + p, u[p], rows_whole_lines(W));
            st_out(wv + p, v[p], rows_whole_lines(W));
        }
    }
}

int launch_hs_copy2(const float *u, const float *v, float *wu, float *wv, int64_t n,
                    hipStream_t s);

// u non-null: also copy (u, v) into (wu, wv) (see hs_setup4_kernel).
int launch_hs_setup(const float *I1, const float *I2, int H, int W, const Taps &dx,
                    const Taps &dy, const Taps &dt, float *Ix, float *Iy, float *It,
                    hipStream_t s, const float *u, const float *v, float *wu, float *wv) {
    if (taps_in_cube(dx) && taps_in_cube(dy) && taps_in_cube(dt) &&
        (H + CUBE_V - 1) / CUBE_V <= 65535) {
        dim3 grid((W + 255) / 256, (H + CUBE_V - 1) / CUBE_V);
        if (u)
            hipLaunchKernelGGL(hs_setup_cube_kernel<true>, grid, dim3(256), 0, s, I1, I2, H, W,
                               dx, dy, dt, Ix, Iy, It, u, v, wu, wv);
        else
            hipLaunchKernelGGL(hs_setup_cube_kernel<false>, grid, dim3(256), 0, s, I1, I2, H, W,
                               dx, dy, dt, Ix, Iy, It, nullptr, nullptr, nullptr, nullptr);
    } else if (dx.nd == 4 && dy.nd == 4 && dt.nd == 4 && (H + SETUP_V - 1) / SETUP_V <= 65535) {
        dim3 grid((W + 255) / 256, (H + SETUP_V - 1) / SETUP_V);
        if (u)
            hipLaunchKernelGGL(hs_setup4_kernel<true>, grid, dim3(256), 0, s, I1, I2, H, W, dx,
                               dy, dt, Ix, Iy, It, u, v, wu, wv);
        else
            hipLaunchKernelGGL(hs_setup4_kernel<false>, grid, dim3(256), 0, s, I1, I2, H, W, dx,
                               dy, dt, Ix, Iy, It, nullptr, nullptr, nullptr, nullptr);
    } else {
        dim3 grid((W + 255) / 256, H < 65535 ? H : 65535);
        hipLaunchKernelGGL(hs_setup_kernel, grid, dim3(256), 0, s, I1, I2, H, W, dx, dy, dt, Ix,
                           Iy, It);
        HMX_LAUNCH_CHECK("hs_setup");
        if (u) return launch_hs_copy2(u, v, wu, wv, (int64_t)H * W, s);
    }
    HMX_LAUNCH_CHECK("hs_setup");
    return HMX_OK;
}

int launch_dia2d(const float *A, const float *x, float *y, int H, int W, const Taps &t,
                 hipStream_t s) {
    dim3 grid((W + 255) / 256, H < 65535 ? H : 65535);
    if (t.nd == 5)
        hipLaunchKernelGGL(dia2d_kernel<5>, grid, dim3(256), 0, s, A, x, y, H, W, t);
    else if (t.nd == 9)
        hipLaunchKernelGGL(dia2d_kernel<9>, grid, dim3(256), 0, s, A, x, y, H, W, t);
    else
        hipLaunchKernelGGL(dia2d_kernel<0>, grid, dim3(256), 0, s, A, x, y, H, W, t);
    HMX_LAUNCH_CHECK("dia2d");
    return HMX_OK;
}

}  // namespace hmx

using namespace hmx;

extern "C" int hmx_stencil2d(const float *x, float *y, int H, int W, int nd, const float *w,
                             const int *offx, const int *offy, void *stream) {
    HMX_REQUIRE(x && y, "hmx_stencil2d: null array");
    HMX_REQUIRE(x != y, "hmx_stencil2d: x and y must be distinct buffers");
    HMX_REQUIRE(H > 0 && W > 0, "hmx_stencil2d: H and W must be positive");
    Taps t;
    int rc = make_taps(t, nd, w, offx, offy);
    if (rc) return rc;
    return launch_stencil2d(x, y, H, W, t, as_stream(stream));
}

extern "C" int hmx_dia2d_spmv(const float *A, int nd, const int *offx, const int *offy,
                              const float *x, float *y, int H, int W, void *stream) {
    HMX_REQUIRE(A && x && y, "hmx_dia2d_spmv: null array");
    HMX_REQUIRE(x != y, "hmx_dia2d_spmv: x and y must be distinct buffers");
    HMX_REQUIRE(H > 0 && W > 0, "hmx_dia2d_spmv: H and W must be positive");
    static const float unused_w[HMX_MAX_TAPS] = {0};  // DIA weights live in A
    Taps t;
    int rc = make_taps(t, nd, unused_w, offx, offy);
    if (rc) return rc;
    return launch_dia2d(A, x, y, H, W, t, as_stream(stream));
}

extern "C" int hmx_dia_spmv(const float *A, int nd, const int *offsets, const float *x,
                            float *y, int n, void *stream) {
    HMX_REQUIRE(A && x && y && offsets, "hmx_dia_spmv: null argument");
    HMX_REQUIRE(x != y, "hmx_dia_spmv: x and y must be distinct buffers");
    HMX_REQUIRE(n > 0, "hmx_dia_spmv: n must be positive");
    HMX_REQUIRE(nd >= 1 && nd <= HMX_MAX_TAPS,
                "hmx_dia_spmv: 1.." + std::to_string(HMX_MAX_TAPS) + " diagonals");
    // a 1 x n grid with no vertical offsets: the 2D DIA rule (coefficient at
    // the neighbour index, out-of-range terms skipped) is the 1-D one
    static const float unused_w[HMX_MAX_TAPS] = {0};
    int zeros[HMX_MAX_TAPS] = {0};
    Taps t;
    int rc = make_taps(t, nd, unused_w, offsets, zeros);
    if (rc) return rc;
    return launch_dia2d(A, x, y, 1, n, t, as_stream(stream));
}

extern "C" int hmx_hs_setup(const float *I1, const float *I2, int H, int W, int ndx,
                            const float *wx, const int *oxx, const int *oyx, int ndy,
                            const float *wy, const int *oxy, const int *oyy, int ndt,
                            const float *wt, const int *oxt, const int *oyt, float *Ix,
                            float *Iy, float *It, void *stream) {
    HMX_REQUIRE(I1 && I2 && Ix && Iy && It, "hmx_hs_setup: null array");
    HMX_REQUIRE(H > 0 && W > 0, "hmx_hs_setup: H and W must be positive");
    Taps dx, dy, dt;
    int rc = make_taps(dx, ndx, wx, oxx, oyx);
    if (!rc) rc = make_taps(dy, ndy, wy, oxy, oyy);
    if (!rc) rc = make_taps(dt, ndt, wt, oxt, oyt);
    if (rc) return rc;
    return launch_hs_setup(I1, I2, H, W, dx, dy, dt, Ix, Iy, It, as_stream(stream), nullptr,
                           nullptr, nullptr, nullptr);
}
