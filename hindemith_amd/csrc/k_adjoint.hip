// a13: adjoint of the Horn-Schunck solve for gfx950 (no counterpart in the
// reference, which has no differentiation).
//
// Forward (per pair):  Ix, Iy, It from the derivative cube of (I1, I2);
//   den = alpha^2 + Ix^2 + Iy^2;  x_{k+1} = F(x_k), x = (u, v), for N steps.
// Reverse step, given g = dL/dx_{k+1} and the state x_k:
//   ubar = A u_k, vbar = A v_k, t = (Ix ubar + Iy vbar + It) / den
//   s = -(Ix gu + Iy gv) / den
//   g_ubar = gu + Ix s,  g_vbar = gv + Iy s
//   dL/dx_k = (A^T g_ubar, A^T g_vbar)
//   gIx += -gu t + s ubar - 2 s t Ix,  gIy += -gv t + s vbar - 2 s t Iy,
//   gIt += s,  g(alpha^2) += sum(-s t)
// then dL/dI1 = Dx^T gIx + Dy^T gIy - Dt^T gIt, dL/dI2 = ... + Dt^T gIt and
// dL/dalpha = 2 alpha g(alpha^2).
//
// A^T of a clamped stencil is not the stencil: a border pixel receives every
// tap that clamps onto it.  Along one axis, clamp(q + o) = p holds for q in
//   [p == 0 ? 0 : p - o,  p == n-1 ? n-1 : p - o]  intersected with [0, n-1],
// so the transpose is a gather over at most (|o|+1)^2 sources per tap.
//
// Memory: states are checkpointed every C iterations in the forward pass; the
// reverse pass recomputes one segment of C states at a time (with the same
// fused kernel, so the recomputed states are bit-identical), then walks it
// backwards.
#include "hmx_common.h"

#include <utility>

#pragma clang fp contract(off)

namespace hmx {

int launch_hs_tb(int variant, const float *Ix, const float *Iy, const float *It,
                 const float *den, float alpha2, int H, int W, const Taps &taps, int k,
                 const float *u_in, const float *v_in, float *u_out, float *v_out,
                 hipStream_t s, int B, const float *alpha2_b);
int hs_jacobi_run(const float *Ix, const float *Iy, const float *It, const float *den,
                  float alpha2, int H, int W, const Taps &taps, int iters, float *u, float *v,
                  float *wu, float *wv, int variant, int k, hipStream_t s, int B,
                  const float *alpha2_b);
void hs_choice(int H, int W, int &variant, int &k);

__device__ inline void tr_range(int p, int o, int n, int &lo, int &hi) {
    lo = (p == 0) ? 0 : p - o;
    hi = (p == n - 1) ? n - 1 : p - o;
    lo = lo < 0 ? 0 : lo;
    hi = hi > n - 1 ? n - 1 : hi;
}

template <typename T>
__device__ inline T tap_apply_t(const T *__restrict__ x, int H, int W, int r, int c,
                                const Taps &t) {
    T acc = 0;
    for (int d = 0; d < t.nd; ++d) {
        int xx = clampi(c + t.ox[d], 0, W - 1), yy = clampi(r + t.oy[d], 0, H - 1);
        acc = fma((T)t.w[d], x[(int64_t)yy * W + xx], acc);
    }
    return acc;
}

// sum over taps d and sources q with clamp(q + o_d) = p of w_d * g[q]
__device__ inline float tap_apply_T(const float *__restrict__ g, int H, int W, int r, int c,
                                    const Taps &t) {
    float acc = 0.0f;
    for (int d = 0; d < t.nd; ++d) {
        int lx, hx, ly, hy;
        tr_range(c, t.ox[d], W, lx, hx);
        tr_range(r, t.oy[d], H, ly, hy);
        for (int qy = ly; qy <= hy; ++qy)
            for (int qx = lx; qx <= hx; ++qx) acc = fmaf(t.w[d], g[(int64_t)qy * W + qx], acc);
    }
    return acc;
}

// ---------------------------------------------------------------------------
// Reverse Jacobi step, fused: one pass reads x_k (halo 1), g (halo 1), Ix, Iy,
// It and the three accumulators, writes dL/dx_k and the accumulators.

struct AdjArgs {
    const float *Ix, *Iy, *It;
    const float *u, *v;    // x_k
    const float *gu, *gv;  // dL/dx_{k+1}
    float *gu_out, *gv_out;
    float *gIx, *gIy, *gIt;
    double *ga2;           // [pair][tile] running sums of -s t
    const float *alpha2_b;
    int H, W, tiles_x, ntiles;
    int64_t bstride;
};

constexpr int ADJ_TW = 64, ADJ_TH = 8, ADJ_NT = 256;

__global__ __launch_bounds__(ADJ_NT) void hs_adj_kernel(AdjArgs a, Taps taps) {
    constexpr int RW = ADJ_TW + 2, RH = ADJ_TH + 2, NPX = RW * RH;
    __shared__ float2 xs[NPX];  // (u_k, v_k) at clamped coordinates
    __shared__ float2 gs[NPX];  // (g_ubar, g_vbar) at real pixels (0 outside)
    __shared__ double red[ADJ_NT / 64];

    const int b = blockIdx.y;
    const int64_t off = (int64_t)b * a.bstride;
    const float *Ix = a.Ix + off, *Iy = a.Iy + off, *It = a.It + off;
    const float *U = a.u + off, *V = a.v + off, *GU = a.gu + off, *GV = a.gv + off;
    const float alpha2 = a.alpha2_b[b];
    const int H = a.H, W = a.W;
    const int t = xcd_remap(blockIdx.x, a.ntiles);
    const int tyi = t / a.tiles_x, txi = t - tyi * a.tiles_x;
    const int x0 = txi * ADJ_TW, y0 = tyi * ADJ_TH;
    const int gx0 = x0 - 1, gy0 = y0 - 1;

    for (int i = threadIdx.x; i < NPX; i += ADJ_NT) {
        const int ry = i / RW, rx = i - ry * RW;
        const int gx = gx0 + rx, gy = gy0 + ry;
        const int64_t gc = (int64_t)clampi(gy, 0, H - 1) * W + clampi(gx, 0, W - 1);
        xs[i] = make_float2(U[gc], V[gc]);
        float2 g = make_float2(0.0f, 0.0f);
        if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
            const float ix = Ix[gc], iy = Iy[gc], gu = GU[gc], gv = GV[gc];
            const float den = (alpha2 + ix * ix) + iy * iy;
            const float s = -(ix * gu + iy * gv) / den;
            g = make_float2(gu + ix * s, gv + iy * s);
        }
        gs[i] = g;
    }
    __syncthreads();

    double acc_a2 = 0.0;
    for (int i = threadIdx.x; i < ADJ_TW * ADJ_TH; i += ADJ_NT) {
        const int ly = i / ADJ_TW, lx = i - ly * ADJ_TW;
        const int gx = x0 + lx, gy = y0 + ly;
        if (gx >= W || gy >= H) continue;
        const int rx = lx + 1, ry = ly + 1, ri = ry * RW + rx;
        float ub = 0.0f, vb = 0.0f;
        for (int d = 0; d < taps.nd; ++d) {
            const float2 q = xs[ri + taps.oy[d] * RW + taps.ox[d]];
            ub = fmaf(taps.w[d], q.x, ub);
            vb = fmaf(taps.w[d], q.y, vb);
        }
        const int64_t p = (int64_t)gy * W + gx;
        const float ix = Ix[p], iy = Iy[p], it = It[p], gu = GU[p], gv = GV[p];
        const float den = (alpha2 + ix * ix) + iy * iy;
        const float tt = ((ix * ub + iy * vb) + it) / den;
        const float s = -(ix * gu + iy * gv) / den;
        const float st2 = 2.0f * s * tt;
        a.gIx[off + p] += (-gu * tt + s * ub) - st2 * ix;
        a.gIy[off + p] += (-gv * tt + s * vb) - st2 * iy;
        a.gIt[off + p] += s;
        acc_a2 += (double)(-s * tt);
        // dL/dx_k = A^T g
        float au = 0.0f, av = 0.0f;
        if (gx > 0 && gx < W - 1 && gy > 0 && gy < H - 1) {
            for (int d = 0; d < taps.nd; ++d) {
                const float2 q = gs[ri - taps.oy[d] * RW - taps.ox[d]];
                au = fmaf(taps.w[d], q.x, au);
                av = fmaf(taps.w[d], q.y, av);
            }
        } else {
            for (int d = 0; d < taps.nd; ++d) {
                int lox, hix, loy, hiy;
                tr_range(gx, taps.ox[d], W, lox, hix);
                tr_range(gy, taps.oy[d], H, loy, hiy);
                for (int qy = loy; qy <= hiy; ++qy)
                    for (int qx = lox; qx <= hix; ++qx) {
                        const float2 q = gs[(qy - gy0) * RW + (qx - gx0)];
                        au = fmaf(taps.w[d], q.x, au);
                        av = fmaf(taps.w[d], q.y, av);
                    }
            }
        }
        a.gu_out[off + p] = au;
        a.gv_out[off + p] = av;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc_a2 += __shfl_xor(acc_a2, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc_a2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int i = 0; i < ADJ_NT / 64; ++i) r += red[i];
        a.ga2[(int64_t)b * a.ntiles + t] += r;
    }
}

// dL/dI1, dL/dI2 from the derivative-cube accumulators.
__global__ __launch_bounds__(256) void hs_setup_T_kernel(const float *__restrict__ gIx,
                                                         const float *__restrict__ gIy,
                                                         const float *__restrict__ gIt, int H,
                                                         int W, int64_t bstride, Taps dx, Taps dy,
                                                         Taps dt, float *__restrict__ gI1,
                                                         float *__restrict__ gI2) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= W) return;
    const int64_t off = (int64_t)blockIdx.z * bstride;
    for (int r = blockIdx.y; r < H; r += gridDim.y) {
        const int64_t p = off + (int64_t)r * W + c;
        const float a = tap_apply_T(gIx + off, H, W, r, c, dx);
        const float b = tap_apply_T(gIy + off, H, W, r, c, dy);
        const float d = tap_apply_T(gIt + off, H, W, r, c, dt);
        const float ab = a + b;
        gI1[p] = ab - d;
        gI2[p] = ab + d;
    }
}

__global__ void alpha_sq_kernel(const float *alpha, float *alpha2, int B) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < B) alpha2[b] = alpha[b] * alpha[b];
}

__global__ void galpha_kernel(const double *__restrict__ ga2, int ntiles, const float *alpha,
                              float *galpha, int B) {
    const int b = blockIdx.x;
    if (b >= B || threadIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < ntiles; ++i) s += ga2[(int64_t)b * ntiles + i];
    galpha[b] = (float)(2.0 * (double)alpha[b] * s);
}

// Batched setup: one launch per pair of the single-pair kernel.
int launch_hs_setup(const float *I1, const float *I2, int H, int W, const Taps &dx,
                    const Taps &dy, const Taps &dt, float *Ix, float *Iy, float *It,
                    hipStream_t s);

static int setup_batch(const float *I1, const float *I2, int B, int H, int W, const Taps &dx,
                       const Taps &dy, const Taps &dt, float *Ix, float *Iy, float *It,
                       hipStream_t s) {
    const int64_t n = (int64_t)H * W;
    for (int b = 0; b < B; ++b) {
        int rc = launch_hs_setup(I1 + b * n, I2 + b * n, H, W, dx, dy, dt, Ix + b * n,
                                 Iy + b * n, It + b * n, s);
        if (rc) return rc;
    }
    return HMX_OK;
}

struct AdjTaps {
    Taps dx, dy, dt, av;
};

static int make_all_taps(AdjTaps &T, int ndx, const float *wx, const int *oxx, const int *oyx,
                         int ndy, const float *wy, const int *oxy, const int *oyy, int ndt,
                         const float *wt, const int *oxt, const int *oyt, int nda,
                         const float *wa, const int *oxa, const int *oya) {
    int rc = make_taps(T.dx, ndx, wx, oxx, oyx);
    if (!rc) rc = make_taps(T.dy, ndy, wy, oxy, oyy);
    if (!rc) rc = make_taps(T.dt, ndt, wt, oxt, oyt);
    if (!rc) rc = make_taps(T.av, nda, wa, oxa, oya);
    if (rc) return rc;
    if (T.av.rx > 1 || T.av.ry > 1 || T.av.nd > 9)
        return fail(HMX_ERR_UNSUPPORTED,
                    "adjoint: averaging stencil must lie within radius 1 with at most 9 taps");
    return HMX_OK;
}

static int64_t adj_tiles(int H, int W) {
    return (int64_t)((W + ADJ_TW - 1) / ADJ_TW) * ((H + ADJ_TH - 1) / ADJ_TH);
}

}  // namespace hmx

using namespace hmx;

extern "C" int hmx_hs_ckpt_count(int iters, int ckpt_every) {
    if (iters <= 0 || ckpt_every <= 0) return 0;
    return (iters + ckpt_every - 1) / ckpt_every;
}

extern "C" int64_t hmx_hs_adjoint_workspace(int B, int H, int W, int ckpt_every) {
    const int64_t n = (int64_t)B * H * W;
    // Ix Iy It | gIx gIy gIt | 2 gradient ping-pong pairs | segment states
    const int64_t floats = 3 * n + 3 * n + 4 * n + 2 * n * (ckpt_every > 1 ? ckpt_every - 1 : 0);
    const int64_t doubles = (int64_t)B * adj_tiles(H, W);
    return floats * 4 + doubles * 8 + (int64_t)B * 4 + 256;
}

extern "C" int64_t hmx_hs_forward_workspace(int B, int H, int W) {
    return (int64_t)B * H * W * 5 * 4 + (int64_t)B * 4 + 256;
}

extern "C" int hmx_hs_forward_ckpt(const float *I1, const float *I2, int B, int H, int W,
                                   const float *alpha, int ndx, const float *wx,
                                   const int *oxx, const int *oyx, int ndy, const float *wy,
                                   const int *oxy, const int *oyy, int ndt, const float *wt,
                                   const int *oxt, const int *oyt, int nda, const float *wa,
                                   const int *oxa, const int *oya, int iters, int ckpt_every,
                                   float *u, float *v, float *ckpt, void *workspace,
                                   void *stream) {
    HMX_REQUIRE(I1 && I2 && alpha && u && v && workspace, "hmx_hs_forward_ckpt: null argument");
    HMX_REQUIRE(B >= 1 && H > 0 && W > 0 && iters >= 0 && ckpt_every >= 1,
                "hmx_hs_forward_ckpt: bad sizes");
    HMX_REQUIRE(ckpt || iters == 0, "hmx_hs_forward_ckpt: null checkpoint buffer");
    HMX_REQUIRE(B <= 65535, "hmx_hs_forward_ckpt: at most 65535 pairs");
    AdjTaps T;
    int rc = make_all_taps(T, ndx, wx, oxx, oyx, ndy, wy, oxy, oyy, ndt, wt, oxt, oyt, nda, wa,
                           oxa, oya);
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    const int64_t n = (int64_t)B * H * W;
    float *ws = (float *)workspace;
    float *Ix = ws, *Iy = ws + n, *It = ws + 2 * n, *wu = ws + 3 * n, *wv = ws + 4 * n;
    float *alpha2 = ws + 5 * n;
    hipLaunchKernelGGL(alpha_sq_kernel, dim3((B + 63) / 64), dim3(64), 0, s, alpha, alpha2, B);
    HMX_LAUNCH_CHECK("alpha_sq");
    rc = setup_batch(I1, I2, B, H, W, T.dx, T.dy, T.dt, Ix, Iy, It, s);
    if (rc) return rc;
    int variant, k;
    hs_choice(H, W, variant, k);
    const int nck = hmx_hs_ckpt_count(iters, ckpt_every);
    for (int c = 0; c < nck; ++c) {
        float *cu = ckpt + (int64_t)(2 * c) * n, *cv = ckpt + (int64_t)(2 * c + 1) * n;
        HMX_HIP_TRY(hipMemcpyAsync(cu, u, n * 4, hipMemcpyDeviceToDevice, s));
        HMX_HIP_TRY(hipMemcpyAsync(cv, v, n * 4, hipMemcpyDeviceToDevice, s));
        const int len = (iters - c * ckpt_every) < ckpt_every ? iters - c * ckpt_every : ckpt_every;
        rc = hs_jacobi_run(Ix, Iy, It, nullptr, 0.0f, H, W, T.av, len, u, v, wu, wv, variant, k, s,
                           B, alpha2);
        if (rc) return rc;
    }
    return HMX_OK;
}

extern "C" int hmx_hs_adjoint(const float *I1, const float *I2, int B, int H, int W,
                              const float *alpha, int ndx, const float *wx, const int *oxx,
                              const int *oyx, int ndy, const float *wy, const int *oxy,
                              const int *oyy, int ndt, const float *wt, const int *oxt,
                              const int *oyt, int nda, const float *wa, const int *oxa,
                              const int *oya, int iters, int ckpt_every, const float *ckpt,
                              const float *gu_N, const float *gv_N, float *gI1, float *gI2,
                              float *galpha, float *gu0, float *gv0, void *workspace,
                              void *stream) {
    HMX_REQUIRE(I1 && I2 && alpha && gu_N && gv_N && gI1 && gI2 && galpha && workspace,
                "hmx_hs_adjoint: null argument");
    HMX_REQUIRE(B >= 1 && B <= 65535 && H > 0 && W > 0 && iters >= 0 && ckpt_every >= 1,
                "hmx_hs_adjoint: bad sizes");
    HMX_REQUIRE(ckpt || iters == 0, "hmx_hs_adjoint: null checkpoints");
    AdjTaps T;
    int rc = make_all_taps(T, ndx, wx, oxx, oyx, ndy, wy, oxy, oyy, ndt, wt, oxt, oyt, nda, wa,
                           oxa, oya);
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    const int64_t n = (int64_t)B * H * W;
    const int64_t ntiles = adj_tiles(H, W);
    HMX_REQUIRE(ntiles < (1 << 30), "hmx_hs_adjoint: image too large");
    float *ws = (float *)workspace;
    float *Ix = ws, *Iy = ws + n, *It = ws + 2 * n;
    float *gIx = ws + 3 * n, *gIy = ws + 4 * n, *gIt = ws + 5 * n;
    float *g[2][2] = {{ws + 6 * n, ws + 7 * n}, {ws + 8 * n, ws + 9 * n}};
    float *seg = ws + 10 * n;
    const int nseg = ckpt_every > 1 ? ckpt_every - 1 : 0;
    double *ga2 = (double *)(((uintptr_t)(seg + 2 * n * nseg) + 255) & ~(uintptr_t)255);
    float *alpha2 = (float *)(ga2 + (int64_t)B * ntiles);

    hipLaunchKernelGGL(alpha_sq_kernel, dim3((B + 63) / 64), dim3(64), 0, s, alpha, alpha2, B);
    HMX_LAUNCH_CHECK("alpha_sq");
    rc = setup_batch(I1, I2, B, H, W, T.dx, T.dy, T.dt, Ix, Iy, It, s);
    if (!rc) rc = launch_fill_f32(gIx, 0.0f, 3 * n, s);
    if (!rc) rc = launch_fill_u32((unsigned *)ga2, 0u, 2 * (int64_t)B * ntiles, s);
    if (rc) return rc;
    HMX_HIP_TRY(hipMemcpyAsync(g[0][0], gu_N, n * 4, hipMemcpyDeviceToDevice, s));
    HMX_HIP_TRY(hipMemcpyAsync(g[0][1], gv_N, n * 4, hipMemcpyDeviceToDevice, s));
    int cur = 0;

    AdjArgs a;
    a.Ix = Ix;
    a.Iy = Iy;
    a.It = It;
    a.gIx = gIx;
    a.gIy = gIy;
    a.gIt = gIt;
    a.ga2 = ga2;
    a.alpha2_b = alpha2;
    a.H = H;
    a.W = W;
    a.tiles_x = (W + ADJ_TW - 1) / ADJ_TW;
    a.ntiles = (int)ntiles;
    a.bstride = (int64_t)H * W;

    const int nck = hmx_hs_ckpt_count(iters, ckpt_every);
    for (int c = nck - 1; c >= 0; --c) {
        const int len = (iters - c * ckpt_every) < ckpt_every ? iters - c * ckpt_every : ckpt_every;
        auto state_u = [&](int j) -> float * {
            return j == 0 ? (float *)ckpt + (int64_t)(2 * c) * n : seg + (int64_t)(2 * (j - 1)) * n;
        };
        auto state_v = [&](int j) -> float * {
            return j == 0 ? (float *)ckpt + (int64_t)(2 * c + 1) * n
                          : seg + (int64_t)(2 * (j - 1) + 1) * n;
        };
        for (int j = 1; j < len; ++j) {  // recompute the segment (one step per launch)
            rc = launch_hs_tb(0, Ix, Iy, It, nullptr, 0.0f, H, W, T.av, 1, state_u(j - 1),
                              state_v(j - 1), state_u(j), state_v(j), s, B, alpha2);
            if (rc) return rc;
        }
        for (int j = len - 1; j >= 0; --j) {
            a.u = state_u(j);
            a.v = state_v(j);
            a.gu = g[cur][0];
            a.gv = g[cur][1];
            a.gu_out = g[cur ^ 1][0];
            a.gv_out = g[cur ^ 1][1];
            hipLaunchKernelGGL(hs_adj_kernel, dim3((unsigned)ntiles, B), dim3(ADJ_NT), 0, s, a,
                               T.av);
            HMX_LAUNCH_CHECK("hs_adj");
            cur ^= 1;
        }
    }
    if (gu0) HMX_HIP_TRY(hipMemcpyAsync(gu0, g[cur][0], n * 4, hipMemcpyDeviceToDevice, s));
    if (gv0) HMX_HIP_TRY(hipMemcpyAsync(gv0, g[cur][1], n * 4, hipMemcpyDeviceToDevice, s));
    dim3 grid((W + 255) / 256, H < 65535 ? H : 65535, B);
    hipLaunchKernelGGL(hs_setup_T_kernel, grid, dim3(256), 0, s, gIx, gIy, gIt, H, W,
                       (int64_t)H * W, T.dx, T.dy, T.dt, gI1, gI2);
    HMX_LAUNCH_CHECK("hs_setup_T");
    hipLaunchKernelGGL(galpha_kernel, dim3(B), dim3(64), 0, s, ga2, (int)ntiles, alpha, galpha, B);
    HMX_LAUNCH_CHECK("galpha");
    return HMX_OK;
}

// ---------------------------------------------------------------------------
// fp64 forward (plain per-iteration kernels), for finite-difference checks of
// the adjoint at full size.

namespace hmx {
__global__ __launch_bounds__(256) void hs_setup_f64_kernel(const double *__restrict__ I1,
                                                           const double *__restrict__ I2, int H,
                                                           int W, Taps dx, Taps dy, Taps dt,
                                                           double *Ix, double *Iy, double *It) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= W) return;
    const int64_t off = (int64_t)blockIdx.z * H * W;
    for (int r = blockIdx.y; r < H; r += gridDim.y) {
        const int64_t p = off + (int64_t)r * W + c;
        Ix[p] = tap_apply_t(I1 + off, H, W, r, c, dx) + tap_apply_t(I2 + off, H, W, r, c, dx);
        Iy[p] = tap_apply_t(I1 + off, H, W, r, c, dy) + tap_apply_t(I2 + off, H, W, r, c, dy);
        It[p] = tap_apply_t(I2 + off, H, W, r, c, dt) - tap_apply_t(I1 + off, H, W, r, c, dt);
    }
}

__global__ __launch_bounds__(256) void hs_step_f64_kernel(const double *__restrict__ Ix,
                                                          const double *__restrict__ Iy,
                                                          const double *__restrict__ It,
                                                          const double *__restrict__ alpha,
                                                          int H, int W, Taps av,
                                                          const double *__restrict__ u,
                                                          const double *__restrict__ v,
                                                          double *un, double *vn) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= W) return;
    const int64_t off = (int64_t)blockIdx.z * H * W;
    const double a2 = alpha[blockIdx.z] * alpha[blockIdx.z];
    for (int r = blockIdx.y; r < H; r += gridDim.y) {
        const int64_t p = off + (int64_t)r * W + c;
        const double ub = tap_apply_t(u + off, H, W, r, c, av);
        const double vb = tap_apply_t(v + off, H, W, r, c, av);
        const double den = a2 + Ix[p] * Ix[p] + Iy[p] * Iy[p];
        const double t = (Ix[p] * ub + Iy[p] * vb + It[p]) / den;
        un[p] = ub - Ix[p] * t;
        vn[p] = vb - Iy[p] * t;
    }
}
}  // namespace hmx

extern "C" int hmx_hs_solve_f64(const double *I1, const double *I2, int B, int H, int W,
                                const double *alpha, int ndx, const float *wx, const int *oxx,
                                const int *oyx, int ndy, const float *wy, const int *oxy,
                                const int *oyy, int ndt, const float *wt, const int *oxt,
                                const int *oyt, int nda, const float *wa, const int *oxa,
                                const int *oya, int iters, double *u, double *v,
                                double *workspace, void *stream) {
    HMX_REQUIRE(I1 && I2 && alpha && u && v && workspace, "hmx_hs_solve_f64: null argument");
    HMX_REQUIRE(B >= 1 && B <= 65535 && H > 0 && W > 0 && iters >= 0, "hmx_hs_solve_f64: bad sizes");
    AdjTaps T;
    int rc = make_all_taps(T, ndx, wx, oxx, oyx, ndy, wy, oxy, oyy, ndt, wt, oxt, oyt, nda, wa,
                           oxa, oya);
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    const int64_t n = (int64_t)B * H * W;
    double *Ix = workspace, *Iy = workspace + n, *It = workspace + 2 * n;
    double *wu = workspace + 3 * n, *wv = workspace + 4 * n;
    dim3 grid((W + 255) / 256, H < 65535 ? H : 65535, B);
    hipLaunchKernelGGL(hs_setup_f64_kernel, grid, dim3(256), 0, s, I1, I2, H, W, T.dx, T.dy, T.dt,
                       Ix, Iy, It);
    HMX_LAUNCH_CHECK("hs_setup_f64");
    double *cu = u, *cv = v, *nu = wu, *nv = wv;
    for (int i = 0; i < iters; ++i) {
        hipLaunchKernelGGL(hs_step_f64_kernel, grid, dim3(256), 0, s, Ix, Iy, It, alpha, H, W,
                           T.av, cu, cv, nu, nv);
        HMX_LAUNCH_CHECK("hs_step_f64");
        std::swap(cu, nu);
        std::swap(cv, nv);
    }
    if (cu != u) {
        HMX_HIP_TRY(hipMemcpyAsync(u, cu, n * 8, hipMemcpyDeviceToDevice, s));
        HMX_HIP_TRY(hipMemcpyAsync(v, cv, n * 8, hipMemcpyDeviceToDevice, s));
    }
    return HMX_OK;
}
