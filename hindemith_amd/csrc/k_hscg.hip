// Operations of the reference's preconditioned-CG Horn-Schunck solver:
//   a15  ApplyHSPreconditioner  (ApplyHSPreconditioner.cpp:46-97)
//   a16  HSMatrixPowers         (HSMatrixPowers.cpp:28-39, 72-126, 191-226)
//
// The HS normal-equation operator on x = (u, v), with I = (Ix, Iy) and L the
// 4-neighbour Laplacian with Neumann (no-flux) borders weighted by lambda:
//   A x = ( Ix (Ix u + Iy v) + lam L u ,  Iy (Ix u + Iy v) + lam L v ).
// Its per-pixel 2x2 diagonal block [[a, b], [b, d]] with a = diag0 + n_nb*w2,
// d = diag1 + n_nb*w2 (n_nb = in-image 4-neighbours, w2 = weight 2 of the
// stencil passed in) is what the preconditioner inverts.
//
// Arithmetic follows the generated OpenCL statement by statement: one
// rounding per operator, the "y += p * q" accumulations of HSMatrixPowers as
// fmas (as the stencil accumulation elsewhere), correctly rounded division.
#include "hmx_common.h"

#include <algorithm>

#pragma clang fp contract(off)

namespace hmx {

// ---------------------------------------------------------------------------
// a15: y = M^-1 x per pixel.  28 B/px (read diag0, diag1, offdiag, x0, x1;
// write y0, y1): HBM bound, one pixel per thread, row-major order.
__global__ __launch_bounds__(256) void hs_precond_kernel(
    const float *__restrict__ diag0, const float *__restrict__ diag1,
    const float *__restrict__ offdiag, const float *__restrict__ x0,
    const float *__restrict__ x1, float *__restrict__ y0, float *__restrict__ y1, int H, int W,
    const float *__restrict__ w_dev, float w_val) {
    const int64_t n = (int64_t)H * W;
    const float s2 = w_dev ? w_dev[2] : w_val;  // the stencil's weight 2 (getData("2"))
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(p / W), c = (int)(p - (int64_t)r * W);
        float a = diag0[p], d = diag1[p];
        // neighbour terms in the order of the generated code
        if (c < W - 1) { a = a + s2; d = d + s2; }
        if (r < H - 1) { a = a + s2; d = d + s2; }
        if (c > 0) { a = a + s2; d = d + s2; }
        if (r > 0) { a = a + s2; d = d + s2; }
        const float b = offdiag[p];
        const float e = x0[p], f = x1[p];
        const float ad = a * d, bc = b * b;
        const float det = ad - bc;
        // 1e-15 is a double literal in the generated code: compare in double
        if (fabs((double)det) > 1e-15) {
            const float t0 = e * d, t1 = b * f;
            const float n0 = t0 - t1;
            const float t2 = (-b) * e, t3 = a * f;
            const float n1 = t2 + t3;
            y0[p] = n0 / det;
            y1[p] = n1 / det;
        } else {
            st_out(y0 + p, e, true);
            st_out(y1 + p, f, true);
        }
    }
}

// ---------------------------------------------------------------------------
// a16: V[0] = r, V[s] = A V[s-1] for s = 1..steps-1, all powers stored.
//
// Temporally blocked like the Jacobi sweep: a work-group loads r on its tile
// grown by m = (powers this launch) pixels, Ix and Iy into registers, and
// computes the m powers on chip; power j is valid on the tile grown by m - j.
// Two LDS windows: power j reads one and writes the other, one barrier per
// power.  Each power's tile is stored as it is made, so HBM traffic per launch
// is 16 B/px read + 8 B/px per power written (the per-power launches of the
// reference move 24 B/px per power).
struct PowArgs {
    const float *Ix, *Iy;
    const float *u0, *v0;  // power s0 - 1 (V[0] = r on the first launch)
    float *V0, *V1;        // steps x H x W
    int s0, m;             // this launch makes powers s0 .. s0+m-1
    int H, W, tiles_x, ntiles;
    int64_t plane;         // H * W
    const void *lam_dev;   // device scalar of dtype lam_dtype, or null: lam_val
    int lam_dtype;
    float lam_val;
    int copy_r;            // also store V[0] = r (the Array2D -> Array3D copy)
};

template <int NT, int TW, int TH, int MMAX>
__global__ __launch_bounds__(NT) void hs_powers_kernel(PowArgs a) {
    constexpr int PITCH = TW + 2 * MMAX, RHM = TH + 2 * MMAX, NPX = PITCH * RHM;
    constexpr int P = (NPX + NT - 1) / NT;
    __shared__ float2 win[2][NPX];
    const int H = a.H, W = a.W, m = a.m;
    const int t = xcd_remap(blockIdx.x, a.ntiles);
    const int tyi = t / a.tiles_x, txi = t - tyi * a.tiles_x;
    const int x0 = txi * TW, y0 = tyi * TH;
    const int x1 = min(x0 + TW, W), y1 = min(y0 + TH, H);
    const int gx0 = max(x0 - m, 0), gy0 = max(y0 - m, 0);
    const int gx1 = min(x1 + m, W), gy1 = min(y1 + m, H);
    const int RW = gx1 - gx0, RH = gy1 - gy0;
    // a scalar of another dtype is rounded to float32 first
    float lam = a.lam_val;
    if (a.lam_dev) {
        switch (a.lam_dtype) {
            case HMX_F64: lam = (float)*(const double *)a.lam_dev; break;
            case HMX_I64: lam = (float)*(const long long *)a.lam_dev; break;
            case HMX_I32: lam = (float)*(const int *)a.lam_dev; break;
            default: lam = *(const float *)a.lam_dev; break;
        }
    }
    const float nlam = -lam;

    // Thread tid owns window cells c = tid + j*NT (the pitch-PITCH grid), so a
    // neighbour of cell c is c +- 1 or c +- PITCH: constant LDS offsets.  pk
    // packs L << 27 | ry << 16 | rx for a cell in the region, L = m - (its
    // distance outside the tile) = the last power that updates it (the
    // region of power j is the tile grown by m - j, clipped to the image);
    // -1 for cells outside the region.
    const bool clipped = (x0 - m < 0) || (y0 - m < 0) || (x1 + m > W) || (y1 + m > H);
    int pk[P];
    float cx[P], cy[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const int c = threadIdx.x + j * NT;
        const int ry = c / PITCH, rx = c - ry * PITCH;
        pk[j] = -1;
        cx[j] = cy[j] = 0.0f;
        if (c < NPX && rx < RW && ry < RH) {
            const int gx = gx0 + rx, gy = gy0 + ry;
            const int dx = max(max(x0 - gx, gx - x1 + 1), 0);
            const int dy = max(max(y0 - gy, gy - y1 + 1), 0);
            pk[j] = ((m - max(dx, dy)) << 27) | (ry << 16) | rx;
            const int64_t g = (int64_t)gy * W + gx;
            cx[j] = a.Ix[g];
            cy[j] = a.Iy[g];
            const float u = a.u0[g], v = a.v0[g];
            win[0][c] = make_float2(u, v);
            if (a.copy_r && gx >= x0 && gx < x1 && gy >= y0 && gy < y1) {
                st_out(a.V0 + g, u, rows_whole_lines(a.W));
                st_out(a.V1 + g, v, rows_whole_lines(a.W));
            }
        }
    }
    __syncthreads();

    for (int j = 1; j <= m; ++j) {
        const float2 *rd = win[(j - 1) & 1];
        float2 *wr = win[j & 1];
        const int act_min = j << 27, tile_min = m << 27;
        float *V0s = a.V0 + (int64_t)(a.s0 + j - 1) * a.plane;
        float *V1s = a.V1 + (int64_t)(a.s0 + j - 1) * a.plane;
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int pkq = pk[q];
            if (pkq < act_min) continue;
            const int c = threadIdx.x + q * NT;
            const float2 mid = rd[c];
            const float ix = cx[q], iy = cy[q];
            const float p1 = ix * mid.x, p2 = iy * mid.y;
            const float dot = p1 + p2;
            float ur = ix * dot, vr = iy * dot;
            if (!clipped) {
                // an updated cell is at least one cell inside the region on
                // every side and the region inside the image: all four
                // neighbours exist; reads issued together
                const float2 nl = rd[c - 1], nr = rd[c + 1], nu = rd[c - PITCH],
                             nd = rd[c + PITCH];
                ur = fmaf(nlam, nl.x, ur);
                vr = fmaf(nlam, nl.y, vr);
                ur = fmaf(lam, mid.x, ur);
                vr = fmaf(lam, mid.y, vr);
                ur = fmaf(nlam, nr.x, ur);
                vr = fmaf(nlam, nr.y, vr);
                ur = fmaf(lam, mid.x, ur);
                vr = fmaf(lam, mid.y, vr);
                ur = fmaf(nlam, nu.x, ur);
                vr = fmaf(nlam, nu.y, vr);
                ur = fmaf(lam, mid.x, ur);
                vr = fmaf(lam, mid.y, vr);
                ur = fmaf(nlam, nd.x, ur);
                vr = fmaf(nlam, nd.y, vr);
                ur = fmaf(lam, mid.x, ur);
                vr = fmaf(lam, mid.y, vr);
            } else {
                int pkc = pkq;
                asm volatile("" : "+v"(pkc));  // keep per-cell values out of registers
                const int gx = gx0 + (pkc & 0xffff), gy = gy0 + ((pkc >> 16) & 0x7ff);
                if (gx > 0) {
                    const float2 nb = rd[c - 1];
                    ur = fmaf(nlam, nb.x, ur);
                    vr = fmaf(nlam, nb.y, vr);
                    ur = fmaf(lam, mid.x, ur);
                    vr = fmaf(lam, mid.y, vr);
                }
                if (gx < W - 1) {
                    const float2 nb = rd[c + 1];
                    ur = fmaf(nlam, nb.x, ur);
                    vr = fmaf(nlam, nb.y, vr);
                    ur = fmaf(lam, mid.x, ur);
                    vr = fmaf(lam, mid.y, vr);
                }
                if (gy > 0) {
                    const float2 nb = rd[c - PITCH];
                    ur = fmaf(nlam, nb.x, ur);
                    vr = fmaf(nlam, nb.y, vr);
                    ur = fmaf(lam, mid.x, ur);
                    vr = fmaf(lam, mid.y, vr);
                }
                if (gy < H - 1) {
                    const float2 nb = rd[c + PITCH];
                    ur = fmaf(nlam, nb.x, ur);
                    vr = fmaf(nlam, nb.y, vr);
                    ur = fmaf(lam, mid.x, ur);
                    vr = fmaf(lam, mid.y, vr);
                }
            }
            wr[c] = make_float2(ur, vr);
            if (pkq >= tile_min) {  // a tile cell (L = m): this power's output
                const int64_t g =
                    (int64_t)(gy0 + ((pkq >> 16) & 0x7ff)) * W + gx0 + (pkq & 0xffff);
                V0s[g] = ur;
                V1s[g] = vr;
            }
        }
        // the window written here was last read by power j-1, before the
        // previous barrier
        __syncthreads();
    }
}

constexpr int POW_NT = 1024, POW_TW = 64, POW_TH = 32, POW_MMAX = 8;

int launch_hs_precond(const float *diag0, const float *diag1, const float *offdiag,
                      const float *x0, const float *x1, float *y0, float *y1, int H, int W,
                      const float *w_dev, float w_val, hipStream_t s) {
    const int64_t n = (int64_t)H * W;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(hs_precond_kernel, dim3((unsigned)blocks), dim3(256), 0, s, diag0, diag1,
                       offdiag, x0, x1, y0, y1, H, W, w_dev, w_val);
    HMX_LAUNCH_CHECK("hs_precond");
    return HMX_OK;
}

// V0, V1: steps planes of H x W.  Powers are made POW_MMAX per launch; each
// launch after the first starts from the last power the previous one stored.
int launch_hs_powers(const float *Ix, const float *Iy, const float *r0, const float *r1,
                     float *V0, float *V1, int H, int W, int steps, const void *lam_dev,
                     int lam_dtype, float lam_val, hipStream_t s) {
    PowArgs a;
    a.Ix = Ix;
    a.Iy = Iy;
    a.V0 = V0;
    a.V1 = V1;
    a.H = H;
    a.W = W;
    a.plane = (int64_t)H * W;
    a.tiles_x = (W + POW_TW - 1) / POW_TW;
    a.ntiles = a.tiles_x * ((H + POW_TH - 1) / POW_TH);
    a.lam_dev = lam_dev;
    a.lam_dtype = lam_dtype;
    a.lam_val = lam_val;
    if (steps == 1) {  // only V[0] = r
        HMX_HIP_TRY(hipMemcpyAsync(V0, r0, a.plane * 4, hipMemcpyDeviceToDevice, s));
        HMX_HIP_TRY(hipMemcpyAsync(V1, r1, a.plane * 4, hipMemcpyDeviceToDevice, s));
        return HMX_OK;
    }
    for (int s0 = 1; s0 < steps; s0 += POW_MMAX) {
        a.s0 = s0;
        a.m = std::min(POW_MMAX, steps - s0);
        a.copy_r = s0 == 1;
        a.u0 = s0 == 1 ? r0 : V0 + (int64_t)(s0 - 1) * a.plane;
        a.v0 = s0 == 1 ? r1 : V1 + (int64_t)(s0 - 1) * a.plane;
        hipLaunchKernelGGL((hs_powers_kernel<POW_NT, POW_TW, POW_TH, POW_MMAX>),
                           dim3(a.ntiles), dim3(POW_NT), 0, s, a);
        HMX_LAUNCH_CHECK("hs_powers");
    }
    return HMX_OK;
}

}  // namespace hmx

using namespace hmx;

extern "C" int hmx_hs_precondition(const float *diag0, const float *diag1, const float *offdiag,
                                   const float *x0, const float *x1, float *y0, float *y1,
                                   int H, int W, int nd, const float *w, void *stream) {
    HMX_REQUIRE(diag0 && diag1 && offdiag && x0 && x1 && y0 && y1 && w,
                "hmx_hs_precondition: null argument");
    HMX_REQUIRE(H > 0 && W > 0, "hmx_hs_precondition: H and W must be positive");
    HMX_REQUIRE(nd >= 3, "hmx_hs_precondition: the stencil needs a weight 2 (>= 3 taps)");
    HMX_REQUIRE(y0 != x0 && y0 != x1 && y1 != x0 && y1 != x1 && y0 != y1,
                "hmx_hs_precondition: outputs must be distinct from the inputs");
    return launch_hs_precond(diag0, diag1, offdiag, x0, x1, y0, y1, H, W, nullptr, w[2],
                             as_stream(stream));
}

extern "C" int hmx_hs_matrix_powers(float lam, const float *Ix, const float *Iy,
                                    const float *r0, const float *r1, float *V0, float *V1,
                                    int H, int W, int steps, void *stream) {
    HMX_REQUIRE(Ix && Iy && r0 && r1 && V0 && V1, "hmx_hs_matrix_powers: null argument");
    HMX_REQUIRE(H > 0 && W > 0, "hmx_hs_matrix_powers: H and W must be positive");
    HMX_REQUIRE(steps >= 1, "hmx_hs_matrix_powers: steps must be >= 1");
    HMX_REQUIRE(H < 32768 && W < 32768, "hmx_hs_matrix_powers: H and W must be < 32768");
    return launch_hs_powers(Ix, Iy, r0, r1, V0, V1, H, W, steps, nullptr, HMX_F32, lam,
                            as_stream(stream));
}
