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
// so the transpose is a gather over at most (|o|+1)^2 sources per tap.  For a
// radius-1 stencil the same sum is the unclamped gather at p plus the
// unclamped gathers at the ghosts of p (the cells just across each image border
// p lies on, whose forward value is p's), with zero outside the image: the
// tiled reverse sweep computes it that way.
//
// Memory: states are checkpointed every C iterations in the forward pass; the
// reverse pass recomputes one segment of C steps at a time with the same fused
// kernel (bit-identical states), keeping for each step only what its reverse
// step needs on the pixel itself -- (ubar, vbar), from which t is recomputed --
// then walks the segment backwards.
#include "hmx_common.h"
#include "hs_tile.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <utility>

#pragma clang fp contract(off)

namespace hmx {

int launch_hs_tb(int variant, const float *Ix, const float *Iy, const float *It,
                 const float *den, float alpha2, int H, int W, const Taps &taps, int k,
                 const float *u_in, const float *v_in, float *u_out, float *v_out,
                 hipStream_t s, int B, const float *alpha2_b, float *lt, float *lub,
                 float *lvb, int64_t lstride, bool rev);
int hs_jacobi_run(const float *Ix, const float *Iy, const float *It, const float *den,
                  float alpha2, int H, int W, const Taps &taps, int iters, float *u, float *v,
                  float *wu, float *wv, int variant, int k, hipStream_t s, int B,
                  const float *alpha2_b, bool start_in_work);
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
// Reverse sweep, temporally blocked: one launch walks m (<= MMAX) reverse
// steps for an output tile.  The gradient (gu, gv) lives on the tile grown by
// m (clipped to the image) and shrinks by one pixel per step, exactly like the
// forward kernel's region, inside an LDS window at a fixed offset from the
// tile whose cells outside the image hold zero; the accumulators gIx, gIy, gIt of the
// tile's own pixels stay in registers and are stored once per launch; each
// step stages only the recomputed (ubar, vbar) of its forward step on the
// tile itself.

struct AdjTbArgs {
    const float *Ix, *Iy, *It;
    const float *gu_in, *gv_in;  // dL/dx_hi
    float *gu_out, *gv_out;      // dL/dx_{hi-m}
    // step j (1..m) reverses forward step hi-j, linearised about (ubar, vbar)
    // at lin_hi - (j-1)*lin_step and + plane (t is recomputed from them)
    const float *lin_hi;
    int64_t lin_step, plane;
    float *gIx, *gIy, *gIt;
    float *ga2;                  // per pixel: running sum over steps of -s t
    const float *alpha2_b;
    int H, W, m, tiles_x, ntiles;
    int64_t bstride;
};

constexpr int ADJ_TW = 64, ADJ_TH = 32, ADJ_NT = 1024, ADJ_MMAX = 8;

// 1/d rounded to nearest: the hardware reciprocal refined by one Newton step
// (hs_recip, the IEEE value wherever it accepts d), the IEEE division for the
// rest (a branch few lanes take)
__device__ __forceinline__ float adj_recip(float d) {
    float r = hs_recip(d);
    if (r != r) {
        float dd = d;
        asm volatile("" : "+v"(dd));
        r = 1.0f / dd;
    }
    return r;
}

// Compiled for 8 waves per SIMD: two 1024-thread work-groups per CU.
// Thread layout: slots 0..PT-1 of a thread are tile pixels in tile order
// (i = tid + j*NT), so the per-pixel accumulators and the step's (t, ubar,
// vbar) exist only there and load coalesced; the remaining PH slots cover the
// ring around the tile of the largest region (the tile grown by MMAX), laid
// out statically so that a slot's cell follows from its index by constant
// divisions.  The gradient g of every slot stays in registers (only its own
// thread reads or writes it); only (g_ubar, g_vbar), which pass B gathers
// from neighbours, lives in LDS, double-buffered, so pass B of step s and
// pass A of step s+1 (same region) run back to back with one barrier per
// step.  The step loop is unrolled over the two buffers, so each buffer's
// offset is a constant.
// NTAP > 0: the averaging stencil has exactly NTAP taps (straight-line
// gather, its reads issued in groups of four so the latency of a group is
// paid once); 0: any count up to HS_MAX_TAPS.  RING: the 8 taps sit at the
// canonical offsets HS_RING_OX/OY in that order (weights at run time), so
// every tap of a gather is a constant LDS offset from the cell's own address:
// no address arithmetic per tap.  Same arithmetic in the same order either
// way.
template <int NT, int TW, int TH, int MMAX, int NTAP, bool RING>
__global__ __launch_bounds__(NT, 8) void hs_adj_tb_kernel(AdjTbArgs a, Taps taps) {
    static_assert(!RING || NTAP == 8, "the ring build has the 8 canonical taps");
    // LDS window: the tile grown by MMAX, plus one column and one row, so the
    // ghost gathers of pass B (up to two pixels past the image) stay inside
    // it on every side
    constexpr int PITCH = TW + 2 * MMAX + 1, RHM = TH + 2 * MMAX + 1, NPX = PITCH * RHM;
    constexpr int BW = TW + 2 * MMAX, BH = TH + 2 * MMAX;  // the largest region
    constexpr int TP = TW * TH;
    static_assert(TP % NT == 0, "tile pixels must split evenly over the threads");
    static_assert((TW & (TW - 1)) == 0, "tile width must be a power of two");
    constexpr int PT = TP / NT;
    constexpr int NTOP = MMAX * BW, NSIDE = TH * 2 * MMAX;
    constexpr int PH = (BW * BH - TP + NT - 1) / NT;
    constexpr int P = PT + PH;
    // Window cell c of buffer b is GB[LEAD + b*NPX + c]: the LEAD cells in
    // front keep the lowest tap of any gathered cell at a non-negative
    // offset from the cell's own address
    constexpr int LEAD = PITCH + 1;
    static_assert(NPX * 8 < (1 << 15), "cell byte offsets must fit 15 bits");
    static_assert(MMAX <= 15, "distances must fit 4 bits");
    __shared__ float2 GB[LEAD + 2 * NPX];

    const int tid = threadIdx.x;
    const int b = blockIdx.y;
    const int64_t off = (int64_t)b * a.bstride;
    const float alpha2 = a.alpha2_b[b];
    const int H = a.H, W = a.W, m = a.m;
    const int t = xcd_remap(blockIdx.x, a.ntiles);
    const int tyi = t / a.tiles_x, txi = t - tyi * a.tiles_x;
    const int x0 = txi * TW, y0 = tyi * TH;
    const int x1 = min(x0 + TW, W), y1 = min(y0 + TH, H);
    const bool clipped = (x0 - m < 0) || (y0 - m < 0) || (x1 + m > W) || (y1 + m > H);
    const int tw = x1 - x0, th = y1 - y0;
    // the window's corner is pixel (x0 - MMAX, y0 - MMAX) for every tile
    const int wx0 = x0 - MMAX, wy0 = y0 - MMAX;
    // pixels are addressed from the tile's corner pixel by byte offsets that
    // fit 32 bits (H, W < 32768): unsigned for the tile's own pixels (the
    // load's uniform base plus one offset register), signed for the ring
    const int64_t corner = off + (int64_t)y0 * W + x0;
    const bool wt = rows_whole_lines(W);
    // tile slot j's pixel is NT / TW rows below slot 0's: one offset register
    // (slot 0's) serves every slot, from a per-slot uniform base
    auto at_c = [&](const float *base, int j = 0) {
        return reinterpret_cast<const char *>(base + corner + (int64_t)j * (NT / TW) * W);
    };
    auto ld = [](const char *p, unsigned o) { return *reinterpret_cast<const float *>(p + o); };

    // Every slot's cell: pk = D << 27 | E << 15 | A, where A is 8 x the window
    // cell index (the cell's byte offset in a buffer), E the image borders the
    // pixel lies on (bit 0 left, 1 right, 2 top, 3 bottom) and D its distance
    // outside the tile (-1: no pixel).  The region of reach e (the tile grown
    // by e, clipped to the image) holds the slot exactly when D <= e, i.e.
    // when (unsigned)pk < (e + 1) << 27: one compare.
    auto border = [&](int gx, int gy) {
        return (gx == 0) | (gx == W - 1) << 1 | (gy == 0) << 2 | (gy == H - 1) << 3;
    };
    int pk[P];
    unsigned ot = (unsigned)(tid / TW) * (unsigned)W * 4u + (unsigned)(tid % TW) * 4u;
    float cx[P], cy[P], cd[P];
    float2 g[P];
    float ax[PT], ay[PT], at[PT], aa[PT], ct[PT];
#pragma unroll
    for (int j = 0; j < PT; ++j) {
        const int i = tid + j * NT;
        const int ly = i / TW, lx = i % TW;
        const char *const pIx = at_c(a.Ix, j), *const pIy = at_c(a.Iy, j);
        pk[j] = -1;
        if (lx < tw && ly < th) {
            pk[j] = (border(x0 + lx, y0 + ly) << 15) |
                    (((ly + MMAX) * PITCH + lx + MMAX) * 8);  // D = 0
            const float ix = ld(pIx, ot), iy = ld(pIy, ot);
            cx[j] = ix;
            cy[j] = iy;
            cd[j] = adj_recip((alpha2 + ix * ix) + iy * iy);
            g[j] = make_float2(ld(at_c(a.gu_in, j), ot), ld(at_c(a.gv_in, j), ot));
            // the accumulators continue from their stored value and take one
            // step's term at a time, so the sums do not depend on how the
            // reverse sweep is cut into launches
            ax[j] = ld(at_c(a.gIx, j), ot);
            ay[j] = ld(at_c(a.gIy, j), ot);
            at[j] = ld(at_c(a.gIt, j), ot);
            aa[j] = ld(at_c(a.ga2, j), ot);
            ct[j] = ld(at_c(a.It, j), ot);
        }
    }
#pragma unroll
    for (int j = 0; j < PH; ++j) {
        // the ring: MMAX rows above the tile, MMAX below (BW wide each), then
        // MMAX columns on each side of the tile's TH rows
        const int h = tid + j * NT;
        int wx = -1, wy = 0;
        if (h < NTOP) {
            wy = h / BW;
            wx = h - wy * BW;
        } else if (h < 2 * NTOP) {
            const int q = h - NTOP, r = q / BW;
            wy = MMAX + TH + r;
            wx = q - r * BW;
        } else if (h < 2 * NTOP + NSIDE) {
            const int q = h - 2 * NTOP, r = q / (2 * MMAX), c = q - r * (2 * MMAX);
            wy = MMAX + r;
            wx = c < MMAX ? c : TW + c;
        }
        const int gx = wx0 + wx, gy = wy0 + wy;
        pk[PT + j] = -1;
        if (wx >= 0 && gx >= 0 && gx < W && gy >= 0 && gy < H) {
            const int dx = max(max(x0 - gx, gx - x1 + 1), 0);
            const int dy = max(max(y0 - gy, gy - y1 + 1), 0);
            const int d = max(dx, dy);
            // only cells within m of the tile are ever in reach
            if (d <= m) {
                pk[PT + j] = (d << 27) | (border(gx, gy) << 15) | ((wy * PITCH + wx) * 8);
                const int o = ((gy - y0) * W + (gx - x0)) * 4;
                const float ix = *reinterpret_cast<const float *>(at_c(a.Ix) + o);
                const float iy = *reinterpret_cast<const float *>(at_c(a.Iy) + o);
                cx[PT + j] = ix;
                cy[PT + j] = iy;
                cd[PT + j] = adj_recip((alpha2 + ix * ix) + iy * iy);
                g[PT + j] = make_float2(*reinterpret_cast<const float *>(at_c(a.gu_in) + o),
                                        *reinterpret_cast<const float *>(at_c(a.gv_in) + o));
            }
        }
    }
    const HsTapTable<PITCH> tp(taps, NTAP);

    // the cell at byte offset A (from buffer 0's first cell) displaced by c
    // cells
    auto cell = [&](int A, int c) -> float2 & {
        return *reinterpret_cast<float2 *>(reinterpret_cast<char *>(GB) + A + c * 8);
    };

    // (ubar, vbar) of forward step hi - step at the tile slots; its t is
    // recomputed as ((Ix ubar + Iy vbar) + It) / den (with the reciprocal:
    // within two ulps of the forward's quotient)
    float lu[PT], lv[PT];
    auto load_lin = [&](int step) {
        const float *U0 = a.lin_hi - (int64_t)(step - 1) * a.lin_step;
        // (opaque per step: otherwise each address becomes a 64-bit induction
        // variable held across the loop)
        asm volatile("" : "+v"(ot));
#pragma unroll
        for (int j = 0; j < PT; ++j) {
            if (pk[j] >= 0) {
                lu[j] = ld(at_c(U0, j), ot);
                lv[j] = ld(at_c(U0 + a.plane, j), ot);
            }
        }
    };

    // pass A of step `step` on R_{m-step+1}: s, (g_ubar, g_vbar) into buffer
    // BUF and the tile's accumulator terms (the alpha term -s t included)
    auto pass_a = [&](int step, auto buf_c) {
        constexpr int BUF = decltype(buf_c)::value;
        const unsigned lim = (unsigned)(m - step + 2) << 27;
#pragma unroll
        for (int j = 0; j < P; ++j) {
            asm volatile("" : "+v"(pk[j]));  // the address is recomputed per step
            const int pkj = pk[j];
            if ((unsigned)pkj >= lim) continue;
            const int A = pkj & 0x7fff;
            // (fused multiply-adds: the adjoint is checked against fp64 and
            // finite differences, not bit for bit)
            const float s = -fmaf(cx[j], g[j].x, cy[j] * g[j].y) * cd[j];
            cell(A, LEAD + BUF * NPX) =
                make_float2(fmaf(cx[j], s, g[j].x), fmaf(cy[j], s, g[j].y));
            if (j < PT) {  // a tile pixel (always inside the region)
                const float tt = (fmaf(cx[j], lu[j], cy[j] * lv[j]) + ct[j]) * cd[j];
                const float st2 = 2.0f * s * tt;
                ax[j] = fmaf(-st2, cx[j], fmaf(s, lu[j], fmaf(-g[j].x, tt, ax[j])));
                ay[j] = fmaf(-st2, cy[j], fmaf(s, lv[j], fmaf(-g[j].y, tt, ay[j])));
                at[j] += s;
                aa[j] = fmaf(-s, tt, aa[j]);
            }
        }
    };

    // A^T applied at the cell displaced by c0 - LEAD cells from A: the sum
    // over taps d of w_d times the cell at -(oy_d, ox_d) from it
    auto gather = [&](int A, int c0, float &au, float &av) {
        if (NTAP > 0) {
            constexpr int G = 4;
#pragma unroll
            for (int d0 = 0; d0 < NTAP; d0 += G) {
                float2 q[G];
#pragma unroll
                for (int i = 0; i < G; ++i)
                    if (d0 + i < NTAP) {
                        const int dl = RING ? HS_RING_OY[(d0 + i) & 7] * PITCH +
                                                  HS_RING_OX[(d0 + i) & 7]
                                            : tp.dl[d0 + i];
                        q[i] = cell(A, c0 - dl);
                    }
#pragma unroll
                for (int i = 0; i < G; ++i)
                    if (d0 + i < NTAP) {
                        au = fmaf(tp.w[d0 + i], q[i].x, au);
                        av = fmaf(tp.w[d0 + i], q[i].y, av);
                    }
                // the next group's reads stay behind this group's use
                asm volatile("" ::: "memory");
            }
        } else {
#pragma unroll
            for (int d = 0; d < HS_MAX_TAPS; ++d) {
                if (d < tp.nd) {
                    const float2 q = cell(A, c0 - tp.dl[d]);
                    au = fmaf(tp.w[d], q.x, au);
                    av = fmaf(tp.w[d], q.y, av);
                }
            }
        }
    };

    // pass B of step `step` on R_{m-step}: g <- A^T g_b, from buffer BUF.  A
    // source q whose tap lands outside the image clamps onto the border pixel
    // p: in the forward map that is a read of a ghost of p (the cell across
    // the border, which holds p's value), so p also collects the unclamped
    // gather at each of its ghosts (one per border it lies on, three at a
    // corner).  The zeros outside the image drop the sources that do not
    // exist.  CLIP (compile time): the tile's regions reach an image border.
    auto pass_b = [&](int step, auto buf_c, auto clip_c) {
        constexpr int BUF = decltype(buf_c)::value;
        constexpr bool CLIP = decltype(clip_c)::value;
        constexpr int C0 = LEAD + BUF * NPX;
        const unsigned lim = (unsigned)(m - step + 1) << 27;
#pragma unroll
        for (int j = 0; j < P; ++j) {
            asm volatile("" : "+v"(pk[j]));
            const int pkj = pk[j];
            if ((unsigned)pkj >= lim) continue;
            const int A = pkj & 0x7fff;
            float au = 0.0f, av = 0.0f;
            gather(A, C0, au, av);
            if (CLIP) {
                const int eb = (pkj >> 15) & 15;
                if (eb) {
                    const bool l = eb & 1, r = eb & 2;
                    if (l) gather(A, C0 - 1, au, av);
                    if (r) gather(A, C0 + 1, au, av);
                    if (eb & 4) {
                        gather(A, C0 - PITCH, au, av);
                        if (l) gather(A, C0 - PITCH - 1, au, av);
                        if (r) gather(A, C0 - PITCH + 1, au, av);
                    }
                    if (eb & 8) {
                        gather(A, C0 + PITCH, au, av);
                        if (l) gather(A, C0 + PITCH - 1, au, av);
                        if (r) gather(A, C0 + PITCH + 1, au, av);
                    }
                }
            }
            g[j] = make_float2(au, av);
        }
    };

    // Window cells outside the image read as zero in both buffers: pass B's
    // unclamped gathers then take no term from a source that does not exist.
    // (Pass A writes only image pixels, so the zeros stay.)
    if (clipped) {
        for (int c = tid; c < NPX; c += NT) {
            const int ry = c / PITCH, rx = c - ry * PITCH;
            const int gx = wx0 + rx, gy = wy0 + ry;
            if (gx < 0 || gx >= W || gy < 0 || gy >= H)
                GB[LEAD + c] = GB[LEAD + NPX + c] = make_float2(0.0f, 0.0f);
        }
    }

    using B0 = std::integral_constant<int, 0>;
    using B1 = std::integral_constant<int, 1>;
    load_lin(1);
    pass_a(1, B0());
    __syncthreads();
    auto run = [&](auto clip_c) {
        // step s reads the buffer pass A of step s wrote, (s - 1) & 1; the
        // other one was last read in pass B of step s - 1, before the
        // previous barrier
        // The next step's (ubar, vbar) are read before the gather, whose
        // work then covers their latency (read after it, every wave waited
        // for them right after the gather: each step paid a full memory
        // latency behind its barrier).
        for (int step = 1; step <= m; step += 2) {
            if (step < m) load_lin(step + 1);
            pass_b(step, B0(), clip_c);
            if (step == m) break;
            pass_a(step + 1, B1());
            __syncthreads();
            if (step + 1 < m) load_lin(step + 2);
            pass_b(step + 1, B1(), clip_c);
            if (step + 1 == m) break;
            pass_a(step + 2, B0());
            __syncthreads();
        }
    };
    if (clipped)
        run(std::true_type());
    else
        run(std::false_type());

    auto st = [&](float *base, int j, float x) {
        st_out(reinterpret_cast<float *>(const_cast<char *>(at_c(base, j)) + ot), x, wt);
    };
#pragma unroll
    for (int j = 0; j < PT; ++j) {
        if (pk[j] < 0) continue;
        st(a.gu_out, j, g[j].x);
        st(a.gv_out, j, g[j].y);
        st(a.gIx, j, ax[j]);
        st(a.gIy, j, ay[j]);
        st(a.gIt, j, at[j]);
        st(a.ga2, j, aa[j]);
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

// dL/dalpha = 2 alpha times the sum of the per-pixel alpha terms, summed in
// fp64 in a fixed order so the result does not depend on timing: block
// (c, b) reduces chunk c of pair b (thread i takes elements i, i + 256, ...
// of the chunk, then a fixed tree), and galpha_finish adds pair b's
// GALPHA_CHUNKS partials in chunk order.
constexpr int GALPHA_CHUNKS = 64;

__global__ __launch_bounds__(256) void galpha_partial_kernel(const float *__restrict__ ga2,
                                                             int64_t n, double *partial) {
    __shared__ double part[256];
    const int c = blockIdx.x, b = blockIdx.y;
    const int64_t per = (n + GALPHA_CHUNKS - 1) / GALPHA_CHUNKS;
    const int64_t lo = c * per, hi = lo + per < n ? lo + per : n;
    const float *g = ga2 + (int64_t)b * n;
    double s = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += 256) s += (double)g[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (unsigned w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(int64_t)b * GALPHA_CHUNKS + c] = part[0];
}

__global__ void galpha_finish_kernel(const double *__restrict__ partial, const float *alpha,
                                     float *galpha, int B) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    double s = 0.0;
    for (int c = 0; c < GALPHA_CHUNKS; ++c) s += partial[(int64_t)b * GALPHA_CHUNKS + c];
    galpha[b] = (float)(2.0 * (double)alpha[b] * s);
}

// Batched setup: one launch per pair of the single-pair kernel.
int launch_hs_setup(const float *I1, const float *I2, int H, int W, const Taps &dx,
                    const Taps &dy, const Taps &dt, float *Ix, float *Iy, float *It,
                    hipStream_t s, const float *u, const float *v, float *wu, float *wv);

static int setup_batch(const float *I1, const float *I2, int B, int H, int W, const Taps &dx,
                       const Taps &dy, const Taps &dt, float *Ix, float *Iy, float *It,
                       hipStream_t s) {
    const int64_t n = (int64_t)H * W;
    for (int b = 0; b < B; ++b) {
        int rc = launch_hs_setup(I1 + b * n, I2 + b * n, H, W, dx, dy, dt, Ix + b * n,
                                 Iy + b * n, It + b * n, s, nullptr, nullptr, nullptr, nullptr);
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

// Variant of the forward kernel used to recompute a checkpoint segment: it
// writes every intermediate state of each k-step launch.
constexpr int ADJ_RECOMPUTE_VARIANT = 3;  // 1024 threads, 64x32 tiles, k <= 8
constexpr int ADJ_RECOMPUTE_K = 8;

int launch_hs_copy2(const float *u, const float *v, float *wu, float *wv, int64_t n,
                    hipStream_t s);

// len iterations from the state in (u, v) with the linearisation-output
// launches (ADJ_RECOMPUTE_K per launch), storing iteration j's (ubar, vbar) at
// lin + 2*j*n and + n, the state ending in (u, v): the launches ping-pong
// through (wu, wv), the first reading (wu, wv) after a copy when their count
// is odd.  The same launches, in the same order, as the adjoint's segment
// recompute, so the states and the stored values are the same bits.
static int lin_segment(const float *Ix, const float *Iy, const float *It, int H, int W,
                       const Taps &av, int len, float *u, float *v, float *wu, float *wv,
                       float *lin, int64_t n, hipStream_t s, int B, const float *alpha2) {
    const int nl = (len + ADJ_RECOMPUTE_K - 1) / ADJ_RECOMPUTE_K;
    float *su = u, *sv = v, *du = wu, *dv = wv;
    if (nl & 1) {
        int rc = launch_hs_copy2(u, v, wu, wv, n, s);
        if (rc) return rc;
        std::swap(su, du);
        std::swap(sv, dv);
    }
    for (int j0 = 0; j0 < len;) {
        const int kk = std::min(ADJ_RECOMPUTE_K, len - j0);
        float *lub = lin + (int64_t)2 * j0 * n;
        int rc = launch_hs_tb(ADJ_RECOMPUTE_VARIANT, Ix, Iy, It, nullptr, 0.0f, H, W, av, kk, su,
                              sv, du, dv, s, B, alpha2, nullptr, lub, lub + n, 2 * n,
                              (j0 / ADJ_RECOMPUTE_K) & 1);
        if (rc) return rc;
        std::swap(su, du);
        std::swap(sv, dv);
        j0 += kk;
    }
    return HMX_OK;
}

}  // namespace hmx

using namespace hmx;

extern "C" int hmx_hs_ckpt_count(int iters, int ckpt_every) {
    if (iters <= 0 || ckpt_every <= 0) return 0;
    return (iters + ckpt_every - 1) / ckpt_every;
}

extern "C" int64_t hmx_hs_adjoint_workspace(int B, int H, int W, int ckpt_every) {
    const int64_t n = (int64_t)B * H * W;
    // Ix Iy It | gIx gIy gIt | 2 gradient ping-pong pairs | 2 recompute state
    // ping-pong pairs | per-pixel alpha terms | (ubar, vbar) of every step of
    // one segment | alpha^2 per pair | fp64 partial sums of the alpha terms
    const int64_t floats = 3 * n + 3 * n + 4 * n + 4 * n + n + 2 * n * std::max(ckpt_every, 1);
    return floats * 4 + (int64_t)B * 4 + 256 + (int64_t)B * GALPHA_CHUNKS * 8;
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
                                   float *u, float *v, float *ckpt, float *lin_last,
                                   void *workspace, void *stream) {
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
        if (lin_last && c == nck - 1)
            // the last segment also keeps each iteration's (ubar, vbar): the
            // reverse pass then needs no recompute of it
            rc = lin_segment(Ix, Iy, It, H, W, T.av, len, u, v, wu, wv, lin_last, n, s, B, alpha2);
        else
            rc = hs_jacobi_run(Ix, Iy, It, nullptr, 0.0f, H, W, T.av, len, u, v, wu, wv, variant,
                               k, s, B, alpha2, false);
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
                              const float *lin_last, const void *fwd_workspace,
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
    if (fwd_workspace) {  // the forward's derivative planes (its workspace's first three)
        float *f = (float *)const_cast<void *>(fwd_workspace);
        Ix = f;
        Iy = f + n;
        It = f + 2 * n;
    }
    float *gIx = ws + 3 * n, *gIy = ws + 4 * n, *gIt = ws + 5 * n;
    float *g[2][2] = {{ws + 6 * n, ws + 7 * n}, {ws + 8 * n, ws + 9 * n}};
    float *rs[2][2] = {{ws + 10 * n, ws + 11 * n}, {ws + 12 * n, ws + 13 * n}};
    float *ga2 = ws + 14 * n;  // per-pixel running sums of the alpha term
    float *lin = ws + 15 * n;  // slot j: ubar, vbar of the segment's step j
    float *alpha2 = lin + 2 * n * ckpt_every;
    double *partial = (double *)(((uintptr_t)(alpha2 + B) + 255) & ~(uintptr_t)255);

    hipLaunchKernelGGL(alpha_sq_kernel, dim3((B + 63) / 64), dim3(64), 0, s, alpha, alpha2, B);
    HMX_LAUNCH_CHECK("alpha_sq");
    rc = fwd_workspace ? HMX_OK : setup_batch(I1, I2, B, H, W, T.dx, T.dy, T.dt, Ix, Iy, It, s);
    if (!rc) rc = launch_fill_f32(gIx, 0.0f, 3 * n, s);
    if (!rc) rc = launch_fill_f32(ga2, 0.0f, n, s);
    if (rc) return rc;
    // the gradient in: dL/dx_N itself for the first reverse launch, then the
    // ping-pong pair g[0], g[1] (the launch writing g[cur ^ 1] reads the other)
    const float *gin_u = gu_N, *gin_v = gv_N;
    int cur = 1;

    AdjTbArgs a;
    std::memset(&a, 0, sizeof(a));
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

    // the reverse sweep's ring build for the canonical 8-neighbour stencil
    // (HMX_ADJ_GENERIC=1 takes the generic build, same results bit for bit)
    const char *gen = std::getenv("HMX_ADJ_GENERIC");
    const bool ring = hs_ring_taps(T.av) && !(gen && gen[0] == '1');
    const int nck = hmx_hs_ckpt_count(iters, ckpt_every);
    for (int c = nck - 1; c >= 0; --c) {
        const int len = (iters - c * ckpt_every) < ckpt_every ? iters - c * ckpt_every : ckpt_every;
        // recompute the segment's len steps from its checkpoint, k steps per
        // launch, keeping each step's (ubar, vbar); the states themselves
        // are bit-identical to the forward pass and only ping-pong here
        const float *cu = ckpt + (int64_t)(2 * c) * n, *cv = ckpt + (int64_t)(2 * c + 1) * n;
        // the last segment's (ubar, vbar) may come from the forward pass
        const bool kept = lin_last && c == nck - 1;
        float *seg_lin = kept ? const_cast<float *>(lin_last) : lin;
        int rcur = 0;
        for (int j0 = 0; j0 < len && !kept;) {
            const int kk = std::min(ADJ_RECOMPUTE_K, len - j0);
            float *lub = lin + (int64_t)2 * j0 * n;
            rc = launch_hs_tb(ADJ_RECOMPUTE_VARIANT, Ix, Iy, It, nullptr, 0.0f, H, W, T.av, kk,
                              j0 == 0 ? cu : rs[rcur][0], j0 == 0 ? cv : rs[rcur][1],
                              rs[rcur ^ 1][0], rs[rcur ^ 1][1], s, B, alpha2, nullptr, lub,
                              lub + n, 2 * n, (j0 / ADJ_RECOMPUTE_K) & 1);
            if (rc) return rc;
            rcur ^= 1;
            j0 += kk;
        }
        // reverse steps len-1 .. 0, up to ADJ_MMAX per launch
        for (int hi = len; hi > 0;) {
            const int mm = std::min(ADJ_MMAX, hi);
            a.m = mm;
            a.lin_hi = seg_lin + (int64_t)2 * (hi - 1) * n;
            a.lin_step = (int64_t)2 * n;
            a.plane = n;
            a.gu_in = gin_u;
            a.gv_in = gin_v;
            a.gu_out = g[cur][0];
            a.gv_out = g[cur][1];
            const dim3 grid((unsigned)ntiles, B);
            if (ring)
                hipLaunchKernelGGL((hs_adj_tb_kernel<ADJ_NT, ADJ_TW, ADJ_TH, ADJ_MMAX, 8, true>),
                                   grid, dim3(ADJ_NT), 0, s, a, T.av);
            else if (T.av.nd == 8)
                hipLaunchKernelGGL((hs_adj_tb_kernel<ADJ_NT, ADJ_TW, ADJ_TH, ADJ_MMAX, 8, false>),
                                   grid, dim3(ADJ_NT), 0, s, a, T.av);
            else
                hipLaunchKernelGGL((hs_adj_tb_kernel<ADJ_NT, ADJ_TW, ADJ_TH, ADJ_MMAX, 0, false>),
                                   grid, dim3(ADJ_NT), 0, s, a, T.av);
            HMX_LAUNCH_CHECK("hs_adj_tb");
            gin_u = g[cur][0];
            gin_v = g[cur][1];
            cur ^= 1;
            hi -= mm;
        }
    }
    // dL/dx_0: the last launch's output (dL/dx_N itself when iters == 0)
    if (gu0) HMX_HIP_TRY(hipMemcpyAsync(gu0, gin_u, n * 4, hipMemcpyDeviceToDevice, s));
    if (gv0) HMX_HIP_TRY(hipMemcpyAsync(gv0, gin_v, n * 4, hipMemcpyDeviceToDevice, s));
    dim3 grid((W + 255) / 256, H < 65535 ? H : 65535, B);
    hipLaunchKernelGGL(hs_setup_T_kernel, grid, dim3(256), 0, s, gIx, gIy, gIt, H, W,
                       (int64_t)H * W, T.dx, T.dy, T.dt, gI1, gI2);
    HMX_LAUNCH_CHECK("hs_setup_T");
    hipLaunchKernelGGL(galpha_partial_kernel, dim3(GALPHA_CHUNKS, B), dim3(256), 0, s, ga2,
                       (int64_t)H * W, partial);
    HMX_LAUNCH_CHECK("galpha_partial");
    hipLaunchKernelGGL(galpha_finish_kernel, dim3((B + 63) / 64), dim3(64), 0, s, partial, alpha,
                       galpha, B);
    HMX_LAUNCH_CHECK("galpha_finish");
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
