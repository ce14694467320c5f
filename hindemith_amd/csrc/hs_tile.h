// Tile body of the fused, temporally blocked Horn-Schunck Jacobi sweep,
// shared by the single-image launch (k_hs.hip) and the persistent batched
// launch (k_batched.hip) so both run the same per-pixel code.
#pragma once

#include <type_traits>

#include "hmx_common.h"
#include "hs_div.h"

#pragma clang fp contract(off)

namespace hmx {

// The canonical 8-tap averaging stencil's offsets, in its tap order (edges,
// then corners).  Builds for it (ring-ordered tiles, RING tile bodies, the
// adjoint's reverse sweep) take each tap at a compile-time LDS offset; the
// weights stay run-time values.
constexpr int HS_RING_OX[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
constexpr int HS_RING_OY[8] = {0, 0, -1, 1, -1, -1, 1, 1};

__host__ __device__ inline bool hs_ring_taps(const Taps &t) {
    if (t.nd != 8) return false;
    for (int d = 0; d < 8; ++d)
        if (t.ox[d] != HS_RING_OX[d] || t.oy[d] != HS_RING_OY[d]) return false;
    return true;
}

// One image of the sweep: read (u_in, v_in), write (u_out, v_out).
struct HsImage {
    const float *Ix, *Iy, *It, *den;  // den null: recompute from alpha2
    const float *u_in, *v_in;
    float *u_out, *v_out;
    float alpha2;
    int H, W, tiles_x;
    // Optional (kernels built with LOUT): for each step j (1..k), the values
    // the step linearises about on the output tile -- t and (ubar, vbar) of
    // the state before the step -- are written to (lt, lub, lvb) +
    // (j-1)*lstride (the adjoint's recompute of a checkpoint segment); lt may
    // be null (t not stored).
    float *lt = nullptr, *lub = nullptr, *lvb = nullptr;
    int64_t lstride = 0;
};

// Largest averaging stencil the fused kernel takes: any radius-1 stencil with
// distinct offsets has at most 9 taps.
constexpr int HS_MAX_TAPS = 9;

// Tap table of a launch, hoisted into registers (uniform -> SGPRs).
template <int PITCH>
struct HsTapTable {
    float w[HS_MAX_TAPS];
    int dl[HS_MAX_TAPS];  // LDS delta oy*PITCH + ox
    int ox[HS_MAX_TAPS], oy[HS_MAX_TAPS];
    int pos[HS_MAX_TAPS];  // 3x3 window position (oy+1)*3 + (ox+1)
    int nd;
    // nfix > 0: the launch's tap count is known to be nfix (a kernel built
    // for that count), so the table needs no per-tap test and its weights
    // stay uniform scalars
    __device__ __forceinline__ explicit HsTapTable(const Taps &t, int nfix = 0) {
        nd = nfix > 0 ? nfix : t.nd;
#pragma unroll
        for (int d = 0; d < HS_MAX_TAPS; ++d) {
            const bool on = d < nd;
            w[d] = on ? t.w[d] : 0.0f;
            ox[d] = on ? t.ox[d] : 0;
            oy[d] = on ? t.oy[d] : 0;
            dl[d] = oy[d] * PITCH + ox[d];
            pos[d] = (oy[d] + 1) * 3 + (ox[d] + 1);
        }
    }
};

// The per-pixel update of the canonical program, one rounding per op:
//   t = ((Ix*ubar + Iy*vbar) + It) / den;  u' = ubar - Ix*t;  v' = vbar - Iy*t
// FDIV: the quotient comes from rcp = hs_recip(den) by hs_div_fast (the same
// correctly rounded value), with the IEEE division for the lanes it declines.
template <bool FDIV = false>
__device__ __forceinline__ float hs_update(float ix, float iy, float it, float den, float rcp,
                                           float ub, float vb, float &un, float &vn) {
    const float m1 = ix * ub;
    const float m2 = iy * vb;
    const float s1 = m1 + m2;
    const float s2 = s1 + it;
    float tt;
    if (!FDIV) {
        tt = s2 / den;
    } else if (!hs_div_fast(s2, den, rcp, tt)) {
        // opaque operand: keeps the division inside the (rarely taken)
        // branch instead of computed for every lane and selected
        float s2o = s2;
        asm volatile("" : "+v"(s2o));
        tt = s2o / den;
    }
    const float p1 = ix * tt;
    const float p2 = iy * tt;
    un = ub - p1;
    vn = vb - p2;
    return tt;
}

// Copies the new value q of window cell c into the ghost cells that clamp to
// it: the neighbours across each image border the pixel lies on (pk's border
// flags, hs_tile), corners included.  Each ghost clamps to exactly one image
// pixel, so no two cells store to the same ghost.
template <int PITCH>
__device__ __forceinline__ void hs_ghost_store(float2 *w, int c, int pk, float2 q) {
    const int e = (pk >> 12) & 15;
    if (e == 0) return;
    const bool l = e & 1, r = e & 2;
    if (l) w[c - 1] = q;
    if (r) w[c + 1] = q;
    if (e & 4) {
        w[c - PITCH] = q;
        if (l) w[c - PITCH - 1] = q;
        if (r) w[c - PITCH + 1] = q;
    }
    if (e & 8) {
        w[c + PITCH] = q;
        if (l) w[c + PITCH - 1] = q;
        if (r) w[c + PITCH + 1] = q;
    }
}

// |a - b| + c on unsigned values, the form the compiler emits as v_sad_u32
__device__ __forceinline__ unsigned hs_usad(unsigned a, unsigned b, unsigned c) {
    return (max(a, b) - min(a, b)) + c;
}

// LDS of one tile window (two with DB): PAD cells before and after it, so a
// thread may read the taps of any cell it owns, updated or not, without
// leaving the allocation.  hs_tile takes uv = the allocation + PAD.
template <int NT, int TW, int TH, int KMAX, bool DB>
struct HsWindow {
    static constexpr int PITCH = TW + 2 * KMAX;
    static constexpr int NPX = PITCH * (TH + 2 * KMAX);
    static constexpr int OWNED = (NPX + NT - 1) / NT * NT;  // cells the threads own
    static constexpr int PAD = PITCH + 1;
    static constexpr int SIZE = PAD + (DB ? NPX : 0) + OWNED + PAD;
};

// Advance output tile t of image a by k (<= KMAX) Jacobi steps.  uv is the
// work-group's LDS window of (TW + 2*KMAX) x (TH + 2*KMAX) float2 (padded as
// HsWindow), stored with the fixed row pitch TW + 2*KMAX so every tap is a
// constant LDS delta.
// NTAP > 0: the stencil has exactly NTAP taps (straight-line code, batched LDS
// reads); NTAP == 0: any count up to HS_MAX_TAPS, tested per tap.
// LOUT: also write each step's (t, ubar, vbar) of the output tile to a.lt,
// a.lub, a.lvb (a compile-time switch: the extra work stays out of the
// kernels that do not need it).  Two-barrier builds stage them through lin
// (3 x TW x TH floats of LDS) and store them in tile order between the step's
// barriers; double-buffered (RING) builds store them from registers as each
// tile cell is computed, at a per-cell offset found once.
// DB: uv holds two windows; step j reads window (j-1)&1 and writes its result
// straight into the other, so a step needs one barrier and no new values are
// held in registers across it.
// SHARED: the tap addresses are per-thread values shared by all owned cells
// (one register per tap, the cell a constant offset); otherwise each cell
// computes its own, for builds with no registers to spare (LOUT, and more
// than 8 cells per thread).
// FDIV: divide by a per-cell reciprocal (hs_update), one more register per
// cell.
// PF: read the next cell's taps ahead (see the step loop); needs SHARED, a
// fixed tap count and no LOUT.  On by default where it fits the registers of
// two 1024-thread work-groups per CU: the double-buffered builds without
// FDIV, whose second barrier's held results free the registers it needs.
// COH: u, v are read with system-scope loads (sc0 sc1), which the hardware
// keeps coherent across the XCDs' L2 caches; for a launch whose tiles read
// other tiles' output of the same launch (hs_persist_kernel).
// RING: the stencil is the canonical 8 taps (hs_ring_taps), so tap d's LDS
// delta is a compile-time constant: a cell's reads are one address register
// plus immediate offsets, and the shared form needs no register per tap.
template <bool COH>
__device__ __forceinline__ float hs_ld_uv(const char *p) {
    if (COH)
        return __hip_atomic_load((const gfloat *)p, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_SYSTEM);
    return ld_g(reinterpret_cast<const float *>(p));
}
// The output tile is stored through the L2 (st_out) when the image's rows
// fill whole cache lines, as every kernel's output here (1080p launches 31.9
// -> 31.3 us on one box; non-temporal stores were slower, 32.4 us); COH
// stores always use system scope.
template <bool COH>
__device__ __forceinline__ void hs_st_uv(char *p, float x, bool wt) {
    st_out(reinterpret_cast<float *>(p), x, COH || wt);
}

template <int NT, int TW, int TH, int KMAX, int NTAP, bool LOUT = false, bool DB = false,
          bool FDIV = false,
          bool SHARED = !LOUT && ((TW + 2 * KMAX) * (TH + 2 * KMAX) + NT - 1) / NT <= 8,
          bool PF = SHARED && DB, bool COH = false, bool RING = false>
__device__ __forceinline__ void hs_tile(const HsImage &a, int t, int k,
                                        const HsTapTable<TW + 2 * KMAX> &tp, float2 *uv,
                                        float *lin = nullptr, int tid_in = -1) {
    static_assert(!(LOUT && DB) || RING,
                  "the linearisation output of a double-buffered step needs the ring taps");
    static_assert(!RING || NTAP == 8, "the ring taps are 8");
    constexpr bool PREFETCH = PF && SHARED && NTAP > 0 && !LOUT;
    // linearisation output stored from registers (no LDS staging)
    constexpr bool LREG = LOUT && DB;
    constexpr int PITCH = TW + 2 * KMAX, RHM = TH + 2 * KMAX;
    // tap d's LDS delta oy*PITCH + ox
    auto dl = [&](int d) -> int {
        if constexpr (RING)
            return HS_RING_OY[d & 7] * PITCH + HS_RING_OX[d & 7];
        else
            return tp.dl[d];
    };
    constexpr int NPX = PITCH * RHM;
    constexpr int P = (NPX + NT - 1) / NT;
    // tid_in: the thread index as a caller that loops over tiles passes it,
    // made opaque each iteration so that nothing derived from it is hoisted
    // out of the caller's loop (and kept live across tiles)
    const int tid = tid_in >= 0 ? tid_in : (int)threadIdx.x;
    __builtin_assume(tid >= 0 && tid < NT);
    const int tyi = t / a.tiles_x, txi = t - tyi * a.tiles_x;
    const int x0 = txi * TW, y0 = tyi * TH;
    const int x1 = min(x0 + TW, a.W), y1 = min(y0 + TH, a.H);
    // The window's corner is image pixel (x0 - k, y0 - k) for every tile, and
    // the region is the output tile grown by k on each side.  Where the
    // region reaches past the image border (a clipped tile), the region
    // cells outside the image are ghosts: each holds the value of the image
    // pixel its coordinates clamp to, loaded here and refreshed after every
    // step by that pixel (hs_ghost_store).  A tap that leaves the image then
    // reads the value the reference's clamped index reads, at the window
    // offset of any other tap, so border tiles run the same code as interior
    // ones.  (A radius-1 stencil reads only the ghost ring next to the
    // image; the rest of the ghosts are loaded and never read.)
    const int gx0 = x0 - k, gy0 = y0 - k;
    const int RW = x1 - x0 + 2 * k, RH = y1 - y0 + 2 * k;
    const bool clipped = (gx0 < 0) || (gy0 < 0) || (x1 + k > a.W) || (y1 + k > a.H);

    // Thread tid owns the window cells c = tid + j*NT (j < P), so the LDS
    // address of tap d of any owned pixel is (tid + dl[d]) + j*NT: one
    // per-thread base per tap, the slot a compile-time offset.  pk packs, for
    // a region cell inside the image, L << 27 | ry << 16 | E << 12 | rx (ry
    // and rx only in LOUT builds, the one user of them), where
    // L = k - (the cell's distance outside the output tile) is the last step
    // that updates it (the region of step it is the tile grown by k - it, so
    // the cell is updated exactly when it <= L), and E flags the image
    // borders the pixel lies on (bit 0 left, 1 right, 2 top, 3 bottom).
    // Ghosts and cells outside the region hold -1.
    float cx[P], cy[P], ct[P], cd[P], cr[FDIV ? P : 1];
    int pk[P];
    if (!clipped) {
        // Unclipped tile (most of them): the region is the window's top-left
        // (TW + 2k) x (TH + 2k) cells and every one of them is an image pixel,
        // so a cell's offset, last step and border flags (none) follow from
        // its window position alone.  All of a thread's loads are issued
        // before any of them is used.
        // Byte offsets from the region's corner fit 32 bits (H, W < 32768),
        // so each load is a uniform base plus one per-cell offset register.
        const int64_t corner = ((int64_t)gy0 * a.W + gx0) * (int64_t)sizeof(float);
        const char *const pIx = reinterpret_cast<const char *>(a.Ix) + corner;
        const char *const pIy = reinterpret_cast<const char *>(a.Iy) + corner;
        const char *const pIt = reinterpret_cast<const char *>(a.It) + corner;
        const char *const pu = reinterpret_cast<const char *>(a.u_in) + corner;
        const char *const pv = reinterpret_cast<const char *>(a.v_in) + corner;
        const char *const pd =
            a.den ? reinterpret_cast<const char *>(a.den) + corner : nullptr;
        const unsigned rowb = (unsigned)a.W * (unsigned)sizeof(float);
        float ui[P], vi[P];
        bool in[P];
        int rxs[P], rys[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int c = tid + j * NT;
            const int ry = c / PITCH, rx = c - ry * PITCH;
            rxs[j] = rx;
            rys[j] = ry;
            in[j] = ((j + 1) * NT <= NPX || c < NPX) && rx < RW && ry < RH;
            // a cell outside the region is never active (pk < 0): its
            // registers are left unset
            if (in[j]) {
                const uint64_t o = (unsigned)ry * rowb + (unsigned)rx * (unsigned)sizeof(float);
                cx[j] = ld_g(reinterpret_cast<const float *>(pIx + o));
                cy[j] = ld_g(reinterpret_cast<const float *>(pIy + o));
                ct[j] = ld_g(reinterpret_cast<const float *>(pIt + o));
                ui[j] = hs_ld_uv<COH>(pu + o);
                vi[j] = hs_ld_uv<COH>(pv + o);
                if (pd) cd[j] = ld_g(reinterpret_cast<const float *>(pd + o));
            }
        }
#pragma unroll
        for (int j = 0; j < P; ++j) {
            if (FDIV) cr[j] = 1.0f;
            pk[j] = -1;
            if (in[j]) {
                const int rx = rxs[j], ry = rys[j];
                // twice the cell's distance outside the tile along an axis,
                // from its distance to the tile's centre line (at 2k + T - 1
                // in doubled coordinates): |2r - (2k + T - 1)| - (T - 1), one
                // v_sad_u32 per axis
                const int ax = (int)hs_usad(2u * rx, 2u * k + (TW - 1), 0u - (TW - 1));
                const int ay = (int)hs_usad(2u * ry, 2u * k + (TH - 1), 0u - (TH - 1));
                const int m2 = max(max(ax, ay), 0);  // 2 x (distance outside)
                pk[j] = ((2 * k - m2) << 26) | (LOUT ? (ry << 16) | rx : 0);
                if (!pd) {
                    const float m1 = cx[j] * cx[j];
                    const float s1 = a.alpha2 + m1;
                    const float m2 = cy[j] * cy[j];
                    cd[j] = s1 + m2;
                }
                if (FDIV) cr[j] = hs_recip(cd[j]);
                uv[tid + j * NT] = make_float2(ui[j], vi[j]);
            }
        }
    } else {
        // Clipped tile: region cells outside the image are ghosts and load the
        // pixel their coordinates clamp to; image cells also load their
        // coefficients.  Byte offsets from the image start fit 32 bits
        // (H, W < 32768).  As above, every load is issued before any use.
        const char *const bIx = reinterpret_cast<const char *>(a.Ix);
        const char *const bIy = reinterpret_cast<const char *>(a.Iy);
        const char *const bIt = reinterpret_cast<const char *>(a.It);
        const char *const bu = reinterpret_cast<const char *>(a.u_in);
        const char *const bv = reinterpret_cast<const char *>(a.v_in);
        const char *const bd = reinterpret_cast<const char *>(a.den);
        const unsigned rowb = (unsigned)a.W * (unsigned)sizeof(float);
        float ui[P], vi[P];
        bool in[P], img[P];
        int gxs[P], gys[P], rxs[P], rys[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int c = tid + j * NT;
            const int ry = c / PITCH, rx = c - ry * PITCH;
            const int gx = gx0 + rx, gy = gy0 + ry;
            rxs[j] = rx;
            rys[j] = ry;
            gxs[j] = gx;
            gys[j] = gy;
            in[j] = ((j + 1) * NT <= NPX || c < NPX) && rx < RW && ry < RH;
            img[j] = in[j] && gx >= 0 && gx < a.W && gy >= 0 && gy < a.H;
            const unsigned o = (unsigned)clampi(gy, 0, a.H - 1) * rowb +
                               (unsigned)clampi(gx, 0, a.W - 1) * (unsigned)sizeof(float);
            if (in[j]) {
                ui[j] = hs_ld_uv<COH>(bu + o);
                vi[j] = hs_ld_uv<COH>(bv + o);
            }
            if (img[j]) {
                cx[j] = ld_g(reinterpret_cast<const float *>(bIx + o));
                cy[j] = ld_g(reinterpret_cast<const float *>(bIy + o));
                ct[j] = ld_g(reinterpret_cast<const float *>(bIt + o));
                if (bd) cd[j] = ld_g(reinterpret_cast<const float *>(bd + o));
            }
        }
#pragma unroll
        for (int j = 0; j < P; ++j) {
            if (FDIV) cr[j] = 1.0f;
            pk[j] = -1;
            if (img[j]) {
                const int gx = gxs[j], gy = gys[j];
                const int dx = max(max(x0 - gx, gx - x1 + 1), 0);
                const int dy = max(max(y0 - gy, gy - y1 + 1), 0);
                const int e = (gx == 0) | (gx == a.W - 1) << 1 | (gy == 0) << 2 |
                              (gy == a.H - 1) << 3;
                // an edge pixel's one ghost as a window offset in 9 signed
                // bits (CLIP 1 below)
                const int g = (e & 1) ? -1 : (e & 2) ? 1 : (e & 4) ? -PITCH : (e & 8) ? PITCH : 0;
                pk[j] = ((k - max(dx, dy)) << 27) | (e << 12) |
                        (LOUT ? (rys[j] << 16) | rxs[j] : SHARED && !COH ? (g & 0x1ff) : 0);
                if (!bd) {
                    const float m1 = cx[j] * cx[j];
                    const float s1 = a.alpha2 + m1;
                    const float m2 = cy[j] * cy[j];
                    cd[j] = s1 + m2;
                }
                if (FDIV) cr[j] = hs_recip(cd[j]);
            }
            if (in[j]) uv[tid + j * NT] = make_float2(ui[j], vi[j]);
        }
    }
    // LREG: each slot's byte offset from the output tile's corner when its
    // cell is a pixel of the output tile (stored every step), else ~0u
    unsigned lo[LREG ? P : 1];
    const int64_t lcorner = (int64_t)y0 * a.W + x0;
    const bool lwt = rows_whole_lines(a.W);
    if (LREG) {
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int lx = (pk[j] & 0xfff) - k, ly = ((pk[j] >> 16) & 0x7ff) - k;
            const bool tile = pk[j] >= 0 && (unsigned)lx < (unsigned)(x1 - x0) &&
                              (unsigned)ly < (unsigned)(y1 - y0);
            lo[j] = tile ? ((unsigned)ly * (unsigned)a.W + (unsigned)lx) * 4u : ~0u;
        }
    }
    __syncthreads();

    // One Jacobi step.  PAR (compile time, SHARED builds) is the window step
    // it reads with DB, (it - 1) & 1, so both windows' tap addresses are
    // constants; other builds pass 0 and select the window at run time.
    // CLIP (compile time): the tile's ghost cells to refresh.  0: none (an
    // unclipped tile: a loop without the ghost stores, whose per-cell border
    // masks would otherwise be set up for every tile).  1 (a tile clipped
    // along one axis): an updated pixel has at most one ghost, at the offset
    // g that pk carries, so the refresh is one masked store.  2: any ghosts,
    // through the border flags (hs_ghost_store: a corner pixel has three).
    auto step = [&](const int it, auto par_c, auto clip_c) {
        constexpr int PAR = decltype(par_c)::value;
        constexpr int CLIP = decltype(clip_c)::value;
        auto ghost = [&](float2 *w, int c, int pkc, float2 q) {
            if constexpr (CLIP == 1) {
                const int g = (pkc << 23) >> 23;
                if (g != 0) w[c + g] = q;
            } else if constexpr (CLIP == 2) {
                hs_ghost_store<PITCH>(w, c, pkc, q);
            }
        };
        const int act_min = it << 27;
        const int par = SHARED ? PAR : (it - 1) & 1;
        const float2 *uvr = DB ? uv + par * NPX : uv;
        float2 *uvw = DB ? uv + (par ^ 1) * NPX : uv;
        if (LREG) {
            // (opaque per step: otherwise each store address becomes a
            // 64-bit induction variable held across the loop)
#pragma unroll
            for (int j = 0; j < P; ++j) asm volatile("" : "+v"(lo[j]));
        }
        float nu[P], nv[P];
        bool act[P];
        if (PREFETCH) {
            // Software-pipelined form: the first NPF taps of cell j+1 are read
            // (updated or not: the window is padded) before cell j is
            // computed, so LDS latency overlaps arithmetic instead of stalling
            // the wave; the rest of a cell's taps are read as it starts.
            constexpr int NPF = !PREFETCH ? 1 : FDIV ? (NTAP + 3) / 4 : (NTAP + 1) / 2;
            const float2 *b0 = uvr + tid;
            float2 qa[NPF];
#pragma unroll
            for (int d = 0; d < NPF; ++d) qa[d] = b0[dl(d)];
#pragma unroll
            for (int j = 0; j < P; ++j) {
                float2 qr[PREFETCH ? NTAP - NPF + 1 : 1], qn[NPF];
#pragma unroll
                for (int d = NPF; d < NTAP; ++d) qr[d - NPF] = b0[j * NT + dl(d)];
                if (j + 1 < P) {
#pragma unroll
                    for (int d = 0; d < NPF; ++d) qn[d] = b0[(j + 1) * NT + dl(d)];
                }
                act[j] = pk[j] >= act_min;
                if (act[j]) {
                    float ub = 0.0f, vb = 0.0f;
#pragma unroll
                    for (int d = 0; d < NTAP; ++d) {
                        const float2 qd = d < NPF ? qa[d] : qr[d - NPF];
                        ub = fmaf(tp.w[d], qd.x, ub);
                        vb = fmaf(tp.w[d], qd.y, vb);
                    }
                    (void)hs_update<FDIV>(cx[j], cy[j], ct[j], cd[j], FDIV ? cr[j] : 0.0f, ub,
                                          vb, nu[j], nv[j]);
                    if (DB) {
                        const float2 q = make_float2(nu[j], nv[j]);
                        uvw[tid + j * NT] = q;
                        ghost(uvw, tid + j * NT, pk[j], q);
                    }
                }
#pragma unroll
                for (int d = 0; d < NPF; ++d) qa[d] = qn[d];
            }
        } else {
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int pkj = pk[j];
            act[j] = pkj >= act_min;
            if (act[j]) {
                float ub = 0.0f, vb = 0.0f;
                // an updated cell is at least one cell inside the window on
                // every side, so every tap stays inside it
                int c = tid + j * NT;
                if (!SHARED) asm volatile("" : "+v"(c));
                const float2 *base = uvr + c;
#pragma unroll
                for (int d = 0; d < (NTAP ? NTAP : HS_MAX_TAPS); ++d) {
                    if (NTAP || d < tp.nd) {
                        const float2 q = base[dl(d)];
                        ub = fmaf(tp.w[d], q.x, ub);
                        vb = fmaf(tp.w[d], q.y, vb);
                    }
                }
                const float tt = hs_update<FDIV>(cx[j], cy[j], ct[j], cd[j],
                                                 FDIV ? cr[j] : 0.0f, ub, vb, nu[j], nv[j]);
                if (DB) {
                    const float2 q = make_float2(nu[j], nv[j]);
                    uvw[tid + j * NT] = q;
                    ghost(uvw, tid + j * NT, pkj, q);
                }
                if (LREG && lo[j] != ~0u) {
                    const int64_t so = (int64_t)(it - 1) * a.lstride + lcorner;
                    auto at = [&](float *plane) {
                        return reinterpret_cast<float *>(reinterpret_cast<char *>(plane + so) +
                                                         lo[j]);
                    };
                    if (a.lt) st_out(at(a.lt), tt, lwt);  // (t is optional)
                    st_out(at(a.lub), ub, lwt);
                    st_out(at(a.lvb), vb, lwt);
                }
                if (LOUT && !LREG) {
                    const int rx = pkj & 0xfff, ry = (pkj >> 16) & 0x7ff;
                    const int lx = gx0 + rx - x0, ly = gy0 + ry - y0;
                    if (lx >= 0 && lx < TW && ly >= 0 && ly < TH) {
                        const int li = ly * TW + lx;
                        lin[li] = tt;
                        lin[TW * TH + li] = ub;
                        lin[2 * TW * TH + li] = vb;
                    }
                }
            }
        }
        }
        if (DB) {
            // the window written here was last read in step it-1, before the
            // previous barrier
            __syncthreads();
            return;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < P; ++j) {
            if (act[j]) {
                const float2 q = make_float2(nu[j], nv[j]);
                uv[tid + j * NT] = q;
                ghost(uv, tid + j * NT, pk[j], q);
            }
        }
        if (LOUT && !LREG) {
            // lin is rewritten only in the next step's compute, after the
            // barrier below
            const int64_t so = (int64_t)(it - 1) * a.lstride;
            for (int i = tid; i < TW * TH; i += NT) {
                const int gy = y0 + i / TW, gx = x0 + i % TW;
                if (gx < x1 && gy < y1) {
                    const int64_t g = so + (int64_t)gy * a.W + gx;
                    if (a.lt) st_out(a.lt + g, lin[i], rows_whole_lines(a.W));  // (t is optional)
                    st_out(a.lub + g, lin[TW * TH + i], rows_whole_lines(a.W));
                    st_out(a.lvb + g, lin[2 * TW * TH + i], rows_whole_lines(a.W));
                }
            }
        }
        __syncthreads();
    };
    auto run = [&](auto clip_c) {
        if (DB && SHARED) {
            for (int it = 1; it <= k; it += 2) {
                step(it, std::integral_constant<int, 0>(), clip_c);
                if (it < k) step(it + 1, std::integral_constant<int, 1>(), clip_c);
            }
        } else {
            for (int it = 1; it <= k; ++it) step(it, std::integral_constant<int, 0>(), clip_c);
        }
    };
    if (!clipped) {
        run(std::integral_constant<int, 0>());
    } else {
        // (LOUT: pk's low bits hold the cell's window position; per-cell tap
        // addresses and the persistent form (COH): no registers to spare for
        // a third loop body)
        if constexpr (LOUT || !SHARED || COH) {
            run(std::integral_constant<int, 2>());
        } else {
            // An updated pixel carrying two border flags lies in a tile
            // clipped along both axes (an image corner pixel: an updated
            // pixel's ghost is inside the region only on a side where the
            // region leaves the image) or in an image one pixel wide or
            // tall; those tiles take the general form.
            if (a.W == 1 || a.H == 1 ||
                (((gx0 < 0) || (x1 + k > a.W)) && ((gy0 < 0) || (y1 + k > a.H))))
                run(std::integral_constant<int, 2>());
            else
                run(std::integral_constant<int, 1>());
        }
    }

    // Store the output tile from the final window (complete: every step ends
    // with a barrier) in tile order: thread tid stores tile cells tid + j*NT,
    // so a wavefront writes whole row segments, at byte offsets from the
    // tile's corner (32 bits: H, W < 32768).
    {
        const float2 *const fin = (DB ? uv + (k & 1) * NPX : uv) + k * PITCH + k;
        const int64_t corner = ((int64_t)y0 * a.W + x0) * (int64_t)sizeof(float);
        char *const pu = reinterpret_cast<char *>(a.u_out) + corner;
        char *const pv = reinterpret_cast<char *>(a.v_out) + corner;
        const unsigned rowb = (unsigned)a.W * (unsigned)sizeof(float);
        const int tw = x1 - x0, th = y1 - y0;
        const bool wt = rows_whole_lines(a.W);
        constexpr int NO = (TW * TH + NT - 1) / NT;
        // all of a thread's cells are read before the first store, so no
        // store waits for an earlier one to leave the registers it reads
        float2 q[NO];
        unsigned o[NO];
        bool st[NO];
#pragma unroll
        for (int j = 0; j < NO; ++j) {
            const int i = tid + j * NT;
            const int ty = i / TW, tx = i - ty * TW;
            st[j] = ((j + 1) * NT <= TW * TH || i < TW * TH) && tx < tw && ty < th;
            o[j] = (unsigned)ty * rowb + (unsigned)tx * (unsigned)sizeof(float);
            if (st[j]) q[j] = fin[ty * PITCH + tx];
        }
#pragma unroll
        for (int j = 0; j < NO; ++j) {
            if (st[j]) {
                hs_st_uv<COH>(pu + o[j], q[j].x, wt);
                hs_st_uv<COH>(pv + o[j], q[j].y, wt);
            }
        }
    }
    __syncthreads();  // the window is reused by the work-group's next tile
}


// Ring-ordered tile body for 64 x 32 tiles, k <= 8, double-buffered window,
// the 8-tap averaging stencil at the canonical offsets (HS_RING_OX/OY, in
// that tap order; its weights stay run-time values).
//
// The cells a Jacobi step updates are the output tile grown by k - it, so a
// cell's last step is k - (its distance from the tile).  hs_tile gives thread
// tid the window cells tid + j*NT in row-major order; a wavefront's 64 cells
// then straddle rows and distances, and 12 % of the lanes of the wave-slots
// that run are idle.  Here cells are owned by distance instead: wave-slot S
// (wave S % 16 of the work-group, cell slot S / 16 of each of its threads)
// holds core row S of the tile (S < 32, last step k) or a quarter of the ring
// at distance m (S = 32 + 4(m - 1) .. +3, last step k - m), so nearly every
// lane of a running wave-slot is active: 368 wave-slot steps per tile at
// k = 8 instead of 396.  A cell's LDS address is then its own register and
// its taps are compile-time offsets from it, which is why the stencil offsets
// are fixed.  (u, v) of the whole region are still loaded in row-major order
// (coalesced); each owner loads its own coefficients.  Same per-pixel
// arithmetic, tap order, ghost cells and results as hs_tile.
// The window's row pitch is odd (81 float2 = 162 dwords): the wave-slots on
// the left and right rings hold 64 cells of one column, and with the even
// pitch 80 (160 dwords, 32 mod the 64 LDS banks) every lane of such a read
// fell on the same two bank pairs.  At 162 dwords the 16 lanes served in one
// LDS cycle touch 16 distinct bank pairs.
constexpr int HS_RING_PITCH = 64 + 2 * 8 + 1;
constexpr int HS_RING_NPX = HS_RING_PITCH * (32 + 2 * 8);
// LDS of the ring form's two windows, with PITCH + 1 cells before and after
// them for the taps of spare lanes (the kernel passes the allocation + PAD)
constexpr int HS_RING_PAD = HS_RING_PITCH + 1;
constexpr int HS_RING_LDS = HS_RING_PAD + 2 * HS_RING_NPX + HS_RING_PAD;

template <bool FDIV>
__device__ __forceinline__ void hs_tile_ring(const HsImage &a, int t, int k,
                                             const HsTapTable<HS_RING_PITCH> &tp, float2 *uv) {
    constexpr int NT = 1024, TW = 64, TH = 32;  // k <= 8 (the window's ring width)
    constexpr int PITCH = HS_RING_PITCH, NPX = HS_RING_NPX;
    constexpr int P = 4;  // cell slots per thread: 32 core rows + 4 per ring, <= 64 slots
    static_assert(NPX <= P * NT, "window larger than the loader pass");
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tyi = t / a.tiles_x, txi = t - tyi * a.tiles_x;
    const int x0 = txi * TW, y0 = tyi * TH;
    const int x1 = min(x0 + TW, a.W), y1 = min(y0 + TH, a.H);
    const int gx0 = x0 - k, gy0 = y0 - k;
    const int RW = x1 - x0 + 2 * k, RH = y1 - y0 + 2 * k;
    const bool clipped = (gx0 < 0) || (gy0 < 0) || (x1 + k > a.W) || (y1 + k > a.H);

    // Byte offsets from the image start fit 32 bits (H, W < 32768): every
    // load below is a uniform base plus one offset register.
    const unsigned rowb = (unsigned)a.W * (unsigned)sizeof(float);
    const char *const bIx = reinterpret_cast<const char *>(a.Ix);
    const char *const bIy = reinterpret_cast<const char *>(a.Iy);
    const char *const bIt = reinterpret_cast<const char *>(a.It);
    const char *const bd = reinterpret_cast<const char *>(a.den);
    const char *const bu = reinterpret_cast<const char *>(a.u_in);
    const char *const bv = reinterpret_cast<const char *>(a.v_in);

    // (u, v) of every region cell into window 0, row-major (coalesced); cells
    // outside the image load the pixel they clamp to (ghosts, see hs_tile)
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const int c = tid + j * NT;
        const int ry = c / PITCH, rx = c - ry * PITCH;
        if (((j + 1) * NT <= NPX || c < NPX) && rx < RW && ry < RH) {
            const unsigned o = (unsigned)clampi(gy0 + ry, 0, a.H - 1) * rowb +
                               (unsigned)clampi(gx0 + rx, 0, a.W - 1) * (unsigned)sizeof(float);
            uv[c] = make_float2(ld_g(reinterpret_cast<const float *>(bu + o)),
                                ld_g(reinterpret_cast<const float *>(bv + o)));
        }
    }

    // owned cells: window position, last step, border flags, coefficients.
    // pk as in hs_tile (last step << 27 | ry << 16 | border flags << 12 | rx;
    // -1: not updated); ca = the cell's window index (0 for a spare lane, whose
    // taps then stay inside the padded allocation and are never used).  Slot
    // S's ring distance and quarter are uniform over the wave; only the
    // position along the ring is per lane.
    float cx[P], cy[P], ct[P], cd[P], cr[FDIV ? P : 1];
    int pk[P], ca[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const int S = wave + 16 * j;
        int rx, ry, L;
        if (S < TH) {  // core row S
            rx = k + lane;
            ry = k + S;
            L = k;
        } else {
            const int m = (S - TH) / 4 + 1;  // ring distance (uniform)
            const int i = ((S - TH) & 3) * 64 + lane;
            const int top = TW + 2 * m, side = TH + 2 * m - 2;
            const int i1 = i - top, i2 = i - 2 * top, i3 = i2 - side;
            L = (m < k && i3 < side) ? k - m : -1;
            rx = i < top ? k - m + i : i1 < top ? k - m + i1 : i3 < 0 ? k - m : k + TW - 1 + m;
            ry = i < top ? k - m : i1 < top ? k + TH - 1 + m : k - m + 1 + (i3 < 0 ? i2 : i3);
        }
        const int gx = gx0 + rx, gy = gy0 + ry;
        // an unclipped tile's cells are all image pixels off the border
        const bool img =
            L >= 0 && (!clipped || ((unsigned)gx < (unsigned)a.W && (unsigned)gy < (unsigned)a.H));
        ca[j] = L >= 0 ? ry * PITCH + rx : 0;
        pk[j] = -1;
        if (img) {
            const int e = !clipped ? 0
                                   : (gx == 0) | (gx == a.W - 1) << 1 | (gy == 0) << 2 |
                                         (gy == a.H - 1) << 3;
            pk[j] = (L << 27) | (e << 12);
            const unsigned o = (unsigned)gy * rowb + (unsigned)gx * (unsigned)sizeof(float);
            cx[j] = ld_g(reinterpret_cast<const float *>(bIx + o));
            cy[j] = ld_g(reinterpret_cast<const float *>(bIy + o));
            ct[j] = ld_g(reinterpret_cast<const float *>(bIt + o));
            if (bd) cd[j] = ld_g(reinterpret_cast<const float *>(bd + o));
        }
    }
#pragma unroll
    for (int j = 0; j < P; ++j) {
        if (FDIV) cr[j] = 1.0f;
        if (pk[j] >= 0) {
            if (!bd) {
                const float m1 = cx[j] * cx[j];
                const float s1 = a.alpha2 + m1;
                const float m2 = cy[j] * cy[j];
                cd[j] = s1 + m2;
            }
            if (FDIV) cr[j] = hs_recip(cd[j]);
        }
    }
    __syncthreads();

    // One Jacobi step reading window PAR and writing the other; the first
    // taps of the next cell are read before the current one is computed.
    auto step = [&](const int it, auto par_c) {
        constexpr int PAR = decltype(par_c)::value;
        const int act_min = it << 27;
        const float2 *uvr = uv + PAR * NPX;
        float2 *uvw = uv + (PAR ^ 1) * NPX;
        constexpr int NPF = FDIV ? 2 : 4;
        float2 qa[NPF];
#pragma unroll
        for (int d = 0; d < NPF; ++d) qa[d] = uvr[ca[0] + HS_RING_OY[d] * PITCH + HS_RING_OX[d]];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            float2 qr[8 - NPF], qn[NPF];
#pragma unroll
            for (int d = NPF; d < 8; ++d)
                qr[d - NPF] = uvr[ca[j] + HS_RING_OY[d] * PITCH + HS_RING_OX[d]];
            if (j + 1 < P) {
#pragma unroll
                for (int d = 0; d < NPF; ++d)
                    qn[d] = uvr[ca[j + 1] + HS_RING_OY[d] * PITCH + HS_RING_OX[d]];
            }
            if (pk[j] >= act_min) {
                float ub = 0.0f, vb = 0.0f;
#pragma unroll
                for (int d = 0; d < 8; ++d) {
                    const float2 qd = d < NPF ? qa[d] : qr[d - NPF];
                    ub = fmaf(tp.w[d], qd.x, ub);
                    vb = fmaf(tp.w[d], qd.y, vb);
                }
                float nu, nv;
                (void)hs_update<FDIV>(cx[j], cy[j], ct[j], cd[j], FDIV ? cr[j] : 0.0f, ub, vb,
                                      nu, nv);
                const float2 q = make_float2(nu, nv);
                uvw[ca[j]] = q;
                if (clipped) hs_ghost_store<PITCH>(uvw, ca[j], pk[j], q);
            }
#pragma unroll
            for (int d = 0; d < NPF; ++d) qa[d] = qn[d];
        }
        __syncthreads();
    };
    for (int it = 1; it <= k; it += 2) {
        step(it, std::integral_constant<int, 0>());
        if (it < k) step(it + 1, std::integral_constant<int, 1>());
    }

    // the core rows' owners store the output tile from the last window
    const float2 *fin = uv + (k & 1) * NPX;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int S = wave + 16 * j;
        const int gx = x0 + lane, gy = y0 + S;
        if (gx < x1 && gy < y1) {
            const int64_t g = (int64_t)gy * a.W + gx;
            const float2 q = fin[(k + S) * PITCH + k + lane];
            a.u_out[g] = q.x;
            a.v_out[g] = q.y;
        }
    }
    __syncthreads();  // the window is reused by the work-group's next tile
}

// Column-strip tile body.  Thread q of the work-group owns S vertically
// adjacent pixels of one column of the region (strips are laid out column-
// fastest so a wavefront reads consecutive LDS words).  Per step it reads the
// 3 x (S+2) float2 window around its strip once -- 3(S+2)/S LDS reads per
// pixel instead of one per tap -- and applies the taps from registers.  Tap
// offsets are runtime values; a uniform switch per tap selects the window
// position, so register indices stay compile-time and the fma order is still
// the user's tap order.  Window coordinates are clamped to the region, which
// equals the reference's image clamp on sides where the region meets the image
// border and is never reached by an updated pixel on the other sides.
template <int NT, int TW, int TH, int KMAX, int S, int NTAP>
__device__ __forceinline__ void hs_tile_strip(const HsImage &a, int t, int k,
                                              const HsTapTable<TW + 2 * KMAX> &tp,
                                              float2 *uv) {
    constexpr int PITCH = TW + 2 * KMAX, RHM = TH + 2 * KMAX;
    constexpr int SR = (RHM + S - 1) / S;
    constexpr int NS = PITCH * SR;
    constexpr int M = (NS + NT - 1) / NT;
    const int tid = threadIdx.x;
    const int tyi = t / a.tiles_x, txi = t - tyi * a.tiles_x;
    const int x0 = txi * TW, y0 = tyi * TH;
    const int x1 = min(x0 + TW, a.W), y1 = min(y0 + TH, a.H);
    const int gx0 = max(x0 - k, 0), gy0 = max(y0 - k, 0);
    const int gx1 = min(x1 + k, a.W), gy1 = min(y1 + k, a.H);
    const int RW = gx1 - gx0, RH = gy1 - gy0;

    float cx[M][S], cy[M][S], ct[M][S], cd[M][S];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int q = tid + m * NT;
        const int col = q % PITCH, srow = q / PITCH;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int ry = srow * S + s;
            cx[m][s] = cy[m][s] = ct[m][s] = 0.0f;
            cd[m][s] = 1.0f;
            if (q < NS && col < RW && ry < RH) {
                const int64_t g = (int64_t)(gy0 + ry) * a.W + gx0 + col;
                const float ix = a.Ix[g], iy = a.Iy[g];
                cx[m][s] = ix;
                cy[m][s] = iy;
                ct[m][s] = a.It[g];
                if (a.den) {
                    cd[m][s] = a.den[g];
                } else {
                    const float m1 = ix * ix;
                    const float s1 = a.alpha2 + m1;
                    const float m2 = iy * iy;
                    cd[m][s] = s1 + m2;
                }
                uv[ry * PITCH + col] = make_float2(a.u_in[g], a.v_in[g]);
            }
        }
    }
    __syncthreads();

    for (int it = 1; it <= k; ++it) {
        const int e = k - it;
        const int rx0 = max(x0 - e, 0) - gx0, ry0 = max(y0 - e, 0) - gy0;
        const int rx1 = min(x1 + e, a.W) - gx0, ry1 = min(y1 + e, a.H) - gy0;
        float nu[M][S], nv[M][S];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            int q = tid + m * NT;
            asm volatile("" : "+v"(q));  // keep the window addresses inside the step loop
            const int col = q % PITCH, srow = q / PITCH;
            const int cm = min(max(col - 1, 0), RW - 1), c0 = min(col, RW - 1);
            const int cp = min(col + 1, RW - 1);
            float2 win[S + 2][3];
#pragma unroll
            for (int j = 0; j < S + 2; ++j) {
                const int r = min(max(srow * S - 1 + j, 0), RH - 1) * PITCH;
                win[j][0] = uv[r + cm];
                win[j][1] = uv[r + c0];
                win[j][2] = uv[r + cp];
            }
            float ub[S], vb[S];
#pragma unroll
            for (int s = 0; s < S; ++s) ub[s] = vb[s] = 0.0f;
#pragma unroll
            for (int d = 0; d < (NTAP ? NTAP : HS_MAX_TAPS); ++d) {
                if (NTAP || d < tp.nd) {
                    const float w = tp.w[d];
#define HS_STRIP_TAP(P)                                                       \
    case P:                                                                   \
        _Pragma("unroll") for (int s = 0; s < S; ++s) {                      \
            ub[s] = fmaf(w, win[s + (P) / 3][(P) % 3].x, ub[s]);             \
            vb[s] = fmaf(w, win[s + (P) / 3][(P) % 3].y, vb[s]);             \
        }                                                                     \
        break;
                    switch (tp.pos[d]) {
                        HS_STRIP_TAP(0)
                        HS_STRIP_TAP(1)
                        HS_STRIP_TAP(2)
                        HS_STRIP_TAP(3)
                        HS_STRIP_TAP(4)
                        HS_STRIP_TAP(5)
                        HS_STRIP_TAP(6)
                        HS_STRIP_TAP(7)
                        HS_STRIP_TAP(8)
                        default: break;
                    }
#undef HS_STRIP_TAP
                }
            }
#pragma unroll
            for (int s = 0; s < S; ++s)
                (void)hs_update(cx[m][s], cy[m][s], ct[m][s], cd[m][s], 0.0f, ub[s], vb[s],
                                nu[m][s], nv[m][s]);
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const int q = tid + m * NT;
            const int col = q % PITCH, srow = q / PITCH;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int ry = srow * S + s;
                if (q < NS && col >= rx0 && col < rx1 && ry >= ry0 && ry < ry1)
                    uv[ry * PITCH + col] = make_float2(nu[m][s], nv[m][s]);
            }
        }
        __syncthreads();
    }

#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int q = tid + m * NT;
        const int col = q % PITCH, srow = q / PITCH;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int ry = srow * S + s;
            const int gx = gx0 + col, gy = gy0 + ry;
            if (q < NS && gx >= x0 && gx < x1 && gy >= y0 && gy < y1 && ry < RH) {
                const int64_t g = (int64_t)gy * a.W + gx;
                const float2 v = uv[ry * PITCH + col];
                a.u_out[g] = v.x;
                a.v_out[g] = v.y;
            }
        }
    }
    __syncthreads();
}

}  // namespace hmx
