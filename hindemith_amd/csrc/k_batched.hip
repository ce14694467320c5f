// a14: many Horn-Schunck pairs of uneven sizes in one captured launch
// sequence (no counterpart in the reference, which runs one program call at a
// time).  Two forms share the tile list and the tile body; the autotuner
// picks one per size signature:
//   * stepped (variants 8-11, the default): one launch per phase over every
//     tile of every pair (see hs_multi_phase_kernel);
//   * persistent (variants 0-7), below.
//
// The tiles of all pairs form one list.  A grid no larger than what the GPU
// holds at once walks that list: first the derivative setup of every tile,
// then ceil(iters / k) phases of k fused Jacobi steps (the same tile body as
// the single-image kernel, hs_tile.h), with a grid-wide barrier between
// phases so every tile of every pair sees the previous phase's halo.  Pairs
// ping-pong between (u, v) and (wu, wv) so that the last phase writes (u, v):
// with an odd phase count the setup also copies (u, v) into (wu, wv) and the
// first phase reads from there.  (The stepped form copies only for a single
// phase: it alternates with a third plane pair instead, phase_image_stepped.)
//
// The barrier is an arrival counter in device memory: each work-group's
// first thread releases its writes (agent-scope fence), increments the
// counter and waits until all work-groups of the grid have arrived, then
// acquires.  The wait is bounded: after spin_limit polls, or as soon as any
// work-group has given up, it sets *status and every work-group returns, so
// the grid always drains.  The launch is captured in a hipGraph (counter and
// status reset by a fill kernel, never a memset node) and replayed.
#include "hs_tile.h"

#include <algorithm>
#include <cstring>
#include <list>
#include <mutex>
#include <vector>

#pragma clang fp contract(off)

namespace hmx {

struct PairDesc {
    const float *I1, *I2;
    float *Ix, *Iy, *It;
    float *u, *v, *wu, *wv;
    float *zu, *zv;  // stepped form: third plane pair (see phase_image_stepped)
    float alpha2;
    int H, W, tiles_x, tile_base, ntiles;
};

struct BatchArgs {
    // the descriptor of each tile's pair, one copy per tile of the list, so a
    // work-group finds its pair with one load whose address it knows at once
    // (a pair index followed by the pair's descriptor would be two dependent
    // loads ahead of every tile)
    const PairDesc *tiles;
    unsigned *claims;      // per (phase, XCD) tile-claim counters, kClaimStride apart
    unsigned *barrier;
    int *status;
    long long spin_limit;
    int npairs, total_tiles, iters, k;
    // 1: the setup copies (u, v) into (wu, wv) and phase 0 reads them from
    // there (persistent form: odd phase count; stepped: one phase)
    int phase0;
    int nph;  // phase count
};

struct SetupTaps {
    Taps dx, dy, dt;
};

// Work-groups are dealt round-robin over the 8 XCDs.  XCD x owns a contiguous
// share [base, base + cnt) of the tile list, so neighbouring tiles of a pair
// meet in one L2; inside a phase its work-groups claim tiles from that share
// with an atomic counter, so cheap (clipped, small) and costly tiles balance.
constexpr int kXcds = 8;
// Default barrier polls before a work-group gives up (~190 ns each: ~3 s, far
// above one phase of any batch that fits in memory).
constexpr long long kDefaultSpinLimit = 1LL << 24;
constexpr int kClaimStride = 32;  // counters 128 B apart

__device__ __forceinline__ void xcd_share(int total, int x, int &base, int &cnt) {
    const int per = total / kXcds, rem = total % kXcds;
    base = x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per;
    cnt = per + (x < rem ? 1 : 0);
}

__device__ __forceinline__ bool grid_sync(unsigned *ctr, unsigned target, int *status,
                                          long long limit) {
    __shared__ int ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(ctr, 1u);
        int good = 1;
        long long n = 0;
        while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                ++n > limit) {
                good = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (!good) atomicExch(status, 1);
        __threadfence();
        ok = good;
    }
    __syncthreads();
    return ok != 0;
}

__device__ inline float setup_tap(const float *__restrict__ x, int H, int W, int r, int c,
                                  const Taps &t) {
    float acc = 0.0f;
    for (int d = 0; d < t.nd; ++d) {
        const int xx = clampi(c + t.ox[d], 0, W - 1), yy = clampi(r + t.oy[d], 0, H - 1);
        acc = fmaf(t.w[d], ld_g(x + (int64_t)yy * W + xx), acc);
    }
    return acc;
}

// 4-tap form: the reads issued before the fma chain.
__device__ __forceinline__ float setup_tap4(const float *__restrict__ x, int H, int W, int r,
                                            int c, const Taps &t) {
    float v[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        const int xx = clampi(c + t.ox[d], 0, W - 1), yy = clampi(r + t.oy[d], 0, H - 1);
        v[d] = ld_g(x + (int64_t)yy * W + xx);
    }
    float acc = 0.0f;
#pragma unroll
    for (int d = 0; d < 4; ++d) acc = fmaf(t.w[d], v[d], acc);
    return acc;
}

// Derivative setup of tile w of the list for stencils inside the cube
// (taps_in_cube): thread tid takes column tid % TW of the tile and RPT
// consecutive rows, and reads the (RPT + 1)-row, 2-column neighbourhood of
// each frame once (row r+1 of one output row is row r of the next), 6 reads
// per pixel instead of 24; each tap is then a register select, accumulated in
// tap order from 0 exactly as setup_tile does, so the results are the same.
template <int NT, int TW, int TH>
__device__ __forceinline__ void setup_tile_cube(const BatchArgs &b, const SetupTaps &st, int w) {
    static_assert(NT % TW == 0 && (TW * TH) % NT == 0, "whole columns of rows per thread");
    constexpr int RPT = TW * TH / NT;
    const PairDesc &P = b.tiles[w];
    const int t = w - P.tile_base;
    const int tyi = t / P.tiles_x, txi = t - tyi * P.tiles_x;
    const int c = txi * TW + (int)(threadIdx.x % TW);
    const int r0 = tyi * TH + (int)(threadIdx.x / TW) * RPT;
    if (c >= P.W || r0 >= P.H) return;
    const int c1 = min(c + 1, P.W - 1);
    const bool wt = rows_whole_lines(P.W);
    // every read is issued before the first store (the descriptor's pointers
    // may alias as far as the compiler knows, so a read after a store would
    // wait for it)
    float a0[RPT + 1], a1[RPT + 1], b0[RPT + 1], b1[RPT + 1], uu[RPT], vv[RPT];
#pragma unroll
    for (int j = 0; j <= RPT; ++j) {
        const int64_t row = (int64_t)min(r0 + j, P.H - 1) * P.W;
        a0[j] = ld_g(P.I1 + row + c);
        a1[j] = ld_g(P.I1 + row + c1);
        b0[j] = ld_g(P.I2 + row + c);
        b1[j] = ld_g(P.I2 + row + c1);
    }
    if (b.phase0) {  // the first phase reads the start state from (wu, wv)
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int64_t p = (int64_t)min(r0 + q, P.H - 1) * P.W + c;
            uu[q] = ld_g(P.u + p);
            vv[q] = ld_g(P.v + p);
        }
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int r = r0 + q;
        if (r >= P.H) break;
        const int64_t p = (int64_t)r * P.W + c;
        if (b.phase0) {
            st_out(P.wu + p, uu[q], wt);
            st_out(P.wv + p, vv[q], wt);
        }
        const float e1 = cube_apply(st.dx, a0[q], a1[q], a0[q + 1], a1[q + 1]);
        const float e2 = cube_apply(st.dx, b0[q], b1[q], b0[q + 1], b1[q + 1]);
        const float f1 = cube_apply(st.dy, a0[q], a1[q], a0[q + 1], a1[q + 1]);
        const float f2 = cube_apply(st.dy, b0[q], b1[q], b0[q + 1], b1[q + 1]);
        const float g2 = cube_apply(st.dt, b0[q], b1[q], b0[q + 1], b1[q + 1]);
        const float g1 = cube_apply(st.dt, a0[q], a1[q], a0[q + 1], a1[q + 1]);
        st_out(P.Ix + p, e1 + e2, wt);
        st_out(P.Iy + p, f1 + f2, wt);
        st_out(P.It + p, g2 - g1, wt);
    }
}

// Derivative setup of tile w of the list (Ix = Dx*I1 + Dx*I2, ...), one
// rounding per op.  TAP4: all three stencils have 4 taps (the canonical
// cube), taken by the unrolled form.
template <int NT, int TW, int TH, bool TAP4 = false>
__device__ __forceinline__ void setup_tile(const BatchArgs &b, const SetupTaps &st, int w) {
    const PairDesc &P = b.tiles[w];
    const int t = w - P.tile_base;
    const int tyi = t / P.tiles_x, txi = t - tyi * P.tiles_x;
    for (int i = threadIdx.x; i < TW * TH; i += NT) {
        const int r = tyi * TH + i / TW, c = txi * TW + i % TW;
        if (r >= P.H || c >= P.W) continue;
        const int64_t p = (int64_t)r * P.W + c;
        if (b.phase0) {  // the first phase reads the start state from (wu, wv)
            st_out(P.wu + p, ld_g(P.u + p), rows_whole_lines(P.W));
            st_out(P.wv + p, ld_g(P.v + p), rows_whole_lines(P.W));
        }
        if (TAP4) {
            const float a1 = setup_tap4(P.I1, P.H, P.W, r, c, st.dx);
            const float a2 = setup_tap4(P.I2, P.H, P.W, r, c, st.dx);
            const float b1 = setup_tap4(P.I1, P.H, P.W, r, c, st.dy);
            const float b2 = setup_tap4(P.I2, P.H, P.W, r, c, st.dy);
            const float t2 = setup_tap4(P.I2, P.H, P.W, r, c, st.dt);
            const float t1 = setup_tap4(P.I1, P.H, P.W, r, c, st.dt);
            st_out(P.Ix + p, a1 + a2, rows_whole_lines(P.W));
            st_out(P.Iy + p, b1 + b2, rows_whole_lines(P.W));
            st_out(P.It + p, t2 - t1, rows_whole_lines(P.W));
            continue;
        }
        const float a1 = setup_tap(P.I1, P.H, P.W, r, c, st.dx);
        const float a2 = setup_tap(P.I2, P.H, P.W, r, c, st.dx);
        st_out(P.Ix + p, a1 + a2, rows_whole_lines(P.W));
        const float b1 = setup_tap(P.I1, P.H, P.W, r, c, st.dy);
        const float b2 = setup_tap(P.I2, P.H, P.W, r, c, st.dy);
        st_out(P.Iy + p, b1 + b2, rows_whole_lines(P.W));
        const float t2 = setup_tap(P.I2, P.H, P.W, r, c, st.dt);
        const float t1 = setup_tap(P.I1, P.H, P.W, r, c, st.dt);
        st_out(P.It + p, t2 - t1, rows_whole_lines(P.W));
    }
}

// The image phase ph reads and writes: (u, v) -> (wu, wv) when ph + phase0 is
// even, back otherwise, so the last phase always writes (u, v).
__device__ __forceinline__ HsImage phase_image(const PairDesc &P, int ph, int phase0) {
    HsImage im;
    im.Ix = P.Ix;
    im.Iy = P.Iy;
    im.It = P.It;
    im.den = nullptr;
    const bool even = ((ph + phase0) & 1) == 0;
    im.u_in = even ? P.u : P.wu;
    im.v_in = even ? P.v : P.wv;
    im.u_out = even ? P.wu : P.u;
    im.v_out = even ? P.wv : P.v;
    im.alpha2 = P.alpha2;
    im.H = P.H;
    im.W = P.W;
    im.tiles_x = P.tiles_x;
    return im;
}

// The stepped form's planes: phase ph reads X_ph and writes X_ph+1, where
// X_0 = X_nph = (u, v) and, between them, X_p = (wu, wv) for odd p and
// (zu, zv) for even p.  With an even phase count (zu, zv) is (u, v) itself;
// with an odd one above 1 it is a plane pair of the plan's own, so no phase
// reads what it writes and no copy is needed.  One phase (nph = 1) reads the
// copy the setup made in (wu, wv) (phase0).
__device__ __forceinline__ HsImage phase_image_stepped(const PairDesc &P, int ph, int nph,
                                                       int phase0) {
    HsImage im = phase_image(P, 0, 0);
    auto plane = [&](int p, float *&pu, float *&pv) {
        if (p == nph || (p == 0 && !phase0)) {
            pu = P.u;
            pv = P.v;
        } else if (p & 1 || p == 0) {
            pu = P.wu;
            pv = P.wv;
        } else {
            pu = P.zu;
            pv = P.zv;
        }
    };
    float *iu, *iv, *ou, *ov;
    plane(ph, iu, iv);
    plane(ph + 1, ou, ov);
    im.u_in = iu;
    im.v_in = iv;
    im.u_out = ou;
    im.v_out = ov;
    return im;
}

// WPE: waves per SIMD the kernel is compiled to allow (the compiler must fit
// its VGPRs and SGPRs to it), from which the resident grid is derived.
// DB: two LDS windows, one barrier per Jacobi step (see hs_tile).
template <int NT, int TW, int TH, int KMAX, int NTAP, int WPE, bool DB = false,
          bool FDIV = false>
__global__ __launch_bounds__(NT, WPE) void hs_batched_kernel(BatchArgs b, SetupTaps st,
                                                             Taps taps) {
    using Win = HsWindow<NT, TW, TH, KMAX, DB>;
    __shared__ float2 uvm[Win::SIZE];
    float2 *const uv = uvm + Win::PAD;
    __shared__ int s_claim;
    const HsTapTable<TW + 2 * KMAX> tp(taps, NTAP);
    unsigned arrivals = 0;
    const int xcd = blockIdx.x % kXcds, slot = blockIdx.x / kXcds;
    const int per_xcd = gridDim.x / kXcds;  // the host launches a multiple of kXcds
    int base, cnt;
    xcd_share(b.total_tiles, xcd, base, cnt);

    // derivative setup of every tile
    for (int i = slot; i < cnt; i += per_xcd) setup_tile<NT, TW, TH>(b, st, base + i);
    arrivals += gridDim.x;
    if (!grid_sync(b.barrier, arrivals, b.status, b.spin_limit)) return;

    const int nph = (b.iters + b.k - 1) / b.k;
    for (int ph = 0; ph < nph; ++ph) {
        const int kk = min(b.k, b.iters - ph * b.k);
        unsigned *ctr = b.claims + (ph * kXcds + xcd) * kClaimStride;
        unsigned next = 0;
        if (threadIdx.x == 0) next = atomicAdd(ctr, 1u);
        for (;;) {
            // hs_tile ends with a barrier, so every thread has read the
            // previous claim before thread 0 overwrites it
            if (threadIdx.x == 0) s_claim = (int)min(next, (unsigned)cnt);
            __syncthreads();
            const int i = s_claim;
            if (i >= cnt) break;
            if (threadIdx.x == 0) next = atomicAdd(ctr, 1u);  // overlaps this tile
            const int w = base + i;
            const PairDesc &P = b.tiles[w];
            const HsImage im = phase_image(P, ph, b.phase0);
            // per-cell tap addresses: the persistent kernel's own state leaves
            // no registers for the shared ones
            hs_tile<NT, TW, TH, KMAX, NTAP, false, DB, FDIV, false, false>(im, w - P.tile_base, kk, tp, uv);
        }
        arrivals += gridDim.x;
        if (!grid_sync(b.barrier, arrivals, b.status, b.spin_limit)) return;
    }
}

// ---------------------------------------------------------------------------


// ---------------------------------------------------------------------------
// Stepped form: the same tile list, one ordinary launch per phase (plus one
// for the setup).  Kernel
// boundaries order the phases, so there is no grid barrier and no residency
// limit, and the tile body is the single-image kernel's full form (shared tap
// addresses, read-ahead, reciprocal quotient).  One work-group per tile; the
// grid is the whole list, dealt so each XCD gets a contiguous range.

// The setup's work-group size, whatever the phase kernel's: a tile's setup is
// a short streaming pass, and 1024-thread work-groups with two pixels per
// thread spent more time being dispatched than moving data (64 uneven pairs:
// 195 us; 256 threads with eight rows each read more per thread in flight).
constexpr int kSetupNT = 256;

template <int NT, int TW, int TH>
__global__ __launch_bounds__(NT) void hs_multi_setup_kernel(BatchArgs b, SetupTaps st) {
    const int w = xcd_remap(blockIdx.x, b.total_tiles);
    if (taps_in_cube(st.dx) && taps_in_cube(st.dy) && taps_in_cube(st.dt))
        setup_tile_cube<NT, TW, TH>(b, st, w);
    else if (st.dx.nd == 4 && st.dy.nd == 4 && st.dt.nd == 4)
        setup_tile<NT, TW, TH, true>(b, st, w);
    else
        setup_tile<NT, TW, TH>(b, st, w);
}

// At most 74 scalar registers: the reciprocal-quotient build compiled to 76,
// and at 76 the hardware kept one 1024-thread work-group per CU instead of
// two (half the waves in flight: 1.5x the cycles of the same body in
// hs_tb_kernel, which compiles to 72); the division build (74) kept two.
template <int NT, int TW, int TH, int KMAX, int NTAP, bool DB, bool FDIV>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_num_sgpr(74))) void hs_multi_phase_kernel(
    BatchArgs b, Taps taps, int ph, int kk) {
    using Win = HsWindow<NT, TW, TH, KMAX, DB>;
    __shared__ float2 uvm[Win::SIZE];
    const HsTapTable<TW + 2 * KMAX> tp(taps, NTAP);
    // phases alternate the walk over each XCD's share of the list (the setup
    // goes forward, phase 0 backward), so a phase starts on the tiles whose
    // data the previous launch left in that XCD's L2 and in the MALL
    const int w = xcd_remap(blockIdx.x, b.total_tiles, (ph & 1) == 0);
    const PairDesc &P = b.tiles[w];
    const HsImage im = phase_image_stepped(P, ph, b.nph, b.phase0);
    hs_tile<NT, TW, TH, KMAX, NTAP, false, DB, FDIV>(im, w - P.tile_base, kk, tp,
                                                      uvm + Win::PAD);
}

// ---------------------------------------------------------------------------
struct BVariant {
    int nt, tw, th, kmax, wpe;
    void (*kernel[2])(BatchArgs, SetupTaps, Taps);  // persistent: 8 taps, any count
    // stepped form (kernel[] null): setup, phase (8 taps, any count)
    void (*setup)(BatchArgs, SetupTaps);
    void (*phase[2])(BatchArgs, Taps, int, int);
};

#define HB_V(NT, TW, TH, K, WPE)                                                          \
    {NT, TW, TH, K, WPE,                                                                  \
     {hs_batched_kernel<NT, TW, TH, K, 8, WPE>, hs_batched_kernel<NT, TW, TH, K, 0, WPE>}, \
     nullptr, {nullptr, nullptr}}
#define HB_D(NT, TW, TH, K, WPE)                                                          \
    {NT, TW, TH, K, WPE,                                                                  \
     {hs_batched_kernel<NT, TW, TH, K, 8, WPE, true>,                                     \
      hs_batched_kernel<NT, TW, TH, K, 0, WPE, true>},                                    \
     nullptr, {nullptr, nullptr}}
#define HB_F(NT, TW, TH, K, WPE)                                                          \
    {NT, TW, TH, K, WPE,                                                                  \
     {hs_batched_kernel<NT, TW, TH, K, 8, WPE, true, true>,                               \
      hs_batched_kernel<NT, TW, TH, K, 0, WPE, true, true>},                              \
     nullptr, {nullptr, nullptr}}
// stepped form (DB: double-buffered window, FDIV: reciprocal quotient)
#define HB_S(NT, TW, TH, K, DB, FDIV)                                                     \
    {NT, TW, TH, K, 0, {nullptr, nullptr}, hs_multi_setup_kernel<kSetupNT, TW, TH>,      \
     {hs_multi_phase_kernel<NT, TW, TH, K, 8, DB, FDIV>,                                  \
      hs_multi_phase_kernel<NT, TW, TH, K, 0, DB, FDIV>}}
static const BVariant kBVariants[] = {
    HB_V(1024, 64, 32, 4, 8),   // 0: two work-groups per CU
    HB_V(1024, 64, 32, 8, 8),   // 1: two work-groups per CU
    HB_V(1024, 128, 32, 8, 4),  // 2: one
    HB_V(512, 64, 32, 4, 4),    // 3: two
    HB_V(1024, 64, 64, 8, 4),   // 4: one
    HB_D(1024, 64, 32, 4, 8),   // 5: double-buffered window, two per CU
    HB_D(1024, 64, 32, 8, 8),   // 6: double-buffered window, two per CU
    HB_F(1024, 64, 32, 4, 8),   // 7: 5 with the reciprocal quotient (k = 8 would spill)
    HB_S(1024, 64, 32, 8, true, true),   // 8: stepped, as single-image variant 24
    HB_S(1024, 64, 32, 4, true, true),   // 9: stepped, as variant 25
    HB_S(1024, 64, 32, 8, true, false),  // 10: stepped, as variant 17
    HB_S(512, 64, 32, 4, true, true),    // 11: stepped, as variant 26
};
#undef HB_V
#undef HB_D
#undef HB_F
#undef HB_S
static constexpr int kNumBVariants = sizeof(kBVariants) / sizeof(kBVariants[0]);

// A cached launch: device descriptors, barrier words and (optionally) the
// captured graph, keyed by everything the launch bakes in.
struct BatchPlan {
    std::vector<char> key;
    int device = 0;
    int variant = 0, k = 1, grid = 0, total_tiles = 0;
    PairDesc *d_pairs = nullptr;  // one pair descriptor per tile of the list
    unsigned *d_sync = nullptr;  // [0] barrier counter, [1] status, claims from [64]
    int64_t sync_words = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    float *d_z = nullptr;  // stepped form, odd phase count above 1: every pair's (zu, zv)
};

static std::mutex g_bmu;
static std::list<BatchPlan> g_plans;  // most recent first
static constexpr size_t kMaxPlans = 8;
static std::vector<std::pair<std::vector<int>, std::pair<int, int>>> g_btuned;

static void free_plan(BatchPlan &p) {
    if (p.exec) (void)hipGraphExecDestroy(p.exec);
    if (p.graph) (void)hipGraphDestroy(p.graph);
    if (p.d_pairs) (void)hipFree(p.d_pairs);
    if (p.d_sync) (void)hipFree(p.d_sync);
    if (p.d_z) (void)hipFree(p.d_z);
    p = BatchPlan();
}

template <typename T>
static void append(std::vector<char> &k, const T &v) {
    const char *c = reinterpret_cast<const char *>(&v);
    k.insert(k.end(), c, c + sizeof(T));
}

static int launch_batch(const BatchPlan &P, const SetupTaps &st, const Taps &av, int npairs,
                        int iters, long long spin_limit, hipStream_t s) {
    int rc = launch_fill_u32(P.d_sync, 0u, P.sync_words, s);
    if (rc) return rc;
    BatchArgs a;
    a.tiles = P.d_pairs;
    a.claims = P.d_sync + 64;
    a.barrier = P.d_sync;
    a.status = (int *)(P.d_sync + 1);
    a.spin_limit = spin_limit;
    a.npairs = npairs;
    a.total_tiles = P.total_tiles;
    a.iters = iters;
    a.k = P.k;
    const int nph = (iters + P.k - 1) / P.k;
    const BVariant &V = kBVariants[P.variant];
    a.nph = nph;
    a.phase0 = V.kernel[0] ? nph & 1 : nph == 1;
    if (V.kernel[0]) {
        hipLaunchKernelGGL(V.kernel[av.nd == 8 ? 0 : 1], dim3(P.grid), dim3(V.nt), 0, s, a, st,
                           av);
        HMX_LAUNCH_CHECK("hs_batched");
        return HMX_OK;
    }
    const dim3 grid(P.total_tiles), block(V.nt);
    hipLaunchKernelGGL(V.setup, grid, dim3(kSetupNT), 0, s, a, st);
    HMX_LAUNCH_CHECK("hs_multi_setup");
    for (int ph = 0; ph < nph; ++ph) {
        const int kk = std::min(P.k, iters - ph * P.k);
        hipLaunchKernelGGL(V.phase[av.nd == 8 ? 0 : 1], grid, block, 0, s, a, av, ph, kk);
        HMX_LAUNCH_CHECK("hs_multi_phase");
    }
    return HMX_OK;
}

// Work-groups of the variant that can be resident on one CU: the smaller of
// the runtime's occupancy figure and a bound from the kernel's own resources
// (the waves per SIMD it was compiled for, 512 VGPRs per lane per SIMD in
// granules of 8, the CU's LDS; one work-group if it uses scratch).  Every work-group of the grid must
// be resident at once, or the grid barrier waits until its spin limit.
static int resident_per_cu(int variant, int nd, int device, int &api, int &own, int &cus) {
    const BVariant &V = kBVariants[variant];
    const void *kern = V.kernel[0] ? reinterpret_cast<const void *>(V.kernel[nd == 8 ? 0 : 1])
                                   : reinterpret_cast<const void *>(V.phase[nd == 8 ? 0 : 1]);
    HMX_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&api, kern, V.nt, 0));
    if (!V.kernel[0]) {  // stepped form: no residency requirement
        HMX_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        own = api;
        return HMX_OK;
    }
    hipFuncAttributes fa;
    HMX_HIP_TRY(hipFuncGetAttributes(&fa, kern));
    int lds = 0;
    HMX_HIP_TRY(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                      device));
    HMX_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    const int waves_per_simd_per_block = ((V.nt + 63) / 64 + 3) / 4;
    const int vgpr = std::max(8, (fa.numRegs + 7) / 8 * 8);
    // the declared waves per SIMD bound every register kind (SGPRs included,
    // which the runtime's figure does not see); registers can only lower it
    const int waves_per_simd = std::min(V.wpe, 512 / vgpr);
    own = waves_per_simd / waves_per_simd_per_block;
    if (fa.sharedSizeBytes > 0 && lds > 0)
        own = std::min(own, (int)(lds / fa.sharedSizeBytes));
    if (fa.localSizeBytes > 0) own = std::min(own, 1);
    return HMX_OK;
}

static int resident_grid(int variant, int nd, int total_tiles, int device, int &grid) {
    if (!kBVariants[variant].kernel[0]) {  // stepped form: one work-group per tile
        grid = total_tiles;
        return HMX_OK;
    }
    int api = 0, own = 0, cus = 0;
    int rc = resident_per_cu(variant, nd, device, api, own, cus);
    if (rc) return rc;
    const int cap = std::min(api, own) * cus / kXcds * kXcds;  // a multiple of the XCD count
    if (cap < kXcds) return fail(HMX_ERR_UNSUPPORTED, "batched kernel cannot be resident");
    grid = std::min(cap, (total_tiles + kXcds - 1) / kXcds * kXcds);
    return HMX_OK;
}

}  // namespace hmx

using namespace hmx;

extern "C" int hmx_hs_batched_variant_count(void) { return kNumBVariants; }

extern "C" int hmx_hs_batched_resident(int variant, int ntaps, int *api_per_cu,
                                       int *own_per_cu, int *cus) {
    HMX_REQUIRE(variant >= 0 && variant < kNumBVariants, "unknown batched variant");
    int device = 0;
    HMX_HIP_TRY(hipGetDevice(&device));
    int a = 0, o = 0, c = 0;
    int rc = resident_per_cu(variant, ntaps, device, a, o, c);
    if (rc) return rc;
    if (api_per_cu) *api_per_cu = a;
    if (own_per_cu) *own_per_cu = o;
    if (cus) *cus = c;
    return HMX_OK;
}

extern "C" int hmx_hs_batched_variant_info(int variant, int *tile_w, int *tile_h, int *threads,
                                           int *kmax) {
    HMX_REQUIRE(variant >= 0 && variant < kNumBVariants, "unknown batched variant");
    const BVariant &V = kBVariants[variant];
    if (tile_w) *tile_w = V.tw;
    if (tile_h) *tile_h = V.th;
    if (threads) *threads = V.nt;
    if (kmax) *kmax = V.kmax;
    return HMX_OK;
}

// Device tables of a new plan: the pair descriptors followed by the tile ->
// pair table, and the barrier / status / claim words.  On failure the caller
// frees what was allocated (free_plan).
static int build_plan(BatchPlan &P, std::vector<PairDesc> desc, int ntaps, int iters) {
    const int npairs = (int)desc.size();
    const int tiles = P.total_tiles;
    int rc = resident_grid(P.variant, ntaps, tiles, P.device, P.grid);
    if (rc) return rc;
    const int64_t nph = (iters + P.k - 1) / P.k;
    if (!kBVariants[P.variant].kernel[0] && (nph & 1) && nph > 1) {
        // the stepped form's third plane pair (phase_image_stepped)
        int64_t total = 0;
        for (const PairDesc &d : desc) total += 2 * (int64_t)d.H * d.W;
        HMX_HIP_TRY(hipMalloc(&P.d_z, sizeof(float) * total));
        int64_t off = 0;
        for (PairDesc &d : desc) {
            const int64_t n = (int64_t)d.H * d.W;
            d.zu = P.d_z + off;
            d.zv = P.d_z + off + n;
            off += 2 * n;
        }
    }
    std::vector<PairDesc> tile_desc(tiles);
    for (int i = 0; i < npairs; ++i)
        std::fill(tile_desc.begin() + desc[i].tile_base,
                  tile_desc.begin() + desc[i].tile_base + desc[i].ntiles, desc[i]);
    HMX_HIP_TRY(hipMalloc(&P.d_pairs, sizeof(PairDesc) * tiles));
    P.sync_words = 64 + (nph + 1) * kXcds * kClaimStride;
    HMX_HIP_TRY(hipMalloc(&P.d_sync, sizeof(unsigned) * P.sync_words));
    HMX_HIP_TRY(hipMemcpy(P.d_pairs, tile_desc.data(), sizeof(PairDesc) * tiles,
                          hipMemcpyHostToDevice));
    return HMX_OK;
}

// Captures the plan's launch sequence (fill kernel, then the persistent kernel
// or the setup and phase launches) into a graph on a private stream.
static int capture_plan(BatchPlan &P, const SetupTaps &st, const Taps &av, int npairs, int iters,
                        long long spin_limit) {
    hipStream_t cs;
    HMX_HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(cs);
        return fail(HMX_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
    }
    const int rc = launch_batch(P, st, av, npairs, iters, spin_limit, cs);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : fail(HMX_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(HMX_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    P.graph = g;
    P.exec = x;
    return HMX_OK;
}

static int batched_impl(const hmx_pair *pairs, int npairs, const SetupTaps &st, const Taps &av,
                        int iters, int variant, int k, int flags, long long spin_limit,
                        hipStream_t s, hipEvent_t done = nullptr) {
    HMX_REQUIRE(pairs && npairs >= 1, "hmx_hs_batched: no pairs");
    HMX_REQUIRE(iters >= 0, "hmx_hs_batched: iters must be >= 0");
    HMX_REQUIRE(av.rx <= 1 && av.ry <= 1 && av.nd <= HS_MAX_TAPS,
                "hmx_hs_batched: averaging stencil must lie within radius 1 with <= 9 taps");
    if (variant < 0) {  // autotuned choice for this size signature, else default
        variant = 8;  // stepped, double-buffered, reciprocal quotient
        k = 8;
        std::vector<int> sig;
        for (int i = 0; i < npairs; ++i) {
            sig.push_back(pairs[i].H);
            sig.push_back(pairs[i].W);
        }
        std::lock_guard<std::mutex> lk(g_bmu);
        for (auto &e : g_btuned)
            if (e.first == sig) {
                variant = e.second.first;
                k = e.second.second;
            }
    }
    HMX_REQUIRE(variant < kNumBVariants, "hmx_hs_batched: unknown variant");
    HMX_REQUIRE(k >= 1 && k <= kBVariants[variant].kmax, "hmx_hs_batched: k out of range");
    const BVariant &V = kBVariants[variant];
    std::vector<PairDesc> desc(npairs);
    int tiles = 0;
    for (int i = 0; i < npairs; ++i) {
        const hmx_pair &p = pairs[i];
        HMX_REQUIRE(p.I1 && p.I2 && p.u && p.v && p.work, "hmx_hs_batched: null pair buffer");
        HMX_REQUIRE(p.H > 0 && p.W > 0 && p.H < 32768 && p.W < 32768,
                    "hmx_hs_batched: pair sizes must be in 1..32767");
        const int64_t n = (int64_t)p.H * p.W;
        PairDesc &d = desc[i];
        d.I1 = p.I1;
        d.I2 = p.I2;
        d.Ix = p.work;
        d.Iy = p.work + n;
        d.It = p.work + 2 * n;
        d.wu = p.work + 3 * n;
        d.wv = p.work + 4 * n;
        d.u = p.u;
        d.v = p.v;
        d.zu = p.u;  // (the plan's own planes where the phase count needs them)
        d.zv = p.v;
        d.alpha2 = p.alpha * p.alpha;
        d.H = p.H;
        d.W = p.W;
        d.tiles_x = (p.W + V.tw - 1) / V.tw;
        d.ntiles = d.tiles_x * ((p.H + V.th - 1) / V.th);
        d.tile_base = tiles;
        HMX_REQUIRE((int64_t)tiles + d.ntiles < (1 << 30), "hmx_hs_batched: too many tiles");
        tiles += d.ntiles;
    }
    int device = 0;
    HMX_HIP_TRY(hipGetDevice(&device));
    std::vector<char> key;
    append(key, device);
    append(key, variant);
    append(key, k);
    append(key, iters);
    append(key, spin_limit);
    append(key, st);
    append(key, av);
    append(key, (void *)s);
    key.insert(key.end(), (const char *)desc.data(),
               (const char *)desc.data() + sizeof(PairDesc) * desc.size());

    std::lock_guard<std::mutex> lk(g_bmu);
    BatchPlan *plan = nullptr;
    for (auto it = g_plans.begin(); it != g_plans.end(); ++it)
        if (it->key == key) {
            g_plans.splice(g_plans.begin(), g_plans, it);
            plan = &g_plans.front();
            break;
        }
    if (!plan) {
        BatchPlan P;
        P.key = key;
        P.device = device;
        P.variant = variant;
        P.k = k;
        P.total_tiles = tiles;
        int rc = build_plan(P, desc, av.nd, iters);
        if (rc) {
            free_plan(P);  // everything allocated before the failure
            return rc;
        }
        if (g_plans.size() >= kMaxPlans) {
            free_plan(g_plans.back());
            g_plans.pop_back();
        }
        g_plans.push_front(std::move(P));
        plan = &g_plans.front();
    }
    // A plan first made by an eager call (the tuner, a sync run) is captured
    // when a graph call first asks for it.
    if ((flags & HMX_BATCH_GRAPH) && !plan->exec && iters > 0) {
        int rc = capture_plan(*plan, st, av, npairs, iters, spin_limit);
        if (rc) return rc;
    }
    if (iters == 0) return HMX_OK;
    if ((flags & HMX_BATCH_GRAPH) && plan->exec) {
        HMX_HIP_TRY(hipGraphLaunch(plan->exec, s));
    } else {
        int rc = launch_batch(*plan, st, av, npairs, iters, spin_limit, s);
        if (rc) return rc;
    }
    if (done) HMX_HIP_TRY(hipEventRecord(done, s));
    // The persistent form's grid barrier can give up (a grid that is not
    // co-resident); its status is read back after every call, so a partial
    // result never goes unreported.  The stepped form has no barrier.
    if ((flags & HMX_BATCH_SYNC) || kBVariants[plan->variant].kernel[0]) {
        unsigned status = 0;
        HMX_HIP_TRY(hipStreamSynchronize(s));
        HMX_HIP_TRY(hipMemcpy(&status, plan->d_sync + 1, 4, hipMemcpyDeviceToHost));
        if (status) return fail(HMX_ERR_TIMEOUT, "hmx_hs_batched: a grid barrier wait gave up");
    }
    return HMX_OK;
}

static int make_batch_taps(SetupTaps &st, Taps &av, int ndx, const float *wx, const int *oxx,
                           const int *oyx, int ndy, const float *wy, const int *oxy,
                           const int *oyy, int ndt, const float *wt, const int *oxt,
                           const int *oyt, int nda, const float *wa, const int *oxa,
                           const int *oya) {
    int rc = make_taps(st.dx, ndx, wx, oxx, oyx);
    if (!rc) rc = make_taps(st.dy, ndy, wy, oxy, oyy);
    if (!rc) rc = make_taps(st.dt, ndt, wt, oxt, oyt);
    if (!rc) rc = make_taps(av, nda, wa, oxa, oya);
    return rc;
}

extern "C" int hmx_hs_batched(const hmx_pair *pairs, int npairs, int ndx, const float *wx,
                              const int *oxx, const int *oyx, int ndy, const float *wy,
                              const int *oxy, const int *oyy, int ndt, const float *wt,
                              const int *oxt, const int *oyt, int nda, const float *wa,
                              const int *oxa, const int *oya, int iters, int variant, int k,
                              int flags, long long spin_limit, void *stream) {
    SetupTaps st;
    Taps av;
    int rc = make_batch_taps(st, av, ndx, wx, oxx, oyx, ndy, wy, oxy, oyy, ndt, wt, oxt, oyt,
                             nda, wa, oxa, oya);
    if (rc) return rc;
    if (spin_limit <= 0) spin_limit = kDefaultSpinLimit;
    return batched_impl(pairs, npairs, st, av, iters, variant, k, flags, spin_limit,
                        as_stream(stream));
}

extern "C" int hmx_hs_batched_tune(const hmx_pair *pairs, int npairs, int ndx, const float *wx,
                                   const int *oxx, const int *oyx, int ndy, const float *wy,
                                   const int *oxy, const int *oyy, int ndt, const float *wt,
                                   const int *oxt, const int *oyt, int nda, const float *wa,
                                   const int *oxa, const int *oya, int iters, void *stream,
                                   int *variant_out, int *k_out) {
    SetupTaps st;
    Taps av;
    int rc = make_batch_taps(st, av, ndx, wx, oxx, oyx, ndy, wy, oxy, oyy, ndt, wt, oxt, oyt,
                             nda, wa, oxa, oya);
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    hipEvent_t e0, e1;
    HMX_HIP_TRY(hipEventCreate(&e0));
    HMX_HIP_TRY(hipEventCreate(&e1));
    float best = 1e30f;
    int bv = 8, bk = 8;
    for (int vi = 0; vi < kNumBVariants && !rc; ++vi) {
        for (int k = 2; k <= kBVariants[vi].kmax && !rc; k *= 2) {
            rc = batched_impl(pairs, npairs, st, av, iters, vi, k, HMX_BATCH_SYNC,
                              kDefaultSpinLimit, s);
            if (rc == HMX_ERR_TIMEOUT) {  // this grid did not stay resident: skip it
                rc = HMX_OK;
                continue;
            }
            if (rc) break;
            // e1 is recorded right after the launches, before the status
            // read-back of the persistent form; a timed run whose barrier
            // gave up is discarded like a warm-up that did
            HMX_HIP_TRY(hipEventRecord(e0, s));
            rc = batched_impl(pairs, npairs, st, av, iters, vi, k, 0, kDefaultSpinLimit, s, e1);
            if (rc == HMX_ERR_TIMEOUT) {
                rc = HMX_OK;
                continue;
            }
            if (rc) break;
            HMX_HIP_TRY(hipEventSynchronize(e1));
            float ms = 0;
            HMX_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) {
                best = ms;
                bv = vi;
                bk = k;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    std::vector<int> sig;
    for (int i = 0; i < npairs; ++i) {
        sig.push_back(pairs[i].H);
        sig.push_back(pairs[i].W);
    }
    {
        std::lock_guard<std::mutex> lk(g_bmu);
        g_btuned.push_back({sig, {bv, bk}});
    }
    if (variant_out) *variant_out = bv;
    if (k_out) *k_out = bk;
    return HMX_OK;
}

extern "C" int hmx_hs_batched_release(void) {
    std::lock_guard<std::mutex> lk(g_bmu);
    for (auto &p : g_plans) free_plan(p);
    g_plans.clear();
    return HMX_OK;
}
