/*
 * hmx.h -- C ABI of libhmx, the MI355X (gfx950) backend for Hindemith's
 * Horn-Schunck hot path.
 *
 * Two groups of entry points:
 *
 *  1. The record API replaces the reference's Python-2 extension module
 *     `backend` (backend.cpp) and everything under it (sm.cpp, oclgen.cpp,
 *     fusion.cpp, lm.cpp, execute.cpp, clhelp.cpp).  It keeps the reference's
 *     XML wire format so the Python front end stays a faithful port.
 *  2. Device-pointer entry points: one per hot-path kernel, for torch callers
 *     and benchmarks.  Pointers are device pointers unless stated; `stream` is
 *     a hipStream_t (NULL = legacy default stream).  Nothing allocates device
 *     memory behind the caller's back except where a call says so.
 *
 * Every function returns HMX_OK (0) or an HMX_ERR_* code; hmx_last_error()
 * gives the message (thread-local).  Plain C types only.
 */
#ifndef HMX_H
#define HMX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HMX_MAX_TAPS 32

enum hmx_status {
    HMX_OK = 0,
    HMX_ERR_INVALID = 1,     /* bad argument, shape, dtype or program */
    HMX_ERR_HIP = 2,         /* HIP runtime error */
    HMX_ERR_UNSUPPORTED = 3, /* valid for the reference, not built here */
    HMX_ERR_NOT_FOUND = 4,   /* unknown record / symbol / result */
    HMX_ERR_STATE = 5,       /* call out of order (e.g. run before compile) */
    HMX_ERR_TIMEOUT = 6      /* a bounded device-side wait gave up */
};

/* dtypes, as Types/primitives.cpp:4-44 names them */
enum hmx_dtype { HMX_F32 = 0, HMX_F64 = 1, HMX_I64 = 2, HMX_I32 = 3 };

/* symbol kinds (sm.h:17-263 Variable subclasses on the hot path) */
enum hmx_kind {
    HMX_KIND_ARRAY2D = 1,
    HMX_KIND_ARRAY = 2,
    HMX_KIND_SCALAR = 3,
    HMX_KIND_STENCIL = 4,
    HMX_KIND_SPMAT2D = 5,
    HMX_KIND_CSRMAT = 6,
    HMX_KIND_ARRAY3D = 7, /* only as the HSMatrixPowers result (dim0 = STEPS planes) */
    HMX_KIND_SPMAT = 8    /* 1-D DIA matrix (scipy dia_matrix): data, offsets */
};

const char *hmx_last_error(void);
const char *hmx_version(void);

/* ====================================================================== */
/* 1. Record API                                                           */
/* ====================================================================== */

/* Select the device for records.  Replaces initbackend -> initialize_ocl
 * (backend.cpp:669-675, clhelp.cpp:33-140); the HM_* environment mapping is
 * done by the Python front end.  device = -1 selects plan-only mode: records
 * compile and report their plan (hmx_plan_text) but cannot run; no HIP call
 * is made, so it works on hosts without a GPU. */
int hmx_init(int device);

/* Parse symbol table, semantic model and constants; plan kernels; allocate
 * device buffers.  Replaces backend.xmlcompile (backend.cpp:70-159). */
int hmx_compile(int rec, const char *symtab_xml, const char *sm_xml,
                const char *defines_xml);

/* Point a Header symbol at caller memory for the next hmx_run.  component:
 * 0 = data (Array/Array2D/Scalar/Stencil/SpMat2D/CSRMat), 1 = offx (Stencil,
 * SpMat2D) or indices (CSRMat), 2 = offy or indptr.  nbytes must equal the
 * symbol's size.  Replaces the binding loop of backend.run
 * (backend.cpp:384-541); the caller keeps the memory alive until hmx_run
 * returns. */
int hmx_bind(int rec, const char *name, int component, const void *host, int64_t nbytes);

/* Upload Header symbols, execute the plan, download Return values (then
 * synchronise).  Replaces backend.run (backend.cpp:320-638). */
int hmx_run(int rec);

int hmx_result_count(int rec, int *n);
/* shape has room for 3 entries: ndim 1 or 2 -> (dim0, dim1, col_mjr flag);
 * ndim 3 -> (dim0, dim1, dim2), C order.  *name stays valid until hmx_release. */
int hmx_result_info(int rec, int i, int *kind, int *dtype, int *ndim, int64_t *shape,
                    const char **name);
/* Copy return value i of the last run into host memory (createPyObject,
 * backend.cpp:160-313, which instead wrapped backend memory). */
int hmx_result_read(int rec, int i, void *host, int64_t nbytes);
/* Human-readable plan: one line per launch group. */
int hmx_plan_text(int rec, char *buf, int64_t cap);
/* HIP source of the kernel generated for each fused group of the record (the
 * reference's generated OpenCL, oclgen.cpp:95-154, which its compile log
 * showed); compile != 0 also builds each with hipRTC for gfx950 (no GPU
 * needed) and fails with the compiler's log if one does not build. */
int hmx_program_source(int rec, int compile, char *buf, int64_t cap);
/* Free a record's device buffers and graphs (free_ocl, backend.cpp:646-651,
 * was a no-op). */
int hmx_release(int rec);

/* ====================================================================== */
/* 2. Device-pointer entry points                                          */
/* ====================================================================== */

/* a1. y = S * x with clamped (replicated) borders.  Stencil2DMVM.cpp:36-58.
 *     x, y: H x W row-major fp32 (distinct buffers).  w/offx/offy: host
 *     arrays of nd taps, accumulated in tap order. */
int hmx_stencil2d(const float *x, float *y, int H, int W, int nd, const float *w,
                  const int *offx, const int *offy, void *stream);

/* a2. y = A * x, 2D DIA with the coefficient stored at the neighbour index
 *     and out-of-range terms skipped.  SpmvDIA2D.cpp:35-55, SpMat2D.cpp:211-216.
 *     A: nd x H x W fp32 (device); offx/offy: host int arrays of nd. */
int hmx_dia2d_spmv(const float *A, int nd, const int *offx, const int *offy,
                   const float *x, float *y, int H, int W, void *stream);

/* a3. y = A * x, CSR with int32 indices.  SpmvCSR.cpp:32-77.
 *     lanes_per_row: 0 = choose from nnz/M, else 1,2,4,...,64. */
int hmx_csr_spmv(int M, int N, int64_t nnz, const int *indptr, const int *indices,
                 const float *data, const float *x, float *y, int lanes_per_row,
                 void *stream);

/* Row-binned CSR SpMV for irregular (e.g. power-law) row lengths.  bins is a
 * device int array of M row ids sorted into 4 length classes whose counts are
 * given by bin_counts (host, 4 entries): rows with <=4, <=16, <=64 and >64
 * entries.  Build it once per matrix with hmx_csr_bin_rows. */
int hmx_csr_bin_rows(int M, const int *indptr_host, int *bins_host, int *bin_counts);
int hmx_csr_spmv_binned(int M, const int *indptr, const int *indices, const float *data,
                        const float *x, float *y, const int *bins, const int *bin_counts,
                        void *stream);

/* Row-block CSR SpMV for irregular row lengths (same product as SpmvCSR).
 * hmx_csr_row_blocks cuts the rows, on the host, into blocks of consecutive
 * rows with at most 2048 entries and 256 rows (a longer row alone):
 * blocks_host (capacity M + 1) receives the nblocks + 1 block start rows.
 * hmx_csr_spmv_rowblocks runs one work-group per block (blocks: the same
 * array on the device); entries are streamed coalesced and each row's
 * products are added in a fixed order (error within the multi-lane bound). */
int hmx_csr_row_blocks(int M, const int *indptr_host, int *blocks_host, int *nblocks);
int hmx_csr_spmv_rowblocks(int M, const int *indptr, const int *indices, const float *data,
                           const float *x, float *y, const int *blocks, int nblocks,
                           void *stream);

/* Column-sliced CSR SpMV for irregular rows with columns spread over all of x
 * (same product as SpmvCSR, SpmvCSR.cpp:30-47; its sums regrouped).  The
 * columns are cut into hmx_csr_slice_count() = 8 equal ranges, one per XCD of
 * the MI355X, and each range's entries form their own CSR over all M rows,
 * so the gathers of x one XCD makes stay within a range that fits its L2.
 * hmx_csr_slice_build (host, once per sparsity pattern): sptr_host
 * (8 * (M + 1) ints) receives each slice's row offsets into the sliced entry
 * arrays, sidx_host (nnz ints) the column indices in slice order, perm_host
 * (nnz ints) the original position of each sliced entry; a row keeps its
 * entry order inside each slice.  hmx_csr_slice_values (device, once per
 * matrix values): sdata[k] = data[perm[k]].  hmx_csr_spmv_sliced: y = A * x;
 * partial is device scratch of 8 * M floats; each row's products are added
 * in entry order within a slice and the 8 slice sums in slice order
 * (deterministic; error within the multi-lane bound). */
int hmx_csr_slice_count(void);
int hmx_csr_slice_build(int M, int N, const int *indptr_host, const int *indices_host,
                        int *sptr_host, int *sidx_host, int *perm_host);
int hmx_csr_slice_values(int64_t nnz, const int *perm, const float *data, float *sdata,
                         void *stream);
int hmx_csr_spmv_sliced(int M, const int *sptr, const int *sidx, const float *sdata,
                        const float *x, float *partial, float *y, void *stream);

/* a7. out[0] = sum of x[0..n) (sum2d, Array2DSum.cpp:12-250).  Deterministic
 *     two-pass reduction with fp64 accumulation.  partials: device scratch of
 *     HMX_SUM_PARTIALS doubles.  out: device float. */
#define HMX_SUM_PARTIALS 1024
int hmx_sum2d(const float *x, int64_t n, double *partials, float *out, void *stream);

/* Allocation-free form of hmx_sum2d.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The pass-1
 * partials live in a file-scope device array of HMX_SUM_PARTIALS doubles
 * (8 KiB per GPU, held until the process ends); its device address is taken
 * once per device, on the first call with that device current, so make that
 * first call outside graph capture.  The partials grid is fixed, so every n
 * fits the array.  All calls on one GPU share its array: calls that overlap
 * on one GPU (other streams or threads) race.  Same result as hmx_sum2d, bit
 * for bit.  out: device float. */
int hmx_sum2d_noalloc(const float *x, int64_t n, float *out, void *stream);

/* a4-a8. Fused elementwise program over n elements (Array2DOp,
 * Array2DScalarOp, ScalarArray2DOp, Array2DCopy, Array2DSqrt and stencil /
 * DIA taps of Stencil2DMVM / SpmvDIA2D).  The record runtime encodes its
 * fused programs with these instructions (registers live in LDS). */
struct hmx_instr {
    int32_t op;   /* HMX_OP_* */
    int32_t dst;  /* register */
    int32_t a;    /* register or slot */
    int32_t b;    /* register or slot */
};
enum hmx_op {
    HMX_OP_LOAD = 1,    /* r[dst] = array[a][i] */
    HMX_OP_LOADS = 2,   /* r[dst] = scalar[a] */
    HMX_OP_STENCIL = 3, /* r[dst] = stencil[b] applied to array[a] (clamp) */
    HMX_OP_DIA = 4,     /* r[dst] = spmat2d[b] applied to array[a] (skip) */
    HMX_OP_ADD = 5,
    HMX_OP_SUB = 6,
    HMX_OP_MUL = 7,
    HMX_OP_DIV = 8,
    HMX_OP_SQRT = 9,    /* r[dst] = sqrt(r[a]) */
    HMX_OP_STORE = 10,  /* out[dst][i] = r[a] */
    HMX_OP_REDUCE = 11, /* reduction slot dst += r[a] */
    HMX_OP_RED = 12,    /* r[dst] = (x + y) odd ? 0 : r[a]   (RedArray2D.cpp:20-24) */
    HMX_OP_BLACK = 13   /* r[dst] = (x + y) odd ? r[a] : 0   (BlackArray2D.cpp:20-24) */
};

/* a10. Canonical HS setup: Ix = Dx*I1 + Dx*I2, Iy = Dy*I1 + Dy*I2,
 *      It = Dt*I2 - Dt*I1 (clamped stencils; one rounding per op). */
int hmx_hs_setup(const float *I1, const float *I2, int H, int W,
                 int ndx, const float *wx, const int *oxx, const int *oyx,
                 int ndy, const float *wy, const int *oxy, const int *oyy,
                 int ndt, const float *wt, const int *oxt, const int *oyt,
                 float *Ix, float *Iy, float *It, void *stream);

/* a9. `iters` fused Jacobi iterations of the canonical loop body
 *        ubar = Avg*u; vbar = Avg*v; t = (Ix*ubar + Iy*vbar + It)/den;
 *        u = ubar - Ix*t; v = vbar - Iy*t
 *     with den = alpha2 + Ix*Ix + Iy*Iy recomputed in registers when den is
 *     NULL (else read).  u, v are updated in place; work_u/work_v are H*W
 *     device scratch planes (ping-pong).  Avg taps must lie within radius 1.
 *     Kernel variant and iterations per launch come from the autotuner cache
 *     (hmx_hs_tune) or a size heuristic. */
int hmx_hs_jacobi(const float *Ix, const float *Iy, const float *It, const float *den,
                  float alpha2, int H, int W, int nda, const float *wa, const int *oxa,
                  const int *oya, int iters, float *u, float *v, float *work_u,
                  float *work_v, void *stream);

/* Same, with an explicit kernel variant (0..hmx_hs_variant_count()-1) and
 * iterations per launch k (1..max for that variant). */
int hmx_hs_jacobi_cfg(const float *Ix, const float *Iy, const float *It, const float *den,
                      float alpha2, int H, int W, int nda, const float *wa, const int *oxa,
                      const int *oya, int iters, float *u, float *v, float *work_u,
                      float *work_v, int variant, int k, void *stream);
int hmx_hs_variant_count(void);
/* Tile width/height, threads and max k of a variant. */
int hmx_hs_variant_info(int variant, int *tile_w, int *tile_h, int *threads, int *kmax);
/* Rows per thread per column of a column-strip variant (hs_strip_kernel), 0 for a
 * rectangular-tile variant (hs_tb_kernel), or a negative HMX_ERR_* code. */
int hmx_hs_variant_strip(int variant);
/* Self-check of the per-pixel reciprocal the fused kernel's reciprocal
 * quotient starts from (on the device: the hardware reciprocal and one Newton
 * step): every b = +-(1.m) * 2^e, e_lo <= e <= e_hi, is compared with the
 * correctly rounded 1/b (NaN outside [2^-24, 2^24]), and the number of
 * mismatches is added to *bad (a device counter the caller zeroes).  Enqueued
 * on stream.  No counterpart in the reference: it pins the bit-exactness of
 * the quotient the reference computes with a division. */
int hmx_hs_recip_check(int e_lo, int e_hi, unsigned long long *bad, void *stream);
/* The (variant, k) hmx_hs_jacobi and hmx_hs_solve use for an H x W image:
 * the tuned one if hmx_hs_tune ran for that size in this process, else the
 * default.  No counterpart in the reference. */
int hmx_hs_config(int H, int W, int *variant, int *k);
/* Time every (variant, k) candidate on the given buffers (their contents are
 * overwritten) and cache the best for this H x W; returns the choice.  The
 * default configuration is kept unless a candidate is more than 2% faster. */
int hmx_hs_tune(const float *Ix, const float *Iy, const float *It, float alpha2, int H, int W,
                int nda, const float *wa, const int *oxa, const int *oya, float *u, float *v,
                float *work_u, float *work_v, void *stream, int *variant, int *k);

/* a9+a10. Whole solve: setup, then iters Jacobi iterations from the (u, v)
 *     passed in.  workspace: 5*H*W device floats (Ix, Iy, It, 2 ping-pong). */
int hmx_hs_solve(const float *I1, const float *I2, int H, int W, float alpha2,
                 int ndx, const float *wx, const int *oxx, const int *oyx,
                 int ndy, const float *wy, const int *oxy, const int *oyy,
                 int ndt, const float *wt, const int *oxt, const int *oyt,
                 int nda, const float *wa, const int *oxa, const int *oya,
                 int iters, float *u, float *v, float *workspace, void *stream);

/* ---------------------------------------------------------------------- */
/* a13. Adjoint (no counterpart in the reference, which has no gradients).  */
/* B equally sized pairs are stacked as B x H x W planes; alpha is a device  */
/* array of B floats (den = alpha^2 + Ix^2 + Iy^2 per pair).                 */

/* Number of checkpoints ceil(iters / ckpt_every). */
int hmx_hs_ckpt_count(int iters, int ckpt_every);
/* Workspace bytes for hmx_hs_forward_ckpt / hmx_hs_adjoint. */
int64_t hmx_hs_forward_workspace(int B, int H, int W);
int64_t hmx_hs_adjoint_workspace(int B, int H, int W, int ckpt_every);

/* Forward solve that stores the state (u, v) every ckpt_every iterations:
 * ckpt holds hmx_hs_ckpt_count() x 2 x B x H x W floats (state before
 * iterations 0, C, 2C, ...).  u, v: initial state in, final state out.
 * lin_last (may be null): receives the last checkpoint segment's
 * linearisation, 2 x L x B x H x W floats for its L iterations (iteration j's
 * ubar plane at 2j, vbar at 2j + 1), which lets hmx_hs_adjoint skip that
 * segment's recompute.  workspace keeps the derivative planes Ix, Iy, It (its
 * first 3 x B x H x W floats), which hmx_hs_adjoint may read instead of
 * recomputing them. */
int hmx_hs_forward_ckpt(const float *I1, const float *I2, int B, int H, int W,
                        const float *alpha,
                        int ndx, const float *wx, const int *oxx, const int *oyx,
                        int ndy, const float *wy, const int *oxy, const int *oyy,
                        int ndt, const float *wt, const int *oxt, const int *oyt,
                        int nda, const float *wa, const int *oxa, const int *oya,
                        int iters, int ckpt_every, float *u, float *v, float *ckpt,
                        float *lin_last, void *workspace, void *stream);

/* Reverse pass: given dL/du_N, dL/dv_N, writes dL/dI1, dL/dI2 (B x H x W),
 * dL/dalpha (B, device) and, if non-null, dL/du_0, dL/dv_0.  Recomputes one
 * checkpoint segment at a time; averaging stencil within radius 1.  Optional
 * (null: recomputed here), from the hmx_hs_forward_ckpt call that wrote ckpt:
 * lin_last, its last segment's linearisation, and fwd_workspace, its
 * workspace (derivative planes).  The gradients are the same bits either way. */
int hmx_hs_adjoint(const float *I1, const float *I2, int B, int H, int W, const float *alpha,
                   int ndx, const float *wx, const int *oxx, const int *oyx,
                   int ndy, const float *wy, const int *oxy, const int *oyy,
                   int ndt, const float *wt, const int *oxt, const int *oyt,
                   int nda, const float *wa, const int *oxa, const int *oya,
                   int iters, int ckpt_every, const float *ckpt, const float *lin_last,
                   const void *fwd_workspace, const float *gu_N, const float *gv_N,
                   float *gI1, float *gI2, float *galpha, float *gu0, float *gv0,
                   void *workspace, void *stream);

/* fp64 forward solve (plain per-iteration kernels) of B pairs, for
 * finite-difference checks.  alpha: device doubles; workspace 5*B*H*W doubles. */
int hmx_hs_solve_f64(const double *I1, const double *I2, int B, int H, int W,
                     const double *alpha,
                     int ndx, const float *wx, const int *oxx, const int *oyx,
                     int ndy, const float *wy, const int *oxy, const int *oyy,
                     int ndt, const float *wt, const int *oxt, const int *oyt,
                     int nda, const float *wa, const int *oxa, const int *oya,
                     int iters, double *u, double *v, double *workspace, void *stream);

/* ---------------------------------------------------------------------- */
/* a14. Batched solve of many pairs of uneven sizes (no counterpart in the  */
/* reference).                                                              */

/* One pair: device frames I1, I2 (H x W), initial/final flow u, v, and a
 * work area of 5*H*W floats (Ix, Iy, It and the ping-pong planes). */
struct hmx_pair {
    const float *I1;
    const float *I2;
    float *u;
    float *v;
    float *work;
    int H;
    int W;
    float alpha; /* den = alpha^2 + Ix^2 + Iy^2 */
};

#define HMX_BATCH_GRAPH 1 /* capture the launch in a hipGraph, replay on later calls */
#define HMX_BATCH_SYNC 2  /* synchronise and report a barrier time-out */

/* Setup + iters Jacobi iterations of every pair over one list of all pairs'
 * tiles, in phases of k iterations.  Two forms:
 *   stepped (variants 8-11, the default): one launch per phase over every
 *     tile, plus a setup launch.  Phases alternate (u, v) with the work
 *     planes; an odd number of phases above one also uses a plane pair the
 *     launch state owns (2 x H x W floats per pair), so the last phase writes
 *     (u, v) and nothing is copied; a single phase reads a copy of (u, v)
 *     the setup makes in the work planes;
 *   persistent (variants 0-7): ONE launch whose grid, no larger than the GPU
 *     holds at once, walks the list with a grid-wide barrier between phases;
 *     each of the 8 XCDs owns a contiguous share of the list and its
 *     work-groups claim tiles from it through a per-phase counter.  Barrier
 *     waits give up after spin_limit polls (<= 0: 2^24, about 3 s) and then
 *     every work-group exits.  A persistent call always synchronises its
 *     stream and returns HMX_ERR_TIMEOUT in that case (so a partial result
 *     is never silent); HMX_BATCH_SYNC makes a stepped call synchronise too.
 * variant = -1 uses the autotuned (or default) variant and k.
 * Launch state (descriptors, barrier words, graph, the stepped form's extra
 * plane pairs) is cached per signature, up to 8 signatures;
 * a graph call on a signature first run eagerly captures the graph then.
 * Calls that overlap on one stream must not share pairs. */
int hmx_hs_batched(const struct hmx_pair *pairs, int npairs,
                   int ndx, const float *wx, const int *oxx, const int *oyx,
                   int ndy, const float *wy, const int *oxy, const int *oyy,
                   int ndt, const float *wt, const int *oxt, const int *oyt,
                   int nda, const float *wa, const int *oxa, const int *oya,
                   int iters, int variant, int k, int flags, long long spin_limit,
                   void *stream);
/* Time every (variant, k) on these pairs (their u, v are overwritten) and
 * cache the best for this list of sizes. */
int hmx_hs_batched_tune(const struct hmx_pair *pairs, int npairs,
                        int ndx, const float *wx, const int *oxx, const int *oyx,
                        int ndy, const float *wy, const int *oxy, const int *oyy,
                        int ndt, const float *wt, const int *oxt, const int *oyt,
                        int nda, const float *wa, const int *oxa, const int *oya,
                        int iters, void *stream, int *variant, int *k);
int hmx_hs_batched_variant_count(void);
int hmx_hs_batched_variant_info(int variant, int *tile_w, int *tile_h, int *threads, int *kmax);
/* Work-groups of a batched variant (averaging stencil of ntaps taps) resident per CU:
 * the runtime's occupancy figure, the bound from the kernel's registers and LDS, and
 * the CU count.  The persistent grid is min(api, own) x cus rounded down to a
 * multiple of 8 (the XCD count); the stepped variants report own = api (they have no
 * residency requirement). */
int hmx_hs_batched_resident(int variant, int ntaps, int *api_per_cu, int *own_per_cu,
                            int *cus);
/* Free every cached batched launch (descriptors, graphs). */
int hmx_hs_batched_release(void);

/* ---------------------------------------------------------------------------
 * Preconditioned-CG Horn-Schunck operations (device pointers, fp32, H x W
 * row-major planes, the caller's stream).
 */

/* a15. (y0, y1) = M^-1 (x0, x1) per pixel, M = [[a, b], [b, d]] with
 * a = diag0 + n*w[2], d = diag1 + n*w[2], b = offdiag, n the number of in-image
 * 4-neighbours; where |det M| <= 1e-15, y = x.  w: the host weights of the
 * stencil passed to applyHSPreconditioner (nd >= 3).
 * Replaces ApplyHSPreconditioner (ApplyHSPreconditioner.cpp:46-97). */
int hmx_hs_precondition(const float *diag0, const float *diag1, const float *offdiag,
                        const float *x0, const float *x1, float *y0, float *y1, int H, int W,
                        int nd, const float *w, void *stream);
/* a16. V0, V1 (steps planes of H x W each): V[0] = (r0, r1), V[s] = A V[s-1],
 * A the HS operator Ix(Ix u + Iy v) + lam*L u (and the v twin) with L the
 * 4-neighbour Neumann Laplacian.  Replaces HSMatrixPowers
 * (HSMatrixPowers.cpp:28-39, 72-126, 191-226). */
int hmx_hs_matrix_powers(float lam, const float *Ix, const float *Iy, const float *r0,
                         const float *r1, float *V0, float *V1, int H, int W, int steps,
                         void *stream);

/* a19. Brox SOR building blocks (device pointers, fp32, H x W planes).
 * hmx_brox_set_matrix: planes 0..4 of A (offsets (0,-1), (-1,0), (0,0), (1,0),
 * (0,1)) get the psi-weighted 4-neighbour Laplacian, each coefficient at the
 * neighbour's index; replaces SetBroxMatrix (SetBroxMatrix.cpp:34-79).
 * hmx_brox_set_redblack_matrix: the off-diagonals into planes 0..3 (top, left,
 * right, bottom) and the diagonal into Adiag; replaces SetBroxRedblackMatrix
 * (SetBroxRedblackMatrix.cpp:37-81).
 * hmx_brox_precondition: per-pixel 2x2 solve with a = diag0 + A[2], d = diag1 +
 * A[2], b = offdiag, x passed through where |det| <= 1e-4; replaces
 * ApplyBroxPreconditioner (ApplyBroxPreconditioner.cpp:46-72).  nd: A's plane
 * count. */
int hmx_brox_set_matrix(float *A, int nd, const float *psi, float alpha, int H, int W,
                        void *stream);
int hmx_brox_set_redblack_matrix(float *Adiag, float *A, int nd, const float *psi, float alpha,
                                 int H, int W, void *stream);
int hmx_brox_precondition(const float *A, int nd, const float *diag0, const float *diag1,
                          const float *offdiag, const float *x0, const float *x1, float *y0,
                          float *y1, int H, int W, void *stream);

/* 1-D DIA SpMV (scipy dia_matrix, square n x n, data nd x n):
 * y[i] = sum_d [0 <= i+off_d < n] A[d][i+off_d] * x[i+off_d], fma'd in
 * diagonal order from 0.  Replaces SpmvDIA (SpmvDIA.cpp:30-50). */
int hmx_dia_spmv(const float *A, int nd, const int *offsets, const float *x, float *y, int n,
                 void *stream);

/* a17. dst = src bilinearly sampled at (x + offx, y + offy) where all four taps
 * are in the image, else src (x, y); offsets split by truncation, one step down
 * when negative.  Offsets must be finite with magnitude below 2^31.
 * Replaces WarpImg2D (WarpImg2D.cpp:33-75). */
int hmx_warp_img2d(const float *src, const float *offx, const float *offy, float *dst, int H,
                   int W, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* HMX_H */
