/*
 * hm_oracle.c -- CPU restatement of Hindemith's Horn-Schunck hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the
 * smoke entry point and the benchmark's cpu_baseline leg load it.  The
 * product library (libhmx.so) never links, calls or falls back to it.
 *
 * Every function restates a kernel fragment of the reference (chick/hindemith)
 * and cites the file:line it follows.  Arithmetic conventions:
 *
 *  - fp32 functions reproduce the generated OpenCL's rounding order: every
 *    Python binary op of the user program is its own statement (one rounding,
 *    no contraction across statements), while a stencil / DIA / CSR
 *    accumulation "y += w * x" is a single OpenCL expression that the vendor
 *    compiler may contract, restated here as fmaf(w, x, y) in tap order,
 *    starting from y = 0.0.  Built with -ffp-contract=off so the compiler adds
 *    no contraction of its own.
 *  - fp64 functions run the same operation order in double; they are the
 *    "exact" reference for endpoint-error and gradient checks.
 *
 * Parity pins (see tests/test_oracle_*.py): scipy.ndimage.correlate
 * (mode='nearest'), scipy.sparse products, closed forms (constant field,
 * linear ramp, 1x1 image), the HS fixed-point / zero / linearity identities,
 * brute-force Python loops on tiny inputs, transpose identities and central
 * finite differences for the gradient.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------- */
/* a1  Stencil2DMVM: y[p] = sum_d w[d] * x[clampX(p.x+offx[d]), clampY(p.y+offy[d])]
 *     Stencil2DMVM.cpp:36-47 (y = 0.0; for diag: clamp, y += D*x).  Row-major
 *     H x W with element_id0 = column (Array2D.cpp:297-310).                  */
void or_stencil2d_f32(const float *x, int H, int W, int nd, const float *w,
                      const int *ox, const int *oy, float *y)
{
    #pragma omp parallel for schedule(static)
    for (int r = 0; r < H; ++r) {
        for (int c = 0; c < W; ++c) {
            float acc = 0.0f;
            for (int d = 0; d < nd; ++d) {
                int xx = clampi(c + ox[d], 0, W - 1);
                int yy = clampi(r + oy[d], 0, H - 1);
                acc = fmaf(w[d], x[(int64_t)yy * W + xx], acc);
            }
            y[(int64_t)r * W + c] = acc;
        }
    }
}

void or_stencil2d_f64(const double *x, int H, int W, int nd, const double *w,
                      const int *ox, const int *oy, double *y)
{
    #pragma omp parallel for schedule(static)
    for (int r = 0; r < H; ++r) {
        for (int c = 0; c < W; ++c) {
            double acc = 0.0;
            for (int d = 0; d < nd; ++d) {
                int xx = clampi(c + ox[d], 0, W - 1);
                int yy = clampi(r + oy[d], 0, H - 1);
                acc += w[d] * x[(int64_t)yy * W + xx];
            }
            y[(int64_t)r * W + c] = acc;
        }
    }
}

/* Transpose of the clamped stencil, written as the literal scatter
 * out[clamp(p + o_d)] += w_d * g[p] (serial, fp64).  Used by the gradient
 * oracle; the GPU adjoint uses a gather form, so the two are independent.  */
void or_stencil2d_T_f64(const double *g, int H, int W, int nd, const double *w,
                        const int *ox, const int *oy, double *out)
{
    memset(out, 0, sizeof(double) * (size_t)H * W);
    for (int r = 0; r < H; ++r)
        for (int c = 0; c < W; ++c)
            for (int d = 0; d < nd; ++d) {
                int xx = clampi(c + ox[d], 0, W - 1);
                int yy = clampi(r + oy[d], 0, H - 1);
                out[(int64_t)yy * W + xx] += w[d] * g[(int64_t)r * W + c];
            }
}

/* ------------------------------------------------------------------------- */
/* a2  SpmvDIA2D: y[p] = sum_d [p+o_d in bounds] A[d][p+o_d] * x[p+o_d]
 *     SpmvDIA2D.cpp:35-45; coefficient index SpMat2D.cpp:211-216
 *     (element_id0 + element_id1*W + diag*W*H + offx + offy*W).             */
void or_dia2d_f32(const float *A, int nd, const int *ox, const int *oy,
                  const float *x, int H, int W, float *y)
{
    const int64_t plane = (int64_t)H * W;
    #pragma omp parallel for schedule(static)
    for (int r = 0; r < H; ++r) {
        for (int c = 0; c < W; ++c) {
            float acc = 0.0f;
            for (int d = 0; d < nd; ++d) {
                int xx = c + ox[d], yy = r + oy[d];
                if (xx >= 0 && xx < W && yy >= 0 && yy < H) {
                    int64_t q = (int64_t)yy * W + xx;
                    acc = fmaf(A[d * plane + q], x[q], acc);
                }
            }
            y[(int64_t)r * W + c] = acc;
        }
    }
}

void or_dia2d_f64(const double *A, int nd, const int *ox, const int *oy,
                  const double *x, int H, int W, double *y)
{
    const int64_t plane = (int64_t)H * W;
    #pragma omp parallel for schedule(static)
    for (int r = 0; r < H; ++r) {
        for (int c = 0; c < W; ++c) {
            double acc = 0.0;
            for (int d = 0; d < nd; ++d) {
                int xx = c + ox[d], yy = r + oy[d];
                if (xx >= 0 && xx < W && yy >= 0 && yy < H) {
                    int64_t q = (int64_t)yy * W + xx;
                    acc += A[d * plane + q] * x[q];
                }
            }
            y[(int64_t)r * W + c] = acc;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* a3  SpmvCSR: y[i] = sum_{j in [indptr[i], indptr[i+1])} data[j]*x[indices[j]]
 *     SpmvCSR.cpp:32-38 (one work-item per row, sequential j).  int32
 *     indices/indptr (CSRMat.cpp:26-27).                                     */
void or_csr_f32(int M, const int *indptr, const int *indices, const float *data,
                const float *x, float *y)
{
    #pragma omp parallel for schedule(dynamic, 1024)
    for (int i = 0; i < M; ++i) {
        float acc = 0.0f;
        for (int j = indptr[i]; j < indptr[i + 1]; ++j)
            acc = fmaf(data[j], x[indices[j]], acc);
        y[i] = acc;
    }
}

void or_csr_f64(int M, const int *indptr, const int *indices, const double *data,
                const double *x, double *y)
{
    #pragma omp parallel for schedule(dynamic, 1024)
    for (int i = 0; i < M; ++i) {
        double acc = 0.0;
        for (int j = indptr[i]; j < indptr[i + 1]; ++j)
            acc += data[j] * x[indices[j]];
        y[i] = acc;
    }
}

/* ------------------------------------------------------------------------- */
/* a7  Array2DSum (sum2d): exact-order-free fp64 sum (Array2DSum.cpp:12-119
 *     reduces in fp32 over a work-group tree; the oracle gives the value the
 *     tree approximates, with |err| <= n*eps32*sum|x| as the test bound).    */
double or_sum_f64(const float *x, int64_t n)
{
    double s = 0.0;
    #pragma omp parallel for reduction(+:s) schedule(static)
    for (int64_t i = 0; i < n; ++i) s += (double)x[i];
    return s;
}

/* ------------------------------------------------------------------------- */
/* a10 HS setup of the canonical program:
 *       Ix = Dx*I1 + Dx*I2 ; Iy = Dy*I1 + Dy*I2 ; It = Dt*I2 - Dt*I1
 *     each "S * X" a Stencil2DMVM (clamp), each +/- an Array2DOp
 *     (Array2DOp.cpp:27-53), one rounding per op.                            */
typedef struct {
    int nd;
    const float *w;
    const int *ox;
    const int *oy;
} or_st32;

static inline float st_apply_f32(const float *x, int H, int W, int r, int c,
                                 int nd, const float *w, const int *ox, const int *oy)
{
    float acc = 0.0f;
    for (int d = 0; d < nd; ++d) {
        int xx = clampi(c + ox[d], 0, W - 1);
        int yy = clampi(r + oy[d], 0, H - 1);
        acc = fmaf(w[d], x[(int64_t)yy * W + xx], acc);
    }
    return acc;
}

static inline double st_apply_f64(const double *x, int H, int W, int r, int c,
                                  int nd, const double *w, const int *ox, const int *oy)
{
    double acc = 0.0;
    for (int d = 0; d < nd; ++d) {
        int xx = clampi(c + ox[d], 0, W - 1);
        int yy = clampi(r + oy[d], 0, H - 1);
        acc += w[d] * x[(int64_t)yy * W + xx];
    }
    return acc;
}

void or_hs_setup_f32(const float *I1, const float *I2, int H, int W,
                     int ndx, const float *wx, const int *oxx, const int *oyx,
                     int ndy, const float *wy, const int *oxy, const int *oyy,
                     int ndt, const float *wt, const int *oxt, const int *oyt,
                     float *Ix, float *Iy, float *It)
{
    #pragma omp parallel for schedule(static)
    for (int r = 0; r < H; ++r)
        for (int c = 0; c < W; ++c) {
            int64_t p = (int64_t)r * W + c;
            float a1 = st_apply_f32(I1, H, W, r, c, ndx, wx, oxx, oyx);
            float a2 = st_apply_f32(I2, H, W, r, c, ndx, wx, oxx, oyx);
            Ix[p] = a1 + a2;
            float b1 = st_apply_f32(I1, H, W, r, c, ndy, wy, oxy, oyy);
            float b2 = st_apply_f32(I2, H, W, r, c, ndy, wy, oxy, oyy);
            Iy[p] = b1 + b2;
            float t2 = st_apply_f32(I2, H, W, r, c, ndt, wt, oxt, oyt);
            float t1 = st_apply_f32(I1, H, W, r, c, ndt, wt, oxt, oyt);
            It[p] = t2 - t1;
        }
}

void or_hs_setup_f64(const double *I1, const double *I2, int H, int W,
                     int ndx, const double *wx, const int *oxx, const int *oyx,
                     int ndy, const double *wy, const int *oxy, const int *oyy,
                     int ndt, const double *wt, const int *oxt, const int *oyt,
                     double *Ix, double *Iy, double *It)
{
    #pragma omp parallel for schedule(static)
    for (int r = 0; r < H; ++r)
        for (int c = 0; c < W; ++c) {
            int64_t p = (int64_t)r * W + c;
            Ix[p] = st_apply_f64(I1, H, W, r, c, ndx, wx, oxx, oyx) +
                    st_apply_f64(I2, H, W, r, c, ndx, wx, oxx, oyx);
            Iy[p] = st_apply_f64(I1, H, W, r, c, ndy, wy, oxy, oyy) +
                    st_apply_f64(I2, H, W, r, c, ndy, wy, oxy, oyy);
            It[p] = st_apply_f64(I2, H, W, r, c, ndt, wt, oxt, oyt) -
                    st_apply_f64(I1, H, W, r, c, ndt, wt, oxt, oyt);
        }
}

/* a9  HS Jacobi iterations of the canonical program (Templates.py:57-71 loop,
 *     body lowered to Stencil2DMVM + Array2DOp chain + Array2DCopy):
 *        den  = alpha2 + Ix*Ix + Iy*Iy          (before the loop)
 *        ubar = Avg*u ; vbar = Avg*v
 *        t    = (Ix*ubar + Iy*vbar + It) / den
 *        u    = ubar - Ix*t ; v = vbar - Iy*t
 *     Jacobi semantics (all stencil reads see the previous iterate) as the
 *     reference's kernel split enforces.  u, v are updated in place.
 *     Python precedence: ((Ix*ubar) + (Iy*vbar)) + It and
 *     (alpha2 + (Ix*Ix)) + (Iy*Iy).                                          */
void or_hs_jacobi_f32(const float *Ix, const float *Iy, const float *It,
                      float alpha2, int H, int W, int nda, const float *wa,
                      const int *oxa, const int *oya, int iters,
                      float *u, float *v)
{
    size_t n = (size_t)H * W;
    float *un = (float *)malloc(n * sizeof(float));
    float *vn = (float *)malloc(n * sizeof(float));
    float *den = (float *)malloc(n * sizeof(float));
    #pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < (int64_t)n; ++p) {
        float a = Ix[p] * Ix[p];
        float b = alpha2 + a;
        float c = Iy[p] * Iy[p];
        den[p] = b + c;
    }
    for (int it = 0; it < iters; ++it) {
        #pragma omp parallel for schedule(static)
        for (int r = 0; r < H; ++r)
            for (int c = 0; c < W; ++c) {
                int64_t p = (int64_t)r * W + c;
                float ub = st_apply_f32(u, H, W, r, c, nda, wa, oxa, oya);
                float vb = st_apply_f32(v, H, W, r, c, nda, wa, oxa, oya);
                float m1 = Ix[p] * ub;
                float m2 = Iy[p] * vb;
                float s1 = m1 + m2;
                float s2 = s1 + It[p];
                float t = s2 / den[p];
                float ix_t = Ix[p] * t;
                float iy_t = Iy[p] * t;
                un[p] = ub - ix_t;
                vn[p] = vb - iy_t;
            }
        memcpy(u, un, n * sizeof(float));
        memcpy(v, vn, n * sizeof(float));
    }
    free(un); free(vn); free(den);
}

void or_hs_jacobi_f64(const double *Ix, const double *Iy, const double *It,
                      double alpha2, int H, int W, int nda, const double *wa,
                      const int *oxa, const int *oya, int iters,
                      double *u, double *v)
{
    size_t n = (size_t)H * W;
    double *un = (double *)malloc(n * sizeof(double));
    double *vn = (double *)malloc(n * sizeof(double));
    for (int it = 0; it < iters; ++it) {
        #pragma omp parallel for schedule(static)
        for (int r = 0; r < H; ++r)
            for (int c = 0; c < W; ++c) {
                int64_t p = (int64_t)r * W + c;
                double den = alpha2 + Ix[p] * Ix[p] + Iy[p] * Iy[p];
                double ub = st_apply_f64(u, H, W, r, c, nda, wa, oxa, oya);
                double vb = st_apply_f64(v, H, W, r, c, nda, wa, oxa, oya);
                double t = (Ix[p] * ub + Iy[p] * vb + It[p]) / den;
                un[p] = ub - Ix[p] * t;
                vn[p] = vb - Iy[p] * t;
            }
        memcpy(u, un, n * sizeof(double));
        memcpy(v, vn, n * sizeof(double));
    }
    free(un); free(vn);
}

/* ------------------------------------------------------------------------- */
/* a13 Reverse-mode gradient of  L(u_N, v_N)  for the HS solve, in fp64.
 *     Forward:  setup (a10), den = alpha^2 + Ix^2 + Iy^2, N Jacobi steps
 *     from (u0, v0).  Given gu = dL/du_N, gv = dL/dv_N, returns dL/dI1,
 *     dL/dI2, dL/dalpha (and dL/du0, dL/dv0 in gu, gv on exit).
 *     All intermediate iterates are stored (the oracle is only run on sizes
 *     where that fits).  Per step, with s = -(Ix*gu + Iy*gv)/den:
 *        g_ubar = gu + Ix*s ; g_vbar = gv + Iy*s
 *        gIx += -gu*t + s*ubar - 2*s*t*Ix ; gIy likewise ; gIt += s
 *        galpha2 += sum(-s*t) ;  gu <- Avg^T g_ubar ; gv <- Avg^T g_vbar
 *     then gI1 = Dx^T gIx + Dy^T gIy - Dt^T gIt, gI2 = ... + Dt^T gIt.       */
void or_hs_grad_f64(const double *I1, const double *I2, int H, int W, double alpha,
                    int ndx, const double *wx, const int *oxx, const int *oyx,
                    int ndy, const double *wy, const int *oxy, const int *oyy,
                    int ndt, const double *wt, const int *oxt, const int *oyt,
                    int nda, const double *wa, const int *oxa, const int *oya,
                    int iters, const double *u0, const double *v0,
                    double *gu, double *gv,
                    double *gI1, double *gI2, double *galpha)
{
    size_t n = (size_t)H * W;
    double *Ix = (double *)malloc(n * sizeof(double));
    double *Iy = (double *)malloc(n * sizeof(double));
    double *It = (double *)malloc(n * sizeof(double));
    double *den = (double *)malloc(n * sizeof(double));
    double *us = (double *)malloc(n * sizeof(double) * (size_t)(iters + 1));
    double *vs = (double *)malloc(n * sizeof(double) * (size_t)(iters + 1));
    double *gIx = (double *)calloc(n, sizeof(double));
    double *gIy = (double *)calloc(n, sizeof(double));
    double *gIt = (double *)calloc(n, sizeof(double));
    double *gub = (double *)malloc(n * sizeof(double));
    double *gvb = (double *)malloc(n * sizeof(double));
    double *tmp = (double *)malloc(n * sizeof(double));
    double alpha2 = alpha * alpha;
    or_hs_setup_f64(I1, I2, H, W, ndx, wx, oxx, oyx, ndy, wy, oxy, oyy, ndt, wt, oxt, oyt,
                    Ix, Iy, It);
    for (size_t p = 0; p < n; ++p) den[p] = alpha2 + Ix[p] * Ix[p] + Iy[p] * Iy[p];
    memcpy(us, u0, n * sizeof(double));
    memcpy(vs, v0, n * sizeof(double));
    for (int k = 0; k < iters; ++k) {
        memcpy(us + n * (k + 1), us + n * k, n * sizeof(double));
        memcpy(vs + n * (k + 1), vs + n * k, n * sizeof(double));
        or_hs_jacobi_f64(Ix, Iy, It, alpha2, H, W, nda, wa, oxa, oya, 1,
                         us + n * (k + 1), vs + n * (k + 1));
    }
    double ga2 = 0.0;
    for (int k = iters - 1; k >= 0; --k) {
        const double *uk = us + n * k, *vk = vs + n * k;
        for (int r = 0; r < H; ++r)
            for (int c = 0; c < W; ++c) {
                size_t p = (size_t)r * W + c;
                double ub = st_apply_f64(uk, H, W, r, c, nda, wa, oxa, oya);
                double vb = st_apply_f64(vk, H, W, r, c, nda, wa, oxa, oya);
                double t = (Ix[p] * ub + Iy[p] * vb + It[p]) / den[p];
                double s = -(Ix[p] * gu[p] + Iy[p] * gv[p]) / den[p];
                gub[p] = gu[p] + Ix[p] * s;
                gvb[p] = gv[p] + Iy[p] * s;
                gIx[p] += -gu[p] * t + s * ub - 2.0 * s * t * Ix[p];
                gIy[p] += -gv[p] * t + s * vb - 2.0 * s * t * Iy[p];
                gIt[p] += s;
                ga2 += -s * t;
            }
        or_stencil2d_T_f64(gub, H, W, nda, wa, oxa, oya, gu);
        or_stencil2d_T_f64(gvb, H, W, nda, wa, oxa, oya, gv);
    }
    /* den = alpha^2 + ... : dL/dalpha = 2*alpha * dL/dalpha2 */
    *galpha = 2.0 * alpha * ga2;
    or_stencil2d_T_f64(gIx, H, W, ndx, wx, oxx, oyx, gI1);
    or_stencil2d_T_f64(gIy, H, W, ndy, wy, oxy, oyy, tmp);
    for (size_t p = 0; p < n; ++p) gI1[p] += tmp[p];
    memcpy(gI2, gI1, n * sizeof(double));
    or_stencil2d_T_f64(gIt, H, W, ndt, wt, oxt, oyt, tmp);
    for (size_t p = 0; p < n; ++p) { gI1[p] -= tmp[p]; gI2[p] += tmp[p]; }
    free(Ix); free(Iy); free(It); free(den); free(us); free(vs);
    free(gIx); free(gIy); free(gIt); free(gub); free(gvb); free(tmp);
}

/* ------------------------------------------------------------------------- */
/* Full fp32 solve as one call: setup (a10) + den + N iterations (a9), u/v in
 * place.  This is the CPU baseline the benchmark times (OpenMP over rows). */
void or_hs_solve_f32(const float *I1, const float *I2, int H, int W, float alpha2,
                     int ndx, const float *wx, const int *oxx, const int *oyx,
                     int ndy, const float *wy, const int *oxy, const int *oyy,
                     int ndt, const float *wt, const int *oxt, const int *oyt,
                     int nda, const float *wa, const int *oxa, const int *oya,
                     int iters, float *u, float *v)
{
    size_t n = (size_t)H * W;
    float *Ix = (float *)malloc(n * sizeof(float));
    float *Iy = (float *)malloc(n * sizeof(float));
    float *It = (float *)malloc(n * sizeof(float));
    or_hs_setup_f32(I1, I2, H, W, ndx, wx, oxx, oyx, ndy, wy, oxy, oyy, ndt, wt, oxt, oyt,
                    Ix, Iy, It);
    or_hs_jacobi_f32(Ix, Iy, It, alpha2, H, W, nda, wa, oxa, oya, iters, u, v);
    free(Ix); free(Iy); free(It);
}

/* ------------------------------------------------------------------------
 * Preconditioned-CG Horn-Schunck operations.
 *
 * or_hs_precond_*: ApplyHSPreconditioner.cpp:46-97.  Per pixel, with s2 the
 * stencil's weight 2: a = diag0 + s2 per in-image 4-neighbour (added in the
 * order x<W-1, y<H-1, x>0, y>0), d likewise from diag1, b = c = offdiag;
 * det = a*d - b*c; if |det| > 1e-15 (a double literal):
 * y0 = (e*d - b*f)/det, y1 = (-c*e + a*f)/det, else y = (e, f).
 */
void or_hs_precond_f32(const float *diag0, const float *diag1, const float *offdiag,
                       const float *x0, const float *x1, float *y0, float *y1, int H, int W,
                       float s2)
{
#pragma omp parallel for schedule(static)
    for (int r = 0; r < H; ++r)
        for (int c = 0; c < W; ++c) {
            const int64_t p = (int64_t)r * W + c;
            float a = diag0[p], d = diag1[p];
            if (c < W - 1) { a = a + s2; d = d + s2; }
            if (r < H - 1) { a = a + s2; d = d + s2; }
            if (c > 0) { a = a + s2; d = d + s2; }
            if (r > 0) { a = a + s2; d = d + s2; }
            const float b = offdiag[p], e = x0[p], f = x1[p];
            const float ad = a * d, bc = b * b;
            const float det = ad - bc;
            if (fabs((double)det) > 1e-15) {
                const float t0 = e * d, t1 = b * f;
                const float n0 = t0 - t1;
                const float t2 = (-b) * e, t3 = a * f;
                const float n1 = t2 + t3;
                y0[p] = n0 / det;
                y1[p] = n1 / det;
            } else {
                y0[p] = e;
                y1[p] = f;
            }
        }
}

void or_hs_precond_f64(const double *diag0, const double *diag1, const double *offdiag,
                       const double *x0, const double *x1, double *y0, double *y1, int H, int W,
                       double s2)
{
#pragma omp parallel for schedule(static)
    for (int r = 0; r < H; ++r)
        for (int c = 0; c < W; ++c) {
            const int64_t p = (int64_t)r * W + c;
            double a = diag0[p], d = diag1[p];
            if (c < W - 1) { a += s2; d += s2; }
            if (r < H - 1) { a += s2; d += s2; }
            if (c > 0) { a += s2; d += s2; }
            if (r > 0) { a += s2; d += s2; }
            const double b = offdiag[p], e = x0[p], f = x1[p];
            const double det = a * d - b * b;
            if (fabs(det) > 1e-15) {
                y0[p] = (e * d - b * f) / det;
                y1[p] = (-b * e + a * f) / det;
            } else {
                y0[p] = e;
                y1[p] = f;
            }
        }
}

/*
 * or_hs_powers_*: HSMatrixPowers.cpp:28-39 (V[0] = r), :72-126 (one power)
 * and :206-226 (s = 1..steps-1).  V0, V1 hold steps planes of H x W.  Per
 * pixel of power s from (u, v) = power s-1:
 *   U = Ix * (Ix*u + Iy*v);  V = Iy * (Ix*u + Iy*v)
 *   then for each in-image neighbour, in the order left, right, up, down:
 *   U += -lam * u_nb;  V += -lam * v_nb;  U += lam * u;  V += lam * v
 * (the "+=" statements as fmas in the fp32 restatement).
 */
void or_hs_powers_f32(float lam, const float *Ix, const float *Iy, const float *r0,
                      const float *r1, float *V0, float *V1, int H, int W, int steps)
{
    const int64_t n = (int64_t)H * W;
    memcpy(V0, r0, sizeof(float) * n);
    memcpy(V1, r1, sizeof(float) * n);
    const float nlam = -lam;
    for (int s = 1; s < steps; ++s) {
        const float *u = V0 + (int64_t)(s - 1) * n, *v = V1 + (int64_t)(s - 1) * n;
        float *U = V0 + (int64_t)s * n, *V = V1 + (int64_t)s * n;
#pragma omp parallel for schedule(static)
        for (int r = 0; r < H; ++r)
            for (int c = 0; c < W; ++c) {
                const int64_t p = (int64_t)r * W + c;
                const float ix = Ix[p], iy = Iy[p], um = u[p], vm = v[p];
                const float p1 = ix * um, p2 = iy * vm;
                const float dot = p1 + p2;
                float ur = ix * dot, vr = iy * dot;
                const int64_t nb[4] = {p - 1, p + 1, p - W, p + W};
                const int ok[4] = {c > 0, c < W - 1, r > 0, r < H - 1};
                for (int k = 0; k < 4; ++k)
                    if (ok[k]) {
                        ur = fmaf(nlam, u[nb[k]], ur);
                        vr = fmaf(nlam, v[nb[k]], vr);
                        ur = fmaf(lam, um, ur);
                        vr = fmaf(lam, vm, vr);
                    }
                U[p] = ur;
                V[p] = vr;
            }
    }
}

void or_hs_powers_f64(double lam, const double *Ix, const double *Iy, const double *r0,
                      const double *r1, double *V0, double *V1, int H, int W, int steps)
{
    const int64_t n = (int64_t)H * W;
    memcpy(V0, r0, sizeof(double) * n);
    memcpy(V1, r1, sizeof(double) * n);
    for (int s = 1; s < steps; ++s) {
        const double *u = V0 + (int64_t)(s - 1) * n, *v = V1 + (int64_t)(s - 1) * n;
        double *U = V0 + (int64_t)s * n, *V = V1 + (int64_t)s * n;
#pragma omp parallel for schedule(static)
        for (int r = 0; r < H; ++r)
            for (int c = 0; c < W; ++c) {
                const int64_t p = (int64_t)r * W + c;
                const double dot = Ix[p] * u[p] + Iy[p] * v[p];
                double ur = Ix[p] * dot, vr = Iy[p] * dot;
                const int64_t nb[4] = {p - 1, p + 1, p - W, p + W};
                const int ok[4] = {c > 0, c < W - 1, r > 0, r < H - 1};
                for (int k = 0; k < 4; ++k)
                    if (ok[k]) {
                        ur += lam * (u[p] - u[nb[k]]);
                        vr += lam * (v[p] - v[nb[k]]);
                    }
                U[p] = ur;
                V[p] = vr;
            }
    }
}

/* ------------------------------------------------------------------------
 * or_warp_img2d_f32: WarpImg2D.cpp:33-75, statement by statement in the types C
 * gives them (double literals 1.0; each "out += term" one fma in the term's
 * type, rounded to float).  or_warp_img2d_f64: the same split of the offset
 * (truncate, one step down when negative) and the same bounds, in double.
 */
void or_warp_img2d_f32(const float *src, const float *offx, const float *offy, float *dst, int H,
                       int W)
{
#pragma omp parallel for schedule(static)
    for (int r = 0; r < H; ++r)
        for (int c = 0; c < W; ++c) {
            const int64_t p = (int64_t)r * W + c;
            const float ox = offx[p], oy = offy[p];
            int mx = (int)ox, my = (int)oy;
            float xf = ox - (float)mx, yf = oy - (float)my;
            if (ox < 0.0f) { --mx; xf = (float)(1.0 + (double)xf); }
            if (oy < 0.0f) { --my; yf = (float)(1.0 + (double)yf); }
            const int64_t sx = (int64_t)c + mx, sy = (int64_t)r + my;
            if (sx >= 0 && sx + 1 < W && sy >= 0 && sy + 1 < H) {
                const int64_t q = sy * W + sx;
                const double ax = 1.0 - (double)xf, ay = 1.0 - (double)yf;
                float out = 0.0f;
                out = (float)fma((double)src[q] * ax, ay, (double)out);
                const float p10 = src[q + 1] * xf;
                out = (float)fma((double)p10, ay, (double)out);
                out = (float)fma((double)src[q + W] * ax, (double)yf, (double)out);
                const float p11 = src[q + W + 1] * xf;
                out = fmaf(p11, yf, out);
                dst[p] = out;
            } else {
                dst[p] = src[p];
            }
        }
}

void or_warp_img2d_f64(const double *src, const double *offx, const double *offy, double *dst,
                       int H, int W)
{
#pragma omp parallel for schedule(static)
    for (int r = 0; r < H; ++r)
        for (int c = 0; c < W; ++c) {
            const int64_t p = (int64_t)r * W + c;
            const double ox = offx[p], oy = offy[p];
            int64_t mx = (int64_t)ox, my = (int64_t)oy;
            double xf = ox - (double)mx, yf = oy - (double)my;
            if (ox < 0.0) { --mx; xf += 1.0; }
            if (oy < 0.0) { --my; yf += 1.0; }
            const int64_t sx = (int64_t)c + mx, sy = (int64_t)r + my;
            if (sx >= 0 && sx + 1 < W && sy >= 0 && sy + 1 < H) {
                const int64_t q = sy * W + sx;
                dst[p] = src[q] * (1 - xf) * (1 - yf) + src[q + 1] * xf * (1 - yf) +
                         src[q + W] * (1 - xf) * yf + src[q + W + 1] * xf * yf;
            } else {
                dst[p] = src[p];
            }
        }
}

/* ------------------------------------------------------------------------
 * Brox SOR building blocks.
 * or_brox_set_f32: SetBroxMatrix.cpp:34-79 (rb = 0: planes top, left, mid,
 * right, bottom) and SetBroxRedblackMatrix.cpp:37-81 (rb = 1: planes top, left,
 * right, bottom; the diagonal into diag).  Off-diagonal coefficients at the
 * neighbour's index: -alpha*psi_e; the diagonal accumulates alpha*psi_e (fma)
 * in the order up, left, right, down.  Entries no statement writes keep their
 * value.
 */
void or_brox_set_f32(float *A, float *diag, const float *psi, int H, int W, float alpha, int rb)
{
    const int64_t HW = (int64_t)H * W;
    const int pt = 0, pl = 1, pr = rb ? 2 : 3, pb = rb ? 3 : 4;
    const float nalpha = -alpha;
    for (int r = 0; r < H; ++r)
        for (int c = 0; c < W; ++c) {
            const int64_t p = (int64_t)r * W + c;
            float mid = 0.0f;
            if (r > 0) { A[pt * HW + p - W] = nalpha * psi[p - W]; mid = fmaf(alpha, psi[p - W], mid); }
            if (c > 0) { A[pl * HW + p - 1] = nalpha * psi[p - 1]; mid = fmaf(alpha, psi[p - 1], mid); }
            if (c < W - 1) { A[pr * HW + p + 1] = nalpha * psi[p]; mid = fmaf(alpha, psi[p], mid); }
            if (r < H - 1) { A[pb * HW + p + W] = nalpha * psi[p]; mid = fmaf(alpha, psi[p], mid); }
            if (rb) diag[p] = mid;
            else A[2 * HW + p] = mid;
        }
}

/* or_brox_precond_*: ApplyBroxPreconditioner.cpp:46-72: a = diag0 + A2, d = diag1
 * + A2 (A2 = the matrix's plane 2 at the pixel), threshold |det| > 1e-4. */
void or_brox_precond_f32(const float *A2, const float *diag0, const float *diag1,
                         const float *offdiag, const float *x0, const float *x1, float *y0,
                         float *y1, int64_t n)
{
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p) {
        const float a = diag0[p] + A2[p], d = diag1[p] + A2[p];
        const float b = offdiag[p], e = x0[p], f = x1[p];
        const float ad = a * d, bc = b * b;
        const float det = ad - bc;
        if (fabs((double)det) > 1e-4) {
            const float t0 = e * d, t1 = b * f;
            const float n0 = t0 - t1;
            const float t2 = (-b) * e, t3 = a * f;
            const float n1 = t2 + t3;
            y0[p] = n0 / det;
            y1[p] = n1 / det;
        } else {
            y0[p] = e;
            y1[p] = f;
        }
    }
}

void or_brox_precond_f64(const double *A2, const double *diag0, const double *diag1,
                         const double *offdiag, const double *x0, const double *x1, double *y0,
                         double *y1, int64_t n)
{
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p) {
        const double a = diag0[p] + A2[p], d = diag1[p] + A2[p], b = offdiag[p];
        const double det = a * d - b * b;
        if (fabs(det) > 1e-4) {
            y0[p] = (x0[p] * d - b * x1[p]) / det;
            y1[p] = (-b * x0[p] + a * x1[p]) / det;
        } else {
            y0[p] = x0[p];
            y1[p] = x1[p];
        }
    }
}
