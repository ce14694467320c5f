// The reference's Brox SOR building blocks (a19):
//   SetBroxMatrix          (SetBroxMatrix.cpp:34-79)
//   SetBroxRedblackMatrix  (SetBroxRedblackMatrix.cpp:37-81)
//   ApplyBroxPreconditioner(ApplyBroxPreconditioner.cpp:46-72)
// (red / black masks are opcodes of the fused elementwise programs.)
//
// SetBroxMatrix writes the psi-weighted 4-neighbour Laplacian into the
// diagonal planes of an SpMat2D, each coefficient at the neighbour's index
// (the SpMat2D convention, SpMat2D.cpp:211-216), so that A * x with offsets
// (0,-1), (-1,0), (0,0), (1,0), (0,1) applies
//   (A x)_p = sum_nb alpha*psi_e (x_p - x_nb),  psi_e = psi at the edge's
//   upper / left pixel.
// The red-black form stores the diagonal in its own Array2D and the four
// off-diagonals in planes 0..3.
#include "hmx_common.h"

#include <algorithm>

#pragma clang fp contract(off)

namespace hmx {

__device__ __forceinline__ float load_scalar_any(const void *p, int dt, float val) {
    if (!p) return val;
    switch (dt) {
        case HMX_F64: return (float)*(const double *)p;
        case HMX_I64: return (float)*(const long long *)p;
        case HMX_I32: return (float)*(const int *)p;
        default: return *(const float *)p;
    }
}

// 8 B/px read (psi twice through the cache: p and its up/left neighbour) +
// 20 B/px written (5 planes) for the plain form; the red-black form writes 4
// planes and the diagonal array.
__global__ __launch_bounds__(256) void brox_set_kernel(float *__restrict__ A,
                                                       float *__restrict__ diag,
                                                       const float *__restrict__ psi, int H,
                                                       int W, const void *alpha_dev,
                                                       int alpha_dt, float alpha_val, int rb) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= W) return;
    const float alpha = load_scalar_any(alpha_dev, alpha_dt, alpha_val);
    const float nalpha = -alpha;
    const int64_t HW = (int64_t)H * W;
    // plane indices: top, left, right, bottom (and mid for the plain form)
    const int pt = 0, pl = 1, pr = rb ? 2 : 3, pb = rb ? 3 : 4;
    for (int r = blockIdx.y; r < H; r += gridDim.y) {
        const int64_t p = (int64_t)r * W + c;
        float mid = 0.0f;
        const float ps = psi[p];
        if (r > 0) {
            const float pu = psi[p - W];
            st_out(A + (pt * HW + p - W), nalpha * pu, rows_whole_lines(W));
            mid = fmaf(alpha, pu, mid);
        }
        if (c > 0) {
            const float pl_ = psi[p - 1];
            st_out(A + (pl * HW + p - 1), nalpha * pl_, rows_whole_lines(W));
            mid = fmaf(alpha, pl_, mid);
        }
        if (c < W - 1) {
            st_out(A + (pr * HW + p + 1), nalpha * ps, rows_whole_lines(W));
            mid = fmaf(alpha, ps, mid);
        }
        if (r < H - 1) {
            st_out(A + (pb * HW + p + W), nalpha * ps, rows_whole_lines(W));
            mid = fmaf(alpha, ps, mid);
        }
        if (rb) st_out(diag + p, mid, rows_whole_lines(W));
        else st_out(A + (2 * HW + p), mid, rows_whole_lines(W));
    }
}

// 2x2 solve with a = diag0 + A[2], d = diag1 + A[2] (A's diagonal plane at the
// pixel), b = offdiag; |det| <= 1e-4 (a double literal) passes x through.
// 32 B/px: read A[2], diag0, diag1, offdiag, x0, x1; write y0, y1.
__global__ __launch_bounds__(256) void brox_precond_kernel(
    const float *__restrict__ A2, const float *__restrict__ diag0,
    const float *__restrict__ diag1, const float *__restrict__ offdiag,
    const float *__restrict__ x0, const float *__restrict__ x1, float *__restrict__ y0,
    float *__restrict__ y1, int64_t n) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * blockDim.x) {
        const float s = A2[p];
        const float a = diag0[p] + s, d = diag1[p] + s;
        const float b = offdiag[p], e = x0[p], f = x1[p];
        const float ad = a * d, bc = b * b;
        const float det = ad - bc;
        if (fabs((double)det) > 1e-4) {
            const float t0 = e * d, t1 = b * f;
            const float n0 = t0 - t1;
            const float t2 = (-b) * e, t3 = a * f;
            const float n1 = t2 + t3;
            st_out(y0 + p, n0 / det, true);
            st_out(y1 + p, n1 / det, true);
        } else {
            st_out(y0 + p, e, true);
            st_out(y1 + p, f, true);
        }
    }
}

int launch_brox_set(float *A, float *diag, const float *psi, int H, int W, const void *alpha_dev,
                    int alpha_dt, float alpha_val, int rb, hipStream_t s) {
    dim3 grid((W + 255) / 256, std::min(H, 65535));
    hipLaunchKernelGGL(brox_set_kernel, grid, dim3(256), 0, s, A, diag, psi, H, W, alpha_dev,
                       alpha_dt, alpha_val, rb);
    HMX_LAUNCH_CHECK("brox_set");
    return HMX_OK;
}

int launch_brox_precond(const float *A2, const float *diag0, const float *diag1,
                        const float *offdiag, const float *x0, const float *x1, float *y0,
                        float *y1, int H, int W, hipStream_t s) {
    const int64_t n = (int64_t)H * W;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(brox_precond_kernel, dim3((unsigned)blocks), dim3(256), 0, s, A2, diag0,
                       diag1, offdiag, x0, x1, y0, y1, n);
    HMX_LAUNCH_CHECK("brox_precond");
    return HMX_OK;
}

}  // namespace hmx

using namespace hmx;

extern "C" int hmx_brox_set_matrix(float *A, int nd, const float *psi, float alpha, int H, int W,
                                   void *stream) {
    HMX_REQUIRE(A && psi, "hmx_brox_set_matrix: null argument");
    HMX_REQUIRE(H > 0 && W > 0, "hmx_brox_set_matrix: H and W must be positive");
    HMX_REQUIRE(nd >= 5, "hmx_brox_set_matrix: the matrix needs 5 diagonal planes");
    return launch_brox_set(A, nullptr, psi, H, W, nullptr, HMX_F32, alpha, 0, as_stream(stream));
}

extern "C" int hmx_brox_set_redblack_matrix(float *Adiag, float *A, int nd, const float *psi,
                                            float alpha, int H, int W, void *stream) {
    HMX_REQUIRE(Adiag && A && psi, "hmx_brox_set_redblack_matrix: null argument");
    HMX_REQUIRE(H > 0 && W > 0, "hmx_brox_set_redblack_matrix: H and W must be positive");
    HMX_REQUIRE(nd >= 4, "hmx_brox_set_redblack_matrix: the matrix needs 4 diagonal planes");
    return launch_brox_set(A, Adiag, psi, H, W, nullptr, HMX_F32, alpha, 1, as_stream(stream));
}

extern "C" int hmx_brox_precondition(const float *A, int nd, const float *diag0,
                                     const float *diag1, const float *offdiag, const float *x0,
                                     const float *x1, float *y0, float *y1, int H, int W,
                                     void *stream) {
    HMX_REQUIRE(A && diag0 && diag1 && offdiag && x0 && x1 && y0 && y1,
                "hmx_brox_precondition: null argument");
    HMX_REQUIRE(H > 0 && W > 0, "hmx_brox_precondition: H and W must be positive");
    HMX_REQUIRE(nd >= 3, "hmx_brox_precondition: the matrix needs a diagonal plane 2");
    HMX_REQUIRE(y0 != x0 && y0 != x1 && y1 != x0 && y1 != x1 && y0 != y1,
                "hmx_brox_precondition: outputs must be distinct from the inputs");
    return launch_brox_precond(A + 2 * (int64_t)H * W, diag0, diag1, offdiag, x0, x1, y0, y1, H,
                               W, as_stream(stream));
}
