// Interpreter kernels for fused elementwise programs and scalar statements.
//
// Rounding matches the reference program: each Array2DOp / ScalarArray2DOp /
// Array2DScalarOp is one IEEE fp32 op (Array2DOp.cpp:30 "a = b op c"),
// division and sqrt are correctly rounded, stencil taps accumulate with fma
// in tap order from 0 (Stencil2DMVM.cpp:37-47), DIA taps likewise with the
// coefficient read at the neighbour (SpmvDIA2D.cpp:41-43, SpMat2D.cpp:214).
#include "program.h"

#pragma clang fp contract(off)

namespace hmx {

__device__ inline float load_scalar(const void *p, int dt) {
    switch (dt) {
        case HMX_F64: return (float)*(const double *)p;
        case HMX_I64: return (float)*(const long long *)p;
        case HMX_I32: return (float)*(const int *)p;
        default: return *(const float *)p;
    }
}

__global__ __launch_bounds__(PROG_NT) void program_kernel(ProgArgs a) {
    __shared__ float reg[PROG_MAX_REG][PROG_NT];
    __shared__ double red_lds[PROG_NT / 64];
    const int tid = threadIdx.x;
    const int c = blockIdx.x * PROG_NT + tid;
    double acc[PROG_MAX_RED] = {0.0, 0.0};
    if (c < a.W) {
        for (int r = blockIdx.y; r < a.H; r += gridDim.y) {
            const int64_t e = (int64_t)r * a.W + c;
            for (int pc = 0; pc < a.ninstr; ++pc) {
                const hmx_instr in = a.ins[pc];
                float v = 0.0f;
                switch (in.op) {
                    case HMX_OP_LOAD: v = a.in[in.a][e]; break;
                    case HMX_OP_LOADS: v = load_scalar(a.scal[in.a], a.scal_dt[in.a]); break;
                    case HMX_OP_STENCIL: {
                        const ProgOperand &o = a.opnd[in.b];
                        const float *x = a.in[in.a];
                        float s = 0.0f;
                        for (int d = 0; d < o.nd; ++d) {
                            const int xx = clampi(c + o.ox[d], 0, a.W - 1);
                            const int yy = clampi(r + o.oy[d], 0, a.H - 1);
                            s = fmaf(o.w[d], x[(int64_t)yy * a.W + xx], s);
                        }
                        v = s;
                        break;
                    }
                    case HMX_OP_DIA: {
                        const ProgOperand &o = a.opnd[in.b];
                        const float *x = a.in[in.a];
                        const int64_t HW = (int64_t)a.H * a.W;
                        float s = 0.0f;
                        for (int d = 0; d < o.nd; ++d) {
                            const int xx = c + o.ox[d], yy = r + o.oy[d];
                            if (xx >= 0 && xx < a.W && yy >= 0 && yy < a.H) {
                                const int64_t q = (int64_t)yy * a.W + xx;
                                s = fmaf(o.w[d * HW + q], x[q], s);
                            }
                        }
                        v = s;
                        break;
                    }
                    case HMX_OP_ADD: v = reg[in.a][tid] + reg[in.b][tid]; break;
                    case HMX_OP_SUB: v = reg[in.a][tid] - reg[in.b][tid]; break;
                    case HMX_OP_MUL: v = reg[in.a][tid] * reg[in.b][tid]; break;
                    case HMX_OP_DIV: v = reg[in.a][tid] / reg[in.b][tid]; break;
                    case HMX_OP_SQRT: v = sqrtf(reg[in.a][tid]); break;
                    case HMX_OP_RED: v = ((c + r) & 1) ? 0.0f : reg[in.a][tid]; break;
                    case HMX_OP_BLACK: v = ((c + r) & 1) ? reg[in.a][tid] : 0.0f; break;
                    case HMX_OP_STORE: a.out[in.dst][e] = reg[in.a][tid]; continue;
                    case HMX_OP_REDUCE:
                        if (in.dst == 0) acc[0] += (double)reg[in.a][tid];
                        else acc[1] += (double)reg[in.a][tid];
                        continue;
                    default: continue;
                }
                reg[in.dst][tid] = v;
            }
        }
    }
    // deterministic per-block partials of each reduction slot
    for (int k = 0; k < a.nred; ++k) {
        double v = acc[k == 0 ? 0 : 1];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        __syncthreads();
        if ((tid & 63) == 0) red_lds[tid >> 6] = v;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
            for (int i = 0; i < PROG_NT / 64; ++i) t += red_lds[i];
            a.red_part[k][blockIdx.y * gridDim.x + blockIdx.x] = t;
        }
    }
}

int program_grid_blocks(int H, int W) {
    const int gx = (W + PROG_NT - 1) / PROG_NT;
    const int gy = H < 4096 ? H : 4096;
    return gx * gy;
}

int launch_program(const ProgArgs &a, hipStream_t s, int *nblocks_out) {
    dim3 grid((a.W + PROG_NT - 1) / PROG_NT, a.H < 4096 ? a.H : 4096);
    if (nblocks_out) *nblocks_out = grid.x * grid.y;
    hipLaunchKernelGGL(program_kernel, grid, dim3(PROG_NT), 0, s, a);
    HMX_LAUNCH_CHECK("program");
    return HMX_OK;
}

__global__ __launch_bounds__(256) void reduce_finalize_kernel(const double *__restrict__ part,
                                                              int np, void *dst, int dt) {
    __shared__ double lds[4];
    double v = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) v += part[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double t = lds[0] + lds[1] + lds[2] + lds[3];
        if (dt == HMX_F64) *(double *)dst = t;
        else *(float *)dst = (float)t;
    }
}

int launch_reduce_finalize(const double *part, int np, void *dst, int dt, hipStream_t s) {
    hipLaunchKernelGGL(reduce_finalize_kernel, dim3(1), dim3(256), 0, s, part, np, dst, dt);
    HMX_LAUNCH_CHECK("reduce_finalize");
    return HMX_OK;
}

__device__ inline double sload(const void *p, int dt) {
    switch (dt) {
        case HMX_F64: return *(const double *)p;
        case HMX_I64: return (double)*(const long long *)p;
        case HMX_I32: return (double)*(const int *)p;
        default: return (double)*(const float *)p;
    }
}

// ScalarOp.cpp:31: "*a = *b op *c" in OpenCL C: float op float stays float,
// anything with a double operand is evaluated in double; the result converts
// to the destination type.
__global__ void scalar_program_kernel(ScalarProg p) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int i = 0; i < p.n; ++i) {
        const ScalarInstr &in = p.ins[i];
        const bool wide = in.dt_a == HMX_F64 || (in.b && in.dt_b == HMX_F64) ||
                          in.dt_dst == HMX_F64 || in.dt_a == HMX_I64;
        double r;
        if (wide) {
            const double x = sload(in.a, in.dt_a);
            const double y = in.b ? sload(in.b, in.dt_b) : 0.0;
            switch (in.op) {
                case HMX_OP_ADD: r = x + y; break;
                case HMX_OP_SUB: r = x - y; break;
                case HMX_OP_MUL: r = x * y; break;
                case HMX_OP_DIV: r = x / y; break;
                case HMX_OP_SQRT: r = sqrt(x); break;
                default: r = x; break;
            }
        } else {
            const float x = (float)sload(in.a, in.dt_a);
            const float y = in.b ? (float)sload(in.b, in.dt_b) : 0.0f;
            float f;
            switch (in.op) {
                case HMX_OP_ADD: f = x + y; break;
                case HMX_OP_SUB: f = x - y; break;
                case HMX_OP_MUL: f = x * y; break;
                case HMX_OP_DIV: f = x / y; break;
                case HMX_OP_SQRT: f = sqrtf(x); break;
                default: f = x; break;
            }
            r = f;
        }
        switch (in.dt_dst) {
            case HMX_F64: *(double *)in.dst = r; break;
            case HMX_I64: *(long long *)in.dst = (long long)r; break;
            case HMX_I32: *(int *)in.dst = (int)r; break;
            default: *(float *)in.dst = (float)r; break;
        }
    }
}

int launch_scalar_program(const ScalarProg &p, hipStream_t s) {
    if (p.n == 0) return HMX_OK;
    hipLaunchKernelGGL(scalar_program_kernel, dim3(1), dim3(64), 0, s, p);
    HMX_LAUNCH_CHECK("scalar_program");
    return HMX_OK;
}

}  // namespace hmx
