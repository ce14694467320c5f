// Fused elementwise programs: the generic kernel behind Array2DOp,
// Array2DScalarOp, ScalarArray2DOp, Array2DCopy, Array2DSqrt, Array2DSum and
// the Stencil2DMVM / SpmvDIA2D taps when they fuse with them (the reference's
// fusion pass, fusion.cpp:113-303, fused those into one OpenCL kernel per
// block; the planner here builds one program per fusible run of statements).
//
// A program is a straight-line list of hmx_instr over "registers".  Each
// thread evaluates it for one element at a time; the register file lives in
// LDS (reg[r][threadIdx.x]) so the dynamic register index costs an address
// computation instead of a private-memory spill.
#pragma once

#include "hmx_common.h"

#include <string>
#include <vector>

namespace hmx {

constexpr int PROG_MAX_INSTR = 64;
constexpr int PROG_MAX_IN = 16;
constexpr int PROG_MAX_OUT = 8;
constexpr int PROG_MAX_SCAL = 8;
constexpr int PROG_MAX_OPND = 4;   // stencil / DIA operands
constexpr int PROG_MAX_RED = 2;
constexpr int PROG_MAX_REG = 24;
constexpr int PROG_NT = 256;

// A stencil (kind 0) or 2D DIA matrix (kind 1) operand, all in device memory.
struct ProgOperand {
    const float *w;   // stencil weights [nd] or DIA planes [nd][H][W]
    const int *ox, *oy;
    int nd, kind;
};

struct ProgArgs {
    int ninstr;
    int nred;
    int H, W;                       // element (r, c) -> r * W + c (1D arrays: H = 1)
    hmx_instr ins[PROG_MAX_INSTR];
    const float *in[PROG_MAX_IN];
    float *out[PROG_MAX_OUT];
    const void *scal[PROG_MAX_SCAL];
    int scal_dt[PROG_MAX_SCAL];
    ProgOperand opnd[PROG_MAX_OPND];
    double *red_part[PROG_MAX_RED]; // per-block partial sums
};

// Scalar statements (ScalarOp / ScalarCopy / ScalarSqrt, ScalarOp.cpp:31-73)
// run by a single thread on device-resident scalars.
constexpr int SPROG_MAX = 32;
struct ScalarInstr {
    int op;          // HMX_OP_ADD.. HMX_OP_SQRT, or 0 = copy
    void *dst;
    const void *a, *b;
    int dt_dst, dt_a, dt_b;
};
struct ScalarProg {
    int n;
    ScalarInstr ins[SPROG_MAX];
};

int launch_program(const ProgArgs &a, hipStream_t s, int *nblocks_out);

// What a generated kernel is specialised on: the instruction list, the slot
// counts, each scalar slot's dtype and each stencil / DIA operand's tap count
// (fixed when the record is compiled; weights and offsets stay device data).
struct ProgSpec {
    std::vector<hmx_instr> code;
    int H = 0;  // rows the launch covers (0: unknown)
    int nin = 0, nout = 0, nscal = 0, nopnd = 0, nred = 0;
    int scal_dt[PROG_MAX_SCAL] = {};
    int opnd_nd[PROG_MAX_OPND] = {};
};
// Generated kernels (progjit.hip) unless HMX_PROGRAM_INTERP is set (then the
// interpreter, program_kernel, runs every program).
bool program_jit_enabled();
// HIP source of the generated kernel for a program.
int program_source(const ProgSpec &sp, std::string &src, std::string &err);
// The compiled kernel for a program on a device (compiled with hipRTC on
// first use, then cached for the life of the process).
int program_jit_get(int device, const ProgSpec &sp, void **fn, std::string &err);
// Compile generated source with hipRTC without loading it (no GPU needed).
int program_jit_compile_only(const std::string &src, std::string &err);
// Launch a generated kernel with the interpreter's arguments and grid.
int launch_program_jit(void *fn, const ProgArgs &a, hipStream_t s);
int program_grid_blocks(int H, int W);
int launch_reduce_finalize(const double *part, int np, void *dst, int dt, hipStream_t s);
int launch_scalar_program(const ScalarProg &p, hipStream_t s);

}  // namespace hmx
