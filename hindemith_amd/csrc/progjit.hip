// Generated kernels for fused elementwise programs.
//
// The reference turns each fused group of statements into OpenCL C at
// xmlcompile time and builds it with the vendor compiler (oclgen.cpp:95-154,
// clhelp.cpp:160-200).  The equivalent here: a group's instruction list
// (program.h) becomes straight-line HIP source -- one local variable per
// program register, the stencil / DIA tap loops unrolled to their compiled
// tap count, scalar operands read once per thread -- compiled with hipRTC for
// gfx950 when the record is compiled, and launched in place of the
// interpreter (program_kernel).
//
// The generated kernel computes exactly what the interpreter computes: the
// same grid, the same element order per thread, the same fp32 operation per
// instruction (no contraction, correctly rounded division and sqrt), the same
// fma chain per stencil / DIA operand and the same reduction tree, so the two
// are bit-identical.  Compiled code is cached per (device, source) for the
// life of the process.
#include "program.h"

#include <hip/hiprtc.h>

#include <cstdlib>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

namespace hmx {

namespace {

// The argument block of a generated kernel.  kJitArgsSrc declares the same
// layout in the generated source (the sizes are spelled out there, and the
// static_asserts below tie them to the host constants).
const char *const kJitArgsSrc = R"(
struct ProgOperand { const float *w; const int *ox; const int *oy; int nd; int kind; };
struct JitArgs {
    int H, W;
    const float *in[16];
    float *out[8];
    const void *scal[8];
    ProgOperand opnd[4];
    double *red_part[2];
};
)";

struct JitArgs {
    int H, W;
    const float *in[PROG_MAX_IN];
    float *out[PROG_MAX_OUT];
    const void *scal[PROG_MAX_SCAL];
    ProgOperand opnd[PROG_MAX_OPND];
    double *red_part[PROG_MAX_RED];
};
static_assert(PROG_MAX_IN == 16 && PROG_MAX_OUT == 8 && PROG_MAX_SCAL == 8 &&
                  PROG_MAX_OPND == 4 && PROG_MAX_RED == 2,
              "kJitArgsSrc spells these sizes out");
static_assert(sizeof(ProgOperand) == 32, "ProgOperand layout");

std::mutex g_jit_mu;
// A compiled program: its kernel and the image rows each thread handles.
struct JitKernel {
    hipFunction_t f;
    int rows;
};
// (device, source) -> kernel; entries live for the life of the process
std::map<std::pair<int, std::string>, JitKernel *> g_jit_cache;

std::string scalar_expr(int slot, int dt) {
    const std::string p = "a.scal[" + std::to_string(slot) + "]";
    switch (dt) {
        case HMX_F64: return "(float)*(const double *)" + p;
        case HMX_I64: return "(float)*(const long long *)" + p;
        case HMX_I32: return "(float)*(const int *)" + p;
        default: return "*(const float *)" + p;
    }
}

// Rows of the image one thread handles.  A program without a reduction takes
// kRowsPerThread consecutive rows, every row's loads issued before any store
// (the fused group never reads through a stencil an array it writes, and an
// element it both reads and writes is read first), which amortises the
// per-wave set-up and keeps more loads in flight; a program with a reduction
// keeps one row per thread, the interpreter's partition, so its partial sums
// are the interpreter's bit for bit.
constexpr int kRowsPerThread = 4;
int rows_per_thread(const ProgSpec &sp) { return sp.nred > 0 ? 1 : kRowsPerThread; }

// Emit the body of one row (index i of the thread's rows): loads and
// arithmetic into q<reg>_<i>, stores into st<k>_<i> (written out later).
bool emit_row(std::ostringstream &o, const ProgSpec &sp, int nreg, int i, std::string &err) {
    const std::string sfx = "_" + std::to_string(i);
    const std::string r = "y" + std::to_string(i);
    const std::string row = "row" + std::to_string(i);
    o << "            const long long " << row << " = (long long)" << r << " * a.W;\n";
    for (int k = 0; k < nreg; ++k) o << "            float q" << k << sfx << " = 0.0f;\n";
    int nst = 0;
    for (const hmx_instr &in : sp.code) {
        const std::string d = "q" + std::to_string(in.dst) + sfx;
        const std::string A = "q" + std::to_string(in.a) + sfx;
        const std::string B = "q" + std::to_string(in.b) + sfx;
        o << "            ";
        switch (in.op) {
            case HMX_OP_LOAD:
                o << d << " = *(const float *)((const char *)(a.in[" << in.a << "] + " << row
                  << ") + cb);\n";
                break;
            case HMX_OP_LOADS:
                if (in.a < 0 || in.a >= sp.nscal) { err = "scalar slot out of range"; return false; }
                o << d << " = sc" << in.a << ";\n";
                break;
            case HMX_OP_STENCIL:
            case HMX_OP_DIA: {
                if (in.b < 0 || in.b >= sp.nopnd) { err = "operand slot out of range"; return false; }
                const int nd = sp.opnd_nd[in.b];
                o << "{\n"
                  << "                const float *x = a.in[" << in.a << "];\n"
                  << "                float s = 0.0f;\n";
                for (int k = 0; k < nd; ++k) {
                    const std::string t = std::to_string(in.b) + "_" + std::to_string(k);
                    if (in.op == HMX_OP_STENCIL) {
                        // clamped taps in tap order (Stencil2DMVM.cpp:37-47)
                        o << "                s = fmaf(w" << t
                          << ", *(const float *)((const char *)(x + (long long)min(max(" << r
                          << " + oy" << t << ", 0), a.H - 1) * a.W) + (unsigned)min(max(c + ox"
                          << t << ", 0), a.W - 1) * 4u), s);\n";
                    } else {
                        // coefficient at the neighbour, out-of-range taps
                        // skipped (SpmvDIA2D.cpp:41-43, SpMat2D.cpp:214)
                        o << "                {\n"
                          << "                    const int xx = c + ox" << t << ", yy = " << r
                          << " + oy" << t << ";\n"
                          << "                    if (xx >= 0 && xx < a.W && yy >= 0 && yy < a.H) {\n"
                          << "                        const long long q = (long long)yy * a.W + xx;\n"
                          << "                        s = fmaf(a.opnd[" << in.b << "].w[" << k
                          << " * ((long long)a.H * a.W) + q], x[q], s);\n"
                          << "                    }\n"
                          << "                }\n";
                    }
                }
                o << "                " << d << " = s;\n            }\n";
                break;
            }
            case HMX_OP_ADD: o << d << " = " << A << " + " << B << ";\n"; break;
            case HMX_OP_SUB: o << d << " = " << A << " - " << B << ";\n"; break;
            case HMX_OP_MUL: o << d << " = " << A << " * " << B << ";\n"; break;
            case HMX_OP_DIV: o << d << " = " << A << " / " << B << ";\n"; break;
            case HMX_OP_SQRT: o << d << " = sqrtf(" << A << ");\n"; break;
            case HMX_OP_RED: o << d << " = ((c + " << r << ") & 1) ? 0.0f : " << A << ";\n"; break;
            case HMX_OP_BLACK: o << d << " = ((c + " << r << ") & 1) ? " << A << " : 0.0f;\n"; break;
            case HMX_OP_STORE: o << "st" << nst++ << sfx << " = " << A << ";\n"; break;
            case HMX_OP_REDUCE:
                o << (in.dst == 0 ? "acc0" : "acc1") << " += (double)" << A << ";\n";
                break;
            default: err = "unknown program instruction " + std::to_string(in.op); return false;
        }
    }
    return true;
}

// HIP source of one program, or "" with err set.
std::string generate(const ProgSpec &sp, std::string &err) {
    int nreg = 0;
    std::vector<int> store_slot;  // output slot of each store, in program order
    for (const hmx_instr &in : sp.code) {
        if (in.op == HMX_OP_STORE) store_slot.push_back(in.dst);
        else if (in.op != HMX_OP_REDUCE) nreg = std::max(nreg, in.dst + 1);
    }
    const int R = rows_per_thread(sp);
    const long long gy_rows = (sp.H + R - 1) / R;  // block rows the grid needs
    std::ostringstream o;
    o << "#pragma clang fp contract(off)\n" << kJitArgsSrc;
    o << "extern \"C\" __global__ __launch_bounds__(" << PROG_NT
      << ") void hmx_prog(JitArgs a) {\n"
      << "    __shared__ double red_lds[" << PROG_NT / 64 << "];\n"
      << "    const int tid = threadIdx.x;\n"
      << "    const int c = blockIdx.x * " << PROG_NT << " + tid;\n"
      << "    double acc0 = 0.0, acc1 = 0.0;\n";
    for (int k = 0; k < sp.nscal; ++k)
        o << "    const float sc" << k << " = " << scalar_expr(k, sp.scal_dt[k]) << ";\n";
    // tap offsets (and stencil weights) of every operand the program uses,
    // read before the first store so they stay uniform scalar loads
    std::vector<int> used_as(sp.nopnd, 0);  // HMX_OP_STENCIL or HMX_OP_DIA
    for (const hmx_instr &in : sp.code)
        if ((in.op == HMX_OP_STENCIL || in.op == HMX_OP_DIA) && in.b >= 0 && in.b < sp.nopnd)
            used_as[in.b] = in.op;
    for (int b = 0; b < sp.nopnd; ++b) {
        if (!used_as[b]) continue;
        for (int d = 0; d < sp.opnd_nd[b]; ++d) {
            const std::string t = std::to_string(b) + "_" + std::to_string(d);
            o << "    const int ox" << t << " = a.opnd[" << b << "].ox[" << d << "], oy" << t
              << " = a.opnd[" << b << "].oy[" << d << "];\n";
            if (used_as[b] == HMX_OP_STENCIL)
                o << "    const float w" << t << " = a.opnd[" << b << "].w[" << d << "];\n";
        }
    }
    o << "    if (c < a.W) {\n"
      << "        // element (y, c): a uniform row offset plus a 32-bit byte offset\n"
      << "        const unsigned cb = (unsigned)c * 4u;\n";
    // the grid has one block row per R image rows, up to 4096 block rows;
    // when that covers the image the row loop runs once and is left out
    if (sp.H > 0 && gy_rows <= 4096)
        o << "        {\n"
          << "            const int ybase = blockIdx.y * " << R << ";\n";
    else
        o << "        for (int ybase = blockIdx.y * " << R << "; ybase < a.H; ybase += gridDim.y * "
          << R << ") {\n";
    for (int i = 0; i < R; ++i)
        for (size_t k = 0; k < store_slot.size(); ++k)
            o << "            float st" << k << "_" << i << " = 0.0f;\n";
    // every row is computed without a branch (a row past the image, in the
    // last block row, is computed on the image's last row and not stored), so
    // the compiler can issue all rows' loads before the first arithmetic; with
    // one row per thread the row is always inside the image
    for (int i = 0; i < R; ++i) {
        o << "            const int y" << i << " = "
          << (R == 1 ? std::string("ybase") : "min(ybase + " + std::to_string(i) + ", a.H - 1)")
          << ";\n";
        if (!emit_row(o, sp, nreg, i, err)) return "";
    }
    // outputs are written through the L2 (system-scope stores) when the rows
    // fill whole 128-byte lines, as the library's own kernels store theirs
    if (!store_slot.empty()) o << "            const bool wt = (a.W & 31) == 0;\n";
    for (int i = 0; i < R; ++i) {
        if (store_slot.empty()) break;
        o << "            if (ybase + " << i << " < a.H) {\n";
        for (size_t k = 0; k < store_slot.size(); ++k) {
            const std::string ptr = "(float *)((char *)(a.out[" + std::to_string(store_slot[k]) +
                                    "] + (long long)y" + std::to_string(i) + " * a.W) + cb)";
            const std::string val = "st" + std::to_string(k) + "_" + std::to_string(i);
            o << "                if (wt) __hip_atomic_store(" << ptr << ", " << val
              << ", __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);\n"
              << "                else *" << ptr << " = " << val << ";\n";
        }
        o << "            }\n";
    }
    o << "        }\n    }\n";
    // per-block partials of each reduction slot, in the interpreter's order
    for (int k = 0; k < sp.nred; ++k) {
        o << "    {\n"
          << "        double v = " << (k == 0 ? "acc0" : "acc1") << ";\n"
          << "#pragma unroll\n"
          << "        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);\n"
          << "        __syncthreads();\n"
          << "        if ((tid & 63) == 0) red_lds[tid >> 6] = v;\n"
          << "        __syncthreads();\n"
          << "        if (tid == 0) {\n"
          << "            double t = 0.0;\n"
          << "            for (int i = 0; i < " << PROG_NT / 64 << "; ++i) t += red_lds[i];\n"
          << "            a.red_part[" << k << "][blockIdx.y * gridDim.x + blockIdx.x] = t;\n"
          << "        }\n"
          << "    }\n";
    }
    o << "}\n";
    return o.str();
}

int compile_source(const std::string &src, std::vector<char> &code, std::string &err) {
    hiprtcProgram p;
    if (hiprtcCreateProgram(&p, src.c_str(), "hmx_prog.hip", 0, nullptr, nullptr) !=
        HIPRTC_SUCCESS) {
        err = "hiprtcCreateProgram failed";
        return HMX_ERR_HIP;
    }
    // the flags of the library's own build: fp32 ops round one by one,
    // division and sqrt are correctly rounded, no packed fp32 ops
    const char *opts[] = {"--offload-arch=gfx950",
                          "-O3",
                          "-ffp-contract=off",
                          "-fhip-fp32-correctly-rounded-divide-sqrt",
                          "-Xclang",
                          "-target-feature",
                          "-Xclang",
                          "-packed-fp32-ops"};
    const hiprtcResult r = hiprtcCompileProgram(p, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    int rc = HMX_OK;
    if (r != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(p, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(p, &log[0]);
        err = std::string("hipRTC: ") + hiprtcGetErrorString(r) + "\n" + log;
        rc = HMX_ERR_HIP;
    } else {
        size_t n = 0;
        if (hiprtcGetCodeSize(p, &n) != HIPRTC_SUCCESS || n == 0) {
            err = "hipRTC returned no code";
            rc = HMX_ERR_HIP;
        } else {
            code.resize(n);
            if (hiprtcGetCode(p, code.data()) != HIPRTC_SUCCESS) {
                err = "hiprtcGetCode failed";
                rc = HMX_ERR_HIP;
            }
        }
    }
    hiprtcDestroyProgram(&p);
    return rc;
}

}  // namespace

bool program_jit_enabled() {
    const char *e = std::getenv("HMX_PROGRAM_INTERP");
    return !(e && e[0] && e[0] != '0');
}

int program_source(const ProgSpec &sp, std::string &src, std::string &err) {
    if ((int)sp.code.size() > PROG_MAX_INSTR || sp.nin > PROG_MAX_IN || sp.nout > PROG_MAX_OUT ||
        sp.nscal > PROG_MAX_SCAL || sp.nopnd > PROG_MAX_OPND || sp.nred > PROG_MAX_RED) {
        err = "program exceeds the fused-program limits";
        return HMX_ERR_INVALID;
    }
    src = generate(sp, err);
    return src.empty() ? HMX_ERR_INVALID : HMX_OK;
}

int program_jit_compile_only(const std::string &src, std::string &err) {
    std::vector<char> code;
    return compile_source(src, code, err);
}

int program_jit_get(int device, const ProgSpec &sp, void **fn, std::string &err) {
    std::string src;
    int rc = program_source(sp, src, err);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_jit_mu);
    auto key = std::make_pair(device, src);
    auto it = g_jit_cache.find(key);
    if (it != g_jit_cache.end()) {
        *fn = it->second;
        return HMX_OK;
    }
    std::vector<char> code;
    rc = compile_source(src, code, err);
    if (rc) return rc;
    // the module stays loaded for the life of the process (cached)
    hipModule_t mod;
    hipError_t e = hipModuleLoadData(&mod, code.data());
    if (e != hipSuccess) {
        err = std::string("hipModuleLoadData: ") + hipGetErrorString(e);
        return HMX_ERR_HIP;
    }
    hipFunction_t f;
    e = hipModuleGetFunction(&f, mod, "hmx_prog");
    if (e != hipSuccess) {
        err = std::string("hipModuleGetFunction: ") + hipGetErrorString(e);
        return HMX_ERR_HIP;
    }
    JitKernel *jk = new JitKernel{f, rows_per_thread(sp)};
    g_jit_cache.emplace(key, jk);
    *fn = jk;
    return HMX_OK;
}

int launch_program_jit(void *fn, const ProgArgs &a, hipStream_t s) {
    JitArgs j;
    j.H = a.H;
    j.W = a.W;
    for (int i = 0; i < PROG_MAX_IN; ++i) j.in[i] = a.in[i];
    for (int i = 0; i < PROG_MAX_OUT; ++i) j.out[i] = a.out[i];
    for (int i = 0; i < PROG_MAX_SCAL; ++i) j.scal[i] = a.scal[i];
    for (int i = 0; i < PROG_MAX_OPND; ++i) j.opnd[i] = a.opnd[i];
    for (int i = 0; i < PROG_MAX_RED; ++i) j.red_part[i] = a.red_part[i];
    // one block row per `rows` image rows, at most 4096 block rows; with one
    // row per thread (programs with a reduction) this is the interpreter's
    // grid (program_grid_blocks), so the partials have its count and order
    const JitKernel *jk = static_cast<const JitKernel *>(fn);
    const unsigned gx = (unsigned)((a.W + PROG_NT - 1) / PROG_NT);
    const int brows = (a.H + jk->rows - 1) / jk->rows;
    const unsigned gy = (unsigned)(brows < 4096 ? brows : 4096);
    void *params[] = {&j};
    const hipError_t e = hipModuleLaunchKernel(jk->f, gx, gy, 1, PROG_NT, 1, 1, 0, s, params,
                                               nullptr);
    if (e != hipSuccess) {
        set_error(std::string("generated program launch: ") + hipGetErrorString(e));
        return HMX_ERR_HIP;
    }
    return HMX_OK;
}

}  // namespace hmx
