// Record runtime: the C++ replacement for the reference's backend module.
//
//   hmx_compile  <- backend.xmlcompile (backend.cpp:70-159): smparse of the
//                   three XML documents (sm.cpp:245-266), symbol table
//                   (sm.cpp:268-313), then planning instead of
//                   generate_ocl (oclgen.cpp:95-154).
//   hmx_run      <- backend.run (backend.cpp:320-638): upload Header
//                   symbols, execute, download Return values.
//
// Planning replaces the dataflow / fusion / OpenCL code generation pipeline
// (dataflow.cpp, fusion.cpp, lm.cpp) with a fixed set of hand-written HIP
// kernels:
//   * runs of fusible array statements become one elementwise program
//     (program.h) -- register-resident temporaries, one HBM pass;
//   * a stencil that reads an array written earlier in the run starts a new
//     program (read-after-write through a neighbour, the case that made the
//     reference split kernels, lm.cpp:170-205); an array written after a
//     stencil read it in the same program is written to a second buffer and
//     the two are swapped afterwards (Jacobi ping-pong, no copy);
//   * scalar statements become a one-thread scalar program, CSR products a
//     CSR launch;
//   * a loop whose body is exactly the Horn-Schunck Jacobi update is replaced
//     by the fused, temporally blocked HS kernel (k_hs.hip).
#include "program.h"

#include <algorithm>
#include <cctype>
#include <functional>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <sstream>
#include <vector>

namespace hmx {
int launch_csr(int M, int64_t nnz, const int *indptr, const int *indices, const float *data,
               const float *x, float *y, int L, hipStream_t s);
int hs_jacobi_run(const float *Ix, const float *Iy, const float *It, const float *den,
                  float alpha2, int H, int W, const Taps &taps, int iters, float *u, float *v,
                  float *wu, float *wv, int variant, int k, hipStream_t s);
void hs_choice(int H, int W, int &variant, int &k);
int launch_hs_precond(const float *diag0, const float *diag1, const float *offdiag,
                      const float *x0, const float *x1, float *y0, float *y1, int H, int W,
                      const float *w_dev, float w_val, hipStream_t s);
int launch_hs_powers(const float *Ix, const float *Iy, const float *r0, const float *r1,
                     float *V0, float *V1, int H, int W, int steps, const void *lam_dev,
                     int lam_dtype, float lam_val, hipStream_t s);
int launch_warp_img2d(const float *src, const float *offx, const float *offy, float *dst, int H,
                      int W, hipStream_t s);
int launch_brox_set(float *A, float *diag, const float *psi, int H, int W, const void *alpha_dev,
                    int alpha_dt, float alpha_val, int rb, hipStream_t s);
int launch_brox_precond(const float *A2, const float *diag0, const float *diag1,
                        const float *offdiag, const float *x0, const float *x1, float *y0,
                        float *y1, int H, int W, hipStream_t s);

namespace rt {

// ===========================================================================
// Minimal XML reader for the wire format of PyAST.py:61-133: elements with
// attributes and children, no text content.
struct XNode {
    std::string tag;
    std::vector<std::pair<std::string, std::string>> at;
    std::vector<XNode> kids;
    const std::string *get(const char *k) const {
        for (auto &p : at)
            if (p.first == k) return &p.second;
        return nullptr;
    }
};

class XReader {
  public:
    explicit XReader(const char *s) : s_(s), n_(std::strlen(s)) {}
    bool parse(XNode &root, std::string &err) {
        skip_misc();
        if (i_ >= n_ || s_[i_] != '<') {
            err = "XML: expected a root element";
            return false;
        }
        if (!elem(root, err)) return false;
        skip_misc();
        if (i_ != n_) {
            err = "XML: trailing content after the root element";
            return false;
        }
        return true;
    }

  private:
    const char *s_;
    size_t n_, i_ = 0;
    void ws() {
        while (i_ < n_ && std::isspace((unsigned char)s_[i_])) ++i_;
    }
    void skip_misc() {
        for (;;) {
            ws();
            if (i_ + 1 < n_ && s_[i_] == '<' && (s_[i_ + 1] == '?' || s_[i_ + 1] == '!')) {
                const char *end = s_[i_ + 1] == '?' ? "?>" : ">";
                if (std::strncmp(s_ + i_, "<!--", 4) == 0) end = "-->";
                const char *p = std::strstr(s_ + i_, end);
                i_ = p ? (size_t)(p - s_) + std::strlen(end) : n_;
                continue;
            }
            return;
        }
    }
    static std::string unescape(const std::string &v) {
        std::string o;
        for (size_t k = 0; k < v.size(); ++k) {
            if (v[k] == '&') {
                static const char *ent[][2] = {{"&lt;", "<"}, {"&gt;", ">"}, {"&amp;", "&"},
                                               {"&quot;", "\""}, {"&apos;", "'"}};
                bool hit = false;
                for (auto &e : ent) {
                    size_t L = std::strlen(e[0]);
                    if (v.compare(k, L, e[0]) == 0) {
                        o += e[1];
                        k += L - 1;
                        hit = true;
                        break;
                    }
                }
                if (!hit) o += '&';
            } else {
                o += v[k];
            }
        }
        return o;
    }
    bool name(std::string &out) {
        size_t b = i_;
        while (i_ < n_ && (std::isalnum((unsigned char)s_[i_]) || s_[i_] == '_' ||
                           s_[i_] == '-' || s_[i_] == '.' || s_[i_] == ':'))
            ++i_;
        out.assign(s_ + b, i_ - b);
        return !out.empty();
    }
    bool elem(XNode &e, std::string &err) {
        ++i_;  // '<'
        if (!name(e.tag)) {
            err = "XML: bad element name at offset " + std::to_string(i_);
            return false;
        }
        for (;;) {
            ws();
            if (i_ >= n_) {
                err = "XML: unterminated tag <" + e.tag + ">";
                return false;
            }
            if (s_[i_] == '/') {
                if (i_ + 1 < n_ && s_[i_ + 1] == '>') {
                    i_ += 2;
                    return true;
                }
                err = "XML: bad '/' in <" + e.tag + ">";
                return false;
            }
            if (s_[i_] == '>') {
                ++i_;
                break;
            }
            std::string k;
            if (!name(k)) {
                err = "XML: bad attribute in <" + e.tag + ">";
                return false;
            }
            ws();
            if (i_ >= n_ || s_[i_] != '=') {
                err = "XML: expected '=' after " + k;
                return false;
            }
            ++i_;
            ws();
            if (i_ >= n_ || (s_[i_] != '"' && s_[i_] != '\'')) {
                err = "XML: expected a quoted value for " + k;
                return false;
            }
            char q = s_[i_++];
            size_t b = i_;
            while (i_ < n_ && s_[i_] != q) ++i_;
            if (i_ >= n_) {
                err = "XML: unterminated value for " + k;
                return false;
            }
            e.at.push_back({k, unescape(std::string(s_ + b, i_ - b))});
            ++i_;
        }
        for (;;) {
            skip_misc();
            if (i_ + 1 < n_ && s_[i_] == '<' && s_[i_ + 1] == '/') {
                i_ += 2;
                std::string nm;
                name(nm);
                ws();
                if (nm != e.tag || i_ >= n_ || s_[i_] != '>') {
                    err = "XML: mismatched end tag for <" + e.tag + ">";
                    return false;
                }
                ++i_;
                return true;
            }
            if (i_ < n_ && s_[i_] == '<') {
                e.kids.emplace_back();
                if (!elem(e.kids.back(), err)) return false;
                continue;
            }
            // text content is not part of the format: skip it
            while (i_ < n_ && s_[i_] != '<') ++i_;
            if (i_ >= n_) {
                err = "XML: unterminated element <" + e.tag + ">";
                return false;
            }
        }
    }
};

// ===========================================================================
// Symbols (sm.h:17-263 Variable hierarchy, restricted to the hot path types)
struct Sym {
    std::string name;
    int kind = 0;
    int dtype = HMX_F32;
    bool header = false;
    int64_t dim0 = 0, dim1 = 0, dim2 = 0, ndiags = 0, nnz = 0;
    bool col_mjr = false;
    // element grid the kernels see (memory order, Array2D.cpp:297-310)
    int H = 0, W = 0;
    // host binding for the next run
    const void *host[3] = {nullptr, nullptr, nullptr};
    int64_t host_bytes[3] = {0, 0, 0};
    // device storage
    void *dev[3] = {nullptr, nullptr, nullptr};
    int64_t bytes[3] = {0, 0, 0};
    void *alt = nullptr;  // second buffer for renamed arrays
    int cur = 0;          // 0: dev[0] holds the value, 1: alt does
    bool renamed = false;
    bool needs_buffer = false;
    float *cur_buf() const { return (float *)(cur ? alt : dev[0]); }
    float *other_buf() const { return (float *)(cur ? dev[0] : alt); }
};

static int dtype_of(const std::string &s) {
    if (s == "float32") return HMX_F32;
    if (s == "float64") return HMX_F64;
    if (s == "int64") return HMX_I64;
    if (s == "int32") return HMX_I32;
    return -1;
}
static int dtype_size(int dt) { return (dt == HMX_F64 || dt == HMX_I64) ? 8 : 4; }

// Semantic model statements
struct Stmt {
    bool loop = false;
    std::string filter, op;
    std::vector<int> args;  // symbol ids, outputs first (Stencil2DMVM.cpp:64-66)
    int ub = 0;
    std::vector<Stmt> body;
    int index = 0;  // program-order position (loop bodies inline)
};

// Plan nodes
struct Group {
    int H = 0, W = 0;
    int dims_set = 0;
    std::vector<hmx_instr> code;
    std::vector<int> in_syms, out_syms, scal_syms, opnd_syms, red_syms;
    std::map<int, int> in_slot, scal_slot, opnd_slot;
    std::map<int, int> cur_reg;
    std::set<int> written, nb_read, red_out;
    std::map<int, int> first_write;  // sym -> statement index
    int s0 = -1, s1 = -1;
    int nreg = 0;
    int nblocks = 0;
    double *partials = nullptr;
    void *jit = nullptr;  // generated kernel (progjit.hip), or null: the interpreter
};

struct Node {
    enum T {
        GROUP, SCALAR, CSR, LOOP, HS, RETURN, PRECOND, POWERS, COPY3D, WARP, BROX_SET,
        BROX_SET_RB, BROX_PRECOND
    } t;
    int idx = 0;                 // GROUP: groups[], SCALAR: sprogs[]
    int ub = 0;                  // LOOP / HS
    std::vector<Node> body;      // LOOP (and HS fallback)
    std::vector<int> syms;       // CSR: y, A, x ; HS: u, v, Ix, Iy, It, den, Avg ; RETURN ;
                                 // PRECOND/POWERS/COPY3D/WARP/BROX_*: the statement's
                                 // operands
};

struct ScalarStmt {
    int op, dst, a, b;  // symbol ids (b = -1 for unary/copy)
};

struct Record {
    int id = 0, device = 0;
    hipStream_t stream = nullptr;
    std::vector<Sym> syms;
    std::map<std::string, int> by_name;
    std::map<std::string, std::string> defines;
    std::vector<Stmt> prog;
    std::vector<Group> groups;
    std::vector<std::vector<ScalarStmt>> sprogs;
    std::vector<Node> plan;
    std::vector<int> returns;
    std::vector<std::vector<char>> results;
    std::map<int, std::vector<int>> reads;  // sym -> statement indices reading it
    std::set<int> renamed;
    float *hs_work = nullptr;
    int64_t hs_work_elems = 0;
    std::vector<void *> allocs;
    std::mutex mu;
    bool ran = false;
};

static std::mutex g_mu;
static std::map<int, std::unique_ptr<Record>> g_records;
static int g_device = 0;

static Record *find_record(int rec) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_records.find(rec);
    return it == g_records.end() ? nullptr : it->second.get();
}

// ---------------------------------------------------------------------------
// Parsing

static int64_t attr_i(const XNode &n, const char *k, int64_t def = 0) {
    const std::string *v = n.get(k);
    if (!v || v->empty()) return def;
    return std::strtoll(v->c_str(), nullptr, 10);
}

static int parse_symtab(Record &R, const XNode &root) {
    if (root.tag != "SymbolTable") return fail(HMX_ERR_INVALID, "symbol table root must be <SymbolTable>");
    for (const XNode &n : root.kids) {
        Sym s;
        const std::string *nm = n.get("name");
        if (!nm) return fail(HMX_ERR_INVALID, "symbol without a name: <" + n.tag + ">");
        s.name = *nm;
        const std::string *dt = n.get("dtype");
        s.dtype = dt ? dtype_of(*dt) : HMX_F32;
        const std::string *src = n.get("source");
        s.header = src && *src == "Header";
        if (n.tag == "Array2D") {
            s.kind = HMX_KIND_ARRAY2D;
            s.dim0 = attr_i(n, "dim0");
            s.dim1 = attr_i(n, "dim1");
            const std::string *cm = n.get("col_mjr");
            s.col_mjr = cm && *cm == "True";
            // the contiguous index is element_id0 (Stencil2DMVM.cpp:83-92)
            s.W = (int)(s.col_mjr ? s.dim0 : s.dim1);
            s.H = (int)(s.col_mjr ? s.dim1 : s.dim0);
        } else if (n.tag == "Array") {
            s.kind = HMX_KIND_ARRAY;
            s.dim0 = attr_i(n, "dim0");
            s.H = 1;
            s.W = (int)s.dim0;
        } else if (n.tag == "Array3D" && !s.header) {
            // the HSMatrixPowers result: dim0 planes of dim1 x dim2 (Array3D.cpp:329-341)
            s.kind = HMX_KIND_ARRAY3D;
            s.dim0 = attr_i(n, "dim0");
            s.dim1 = attr_i(n, "dim1");
            s.dim2 = attr_i(n, "dim2");
            s.H = (int)s.dim1;
            s.W = (int)s.dim2;
            if (s.dim0 < 1 || s.dim1 < 1 || s.dim2 < 1 || s.dim1 > 1 << 30 || s.dim2 > 1 << 30)
                return fail(HMX_ERR_INVALID, "symbol " + s.name + ": bad dimensions");
        } else if (n.tag == "Scalar") {
            s.kind = HMX_KIND_SCALAR;
        } else if (n.tag == "Stencil") {
            s.kind = HMX_KIND_STENCIL;
            s.ndiags = attr_i(n, "ndiags");
        } else if (n.tag == "SpMat2D") {
            s.kind = HMX_KIND_SPMAT2D;
            s.ndiags = attr_i(n, "ndiags");
            s.dim0 = attr_i(n, "dim0");
            s.dim1 = attr_i(n, "dim1");
        } else if (n.tag == "SpMat") {
            // scipy dia_matrix, square (PyAST.py:25-26, SpMat.cpp:147-159)
            s.kind = HMX_KIND_SPMAT;
            s.ndiags = attr_i(n, "ndiags");
            s.dim0 = attr_i(n, "dim");
            s.H = 1;
            s.W = (int)s.dim0;
            if (s.ndiags < 1 || s.ndiags > HMX_MAX_TAPS || s.dim0 < 1 || s.dim0 > 1 << 30)
                return fail(HMX_ERR_UNSUPPORTED, "dia_matrix " + s.name + " must have 1.." +
                                                     std::to_string(HMX_MAX_TAPS) +
                                                     " diagonals");
        } else if (n.tag == "CSRMat") {
            s.kind = HMX_KIND_CSRMAT;
            s.nnz = attr_i(n, "nnz");
            s.dim0 = attr_i(n, "dim0");
            s.dim1 = attr_i(n, "dim1");
        } else {
            return fail(HMX_ERR_UNSUPPORTED,
                        "symbol type <" + n.tag + "> (" + s.name +
                            ") is not on the Horn-Schunck path and is not built by this backend");
        }
        if (s.dtype < 0)
            return fail(HMX_ERR_UNSUPPORTED, "symbol " + s.name + ": unsupported dtype " +
                                                 (dt ? *dt : std::string("?")));
        if (s.kind != HMX_KIND_SCALAR && s.dtype != HMX_F32)
            return fail(HMX_ERR_UNSUPPORTED, "symbol " + s.name +
                                                 ": array, stencil and matrix data must be "
                                                 "float32 on this backend");
        if ((s.kind == HMX_KIND_ARRAY2D || s.kind == HMX_KIND_ARRAY) &&
            (s.dim0 < 0 || s.dim1 < 0 || s.H > 1 << 30 || s.W > 1 << 30))
            return fail(HMX_ERR_INVALID, "symbol " + s.name + ": bad dimensions");
        if (s.kind == HMX_KIND_STENCIL && (s.ndiags < 1 || s.ndiags > HMX_MAX_TAPS))
            return fail(HMX_ERR_UNSUPPORTED, "stencil " + s.name + " must have 1.." +
                                                 std::to_string(HMX_MAX_TAPS) + " taps");
        if (R.by_name.count(s.name)) return fail(HMX_ERR_INVALID, "duplicate symbol " + s.name);
        R.by_name[s.name] = (int)R.syms.size();
        R.syms.push_back(s);
    }
    return HMX_OK;
}

static const std::set<std::string> kFilters = {
    "Array2DOp", "Array2DScalarOp", "ScalarArray2DOp", "Array2DCopy", "Array2DSqrt",
    "Array2DSum", "Stencil2DMVM", "SpmvDIA2D", "SpmvCSR", "ArrayOp", "ArrayScalarOp",
    "ScalarArrayOp", "ArrayCopy", "ArraySqrt", "ArraySum", "ScalarOp", "ScalarCopy",
    "ScalarSqrt", "Return", "ApplyHSPreconditioner", "HSMatrixPowers", "Array3DCopy",
    "WarpImg2D", "SpmvDIA", "SetBroxMatrix", "SetBroxRedblackMatrix",
    "ApplyBroxPreconditioner", "RedArray2D", "BlackArray2D"};

static int parse_block(Record &R, const XNode &paf, std::vector<Stmt> &out, int &counter) {
    for (const XNode &n : paf.kids) {
        Stmt st;
        if (n.tag == "Iterator") {
            st.loop = true;
            st.ub = (int)attr_i(n, "ub", -1);
            if (st.ub < 0) return fail(HMX_ERR_INVALID, "Iterator without a valid ub");
            st.index = counter++;
            for (const XNode &k : n.kids)
                if (k.tag == "PipeAndFilter") {
                    int rc = parse_block(R, k, st.body, counter);
                    if (rc) return rc;
                }
        } else if (n.tag == "Filter") {
            const std::string *nm = n.get("name");
            if (!nm) return fail(HMX_ERR_INVALID, "Filter without a name");
            st.filter = *nm;
            if (!kFilters.count(st.filter))
                return fail(HMX_ERR_UNSUPPORTED,
                            "operation " + st.filter +
                                " is not on the Horn-Schunck path and is not built by this "
                                "backend");
            const std::string *op = n.get("op");
            if (op) st.op = *op;
            for (const XNode &k : n.kids) {
                const std::string *a = k.get("name");
                if (!a) return fail(HMX_ERR_INVALID, "operand of " + st.filter + " has no name");
                auto it = R.by_name.find(*a);
                if (it == R.by_name.end())
                    return fail(HMX_ERR_INVALID, "operand " + *a + " of " + st.filter +
                                                     " is not in the symbol table");
                st.args.push_back(it->second);
            }
            st.index = counter++;
        } else if (n.tag == "PipeAndFilter") {
            int rc = parse_block(R, n, out, counter);
            if (rc) return rc;
            continue;
        } else {
            return fail(HMX_ERR_UNSUPPORTED, "semantic model node <" + n.tag + "> not supported");
        }
        out.push_back(std::move(st));
    }
    return HMX_OK;
}

static int parse_sm(Record &R, const XNode &root) {
    const XNode *fd = nullptr;
    if (root.tag == "SemanticModel") {
        for (const XNode &k : root.kids)
            if (k.tag == "FunctionDef") fd = &k;
    } else if (root.tag == "FunctionDef") {
        fd = &root;
    }
    if (!fd) return fail(HMX_ERR_INVALID, "semantic model has no FunctionDef");
    int counter = 0;
    for (const XNode &k : fd->kids)
        if (k.tag == "PipeAndFilter") {
            int rc = parse_block(R, k, R.prog, counter);
            if (rc) return rc;
        }
    return HMX_OK;
}

// ---------------------------------------------------------------------------
// Validation of operand kinds (Templates.py dispatch guarantees them; the C
// side re-checks because the XML can come from anywhere).

static bool is_arr(const Sym &s) { return s.kind == HMX_KIND_ARRAY2D || s.kind == HMX_KIND_ARRAY; }

static int check_stmt(Record &R, const Stmt &st) {
    auto K = [&](int i) { return R.syms[st.args[i]].kind; };
    auto need = [&](size_t n) -> int {
        if (st.args.size() != n)
            return fail(HMX_ERR_INVALID, st.filter + " expects " + std::to_string(n) +
                                             " operands, got " + std::to_string(st.args.size()));
        return HMX_OK;
    };
    const std::string &f = st.filter;
    int rc = HMX_OK;
    if (f == "Return") return st.args.empty() ? fail(HMX_ERR_INVALID, "empty Return") : HMX_OK;
    if (f == "Array2DOp" || f == "ArrayOp") {
        if ((rc = need(3))) return rc;
        if (!is_arr(R.syms[st.args[0]]) || !is_arr(R.syms[st.args[1]]) || !is_arr(R.syms[st.args[2]]))
            return fail(HMX_ERR_INVALID, f + " operands must be arrays");
    } else if (f == "Array2DScalarOp" || f == "ArrayScalarOp") {
        if ((rc = need(3))) return rc;
        if (!is_arr(R.syms[st.args[1]]) || K(2) != HMX_KIND_SCALAR)
            return fail(HMX_ERR_INVALID, f + " expects (array, scalar)");
    } else if (f == "ScalarArray2DOp" || f == "ScalarArrayOp") {
        if ((rc = need(3))) return rc;
        if (K(1) != HMX_KIND_SCALAR || !is_arr(R.syms[st.args[2]]))
            return fail(HMX_ERR_INVALID, f + " expects (scalar, array)");
    } else if (f == "Array2DCopy" || f == "ArrayCopy" || f == "Array2DSqrt" || f == "ArraySqrt") {
        if ((rc = need(2))) return rc;
        if (!is_arr(R.syms[st.args[0]]) || !is_arr(R.syms[st.args[1]]))
            return fail(HMX_ERR_INVALID, f + " operands must be arrays");
    } else if (f == "Array2DSum" || f == "ArraySum") {
        if ((rc = need(2))) return rc;
        if (K(0) != HMX_KIND_SCALAR || !is_arr(R.syms[st.args[1]]))
            return fail(HMX_ERR_INVALID, f + " expects (scalar <- array)");
    } else if (f == "Stencil2DMVM") {
        if ((rc = need(3))) return rc;
        if (K(1) != HMX_KIND_STENCIL || K(2) != HMX_KIND_ARRAY2D || K(0) != HMX_KIND_ARRAY2D)
            return fail(HMX_ERR_INVALID, "Stencil2DMVM expects (Array2D <- Stencil, Array2D)");
    } else if (f == "SpmvDIA2D") {
        if ((rc = need(3))) return rc;
        if (K(1) != HMX_KIND_SPMAT2D || K(2) != HMX_KIND_ARRAY2D || K(0) != HMX_KIND_ARRAY2D)
            return fail(HMX_ERR_INVALID, "SpmvDIA2D expects (Array2D <- SpMat2D, Array2D)");
        const Sym &A = R.syms[st.args[1]], &x = R.syms[st.args[2]];
        if (A.dim0 != x.dim0 || A.dim1 != x.dim1)
            return fail(HMX_ERR_INVALID, "SpmvDIA2D: matrix planes and vector differ in shape");
    } else if (f == "SpmvDIA") {
        if ((rc = need(3))) return rc;
        if (K(1) != HMX_KIND_SPMAT || K(2) != HMX_KIND_ARRAY || K(0) != HMX_KIND_ARRAY)
            return fail(HMX_ERR_INVALID, "SpmvDIA expects (Array <- SpMat, Array)");
        const Sym &A = R.syms[st.args[1]], &x = R.syms[st.args[2]];
        if (A.dim0 != x.dim0)
            return fail(HMX_ERR_INVALID, "SpmvDIA: matrix and vector sizes do not match");
    } else if (f == "SpmvCSR") {
        if ((rc = need(3))) return rc;
        if (K(1) != HMX_KIND_CSRMAT || K(2) != HMX_KIND_ARRAY || K(0) != HMX_KIND_ARRAY)
            return fail(HMX_ERR_INVALID, "SpmvCSR expects (Array <- CSRMat, Array)");
        const Sym &A = R.syms[st.args[1]], &x = R.syms[st.args[2]], &y = R.syms[st.args[0]];
        if (A.dim1 != x.dim0 || A.dim0 != y.dim0)
            return fail(HMX_ERR_INVALID, "SpmvCSR: matrix and vector sizes do not match");
    } else if (f == "ApplyHSPreconditioner") {
        // (y0, y1 <- D, diag0, diag1, offdiag, x0, x1), Templates.py:412-430
        if ((rc = need(8))) return rc;
        if (K(2) != HMX_KIND_STENCIL)
            return fail(HMX_ERR_INVALID, "ApplyHSPreconditioner: the first argument must be a "
                                         "Stencil");
        if (R.syms[st.args[2]].ndiags < 3)
            return fail(HMX_ERR_INVALID, "ApplyHSPreconditioner: the stencil needs a weight 2 "
                                         "(at least 3 taps)");
        const Sym &x = R.syms[st.args[6]];
        for (int i : {0, 1, 3, 4, 5, 6, 7}) {
            const Sym &a = R.syms[st.args[i]];
            if (a.kind != HMX_KIND_ARRAY2D)
                return fail(HMX_ERR_INVALID, "ApplyHSPreconditioner: " + a.name +
                                                 " must be an Array2D");
            if (a.H != x.H || a.W != x.W)
                return fail(HMX_ERR_INVALID, "ApplyHSPreconditioner: " + a.name +
                                                 " differs in shape from " + x.name);
        }
    } else if (f == "HSMatrixPowers") {
        // (V0, V1 <- lam, Ix, Iy, r0, r1), Templates.py:277-300
        if ((rc = need(7))) return rc;
        if (K(0) != HMX_KIND_ARRAY3D || K(1) != HMX_KIND_ARRAY3D || K(2) != HMX_KIND_SCALAR)
            return fail(HMX_ERR_INVALID, "HSMatrixPowers expects (Array3D, Array3D <- Scalar, "
                                         "Array2D x 4)");
        const Sym &r = R.syms[st.args[5]], &V = R.syms[st.args[0]], &V1 = R.syms[st.args[1]];
        for (int i : {3, 4, 5, 6}) {
            const Sym &a = R.syms[st.args[i]];
            if (a.kind != HMX_KIND_ARRAY2D || a.H != r.H || a.W != r.W)
                return fail(HMX_ERR_INVALID, "HSMatrixPowers: " + a.name +
                                                 " must be an Array2D shaped like " + r.name);
        }
        if (V.H != r.H || V.W != r.W || V.dim0 != V1.dim0 || V1.H != r.H || V1.W != r.W)
            return fail(HMX_ERR_INVALID, "HSMatrixPowers: result planes differ from " + r.name);
        auto it = R.defines.find("STEPS");
        if (it == R.defines.end())
            return fail(HMX_ERR_INVALID, "HSMatrixPowers needs defines['STEPS'] "
                                         "(HSMatrixPowers.cpp:153)");
        if (std::strtoll(it->second.c_str(), nullptr, 10) != V.dim0)
            return fail(HMX_ERR_INVALID, "HSMatrixPowers: result has " + std::to_string(V.dim0) +
                                             " planes, defines['STEPS'] is " + it->second);
    } else if (f == "SetBroxMatrix" || f == "SetBroxRedblackMatrix") {
        // (A, psi, alpha) / (Adiag <- A, psi, alpha), Templates.py:311-350
        const bool rb = f == "SetBroxRedblackMatrix";
        if ((rc = need(rb ? 4 : 3))) return rc;
        const int o = rb ? 1 : 0;
        const Sym &A = R.syms[st.args[o]], &psi = R.syms[st.args[o + 1]];
        if (A.kind != HMX_KIND_SPMAT2D || psi.kind != HMX_KIND_ARRAY2D ||
            R.syms[st.args[o + 2]].kind != HMX_KIND_SCALAR)
            return fail(HMX_ERR_INVALID, f + " expects (SpMat2D, Array2D, Scalar)");
        if (A.ndiags < (rb ? 4 : 5))
            return fail(HMX_ERR_INVALID, f + ": " + A.name + " needs " +
                                             std::to_string(rb ? 4 : 5) + " diagonal planes");
        if (A.dim0 * A.dim1 != (int64_t)psi.H * psi.W || psi.col_mjr)
            return fail(HMX_ERR_INVALID, f + ": " + A.name + " and " + psi.name +
                                             " differ in shape (or psi is column-major)");
        if (rb && (R.syms[st.args[0]].H != psi.H || R.syms[st.args[0]].W != psi.W))
            return fail(HMX_ERR_INVALID, f + ": the diagonal must be shaped like psi");
    } else if (f == "ApplyBroxPreconditioner") {
        // (y0, y1 <- A, diag0, diag1, offdiag, x0, x1), Templates.py:375-400
        if ((rc = need(8))) return rc;
        const Sym &A = R.syms[st.args[2]], &x = R.syms[st.args[6]];
        if (A.kind != HMX_KIND_SPMAT2D || A.ndiags < 3)
            return fail(HMX_ERR_INVALID, "ApplyBroxPreconditioner: the first argument must be "
                                         "an SpMat2D with a diagonal plane 2");
        if (A.dim0 * A.dim1 != (int64_t)x.H * x.W)
            return fail(HMX_ERR_INVALID, "ApplyBroxPreconditioner: matrix and arrays differ");
        for (int i : {0, 1, 3, 4, 5, 6, 7}) {
            const Sym &a = R.syms[st.args[i]];
            if (a.kind != HMX_KIND_ARRAY2D || a.H != x.H || a.W != x.W)
                return fail(HMX_ERR_INVALID, "ApplyBroxPreconditioner: " + a.name +
                                                 " must be an Array2D shaped like " + x.name);
        }
    } else if (f == "RedArray2D" || f == "BlackArray2D") {
        if ((rc = need(2))) return rc;
        if (!is_arr(R.syms[st.args[0]]) || !is_arr(R.syms[st.args[1]]) ||
            R.syms[st.args[1]].kind != HMX_KIND_ARRAY2D)
            return fail(HMX_ERR_INVALID, f + " operands must be Array2Ds");
    } else if (f == "WarpImg2D") {
        // (dst <- src, offx, offy), Templates.py:257-274
        if ((rc = need(4))) return rc;
        const Sym &x = R.syms[st.args[1]];
        for (int i = 0; i < 4; ++i) {
            const Sym &a = R.syms[st.args[i]];
            if (a.kind != HMX_KIND_ARRAY2D || a.H != x.H || a.W != x.W || a.col_mjr != x.col_mjr)
                return fail(HMX_ERR_INVALID, "WarpImg2D: " + a.name +
                                                 " must be an Array2D shaped like " + x.name);
        }
    } else if (f == "Array3DCopy") {
        if ((rc = need(2))) return rc;
        const Sym &a = R.syms[st.args[0]], &b = R.syms[st.args[1]];
        if (a.kind != HMX_KIND_ARRAY3D || b.kind != HMX_KIND_ARRAY3D || a.dim0 != b.dim0 ||
            a.H != b.H || a.W != b.W)
            return fail(HMX_ERR_INVALID, "Array3DCopy operands must be Array3Ds of one shape");
    } else if (f == "ScalarOp") {
        if ((rc = need(3))) return rc;
        if (K(0) != HMX_KIND_SCALAR || K(1) != HMX_KIND_SCALAR || K(2) != HMX_KIND_SCALAR)
            return fail(HMX_ERR_INVALID, "ScalarOp operands must be scalars");
    } else if (f == "ScalarCopy" || f == "ScalarSqrt") {
        if ((rc = need(2))) return rc;
        if (K(0) != HMX_KIND_SCALAR || K(1) != HMX_KIND_SCALAR)
            return fail(HMX_ERR_INVALID, f + " operands must be scalars");
    }
    if (!st.op.empty() && st.op != "Add" && st.op != "Sub" && st.op != "Mult" && st.op != "Div")
        return fail(HMX_ERR_UNSUPPORTED, f + ": operator " + st.op + " not supported");
    return HMX_OK;
}

static void collect_reads(Record &R, const std::vector<Stmt> &b) {
    for (const Stmt &st : b) {
        if (st.loop) {
            collect_reads(R, st.body);
            continue;
        }
        size_t first = st.filter == "Return"
                           ? 0
                           : ((st.filter == "ApplyHSPreconditioner" ||
                               st.filter == "HSMatrixPowers" ||
                               st.filter == "ApplyBroxPreconditioner") ? 2 : 1);
        for (size_t i = first; i < st.args.size(); ++i) R.reads[st.args[i]].push_back(st.index);
    }
}

// ---------------------------------------------------------------------------
// Planner

static int op_code(const std::string &op) {
    if (op == "Add") return HMX_OP_ADD;
    if (op == "Sub") return HMX_OP_SUB;
    if (op == "Mult") return HMX_OP_MUL;
    return HMX_OP_DIV;
}

class Planner {
  public:
    explicit Planner(Record &R) : R_(R) {}

    int plan_block(const std::vector<Stmt> &b, std::vector<Node> &out, bool top) {
        for (const Stmt &st : b) {
            int rc;
            if (st.loop) {
                if ((rc = flush(out))) return rc;
                Node loop;
                loop.t = Node::LOOP;
                loop.ub = st.ub;
                if ((rc = plan_block(st.body, loop.body, false))) return rc;
                if ((rc = flush(loop.body))) return rc;
                out.push_back(try_hs(loop));
                continue;
            }
            if ((rc = check_stmt(R_, st))) return rc;
            const std::string &f = st.filter;
            if (f == "Return") {
                if (!top)
                    return fail(HMX_ERR_UNSUPPORTED, "Return inside a loop is not supported");
                if ((rc = flush(out))) return rc;
                Node n;
                n.t = Node::RETURN;
                n.syms = st.args;
                out.push_back(n);
                continue;
            }
            if (f == "ScalarOp" || f == "ScalarCopy" || f == "ScalarSqrt") {
                if ((rc = flush(out))) return rc;
                ScalarStmt ss;
                ss.dst = st.args[0];
                ss.a = st.args[1];
                ss.b = st.args.size() > 2 ? st.args[2] : -1;
                ss.op = f == "ScalarOp" ? op_code(st.op) : (f == "ScalarSqrt" ? HMX_OP_SQRT : 0);
                if (!out.empty() && out.back().t == Node::SCALAR &&
                    R_.sprogs[out.back().idx].size() < (size_t)SPROG_MAX) {
                    R_.sprogs[out.back().idx].push_back(ss);
                } else {
                    Node n;
                    n.t = Node::SCALAR;
                    n.idx = (int)R_.sprogs.size();
                    R_.sprogs.push_back({ss});
                    out.push_back(n);
                }
                continue;
            }
            if (f == "SpmvCSR") {
                if ((rc = flush(out))) return rc;
                Node n;
                n.t = Node::CSR;
                n.syms = st.args;
                out.push_back(n);
                continue;
            }
            static const std::map<std::string, Node::T> kNodeOf = {
                {"ApplyHSPreconditioner", Node::PRECOND}, {"HSMatrixPowers", Node::POWERS},
                {"Array3DCopy", Node::COPY3D},           {"WarpImg2D", Node::WARP},
                {"SetBroxMatrix", Node::BROX_SET},       {"SetBroxRedblackMatrix", Node::BROX_SET_RB},
                {"ApplyBroxPreconditioner", Node::BROX_PRECOND}};
            auto kn = kNodeOf.find(f);
            if (kn != kNodeOf.end()) {
                if ((rc = flush(out))) return rc;
                Node n;
                n.t = kn->second;
                n.syms = st.args;
                out.push_back(n);
                continue;
            }
            // array statement: fuse into the current program if legal
            if ((rc = add_array_stmt(st, out))) return rc;
        }
        return HMX_OK;
    }

    int flush(std::vector<Node> &out) {
        if (!cur_) return HMX_OK;
        Group &g = *cur_;
        // stores of live-out values
        for (int s : g.written) {
            if (!live_out(s, g)) continue;
            if (g.out_syms.size() >= (size_t)PROG_MAX_OUT)
                return fail(HMX_ERR_UNSUPPORTED, "program needs too many outputs");
            hmx_instr in{HMX_OP_STORE, (int)g.out_syms.size(), g.cur_reg[s], 0};
            g.out_syms.push_back(s);
            g.code.push_back(in);
        }
        ssa_[(int)R_.groups.size()] = g.code;  // kept for loop matching
        int rc = regalloc(g);
        if (rc) return rc;
        Node n;
        n.t = Node::GROUP;
        n.idx = (int)R_.groups.size();
        R_.groups.push_back(std::move(g));
        out.push_back(n);
        cur_.reset();
        return HMX_OK;
    }

  private:
    Record &R_;
    std::unique_ptr<Group> cur_;

    bool live_out(int s, const Group &g) {
        const Sym &S = R_.syms[s];
        if (S.header) return true;
        auto fw = g.first_write.find(s);
        auto it = R_.reads.find(s);
        if (it == R_.reads.end()) return false;
        for (int r : it->second)
            if (r <= fw->second || r > g.s1) return true;
        return false;
    }

    size_t count_live(const Group &g) {
        size_t n = 0;
        for (int s : g.written) n += live_out(s, g);
        return n;
    }

    int new_reg(Group &g) { return g.nreg++; }

    int emit(Group &g, int op, int a, int b) {
        int r = new_reg(g);
        g.code.push_back(hmx_instr{op, r, a, b});
        return r;
    }

    int slot_in(Group &g, int s) {
        auto it = g.in_slot.find(s);
        if (it != g.in_slot.end()) return it->second;
        int k = (int)g.in_syms.size();
        g.in_syms.push_back(s);
        g.in_slot[s] = k;
        return k;
    }

    int reg_of(Group &g, int s) {
        auto it = g.cur_reg.find(s);
        if (it != g.cur_reg.end()) return it->second;
        int r = emit(g, HMX_OP_LOAD, slot_in(g, s), 0);
        g.cur_reg[s] = r;
        return r;
    }

    int sreg_of(Group &g, int s) {
        int k;
        auto it = g.scal_slot.find(s);
        if (it != g.scal_slot.end()) {
            k = it->second;
        } else {
            k = (int)g.scal_syms.size();
            g.scal_syms.push_back(s);
            g.scal_slot[s] = k;
        }
        return emit(g, HMX_OP_LOADS, k, 0);
    }

    int opnd_of(Group &g, int s) {
        auto it = g.opnd_slot.find(s);
        if (it != g.opnd_slot.end()) return it->second;
        int k = (int)g.opnd_syms.size();
        g.opnd_syms.push_back(s);
        g.opnd_slot[s] = k;
        return k;
    }

    void write(Group &g, int s, int r, int idx) {
        g.cur_reg[s] = r;
        if (!g.written.count(s)) g.first_write[s] = idx;
        g.written.insert(s);
        if (g.nb_read.count(s)) {
            R_.syms[s].renamed = true;
            R_.renamed.insert(s);
        }
    }

    int add_array_stmt(const Stmt &st, std::vector<Node> &out) {
        const std::string &f = st.filter;
        // element grid of the statement
        const Sym &src = R_.syms[st.args[f == "ScalarArray2DOp" || f == "ScalarArrayOp" ? 2 : 1]];
        const Sym &dst = R_.syms[st.args[0]];
        int H = is_arr(dst) ? dst.H : src.H, W = is_arr(dst) ? dst.W : src.W;
        // read operands of the statement
        std::vector<int> rd(st.args.begin() + 1, st.args.end());
        const bool nb = f == "Stencil2DMVM" || f == "SpmvDIA2D" || f == "SpmvDIA";
        for (int a : rd)
            if (is_arr(R_.syms[a]) && (R_.syms[a].H != H || R_.syms[a].W != W))
                return fail(HMX_ERR_INVALID, f + ": operand " + R_.syms[a].name +
                                                 " has a different shape than " + dst.name);
        bool split = false;
        if (cur_) {
            Group &g = *cur_;
            if (g.H != H || g.W != W) split = true;
            if (nb && g.written.count(st.args[2])) split = true;
            for (int a : rd)
                if (R_.syms[a].kind == HMX_KIND_SCALAR && g.red_out.count(a)) split = true;
            const size_t nlive = count_live(g);
            if (g.code.size() + nlive + 8 > (size_t)PROG_MAX_INSTR) split = true;
            if (g.in_syms.size() + 3 > (size_t)PROG_MAX_IN) split = true;
            if (g.scal_syms.size() + 2 > (size_t)PROG_MAX_SCAL) split = true;
            if (nb && g.opnd_syms.size() + 1 > (size_t)PROG_MAX_OPND && !g.opnd_slot.count(st.args[1]))
                split = true;
            if ((f == "Array2DSum" || f == "ArraySum") && g.red_syms.size() >= (size_t)PROG_MAX_RED)
                split = true;
            if (nlive + 1 > (size_t)PROG_MAX_OUT) split = true;
        }
        if (split) {
            int rc = flush(out);
            if (rc) return rc;
        }
        if (!cur_) {
            cur_.reset(new Group());
            cur_->H = H;
            cur_->W = W;
            cur_->s0 = st.index;
        }
        Group &g = *cur_;
        g.s1 = st.index;
        const int idx = st.index;
        if (f == "Array2DOp" || f == "ArrayOp") {
            int a = reg_of(g, st.args[1]), b = reg_of(g, st.args[2]);
            write(g, st.args[0], emit(g, op_code(st.op), a, b), idx);
        } else if (f == "Array2DScalarOp" || f == "ArrayScalarOp") {
            int a = reg_of(g, st.args[1]), b = sreg_of(g, st.args[2]);
            write(g, st.args[0], emit(g, op_code(st.op), a, b), idx);
        } else if (f == "ScalarArray2DOp" || f == "ScalarArrayOp") {
            int a = sreg_of(g, st.args[1]), b = reg_of(g, st.args[2]);
            write(g, st.args[0], emit(g, op_code(st.op), a, b), idx);
        } else if (f == "Array2DCopy" || f == "ArrayCopy") {
            write(g, st.args[0], reg_of(g, st.args[1]), idx);
        } else if (f == "Array2DSqrt" || f == "ArraySqrt") {
            write(g, st.args[0], emit(g, HMX_OP_SQRT, reg_of(g, st.args[1]), 0), idx);
        } else if (f == "RedArray2D" || f == "BlackArray2D") {
            write(g, st.args[0],
                  emit(g, f == "RedArray2D" ? HMX_OP_RED : HMX_OP_BLACK, reg_of(g, st.args[1]), 0),
                  idx);
        } else if (f == "Array2DSum" || f == "ArraySum") {
            int r = reg_of(g, st.args[1]);
            int slot = (int)g.red_syms.size();
            g.red_syms.push_back(st.args[0]);
            g.red_out.insert(st.args[0]);
            g.code.push_back(hmx_instr{HMX_OP_REDUCE, slot, r, 0});
        } else if (nb) {
            int x = st.args[2];
            g.nb_read.insert(x);
            int r = emit(g, f == "Stencil2DMVM" ? HMX_OP_STENCIL : HMX_OP_DIA, slot_in(g, x),
                         opnd_of(g, st.args[1]));
            write(g, st.args[0], r, idx);
        }
        return HMX_OK;
    }

    // Linear-scan compaction of SSA registers into the LDS register file.
    int regalloc(Group &g) {
        std::vector<int> last(g.nreg, -1);
        auto uses = [](const hmx_instr &in, int &u0, int &u1) {
            u0 = u1 = -1;
            switch (in.op) {
                case HMX_OP_ADD: case HMX_OP_SUB: case HMX_OP_MUL: case HMX_OP_DIV:
                    u0 = in.a; u1 = in.b; break;
                case HMX_OP_SQRT: case HMX_OP_STORE: case HMX_OP_REDUCE: case HMX_OP_RED:
                case HMX_OP_BLACK:
                    u0 = in.a;
                    break;
                default: break;
            }
        };
        auto defines = [](const hmx_instr &in) {
            return in.op != HMX_OP_STORE && in.op != HMX_OP_REDUCE;
        };
        for (size_t i = 0; i < g.code.size(); ++i) {
            int u0, u1;
            uses(g.code[i], u0, u1);
            if (u0 >= 0) last[u0] = (int)i;
            if (u1 >= 0) last[u1] = (int)i;
        }
        std::vector<int> phys(g.nreg, -1);
        std::vector<int> free_regs;
        for (int r = PROG_MAX_REG - 1; r >= 0; --r) free_regs.push_back(r);
        std::vector<hmx_instr> out;
        for (size_t i = 0; i < g.code.size(); ++i) {
            hmx_instr in = g.code[i];
            int u0, u1;
            uses(in, u0, u1);
            hmx_instr o = in;
            if (u0 >= 0) o.a = phys[u0];
            if (u1 >= 0) o.b = phys[u1];
            // release operands whose last use is this instruction
            std::set<int> rel;
            if (u0 >= 0 && last[u0] == (int)i) rel.insert(phys[u0]);
            if (u1 >= 0 && last[u1] == (int)i) rel.insert(phys[u1]);
            for (int p : rel) free_regs.push_back(p);
            if (defines(in)) {
                if (last[in.dst] < 0) continue;  // dead value
                if (free_regs.empty())
                    return fail(HMX_ERR_UNSUPPORTED,
                                "fused program needs more than " + std::to_string(PROG_MAX_REG) +
                                    " live temporaries");
                phys[in.dst] = free_regs.back();
                free_regs.pop_back();
                o.dst = phys[in.dst];
            }
            out.push_back(o);
        }
        g.code = out;
        return HMX_OK;
    }

    // ---- Horn-Schunck loop recognition --------------------------------
    // Body == one program whose only live-out stores are
    //   U <- SUB(S(U), MUL(Ix, T)),  V <- SUB(S(V), MUL(Iy, T)),
    //   T  = DIV(ADD(ADD(MUL(Ix, S(U)), MUL(Iy, S(V))), It), Den)
    // with S the same radius-1 stencil, Ix/Iy/It/Den loaded arrays not
    // written in the body.  Matching runs on the SSA code before register
    // allocation, so it is kept in a side copy.
    Node try_hs(Node &loop) {
        if (loop.body.size() != 1 || loop.body[0].t != Node::GROUP) return loop;
        const Group &g = R_.groups[loop.body[0].idx];
        if (!g.red_syms.empty() || g.out_syms.size() != 2) return loop;
        const std::vector<hmx_instr> &c = ssa_[loop.body[0].idx];
        std::map<int, const hmx_instr *> def;
        for (const hmx_instr &in : c)
            if (in.op != HMX_OP_STORE && in.op != HMX_OP_REDUCE) def[in.dst] = &in;
        auto D = [&](int r) -> const hmx_instr * {
            auto it = def.find(r);
            return it == def.end() ? nullptr : it->second;
        };
        auto loaded = [&](int r) -> int {  // sym loaded into r, or -1
            const hmx_instr *d = D(r);
            return d && d->op == HMX_OP_LOAD ? g.in_syms[d->a] : -1;
        };
        int su = -1, sv = -1, ru = -1, rv = -1;
        for (const hmx_instr &in : c)
            if (in.op == HMX_OP_STORE) {
                if (su < 0) { su = g.out_syms[in.dst]; ru = in.a; }
                else { sv = g.out_syms[in.dst]; rv = in.a; }
            }
        if (su < 0 || sv < 0 || su == sv) return loop;
        // MUL(LOAD X, r) in either operand order (fp32 multiply commutes)
        auto mul_load = [&](int r, int &X, int &other) -> bool {
            const hmx_instr *m = D(r);
            if (!m || m->op != HMX_OP_MUL) return false;
            if (loaded(m->a) >= 0 && loaded(m->b) < 0) { X = loaded(m->a); other = m->b; return true; }
            if (loaded(m->b) >= 0 && loaded(m->a) < 0) { X = loaded(m->b); other = m->a; return true; }
            return false;
        };
        // U <- SUB(S(U), MUL(X, T))
        auto match_upd = [&](int r, int self, int &stencil, int &X, int &T, int &Sreg) -> bool {
            const hmx_instr *sub = D(r);
            if (!sub || sub->op != HMX_OP_SUB) return false;
            const hmx_instr *st = D(sub->a);
            if (!st || st->op != HMX_OP_STENCIL || g.in_syms[st->a] != self) return false;
            if (!mul_load(sub->b, X, T)) return false;
            stencil = g.opnd_syms[st->b];
            Sreg = sub->a;
            return true;
        };
        int stu, stv, Xu, Xv, Tu, Tv, Su, Sv;
        if (!match_upd(ru, su, stu, Xu, Tu, Su) || !match_upd(rv, sv, stv, Xv, Tv, Sv))
            return loop;
        if (stu != stv || Tu != Tv) return loop;
        // T = DIV(ADD(ADD(MUL(Xu, Su), MUL(Xv, Sv)), It), Den), adds in either order
        const hmx_instr *dv = D(Tu);
        if (!dv || dv->op != HMX_OP_DIV) return loop;
        int den = loaded(dv->b);
        const hmx_instr *a2 = D(dv->a);
        if (den < 0 || !a2 || a2->op != HMX_OP_ADD) return loop;
        int it = loaded(a2->b), inner = a2->a;
        if (it < 0) { it = loaded(a2->a); inner = a2->b; }
        const hmx_instr *a1 = D(inner);
        if (it < 0 || !a1 || a1->op != HMX_OP_ADD) return loop;
        auto term = [&](int r, int X, int S) {
            int x, o;
            return mul_load(r, x, o) && x == X && o == S;
        };
        const bool sums = (term(a1->a, Xu, Su) && term(a1->b, Xv, Sv)) ||
                          (term(a1->b, Xu, Su) && term(a1->a, Xv, Sv));
        if (!sums) return loop;
        std::set<int> ro = {Xu, Xv, it, den};
        if (ro.count(su) || ro.count(sv) || g.written.count(Xu) || g.written.count(Xv) ||
            g.written.count(it) || g.written.count(den))
            return loop;
        if (ro.size() != 4) return loop;  // distinct Ix, Iy, It, den arrays
        if (R_.syms[su].H != R_.syms[Xu].H || R_.syms[su].W != R_.syms[Xu].W) return loop;
        Node hs;
        hs.t = Node::HS;
        hs.ub = loop.ub;
        hs.syms = {su, sv, Xu, Xv, it, den, stu};
        hs.body = loop.body;  // generic fallback if the stencil turns out wider than radius 1
        return hs;
    }

    std::map<int, std::vector<hmx_instr>> ssa_;  // group index -> SSA code
};

// ---------------------------------------------------------------------------
// Device memory

static int dalloc(Record &R, void **p, int64_t bytes) {
    if (bytes <= 0) bytes = 4;
    HMX_HIP_TRY(hipMalloc(p, (size_t)bytes));
    HMX_HIP_TRY(hipMemsetAsync(*p, 0, (size_t)bytes, R.stream));
    R.allocs.push_back(*p);
    return HMX_OK;
}

static void mark_buffers(Record &R, const std::vector<Node> &nodes) {
    for (const Node &n : nodes) {
        switch (n.t) {
            case Node::GROUP: {
                const Group &g = R.groups[n.idx];
                for (int s : g.in_syms) R.syms[s].needs_buffer = true;
                for (int s : g.out_syms) R.syms[s].needs_buffer = true;
                break;
            }
            case Node::CSR: case Node::RETURN: case Node::HS: case Node::PRECOND:
            case Node::POWERS: case Node::COPY3D: case Node::WARP: case Node::BROX_SET:
            case Node::BROX_SET_RB: case Node::BROX_PRECOND:
                for (int s : n.syms) R.syms[s].needs_buffer = true;
                mark_buffers(R, n.body);
                break;
            case Node::LOOP: mark_buffers(R, n.body); break;
            default: break;
        }
    }
}

// Byte size of each component (data, offx/indices, offy/indptr).
static void size_symbols(Record &R) {
    for (Sym &s : R.syms) {
        switch (s.kind) {
            case HMX_KIND_ARRAY2D: case HMX_KIND_ARRAY: s.bytes[0] = (int64_t)s.H * s.W * 4; break;
            case HMX_KIND_ARRAY3D: s.bytes[0] = s.dim0 * s.H * s.W * 4; break;
            case HMX_KIND_SCALAR: s.bytes[0] = dtype_size(s.dtype); break;
            case HMX_KIND_STENCIL:
                s.bytes[0] = s.bytes[1] = s.bytes[2] = s.ndiags * 4;
                break;
            case HMX_KIND_SPMAT2D:
                s.bytes[0] = s.ndiags * s.dim0 * s.dim1 * 4;
                s.bytes[1] = s.bytes[2] = s.ndiags * 4;
                break;
            case HMX_KIND_SPMAT:  // data, offsets, and zero vertical offsets (internal)
                s.bytes[0] = s.ndiags * s.dim0 * 4;
                s.bytes[1] = s.bytes[2] = s.ndiags * 4;
                break;
            case HMX_KIND_CSRMAT:
                s.bytes[0] = s.bytes[1] = s.nnz * 4;
                s.bytes[2] = (s.dim0 + 1) * 4;
                break;
        }
    }
}

static int allocate(Record &R) {
    for (Sym &s : R.syms) {
        int rc = HMX_OK;
        switch (s.kind) {
            case HMX_KIND_ARRAY2D: case HMX_KIND_ARRAY:
                if (s.header || s.needs_buffer) rc = dalloc(R, &s.dev[0], s.bytes[0]);
                if (!rc && s.renamed) rc = dalloc(R, &s.alt, s.bytes[0]);
                break;
            case HMX_KIND_SCALAR: rc = dalloc(R, &s.dev[0], 8); break;
            case HMX_KIND_ARRAY3D:
                if (s.needs_buffer) rc = dalloc(R, &s.dev[0], s.bytes[0]);
                break;
            case HMX_KIND_SPMAT:
                for (int k = 0; k < 3 && !rc; ++k) rc = dalloc(R, &s.dev[k], s.bytes[k]);
                // the vertical offsets of the 1 x n grid are all zero
                if (!rc && hipMemset(s.dev[2], 0, (size_t)s.bytes[2]) != hipSuccess)
                    rc = fail(HMX_ERR_HIP, "clearing the offsets of " + s.name);
                break;
            default:
                for (int k = 0; k < 3 && !rc; ++k) rc = dalloc(R, &s.dev[k], s.bytes[k]);
                break;
        }
        if (rc) return rc;
    }
    return HMX_OK;
}

static void release_record(Record &R) {
    if (R.stream) (void)hipStreamSynchronize(R.stream);
    for (void *p : R.allocs) (void)hipFree(p);
    R.allocs.clear();
    for (Group &g : R.groups) g.partials = nullptr;
    if (R.stream) (void)hipStreamDestroy(R.stream);
    R.stream = nullptr;
}

// ---------------------------------------------------------------------------
// Execution

// What a group's generated kernel is specialised on.
static ProgSpec group_spec(const Record &R, const Group &g) {
    ProgSpec sp;
    sp.code = g.code;
    sp.H = g.H;
    sp.nin = (int)g.in_syms.size();
    sp.nout = (int)g.out_syms.size();
    sp.nscal = (int)g.scal_syms.size();
    sp.nopnd = (int)g.opnd_syms.size();
    sp.nred = (int)g.red_syms.size();
    for (int i = 0; i < sp.nscal && i < PROG_MAX_SCAL; ++i)
        sp.scal_dt[i] = R.syms[g.scal_syms[i]].dtype;
    for (int i = 0; i < sp.nopnd && i < PROG_MAX_OPND; ++i)
        sp.opnd_nd[i] = (int)R.syms[g.opnd_syms[i]].ndiags;
    return sp;
}

static int run_group(Record &R, Group &g) {
    ProgArgs a;
    std::memset(&a, 0, sizeof(a));
    a.ninstr = (int)g.code.size();
    a.nred = (int)g.red_syms.size();
    a.H = g.H;
    a.W = g.W;
    for (size_t i = 0; i < g.code.size(); ++i) a.ins[i] = g.code[i];
    for (size_t i = 0; i < g.in_syms.size(); ++i) a.in[i] = R.syms[g.in_syms[i]].cur_buf();
    for (size_t i = 0; i < g.out_syms.size(); ++i) {
        Sym &s = R.syms[g.out_syms[i]];
        a.out[i] = s.renamed ? s.other_buf() : s.cur_buf();
    }
    for (size_t i = 0; i < g.scal_syms.size(); ++i) {
        a.scal[i] = R.syms[g.scal_syms[i]].dev[0];
        a.scal_dt[i] = R.syms[g.scal_syms[i]].dtype;
    }
    for (size_t i = 0; i < g.opnd_syms.size(); ++i) {
        const Sym &s = R.syms[g.opnd_syms[i]];
        a.opnd[i].w = (const float *)s.dev[0];
        a.opnd[i].ox = (const int *)s.dev[1];
        a.opnd[i].oy = (const int *)s.dev[2];
        a.opnd[i].nd = (int)s.ndiags;
        a.opnd[i].kind = s.kind == HMX_KIND_SPMAT2D || s.kind == HMX_KIND_SPMAT;
    }
    for (size_t k = 0; k < g.red_syms.size(); ++k) a.red_part[k] = g.partials + k * g.nblocks;
    int rc = g.jit ? launch_program_jit(g.jit, a, R.stream) : launch_program(a, R.stream, nullptr);
    if (rc) return rc;
    for (size_t k = 0; k < g.red_syms.size(); ++k) {
        Sym &s = R.syms[g.red_syms[k]];
        rc = launch_reduce_finalize(g.partials + k * g.nblocks, g.nblocks, s.dev[0], s.dtype,
                                    R.stream);
        if (rc) return rc;
    }
    for (int s : g.out_syms)
        if (R.syms[s].renamed) R.syms[s].cur ^= 1;
    return HMX_OK;
}

static int run_scalar(Record &R, const std::vector<ScalarStmt> &p) {
    ScalarProg sp;
    std::memset(&sp, 0, sizeof(sp));
    sp.n = (int)p.size();
    for (size_t i = 0; i < p.size(); ++i) {
        const ScalarStmt &s = p[i];
        ScalarInstr &in = sp.ins[i];
        in.op = s.op;
        in.dst = R.syms[s.dst].dev[0];
        in.dt_dst = R.syms[s.dst].dtype;
        in.a = R.syms[s.a].dev[0];
        in.dt_a = R.syms[s.a].dtype;
        in.b = s.b >= 0 ? R.syms[s.b].dev[0] : nullptr;
        in.dt_b = s.b >= 0 ? R.syms[s.b].dtype : HMX_F32;
    }
    return launch_scalar_program(sp, R.stream);
}

static int host_taps(const Sym &st, Taps &t) {
    if (!st.host[0] || !st.host[1] || !st.host[2])
        return fail(HMX_ERR_STATE, "stencil " + st.name + " was not bound");
    return make_taps(t, (int)st.ndiags, (const float *)st.host[0], (const int *)st.host[1],
                     (const int *)st.host[2]);
}

static int exec_nodes(Record &R, std::vector<Node> &nodes);

static int exec_node(Record &R, Node &n) {
    switch (n.t) {
        case Node::GROUP: return run_group(R, R.groups[n.idx]);
        case Node::SCALAR: return run_scalar(R, R.sprogs[n.idx]);
        case Node::CSR: {
            Sym &y = R.syms[n.syms[0]], &A = R.syms[n.syms[1]], &x = R.syms[n.syms[2]];
            int rc = launch_csr((int)A.dim0, A.nnz, (const int *)A.dev[2], (const int *)A.dev[1],
                                (const float *)A.dev[0], x.cur_buf(),
                                y.renamed ? y.other_buf() : y.cur_buf(), 0, R.stream);
            if (!rc && y.renamed) y.cur ^= 1;
            return rc;
        }
        case Node::LOOP:
            for (int i = 0; i < n.ub; ++i) {
                int rc = exec_nodes(R, n.body);
                if (rc) return rc;
            }
            return HMX_OK;
        case Node::PRECOND: {
            Sym &y0 = R.syms[n.syms[0]], &y1 = R.syms[n.syms[1]], &D = R.syms[n.syms[2]];
            const Sym &x = R.syms[n.syms[6]];
            int rc = launch_hs_precond(R.syms[n.syms[3]].cur_buf(), R.syms[n.syms[4]].cur_buf(),
                                       R.syms[n.syms[5]].cur_buf(), x.cur_buf(),
                                       R.syms[n.syms[7]].cur_buf(),
                                       y0.renamed ? y0.other_buf() : y0.cur_buf(),
                                       y1.renamed ? y1.other_buf() : y1.cur_buf(), x.H, x.W,
                                       (const float *)D.dev[0], 0.0f, R.stream);
            if (!rc && y0.renamed) y0.cur ^= 1;
            if (!rc && y1.renamed) y1.cur ^= 1;
            return rc;
        }
        case Node::POWERS: {
            Sym &V0 = R.syms[n.syms[0]], &V1 = R.syms[n.syms[1]], &lam = R.syms[n.syms[2]];
            const Sym &r = R.syms[n.syms[5]];
            return launch_hs_powers(R.syms[n.syms[3]].cur_buf(), R.syms[n.syms[4]].cur_buf(),
                                    r.cur_buf(), R.syms[n.syms[6]].cur_buf(),
                                    (float *)V0.dev[0], (float *)V1.dev[0], r.H, r.W,
                                    (int)V0.dim0, lam.dev[0], lam.dtype, 0.0f, R.stream);
        }
        case Node::BROX_SET: case Node::BROX_SET_RB: {
            const bool rb = n.t == Node::BROX_SET_RB;
            const int o = rb ? 1 : 0;
            Sym &A = R.syms[n.syms[o]], &psi = R.syms[n.syms[o + 1]], &al = R.syms[n.syms[o + 2]];
            float *diag = nullptr;
            Sym *D = rb ? &R.syms[n.syms[0]] : nullptr;
            if (D) diag = D->renamed ? D->other_buf() : D->cur_buf();
            int rc = launch_brox_set((float *)A.dev[0], diag, psi.cur_buf(), psi.H, psi.W,
                                     al.dev[0], al.dtype, 0.0f, rb ? 1 : 0, R.stream);
            if (!rc && D && D->renamed) D->cur ^= 1;
            return rc;
        }
        case Node::BROX_PRECOND: {
            Sym &y0 = R.syms[n.syms[0]], &y1 = R.syms[n.syms[1]], &A = R.syms[n.syms[2]];
            const Sym &x = R.syms[n.syms[6]];
            int rc = launch_brox_precond((const float *)A.dev[0] + 2 * (int64_t)x.H * x.W,
                                         R.syms[n.syms[3]].cur_buf(), R.syms[n.syms[4]].cur_buf(),
                                         R.syms[n.syms[5]].cur_buf(), x.cur_buf(),
                                         R.syms[n.syms[7]].cur_buf(),
                                         y0.renamed ? y0.other_buf() : y0.cur_buf(),
                                         y1.renamed ? y1.other_buf() : y1.cur_buf(), x.H, x.W,
                                         R.stream);
            if (!rc && y0.renamed) y0.cur ^= 1;
            if (!rc && y1.renamed) y1.cur ^= 1;
            return rc;
        }
        case Node::WARP: {
            Sym &d = R.syms[n.syms[0]], &src = R.syms[n.syms[1]];
            int rc = launch_warp_img2d(src.cur_buf(), R.syms[n.syms[2]].cur_buf(),
                                       R.syms[n.syms[3]].cur_buf(),
                                       d.renamed ? d.other_buf() : d.cur_buf(), src.H, src.W,
                                       R.stream);
            if (!rc && d.renamed) d.cur ^= 1;
            return rc;
        }
        case Node::COPY3D: {
            Sym &d = R.syms[n.syms[0]], &src = R.syms[n.syms[1]];
            if (d.dev[0] != src.dev[0])
                HMX_HIP_TRY(hipMemcpyAsync(d.dev[0], src.dev[0], (size_t)d.bytes[0],
                                           hipMemcpyDeviceToDevice, R.stream));
            return HMX_OK;
        }
        case Node::HS: {
            Sym &U = R.syms[n.syms[0]], &V = R.syms[n.syms[1]];
            Taps t;
            int rc = host_taps(R.syms[n.syms[6]], t);
            if (rc) return rc;
            if (t.rx > 1 || t.ry > 1 || t.nd > 9) {  // beyond what the fused kernel takes
                for (int i = 0; i < n.ub && !rc; ++i) rc = exec_nodes(R, n.body);
                return rc;
            }
            int variant, k;
            hs_choice(U.H, U.W, variant, k);
            const int64_t np = (int64_t)U.H * U.W;
            return hs_jacobi_run(R.syms[n.syms[2]].cur_buf(), R.syms[n.syms[3]].cur_buf(),
                                 R.syms[n.syms[4]].cur_buf(), R.syms[n.syms[5]].cur_buf(), 0.0f,
                                 U.H, U.W, t, n.ub, U.cur_buf(), V.cur_buf(), R.hs_work,
                                 R.hs_work + np, variant, k, R.stream);
        }
        case Node::RETURN: {
            R.returns = n.syms;
            R.results.assign(n.syms.size(), {});
            for (size_t i = 0; i < n.syms.size(); ++i) {
                Sym &s = R.syms[n.syms[i]];
                if (s.kind != HMX_KIND_ARRAY2D && s.kind != HMX_KIND_ARRAY &&
                    s.kind != HMX_KIND_SCALAR && s.kind != HMX_KIND_ARRAY3D)
                    return fail(HMX_ERR_UNSUPPORTED, "returning " + s.name +
                                                         " (a matrix/stencil) is not supported");
                R.results[i].resize((size_t)s.bytes[0]);
                const void *src = (s.kind == HMX_KIND_SCALAR || s.kind == HMX_KIND_ARRAY3D)
                                      ? s.dev[0]
                                      : (void *)s.cur_buf();
                HMX_HIP_TRY(hipMemcpyAsync(R.results[i].data(), src, (size_t)s.bytes[0],
                                           hipMemcpyDeviceToHost, R.stream));
            }
            return HMX_OK;
        }
    }
    return HMX_OK;
}

static int exec_nodes(Record &R, std::vector<Node> &nodes) {
    for (Node &n : nodes) {
        int rc = exec_node(R, n);
        if (rc) return rc;
        if (n.t == Node::RETURN) break;
    }
    return HMX_OK;
}

static void describe(const Record &R, const std::vector<Node> &nodes, int depth,
                     std::ostringstream &os) {
    std::string pad(depth * 2, ' ');
    for (const Node &n : nodes) {
        switch (n.t) {
            case Node::GROUP: {
                const Group &g = R.groups[n.idx];
                int nst = 0;
                for (auto &in : g.code) nst += in.op == HMX_OP_STENCIL || in.op == HMX_OP_DIA;
                os << pad << "program " << g.H << "x" << g.W << ": " << g.code.size()
                   << " instr, " << nst << " stencil/dia taps, loads {";
                for (size_t i = 0; i < g.in_syms.size(); ++i)
                    os << (i ? "," : "") << R.syms[g.in_syms[i]].name;
                os << "} stores {";
                for (size_t i = 0; i < g.out_syms.size(); ++i)
                    os << (i ? "," : "") << R.syms[g.out_syms[i]].name
                       << (R.syms[g.out_syms[i]].renamed ? "*" : "");
                os << "}";
                if (!g.red_syms.empty()) os << " reduce " << g.red_syms.size();
                os << "\n";
                break;
            }
            case Node::SCALAR: os << pad << "scalar program: " << R.sprogs[n.idx].size() << " ops\n"; break;
            case Node::CSR:
                os << pad << "csr spmv " << R.syms[n.syms[0]].name << " = " << R.syms[n.syms[1]].name
                   << " * " << R.syms[n.syms[2]].name << "\n";
                break;
            case Node::LOOP:
                os << pad << "loop x" << n.ub << "\n";
                describe(R, n.body, depth + 1, os);
                break;
            case Node::HS:
                os << pad << "hs_jacobi_fused x" << n.ub << " (u=" << R.syms[n.syms[0]].name
                   << ", v=" << R.syms[n.syms[1]].name << ", Ix=" << R.syms[n.syms[2]].name
                   << ", Iy=" << R.syms[n.syms[3]].name << ", It=" << R.syms[n.syms[4]].name
                   << ", den=" << R.syms[n.syms[5]].name << ", S=" << R.syms[n.syms[6]].name
                   << ")\n";
                break;
            case Node::PRECOND:
                os << pad << "hs_precondition (" << R.syms[n.syms[0]].name << ", "
                   << R.syms[n.syms[1]].name << ") = M^-1 (" << R.syms[n.syms[6]].name << ", "
                   << R.syms[n.syms[7]].name << "), D=" << R.syms[n.syms[2]].name << "\n";
                break;
            case Node::POWERS:
                os << pad << "hs_matrix_powers x" << R.syms[n.syms[0]].dim0 << " ("
                   << R.syms[n.syms[0]].name << ", " << R.syms[n.syms[1]].name << ") from ("
                   << R.syms[n.syms[5]].name << ", " << R.syms[n.syms[6]].name
                   << "), lam=" << R.syms[n.syms[2]].name << "\n";
                break;
            case Node::BROX_SET:
                os << pad << "brox_set_matrix " << R.syms[n.syms[0]].name << " from psi="
                   << R.syms[n.syms[1]].name << "\n";
                break;
            case Node::BROX_SET_RB:
                os << pad << "brox_set_redblack_matrix " << R.syms[n.syms[1]].name << ", diag "
                   << R.syms[n.syms[0]].name << " from psi=" << R.syms[n.syms[2]].name << "\n";
                break;
            case Node::BROX_PRECOND:
                os << pad << "brox_precondition (" << R.syms[n.syms[0]].name << ", "
                   << R.syms[n.syms[1]].name << ") = M^-1 (" << R.syms[n.syms[6]].name << ", "
                   << R.syms[n.syms[7]].name << "), A=" << R.syms[n.syms[2]].name << "\n";
                break;
            case Node::WARP:
                os << pad << "warp_img2d " << R.syms[n.syms[0]].name << " = "
                   << R.syms[n.syms[1]].name << " at +(" << R.syms[n.syms[2]].name << ", "
                   << R.syms[n.syms[3]].name << ")\n";
                break;
            case Node::COPY3D:
                os << pad << "copy " << R.syms[n.syms[0]].name << " = " << R.syms[n.syms[1]].name
                   << "\n";
                break;
            case Node::RETURN:
                os << pad << "return";
                for (int s : n.syms) os << " " << R.syms[s].name;
                os << "\n";
                break;
        }
    }
}

static const std::vector<std::string> kDef2D = {"nd20", "nd21", "nt20", "nt21", "ne20", "ne21"};
static const std::vector<std::string> kDef1D = {"nd1", "nt1", "ne1"};

static int check_defines(Record &R, const std::vector<Stmt> &b, bool &need2d, bool &need1d) {
    for (const Stmt &st : b) {
        if (st.loop) {
            int rc = check_defines(R, st.body, need2d, need1d);
            if (rc) return rc;
            continue;
        }
        const bool scalar_only = st.filter.rfind("Scalar", 0) == 0 &&
                                 st.filter.find("Array") == std::string::npos;
        if (st.filter == "Return" || scalar_only) continue;
        // the preconditioned-CG operations launch 2D (ApplyHSPreconditioner.cpp:111-116,
        // HSMatrixPowers.cpp:147-153)
        if (st.filter.find("2D") != std::string::npos || st.filter.find("3D") != std::string::npos ||
            st.filter == "ApplyHSPreconditioner" || st.filter == "HSMatrixPowers" ||
            st.filter.find("Brox") != std::string::npos)
            need2d = true;
        else need1d = true;
    }
    return HMX_OK;
}

}  // namespace rt
}  // namespace hmx

using namespace hmx;
using namespace hmx::rt;

extern "C" int hmx_init(int device) {
    if (device == -1) {  // plan-only mode: compile and describe, never run
        std::lock_guard<std::mutex> lk(g_mu);
        g_device = -1;
        return HMX_OK;
    }
    int n = 0;
    HMX_HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return fail(HMX_ERR_INVALID, "device " + std::to_string(device) + " not present (" +
                                         std::to_string(n) + " visible)");
    std::lock_guard<std::mutex> lk(g_mu);
    g_device = device;
    return HMX_OK;
}

extern "C" int hmx_compile(int rec, const char *symtab_xml, const char *sm_xml,
                           const char *defines_xml) {
    if (!symtab_xml || !sm_xml || !defines_xml) return fail(HMX_ERR_INVALID, "null XML");
    std::unique_ptr<Record> R(new Record());
    R->id = rec;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        R->device = g_device;
    }
    XNode st, sm, df;
    std::string err;
    if (!XReader(symtab_xml).parse(st, err)) return fail(HMX_ERR_INVALID, "symbol table: " + err);
    if (!XReader(sm_xml).parse(sm, err)) return fail(HMX_ERR_INVALID, "semantic model: " + err);
    if (!XReader(defines_xml).parse(df, err)) return fail(HMX_ERR_INVALID, "constants: " + err);
    int rc = parse_symtab(*R, st);
    if (!rc) rc = parse_sm(*R, sm);
    if (rc) return rc;
    for (const XNode &k : df.kids) {
        const std::string *n = k.get("name"), *v = k.get("value");
        if (n && v) R->defines[*n] = *v;
    }
    bool need2d = false, need1d = false;
    check_defines(*R, R->prog, need2d, need1d);
    std::string missing;
    if (need2d)
        for (auto &k : kDef2D)
            if (!R->defines.count(k)) missing += " " + k;
    if (need1d)
        for (auto &k : kDef1D)
            if (!R->defines.count(k)) missing += " " + k;
    if (!missing.empty())
        return fail(HMX_ERR_INVALID, "defines is missing required constants:" + missing +
                                         " (the reference dereferences them, "
                                         "Stencil2DMVM.cpp:68-73)");
    if (need2d && (std::atoi(R->defines["nd20"].c_str()) * std::atoi(R->defines["nd21"].c_str()) != 1))
        return fail(HMX_ERR_UNSUPPORTED, "nd20*nd21 must be 1: this backend runs on one GPU");
    if (need1d && std::atoi(R->defines["nd1"].c_str()) != 1)
        return fail(HMX_ERR_UNSUPPORTED, "nd1 must be 1: this backend runs on one GPU");

    collect_reads(*R, R->prog);
    size_symbols(*R);
    Planner P(*R);
    rc = P.plan_block(R->prog, R->plan, true);
    if (!rc) rc = P.flush(R->plan);
    if (rc) return rc;
    if (R->device < 0) {  // plan-only
        std::lock_guard<std::mutex> lk(g_mu);
        g_records[rec] = std::move(R);
        return HMX_OK;
    }
    HMX_HIP_TRY(hipSetDevice(R->device));
    HMX_HIP_TRY(hipStreamCreateWithFlags(&R->stream, hipStreamNonBlocking));
    mark_buffers(*R, R->plan);
    rc = allocate(*R);
    const bool jit = program_jit_enabled();
    for (Group &g : R->groups) {
        if (rc) break;
        g.nblocks = program_grid_blocks(g.H, g.W);
        if (!g.red_syms.empty())
            rc = dalloc(*R, (void **)&g.partials, (int64_t)g.nblocks * 8 * g.red_syms.size());
        if (!rc && jit) {
            // each fused group becomes its own kernel, as the reference
            // generates and builds one OpenCL kernel per group at xmlcompile
            std::string err;
            rc = program_jit_get(R->device, group_spec(*R, g), &g.jit, err);
            if (rc) rc = fail(rc, "generated program: " + err);
        }
    }
    // HS nodes need two ping-pong planes
    std::function<void(const std::vector<Node> &)> hs_need;
    int64_t need = 0;
    hs_need = [&](const std::vector<Node> &ns) {
        for (const Node &n : ns) {
            if (n.t == Node::HS) {
                const Sym &U = R->syms[n.syms[0]];
                need = std::max<int64_t>(need, 2 * (int64_t)U.H * U.W);
            }
            hs_need(n.body);
        }
    };
    hs_need(R->plan);
    if (!rc && need > 0) {
        rc = dalloc(*R, (void **)&R->hs_work, need * 4);
        R->hs_work_elems = need;
    }
    if (!rc) {
        hipError_t e = hipStreamSynchronize(R->stream);
        if (e != hipSuccess) rc = fail(HMX_ERR_HIP, hipGetErrorString(e));
    }
    if (rc) {
        release_record(*R);
        return rc;
    }
    std::unique_ptr<Record> old;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_records.find(rec);
        if (it != g_records.end()) old = std::move(it->second);
        g_records[rec] = std::move(R);
    }
    if (old) {
        std::lock_guard<std::mutex> lk2(old->mu);
        release_record(*old);
    }
    return HMX_OK;
}

extern "C" int hmx_bind(int rec, const char *name, int component, const void *host,
                        int64_t nbytes) {
    Record *R = find_record(rec);
    if (!R) return fail(HMX_ERR_NOT_FOUND, "no record " + std::to_string(rec));
    if (!name) return fail(HMX_ERR_INVALID, "null name");
    std::lock_guard<std::mutex> lk(R->mu);
    auto it = R->by_name.find(name);
    if (it == R->by_name.end()) return fail(HMX_ERR_NOT_FOUND, std::string("no symbol ") + name);
    Sym &s = R->syms[it->second];
    if (!s.header) return fail(HMX_ERR_INVALID, s.name + " is not a Header (kwarg) symbol");
    if (component < 0 || component > 2 || (component > 0 && (s.kind == HMX_KIND_ARRAY ||
                                                            s.kind == HMX_KIND_ARRAY2D ||
                                                            s.kind == HMX_KIND_SCALAR)) ||
        (component > 1 && s.kind == HMX_KIND_SPMAT))
        return fail(HMX_ERR_INVALID, s.name + ": bad component " + std::to_string(component));
    if (nbytes != s.bytes[component])
        return fail(HMX_ERR_INVALID, s.name + "[" + std::to_string(component) + "]: expected " +
                                         std::to_string(s.bytes[component]) + " bytes, got " +
                                         std::to_string(nbytes));
    if (!host && nbytes > 0) return fail(HMX_ERR_INVALID, s.name + ": null host pointer");
    s.host[component] = host;
    s.host_bytes[component] = nbytes;
    return HMX_OK;
}

extern "C" int hmx_run(int rec) {
    Record *R = find_record(rec);
    if (!R) return fail(HMX_ERR_NOT_FOUND, "no record " + std::to_string(rec));
    std::lock_guard<std::mutex> lk(R->mu);
    if (R->device < 0) return fail(HMX_ERR_STATE, "record was compiled in plan-only mode");
    HMX_HIP_TRY(hipSetDevice(R->device));
    // upload every Header symbol (the reference re-uploads on each call,
    // Array2D.cpp:149-200)
    for (Sym &s : R->syms) {
        if (!s.header) continue;
        int ncomp = (s.kind == HMX_KIND_STENCIL || s.kind == HMX_KIND_SPMAT2D ||
                     s.kind == HMX_KIND_CSRMAT) ? 3 : (s.kind == HMX_KIND_SPMAT ? 2 : 1);
        for (int k = 0; k < ncomp; ++k) {
            if (s.bytes[k] == 0) continue;
            if (!s.host[k])
                return fail(HMX_ERR_STATE, "Header symbol " + s.name + " was not bound");
            HMX_HIP_TRY(hipMemcpyAsync(s.dev[k], s.host[k], (size_t)s.bytes[k],
                                       hipMemcpyHostToDevice, R->stream));
        }
        s.cur = 0;
    }
    for (Sym &s : R->syms) s.cur = 0;
    R->results.clear();
    R->returns.clear();
    int rc = exec_nodes(*R, R->plan);
    hipError_t e = hipStreamSynchronize(R->stream);
    if (rc) return rc;
    if (e != hipSuccess) return fail(HMX_ERR_HIP, std::string("run: ") + hipGetErrorString(e));
    // bindings are valid for one run only
    for (Sym &s : R->syms)
        for (int k = 0; k < 3; ++k) s.host[k] = nullptr;
    R->ran = true;
    return HMX_OK;
}

extern "C" int hmx_result_count(int rec, int *n) {
    Record *R = find_record(rec);
    if (!R) return fail(HMX_ERR_NOT_FOUND, "no record " + std::to_string(rec));
    if (!n) return fail(HMX_ERR_INVALID, "null n");
    *n = (int)R->results.size();
    return HMX_OK;
}

extern "C" int hmx_result_info(int rec, int i, int *kind, int *dtype, int *ndim,
                               int64_t *shape, const char **name) {
    Record *R = find_record(rec);
    if (!R) return fail(HMX_ERR_NOT_FOUND, "no record " + std::to_string(rec));
    if (i < 0 || i >= (int)R->returns.size()) return fail(HMX_ERR_NOT_FOUND, "no such result");
    const Sym &s = R->syms[R->returns[i]];
    if (kind) *kind = s.kind;
    if (dtype) *dtype = s.dtype;
    int nd = s.kind == HMX_KIND_ARRAY3D
                 ? 3
                 : (s.kind == HMX_KIND_ARRAY2D ? 2 : (s.kind == HMX_KIND_ARRAY ? 1 : 0));
    if (ndim) *ndim = nd;
    if (shape) {
        shape[0] = s.dim0;
        shape[1] = s.dim1;
        shape[2] = nd == 3 ? s.dim2 : (s.col_mjr ? 1 : 0);
    }
    if (name) *name = s.name.c_str();
    return HMX_OK;
}

extern "C" int hmx_result_read(int rec, int i, void *host, int64_t nbytes) {
    Record *R = find_record(rec);
    if (!R) return fail(HMX_ERR_NOT_FOUND, "no record " + std::to_string(rec));
    if (i < 0 || i >= (int)R->results.size()) return fail(HMX_ERR_NOT_FOUND, "no such result");
    if (nbytes != (int64_t)R->results[i].size())
        return fail(HMX_ERR_INVALID, "result " + std::to_string(i) + ": expected " +
                                         std::to_string(R->results[i].size()) + " bytes");
    if (!host) return fail(HMX_ERR_INVALID, "null destination");
    std::memcpy(host, R->results[i].data(), (size_t)nbytes);
    return HMX_OK;
}

extern "C" int hmx_plan_text(int rec, char *buf, int64_t cap) {
    Record *R = find_record(rec);
    if (!R) return fail(HMX_ERR_NOT_FOUND, "no record " + std::to_string(rec));
    std::ostringstream os;
    describe(*R, R->plan, 0, os);
    std::string s = os.str();
    if (!buf || cap <= 0) return fail(HMX_ERR_INVALID, "null buffer");
    size_t n = std::min<size_t>(s.size(), (size_t)cap - 1);
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
    return HMX_OK;
}

extern "C" int hmx_program_source(int rec, int compile, char *buf, int64_t cap) {
    Record *R = find_record(rec);
    if (!R) return fail(HMX_ERR_NOT_FOUND, "no record " + std::to_string(rec));
    if (!buf || cap <= 0) return fail(HMX_ERR_INVALID, "null buffer");
    std::string all;
    for (size_t i = 0; i < R->groups.size(); ++i) {
        const ProgSpec sp = group_spec(*R, R->groups[i]);
        std::string src, err;
        int rc = program_source(sp, src, err);
        if (!rc && compile) rc = program_jit_compile_only(src, err);
        if (rc) return fail(rc, "group " + std::to_string(i) + ": " + err);
        all += "// group " + std::to_string(i) + "\n" + src;
    }
    size_t n = std::min<size_t>(all.size(), (size_t)cap - 1);
    std::memcpy(buf, all.data(), n);
    buf[n] = 0;
    return HMX_OK;
}

extern "C" int hmx_release(int rec) {
    std::unique_ptr<Record> R;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_records.find(rec);
        if (it == g_records.end()) return fail(HMX_ERR_NOT_FOUND, "no record " + std::to_string(rec));
        R = std::move(it->second);
        g_records.erase(it);
    }
    std::lock_guard<std::mutex> lk(R->mu);
    release_record(*R);
    return HMX_OK;
}
