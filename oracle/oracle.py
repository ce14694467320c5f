"""ctypes wrapper over liboracle.so plus independent numpy restatements.

TEST INFRASTRUCTURE ONLY (see package docstring).  Each wrapper names the
reference file:line its C function follows.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

__all__ = [
    "build_oracle", "load_oracle", "stencil2d", "stencil2d_T", "dia2d", "csr_spmv",
    "sum2d", "hs_setup", "hs_jacobi", "hs_solve", "hs_grad", "num_threads",
    "set_num_threads", "np_stencil2d", "np_dia2d_csr", "np_hs_solve_f64",
    "endpoint_error", "hs_precondition", "hs_matrix_powers", "warp_img2d", "dia1d",
    "brox_set_matrix", "brox_precondition", "red", "black",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lib = None


def build_oracle(force=False):
    """Compile hm_oracle.c with the committed Makefile (gcc)."""
    src = os.path.join(_HERE, "hm_oracle.c")
    if force or not os.path.exists(_SO) or os.path.getmtime(_SO) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _SO


def load_oracle():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build_oracle()
        lib = ctypes.CDLL(_SO)
        lib.or_sum_f64.restype = ctypes.c_double
        lib.or_num_threads.restype = ctypes.c_int
        _lib = lib
    return _lib


def num_threads():
    return load_oracle().or_num_threads()


def set_num_threads(n):
    load_oracle().or_set_num_threads(int(n))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _taps(st, dt):
    w, ox, oy = st
    cast = _f32 if dt == np.float32 else _f64
    return cast(w), _i32(ox), _i32(oy)


def stencil2d(x, st, dtype=np.float32):
    """Stencil2DMVM.cpp:36-47 -- clamped stencil, fma in tap order (fp32)."""
    lib = load_oracle()
    cast = _f32 if dtype == np.float32 else _f64
    x = cast(x)
    H, W = x.shape
    w, ox, oy = _taps(st, dtype)
    y = np.empty_like(x)
    fn = lib.or_stencil2d_f32 if dtype == np.float32 else lib.or_stencil2d_f64
    fn(_p(x), H, W, len(w), _p(w), _p(ox), _p(oy), _p(y))
    return y


def stencil2d_T(g, st):
    """Transpose of the clamped stencil (fp64 scatter form)."""
    lib = load_oracle()
    g = _f64(g)
    H, W = g.shape
    w, ox, oy = _taps(st, np.float64)
    out = np.empty_like(g)
    lib.or_stencil2d_T_f64(_p(g), H, W, len(w), _p(w), _p(ox), _p(oy), _p(out))
    return out


def dia2d(A, offx, offy, x, dtype=np.float32):
    """SpmvDIA2D.cpp:35-45 + SpMat2D.cpp:211-216 -- coefficient at p+o, skip border."""
    lib = load_oracle()
    cast = _f32 if dtype == np.float32 else _f64
    A = cast(A)
    x = cast(x)
    nd, H, W = A.shape
    ox, oy = _i32(offx), _i32(offy)
    y = np.empty_like(x)
    fn = lib.or_dia2d_f32 if dtype == np.float32 else lib.or_dia2d_f64
    fn(_p(A), nd, _p(ox), _p(oy), _p(x), H, W, _p(y))
    return y


def dia1d(data, offsets, x, dtype=np.float32):
    """SpmvDIA.cpp:30-50 (1-D DIA, square, data ndiags x n): the 2D DIA rule on a
    1 x n grid with zero vertical offsets (coefficient at i+o, out-of-range terms
    skipped, fma in diagonal order)."""
    data = np.asarray(data)
    nd, n = data.shape
    y = dia2d(data.reshape(nd, 1, n), offsets, np.zeros(nd, np.int32),
              np.asarray(x).reshape(1, n), dtype)
    return y.reshape(n)


def csr_spmv(indptr, indices, data, x, dtype=np.float32):
    """SpmvCSR.cpp:32-38 -- one row at a time, fma in index order."""
    lib = load_oracle()
    cast = _f32 if dtype == np.float32 else _f64
    indptr, indices = _i32(indptr), _i32(indices)
    data, x = cast(data), cast(x)
    M = len(indptr) - 1
    y = np.empty(M, dtype=dtype)
    fn = lib.or_csr_f32 if dtype == np.float32 else lib.or_csr_f64
    fn(M, _p(indptr), _p(indices), _p(data), _p(x), _p(y))
    return y


def sum2d(x):
    """Array2DSum.cpp:12-119 value, accumulated exactly enough in fp64."""
    x = _f32(x)
    return load_oracle().or_sum_f64(_p(x), ctypes.c_int64(x.size))


def _st_args(sts, dtype):
    args = []
    keep = []
    for st in sts:
        w, ox, oy = _taps(st, dtype)
        keep += [w, ox, oy]
        args += [len(w), _p(w), _p(ox), _p(oy)]
    return args, keep


def hs_setup(I1, I2, Dx, Dy, Dt, dtype=np.float32):
    """Canonical setup Ix = Dx*I1 + Dx*I2, Iy = Dy*I1 + Dy*I2, It = Dt*I2 - Dt*I1."""
    lib = load_oracle()
    cast = _f32 if dtype == np.float32 else _f64
    I1, I2 = cast(I1), cast(I2)
    H, W = I1.shape
    Ix, Iy, It = np.empty_like(I1), np.empty_like(I1), np.empty_like(I1)
    args, keep = _st_args((Dx, Dy, Dt), dtype)
    fn = lib.or_hs_setup_f32 if dtype == np.float32 else lib.or_hs_setup_f64
    fn(_p(I1), _p(I2), H, W, *args, _p(Ix), _p(Iy), _p(It))
    return Ix, Iy, It


def hs_jacobi(Ix, Iy, It, alpha2, Avg, iters, u=None, v=None, dtype=np.float32):
    """N Jacobi iterations of the canonical loop body; returns new (u, v)."""
    lib = load_oracle()
    cast = _f32 if dtype == np.float32 else _f64
    Ix, Iy, It = cast(Ix), cast(Iy), cast(It)
    H, W = Ix.shape
    u = np.zeros_like(Ix) if u is None else cast(u).copy()
    v = np.zeros_like(Ix) if v is None else cast(v).copy()
    args, keep = _st_args((Avg,), dtype)
    if dtype == np.float32:
        lib.or_hs_jacobi_f32(_p(Ix), _p(Iy), _p(It), ctypes.c_float(alpha2), H, W, *args,
                             int(iters), _p(u), _p(v))
    else:
        lib.or_hs_jacobi_f64(_p(Ix), _p(Iy), _p(It), ctypes.c_double(alpha2), H, W, *args,
                             int(iters), _p(u), _p(v))
    return u, v


def hs_solve(I1, I2, alpha2, taps, iters, u=None, v=None):
    """fp32 setup + iterations in one call (the CPU baseline)."""
    lib = load_oracle()
    I1, I2 = _f32(I1), _f32(I2)
    H, W = I1.shape
    u = np.zeros_like(I1) if u is None else _f32(u).copy()
    v = np.zeros_like(I1) if v is None else _f32(v).copy()
    args, keep = _st_args((taps["Dx"], taps["Dy"], taps["Dt"], taps["Avg"]), np.float32)
    lib.or_hs_solve_f32(_p(I1), _p(I2), H, W, ctypes.c_float(alpha2), *args, int(iters),
                        _p(u), _p(v))
    return u, v


def hs_grad(I1, I2, alpha, taps, iters, gu, gv, u0=None, v0=None):
    """fp64 reverse-mode gradient of L(u_N, v_N) given dL/du_N, dL/dv_N.

    Returns (gI1, gI2, galpha, gu0, gv0)."""
    lib = load_oracle()
    I1, I2 = _f64(I1), _f64(I2)
    H, W = I1.shape
    u0 = np.zeros_like(I1) if u0 is None else _f64(u0)
    v0 = np.zeros_like(I1) if v0 is None else _f64(v0)
    gu = _f64(gu).copy()
    gv = _f64(gv).copy()
    gI1, gI2 = np.empty_like(I1), np.empty_like(I1)
    ga = ctypes.c_double(0.0)
    args, keep = _st_args((taps["Dx"], taps["Dy"], taps["Dt"], taps["Avg"]), np.float64)
    lib.or_hs_grad_f64(_p(I1), _p(I2), H, W, ctypes.c_double(alpha), *args, int(iters),
                       _p(u0), _p(v0), _p(gu), _p(gv), _p(gI1), _p(gI2), ctypes.byref(ga))
    return gI1, gI2, ga.value, gu, gv


# --------------------------------------------------------------------------
# Independent numpy restatements (used to pin the C code, never the reverse)

def np_stencil2d(x, st):
    """fp64 clamped stencil via edge padding + shifted slices (hlib.py:173-191
    idea, with the slice sizing written for asymmetric offsets)."""
    w, ox, oy = st
    x = np.asarray(x, dtype=np.float64)
    H, W = x.shape
    R = int(max(1, np.max(np.abs(ox)), np.max(np.abs(oy))))
    xp = np.pad(x, R, mode="edge")
    y = np.zeros((H, W))
    for d in range(len(w)):
        y += float(w[d]) * xp[R + oy[d]:R + oy[d] + H, R + ox[d]:R + ox[d] + W]
    return y


def np_dia2d_csr(A, offx, offy):
    """scipy CSR matrix with entries (p, p+o, A[d][p+o]) for in-bounds p+o."""
    import scipy.sparse as sp
    nd, H, W = A.shape
    rows, cols, vals = [], [], []
    rr, cc = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    for d in range(nd):
        r2, c2 = rr + offy[d], cc + offx[d]
        m = (r2 >= 0) & (r2 < H) & (c2 >= 0) & (c2 < W)
        p = (rr * W + cc)[m]
        q = (r2 * W + c2)[m]
        rows.append(p)
        cols.append(q)
        vals.append(A[d].reshape(-1)[q].astype(np.float64))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(H * W, H * W))


def np_hs_solve_f64(I1, I2, alpha2, taps, iters, u=None, v=None):
    """Textbook Horn-Schunck (1981) iteration in numpy fp64."""
    I1 = np.asarray(I1, np.float64)
    I2 = np.asarray(I2, np.float64)
    Ix = np_stencil2d(I1, taps["Dx"]) + np_stencil2d(I2, taps["Dx"])
    Iy = np_stencil2d(I1, taps["Dy"]) + np_stencil2d(I2, taps["Dy"])
    It = np_stencil2d(I2, taps["Dt"]) - np_stencil2d(I1, taps["Dt"])
    den = alpha2 + Ix * Ix + Iy * Iy
    u = np.zeros_like(I1) if u is None else np.asarray(u, np.float64).copy()
    v = np.zeros_like(I1) if v is None else np.asarray(v, np.float64).copy()
    for _ in range(iters):
        ub = np_stencil2d(u, taps["Avg"])
        vb = np_stencil2d(v, taps["Avg"])
        t = (Ix * ub + Iy * vb + It) / den
        u = ub - Ix * t
        v = vb - Iy * t
    return u, v


def hs_precondition(s2, diag0, diag1, offdiag, x0, x1, dtype=np.float32):
    """ApplyHSPreconditioner.cpp:46-97: (y0, y1) = M^-1 (x0, x1) per pixel; s2 is
    weight 2 of the stencil operand."""
    lib = load_oracle()
    cast = _f32 if dtype == np.float32 else _f64
    d0, d1, od, a, b = (cast(t) for t in (diag0, diag1, offdiag, x0, x1))
    H, W = a.shape
    y0, y1 = np.empty_like(a), np.empty_like(a)
    if dtype == np.float32:
        lib.or_hs_precond_f32(_p(d0), _p(d1), _p(od), _p(a), _p(b), _p(y0), _p(y1), H, W,
                              ctypes.c_float(s2))
    else:
        lib.or_hs_precond_f64(_p(d0), _p(d1), _p(od), _p(a), _p(b), _p(y0), _p(y1), H, W,
                              ctypes.c_double(s2))
    return y0, y1


def hs_matrix_powers(lam, Ix, Iy, r0, r1, steps, dtype=np.float32):
    """HSMatrixPowers.cpp:28-39, 72-126, 206-226: V[0] = r, V[s] = A V[s-1];
    returns (V0, V1), each steps x H x W."""
    lib = load_oracle()
    cast = _f32 if dtype == np.float32 else _f64
    ix, iy, a, b = (cast(t) for t in (Ix, Iy, r0, r1))
    H, W = a.shape
    V0 = np.empty((steps, H, W), dtype=dtype)
    V1 = np.empty_like(V0)
    if dtype == np.float32:
        lib.or_hs_powers_f32(ctypes.c_float(lam), _p(ix), _p(iy), _p(a), _p(b), _p(V0), _p(V1),
                             H, W, int(steps))
    else:
        lib.or_hs_powers_f64(ctypes.c_double(lam), _p(ix), _p(iy), _p(a), _p(b), _p(V0),
                             _p(V1), H, W, int(steps))
    return V0, V1


def warp_img2d(src, offx, offy, dtype=np.float32):
    """WarpImg2D.cpp:33-75: bilinear sample at (x + offx, y + offy) if all four taps
    are inside, else src(x, y).  The offset's integer part is truncated and moved
    down by one when negative, in both forms."""
    lib = load_oracle()
    cast = _f32 if dtype == np.float32 else _f64
    s, ox, oy = cast(src), cast(offx), cast(offy)
    H, W = s.shape
    d = np.empty_like(s)
    fn = lib.or_warp_img2d_f32 if dtype == np.float32 else lib.or_warp_img2d_f64
    fn(_p(s), _p(ox), _p(oy), _p(d), H, W)
    return d


def brox_set_matrix(A, psi, alpha, redblack=False):
    """SetBroxMatrix.cpp:34-79 / SetBroxRedblackMatrix.cpp:37-81 (fp32): returns the
    updated planes (a copy of A) and, for the red-black form, the diagonal."""
    lib = load_oracle()
    A = _f32(A).copy()
    psi = _f32(psi)
    H, W = psi.shape
    diag = np.zeros((H, W), np.float32)
    lib.or_brox_set_f32(_p(A), _p(diag), _p(psi), H, W, ctypes.c_float(alpha),
                        1 if redblack else 0)
    return (A, diag) if redblack else A


def brox_precondition(A, diag0, diag1, offdiag, x0, x1, dtype=np.float32):
    """ApplyBroxPreconditioner.cpp:46-72 with A2 = plane 2 of the matrix A."""
    lib = load_oracle()
    cast = _f32 if dtype == np.float32 else _f64
    a2 = cast(np.asarray(A)[2])
    d0, d1, od, e, f = (cast(t) for t in (diag0, diag1, offdiag, x0, x1))
    y0, y1 = np.empty_like(e), np.empty_like(e)
    fn = lib.or_brox_precond_f32 if dtype == np.float32 else lib.or_brox_precond_f64
    fn(_p(a2), _p(d0), _p(d1), _p(od), _p(e), _p(f), _p(y0), _p(y1), ctypes.c_int64(e.size))
    return y0, y1


def red(x):
    """RedArray2D.cpp:20-24: x where (column + row) is even, else 0."""
    x = np.asarray(x)
    rr, cc = np.indices(x.shape)
    return np.where((rr + cc) & 1, 0, x).astype(x.dtype)


def black(x):
    """BlackArray2D.cpp:20-24: x where (column + row) is odd, else 0."""
    x = np.asarray(x)
    rr, cc = np.indices(x.shape)
    return np.where((rr + cc) & 1, x, 0).astype(x.dtype)


def endpoint_error(u, v, fx, fy, margin=8):
    """Mean endpoint error over the interior (``margin`` px from each border)."""
    sl = (slice(margin, -margin), slice(margin, -margin))
    du = np.asarray(u, np.float64)[sl] - np.asarray(fx, np.float64)[sl]
    dv = np.asarray(v, np.float64)[sl] - np.asarray(fy, np.float64)[sl]
    return float(np.mean(np.sqrt(du * du + dv * dv)))
