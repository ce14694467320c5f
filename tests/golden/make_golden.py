"""Write the golden fixtures under tests/golden/ (run: python tests/golden/make_golden.py).

The reference ships no tests, fixtures or golden vectors, and cannot run here
(Python-2 C extension + OpenCL without a device).  These fixtures are computed
with numpy / scipy only, never with the C oracle, so they are an independent
anchor for both the oracle (tests/test_golden.py) and the GPU kernels
(tests/test_gpu_golden.py).  All expected values are fp64.

Each fixture follows the reference kernel it pins:
  stencil   Stencil2DMVM.cpp:36-47  (clamped taps)   -> scipy.ndimage.correlate(mode="nearest")
  dia2d     SpmvDIA2D.cpp:35-45, SpMat2D.cpp:211-216 -> scipy.sparse matrix from (p, p+o, A[d][p+o])
  csr       SpmvCSR.cpp:32-38                        -> scipy.sparse.csr_matrix @ x
  sum2d     Array2DSum.cpp / hlib.py:83-84           -> math.fsum
  hs        canonical HS program (Jacobi)            -> numpy textbook iteration
"""
import math
import os
import sys

import numpy as np
import scipy.ndimage as ndi
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from hindemith_amd import synthetic as S  # noqa: E402


def kernel3(st, R):
    w, ox, oy = st
    K = np.zeros((2 * R + 1, 2 * R + 1))
    for d in range(len(w)):
        K[R + oy[d], R + ox[d]] += float(w[d])
    return K


def corr(x, st, R=1):
    return ndi.correlate(np.asarray(x, np.float64), kernel3(st, R), mode="nearest")


def main():
    rng = np.random.default_rng(20261202)
    taps = S.hs_stencil_taps()
    out = {}
    # stencils on an odd-sized image
    x = rng.uniform(-1, 1, (9, 13)).astype(np.float32)
    out["stencil_x"] = x
    for name in ("Avg", "Dx", "Dy", "Dt"):
        out[f"stencil_{name}_y"] = corr(x, taps[name])
    lap = S.laplacian5_taps()
    out["stencil_Lap5_y"] = corr(x, lap)
    # 2D DIA with arbitrary planes
    H, W = 6, 7
    A = rng.normal(size=(5, H, W)).astype(np.float32)
    xd = rng.uniform(-1, 1, (H, W)).astype(np.float32)
    w, ox, oy = lap
    rows, cols, vals = [], [], []
    for r in range(H):
        for c in range(W):
            for d in range(5):
                r2, c2 = r + oy[d], c + ox[d]
                if 0 <= r2 < H and 0 <= c2 < W:
                    rows.append(r * W + c)
                    cols.append(r2 * W + c2)
                    vals.append(float(A[d, r2, c2]))
    M = sp.csr_matrix((vals, (rows, cols)), shape=(H * W, H * W))
    out.update(dia_A=A, dia_offx=ox, dia_offy=oy, dia_x=xd,
               dia_y=(M @ xd.reshape(-1).astype(np.float64)).reshape(H, W))
    # CSR with empty rows
    C = sp.random(40, 30, density=0.15, random_state=3, format="csr", dtype=np.float32)
    C[5, :] = 0
    C.eliminate_zeros()
    xc = rng.uniform(-1, 1, 30).astype(np.float32)
    out.update(csr_indptr=C.indptr.astype(np.int32), csr_indices=C.indices.astype(np.int32),
               csr_data=C.data.astype(np.float32), csr_x=xc,
               csr_y=C.astype(np.float64) @ xc.astype(np.float64))
    # sum2d
    s = rng.uniform(-1, 1, (33, 17)).astype(np.float32)
    out.update(sum_x=s, sum_y=np.array(math.fsum(s.astype(np.float64).reshape(-1))))
    # Horn-Schunck: 24 x 32 blob pair, alpha^2 = 1, 20 iterations
    I1, I2, fx, fy = S.hs_pair(24, 32, seed=7)
    I1d, I2d = I1.astype(np.float64), I2.astype(np.float64)
    Ix = corr(I1d, taps["Dx"]) + corr(I2d, taps["Dx"])
    Iy = corr(I1d, taps["Dy"]) + corr(I2d, taps["Dy"])
    It = corr(I2d, taps["Dt"]) - corr(I1d, taps["Dt"])
    den = 1.0 + Ix * Ix + Iy * Iy
    u = np.zeros_like(Ix)
    v = np.zeros_like(Ix)
    for _ in range(20):
        ub, vb = corr(u, taps["Avg"]), corr(v, taps["Avg"])
        t = (Ix * ub + Iy * vb + It) / den
        u, v = ub - Ix * t, vb - Iy * t
    out.update(hs_I1=I1, hs_I2=I2, hs_Ix=Ix, hs_Iy=Iy, hs_It=It, hs_u=u, hs_v=v,
               hs_alpha2=np.array(1.0), hs_iters=np.array(20))
    np.savez_compressed(os.path.join(HERE, "golden_hs.npz"), **out)
    print("wrote", os.path.join(HERE, "golden_hs.npz"), len(out), "arrays")


if __name__ == "__main__":
    main()
