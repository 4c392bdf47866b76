"""FP64 ground truth for the 481x321 workload (synthetic image: 40 Voronoi
cells, seed 20261202), written to tests/golden/truth_481x321.npz; with a
seed argument (e.g. 20261204) to tests/golden/truth_481x321_s<seed>.npz.

The CPU oracle produces nmax, the affinity matrix and L_sym (bit-identical to
the GPU path).  Then:
  * the 10 smallest eigenpairs of L_sym come from scipy's shift-invert
    eigsh in FP64 (sigma = -1e-3, tol 1e-12), an eigensolver independent of
    both the oracle's Lanczos and the GPU's;
  * gPb comes from those eigenvectors through the oracle's finishing stages
    (D^-1/2 scaling, min-max, spectral Pb, gPb).

Saved: eigenvalues (10, FP64); u_1..u_8 (the non-trivial eigenvectors of
L_sym, unit norm, FP32); sqrt(d) (row sums, FP32); gPb (FP32).

    python tests/golden/make_truth.py [seed]      (about 6 minutes on 8 cores)
"""
import os
import sys
import time

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import damascene_amd as dm  # noqa: E402  (host-side generator only)
from oracle import pyoracle as O  # noqa: E402

W, H, CELLS, SEED, NVEC = 481, 321, 40, 20261202, 9


def main():
    t0 = time.time()
    rgb = dm.generate_image(W, H, CELLS, SEED)
    g = O.rgb_to_grey(rgb)
    tex, _ = O.find_textons(g)
    L, a, b = O.rgb_to_lab(rgb)
    mpb, cg = O.combine(*O.local_cues(L, a, b, tex))
    nm = O.nonmax(mpb)
    Wm = O.intervene(nm)
    lsym, d, rs = O.convert_matrix(W, H, Wm)
    n = W * H
    offs, _ = O.stencil(5, W)
    lin = np.cumsum(offs[:81])
    rows, cols, vals = [], [], []
    for k in range(81):
        r = np.arange(n)
        c = r + lin[k]
        ok = (c >= 0) & (c < n)
        rows.append(r[ok])
        cols.append(c[ok])
        vals.append(lsym[k][ok].astype(np.float64))
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    ev, U = sla.eigsh(A, k=NVEC + 1, sigma=-1e-3, which="LM", tol=1e-12)
    o = np.argsort(ev)
    ev, U = ev[o], U[:, o]
    res = np.abs(A @ U - U * ev).max()
    vecs = O.finish_eigvecs(W, H, rs, U[:, :NVEC].T.astype(np.float32).copy())
    spb = O.spectral_pb(ev[:NVEC].astype(np.float32), vecs, W, H)
    _, gpb = O.gpb(cg, spb, nm)
    name = "truth_481x321.npz" if SEED == 20261202 else f"truth_481x321_s{SEED}.npz"
    np.savez_compressed(os.path.join(HERE, name), eigvals=ev, u=U[:, 1:NVEC].T.astype(np.float32),
                        sqrt_d=np.sqrt(d.astype(np.float64)).astype(np.float32), gpb=gpb.astype(np.float32),
                        meta=np.array([W, H, CELLS, SEED, NVEC]))
    print(f"eigenvalues {ev}\nmax residual {res:.2e}; {time.time() - t0:.1f} s")


if __name__ == "__main__":
    if len(sys.argv) > 1:
        SEED = int(sys.argv[1])
    main()
