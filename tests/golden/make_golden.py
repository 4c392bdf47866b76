"""Write the golden fixtures under tests/golden/ (run in the build container,
where the reference tree is mounted; the fixtures are committed and the
tests read only them).

    python tests/golden/make_golden.py [--reference /root/reference]

Fixtures (numpy .npz, no pickles):
  filters_h.npz   -- the 9,010 texton coefficients and 34 radii the reference
                     ships as a table (include/filters.h:22-23).
  savgol.npz      -- the Savitzky-Golay parabola filters the reference ships
                     as data (damascene/data/savgol_{3,5,10,20}_{1..8}.dat).
  ref_build.npz   -- outputs of the reference's own host code, compiled from
                     its sources into oracle/_ref by `make -C oracle ref`:
                     createTextonFilters (textons/filters.cpp:404-454), the
                     eight spectralPb kernels built the way
                     sPb/spectralPb.cu:708-711 builds them (gaussian_2D,
                     filters.cpp:261-336), and Stencil offsets/diagonal maps
                     (stencilMatrixMultiply/Stencil.cpp:6-67).
  combine_tables.npz -- the cue weights of combine/combine.cu:16-21 for 32
                     and 64 textons (mPb coefficients, combinedG weights;
                     the weights arrays carry a 13th entry, the sPb weight,
                     kept as read).
"""
import argparse
import ctypes
import math
import os
import re
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def parse_c_array(text, name, dtype):
    m = re.search(name + r"\[\]\s*=\s*\{(.*?)\}", text, re.S)
    if not m:
        raise SystemExit(f"{name} not found")
    vals = [v for v in m.group(1).replace("\n", " ").split(",") if v.strip()]
    return np.array([float(v) for v in vals]).astype(dtype)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--only", choices=["combine"], help="write one fixture only")
    a = ap.parse_args()
    ref = a.reference
    if not os.path.isdir(ref):
        raise SystemExit(f"reference tree {ref} not present; fixtures are already committed")

    text = open(os.path.join(ref, "combine", "combine.cu")).read()
    tables = {k: parse_c_array(text, k, np.float32)
              for k in ("hostCoefficients", "hostCoefficients_64", "hostweights", "hostweights_64")}
    np.savez_compressed(os.path.join(HERE, "combine_tables.npz"), **tables)
    if a.only == "combine":
        return 0

    text = open(os.path.join(ref, "include", "filters.h")).read()
    coeffs = parse_c_array(text, "hCoefficients", np.float32)
    radii = parse_c_array(text, "hRadii", np.int32)  # given in a comment in the header
    assert coeffs.size == 9010 and radii.size == 34, (coeffs.size, radii.size)
    np.savez_compressed(os.path.join(HERE, "filters_h.npz"), coeffs=coeffs, radii=radii)

    sg = {}
    for r in (3, 5, 10, 20):
        for k in range(1, 9):
            d = np.fromfile(os.path.join(ref, "damascene", "data", f"savgol_{r}_{k}.dat"), np.float32)
            assert d.size == (2 * r + 1) ** 2
            sg[f"r{r}_k{k}"] = d.reshape(2 * r + 1, 2 * r + 1)
    np.savez_compressed(os.path.join(HERE, "savgol.npz"), **sg)

    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", f"REF={ref}"])
    L = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_host.so"))
    P = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    tc = np.zeros(9010, np.float32)
    tr = np.zeros(34, np.int32)
    L.ref_texton_filters(P(tc), P(tr))
    L.ref_gaussian_2d.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_uint, ctypes.c_int, ctypes.c_ulong, ctypes.c_ulong]
    spb = np.zeros((8, 49), np.float32)
    for ori in range(7, -1, -1):
        f = np.zeros(49, np.float32)
        L.ref_gaussian_2d(P(f), 1.0, 1.0, math.pi / 2 + ori * math.pi / 8, 1, 0, 3, 3)
        spb[7 - ori] = f
    st = {}
    for w in (7, 33, 160, 481):
        offs = np.zeros(256, np.int32)
        diag = np.zeros(121, np.int32)
        area = L.ref_stencil(5, w, 11, P(offs), P(diag))
        st[f"offsets_w{w}"] = offs[:area]
        st[f"diag_w{w}"] = diag
    np.savez_compressed(os.path.join(HERE, "ref_build.npz"), texton_coeffs=tc, texton_radii=tr, spb=spb, **st)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    sys.exit(main())
