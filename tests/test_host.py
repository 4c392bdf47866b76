"""CPU tests of the product library's host side (no GPU needed): the C ABI
loads and exports every function include/dmsc.h declares, its host-built
tables are bit-identical to the oracle's, the synthetic image generator is
deterministic and shaped as specified, and calls that need a device fail
with a status code instead of crashing.
"""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import damascene_amd as dm
from damascene_amd._lib import lib
from oracle import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
P = O.P


@pytest.fixture(scope="module", autouse=True)
def built():
    O.build()
    if not os.path.exists(os.path.join(ROOT, "damascene_amd", "libdmsc.so")):
        pytest.fail("libdmsc.so is not built (python -m damascene_amd.build)")


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "dmsc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dmsc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_symbols()
    assert len(names) >= 30
    l = ctypes.CDLL(os.path.join(ROOT, "damascene_amd", "libdmsc.so"))
    missing = [n for n in names if not hasattr(l, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "damascene_amd", "libdmsc.so")],
                         capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (dmsc_[a-z0-9_]+)", out))
    assert set(names) <= exported


def test_version_and_pitch():
    assert lib().dmsc_version().startswith(b"dmsc")
    for n in (1, 63, 64, 65, 19200, 154401):
        p = dm.pitch_in_floats(n)
        assert p >= n and p % 64 == 0 and p - n < 64


def test_no_device_is_an_error_not_a_crash():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    c = ctypes.c_void_p()
    rc = lib().dmsc_ctx_create(ctypes.byref(c), 0, None)
    assert rc != 0
    assert b"device" in lib().dmsc_last_error_string().lower()


def test_texton_and_parabola_and_spb_tables_match_oracle():
    c1, r1, n1 = O.texton_filters()
    c2 = np.zeros(9010, np.float32)
    r2 = np.zeros(34, np.int32)
    assert lib().dmsc_texton_filters(P(c2), P(r2)) == n1
    np.testing.assert_array_equal(c1.view(np.uint32), c2.view(np.uint32))
    np.testing.assert_array_equal(r1, r2)
    for r in (3, 5, 10, 20):
        a = O.parabola_filters(r)
        b = np.zeros_like(a)
        lib().dmsc_parabola_filters(r, P(b))
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    a = O.spb_filters()
    b = np.zeros_like(a)
    lib().dmsc_spb_filters(P(b))
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("W,H", [(541, 381), (220, 180), (1084, 828), (100, 100)])
def test_rotation_and_halfdisc_tables_match_oracle(W, H):
    for o in range(4):
        tp = -float(o) / 8
        nw, nh, bx, by, gx, gy = O.rotation_tables(W, H, tp)
        n = max(W, H) * 4 + 16
        B = [np.zeros(n, np.int32) for _ in range(4)]
        nw2, nh2 = ctypes.c_int(), ctypes.c_int()
        assert lib().dmsc_rotation_tables(W, H, ctypes.c_float(tp), ctypes.byref(nw2), ctypes.byref(nh2),
                                          *[P(x) for x in B]) == 0
        assert (nw, nh) == (nw2.value, nh2.value)
        for x, y in zip(B, (bx, by, gx, gy)):
            np.testing.assert_array_equal(x[:len(y)], y)
        for radius in (3, 5, 10, 20):
            want = np.zeros(8, np.int32)
            O.lib().orc_halfdisc_offsets(W - 60, H - 60, ctypes.c_float(tp), radius, P(bx), P(by), P(gx), P(gy),
                                         nw, P(want))
            got = np.zeros(8, np.int32)
            assert lib().dmsc_halfdisc_offsets(W - 60, H - 60, ctypes.c_float(tp), radius, P(B[0]), P(B[1]),
                                               P(B[2]), P(B[3]), nw, P(got)) == 0
            np.testing.assert_array_equal(got, want)


def test_blur_gamma_stencil_tables_match_oracle():
    for s in (2.0, 1.0):
        k, r = O.blur_kernel(s)
        k2 = np.zeros(64, np.float32)
        rr = ctypes.c_int()
        n = lib().dmsc_blur_kernel(ctypes.c_float(s), P(k2), ctypes.byref(rr))
        assert n == k.size and rr.value == r
        np.testing.assert_array_equal(k, k2[:n])
    g1 = O.gamma_lut(2.5)
    g2 = np.zeros(256, np.float32)
    lib().dmsc_gamma_lut(ctypes.c_float(2.5), P(g2))
    np.testing.assert_array_equal(g1.view(np.uint32), g2.view(np.uint32))
    for w in (7, 160, 481):
        o1, d1 = O.stencil(5, w)
        o2 = np.zeros(128, np.int32)
        d2 = np.zeros(121, np.int32)
        assert lib().dmsc_stencil(5, w, P(o2), P(d2)) == 81
        np.testing.assert_array_equal(o1[:81], o2[:81])
        np.testing.assert_array_equal(d1[:121], d2)


def test_generator_is_deterministic_and_shaped():
    a, cells = dm.generate_image(160, 120, 12, 20261202, with_cells=True)
    b = dm.generate_image(160, 120, 12, 20261202)
    c = dm.generate_image(160, 120, 12, 20261203)
    assert a.shape == (120, 160) and a.dtype == np.uint32
    np.testing.assert_array_equal(a, b)
    assert (a != c).any()
    assert (a >> 24 == 0).all()  # RGB8 packed in the low 24 bits
    assert set(np.unique(cells)) == set(range(12))
    m = dm.boundary_mask(cells)
    # boundaries = cell changes between 4-neighbours, dilated by 1 px
    edge = np.zeros_like(cells, bool)
    edge[:, 1:] |= cells[:, 1:] != cells[:, :-1]
    edge[1:, :] |= cells[1:, :] != cells[:-1, :]
    assert m[edge].all()
    assert 0.02 < m.mean() < 0.5


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stdout


def test_ppm_loader_packs_rgb(tmp_path):
    w, h = 7, 5
    px = np.random.default_rng(3).integers(0, 256, size=(h, w, 3), dtype=np.uint8)
    path = tmp_path / "t.ppm"
    path.write_bytes(b"P6\n# comment\n7 5\n255\n" + px.tobytes())
    W, H = ctypes.c_int(), ctypes.c_int()
    assert lib().dmsc_load_ppm(str(path).encode(), ctypes.byref(W), ctypes.byref(H), None) == 0
    assert (W.value, H.value) == (w, h)
    out = np.zeros((h, w), np.uint32)
    assert lib().dmsc_load_ppm(str(path).encode(), ctypes.byref(W), ctypes.byref(H), P(out)) == 0
    want = px[..., 0].astype(np.uint32) | (px[..., 1].astype(np.uint32) << 8) | (px[..., 2].astype(np.uint32) << 16)
    np.testing.assert_array_equal(out, want)
    bad = tmp_path / "bad.ppm"
    bad.write_bytes(b"P5\n7 5\n255\n" + px.tobytes())
    assert lib().dmsc_load_ppm(str(bad).encode(), ctypes.byref(W), ctypes.byref(H), None) != 0
    np.testing.assert_array_equal(dm.load_ppm(path), want)  # the Python wrapper
    with pytest.raises(dm.DmscError):
        dm.load_ppm(bad)


STAGE_API_PROGRAM = r"""
#include <cstdio>
#include <stdexcept>
#include "Stencil.h"
#include "combine.h"
#include "convert.h"
#include "globalPb.h"
#include "intervening.h"
#include "kmeans.h"
#include "lanczos.h"
#include "localcues.h"
#include "nonmax.h"
#include "skeleton.h"
#include "spectralPb.h"
#include "stencilMVM.h"
#include "texton.h"
int main() {
  Stencil s(5, 481, 321, findPitchInFloats(481 * 321));
  static int offs[STENCILAREAMAX], diag[CONSTSPACE];
  s.copyOffsets(offs);
  s.copyDiagonalOffsets(diag);
  if (offs[s.getStencilArea()] != -2000000000 || offs[STENCILAREAMAX - 1] != -2000000000) return 3;
  long lin = 0;
  for (unsigned i = 0; i < s.getStencilArea(); i++) lin += offs[i];
  // last diagonal is (dy, dx) = (5, 0); entry (row, row + 5w) sits on it
  int idx = s.matrixIndex(1000, 1000 + 5 * 481);
  std::printf("area %u pitch %d last %ld index %d\n", s.getStencilArea(), s.getMatrixPitchInFloats(), lin, idx);
  try {
    int* t = nullptr;
    findTextons(481, 321, nullptr, &t, 1);
  } catch (const std::runtime_error& e) {
    std::printf("error: %s\n", e.what());
    return 0;
  }
  return 1;
}
"""


def test_reference_stage_api_compiles_links_and_reports_errors(tmp_path):
    """The reference-named C++ headers build with the host compiler alone and
    link against libdmsc.so; without a device a stage call throws."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    src = tmp_path / "use_api.cpp"
    src.write_text(STAGE_API_PROGRAM)
    exe = tmp_path / "use_api"
    libdir = os.path.join(ROOT, "damascene_amd")
    cmd = ["g++", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"), "-I",
           "/opt/rocm/include", str(src), "-o", str(exe), "-L", libdir, "-ldmsc", f"-Wl,-rpath,{libdir}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    pitch = dm.pitch_in_floats(481 * 321)
    assert f"area 81 pitch {pitch} last {5 * 481} index {pitch * 80 + 1000}" in r.stdout
    assert "error:" in r.stdout


READ_SMA_PROGRAM = r"""
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include "Stencil.h"
int main(int argc, char** argv) {
  Stencil s(5, 12, 10, 128);
  try {
    float* m = s.readStencilMatrix(argv[1]);
    FILE* f = std::fopen(argv[2], "wb");
    std::fwrite(m, sizeof(float), s.getStencilArea() * 128, f);
    std::fclose(f);
    std::free(m);
  } catch (const std::runtime_error& e) {
    std::printf("error: %s\n", e.what());
    return 1;
  }
  return 0;
}
"""


def write_sma(path, rows):
    """Stencil.cpp:92-145 .sma layout: int n, int nnz, then per row: int
    count, double values[count], int columns[count]."""
    nnz = sum(len(r) for r in rows)
    with open(path, "wb") as f:
        f.write(np.array([len(rows), nnz], np.int32).tobytes())
        for r in rows:
            f.write(np.array([len(r)], np.int32).tobytes())
            f.write(np.array([v for _, v in r], np.float64).tobytes())
            f.write(np.array([c for c, _ in r], np.int32).tobytes())


def test_read_stencil_matrix(tmp_path):
    """Stencil::readStencilMatrix places each (row, column, value) of a .sma
    file at its DIA slot (the diagonal of offset (dy, dx) = column - row split
    across row wrap, Stencil.cpp:41-59) as a float, zero elsewhere, and
    rejects a file whose row count is not width * height."""
    w, h, pitch = 12, 10, 128
    n = w * h
    offs = [(dy, dx) for dy in range(-5, 6) for dx in range(-5, 6) if dy * dy + dx * dx <= 25]
    rng = np.random.default_rng(17)
    rows, want = [], np.zeros((81, pitch), np.float32)
    for row in range(n):
        entries = []
        for d in rng.choice(81, size=6, replace=False):
            dy, dx = offs[d]
            col = row + dy * w + dx  # linear column: wraps across image rows like the reference
            if 0 <= col < n:
                v = float(rng.standard_normal())
                entries.append((col, v))
                want[d, row] = np.float32(v)
        rows.append(entries)
    sma = tmp_path / "m.sma"
    write_sma(sma, rows)
    src = tmp_path / "read_sma.cpp"
    src.write_text(READ_SMA_PROGRAM)
    exe = tmp_path / "read_sma"
    libdir = os.path.join(ROOT, "damascene_amd")
    subprocess.run(["g++", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"), "-I",
                    "/opt/rocm/include", str(src), "-o", str(exe), "-L", libdir, "-ldmsc", f"-Wl,-rpath,{libdir}"],
                   check=True, capture_output=True, text=True)
    out = tmp_path / "m.f32"
    r = subprocess.run([str(exe), str(sma), str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(out, np.float32).reshape(81, pitch)
    np.testing.assert_array_equal(got, want)
    write_sma(sma, rows[:-1])
    r = subprocess.run([str(exe), str(sma), str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "rows for a 12x10 stencil" in r.stdout


def test_result_file_formats(tmp_path):
    """damascene.cu writeFile (.pb) and writeArray (GpbAll.ary: orientation
    order {3,2,1,0,7,6,5,4}, planes transposed), and the 8-bit PGM."""
    w, h = 5, 3
    rng = np.random.default_rng(9)
    g = rng.random((h, w), dtype=np.float32)
    pb = tmp_path / "x.pb"
    assert lib().dmsc_write_pb(str(pb).encode(), w, h, P(g)) == 0
    raw = pb.read_bytes()
    assert np.frombuffer(raw[:8], np.int32).tolist() == [w, h]
    np.testing.assert_array_equal(np.frombuffer(raw[8:], np.float32).reshape(h, w), g)
    pitch = 64
    planes = np.zeros((8, pitch), np.float32)
    planes[:, :w * h] = rng.random((8, w * h), dtype=np.float32)
    ary = tmp_path / "GpbAll.ary"
    assert lib().dmsc_write_gpb_all(str(ary).encode(), w, h, pitch, P(planes)) == 0
    raw = ary.read_bytes()
    hdr = np.frombuffer(raw[:16], np.int32)
    assert hdr.tolist() == [3, 8, w, h]
    data = np.frombuffer(raw[16:], np.float32).reshape(8, w, h)
    for i, o in enumerate([3, 2, 1, 0, 7, 6, 5, 4]):
        np.testing.assert_array_equal(data[i], planes[o, :w * h].reshape(h, w).T)
    pgm = tmp_path / "Pb.pgm"
    g[0, 0], g[0, 1] = 1.5, -0.2
    assert lib().dmsc_write_pgm(str(pgm).encode(), w, h, P(g)) == 0
    raw = pgm.read_bytes()
    head = b"P5\n5 3\n255\n"
    assert raw.startswith(head)
    px = np.frombuffer(raw[len(head):], np.uint8).reshape(h, w)
    np.testing.assert_array_equal(px, (np.clip(g, 0, 1) * np.float32(255)).astype(np.uint8))


def test_combine_weights_match_reference_tables():
    """dmsc_combine_weights (the kernels' tables) against the reference's own
    combine.cu:16-21 arrays (fixture read from the source text)."""
    t = np.load(os.path.join(ROOT, "tests", "golden", "combine_tables.npz"))
    for choice, suffix in ((dm.TEXTON32, ""), (dm.TEXTON64, "_64")):
        coef, wts = (ctypes.c_float * 12)(), (ctypes.c_float * 12)()
        assert lib().dmsc_combine_weights(choice, coef, wts) == 0
        np.testing.assert_array_equal(np.array(coef[:], np.float32), t["hostCoefficients" + suffix].astype(np.float32))
        np.testing.assert_array_equal(np.array(wts[:], np.float32), t["hostweights" + suffix][:12].astype(np.float32))
    assert lib().dmsc_combine_weights(3, coef, wts) == 1  # DMSC_ERR_ARG


def test_new_entry_points_reject_bad_arguments_without_a_device():
    """Argument checks come before any device work: a null context or a bad
    tuning value is an error status on a machine without a GPU."""
    L = lib()
    assert L.dmsc_gpb_batch(None, 0, None, None, 0) == 1
    assert L.dmsc_spb_combine_forward(None, 64, 64, 4096, None, None, None, None, 9, None, None, None, 0, None,
                                      None) == 1
    assert L.dmsc_autotune_eigensolver(None, 64, 64, None, 4096, 0, 100, None, None, 0, None) == 1
    assert L.dmsc_kmeans_rechecks(None, None) == 1
    assert L.dmsc_chebyshev_time_products(None, 64, 64, None, 4096, 10, None) == 1
    assert L.dmsc_autotune_chebyshev(None, 64, 64, None, 4096, 10, None, None, None, 0, None) == 1
    assert L.dmsc_chebyshev_tile(None, 64, 64, None) == 1
    for key, bad in ((b"lanczos_tile", 6405), (b"lanczos_max_iter", 10), (b"lanczos_mixed_stop", 9),
                     (b"lanczos_check_max", 30), (b"batch_lanes", 0), (b"eigensolver", 2),
                     (b"chebyshev_shape", 44), (b"chebyshev_shape", 346), (b"integral_pass", 2),
                     (b"chebyshev_start", 2), (b"chebyshev_first_degree", 1), (b"chebyshev_first_degree", 301)):
        assert L.dmsc_tune_set(key, bad) == 1, key
    v = ctypes.c_int()
    for key, want in ((b"lanczos_tile", 3208), (b"lanczos_max_iter", 6000), (b"lanczos_mixed_stop", 4),
                      (b"lanczos_check_max", 0), (b"batch_lanes", 4), (b"eigensolver", 1),
                      (b"chebyshev_shape", 0), (b"integral_pass", 1), (b"chebyshev_reduction", 10),
                      (b"chebyshev_start", 1), (b"chebyshev_first_degree", 50)):
        assert L.dmsc_tune_get(key, ctypes.byref(v)) == 0 and v.value == want, key
    assert L.dmsc_kmeans_noalloc(None, 64, 64, 32, 34, None, None, 15, 0, 0, None, None) == 1



@pytest.mark.parametrize("n,graded,seed", [(1, False, 0), (2, False, 1), (5, True, 2), (32, False, 3), (32, True, 4),
                                            (32, True, 5), (64, False, 6)])
def test_symmetric_eigen_against_numpy(n, graded, seed):
    """The block eigensolver's Rayleigh-Ritz eigensolve (Householder +
    implicit symmetric QR, host FP64) against numpy's LAPACK eigh: sorted
    eigenvalues, eigenvector residuals |A z - lam z| and orthonormality, on
    random symmetric matrices and on graded ones like the solver's (diagonal
    from 1e-8 to 1 with small couplings, and a repeated eigenvalue)."""
    rng = np.random.default_rng(seed)
    if graded:
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        lam = np.logspace(-8, 0, n)
        if n > 3:
            lam[2] = lam[1]  # a double eigenvalue
        A = (Q * lam) @ Q.T
    else:
        A = rng.standard_normal((n, n))
    A = 0.5 * (A + A.T)
    ev = np.zeros(n)
    Z = np.zeros((n, n))
    assert lib().dmsc_symmetric_eigen(n, A.ctypes.data, ev.ctypes.data, Z.ctypes.data) == 0
    ref = np.linalg.eigvalsh(A)
    scale = np.abs(ref).max()
    assert np.max(np.abs(np.sort(ev) - ref)) <= 1e-13 * scale
    assert np.max(np.linalg.norm(A @ Z - Z * ev[None, :], axis=0)) <= 1e-13 * scale
    assert np.max(np.abs(Z.T @ Z - np.eye(n))) <= 1e-13


def test_symmetric_eigen_rejects_bad_arguments():
    x = np.zeros(4)
    assert lib().dmsc_symmetric_eigen(0, x.ctypes.data, x.ctypes.data, x.ctypes.data) == 1
    assert lib().dmsc_symmetric_eigen(2, None, x.ctypes.data, x.ctypes.data) == 1
