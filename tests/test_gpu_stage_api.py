"""The reference-named C++ stage API on the GPU.

tests/stage_api_gpb (built from tests/stage_api_gpb.cpp by build()) runs
damascene.cu:269-470's call sequence -- loadPPM_rgbU, rgbUtoGreyF,
findTextons, rgbUtoLab3F, normalizeLab, localCues, combine,
nonMaxSuppression, Stencil, intervene, generalizedEigensolve, spectralPb,
StartCalcGPb -- with the reference's signatures and ownership (callee
hipMallocs the ** outputs, host eigenvalues), in a process of its own.  Its
textons, nonmax mask, eigenvalues, gPbAll and gPb must equal dmsc_gpb's
(plain FP32 path, the reference algorithm) on the same image bit for bit.
"""
import os
import subprocess

import numpy as np
import pytest
import torch

import damascene_amd as dm

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "stage_api_gpb")


def write_ppm(path, rgb):
    h, w = rgb.shape
    px = np.stack([rgb & 0xFF, (rgb >> 8) & 0xFF, (rgb >> 16) & 0xFF], axis=-1).astype(np.uint8)
    with open(path, "wb") as f:
        f.write(f"P6\n{w} {h}\n255\n".encode() + px.tobytes())


@pytest.mark.parametrize("w,h,cells,seed", [(160, 120, 12, 20261202), (481, 321, 40, 20261202)])
def test_stage_api_program_matches_dmsc_gpb(ctx, tmp_path, w, h, cells, seed):
    assert os.path.exists(EXE), "tests/stage_api_gpb is not built (run __graft_entry__.build())"
    print(f"case: stage API program {w}x{h} seed {seed}", flush=True)
    rgb = dm.generate_image(w, h, cells, seed)
    ppm = tmp_path / "in.ppm"
    write_ppm(ppm, rgb)
    assert np.array_equal(dm.load_ppm(ppm), rgb)
    prefix = str(tmp_path / "out")
    r = subprocess.run([EXE, str(ppm), prefix], capture_output=True, text=True, timeout=180)
    print(r.stdout.strip(), r.stderr.strip()[-500:], flush=True)
    assert r.returncode == 0
    n = w * h
    got = {
        "textons": np.fromfile(prefix + ".textons", np.int32).reshape(h, w),
        "nmax": np.fromfile(prefix + ".nmax", np.float32).reshape(h, w),
        "gpb": np.fromfile(prefix + ".gpb", np.float32).reshape(h, w),
        "gpb_all": np.fromfile(prefix + ".gpball", np.float32).reshape(8, n),
        "eigvals": np.fromfile(prefix + ".eigvals", np.float32),
    }
    out = ctx.gpb(ctx.upload(rgb), precision=dm.PREC_FP32)
    torch.cuda.synchronize()
    want = {
        "textons": out["textons"].cpu().numpy(),
        "nmax": out["nmax"].cpu().numpy(),
        "gpb": out["gpb"].cpu().numpy(),
        "gpb_all": out["gpb_all"].cpu().numpy()[:, :n],
        "eigvals": out["eigvals"],
    }
    for k in want:
        a, b = got[k], want[k]
        assert a.shape == b.shape, k
        bad = int(np.count_nonzero(a.view(np.uint32) != b.view(np.uint32)))
        assert bad == 0, f"{k}: {bad} of {a.size} values differ from dmsc_gpb"
