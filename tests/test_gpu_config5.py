"""The mode-2 gradient at survey scale (160x160x128 multigrid, 201x201x126
input model, three frequencies) on the GPU against the reference.

The reference's own sources, run as one rank of its mode-2 job
(oracle/ref_runner task=fwi: fg_fwi_init and fg_fwi of src/fg_fwi.c:43-285
on one rank, addnoise=0) for shot 9 of the survey, solved to a relative
residual of 1e-10, wrote tests/golden/config5/ (make_survey_golden.py): the
observed data emf_0009.txt (the reference's converged receiver data of the
true earth) and, on the starting model, the synthetic data, the misfit and
both gradient volumes (sampled every 4th cell per axis, the y = 0 planes and
the full-volume norms).  emmi_main runs the same rank with FP32 cycles inside
FP64 refinement (the default) to the same tolerance; the converged gradients
agree within 1e-5 of their maximum."""
import gzip
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from libemmi_mgfd_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "libemmi_mgfd_amd", "emmi_main")
GOLD = os.path.join(ROOT, "tests", "golden", "config5")

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not os.path.exists(EXE), reason="emmi_main not built")]


def test_gradient_160x160x128_matches_reference(tmp_path):
    d = np.load(os.path.join(GOLD, "config5.npz"))
    shot = int(d["shot"])
    s = configs.config5(true_model=False)
    files = configs.write_reference_inputs(s, str(tmp_path / "in"))
    run = tmp_path / "run"
    run.mkdir()
    name = f"emf_{shot:04d}.txt"
    with gzip.open(os.path.join(GOLD, name + ".gz"), "rb") as src, open(run / name, "wb") as dst:
        shutil.copyfileobj(src, dst)
    args = [EXE, "mode=2"] + s.getpar_args(files) + [f"shots={shot}", "chrec=Ex,Ey,Hx,Hy", "addnoise=0",
                                                       "tol=1e-10", "ncycle=200", "precision=1"]
    print(f"emmi_main mode=2, shot {shot}, MG {s.n1}x{s.n2}x{s.n3}, FP32 cycles in FP64 refinement", flush=True)
    r = subprocess.run(args, cwd=run, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "not converged" not in r.stdout
    m = re.search(r"isrc=%d, fcost=(\S+)," % shot, r.stdout)
    fcost = float(m.group(1))
    e_f = abs(fcost - float(d["fcost"][0])) / float(d["fcost"][0])
    print(f"misfit {fcost:.7g} vs reference {float(d['fcost'][0]):.7g}: rel {e_f:.2e}", flush=True)
    assert e_f < 1e-5
    # synthetic data: receivers whose amplitude is above 1e-3 of the
    # largest of their field (E or H); the inline Ey and Hx are near the
    # null of the model's near-symmetry about y = 0 and mostly below it
    syn = np.loadtxt(run / f"syn_{shot:04d}.txt", skiprows=1, usecols=(4, 5))
    syn = (syn[:, 0] + 1j * syn[:, 1]).reshape(4, 3, -1)
    ref = d["syn"].reshape(4, 3, -1)
    for ch, name in enumerate(("Ex", "Ey", "Hx", "Hy")):
        amp = np.abs(ref[ch])
        field = np.abs(ref[0:2] if ch < 2 else ref[2:4]).max()
        strong = amp > 1e-3 * field
        err = np.abs(syn[ch] - ref[ch]) / np.maximum(amp, 1e-30)
        e = float(err[strong].max()) if strong.any() else 0.0
        print(f"  syn {name}: max |d| {amp.max():.3e} (field max {field:.3e}), {int(strong.sum())} strong, "
              f"max rel err {e:.2e}", flush=True)
        assert e < 1e-5
    nz, ny, nx = s.nz, s.ny, s.nx
    for k in ("Rv", "Rh"):
        g = np.fromfile(run / f"gradient_dat_{k}", np.float32)
        assert g.size == nx * ny * nz and np.all(np.isfinite(g))
        g = g.reshape(nz, ny, nx)
        gmax = float(d[f"grad_{k}_maxabs"])
        e_sub = float(np.abs(g[::4, ::4, ::4] - d[f"grad_{k}_sub4"]).max()) / gmax
        e_y0 = float(np.abs(g[:, ny // 2, :] - d[f"grad_{k}_y0"]).max()) / gmax
        nrm = float(np.linalg.norm(g.astype(np.float64)))
        e_n = abs(nrm - float(d[f"grad_{k}_norm"])) / float(d[f"grad_{k}_norm"])
        print(f"  gradient {k}: max |diff| / max |g| {e_sub:.2e} (every 4th cell), {e_y0:.2e} (y = 0 plane); "
              f"norm rel {e_n:.2e}", flush=True)
        assert e_sub < 1e-5 and e_y0 < 1e-5 and e_n < 1e-5
