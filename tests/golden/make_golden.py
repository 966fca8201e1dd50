"""Write the golden fixtures of tests/golden/ from the compiled reference.

The reference's own sources are compiled in place by ``make -C oracle ref``
into oracle/_ref/ref_runner (never committed); this script runs it on the
synthetic surveys of libemmi_mgfd_amd.configs and keeps only small outputs:

  tiny_units.npz   kernel known answers on a 8x6x14 grid (inputs and outputs of
                   residual, line/point Gauss-Seidel, transfer, curl)
  tiny_model.npz   the 12x10x16 two-frequency survey: grid, indices, source
                   term, E/H of the first frequency, receiver data, the E and
                   H volumes on the input grid (icheck=1: extract_electric_field
                   and extract_magnetic_field, both frequencies) and residual
                   histories at the default tolerance and at tol=1e-10
  tiny_cycles.npz  the same survey solved with the V-cycle (cycleopt=1), the
                   W-cycle (cycleopt=3) and the F-cycle without
                   semi-coarsening (isemicoarsen=0: point Gauss-Seidel on
                   every level): E/H of the first frequency and residual
                   histories
  tiny_emf_0001.txt the reference's receiver-data text file for that survey
  tiny_grad.npz    one mode-2 gradient (forward + adjoint) on the same survey
  config1.npz      grid arrays, medium hashes, receiver data and residual
                   histories of the 64^3 survey (default and tol=1e-10), and
                   the y = 0 planes of its E and H volumes (icheck=1)
  config2.npz      (with --config2) the 128^3 three-frequency survey: grid,
                   medium hashes, receiver data and residual histories at the
                   default tolerance and at tol=1e-10

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from libemmi_mgfd_amd import configs  # noqa: E402

RUNNER = os.path.join(ROOT, "oracle", "_ref", "ref_runner")
GRID_KEYS = ["xx", "yy", "zz", "x1", "x2", "x3", "x1s", "x2s", "x3s", "d1", "d2", "d3", "d1s", "d2s", "d3s"]


def run(s, task, workdir, extra=()):
    files = configs.write_reference_inputs(s, os.path.join(workdir, "in"))
    out = os.path.join(workdir, "out")
    args = [RUNNER, f"task={task}", f"outdir={out}"] + s.getpar_args(files) + list(extra)
    r = subprocess.run(args, capture_output=True, text=True, check=True)
    hist = [[]]
    for line in r.stdout.splitlines():
        if line.startswith("solve_begin"):
            hist.append([])
        m = re.match(r"icycle=(\d+) rnorm=(\S+)", line)
        if m:
            hist[-1].append(float(m.group(2)))
    return out, [h for h in hist if h], files


def ld(d, name, dt=np.complex128):
    return np.fromfile(os.path.join(d, name + ".bin"), dtype=dt)


def grid_dict(d, s):
    g = {k: ld(d, k, np.float32) for k in GRID_KEYS}
    for k in ("sigma11", "sigma22", "sigma33"):
        g[k] = ld(d, k, np.float32)
    return g


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def pad_hist(hs):
    m = max(len(h) for h in hs)
    out = np.full((len(hs), m), np.nan)
    for i, h in enumerate(hs):
        out[i, :len(h)] = h
    return out


def tiny_units(tmp):
    s = configs.tiny(n1=8, n2=6, n3=14)
    d, _, _ = run(s, "units", os.path.join(tmp, "units"), ["seed=7"])
    out = {"n": np.array([s.n1, s.n2, s.n3]), "freq": np.array(s.freqs[0])}
    out.update(grid_dict(d, s))
    names = ["u", "f", "r", "H", "f2"]
    names += [f"lgs_{a}_it{i}" for a in "xyz" for i in (0, 1)] + [f"pgs_it{i}" for i in (0, 1)]
    names += [f"{p}_c{c}" for c in range(3) for p in ("rf", "fc", "uc", "uf")]
    for nm in names:
        out[nm] = ld(d, nm)
    for c in range(3):
        out[f"lev1_dims_c{c}"] = ld(d, f"lev1_dims_c{c}", np.int32)
        out[f"lev1_sigma11_c{c}"] = ld(d, f"lev1_sigma11_c{c}", np.float64)
        out[f"lev1_sigma33_c{c}"] = ld(d, f"lev1_sigma33_c{c}", np.float64)
    for q in ("11", "22", "33"):
        out[f"lev0_sigma{q}"] = ld(d, f"lev0_sigma{q}", np.float64)
    np.savez_compressed(os.path.join(HERE, "tiny_units.npz"), **out)


def tiny_model(tmp):
    s = configs.tiny()
    d, hist, _ = run(s, "model", os.path.join(tmp, "model"), ["dumpfields=1", "icheck=1"])
    out = {"n": np.array([s.n1, s.n2, s.n3])}
    for k in ("E1", "E2", "E3", "H1", "H2", "H3"):
        out[k] = ld(d, k, np.complex64)   # [nfreq][nz][ny][nx]
    out.update(grid_dict(d, s))
    out["rec_idx"] = ld(d, "rec_idx", np.int32)
    out["rec_w"] = ld(d, "rec_w", np.float32)
    out["b_f0"] = ld(d, "b_f0")
    out["b_f1"] = ld(d, "b_f1")
    out["E_f0"] = ld(d, "E_f0")
    out["H_f0"] = ld(d, "H_f0")
    out["dcal"] = ld(d, "dcal", np.complex64)
    out["hist"] = pad_hist(hist)
    with open(os.path.join(d, "emf_0001.txt")) as fp:
        txt = fp.read()
    with open(os.path.join(HERE, "tiny_emf_0001.txt"), "w") as fp:
        fp.write(txt)
    d2, hist2, _ = run(s, "model", os.path.join(tmp, "model_tight"), ["tol=1e-10", "ncycle=200"])
    out["dcal_tight"] = ld(d2, "dcal", np.complex64)
    out["hist_tight"] = pad_hist(hist2)
    np.savez_compressed(os.path.join(HERE, "tiny_model.npz"), **out)
    return out


CYCLE_VARIANTS = {"v": ["cycleopt=1"], "w": ["cycleopt=3"], "f_point": ["cycleopt=2", "isemicoarsen=0"]}


def tiny_cycles(tmp):
    """The other cycle types and the point smoother (src/gmg.c:1586-1668,
    116-336, 1301-1313) on the tiny survey, default tolerance."""
    s = configs.tiny()
    out = {}
    for name, extra in CYCLE_VARIANTS.items():
        d, hist, _ = run(s, "model", os.path.join(tmp, "cyc_" + name), ["dumpfields=1"] + extra)
        out[f"{name}_E_f0"] = ld(d, "E_f0")
        out[f"{name}_H_f0"] = ld(d, "H_f0")
        out[f"{name}_hist"] = pad_hist(hist)
        out[f"{name}_dcal"] = ld(d, "dcal", np.complex64)
    np.savez_compressed(os.path.join(HERE, "tiny_cycles.npz"), **out)


def tiny_grad(tmp, model):
    """Observed data: the tight receiver data of the true tiny model scaled by
    a smooth factor; the starting model is a 1 ohm-m half-space."""
    s = configs.tiny()
    nch, nf = len(s.chrec), len(s.freqs)
    dobs = model["dcal_tight"].reshape(nch, nf, -1) * np.complex64(1.0 + 0.05j)
    delta = (np.abs(dobs) * 0.03 + 1e-15).astype(np.float32)
    s0 = configs.tiny()
    sea = s0.rho11 <= configs.RHO_WATER + 1e-6
    s0.rho11 = np.where(sea, s0.rho11, np.float32(1.0)).astype(np.float32)
    s0.rho22 = s0.rho11.copy()
    s0.rho33 = s0.rho11.copy()
    w = os.path.join(tmp, "grad")
    os.makedirs(w, exist_ok=True)
    pd, pl = os.path.join(w, "dobs.bin"), os.path.join(w, "delta.bin")
    dobs.astype(np.complex64).tofile(pd)
    delta.tofile(pl)
    d, _, _ = run(s0, "grad", w, [f"dobs={pd}", f"delta={pl}", "dsmute=300", "drmute=150"])
    out = {"dobs": dobs.reshape(-1), "delta": delta.reshape(-1), "rho": s0.rho11}
    for k in ("dcal", "dres"):
        out[k] = ld(d, k, np.complex64)
    out["fcost"] = ld(d, "fcost", np.float32)
    out["badj_f0"] = ld(d, "badj_f0")
    for k in ("E1f", "E2f", "E3f", "E1a", "E2a", "E3a"):
        out[k] = ld(d, k, np.complex64)
    out["grad"] = ld(d, "grad", np.float32)
    out["phess"] = ld(d, "phess", np.float32)
    np.savez_compressed(os.path.join(HERE, "tiny_grad.npz"), **out)


def config1(tmp):
    s = configs.config1()
    d, hist, _ = run(s, "model", os.path.join(tmp, "c1"), ["icheck=1"])
    out = {"n": np.array([s.n1, s.n2, s.n3])}
    for k in ("E1", "E2", "E3", "H1", "H2", "H3"):
        # the y = 0 plane (iy = ny // 2) of the one frequency: [nz][nx]
        out[k + "_y0"] = ld(d, k, np.complex64).reshape(s.nz, s.ny, s.nx)[:, s.ny // 2, :].copy()
    g = grid_dict(d, s)
    for k in GRID_KEYS:
        out[k] = g[k]
    for k in ("sigma11", "sigma22", "sigma33"):
        out[k + "_sha256"] = np.array(sha(g[k]))
        out[k + "_slice"] = g[k].reshape(s.n3, s.n2, s.n1)[:, s.n2 // 2, :]
    out["rec_idx"] = ld(d, "rec_idx", np.int32)
    out["rec_w"] = ld(d, "rec_w", np.float32)
    out["dcal"] = ld(d, "dcal", np.complex64)
    out["hist"] = pad_hist(hist)
    d2, hist2, _ = run(s, "model", os.path.join(tmp, "c1_tight"), ["tol=1e-10", "ncycle=200"])
    out["dcal_tight"] = ld(d2, "dcal", np.complex64)
    out["hist_tight"] = pad_hist(hist2)
    np.savez_compressed(os.path.join(HERE, "config1.npz"), **out)


def config2(tmp):
    """Headline survey: receiver data and residual histories (default tol)."""
    s = configs.config2()
    d, hist, _ = run(s, "model", os.path.join(tmp, "c2"))
    out = {"n": np.array([s.n1, s.n2, s.n3])}
    g = grid_dict(d, s)
    for k in GRID_KEYS:
        out[k] = g[k]
    for k in ("sigma11", "sigma22", "sigma33"):
        out[k + "_sha256"] = np.array(sha(g[k]))
    out["dcal"] = ld(d, "dcal", np.complex64)
    out["hist"] = pad_hist(hist)
    # near-exact discrete solution: the same solve to tol=1e-10
    d2, hist2, _ = run(s, "model", os.path.join(tmp, "c2_tight"), ["tol=1e-10", "ncycle=200"])
    out["dcal_tight"] = ld(d2, "dcal", np.complex64)
    out["hist_tight"] = pad_hist(hist2)
    np.savez_compressed(os.path.join(HERE, "config2.npz"), **out)


def main():
    if not os.path.exists(RUNNER):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
    with tempfile.TemporaryDirectory() as tmp:
        if "--only-config2" in sys.argv:
            config2(tmp)
            return
        tiny_units(tmp)
        model = tiny_model(tmp)
        tiny_cycles(tmp)
        tiny_grad(tmp, model)
        if "--skip-config1" not in sys.argv:
            config1(tmp)
        if "--config2" in sys.argv:
            config2(tmp)
    for f in sorted(os.listdir(HERE)):
        print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
