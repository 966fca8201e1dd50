"""Write the survey-scale golden fixtures from the compiled reference.

The reference's own sources are compiled in place by ``make -C oracle ref``
into oracle/_ref/ref_runner (never committed).  Every reference run here is
one rank of the reference's job (``shots=<k>``), solved to a relative
residual of 1e-10 (``tol=1e-10 ncycle=200``), so that its fields stand for
the converged discrete solution.

  config3.npz   four (shot, frequency) right-hand sides of the batched survey,
                one per multigrid size (96^3, 128^3, 128^3, 160x160x128):
                grid node arrays, receiver data, residual histories
  config5/      one shot (9) of the survey-scale gradient on the 160x160x128
                grid: emf_0009.txt (observed data: the reference's converged
                receiver data of the true earth, three frequencies), and
                config5.npz with the reference's mode-2 outputs on the
                starting model (addnoise=0): receiver data of syn_0009.txt,
                the misfit, both gradient volumes sampled every 4th cell per
                axis, their y = 0 planes and full-volume norms

Runs in parallel processes (about 1.5 GB each) and takes about an hour.
Run from the repository root:  python tests/golden/make_survey_golden.py [config3] [config5]
"""
from __future__ import annotations

import gzip
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from libemmi_mgfd_amd import configs  # noqa: E402

RUNNER = os.path.join(ROOT, "oracle", "_ref", "ref_runner")
TIGHT = ["tol=1e-10", "ncycle=200"]
CONFIG3_PAIRS = [(7, 0.1), (13, 0.25), (19, 0.5), (25, 1.0)]
CONFIG5_SHOT = 9
CHREC = ("Ex", "Ey", "Hx", "Hy")


def _hist(text):
    return [float(m.group(1)) for m in re.finditer(r"icycle=\d+ rnorm=(\S+)", text)]


def _start(args, cwd, log):
    fp = open(log, "w")
    return subprocess.Popen(args, cwd=cwd, stdout=fp, stderr=subprocess.STDOUT), fp


def _wait(jobs):
    for p, fp in jobs:
        if p.wait() != 0:
            raise RuntimeError(f"reference run failed: {' '.join(p.args[:3])} (log {fp.name})")
        fp.close()


def _ld(d, name, dt):
    return np.fromfile(os.path.join(d, name + ".bin"), dtype=dt)


def config3(tmp):
    s = configs.config3()
    files = configs.write_reference_inputs(s, os.path.join(tmp, "c3in"))
    jobs, outs = [], []
    for shot, f in CONFIG3_PAIRS:
        n1, n2, n3 = s.extra["grids"][f]
        out = os.path.join(tmp, f"c3_s{shot}")
        os.makedirs(out, exist_ok=True)
        args = [RUNNER, "task=model", f"outdir={out}"] + s.getpar_args(files) + [
            f"shots={shot}", f"freqs={f:g}", f"n1={n1}", f"n2={n2}", f"n3={n3}", "chrec=" + ",".join(CHREC)] + TIGHT
        jobs.append(_start(args, tmp, out + ".log"))
        outs.append(out)
    _wait(jobs)
    res = {"pairs": np.array([(p[0], p[1]) for p in CONFIG3_PAIRS])}
    for (shot, f), out in zip(CONFIG3_PAIRS, outs):
        key = f"s{shot}"
        res[key + "_n"] = np.array(s.extra["grids"][f])
        for g in ("x1", "x2", "x3"):
            res[f"{key}_{g}"] = _ld(out, g, np.float32)
        res[key + "_dcal"] = _ld(out, "dcal", np.complex64)   # [chrec][1][nrec]
        res[key + "_hist"] = np.array(_hist(open(out + ".log").read()))
    np.savez_compressed(os.path.join(HERE, "config3.npz"), **res)


def _write_emf(path, shot, freqs, dcal):
    """The reference's data-file format (src/read_write_data.c:14-35)."""
    with open(path, "w") as fp:
        fp.write("iTx \tiRx    chrec  frequency/Hz    Real{E/H}     Imag{E/H}\n")
        for ic, ch in enumerate(CHREC):
            for jf, f in enumerate(freqs):
                for ir in range(dcal.shape[2]):
                    v = dcal[ic, jf, ir]
                    fp.write("%d \t %d \t %s \t %g \t %e \t %e\n" % (shot, ir + 1, ch, f, v.real, v.imag))


def config5(tmp):
    true = configs.config5(True)
    start = configs.config5(False)
    ftrue = configs.write_reference_inputs(true, os.path.join(tmp, "c5true"))
    fstart = configs.write_reference_inputs(start, os.path.join(tmp, "c5start"))
    # observed data: the converged data of the true earth, one process per frequency
    jobs, outs = [], []
    for f in true.freqs:
        out = os.path.join(tmp, f"c5obs_{f:g}")
        os.makedirs(out, exist_ok=True)
        args = [RUNNER, "task=model", f"outdir={out}"] + true.getpar_args(ftrue) + [
            f"shots={CONFIG5_SHOT}", f"freqs={f:g}", "chrec=" + ",".join(CHREC)] + TIGHT
        jobs.append(_start(args, tmp, out + ".log"))
        outs.append(out)
    _wait(jobs)
    nrec = len(true.receivers)
    dobs = np.stack([_ld(o, "dcal", np.complex64).reshape(len(CHREC), 1, nrec) for o in outs], axis=1)[:, :, 0, :]
    d5 = os.path.join(HERE, "config5")
    os.makedirs(d5, exist_ok=True)
    name = f"emf_{CONFIG5_SHOT:04d}.txt"
    run = os.path.join(tmp, "c5grad")
    os.makedirs(run, exist_ok=True)
    _write_emf(os.path.join(run, name), CONFIG5_SHOT, true.freqs, dobs)
    with open(os.path.join(run, name), "rb") as src, gzip.open(os.path.join(d5, name + ".gz"), "wb") as dst:
        shutil.copyfileobj(src, dst)
    # one rank of the reference's mode-2 run on the starting model
    args = [RUNNER, "task=fwi", f"outdir={run}"] + start.getpar_args(fstart) + [
        f"shots={CONFIG5_SHOT}", "chrec=" + ",".join(CHREC), "addnoise=0"] + TIGHT
    _wait([_start(args, run, os.path.join(tmp, "c5grad.log"))])
    nxyz = start.nx * start.ny * start.nz
    res = {"shot": np.array(CONFIG5_SHOT), "hist": np.array(_hist(open(os.path.join(tmp, "c5grad.log")).read()))}
    syn = np.loadtxt(os.path.join(run, f"syn_{CONFIG5_SHOT:04d}.txt"), skiprows=1, usecols=(4, 5))
    res["syn"] = (syn[:, 0] + 1j * syn[:, 1]).astype(np.complex64)
    res["fcost"] = _ld(run, "fcost", np.float32)
    for k in ("Rv", "Rh"):
        g = np.fromfile(os.path.join(run, f"gradient_dat_{k}"), np.float32)
        assert g.size == nxyz
        g = g.reshape(start.nz, start.ny, start.nx)
        res[f"grad_{k}_sub4"] = np.ascontiguousarray(g[::4, ::4, ::4])
        res[f"grad_{k}_y0"] = np.ascontiguousarray(g[:, start.ny // 2, :])
        res[f"grad_{k}_norm"] = np.array(float(np.linalg.norm(g.astype(np.float64))))
        res[f"grad_{k}_maxabs"] = np.array(float(np.abs(g).max()))
    np.savez_compressed(os.path.join(d5, "config5.npz"), **res)


def main():
    if not os.path.exists(RUNNER):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
    which = [a for a in sys.argv[1:] if a in ("config3", "config5")] or ["config3", "config5"]
    with tempfile.TemporaryDirectory(dir=os.environ.get("EMMI_GOLDEN_TMP")) as tmp:
        for w in which:
            globals()[w](tmp)
            print(w, "done", flush=True)


if __name__ == "__main__":
    main()
