"""Write the golden fixtures of the command-line program's host side.

Run in a tree that has the reference (it builds oracle/_ref/cli_runner from
the reference's own src/cstd.c, src/emf_init_free.c, src/acq_init_free.c,
src/read_write_data.c, src/cal_uncertainty_noise.c, src/frannor.c,
src/regularization.c, src/triangle_smoothing.c, src/precondition.c and
src/fg_mod_reg.c, compiled
where they lie):

    python tests/golden/make_cli_fixtures.py

Outputs (all data, no reference source):
  cli_fixtures.json      what the reference's functions compute for the
                         inputs of tests/helpers.py: frannor sequences,
                         getpar lookups, emf_init/acq_init results, the
                         uncertainty, noise and misfit of a 6-channel survey,
                         and regularisation, smoothing, preconditioning and
                         fg_mod_reg on seeded random volumes
  run_inversion_0001/    the reference's own example data of shot 1 of
                         run_inversion/ (its Ex lines of emf_0001.txt, which
                         are all that run's chrec=Ex reads, and obs_0001.txt,
                         the same data with the noise of
                         src/cal_uncertainty_noise.c), gzipped
"""
from __future__ import annotations

import gzip
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

REF = "/root/reference"
RUNNER = os.path.join(ROOT, "oracle", "_ref", "cli_runner")

# getpar cases: (command-line words, parameter-file text or None, keys, types)
PAR_CASES = [
    (["a=1,2,,3", "b=0.1", "c=Ex,Ey", "d=", "e=1,", "f=,5"], None, "a,b,c,d,e,f", "i,f,a,i,i,i"),
    (["x=1", "x=2", "y=3.5e-3", "z=-7"], None, "x,y,z,w", "i,d,i,i"),
    (["freqs=0.25", ",0.75", "chrec=Ex", ",Ey,Hx"], None, "freqs,chrec", "f,a"),
    (["k=9"], "k=1 m=2.5\n# note=a b\nq='a b c' r=\"x y\"\ns=3,4\tt=5\n", "k,m,q,r,s,t,note", "i,f,s,s,i,i,s"),
    (["m=7"], "m=1\nm=2 n=3.25,4.5", "m,n", "i,f"),
    (["v=1e3", "u=12abc", "w=0x10"], None, "v,u,w", "f,i,d"),
]


def run(args, cwd=None) -> str:
    r = subprocess.run([RUNNER] + args, capture_output=True, text=True, cwd=cwd)
    if r.returncode != 0:
        raise RuntimeError(f"cli_runner {args} failed: {r.stderr}")
    return r.stdout


def after_dump(text: str) -> list:
    lines = text.splitlines()
    return lines[lines.index("---- dump") + 1:]


MODEL_SHAPE = (9, 7, 5)   # n1 (x), n2 (y), n3 (z)


def model_side_cases(d: str) -> list:
    """Regularisation, smoothing, preconditioning and fg_mod_reg of the
    reference on seeded random volumes (inputs kept in the fixture)."""
    import numpy as np
    rng = np.random.default_rng(5)
    n1, n2, n3 = MODEL_SHAPE
    nxyz = n1 * n2 * n3

    def put(name, a):
        a = np.ascontiguousarray(a, np.float32)
        a.tofile(os.path.join(d, name + ".bin"))
        return [float(v).hex() for v in a.reshape(-1)]

    x1 = rng.standard_normal(nxyz).astype(np.float32)
    x2 = (rng.standard_normal(2 * nxyz) * 0.7 + 1.0).astype(np.float32)
    xref = (x2 + rng.standard_normal(2 * nxyz).astype(np.float32) * np.float32(0.2)).astype(np.float32)
    g2 = rng.standard_normal(2 * nxyz).astype(np.float32) * np.float32(1e-3)
    phess = np.abs(rng.standard_normal(2 * nxyz)).astype(np.float32) * np.float32(1e-2)
    phess[::7] = 0
    bathy = rng.uniform(20.0, 180.0, n1 * n2).astype(np.float32)
    inputs = {"x1": put("x1", x1), "x2": put("x2", x2), "xref": put("xref", xref), "g2": put("g2", g2),
              "phess": put("phess", phess), "bathy": put("bathy", bathy)}
    dims = [f"n1={n1}", f"n2={n2}", f"n3={n3}"]
    files = {k: os.path.join(d, k + ".bin") for k in inputs}
    cases = []
    spec = [
        ("reg", ["npar=1", "fx=" + files["x1"]]),
        ("smooth", ["npar=1", "fx=" + files["x1"], "r1=3", "r2=2", "r3=1", "repeat=2"]),
        ("smooth", ["npar=1", "fx=" + files["x1"], "r1=1", "r2=4", "r3=2", "repeat=1"]),
    ]
    common = ["npar=2", "fx=" + files["x2"], "fxref=" + files["xref"], "fg=" + files["g2"],
              "fphess=" + files["phess"], "fbathy=" + files["bathy"], "dz=50", "oz=0", "freq=0.5"]
    for preco in (1, 2, 3):
        spec.append(("precond", common + [f"preco={preco}", "r1=2", "r2=3", "r3=2", "repeat=1"]))
    spec.append(("modreg", common + ["iter=2", "gamma1=10", "gamma2=0.5"]))
    spec.append(("modreg", common + ["iter=0", "gamma1=0", "gamma2=2"]))
    for task, args in spec:
        cases.append({"task": task, "args": [a for a in args if not a.startswith("f")] ,
                      "lines": run([f"task={task}"] + dims + args).splitlines()})
    return {"shape": list(MODEL_SHAPE), "inputs": inputs, "cases": cases}


def main() -> None:
    from helpers import cli_args, cli_write_inputs
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "cli"])
    out = {"rannor": {}, "pars": [], "init": {}, "noise": {}}
    for key, args in (("fresh", ["n=6000"]), ("seed7", ["n=1500", "seed=7"]), ("seed-12345", ["n=1500", "seed=-12345"])):
        out["rannor"][key] = run(["task=rannor"] + args).split()
    with tempfile.TemporaryDirectory() as d:
        for words, partext, keys, types in PAR_CASES:
            args = list(words)
            if partext is not None:
                p = os.path.join(d, "p.par")
                with open(p, "w") as fp:
                    fp.write(partext)
                args = ["par=" + p] + args
            out["pars"].append({"words": words, "parfile": partext, "keys": keys, "types": types,
                                "lines": run(["task=pars", f"keys={keys}", f"types={types}"] + args).splitlines()})
        files = cli_write_inputs(d)
        rel = {k: os.path.basename(v) for k, v in files.items()}
        base = cli_args(rel)
        cases = {
            "rank0": ["rank=0", "nranks=1"],
            "rank1of3": ["rank=1", "nranks=3"],
            "shots": ["rank=1", "nranks=2", "shots=3,1"],
            "reciprocity": ["rank=0", "nranks=1", "reciprocity=1", "shots=1"],
        }
        for name, extra in cases.items():
            out["init"][name] = after_dump(run(["task=init"] + base + extra, cwd=d))
        noise_cases = {
            "shot1": ["rank=0", "nranks=1", "delta_phi=0.05", "amp_perc=0.05"],
            "shot2": ["rank=1", "nranks=2", "delta_phi=0.02"],
            "shot3_nonoise": ["rank=0", "nranks=1", "shots=3", "addnoise=0", "noisefloorE=1e-13"],
        }
        for name, extra in noise_cases.items():
            lines = after_dump(run(["task=noise"] + base + extra, cwd=d))
            shot = {"shot1": 1, "shot2": 2, "shot3_nonoise": 3}[name]
            with open(os.path.join(d, f"obs_{shot:04d}.txt")) as fp:
                obs = fp.read()
            with open(os.path.join(d, f"rmse_{shot:04d}.txt")) as fp:
                rmse = fp.read()
            out["noise"][name] = {"extra": extra, "lines": lines, "obs": obs, "rmse": rmse}
        out["args"] = base
        out["model_side"] = model_side_cases(d)
    with open(os.path.join(HERE, "cli_fixtures.json"), "w") as fp:
        json.dump(out, fp, indent=0)

    dst = os.path.join(HERE, "run_inversion_0001")
    os.makedirs(dst, exist_ok=True)
    with open(os.path.join(REF, "run_inversion", "emf_0001.txt")) as fp:
        lines = fp.read().splitlines(True)
    ex = [lines[0]] + [ln for ln in lines[1:] if ln.split()[2] == "Ex"]
    with gzip.open(os.path.join(dst, "emf_0001_Ex.txt.gz"), "wt") as fp:
        fp.write("".join(ex))
    with open(os.path.join(REF, "run_inversion", "obs_0001.txt")) as fp, \
            gzip.open(os.path.join(dst, "obs_0001.txt.gz"), "wt") as gp:
        gp.write(fp.read())
    print("wrote", os.path.join(HERE, "cli_fixtures.json"), dst)


if __name__ == "__main__":
    main()
