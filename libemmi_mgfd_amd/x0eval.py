"""What the learned initial guess gains: V-cycles to relative residual 1e-6
on the config-4 test earths (seeds 2000-2199, configs.config4), from a zero
guess and from the network's guess.

    python -m libemmi_mgfd_amd.x0eval --weights x0net_weights_96.npz --out report.json

Per chunk of models one survey solver (FP32 cycles in FP64 refinement, one
right-hand side per model, each on its own 96^3 grid) runs, after a warm-up
solve that builds its line factors:

* ``zero``: the solve from a zero guess;
* ``x1``: one cycle from zero, then the solve restarted from that field
  (the learned path without the network);
* ``learned``: one cycle from zero, the network's refinement of that field
  (features, U-Net, reconstruction on the edges), then the solve from the
  refined field.  The solver judges the guess before its first cycle: a
  guess whose residual is not below the zero guess's ||f|| is dropped and
  the slot starts from zero (``rejected``); the cycle that made x1 is charged
  either way.
* ``learned`` with a ``*_blend`` reconstruction: the same, but before the
  solve the solver blends the guess with x1, x1 + beta (guess - x1) with the
  beta that minimises the residual (emmi_solver_blend_guess), and the solve
  continues x1's pattern sequence (the blend's residual is never above x1's,
  so it is never rejected).

Cycle counts are per right-hand side (each retires at its own convergence
test); ``saved`` = zero - learned.  Wall times are per chunk (all its
right-hand sides), the learned path including the network.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import configs, native, x0data, x0net

WEIGHTS_96 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "x0net_weights_96.npz")


SIG_AIR = 1e-6   # S/m: cells below it are air (1e-8 in these surveys)
RECONS = ("cell", "avg", "avg_air", "cell_blend", "avg_air_blend")


def edges_of(net_out, grid, x1, recon, s11=None):
    """Device guess (complex double edge vector) from the network output:
    ``cell`` takes each edge's correction from one cell, ``avg`` averages it
    over the four cells that share the edge, ``avg_air`` also leaves the air
    cells uncorrected.  (The ``*_blend`` paths of the evaluation build one of
    these and then let the solver blend it with x1.)"""
    x0 = native.DeviceBuffer(16 * grid.sc.nvec)
    n1, n2, n3 = grid.n
    L = x0net._lib()
    if recon == "cell":
        rc = L.emmi_x0_out_to_edges(net_out.ptr, x1.ptr, n1, n2, n3, x0.ptr, None)
    elif recon in ("avg", "avg_air"):
        rc = L.emmi_x0_out_to_edges_avg(net_out.ptr, x1.ptr, s11.ptr if recon == "avg_air" else None, SIG_AIR,
                                        n1, n2, n3, x0.ptr, None)
    else:
        raise ValueError(recon)
    x0net._ok(rc, "out_to_edges")
    return x0


def evaluate_chunk(net, grids, tol=1e-6, ncycle=60, recons=("avg_air",), first_cycle=0):
    """Cycles of the zero, x1 and learned paths (one per reconstruction in
    ``recons``) for one chunk of models; the solves from x1 or a guess
    continue the semi-coarsening pattern sequence at ``first_cycle``."""
    n = grids[0].n
    for g in grids:
        if g.dev is None:
            g.upload()
    ds = [(x0net.Buf.of(g.g["sigma11"]), x0net.Buf.of(g.g["sigma33"])) for g in grids]
    gl = [tuple(g.sc.x_arrays()) + (a.ptr, a.ptr, b.ptr) for g, (a, b) in zip(grids, ds)]
    S = native.Solver.survey(n, gl, list(range(len(grids))), [g.omega for g in grids], sig_on_device=True,
                             precision=1, tol=tol, ncycle=ncycle)
    L = native.lib()
    nb = 16 * grids[0].sc.nvec
    bs = []
    for g in grids:
        b = native.DeviceBuffer(nb)
        if L.emmi_inject_source_dev(g.sc.acq, g.sc.emf, 0, b.ptr, None):
            raise native.NativeError(L.emmi_path_last_error().decode())
        bs.append(b)
    hs = [native.DeviceBuffer(nb) for _ in grids]
    xo = [native.DeviceBuffer(nb) for _ in grids]
    x1 = [native.DeviceBuffer(nb) for _ in grids]

    def run(x0s, out, first=0, cycles=ncycle):
        S.set_first_cycle(first)
        S.set_ncycle(cycles)
        x0data.copy_all(hs, bs)
        st = S.solve_batch_device([h.ptr for h in hs], [x.ptr for x in out], [h.ptr for h in hs], x0s)
        L.emmi_device_sync()
        return st

    run(None, xo)                      # warm-up: line factors of every level
    t = time.perf_counter()
    st0 = run(None, xo)
    times = {"zero_s": time.perf_counter() - t}
    t = time.perf_counter()
    run(None, x1, 0, 1)
    stx = run([x.ptr for x in x1], xo, first_cycle)
    times["x1_s"] = time.perf_counter() - t
    stl, betas = {}, {}
    for rc in recons:
        t = time.perf_counter()
        run(None, x1, 0, 1)
        blend = rc.endswith("_blend")
        guesses = []
        for g, (a, b), x in zip(grids, ds, x1):
            y = net.forward(x0data.features(g, a, b, g.freq, x))
            guesses.append(edges_of(y, g, x, rc[:-len("_blend")] if blend else rc, a))
        if blend:
            # the solver takes the multiple of the network's correction of x1
            # that minimises the residual; the blend is never rejected, so the
            # solve continues x1's pattern sequence
            betas[rc] = S.blend_guess([b.ptr for b in bs], [x.ptr for x in x1], [p.ptr for p in guesses])
        stl[rc] = run([p.ptr for p in guesses], xo, 1 if blend else first_cycle)
        times[f"learned_{rc}_s"] = time.perf_counter() - t
    S.close()
    rows = []
    for k, g in enumerate(grids):
        r = {"seed": int(g.survey.extra["seed"]), "freq": g.freq, "cycles_zero": int(st0[k].cycles),
             "cycles_x1": 1 + int(stx[k].cycles)}
        status = [int(st0[k].status), int(stx[k].status)]
        for rc in recons:
            c = stl[rc][k]
            r[f"cycles_learned_{rc}"] = 1 + int(c.cycles)
            r[f"rejected_{rc}"] = int(c.x0_rejected)
            r[f"guess_rel_residual_{rc}"] = float(c.x0_rnorm / c.rnorm0)
            if rc in betas:
                r[f"beta_{rc}"] = [betas[rc][k].real, betas[rc][k].imag]
            status.append(int(c.status))
        r["status"] = status
        rows.append(r)
    return rows, times


def summarize(rows, seeds, tol, recons, times):
    z = np.array([r["cycles_zero"] for r in rows], float)
    x1 = np.array([r["cycles_x1"] for r in rows], float)
    out = {"models": len(rows), "seeds": [int(min(seeds)), int(max(seeds))], "tol": tol,
           "cycles_zero_mean": float(z.mean()), "cycles_x1_mean": float(x1.mean()),
           "all_converged": bool(all(s == 0 for r in rows for s in r["status"])), "wall_s": times}
    for rc in recons:
        ln = np.array([r[f"cycles_learned_{rc}"] for r in rows], float)
        saved = z - ln
        out[rc] = {"cycles_learned_mean": float(ln.mean()), "vcycles_saved_mean": float(saved.mean()),
                   "vcycles_saved_total": int(saved.sum()), "models_faster": int((saved > 0).sum()),
                   "models_slower": int((saved < 0).sum()),
                   "rejected": int(sum(r[f"rejected_{rc}"] for r in rows))}
    return out


def evaluate(net, seeds, chunk=16, tol=1e-6, log=print, recons=("avg_air",), first_cycle=0):
    """Report (dict) over the config-4 models ``seeds``."""
    rows, times = [], {}
    for ss, grids in x0data.config4_chunks(seeds, chunk):
        r, t = evaluate_chunk(net, grids, tol=tol, recons=recons, first_cycle=first_cycle)
        for g in grids:
            g.close()
        rows += r
        for k, v in t.items():
            times[k] = times.get(k, 0.0) + v
        msg = ", ".join(f"{rc} {np.mean([x[f'cycles_learned_{rc}'] for x in r]):.2f} "
                        f"(rejected {sum(x[f'rejected_{rc}'] for x in r)})" for rc in recons)
        log(f"eval seeds {ss[0]}-{ss[-1]}: zero {np.mean([x['cycles_zero'] for x in r]):.2f} "
            f"x1 {np.mean([x['cycles_x1'] for x in r]):.2f} learned {msg}")
    return {"summary": summarize(rows, seeds, tol, recons, times), "rows": rows,
            "recons": list(recons), "first_cycle": first_cycle}


def main(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("--weights", default=WEIGHTS_96)
    p.add_argument("--models", type=int, default=200, help="first N test seeds (2000, 2001, ...)")
    p.add_argument("--chunk", type=int, default=16)
    p.add_argument("--out", default=None)
    p.add_argument("--recons", default="avg_air", help="comma list of cell, avg, avg_air")
    p.add_argument("--first-cycle", type=int, default=0,
                   help="pattern index of the first cycle of the solves from x1 or a guess (1 continues the "
                        "sequence after x1's cycle; 0, the default, restarts it as a zero-guess solve does)")
    a = p.parse_args(argv)
    d = np.load(a.weights)
    net = x0net.UNet((96, 96, 96), levels=tuple(int(v) for v in d["levels"]))
    net.load(d)
    rep = evaluate(net, list(configs.CONFIG4_TEST)[:a.models], chunk=a.chunk,
                   log=lambda m: print(m, flush=True), recons=tuple(a.recons.split(",")),
                   first_cycle=a.first_cycle)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(rep, f, indent=1)
    print(json.dumps(rep["summary"]), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
