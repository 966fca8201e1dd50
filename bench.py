#!/usr/bin/env python
"""Throughput of the multigrid Maxwell solve on one MI355X.

Metric: converged solves per second at 128^3 to relative residual 1e-6,
plain FP64 versus FP32 cycles in FP64 refinement, batched, from a learned
initial guess.

Workload (one step): the 128^3 survey (201x201x101 input model, 50/50/20 m
cells, anisotropic sediment, sinusoidal seafloor, one resistive block;
multigrid 128^3; one x-dipole shot; 201 inline receivers; 0.25, 0.75 and
1.25 Hz).  A step runs, for each frequency, the source injection into the
device field vector, the solve to ||f - Au|| < 1e-6 ||f|| and the receiver
extraction.  Every variant has the same per-step work: the model, the
regridded medium and the solver hierarchies are built before the timed
region (``setup_s``), and every variant keeps one right-hand-side slot per
frequency, so the line factors and captured cycle graphs of the warmup steps
serve the timed steps of every variant alike.  Data are synthetic.

Before its timed steps each variant runs the autotuner on its own launch
shapes (``Solver.autotune``, part of ``setup_s``): every line-sweep shape
gets the substitution kernel measured fastest (``line_autotune`` lists the
table; ``--no-autotune`` keeps the built-in rules).

Variants: ``plain`` (complex FP64 cycles, one solver per frequency, solved
one after another), ``mixed`` (FP32-storage cycles inside FP64 refinement,
one solver per frequency), ``mixed_batched`` (the three frequencies as one
batched solve) and ``mixed_batched_learned`` (the same from the network's
initial guess, blended by the solver with the one-cycle field it refines,
charged the cycle and the network that make it).  ``value``
is the fastest variant whose solves all converged (status 0) with finite
receiver data.

Further lines of the JSON object: per-kernel rates of the plain FP64
level-0 kernels (``plain_kernels``); the batched survey (25 shots x 4
frequencies, each frequency on its own multigrid size, ``survey``); one
mode-2 gradient at survey scale (25 shots x 3 frequencies on 160x160x128,
``gradient``); the rates of the initial-guess network's BF16 MFMA
convolutions against the dense BF16 peak (``x0net_conv``); the reference's
CPU path on a bounded sample (``cpu_baseline``).

Prints one JSON line.  ``--gpus`` above 1 is refused: this build runs on one
GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/s at 128^3 to rel-res 1e-6: plain FP64 vs mixed+batched+learned x0"
HBM_PEAK_GBS = 8000.0
# HBM bytes moved per line-sweep launch over its compulsory bytes, measured
# with two rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE) on the
# mixed+batched 128^3 step, each scaled by the same counter on the run's
# 103 MB device-to-device copies (FETCH_SIZE x 2.00, WRITE_SIZE x 1.00):
# 87.30 MB per launch against 8.97 MB compulsory (the stored FP64 factor
# records dominate)
LINE_SWEEP_TRAFFIC_RATIO = 87.30 / 8.97


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--config", default="config2")
    p.add_argument("--variants", default="plain,mixed,mixed_batched,mixed_batched_learned")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--no-survey", action="store_true", help="skip the batched survey leg")
    p.add_argument("--no-gradient", action="store_true", help="skip the survey-scale gradient leg")
    p.add_argument("--no-kernels", action="store_true", help="skip the per-kernel rates")
    p.add_argument("--cpu-cycles", type=int, default=2, help="F-cycles of the bounded CPU sample")
    p.add_argument("--no-autotune", action="store_true",
                   help="keep the built-in rules for the line-sweep kernels instead of timing them per launch shape")
    p.add_argument("--no-graphs", action="store_true",
                   help="launch every kernel eagerly instead of replaying captured cycle graphs")
    p.add_argument("--verbose", action="store_true")
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class ShotRun:
    """One shot on the GPU: device-resident medium, solvers and vectors.
    batched: one solver with a slot per frequency, all frequencies in one
    solve; otherwise one single-slot solver per frequency, solved one after
    another."""

    def __init__(self, survey, precision, batched=False, shot=0, graphs=True, learned=False, autotune=True):
        from libemmi_mgfd_amd import carriers, native
        self.native = native
        self.batched = batched
        t0 = time.perf_counter()
        self.sc = carriers.ShotCarriers(survey, shot)
        L = native.lib()
        L.regrid_init(self.sc.acq, self.sc.emf)
        e = self.sc.emf
        self.n = (e.n1, e.n2, e.n3)
        ncell = e.n1 * e.n2 * e.n3
        g = self.sc.grid()
        self.dsig = []
        for k in ("sigma11", "sigma22", "sigma33"):
            d = native.DeviceBuffer(4 * ncell).upload(np.ascontiguousarray(g[k], np.float32))
            self.dsig.append(d)
        x1, x2, x3 = self.sc.x_arrays()
        nf = e.nfreq
        kw = dict(sig_on_device=True, precision=precision, tol=1e-6, ncycle=60, graphs=1 if graphs else 0)
        sig = [d.ptr for d in self.dsig]
        if batched:
            self.solvers = [native.Solver(self.n, x1, x2, x3, *sig, list(self.sc.omegas[:nf]), **kw)]
        else:
            self.solvers = [native.Solver(self.n, x1, x2, x3, *sig, self.sc.omegas[f], **kw) for f in range(nf)]
        if autotune:
            # time the line-sweep kernels on this variant's launch shapes
            # (the table is process-wide: shapes already tuned are skipped)
            for S in self.solvers:
                S.autotune()
        nb = 16 * self.sc.nvec
        self.b = [native.DeviceBuffer(nb) for _ in range(nf)]
        self.x = [native.DeviceBuffer(nb) for _ in range(nf)]
        self.guesser = None
        if learned:
            from libemmi_mgfd_amd import x0guess
            self.guesser = x0guess.Guesser(self.n, g, survey.sources[shot][:3], self.sc.emf.rho_water)
            self.x0 = [native.DeviceBuffer(nb) for _ in range(nf)]
        native.lib().emmi_device_sync()
        self.setup_s = time.perf_counter() - t0

    def _inject(self, f):
        L = self.native.lib()
        if L.emmi_inject_source_dev(self.sc.acq, self.sc.emf, f, self.b[f].ptr, None):
            raise RuntimeError(L.emmi_path_last_error().decode())

    def _extract(self, f):
        L = self.native.lib()
        if L.emmi_extract_emf_data_dev(self.sc.acq, self.sc.emf, f, self.x[f].ptr, self.b[f].ptr, None):
            raise RuntimeError(L.emmi_path_last_error().decode())

    def solve_all(self):
        """Inject, solve and extract every frequency; returns (status,
        cycles, rel-res, x0 rejected) per frequency."""
        nf = self.sc.emf.nfreq
        stats = []
        if self.batched:
            S = self.solvers[0]
            for f in range(nf):
                self._inject(f)
            x0 = None
            extra = 0
            if self.guesser is not None:
                # one cycle from zero, the network's refinement of it, then
                # the solve from the refined fields (charged the extra cycle)
                S.set_ncycle(1)
                S.solve_batch_device([b.ptr for b in self.b], [x.ptr for x in self.x], [b.ptr for b in self.b])
                for f in range(nf):
                    self.guesser.guess(self.dsig[0].ptr, self.dsig[2].ptr, float(self.sc.freqs[f]),
                                       self.x[f].ptr, self.x0[f].ptr)
                S.set_ncycle(60)
                for f in range(nf):
                    self._inject(f)
                # the solver blends each guess with x1 (the multiple of the
                # network's correction that minimises the residual), so the
                # solve from the blend continues x1's pattern sequence
                S.blend_guess([b.ptr for b in self.b], [x.ptr for x in self.x], [p.ptr for p in self.x0])
                S.set_first_cycle(1)
                x0 = [p.ptr for p in self.x0]
                extra = 1
            sts = S.solve_batch_device([b.ptr for b in self.b], [x.ptr for x in self.x], [b.ptr for b in self.b], x0)
            S.set_first_cycle(0)
            for f in range(nf):
                self._extract(f)
                st = sts[f]
                stats.append((st.status, st.cycles + extra, st.rnorm / st.rnorm0, st.x0_rejected))
            return stats
        for f in range(nf):
            self._inject(f)
            st = self.solvers[f].solve_device(self.b[f].ptr, self.x[f].ptr, self.b[f].ptr)
            self._extract(f)
            stats.append((st.status, st.cycles, st.rnorm / st.rnorm0, 0))
        return stats

    def close(self):
        for S in self.solvers:
            S.close()


VARIANTS = {"plain": (0, False, False), "mixed": (1, False, False), "batched": (0, True, False),
            "mixed_batched": (1, True, False), "mixed_batched_learned": (1, True, True)}


def time_variant(survey, variant, steps, warmup, graphs=True, autotune=True):
    """Warmup, then `steps` timed steps (cycle bodies replayed as HIP graphs);
    then one more step runs eagerly with per-launch HIP events for the
    kernel-level figures (the same kernels with the same arguments)."""
    from libemmi_mgfd_amd import native
    precision, batched, learned = VARIANTS[variant]
    run = ShotRun(survey, precision, batched, graphs=graphs, learned=learned, autotune=autotune)
    for _ in range(warmup):
        run.solve_all()
    native.lib().emmi_device_sync()
    t0 = time.perf_counter()
    stats = []
    for _ in range(steps):
        stats += run.solve_all()
    native.lib().emmi_device_sync()
    dt = time.perf_counter() - t0
    native.timing_enable(True)
    run.solve_all()
    native.lib().emmi_device_sync()
    kt = native.timing_read()
    native.timing_enable(False)
    nsolve = steps * len(survey.freqs)
    out = {"solves_per_s": nsolve / dt, "ms_per_step": 1e3 * dt / steps, "setup_s": run.setup_s,
           "cycles": [c for _, c, _, _ in stats[:len(survey.freqs)]],
           "status": sorted(set(s for s, _, _, _ in stats)), "rel_res": max(r for _, _, r, _ in stats),
           "x0_rejected": [x for _, _, _, x in stats[:len(survey.freqs)]],
           "kernels": kt, "dcal": run.sc.dcal.copy()}
    run.close()
    return out


def kernel_rates(survey, precision):
    """Level-0 kernels of the 128^3 grid one by one (emmi_solver_kernel_rates)."""
    from libemmi_mgfd_amd import native
    run = ShotRun(survey, precision, batched=False)
    r = run.solvers[0].kernel_rates(reps=20)
    run.close()
    out = {}
    for k, (ms, by) in r.items():
        gbs = by / (ms * 1e-3) / 1e9
        out[k] = {"us": round(1e3 * ms, 2), "compulsory_MB": round(by / 1e6, 2), "GB_s": round(gbs, 1),
                  "frac": round(gbs / HBM_PEAK_GBS, 3)}
    return out


def survey_leg(autotune=True):
    """Batched survey: 25 shots x 4 frequencies, each frequency on its own
    multigrid size, FP32 cycles in FP64 refinement, rel-res 1e-6."""
    from libemmi_mgfd_amd import configs, survey
    s = configs.config3()
    run = survey.SurveyRun(s, grids=configs.CONFIG3_GRIDS, precision=1, tol=1e-6, ncycle=60, max_slots=25,
                           autotune=autotune)
    out = run.run(repeats=1)
    st = sorted({r["status"] for r in run.results.values()})
    cy = [r["cycles"] for r in run.results.values()]
    return {"workload": "config3: 161x161x71 input model, 25 shots x 4 freqs (0.1/0.25/0.5/1 Hz on "
                        "96^3/128^3/128^3/160x160x128), 405 receivers per shot, rel-res 1e-6",
            "rhs": out["rhs"], "solves_per_s": out["rhs"] / out["timed_s"], "seconds": out["timed_s"],
            "first_pass_s": out["first_s"], "setup_s": out["setup_s"], "status": st,
            "cycles_mean": float(np.mean(cy)), "cycles_max": int(max(cy)),
            "chunks": [{k: c[k] for k in ("n", "rhs", "timed_s")} for c in out["chunks"]]}


def gradient_leg():
    """One mode-2 evaluation (misfit and gradient) of the 25-shot,
    3-frequency survey on 160x160x128 through the command-line program:
    observed data from a mode-0 run on the true earth, then the timed
    mode-2 run on the starting model (wall clock of the whole program:
    parameter and data files, regridding, 75 forward and 75 adjoint solves,
    gradient, output files)."""
    from libemmi_mgfd_amd import configs
    exe = os.path.join(ROOT, "libemmi_mgfd_amd", "emmi_main")
    true, start = configs.config5(True), configs.config5(False)
    with tempfile.TemporaryDirectory() as tmp:
        ft = configs.write_reference_inputs(true, os.path.join(tmp, "true"))
        fs = configs.write_reference_inputs(start, os.path.join(tmp, "start"))
        obs = os.path.join(tmp, "obs")
        os.makedirs(obs)
        common = ["nproc=25", "chrec=Ex,Ey,Hx,Hy"]
        t0 = time.perf_counter()
        r = subprocess.run([exe, "mode=0"] + true.getpar_args(ft) + common, cwd=obs, capture_output=True,
                           text=True, timeout=900)
        if r.returncode != 0:
            raise RuntimeError("mode 0 of the gradient leg failed: " + r.stdout[-2000:] + r.stderr[-2000:])
        t_obs = time.perf_counter() - t0
        run = os.path.join(tmp, "grad")
        os.makedirs(run)
        for k in range(1, 26):
            shutil.copy(os.path.join(obs, f"emf_{k:04d}.txt"), run)
        t0 = time.perf_counter()
        r = subprocess.run([exe, "mode=2"] + start.getpar_args(fs) + common + ["addnoise=0"], cwd=run,
                           capture_output=True, text=True, timeout=900)
        t_grad = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError("mode 2 of the gradient leg failed: " + r.stdout[-2000:] + r.stderr[-2000:])
        g = np.fromfile(os.path.join(run, "gradient_dat_Rv"), np.float32)
        nc = len(re.findall(r"not converged", r.stdout))
    return {"workload": "config5: 201x201x126 input model, 25 shots x 3 freqs (0.25/0.75/1.25 Hz), MG "
                        "160x160x128, 201 receivers, forward + adjoint to rel-res 1e-6, FP32 cycles in FP64 "
                        "refinement, 8 shots per survey solve",
            "seconds_per_gradient": t_grad, "solves": 150, "solves_per_s": 150 / t_grad,
            "observed_data_mode0_s": t_obs, "gradient_finite": bool(np.all(np.isfinite(g))),
            "not_converged_solves": nc}


def cpu_baseline(survey, cycles):
    """The reference (compiled from its own sources into oracle/_ref) on one
    host core: a bounded sample of `cycles` F-cycles of the first frequency;
    a cycle costs the same at every frequency (same grid), so the time of
    the step's three solves is the cycle time times the reference's own
    cycle counts of the three frequencies."""
    runner = os.path.join(ROOT, "oracle", "_ref", "ref_runner")
    if not os.path.exists(runner):
        return None
    from libemmi_mgfd_amd import configs
    s1 = configs.CONFIGS[survey.name]()
    s1.freqs = s1.freqs[:1]
    with tempfile.TemporaryDirectory() as tmp:
        files = configs.write_reference_inputs(s1, os.path.join(tmp, "in"))
        args = [runner, "task=model", f"outdir={tmp}/out", f"ncycle={cycles}", "verb=0"] + s1.getpar_args(files)
        r = subprocess.run(args, capture_output=True, text=True, timeout=600)
    m = re.search(r"solve_end ifreq=0 seconds=(\S+)", r.stdout)
    if not m:
        return None
    t = float(m.group(1))
    gold = os.path.join(ROOT, "tests", "golden", f"{survey.name}.npz")
    if not os.path.exists(gold):
        return None
    # residual norms printed per cycle: the last one passes the test, so the
    # cycles run are one fewer than the norms
    ncyc = [int(np.sum(np.isfinite(h))) - 1 for h in np.load(gold)["hist"]]
    per_cycle = t / cycles
    return {"value": len(ncyc) / (per_cycle * sum(ncyc)), "unit": "solves/s", "cores": 1, "kind": "reference",
            "sample": f"{cycles} F-cycles of the {s1.freqs[0]} Hz solve at {s1.n1}x{s1.n2}x{s1.n3} "
                      f"({t:.1f} s, single-threaded); the step's {len(ncyc)} solves take the reference's own "
                      f"{ncyc} cycles to rel-res 1e-6"}


def receiver_errors(survey, dcal, dref):
    out = {}
    for ch, name in enumerate(survey.chrec):
        if name not in ("Ex", "Hy"):
            continue
        amp = np.abs(dref[ch])
        m = amp > 1e-15
        e = np.abs(dcal[ch] - dref[ch])[m] / amp[m]
        out[name] = {"max": float(e.max()), "median": float(np.median(e))}
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    from libemmi_mgfd_amd import configs, native
    native.lib()
    survey = configs.CONFIGS[a.config]()
    gold = os.path.join(ROOT, "tests", "golden", f"{survey.name}.npz")
    gd = np.load(gold) if os.path.exists(gold) else None
    res, errors = {}, {}
    for v in a.variants.split(","):
        if VARIANTS[v][2]:
            from libemmi_mgfd_amd import x0guess
            if not os.path.exists(x0guess.WEIGHTS):
                log(f"bench.py: {x0guess.WEIGHTS} missing, variant {v} skipped")
                continue
        try:
            res[v] = time_variant(survey, v, a.steps, a.warmup, graphs=not a.no_graphs, autotune=not a.no_autotune)
        except Exception as e:   # a failed variant is reported, the others still run
            errors[v] = f"{type(e).__name__}: {e}"
            log(f"bench.py: variant {v} failed: {errors[v]}")
            continue
        log(v, {k: res[v][k] for k in ("solves_per_s", "ms_per_step", "cycles", "status", "rel_res")})
    # receiver data of the signal channels (inline Ex, Hy; inline Ey and Hx
    # vanish by symmetry) against the reference's converged FP64 solve
    # (tol 1e-10); the reference's own solve at the bench tolerance is given
    # for scale
    rec = {}
    if res and gd is not None and "dcal_tight" in gd.files:
        shp = next(iter(res.values()))["dcal"].shape
        dref = gd["dcal_tight"].reshape(shp)
        rec = {k: receiver_errors(survey, v["dcal"], dref) for k, v in res.items()}
    # the headline: the fastest variant whose solves all converged with
    # finite receiver data
    ok = [k for k, v in res.items() if v["status"] == [0] and np.all(np.isfinite(v["dcal"]))
          and all(np.isfinite(e["max"]) for e in rec.get(k, {}).values())]
    if not ok:
        raise SystemExit("bench.py: no variant converged with finite receiver data")
    best = max(ok, key=lambda k: res[k]["solves_per_s"])
    r = res[best]
    # dominant kernel class of the reported variant: the line sweep
    # (substitution launches and their separate right-hand-side launches)
    kt = r["kernels"]
    ms = kt["line_solve"][0] + kt["line_rhs"][0]
    cnt = kt["line_solve"][1]
    comp = kt["line_solve"][2] + kt["line_rhs"][2]
    lay = kt["line_solve"][3] + kt["line_rhs"][3]
    achieved = comp / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": r["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if VARIANTS[best][0] == 0 else "f64 arithmetic, c64-stored cycles",
        "data": "synthetic",
        "config": {"workload": f"{survey.name}: {survey.nx}x{survey.ny}x{survey.nz} input model, MG "
                               f"{survey.n1}x{survey.n2}x{survey.n3}, 1 shot x {len(survey.freqs)} freqs "
                               f"{survey.freqs} Hz, {len(survey.receivers)} receivers, rel-res 1e-6",
                   "solves_per_step": len(survey.freqs), "variant": best},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (LINE_SWEEP_TRAFFIC_RATIO * comp / max(cnt, 1)
                                 if LINE_SWEEP_TRAFFIC_RATIO else None),
                     "kernel": "line sweep (k_line_solve*, k_line_pcr_solve, k_line_rhs*)",
                     "bytes": "compulsory: a full sweep (4 colours) reads u and f and sigma-bar and writes u once, "
                              "3 Ne s + 3 Nc c; one colour launch a quarter",
                     "avg_launch_us": 1e3 * ms / max(cnt, 1), "launches": cnt,
                     "achieved_with_record_layout_GB_s": lay / (ms * 1e-3) / 1e9 if ms > 0 else 0.0},
        "variants": {k: {"solves_per_s": v["solves_per_s"], "ms_per_step": v["ms_per_step"],
                         "cycles": v["cycles"], "status": v["status"], "max_rel_res": v["rel_res"],
                         "x0_rejected": v["x0_rejected"], "setup_s": v["setup_s"],
                         "kernels": {c: {"ms": round(t[0], 3), "launches": t[1],
                                         "compulsory_GB_s": round(t[2] / (t[0] * 1e-3) / 1e9, 1) if t[0] else 0.0}
                                     for c, t in v["kernels"].items() if t[1]}}
                     for k, v in res.items()},
    }
    if errors:
        out["variant_errors"] = errors
    if not a.no_autotune:
        tab = native.line_tuning()
        out["line_autotune"] = {
            "shapes": len(tab),
            "changed": sum(1 for t in tab if t["us"] < t["us_rule"] * 0.999),
            "us_saved_per_launch_sum": sum(t["us_rule"] - t["us"] for t in tab),
            "choices": {c: sum(1 for t in tab if t["choice"] == c) for c in native.LINE_CHOICES},
            "table": [[t["esize"], t["axis"], t["n"], t["lines"], t["jobs"], t["choice"], round(t["us"], 1),
                       round(t["us_rule"], 1)] for t in tab],
            "columns": ["value bytes", "axis", "line cells", "lines per colour", "rhs per launch", "choice",
                        "us", "us of the rule"]}
    if "plain" in res:
        for k in res:
            if k != "plain":
                out[f"speedup_{k}_vs_plain"] = res[k]["solves_per_s"] / res["plain"]["solves_per_s"]
    if rec:
        shp = r["dcal"].shape
        out["receiver_rel_err_vs_converged_reference"] = rec
        out["reference_at_tol_1e-6_vs_converged_reference"] = receiver_errors(
            survey, gd["dcal"].reshape(shp), gd["dcal_tight"].reshape(shp))
    def leg(key, fn):
        """An auxiliary leg; its failure is reported under its key and does
        not cost the headline line."""
        try:
            out[key] = fn()
        except Exception as e:
            out[key] = {"error": f"{type(e).__name__}: {e}"}
            log(f"bench.py: leg {key} failed: {out[key]['error']}")

    if not a.no_kernels:
        leg("plain_kernels", lambda: {"level": "128^3 level 0, complex FP64", **kernel_rates(survey, 0)})
        leg("mixed_kernels", lambda: {"level": "128^3 level 0, FP32 storage", **kernel_rates(survey, 1)})
        # the initial-guess network's BF16 MFMA convolutions, one training
        # step (forward, input and weight gradients) on the 128^3 grid
        from libemmi_mgfd_amd import x0net
        leg("x0net_conv", lambda: x0net.conv_rates((survey.n3, survey.n2, survey.n1)))
    if not a.no_survey:
        leg("survey", lambda: survey_leg(autotune=not a.no_autotune))
        log("survey", out["survey"].get("solves_per_s"))
    if not a.no_gradient:
        leg("gradient", gradient_leg)
        log("gradient", out["gradient"].get("seconds_per_gradient"))
    if not a.no_cpu:
        leg("cpu_baseline", lambda: cpu_baseline(survey, a.cpu_cycles))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
