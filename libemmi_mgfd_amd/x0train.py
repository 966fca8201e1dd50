"""Train the initial-guess network on solver-generated fields and evaluate it
through the solver.

    python -m libemmi_mgfd_amd.x0train --models 100 --steps 3000 --out weights.npz

Data: libemmi_mgfd_amd.x0data (random earths on the 128^3 survey grid, three
frequencies each).  Training: the U-Net of libemmi_mgfd_amd.x0net on its own
HIP kernels, Adam with cosine decay, one sample per step.  Evaluation: held-out
models solved to relative residual 1e-6 from a zero guess and from the
network's prediction, which refines the field after one multigrid cycle (FP32
cycles in FP64 refinement, all three frequencies batched); the learned path
is charged that extra cycle.  The solver judges the prediction before its
first cycle and falls back to zero when its residual is not below ||f||.

    python -m libemmi_mgfd_amd.x0train --config4 --models 2000 --steps 12000 \
        --lr 1e-3 --out x0net_weights_96.npz --eval-out x0net_eval_96.json

trains on the config-4 earths instead (each on its own 96^3 multigrid, one
frequency each, online from a replay buffer; train_config4) and evaluates on
the test seeds with libemmi_mgfd_amd.x0eval.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import native, x0data, x0net


def predict(net, grid, feat, x1):
    """The network's refinement of the one-cycle field x1, as a device edge
    vector (complex double)."""
    y = net.forward(feat)
    x0 = native.DeviceBuffer(16 * grid.sc.nvec)
    n1, n2, n3 = grid.n
    x0net._ok(x0net._lib().emmi_x0_out_to_edges(y.ptr, x1.ptr, n1, n2, n3, x0.ptr, None), "out_to_edges")
    native.lib().emmi_device_sync()
    return x0


def evaluate(net, grid, nmodels, seed, log=print):
    """Cycles to rel-res 1e-6 from zero and from the prediction on fresh
    random models (three frequencies per model, batched)."""
    rng = np.random.default_rng(seed)
    rows = []
    for m in range(nmodels):
        s11, s33 = x0data.random_model(grid, rng)
        freqs = list(np.exp(rng.uniform(np.log(0.1), np.log(2.0), 3)))
        rows += solve_pair(net, grid, s11, s33, freqs)
        log(f"eval model {m}: " + ", ".join(f"f={r['freq']:.3f} zero {r['cycles_zero']} net {r['cycles_x0']}"
                                            f"{' (rejected)' if r['rejected'] else ''}" for r in rows[-3:]))
    return rows


def solve_pair(net, grid, s11, s33, freqs, tol=1e-6):
    """Cycles from zero, and for the learned path 1 + cycles from the network's
    refinement of the one-cycle field, for one medium at several frequencies
    (batched)."""
    n = grid.n
    ds11, ds33 = x0net.Buf.of(s11), x0net.Buf.of(s33)
    xs = grid.sc.x_arrays()
    om = [np.float32(2.0 * math.pi * float(np.float32(f))) for f in freqs]
    S = native.Solver(n, xs[0], xs[1], xs[2], ds11.ptr, ds11.ptr, ds33.ptr, om, sig_on_device=True, precision=1,
                      tol=tol, ncycle=60)
    nb = 16 * grid.sc.nvec
    bs = x0data.source_terms(grid, om)
    hs = [native.DeviceBuffer(nb) for _ in om]
    xo = [native.DeviceBuffer(nb) for _ in om]

    def run(x0s):
        x0data.copy_all(hs, bs)
        return S.solve_batch_device([h.ptr for h in hs], [x.ptr for x in xo], [h.ptr for h in hs], x0s)

    st0 = run(None)
    x0 = learned_guess(S, net, grid, ds11, ds33, freqs, run, xo)
    st1 = run([p.ptr for p in x0])
    S.close()
    return [{"freq": float(f), "cycles_zero": a.cycles, "cycles_x0": 1 + b.cycles, "rejected": int(b.x0_rejected),
             "status": [a.status, b.status]} for f, a, b in zip(freqs, st0, st1)]


def learned_guess(S, net, grid, ds11, ds33, freqs, run, xo, ncycle=60):
    """One batched cycle from zero, then the network's refinement of it."""
    S.set_ncycle(1)
    run(None)
    S.set_ncycle(ncycle)
    return [predict(net, grid, x0data.features(grid, ds11, ds33, f, x), x) for f, x in zip(freqs, xo)]


def train(net, samples, steps, lr0, log=print, val=None):
    rng = np.random.default_rng(0)
    hist = []
    t0 = time.perf_counter()
    for it in range(steps):
        sm = samples[int(rng.integers(len(samples)))]
        lr = lr0 * 0.5 * (1 + math.cos(math.pi * it / steps)) + 0.02 * lr0
        net.forward(sm.feat)
        loss = net.loss_and_grad(sm.tgt, sm.mask, sm.count)
        net.backward()
        net.adam(lr=lr)
        hist.append(loss)
        if it % 100 == 0 or it == steps - 1:
            msg = f"step {it} loss {np.mean(hist[-100:]):.4f} lr {lr:.2e} {time.perf_counter() - t0:.0f}s"
            if val:
                vl = []
                for v in val:
                    net.forward(v.feat)
                    vl.append(net.loss_and_grad(v.tgt, v.mask, v.count))
                msg += f" val {np.mean(vl):.4f}"
            log(msg)
    return hist


def train_config4(net, seeds, steps, lr0, chunk=16, buffer=384, log=print):
    """Online training on the config-4 earths of ``seeds``: each chunk of
    models is solved (x1 and the target field, x0data.solve_models), its
    samples join a replay buffer of the last ``buffer`` samples (about 120 MB
    each at 96^3), and the chunk's share of the ``steps`` Adam steps follows
    on samples drawn from the buffer.  Returns the loss history and the
    solver's cycle counts to the targets."""
    from collections import deque
    rng = np.random.default_rng(0)
    buf = deque(maxlen=buffer)
    seeds = list(seeds)
    nchunks = max(1, (len(seeds) + chunk - 1) // chunk)
    hist, cycles = [], []
    it = 0
    t0 = time.perf_counter()
    for k, (ss, grids) in enumerate(x0data.config4_chunks(seeds, chunk)):
        solved = x0data.solve_models(grids)
        for ms in solved:
            smp = x0data.sample_of(ms)
            cycles.append(ms.stats.cycles)
            if smp is None:
                log(f"model {ms.grid.survey.extra['seed']}: solve status {ms.stats.status}, skipped")
            else:
                buf.append(smp)
        for g in grids:
            g.close()
        del solved, grids
        nsteps = round((k + 1) * steps / nchunks) - round(k * steps / nchunks)
        for _ in range(nsteps):
            sm = buf[int(rng.integers(len(buf)))]
            lr = lr0 * 0.5 * (1 + math.cos(math.pi * it / steps)) + 0.02 * lr0
            net.forward(sm.feat)
            hist.append(net.loss_and_grad(sm.tgt, sm.mask, sm.count))
            net.backward()
            net.adam(lr=lr)
            it += 1
        log(f"chunk {k + 1}/{nchunks} seeds {ss[0]}-{ss[-1]}: buffer {len(buf)}, step {it}, "
            f"loss {np.mean(hist[-max(nsteps, 1):]) if hist else float('nan'):.4f}, "
            f"cycles to 1e-7 {np.mean(cycles[-len(ss):]):.1f}, {time.perf_counter() - t0:.0f}s")
    return hist, cycles


def main(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("--models", type=int, default=60)
    p.add_argument("--val-models", type=int, default=6)
    p.add_argument("--steps", type=int, default=2000)
    p.add_argument("--lr", type=float, default=2e-3)
    p.add_argument("--eval-models", type=int, default=4)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--out", default="x0net_weights.npz")
    p.add_argument("--init", default=None, help="start from these weights")
    p.add_argument("--config4", action="store_true",
                   help="train on the config-4 earths (96^3, one grid per model, seeds 0..models-1) and "
                        "evaluate on the test seeds (x0eval)")
    p.add_argument("--chunk", type=int, default=16, help="config-4 models per batched solve")
    p.add_argument("--buffer", type=int, default=384, help="config-4 replay buffer (samples)")
    p.add_argument("--eval-out", default=None, help="config-4 evaluation report (JSON)")
    a = p.parse_args(argv)

    def log(m):
        print(m, flush=True)

    if a.config4:
        from . import configs, x0eval
        seeds = list(configs.CONFIG4_TRAIN)[:a.models]
        net = x0net.UNet((96, 96, 96), seed=a.seed)
        if a.init:
            net.load(np.load(a.init))
        t0 = time.perf_counter()
        hist, cycles = train_config4(net, seeds, a.steps, a.lr, chunk=a.chunk, buffer=a.buffer, log=log)
        state = net.state()
        state["train_loss"] = np.array(hist, np.float32)
        np.savez_compressed(a.out, **state)
        log(f"saved {a.out} after {time.perf_counter() - t0:.0f}s")
        test = list(configs.CONFIG4_TEST)[:a.eval_models]
        rep = x0eval.evaluate(net, test, chunk=a.chunk, log=log)
        rep["training"] = {"models": len(seeds), "steps": a.steps, "lr": a.lr, "buffer": a.buffer,
                           "final_loss": float(np.mean(hist[-200:])) if hist else None,
                           "target_cycles_mean": float(np.mean(cycles)) if cycles else None}
        if a.eval_out:
            with open(a.eval_out, "w") as f:
                json.dump(rep, f, indent=1)
        log(json.dumps(rep["summary"]))
        return 0

    grid = x0data.Grid()
    t0 = time.perf_counter()
    samples = x0data.generate(grid, a.models, seed=a.seed, log=log)
    val = x0data.generate(grid, a.val_models, seed=10_000 + a.seed, log=log)
    log(f"data: {len(samples)} training and {len(val)} validation samples in {time.perf_counter() - t0:.0f}s")
    net = x0net.UNet((grid.n[2], grid.n[1], grid.n[0]), seed=a.seed)
    if a.init:
        net.load(np.load(a.init))
    hist = train(net, samples, a.steps, a.lr, log=log, val=val)
    state = net.state()
    state["train_loss"] = np.array(hist, np.float32)
    np.savez_compressed(a.out, **state)
    log(f"saved {a.out}")
    del samples
    rows = evaluate(net, grid, a.eval_models, seed=20_000 + a.seed, log=log)
    z = np.mean([r["cycles_zero"] for r in rows])
    x = np.mean([r["cycles_x0"] for r in rows])
    log(json.dumps({"eval_cycles_zero": z, "eval_cycles_x0": x, "rows": rows}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
