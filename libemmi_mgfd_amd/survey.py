"""Batched survey solves: every (shot, frequency) right-hand side of a survey
on one GPU.

The reference runs one MPI rank per shot and, inside a rank, one
gmg_init/gmg_apply/gmg_free per frequency (src/acq_init_free.c:120-142,
src/do_modelling.c:49-67); each shot is regridded about its own source
(src/regrid_init_free.c:158-204), so no two shots share an operator.  Here the
right-hand sides are grouped by multigrid size (a survey may give each
frequency its own size) and each group is solved by survey solvers
(emmi_solver_create_survey): one solver holds the grids and media of its
shots and up to 64 right-hand sides, and every kernel launch of a cycle serves
all of its unconverged right-hand sides whatever their grid.  Each
right-hand side retires at its own convergence test, so it gets exactly the
iterates it would get alone.  Source injection and receiver extraction run on
the device (emmi_inject_source_dev, emmi_extract_emf_data_dev).

With ``autotune`` each chunk's solver first times the line-sweep kernels on
its launch shapes (emmi_solver_autotune, counted in ``setup_s``).

Groups are solved one after another, each in chunks of at most ``max_slots``
right-hand sides, so that the device memory of one chunk (about 2.5 GB per
128^3 right-hand side in mixed precision: factor records, vectors of all
levels) bounds the footprint.
"""
from __future__ import annotations

import dataclasses
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import carriers, configs, native


@dataclasses.dataclass
class Chunk:
    n: Tuple[int, int, int]
    pairs: List[Tuple[int, int]]           # (shot, frequency index in the survey)
    shots: List[int]                       # the grids (media) of the chunk, by shot
    carriers: Dict[int, carriers.ShotCarriers]
    group_freqs: List[float]               # frequencies of the group (carriers index these)


class SurveyRun:
    """All (shot, frequency) pairs of ``survey`` (or ``pairs``, 0-based shot
    and frequency indices), frequency f on the multigrid size grids[f]
    (default: the survey's n1, n2, n3)."""

    def __init__(self, survey: configs.Survey, grids: Optional[Dict[float, Sequence[int]]] = None,
                 pairs: Optional[Sequence[Tuple[int, int]]] = None, precision: int = 1, tol: float = 1e-6,
                 ncycle: int = 60, max_slots: int = 25, graphs: bool = True, autotune: bool = False):
        self.survey = survey
        self.kw = dict(precision=precision, tol=tol, ncycle=ncycle, graphs=1 if graphs else 0)
        if grids is None:
            grids = {f: (survey.n1, survey.n2, survey.n3) for f in survey.freqs}
        self.grids = {float(f): tuple(int(v) for v in n) for f, n in grids.items()}
        if pairs is None:
            pairs = [(s, f) for f in range(len(survey.freqs)) for s in range(survey.nshots)]
        self.pairs = [(int(s), int(f)) for s, f in pairs]
        if not 1 <= max_slots <= 64:
            raise ValueError("max_slots must be 1..64")
        self.max_slots = max_slots
        self.autotune = autotune
        # group by multigrid size, keep the survey's frequency order
        groups: Dict[Tuple[int, int, int], List[Tuple[int, int]]] = {}
        for s, f in self.pairs:
            groups.setdefault(self.grids[float(survey.freqs[f])], []).append((s, f))
        self.chunks: List[Chunk] = []
        for n, ps in groups.items():
            gf = sorted({f for _, f in ps})
            for i in range(0, len(ps), max_slots):
                cp = ps[i:i + max_slots]
                shots = sorted({s for s, _ in cp})
                self.chunks.append(Chunk(n, cp, shots, {}, [survey.freqs[f] for f in gf]))
        self.setup_s = 0.0
        self.results: Dict[Tuple[int, int], dict] = {}

    def _regrid(self, ch: Chunk):
        """Carriers of each shot of the chunk on the chunk's grid size, with
        the group's frequencies, after regrid_init (host, bit-exact)."""
        s = dataclasses.replace(self.survey, n1=ch.n[0], n2=ch.n[1], n3=ch.n[2], freqs=list(ch.group_freqs))
        L = native.lib()
        for shot in ch.shots:
            sc = carriers.ShotCarriers(s, shot)
            L.regrid_init(sc.acq, sc.emf)
            ch.carriers[shot] = sc

    def _build(self, ch: Chunk):
        t0 = time.perf_counter()
        if not ch.carriers:
            self._regrid(ch)
        grids = []
        for shot in ch.shots:
            sc = ch.carriers[shot]
            g = sc.grid()
            x1, x2, x3 = sc.x_arrays()
            grids.append((x1, x2, x3, g["sigma11"], g["sigma22"], g["sigma33"]))
        med = [ch.shots.index(s) for s, _ in ch.pairs]
        fidx = [ch.group_freqs.index(self.survey.freqs[f]) for _, f in ch.pairs]
        om = [ch.carriers[s].omegas[fi] for (s, _), fi in zip(ch.pairs, fidx)]
        solver = native.Solver.survey(ch.n, grids, med, om, **self.kw)
        if self.autotune:
            solver.autotune()
        nb = 16 * 3 * (ch.n[0] + 1) * (ch.n[1] + 1) * (ch.n[2] + 1)
        bufs = {"b": [native.DeviceBuffer(nb) for _ in ch.pairs], "x": [native.DeviceBuffer(nb) for _ in ch.pairs]}
        native.lib().emmi_device_sync()
        self.setup_s += time.perf_counter() - t0
        return solver, bufs, fidx

    def _solve_chunk(self, ch: Chunk, solver, bufs, fidx):
        L = native.lib()
        for k, (s, _) in enumerate(ch.pairs):
            sc = ch.carriers[s]
            if L.emmi_inject_source_dev(sc.acq, sc.emf, fidx[k], bufs["b"][k].ptr, None):
                raise native.NativeError(L.emmi_path_last_error().decode())
        sts = solver.solve_batch_device([b.ptr for b in bufs["b"]], [x.ptr for x in bufs["x"]],
                                        [b.ptr for b in bufs["b"]])
        for k, (s, f) in enumerate(ch.pairs):
            sc = ch.carriers[s]
            if L.emmi_extract_emf_data_dev(sc.acq, sc.emf, fidx[k], bufs["x"][k].ptr, bufs["b"][k].ptr, None):
                raise native.NativeError(L.emmi_path_last_error().decode())
            self.results[(s, f)] = {"dcal": sc.dcal[:, fidx[k], :].copy(), "status": sts[k].status,
                                    "cycles": sts[k].cycles, "rel_res": sts[k].rnorm / sts[k].rnorm0,
                                    "n": ch.n}
        return sts

    def run(self, repeats: int = 0) -> dict:
        """Solves every chunk once (the first solve of a chunk builds its
        line factors and captures its cycle graphs), then ``repeats`` more
        times timed.  Returns the timing summary; receiver data and stats per
        pair are in ``results``."""
        total_first, total_timed, nrhs = 0.0, 0.0, 0
        per_chunk = []
        for ch in self.chunks:
            solver, bufs, fidx = self._build(ch)
            L = native.lib()
            t0 = time.perf_counter()
            self._solve_chunk(ch, solver, bufs, fidx)
            L.emmi_device_sync()
            t1 = time.perf_counter()
            for _ in range(repeats):
                self._solve_chunk(ch, solver, bufs, fidx)
            L.emmi_device_sync()
            t2 = time.perf_counter()
            total_first += t1 - t0
            total_timed += t2 - t1
            nrhs += len(ch.pairs)
            per_chunk.append({"n": ch.n, "rhs": len(ch.pairs), "shots": len(ch.shots), "first_s": t1 - t0,
                              "timed_s": (t2 - t1) / max(repeats, 1)})
            solver.close()
            del bufs
            ch.carriers.clear()
        return {"rhs": nrhs, "first_s": total_first, "timed_s": total_timed / max(repeats, 1),
                "setup_s": self.setup_s, "chunks": per_chunk}
