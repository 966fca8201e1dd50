"""Learned initial guesses for the solves of one shot.

The network (libemmi_mgfd_amd.x0net) refines x1, the field after one
multigrid cycle from zero, towards the converged field; the solve then
iterates from the refined field to its tolerance, and the solver discards
the guess before its first cycle, restarting from zero, when the guess's
residual ||f - A x0|| is not below the zero guess's ||f||.  The solve from
the guess starts the semi-coarsening pattern sequence afresh (cycle 0): the
right-hand sides of a batch share one sequence, and a rejected guess then
restarts exactly as the zero-guess solve (continuing at pattern 1, as
Solver.set_first_cycle(1) allows, would suit accepted guesses only).  The
weights are trained for the multigrid of the 128^3 survey
(libemmi_mgfd_amd.x0train); other grid shapes are refused.
"""
from __future__ import annotations

import math
import os

import numpy as np

from . import native, x0net

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "x0net_weights.npz")
MU0 = 4e-7 * math.pi


def skin_depth(freq, rho):
    return math.sqrt(2.0 * rho / (2 * math.pi * freq * MU0))


class Guesser:
    """Features, network and reconstruction for one shot's grid.

    grid: dict of the multigrid's float arrays (x1s, x2s, x3s, d1s, d2s, d3s,
    n+1 entries each, as carriers.ShotCarriers.grid returns them); src: source
    position; rho_src: resistivity of the source medium."""

    def __init__(self, n, grid, src, rho_src, weights=WEIGHTS):
        self.n = tuple(int(v) for v in n)
        d = np.load(weights)
        shape = tuple(int(v) for v in d["shape"])
        if shape != (self.n[2], self.n[1], self.n[0]):
            raise ValueError(f"the network was trained for grid {shape[::-1]}, not {self.n}")
        self.net = x0net.UNet(shape, levels=tuple(int(v) for v in d["levels"]))
        self.net.load(d)
        self.dev = {k: x0net.Buf.of(np.ascontiguousarray(grid[k], np.float32)) for k in
                    ("x1s", "x2s", "x3s", "d1s", "d2s", "d3s")}
        self.src = tuple(float(v) for v in src)
        self.rho_src = float(rho_src)
        x = [np.asarray(grid[k], np.float64) for k in ("x1", "x2", "x3")]
        ds = [np.asarray(grid[k], np.float64) for k in ("d1s", "d2s", "d3s")]
        ijk = [int(np.clip(np.searchsorted(x[a], self.src[a]) - 1, 0, self.n[a] - 1)) for a in range(3)]
        self.hsrc = float((ds[0][ijk[0]] * ds[1][ijk[1]] * ds[2][ijk[2]]) ** (1.0 / 3.0))
        self.nv = self.n[0] * self.n[1] * self.n[2]
        self.nvec = 3 * (self.n[0] + 1) * (self.n[1] + 1) * (self.n[2] + 1)
        self.feat = x0net.Buf(2 * self.nv * x0net.CIN, zero=False)

    def features(self, s11_ptr, s33_ptr, freq, x1_ptr):
        d = self.dev
        n1, n2, n3 = self.n
        x0net._ok(x0net._lib().emmi_x0_features(s11_ptr, s33_ptr, d["x1s"].ptr, d["x2s"].ptr, d["x3s"].ptr,
                                                d["d1s"].ptr, d["d2s"].ptr, d["d3s"].ptr, n1, n2, n3,
                                                self.src[0], self.src[1], self.src[2],
                                                skin_depth(freq, self.rho_src), self.hsrc, x1_ptr, x0net.EREF,
                                                self.feat.ptr, None), "features")
        return self.feat

    def guess(self, s11_ptr, s33_ptr, freq, x1_ptr, out):
        """Write the refined field for frequency freq into the device vector
        out (complex double, 3(n1+1)(n2+1)(n3+1))."""
        y = self.net.forward(self.features(s11_ptr, s33_ptr, freq, x1_ptr))
        n1, n2, n3 = self.n
        x0net._ok(x0net._lib().emmi_x0_out_to_edges(y.ptr, x1_ptr, n1, n2, n3, out, None), "out_to_edges")
        return out
