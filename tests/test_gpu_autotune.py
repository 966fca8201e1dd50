"""The line-sweep autotuner (emmi_solver_autotune; no counterpart in the
reference, which has one smoother implementation).

For each launch shape of a solve the autotuner times the three substitution
kernels (one lane per line half; four lanes with the right-hand sides in the
same launch; four lanes with a separate right-hand-side launch) and records
the fastest in a process-wide table that later launches of that shape use.
Checked here on the 64^3 survey (six levels, every shape class of the line
sweeps) with three right-hand sides batched:

* the table gets one entry per tuned shape, each choice a valid kernel;
  on the launches long enough to time (20 us and more) its time is no
  worse than the built-in rule's;
* a second call finds every shape tuned already;
* solves after tuning converge in the same cycles to the same magnetic
  field as before (the kernels differ only in rounding order).

The table is cleared afterwards, so the bit-identity tests of batched and
single solves elsewhere run on the built-in rules.
"""
import numpy as np
import pytest

from libemmi_mgfd_amd import carriers, configs, native

pytestmark = pytest.mark.gpu


def test_autotune_table_and_solves():
    s = configs.config1()
    s.freqs = [0.25, 0.75, 1.25]
    sc = carriers.ShotCarriers(s)
    L = native.lib()
    L.regrid_init(sc.acq, sc.emf)
    g = sc.grid()
    xs = sc.x_arrays()
    n = (s.n1, s.n2, s.n3)
    native.line_tuning_clear()
    try:
        def solve():
            S = native.Solver(n, xs[0], xs[1], xs[2], g["sigma11"], g["sigma22"], g["sigma33"], list(sc.omegas),
                              precision=1, tol=1e-8, ncycle=80)
            nb = 16 * sc.nvec
            bs = [native.DeviceBuffer(nb) for _ in s.freqs]
            xo = [native.DeviceBuffer(nb) for _ in s.freqs]
            for f, b in enumerate(bs):
                assert L.emmi_inject_source_dev(sc.acq, sc.emf, f, b.ptr, None) == 0
            return S, bs, xo

        S0, bs, xo = solve()
        st0 = S0.solve_batch_device([b.ptr for b in bs], [x.ptr for x in xo], [b.ptr for b in bs])
        H0 = [b.download(np.complex128) for b in bs]
        S0.close()

        S1, bs, xo = solve()
        print("autotune 64^3, 3 right-hand sides", flush=True)
        ntuned = S1.autotune(reps=5)
        tab = native.line_tuning()
        for t in tab:
            print(t)
        assert ntuned > 0 and len(tab) == ntuned
        for t in tab:
            assert t["choice"] in native.LINE_CHOICES and t["esize"] == 4 and t["jobs"] == 3
            assert t["us"] > 0 and t["us_rule"] > 0
            # launches of a few microseconds vary by more than the kernels
            # differ; the longer ones must not have picked a slower kernel
            if t["us_rule"] >= 20.0:
                assert t["us"] <= 1.15 * t["us_rule"], t
        assert S1.autotune(reps=5) == 0      # every shape of this solver is in the table
        st1 = S1.solve_batch_device([b.ptr for b in bs], [x.ptr for x in xo], [b.ptr for b in bs])
        print("cycles before", [t.cycles for t in st0], "after", [t.cycles for t in st1], flush=True)
        for a, b, h0, hb in zip(st0, st1, H0, bs):
            assert a.status == 0 and b.status == 0 and a.cycles == b.cycles
            h1 = hb.download(np.complex128)
            err = np.abs(h1 - h0).max() / np.abs(h0).max()
            print(f"  H max rel diff {err:.2e}", flush=True)
            assert err <= 1e-6
        S1.close()
    finally:
        native.line_tuning_clear()
    assert native.line_tuning() == []
