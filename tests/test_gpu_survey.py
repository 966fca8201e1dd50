"""Batched survey solves on the GPU (libemmi_mgfd_amd/survey.py,
emmi_solver_create_survey): right-hand sides of different shots -- each
shot regridded about its own source, so each with its own grid and medium --
share every launch of a cycle.

* Bit-identical to solving each (shot, frequency) alone: the batch changes
  which launch serves a right-hand side, not its arithmetic.
* The batched survey's (shot, frequency) pairs at their own multigrid sizes
  (96^3, 128^3, 128^3, 160x160x128) against the reference's converged
  receiver data (tests/golden/config3.npz, make_survey_golden.py): both
  solved to a relative residual of 1e-10, FP32 cycles in FP64 refinement on
  the GPU.
"""
import os

import numpy as np
import pytest

from helpers import cli_survey
from libemmi_mgfd_amd import carriers, configs, native, survey

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("precision", [0, 1])
def test_survey_batch_equals_single_solves(precision):
    s = cli_survey()
    s.freqs = [0.5, 1.25]
    print(f"3 shots x 2 frequencies on the tiny model, precision {precision}", flush=True)
    run = survey.SurveyRun(s, precision=precision, tol=1e-8, ncycle=80, max_slots=6)
    assert len(run.chunks) == 1 and len(run.chunks[0].shots) == 3
    run.run()
    L = native.lib()
    for shot in range(3):
        sc = carriers.ShotCarriers(s, shot)
        L.regrid_init(sc.acq, sc.emf)
        g = sc.grid()
        x1, x2, x3 = sc.x_arrays()
        nb = 16 * sc.nvec
        for f in range(2):
            S = native.Solver((s.n1, s.n2, s.n3), x1, x2, x3, g["sigma11"], g["sigma22"], g["sigma33"],
                              sc.omegas[f], precision=precision, tol=1e-8, ncycle=80)
            b = native.DeviceBuffer(nb)
            x = native.DeviceBuffer(nb)
            assert L.emmi_inject_source_dev(sc.acq, sc.emf, f, b.ptr, None) == 0
            st = S.solve_device(b.ptr, x.ptr, b.ptr)
            assert L.emmi_extract_emf_data_dev(sc.acq, sc.emf, f, x.ptr, b.ptr, None) == 0
            r = run.results[(shot, f)]
            assert r["status"] == 0 and r["cycles"] == st.cycles, (shot, f, r["cycles"], st.cycles)
            np.testing.assert_array_equal(r["dcal"], sc.dcal[:, f, :])


def test_config3_pairs_match_converged_reference():
    d = np.load(os.path.join(GOLD, "config3.npz"))
    s = configs.config3()
    pairs = [(int(p[0]) - 1, s.freqs.index(float(p[1]))) for p in d["pairs"]]
    print("config3 pairs (shot, freq):", [(p + 1, s.freqs[f]) for p, f in pairs], flush=True)
    run = survey.SurveyRun(s, grids=configs.CONFIG3_GRIDS, pairs=pairs, precision=1, tol=1e-10, ncycle=200)
    run.run()
    worst = 0.0
    for shot, f in pairs:
        r = run.results[(shot, f)]
        ref = d[f"s{shot + 1}_dcal"].reshape(len(s.chrec), -1)
        got = r["dcal"]
        print(f"  shot {shot + 1} {s.freqs[f]} Hz on {r['n']}: status {r['status']}, {r['cycles']} cycles "
              f"(reference {len(d[f's{shot + 1}_hist'])})", flush=True)
        assert r["status"] == 0
        for ch, name in enumerate(s.chrec):
            amp = np.abs(ref[ch])
            strong = amp > 1e-3 * amp.max()
            floor = amp > 1e-15
            e_strong = float(np.max(np.abs(got[ch] - ref[ch])[strong] / amp[strong]))
            e_floor = float(np.max(np.abs(got[ch] - ref[ch])[floor] / amp[floor])) if floor.any() else 0.0
            print(f"    {name}: max rel err {e_strong:.2e} (|d| > 1e-3 max), {e_floor:.2e} (|d| > 1e-15)", flush=True)
            assert e_strong < 1e-5, (shot, f, name)
            worst = max(worst, e_strong)
    print(f"  worst {worst:.2e}", flush=True)


def test_survey_packs_of_sixteen_equal_single_solves():
    """18 right-hand sides (9 shots x 2 frequencies) in one solver: each
    launch serves up to 16 of them, so the cycles run one full pack and one
    of 2; every right-hand side still gets its single-solve iterates."""
    s = cli_survey()
    xs = np.linspace(-1500.0, 1500.0, 3)
    s.sources = [(float(x), float(y), 450.0, 15.0 * k, 2.0) for k, (x, y) in
                 enumerate((x, y) for y in (-300.0, 0.0, 300.0) for x in xs)]
    s.table = [(i + 1, r + 1) for i in range(len(s.sources)) for r in range(len(s.receivers))]
    s.freqs = [0.5, 1.25]
    print("9 shots x 2 frequencies, one survey solver", flush=True)
    run = survey.SurveyRun(s, precision=1, tol=1e-8, ncycle=80, max_slots=18)
    assert len(run.chunks) == 1 and len(run.chunks[0].pairs) == 18
    run.run()
    L = native.lib()
    for shot in (0, 4, 8):
        sc = carriers.ShotCarriers(s, shot)
        L.regrid_init(sc.acq, sc.emf)
        g = sc.grid()
        x1, x2, x3 = sc.x_arrays()
        nb = 16 * sc.nvec
        for f in range(2):
            S = native.Solver((s.n1, s.n2, s.n3), x1, x2, x3, g["sigma11"], g["sigma22"], g["sigma33"],
                              sc.omegas[f], precision=1, tol=1e-8, ncycle=80)
            b = native.DeviceBuffer(nb)
            x = native.DeviceBuffer(nb)
            assert L.emmi_inject_source_dev(sc.acq, sc.emf, f, b.ptr, None) == 0
            st = S.solve_device(b.ptr, x.ptr, b.ptr)
            assert L.emmi_extract_emf_data_dev(sc.acq, sc.emf, f, x.ptr, b.ptr, None) == 0
            r = run.results[(shot, f)]
            assert r["status"] == 0 and r["cycles"] == st.cycles
            np.testing.assert_array_equal(r["dcal"], sc.dcal[:, f, :])
