"""The learned-initial-guess data set on the host, without a GPU.

* configs.config4(seed): one random earth per seed, reproducible, drawn from
  the stated ranges (seafloor depth 500-2000 m plus at most 150 m of relief,
  3-6 layers of 0.5-5 ohm-m with anisotropy 1-3, 0-3 boxes of 20-200 ohm-m,
  an x-dipole 50 m above the seafloor, one frequency in 0.1-2 Hz); seeds
  0-1999 train, 2000-2199 test.
* Each model's multigrid is 96^3, stretched about its own source by the
  product's regrid_init, and x0data.config4_chunks builds the grids in worker
  threads in seed order.
* The committed evaluation report (libemmi_mgfd_amd/x0net_eval_96.json,
  written by ``python -m libemmi_mgfd_amd.x0train --config4``) covers the 200
  test seeds and its summary is the summary of its own rows.
"""
import json
import os

import numpy as np
import pytest

from libemmi_mgfd_amd import configs, x0data

REPORT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "libemmi_mgfd_amd",
                      "x0net_eval_96.json")


def test_split():
    tr, te = list(configs.CONFIG4_TRAIN), list(configs.CONFIG4_TEST)
    assert len(tr) == 2000 and len(te) == 200 and not set(tr) & set(te)


@pytest.mark.parametrize("seed", [0, 7, 1999, 2000, 2199])
def test_model_reproducible_and_in_range(seed):
    a, b = configs.config4(seed), configs.config4(seed)
    np.testing.assert_array_equal(a.rho11, b.rho11)
    np.testing.assert_array_equal(a.rho33, b.rho33)
    assert a.sources == b.sources and a.freqs == b.freqs
    assert (a.n1, a.n2, a.n3) == (96, 96, 96)
    # water above the seafloor, sediment and resistors below
    zc = (np.arange(a.nz) + 0.5) * a.dz
    water = a.rho11 == np.float32(configs.RHO_WATER)
    sea = a.bathy
    assert 500.0 - 150.0 <= sea.min() and sea.max() <= 2000.0 + 150.0
    below = zc[:, None, None] >= sea[None]
    assert np.all(water[~below])
    rh, rv = a.rho11[below], a.rho33[below]
    assert rh.min() >= 0.5 - 1e-6 and rh.max() <= 200.0 + 1e-3
    sed = rh <= 5.0 + 1e-6
    ratio = rv[sed] / rh[sed]
    assert ratio.min() >= 1.0 - 1e-6 and ratio.max() <= 3.0 + 1e-6
    assert np.all(rv[~sed] == rh[~sed])    # resistive boxes are isotropic
    # one x-dipole 50 m above the seafloor inside +-2 km, one frequency
    (sx, sy, sz, az, dip), = a.sources
    assert abs(sx) <= 2000 and abs(sy) <= 2000 and az == 0 and dip == 0
    assert len(a.freqs) == 1 and 0.1 <= a.freqs[0] <= 2.0
    assert len(a.receivers) == 31


def test_models_differ():
    a, b = configs.config4(1), configs.config4(2)
    assert a.sources != b.sources and a.freqs != b.freqs
    assert not np.array_equal(a.rho11, b.rho11)


def test_chunks_regrid_about_each_source():
    seen = []
    for seeds, grids in x0data.config4_chunks([3, 4, 5], chunk=2, threads=2):
        for s, g in zip(seeds, grids):
            seen.append(s)
            assert g.n == (96, 96, 96)
            assert g.survey.extra["seed"] == s
            # the source lies in the grid, horizontally in its finest cells
            for a in range(3):
                i = int(np.searchsorted(g.x[a], g.src[a]) - 1)
                assert 0 <= i < g.n[a]
                if a < 2:
                    assert g.ds[a][i] <= 1.01 * g.ds[a].min()
            g.close()
    assert seen == [3, 4, 5]


@pytest.mark.skipif(not os.path.exists(REPORT), reason="no committed evaluation report")
def test_committed_report_is_consistent():
    rep = json.load(open(REPORT))
    rows, s = rep["rows"], rep["summary"]
    assert [r["seed"] for r in rows] == list(configs.CONFIG4_TEST)
    z = np.array([r["cycles_zero"] for r in rows], float)
    x1 = np.array([r["cycles_x1"] for r in rows], float)
    assert s["models"] == 200
    assert s["cycles_zero_mean"] == pytest.approx(z.mean())
    assert s["cycles_x1_mean"] == pytest.approx(x1.mean())
    assert s["all_converged"] and all(st == 0 for r in rows for st in r["status"])
    assert rep["recons"]
    for rc in rep["recons"]:
        ln = np.array([r[f"cycles_learned_{rc}"] for r in rows], float)
        t = s[rc]
        assert t["cycles_learned_mean"] == pytest.approx(ln.mean())
        assert t["vcycles_saved_mean"] == pytest.approx((z - ln).mean())
        assert t["vcycles_saved_total"] == int((z - ln).sum())
        assert t["models_faster"] == int((z > ln).sum()) and t["models_slower"] == int((z < ln).sum())
        assert t["rejected"] == sum(r[f"rejected_{rc}"] for r in rows)
        # the learned path is charged the cycle that makes its input
        assert ln.min() >= 1
        # a rejected guess restarts from zero: the guess's residual was not
        # below ||f||
        for r in rows:
            assert r[f"rejected_{rc}"] == int(not r[f"guess_rel_residual_{rc}"] < 1.0)
    # frequencies are each test seed's own
    for r in rows[:5]:
        assert r["freq"] == configs.config4(r["seed"]).freqs[0]
