"""Survey-scale grids on the host, without a GPU: the product's regrid_init
(src/regrid_init_free.c:37-232) builds, bit for bit, the reference's
multigrid node arrays for the batched survey's (shot, frequency) pairs, each
on its frequency's own multigrid size (96^3, 128^3, 128^3, 160x160x128), and
for the survey-scale gradient's shot (160x160x128).  Fixtures:
tests/golden/config3.npz and tests/golden/config5/ (make_survey_golden.py)."""
import dataclasses
import os

import numpy as np
import pytest

from libemmi_mgfd_amd import carriers, configs, native, survey

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def regridded(s, shot, n, freqs):
    s2 = dataclasses.replace(s, n1=n[0], n2=n[1], n3=n[2], freqs=list(freqs))
    sc = carriers.ShotCarriers(s2, shot)
    native.lib().regrid_init(sc.acq, sc.emf)
    return sc


@pytest.mark.parametrize("shot,freq", [(7, 0.1), (13, 0.25), (19, 0.5), (25, 1.0)])
def test_config3_grids_bitexact(shot, freq):
    d = np.load(os.path.join(GOLD, "config3.npz"))
    s = configs.config3()
    n = tuple(int(v) for v in d[f"s{shot}_n"])
    assert n == configs.CONFIG3_GRIDS[freq]
    sc = regridded(s, shot - 1, n, [freq])
    x1, x2, x3 = sc.x_arrays()
    np.testing.assert_array_equal(x1, d[f"s{shot}_x1"])
    np.testing.assert_array_equal(x2, d[f"s{shot}_x2"])
    np.testing.assert_array_equal(x3, d[f"s{shot}_x3"])


def test_config3_survey_layout():
    """25 shots x 4 frequencies = 100 right-hand sides in chunks of one
    multigrid size each; 405 receivers per shot."""
    s = configs.config3()
    assert s.nshots == 25 and len(s.receivers) == 405 and len(s.freqs) == 4
    run = survey.SurveyRun(s, grids=configs.CONFIG3_GRIDS, max_slots=25)
    assert sum(len(c.pairs) for c in run.chunks) == 100
    assert sorted({c.n for c in run.chunks}) == [(96, 96, 96), (128, 128, 128), (160, 160, 128)]
    for c in run.chunks:
        assert len(c.pairs) <= 25
        assert all(configs.CONFIG3_GRIDS[s.freqs[f]] == c.n for _, f in c.pairs)
