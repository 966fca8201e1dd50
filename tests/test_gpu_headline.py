"""Parity on the 128^3 survey (201x201x101 input model, three frequencies):
the GPU solve driven through the product's own host path (regrid_init,
device injection, batched solve, device extraction) reproduces the receiver
data of the reference's converged FP64 solve (tol 1e-10, in
tests/golden/config2.npz) when it is converged as far.

Channels: inline Ex and Hy carry the signal; inline Ey and Hx vanish by the
model's y-symmetry (|Ey| < 1e-16, |Hx| < 2e-15 against |Hy| ~ 1e-5), so only
their smallness is checked.  Bound: 1e-5 relative above a 1e-15 floor."""
import os

import numpy as np
import pytest

from libemmi_mgfd_amd import carriers, configs, native

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config2.npz")


def solve_config2(precision, tol, survey=None):
    """The product's path on config2 (or another survey): grid, device
    injection, batched solve of all frequencies, device extraction.  Returns
    the carriers (with dcal filled), the survey and the solve stats."""
    s = survey or configs.config2()
    sc = carriers.ShotCarriers(s)
    L = native.lib()
    L.regrid_init(sc.acq, sc.emf)
    g = sc.grid()
    xs = sc.x_arrays()
    S = native.Solver((s.n1, s.n2, s.n3), xs[0], xs[1], xs[2], g["sigma11"], g["sigma22"], g["sigma33"],
                      list(sc.omegas), precision=precision, tol=tol, ncycle=200)
    nb = 16 * sc.nvec
    bs = [native.DeviceBuffer(nb) for _ in s.freqs]
    xo = [native.DeviceBuffer(nb) for _ in s.freqs]
    for f, b in enumerate(bs):
        assert L.emmi_inject_source_dev(sc.acq, sc.emf, f, b.ptr, None) == 0
    sts = S.solve_batch_device([b.ptr for b in bs], [x.ptr for x in xo], [b.ptr for b in bs])
    for f in range(len(s.freqs)):
        assert L.emmi_extract_emf_data_dev(sc.acq, sc.emf, f, xo[f].ptr, bs[f].ptr, None) == 0
    return sc, s, sts


@pytest.fixture(params=[0, 1, 2], ids=["auto", "one_lane", "quad_lane"])
def line_kernel(request):
    L = native.lib()
    prev = L.emmi_set_line_kernel(request.param)
    yield request.param
    L.emmi_set_line_kernel(prev)


def test_config2_mixed_matches_fp64_at_bench_tolerance():
    """FP32-storage cycles (one-lane substitution on the two finest levels,
    quad-lane below) against the FP64 solve at the benchmark's tolerance:
    the signal channels agree within 1e-5."""
    print("config2 batched solves, precision 1 and 0, tol 1e-6", flush=True)
    sc1, s, st1 = solve_config2(1, 1e-6)
    d1 = sc1.dcal.copy()
    sc0, _, st0 = solve_config2(0, 1e-6)
    print("cycles mixed", [t.cycles for t in st1], "fp64", [t.cycles for t in st0])
    assert all(t.status == 0 for t in st1 + st0)
    for ch, name in enumerate(s.chrec):
        if name in ("Ey", "Hx"):
            continue
        a = np.abs(sc0.dcal[ch])
        m = a > 1e-15
        err = np.abs(d1[ch] - sc0.dcal[ch])[m] / a[m]
        print(f"{name}: mixed vs fp64 max rel {err.max():.2e}")
        assert err.max() < 1e-5, name


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 1e-10)])
def test_config2_receivers_match_converged_reference(precision, tol, line_kernel):
    s = configs.config2()
    sc = carriers.ShotCarriers(s)
    L = native.lib()
    L.regrid_init(sc.acq, sc.emf)
    g = sc.grid()
    ref_grid = np.load(GOLD)
    for k in ("x1", "x2", "x3", "x1s", "x2s", "x3s", "d1", "d2", "d3"):
        np.testing.assert_array_equal(g[k], ref_grid[k], err_msg=k)
    n = (s.n1, s.n2, s.n3)
    xs = sc.x_arrays()
    S = native.Solver(n, xs[0], xs[1], xs[2], g["sigma11"], g["sigma22"], g["sigma33"], list(sc.omegas),
                      precision=precision, tol=tol, ncycle=200)
    nb = 16 * sc.nvec
    bs = [native.DeviceBuffer(nb) for _ in s.freqs]
    xo = [native.DeviceBuffer(nb) for _ in s.freqs]
    for f, b in enumerate(bs):
        assert L.emmi_inject_source_dev(sc.acq, sc.emf, f, b.ptr, None) == 0
    print(f"config2 batched solve precision {precision} tol {tol} line kernel {line_kernel}", flush=True)
    sts = S.solve_batch_device([b.ptr for b in bs], [x.ptr for x in xo], [b.ptr for b in bs])
    print("cycles", [st.cycles for st in sts], "status", [st.status for st in sts])
    assert all(st.status == 0 for st in sts)
    for f in range(len(s.freqs)):
        assert L.emmi_extract_emf_data_dev(sc.acq, sc.emf, f, xo[f].ptr, bs[f].ptr, None) == 0
    ref = ref_grid["dcal_tight"].reshape(sc.dcal.shape)
    for ch, name in enumerate(s.chrec):
        a = np.abs(ref[ch])
        if name in ("Ey", "Hx"):
            assert np.abs(sc.dcal[ch]).max() < 1e-13 * max(np.abs(ref[3]).max(), 1e-30) + 5e-15
            continue
        m = a > 1e-15
        err = np.abs(sc.dcal[ch] - ref[ch])[m] / a[m]
        print(f"{name}: max rel {err.max():.2e} median {np.median(err):.2e}")
        assert err.max() < 1e-5, name


def test_config1_receivers_match_converged_reference():
    """The 64^3 survey (six levels, 101 inline receivers, 0.25 Hz) solved to
    1e-10 on the GPU against the reference's own solve to 1e-10
    (tests/golden/config1.npz): signal channels within 1e-5, the inline
    null channels (Ey, Hx) at symmetry-noise level."""
    gold = np.load(os.path.join(os.path.dirname(GOLD), "config1.npz"))
    print("config1 solve, precision 0, tol 1e-10", flush=True)
    sc, s, sts = solve_config2(0, 1e-10, configs.config1())
    print("cycles", [t.cycles for t in sts])
    assert all(t.status == 0 for t in sts)
    ref = gold["dcal_tight"].reshape(sc.dcal.shape)
    sig = max(np.abs(ref[ch]).max() for ch, name in enumerate(s.chrec) if name not in ("Ey", "Hx"))
    for ch, name in enumerate(s.chrec):
        if name in ("Ey", "Hx"):
            assert np.abs(sc.dcal[ch]).max() < 1e-6 * sig, name
            continue
        a = np.abs(ref[ch])
        m = a > 1e-15
        err = np.abs(sc.dcal[ch] - ref[ch])[m] / a[m]
        print(f"{name}: max rel {err.max():.2e} median {np.median(err):.2e}")
        assert err.max() < 1e-5, name
