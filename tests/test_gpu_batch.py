"""Batched solves: slots that share one grid and medium, each at its own
frequency, solved together must give every slot exactly the iterates of a
single solve at that frequency (same kernels on the same inputs), retire at
their own cycle counts, and honour an initial guess."""
import numpy as np
import pytest

import pyoracle as O
from helpers import rel
from libemmi_mgfd_amd import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    s = configs.tiny()
    g = O.regrid(s)
    rw = float(min(s.rho11.min(), s.rho22.min(), s.rho33.min()))
    mg = O.mg_struct(g, rw)
    bs = [O.inject_source(mg, s.sources, s.chsrc, fr) for fr in s.freqs]
    return s, g, bs


def _solver(s, g, omegas, precision, **kw):
    from libemmi_mgfd_amd import native
    return native.Solver((s.n1, s.n2, s.n3), g["x1"], g["x2"], g["x3"], g["sigma11"], g["sigma22"], g["sigma33"],
                         omegas, precision=precision, **kw)


def _batch(S, bs, x0s=None):
    from libemmi_mgfd_amd import native
    nb = 16 * S.nvec
    db = [native.DeviceBuffer(nb).upload(np.ascontiguousarray(b, np.complex128)) for b in bs]
    dx = [native.DeviceBuffer(nb) for _ in bs]
    d0 = None
    if x0s is not None:
        d0 = [native.DeviceBuffer(nb).upload(np.ascontiguousarray(x, np.complex128)) if x is not None else None
              for x in x0s]
    sts = S.solve_batch_device([b.ptr for b in db], [x.ptr for x in dx], [b.ptr for b in db],
                               None if d0 is None else [d.ptr if d is not None else None for d in d0])
    return [x.download(np.complex128) for x in dx], [b.download(np.complex128) for b in db], sts


@pytest.mark.parametrize("precision,tol", [(0, 1e-10), (1, 1e-7)])
def test_batch_equals_single_solves(case, precision, tol):
    s, g, bs = case
    om = [O.omega_f32(f) for f in s.freqs]
    print(f"batch of {len(om)} slots, precision {precision}, tol {tol}", flush=True)
    B = _solver(s, g, om, precision, tol=tol, ncycle=200)
    Eb, Hb, sts = _batch(B, bs)
    single = _solver(s, g, om[0], precision, tol=tol, ncycle=200)
    for k, fr in enumerate(s.freqs):
        single.set_omega(om[k])
        E, H, st = single.solve(bs[k])
        print(f"slot {k}: batch cycles {sts[k].cycles} status {sts[k].status}; single cycles {st.cycles}")
        assert sts[k].status == 0
        assert sts[k].cycles == st.cycles
        np.testing.assert_array_equal(B.history(k), single.history())
        assert rel(Eb[k], E) == 0.0
        assert rel(Hb[k], H) == 0.0


def test_batch_retires_slots_independently(case):
    """Slot 1 starts from its converged field and retires at cycle 0 while
    slot 0 keeps iterating; slot 0's result is unchanged by its neighbour."""
    s, g, bs = case
    om = [O.omega_f32(f) for f in s.freqs]
    single = _solver(s, g, om[1], 0, tol=1e-9, ncycle=200)
    E1, _, _ = single.solve(bs[1])
    B = _solver(s, g, om, 0, tol=1e-9, ncycle=200)
    Eb, _, sts = _batch(B, bs, x0s=[None, E1])
    print("cycles", [st.cycles for st in sts], "status", [st.status for st in sts])
    assert sts[1].status == 0 and sts[1].cycles == 0
    assert sts[0].status == 0 and sts[0].cycles > 0
    ref = _solver(s, g, om[0], 0, tol=1e-9, ncycle=200)
    E0, _, st0 = ref.solve(bs[0])
    assert sts[0].cycles == st0.cycles
    assert rel(Eb[0], E0) == 0.0
    assert rel(Eb[1], E1) < 1e-14


def test_batch_rejects_oversized_batch(case):
    from libemmi_mgfd_amd import native
    s, g, bs = case
    B = _solver(s, g, [O.omega_f32(s.freqs[0])], 0)
    with pytest.raises(native.NativeError):
        _batch(B, bs)


@pytest.mark.parametrize("precision,tol", [(0, 1e-9), (1, 1e-7)])
def test_graph_replay_matches_eager(case, precision, tol):
    """Cycle bodies replayed from captured HIP graphs give bit-identical
    fields and residual histories to eager launches, over repeated solves,
    a change of frequencies (graphs dropped and recaptured) and slots that
    retire at different cycles (new active-set graphs)."""
    s, g, bs = case
    om = [O.omega_f32(f) for f in s.freqs]
    G = _solver(s, g, om, precision, tol=tol, ncycle=200, graphs=1)
    E = _solver(s, g, om, precision, tol=tol, ncycle=200, graphs=0)
    for rep in range(3):
        if rep == 2:
            for k, w in enumerate(reversed(om)):
                G.set_omega(w, slot=k)
                E.set_omega(w, slot=k)
        print(f"precision {precision} solve {rep}", flush=True)
        Eg, Hg, sg = _batch(G, bs)
        Ee, He, se = _batch(E, bs)
        for k in range(len(bs)):
            assert sg[k].cycles == se[k].cycles and sg[k].status == se[k].status == 0
            np.testing.assert_array_equal(G.history(k), E.history(k))
            assert rel(Eg[k], Ee[k]) == 0.0
            assert rel(Hg[k], He[k]) == 0.0
