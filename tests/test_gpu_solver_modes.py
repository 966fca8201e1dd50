"""The GPU solve in every mode the reference offers, and its failure contract.

* The V- and W-cycles (cycleopt=1, 3), src/gmg.c:1586-1668: each converges
  to the field of the oracle's converged solve (the oracle is bit-exact with
  the reference in all these modes, test_oracle_golden.py).  The smoothers
  order their updates differently, so the iterates differ; the converged
  field does not.
* Full coarsening with the vertex-patch point smoother (isemicoarsen=0,
  src/gmg.c:116-336, 1301-1313), V, F and W: the same iterates as the
  oracle run in the GPU's colour order.
* The fused residual norm (src/gmg.c:1576-1583): the norms the GPU reports
  equal ||f - A u|| that the oracle computes for the GPU's own iterate, and
  ||f|| for the zero guess.
* rnorm0 of the FP64 V- and W-cycles is the first in-cycle norm, after the
  pre-smoothing of cycle 0 (src/gmg.c:1594-1599, 1651-1656).
* Failure contract (no counterpart in the reference, which is silent): an
  unreachable tolerance returns NOT_CONVERGED with the last iterate; a NaN in
  the conductivity or the right-hand side returns NONFINITE; in a batch the
  other slots converge exactly as alone.
* Initial guesses: a guess with a larger residual than the zero guess is
  dropped before the first cycle (the solve is then bit-identical to the
  zero-guess solve); a good guess is kept.
"""
import numpy as np
import pytest

import pyoracle as O
from helpers import interior_mask
from libemmi_mgfd_amd import configs, native

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tiny():
    s = configs.tiny()
    g = O.regrid(s)
    n = (s.n1, s.n2, s.n3)
    L0 = O.Level.level0(n, g["x1"], g["x2"], g["x3"], g["sigma11"], g["sigma22"], g["sigma33"])
    rw = float(min(s.rho11.min(), s.rho22.min(), s.rho33.min()))
    mg = O.mg_struct(g, rw)
    b = O.inject_source(mg, s.sources, s.chsrc, s.freqs[0])
    E, H, hist = O.gmg_solve(n, g["x1"], g["x2"], g["x3"], g["sigma11"], g["sigma22"], g["sigma33"],
                             s.freqs[0], b, tol=1e-12, ncycle=300)
    return s, g, n, L0, b, E, H


def solver(s, g, omega=None, sig=None, **kw):
    n = (s.n1, s.n2, s.n3)
    sig = sig or (g["sigma11"], g["sigma22"], g["sigma33"])
    om = O.omega_f32(s.freqs[0]) if omega is None else omega
    return native.Solver(n, g["x1"], g["x2"], g["x3"], *sig, om, **kw)


def dual_scaled(L, n, b):
    """f = b times the dual volume of each edge (src/gmg.c:1899-1910)."""
    W = [L.axis(a, 3) for a in range(3)]
    D = [L.axis(a, 2) for a in range(3)]
    k, j, i = np.meshgrid(np.arange(n[2] + 1), np.arange(n[1] + 1), np.arange(n[0] + 1), indexing="ij")
    sc = np.stack([W[0][i] * D[1][j] * D[2][k], D[0][i] * W[1][j] * D[2][k], D[0][i] * D[1][j] * W[2][k]])
    return b * sc.reshape(-1)


def below_air(g, n):
    """Edges at or below z = 0 (the air layer is z < 0): there E is fixed by
    the conductive earth; in the air its gradient part is barely constrained
    and converges last."""
    x3 = np.asarray(g["x3"], np.float64)
    zc = 0.5 * (x3[:-1] + x3[1:])
    m = np.zeros((3, n[2] + 1, n[1] + 1, n[0] + 1), bool)
    m[0] = (x3 >= 0)[:, None, None]
    m[1] = (x3 >= 0)[:, None, None]
    m[2, :-1] = (zc >= 0)[:, None, None]
    return (m & interior_mask(n)).reshape(-1)


def relmax(a, b, m=None):
    if m is not None:
        a, b = a[m], b[m]
    return float(np.abs(a - b).max() / np.abs(b).max())


MODES = {"v": dict(cycleopt=1), "w": dict(cycleopt=3), "f": dict(cycleopt=2)}


@pytest.mark.parametrize("mode", list(MODES))
def test_cycle_modes_converge_to_oracle_field(tiny, mode):
    s, g, n, L0, b, E_ref, H_ref = tiny
    S = solver(s, g, precision=0, tol=1e-11, ncycle=400, **MODES[mode])
    print(f"tiny survey {n}, mode {mode} {MODES[mode]}", flush=True)
    E, H, st = S.solve(b)
    h = S.history()
    print(f"  status {st.status} cycles {st.cycles} rel-res {st.rnorm / st.rnorm0:.2e}", flush=True)
    assert st.status == 0, (st.status, st.cycles)
    m = below_air(g, n)
    eE, eH = relmax(E, E_ref, m), relmax(H, H_ref)
    print(f"  E rel (z >= 0) {eE:.2e}, H rel {eH:.2e}", flush=True)
    assert eE < 1e-8 and eH < 1e-8
    # the reported norms are the true residual norms of the iterates
    f = dual_scaled(L0, n, b)
    omu = O.omega_mu0(s.freqs[0])
    r = np.linalg.norm(L0.residual(omu, E, f))
    assert abs(h[-1] - r) <= 1e-6 * r, (h[-1], r)
    assert abs(st.rnorm - r) <= 1e-6 * r
    if MODES[mode]["cycleopt"] == 2:
        assert abs(st.rnorm0 - np.linalg.norm(f)) <= 1e-12 * np.linalg.norm(f)
        assert h[0] == st.rnorm0
    else:
        # rnorm0 is the first in-cycle norm, taken after one pre-smoothing
        assert h[0] == st.rnorm0 and st.rnorm0 < np.linalg.norm(f)


@pytest.mark.parametrize("cycleopt", [1, 2, 3])
def test_full_coarsening_point_smoother_follows_oracle(tiny, cycleopt):
    """isemicoarsen=0: full coarsening and the vertex-patch point smoother on
    every level, as the reference runs it (src/gmg.c:1301-1313).  The GPU
    sweeps its nodes in 8 colours, so the oracle runs the same algorithm in
    the same colour order (colour_order=1): the residual histories of 30
    cycles agree to the rounding that the near-singular air blocks amplify.
    The smoother is slow on these stretched grids (the reference's own rate,
    about 0.9 per cycle here)."""
    s, g, n, L0, b, _, _ = tiny
    kw = dict(cycleopt=cycleopt, isemicoarsen=0)
    S = solver(s, g, precision=0, tol=1e-30, ncycle=30, **kw)
    print(f"tiny survey {n}, full coarsening, cycleopt {cycleopt}", flush=True)
    E, H, st = S.solve(b)
    h = S.history()
    _, _, ho = O.gmg_solve(n, g["x1"], g["x2"], g["x3"], g["sigma11"], g["sigma22"], g["sigma33"], s.freqs[0], b,
                           tol=1e-30, ncycle=30, colour_order=1, **kw)
    print("  gpu   ", (h[:6] / h[0]).tolist(), (h[-1] / h[0]), flush=True)
    print("  oracle", (ho[:6] / ho[0]).tolist(), (ho[-1] / ho[0]), flush=True)
    assert st.status == 1 and st.cycles == 30 and np.all(np.isfinite(E))
    m = min(len(h), len(ho))
    assert m >= 29
    np.testing.assert_allclose(h[:m], ho[:m], rtol=1e-9)
    assert h[m - 1] < 0.5 * h[0]


@pytest.mark.parametrize("precision", [0, 1])
def test_fused_norm_equals_oracle_norm(tiny, precision):
    """The F-cycle tests the iterate at the start of each cycle: a solve
    limited to 4 cycles returns u_4, and a solve limited to 5 reports
    ||f - A u_4|| as its fifth norm.  That norm equals the oracle's FP64
    residual norm of the returned u_4; the first norm is ||f||."""
    s, g, n, L0, b, _, _ = tiny
    S4 = solver(s, g, precision=precision, tol=1e-30, ncycle=4)
    E4, _, st4 = S4.solve(b)
    h4 = S4.history()
    S5 = solver(s, g, precision=precision, tol=1e-30, ncycle=5)
    S5.solve(b)
    h5 = S5.history()
    f = dual_scaled(L0, n, b)
    omu = O.omega_mu0(s.freqs[0])
    assert abs(h5[0] - np.linalg.norm(f)) <= 1e-12 * np.linalg.norm(f)
    np.testing.assert_array_equal(h5[:4], h4)
    r = np.linalg.norm(L0.residual(omu, E4, f))
    print(f"precision {precision}: history {h5.tolist()}, oracle norm of u_4 {r:.6e}", flush=True)
    assert abs(h5[4] - r) <= 1e-9 * r
    # the three semi-coarsening patterns take turns: the norm falls over the
    # cycles, not necessarily in every one
    assert h5[4] < 1e-2 * h5[0]
    assert st4.status == 1 and st4.cycles == 4


def test_unreachable_tolerance_returns_not_converged(tiny):
    s, g, n, L0, b, _, _ = tiny
    S = solver(s, g, precision=0, tol=1e-30, ncycle=6)
    E, H, st = S.solve(b)
    print(f"tol 1e-30: status {st.status}, cycles {st.cycles}", flush=True)
    assert st.status == 1 and st.cycles == 6
    assert np.all(np.isfinite(E)) and np.abs(E).max() > 0 and np.all(np.isfinite(H))
    # the returned field is the iterate after the last cycle: a solve with
    # one cycle more tests it first
    S7 = solver(s, g, precision=0, tol=1e-30, ncycle=7)
    S7.solve(b)
    f = dual_scaled(L0, n, b)
    r = np.linalg.norm(L0.residual(O.omega_mu0(s.freqs[0]), E, f))
    assert abs(S7.history()[6] - r) <= 1e-9 * r and st.rnorm == S.history()[-1]


@pytest.mark.parametrize("precision", [0, 1])
def test_nan_conductivity_returns_nonfinite(tiny, precision):
    s, g, n, L0, b, _, _ = tiny
    s11 = np.array(g["sigma11"], np.float32, copy=True)
    s11.reshape(-1)[s11.size // 2] = np.nan
    S = solver(s, g, sig=(s11, g["sigma22"], g["sigma33"]), precision=precision, tol=1e-6, ncycle=10)
    _, _, st = S.solve(b)
    print(f"NaN conductivity, precision {precision}: status {st.status}", flush=True)
    assert st.status == 2


def test_nan_rhs_returns_nonfinite_and_batch_neighbours_unaffected(tiny):
    s, g, n, L0, b, _, _ = tiny
    om = [O.omega_f32(f) for f in s.freqs]
    mg = O.mg_struct(g, float(min(s.rho11.min(), s.rho22.min(), s.rho33.min())))
    b1 = O.inject_source(mg, s.sources, s.chsrc, s.freqs[1])
    bad = b1.copy()
    bad[len(bad) // 3 + 7] = np.nan
    nb = 16 * b.size
    B = solver(s, g, omega=om, precision=0, tol=1e-8, ncycle=60)
    bufs = [native.DeviceBuffer(nb).upload(v) for v in (b, bad)]
    xs = [native.DeviceBuffer(nb) for _ in range(2)]
    sts = B.solve_batch_device([q.ptr for q in bufs], [q.ptr for q in xs], [q.ptr for q in bufs])
    print(f"batch with a NaN right-hand side: statuses {[t.status for t in sts]}", flush=True)
    assert sts[0].status == 0 and sts[1].status == 2
    # slot 0 alone
    A = solver(s, g, omega=om[0], precision=0, tol=1e-8, ncycle=60)
    E0, H0, st0 = A.solve(b)
    assert st0.cycles == sts[0].cycles
    np.testing.assert_array_equal(xs[0].download(np.complex128), E0)
    np.testing.assert_array_equal(bufs[0].download(np.complex128), H0)


@pytest.mark.parametrize("precision", [0, 1])
def test_bad_initial_guess_is_rejected(tiny, precision):
    s, g, n, L0, b, E_ref, _ = tiny
    S0 = solver(s, g, precision=precision, tol=1e-7, ncycle=80)
    E0, H0, st0 = S0.solve(b)
    rng = np.random.default_rng(5)
    scale = np.abs(E_ref).max()
    bad = scale * (rng.standard_normal(E_ref.size) + 1j * rng.standard_normal(E_ref.size))
    bad = np.where(interior_mask(n).reshape(-1), bad, 0)
    S1 = solver(s, g, precision=precision, tol=1e-7, ncycle=80)
    E1, H1, st1 = S1.solve(b, x0=bad)
    print(f"bad guess, precision {precision}: rejected {st1.x0_rejected}, guess residual "
          f"{st1.x0_rnorm:.3e} vs ||f|| {st1.rnorm0:.3e}, cycles {st1.cycles} (zero guess {st0.cycles})", flush=True)
    assert st1.x0_rejected == 1 and st1.x0_rnorm > st1.rnorm0
    assert st1.cycles == st0.cycles and st1.status == 0
    np.testing.assert_array_equal(E1, E0)
    np.testing.assert_array_equal(H1, H0)


@pytest.mark.parametrize("precision", [0, 1])
def test_good_initial_guess_is_kept(tiny, precision):
    s, g, n, L0, b, E_ref, _ = tiny
    S0 = solver(s, g, precision=precision, tol=1e-8, ncycle=80)
    _, _, st0 = S0.solve(b)
    S1 = solver(s, g, precision=precision, tol=1e-8, ncycle=80)
    E1, _, st1 = S1.solve(b, x0=E_ref)
    print(f"good guess, precision {precision}: rejected {st1.x0_rejected}, guess residual "
          f"{st1.x0_rnorm / st1.rnorm0:.2e} ||f||, cycles {st1.cycles} (zero guess {st0.cycles})", flush=True)
    assert st1.x0_rejected == 0 and st1.x0_rnorm < 1e-6 * st1.rnorm0
    assert st1.status == 0 and st1.cycles <= 1 < st0.cycles
    assert relmax(E1, E_ref, below_air(g, n)) < 1e-7


@pytest.mark.parametrize("precision", [0, 1])
def test_continuation_from_first_cycle_iterate(tiny, precision):
    """A solve from the iterate after one cycle, started at cycle index 1
    (Solver.set_first_cycle: the pattern sequence 1, 2, 0, ... continues),
    reproduces the zero-guess solve: one cycle fewer and the same fields and
    residual history from cycle 1 on (to rounding: the continued solve
    recomputes the residual norm of its guess)."""
    s, g, n, L0, b, E_ref, _ = tiny
    S = solver(s, g, precision=precision, tol=1e-8, ncycle=80)
    E0, H0, st0 = S.solve(b)
    h0 = S.history()
    S.set_ncycle(1)
    x1, _, _ = S.solve(b)
    S.set_ncycle(80)
    S.set_first_cycle(1)
    E1, H1, st1 = S.solve(b, x0=x1)
    h1 = S.history()
    S.set_first_cycle(0)
    print(f"precision {precision}: zero guess {st0.cycles} cycles, continued 1 + {st1.cycles}; "
          f"E max rel diff {relmax(E1, E0):.2e}", flush=True)
    assert st0.status == 0 and st1.status == 0 and st1.x0_rejected == 0
    assert st1.cycles + 1 == st0.cycles
    np.testing.assert_allclose(h1, h0[1:len(h1) + 1], rtol=1e-10)
    assert relmax(E1, E0) < 1e-10 and relmax(H1, H0) < 1e-10


def test_blend_guess_never_raises_the_residual(tiny):
    """emmi_solver_blend_guess: x1 + beta (xg - x1) with the residual-minimising
    beta.  With xg the converged field, beta is 1 to rounding and the blend's
    residual is the converged one; with xg = x1 plus noise larger than the
    field, the blend's residual stays at or below x1's; in both cases the
    solver accepts the blend and a solve from it converges."""
    s, g, n, L0, b, E_ref, _ = tiny
    S = solver(s, g, precision=1, tol=1e-8, ncycle=80)
    nb = 16 * S.nvec
    db = native.DeviceBuffer(nb).upload(b)
    dx1 = native.DeviceBuffer(nb)
    dh = native.DeviceBuffer(nb)
    S.set_ncycle(1)
    S.solve_device(db.ptr, dx1.ptr, dh.ptr)
    S.set_ncycle(80)
    x1 = dx1.download(np.complex128)
    dout = native.DeviceBuffer(nb)
    # residual of x1 itself, as the solver reports it for a guess
    st_x1 = S.solve_device(db.ptr, dout.ptr, dh.ptr, dx1.ptr)
    rng = np.random.default_rng(3)
    noise = np.abs(E_ref).max() * (rng.standard_normal(E_ref.size) + 1j * rng.standard_normal(E_ref.size))
    noise = np.where(interior_mask(n).reshape(-1), noise, 0)
    for name, xg in (("converged field", E_ref), ("x1 plus noise", x1 + noise)):
        dg = native.DeviceBuffer(nb).upload(np.ascontiguousarray(xg, np.complex128))
        beta, = S.blend_guess([db.ptr], [dx1.ptr], [dg.ptr])
        S.set_first_cycle(1)
        st = S.solve_device(db.ptr, dout.ptr, dh.ptr, dg.ptr)
        S.set_first_cycle(0)
        print(f"{name}: beta {beta:.3e}, blend residual {st.x0_rnorm / st.rnorm0:.3e} ||f|| "
              f"(x1: {st_x1.x0_rnorm / st_x1.rnorm0:.3e}), cycles {st.cycles}", flush=True)
        assert st.x0_rejected == 0 and st.status == 0
        assert st.x0_rnorm <= st_x1.x0_rnorm * (1 + 1e-12)
        if name == "converged field":
            assert abs(beta - 1) < 1e-6 and st.x0_rnorm < 1e-6 * st.rnorm0 and st.cycles <= 1
