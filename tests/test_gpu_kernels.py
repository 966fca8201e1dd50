"""Element-wise parity of the HIP kernels against the oracle on the same inputs.

The oracle is itself pinned bit-exactly to the compiled reference (see
test_oracle_golden.py), so these compare the GPU path with the reference's
own residual, line block Gauss-Seidel, transfer and curl kernels.
Tolerances: FP64 kernels 1e-12 relative (max-norm), FP32 kernels 2e-5.
"""
import numpy as np
import pytest

import pyoracle as O
from helpers import interior_mask, random_field, rel
from libemmi_mgfd_amd import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tiny():
    s = configs.tiny()
    g = O.regrid(s)
    n = (s.n1, s.n2, s.n3)
    L0 = O.Level.level0(n, g["x1"], g["x2"], g["x3"], g["sigma11"], g["sigma22"], g["sigma33"])
    return s, g, n, L0


@pytest.fixture(params=[1, 2], ids=["one_lane", "quad_lane"])
def line_kernel(request):
    """Runs a test with each line-substitution kernel (one lane or four lanes
    per line half), whatever the line count would select."""
    from libemmi_mgfd_amd import native
    L = native.lib()
    prev = L.emmi_set_line_kernel(request.param)
    assert prev >= 0
    yield request.param
    L.emmi_set_line_kernel(prev)


def make_solver(s, g, precision, freq_index=0, **kw):
    from libemmi_mgfd_amd import native
    return native.Solver((s.n1, s.n2, s.n3), g["x1"], g["x2"], g["x3"], g["sigma11"], g["sigma22"], g["sigma33"],
                         O.omega_f32(s.freqs[freq_index]), precision=precision, **kw)


@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 2e-5)])
def test_residual(tiny, precision, tol):
    s, g, n, L0 = tiny
    S = make_solver(s, g, precision)
    omu = O.omega_mu0(s.freqs[0])
    u = random_field(n, 1)
    f = random_field(n, 2, interior_only=False)
    ref = L0.residual(omu, u, f)
    out = S.check_op(0, 0, 0, 0, 0, u, f)
    assert rel(out, ref) < tol


@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 2e-5)])
def test_residual_large_level(precision, tol):
    """A level of more than 2^20 points takes the residual kernel's large-level
    form (five planes per thread with nontemporal streams in FP64, six in
    FP32); checked against the oracle on a stretched 128 x 128 x 64 grid with
    an air layer, a sea layer and a layered, anisotropic earth."""
    from libemmi_mgfd_amd import native
    n = (128, 128, 64)
    rng = np.random.default_rng(5)
    xs = [np.cumsum(np.r_[0.0, rng.uniform(20.0, 80.0, m)]).astype(np.float32) for m in n]
    s11 = np.empty((n[2], n[1], n[0]), np.float32)
    s11[:8] = 1e-8                      # air
    s11[8:20] = 3.3                     # sea
    s11[20:] = rng.uniform(0.01, 2.0, (n[2] - 20, 1, 1)).astype(np.float32)
    s33 = (s11 * np.float32(0.5)).astype(np.float32)
    L0 = O.Level.level0(n, *xs, s11, s11, s33)
    freq = 0.75
    S = native.Solver(n, *xs, s11, s11, s33, O.omega_f32(freq), precision=precision)
    u = random_field(n, 3)
    f = random_field(n, 4, interior_only=False)
    ref = L0.residual(O.omega_mu0(freq), u, f)
    out = S.check_op(0, 0, 0, 0, 0, u, f)
    err = rel(out, ref)
    print(f"precision {precision}: residual of a {L0.nvec // 3}-point level, max rel diff {err:.2e}", flush=True)
    assert err < tol
    S.close()


def conductive_level(s, g, n):
    """Same grid, uniform 1 S/m medium: well-conditioned line blocks."""
    one = np.ones((n[2], n[1], n[0]), np.float32)
    return one, O.Level.level0(n, g["x1"], g["x2"], g["x3"], one, one, one)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("it", [0, 1])
def test_line_gs_fp64_conductive(tiny, axis, it, line_kernel):
    """Exact line solves agree to rounding where the blocks are well conditioned."""
    from libemmi_mgfd_amd import native
    s, g, n, _ = tiny
    one, L = conductive_level(s, g, n)
    S = native.Solver(n, g["x1"], g["x2"], g["x3"], one, one, one, O.omega_f32(s.freqs[0]), precision=0)
    omu = O.omega_mu0(s.freqs[0])
    u = random_field(n, 3)
    f = random_field(n, 4)
    ref = L.line_gs(omu, axis, it, u, f, colour=1)
    out = S.check_op(1, 0, 0, axis, it, u, f)
    assert rel(out, ref) < 1e-11


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_line_gs_fp64_mild_air(axis, line_kernel):
    """With a mildly resistive 'air' (100 ohm-m) the blocks are well
    conditioned and the exact solves agree to rounding."""
    s = configs.tiny(rho_air=1e2)
    g = O.regrid(s)
    n = (s.n1, s.n2, s.n3)
    L = O.Level.level0(n, g["x1"], g["x2"], g["x3"], g["sigma11"], g["sigma22"], g["sigma33"])
    S = make_solver(s, g, 0)
    omu = O.omega_mu0(s.freqs[0])
    u = random_field(n, 11)
    f = random_field(n, 12)
    ref = L.line_gs(omu, axis, 0, u, f, colour=1)
    out = S.check_op(1, 0, 0, axis, 0, u, f)
    print("mild air rel", rel(out, ref))
    assert rel(out, ref) < 1e-10


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("it", [0, 1])
def test_line_gs_fp64(tiny, axis, it, line_kernel):
    """With 1e8 ohm-m air the line blocks are near-singular along discrete
    gradients (omega mu0 sigma_air ~ 1e-10 of the curl-curl scale): a random
    right-hand side drives the solution to ~1e6 and two exact solvers differ
    by cond*eps ~ 1e-5.  Assert that bound and that the GPU sweep leaves a
    residual no larger than the reference ordering's own."""
    s, g, n, L0 = tiny
    S = make_solver(s, g, 0)
    omu = O.omega_mu0(s.freqs[0])
    u = random_field(n, 3)
    f = random_field(n, 4)
    ref = L0.line_gs(omu, axis, it, u, f, colour=1)
    out = S.check_op(1, 0, 0, axis, it, u, f)
    r_ref = np.abs(L0.residual(omu, ref, f))
    r_out = np.abs(L0.residual(omu, out, f))
    print("air rel", rel(out, ref), "res", r_out.max(), r_ref.max())
    assert rel(out, ref) < 1e-4
    assert r_out.max() < 1.01 * r_ref.max()


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_line_gs_fp32(axis, line_kernel):
    """FP32 storage with FP64 arithmetic on better-conditioned blocks (100
    ohm-m air): agreement to the FP32 rounding of the stored vectors times
    the remaining block conditioning."""
    s = configs.tiny(rho_air=1e2)
    g = O.regrid(s)
    n = (s.n1, s.n2, s.n3)
    L = O.Level.level0(n, g["x1"], g["x2"], g["x3"], g["sigma11"], g["sigma22"], g["sigma33"])
    S = make_solver(s, g, 1)
    omu = O.omega_mu0(s.freqs[0])
    u = random_field(n, 5)
    f = random_field(n, 6)
    ref = L.line_gs(omu, axis, 0, u, f, colour=1)
    out = S.check_op(1, 0, 0, axis, 0, u, f)
    print("fp32 rel", rel(out, ref))
    assert rel(out, ref) < 1e-4


def test_line_gs_fp32_air_finite(tiny, line_kernel):
    """In 1e8 ohm-m air the FP32-stored blocks are singular to working
    precision; the FP64 arithmetic must still return finite values."""
    s, g, n, L0 = tiny
    S = make_solver(s, g, 1)
    for axis in range(3):
        out = S.check_op(1, 0, 0, axis, 0, random_field(n, 5), random_field(n, 6))
        assert np.all(np.isfinite(out))


@pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 1e-6)])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_line_kernels_agree(tiny, axis, precision, tol):
    """The one-lane and quad-lane substitutions solve the same line systems
    exactly: on well-conditioned blocks they agree to rounding (FP64) or to
    the FP32 rounding of the stored records and vectors."""
    from libemmi_mgfd_amd import native
    s, g, n, _ = tiny
    one, _ = conductive_level(s, g, n)
    L = native.lib()
    outs = []
    prev = L.emmi_set_line_kernel(1)
    try:
        for mode in (1, 2):
            L.emmi_set_line_kernel(mode)
            S = native.Solver(n, g["x1"], g["x2"], g["x3"], one, one, one, O.omega_f32(s.freqs[0]),
                              precision=precision)
            outs.append(S.check_op(1, 0, 0, axis, 1, random_field(n, 21), random_field(n, 22)))
    finally:
        L.emmi_set_line_kernel(prev)
    print("one-lane vs quad-lane rel", rel(outs[1], outs[0]))
    assert rel(outs[1], outs[0]) < tol


@pytest.mark.parametrize("n", [(128, 128, 64), (64, 96, 48)], ids=["128x128x64", "64x96x48"])
@pytest.mark.parametrize("precision", [0, 1])
def test_line_rhs_march_matches_node_records(precision, n):
    """The fused quad-lane launch builds its lines' right-hand sides by
    marching along them (values shared by neighbouring nodes carried in
    registers); the separate right-hand-side launch builds each node on its
    own.  Both evaluate every operator row with the same arithmetic, so a
    sweep gives the same result to the last bit.  Lines of 128, 96, 64 and 48
    nodes (records in LDS and in memory; the march on the 64- and 128-node
    lines along y and z), with an air layer and a layered earth."""
    from libemmi_mgfd_amd import native
    rng = np.random.default_rng(11)
    xs = [np.cumsum(np.r_[0.0, rng.uniform(20.0, 80.0, m)]).astype(np.float32) for m in n]
    s11 = np.empty((n[2], n[1], n[0]), np.float32)
    s11[:8] = 1e-8
    s11[8:] = rng.uniform(0.01, 2.0, (n[2] - 8, 1, 1)).astype(np.float32)
    L = native.lib()
    u, f = random_field(n, 31), random_field(n, 32)
    S = native.Solver(n, *xs, s11, s11, s11, O.omega_f32(0.5), precision=precision)
    prev = L.emmi_set_line_kernel(3)
    try:
        for axis in range(3):
            print(f"line sweep {n} precision {precision} axis {axis}: fused vs separate records", flush=True)
            outs = []
            for mode in (3, 4):
                L.emmi_set_line_kernel(mode)
                outs.append(S.check_op(1, 0, 0, axis, 0, u, f))
            assert np.all(np.isfinite(outs[0]))
            assert np.array_equal(outs[0], outs[1]), f"axis {axis}: max rel diff {rel(outs[0], outs[1]):.2e}"
    finally:
        L.emmi_set_line_kernel(prev)
        S.close()


@pytest.mark.parametrize("pattern", [0, 1, 2])
def test_transfer_fp64(tiny, pattern):
    s, g, n, L0 = tiny
    S = make_solver(s, g, 0)
    L1 = L0.coarsen(pattern)
    assert S.dims(pattern, 1) == L1.n
    r = random_field(n, 7)
    fc_ref = O.restrict(L0, L1, r)
    fc = S.check_op(2, pattern, 0, 0, 0, r, out_len=L1.nvec)
    mask = __import__("helpers").interior_mask(L1.n).reshape(-1)
    # coarse boundary entries are never read by the solver: compare interior
    assert rel(fc[mask], fc_ref[mask]) < 1e-13
    assert np.all(fc[~mask] == 0)
    uc = random_field(L1.n, 8)
    uf = random_field(n, 9)
    uf_ref = O.prolong(L0, L1, uc, uf)
    out = S.check_op(3, pattern, 0, 0, 0, uc, uf)
    assert rel(out, uf_ref) < 1e-13


def test_curl_fp64(tiny):
    s, g, n, L0 = tiny
    S = make_solver(s, g, 0)
    omu = O.omega_mu0(s.freqs[0])
    u = random_field(n, 10)
    assert rel(S.check_op(4, 0, 0, 0, 0, u), L0.curl_h(omu, u)) < 1e-13


@pytest.mark.parametrize("precision,tol", [(0, 1e-11), (1, 1e-11)])
def test_solve_matches_oracle(tiny, precision, tol, line_kernel):
    """Converged fields agree: both solvers run to relative residual 1e-11
    (FP64 cycles, and FP32-storage cycles inside FP64 refinement), H
    everywhere and E on the edges at and below the sea surface (in the air
    the gradient part of E is barely constrained and converges last)."""
    s, g, n, L0 = tiny
    rw = float(min(s.rho11.min(), s.rho22.min(), s.rho33.min()))
    mg = O.mg_struct(g, rw)
    for fi, fr in enumerate(s.freqs):
        b = O.inject_source(mg, s.sources, s.chsrc, fr)
        E_ref, H_ref, hist = O.gmg_solve(n, g["x1"], g["x2"], g["x3"], g["sigma11"], g["sigma22"], g["sigma33"],
                                         fr, b, tol=1e-11, ncycle=200)
        S = make_solver(s, g, precision, fi, tol=tol, ncycle=200)
        E, H, st = S.solve(b)
        h = S.history()
        print("precision", precision, "status", st.status, "cycles", st.cycles, "oracle cycles", len(hist))
        print("  gpu hist", (h[:12] / h[0]).tolist())
        print("  ref hist", (hist[:12] / hist[0]).tolist())
        print("  gpu tail", (h[-6:] / h[0]).tolist())
        print("  E rel", rel(E, E_ref), "H rel", rel(H, H_ref))
        assert st.status == 0, (st.status, st.cycles)
        assert rel(H, H_ref) < 100 * tol
        x3 = np.asarray(g["x3"], np.float64)
        zc = 0.5 * (x3[:-1] + x3[1:])
        m = np.zeros((3, n[2] + 1, n[1] + 1, n[0] + 1), bool)
        m[0] = m[1] = (x3 >= 0)[:, None, None]
        m[2, :-1] = (zc >= 0)[:, None, None]
        m = m.reshape(-1)
        e_sea = float(np.abs(E - E_ref)[m].max() / np.abs(E_ref)[m].max())
        print("  E rel below the sea surface", e_sea)
        assert e_sea < 1e4 * tol


@pytest.mark.parametrize("it", [0, 1])
def test_point_gs_fp64_conductive(tiny, it):
    """Vertex-patch Gauss-Seidel (src/gmg.c:116-336) in 8 node colours
    against the oracle's sweep in the same colour order: the 6x6 patch
    solves agree to rounding where they are well conditioned."""
    from libemmi_mgfd_amd import native
    s, g, n, _ = tiny
    one, L = conductive_level(s, g, n)
    S = native.Solver(n, g["x1"], g["x2"], g["x3"], one, one, one, O.omega_f32(s.freqs[0]), precision=0)
    omu = O.omega_mu0(s.freqs[0])
    u = random_field(n, 31)
    f = random_field(n, 32)
    ref = L.point_gs(omu, it, u, f, colour=1)
    out = S.check_op(6, 0, 0, 0, it, u, f)
    print("point GS conductive rel", rel(out, ref))
    assert rel(out, ref) < 1e-12
    # the sweep changed every interior edge
    assert np.all(out[interior_mask(n).reshape(-1)] != u[interior_mask(n).reshape(-1)])


@pytest.mark.parametrize("it", [0, 1])
def test_point_gs_fp64_air(tiny, it):
    """With 1e8 ohm-m air the patch blocks are near-singular along the
    node's discrete gradient; as for the line solves, the two exact patch
    solvers differ by about cond*eps and the GPU sweep's residual is no
    larger than the oracle's."""
    s, g, n, L0 = tiny
    S = make_solver(s, g, 0)
    omu = O.omega_mu0(s.freqs[0])
    u = random_field(n, 33)
    f = random_field(n, 34)
    ref = L0.point_gs(omu, it, u, f, colour=1)
    out = S.check_op(6, 0, 0, 0, it, u, f)
    r_ref = np.abs(L0.residual(omu, ref, f))
    r_out = np.abs(L0.residual(omu, out, f))
    print("point GS air rel", rel(out, ref), "res", r_out.max(), r_ref.max())
    assert rel(out, ref) < 1e-4
    assert r_out.max() < 1.01 * r_ref.max()


def test_point_gs_fp32_mild_air():
    s = configs.tiny(rho_air=1e2)
    g = O.regrid(s)
    n = (s.n1, s.n2, s.n3)
    L = O.Level.level0(n, g["x1"], g["x2"], g["x3"], g["sigma11"], g["sigma22"], g["sigma33"])
    S = make_solver(s, g, 1)
    omu = O.omega_mu0(s.freqs[0])
    u = random_field(n, 35)
    f = random_field(n, 36)
    ref = L.point_gs(omu, 0, u, f, colour=1)
    out = S.check_op(6, 0, 0, 0, 0, u, f)
    print("point GS fp32 rel", rel(out, ref))
    assert rel(out, ref) < 1e-5
