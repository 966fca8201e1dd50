"""Properties that fix the oracle's arithmetic independently of the reference:
what the mathematics of the finite-integration curl-curl discretisation
demands, checked on small non-uniform grids.  A dropped term, a wrong sign,
index or width in the operator, the transfers or the curl fails one of these.

  * the operator is complex symmetric: u^T A v = v^T A u
  * discrete curl annihilates discrete gradients: with omega = 0, A G phi = 0,
    and curl_h(G phi) = 0
  * discrete div of H = curl E vanishes on every cell
  * restriction is the exact transpose of prolongation
  * prolongation reproduces fields constant along the edge and bilinear across
  * one exact line solve zeroes the residual rows of that line
  * the multigrid solve converges to the dense direct solution (brute force)
  * grid helpers: find_index brackets, create_nugrid is geometric with the
    requested first cell and total length
"""
import numpy as np
import pytest

import pyoracle as O
from helpers import interior_mask, random_field

N = (6, 4, 8)


def grid(n, seed=0):
    rng = np.random.default_rng(seed)
    xs = []
    for m in n:
        w = rng.uniform(50.0, 200.0, m)
        xs.append(np.concatenate([[0.0], np.cumsum(w)]).astype(np.float32))
    return xs


def medium(n, seed=1, air_rows=0):
    rng = np.random.default_rng(seed)
    shape = (n[2], n[1], n[0])
    s = [np.exp(rng.uniform(-1.0, 1.0, shape)).astype(np.float32) for _ in range(3)]
    if air_rows:
        for a in s:
            a[:air_rows] = 1e-8
    return s


@pytest.fixture(scope="module")
def lev():
    x = grid(N)
    s = medium(N)
    return O.Level.level0(N, *x, *s), x, s


OMU = float(O.omega_f32(0.5)) * O.MU0


def nodal_gradient(L, phi):
    """Edge field E_a(p) = (phi(p + e_a) - phi(p)) / W_a(p_a) of a nodal phi
    that vanishes on the boundary nodes."""
    n = L.n
    E = np.zeros((3, n[2] + 1, n[1] + 1, n[0] + 1), np.complex128)
    W = [L.axis(a, 3) for a in range(3)]
    # phi indexed [k][j][i]
    E[0, :, :, :-1] = (phi[:, :, 1:] - phi[:, :, :-1]) / W[0][None, None, :n[0]]
    E[1, :, :-1, :] = (phi[:, 1:, :] - phi[:, :-1, :]) / W[1][None, :n[1], None]
    E[2, :-1, :, :] = (phi[1:, :, :] - phi[:-1, :, :]) / W[2][:n[2], None, None]
    return E.reshape(-1)


def random_phi(n, seed):
    rng = np.random.default_rng(seed)
    phi = rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1)) + 1j * rng.standard_normal((n[2] + 1, n[1] + 1, n[0] + 1))
    phi[0] = phi[-1] = 0
    phi[:, 0] = phi[:, -1] = 0
    phi[:, :, 0] = phi[:, :, -1] = 0
    return phi


def test_operator_complex_symmetric(lev):
    L, _, _ = lev
    u = random_field(N, 11)
    v = random_field(N, 12)
    a = np.sum(u * L.apply(OMU, v))
    b = np.sum(v * L.apply(OMU, u))
    assert abs(a - b) <= 1e-12 * abs(a)


def test_operator_not_hermitian(lev):
    """The mass term is imaginary: A differs from its conjugate transpose."""
    L, _, _ = lev
    u = random_field(N, 11)
    v = random_field(N, 12)
    a = np.sum(np.conj(u) * L.apply(OMU, v))
    b = np.sum(np.conj(L.apply(OMU, u)) * v)
    assert abs(a - b) > 1e-6 * abs(a)


def test_curl_curl_annihilates_gradients(lev):
    L, _, _ = lev
    mask = interior_mask(N).reshape(-1)
    E = nodal_gradient(L, random_phi(N, 3))
    # a gradient of a phi that vanishes on the boundary has no tangential boundary edges
    assert np.abs(E[~mask]).max() == 0.0
    AE = L.apply(0.0, E)
    scale = np.abs(L.apply(0.0, random_field(N, 4))).max()
    assert np.abs(AE).max() <= 1e-12 * scale
    # with omega a gradient sees only the mass term -i omega mu0 sbar: a
    # diagonal, purely imaginary, negative factor that does not depend on phi
    E2 = nodal_gradient(L, random_phi(N, 13))
    d1 = L.apply(OMU, E)
    d2 = L.apply(OMU, E2)
    sel = mask & (np.abs(E) > 1e-6 * np.abs(E).max()) & (np.abs(E2) > 1e-6 * np.abs(E2).max())
    q1, q2 = d1[sel] / E[sel], d2[sel] / E2[sel]
    assert np.abs(q1 - q2).max() <= 1e-9 * np.abs(q1).max()
    assert np.abs(q1.real).max() <= 1e-9 * np.abs(q1).max()
    assert (q1.imag < 0).all()


def test_curl_h_of_gradient_is_zero(lev):
    L, _, _ = lev
    E = nodal_gradient(L, random_phi(N, 5))
    H = L.curl_h(OMU, E)
    ref = np.abs(L.curl_h(OMU, random_field(N, 6))).max()
    assert np.abs(H).max() <= 1e-12 * ref


def test_div_of_h_vanishes(lev):
    L, _, _ = lev
    n = N
    H = L.curl_h(OMU, random_field(N, 7)).reshape(3, n[2] + 1, n[1] + 1, n[0] + 1)
    W = [L.axis(a, 3) for a in range(3)]
    k, j, i = np.meshgrid(np.arange(n[2] - 1), np.arange(n[1] - 1), np.arange(n[0] - 1), indexing="ij")
    div = ((H[0, k, j, i + 1] - H[0, k, j, i]) / W[0][i]
           + (H[1, k, j + 1, i] - H[1, k, j, i]) / W[1][j]
           + (H[2, k + 1, j, i] - H[2, k, j, i]) / W[2][k])
    terms = np.abs(H[0, k, j, i]) / W[0][i]
    assert np.abs(div).max() <= 1e-12 * terms.max()


@pytest.mark.parametrize("pattern", [0, 1, 2])
def test_restriction_is_transpose_of_prolongation(lev, pattern):
    L, _, _ = lev
    C = L.coarsen(pattern)
    r = random_field(N, 21)
    uc = random_field(C.n, 22)
    Ruc = np.sum(O.restrict(L, C, r) * uc)
    Pu = O.prolong(L, C, uc, np.zeros_like(r))
    rPu = np.sum(r * Pu)
    assert abs(Ruc - rPu) <= 1e-12 * abs(Ruc)


@pytest.mark.parametrize("pattern", [0, 1, 2])
def test_prolongation_reproduces_bilinear_fields(lev, pattern):
    """u_a = alpha + beta x_b + gamma x_c (+ delta x_b x_c) on the coarse grid,
    constant along the edge, is reproduced exactly on the fine edges."""
    L, _, _ = lev
    C = L.coarsen(pattern)
    cn, fn = C.n, L.n

    def field(lv, n):
        X = [lv.axis(a, 0) for a in range(3)]
        out = np.zeros((3, n[2] + 1, n[1] + 1, n[0] + 1), np.complex128)
        k, j, i = np.meshgrid(np.arange(n[2] + 1), np.arange(n[1] + 1), np.arange(n[0] + 1), indexing="ij")
        p = (i, j, k)
        for a in range(3):
            b, c = (a + 1) % 3, (a + 2) % 3
            xb, xc = X[b][p[b]], X[c][p[c]]
            out[a] = (1.0 + 2.0j) + 0.003 * xb - 0.002j * xc + 1e-6 * xb * xc
            out[a][p[a] >= n[a]] = 0
        return out.reshape(-1)

    uc = field(C, cn) * interior_mask(cn).reshape(-1)
    uf = O.prolong(L, C, uc, np.zeros(L.nvec, np.complex128))
    want = field(L, fn)
    m = interior_mask(fn).reshape(-1)
    # fine edges whose coarse neighbours across the edge are all interior
    X = [L.axis(a, 0) for a in range(3)]
    Xc = [C.axis(a, 0) for a in range(3)]
    k, j, i = np.meshgrid(np.arange(fn[2] + 1), np.arange(fn[1] + 1), np.arange(fn[0] + 1), indexing="ij")
    p = (i, j, k)
    inner = np.zeros((3,) + i.shape, bool)
    for a in range(3):
        b, c = (a + 1) % 3, (a + 2) % 3
        ok = np.ones(i.shape, bool)
        for q in (b, c):
            ok &= (X[q][p[q]] > Xc[q][1] - 1e-3) & (X[q][p[q]] < Xc[q][cn[q] - 1] + 1e-3)
        inner[a] = ok
    sel = m & inner.reshape(-1)
    assert sel.sum() > 10
    assert np.abs(uf[sel] - want[sel]).max() <= 1e-10 * np.abs(want[sel]).max()


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_line_solve_zeroes_its_rows(axis):
    """After the exact solve of line (axis; pb, pc) the residual vanishes on
    the line's along-edges and on the four perpendicular edges of each of its
    nodes (src/gmg.c:340-659: the banded LU is exact)."""
    x = grid(N, 3)
    s = medium(N, 4, air_rows=2)
    L = O.Level.level0(N, *x, *s)
    u = random_field(N, 31)
    f = random_field(N, 32)
    b, c = (axis + 1) % 3, (axis + 2) % 3
    pb, pc = 2, 1
    u1 = L.line_solve(OMU, axis, pb, pc, u, f)
    r0 = np.abs(L.residual(OMU, u, f)).max()
    r = L.residual(OMU, u1, f).reshape(3, N[2] + 1, N[1] + 1, N[0] + 1)
    rows = []
    for e in range(N[axis]):
        q = [0, 0, 0]
        q[axis], q[b], q[c] = e, pb, pc
        rows.append((axis, q))
    for pnode in range(1, N[axis]):
        for comp, db, dc in ((b, -1, 0), (b, 0, 0), (c, 0, -1), (c, 0, 0)):
            q = [0, 0, 0]
            q[axis], q[b], q[c] = pnode, pb + db, pc + dc
            rows.append((comp, q))
    vals = [abs(r[comp, q[2], q[1], q[0]]) for comp, q in rows]
    assert max(vals) <= 1e-9 * r0
    # other rows are changed but not zeroed
    assert np.abs(r).max() > 1e-3 * r0


def dense_operator(L, omu):
    m = interior_mask(L.n).reshape(-1)
    idx = np.flatnonzero(m)
    A = np.zeros((idx.size, idx.size), np.complex128)
    e = np.zeros(L.nvec, np.complex128)
    for col, k in enumerate(idx):
        e[k] = 1.0
        A[:, col] = L.apply(omu, e)[idx]
        e[k] = 0.0
    return A, idx


@pytest.mark.parametrize("cycleopt,isemi", [(1, 1), (2, 1), (3, 1), (1, 0), (2, 0), (3, 0)])
def test_solve_matches_dense_direct_solution(cycleopt, isemi):
    """Brute force: the multigrid solve to 1e-12 equals A^{-1} (D b) from a
    dense LU, D the dual-volume scaling of the right-hand side
    (src/gmg.c:1899-1910), for the V-, F- and W-cycles (src/gmg.c:1586-1668)
    with semi-coarsening and line smoothing and with full coarsening and the
    vertex-patch point smoother (src/gmg.c:116-336, 1301-1313)."""
    n = (4, 4, 6)
    x = grid(n, 5)
    s = medium(n, 6)
    L = O.Level.level0(n, *x, *s)
    freq = 0.5
    omu = O.omega_mu0(freq)
    A, idx = dense_operator(L, omu)
    rng = np.random.default_rng(9)
    bvec = np.zeros(L.nvec, np.complex128)
    bvec[idx] = rng.standard_normal(idx.size) + 1j * rng.standard_normal(idx.size)
    # dual-volume scaling: E_a rows carry W_a * D_b * D_c
    W = [L.axis(a, 3) for a in range(3)]
    D = [L.axis(a, 2) for a in range(3)]
    k, j, i = np.meshgrid(np.arange(n[2] + 1), np.arange(n[1] + 1), np.arange(n[0] + 1), indexing="ij")
    scale = np.stack([W[0][i] * D[1][j] * D[2][k], D[0][i] * W[1][j] * D[2][k], D[0][i] * D[1][j] * W[2][k]]).reshape(-1)
    xd = np.linalg.solve(A, (bvec * scale)[idx])
    E, H, hist = O.gmg_solve(n, *x, *s, freq, bvec, tol=1e-13, ncycle=3000, cycleopt=cycleopt,
                             isemicoarsen=isemi)
    assert hist[-1] < 1e-12 * hist[0], (len(hist), hist[-1] / hist[0])
    assert np.abs(E[idx] - xd).max() <= 1e-8 * np.abs(xd).max()
    assert np.abs(np.delete(E, idx)).max() == 0.0


def test_find_index_brackets():
    rng = np.random.default_rng(0)
    x = np.sort(rng.uniform(-100, 100, 50)).astype(np.float32)
    for v in rng.uniform(x[0], x[-1], 200).astype(np.float32):
        i = O.lib().emo_find_index(x.size, x, np.float32(v))
        assert x[i] <= v and (i == x.size - 1 or v < x[i + 1] or x[i] == v)


def test_create_nugrid_geometric():
    lib = O.lib()
    n, dx, length = 20, 50.0, 20000.0
    x = np.zeros(n + 1, np.float32)
    q = lib.emo_create_nugrid(n, length, dx, x)
    w = np.diff(x.astype(np.float64))
    assert abs(w[0] - dx) <= 1e-3 * dx
    assert abs(x[-1] - length) <= 1e-3 * length
    assert np.allclose(w[1:] / w[:-1], q, rtol=1e-3)
    # uniform when n*dx already covers the length
    x2 = np.zeros(n + 1, np.float32)
    lib.emo_create_nugrid(n, 500.0, dx, x2)
    assert np.allclose(np.diff(x2), dx, rtol=1e-5)


# ---- the initial-guess network's oracle (oracle/x0net_ref.py) -------------
def test_x0net_oracle_adjoints_and_brute_force():
    """Convolution, its input and weight gradients are mutual adjoints;
    pooling and upsampling are adjoints; conv_fwd equals a direct sum over
    taps and channels."""
    import x0net_ref as R
    rng = np.random.default_rng(0)
    x = rng.standard_normal((3, 4, 5, 6))
    w = rng.standard_normal((7, 27 * 6))
    g = rng.standard_normal((3, 4, 5, 7))
    a = np.sum(R.conv_fwd(x, w) * g)
    assert abs(a - np.sum(x * R.conv_dx(g, w, 6))) <= 1e-10 * abs(a)
    assert abs(a - np.sum(w * R.conv_dw(x, g))) <= 1e-10 * abs(a)
    p = rng.standard_normal((4, 4, 6, 3))
    q = rng.standard_normal((2, 2, 3, 3))
    assert abs(np.sum(R.pool2(p) * q) - np.sum(p * R.pool2_bwd(q))) <= 1e-12
    assert abs(np.sum(R.up2(q) * p) - np.sum(q * R.up2_bwd(p))) <= 1e-12
    # brute force at one output point, with the tap numbering 9 dz + 3 dy + dx
    y = R.conv_fwd(x, w)
    z0, y0, x0, co = 1, 0, 4, 3
    s = 0.0
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                zz, yy, xx = z0 + dz - 1, y0 + dy - 1, x0 + dx - 1
                if 0 <= zz < 3 and 0 <= yy < 4 and 0 <= xx < 5:
                    t = 9 * dz + 3 * dy + dx
                    s += np.dot(x[zz, yy, xx], w[co, t * 6:(t + 1) * 6])
    assert abs(s - y[z0, y0, x0, co]) <= 1e-12 * max(1.0, abs(s))


def test_x0net_oracle_bf16_rounding():
    import x0net_ref as R
    v = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 0.0, 1e-30], np.float32)
    b = R.to_bf16_bits(v)
    back = R.from_bf16_bits(b)
    # 1 + 2^-8 is half way between 1 and 1 + 2^-7: ties go to the even mantissa
    assert back[0] == 1.0 and back[1] == 1.0 and back[2] == 1.0078125
    assert np.all(np.abs(back - v) <= np.abs(v) * 2.0 ** -8 + 1e-38)
