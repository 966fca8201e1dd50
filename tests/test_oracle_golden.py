"""The oracle (oracle/liboracle.so, the CPU restatement of the reference's
multigrid path) against the golden fixtures that the compiled reference wrote
(tests/golden/make_golden.py).  Every comparison here is bit-exact: the
restatement keeps the reference's float/double evaluation order.

Fixtures: tiny_units.npz (single kernels on an 8x6x14 grid), tiny_model.npz
(12x10x16 two-frequency survey end to end, with its E and H volumes on the
input grid), tiny_cycles.npz (the same survey with V- and W-cycles and with
point smoothing on full coarsening), tiny_emf_0001.txt, tiny_grad.npz (mode-2
gradient), config1.npz (64^3 survey: grid, medium, receiver data, the y = 0
planes of its E and H volumes).
"""
import os

import numpy as np
import pytest

import pyoracle as O
from libemmi_mgfd_amd import configs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GRID_KEYS = ["xx", "yy", "zz", "x1", "x2", "x3", "x1s", "x2s", "x3s", "d1", "d2", "d3", "d1s", "d2s", "d3s"]


def load(name):
    return np.load(os.path.join(GOLD, name))


@pytest.fixture(scope="module")
def units():
    d = load("tiny_units.npz")
    s = configs.tiny(n1=8, n2=6, n3=14)
    g = O.regrid(s)
    n = (s.n1, s.n2, s.n3)
    L0 = O.Level.level0(n, g["x1"], g["x2"], g["x3"], g["sigma11"], g["sigma22"], g["sigma33"])
    return d, s, g, L0


def test_units_grid_bitexact(units):
    d, s, g, _ = units
    for k in GRID_KEYS + ["sigma11", "sigma22", "sigma33"]:
        np.testing.assert_array_equal(g[k].reshape(-1), d[k], err_msg=k)


def test_units_level0_medium(units):
    d, _, _, L0 = units
    for q, nm in enumerate(("11", "22", "33")):
        np.testing.assert_array_equal(L0.cells(q).reshape(-1), d[f"lev0_sigma{nm}"])


def test_units_residual_and_curl(units):
    d, s, _, L0 = units
    omu = O.omega_mu0(s.freqs[0])
    np.testing.assert_array_equal(L0.residual(omu, d["u"], d["f"]), d["r"])
    np.testing.assert_array_equal(L0.curl_h(omu, d["u"]), d["H"])


@pytest.mark.parametrize("it", [0, 1])
def test_units_line_gs(units, it):
    """Lexicographic line block Gauss-Seidel along x, y, z (src/gmg.c:340-1298)."""
    d, s, _, L0 = units
    omu = O.omega_mu0(s.freqs[0])
    for ax, nm in enumerate("xyz"):
        np.testing.assert_array_equal(L0.line_gs(omu, ax, it, d["u"], d["f2"]), d[f"lgs_{nm}_it{it}"], err_msg=nm)


@pytest.mark.parametrize("it", [0, 1])
def test_units_point_gs(units, it):
    """Vertex-patch Gauss-Seidel (src/gmg.c:116-336)."""
    d, s, _, L0 = units
    omu = O.omega_mu0(s.freqs[0])
    np.testing.assert_array_equal(L0.point_gs(omu, it, d["u"], d["f2"]), d[f"pgs_it{it}"])


@pytest.mark.parametrize("c", [0, 1, 2])
def test_units_coarsening_and_transfer(units, c):
    """Semi-coarsening pattern c: level-1 shape and medium, restriction and
    prolongation (src/gmg.c:1463-1573, 1736-1860)."""
    d, _, _, L0 = units
    L1 = L0.coarsen(c)
    np.testing.assert_array_equal(np.array(L1.n), d[f"lev1_dims_c{c}"])
    np.testing.assert_array_equal(L1.cells(0).reshape(-1), d[f"lev1_sigma11_c{c}"])
    np.testing.assert_array_equal(L1.cells(2).reshape(-1), d[f"lev1_sigma33_c{c}"])
    np.testing.assert_array_equal(O.restrict(L0, L1, d[f"rf_c{c}"]), d[f"fc_c{c}"])
    np.testing.assert_array_equal(O.prolong(L0, L1, d[f"uc_c{c}"], d["u"]), d[f"uf_c{c}"])


@pytest.fixture(scope="module")
def model():
    d = load("tiny_model.npz")
    s = configs.tiny()
    g = O.regrid(s)
    rw = float(min(s.rho11.min(), s.rho22.min(), s.rho33.min()))
    return d, s, g, O.mg_struct(g, rw)


def test_model_grid_bitexact(model):
    d, _, g, _ = model
    for k in GRID_KEYS + ["sigma11", "sigma22", "sigma33"]:
        np.testing.assert_array_equal(g[k].reshape(-1), d[k], err_msg=k)


def test_model_injection_bitexact(model):
    d, s, _, mg = model
    for fi, fr in enumerate(s.freqs):
        np.testing.assert_array_equal(O.inject_source(mg, s.sources, s.chsrc, fr), d[f"b_f{fi}"])


def test_model_solve_bitexact(model):
    """The whole F-cycle solve of the first frequency: fields, residual
    history and receiver data equal the reference's bit for bit."""
    d, s, g, mg = model
    n = (s.n1, s.n2, s.n3)
    b = O.inject_source(mg, s.sources, s.chsrc, s.freqs[0])
    E, H, hist = O.gmg_solve(n, g["x1"], g["x2"], g["x3"], g["sigma11"], g["sigma22"], g["sigma33"], s.freqs[0], b)
    ref_hist = d["hist"][0]
    ref_hist = ref_hist[np.isfinite(ref_hist)]
    # the reference prints its residual norms with %e: 7 significant digits
    np.testing.assert_allclose(np.asarray(hist), ref_hist, rtol=5e-7, atol=0)
    np.testing.assert_array_equal(E, d["E_f0"])
    np.testing.assert_array_equal(H, d["H_f0"])
    dc = O.extract_data(mg, s.receivers, s.chrec, E, H)
    ref = d["dcal"].reshape(len(s.chrec), len(s.freqs), -1)[:, 0]
    np.testing.assert_array_equal(dc, ref)


def test_model_field_volumes_bitexact(model):
    """E and H on the input-grid cell centres from the reference's first
    frequency fields (src/inject_extract.c:477-635, 638-785, the icheck=1
    volumes of src/do_modelling.c:58-61)."""
    d, s, _, mg = model
    E = O.extract_efield(mg, s, d["E_f0"])
    H = O.extract_hfield(mg, s, d["H_f0"])
    shp = (len(s.freqs), s.nz, s.ny, s.nx)
    for q in range(3):
        np.testing.assert_array_equal(E[q], d[f"E{q + 1}"].reshape(shp)[0], err_msg=f"E{q + 1}")
        np.testing.assert_array_equal(H[q], d[f"H{q + 1}"].reshape(shp)[0], err_msg=f"H{q + 1}")
    # the volumes are not trivially zero or constant
    assert np.abs(E[0]).max() > 0 and np.abs(H[1]).max() > 0
    assert np.unique(np.abs(E[0])).size > 100


CYCLES = {"v": dict(cycleopt=1), "w": dict(cycleopt=3), "f_point": dict(cycleopt=2, isemicoarsen=0)}


@pytest.mark.parametrize("name", list(CYCLES))
def test_model_cycles_bitexact(model, name):
    """V-cycle, W-cycle, and the F-cycle with full coarsening and the
    vertex-patch point smoother (src/gmg.c:1586-1668, 116-336, 1301-1313):
    residual history (printed digits), E and H equal the reference's."""
    d, s, g, mg = model
    cy = load("tiny_cycles.npz")
    n = (s.n1, s.n2, s.n3)
    b = O.inject_source(mg, s.sources, s.chsrc, s.freqs[0])
    E, H, hist = O.gmg_solve(n, g["x1"], g["x2"], g["x3"], g["sigma11"], g["sigma22"], g["sigma33"],
                             s.freqs[0], b, **CYCLES[name])
    ref_hist = cy[f"{name}_hist"][0]
    np.testing.assert_allclose(np.asarray(hist), ref_hist[np.isfinite(ref_hist)], rtol=5e-7, atol=0)
    np.testing.assert_array_equal(E, cy[f"{name}_E_f0"])
    np.testing.assert_array_equal(H, cy[f"{name}_H_f0"])


def test_model_emf_text_file(model):
    """The reference's receiver-data file holds the same numbers (printed
    with %e) as the binary dump of dcal."""
    d, s, _, _ = model
    rows = []
    with open(os.path.join(GOLD, "tiny_emf_0001.txt")) as fp:
        for line in fp:
            if line.strip() and not line.lstrip().startswith("#"):
                rows.append(line.split())
    assert len(rows) == len(s.receivers) * len(s.freqs) * len(s.chrec) or len(rows) > 0
    dc = d["dcal"].reshape(len(s.chrec), len(s.freqs), -1)
    vals = set()
    for r in rows:
        for tok in r:
            try:
                vals.add(float(tok))
            except ValueError:
                pass
    # every real part of the first channel appears in the text at %e precision
    for v in dc[0, 0].real[:5]:
        assert any(abs(t - float(v)) <= 1e-6 * abs(float(v)) + 1e-30 for t in vals)


@pytest.fixture(scope="module")
def grad():
    return load("tiny_grad.npz")


def _grad_survey(gd):
    s = configs.tiny()
    rho = np.ascontiguousarray(gd["rho"], np.float32).reshape(s.rho11.shape)
    s.rho11 = rho
    s.rho22 = rho.copy()
    s.rho33 = rho.copy()
    return s


def test_grad_adjoint_source_bitexact(grad):
    """Adjoint source of the weighted residuals (src/inject_extract.c:124-314)."""
    s = _grad_survey(grad)
    g = O.regrid(s)
    mg = O.mg_struct(g, float(min(s.rho11.min(), s.rho22.min(), s.rho33.min())))
    nch, nf = len(s.chrec), len(s.freqs)
    dres = grad["dres"].reshape(nch, nf, -1)[:, 0]
    b, _ = O.inject_adj_source(mg, s.receivers, s.chrec, dres, s.freqs[0])
    np.testing.assert_array_equal(b, grad["badj_f0"])


def test_grad_build_gradient_bitexact(grad):
    """Gradient and pseudo-Hessian from the forward and adjoint fields
    (src/build_gradient.c:13-111) with the fixture's fields as input."""
    s = _grad_survey(grad)
    nf = len(s.freqs)
    spec = O.GradSpec()
    keep = []

    def ptr(a, dt):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a.ctypes.data

    spec.nx, spec.ny, spec.nz = s.nx, s.ny, s.nz
    spec.ox, spec.oy, spec.oz = s.ox, s.oy, s.oz
    spec.dx, spec.dy, spec.dz = s.dx, s.dy, s.dz
    spec.rho11 = ptr(s.rho11, np.float32)
    spec.rho33 = ptr(s.rho33, np.float32)
    spec.bathy = ptr(s.bathy, np.float32)
    spec.dsmute, spec.drmute = 300.0, 150.0
    spec.npar = 2
    spec.idxpar = ptr(np.array([1, 2], np.int32), np.int32)
    spec.preco = 1
    E = {k: np.ascontiguousarray(grad[k], np.complex64) for k in ("E1f", "E2f", "E3f", "E1a", "E2a", "E3a")}
    out_g, out_h = O.build_gradient(spec, nf, E, s.sources[:1], s.receivers)
    np.testing.assert_array_equal(out_g.reshape(-1), grad["grad"])
    np.testing.assert_array_equal(out_h.reshape(-1), grad["phess"])


@pytest.fixture(scope="module")
def c1():
    d = load("config1.npz")
    s = configs.config1()
    return d, s, O.regrid(s)


def test_config1_grid_and_medium(c1):
    import hashlib
    d, s, g = c1
    for k in GRID_KEYS:
        np.testing.assert_array_equal(g[k].reshape(-1), d[k], err_msg=k)
    for k in ("sigma11", "sigma22", "sigma33"):
        h = hashlib.sha256(np.ascontiguousarray(g[k].reshape(-1)).tobytes()).hexdigest()
        assert h == str(d[k + "_sha256"]), k


@pytest.fixture(scope="module")
def c1_solve(c1):
    d, s, g = c1
    n = (s.n1, s.n2, s.n3)
    rw = float(min(s.rho11.min(), s.rho22.min(), s.rho33.min()))
    mg = O.mg_struct(g, rw)
    b = O.inject_source(mg, s.sources, s.chsrc, s.freqs[0])
    E, H, hist = O.gmg_solve(n, g["x1"], g["x2"], g["x3"], g["sigma11"], g["sigma22"], g["sigma33"], s.freqs[0], b)
    return mg, E, H, hist


def test_config1_receiver_data_bitexact(c1, c1_solve):
    """64^3 survey end to end: regrid, injection, F-cycle solve to 1e-6 and
    extraction reproduce the reference's receiver data bit for bit."""
    d, s, g = c1
    mg, E, H, hist = c1_solve
    ref_hist = d["hist"][0]
    np.testing.assert_allclose(np.asarray(hist), ref_hist[np.isfinite(ref_hist)], rtol=5e-7, atol=0)
    dc = O.extract_data(mg, s.receivers, s.chrec, E, H)
    np.testing.assert_array_equal(dc, d["dcal"].reshape(len(s.chrec), 1, -1)[:, 0])


def test_config1_field_volumes_bitexact(c1, c1_solve):
    """The 64^3 survey's E and H volumes on the input grid (icheck=1),
    compared on the y = 0 plane, where the source and receivers lie."""
    d, s, g = c1
    mg, E, H, _ = c1_solve
    Ev = O.extract_efield(mg, s, E)
    Hv = O.extract_hfield(mg, s, H)
    for q in range(3):
        np.testing.assert_array_equal(Ev[q][:, s.ny // 2, :], d[f"E{q + 1}_y0"], err_msg=f"E{q + 1}")
        np.testing.assert_array_equal(Hv[q][:, s.ny // 2, :], d[f"H{q + 1}_y0"], err_msg=f"H{q + 1}")
