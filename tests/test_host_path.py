"""The product library's host entry points (libemmi_mgfd_amd/libemmi_mgfd.so,
CPU side: regrid_init, inject_source, inject_adj_source, extract_emf_data,
extract_electric_field, extract_magnetic_field, build_gradient, find_index,
create_nugrid) called through the C-ABI with layout-identical emf_t/acq_t/
fwi_t carriers, compared bit for bit with the oracle and with the reference's
golden fixtures.  No GPU work happens here."""
import os

import numpy as np
import pytest

import pyoracle as O
from libemmi_mgfd_amd import carriers, configs, native

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def regridded(s):
    sc = carriers.ShotCarriers(s)
    native.lib().regrid_init(sc.acq, sc.emf)
    return sc


@pytest.mark.parametrize("name", ["tiny", "config1"])
def test_regrid_bitexact(name):
    s = configs.CONFIGS[name]()
    sc = regridded(s)
    g, go = sc.grid(), O.regrid(s)
    for k in g:
        np.testing.assert_array_equal(g[k].reshape(-1), go[k].reshape(-1), err_msg=k)
    native.lib().regrid_free(sc.emf)


@pytest.mark.parametrize("name", ["tiny", "config1"])
def test_inject_source_bitexact(name):
    s = configs.CONFIGS[name]()
    sc = regridded(s)
    mg = O.mg_struct(O.regrid(s), sc.emf.rho_water)
    for fi, fr in enumerate(s.freqs):
        b = np.zeros(sc.nvec, np.complex128)
        native.lib().inject_source(sc.acq, sc.emf, b.ctypes.data, fi)
        np.testing.assert_array_equal(b, O.inject_source(mg, s.sources, s.chsrc, fr))


def test_inject_source_matches_reference_fixture():
    d = np.load(os.path.join(GOLD, "tiny_model.npz"))
    s = configs.tiny()
    sc = regridded(s)
    for fi in range(len(s.freqs)):
        b = np.zeros(sc.nvec, np.complex128)
        native.lib().inject_source(sc.acq, sc.emf, b.ctypes.data, fi)
        np.testing.assert_array_equal(b, d[f"b_f{fi}"])


def test_extract_emf_data_matches_reference_fixture():
    """Receiver data gathered from the reference's own E and H."""
    d = np.load(os.path.join(GOLD, "tiny_model.npz"))
    s = configs.tiny()
    sc = regridded(s)
    E = np.ascontiguousarray(d["E_f0"])
    H = np.ascontiguousarray(d["H_f0"])
    native.lib().extract_emf_data(sc.acq, sc.emf, E.ctypes.data, H.ctypes.data, 0)
    ref = d["dcal"].reshape(len(s.chrec), len(s.freqs), -1)[:, 0]
    np.testing.assert_array_equal(sc.dcal[:, 0, :], ref)


def _grad_survey():
    gd = np.load(os.path.join(GOLD, "tiny_grad.npz"))
    s = configs.tiny()
    rho = np.ascontiguousarray(gd["rho"], np.float32).reshape(s.rho11.shape)
    s.rho11, s.rho22, s.rho33 = rho, rho.copy(), rho.copy()
    return s, gd


def test_inject_adj_source_matches_reference_fixture():
    s, gd = _grad_survey()
    sc = regridded(s)
    nch, nf = len(s.chrec), len(s.freqs)
    sc.dres[...] = gd["dres"].reshape(nch, nf, -1)
    b = np.zeros(sc.nvec, np.complex128)
    x = np.zeros(sc.nvec, np.complex128)
    native.lib().inject_adj_source(sc.acq, sc.emf, x.ctypes.data, b.ctypes.data, 0)
    np.testing.assert_array_equal(b, gd["badj_f0"])


def _field_tables(sc, s, names):
    nf = len(s.freqs)
    arrs = {}
    for nm in names:
        a = np.zeros((nf, s.nz, s.ny, s.nx), np.complex64)
        arrs[nm] = a
        setattr(sc.emf, nm, sc.t.c4(a))
    return arrs


def test_extract_fields_match_reference_fixture():
    """E and H on the input-grid cell centres (src/inject_extract.c:477-785):
    the reference's own icheck=1 volumes of the fixture's first frequency,
    and the oracle."""
    d = np.load(os.path.join(GOLD, "tiny_model.npz"))
    s = configs.tiny()
    sc = regridded(s)
    arrs = _field_tables(sc, s, ["E1", "E2", "E3", "H1", "H2", "H3"])
    E = np.ascontiguousarray(d["E_f0"])
    H = np.ascontiguousarray(d["H_f0"])
    native.lib().extract_electric_field(sc.acq, sc.emf, E.ctypes.data, 0)
    native.lib().extract_magnetic_field(sc.acq, sc.emf, H.ctypes.data, 0)
    mg = O.mg_struct(O.regrid(s), sc.emf.rho_water)
    Eo = O.extract_efield(mg, s, E)
    Ho = O.extract_hfield(mg, s, H)
    shp = (len(s.freqs), s.nz, s.ny, s.nx)
    for i in range(3):
        np.testing.assert_array_equal(arrs[f"E{i + 1}"][0], d[f"E{i + 1}"].reshape(shp)[0])
        np.testing.assert_array_equal(arrs[f"H{i + 1}"][0], d[f"H{i + 1}"].reshape(shp)[0])
        np.testing.assert_array_equal(arrs[f"E{i + 1}"][0], Eo[i])
        np.testing.assert_array_equal(arrs[f"H{i + 1}"][0], Ho[i])
    assert np.abs(arrs["E1"][0]).max() > 0


def test_build_gradient_matches_reference_fixture():
    s, gd = _grad_survey()
    sc = carriers.ShotCarriers(s)
    e = sc.emf
    e.dsmute, e.drmute = 300.0, 150.0
    nf = len(s.freqs)
    shp = (nf, s.nz, s.ny, s.nx)
    E = {k: np.ascontiguousarray(gd[k].reshape(shp), np.complex64) for k in ("E1f", "E2f", "E3f", "E1a", "E2a", "E3a")}
    tabs = [sc.t.c4(E[k]) for k in ("E1f", "E2f", "E3f", "E1a", "E2a", "E3a")]
    fwi = carriers.Fwi()
    fwi.npar = 2
    fwi.n = 2 * s.nx * s.ny * s.nz
    fwi.preco = 1
    idx = np.array([1, 2], np.int32)
    fwi.idxpar = idx.ctypes.data_as(carriers.C.POINTER(carriers.C.c_int))
    grad = np.zeros((2, s.nz, s.ny, s.nx), np.float32)
    phess = np.zeros((2, s.nz, s.ny, s.nx), np.float32)
    fwi.grad = sc.t.f4(grad)
    fwi.phess = sc.t.f4(phess)
    native.lib().build_gradient(sc.acq, e, carriers.C.byref(fwi), *tabs)
    np.testing.assert_array_equal(grad.reshape(-1), gd["grad"])
    np.testing.assert_array_equal(phess.reshape(-1), gd["phess"])


def test_find_index_and_create_nugrid_match_oracle():
    L = native.lib()
    rng = np.random.default_rng(3)
    x = np.sort(rng.uniform(-1e4, 1e4, 101)).astype(np.float32)
    for v in rng.uniform(-1.1e4, 1.1e4, 300).astype(np.float32):
        assert L.find_index(x.size, x.ctypes.data, np.float32(v)) == O.lib().emo_find_index(x.size, x, np.float32(v))
    for n, ln, dx in ((10, 8000.0, 100.0), (20, 50000.0, 50.0), (10, 500.0, 100.0)):
        a = np.zeros(n + 1, np.float32)
        b = np.zeros(n + 1, np.float32)
        qa = L.create_nugrid(n, ln, dx, a.ctypes.data)
        qb = O.lib().emo_create_nugrid(n, ln, dx, b)
        np.testing.assert_array_equal(a, b)
        assert np.float32(qa) == np.float32(qb)
