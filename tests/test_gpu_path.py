"""The device-resident steps of the gradient path against the reference's
fixtures and the product's host entry points (themselves bit-exact with the
fixtures, test_host_path.py): the adjoint source
(src/inject_extract.c:124-314), E and H on the model's cell centres
(src/inject_extract.c:477-785, the reference's icheck=1 volumes) and the
gradient with its pseudo-Hessian and mutes (src/build_gradient.c:13-111).
The adjoint source, E, H and the gradient must match bit for bit; the
pseudo-Hessian goes through the device hypot and is held to float rounding
(2 ulp)."""
import ctypes as C
import os

import numpy as np
import pytest

from libemmi_mgfd_amd import carriers, configs, native
from test_host_path import GOLD, _field_tables, _grad_survey, regridded

pytestmark = pytest.mark.gpu


def _check(rc):
    if rc != 0:
        raise native.NativeError(native.lib().emmi_path_last_error().decode())


def test_inject_adj_source_dev_matches_reference_fixture():
    s, gd = _grad_survey()
    sc = regridded(s)
    nch, nf = len(s.chrec), len(s.freqs)
    sc.dres[...] = gd["dres"].reshape(nch, nf, -1)
    b = np.zeros(sc.nvec, np.complex128)
    x = np.zeros(sc.nvec, np.complex128)
    native.lib().inject_adj_source(sc.acq, sc.emf, x.ctypes.data, b.ctypes.data, 0)
    bd, xd = native.DeviceBuffer(b.nbytes), native.DeviceBuffer(x.nbytes)
    print(f"adjoint source, tiny survey, {len(s.receivers)} receivers", flush=True)
    _check(native.lib().emmi_inject_adj_source_dev(sc.acq, sc.emf, 0, xd.ptr, bd.ptr, None))
    np.testing.assert_array_equal(bd.download(np.complex128), b)
    np.testing.assert_array_equal(bd.download(np.complex128), gd["badj_f0"])
    np.testing.assert_array_equal(xd.download(np.complex128), x)


def test_extract_electric_field_dev_matches_host():
    d = np.load(os.path.join(GOLD, "tiny_model.npz"))
    s = configs.tiny()
    sc = regridded(s)
    arrs = _field_tables(sc, s, ["E1", "E2", "E3"])
    x = np.ascontiguousarray(d["E_f0"])
    native.lib().extract_electric_field(sc.acq, sc.emf, x.ctypes.data, 0)
    xd = native.DeviceBuffer(x.nbytes).upload(x)
    nxyz = s.nx * s.ny * s.nz
    ed = native.DeviceBuffer(3 * nxyz * 8)
    print(f"E on the model grid, tiny survey {s.nx}x{s.ny}x{s.nz}", flush=True)
    _check(native.lib().emmi_extract_electric_field_dev(sc.acq, sc.emf, 0, xd.ptr, ed.ptr, None))
    e = ed.download(np.complex64, (3, s.nz, s.ny, s.nx))
    shp = (len(s.freqs), s.nz, s.ny, s.nx)
    for c in range(3):
        np.testing.assert_array_equal(e[c], arrs[f"E{c + 1}"][0], err_msg=f"E{c + 1}")
        np.testing.assert_array_equal(e[c], d[f"E{c + 1}"].reshape(shp)[0], err_msg=f"E{c + 1} vs reference")
    assert np.abs(e).max() > 0


def test_extract_magnetic_field_dev_matches_reference_fixture():
    d = np.load(os.path.join(GOLD, "tiny_model.npz"))
    s = configs.tiny()
    sc = regridded(s)
    arrs = _field_tables(sc, s, ["H1", "H2", "H3"])
    h = np.ascontiguousarray(d["H_f0"])
    native.lib().extract_magnetic_field(sc.acq, sc.emf, h.ctypes.data, 0)
    hd = native.DeviceBuffer(h.nbytes).upload(h)
    nxyz = s.nx * s.ny * s.nz
    ed = native.DeviceBuffer(3 * nxyz * 8)
    print(f"H on the model grid, tiny survey {s.nx}x{s.ny}x{s.nz}", flush=True)
    _check(native.lib().emmi_extract_magnetic_field_dev(sc.acq, sc.emf, 0, hd.ptr, ed.ptr, None))
    e = ed.download(np.complex64, (3, s.nz, s.ny, s.nx))
    shp = (len(s.freqs), s.nz, s.ny, s.nx)
    for c in range(3):
        np.testing.assert_array_equal(e[c], d[f"H{c + 1}"].reshape(shp)[0], err_msg=f"H{c + 1} vs reference")
        np.testing.assert_array_equal(e[c], arrs[f"H{c + 1}"][0], err_msg=f"H{c + 1}")
    assert np.abs(e).max() > 0


def test_build_gradient_dev_matches_reference_fixture():
    s, gd = _grad_survey()
    sc = carriers.ShotCarriers(s)
    e = sc.emf
    e.dsmute, e.drmute = 300.0, 150.0
    nf = len(s.freqs)
    shp = (nf, s.nz, s.ny, s.nx)
    E = {k: np.ascontiguousarray(gd[k].reshape(shp), np.complex64) for k in ("E1f", "E2f", "E3f", "E1a", "E2a", "E3a")}
    fwd = [native.DeviceBuffer(3 * E["E1f"][f].nbytes).upload(np.stack([E["E1f"][f], E["E2f"][f], E["E3f"][f]]))
           for f in range(nf)]
    adj = [native.DeviceBuffer(3 * E["E1a"][f].nbytes).upload(np.stack([E["E1a"][f], E["E2a"][f], E["E3a"][f]]))
           for f in range(nf)]
    fwi = carriers.Fwi()
    fwi.npar = 2
    fwi.n = 2 * s.nx * s.ny * s.nz
    fwi.preco = 1
    idx = np.array([1, 2], np.int32)
    fwi.idxpar = idx.ctypes.data_as(C.POINTER(C.c_int))
    nbytes = 2 * s.nx * s.ny * s.nz * 4
    gdev, hdev = native.DeviceBuffer(nbytes), native.DeviceBuffer(nbytes)
    pf = (C.c_void_p * nf)(*[b.ptr for b in fwd])
    pa = (C.c_void_p * nf)(*[b.ptr for b in adj])
    print(f"gradient, tiny survey, {nf} frequencies, 2 parameters", flush=True)
    _check(native.lib().emmi_build_gradient_dev(sc.acq, e, C.byref(fwi), pf, pa, gdev.ptr, hdev.ptr, None))
    grad = gdev.download(np.float32)
    phess = hdev.download(np.float32)
    np.testing.assert_array_equal(grad, gd["grad"])
    ref = gd["phess"]
    ulp = np.spacing(np.abs(ref).astype(np.float32))
    assert np.all(np.abs(phess - ref) <= 2 * ulp), float(np.max(np.abs(phess - ref) / np.maximum(ulp, 1e-45)))
    assert np.count_nonzero(grad) > 0 and np.count_nonzero(grad) < grad.size   # mutes and bathymetry applied


def test_build_gradient_dev_rejects_bad_sizes():
    s, gd = _grad_survey()
    sc = carriers.ShotCarriers(s)
    fwi = carriers.Fwi()
    fwi.npar = 9
    fwi.preco = 0
    idx = np.ones(9, np.int32)
    fwi.idxpar = idx.ctypes.data_as(C.POINTER(C.c_int))
    nf = len(s.freqs)
    null = (C.c_void_p * nf)()
    assert native.lib().emmi_build_gradient_dev(sc.acq, sc.emf, C.byref(fwi), null, null, None, None, None) == -1
    assert b"npar" in native.lib().emmi_path_last_error()
