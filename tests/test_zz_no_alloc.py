"""emmi_inject_adj_source_dev_noalloc -- the adjoint source with its device
tables in the library's static scratch -- against emmi_inject_adj_source_dev
through the C-ABI: identical b and x (bit for bit) on two calls in a row, the
second with fewer receivers and a smaller grid than the first (the scratch
keeps the first call's contents, which must not leak into the second), and
a clear error for tables beyond the scratch."""
import numpy as np
import pytest

from libemmi_mgfd_amd import carriers, configs, native

pytestmark = pytest.mark.gpu


def _survey(nrec, n):
    s = configs.tiny(n1=n[0], n2=n[1], n3=n[2], nrec=nrec)
    s.chrec = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
    return s


def _residuals(sc, seed):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal(sc.dres.shape) + 1j * rng.standard_normal(sc.dres.shape)
    sc.dres[...] = (d * 1e-12).astype(np.complex64)


def _adjoint(fn, sc, f):
    nb = 16 * sc.nvec
    b, x = native.DeviceBuffer(nb), native.DeviceBuffer(nb)
    # stale values in the outputs: the entry points overwrite them
    b.upload(np.full(sc.nvec, 7.0 + 1j, np.complex128))
    x.upload(np.full(sc.nvec, -3.0 + 2j, np.complex128))
    rc = fn(sc.acq, sc.emf, f, x.ptr, b.ptr, None)
    if rc != 0:
        raise native.NativeError(native.lib().emmi_path_last_error().decode())
    return b.download(np.complex128), x.download(np.complex128)


def test_noalloc_adjoint_source_matches_allocating_entry_point():
    L = native.lib()
    cases = [(9, (16, 12, 20), 0, 11), (3, (12, 10, 16), 1, 12)]
    for nrec, n, f, seed in cases:
        s = _survey(nrec, n)
        sc = carriers.ShotCarriers(s)
        L.regrid_init(sc.acq, sc.emf)
        _residuals(sc, seed)
        print(f"adjoint source: {nrec} receivers, 6 channels, grid {n}, frequency {f}", flush=True)
        b0, x0 = _adjoint(L.emmi_inject_adj_source_dev, sc, f)
        b1, x1 = _adjoint(L.emmi_inject_adj_source_dev_noalloc, sc, f)
        np.testing.assert_array_equal(b1, b0)
        np.testing.assert_array_equal(x1, x0)
        assert np.count_nonzero(b0) > 10 * nrec and np.count_nonzero(x0) > 10 * nrec
        # host entry point of the reference's arithmetic
        bh = np.zeros(sc.nvec, np.complex128)
        xh = np.zeros(sc.nvec, np.complex128)
        L.inject_adj_source(sc.acq, sc.emf, xh.ctypes.data, bh.ctypes.data, f)
        np.testing.assert_array_equal(b1, bh)


def test_noalloc_adjoint_source_refuses_tables_beyond_the_scratch():
    """19200 receivers on a 40x40x12 lattice through the 128^3 survey's
    earth: about 400k distinct table entries (10 MB), beyond the 2 MiB
    scratch."""
    L = native.lib()
    s = configs.config2()
    xy = np.linspace(-4800.0, 4800.0, 40)
    zz = np.linspace(100.0, 1900.0, 12)
    s.receivers = [(float(x), float(y), float(z), 0.0, 0.0) for z in zz for y in xy for x in xy]
    s.table = [(1, i + 1) for i in range(len(s.receivers))]
    s.freqs = [0.25]
    s.chrec = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
    sc = carriers.ShotCarriers(s)
    L.regrid_init(sc.acq, sc.emf)
    _residuals(sc, 3)
    nb = 16 * sc.nvec
    b, x = native.DeviceBuffer(nb), native.DeviceBuffer(nb)
    rc = L.emmi_inject_adj_source_dev_noalloc(sc.acq, sc.emf, 0, x.ptr, b.ptr, None)
    msg = L.emmi_path_last_error().decode()
    print("19200 receivers:", rc, msg, flush=True)
    assert rc == -1 and "static scratch" in msg
    # the allocating entry point takes it
    assert L.emmi_inject_adj_source_dev(sc.acq, sc.emf, 0, x.ptr, b.ptr, None) == 0
