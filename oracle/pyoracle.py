"""ctypes binding of the oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
c128p = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
c64p = np.ctypeslib.ndpointer(dtype=np.complex64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Model(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nb", C.c_int),
                ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
                ("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float),
                ("lextend", C.c_float), ("rho_air", C.c_float),
                ("n1", C.c_int), ("n2", C.c_int), ("n3", C.c_int),
                ("d1min", C.c_float), ("d2min", C.c_float), ("d3min", C.c_float),
                ("rho11", C.c_void_p), ("rho22", C.c_void_p), ("rho33", C.c_void_p)]


class Grid(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("xx", "yy", "zz", "x1", "x2", "x3", "x1s", "x2s", "x3s",
                                          "d1", "d2", "d3", "d1s", "d2s", "d3s",
                                          "sigma11", "sigma22", "sigma33")]


class Params(C.Structure):
    _fields_ = [("cycleopt", C.c_int), ("ncycle", C.c_int), ("v1", C.c_int), ("v2", C.c_int),
                ("isemicoarsen", C.c_int), ("tol", C.c_double), ("colour_order", C.c_int),
                ("verb", C.c_int)]


class MG(C.Structure):
    _fields_ = [("n1", C.c_int), ("n2", C.c_int), ("n3", C.c_int)] + \
               [(n, C.c_void_p) for n in ("x1", "x2", "x3", "x1s", "x2s", "x3s", "d1", "d2", "d3", "sigma33")] + \
               [("rho_water", C.c_float)]


class GradSpec(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float),
                ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
                ("rho11", C.c_void_p), ("rho33", C.c_void_p), ("bathy", C.c_void_p),
                ("dsmute", C.c_float), ("drmute", C.c_float), ("npar", C.c_int),
                ("idxpar", C.c_void_p), ("preco", C.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])
        L = C.CDLL(path)
        L.emo_find_index.restype = C.c_int
        L.emo_find_index.argtypes = [C.c_int, f32p, C.c_float]
        L.emo_create_nugrid.restype = C.c_float
        L.emo_create_nugrid.argtypes = [C.c_int, C.c_float, C.c_float, f32p]
        L.emo_regrid.argtypes = [C.POINTER(Model), C.c_float, C.c_float, C.c_float, C.POINTER(Grid)]
        L.emo_level0_new.restype = C.c_void_p
        L.emo_level0_new.argtypes = [C.c_int] * 3 + [f32p] * 6
        L.emo_level_coarsen.restype = C.c_void_p
        L.emo_level_coarsen.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.emo_level_free.argtypes = [C.c_void_p]
        L.emo_level_dims.argtypes = [C.c_void_p, i32p, i32p]
        L.emo_level_cells.restype = C.POINTER(C.c_double)
        L.emo_level_cells.argtypes = [C.c_void_p, C.c_int]
        L.emo_level_axis.restype = C.POINTER(C.c_double)
        L.emo_level_axis.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.emo_residual.argtypes = [C.c_void_p, C.c_double, c128p, c128p, c128p]
        L.emo_apply.argtypes = [C.c_void_p, C.c_double, c128p, c128p]
        L.emo_line_gs.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int, c128p, c128p]
        L.emo_line_solve.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int, c128p, c128p]
        L.emo_point_gs.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, c128p, c128p]
        L.emo_restrict.argtypes = [C.c_void_p, C.c_void_p, c128p, c128p]
        L.emo_prolong.argtypes = [C.c_void_p, C.c_void_p, c128p, c128p]
        L.emo_curl_h.argtypes = [C.c_void_p, C.c_double, c128p, c128p]
        L.emo_default_params.argtypes = [C.POINTER(Params)]
        L.emo_gmg_solve.restype = C.c_int
        L.emo_gmg_solve.argtypes = [C.c_int] * 3 + [f32p] * 6 + [C.c_float, C.POINTER(Params), c128p, c128p, f64p]
        L.emo_inject_source.argtypes = [C.POINTER(MG), C.c_int] + [f32p] * 5 + [C.c_int, i32p, C.c_float, c128p]
        L.emo_inject_adj_source.argtypes = [C.POINTER(MG), C.c_int] + [f32p] * 5 + [C.c_int, i32p, c64p, C.c_float, c128p, c128p]
        L.emo_extract_data.argtypes = [C.POINTER(MG), C.c_int] + [f32p] * 5 + [C.c_int, i32p, c128p, c128p, c64p]
        L.emo_extract_efield.argtypes = [C.POINTER(MG), C.c_int, C.c_int, C.c_int] + [C.c_float] * 6 + [f32p, c128p, c64p, c64p, c64p]
        L.emo_extract_hfield.argtypes = [C.POINTER(MG), C.c_int, C.c_int, C.c_int] + [C.c_float] * 6 + [c128p, c64p, c64p, c64p]
        L.emo_build_gradient.argtypes = [C.POINTER(GradSpec), C.c_int] + [c64p] * 6 + \
            [C.c_int, f32p, f32p, f32p, C.c_int, f32p, f32p, f32p, f32p, f32p]
        _LIB = L
    return _LIB


MU0 = 4.0 * 3.14159265358979323846264 * 1e-7
CHANNELS = {"Ex": 0, "Ey": 1, "Ez": 2, "Hx": 3, "Hy": 4, "Hz": 5}


def omega_mu0(freq: float) -> float:
    """(double)(float)(2 pi f) * mu0, the reference's i*omega*mu0 imaginary part."""
    return float(omega_f32(freq)) * MU0


def omega_f32(freq: float) -> np.float32:
    """(float)(2.*PI*freq) with freq parsed as float (src/emf_init_free.c:23)."""
    return np.float32(2.0 * 3.14159265358979323846264 * float(np.float32(freq)))


class RegridResult(dict):
    pass


def regrid(s) -> RegridResult:
    """Run the oracle's regridding on a configs.Survey (first source)."""
    L = lib()
    nb = s.nb
    out = RegridResult()
    out["xx"] = np.zeros(s.nx + 2 * nb + 1, np.float32)
    out["yy"] = np.zeros(s.ny + 2 * nb + 1, np.float32)
    out["zz"] = np.zeros(s.nz + 2 * nb + 1, np.float32)
    for a, n in zip("123", (s.n1, s.n2, s.n3)):
        for k in ("x", "xs", "d", "ds"):
            name = k[0] + a + k[1:]
            out[name] = np.zeros(n + 1, np.float32)
    for k in ("sigma11", "sigma22", "sigma33"):
        out[k] = np.zeros((s.n3, s.n2, s.n1), np.float32)
    r11 = np.ascontiguousarray(s.rho11, np.float32)
    r22 = np.ascontiguousarray(s.rho22, np.float32)
    r33 = np.ascontiguousarray(s.rho33, np.float32)
    m = Model(s.nx, s.ny, s.nz, s.nb, s.dx, s.dy, s.dz, s.ox, s.oy, s.oz, s.lextend, s.rho_air,
              s.n1, s.n2, s.n3, s.d1min, s.d2min, s.d3min,
              r11.ctypes.data, r22.ctypes.data, r33.ctypes.data)
    g = Grid(*[out[k].ctypes.data for k in ("xx", "yy", "zz", "x1", "x2", "x3", "x1s", "x2s", "x3s",
                                             "d1", "d2", "d3", "d1s", "d2s", "d3s",
                                             "sigma11", "sigma22", "sigma33")])
    src = s.sources[0]
    L.emo_regrid(C.byref(m), np.float32(src[0]), np.float32(src[1]), np.float32(src[2]), C.byref(g))
    out["_keep"] = (r11, r22, r33)
    return out


class Level:
    """An oracle multigrid level."""

    def __init__(self, handle, parent=None):
        self.h = handle
        self.parent = parent
        dims = np.zeros(3, np.int32)
        sc = np.zeros(3, np.int32)
        lib().emo_level_dims(self.h, dims, sc)
        self.n = tuple(int(v) for v in dims)
        self.sc = tuple(int(v) for v in sc)

    @classmethod
    def level0(cls, n, x1, x2, x3, s11, s22, s33):
        f = lambda a: np.ascontiguousarray(a, np.float32)
        keep = [f(a) for a in (x1, x2, x3, s11, s22, s33)]
        h = lib().emo_level0_new(n[0], n[1], n[2], *keep)
        lv = cls(h)
        lv._keep = keep
        return lv

    def coarsen(self, icycle, isemi=1):
        return Level(lib().emo_level_coarsen(self.h, icycle, isemi), parent=self)

    @property
    def nvec(self):
        n1, n2, n3 = self.n
        return 3 * (n1 + 1) * (n2 + 1) * (n3 + 1)

    @property
    def shape(self):
        n1, n2, n3 = self.n
        return (3, n3 + 1, n2 + 1, n1 + 1)

    def cells(self, which):
        n1, n2, n3 = self.n
        p = lib().emo_level_cells(self.h, which)
        return np.ctypeslib.as_array(p, shape=(n3, n2, n1)).copy()

    def axis(self, a, kind):
        p = lib().emo_level_axis(self.h, a, kind)
        return np.ctypeslib.as_array(p, shape=(self.n[a] + 1,)).copy()

    def zeros(self):
        return np.zeros(self.nvec, np.complex128)

    def residual(self, omu, u, f):
        r = self.zeros()
        lib().emo_residual(self.h, omu, _c(u), _c(f), r)
        return r

    def apply(self, omu, u):
        y = self.zeros()
        lib().emo_apply(self.h, omu, _c(u), y)
        return y

    def line_gs(self, omu, axis, it, u, f, colour=0):
        u = _c(u).copy()
        lib().emo_line_gs(self.h, omu, axis, it, colour, u, _c(f))
        return u

    def line_solve(self, omu, axis, pb, pc, u, f):
        """Exact solve of the single line (axis; pb, pc) for the current residual."""
        u = _c(u).copy()
        lib().emo_line_solve(self.h, omu, axis, pb, pc, u, _c(f))
        return u

    def point_gs(self, omu, it, u, f, colour=0):
        u = _c(u).copy()
        lib().emo_point_gs(self.h, omu, it, colour, u, _c(f))
        return u

    def curl_h(self, omu, u):
        h = self.zeros()
        lib().emo_curl_h(self.h, omu, _c(u), h)
        return h

    def __del__(self):
        try:
            lib().emo_level_free(self.h)
        except Exception:
            pass


def restrict(fine: Level, coarse: Level, r):
    fc = coarse.zeros()
    lib().emo_restrict(fine.h, coarse.h, _c(r), fc)
    return fc


def prolong(fine: Level, coarse: Level, uc, uf):
    uf = _c(uf).copy()
    lib().emo_prolong(fine.h, coarse.h, _c(uc), uf)
    return uf


def _c(a):
    return np.ascontiguousarray(a, np.complex128).reshape(-1)


def params(**kw) -> Params:
    p = Params()
    lib().emo_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def gmg_solve(n, x1, x2, x3, s11, s22, s33, freq, b, **kw):
    """Full reference-algorithm solve; returns (E, H, residual history)."""
    p = params(**kw)
    f = lambda a: np.ascontiguousarray(a, np.float32)
    b = _c(b).copy()
    x = np.zeros_like(b)
    hist = np.zeros(max(p.ncycle, 1) + 1, np.float64)
    nh = lib().emo_gmg_solve(n[0], n[1], n[2], f(x1), f(x2), f(x3), f(s11), f(s22), f(s33),
                             omega_f32(freq), C.byref(p), b, x, hist)
    return x, b, hist[:nh]


def mg_struct(g, rho_water):
    """emo_mg_t over a regrid() result; keeps the arrays alive on the struct."""
    m = MG(g["x1"].size - 1, g["x2"].size - 1, g["x3"].size - 1,
           *[g[k].ctypes.data for k in ("x1", "x2", "x3", "x1s", "x2s", "x3s", "d1", "d2", "d3", "sigma33")],
           rho_water)
    m._keep = g
    return m


def deg2rad_f32(v) -> np.float32:
    """float(PI * (float)deg / 180.) as src/acq_init_free.c:220-247 converts."""
    return np.float32(3.14159265358979323846264 * float(np.float32(v)) / 180.0)


def _acq_arrays(points):
    a = np.asarray(points, np.float64)
    x, y, z = (np.ascontiguousarray(a[:, i], np.float32) for i in range(3))
    az = np.ascontiguousarray([deg2rad_f32(v) for v in a[:, 3]], np.float32)
    dip = np.ascontiguousarray([deg2rad_f32(v) for v in a[:, 4]], np.float32)
    return x, y, z, az, dip


def inject_source(mg: MG, sources, chsrc, freq):
    n = 3 * (mg.n1 + 1) * (mg.n2 + 1) * (mg.n3 + 1)
    b = np.zeros(n, np.complex128)
    x, y, z, az, dip = _acq_arrays(sources)
    ch = np.array([CHANNELS[c] for c in chsrc], np.int32)
    lib().emo_inject_source(C.byref(mg), len(x), x, y, z, az, dip, len(ch), ch, omega_f32(freq), b)
    return b


def inject_adj_source(mg: MG, receivers, chrec, dres, freq):
    n = 3 * (mg.n1 + 1) * (mg.n2 + 1) * (mg.n3 + 1)
    b = np.zeros(n, np.complex128)
    xs = np.zeros(n, np.complex128)
    x, y, z, az, dip = _acq_arrays(receivers)
    ch = np.array([CHANNELS[c] for c in chrec], np.int32)
    d = np.ascontiguousarray(dres, np.complex64)
    lib().emo_inject_adj_source(C.byref(mg), len(x), x, y, z, az, dip, len(ch), ch, d, omega_f32(freq), xs, b)
    return b, xs


def extract_data(mg: MG, receivers, chrec, E, H):
    x, y, z, az, dip = _acq_arrays(receivers)
    ch = np.array([CHANNELS[c] for c in chrec], np.int32)
    out = np.zeros((len(ch), len(x)), np.complex64)
    lib().emo_extract_data(C.byref(mg), len(x), x, y, z, az, dip, len(ch), ch, _c(E), _c(H), out)
    return out


def extract_efield(mg: MG, s, x):
    """E on the input-grid cell centres of survey s: (E1, E2, E3) complex64 [nz][ny][nx]."""
    shp = (s.nz, s.ny, s.nx)
    E = [np.zeros(shp, np.complex64) for _ in range(3)]
    rho33 = np.ascontiguousarray(s.rho33, np.float32)
    lib().emo_extract_efield(C.byref(mg), s.nx, s.ny, s.nz, s.ox, s.oy, s.oz, s.dx, s.dy, s.dz, rho33, _c(x), *E)
    return E


def extract_hfield(mg: MG, s, h):
    shp = (s.nz, s.ny, s.nx)
    H = [np.zeros(shp, np.complex64) for _ in range(3)]
    lib().emo_extract_hfield(C.byref(mg), s.nx, s.ny, s.nz, s.ox, s.oy, s.oz, s.dx, s.dy, s.dz, _c(h), *H)
    return H


def build_gradient(spec: GradSpec, nfreq, E, sources, receivers):
    """grad, phess [npar][nz][ny][nx] from the forward (E?f) and adjoint (E?a)
    fields, complex64 [nfreq][nz][ny][nx] each, in dict E."""
    n = spec.nx * spec.ny * spec.nz
    grad = np.zeros(spec.npar * n, np.float32)
    phess = np.zeros(spec.npar * n, np.float32)
    sx, sy, sz, _, _ = _acq_arrays(sources)
    rx, ry, rz, _, _ = _acq_arrays(receivers)
    f = [np.ascontiguousarray(E[k], np.complex64).reshape(-1) for k in ("E1f", "E2f", "E3f", "E1a", "E2a", "E3a")]
    lib().emo_build_gradient(C.byref(spec), nfreq, *f, len(sx), sx, sy, sz, len(rx), rx, ry, rz, grad, phess)
    return grad.reshape(spec.npar, spec.nz, spec.ny, spec.nx), phess.reshape(spec.npar, spec.nz, spec.ny, spec.nx)
