"""ctypes mirrors of the reference's carriers emf_t, acq_t and fwi_t.

Field order and types follow include/emf.h:4-48, include/acq.h:4-19 and
include/fwi.h:4-28 of libEMMI_MGFD (mirrored in include/emmi_mgfd.h).  3D
and 4D arrays are pointer tables over one contiguous numpy buffer, so
``&a[0][0][0]`` is the buffer, exactly as the reference's allocators lay
them out.  Builders turn a configs.Survey shot into filled carriers the
way src/emf_init_free.c and src/acq_init_free.c fill them.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import configs

FP = C.POINTER(C.c_float)
FPP = C.POINTER(FP)
FPPP = C.POINTER(FPP)
FPPPP = C.POINTER(FPPP)


class C64(C.Structure):
    _fields_ = [("re", C.c_float), ("im", C.c_float)]


CP = C.POINTER(C64)
CPP = C.POINTER(CP)
CPPP = C.POINTER(CPP)
CPPPP = C.POINTER(CPPP)


class Emf(C.Structure):
    _fields_ = [
        ("mode", C.c_int), ("verb", C.c_int), ("reciprocity", C.c_int), ("nfreq", C.c_int),
        ("freqs", FP), ("omegas", FP),
        ("rho_air", C.c_float), ("rho_water", C.c_float), ("lextend", C.c_float),
        ("nchsrc", C.c_int), ("nchrec", C.c_int),
        ("chsrc", C.POINTER(C.c_char_p)), ("chrec", C.POINTER(C.c_char_p)),
        ("x1min", C.c_float), ("x1max", C.c_float), ("x2min", C.c_float), ("x2max", C.c_float),
        ("x3min", C.c_float), ("x3max", C.c_float),
        ("xx", FP), ("yy", FP), ("zz", FP),
        ("n1", C.c_int), ("n2", C.c_int), ("n3", C.c_int),
        ("d1min", C.c_float), ("d2min", C.c_float), ("d3min", C.c_float),
        ("x1", FP), ("x2", FP), ("x3", FP), ("x1s", FP), ("x2s", FP), ("x3s", FP),
        ("d1", FP), ("d2", FP), ("d3", FP), ("d1s", FP), ("d2s", FP), ("d3s", FP),
        ("sigma11", FPPP), ("sigma22", FPPP), ("sigma33", FPPP), ("invmur", FPPP),
        ("dcal_fd", CPPP), ("dobs_fd", CPPP), ("dres_fd", CPPP),
        ("E1", CPPPP), ("E2", CPPPP), ("E3", CPPPP), ("H1", CPPPP), ("H2", CPPPP), ("H3", CPPPP),
        ("mur", FPPP), ("rho11", FPPP), ("rho22", FPPP), ("rho33", FPPP),
        ("bathy", FPP),
        ("rhomin", C.c_float), ("rhomax", C.c_float), ("depthmax", C.c_float),
        ("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float),
        ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nb", C.c_int),
        ("addnoise", C.c_int),
        ("amp_perc", C.c_float), ("delta_phi", C.c_float), ("delta_theta", C.c_float),
        ("noisefloorE", C.c_float), ("noisefloorH", C.c_float),
        ("delta_emf", FPPP), ("rmse", FPPP),
        ("dsmute", C.c_float), ("drmute", C.c_float),
    ]


class Acq(C.Structure):
    _fields_ = [
        ("nsrc", C.c_int), ("nrec", C.c_int), ("nsubsrc", C.c_int), ("nsubrec", C.c_int),
        ("lensrc", C.c_float), ("lenrec", C.c_float),
        ("src_x1", FP), ("src_x2", FP), ("src_x3", FP), ("src_azimuth", FP), ("src_dip", FP),
        ("rec_x1", FP), ("rec_x2", FP), ("rec_x3", FP), ("rec_azimuth", FP), ("rec_dip", FP),
        ("x1min", C.c_float), ("x1max", C.c_float), ("x2min", C.c_float), ("x2max", C.c_float),
        ("x3min", C.c_float), ("x3max", C.c_float),
        ("nsrc_total", C.c_int), ("nrec_total", C.c_int),
        ("shot_idx", C.POINTER(C.c_int)),
    ]


class Fwi(C.Structure):
    _fields_ = [
        ("npar", C.c_int), ("n", C.c_int), ("preco", C.c_int), ("r1", C.c_int), ("r2", C.c_int),
        ("r3", C.c_int), ("repeat", C.c_int),
        ("idxpar", C.POINTER(C.c_int)), ("minpar", FP), ("maxpar", FP),
        ("iter", C.c_int), ("niter", C.c_int), ("xref", FP),
        ("gamma1", C.c_float), ("gamma2", C.c_float), ("firstgrad", C.c_int), ("alpha", C.c_float),
        ("fcost_dat", C.c_float), ("fcost_mod", C.c_float), ("fcost", C.c_float), ("fcost_list", FP),
        ("grad", FPPPP), ("phess", FPPPP), ("Hv", FPPP),
        ("mse", C.c_float),
    ]


class Tables:
    """Pointer tables over numpy buffers; keeps every buffer alive."""

    def __init__(self):
        self.keep = []

    def f1(self, a) -> FP:
        a = np.ascontiguousarray(a, np.float32)
        self.keep.append(a)
        return a.ctypes.data_as(FP)

    def _tables(self, base_addr, shape, itemsize, elem_ptr, levels):
        # rows of the innermost dimension, then each outer level
        shape = tuple(shape)
        nrows = int(np.prod(shape[:-1]))
        rows = (elem_ptr * nrows)(*[C.cast(base_addr + r * shape[-1] * itemsize, elem_ptr) for r in range(nrows)])
        self.keep.append(rows)
        cur, cur_type = rows, elem_ptr
        for d in range(len(shape) - 2, 0, -1):
            n = int(np.prod(shape[:d]))
            ptype = C.POINTER(cur_type)
            tab = (ptype * n)(*[C.cast(C.addressof(cur) + i * shape[d] * C.sizeof(cur_type), ptype) for i in range(n)])
            self.keep.append(tab)
            cur, cur_type = tab, ptype
        return C.cast(C.addressof(cur), C.POINTER(cur_type))

    def f3(self, a) -> FPPP:
        a = np.ascontiguousarray(a, np.float32)
        self.keep.append(a)
        return self._tables(a.ctypes.data, a.shape, 4, FP, 3)

    def f2(self, a) -> FPP:
        a = np.ascontiguousarray(a, np.float32)
        self.keep.append(a)
        return self._tables(a.ctypes.data, a.shape, 4, FP, 2)

    def f4(self, a) -> FPPPP:
        assert a.dtype == np.float32 and a.flags.c_contiguous
        self.keep.append(a)
        return self._tables(a.ctypes.data, a.shape, 4, FP, 4)

    def c3(self, a) -> CPPP:
        assert a.dtype == np.complex64 and a.flags.c_contiguous
        self.keep.append(a)
        return self._tables(a.ctypes.data, a.shape, 8, CP, 3)

    def c4(self, a) -> CPPPP:
        assert a.dtype == np.complex64 and a.flags.c_contiguous
        self.keep.append(a)
        return self._tables(a.ctypes.data, a.shape, 8, CP, 4)

    def strings(self, names):
        arr = (C.c_char_p * len(names))(*[n.encode() for n in names])
        self.keep.append(arr)
        return C.cast(arr, C.POINTER(C.c_char_p))


def deg2rad_f32(v) -> np.float32:
    """float(PI * (float)deg / 180.) as src/acq_init_free.c:220-247 converts."""
    return np.float32(3.14159265358979323846264 * float(np.float32(v)) / 180.0)


def omega_f32(freq) -> np.float32:
    """(float)(2 PI f) with f parsed as float (src/emf_init_free.c:23)."""
    return np.float32(2.0 * 3.14159265358979323846264 * float(np.float32(freq)))


class ShotCarriers:
    """emf_t/acq_t of one shot of a survey, filled as emf_init/acq_init do."""

    def __init__(self, s: configs.Survey, shot: int = 0, reciprocity: bool = False):
        t = self.t = Tables()
        e = self.emf = Emf()
        a = self.acq = Acq()
        self.survey = s
        e.mode = 0
        e.verb = 0
        e.reciprocity = 0
        e.nfreq = len(s.freqs)
        self.freqs = np.array([np.float32(f) for f in s.freqs], np.float32)
        self.omegas = np.array([omega_f32(f) for f in s.freqs], np.float32)
        e.freqs = t.f1(self.freqs)
        e.omegas = t.f1(self.omegas)
        e.rho_air = s.rho_air
        e.lextend = s.lextend
        e.nchsrc = len(s.chsrc)
        e.nchrec = len(s.chrec)
        e.chsrc = t.strings(list(s.chsrc))
        e.chrec = t.strings(list(s.chrec))
        e.nx, e.ny, e.nz = s.nx, s.ny, s.nz
        e.dx, e.dy, e.dz = s.dx, s.dy, s.dz
        e.ox, e.oy, e.oz = s.ox, s.oy, s.oz
        e.d1min, e.d2min, e.d3min = s.d1min, s.d2min, s.d3min
        e.n1, e.n2, e.n3 = s.n1, s.n2, s.n3
        e.nb = s.nb
        self.rho11 = np.ascontiguousarray(s.rho11, np.float32)
        self.rho22 = np.ascontiguousarray(s.rho22, np.float32)
        self.rho33 = np.ascontiguousarray(s.rho33, np.float32)
        e.rho11 = t.f3(self.rho11)
        e.rho22 = t.f3(self.rho22)
        e.rho33 = t.f3(self.rho33)
        # rho_water = the global minimum resistivity (src/emf_init_free.c:113-128)
        e.rho_water = float(min(self.rho11.min(), self.rho22.min(), self.rho33.min()))
        if s.bathy is not None:
            self.bathy = np.ascontiguousarray(s.bathy, np.float32)
            e.bathy = t.f2(self.bathy)
        # acquisition of this shot: the source row and the receivers of its
        # table rows (src/acq_init_free.c:144-262)
        src = s.sources[shot]
        recs = [s.receivers[irec - 1] for isrc, irec in s.table if isrc == shot + 1]
        a.nsrc = 1
        a.nrec = len(recs)
        a.nsubsrc = a.nsubrec = 1
        a.nsrc_total = len(s.sources)
        a.nrec_total = len(s.receivers)
        self.src = np.array([[np.float32(src[0])], [np.float32(src[1])], [np.float32(src[2])],
                             [deg2rad_f32(src[3])], [deg2rad_f32(src[4])]], np.float32)
        a.src_x1, a.src_x2, a.src_x3, a.src_azimuth, a.src_dip = [t.f1(self.src[i]) for i in range(5)]
        r = np.array(recs, np.float64).reshape(-1, 5)
        self.rec = np.stack([r[:, 0].astype(np.float32), r[:, 1].astype(np.float32), r[:, 2].astype(np.float32),
                             np.array([deg2rad_f32(v) for v in r[:, 3]], np.float32),
                             np.array([deg2rad_f32(v) for v in r[:, 4]], np.float32)])
        a.rec_x1, a.rec_x2, a.rec_x3, a.rec_azimuth, a.rec_dip = [t.f1(self.rec[i]) for i in range(5)]
        self.shot_idx = np.array([shot + 1], np.int32)
        t.keep.append(self.shot_idx)
        a.shot_idx = self.shot_idx.ctypes.data_as(C.POINTER(C.c_int))
        nch, nf, nr = e.nchrec, e.nfreq, max(a.nrec, 1)
        self.dcal = np.zeros((nch, nf, nr), np.complex64)
        self.dres = np.zeros((nch, nf, nr), np.complex64)
        e.dcal_fd = t.c3(self.dcal)
        e.dres_fd = t.c3(self.dres)
        self.regridded = False

    # -- MG grid views after regrid_init ---------------------------------------
    def grid(self):
        e = self.emf
        n = (e.n1, e.n2, e.n3)
        g = {}
        for a, nn in zip("123", n):
            for k in ("x", "xs", "d", "ds"):
                name = k[0] + a + k[1:]
                g[name] = np.ctypeslib.as_array(getattr(e, name), shape=(nn + 1,)).copy()
        for k in ("sigma11", "sigma22", "sigma33"):
            p = getattr(e, k)
            g[k] = np.ctypeslib.as_array(p[0][0], shape=(n[2] * n[1] * n[0],)).reshape(n[2], n[1], n[0]).copy()
        nb = e.nb
        g["xx"] = np.ctypeslib.as_array(e.xx, shape=(e.nx + 2 * nb + 1,)).copy()
        g["yy"] = np.ctypeslib.as_array(e.yy, shape=(e.ny + 2 * nb + 1,)).copy()
        g["zz"] = np.ctypeslib.as_array(e.zz, shape=(e.nz + 2 * nb + 1,)).copy()
        return g

    def sigma_ptrs(self):
        e = self.emf
        return [C.cast(getattr(e, k)[0][0], C.c_void_p).value for k in ("sigma11", "sigma22", "sigma33")]

    def x_arrays(self):
        e = self.emf
        return [np.ctypeslib.as_array(getattr(e, k), shape=(n + 1,)).copy()
                for k, n in (("x1", e.n1), ("x2", e.n2), ("x3", e.n3))]

    @property
    def nvec(self):
        e = self.emf
        return 3 * (e.n1 + 1) * (e.n2 + 1) * (e.n3 + 1)
