"""ctypes binding of libemmi_mgfd.so (the HIP library of this package).

The product path is the native library; nothing here computes on the CPU.
Loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libemmi_mgfd.so")
_LIB = None

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
c128p = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")


class Params(C.Structure):
    _fields_ = [("cycleopt", C.c_int), ("ncycle", C.c_int), ("v1", C.c_int), ("v2", C.c_int),
                ("isemicoarsen", C.c_int), ("tol", C.c_double), ("precision", C.c_int),
                ("coarse_sweeps", C.c_int), ("verb", C.c_int), ("graphs", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("cycles", C.c_int), ("status", C.c_int), ("rnorm0", C.c_double), ("rnorm", C.c_double),
                ("x0_rejected", C.c_int), ("x0_rnorm", C.c_double)]


STATUS = {0: "converged", 1: "not converged", 2: "non-finite residual", -1: "error"}


class NativeError(RuntimeError):
    pass


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise NativeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (or python -m libemmi_mgfd_amd.build)")
        L = C.CDLL(LIB_PATH)
        L.emmi_default_params.argtypes = [C.POINTER(Params)]
        L.emmi_last_error.restype = C.c_char_p
        L.emmi_solver_create.restype = C.c_void_p
        L.emmi_solver_create.argtypes = [C.c_int] * 3 + [C.c_void_p] * 6 + [C.c_int, C.c_float, C.POINTER(Params)]
        L.emmi_solver_solve.restype = C.c_int
        L.emmi_solver_solve.argtypes = [C.c_void_p] * 5 + [C.POINTER(Stats)]
        L.emmi_solver_set_omega.restype = C.c_int
        L.emmi_solver_set_omega.argtypes = [C.c_void_p, C.c_float]
        L.emmi_solver_create_batch.restype = C.c_void_p
        L.emmi_solver_create_batch.argtypes = ([C.c_int] * 3 + [C.c_void_p] * 6 +
                                               [C.c_int, C.c_void_p, C.c_int, C.POINTER(Params)])
        L.emmi_solver_create_survey.restype = C.c_void_p
        L.emmi_solver_create_survey.argtypes = ([C.c_int] * 4 + [C.c_void_p] * 6 +
                                                [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(Params)])
        L.emmi_solver_solve_batch.restype = C.c_int
        L.emmi_solver_solve_batch.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 4 + [C.POINTER(Stats)]
        L.emmi_solver_set_omega_slot.restype = C.c_int
        L.emmi_solver_set_omega_slot.argtypes = [C.c_void_p, C.c_int, C.c_float]
        L.emmi_solver_history_slot.restype = C.c_int
        L.emmi_solver_history_slot.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_int]
        L.emmi_solver_set_ncycle.restype = C.c_int
        L.emmi_solver_set_ncycle.argtypes = [C.c_void_p, C.c_int]
        L.emmi_solver_set_first_cycle.restype = C.c_int
        L.emmi_solver_set_first_cycle.argtypes = [C.c_void_p, C.c_int]
        L.emmi_solver_blend_guess.restype = C.c_int
        L.emmi_solver_blend_guess.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.POINTER(C.c_double)]
        L.emmi_solver_autotune.restype = C.c_int
        L.emmi_solver_autotune.argtypes = [C.c_void_p, C.c_int]
        L.emmi_line_tuning.restype = C.c_int
        L.emmi_line_tuning.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_int]
        L.emmi_line_tuning_clear.restype = None
        L.emmi_line_tuning_clear.argtypes = []
        L.emmi_solver_slots.restype = C.c_int
        L.emmi_solver_slots.argtypes = [C.c_void_p]
        from . import carriers as K
        PE, PA, PF = C.POINTER(K.Emf), C.POINTER(K.Acq), C.POINTER(K.Fwi)
        L.regrid_init.argtypes = [PA, PE]
        L.regrid_free.argtypes = [PE]
        L.inject_source.argtypes = [PA, PE, C.c_void_p, C.c_int]
        L.inject_adj_source.argtypes = [PA, PE, C.c_void_p, C.c_void_p, C.c_int]
        L.extract_emf_data.argtypes = [PA, PE, C.c_void_p, C.c_void_p, C.c_int]
        L.extract_electric_field.argtypes = [PA, PE, C.c_void_p, C.c_int]
        L.extract_magnetic_field.argtypes = [PA, PE, C.c_void_p, C.c_int]
        L.build_gradient.argtypes = [PA, PE, PF] + [K.CPPPP] * 6
        L.gmg_init.argtypes = [PE, C.c_int]
        L.emmi_inject_source_dev.restype = C.c_int
        L.emmi_inject_source_dev.argtypes = [PA, PE, C.c_int, C.c_void_p, C.c_void_p]
        L.emmi_extract_emf_data_dev.restype = C.c_int
        L.emmi_extract_emf_data_dev.argtypes = [PA, PE, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.emmi_path_last_error.restype = C.c_char_p
        L.emmi_inject_adj_source_dev.restype = C.c_int
        L.emmi_inject_adj_source_dev.argtypes = [PA, PE, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.emmi_inject_adj_source_dev_noalloc.restype = C.c_int
        L.emmi_inject_adj_source_dev_noalloc.argtypes = [PA, PE, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.emmi_extract_electric_field_dev.restype = C.c_int
        L.emmi_extract_electric_field_dev.argtypes = [PA, PE, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.emmi_extract_magnetic_field_dev.restype = C.c_int
        L.emmi_extract_magnetic_field_dev.argtypes = [PA, PE, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.emmi_build_gradient_dev.restype = C.c_int
        L.emmi_build_gradient_dev.argtypes = [PA, PE, C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                              C.c_void_p, C.c_void_p, C.c_void_p]
        L.gmg_apply.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
        L.gmg_free.restype = None
        L.gmg_free.argtypes = []
        L.find_index.restype = C.c_int
        L.find_index.argtypes = [C.c_int, C.c_void_p, C.c_float]
        L.create_nugrid.restype = C.c_float
        L.create_nugrid.argtypes = [C.c_int, C.c_float, C.c_float, C.c_void_p]
        L.emmi_solver_history.restype = C.c_int
        L.emmi_solver_history.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int]
        L.emmi_solver_destroy.argtypes = [C.c_void_p]
        L.emmi_solver_levels.restype = C.c_int
        L.emmi_solver_levels.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.emmi_solver_dims.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int)]
        L.emmi_solver_check_op.restype = C.c_int
        L.emmi_solver_check_op.argtypes = [C.c_void_p] + [C.c_int] * 5 + [C.c_void_p] * 3
        L.emmi_device_alloc.restype = C.c_void_p
        L.emmi_device_alloc.argtypes = [C.c_size_t]
        L.emmi_device_free.argtypes = [C.c_void_p]
        L.emmi_memcpy_h2d.restype = C.c_int
        L.emmi_memcpy_h2d.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.emmi_memcpy_d2h.restype = C.c_int
        L.emmi_memcpy_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.emmi_memcpy_d2d.restype = C.c_int
        L.emmi_memcpy_d2d.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.emmi_device_sync.restype = C.c_int
        L.emmi_set_line_kernel.restype = C.c_int
        L.emmi_set_line_kernel.argtypes = [C.c_int]
        L.emmi_timing_enable.argtypes = [C.c_int]
        L.emmi_solver_kernel_rates.restype = C.c_int
        L.emmi_solver_kernel_rates.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                               C.c_int]
        L.emmi_timing_read.restype = C.c_int
        L.emmi_timing_read.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_long), C.POINTER(C.c_double), C.c_int]
        L.emmi_timing_read_layout.restype = C.c_int
        L.emmi_timing_read_layout.argtypes = [C.POINTER(C.c_double), C.c_int]
        _LIB = L
    return _LIB


def last_error() -> str:
    return lib().emmi_last_error().decode()


def params(**kw) -> Params:
    p = Params()
    lib().emmi_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


class DeviceBuffer:
    """A block of device memory owned by this object."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = lib().emmi_device_alloc(self.nbytes)
        if not self.ptr:
            raise NativeError(last_error())

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        if lib().emmi_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes):
            raise NativeError("host to device copy failed")
        return self

    def download(self, dtype, shape=None) -> np.ndarray:
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        if lib().emmi_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes):
            raise NativeError("device to host copy failed")
        return out.reshape(shape) if shape is not None else out

    def free(self):
        if self.ptr:
            lib().emmi_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Solver:
    """A multigrid solver for one (MG grid, medium) on the GPU with one
    right-hand-side slot per frequency in `omega` (a number or a sequence)."""

    def __init__(self, n, x1, x2, x3, sig11, sig22, sig33, omega, sig_on_device=False, **kw):
        self.n = tuple(int(v) for v in n)
        self.p = params(**kw)
        self._keep = [np.ascontiguousarray(a, np.float32) for a in (x1, x2, x3)]
        if sig_on_device:
            ptrs = [int(s) for s in (sig11, sig22, sig33)]
        else:
            sig = [np.ascontiguousarray(a, np.float32) for a in (sig11, sig22, sig33)]
            self._keep += sig
            ptrs = [a.ctypes.data for a in sig]
        om = np.atleast_1d(np.asarray(omega, np.float32)).copy()
        self._keep.append(om)
        self.nslots = int(om.size)
        self.h = lib().emmi_solver_create_batch(self.n[0], self.n[1], self.n[2],
                                                self._keep[0].ctypes.data, self._keep[1].ctypes.data,
                                                self._keep[2].ctypes.data, ptrs[0], ptrs[1], ptrs[2],
                                                1 if sig_on_device else 0, om.ctypes.data, self.nslots,
                                                C.byref(self.p))
        if not self.h:
            raise NativeError(last_error())

    @classmethod
    def survey(cls, n, grids, slot_med, omegas, sig_on_device=False, **kw):
        """A solver over several grids and media of one multigrid size:
        grids[m] = (x1, x2, x3, sig11, sig22, sig33) of grid m (host arrays,
        or device pointers for the conductivities with sig_on_device); slot k
        solves on grid slot_med[k] at omegas[k]."""
        self = cls.__new__(cls)
        self.n = tuple(int(v) for v in n)
        self.p = params(**kw)
        self._keep = []
        nmed = len(grids)
        tabs = [(C.c_void_p * nmed)() for _ in range(6)]
        for m, g in enumerate(grids):
            for q in range(3):
                a = np.ascontiguousarray(g[q], np.float32)
                self._keep.append(a)
                tabs[q][m] = a.ctypes.data
            for q in range(3, 6):
                if sig_on_device:
                    tabs[q][m] = int(g[q])
                else:
                    a = np.ascontiguousarray(g[q], np.float32)
                    self._keep.append(a)
                    tabs[q][m] = a.ctypes.data
        med = np.ascontiguousarray(slot_med, np.int32)
        om = np.ascontiguousarray(omegas, np.float32)
        self._keep += [med, om] + tabs
        self.nslots = int(om.size)
        self.h = lib().emmi_solver_create_survey(self.n[0], self.n[1], self.n[2], nmed,
                                                 *[C.cast(t, C.c_void_p) for t in tabs],
                                                 1 if sig_on_device else 0, med.ctypes.data, om.ctypes.data,
                                                 self.nslots, C.byref(self.p))
        if not self.h:
            raise NativeError(last_error())
        return self

    @property
    def nvec(self) -> int:
        n1, n2, n3 = self.n
        return 3 * (n1 + 1) * (n2 + 1) * (n3 + 1)

    def solve_device(self, b_ptr, x_ptr, h_ptr, x0_ptr=None) -> Stats:
        st = Stats()
        rc = lib().emmi_solver_solve(self.h, b_ptr, x_ptr, h_ptr, x0_ptr, C.byref(st))
        if rc < 0:
            raise NativeError(last_error())
        return st

    def solve(self, b: np.ndarray, x0: np.ndarray | None = None):
        """Host convenience: returns (E, H, stats)."""
        nb = 16 * self.nvec
        db = DeviceBuffer(nb).upload(np.ascontiguousarray(b, np.complex128).reshape(-1))
        dx = DeviceBuffer(nb)
        d0 = DeviceBuffer(nb).upload(np.ascontiguousarray(x0, np.complex128).reshape(-1)) if x0 is not None else None
        st = self.solve_device(db.ptr, dx.ptr, db.ptr, d0.ptr if d0 else None)
        return dx.download(np.complex128), db.download(np.complex128), st

    def solve_batch_device(self, b_ptrs, x_ptrs, h_ptrs, x0_ptrs=None):
        """Solves slots 0..len(b_ptrs)-1 together; returns one Stats per slot."""
        n = len(b_ptrs)
        P = C.c_void_p * n
        st = (Stats * n)()
        x0 = P(*x0_ptrs) if x0_ptrs is not None else None
        rc = lib().emmi_solver_solve_batch(self.h, n, P(*b_ptrs), P(*x_ptrs), P(*h_ptrs), x0, st)
        if rc < 0:
            raise NativeError(last_error())
        return list(st)

    def set_ncycle(self, ncycle):
        if lib().emmi_solver_set_ncycle(self.h, int(ncycle)):
            raise NativeError(last_error())

    def autotune(self, reps=5):
        """Times the line-sweep kernels on every launch shape of this solver
        not yet in the process-wide table and records the fastest
        (emmi_solver_autotune); returns the number of shapes tuned."""
        n = lib().emmi_solver_autotune(self.h, int(reps))
        if n < 0:
            raise NativeError(last_error())
        return n

    def blend_guess(self, b_ptrs, x1_ptrs, xg_ptrs):
        """xg <- x1 + beta (xg - x1) per slot with the beta that minimises the
        residual (emmi_solver_blend_guess); returns the betas (complex)."""
        n = len(b_ptrs)
        P = C.c_void_p * n
        beta = (C.c_double * (2 * max(n, 1)))()
        if lib().emmi_solver_blend_guess(self.h, n, P(*b_ptrs), P(*x1_ptrs), P(*xg_ptrs), beta):
            raise NativeError(last_error())
        return [complex(beta[2 * k], beta[2 * k + 1]) for k in range(n)]

    def set_first_cycle(self, first):
        """Cycle k of the following solves uses pattern (first + k) % 3."""
        if lib().emmi_solver_set_first_cycle(self.h, int(first)):
            raise NativeError(last_error())

    def set_omega(self, omega, slot=0):
        if lib().emmi_solver_set_omega_slot(self.h, int(slot), np.float32(omega)):
            raise NativeError(last_error())

    def history(self, slot=0) -> np.ndarray:
        buf = (C.c_double * 4096)()
        n = lib().emmi_solver_history_slot(self.h, int(slot), buf, 4096)
        return np.array(buf[:min(n, 4096)])

    KERNEL_RATES = ["residual", "restriction", "prolongation", "line_sweep_x", "line_sweep_y", "line_sweep_z"]

    def kernel_rates(self, reps=20):
        """{kernel: (ms per launch, compulsory bytes)} of the level-0 kernels
        of slot 0 (emmi_solver_kernel_rates)."""
        ms = (C.c_double * 8)()
        by = (C.c_double * 8)()
        n = lib().emmi_solver_kernel_rates(self.h, int(reps), ms, by, 8)
        if n < 0:
            raise NativeError(last_error())
        return {k: (ms[i], by[i]) for i, k in enumerate(self.KERNEL_RATES[:n])}

    def levels(self):
        npat = C.c_int()
        lmax = lib().emmi_solver_levels(self.h, C.byref(npat))
        return lmax, npat.value

    def dims(self, pattern, lev):
        d = (C.c_int * 3)()
        lib().emmi_solver_dims(self.h, pattern, lev, d)
        return tuple(d)

    def check_op(self, op, pattern, lev, axis, it, in1, in2=None, out_len=None):
        """Run one kernel on host vectors (see emmi_solver_check_op)."""
        def nv(d):
            return 3 * (d[0] + 1) * (d[1] + 1) * (d[2] + 1)
        lmax, npat = self.levels()
        if not (0 <= pattern < npat and 0 <= lev < lmax):
            raise ValueError("no such level")
        nf = nv(self.dims(pattern, lev))
        nc = nv(self.dims(pattern, lev + 1)) if op in (2, 3) and lev + 1 < lmax else None
        in1 = np.ascontiguousarray(in1, np.complex128).reshape(-1)
        in2 = None if in2 is None else np.ascontiguousarray(in2, np.complex128).reshape(-1)
        want = {0: (nf, nf, nf), 1: (nf, nf, nf), 2: (nf, None, nc), 3: (nc, nf, nf), 4: (nf, None, nf), 5: (nf, nf, nf),
                6: (nf, nf, nf)}[op]
        if want[0] is None or in1.size != want[0] or (want[1] is not None and (in2 is None or in2.size != want[1])):
            raise ValueError("vector sizes do not match the level")
        out = np.zeros(want[2], np.complex128)
        rc = lib().emmi_solver_check_op(self.h, op, pattern, lev, axis, it, in1.ctypes.data,
                                        None if in2 is None else in2.ctypes.data, out.ctypes.data)
        if rc != 0:
            raise NativeError(last_error())
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().emmi_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def timing_enable(on=True):
    lib().emmi_timing_enable(1 if on else 0)


LINE_CHOICES = ["one_lane", "quad_fused", "quad_separate"]


def line_tuning():
    """The autotuner's table: one dict per launch shape."""
    n = lib().emmi_line_tuning(None, None, 0)
    rows = (C.c_int * (6 * max(n, 1)))()
    us = (C.c_double * (2 * max(n, 1)))()
    n = min(n, lib().emmi_line_tuning(rows, us, n))
    out = []
    for k in range(n):
        r = rows[6 * k:6 * k + 6]
        out.append({"esize": r[0], "axis": r[1], "n": r[2], "lines": r[3], "jobs": r[4],
                    "choice": LINE_CHOICES[r[5]], "us": us[2 * k], "us_rule": us[2 * k + 1]})
    return out


def line_tuning_clear():
    lib().emmi_line_tuning_clear()


TIMING_CLASSES = ["residual", "line_solve", "restriction", "prolongation", "other", "line_rhs", "conv", "conv_dw"]


def timing_read():
    """{class: (total ms, launches, compulsory bytes, layout bytes)} since
    timing_enable: compulsory bytes read each input once and write each
    output once at the launch's level shape; layout bytes add what the
    kernels' own data layout moves (the line kernels' stored records).  For
    the network's convolutions (conv: forward and input gradient; conv_dw:
    weight gradient) the byte slots hold algorithmic flops."""
    ms = (C.c_double * 8)()
    cnt = (C.c_long * 8)()
    by = (C.c_double * 8)()
    lay = (C.c_double * 8)()
    n = lib().emmi_timing_read(ms, cnt, by, 8)
    lib().emmi_timing_read_layout(lay, 8)
    return {TIMING_CLASSES[i]: (ms[i], cnt[i], by[i], lay[i]) for i in range(min(n, len(TIMING_CLASSES)))}
