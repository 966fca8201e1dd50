"""ctypes bindings of libemmi_cli.so, the host side of the command-line
program (parameters, survey and data files, data weighting), and the path of
the program emmi_main.  Both are built in-tree by ``build.py``.

The functions keep the reference's names and argument meaning
(``src/cstd.c`` getpar, ``src/emf_init_free.c``, ``src/acq_init_free.c``,
``src/read_write_data.c``, ``src/cal_uncertainty_noise.c``,
``src/frannor.c``); the carriers are those of :mod:`carriers`.
"""
from __future__ import annotations

import ctypes as C
import os

from . import carriers

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libemmi_cli.so")
EXE = os.path.join(HERE, "emmi_main")

_lib = None
_argv_keep = []


def lib() -> C.CDLL:
    """Load libemmi_cli.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} is missing: build it with libemmi_mgfd_amd/build.py")
    L = C.CDLL(LIB)
    P = C.POINTER
    L.initargs.argtypes = [C.c_int, P(C.c_char_p)]
    L.initargs.restype = None
    for name, t in (("getparint", C.c_int), ("getparfloat", C.c_float), ("getpardouble", C.c_double)):
        f = getattr(L, name)
        f.argtypes = [C.c_char_p, P(t)]
        f.restype = C.c_int
    L.getparstring.argtypes = [C.c_char_p, P(C.c_char_p)]
    L.getparstring.restype = C.c_int
    L.getparstringarray.argtypes = [C.c_char_p, P(C.c_char_p)]
    L.getparstringarray.restype = C.c_int
    L.countparval.argtypes = [C.c_char_p]
    L.countparval.restype = C.c_int
    pe, pa = P(carriers.Emf), P(carriers.Acq)
    for name in ("emf_init", "emf_free"):
        getattr(L, name).argtypes = [pe]
        getattr(L, name).restype = None
    L.acq_init.argtypes = [pa, pe]
    L.acq_init.restype = None
    L.acq_free.argtypes = [pa]
    L.acq_free.restype = None
    L.write_data.argtypes = [pa, pe, C.c_char_p, C.c_void_p]
    L.write_data.restype = None
    for name in ("read_data", "write_rmse"):
        getattr(L, name).argtypes = [pa, pe]
        getattr(L, name).restype = None
    L.frannor.argtypes = []
    L.frannor.restype = C.c_float
    L.srannor.argtypes = [C.c_int]
    L.srannor.restype = None
    L.emmi_rannor_reset.argtypes = []
    L.emmi_rannor_reset.restype = None
    for name in ("cal_uncertainty_noise", "emmi_data_misfit"):
        getattr(L, name).argtypes = [pa, pe]
        getattr(L, name).restype = C.c_float
    _lib = L
    return L


def set_args(args) -> None:
    """initargs over ["emmi_main", *args] (the strings are kept alive)."""
    L = lib()
    words = [b"emmi_main"] + [a.encode() if isinstance(a, str) else a for a in args]
    arr = (C.c_char_p * (len(words) + 1))(*words, None)
    _argv_keep[:] = [arr, words]
    L.initargs(len(words), arr)


def set_rank(rank: int, nranks: int) -> None:
    L = lib()
    C.c_int.in_dll(L, "iproc").value = rank
    C.c_int.in_dll(L, "nproc").value = nranks
