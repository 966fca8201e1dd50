"""Host side of the command-line program (libemmi_cli.so) against the
reference: getpar lookups, emf_init/acq_init, the data files, the noise
generator, the data uncertainty and the misfit, bit for bit.

Pins, in order of independence from this build:
  * the reference's own example data (run_inversion/emf_0001.txt and its
    noisy copy obs_0001.txt, which the reference's run wrote with
    src/cal_uncertainty_noise.c; the target rmse 1.021664e+00 its log
    printed): the whole read -> noise -> write chain, byte for byte;
  * fixtures written by the reference's own host sources compiled in place
    (tests/golden/make_cli_fixtures.py, oracle/cli_runner.c) for inputs that
    take every branch (6 channels, varied azimuths, delta_phi > 0, shots=,
    reciprocity, parameter files);
  * the normal distribution itself for frannor (moments and a
    Kolmogorov-Smirnov test), independent of any fixture.
"""
from __future__ import annotations

import ctypes as C
import gzip
import json
import os
import subprocess

import numpy as np
import pytest

from helpers import CLI_CHANNELS, CLI_FREQS, cli_args, cli_write_inputs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def L():
    from libemmi_mgfd_amd import cli
    return cli.lib()


@pytest.fixture(scope="module")
def fx():
    with open(os.path.join(GOLD, "cli_fixtures.json")) as fp:
        return json.load(fp)


def f32(h: str) -> np.float32:
    return np.float32(float.fromhex(h))


# ---------------------------------------------------------------------------
# frannor
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("key,seed", [("fresh", None), ("seed7", 7), ("seed-12345", -12345)])
def test_frannor_matches_reference_sequence(L, fx, key, seed):
    want = np.array([f32(h) for h in fx["rannor"][key]], np.float32)
    L.emmi_rannor_reset()
    if seed is not None:
        L.srannor(seed)
    got = np.array([L.frannor() for _ in range(len(want))], np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), int(np.argmax(got != want))
    if key == "fresh":
        # the sequence reaches the wedge and the tail of the method
        assert np.sum(np.abs(want) > 2.776994) > 5


def test_frannor_is_standard_normal(L):
    from scipy import stats
    L.emmi_rannor_reset()
    L.srannor(2024)
    v = np.array([L.frannor() for _ in range(100000)], np.float64)
    assert abs(v.mean()) < 0.015
    assert abs(v.var() - 1.0) < 0.02
    assert stats.kstest(v, "norm").pvalue > 1e-3
    L.emmi_rannor_reset()


# ---------------------------------------------------------------------------
# getpar
# ---------------------------------------------------------------------------
def _lookup(L, key: str, typ: str):
    k = key.encode()
    n = L.countparval(k)
    if typ == "i":
        buf = (C.c_int * 64)()
        r = L.getparint(k, buf)
        return r, n, [buf[i] for i in range(r)]
    if typ in "fd":
        buf = ((C.c_float if typ == "f" else C.c_double) * 64)()
        r = (L.getparfloat if typ == "f" else L.getpardouble)(k, buf)
        return r, n, [buf[i] for i in range(r)]
    if typ == "s":
        p = C.c_char_p()
        r = L.getparstring(k, C.byref(p))
        return r, n, [p.value.decode()] if r else [""]
    buf = (C.c_char_p * 64)()
    r = L.getparstringarray(k, buf)
    return r, n, [buf[i].decode() for i in range(r)]


def test_getpar_matches_reference(L, fx, tmp_path):
    from libemmi_mgfd_amd import cli
    for case in fx["pars"]:
        words = list(case["words"])
        if case["parfile"] is not None:
            p = tmp_path / "p.par"
            p.write_text(case["parfile"])
            words = ["par=" + str(p)] + words
        cli.set_args(["task=pars", f"keys={case['keys']}", f"types={case['types']}"] + words)
        for line in case["lines"]:
            t = line.split(" ")
            key, typ, r, n = t[0], t[1], int(t[2]), int(t[3])
            gr, gn, gv = _lookup(L, key, typ)
            assert (gr, gn) == (r, n), (case, line)
            if typ == "i":
                assert gv == [int(x) for x in t[4:]], line
            elif typ == "f":
                assert [np.float32(x) for x in gv] == [f32(x) for x in t[4:]], line
            elif typ == "d":
                assert gv == [float.fromhex(x) for x in t[4:]], line
            else:
                want = line.split(" ", 4)[4] if len(t) > 4 else ""
                assert " ".join(f"[{v}]" for v in gv) == want, line


# ---------------------------------------------------------------------------
# emf_init / acq_init
# ---------------------------------------------------------------------------
def _init(L, tmp_path, monkeypatch, extra):
    from libemmi_mgfd_amd import carriers, cli
    files = cli_write_inputs(str(tmp_path))
    monkeypatch.chdir(tmp_path)
    rank = nranks = None
    words = []
    for w in extra:
        k, v = w.split("=", 1)
        if k == "rank":
            rank = int(v)
        elif k == "nranks":
            nranks = int(v)
        else:
            words.append(w)
    cli.set_args(cli_args({k: os.path.basename(v) for k, v in files.items()}, words))
    cli.set_rank(rank, nranks)
    emf, acq = carriers.Emf(), carriers.Acq()
    L.emf_init(C.byref(emf))
    L.acq_init(C.byref(acq), C.byref(emf))
    return emf, acq, rank


def _farr(p, n):
    return np.array([p[i] for i in range(n)], np.float32)


@pytest.mark.parametrize("case", ["rank0", "rank1of3", "shots", "reciprocity"])
def test_emf_acq_init_match_reference(L, fx, tmp_path, monkeypatch, case):
    from libemmi_mgfd_amd import cli
    extra = {"rank0": ["rank=0", "nranks=1"], "rank1of3": ["rank=1", "nranks=3"],
             "shots": ["rank=1", "nranks=2", "shots=3,1"],
             "reciprocity": ["rank=0", "nranks=1", "reciprocity=1", "shots=1"]}[case]
    emf, acq, rank = _init(L, tmp_path, monkeypatch, extra)
    try:
        want = {ln.split(" ", 1)[0]: ln.split(" ")[1:] for ln in fx["init"][case]}
        assert emf.nfreq == int(want["nfreq"][0])
        for name in ("freqs", "omegas"):
            got = _farr(getattr(emf, name), emf.nfreq)
            assert np.array_equal(got, [f32(h) for h in want[name][1:]]), name
        assert [emf.chsrc[i].decode() for i in range(emf.nchsrc)] == want["chsrc"][1:]
        assert [emf.chrec[i].decode() for i in range(emf.nchrec)] == want["chrec"][1:]
        ints = [emf.reciprocity, emf.nx, emf.ny, emf.nz, emf.n1, emf.n2, emf.n3, emf.nb]
        assert ints == [int(x) for x in want["ints"]]
        fl = [emf.dx, emf.dy, emf.dz, emf.ox, emf.oy, emf.oz, emf.d1min, emf.d2min, emf.d3min, emf.lextend,
              emf.rho_air, emf.rho_water]
        assert [np.float32(x) for x in fl] == [f32(h) for h in want["floats"][:12]]
        ai = [acq.nsrc, acq.nrec, acq.nsubsrc, acq.nsubrec, acq.nsrc_total, acq.nrec_total, acq.shot_idx[rank]]
        assert ai == [int(x) for x in want["acq_ints"]]
        af = [acq.lensrc, acq.lenrec, acq.x1min, acq.x1max, acq.x2min, acq.x2max, acq.x3min, acq.x3max]
        assert [np.float32(x) for x in af] == [f32(h) for h in want["acq_floats"]]
        for name in ("src_x1", "src_x2", "src_x3", "src_azimuth", "src_dip"):
            assert np.array_equal(_farr(getattr(acq, name), acq.nsrc), [f32(h) for h in want[name][1:]]), name
        for name in ("rec_x1", "rec_x2", "rec_x3", "rec_azimuth", "rec_dip"):
            assert np.array_equal(_farr(getattr(acq, name), acq.nrec), [f32(h) for h in want[name][1:]]), name
    finally:
        L.acq_free(C.byref(acq))
        L.emf_free(C.byref(emf))
        cli.set_rank(0, 1)


# ---------------------------------------------------------------------------
# read_data / cal_uncertainty_noise / write_data / misfit / write_rmse
# ---------------------------------------------------------------------------
NOISE_EXTRA = {"shot1": ["rank=0", "nranks=1", "delta_phi=0.05", "amp_perc=0.05"],
               "shot2": ["rank=1", "nranks=2", "delta_phi=0.02"],
               "shot3_nonoise": ["rank=0", "nranks=1", "shots=3", "addnoise=0", "noisefloorE=1e-13"]}


@pytest.mark.parametrize("case", sorted(NOISE_EXTRA))
def test_noise_and_misfit_match_reference(L, fx, tmp_path, monkeypatch, case):
    from libemmi_mgfd_amd import carriers, cli
    extra = NOISE_EXTRA[case]
    emf, acq, rank = _init(L, tmp_path, monkeypatch, extra)
    pars = dict(w.split("=", 1) for w in extra)
    try:
        emf.addnoise = int(pars.get("addnoise", 1))
        emf.amp_perc = float(pars.get("amp_perc", 0.03))
        emf.delta_phi = float(pars.get("delta_phi", 0.0))
        emf.noisefloorE = float(pars.get("noisefloorE", 1e-15))
        emf.noisefloorH = float(pars.get("noisefloorH", 1e-13))
        shape = (emf.nchrec, emf.nfreq, acq.nrec)
        T = carriers.Tables()
        dobs = np.zeros(shape, np.complex64)
        dcal = np.zeros(shape, np.complex64)
        dres = np.zeros(shape, np.complex64)
        delta = np.zeros(shape, np.float32)
        rmse = np.zeros(shape, np.float32)
        emf.dobs_fd, emf.dcal_fd, emf.dres_fd = T.c3(dobs), T.c3(dcal), T.c3(dres)
        emf.delta_emf, emf.rmse = T.f3(delta), T.f3(rmse)
        L.read_data(C.byref(acq), C.byref(emf))
        dcal[...] = dobs
        L.emmi_rannor_reset()
        mse = np.float32(L.cal_uncertainty_noise(C.byref(acq), C.byref(emf)))
        shot = acq.shot_idx[rank]
        L.write_data(C.byref(acq), C.byref(emf), f"obs_{shot:04d}.txt".encode(), emf.dobs_fd)
        fcost = np.float32(L.emmi_data_misfit(C.byref(acq), C.byref(emf)))
        L.write_rmse(C.byref(acq), C.byref(emf))

        want = fx["noise"][case]
        lines = {ln.split(" ", 1)[0]: ln.split(" ")[1:] for ln in want["lines"]}
        assert mse == f32(lines["mse"][0])
        assert np.array_equal(delta.reshape(-1), [f32(h) for h in lines["delta"][1:]])
        assert (tmp_path / f"obs_{shot:04d}.txt").read_text() == want["obs"]
        # the Ez/Hz channels have no uncertainty (delta 0), so their residuals
        # and the misfit are NaN in the reference too: compare bit patterns
        assert np.float32(fcost).view(np.uint32) == f32(lines["fcost"][0]).view(np.uint32)
        got = dres.view(np.float32).reshape(-1)
        assert np.array_equal(got.view(np.uint32), np.array([f32(h) for h in lines["dres"][1:]]).view(np.uint32))
        assert (tmp_path / f"rmse_{shot:04d}.txt").read_text() == want["rmse"]
        assert set(CLI_CHANNELS) == {emf.chrec[i].decode() for i in range(emf.nchrec)}
        assert emf.nfreq == len(CLI_FREQS)
    finally:
        emf.dobs_fd = emf.dcal_fd = emf.dres_fd = None
        emf.delta_emf = emf.rmse = None
        L.acq_free(C.byref(acq))
        L.emf_free(C.byref(emf))
        cli.set_rank(0, 1)


def test_reference_example_obs_file_reproduced(L, tmp_path, monkeypatch):
    """run_inversion/ of the reference: shot 1, chrec=Ex, freqs=0.25,1,2.75,
    810 receivers and the source all at azimuth 30 deg, default uncertainty
    parameters; its obs_0001.txt and the printed target rmse."""
    from libemmi_mgfd_amd import carriers, cli
    d = os.path.join(GOLD, "run_inversion_0001")
    with gzip.open(os.path.join(d, "emf_0001_Ex.txt.gz"), "rt") as fp:
        (tmp_path / "emf_0001.txt").write_text(fp.read())
    with gzip.open(os.path.join(d, "obs_0001.txt.gz"), "rt") as fp:
        want = fp.read()
    monkeypatch.chdir(tmp_path)
    cli.set_rank(0, 1)
    T = carriers.Tables()
    nrec, freqs = 810, [0.25, 1.0, 2.75]
    emf, acq = carriers.Emf(), carriers.Acq()
    emf.nfreq = 3
    emf.freqs = T.f1(np.array(freqs, np.float32))
    emf.nchrec = 1
    emf.chrec = T.strings(["Ex"])
    emf.addnoise, emf.amp_perc, emf.delta_phi = 1, 0.03, 0.0
    emf.noisefloorE, emf.noisefloorH = 1e-15, 1e-13
    shot = (C.c_int * 1)(1)
    acq.shot_idx = C.cast(shot, C.POINTER(C.c_int))
    acq.nrec = nrec
    az = carriers.deg2rad_f32(30.0)
    acq.rec_azimuth = T.f1(np.full(nrec, az, np.float32))
    acq.src_azimuth = T.f1(np.array([az], np.float32))
    dobs = np.zeros((1, 3, nrec), np.complex64)
    delta = np.zeros((1, 3, nrec), np.float32)
    emf.dobs_fd = T.c3(dobs)
    emf.delta_emf = T.f3(delta)
    L.read_data(C.byref(acq), C.byref(emf))
    L.emmi_rannor_reset()
    mse = L.cal_uncertainty_noise(C.byref(acq), C.byref(emf))
    L.write_data(C.byref(acq), C.byref(emf), b"obs_0001.txt", emf.dobs_fd)
    assert (tmp_path / "obs_0001.txt").read_text() == want
    assert "%e" % np.sqrt(np.float32(mse) / (3 * nrec)) == "1.021664e+00"


# ---------------------------------------------------------------------------
# the executable's own checks (no GPU work before them)
# ---------------------------------------------------------------------------
def _exe(args, cwd):
    from libemmi_mgfd_amd import cli
    if not os.path.exists(cli.EXE):
        pytest.skip("emmi_main is not built")
    return subprocess.run([cli.EXE] + args, capture_output=True, text=True, cwd=cwd, timeout=60)


def test_emmi_main_refuses_gauss_newton(tmp_path):
    r = _exe(["mode=1", "method=1", "freqs=1"], str(tmp_path))
    assert r.returncode == 1 and "method=1" in r.stderr


def test_emmi_main_reports_missing_parameters(tmp_path):
    r = _exe(["mode=0", "nx=4"], str(tmp_path))
    assert r.returncode == 1 and "Need freqs= vector" in r.stderr
    files = cli_write_inputs(str(tmp_path))
    args = [w for w in cli_args({k: os.path.basename(v) for k, v in files.items()}) if not w.startswith("fsrc=")]
    r = _exe(["mode=0"] + args, str(tmp_path))
    assert r.returncode == 1 and "Need fsrc=" in r.stderr
    r = _exe(["mode=0", "shots=1,9"] + cli_args({k: os.path.basename(v) for k, v in files.items()}), str(tmp_path))
    assert r.returncode == 1 and "shots[1]=9 out of range [1,3]" in r.stderr


def test_every_declared_cli_function_is_exported():
    import re
    from libemmi_mgfd_amd import cli
    text = open(os.path.join(ROOT, "include", "emmi_cli.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    decl = set(re.findall(r"^[ \t]*(?:const[ \t]+)?[A-Za-z_]\w*[ \t\*]+\**[ \t]*([A-Za-z_]\w*)[ \t]*\(", text, re.M))
    out = subprocess.run(["nm", "-D", "--defined-only", cli.LIB], capture_output=True, text=True, check=True)
    syms = {ln.split()[-1]: ln.split()[-2] for ln in out.stdout.splitlines()}
    assert len(decl) >= 15
    assert not [n for n in decl if syms.get(n) != "T"], decl
    assert syms.get("iproc") in ("B", "D") and syms.get("nproc") in ("B", "D")


# ---------------------------------------------------------------------------
# regularisation, smoothing, preconditioning, fg_mod_reg
# ---------------------------------------------------------------------------
def _hexarr(tokens):
    return np.array([f32(h) for h in tokens], np.float32)


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def _model_side_lib(L):
    from libemmi_mgfd_amd import carriers
    P = C.POINTER
    fp = P(C.c_float)
    L.regularization_tikhonov.argtypes = [fp, fp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float]
    L.regularization_tikhonov.restype = C.c_float
    L.regularization_tv.argtypes = L.regularization_tikhonov.argtypes
    L.regularization_tv.restype = C.c_float
    L.triangle_smoothing.argtypes = [C.c_void_p] + [C.c_int] * 7
    L.triangle_smoothing.restype = None
    L.precondition.argtypes = [P(carriers.Emf), P(carriers.Fwi), fp]
    L.precondition.restype = None
    L.fg_mod_reg.argtypes = [P(carriers.Acq), P(carriers.Emf), P(carriers.Fwi), fp, fp]
    L.fg_mod_reg.restype = None


def test_model_side_matches_reference(L, fx):
    from libemmi_mgfd_amd import carriers
    _model_side_lib(L)
    ms = fx["model_side"]
    n1, n2, n3 = ms["shape"]
    nxyz = n1 * n2 * n3
    inp = {k: _hexarr(v) for k, v in ms["inputs"].items()}
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    for case in ms["cases"]:
        pars = dict(a.split("=", 1) for a in case["args"])
        want = {ln.split(" ", 1)[0]: ln.split(" ")[1:] for ln in case["lines"]}
        task = case["task"]
        if task == "reg":
            x = inp["x1"].copy()
            g = np.zeros(nxyz, np.float32)
            f = np.float32(L.regularization_tikhonov(fp(x), fp(g), n1, n2, n3, 1, 1, 1))
            assert _bits_equal(f, f32(want["tik"][0])) and _bits_equal(g, _hexarr(want["gtik"][1:]))
            f = np.float32(L.regularization_tv(fp(x), fp(g), n1, n2, n3, 1, 1, 1))
            assert _bits_equal(f, f32(want["tv"][0])) and _bits_equal(g, _hexarr(want["gtv"][1:]))
        elif task == "smooth":
            T = carriers.Tables()
            m = inp["x1"].reshape(n3, n2, n1).copy()
            L.triangle_smoothing(T.f3(m), n1, n2, n3, int(pars["r1"]), int(pars["r2"]), int(pars["r3"]),
                                 int(pars["repeat"]))
            assert _bits_equal(T.keep[0].reshape(-1), _hexarr(want["smooth"][1:])), pars
        else:
            T = carriers.Tables()
            emf, fwi, acq = carriers.Emf(), carriers.Fwi(), carriers.Acq()
            emf.nx, emf.ny, emf.nz = n1, n2, n3
            emf.dx = emf.dy = 100.0
            emf.dz, emf.oz = float(pars["dz"]), float(pars["oz"])
            emf.omegas = T.f1(np.array([carriers.omega_f32(0.5)], np.float32))
            emf.bathy = T.f2(inp["bathy"].reshape(n2, n1))
            npar = int(pars["npar"])
            fwi.npar, fwi.n = npar, npar * nxyz
            fwi.preco = int(pars.get("preco", 1))
            fwi.r1, fwi.r2, fwi.r3, fwi.repeat = (int(pars.get(k, d)) for k, d in
                                                   (("r1", 3), ("r2", 3), ("r3", 1), ("repeat", 1)))
            fwi.iter = int(pars.get("iter", 0))
            fwi.gamma1, fwi.gamma2 = float(pars.get("gamma1", 10)), float(pars.get("gamma2", 0))
            grad = np.zeros((npar, n3, n2, n1), np.float32)
            phess = inp["phess"].reshape(npar, n3, n2, n1).copy()
            fwi.grad, fwi.phess = T.f4(grad), T.f4(phess)
            xref = inp["xref"].copy()
            fwi.xref = fp(xref)
            g = inp["g2"].copy()
            x = inp["x2"].copy()
            if task == "precond":
                L.precondition(C.byref(emf), C.byref(fwi), fp(g))
            else:
                L.fg_mod_reg(C.byref(acq), C.byref(emf), C.byref(fwi), fp(x), fp(g))
                assert _bits_equal(np.float32(fwi.fcost_mod), f32(want["fcost_mod"][0])), pars
            assert _bits_equal(g, _hexarr(want["g"][1:])), (task, pars)
