"""Synthetic survey configurations for the multigrid Maxwell solve.

Each configuration is a plain description of the reference's inputs: an input
resistivity model on a uniform grid (the arrays that ``frho11/frho22/frho33``
point to, float32 ``[nz][ny][nx]``), a source list, a receiver list, a
source-receiver table, the frequencies and the multigrid size.  The files the
reference reads (``src/emf_init_free.c:81-107``, ``src/acq_init_free.c:68-262``)
can be written with :func:`write_reference_inputs`; the same description feeds
the MI355X host program and the benchmarks.

Cell-centred classification is used throughout: a cell belongs to a layer
when its centre lies in it.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import List, Sequence, Tuple

import numpy as np

RHO_WATER = 0.3125


@dataclasses.dataclass
class Survey:
    name: str
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    ox: float
    oy: float
    oz: float
    rho11: np.ndarray  # float32 [nz][ny][nx]
    rho22: np.ndarray
    rho33: np.ndarray
    sources: List[Tuple[float, float, float, float, float]]    # x, y, z, azimuth deg, dip deg
    receivers: List[Tuple[float, float, float, float, float]]
    table: List[Tuple[int, int]]                               # 1-based (isrc, irec)
    freqs: List[float]
    n1: int
    n2: int
    n3: int
    chsrc: Sequence[str] = ("Ex",)
    chrec: Sequence[str] = ("Ex", "Ey", "Hx", "Hy")
    d1min: float = 100.0
    d2min: float = 100.0
    d3min: float = 20.0
    nb: int = 10
    lextend: float = 50e3
    rho_air: float = 1e8
    bathy: np.ndarray | None = None  # float32 [ny][nx], depth of seafloor
    extra: dict = dataclasses.field(default_factory=dict)

    @property
    def nshots(self) -> int:
        return len(self.sources)

    def getpar_args(self, files: dict) -> List[str]:
        """key=value arguments in the reference's getpar convention."""
        a = [
            "freqs=" + ",".join(_fmt(f) for f in self.freqs),
            "chsrc=" + ",".join(self.chsrc),
            "chrec=" + ",".join(self.chrec),
            f"nx={self.nx}", f"ny={self.ny}", f"nz={self.nz}",
            f"dx={_fmt(self.dx)}", f"dy={_fmt(self.dy)}", f"dz={_fmt(self.dz)}",
            f"ox={_fmt(self.ox)}", f"oy={_fmt(self.oy)}", f"oz={_fmt(self.oz)}",
            f"n1={self.n1}", f"n2={self.n2}", f"n3={self.n3}",
            f"d1min={_fmt(self.d1min)}", f"d2min={_fmt(self.d2min)}", f"d3min={_fmt(self.d3min)}",
            f"nb={self.nb}", f"lextend={_fmt(self.lextend)}", f"rho_air={_fmt(self.rho_air)}",
            f"frho11={files['frho11']}", f"frho22={files['frho22']}", f"frho33={files['frho33']}",
            f"fsrc={files['fsrc']}", f"frec={files['frec']}", f"fsrcrec={files['fsrcrec']}",
        ]
        if "fbathy" in files:
            a.append(f"fbathy={files['fbathy']}")
        return a


def _fmt(v: float) -> str:
    r = repr(float(v))
    return r[:-2] if r.endswith(".0") else r


def write_reference_inputs(s: Survey, d: str) -> dict:
    """Write the model and acquisition files in the reference's formats."""
    os.makedirs(d, exist_ok=True)
    files = {}
    for key, arr in (("frho11", s.rho11), ("frho22", s.rho22), ("frho33", s.rho33)):
        p = os.path.join(d, key + ".bin")
        np.ascontiguousarray(arr, dtype=np.float32).tofile(p)
        files[key] = p
    p = os.path.join(d, "sources.txt")
    with open(p, "w") as fp:
        fp.write(" x y z azimuth dip isrc\n")
        for i, (x, y, z, az, dip) in enumerate(s.sources):
            fp.write(f"{x!r} {y!r} {z!r} {az!r} {dip!r} {i + 1}\n")
    files["fsrc"] = p
    p = os.path.join(d, "receivers.txt")
    with open(p, "w") as fp:
        fp.write(" x y z azimuth dip irec\n")
        for i, (x, y, z, az, dip) in enumerate(s.receivers):
            fp.write(f"{x!r} {y!r} {z!r} {az!r} {dip!r} {i + 1}\n")
    files["frec"] = p
    p = os.path.join(d, "src_rec_table.txt")
    with open(p, "w") as fp:
        fp.write(" isrc irec\n")
        for isrc, irec in s.table:
            fp.write(f"{isrc} {irec}\n")
    files["fsrcrec"] = p
    if s.bathy is not None:
        p = os.path.join(d, "bathy.bin")
        np.ascontiguousarray(s.bathy, dtype=np.float32).tofile(p)
        files["fbathy"] = p
    return files


def _centres(n, d, o):
    return o + (np.arange(n) + 0.5) * d


def _layered(nx, ny, nz, dx, dy, dz, ox, oy, oz, seafloor, layers, rhov_sed=1.0, blocks=()):
    """Water above ``seafloor(x, y)``, sediment below, plus resistive layers
    ``(ztop, zbot, rho)`` and boxes ``(xc, yc, zc, wx, wy, wz, rho)``."""
    xc = _centres(nx, dx, ox)
    yc = _centres(ny, dy, oy)
    zc = _centres(nz, dz, oz)
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    sf = seafloor(X, Y)
    rh = np.where(Z < sf, RHO_WATER, 1.0)
    rv = np.where(Z < sf, RHO_WATER, rhov_sed)
    for ztop, zbot, rho in layers:
        m = (Z >= ztop) & (Z < zbot) & (Z >= sf)
        rh = np.where(m, rho, rh)
        rv = np.where(m, rho, rv)
    for xcen, ycen, zcen, wx, wy, wz, rho in blocks:
        m = ((np.abs(X - xcen) < 0.5 * wx) & (np.abs(Y - ycen) < 0.5 * wy)
             & (np.abs(Z - zcen) < 0.5 * wz) & (Z >= sf))
        rh = np.where(m, rho, rh)
        rv = np.where(m, rho, rv)
    bathy = seafloor(X[0], Y[0]).astype(np.float32)
    return rh.astype(np.float32), rv.astype(np.float32), bathy


def config1(n=64) -> Survey:
    """Plain FP64 forward modelling, 1 shot at 0.25 Hz, MG grid 64^3."""
    nx = ny = 101
    nz = 51
    dx = dy = 100.0
    dz = 40.0
    ox = oy = -5050.0
    rh, rv, bathy = _layered(nx, ny, nz, dx, dy, dz, ox, oy, 0.0,
                             lambda x, y: np.full_like(x, 1000.0),
                             [(1480.0, 1600.0, 100.0)])
    recs = [(float(x), 0.0, 1000.0, 0.0, 0.0) for x in np.arange(-5000.0, 5000.0 + 1, 100.0)]
    return Survey("config1", nx, ny, nz, dx, dy, dz, ox, oy, 0.0, rh, rh.copy(), rv,
                  [(0.0, 0.0, 950.0, 0.0, 0.0)], recs, [(1, i + 1) for i in range(len(recs))],
                  [0.25], n, n, n, bathy=bathy)


def _seafloor2(x, y):
    return 1000.0 + 100.0 * np.sin(2 * np.pi * x / 4000.0) * np.cos(2 * np.pi * y / 6000.0)


def config2(n=128) -> Survey:
    """Anisotropic sediment, sinusoidal seafloor, one 100 ohm-m block; 3 freqs at MG 128^3."""
    nx = ny = 201
    nz = 101
    dx = dy = 50.0
    dz = 20.0
    ox = oy = -5025.0
    rh, rv, bathy = _layered(nx, ny, nz, dx, dy, dz, ox, oy, 0.0, _seafloor2,
                             [(1480.0, 1600.0, 100.0)], rhov_sed=2.0,
                             blocks=[(1000.0, 0.0, 1550.0, 4000.0, 4000.0, 100.0, 100.0)])
    zs = float(_seafloor2(np.float64(0.0), np.float64(0.0))) - 50.0
    xr = np.arange(-5000.0, 5000.0 + 1, 50.0)
    recs = [(float(x), 0.0, float(_seafloor2(x, 0.0)), 0.0, 0.0) for x in xr]
    return Survey("config2", nx, ny, nz, dx, dy, dz, ox, oy, 0.0, rh, rh.copy(), rv,
                  [(0.0, 0.0, zs, 0.0, 0.0)], recs, [(1, i + 1) for i in range(len(recs))],
                  [0.25, 0.75, 1.25], n, n, n, bathy=bathy)


def tiny(n1=12, n2=10, n3=16, nrec=5, rho_air=1e8) -> Survey:
    """A small anisotropic survey for kernel-level known-answer tests; the
    multigrid sizes are deliberately unequal so a transposed index fails."""
    nx, ny, nz = 24, 20, 10
    dx = dy = 200.0
    dz = 100.0
    ox, oy = -2400.0, -2000.0
    rh, rv, bathy = _layered(nx, ny, nz, dx, dy, dz, ox, oy, 0.0,
                             lambda x, y: 500.0 + 40.0 * np.sin(2 * np.pi * x / 3000.0),
                             [(700.0, 800.0, 50.0)], rhov_sed=1.7)
    recs = [(float(x), 100.0, 520.0, 10.0 * i, 3.0 * i) for i, x in enumerate(np.linspace(-1800, 1800, nrec))]
    return Survey("tiny", nx, ny, nz, dx, dy, dz, ox, oy, 0.0, rh, rh * np.float32(1.1), rv,
                  [(100.0, -50.0, 450.0, 30.0, 5.0)], recs, [(1, i + 1) for i in range(len(recs))],
                  [0.5, 1.5], n1, n2, n3, d1min=200.0, d2min=200.0, d3min=60.0, nb=3,
                  lextend=8000.0, rho_air=rho_air, bathy=bathy)


def _flat1000(x, y):
    return np.full_like(x, 1000.0)


# multigrid size of each frequency of the batched survey: lower frequencies
# diffuse further and get the coarser grid
CONFIG3_GRIDS = {0.1: (96, 96, 96), 0.25: (128, 128, 128), 0.5: (128, 128, 128), 1.0: (160, 160, 128)}


def config3() -> Survey:
    """Batched survey: 25 shots x 4 frequencies, each frequency on its own
    multigrid size (``extra["grids"]``).

    The configuration-1 earth (flat seafloor at 1000 m, 100 ohm-m layer at
    1480-1600 m, 1 ohm-m below) on a 161x161x71 grid of 100/100/40 m cells
    with origin (-8050, -8050, 0).  The receiver lines reach x = +-8 km,
    outside configuration 1's +-5.05 km model, which the reference's
    acquisition check (src/acq_init_free.c:207-212) refuses.  The model is
    71 cells deep (to 2840 m) rather than 51: the reference finds the source
    cell with a bisection over nx+2nb = 181 nodes of the z axis
    (src/regrid_init_free.c:194), whose first probe (node 90) lies past a
    72-node z axis; with 92 nodes every probe stays on the axis.  The extra
    cells repeat the 1 ohm-m half-space that the padding below would
    repeat anyway.  Sources on a 5x5 grid,
    x, y in {-4, -2, 0, 2, 4} km (x fastest), 50 m above the seafloor,
    azimuth 0; 5 receiver lines y in {-4, ..., 4} km of 81 receivers,
    x = -8..8 km every 200 m, on the seafloor; every source records every
    receiver."""
    nx = ny = 161
    nz = 71
    dx = dy = 100.0
    dz = 40.0
    ox = oy = -8050.0
    rh, rv, bathy = _layered(nx, ny, nz, dx, dy, dz, ox, oy, 0.0, _flat1000, [(1480.0, 1600.0, 100.0)])
    km = [-4000.0, -2000.0, 0.0, 2000.0, 4000.0]
    srcs = [(x, y, 950.0, 0.0, 0.0) for y in km for x in km]
    recs = [(float(x), y, 1000.0, 0.0, 0.0) for y in km for x in np.arange(-8000.0, 8000.0 + 1, 200.0)]
    table = [(s + 1, r + 1) for s in range(len(srcs)) for r in range(len(recs))]
    freqs = [0.1, 0.25, 0.5, 1.0]
    s = Survey("config3", nx, ny, nz, dx, dy, dz, ox, oy, 0.0, rh, rh.copy(), rv, srcs, recs, table,
               freqs, 128, 128, 128, bathy=bathy)
    s.extra["grids"] = dict(CONFIG3_GRIDS)
    return s


def config3_single(shot: int, freq: float) -> Survey:
    """One (shot, frequency) right-hand side of :func:`config3` as the
    reference runs it: one source (1-based ``shot``), that shot's receivers,
    one frequency on that frequency's multigrid size."""
    s = config3()
    src = s.sources[shot - 1]
    recs = [s.receivers[r - 1] for (i, r) in s.table if i == shot]
    n1, n2, n3 = s.extra["grids"][freq]
    return Survey(f"config3_s{shot}_f{freq:g}", s.nx, s.ny, s.nz, s.dx, s.dy, s.dz, s.ox, s.oy, s.oz,
                  s.rho11, s.rho22, s.rho33, [src], recs, [(1, i + 1) for i in range(len(recs))],
                  [freq], n1, n2, n3, bathy=s.bathy)


CONFIG5_BOXES = [(-1500.0, 0.0, 1400.0, 3000.0, 2000.0, 150.0, 50.0),
                 (1500.0, 500.0, 2100.0, 3000.0, 2000.0, 150.0, 50.0)]


def config5(true_model: bool = True) -> Survey:
    """Mode-2 gradient at survey scale: 25 shots x 3 frequencies on a
    160x160x128 multigrid grid.

    Configuration-2 acquisition on its 50/50/20 m input grid with origin
    (-5025, -5025, 0): 201 inline seafloor receivers (x = -5..5 km every
    50 m, y = 0) and 25 inline x-dipole sources, x = -4.8..4.8 km every
    400 m, 50 m above the seafloor.  The true earth is the configuration-1
    background (flat seafloor at 1000 m, 1 ohm-m sediment, 100 ohm-m layer
    at 1480-1600 m) plus two 50 ohm-m boxes of 3x2x0.15 km centred at depths
    1400 m and 2100 m (``CONFIG5_BOXES``); the grid is 126 cells deep (to
    2520 m) so that the deeper box lies inside the model.  With
    ``true_model=False`` the model is the starting model of the gradient: a
    1 ohm-m half-space under the seawater."""
    nx = ny = 201
    nz = 126
    dx = dy = 50.0
    dz = 20.0
    ox = oy = -5025.0
    if true_model:
        rh, rv, bathy = _layered(nx, ny, nz, dx, dy, dz, ox, oy, 0.0, _flat1000, [(1480.0, 1600.0, 100.0)],
                                 blocks=CONFIG5_BOXES)
    else:
        rh, rv, bathy = _layered(nx, ny, nz, dx, dy, dz, ox, oy, 0.0, _flat1000, [])
    srcs = [(float(x), 0.0, 950.0, 0.0, 0.0) for x in np.arange(-4800.0, 4800.0 + 1, 400.0)]
    recs = [(float(x), 0.0, 1000.0, 0.0, 0.0) for x in np.arange(-5000.0, 5000.0 + 1, 50.0)]
    table = [(s + 1, r + 1) for s in range(len(srcs)) for r in range(len(recs))]
    return Survey("config5" if true_model else "config5_start", nx, ny, nz, dx, dy, dz, ox, oy, 0.0,
                  rh, rh.copy(), rv, srcs, recs, table, [0.25, 0.75, 1.25], 160, 160, 128, bathy=bathy)


CONFIG4_TRAIN = range(0, 2000)
CONFIG4_TEST = range(2000, 2200)


def config4(seed: int, n: int = 96) -> Survey:
    """One random earth of the learned-initial-guess data set (seeds 0-1999
    train, 2000-2199 test), drawn from numpy's default generator seeded with
    ``seed``.

    Input model 101x101x81 cells of 100/100/50 m, origin (-5050, -5050, 0):
    sea water (0.3125 ohm-m) down to a seafloor at depth d0, uniform in
    500-2000 m, plus a sinusoidal relief of amplitude up to 150 m (wavelengths
    2-12 km, random phases); below it 3-6 flat layers whose interfaces lie
    uniformly between d0 + 100 m and 4 km, each with rho_h log-uniform in
    0.5-5 ohm-m and rho_v / rho_h uniform in 1-3; then 0-3 resistive boxes of
    20-200 ohm-m (log-uniform), 0.5-5 km wide in x and y, 50-300 m thick, with
    their centre at x, y uniform in +-4 km and 200-2000 m below d0 (only cells
    below the seafloor take them).  One x-directed dipole 50 m above the
    seafloor at x, y uniform in +-2 km; one frequency, log-uniform in
    0.1-2 Hz; 31 inline receivers on the seafloor within 3 km of the source
    (every 200 m).  Multigrid n^3 (96^3), stretched about the source."""
    rng = np.random.default_rng(int(seed))
    nx = ny = 101
    nz = 81
    dx = dy = 100.0
    dz = 50.0
    ox = oy = -5050.0
    d0 = rng.uniform(500.0, 2000.0)
    amp = rng.uniform(0.0, 150.0)
    lx, ly = rng.uniform(2000.0, 12000.0, 2)
    px, py = rng.uniform(0.0, 2.0 * np.pi, 2)

    def seafloor(x, y):
        return d0 + amp * np.sin(2.0 * np.pi * x / lx + px) * np.cos(2.0 * np.pi * y / ly + py)

    xc, yc, zc = _centres(nx, dx, ox), _centres(ny, dy, oy), _centres(nz, dz, 0.0)
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    sf = seafloor(X, Y)
    nlay = int(rng.integers(3, 7))
    tops = np.sort(rng.uniform(d0 + 100.0, 4000.0, nlay - 1))
    lay_h = np.exp(rng.uniform(np.log(0.5), np.log(5.0), nlay))
    lay_v = lay_h * rng.uniform(1.0, 3.0, nlay)
    idx = np.searchsorted(tops, Z)
    below = Z >= sf
    rh = np.where(below, lay_h[idx], RHO_WATER)
    rv = np.where(below, lay_v[idx], RHO_WATER)
    for _ in range(int(rng.integers(0, 4))):
        rho = float(np.exp(rng.uniform(np.log(20.0), np.log(200.0))))
        wx, wy = rng.uniform(500.0, 5000.0, 2)
        th = rng.uniform(50.0, 300.0)
        cx, cy = rng.uniform(-4000.0, 4000.0, 2)
        cz = d0 + rng.uniform(200.0, 2000.0)
        m = (np.abs(X - cx) < 0.5 * wx) & (np.abs(Y - cy) < 0.5 * wy) & (np.abs(Z - cz) < 0.5 * th) & below
        rh = np.where(m, rho, rh)
        rv = np.where(m, rho, rv)
    sx, sy = (float(v) for v in rng.uniform(-2000.0, 2000.0, 2))
    freq = float(np.exp(rng.uniform(np.log(0.1), np.log(2.0))))
    zs = float(seafloor(sx, sy)) - 50.0
    recs = [(float(x), sy, float(seafloor(x, sy)), 0.0, 0.0) for x in sx + np.arange(-3000.0, 3000.0 + 1, 200.0)]
    rh = rh.astype(np.float32)
    bathy = seafloor(X[0], Y[0]).astype(np.float32)
    s = Survey(f"config4_{seed}", nx, ny, nz, dx, dy, dz, ox, oy, 0.0, rh, rh.copy(), rv.astype(np.float32),
               [(sx, sy, zs, 0.0, 0.0)], recs, [(1, i + 1) for i in range(len(recs))], [freq], n, n, n,
               bathy=bathy)
    s.extra["seed"] = int(seed)
    return s


CONFIGS = {"config1": config1, "config2": config2, "config3": config3, "config4": config4, "config5": config5,
           "tiny": tiny}
