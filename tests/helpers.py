"""Shared helpers of the test suite (no GPU work here)."""
import os

import numpy as np


def interior_mask(n):
    """Boolean [3][n3+1][n2+1][n1+1] mask of the edges that carry an equation."""
    n1, n2, n3 = n
    k, j, i = np.meshgrid(np.arange(n3 + 1), np.arange(n2 + 1), np.arange(n1 + 1), indexing="ij")
    p = (i, j, k)
    m = np.zeros((3, n3 + 1, n2 + 1, n1 + 1), bool)
    nn = (n1, n2, n3)
    for a in range(3):
        b, c = (a + 1) % 3, (a + 2) % 3
        m[a] = (p[a] < nn[a]) & (p[b] >= 1) & (p[b] <= nn[b] - 1) & (p[c] >= 1) & (p[c] <= nn[c] - 1)
    return m


def random_field(n, seed, interior_only=True):
    rng = np.random.default_rng(seed)
    n1, n2, n3 = n
    v = rng.standard_normal((3, n3 + 1, n2 + 1, n1 + 1)) + 1j * rng.standard_normal((3, n3 + 1, n2 + 1, n1 + 1))
    if interior_only:
        v = np.where(interior_mask(n), v, 0)
    return np.ascontiguousarray(v.reshape(-1))


def rel(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


# ---------------------------------------------------------------------------
# inputs of the command-line program's tests
# ---------------------------------------------------------------------------
CLI_CHANNELS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
CLI_FREQS = (0.5, 1.25)


def cli_survey():
    """A 3-source, 6-receiver survey with varied azimuths and dips (so every
    rotation branch of the data weighting is taken), on the tiny model."""
    from libemmi_mgfd_amd import configs
    s = configs.tiny()
    s.sources = [(100.0, -50.0, 450.0, 30.0, 5.0), (-900.0, 300.0, 440.0, 75.0, 0.0),
                 (1500.0, -800.0, 460.0, -20.0, 2.0)]
    s.receivers = [(float(x), 100.0 - 37.0 * i, 520.0, 10.0 * i - 15.0, 3.0 * i)
                   for i, x in enumerate(np.linspace(-2200, 2100, 6))]
    s.table = [(1, 1), (1, 2), (1, 3), (1, 4), (1, 5), (1, 6), (2, 2), (2, 4), (2, 6), (3, 1), (3, 6)]
    s.freqs = list(CLI_FREQS)
    s.chrec = CLI_CHANNELS
    return s


def cli_write_inputs(d, shots=(1, 2, 3), seed=11):
    """The survey's files in d plus emf_<shot>.txt data files in the
    reference's format (random complex values, one line per channel,
    frequency and receiver of the shot).  Returns the file map."""
    from libemmi_mgfd_amd import configs
    s = cli_survey()
    files = configs.write_reference_inputs(s, d)
    rng = np.random.default_rng(seed)
    for shot in shots:
        recs = [r for (i, r) in s.table if i == shot]
        with open(os.path.join(d, f"emf_{shot:04d}.txt"), "w") as fp:
            fp.write("iTx \tiRx    chrec  frequency/Hz    Real{E/H}     Imag{E/H}\n")
            for ch in CLI_CHANNELS:
                scale = 1e-12 if ch.startswith("E") else 1e-10
                for f in CLI_FREQS:
                    for k in range(len(recs)):
                        re, im = rng.standard_normal(2) * scale
                        fp.write(f"{shot} \t {k + 1} \t {ch} \t {f:g} \t {re:e} \t {im:e}\n")
    return files


def cli_args(files, extra=()):
    """The command-line words of the survey (the reference's key=value)."""
    s = cli_survey()
    words = [f"freqs={','.join(f'{f:g}' for f in s.freqs)}", "chsrc=Ex", f"chrec={','.join(CLI_CHANNELS)}",
             f"nx={s.nx}", f"ny={s.ny}", f"nz={s.nz}", f"dx={s.dx:g}", f"dy={s.dy:g}", f"dz={s.dz:g}",
             f"ox={s.ox:g}", f"oy={s.oy:g}", f"oz={s.oz:g}", f"n1={s.n1}", f"n2={s.n2}", f"n3={s.n3}",
             f"d1min={s.d1min:g}", f"d2min={s.d2min:g}", f"d3min={s.d3min:g}", f"nb={s.nb}",
             f"lextend={s.lextend:g}", f"rho_air={s.rho_air:g}"]
    words += [f"{k}={v}" for k, v in files.items()]
    return words + list(extra)
