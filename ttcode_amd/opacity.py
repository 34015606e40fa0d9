"""Opacity tables of IKAPTYP 4 and 5 (the Freedman, Marley & Lodders
tables as ttcode reads them).

The tables hold log10 of the Planck, Rosseland and absorption mean opacities
on a fixed grid of 42 temperatures x 18 pressures: the reference's DATA
statements (ppmod.f:690-699).  ``read_freedman_table`` restates the
reference's reader, READFREEDMANOPCTABLE (pprphy.f:944-1002), for its text
format, and the library takes the result through ``tt_set_opacity_table``
(``Device.set_opacity_table``).  IKAPTYP 4 reads ``Freedman6117.dat`` and
IKAPTYP 5 ``Freedman6117_noTIVO.dat`` (without TiO/VO); the interpolation is
the same.
"""
from __future__ import annotations

import numpy as np

# table axes: T in K, P in mbar (ppmod.f:690-699)
FREEDMAN_T = (75.0, 100.0, 110.0, 120.0, 130.0, 140.0, 150.0, 160.0, 170.0, 180.0, 190.0, 200.0,
              210.0, 220.0, 230.0, 240.0, 250.0, 275.0, 300.0, 400.0, 500.0, 575.0, 650.0, 725.0,
              800.0, 900.0, 1000.0, 1100.0, 1200.0, 1300.0, 1400.0, 1500.0, 1600.0, 1700.0, 1800.0,
              1900.0, 2000.0, 2300.0, 2600.0, 3000.0, 3500.0, 4000.0)
FREEDMAN_P = (0.001, 0.003, 0.01, 0.03, 0.1, 0.3, 1.0, 3.0, 10.0, 30.0, 100.0, 300.0, 1000.0,
              3000.0, 10000.0, 30000.0, 100000.0, 300000.0)
NT, NP = len(FREEDMAN_T), len(FREEDMAN_P)
MISSING = -7.0        # value of entries the file does not list (pprphy.f:977-979)
N_ENTRIES = 736       # rows the reference reads (pprphy.f:976)
TABLE_FILES = {4: "Freedman6117.dat", 5: "Freedman6117_noTIVO.dat"}


def _records(path):
    """The file's lines as lists of list-directed values (blanks and commas
    separate; a ``d`` exponent is a Fortran double)."""
    with open(path) as f:
        for line in f:
            yield line.replace(",", " ").split()


def _number(tok: str) -> float:
    return float(tok.lower().replace("d", "e"))


def read_freedman_table(path):
    """READFREEDMANOPCTABLE (pprphy.f:944-1002).

    Ten header records, one record with the stellar temperature the
    absorption opacities were computed for, then 736 rows ``T P log_kP
    log_kR log_kABS``.  Each list-directed READ starts a new record, skips
    records without values and takes its values from as many records as it
    needs.  A row's T and P must be grid values exactly; entries the file
    does not list stay -7.

    Returns ``(kap_planck, kap_ross, kap_abs, t_star)``: three arrays of
    42 x 18 = 756 values, temperature fastest (the reference's
    FreedmanXKAP_*(IT, IP) in Fortran order), and T_star.
    """
    recs = (r for r in _records(path) if r)
    for _ in range(10):
        next(recs)
    t_star = _number(next(recs)[0])
    tabs = np.full((3, NP, NT), MISSING)
    tidx = {t: i for i, t in enumerate(FREEDMAN_T)}
    pidx = {p: i for i, p in enumerate(FREEDMAN_P)}
    for row in range(N_ENTRIES):
        vals = []
        while len(vals) < 5:
            try:
                vals += [_number(t) for t in next(recs)]
            except StopIteration:
                raise ValueError(f"{path}: {row} of {N_ENTRIES} table rows") from None
        t, p, kp, kr, ka = vals[:5]
        if t not in tidx or p not in pidx:
            raise ValueError(f"{path}: row {row + 1} (T={t}, P={p}) is not on the table grid")
        tabs[:, pidx[p], tidx[t]] = (kp, kr, ka)
    kp, kr, ka = (tabs[q].reshape(-1) for q in range(3))
    return kp, kr, ka, t_star


def synthetic_table(seed: int = 0):
    """A smooth synthetic table on the reference's grid, for measurements
    where no table file is at hand: log10 kappa_R = -2 + 0.45 log10 P +
    1.2 log10(T / 1500 K) + 0.05 sin(3 log10 P), Planck 0.3 dex above it
    (less 0.1 log10(T / 1500 K)), absorption 0.5 dex below it with a ripple
    in log T, each with seeded 0.02 dex noise; P in mbar.  Returns
    ``(kap_planck, kap_ross, kap_abs)``, temperature fastest, no entry
    missing."""
    rng = np.random.default_rng(seed)
    lt = np.log10(np.asarray(FREEDMAN_T) / 1500.0)[None, :]
    lp = np.log10(np.asarray(FREEDMAN_P))[:, None]
    kr = -2.0 + 0.45 * lp + 1.2 * lt + 0.05 * np.sin(3.0 * lp) + 0.02 * rng.standard_normal((NP, NT))
    kp = kr + 0.3 - 0.1 * lt + 0.02 * rng.standard_normal((NP, NT))
    ka = kr - 0.5 + 0.2 * np.cos(2.0 * np.log10(np.asarray(FREEDMAN_T) / 100.0))[None, :] \
        + 0.02 * rng.standard_normal((NP, NT))
    return kp.reshape(-1), kr.reshape(-1), ka.reshape(-1)
