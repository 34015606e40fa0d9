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


# ---- IKAPTYP 9-12 and 14: the Burrows tables
# Axes (ppmod.f:724-745): log10 T and log10 rho on 50 nodes each, REAL
# constants widened to double.
BURROWS_LOGT = tuple(float(np.float32(v)) for v in (
    1.69896, 1.73978, 1.78061, 1.82143, 1.86225, 1.90308, 1.94390, 1.98473, 2.02555, 2.06637,
    2.10720, 2.14802, 2.18884, 2.22967, 2.27049, 2.31132, 2.35214, 2.39296, 2.43379, 2.47461,
    2.51543, 2.55626, 2.59708, 2.63790, 2.67873, 2.71955, 2.76038, 2.80120, 2.84202, 2.88285,
    2.92367, 2.96406, 3.00488, 3.04571, 3.08653, 3.12735, 3.16818, 3.20900, 3.24983, 3.29065,
    3.33147, 3.37230, 3.41312, 3.45394, 3.49477, 3.53559, 3.57642, 3.61724, 3.65806, 3.69889))
BURROWS_LOGRH = tuple(float(np.float32(v)) for v in (
    -11.9996, -11.7954, -11.5913, -11.3872, -11.1831, -10.9790, -10.7748, -10.5707, -10.3666,
    -10.1625, -9.95837, -9.75425, -9.55014, -9.34602, -9.14190, -8.93778, -8.73366, -8.52954,
    -8.32543, -8.12131, -7.91719, -7.71307, -7.50895, -7.30483, -7.10071, -6.89660, -6.69248,
    -6.48836, -6.28424, -6.08012, -5.87600, -5.67189, -5.46777, -5.26365, -5.05953, -4.85541,
    -4.65129, -4.44718, -4.24479, -4.04068, -3.83656, -3.63244, -3.42875, -3.22464, -3.02052,
    -2.81640, -2.61228, -2.40816, -2.20404, -1.99993))
BNT = BNR = 50
# planet_num 1 (HD209458b) reads these for IKAPTYP 9 and for 10-12, 14
# (OPENOPACITYFILE_WITH_TIOVO / _WOTIOVO, ppplanets.f:90-213)
BURROWS_FILES = {9: "BURROWS6117_wTiVO.dat", 10: "BURROWS6117_woutTiVO.dat", 11: "BURROWS6117_woutTiVO.dat",
                 12: "BURROWS6117_woutTiVO.dat", 14: "BURROWS6117_woutTiVO.dat"}
# SETPLANETKAPEXTRA (ppplanets.f:36-59), REAL constants: the extra absorber
# of IKAPTYP 11 in cm^2/g by planet_num
PLANET_KAP_EXTRA = {1: float(np.float32(0.1)), 2: float(np.float32(0.035)), 3: 0.0, 4: 0.0, 5: 0.0}
RGAS = float(np.float32(8.315e7))  # Rgas (global_constants), a REAL constant


def read_burrows_table(path, ikaptyp, gamma_opc=0.1, mu_gas=2.3, planet_num=1):
    """READBURROWSOPCTABLE (pprphy.f:1227-1520) for IKAPTYP 9-12 and 14.

    Ten header records, one with the stellar temperature, then 2500 rows
    ``T rho log_kP log_kR log_kABS`` read in the order (T outer, rho inner)
    into the tables whatever T and rho the row holds.  Then, per row, as the
    reference does it: IKAPTYP 12 and 14 set log kABS = log kP +
    4 log10(gamma_OPC); 14 sets log kR = log kP; 11 adds the planet's extra
    absorber, log kABS = log10(10^log kABS + kappa_extra), where the row's
    P = Rgas T rho / mu_gas 10^-6 bar is at most 0.03 (a REAL constant).

    Returns ``(kap_planck, kap_ross, kap_abs, t_star)``: 50 x 50 values each,
    temperature fastest (BurrowsXKAP_*(IT, IR) in Fortran order).
    """
    import math
    if ikaptyp not in BURROWS_FILES:
        raise ValueError(f"IKAPTYP {ikaptyp} is not a grey Burrows table (9-12, 14)")
    recs = (r for r in _records(path) if r)
    for _ in range(10):
        next(recs)
    t_star = _number(next(recs)[0])
    tabs = np.full((3, BNR, BNT), MISSING)
    kap_ext = PLANET_KAP_EXTRA[planet_num] if ikaptyp == 11 else 0.0
    lim = float(np.float32(0.03))
    for it in range(BNT):
        for ir in range(BNR):
            vals = []
            while len(vals) < 5:
                try:
                    vals += [_number(t) for t in next(recs)]
                except StopIteration:
                    raise ValueError(f"{path}: {it * BNR + ir} of {BNT * BNR} table rows") from None
            t, rh, kp, kr, ka = vals[:5]
            if ikaptyp in (12, 14):
                ka = kp + 4.0 * math.log10(gamma_opc)
            if ikaptyp == 14:
                kr = kp
            if ikaptyp == 11:
                pg_test = (RGAS * t * rh / mu_gas) * (10.0 ** (-6.0))
                if pg_test <= lim:
                    ka = math.log10(10.0 ** ka + kap_ext)
            tabs[:, ir, it] = (kp, kr, ka)
    kp, kr, ka = (tabs[q].reshape(-1) for q in range(3))
    return kp, kr, ka, t_star


def synthetic_burrows_table(seed: int = 0):
    """A smooth synthetic table on the Burrows grid: log10 kappa_R = -1.5 +
    0.5 log10(rho / 1e-6) + 1.5 log10(T / 1500 K) + 0.05 sin(2 log10 rho),
    Planck 0.25 dex above it, absorption 0.4 dex below it with a ripple in
    log T, each with seeded 0.02 dex noise.  Returns ``(log_t, log_rh,
    kap_planck, kap_ross, kap_abs)``: the grid nodes and three tables,
    temperature fastest."""
    rng = np.random.default_rng(seed)
    lt = np.asarray(BURROWS_LOGT)[None, :]
    lr = np.asarray(BURROWS_LOGRH)[:, None]
    kr = -1.5 + 0.5 * (lr + 6.0) + 1.5 * (lt - np.log10(1500.0)) + 0.05 * np.sin(2.0 * lr) \
        + 0.02 * rng.standard_normal((BNR, BNT))
    kp = kr + 0.25 + 0.02 * rng.standard_normal((BNR, BNT))
    ka = kr - 0.4 + 0.15 * np.cos(3.0 * lt) + 0.02 * rng.standard_normal((BNR, BNT))
    return (np.asarray(BURROWS_LOGT), np.asarray(BURROWS_LOGRH), kp.reshape(-1), kr.reshape(-1),
            ka.reshape(-1))


# ---- IKAPTYP 20-25: the Burrows tables with 30 wavelength bins
NWAVE = 30  # nwave_bins (ppmod.f:489)
# OPENOPACITYFILE_WITH_TIOVO / _WOTIOVO (ppplanets.f:90-217) for planet_num 1
# (HD209458b).  IKAPTYP 22 opens neither (the reference's reader stops on it,
# pprphy.f:1265-1275); 25 falls through to the 5-column grey file for planet 1.
BURROWS_WAVE_FILES = {20: "BURROWS6070_wTiVO_wavelth.dat", 21: "BURROWS6070_woTiVO_wavelth.dat",
                      23: "BURROWS6070_woTiVO_radtransfer.dat", 24: "BURROWS6070_woTiVO_radtransfer.dat"}
BSTAR_FILES = {1: "BSTAR_WAVEBINS_TS6070.dat", 2: "BSTAR_WAVEBINS_TS5040.dat", 5: "BSTAR_WAVEBINS_TS6300.dat"}


def read_burrows_wave_table(path, ikaptyp, planet_num=1):
    """READBURROWSOPCTABLE (pprphy.f:1227-1520) for IKAPTYP 20-25.

    Ten header records, T_star, then 2500 rows in the order (T outer, rho
    inner): ``T rho log_kP log_kR log_kABS`` and the 30 bins' log opacity of
    the stellar irradiation (35 values, IKAPTYP 20 and 21), or those and 60
    more (the local re-radiation opacity and Planck bins of the two-stream
    transfer, which only IRAD 5-7 read; 95 values, 22-25).  IKAPTYP 20, 24 and
    25 add the planet's extra absorber to every irradiation bin, log k =
    log10(10^log k + kappa_extra) (pprphy.f:1438-1469).

    Returns ``(kap_planck, kap_ross, kap_abs, kap_wave, t_star)``: three
    50 x 50 tables, temperature fastest, and 50 x 50 x 30 bins, temperature
    fastest, then rho, then bin (BurrowsXKAP_waveopac(IT, IR, L))."""
    import math
    if ikaptyp not in (20, 21, 22, 23, 24, 25):
        raise ValueError(f"IKAPTYP {ikaptyp} is not a wavelength-binned Burrows table (20-25)")
    if ikaptyp == 22:
        raise ValueError("IKAPTYP 22: the reference's READBURROWSOPCTABLE opens no file for it and stops")
    ncol = 35 if ikaptyp in (20, 21) else 95
    recs = (r for r in _records(path) if r)
    for _ in range(10):
        next(recs)
    t_star = _number(next(recs)[0])
    tabs = np.full((3, BNR, BNT), MISSING)
    wave = np.full((NWAVE, BNR, BNT), MISSING)
    kap_ext = PLANET_KAP_EXTRA[planet_num] if ikaptyp in (20, 24, 25) else 0.0
    for it in range(BNT):
        for ir in range(BNR):
            vals = []
            while len(vals) < ncol:
                try:
                    vals += [_number(t) for t in next(recs)]
                except StopIteration:
                    raise ValueError(f"{path}: {it * BNR + ir} of {BNT * BNR} table rows") from None
            tabs[:, ir, it] = vals[2:5]
            w = vals[5:5 + NWAVE]
            if kap_ext:
                w = [math.log10(10.0 ** x + kap_ext) for x in w]
            wave[:, ir, it] = w
    kp, kr, ka = (tabs[q].reshape(-1) for q in range(3))
    return kp, kr, ka, wave.reshape(-1), t_star


def read_bstar(path):
    """The bins' stellar flux BSTAR_BIN read after the table (pprphy.f:
    1509-1538): three header records, the stellar temperature, the number of
    bins (30), then one row per bin whose second value is BSTAR_BIN."""
    recs = (r for r in _records(path) if r)
    for _ in range(3):
        next(recs)
    next(recs)
    n = int(_number(next(recs)[0]))
    if n != NWAVE:
        raise ValueError(f"{path}: {n} bins, the reference needs {NWAVE}")
    out = []
    while len(out) < NWAVE:
        vals = []
        while len(vals) < 2:
            vals += [_number(t) for t in next(recs)]
        out.append(vals[1])
    return np.asarray(out)


# global_constants (ppmod.f:1074-1083): REAL literals into DOUBLE parameters
_GRAV, _MSOL, _RSOL, _DAY = (float(np.float32(v)) for v in (6.672e-8, 1.989e33, 6.9598e10, 8.64e4))
_PI = 3.141592653589793


def wave_dilution(rstar=1.155, porb=3.52, mstar1=1.119, mstar2=0.000654):
    """(RSTAR RSOL / semi)^2 of the wavelength-binned STELLARINPUT_HEATING
    (pprad.f:444), with semi the semi-major axis of pp.f:204-206 from the
    orbital period (days) and the masses (solar); defaults: the shipped ipp
    (ipp:18, 22-23)."""
    semi = (_GRAV * _MSOL * (mstar1 + mstar2) * ((porb * _DAY) ** 2.0) / (4.0 * _PI * _PI)) ** (1.0 / 3.0)
    return (rstar * _RSOL / semi) ** 2.0


def synthetic_wave_table(seed: int = 0):
    """30 smooth synthetic irradiation bins on the Burrows grid (log10 k of
    bin l: -2.5 + 2 l/29 + 0.3 log10(rho / 1e-6) + 0.2 log10(T / 1500 K), with
    seeded 0.02 dex noise) and a stellar spectrum over the bins whose total
    is sigma (6070 K)^4.  Returns ``(kap_wave, bstar)``."""
    rng = np.random.default_rng(seed)
    lt = np.asarray(BURROWS_LOGT)[None, None, :]
    lr = np.asarray(BURROWS_LOGRH)[None, :, None]
    lb = np.arange(NWAVE)[:, None, None]
    kw = -2.5 + 2.0 * lb / 29.0 + 0.3 * (lr + 6.0) + 0.2 * (lt - np.log10(1500.0)) \
        + 0.02 * rng.standard_normal((NWAVE, BNR, BNT))
    shape = np.exp(-0.5 * ((np.arange(NWAVE) - 12.0) / 6.0) ** 2)
    bstar = shape / shape.sum() * float(np.float32(5.6705e-5)) * 6070.0 ** 4
    return kw.reshape(-1), bstar


# ---- IKAPTYP 6 and 7: the column-depth absorption table
COLUMN_LEVELS = (-26.0, -25.0, -24.0, -23.0, -22.0, -21.0, -20.0, -19.0)  # log10 kappa cut-off (ppmod.f:701-702)
NC = len(COLUMN_LEVELS)
N_COLUMN_ENTRIES = 5888  # rows the reference reads (pprphy.f:1763)
COLUMN_FILES = {6: "Freedman6117_column.dat", 7: "Freedman6117_column_noTIVO.dat"}


def read_column_table(path):
    """READCOLUMNDEPTHOPCTABLE (pprphy.f:1747-1833).

    Ten header records, T_star, the number of levels, the level range (two
    values), then 5888 rows ``T P level log_kABS`` placed by their grid
    values (T, P and the level must be grid values exactly); entries the file
    does not list stay -7.  Then, as the reference does, every level 1-7 of
    each (T, P) is capped at level 8's value.

    Returns ``(kap_abs, t_star)``: 42 x 18 x 8 values, temperature fastest,
    then pressure, then level (ColumnXKAP_abs(IT, IP, IC) in Fortran order).
    """
    recs = (r for r in _records(path) if r)
    for _ in range(10):
        next(recs)
    t_star = _number(next(recs)[0])
    next(recs)  # the number of levels
    lv = []
    while len(lv) < 2:  # the level range
        lv += next(recs)
    tab = np.full((NC, NP, NT), MISSING)
    tidx = {t: i for i, t in enumerate(FREEDMAN_T)}
    pidx = {p: i for i, p in enumerate(FREEDMAN_P)}
    cidx = {c: i for i, c in enumerate(COLUMN_LEVELS)}
    for row in range(N_COLUMN_ENTRIES):
        vals = []
        while len(vals) < 4:
            try:
                vals += [_number(t) for t in next(recs)]
            except StopIteration:
                raise ValueError(f"{path}: {row} of {N_COLUMN_ENTRIES} table rows") from None
        t, p, c, ka = vals[:4]
        if t not in tidx or p not in pidx or c not in cidx:
            raise ValueError(f"{path}: row {row + 1} (T={t}, P={p}, level={c}) is not on the table grid")
        tab[cidx[c], pidx[p], tidx[t]] = ka
    top = tab[NC - 1]
    for ic in range(NC - 2, -1, -1):
        tab[ic] = np.where(tab[ic] > top, top, tab[ic])
    return tab.reshape(-1), t_star
