"""Closed forms that pin parts of the path the reference cannot be run on
(ppsub.f needs hypre's absent bHYPRE_SStructVariable.inc, so SCHRITT has no
reference fixture).  Written from the reference's formulas in numpy, on the
wire-format arrays; shared by the oracle tests and the GPU tests."""
from __future__ import annotations

import numpy as np

from ttcode_amd import params as P

# metric ids (include/ttgpu.h TT_M_*)
M_DXA, M_DYA, M_DZA = 2, 6, 10
M_GEOXG, M_GEOXH, M_GEOZG = 42, 43, 44
M_LO = {M_DXA: 1, M_DYA: 1, M_DZA: 1, M_GEOXG: -1, M_GEOXH: -1, M_GEOZG: -1}


def interior(a, name, p, di=0, dj=0, dk=0):
    """A(i+di, j+dj, k+dk) for the interior i=1..nx, j=1..ny, k=1..nz, as a
    (nz, ny, nx) array of the wire-format array `a` of field `name`."""
    lo, _, _ = P.field_bounds(P.FIELD_IDS[name], p.nx, p.ny, p.nz)
    return a[1 + dk - lo[2]:p.nz + dk - lo[2] + 1,
             1 + dj - lo[1]:p.ny + dj - lo[1] + 1,
             1 + di - lo[0]:p.nx + di - lo[0] + 1]


def metric_at(m, arr, first, n):
    """arr(first .. first+n-1) of a metric array held from its lower bound."""
    s = first - M_LO[m]
    return arr[s:s + n]


def perturbed_metrics(get, rng, amp=0.2):
    """The six metric arrays SCHRITT reads, each entry scaled by a seeded
    factor in [1-amp, 1+amp], so that an index shift or a swapped array
    changes the bound."""
    out = {}
    for m in (M_DXA, M_DYA, M_DZA, M_GEOXG, M_GEOXH, M_GEOZG):
        a = get(m)
        out[m] = a * (1.0 + amp * rng.uniform(-1.0, 1.0, a.shape))
    return out


def schritt_cfl(p, st, met, mutation=None):
    """SCHRITT's Courant bound, ppsub.f:49-78 and :129-131 (IGASTYP=1,
    ICSTYP=0, no viscosity): PG = RGAS T RH / mu (ppphy.f:20), CS =
    sqrt(gamma PG / RH) (ppphy.f:92-98),
        ZC1 = DXA(i+1) / (|V(i+1,j,k)| + CS)
        ZC2 = DYA(j+1) / (|G(i,j+1,k)| + CS / GEOXG(i) / GEOZG(k))
        ZC3 = DZA(k+1) / (|H(i,j,k+1)| + CS / GEOXH(i))
    delt = FDELT min(DTMAX, min over the interior of ZC1, ZC2, ZC3).
    `met` maps the metric ids to their arrays (declared bounds).  `mutation`
    names a planted mistake, to show that the comparison would catch it.
    Returns (delt, the minima of ZC1, ZC2, ZC3)."""
    nx, ny, nz = p.nx, p.ny, p.nz
    rh = interior(st["RH"], "RH", p)
    t = interior(st["T"], "T", p)
    pg = P.RGAS * t * rh / p.mu_gas
    cs = np.sqrt(p.gamma * pg / rh)
    sv = -1 if mutation == "V(i)" else 0
    sg = -1 if mutation == "G(j)" else 0
    sh = -1 if mutation == "H(k)" else 0
    v1 = np.abs(interior(st["V"], "V", p, di=1 + sv))
    v2 = np.abs(interior(st["G"], "G", p, dj=1 + sg))
    v3 = np.abs(interior(st["H"], "H", p, dk=1 + sh))
    dxa = metric_at(M_DXA, met[M_DXA], 1 if mutation == "DXA(i)" else 2, nx)[None, None, :]
    dya = metric_at(M_DYA, met[M_DYA], 1 if mutation == "DYA(j)" else 2, ny)[None, :, None]
    dza = metric_at(M_DZA, met[M_DZA], 1 if mutation == "DZA(k)" else 2, nz)[:, None, None]
    gxg_id, gxh_id = (M_GEOXH, M_GEOXG) if mutation == "GEOXG<->GEOXH" else (M_GEOXG, M_GEOXH)
    gxg = metric_at(gxg_id, met[gxg_id], 1, nx)[None, None, :]
    gxh = metric_at(gxh_id, met[gxh_id], 1, nx)[None, None, :]
    gzg = metric_at(M_GEOZG, met[M_GEOZG], 1, nz)[:, None, None]
    z1 = dxa / (v1 + cs)
    if mutation == "no GEOZG":
        z2 = dya / (v2 + cs / gxg)
    else:
        z2 = dya / (v2 + cs / gxg / gzg)
    z3 = dza / (v3 + cs / gxh)
    m1, m2, m3 = float(z1.min()), float(z2.min()), float(z3.min())
    return min(m1, m2, m3, p.dtmax) * p.fdelt, (m1, m2, m3)


MUTATIONS = ("GEOXG<->GEOXH", "DXA(i)", "DYA(j)", "DZA(k)", "V(i)", "G(j)", "H(k)", "no GEOZG")

# Grids on which each of the three Courant terms is the one that binds: the
# default (the radial one), a band next to the pole (cos(lat) small, so the
# phi term), and a narrow latitude band (the theta term).
SCHRITT_GRIDS = {
    "radial": dict(),
    "phi": dict(zmin=0.8, zmax=0.998),
    "theta": dict(zmin=-0.004, zmax=0.004),
}
