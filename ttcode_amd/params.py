"""Run parameters, workload definitions and the seeded initial-condition
generator for the ttcode hot path.

``TTParams`` mirrors ``tt_params`` of include/ttgpu.h field for field (the
switches of ttcode's control file ``ipp`` that the hot path reads,
ppmod.f:2-78).  ``WORKLOADS`` holds the five workloads this project measures;
``generate_state`` builds their initial state in the reference's wire format
(Fortran order, declared bounds, ghosts included).

Arrays in the wire format are numpy arrays in C order with the axes reversed,
``a[k - lo_k, j - lo_j, i - lo_i]``, which is byte-for-byte the Fortran
column-major layout of ``A(i, j, k)``.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

# ppmod.f:1071-1093.  The reference writes REAL literals into DOUBLE
# PRECISION parameters, so the value in use is the nearest float.
F32 = lambda x: float(np.float32(x))  # noqa: E731
PI = 3.141592653589793
FD = F32(2.99792458e10)
ARAD = F32(7.56596e-15)
RGAS = F32(8.315e7)
SBCONST = F32(5.6705e-5)
RJUP = F32(7.1492e9)


class TTParams(ctypes.Structure):
    """``tt_params`` of include/ttgpu.h."""

    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("bcxmin", ctypes.c_int32), ("ibdrad", ctypes.c_int32),
        ("iflimtyp", ctypes.c_int32), ("f_incident", ctypes.c_int32),
        ("ikaptyp", ctypes.c_int32), ("isigtyp", ctypes.c_int32),
        ("isoth", ctypes.c_int32), ("itmax", ctypes.c_int32), ("igrav", ctypes.c_int32),
        ("xmin", ctypes.c_double), ("xmax", ctypes.c_double),
        ("ymin", ctypes.c_double), ("ymax", ctypes.c_double),
        ("zmin", ctypes.c_double), ("zmax", ctypes.c_double),
        ("gamma", ctypes.c_double), ("mu_gas", ctypes.c_double),
        ("zkapmult", ctypes.c_double), ("kap0", ctypes.c_double),
        ("dkap0", ctypes.c_double), ("kap_rho0", ctypes.c_double),
        ("kap_t0", ctypes.c_double), ("kap_a", ctypes.c_double),
        ("kap_b", ctypes.c_double), ("tirr", ctypes.c_double),
        ("fbottom", ctypes.c_double), ("d_t", ctypes.c_double),
        ("omrot", ctypes.c_double), ("sorparam", ctypes.c_double),
        ("epsmat", ctypes.c_double), ("fdelt", ctypes.c_double),
        ("dtmax", ctypes.c_double), ("gamma_opc", ctypes.c_double),
        ("mstar2", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class SolverOpts(ctypes.Structure):
    """``tt_solver_opts``."""

    _fields_ = [("precision", ctypes.c_int32), ("guess", ctypes.c_int32),
                ("batch", ctypes.c_int32), ("inner_max", ctypes.c_int32),
                ("adjust_omega", ctypes.c_int32), ("inner", ctypes.c_int32),
                ("inner_tol", ctypes.c_double), ("warm_scale", ctypes.c_double),
                ("line_omega", ctypes.c_double)]


class Viscosity(ctypes.Structure):
    """``tt_viscosity``."""

    _fields_ = [("ivis", ctypes.c_int32), ("ivistyp", ctypes.c_int32), ("ieta", ctypes.c_int32),
                ("idiss", ctypes.c_int32), ("xnue", ctypes.c_double), ("iartvis", ctypes.c_int32),
                ("pad0", ctypes.c_int32)]


class Limits(ctypes.Structure):
    """``tt_limits``."""

    _fields_ = [("irho", ctypes.c_int32), ("itemp", ctypes.c_int32), ("ivel", ctypes.c_int32),
                ("pad0", ctypes.c_int32), ("rhmin", ctypes.c_double), ("tempmax", ctypes.c_double),
                ("tempmin", ctypes.c_double), ("velmax", ctypes.c_double), ("hydrotime", ctypes.c_double)]


class SorStats(ctypes.Structure):
    """``tt_sor_stats``."""

    _fields_ = [("sor_iter", ctypes.c_int32), ("outer_iter", ctypes.c_int32),
                ("fp32_sweeps", ctypes.c_int32), ("guess_used", ctypes.c_int32),
                ("rhsnorm", ctypes.c_double), ("rnorm", ctypes.c_double),
                ("r0_guess", ctypes.c_double), ("r0_cold", ctypes.c_double),
                ("guess_scale", ctypes.c_double)]


class EnergyStats(ctypes.Structure):
    """``tt_energy_stats``."""

    _fields_ = [("energy_iter", ctypes.c_int32), ("sor_iter_total", ctypes.c_int32),
                ("sor_iter_last", ctypes.c_int32), ("sor_iter_first", ctypes.c_int32),
                ("fp32_sweeps", ctypes.c_int32), ("fallbacks", ctypes.c_int32),
                ("learned_used", ctypes.c_int32), ("learned_skipped", ctypes.c_int32),
                ("max_dq", ctypes.c_double), ("max_de", ctypes.c_double),
                ("guess_scale_first", ctypes.c_double), ("guess_r0_first", ctypes.c_double),
                ("first_start", ctypes.c_int32), ("first_choice", ctypes.c_int32),
                ("ms_build", ctypes.c_double), ("ms_sor", ctypes.c_double), ("ms_update", ctypes.c_double),
                ("ms_sor_first", ctypes.c_double)]


class StepStats(ctypes.Structure):
    """``tt_step_stats``."""

    _fields_ = [("energy", EnergyStats), ("ms_advect", ctypes.c_double),
                ("ms_force", ctypes.c_double), ("ms_energy", ctypes.c_double),
                ("ms_total", ctypes.c_double)]


# Field ids (include/ttgpu.h)
T, RH, V, G, H, ER, DELTA_ER, GD = 1, 2, 3, 4, 5, 6, 8, 11
PG, CS, CV, RHQX, RHQY, RHQZ = 20, 21, 22, 23, 24, 25
XKAPR, XKAPP, XKAPA, DXKAPP_DT, DXKAPA_DT, SIG = 26, 27, 28, 29, 30, 31
GRAVX, GRAVY, GRAVZ, STELLARINPUT, STELLARTAU, DTAUS, ER_OLD = 32, 33, 34, 35, 36, 37, 38
RHS_ER = 40
SOR_A, SOR_B, SOR_C, SOR_D, SOR_E, SOR_F, SOR_G = range(41, 48)
EVAL1, EVAL2, EVAL3, EVAL4, EVAL5, EVAL6 = range(50, 56)
EVAL7, EVAL8, EVAL_1_6, JACOB_ETA, FQ, FR = 56, 57, 58, 59, 60, 61
FLIMX, FLIMY, FLIMZ, DIFRX, DIFRY, DIFRZ = 62, 63, 64, 65, 66, 67
FLUX_X, FLUX_Y, FLUX_Z, CENTCORX, CENTCORY, CENTCORZ = 68, 69, 70, 71, 72, 73
UPPERBND, MU_STAR, Q = 80, 81, 82
ETA, DIVTX, DIVTY, DIVTZ, DISFN = 83, 84, 85, 86, 87
CONV_FLUX1, CONV_FLUX2, CONV_FLUX3 = 88, 89, 90

FIELD_NAMES = {
    T: "T", RH: "RH", V: "V", G: "G", H: "H", ER: "ER", DELTA_ER: "delta_ER", GD: "GD",
    PG: "PG", CS: "CS", CV: "CV", RHQX: "RHQX", RHQY: "RHQY", RHQZ: "RHQZ",
    XKAPR: "xkapR", XKAPP: "xkapP", XKAPA: "xkapA", DXKAPP_DT: "dxkapP_dT",
    DXKAPA_DT: "dxkapA_dT", SIG: "sig", GRAVX: "gravx", GRAVY: "gravy", GRAVZ: "gravz",
    STELLARINPUT: "STELLARINPUT", STELLARTAU: "STELLARTAU", DTAUS: "DTAUS",
    ER_OLD: "ER_OLD", RHS_ER: "RHS_ER", SOR_A: "SOR_A", SOR_B: "SOR_B", SOR_C: "SOR_C",
    SOR_D: "SOR_D", SOR_E: "SOR_E", SOR_F: "SOR_F", SOR_G: "SOR_G",
    EVAL1: "EVAL1", EVAL2: "EVAL2", EVAL3: "EVAL3", EVAL4: "EVAL4", EVAL5: "EVAL5",
    EVAL6: "EVAL6", EVAL7: "EVAL7", EVAL8: "EVAL8", EVAL_1_6: "EVAL_1_6",
    JACOB_ETA: "Jacob_eta", FQ: "FQ", FR: "FR", FLIMX: "FLIMX", FLIMY: "FLIMY",
    FLIMZ: "FLIMZ", DIFRX: "difrx", DIFRY: "difry", DIFRZ: "difrz", FLUX_X: "flux_x",
    FLUX_Y: "flux_y", FLUX_Z: "flux_z", CENTCORX: "centcorx", CENTCORY: "centcory",
    CENTCORZ: "centcorz", UPPERBND: "upperbnd", MU_STAR: "mu_star", Q: "Q",
    ETA: "eta", DIVTX: "divtx", DIVTY: "divty", DIVTZ: "divtz", DISFN: "DISFN",
    CONV_FLUX1: "conv_flux1", CONV_FLUX2: "conv_flux2", CONV_FLUX3: "conv_flux3",
}
FIELD_IDS = {v: k for k, v in FIELD_NAMES.items()}


def field_bounds(f: int, nx: int, ny: int, nz: int):
    """Declared Fortran bounds (lo, hi inclusive, per axis i, j, k) and dtype.
    Mirror of include/ttgpu_fields.h (checked against it by the tests)."""
    n = (nx, ny, nz)
    lo = [-1, -1, -1]
    hi = [nx + 2, ny + 2, nz + 2]
    if f in (V, RHQX):
        lo[0] = 0
    elif f in (G, RHQY):
        lo[1] = 0
    elif f in (H, RHQZ):
        lo[2] = 0
    elif f in (SIG, DISFN):
        lo, hi = [0, 0, 0], [nx + 1, ny + 1, nz + 1]
    elif f in (DIVTX, DIVTY, DIVTZ):
        lo, hi = [0, 0, 0], [nx + 1, ny + 1, nz + 1]
        lo[f - DIVTX] = 1
    elif f == STELLARINPUT:
        lo[0], hi[0] = 1, nx
    elif f in (STELLARTAU, DTAUS):
        lo[0], hi[0] = 1, nx + 1
    elif f in (RHS_ER, SOR_A, SOR_B, SOR_C, SOR_D, SOR_E, SOR_F, SOR_G, EVAL1, EVAL2,
               EVAL3, EVAL4, EVAL5, EVAL6, EVAL7, EVAL8, EVAL_1_6, JACOB_ETA, FQ, FR,
               CONV_FLUX1, CONV_FLUX2, CONV_FLUX3):
        lo, hi = [1, 1, 1], list(n)
    elif f in (FLIMX, DIFRX, FLUX_X):
        lo, hi = [1, 1, 1], [nx + 1, ny, nz]
    elif f in (FLIMY, DIFRY, FLUX_Y):
        lo, hi = [1, 1, 1], [nx, ny + 1, nz]
    elif f in (FLIMZ, DIFRZ, FLUX_Z):
        lo, hi = [1, 1, 1], [nx, ny, nz + 1]
    elif f in (CENTCORX, CENTCORY, CENTCORZ):
        lo, hi = [1, 1, 1], [nx + 1, ny + 1, nz + 1]
    elif f in (UPPERBND, MU_STAR):
        return (-1, -1), (ny + 2, nz + 2), (np.int32 if f == UPPERBND else np.float64)
    return tuple(lo), tuple(hi), np.float64


def wire_shape(f: int, nx: int, ny: int, nz: int):
    lo, hi, dt = field_bounds(f, nx, ny, nz)
    return tuple(h - l + 1 for l, h in zip(lo, hi))[::-1], lo, dt


def empty_field(f: int, nx: int, ny: int, nz: int):
    shape, _, dt = wire_shape(f, nx, ny, nz)
    return np.zeros(shape, dtype=dt)


def ghosted_view(a: np.ndarray, lo):
    """Index helper: returns a function (i, j, k) -> a[k-lo_k, j-lo_j, i-lo_i]."""
    def at(i, j, k):
        return a[k - lo[2], j - lo[1], i - lo[0]]
    return at


# --------------------------------------------------------------- workloads
@dataclass(frozen=True)
class Workload:
    name: str
    nx: int  # n_r
    ny: int  # n_phi
    nz: int  # n_theta
    epsmat: float
    steps: int
    contrast: float = 0.0  # day-night T contrast (cos phi), C5 only


# n_r x n_theta x n_phi in the workload names; (nx, ny, nz) = (n_r, n_phi, n_theta)
WORKLOADS = {
    "sph32x32x64": Workload("sph32x32x64", 32, 64, 32, 1e-8, 1),
    "sph128x128x256": Workload("sph128x128x256", 128, 256, 128, 1e-10, 100),
    "sph128x128x256-mixed": Workload("sph128x128x256-mixed", 128, 256, 128, 1e-10, 100),
    "sph128x128x256-learned": Workload("sph128x128x256-learned", 128, 256, 128, 1e-10, 100),
    "sph192x256x512": Workload("sph192x256x512", 192, 512, 256, 1e-10, 200, 0.05),
}

# Physical set-up of the generator (values the workloads leave open are
# fixed here and documented: bottom density, opacity normalisation and
# irradiation temperature).
R_P = RJUP
SCALE_HEIGHT = 0.01 * R_P
T0 = 1500.0
RHO0 = 1.0e-4        # g/cm^3 at r = R_p (> 2.2e-5 keeps rho >= 1e-9 in the box)
KAP0 = 1.0e-2        # cm^2/g at (rho0, T0)
KAP_A, KAP_B = 0.5, 2.0
TIRR = 1500.0        # K
GAMMA, MU_GAS = 1.4, 2.3


def make_params(nx, ny, nz, *, epsmat=1e-8, itmax=20000, sorparam=1.7,
                kap0=KAP0, kap_a=KAP_A, kap_b=KAP_B, rho0=RHO0, tirr=TIRR,
                ikaptyp=26, f_incident=2, iflimtyp=2, ibdrad=7, **over) -> TTParams:
    p = TTParams()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.bcxmin = 12
    p.ibdrad = ibdrad
    p.iflimtyp = iflimtyp
    p.f_incident = f_incident
    p.ikaptyp = ikaptyp
    p.isigtyp = 0
    p.isoth = 0
    p.itmax = itmax
    p.xmin, p.xmax = R_P, 1.1 * R_P
    p.ymin, p.ymax = 0.0, 1.0
    p.zmin, p.zmax = -0.5, 0.5
    p.gamma, p.mu_gas = GAMMA, MU_GAS
    p.zkapmult = kap0
    p.gamma_opc = 0.1  # the shipped ipp's gamma_OPC (ipp:11)
    p.igrav = 1        # constant radial gravity, uploaded as gravx (generate_state)
    p.mstar2 = 0.0
    p.kap0 = kap0
    p.kap_rho0 = rho0
    p.kap_t0 = T0
    p.kap_a, p.kap_b = kap_a, kap_b
    p.dkap0 = kap_b * kap0 / T0
    p.tirr = tirr
    p.fbottom = 0.0
    p.d_t = 0.0
    p.omrot = 0.0
    p.sorparam = sorparam
    p.epsmat = epsmat
    p.fdelt = 0.4
    p.dtmax = 1.0e12
    for k, v in over.items():
        setattr(p, k, v)
    return p


def grid_1d(n, lo_val, hi_val):
    """Cell faces a(0:n+2) and centres b(-1:n+2) by the recurrences of GITTER
    (ppinit.f:1083-1098 for r, :1132-1150 for phi, :1159-1177 for theta)."""
    d = (hi_val - lo_val) / n
    a = np.zeros(n + 3)          # index i -> a[i], i = 0..n+2
    a[1] = lo_val + (0 * (hi_val - lo_val) / 1)
    a[0] = a[1] - d
    for i in range(2, n + 3):
        a[i] = a[i - 1] + d
    b = np.zeros(n + 4)          # index i -> b[i+1], i = -1..n+2
    for i in range(0, n + 2):
        b[i + 1] = (a[i] + a[i + 1]) / 2.0
    b[n + 3] = (a[n + 2] + (a[n + 2] + d * 1.0)) / 2.0
    b[0] = (a[0] + (a[0] - d / 1.0)) / 2.0
    return a, b


def generate_state(p: TTParams, seed: int = 1234, contrast: float = 0.0,
                   vel_frac: float = 0.01, rho0: float = RHO0):
    """Initial state of the workloads (isothermal hydrostatic atmosphere with
    seeded velocity noise), as wire-format arrays keyed by field name.

    rho = rho0 exp(-(r - R_p)/H), T = T0 (1 + contrast cos phi), ER = a T^4,
    gravx = -Rgas T0 / (mu H) (IGRAV=1 constant radial gravity, ppfor.f:682),
    V ~ N(0, (vel_frac c_s)^2), G = v_phi / (r cos lat), H = v_theta / r
    (angular velocities, ppadv.f:1082,1283).  Ghost cells are filled from the
    same formulas; the first BOUNDS call imposes the boundary conditions.
    """
    nx, ny, nz = p.nx, p.ny, p.nz
    xa, xb = grid_1d(nx, p.xmin, p.xmax)
    ya, yb = grid_1d(ny, p.ymin * 2.0 * PI, p.ymax * 2.0 * PI)
    za, zb = grid_1d(nz, (p.zmin / 2.0) * PI, (p.zmax / 2.0) * PI)
    rng = np.random.default_rng(seed)
    cs = math.sqrt(p.gamma * RGAS * T0 / p.mu_gas)
    sig = vel_frac * cs

    out = {}
    sh = (nz + 4, ny + 4, nx + 4)
    r = xb[None, None, :]                         # i = -1..nx+2
    phi = yb[None, :, None]
    lat = zb[:, None, None]
    rho = rho0 * np.exp(-(r - R_P) / SCALE_HEIGHT) * np.ones(sh)
    temp = T0 * (1.0 + contrast * np.cos(phi)) * np.ones(sh)
    out["RH"] = np.ascontiguousarray(rho)
    out["T"] = np.ascontiguousarray(temp)
    out["ER"] = ARAD * out["T"] ** 4
    gx = -RGAS * T0 / (p.mu_gas * SCALE_HEIGHT)
    out["gravx"] = np.full(sh, gx)
    # V(0:nx+2, -1:ny+2, -1:nz+2)
    out["V"] = sig * rng.standard_normal((nz + 4, ny + 4, nx + 3))
    # G(-1:nx+2, 0:ny+2, -1:nz+2): at (r_b(i), lat_b(k))
    vphi = sig * rng.standard_normal((nz + 4, ny + 3, nx + 4))
    out["G"] = vphi / (xb[None, None, :] * np.cos(zb)[:, None, None])
    # H(-1:nx+2, -1:ny+2, 0:nz+2)
    vth = sig * rng.standard_normal((nz + 3, ny + 4, nx + 4))
    out["H"] = vth / xb[None, None, :]
    # mu_star (CALCMUSTAR, ppsub.f:156-179) for the reference harness
    ang = np.cos(yb)[None, :] * np.cos(zb)[:, None]
    out["mu_star"] = np.maximum(ang, F32(1e-7))
    return out
