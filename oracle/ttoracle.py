"""ctypes binding of oracle/libttoracle.so, the CPU FP64 restatement of the
ttcode hot path.  TEST INFRASTRUCTURE ONLY: imported by tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg, never by the product
path (ttcode_amd/)."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from ttcode_amd import params as P

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libttoracle.so")


def build(force: bool = False) -> str:
    """Compile oracle/ttoracle.c (gcc, no FMA contraction, OpenMP)."""
    if force or not os.path.exists(LIB):
        subprocess.check_call(["make", "-C", HERE, "-f", "Makefile", "liboracle"])
    return LIB


class Oracle:
    """The oracle holds one state (like the reference's module globals)."""

    _lib = None

    def __init__(self, p: P.TTParams, threads: int = 1):
        if Oracle._lib is None:
            build()
            Oracle._lib = ctypes.CDLL(LIB)
            L = Oracle._lib
            L.tor_last_error.restype = ctypes.c_char_p
            for name in ("tor_advect3d", "tor_forgas3d"):
                getattr(L, name).argtypes = [ctypes.c_double]
            L.tor_coupled_energy.argtypes = [ctypes.c_double, ctypes.c_int,
                                             ctypes.POINTER(P.EnergyStats)]
            L.tor_step.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.POINTER(P.StepStats)]
            L.tor_solve_sor.argtypes = [ctypes.c_int, ctypes.POINTER(P.SorStats)]
            L.tor_adjust_sor_param.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]
            L.tor_get_sorparam.restype = ctypes.c_double
            L.tor_varlimit.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
            L.tor_sphericalvellimit.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
            L.tor_set_limits.argtypes = [ctypes.POINTER(P.Limits)]
            L.tor_get_limits.argtypes = [ctypes.POINTER(P.Limits)]
            L.tor_schritt.argtypes = [ctypes.POINTER(ctypes.c_double)]
            L.tor_diffusion.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.POINTER(ctypes.c_int)]
            L.tor_set_opacity_table.argtypes = [ctypes.c_void_p] * 3
            L.tor_set_burrows_table.argtypes = [ctypes.c_void_p] * 3
            L.tor_set_burrows_wave_table.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double]
            L.tor_set_column_table.argtypes = [ctypes.c_void_p]
            L.tor_set_viscosity.argtypes = [ctypes.POINTER(P.Viscosity)]
            L.tor_get_viscosity.argtypes = [ctypes.POINTER(P.Viscosity)]
            L.tor_get_dt_artvis.argtypes = [ctypes.POINTER(ctypes.c_double)]
        self.L = Oracle._lib
        self.p = p
        self.L.tor_set_threads(int(threads))
        self._check(self.L.tor_init(ctypes.byref(p)), "tor_init")

    def _check(self, rc, what):
        if rc > 0:
            raise RuntimeError(f"{what}: {self.L.tor_last_error().decode()} (rc={rc})")
        return rc

    def set_threads(self, n):
        self.L.tor_set_threads(int(n))

    # -- data movement (wire format) --
    def upload(self, f, a):
        if isinstance(f, str):
            f = P.FIELD_IDS[f]
        shape, _, dt = P.wire_shape(f, self.p.nx, self.p.ny, self.p.nz)
        a = np.ascontiguousarray(a, dtype=dt)
        assert a.shape == shape, (f, a.shape, shape)
        self._check(self.L.tor_upload(f, a.ctypes.data_as(ctypes.c_void_p)), "upload")

    def download(self, f):
        if isinstance(f, str):
            f = P.FIELD_IDS[f]
        out = P.empty_field(f, self.p.nx, self.p.ny, self.p.nz)
        self._check(self.L.tor_download(f, out.ctypes.data_as(ctypes.c_void_p)), "download")
        return out

    def metric(self, m):
        lo, hi = ctypes.c_int32(), ctypes.c_int32()
        self.L.tor_metric_info(m, ctypes.byref(lo), ctypes.byref(hi))
        out = np.zeros(hi.value - lo.value + 1)
        self.L.tor_download_metric(m, out.ctypes.data_as(ctypes.c_void_p))
        return out

    def upload_metric(self, m, a):
        lo, hi = ctypes.c_int32(), ctypes.c_int32()
        self.L.tor_metric_info(m, ctypes.byref(lo), ctypes.byref(hi))
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == (hi.value - lo.value + 1,), (m, a.shape)
        self._check(self.L.tor_upload_metric(m, a.ctypes.data_as(ctypes.c_void_p)), "upload_metric")

    def load_state(self, st: dict):
        for name, a in st.items():
            if name in P.FIELD_IDS:
                self.upload(P.FIELD_IDS[name], a)

    # -- routines --
    def bounds(self):
        return self._check(self.L.tor_bounds(), "BOUNDS")

    def set_opacity_table(self, kap_planck, kap_ross, kap_abs):
        """The log10 tables of IKAPTYP 4/5 on the 42 x 18 (T, P) grid,
        temperature fastest (held across tor_init, like the reference's
        opc_init module)."""
        t = [np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1)) for a in (kap_planck, kap_ross, kap_abs)]
        assert all(a.size == 42 * 18 for a in t)
        self._check(self.L.tor_set_opacity_table(*(a.ctypes.data_as(ctypes.c_void_p) for a in t)),
                    "set_opacity_table")

    def set_column_table(self, kap_abs):
        a = np.ascontiguousarray(np.asarray(kap_abs, np.float64).reshape(-1))
        assert a.size == 42 * 18 * 8
        self._check(self.L.tor_set_column_table(a.ctypes.data_as(ctypes.c_void_p)), "set_column_table")

    def set_burrows_table(self, kap_planck, kap_ross, kap_abs):
        """The log10 tables of IKAPTYP 9-12/14 on the 50 x 50 (log T,
        log rho) grid, temperature fastest."""
        t = [np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1)) for a in (kap_planck, kap_ross, kap_abs)]
        assert all(a.size == 50 * 50 for a in t)
        self._check(self.L.tor_set_burrows_table(*(a.ctypes.data_as(ctypes.c_void_p) for a in t)),
                    "set_burrows_table")

    def set_burrows_wave_table(self, kap_wave, bstar, dilution):
        """IKAPTYP 20-25: log10 of the 30 bins' irradiation opacity on the
        Burrows grid (T fastest, then rho, then bin), the bins' BSTAR_BIN and
        the dilution (RSTAR RSOL / semi)^2."""
        kw = np.ascontiguousarray(np.asarray(kap_wave, np.float64).reshape(-1))
        bs = np.ascontiguousarray(np.asarray(bstar, np.float64).reshape(-1))
        assert kw.size == 50 * 50 * 30 and bs.size == 30
        self._check(self.L.tor_set_burrows_wave_table(kw.ctypes.data_as(ctypes.c_void_p),
                                                      bs.ctypes.data_as(ctypes.c_void_p), float(dilution)),
                    "set_burrows_wave_table")

    def set_viscosity(self, ivis=1, ivistyp=1, ieta=4, idiss=1, xnue=1e9, iartvis=0):
        """ETACALC + VISC inside FORGAS3D (ppfor.f:170-174), off by default;
        iartvis 1 adds NR_ARTIFICIAL_VISC (ppfor.f:374-395)."""
        v = P.Viscosity(ivis, ivistyp, ieta, idiss, xnue, iartvis)
        self._check(self.L.tor_set_viscosity(ctypes.byref(v)), "set_viscosity")

    def dt_artvis(self):
        d = ctypes.c_double()
        self._check(self.L.tor_get_dt_artvis(ctypes.byref(d)), "dt_artvis")
        return d.value

    def kappa(self):
        """KAPPA + SUBSIG (pprphy.f:1-360,567-583) on the current state."""
        return self._check(self.L.tor_kappa(), "KAPPA")

    def pressg(self):
        return self._check(self.L.tor_pressg(), "PRESSG")

    def sound(self):
        return self._check(self.L.tor_sound(), "SOUND")

    def advect3d(self, delt):
        return self._check(self.L.tor_advect3d(delt), "ADVECT3D")

    def forgas3d(self, delt):
        return self._check(self.L.tor_forgas3d(delt), "FORGAS3D")

    def coupled_energy(self, delt, iii=1):
        st = P.EnergyStats()
        rc = self._check(self.L.tor_coupled_energy(delt, iii, ctypes.byref(st)), "COUPLED_ENERGY_EQS")
        return st, rc

    def solve_sor(self, fill=True):
        st = P.SorStats()
        rc = self._check(self.L.tor_solve_sor(1 if fill else 0, ctypes.byref(st)), "SOLVE_WITH_SOR")
        return st, rc

    def adjust_sor_param(self):
        """ADJUST_SOR_PARAM on the current EVALs; returns (omega, counts)."""
        om = ctypes.c_double()
        it = np.zeros(50, np.int32)
        self._check(self.L.tor_adjust_sor_param(ctypes.byref(om), it.ctypes.data_as(ctypes.c_void_p)),
                    "ADJUST_SOR_PARAM")
        return om.value, it

    def set_adjust(self, on):
        """Run ADJUST_SOR_PARAM inside COUPLED_ENERGY_EQS (ppenergy.f:372-374)."""
        self.L.tor_set_adjust(1 if on else 0)

    def get_sorparam(self):
        return self.L.tor_get_sorparam()

    def set_limits(self, irho=0, rhmin=0.0, itemp=0, tempmin=0.0, tempmax=0.0, ivel=0, velmax=0.0,
                   hydrotime=0.0):
        self.L.tor_set_limits(ctypes.byref(P.Limits(irho, itemp, ivel, 0, rhmin, tempmax, tempmin, velmax,
                                                    hydrotime)))

    def get_limits(self):
        lim = P.Limits()
        self.L.tor_get_limits(ctypes.byref(lim))
        return lim

    def varlimit(self, field, vmin, vmax):
        if isinstance(field, str):
            field = P.FIELD_IDS[field]
        return self._check(self.L.tor_varlimit(field, vmin, vmax), "VARLIMIT")

    def sphericalvellimit(self, ivel, velmax, hydrotime=0.0):
        return self._check(self.L.tor_sphericalvellimit(ivel, velmax, hydrotime), "SPHERICALVELLIMIT")

    def set_iii(self, iii):
        """III of the step loop as BOUNDS reads it (step and coupled_energy
        set it from their argument)."""
        self.L.tor_set_iii(int(iii))

    def schritt(self):
        d = ctypes.c_double()
        self._check(self.L.tor_schritt(ctypes.byref(d)), "SCHRITT")
        return d.value

    def diffusion(self, delt, xnue, pr_num):
        it = ctypes.c_int()
        rc = self._check(self.L.tor_diffusion(delt, xnue, pr_num, ctypes.byref(it)), "DIFFUSION")
        return it.value, rc

    def step(self, delt, iii=1):
        st = P.StepStats()
        rc = self._check(self.L.tor_step(delt, iii, ctypes.byref(st)), "step")
        return st, rc

    def field_info(self, f):
        lo = (ctypes.c_int32 * 3)()
        hi = (ctypes.c_int32 * 3)()
        rank, isint = ctypes.c_int32(), ctypes.c_int32()
        rc = self.L.tor_field_info(f, lo, hi, ctypes.byref(rank), ctypes.byref(isint))
        return rc, tuple(lo), tuple(hi), rank.value, isint.value
