"""Host-side binding of libttgpu.so, the gfx950 implementation of ttcode's
per-step hot path.

``Device`` mirrors the reference's routines (ADVECT3D, FORGAS3D,
COUPLED_ENERGY_EQS, SOLVE_WITH_SOR, BOUNDS, SCHRITT, DIFFUSION) on a
device-resident copy of the module globals; arrays cross in the reference's
wire format (see ttcode_amd.params).  There is no CPU fallback: constructing
a Device without the built library or without a GPU raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import params as P

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libttgpu.so")

_lib = None


def load_library():
    """Load the in-tree libttgpu.so (built by ttcode_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m ttcode_amd.build` first")
    L = ctypes.CDLL(LIB_PATH)
    L.tt_last_error.restype = ctypes.c_char_p
    L.tt_init.argtypes = [ctypes.POINTER(P.TTParams)]
    L.tt_upload.argtypes = [ctypes.c_int, ctypes.c_void_p]
    L.tt_download.argtypes = [ctypes.c_int, ctypes.c_void_p]
    L.tt_upload_metric.argtypes = [ctypes.c_int, ctypes.c_void_p]
    L.tt_download_metric.argtypes = [ctypes.c_int, ctypes.c_void_p]
    L.tt_advect3d.argtypes = [ctypes.c_double]
    L.tt_forgas3d.argtypes = [ctypes.c_double]
    L.tt_coupled_energy.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.POINTER(P.EnergyStats)]
    L.tt_step.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.POINTER(P.StepStats)]
    L.tt_solve_sor.argtypes = [ctypes.c_int, ctypes.POINTER(P.SorStats)]
    L.tt_schritt.argtypes = [ctypes.POINTER(ctypes.c_double)]
    L.tt_diffusion.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_int)]
    L.tt_set_solver.argtypes = [ctypes.POINTER(P.SolverOpts)]
    L.tt_get_solver.argtypes = [ctypes.POINTER(P.SolverOpts)]
    L.tt_sweep_timing.argtypes = [ctypes.c_int]
    L.tt_sor_speculation.argtypes = [ctypes.c_double]
    L.tt_sor_speculation_get.argtypes = [ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    L.tt_adjust_sor_param.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]
    L.tt_set_sorparam.argtypes = [ctypes.c_double]
    L.tt_get_line_omega.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)]
    L.tt_set_limits.argtypes = [ctypes.POINTER(P.Limits)]
    L.tt_get_limits.argtypes = [ctypes.POINTER(P.Limits)]
    L.tt_varlimit.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
    L.tt_sphericalvellimit.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
    L.tt_set_opacity_table.argtypes = [ctypes.c_void_p] * 3
    L.tt_set_burrows_table.argtypes = [ctypes.c_void_p] * 3
    L.tt_set_burrows_wave_table.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double]
    L.tt_set_column_table.argtypes = [ctypes.c_void_p]
    L.tt_set_viscosity.argtypes = [ctypes.POINTER(P.Viscosity)]
    L.tt_get_viscosity.argtypes = [ctypes.POINTER(P.Viscosity)]
    L.tt_get_dt_artvis.argtypes = [ctypes.POINTER(ctypes.c_double)]
    L.tt_get_sorparam.restype = ctypes.c_double
    L.tt_device_bytes.restype = ctypes.c_int64
    L.tt_sweep_timing_get.argtypes = [ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double)]
    L.tt_cnn_num_weights.restype = ctypes.c_int64
    L.tt_cnn_num_biases.restype = ctypes.c_int64
    fp = ctypes.POINTER(ctypes.c_float)
    L.tt_cnn_set_weights.argtypes = [fp, ctypes.c_int64, fp, ctypes.c_int64]
    L.tt_cnn_get_weights.argtypes = [fp, ctypes.c_int64, fp, ctypes.c_int64]
    L.tt_cnn_set_norm.argtypes = [fp]
    L.tt_cnn_get_norm.argtypes = [fp]
    L.tt_cnn_eval_host.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, fp, fp, fp, fp, fp, fp]
    L.tt_cnn_reserve.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.tt_cnn_collect.argtypes = [ctypes.c_int]
    L.tt_cnn_train.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint64, fp]
    _lib = L
    return L


class TTError(RuntimeError):
    pass


class Device:
    """One device-resident ttcode state (one rank's whole subdomain)."""

    def __init__(self, p: P.TTParams):
        self.L = load_library()
        self.p = p
        self._check(self.L.tt_init(ctypes.byref(p)), "tt_init")
        self.last_rc = 0

    def close(self):
        self.L.tt_finalize()

    def _check(self, rc, what):
        self.last_rc = rc
        if rc > 0:
            raise TTError(f"{what}: {self.L.tt_last_error().decode()} (rc={rc})")
        return rc

    # ---- data movement
    def upload(self, f, a):
        if isinstance(f, str):
            f = P.FIELD_IDS[f]
        shape, _, dt = P.wire_shape(f, self.p.nx, self.p.ny, self.p.nz)
        a = np.ascontiguousarray(a, dtype=dt)
        if a.shape != shape:
            raise ValueError(f"field {f}: shape {a.shape} != {shape}")
        self._check(self.L.tt_upload(f, a.ctypes.data), "tt_upload")

    def download(self, f):
        if isinstance(f, str):
            f = P.FIELD_IDS[f]
        out = P.empty_field(f, self.p.nx, self.p.ny, self.p.nz)
        self._check(self.L.tt_download(f, out.ctypes.data), "tt_download")
        return out

    def metric(self, m):
        lo, hi = ctypes.c_int32(), ctypes.c_int32()
        self.L.tt_metric_info(m, ctypes.byref(lo), ctypes.byref(hi))
        out = np.zeros(hi.value - lo.value + 1)
        self._check(self.L.tt_download_metric(m, out.ctypes.data), "tt_download_metric")
        return out

    def upload_metric(self, m, a):
        """Replace one 1-D metric array (declared bounds, as tt_metric_info)."""
        lo, hi = ctypes.c_int32(), ctypes.c_int32()
        self.L.tt_metric_info(m, ctypes.byref(lo), ctypes.byref(hi))
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != (hi.value - lo.value + 1,):
            raise ValueError(f"metric {m}: shape {a.shape} != ({hi.value - lo.value + 1},)")
        self._check(self.L.tt_upload_metric(m, a.ctypes.data), "tt_upload_metric")

    def load_state(self, st: dict):
        for name, a in st.items():
            if name in P.FIELD_IDS:
                self.upload(P.FIELD_IDS[name], a)

    # ---- solver options
    def set_solver(self, precision=0, guess=0, batch=16, inner_max=4000, inner_tol=1e-5, warm_scale=0.9,
                   adjust_omega=0, inner=0, line_omega=0.0):
        """tt_set_solver; line_omega 0 lets the first line solve choose the
        line sweeps' over-relaxation by a search (get_line_omega)."""
        o = P.SolverOpts(precision, guess, batch, inner_max, adjust_omega, inner, inner_tol, warm_scale, line_omega)
        self._check(self.L.tt_set_solver(ctypes.byref(o)), "tt_set_solver")

    def adjust_sor_param(self, fill_matrix=False):
        """ADJUST_SOR_PARAM (ppsor.f:332-354); returns (omega, iterations
        of the 50 trial solves)."""
        om = ctypes.c_double()
        it = np.zeros(50, np.int32)
        self._check(self.L.tt_adjust_sor_param(1 if fill_matrix else 0, ctypes.byref(om), it.ctypes.data),
                    "ADJUST_SOR_PARAM")
        return om.value, it

    def set_limits(self, irho=0, rhmin=0.0, itemp=0, tempmin=0.0, tempmax=0.0, ivel=0, velmax=0.0,
                   hydrotime=0.0):
        """Clamps tt_step applies (VARLIMIT / SPHERICALVELLIMIT, pp.f:543-562,
        673-678, 708-710)."""
        lim = P.Limits(irho, itemp, ivel, 0, rhmin, tempmax, tempmin, velmax, hydrotime)
        self._check(self.L.tt_set_limits(ctypes.byref(lim)), "tt_set_limits")

    def get_limits(self):
        lim = P.Limits()
        self._check(self.L.tt_get_limits(ctypes.byref(lim)), "tt_get_limits")
        return lim

    def varlimit(self, field, vmin, vmax):
        if isinstance(field, str):
            field = P.FIELD_IDS[field]
        self._check(self.L.tt_varlimit(field, vmin, vmax), "VARLIMIT")

    def sphericalvellimit(self, ivel, velmax, hydrotime=0.0):
        self._check(self.L.tt_sphericalvellimit(ivel, velmax, hydrotime), "SPHERICALVELLIMIT")

    def set_sorparam(self, omega):
        self._check(self.L.tt_set_sorparam(float(omega)), "tt_set_sorparam")

    def get_sorparam(self):
        return self.L.tt_get_sorparam()

    def get_line_omega(self):
        """(omega of the line sweeps, trial solves of the search that chose it)"""
        om = ctypes.c_double()
        n = ctypes.c_int32()
        self._check(self.L.tt_get_line_omega(ctypes.byref(om), ctypes.byref(n)), "tt_get_line_omega")
        return om.value, n.value

    # ---- the reference's routines
    def bounds(self):
        return self._check(self.L.tt_bounds(), "BOUNDS")

    def set_opacity_table(self, kap_planck, kap_ross, kap_abs):
        """The log10 tables of IKAPTYP 4/5 on the 42 x 18 (T, P) grid,
        temperature fastest (``ttcode_amd.opacity.read_freedman_table``
        reads the reference's table files)."""
        t = [np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1)) for a in (kap_planck, kap_ross, kap_abs)]
        if any(a.size != 42 * 18 for a in t):
            raise ValueError("opacity tables must hold 42 x 18 values")
        self._check(self.L.tt_set_opacity_table(*(a.ctypes.data_as(ctypes.c_void_p) for a in t)),
                    "tt_set_opacity_table")

    def set_column_table(self, kap_abs):
        """The IKAPTYP 6/7 column-depth absorption table, 42 x 18 x 8 log10
        values, temperature fastest then pressure then level
        (``ttcode_amd.opacity.read_column_table``)."""
        a = np.ascontiguousarray(np.asarray(kap_abs, np.float64).reshape(-1))
        if a.size != 42 * 18 * 8:
            raise ValueError("the column table must hold 42 x 18 x 8 values")
        self._check(self.L.tt_set_column_table(a.ctypes.data_as(ctypes.c_void_p)), "tt_set_column_table")

    def set_burrows_table(self, kap_planck, kap_ross, kap_abs):
        """The log10 tables of IKAPTYP 9-12/14 on the 50 x 50 (log T,
        log rho) grid, temperature fastest, as READBURROWSOPCTABLE leaves
        them (``ttcode_amd.opacity.read_burrows_table``)."""
        t = [np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1)) for a in (kap_planck, kap_ross, kap_abs)]
        if any(a.size != 50 * 50 for a in t):
            raise ValueError("Burrows tables must hold 50 x 50 values")
        self._check(self.L.tt_set_burrows_table(*(a.ctypes.data_as(ctypes.c_void_p) for a in t)),
                    "tt_set_burrows_table")

    def set_burrows_wave_table(self, kap_wave, bstar, dilution):
        """IKAPTYP 20-25: log10 of the 30 irradiation bins on the Burrows grid
        (T fastest, then rho, then bin), BSTAR_BIN and the dilution
        (``ttcode_amd.opacity.read_burrows_wave_table``, ``read_bstar``,
        ``wave_dilution``)."""
        kw = np.ascontiguousarray(np.asarray(kap_wave, np.float64).reshape(-1))
        bs = np.ascontiguousarray(np.asarray(bstar, np.float64).reshape(-1))
        if kw.size != 50 * 50 * 30 or bs.size != 30:
            raise ValueError("the wavelength table holds 50 x 50 x 30 values and BSTAR_BIN 30")
        self._check(self.L.tt_set_burrows_wave_table(kw.ctypes.data_as(ctypes.c_void_p),
                                                     bs.ctypes.data_as(ctypes.c_void_p), float(dilution)),
                    "tt_set_burrows_wave_table")

    def set_viscosity(self, ivis=1, ivistyp=1, ieta=4, idiss=1, xnue=1e9, iartvis=0):
        """ETACALC + VISC inside FORGAS3D (ppfor.f:170-174), ivis 0 is off;
        iartvis 1 adds NR_ARTIFICIAL_VISC at its end (ppfor.f:374-395)."""
        v = P.Viscosity(ivis, ivistyp, ieta, idiss, xnue, iartvis)
        self._check(self.L.tt_set_viscosity(ctypes.byref(v)), "tt_set_viscosity")

    def dt_artvis(self):
        """dt_artvis of the last NR_ARTIFICIAL_VISC (DTMAX before the first)."""
        d = ctypes.c_double()
        self._check(self.L.tt_get_dt_artvis(ctypes.byref(d)), "tt_get_dt_artvis")
        return d.value

    def get_viscosity(self):
        v = P.Viscosity()
        self._check(self.L.tt_get_viscosity(ctypes.byref(v)), "tt_get_viscosity")
        return v

    def kappa(self):
        """KAPPA + SUBSIG (pprphy.f:1-360,567-583)."""
        return self._check(self.L.tt_kappa(), "KAPPA")

    def pressg(self):
        return self._check(self.L.tt_pressg(), "PRESSG")

    def sound(self):
        return self._check(self.L.tt_sound(), "SOUND")

    def qdens(self):
        return self._check(self.L.tt_qdens(), "QDENS")

    def advect3d(self, delt):
        return self._check(self.L.tt_advect3d(delt), "ADVECT3D")

    def forgas3d(self, delt):
        return self._check(self.L.tt_forgas3d(delt), "FORGAS3D")

    def coupled_energy(self, delt, iii=1):
        st = P.EnergyStats()
        rc = self._check(self.L.tt_coupled_energy(delt, iii, ctypes.byref(st)), "COUPLED_ENERGY_EQS")
        return st, rc

    def solve_sor(self, fill=True):
        st = P.SorStats()
        rc = self._check(self.L.tt_solve_sor(1 if fill else 0, ctypes.byref(st)), "SOLVE_WITH_SOR")
        return st, rc

    def set_iii(self, iii):
        """III of the step loop as BOUNDS reads it (tt_set_step)."""
        self._check(self.L.tt_set_step(int(iii)), "tt_set_step")

    def schritt(self):
        d = ctypes.c_double()
        self._check(self.L.tt_schritt(ctypes.byref(d)), "SCHRITT")
        return d.value

    def diffusion(self, delt, xnue, pr_num):
        it = ctypes.c_int()
        rc = self._check(self.L.tt_diffusion(delt, xnue, pr_num, ctypes.byref(it)), "DIFFUSION")
        return it.value, rc

    def step(self, delt, iii=1):
        st = P.StepStats()
        rc = self._check(self.L.tt_step(delt, iii, ctypes.byref(st)), "step")
        return st, rc

    def sync(self):
        return self._check(self.L.tt_sync(), "tt_sync")

    def device_bytes(self):
        """Device memory held by the library for this grid."""
        return int(self.L.tt_device_bytes())

    def energy_totals(self):
        """(integral rho c_v T dV, integral E dV, integral STELLARINPUT dV)."""
        out = (ctypes.c_double * 3)()
        self._check(self.L.tt_energy_totals(out), "tt_energy_totals")
        return float(out[0]), float(out[1]), float(out[2])

    # ---- learned initial guess
    @staticmethod
    def _fptr(a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if a is not None else None

    def cnn_set_weights(self, w, b):
        w = np.ascontiguousarray(w, dtype=np.float32)
        b = np.ascontiguousarray(b, dtype=np.float32)
        self._check(self.L.tt_cnn_set_weights(self._fptr(w), w.size, self._fptr(b), b.size), "tt_cnn_set_weights")

    def cnn_get_weights(self):
        w = np.zeros(self.L.tt_cnn_num_weights(), np.float32)
        b = np.zeros(self.L.tt_cnn_num_biases(), np.float32)
        self._check(self.L.tt_cnn_get_weights(self._fptr(w), w.size, self._fptr(b), b.size), "tt_cnn_get_weights")
        return w, b

    def cnn_set_norm(self, norm12):
        a = np.ascontiguousarray(norm12, dtype=np.float32)
        self._check(self.L.tt_cnn_set_norm(self._fptr(a)), "tt_cnn_set_norm")

    def cnn_get_norm(self):
        a = np.zeros(12, np.float32)
        self._check(self.L.tt_cnn_get_norm(self._fptr(a)), "tt_cnn_get_norm")
        return a

    def cnn_eval(self, x0, target=None):
        """x0: (nz, ny, nx, 5) normalised features.  Returns y (nz, ny, nx)
        and, with a target, (loss, grad_w, grad_b)."""
        nz, ny, nx, _ = x0.shape
        x = np.ascontiguousarray(x0, dtype=np.float32)
        y = np.zeros((nz, ny, nx), np.float32)
        if target is None:
            self._check(self.L.tt_cnn_eval_host(nx, ny, nz, self._fptr(x), None, self._fptr(y), None, None, None),
                        "tt_cnn_eval_host")
            return y
        t = np.ascontiguousarray(target, dtype=np.float32)
        loss = np.zeros(1, np.float32)
        gw = np.zeros(self.L.tt_cnn_num_weights(), np.float32)
        gb = np.zeros(self.L.tt_cnn_num_biases(), np.float32)
        self._check(self.L.tt_cnn_eval_host(nx, ny, nz, self._fptr(x), self._fptr(t), self._fptr(y), self._fptr(loss),
                                            self._fptr(gw), self._fptr(gb)), "tt_cnn_eval_host")
        return y, float(loss[0]), gw, gb

    def cnn_reserve(self, max_samples, nx, ny, nz):
        self._check(self.L.tt_cnn_reserve(max_samples, nx, ny, nz), "tt_cnn_reserve")

    def cnn_collect(self, enable=True):
        self._check(self.L.tt_cnn_collect(1 if enable else 0), "tt_cnn_collect")

    def cnn_dataset_size(self):
        return int(self.L.tt_cnn_dataset_size())

    def cnn_fit_norm(self):
        self._check(self.L.tt_cnn_fit_norm(), "tt_cnn_fit_norm")

    def cnn_train(self, epochs, batch=8, lr=1e-3, seed=0):
        losses = np.zeros(epochs, np.float32)
        self._check(self.L.tt_cnn_train(epochs, batch, lr, seed, self._fptr(losses)), "tt_cnn_train")
        return losses

    def cnn_release(self):
        self.L.tt_cnn_release()

    # ---- per-launch timing of the SOR half sweeps
    def sweep_timing(self, period=1):
        """Time every period-th SOR half-sweep launch with HIP events (0 or
        False: off)."""
        self.L.tt_sweep_timing(int(period))

    def sor_speculation(self, margin):
        """Speculative convergence tests of the FP64 solve (margin > 0: the
        factor, 0: always, < 0: off); resets the counters."""
        self.L.tt_sor_speculation(float(margin))

    def sor_speculation_get(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self.L.tt_sor_speculation_get(ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def sweep_timing_get(self):
        n64, n32 = ctypes.c_int64(), ctypes.c_int64()
        m64, m32 = ctypes.c_double(), ctypes.c_double()
        self.L.tt_sweep_timing_get(ctypes.byref(n64), ctypes.byref(m64), ctypes.byref(n32), ctypes.byref(m32))
        return n64.value, m64.value, n32.value, m32.value
