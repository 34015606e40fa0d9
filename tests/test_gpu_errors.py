"""The reference's error paths on the GPU, forced from the host, each
against the oracle on the same input:
- ITMAX reached by SOR and by the Newton loop: not fatal, the reference
  prints and goes on (ppsor.f:161-174, ppenergy.f:750-775) -- a negative
  return, and the state equals the oracle's after the same capped solves;
- a non-positive Q after the damped update (ppenergy.f:411-521), the
  reference's CLEAN_STOP -- a positive return;
- a zero right-hand side: the stopping ratio is 0/0, which no comparison
  passes, so the solve stops after ITMIN = 1 iteration with delta_ER = 0
  (ppsor.f:176)."""
import numpy as np
import pytest

from oracle.ttoracle import Oracle
from tests.golden import compare as C

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ttgpu():
    from ttcode_amd import ttgpu as T
    T.load_library()
    return T


def test_itmax_cap_is_informational(ttgpu):
    g = C.load("energy_const_12x16x12")
    p = C.params_of(g)
    p.itmax = 6
    o = Oracle(p, threads=1)
    o.load_state(C.inputs_of(g))
    so, rco = o.coupled_energy(float(g["delt"]), int(g["iii"]))
    d = ttgpu.Device(p)
    try:
        d.load_state(C.inputs_of(g))
        sg, rcg = d.coupled_energy(float(g["delt"]), int(g["iii"]))
        print(f"oracle rc {rco}, GPU rc {rcg}; Newton {sg.energy_iter}, SOR {sg.sor_iter_total}")
        assert rco < 0 and rcg == rco
        assert sg.energy_iter == so.energy_iter == p.itmax
        assert sg.sor_iter_total == so.sor_iter_total == p.itmax * p.itmax
        for k in ("T", "ER"):
            reg = C.compare_region(k, p.nx, p.ny, p.nz)
            e = C.rel_err(d.download(k)[reg], o.download(k)[reg])
            print(f"  {k}: {e:.2e}")
            assert e <= 1e-12, (k, e)
    finally:
        d.close()


def test_nonpositive_q_stops(ttgpu):
    g = C.load("energy_const_12x16x12")
    p = C.params_of(g)
    inp = dict(C.inputs_of(g))
    T = inp["T"].copy()
    T[6, 8, 5] = -100.0  # cell (i, j, k) = (4, 7, 5): Q = rho c_v T < 0
    inp["T"] = T
    o = Oracle(p, threads=1)
    o.load_state(inp)
    with pytest.raises(RuntimeError, match="DELTA_Q"):
        o.coupled_energy(float(g["delt"]), int(g["iii"]))
    d = ttgpu.Device(p)
    try:
        d.load_state(inp)
        with pytest.raises(ttgpu.TTError, match="DELTA_Q") as e:
            d.coupled_energy(float(g["delt"]), int(g["iii"]))
        assert d.last_rc > 0
        print(e.value)
    finally:
        d.close()


def test_zero_rhs_stops_after_itmin(ttgpu):
    g = C.load("sor_12x16x12")
    p = C.params_of(g)
    sys_in = {k: g["in_" + k] for k in ("EVAL1", "EVAL2", "EVAL3", "EVAL4", "EVAL5", "EVAL6", "EVAL7", "EVAL8",
                                         "EVAL_1_6", "upperbnd")}
    sys_in["RHS_ER"] = np.zeros_like(g["in_RHS_ER"])
    o = Oracle(p, threads=1)
    o.load_state(C.inputs_of(g))
    for k, a in sys_in.items():
        o.upload(k, a)
    so, _ = o.solve_sor(True)
    d = ttgpu.Device(p)
    try:
        d.load_state(C.inputs_of(g))
        for k, a in sys_in.items():
            d.upload(k, a)
        sg, rc = d.solve_sor(True)
        print(f"SOR_ITER oracle {so.sor_iter}, GPU {sg.sor_iter}; rhsnorm {sg.rhsnorm}")
        assert so.sor_iter == sg.sor_iter == 1 and rc == 0
        assert sg.rhsnorm == 0.0
        assert not d.download("delta_ER").any()
    finally:
        d.close()


def test_itmax_cap_in_the_defect_correction(ttgpu):
    """With FP32 line sweeps the outer loop's caps are decided on the device:
    an ITMAX below what the solve needs ends the corrections early, the
    return is the informational cap (negative) and the state stays finite;
    the same solve uncapped converges."""
    g = C.load("energy_const_12x16x12")
    p = C.params_of(g)
    d = ttgpu.Device(p)
    try:
        d.set_solver(precision=1, inner=1)
        d.load_state(C.inputs_of(g))
        sf, rcf = d.coupled_energy(float(g["delt"]), int(g["iii"]))
        d.load_state(C.inputs_of(g))
        p2 = C.params_of(g)
        p2.itmax = 2
        d.close()
        d = ttgpu.Device(p2)
        d.set_solver(precision=1, inner=1)
        d.load_state(C.inputs_of(g))
        sc, rcc = d.coupled_energy(float(g["delt"]), int(g["iii"]))
        print(f"uncapped rc {rcf}, Newton {sf.energy_iter}, FP32 iterations {sf.fp32_sweeps}; "
              f"capped rc {rcc}, Newton {sc.energy_iter}, FP32 iterations {sc.fp32_sweeps}")
        assert rcf == 0 and rcc < 0
        for k in ("T", "ER"):
            assert np.isfinite(d.download(k)).all()
    finally:
        d.close()
