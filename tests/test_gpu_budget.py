"""The energy budget of COUPLED_ENERGY_EQS on the GPU (tests/energy_budget.py):
the per-cell Newton identity, the telescoping of the divergence, and the
closure of the solved region's energy change to the SOR tolerance, at
32x64x32 and at 128x128x256 (size-independent properties), in both solver
precisions."""
import ctypes

import pytest

from tests import energy_budget as EB
from ttcode_amd import params as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ttgpu():
    from ttcode_amd import ttgpu as T
    T.load_library()
    return T


def _run(ttgpu, nx, ny, nz, seed, precision=0, inner=0):
    p = P.make_params(nx, ny, nz, epsmat=1e-10)
    st = P.generate_state(p, seed=seed)
    d = ttgpu.Device(p)
    try:
        d.load_state(st)
        d.bounds()
        d.pressg()
        dt = d.schritt()
        d.load_state(st)
        d.bounds()
        d.pressg()
        d.set_solver(precision=precision, inner=inner)

        def info(m):
            lo, hi = ctypes.c_int32(), ctypes.c_int32()
            d.L.tt_metric_info(m, ctypes.byref(lo), ctypes.byref(hi))
            return lo.value

        q0, e0 = EB.state_q_e(d.download)
        s, rc = d.coupled_energy(dt, 2)
        assert rc == 0 and s.energy_iter == 1, (rc, s.energy_iter)
        return EB.budget(d.download, d.metric, info, nx, ny, nz, dt, q0, e0)
    finally:
        d.close()


@pytest.mark.parametrize("shape,seed,precision,inner", [((32, 64, 32), 3, 0, 0), ((32, 64, 32), 3, 1, 1),
                                                        ((128, 256, 128), 2024, 0, 0), ((128, 256, 128), 2024, 1, 1)],
                         ids=["32x64x32-fp64", "32x64x32-line", "128x256x128-fp64", "128x256x128-line"])
def test_energy_budget_closes(ttgpu, shape, seed, precision, inner):
    b = _run(ttgpu, *shape, seed, precision, inner)
    closure = (b["sum_w_dQdE"] - b["predicted"]) / b["scale"]
    print(f"{shape} prec {precision} inner {inner}: identity {b['identity_err']:.2e}, closure {closure:.2e}, "
          f"residual term {b['sum_res'] / b['scale']:.2e}, outflow dt*Phi {b['dt'] * b['phi_faces']:.3e}, "
          f"source dt*S {b['dt'] * b['S']:.3e}, change in R {b['change_in_R']:.3e}, "
          f"boundary blocks {b['change_outside_R']:.3e}")
    assert b["identity_err"] < 1e-8
    assert abs(b["sum_w_divF"] - b["phi_faces"]) <= 1e-12 * abs(b["phi_faces"])
    assert abs(b["sum_w_FQFR"] - b["dt"] * (b["sum_w_divF"] - b["S"])) <= 1e-12 * abs(b["dt"] * b["sum_w_divF"])
    assert abs(EB.CONV_PARM * b["sum_w_dQdE"] - b["change_in_R"]) <= 1e-12 * abs(b["change_in_R"])
    assert abs(b["sum_w_dQdE"] - b["predicted"] - b["sum_res"]) <= 1e-9 * b["scale"]
    assert b["sum_abs_res_bound"] <= 1e-7 * b["scale"]
