"""Line sweeps with several Newton iterations per COUPLED_ENERGY_EQS.  With
line sweeps the library queues what follows each solve (the update, the
device's Newton test, the final T and delta_ER and tt_step's trailing
BOUNDS/PRESSG) ahead of the solve's control-block read, gated on device
flags; here the Newton loop goes on for several iterations, so those gates
hold the queued work back until the last one.  Against the CPU oracle's
FP64 solves: the same Newton iteration counts, and fields within the
solvers' tolerance (each inner solve reaches the same FP64 residual ratio
by another route, so results agree to about that ratio, not bit for bit)."""
import numpy as np
import pytest

from tests.golden import compare as C
from ttcode_amd import params as P

pytestmark = pytest.mark.gpu

KEYS = ("T", "RH", "V", "G", "H", "ER")


def _setup(tirr):
    from oracle.ttoracle import Oracle
    p = P.make_params(16, 32, 16, epsmat=1e-10, tirr=tirr)
    st = P.generate_state(p, seed=5, contrast=0.05)
    o = Oracle(p, threads=4)
    o.load_state(st)
    o.bounds()
    o.pressg()
    dt = o.schritt()
    o.load_state(st)
    return p, st, o, dt


def test_line_energy_solve_several_newton_iterations():
    from ttcode_amd import ttgpu
    p, st, o, dt = _setup(2000.0)
    o.bounds()
    o.pressg()
    so, rco = o.coupled_energy(10 * dt, 2)
    d = ttgpu.Device(p)
    try:
        d.load_state(st)
        d.set_solver(precision=1, inner=1)
        d.bounds()
        d.pressg()
        sg, rcg = d.coupled_energy(10 * dt, 2)
        print(f"Newton {sg.energy_iter} vs {so.energy_iter}, SOR {sg.sor_iter_total} vs {so.sor_iter_total}")
        assert so.energy_iter >= 3 and sg.energy_iter == so.energy_iter and rcg == rco == 0
        for k in ("T", "ER", "Q"):
            e = C.rel_err(d.download(k), o.download(k))
            print(f"  {k}: {e:.2e}")
            assert e <= 1e-8, (k, e)
        # delta_ER = ER - ER_old cancels where ER barely moves: on ER's scale
        er = o.download("ER")
        e = float(np.max(np.abs(d.download("delta_ER") - o.download("delta_ER")) / np.abs(er).max()))
        print(f"  delta_ER on the scale of ER: {e:.2e}")
        assert e <= 1e-8, e
    finally:
        d.close()


def test_line_steps_several_newton_iterations():
    from ttcode_amd import ttgpu
    p, st, o, dt = _setup(3000.0)
    d = ttgpu.Device(p)
    try:
        d.load_state(st)
        d.set_solver(precision=1, inner=1)
        for s in range(2):
            so, _ = o.step(2 * dt, 2 + s)
            sg, _ = d.step(2 * dt, 2 + s)
            print(f"step {s}: Newton {sg.energy.energy_iter} vs {so.energy.energy_iter}")
            assert so.energy.energy_iter >= 3 and sg.energy.energy_iter == so.energy.energy_iter
            assert np.array_equal(d.download("upperbnd"), o.download("upperbnd"))
            for k in KEYS + ("PG",):
                e = C.rel_err(d.download(k), o.download(k))
                print(f"  {k}: {e:.2e}")
                assert e <= 1e-8, (s, k, e)
    finally:
        d.close()
