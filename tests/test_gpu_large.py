"""The large coupled configuration on one GPU: 192x256x512 (n_r x n_theta x
n_phi, 25.2M cells) with the 5% day-night contrast, learned guess + FP32
sweeps in FP64 defect correction: two steps, the true FP64 residual of the
last system (re-solved) at or below 1e-10, the library's HBM footprint, and
the energy budget of the steps closing (tt_energy_totals)."""
import os

import pytest

from ttcode_amd import params as P

pytestmark = pytest.mark.gpu

NX, NY, NZ = 192, 512, 256
CNN_WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ttcode_amd", "data",
                           "cnn_guess.npz")


def test_large_config_c5():
    from tests.test_gpu_fullsize import _residual_ratio
    from ttcode_amd import cnn_train, ttgpu
    p = P.make_params(NX, NY, NZ, epsmat=1e-10)
    st = P.generate_state(p, seed=2024, contrast=0.05)
    d = ttgpu.Device(p)
    try:
        d.load_state(st)
        d.bounds()
        d.pressg()
        dt = d.schritt()
        d.load_state(st)
        cnn_train.load(d, CNN_WEIGHTS)
        d.set_solver(precision=1, guess=2)
        hbm = d.device_bytes()
        print(f"192x256x512: dt {dt:.4g} s, library HBM footprint {hbm / 1e9:.2f} GB", flush=True)
        assert 10e9 < hbm < 24e9
        for s in range(2):
            sg, rc = d.step(dt, 2 + s)
            e = sg.energy
            print(f"step {s}: {sg.ms_total:.1f} ms, Newton {e.energy_iter}, FP32 sweeps {e.fp32_sweeps}, "
                  f"learned used {e.learned_used}, fallbacks {e.fallbacks}", flush=True)
            assert rc == 0 and e.energy_iter >= 1
        d.set_solver(precision=1, guess=0)
        ss, _ = d.solve_sor(False)
        names = ("SOR_A", "SOR_B", "SOR_C", "SOR_D", "SOR_E", "SOR_F", "SOR_G", "RHS_ER")
        A, B, Cc, D, E, F, G, R = (d.download(k) for k in names)
        ratio = _residual_ratio(A, B, Cc, D, E, F, G, R, d.download("delta_ER"))
        print(f"re-solved last system: {ss.fp32_sweeps} FP32 sweeps, true residual ratio {ratio:.3e}")
        assert ratio <= p.epsmat, ratio
    finally:
        d.cnn_release()
        d.close()
