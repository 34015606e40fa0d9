"""The over-relaxation of the radial-line sweeps chosen by trial solves
(tt_solver_opts.line_omega = 0, the default): the first line solve after
tt_set_solver runs cold-start solves of its system over omega in [1, 1.95]
and keeps the one with the fewest line iterations, as ADJUST_SOR_PARAM
(ppsor.f:332-354) keeps SORPARAM's.  Checked here: the choice is reported,
it takes no more iterations than omega = 1.1 on the same system, the solve
with the chosen omega is bit-identical to a solve with that omega fixed, and
the step meets the FP64 step within the solvers' tolerance."""
import numpy as np
import pytest

from ttcode_amd import params as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ttgpu():
    from ttcode_amd import ttgpu as T
    T.load_library()
    return T


def _device(ttgpu, nx, ny, nz, seed, contrast=0.0):
    p = P.make_params(nx, ny, nz, epsmat=1e-10)
    st = P.generate_state(p, seed=seed, contrast=contrast)
    d = ttgpu.Device(p)
    d.load_state(st)
    d.bounds()
    d.pressg()
    dt = d.schritt()
    return d, st, dt


def _step(d, st, dt, **opts):
    d.load_state(st)
    d.set_solver(precision=1, inner=1, **opts)
    s, _ = d.step(dt, 2)
    return s, {k: d.download(k) for k in ("ER", "T")}


@pytest.mark.parametrize("shape", [(32, 64, 32), (48, 128, 64)])
def test_line_omega_search(ttgpu, shape):
    d, st, dt = _device(ttgpu, *shape, seed=5, contrast=0.05)
    try:
        d.load_state(st)
        d.set_solver(precision=0)
        s64, _ = d.step(dt, 2)
        f64 = {k: d.download(k) for k in ("ER", "T")}

        sa, fa = _step(d, st, dt)                       # the search, then the solve
        om, trials = d.get_line_omega()
        na = sa.energy.sor_iter_total
        print(f"{shape}: chosen omega {om:.2f} after {trials} trial solves, {na} line iterations")
        assert 1.0 <= om <= 1.95 and trials >= 5

        # omega fixed at the chosen value: the same solve, bit for bit
        sf, ff = _step(d, st, dt, line_omega=om)
        assert d.get_line_omega() == (om, 0)
        assert sf.energy.sor_iter_total == na
        for k in ("ER", "T"):
            assert np.array_equal(ff[k], fa[k]), k

        # omega 1.1 on the same system, with the budget of its own previous
        # solve (the second of two), takes at least as many iterations
        _step(d, st, dt, line_omega=1.1)
        s11, _ = _step(d, st, dt, line_omega=1.1)
        print(f"omega 1.1: {s11.energy.sor_iter_total} line iterations")
        assert na <= s11.energy.sor_iter_total

        # the FP64 step, within the solvers' tolerance (both reach the FP64
        # residual ratio 1e-10 of a system whose solution is a small
        # correction of E; measured about 1e-13)
        for k in ("ER", "T"):
            a, b = fa[k][2:-2, 2:-2, 2:-2], f64[k][2:-2, 2:-2, 2:-2]
            assert np.max(np.abs(a - b) / np.abs(b)) <= 1e-11, k
        assert s64.energy.energy_iter == sa.energy.energy_iter
    finally:
        d.close()


def test_line_omega_option_checks(ttgpu):
    d, st, dt = _device(ttgpu, 16, 32, 16, seed=3)
    try:
        with pytest.raises(RuntimeError):
            d.set_solver(precision=1, inner=1, line_omega=-0.5)
        with pytest.raises(RuntimeError):
            d.set_solver(precision=1, inner=1, line_omega=2.0)
        d.set_solver(precision=1, inner=1)
        assert d.get_line_omega() == (0.0, 0)        # before the first line solve
    finally:
        d.close()
