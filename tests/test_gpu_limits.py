"""The clamps of the step loop (VARLIMIT, ppsub.f:140-154; SPHERICALVELLIMIT,
ppsub.f:316-391) on the GPU against the CPU oracle: bit-identical for the
clamps themselves (comparisons and divisions by the same host-computed
metrics); a step with every clamp switched on within the bars of the plain
step test (1e-12 where the SOR counts are equal; where the L1 norm, summed in
another order, moves the stop by one iteration, the solver tolerance)."""
import numpy as np
import pytest

from tests.golden import compare as C
from ttcode_amd import params as P

pytestmark = pytest.mark.gpu

DAY = 86400.0
KEYS = ("T", "RH", "V", "G", "H")


def _state():
    p = P.make_params(16, 32, 16, ikaptyp=1, kap0=1e-1, epsmat=1e-8, tirr=2000.0)
    return p, P.generate_state(p, seed=41)


def _run_oracle(p, st, fn):
    from oracle.ttoracle import Oracle
    o = Oracle(p)
    o.load_state(st)
    out = {"sor": fn(o)}
    out.update({k: o.download(k) for k in KEYS})
    return out


def _run_gpu(p, st, fn):
    from ttcode_amd import ttgpu
    d = ttgpu.Device(p)
    d.load_state(st)
    out = {"sor": fn(d)}
    out.update({k: d.download(k) for k in KEYS})
    d.close()
    return out


@pytest.mark.parametrize("ivel,t", [(1, 0.0), (2, 7.0 * DAY), (2, 30 * DAY), (3, 0.0), (4, 12.5 * DAY)])
def test_clamps_match_oracle(ivel, t):
    p, st = _state()
    vm = 0.4 * float(np.abs(st["V"]).max())
    tlo, thi = float(np.quantile(st["T"], 0.25)), float(np.quantile(st["T"], 0.75))

    def fn(x):
        x.varlimit("T", tlo, thi)
        x.varlimit("RH", float(np.quantile(st["RH"], 0.5)), 1e10)
        x.sphericalvellimit(ivel, vm, t)
    want = _run_oracle(p, st, fn)
    got = _run_gpu(p, st, fn)
    for k in KEYS:
        assert np.array_equal(got[k], want[k]), k
    if ivel in (1, 2):
        assert not np.array_equal(got["G"], st["G"])


def test_step_with_limits_matches_oracle():
    p, st = _state()
    from oracle.ttoracle import Oracle
    o = Oracle(p)
    o.load_state(st)
    o.bounds()
    o.pressg()
    dt = o.schritt()
    lim = dict(irho=1, rhmin=float(np.quantile(st["RH"], 0.3)), itemp=1,
               tempmin=float(np.quantile(st["T"], 0.1)), tempmax=float(np.quantile(st["T"], 0.9)),
               ivel=2, velmax=0.5 * float(np.abs(st["V"]).max()), hydrotime=10 * DAY)

    def fn(x):
        x.set_limits(**lim)
        its = []
        for s in range(2):
            its.append(x.step(dt, 2 + s)[0].energy.sor_iter_total)
        assert x.get_limits().hydrotime == 10 * DAY + dt + dt
        return its
    want = _run_oracle(p, st, fn)
    got = _run_gpu(p, st, fn)
    same = got["sor"] == want["sor"]
    for k in KEYS:
        e = C.rel_err(got[k], want[k])
        print(f"{k}: max rel err {e:.2e} (SOR {got['sor']} vs {want['sor']})")
        assert e <= (1e-12 if same else p.epsmat), (k, e)
