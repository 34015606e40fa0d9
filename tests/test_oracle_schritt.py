"""SCHRITT's Courant bound (ppsub.f:49-78,129-134) in the oracle, pinned by
its closed form (tests/closed_forms.py).  ppsub.f does not build here, so no
reference fixture holds SCHRITT's output; the closed form is written from
the reference's formulas in numpy on the same arrays.

The metrics SCHRITT reads are perturbed cell by cell and the velocities are
half the sound speed, so that index shifts and swapped arrays move the
bound; three grids make each of the three Courant terms the binding one in
turn.  Each planted mistake in the closed form is then shown to change the
bound on at least one grid, i.e. the comparison would catch that mistake in
the oracle."""
import numpy as np
import pytest

from oracle.ttoracle import Oracle
from tests import closed_forms as CF
from ttcode_amd import params as P

BINDS = {"radial": 0, "phi": 1, "theta": 2}
# which grid shows each planted mistake
CATCHES = {"GEOXG<->GEOXH": ("phi", "theta"), "DXA(i)": ("radial",), "DYA(j)": ("phi",),
           "DZA(k)": ("theta",), "V(i)": ("radial",), "G(j)": ("phi",), "H(k)": ("theta",),
           "no GEOZG": ("phi",)}


def _oracle_case(grid, seed=11):
    p = P.make_params(12, 16, 12, **CF.SCHRITT_GRIDS[grid])
    st = P.generate_state(p, seed=seed, vel_frac=0.5)
    o = Oracle(p, threads=1)
    o.load_state(st)
    o.pressg()
    met = CF.perturbed_metrics(o.metric, np.random.default_rng(seed + 100))
    for m, a in met.items():
        o.upload_metric(m, a)
    return p, st, met, o


@pytest.mark.parametrize("grid", list(CF.SCHRITT_GRIDS))
def test_schritt_cfl_closed_form(grid):
    p, st, met, o = _oracle_case(grid)
    got = o.schritt()
    want, mins = CF.schritt_cfl(p, st, met)
    assert int(np.argmin(mins)) == BINDS[grid], mins
    assert got == want, (got, want)


def test_schritt_cfl_closed_form_catches_planted_mistakes():
    cases = {g: _oracle_case(g) for g in CF.SCHRITT_GRIDS}
    dts = {g: c[3].schritt() for g, c in cases.items()}
    for mut in CF.MUTATIONS:
        for g in CATCHES[mut]:
            p, st, met, _ = cases[g]
            bad, _ = CF.schritt_cfl(p, st, met, mutation=mut)
            assert abs(bad - dts[g]) > 1e-6 * dts[g], (mut, g, bad, dts[g])
