"""EVAL1..6 and EVAL_1_6 after COUPLED_ENERGY_EQS.  The library forms them
in registers for the SOR matrix and writes their module arrays only when
they are read, from DIFR and the metrics (ppenergy.f:984-1030).  These tests
hold that to the reference's semantics: the arrays read back equal the
oracle's after the solve, and overwriting DIFR or a metric from the host
afterwards does not change them (the reference's arrays keep the values of
the last FILL_EVALUES)."""
import numpy as np
import pytest

from oracle.ttoracle import Oracle
from tests.golden import compare as C

pytestmark = pytest.mark.gpu

EV = ("EVAL1", "EVAL2", "EVAL3", "EVAL4", "EVAL5", "EVAL6", "EVAL_1_6")


@pytest.fixture(scope="module")
def ttgpu():
    from ttcode_amd import ttgpu as T
    T.load_library()
    return T


def _solved(ttgpu):
    g = C.load("energy_const_12x16x12")
    p = C.params_of(g)
    o = Oracle(p, threads=1)
    o.load_state(C.inputs_of(g))
    o.coupled_energy(float(g["delt"]), int(g["iii"]))
    d = ttgpu.Device(p)
    d.load_state(C.inputs_of(g))
    d.coupled_energy(float(g["delt"]), int(g["iii"]))
    return o, d


def test_evals_read_back_equal_the_oracle(ttgpu):
    o, d = _solved(ttgpu)
    try:
        for k in EV + ("EVAL7", "EVAL8"):
            e = C.rel_err(d.download(k), o.download(k))
            print(f"{k}: {e:.2e}")
            assert e <= 1e-12, (k, e)
    finally:
        d.close()


def test_evals_survive_host_writes_of_their_inputs(ttgpu):
    o, d = _solved(ttgpu)
    try:
        want = {k: o.download(k) for k in EV}
        dx = d.download("difrx")
        d.upload("difrx", 3.0 * dx + 1.0)          # the EVALs still hold the solve's values
        m_xxb = 1  # TT_M_XXB
        d.upload_metric(m_xxb, 1.5 * d.metric(m_xxb))
        for k in EV:
            e = C.rel_err(d.download(k), want[k])
            assert e <= 1e-12, (k, e)
        # the arrays are ordinary module arrays again: an upload sticks
        z = np.zeros_like(want["EVAL3"])
        d.upload("EVAL3", z)
        assert np.array_equal(d.download("EVAL3"), z)
        assert C.rel_err(d.download("EVAL1"), want["EVAL1"]) <= 1e-12
    finally:
        d.close()
