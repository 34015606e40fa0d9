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


def test_centcory_after_an_upload(ttgpu):
    """CENTRIFCORIOLIS sets CENTCORY to zero (ppfor.f:435-486); the library
    writes those zeros only after the host has put something else there, so
    an upload of CENTCORY is undone by the next FORGAS3D, as in the
    reference, and FORGAS3D's fields are the same either way."""
    g = C.load("forgas_12x16x12")
    p = C.params_of(g)
    d = ttgpu.Device(p)
    try:
        d.load_state(C.inputs_of(g))
        d.bounds()
        d.qdens()                 # RHQX/Y/Z, as the parity test of FORGAS3D prepares them
        d.forgas3d(float(g["delt"]))
        want = {k: d.download(k) for k in ("V", "G", "H", "T", "centcory")}
        assert not np.any(want["centcory"])
        assert all(np.all(np.isfinite(want[k])) for k in ("V", "G", "H", "T"))
        d.load_state(C.inputs_of(g))
        d.upload("centcory", np.ones_like(want["centcory"]))
        d.bounds()
        d.qdens()
        d.forgas3d(float(g["delt"]))
        for k, v in want.items():
            got = d.download(k)
            if k == "centcory":
                # declared (1:nx+1, 1:ny+1, 1:nz+1); CENTRIFCORIOLIS sets
                # the cells 1..n and leaves the last plane of each axis
                assert not np.any(got[:-1, :-1, :-1]), k
                assert np.all(got[-1] == 1.0) and np.all(got[:, -1] == 1.0) and np.all(got[:, :, -1] == 1.0), k
            else:
                assert np.array_equal(got, v), k
    finally:
        d.close()
