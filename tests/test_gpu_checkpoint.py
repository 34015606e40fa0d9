"""Checkpoints of the device state in the reference's formats: PRINTOUT(1)
written from the GPU after one energy solve against the reference's own
file, and a global restart file carried through a second device."""
import numpy as np
import pytest

from tests.golden import compare as C
from ttcode_amd import checkpoint as K

pytestmark = pytest.mark.gpu


def _device_after_solve(g):
    from ttcode_amd import ttgpu
    p = C.params_of(g)
    d = ttgpu.Device(p)
    d.load_state(C.inputs_of(g))
    d.coupled_energy(float(g["delt"]), int(g["iii"]))
    d.pressg()
    d.sound()
    return p, d


def _scalars(g):
    sc = {k[5:]: g[k].item() for k in g.files if k.startswith("ckpt_")}
    init = (sc["h_tinitbottom"], sc["h_tinittop"], sc["h_rhinittop"], sc["h_rhinitbottom"],
            sc["h_hydrotime"], sc["h_maccreated"])
    return sc, init


def test_printout_from_the_device(tmp_path):
    """Same records and sizes as the reference's file; hydro fields and
    upperbnd equal, the energy solve's fields within 1e-12."""
    g = C.load("checkpoint_12x16x12")
    sc, init = _scalars(g)
    p, d = _device_after_solve(g)
    try:
        fn = tmp_path / "SNAP_gpu"
        K.write_printout(fn, d, p, sc["h_zeit"], float(g["delt"]), int(sc["h_num_iter"]), init)
    finally:
        d.close()
    assert fn.stat().st_size == g["snap_bytes"].size
    ref_fn = tmp_path / "SNAP_ref"
    ref_fn.write_bytes(g["snap_bytes"].tobytes())
    mine, ref = K.read_printout(fn), K.read_printout(ref_fn)
    assert mine.keys() == ref.keys()
    for k, v in ref.items():
        if isinstance(v, np.ndarray):
            if k in ("V", "G", "H", "RH", "upperbnd", "xxa", "xxb", "dxa", "xya", "xyb", "dya", "xza", "xzb",
                     "dza", "conv_flux_1", "conv_flux_2", "conv_flux_3", "DISFN", "xkapR", "xkapA", "xkapP"):
                np.testing.assert_array_equal(mine[k], v, err_msg=k)
            else:
                e = C.rel_err(mine[k], v)
                print(f"  {k}: {e:.2e}")
                assert e <= 1e-12, (k, e)
        else:
            assert mine[k] == v, k


def test_restart_through_a_second_device(tmp_path):
    """write_restart from one device, read_restart + restart_state (ghost
    cells from the run's initial model) into a fresh one, BOUNDS (as
    LOCALINIT does): the second device steps like the first."""
    from ttcode_amd import ttgpu
    g = C.load("checkpoint_12x16x12")
    sc, init = _scalars(g)
    p, d = _device_after_solve(g)
    try:
        fn = tmp_path / "AKTTD"
        K.write_restart(fn, d, p, sc["h_zeit"], float(g["delt"]), int(sc["h_num_iter"]), init)
        d.bounds()
        st1, _ = d.step(float(g["delt"]), 3)
        after1 = {k: d.download(k) for k in ("T", "ER", "RH", "V")}
    finally:
        d.close()
    r = K.read_restart(fn, p.nx, p.ny, p.nz)
    inp = C.inputs_of(g)
    st = K.restart_state(r, p, base=inp)
    d2 = ttgpu.Device(p)
    try:
        d2.load_state(st)
        d2.bounds()
        d2.pressg()
        st2, _ = d2.step(float(g["delt"]), 3)
        assert st2.energy.energy_iter == st1.energy.energy_iter
        for k, v in after1.items():
            reg = C.compare_region(k, p.nx, p.ny, p.nz)
            e = C.rel_err(d2.download(k)[reg], v[reg])
            print(f"  after a step, {k}: {e:.2e}")
            assert e <= 1e-12, (k, e)
    finally:
        d2.close()


def test_printout_after_a_viscous_step_from_the_device(tmp_path):
    """PRINTOUT(1) from the device after a full viscous step: the records
    of the reference's file, conv_flux (ADVECT3D's thermal fluxes) and DISFN
    (FORGAS3D's dissipation) included, within 1e-12."""
    from ttcode_amd import ttgpu
    g = C.load("checkpoint_visc_step_12x16x12")
    p = C.params_of(g)
    sc, init = _scalars(g)
    d = ttgpu.Device(p)
    try:
        d.load_state(C.inputs_of(g))
        d.set_viscosity(int(g["vis_h_ivis"]), int(g["vis_h_ivistyp"]), int(g["vis_h_ieta"]),
                        int(g["vis_h_idiss"]), float(g["vis_xnue"]))
        st, _ = d.step(float(g["delt"]), int(g["iii"]))
        assert st.energy.energy_iter == int(g["scal_energy_iter"])
        assert st.energy.sor_iter_last == int(g["scal_SOR_ITER"])
        d.pressg()
        d.sound()
        fn = tmp_path / "SNAP_gpu"
        K.write_printout(fn, d, p, sc["h_zeit"], float(g["delt"]), int(sc["h_num_iter"]), init)
    finally:
        d.close()
    ref_fn = tmp_path / "SNAP_ref"
    ref_fn.write_bytes(g["snap_bytes"].tobytes())
    mine, ref = K.read_printout(fn), K.read_printout(ref_fn)
    assert mine.keys() == ref.keys()
    for k, v in ref.items():
        if isinstance(v, np.ndarray):
            if v.dtype.kind == "i":
                np.testing.assert_array_equal(mine[k], v, err_msg=k)
                continue
            e = C.rel_err(mine[k], v)
            print(f"  {k}: {e:.2e}")
            assert e <= 1e-12, (k, e)
        else:
            assert mine[k] == v, k
    assert np.count_nonzero(ref["DISFN"]) and np.count_nonzero(ref["conv_flux_1"])
