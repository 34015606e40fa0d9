"""Checkpoint files in the reference's formats (ttcode_amd/checkpoint.py)
against the reference itself: the fixture holds the PRINTOUT(1) file the
reference wrote after one energy solve, the global restart file this
library wrote from the oracle's state after the same solve, and what the
reference's READRESTART read from it.  The oracle is bit-identical to the
reference, so the files must agree byte for byte."""
import numpy as np
import pytest

from oracle.ttoracle import Oracle
from tests.golden import compare as C
from ttcode_amd import checkpoint as K


@pytest.fixture(scope="module")
def case():
    g = C.load("checkpoint_12x16x12")
    p = C.params_of(g)
    o = Oracle(p, threads=1)
    o.load_state(C.inputs_of(g))
    o.coupled_energy(float(g["delt"]), int(g["iii"]))
    o.pressg()
    o.sound()
    sc = {k[5:]: g[k].item() for k in g.files if k.startswith("ckpt_")}
    init = (sc["h_tinitbottom"], sc["h_tinittop"], sc["h_rhinittop"], sc["h_rhinitbottom"],
            sc["h_hydrotime"], sc["h_maccreated"])
    return g, p, o, sc, init


def test_printout_is_the_references_file(case, tmp_path):
    """PRINTOUT(1) (ppprt.f:13-124): the library's file equals the
    reference's byte for byte."""
    g, p, o, sc, init = case
    fn = tmp_path / "SNAP_000.0001"
    K.write_printout(fn, o, p, sc["h_zeit"], float(g["delt"]), int(sc["h_num_iter"]), init)
    mine = fn.read_bytes()
    ref = g["snap_bytes"].tobytes()
    assert len(mine) == len(ref)
    assert mine == ref


def test_read_printout(case, tmp_path):
    g, p, o, sc, init = case
    fn = tmp_path / "SNAP_ref"
    fn.write_bytes(g["snap_bytes"].tobytes())
    d = K.read_printout(fn)
    assert d["global_n"] == (p.nx, p.ny, p.nz) and d["numprocs"] == 1 and d["proc_dims"] == (1, 1, 1)
    assert d["ZEIT"] == sc["h_zeit"] and d["NUM_ITER"] == sc["h_num_iter"] and d["DELT"] == float(g["delt"])
    assert d["TinitBottom"] == sc["h_tinitbottom"] and d["rhinittop"] == sc["h_rhinittop"]
    assert d["NCOSYS"] == 2
    for name in ("V", "G", "H", "RH", "T", "ER", "PG"):
        np.testing.assert_array_equal(d[name], o.download(name), err_msg=name)
    np.testing.assert_array_equal(d["T"][2:-2, 2:-2, 2:-2], C.outputs_of(g)["T"][2:-2, 2:-2, 2:-2])
    np.testing.assert_array_equal(d["xkapR"], o.download("xkapR")[2:-2, 2:-2, 2:-2])
    np.testing.assert_array_equal(d["upperbnd"], o.download("upperbnd")[2:-2, 2:-2])
    np.testing.assert_array_equal(d["xxb"], o.metric(1))
    assert not d["DISFN"].any() and not d["conv_flux_2"].any()


def test_restart_file_reads_in_the_reference(case, tmp_path):
    """The global restart file the library writes is the one the fixture's
    reference run read (byte for byte), READRESTART (ppinit.f:1563-1598)
    filled its arrays with the oracle's interior cells, and read_restart
    reads back the same values."""
    g, p, o, sc, init = case
    fn = tmp_path / "AKTTD"
    K.write_restart(fn, o, p, sc["h_zeit"], float(g["delt"]), int(sc["h_num_iter"]), init)
    assert fn.read_bytes() == g["akttd_bytes"].tobytes()
    r = K.read_restart(fn, p.nx, p.ny, p.nz)
    for name in ("T", "ER", "RH", "V", "G", "H"):
        ref = g["ref_rs_" + name].reshape(p.nz, p.ny, p.nx)  # Fortran order (1:nx, 1:ny, 1:nz)
        np.testing.assert_array_equal(r[name], ref, err_msg=name)
    np.testing.assert_array_equal(r["upperbnd"], g["ref_rs_upperbnd"].reshape(p.nz, p.ny))
    assert r["ZEIT"] == g["ref_scal_ZEIT"].item() and r["NUM_ITER"] == g["ref_scal_NUM_ITER"].item()
    assert r["DELT"] == g["ref_scal_DELT"].item()
    for k in K.INIT_SCALARS:
        assert r[k] == g["ref_scal_" + k].item(), k
    # LOCALINIT's copy into the ghosted arrays (ppinit.f:23-49)
    st = K.restart_state(r, p)
    for name in ("T", "ER", "RH"):
        np.testing.assert_array_equal(st[name][2:-2, 2:-2, 2:-2], o.download(name)[2:-2, 2:-2, 2:-2])


def test_damaged_files_are_refused(case, tmp_path):
    g, p, o, sc, init = case
    b = g["akttd_bytes"].tobytes()
    fn = tmp_path / "AKTTD"
    fn.write_bytes(b[:-3])
    with pytest.raises(ValueError):
        K.read_restart(fn, p.nx, p.ny, p.nz)
    fn.write_bytes(b)
    with pytest.raises(ValueError, match="grid"):
        K.read_restart(fn, p.nx + 2, p.ny, p.nz)
    bad = bytearray(g["snap_bytes"].tobytes())
    bad[-1] ^= 0xFF  # the last record's trailing length
    fn.write_bytes(bytes(bad))
    with pytest.raises(ValueError, match="trailing length"):
        K.read_printout(fn)


def test_restart_continues_the_run(case, tmp_path):
    """A run restarted from the library's restart file, with the ghost cells
    of its initial model (BOUNDS never sets RH below the bottom for
    BCXMIN=12), steps exactly as the run it was written from."""
    g, p, _, sc, init = case
    dt = float(g["delt"])
    o = Oracle(p, threads=1)  # its own run: the shared one stays at the solve
    o.load_state(C.inputs_of(g))
    o.coupled_energy(dt, int(g["iii"]))
    o.pressg()
    o.sound()
    fn = tmp_path / "AKTTD"
    K.write_restart(fn, o, p, sc["h_zeit"], dt, int(sc["h_num_iter"]), init)
    o.bounds()
    s1, _ = o.step(dt, 3)
    after = {k: o.download(k) for k in ("T", "ER", "RH", "V", "G", "H")}
    o2 = Oracle(p, threads=1)
    o2.load_state(K.restart_state(K.read_restart(fn, p.nx, p.ny, p.nz), p, base=C.inputs_of(g)))
    o2.bounds()
    o2.pressg()
    s2, _ = o2.step(dt, 3)
    assert s2.energy.energy_iter == s1.energy.energy_iter
    for k, v in after.items():
        np.testing.assert_array_equal(o2.download(k), v, err_msg=k)
    # without the initial model the bottom ghost densities are zero, and the
    # step differs
    o3 = Oracle(p, threads=1)
    o3.load_state(K.restart_state(K.read_restart(fn, p.nx, p.ny, p.nz), p,
                                  base={k: C.inputs_of(g)[k] for k in ("gravx", "mu_star")}))
    o3.bounds()
    o3.pressg()
    o3.step(dt, 3)
    assert not np.array_equal(o3.download("RH"), after["RH"])



def _viscous_step(x, g):
    x.load_state(C.inputs_of(g))
    x.set_viscosity(int(g["vis_h_ivis"]), int(g["vis_h_ivistyp"]), int(g["vis_h_ieta"]), int(g["vis_h_idiss"]),
                    float(g["vis_xnue"]))
    x.step(float(g["delt"]), int(g["iii"]))
    x.pressg()
    x.sound()


def test_printout_after_a_viscous_step_is_the_references_file(tmp_path):
    """PRINTOUT(1) after a full step with viscosity and dissipative heating:
    the file carries that step's ADVECT3D fluxes (conv_flux, ppadv.f:857-865)
    and FORGAS3D's dissipation (DISFN); the library's file equals the
    reference's byte for byte, and neither record is zero."""
    g = C.load("checkpoint_visc_step_12x16x12")
    p = C.params_of(g)
    o = Oracle(p, threads=1)
    _viscous_step(o, g)
    for c in (1, 2, 3):
        np.testing.assert_array_equal(o.download(f"conv_flux{c}"), g[f"out_conv_flux{c}"])
    sc = {k[5:]: g[k].item() for k in g.files if k.startswith("ckpt_")}
    init = (sc["h_tinitbottom"], sc["h_tinittop"], sc["h_rhinittop"], sc["h_rhinitbottom"],
            sc["h_hydrotime"], sc["h_maccreated"])
    fn = tmp_path / "SNAP_000.0001"
    K.write_printout(fn, o, p, sc["h_zeit"], float(g["delt"]), int(sc["h_num_iter"]), init)
    assert fn.read_bytes() == g["snap_bytes"].tobytes()
    d = K.read_printout(fn)
    for k in ("DISFN", "conv_flux_1", "conv_flux_2", "conv_flux_3"):
        assert np.count_nonzero(d[k]) > d[k].size // 2, k
