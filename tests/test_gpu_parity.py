"""Parity of the HIP path (through the C ABI of libttgpu.so) with the
reference Fortran's outputs stored in the golden fixtures (same inputs), and
with the CPU oracle at larger sizes.

Bars: advection/force/boundaries/energy fields within
1e-12 relative of the reference (transcendental functions -- pow, exp, cos --
come from the GPU math library, so results are not bit-identical); SOR
iteration counts identical; delta_ER within the solver tolerance.
"""
import numpy as np
import pytest

from tests.golden import compare as C
from ttcode_amd import params as P

pytestmark = pytest.mark.gpu

REL = 1e-12


@pytest.fixture(scope="module")
def ttgpu():
    from ttcode_amd import ttgpu as T
    T.load_library()
    return T


def _dev(ttgpu, g):
    p = C.params_of(g)
    d = ttgpu.Device(p)
    d.load_state(C.inputs_of(g))
    return p, d


# fields that are differences of nearly equal numbers (fluxes, Newton
# residuals, corrections) are compared against the array's own scale
SCALED = {"flux_x", "flux_y", "flux_z", "FQ", "FR", "RHS_ER", "delta_ER", "Jacob_eta"}


def scaled_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# A face flux is -difr * (E_i - E_i-1): where E is nearly uniform across the
# face the difference cancels, and a last-bit difference of E becomes a large
# fraction of the flux.  Its forward-error scale is |difr| * |E|.
FLUX_DIFR = {"flux_x": "difrx", "flux_y": "difry", "flux_z": "difrz"}


def flux_err(g, k, mine, want, reg):
    outs = C.outputs_of(g)
    d = outs.get(FLUX_DIFR[k])
    e = outs.get("ER")
    if d is None or e is None:
        return None
    scale = float(np.max(np.abs(d[reg]))) * float(np.max(np.abs(e)))
    return float(np.max(np.abs(mine - want)) / max(scale, 1e-300))


def _cmp(g, p, d, names=None, rel=REL, skip=(), scaled_tol=None):
    worst = {}
    for k, ref in C.outputs_of(g).items():
        if (names and k not in names) or k in skip:
            continue
        reg = C.compare_region(k, p.nx, p.ny, p.nz)
        mine = d.download(k)[reg]
        want = ref[reg]
        if mine.dtype.kind == "i":
            assert np.array_equal(mine, want), k
            continue
        fin = np.isfinite(want)
        assert np.array_equal(np.isfinite(mine), fin), f"{k}: non-finite pattern differs"
        fe = flux_err(g, k, mine, want, reg) if k in FLUX_DIFR else None
        if fe is not None:
            e, tol = fe, rel
        elif k in SCALED:
            e = scaled_err(mine[fin], want[fin])
            tol = scaled_tol if scaled_tol is not None else rel
        else:
            e = C.rel_err(mine[fin], want[fin])
            tol = rel
        worst[k] = e
        print(f"  {g.fid.name if hasattr(g, 'fid') else ''} {k:12s} err {e:.3e} (tol {tol:.0e})")
        assert e <= tol, f"{k}: max err {e:.3e} > {tol:.1e}"
    return worst


def test_metrics(ttgpu):
    g = C.load("bounds_12x16x12")
    p, d = _dev(ttgpu, g)
    ids = {"xxa": 0, "xxb": 1, "dxa": 2, "dxb": 3, "xyb": 5, "dyb": 7, "xzb": 9, "dzb": 11,
           "surxa": 22, "volxb": 25, "surzb": 35, "volzb": 37, "geozg": 44}
    for n, m in ids.items():
        np.testing.assert_array_equal(d.metric(m), g["out_m_" + n], err_msg=n)
    d.close()


# IBDRAD=3 takes pow(cos*cos, 1/4) from the device math library: the
# transcendental bar; the other versions are pure arithmetic: bit-identical
@pytest.mark.parametrize("name,rel", [("bounds_12x16x12", 0.0), ("bounds_ibdrad3_12x16x12", REL),
                                      ("bounds_ibdrad4_12x16x12", 0.0)])
def test_bounds(ttgpu, name, rel):
    g = C.load(name)
    p, d = _dev(ttgpu, g)
    d.bounds()
    _cmp(g, p, d, rel=rel)
    d.close()


def _burrows_wave(d, g):
    d.set_burrows_table(g["tab_planck"], g["tab_ross"], g["tab_abs"])
    d.set_burrows_wave_table(g["tab_wave"], g["tab_bstar"], float(g["wave_dilution"]))


def test_kappa_burrows21(ttgpu):
    """KAPPA, IKAPTYP 21: the means, and upperbnd-1's values from upperbnd up."""
    g = C.load("kappa_burrows21_12x16x12")
    p, d = _dev(ttgpu, g)
    _burrows_wave(d, g)
    d.bounds()
    d.kappa()
    _cmp(g, p, d)
    d.close()


@pytest.mark.parametrize("name", ["energy_burrows21_12x16x12", "energy_burrows24_12x16x12"])
def test_coupled_energy_burrows_wave(ttgpu, name):
    """COUPLED_ENERGY_EQS with IKAPTYP 21 and 24: the wavelength-binned
    STELLARINPUT (pprad.f:429-447,519-569) within 1e-12 of the reference."""
    g = C.load(name)
    p, d = _dev(ttgpu, g)
    _burrows_wave(d, g)
    st, _ = d.coupled_energy(float(g["delt"]), int(g["iii"]))
    assert st.energy_iter == int(g["scal_energy_iter"])
    assert st.sor_iter_last == int(g["scal_SOR_ITER"])
    _cmp(g, p, d, rel=1e-12, scaled_tol=1e-10)
    d.close()


def test_bounds_bcxmin9(ttgpu):
    """BOUNDXMIN(9) at III = 1 (ppbou.f:1029-1066) on varying opacities."""
    g = C.load("bounds_bcxmin9_12x16x12")
    p, d = _dev(ttgpu, g)
    d.set_iii(int(g["iii"]))
    d.bounds()
    _cmp(g, p, d)
    d.close()


def test_bounds_bcxmin9_iii0_burrows_cell(ttgpu):
    """BOUNDXMIN(9) at III = 0: BurrowsOPC at cell (1,1,1) on the device."""
    g = C.load("bounds_bcxmin9_iii0_12x16x12")
    p, d = _dev(ttgpu, g)
    d.set_burrows_table(g["tab_planck"], g["tab_ross"], g["tab_abs"])
    d.set_iii(int(g["iii"]))
    d.bounds()
    _cmp(g, p, d)
    d.close()


def test_coupled_energy_bcxmin9(ttgpu):
    """COUPLED_ENERGY_EQS with the BCXMIN 9 Q/E bottom blocks and delta_Q
    from i = 2 (ppenergy.f:150-167,390-396,621-638)."""
    g = C.load("energy_bcxmin9_12x16x12")
    p, d = _dev(ttgpu, g)
    st, _ = d.coupled_energy(float(g["delt"]), int(g["iii"]))
    assert st.energy_iter == int(g["scal_energy_iter"])
    assert st.sor_iter_last == int(g["scal_SOR_ITER"])
    _cmp(g, p, d, rel=1e-12, scaled_tol=1e-10)
    d.close()


# *_omrot_*: the shipped ipp's rotating frame, OMROT = 2 PI / (3.52 days)
@pytest.mark.parametrize("name", ["advect_12x16x12", "advect_omrot_12x16x12"])
def test_advect3d(ttgpu, name):
    g = C.load(name)
    p, d = _dev(ttgpu, g)
    d.advect3d(float(g["delt"]))
    _cmp(g, p, d)
    d.close()


@pytest.mark.parametrize("name", ["forgas_12x16x12", "forgas_omrot_12x16x12", "forgas_igrav4_12x16x12"])
def test_forgas3d(ttgpu, name):
    g = C.load(name)
    p, d = _dev(ttgpu, g)
    d.bounds()
    d.qdens()
    d.forgas3d(float(g["delt"]))
    _cmp(g, p, d)
    d.close()


def test_solve_with_sor(ttgpu):
    g = C.load("sor_12x16x12")
    p, d = _dev(ttgpu, g)
    for k in ("EVAL1", "EVAL2", "EVAL3", "EVAL4", "EVAL5", "EVAL6", "EVAL7", "EVAL8",
              "EVAL_1_6", "RHS_ER", "upperbnd"):
        d.upload(k, g["in_" + k])
    st, _ = d.solve_sor(True)
    assert st.sor_iter == int(g["scal_SOR_ITER"])
    assert abs(st.rhsnorm / float(g["scal_global_RHSNORM"]) - 1) < 1e-13
    assert abs(st.rnorm / float(g["scal_global_RNORM"]) - 1) < 1e-10
    # red-black sweeps are order independent: same operations as the reference
    _cmp(g, p, d, rel=1e-13)
    d.close()


def test_adjust_sor_param(ttgpu):
    """ADJUST_SOR_PARAM (ppsor.f:332-354): the 50 trial counts, the chosen
    SORPARAM and the last trial's delta_ER against the reference's."""
    g = C.load("adjust_12x16x12")
    p, d = _dev(ttgpu, g)
    for k in ("EVAL1", "EVAL2", "EVAL3", "EVAL4", "EVAL5", "EVAL6", "EVAL7", "EVAL8",
              "EVAL_1_6", "RHS_ER", "upperbnd"):
        d.upload(k, g["in_" + k])
    om, it = d.adjust_sor_param(fill_matrix=True)
    want = [int(g[f"scal_ADJ_ITER_{i:02d}"]) for i in range(1, 51)]
    print("counts", list(it))
    assert list(it) == want
    assert om == float(g["scal_SORPARAM"])
    assert d.get_sorparam() == om
    _cmp(g, p, d, rel=1e-13)
    d.close()


def test_coupled_energy_adjusts_omega_at_step_one(ttgpu):
    """With adjust_omega on, COUPLED_ENERGY_EQS at III = 1 retunes omega
    before its first solve, as the reference does (ppenergy.f:372-374)."""
    g = C.load("energy_adjust_iii1_12x16x12")
    p, d = _dev(ttgpu, g)
    d.set_solver(adjust_omega=1)
    st, _ = d.coupled_energy(float(g["delt"]), int(g["iii"]))
    assert st.energy_iter == int(g["scal_energy_iter"])
    assert st.sor_iter_last == int(g["scal_SOR_ITER"])
    assert d.get_sorparam() == float(g["scal_SORPARAM"])
    _cmp(g, p, d, rel=1e-12, scaled_tol=1e-10)
    d.close()


@pytest.mark.parametrize("name", ["energy_const_12x16x12", "energy_pow8_12x16x12",
                                  "energy_flim1_sig_noirr_12x16x12"])
def test_coupled_energy(ttgpu, name):
    g = C.load(name)
    p, d = _dev(ttgpu, g)
    st, _ = d.coupled_energy(float(g["delt"]), int(g["iii"]))
    assert st.energy_iter == int(g["scal_energy_iter"])
    assert st.sor_iter_last == int(g["scal_SOR_ITER"])
    # delta_ER is the whole step's change; compare it where it is not tiny
    _cmp(g, p, d, rel=1e-12, scaled_tol=1e-10)
    d.close()


def _table(d, g):
    d.set_opacity_table(g["tab_planck"], g["tab_ross"], g["tab_abs"])


def test_kappa_freedman_table(ttgpu):
    """KAPPA with IKAPTYP 4 (table opacities, pprphy.f:84-120,800-942) and
    SUBSIG against the reference, on a state that reaches both table clamps,
    the -7 entries and the copy above upperbnd (10^x comes from the device
    math library: within 1e-12)."""
    g = C.load("kappa_freedman_12x16x12")
    p, d = _dev(ttgpu, g)
    try:
        _table(d, g)
        d.bounds()
        d.kappa()
        _cmp(g, p, d, rel=1e-12)
    finally:
        d.close()


def test_coupled_energy_freedman_table(ttgpu):
    g = C.load("energy_freedman_12x16x12")
    p, d = _dev(ttgpu, g)
    try:
        _table(d, g)
        st, _ = d.coupled_energy(float(g["delt"]), int(g["iii"]))
        assert st.energy_iter == int(g["scal_energy_iter"])
        assert st.sor_iter_last == int(g["scal_SOR_ITER"])
        _cmp(g, p, d, rel=1e-12, scaled_tol=1e-10)
    finally:
        d.close()


def test_kappa_freedman_table_stops_on_nonpositive_temperature(ttgpu):
    """FreedmanOPC stops on T <= 0 (pprphy.f:825-828): a zero temperature
    planted from the host makes KAPPA return a positive code; every kernel
    still exits normally, and the next call on a good state succeeds."""
    g = C.load("kappa_freedman_12x16x12")
    p, d = _dev(ttgpu, g)
    try:
        _table(d, g)
        d.bounds()
        T = d.download("T")
        T0 = T.copy()
        T[3, 4, 5] = 0.0
        d.upload("T", T)
        with pytest.raises(ttgpu.TTError, match="FREEDMAN"):
            d.kappa()
    finally:
        d.close()
    p, d = _dev(ttgpu, g)
    try:
        _table(d, g)
        d.bounds()
        d.kappa()
    finally:
        d.close()


def _burrows(d, g):
    d.set_burrows_table(g["tab_planck"], g["tab_ross"], g["tab_abs"])


@pytest.mark.parametrize("name", ["kappa_burrows11_12x16x12", "kappa_burrows14_12x16x12"])
def test_kappa_burrows_table(ttgpu, name):
    """KAPPA with IKAPTYP 11 and 14 (BurrowsOPC, pprphy.f:199-274,1005-1225)
    and SUBSIG against the reference, on a state beyond both clamps in T and
    rho (log10 and 10^x from the device math library: within 1e-12)."""
    g = C.load(name)
    p, d = _dev(ttgpu, g)
    try:
        _burrows(d, g)
        d.bounds()
        d.kappa()
        _cmp(g, p, d, rel=1e-12)
    finally:
        d.close()


def test_coupled_energy_burrows_table(ttgpu):
    g = C.load("energy_burrows10_12x16x12")
    p, d = _dev(ttgpu, g)
    try:
        _burrows(d, g)
        st, _ = d.coupled_energy(float(g["delt"]), int(g["iii"]))
        assert st.energy_iter == int(g["scal_energy_iter"])
        assert st.sor_iter_last == int(g["scal_SOR_ITER"])
        _cmp(g, p, d, rel=1e-12, scaled_tol=1e-10)
    finally:
        d.close()


def test_kappa_burrows_table_stops_on_nonpositive_temperature(ttgpu):
    """BurrowsOPC stops on T <= 0 (pprphy.f:1039-1045): a zero temperature
    planted from the host makes KAPPA return a positive code; the kernels
    still exit normally."""
    g = C.load("kappa_burrows14_12x16x12")
    p, d = _dev(ttgpu, g)
    try:
        _burrows(d, g)
        d.bounds()
        T = d.download("T")
        T[3, 4, 5] = 0.0
        d.upload("T", T)
        with pytest.raises(ttgpu.TTError, match="BURROWS"):
            d.kappa()
    finally:
        d.close()


def _column(d, g):
    d.set_opacity_table(g["tab_planck"], g["tab_ross"], g["tab_abs"])
    d.set_column_table(g["tab_col"])


def test_kappa_column_table(ttgpu):
    """KAPPA with IKAPTYP 6 (the column-depth absorption mean, pprphy.f:
    123-175,1543-1745) against the reference (log10 and 10^x from the
    device math library: within 1e-12)."""
    g = C.load("kappa_column6_12x16x12")
    p, d = _dev(ttgpu, g)
    try:
        _column(d, g)
        d.bounds()
        d.kappa()
        _cmp(g, p, d, rel=1e-12)
    finally:
        d.close()


def test_coupled_energy_column_table(ttgpu):
    g = C.load("energy_column6_12x16x12")
    p, d = _dev(ttgpu, g)
    try:
        _column(d, g)
        st, _ = d.coupled_energy(float(g["delt"]), int(g["iii"]))
        assert st.energy_iter == int(g["scal_energy_iter"])
        assert st.sor_iter_last == int(g["scal_SOR_ITER"])
        _cmp(g, p, d, rel=1e-12, scaled_tol=1e-10)
    finally:
        d.close()


def test_kappa_hminus(ttgpu):
    """IKAPTYP 8 (H- bound-free) against the reference on prefilled opacity
    arrays (pow from the device math library: within 1e-12)."""
    g = C.load("kappa_hminus8_12x16x12")
    p, d = _dev(ttgpu, g)
    try:
        d.bounds()
        d.kappa()
        _cmp(g, p, d, rel=1e-12)
    finally:
        d.close()


def test_coupled_energy_hminus(ttgpu):
    g = C.load("energy_hminus8_12x16x12")
    p, d = _dev(ttgpu, g)
    try:
        st, _ = d.coupled_energy(float(g["delt"]), int(g["iii"]))
        assert st.energy_iter == int(g["scal_energy_iter"])
        assert st.sor_iter_last == int(g["scal_SOR_ITER"])
        _cmp(g, p, d, rel=1e-12, scaled_tol=1e-10)
    finally:
        d.close()


def test_coupled_energy_constant_gamma_opacity(ttgpu):
    """IKAPTYP 13 (kappa_A = gamma_OPC^4 ZKAPMULT) against the reference."""
    g = C.load("energy_const13_12x16x12")
    p, d = _dev(ttgpu, g)
    try:
        st, _ = d.coupled_energy(float(g["delt"]), int(g["iii"]))
        assert st.energy_iter == int(g["scal_energy_iter"])
        assert st.sor_iter_last == int(g["scal_SOR_ITER"])
        _cmp(g, p, d, rel=1e-12, scaled_tol=1e-10)
    finally:
        d.close()


def test_step_burrows_table_32x64x32_vs_oracle(ttgpu):
    """One full step with the Burrows tables (IKAPTYP 10, the synthetic
    table) at 32x64x32 with a 5% day-night contrast, against the oracle."""
    from oracle.ttoracle import Oracle
    from ttcode_amd import opacity
    p = P.make_params(32, 64, 32, epsmat=1e-8, ikaptyp=10)
    st = P.generate_state(p, seed=71, contrast=0.05)
    tab = opacity.synthetic_burrows_table(2)[2:]
    o = Oracle(p, threads=1)
    o.set_burrows_table(*tab)
    o.load_state(st)
    o.bounds()
    o.pressg()
    dt = o.schritt()
    o.load_state(st)
    so, _ = o.step(dt, 1)
    d = ttgpu.Device(p)
    try:
        d.set_burrows_table(*tab)
        d.load_state(st)
        sg, _ = d.step(dt, 1)
        assert sg.energy.energy_iter == so.energy.energy_iter
        same = sg.energy.sor_iter_total == so.energy.sor_iter_total
        assert same or abs(sg.energy.sor_iter_total - so.energy.sor_iter_total) <= so.energy.energy_iter
        for k in ("RH", "V", "G", "H", "T", "ER", "xkapR", "dxkapP_dT"):
            reg = C.compare_region(k, p.nx, p.ny, p.nz)
            e = C.rel_err(d.download(k)[reg], o.download(k)[reg])
            tol = 1e-12 if same or k in ("RH", "V", "G", "H") else p.epsmat
            print(f"  {k}: {e:.2e}")
            assert e <= tol, (k, e)
    finally:
        d.close()


def test_step_freedman_table_32x64x32_vs_oracle(ttgpu):
    """One full step with table opacities (IKAPTYP 4, the synthetic table)
    at 32x64x32 with a 5% day-night contrast, against the oracle."""
    from oracle.ttoracle import Oracle
    from ttcode_amd import opacity
    p = P.make_params(32, 64, 32, epsmat=1e-8, ikaptyp=4)
    st = P.generate_state(p, seed=29, contrast=0.05)
    tab = opacity.synthetic_table(1)
    o = Oracle(p, threads=1)
    o.set_opacity_table(*tab)
    o.load_state(st)
    o.bounds()
    o.pressg()
    dt = o.schritt()
    o.load_state(st)
    so, _ = o.step(dt, 1)
    d = ttgpu.Device(p)
    try:
        d.set_opacity_table(*tab)
        d.load_state(st)
        sg, _ = d.step(dt, 1)
        assert sg.energy.energy_iter == so.energy.energy_iter
        same = sg.energy.sor_iter_total == so.energy.sor_iter_total
        assert same or abs(sg.energy.sor_iter_total - so.energy.sor_iter_total) <= so.energy.energy_iter
        for k in ("RH", "V", "G", "H", "T", "ER", "xkapR", "dxkapP_dT"):
            reg = C.compare_region(k, p.nx, p.ny, p.nz)
            e = C.rel_err(d.download(k)[reg], o.download(k)[reg])
            tol = 1e-12 if same or k in ("RH", "V", "G", "H") else p.epsmat
            print(f"  {k}: {e:.2e}")
            assert e <= tol, (k, e)
    finally:
        d.close()


@pytest.mark.parametrize("name", ["forgas_visc_ieta4_12x16x12", "forgas_visc_ieta0_12x16x12",
                                  "forgas_visc_ieta2_12x16x12", "forgas_visc_ieta3_16x16x12"])
def test_forgas3d_viscosity(ttgpu, name):
    """FORGAS3D with ETACALC + VISC (IVISTYP=1, spherical) and dissipative
    heating against the reference (tan from the device math library)."""
    g = C.load(name)
    p, d = _dev(ttgpu, g)
    try:
        d.set_viscosity(int(g["vis_ivis"]), int(g["vis_ivistyp"]), int(g["vis_ieta"]), int(g["vis_idiss"]),
                        float(g["vis_xnue"]))
        if "vis_hydrotime" in g.files:  # IETA 2 ramps with the run's time
            d.set_limits(hydrotime=float(g["vis_hydrotime"]))
        d.bounds()
        d.qdens()
        d.forgas3d(float(g["delt"]))
        _cmp(g, p, d, rel=1e-12)
    finally:
        d.close()


@pytest.mark.parametrize("name", ["forgas_artvis_12x16x12", "forgas_visc_artvis_12x16x12"])
def test_forgas3d_artificial_viscosity(ttgpu, name):
    """FORGAS3D with NR_ARTIFICIAL_VISC against the reference: alone, with
    upperbnd varying between columns, and with the shipped ipp's viscosity;
    dt_artvis exactly (a minimum of quotients), then SCHRITT takes it."""
    g = C.load(name)
    p, d = _dev(ttgpu, g)
    try:
        get = lambda k, v: g[k].item() if k in g.files else v  # noqa: E731
        d.set_viscosity(int(get("vis_ivis", 0)), int(get("vis_ivistyp", 1)), int(get("vis_ieta", 4)),
                        int(get("vis_idiss", 0)), float(get("vis_xnue", 1e9)), 1)
        assert d.dt_artvis() == p.dtmax
        d.bounds()
        d.qdens()
        d.forgas3d(float(g["delt"]))
        _cmp(g, p, d, rel=1e-12)
        dav = d.dt_artvis()
        assert dav == g["scal_dt_artvis"].item()
        d.sound()
        d_all = d.schritt()
        v = d.get_viscosity()
        d.set_viscosity(v.ivis, v.ivistyp, v.ieta, v.idiss, v.xnue, 0)
        d_rest = d.schritt()
        assert abs(d_all - min(d_rest / p.fdelt, dav) * p.fdelt) <= 1e-15 * d_all
    finally:
        d.close()


def test_schritt_viscous_limit_vs_oracle(ttgpu):
    """SCHRITT with IVIS = 1 (the viscous condition, ppsub.f:97-119) and
    IARTVIS = 1 after one viscous FORGAS3D, against the oracle; pow(x, -2)
    from the device math library, so within 1e-14."""
    from oracle.ttoracle import Oracle
    g = C.load("forgas_visc_ieta0_12x16x12")
    p = C.params_of(g)
    dt = float(g["delt"])
    vis = (1, 1, 0, 1, float(g["vis_xnue"]), 1)
    o = Oracle(p, threads=1)
    o.load_state(C.inputs_of(g))
    o.set_viscosity(*vis)
    o.bounds()
    o.L.tor_qdens()
    o.forgas3d(dt)
    want = o.schritt()
    d = ttgpu.Device(p)
    try:
        d.load_state(C.inputs_of(g))
        d.set_viscosity(*vis)
        d.bounds()
        d.qdens()
        d.forgas3d(dt)
        assert d.dt_artvis() == o.dt_artvis()
        got = d.schritt()
        print(f"  delt {got:.17e} vs {want:.17e}")
        assert abs(got - want) <= 1e-14 * want
    finally:
        d.close()


def test_step_artificial_viscosity_32x64x32_vs_oracle(ttgpu):
    """Three steps with the shipped ipp's viscosity and artificial viscosity
    (IVIS=1, IETA=4, IDISS=1, IARTVIS=1), SCHRITT between them taking the
    viscous and artificial bounds, at 32x64x32 against the oracle."""
    from oracle.ttoracle import Oracle
    p = P.make_params(32, 64, 32, epsmat=1e-8)
    st = P.generate_state(p, seed=53, contrast=0.05)
    vis = (1, 1, 4, 1, 1e9, 1)
    o = Oracle(p, threads=1)
    d = ttgpu.Device(p)
    try:
        o.load_state(st)
        o.set_viscosity(*vis)
        d.load_state(st)
        d.set_viscosity(*vis)
        same = True
        for it in range(3):
            o.bounds()
            o.pressg()
            d.bounds()
            d.pressg()
            dto, dtg = o.schritt(), d.schritt()
            assert abs(dtg - dto) <= 1e-13 * dto, (it, dtg, dto)
            so, _ = o.step(dto, it + 1)
            sg, _ = d.step(dto, it + 1)
            assert sg.energy.energy_iter == so.energy.energy_iter
            same = same and sg.energy.sor_iter_total == so.energy.sor_iter_total
            assert abs(d.dt_artvis() - o.dt_artvis()) <= 1e-12 * o.dt_artvis()
        # bit-near while the SOR counts agree; else within the solver tolerance
        tol = 1e-11 if same else 1e-6
        for k in ("RH", "V", "G", "H", "T", "ER"):
            reg = C.compare_region(k, p.nx, p.ny, p.nz)
            e = C.rel_err(d.download(k)[reg], o.download(k)[reg])
            print(f"  {k}: {e:.2e}")
            assert e <= tol, (k, e)
    finally:
        d.close()


def test_step_viscosity_32x64x32_vs_oracle(ttgpu):
    """One full step with viscosity (IETA=0, strong enough to move the
    flow) at 32x64x32 against the oracle."""
    from oracle.ttoracle import Oracle
    p = P.make_params(32, 64, 32, epsmat=1e-8)
    st = P.generate_state(p, seed=37, contrast=0.05)
    vis = (1, 1, 0, 1, 3e14)
    o = Oracle(p, threads=1)
    o.load_state(st)
    o.bounds()
    o.pressg()
    dt = o.schritt()
    o.load_state(st)
    o.set_viscosity(*vis)
    so, _ = o.step(dt, 1)
    d = ttgpu.Device(p)
    try:
        d.load_state(st)
        d.set_viscosity(*vis)
        sg, _ = d.step(dt, 1)
        assert sg.energy.energy_iter == so.energy.energy_iter
        same = sg.energy.sor_iter_total == so.energy.sor_iter_total
        assert same or abs(sg.energy.sor_iter_total - so.energy.sor_iter_total) <= so.energy.energy_iter
        for k in ("RH", "V", "G", "H", "T", "ER", "divtx", "DISFN"):
            reg = C.compare_region(k, p.nx, p.ny, p.nz)
            e = C.rel_err(d.download(k)[reg], o.download(k)[reg])
            tol = 1e-12 if same or k not in ("T", "ER") else p.epsmat
            print(f"  {k}: {e:.2e}")
            assert e <= tol, (k, e)
    finally:
        d.close()


@pytest.mark.parametrize("name", ["step2_12x16x12", "step2_ibdrad9_12x16x12", "step2_omrot_12x16x12",
                                  "step2_igrav4_12x16x12", "step2_bcxmin9_12x16x12"])
def test_two_steps(ttgpu, name):
    g = C.load(name)
    p, d = _dev(ttgpu, g)
    for s in range(int(g["nsteps"])):
        st, _ = d.step(float(g["delt"]), int(g["iii"]) + s)
    assert st.energy.energy_iter == int(g["scal_energy_iter"])
    assert st.energy.sor_iter_last == int(g["scal_SOR_ITER"])
    _cmp(g, p, d, rel=1e-12, scaled_tol=1e-10)
    d.close()


def test_diffusion(ttgpu):
    g = C.load("diffusion_12x16x12")
    p, d = _dev(ttgpu, g)
    d.bounds()
    it, _ = d.diffusion(float(g["delt"]), float(g["xnue"]), float(g["pr_num"]))
    assert it == int(g["scal_SOR_ITER"])
    _cmp(g, p, d, rel=1e-13)
    d.close()


def test_schritt_matches_oracle(ttgpu):
    from oracle.ttoracle import Oracle
    p = P.make_params(32, 64, 32)
    st = P.generate_state(p, seed=3)
    o = Oracle(p)
    o.load_state(st)
    o.bounds()
    o.pressg()
    d = ttgpu.Device(p)
    d.load_state(st)
    d.bounds()
    d.pressg()
    assert d.schritt() == o.schritt()
    d.close()


@pytest.mark.parametrize("grid", ["radial", "phi", "theta"])
def test_schritt_closed_form(ttgpu, grid):
    """k_schritt against SCHRITT's Courant bound written in numpy
    (ppsub.f:49-78,129-131), on perturbed metrics and fast flows, on grids
    where each of the three Courant terms binds in turn."""
    from tests import closed_forms as CF
    p = P.make_params(12, 16, 12, **CF.SCHRITT_GRIDS[grid])
    st = P.generate_state(p, seed=11, vel_frac=0.5)
    d = ttgpu.Device(p)
    d.load_state(st)
    d.pressg()
    met = CF.perturbed_metrics(d.metric, np.random.default_rng(111))
    for m, a in met.items():
        d.upload_metric(m, a)
    got = d.schritt()
    want, _ = CF.schritt_cfl(p, st, met)
    d.close()
    assert got == want, (got, want)


def test_step_32x64x32_vs_oracle(ttgpu):
    """One full step of the CPU-runnable workload (nr=32, ntheta=32, nphi=64)."""
    from oracle.ttoracle import Oracle
    p = P.make_params(32, 64, 32, epsmat=1e-8)
    st = P.generate_state(p, seed=5)
    o = Oracle(p, threads=1)
    o.load_state(st)
    o.bounds()
    o.pressg()
    dt = o.schritt()
    o.load_state(st)
    so, _ = o.step(dt, 1)
    d = ttgpu.Device(p)
    d.load_state(st)
    sg, _ = d.step(dt, 1)
    assert sg.energy.energy_iter == so.energy.energy_iter
    same = sg.energy.sor_iter_total == so.energy.sor_iter_total
    assert same or abs(sg.energy.sor_iter_total - so.energy.sor_iter_total) <= so.energy.energy_iter
    for k in ("RH", "V", "G", "H"):
        e = C.rel_err(d.download(k), o.download(k))
        assert e <= 1e-12, (k, e)
    # equal SOR counts: the same solve, to the north-star bar; a stop one
    # iteration apart (the L1 norm summed in another order) moves E by less
    # than the solver tolerance
    for k in ("T", "ER"):
        reg = C.compare_region(k, p.nx, p.ny, p.nz)
        e = C.rel_err(d.download(k)[reg], o.download(k)[reg])
        print(f"  {k}: {e:.2e} (SOR {sg.energy.sor_iter_total} vs {so.energy.sor_iter_total})")
        assert e <= (1e-12 if same else p.epsmat), (k, e)
    d.close()


@pytest.mark.parametrize("shape", [(40, 20, 18), (14, 10, 7)])
def test_ragged_shapes_vs_oracle(ttgpu, shape):
    """Grids that are not multiples of the kernels' tiles: the advection
    tiles (32 r x 8 phi cells, two theta levels per thread) end part-way in
    r, phi and theta, as do the SOR row blocks.  ADVECT3D + FORGAS3D within
    1e-13 of the oracle, then a full step: equal Newton counts, and T, ER
    within 1e-12 when the SOR counts are equal (the during-sweep L1 norm is
    summed in another order, which may move the stop by one iteration:
    then within the solver tolerance)."""
    from oracle.ttoracle import Oracle
    nx, ny, nz = shape
    p = P.make_params(nx, ny, nz, epsmat=1e-8)
    st = P.generate_state(p, seed=17)
    o = Oracle(p, threads=1)
    o.load_state(st)
    o.bounds()
    o.pressg()
    dt = o.schritt()
    print(f"shape {shape} dt {dt:.4g}", flush=True)
    d = ttgpu.Device(p)
    try:
        for x in (o, d):
            x.load_state(st)
            x.bounds()
            x.pressg()
            x.advect3d(dt)
            x.forgas3d(dt)
        for k in ("RH", "V", "G", "H", "T"):
            reg = C.compare_region(k, nx, ny, nz)
            e = C.rel_err(d.download(k)[reg], o.download(k)[reg])
            print(f"  hydro {k}: {e:.2e}")
            assert e <= 1e-13, (k, e)
        for x in (o, d):
            x.load_state(st)
        so, _ = o.step(dt, 1)
        sg, _ = d.step(dt, 1)
        assert sg.energy.energy_iter == so.energy.energy_iter
        same = sg.energy.sor_iter_total == so.energy.sor_iter_total
        assert same or abs(sg.energy.sor_iter_total - so.energy.sor_iter_total) <= so.energy.energy_iter
        for k in ("RH", "V", "G", "H", "T", "ER"):
            reg = C.compare_region(k, nx, ny, nz)
            e = C.rel_err(d.download(k)[reg], o.download(k)[reg])
            tol = 1e-12 if same or k in ("RH", "V", "G", "H") else p.epsmat
            print(f"  step {k}: {e:.2e} (SOR {sg.energy.sor_iter_total} vs {so.energy.sor_iter_total})")
            assert e <= tol, (k, e)
    finally:
        d.close()


def test_speculative_tests_never_change_results(ttgpu):
    """Speculating on every batch (margin 0) makes deferred tests end solves
    one red half sweep late; each such solve is rerun with tests of its own,
    so two steps agree bit for bit with speculation switched off."""
    p = P.make_params(32, 64, 32, epsmat=1e-8)
    st = P.generate_state(p, seed=5)
    out, stats = {}, {}
    for margin in (-1.0, 0.0, 2.5):
        d = ttgpu.Device(p)
        d.load_state(st)
        d.bounds()
        d.pressg()
        dt = d.schritt()
        d.load_state(st)
        d.sor_speculation(margin)
        try:
            its = [d.step(dt, 2 + s)[0].energy.sor_iter_total for s in range(2)]
            stats[margin] = d.sor_speculation_get()
        finally:
            d.sor_speculation(2.5)
        out[margin] = (its, {k: d.download(k) for k in ("T", "ER", "delta_ER")})
        d.close()
    print("iterations", {m: v[0] for m, v in out.items()}, "speculative batches, reruns", stats)
    assert stats[0.0][0] > 0 and stats[0.0][1] > 0
    assert stats[-1.0] == (0, 0)
    for m in (0.0, 2.5):
        assert out[m][0] == out[-1.0][0]
        for k in ("T", "ER", "delta_ER"):
            assert np.array_equal(out[m][1][k], out[-1.0][1][k]), (m, k)


def test_mixed_precision_solve(ttgpu):
    """FP32 sweeps inside FP64 defect correction reach the FP64 residual target
    and agree with the FP64 solve within it."""
    g = C.load("sor_12x16x12")
    p = C.params_of(g)
    p.epsmat = 1e-10
    d = ttgpu.Device(p)
    d.load_state(C.inputs_of(g))
    for k in ("EVAL1", "EVAL2", "EVAL3", "EVAL4", "EVAL5", "EVAL6", "EVAL7", "EVAL8",
              "EVAL_1_6", "RHS_ER", "upperbnd"):
        d.upload(k, g["in_" + k])
    st64, _ = d.solve_sor(True)
    x64 = d.download("delta_ER")
    d.set_solver(precision=1, inner_tol=1e-4)
    st32, _ = d.solve_sor(True)
    x32 = d.download("delta_ER")
    assert st32.rnorm / st32.rhsnorm <= 1e-10
    assert st32.outer_iter >= 1 and st32.fp32_sweeps > 0
    scale = np.max(np.abs(x64))
    assert np.max(np.abs(x32 - x64)) <= 1e-6 * scale
    d.close()


def test_line_sor_solve(ttgpu):
    """FP32 red-black radial-line sweeps inside FP64 defect correction reach
    the FP64 residual target, agree with the FP64 solve within it, and take
    far fewer iterations; the same system solved twice gives the same bits."""
    p = P.make_params(32, 64, 32, epsmat=1e-10)
    st = P.generate_state(p, seed=5)
    d = ttgpu.Device(p)
    try:
        d.load_state(st)
        d.bounds()
        d.pressg()
        dt = d.schritt()
        d.coupled_energy(dt, 2)                 # leaves its last system on the device
        st64, _ = d.solve_sor(False)
        x64 = d.download("delta_ER")
        d.set_solver(precision=1, inner=1)
        stl, _ = d.solve_sor(False)
        xl = d.download("delta_ER")
        stl2, _ = d.solve_sor(False)
        xl2 = d.download("delta_ER")
        print(f"FP64 {st64.sor_iter} iterations; line {stl.sor_iter} in {stl.outer_iter} corrections, "
              f"ratio {stl.rnorm / stl.rhsnorm:.2e}")
        assert stl.rnorm / stl.rhsnorm <= 1e-10
        assert stl.outer_iter >= 1 and 0 < stl.fp32_sweeps and stl.sor_iter * 5 < st64.sor_iter
        assert np.max(np.abs(xl - x64)) <= 1e-6 * np.max(np.abs(x64))
        assert np.array_equal(xl, xl2) and stl2.sor_iter == stl.sor_iter
    finally:
        d.close()


def test_line_sor_fixture_system(ttgpu):
    """The line solver on the reference's SOR fixture system (12x16x12, whose
    identity rows above upperbnd cut the lines): the FP64 residual target is
    met and the result agrees with the FP64 solve."""
    g = C.load("sor_12x16x12")
    p = C.params_of(g)
    p.epsmat = 1e-10
    d = ttgpu.Device(p)
    try:
        d.load_state(C.inputs_of(g))
        for k in ("EVAL1", "EVAL2", "EVAL3", "EVAL4", "EVAL5", "EVAL6", "EVAL7", "EVAL8",
                  "EVAL_1_6", "RHS_ER", "upperbnd"):
            d.upload(k, g["in_" + k])
        st64, _ = d.solve_sor(True)
        x64 = d.download("delta_ER")
        d.set_solver(precision=1, inner=1)
        stl, _ = d.solve_sor(True)
        xl = d.download("delta_ER")
        print(f"FP64 {st64.sor_iter}, line {stl.sor_iter} ({stl.outer_iter} corrections)")
        assert stl.rnorm / stl.rhsnorm <= 1e-10
        assert np.max(np.abs(xl - x64)) <= 1e-6 * np.max(np.abs(x64))
    finally:
        d.close()


def test_energy_totals_match_host_integral(ttgpu):
    """tt_energy_totals against the same volume integrals in numpy from the
    downloaded fields and the GITTER volume metrics."""
    p = P.make_params(32, 64, 32, epsmat=1e-8)
    st = P.generate_state(p, seed=9)
    d = ttgpu.Device(p)
    d.load_state(st)
    d.bounds()
    d.pressg()
    dt = d.schritt()
    d.step(dt, 2)
    tot = d.energy_totals()
    nx, ny, nz = p.nx, p.ny, p.nz
    inner = (slice(2, nz + 2), slice(2, ny + 2), slice(2, nx + 2))
    # volxb, volyb, volzb have lower bound 0: entries 1..n are the interior
    vx, vy, vz = (d.metric(m)[1:n + 1] for m, n in ((25, nx), (31, ny), (37, nz)))
    dv = vz[:, None, None] * vy[None, :, None] * vx[None, None, :]
    rh, t, er, cv = (d.download(k)[inner] for k in ("RH", "T", "ER", "CV"))
    sin = d.download("STELLARINPUT")[2:nz + 2, 2:ny + 2, :]   # declared (1:nx, -1:ny+2, -1:nz+2)
    d.close()
    for got, want in ((tot[0], np.sum(rh * cv * t * dv)), (tot[1], np.sum(er * dv)), (tot[2], np.sum(sin * dv))):
        assert abs(got / want - 1) < 1e-12, (got, want)
    assert tot[2] > 0.0


@pytest.mark.parametrize("epsmat", [3e-2, 1e-3, 3e-5, 1e-6, 1e-9])
def test_sor_stops_at_the_reference_iteration(ttgpu, epsmat):
    """The convergence test runs once per group of iterations; a stop inside
    a group is replayed from a snapshot (or from the solve's start) so that
    the unknowns equal those of a test after every iteration.  Tolerances
    that stop early, mid-group and late, against the CPU oracle."""
    from oracle.ttoracle import Oracle
    g = C.load("sor_12x16x12")
    p = C.params_of(g)
    p.epsmat = epsmat
    keys = ("EVAL1", "EVAL2", "EVAL3", "EVAL4", "EVAL5", "EVAL6", "EVAL7", "EVAL8", "EVAL_1_6", "RHS_ER", "upperbnd")
    o = Oracle(p, threads=1)
    o.load_state(C.inputs_of(g))
    for k in keys:
        o.upload(k, g["in_" + k])
    so, _ = o.solve_sor(True)
    d = ttgpu.Device(p)
    d.load_state(C.inputs_of(g))
    for k in keys:
        d.upload(k, g["in_" + k])
    sg, _ = d.solve_sor(True)
    x = d.download("delta_ER")
    d.close()
    print(f"epsmat {epsmat}: iterations {sg.sor_iter} (oracle {so.sor_iter})")
    assert sg.sor_iter == so.sor_iter
    want = o.download("delta_ER")
    reg = C.compare_region("delta_ER", p.nx, p.ny, p.nz)
    assert C.rel_err(x[reg], want[reg]) <= 1e-13
