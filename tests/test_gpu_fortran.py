"""The Fortran host path: tt_driver (flang, ISO_C_BINDING layer with the
reference's routine names) runs two steps on the GPU; its results equal the
ctypes path's through the same C ABI."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from ttcode_amd import params as P

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ttcode_amd", "fortran",
                      "tt_driver")


def test_fortran_driver_matches_c_abi():
    from ttcode_amd import ttgpu
    assert os.path.exists(DRIVER), "build the Fortran driver first (python -m ttcode_amd.build)"
    p = P.make_params(16, 32, 16, epsmat=1e-8)
    st = P.generate_state(p, seed=21)
    d = ttgpu.Device(p)
    d.load_state(st)
    d.bounds()
    d.pressg()
    dt = d.schritt()
    d.load_state(st)
    for s in range(2):
        es, _ = d.step(dt, 1 + s)
    want = {k: d.download(k) for k in ("T", "RH", "V", "G", "H", "ER")}
    d.close()
    with tempfile.TemporaryDirectory() as tmp:
        nml = (f"&ttdrv h_nx={p.nx}, h_ny={p.ny}, h_nz={p.nz}, h_itmax={p.itmax}, h_iii=1, h_nsteps=2, "
               f"h_epsmat={p.epsmat!r}, h_delt={dt!r}, h_tirr={p.tirr!r}, h_kap0={p.kap0!r}, "
               f"h_dkap0={p.dkap0!r}, h_rho0={p.kap_rho0!r}, h_t0={p.kap_t0!r}, h_a={p.kap_a!r}, "
               f"h_b={p.kap_b!r} /\n").replace("e-", "d-").replace("e+", "d+")
        open(os.path.join(tmp, "params.nml"), "w").write(nml)
        for k in ("T", "RH", "V", "G", "H", "ER", "gravx"):
            np.ascontiguousarray(st[k]).tofile(os.path.join(tmp, f"in_{k}.bin"))
        r = subprocess.run([DRIVER, tmp], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        for k, ref in want.items():
            got = np.fromfile(os.path.join(tmp, f"out_{k}.bin")).reshape(ref.shape)
            np.testing.assert_array_equal(got, ref, err_msg=k)
        sc = dict(line.split() for line in open(os.path.join(tmp, "out_scalars.txt")))
        assert int(sc["energy_iter"]) == es.energy.energy_iter
