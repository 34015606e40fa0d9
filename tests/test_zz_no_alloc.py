"""tt_schritt_noalloc (SCHRITT's Courant bound on caller-owned device arrays,
allocating nothing) against tt_schritt on the library's own state: equal
bit for bit on two calls in a row, the second on a smaller grid, through
the C ABI; and its refusals.  The caller's device memory comes from the HIP
runtime the library itself uses (hipMalloc through ctypes): a second HIP
runtime in the process, such as torch's, is not needed."""
import ctypes

import numpy as np
import pytest

from ttcode_amd import params as P

pytestmark = pytest.mark.gpu

# metric ids (include/ttgpu.h TT_M_*) in the argument order
METRICS = (2, 6, 10, 42, 43, 44)  # DXA, DYA, DZA, GEOXG, GEOXH, GEOZG


def _box(name, a, p):
    """A wire-format array as the full ghosted box (-1:n+2)^3."""
    out = np.zeros((p.nz + 4, p.ny + 4, p.nx + 4))
    lo, _, _ = P.field_bounds(P.FIELD_IDS[name], p.nx, p.ny, p.nz)
    out[lo[2] + 1:, lo[1] + 1:, lo[0] + 1:] = a
    return out


class _Hip:
    """Caller-side device memory: hipMalloc / hipMemcpy / hipFree of the HIP
    runtime libttgpu.so is linked against."""

    def __init__(self):
        _lib()  # the library, and with it its HIP runtime, loaded first
        with open("/proc/self/maps") as f:
            paths = {ln.split()[-1] for ln in f if "libamdhip64" in ln}
        assert len(paths) == 1, paths
        self.h = ctypes.CDLL(paths.pop())
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), ctypes.c_size_t(a.nbytes)) == 0
        self.bufs.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0
        return p

    def get(self, p, n):
        out = np.empty(n)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _case(hip, nx, ny, nz, seed, edit=None):
    """tt_schritt on the library's state, and that state's arrays in
    caller-owned device memory."""
    from ttcode_amd import ttgpu
    p = P.make_params(nx, ny, nz)
    st = P.generate_state(p, seed=seed, vel_frac=0.5)
    d = ttgpu.Device(p)
    try:
        d.load_state(st)
        d.bounds()
        d.pressg()
        want = d.schritt()
        host = {k: d.download(k) for k in ("RH", "T", "PG", "V", "G", "H")}
        mets = [d.metric(m) for m in METRICS]
    finally:
        d.close()
    boxes = {k: _box(k, a, p) for k, a in host.items()}
    if edit:
        edit(boxes)
    dev = {k: hip.put(a) for k, a in boxes.items()}
    mdev = [hip.put(m) for m in mets]
    return p, want, dev, mdev


def _call(lib, p, dev, mdev, out):
    return lib.tt_schritt_noalloc(p.nx, p.ny, p.nz, *(dev[k] for k in ("RH", "T", "PG", "V", "G", "H")),
                                  *mdev, ctypes.c_double(p.gamma), ctypes.c_double(p.fdelt),
                                  ctypes.c_double(p.dtmax), out, None)


def _lib():
    from ttcode_amd import ttgpu
    lib = ttgpu.load_library()
    lib.tt_schritt_noalloc.restype = ctypes.c_int
    lib.tt_schritt_noalloc.argtypes = ([ctypes.c_int] * 3 + [ctypes.c_void_p] * 12 + [ctypes.c_double] * 3 +
                                       [ctypes.c_void_p, ctypes.c_void_p])
    lib.tt_last_error.restype = ctypes.c_char_p
    return lib


def test_noalloc_matches_tt_schritt_twice_second_smaller():
    lib = _lib()
    hip = _Hip()
    try:
        big = _case(hip, 32, 64, 32, seed=3)
        small = _case(hip, 12, 16, 10, seed=5)
        for p, want, dev, mdev in (big, small):
            out = hip.put(np.full(2, -1.0))
            print(f"tt_schritt_noalloc {p.nx}x{p.ny}x{p.nz}", flush=True)
            assert _call(lib, p, dev, mdev, out) == 0, lib.tt_last_error()
            got = hip.get(out, 2)
            print(f"  {got[0]!r} vs tt_schritt {want!r}, status {got[1]}")
            assert got[1] == 0.0
            assert got[0] == want
    finally:
        hip.free()


def test_noalloc_flags_a_non_positive_state():
    """A cell with T <= 0: out[1] = 1, as SCHRITT stops (ppsub.f:26-46)."""
    lib = _lib()
    hip = _Hip()

    def cold_cell(boxes):
        boxes["T"][5, 6, 7] = -1.0
    try:
        p, _, dev, mdev = _case(hip, 12, 16, 10, seed=7, edit=cold_cell)
        out = hip.put(np.zeros(2))
        assert _call(lib, p, dev, mdev, out) == 0
        assert hip.get(out, 2)[1] == 1.0
    finally:
        hip.free()


def test_noalloc_refuses_sizes_beyond_its_limits():
    lib = _lib()
    null = ctypes.c_void_p(0)
    args = [null] * 12 + [ctypes.c_double(1.4), ctypes.c_double(0.4), ctypes.c_double(1e12), null, None]
    assert lib.tt_schritt_noalloc(0, 16, 16, *args) == 1
    assert b"nx, ny, nz" in lib.tt_last_error()
    assert lib.tt_schritt_noalloc(2000, 2000, 2000, *args) == 1
    assert b"2^31" in lib.tt_last_error()
