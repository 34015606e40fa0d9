"""Checkpoint files in the reference's own formats, so that runs move between
ttcode and this library in both directions.

* PRINTOUT(1) (ppprt.f:13-124) writes each rank's "recent dataset" --
  SNAP_<rank>.<n> during a run and AKTTD_<rank> at its end (pp.f:569-582,
  826-830) -- as Fortran sequential unformatted records: the grid sizes and
  decomposition, ZEIT/DELT/NUM_ITER, the 1-D grid, V, G, H, RH, T, ER, PG
  over their ghosted bounds, CS, upperbnd, the initial-profile scalars, then
  diagnostics (NCOSYS, kappa_R, the x limiter, convective flux, kappa_A,
  STELLARINPUT, viscous dissipation, kappa_P).  ``write_printout`` writes it
  for the one-rank decomposition this library runs (numprocs 1, rank 0).
* READRESTART (ppinit.f:1543-1600) reads the global restart file AKTTD
  (interior cells of the whole grid, no ghosts): ``write_restart`` writes
  it, ``read_restart`` reads it, and ``restart_state`` turns it into the
  wire-format arrays ``Device.load_state`` takes.

A record is a 4-byte little-endian length, the payload, the length again:
the layout flang and gfortran write for sequential unformatted files.
Records of 2 GiB or more are refused (the compilers split those into
subrecords).  The state comes from any object with ``download(name)`` and
``metric(id)`` (``ttgpu.Device``).
"""
from __future__ import annotations

import struct

import numpy as np

from . import params as P

# metric ids (include/ttgpu.h)
_M = {"xxa": 0, "xxb": 1, "dxa": 2, "xya": 4, "xyb": 5, "dya": 6, "xza": 8, "xzb": 9, "dza": 10}
INIT_SCALARS = ("TinitBottom", "Tinittop", "rhinittop", "rhinitbottom", "hydrotime", "Maccreated")
MAX_RECORD = (1 << 31) - 1


def _record(f, payload: bytes):
    n = len(payload)
    if n > MAX_RECORD:
        raise ValueError(f"record of {n} bytes: the 4-byte length markers hold at most {MAX_RECORD}")
    f.write(struct.pack("<i", n))
    f.write(payload)
    f.write(struct.pack("<i", n))


def _f8(a) -> bytes:
    return np.ascontiguousarray(np.asarray(a, dtype="<f8")).tobytes()


def _i4(*v) -> bytes:
    return struct.pack(f"<{len(v)}i", *v)


def _sub(src, name, nx, ny, nz, lo, hi):
    """Cells lo..hi (inclusive, per axis i, j, k) of a field, in Fortran order."""
    f = P.FIELD_IDS[name]
    a = src.download(name)
    flo, _, _ = P.field_bounds(f, nx, ny, nz)
    k0, j0, i0 = (lo[2] - flo[2], lo[1] - flo[1], lo[0] - flo[0])
    return a[k0:k0 + hi[2] - lo[2] + 1, j0:j0 + hi[1] - lo[1] + 1, i0:i0 + hi[0] - lo[0] + 1]


def write_printout(path, src, p: P.TTParams, zeit: float, delt: float, num_iter: int, init=(0.0,) * 6):
    """PRINTOUT(1) of ``src``'s state to ``path`` (ppprt.f:15-123): one rank
    holding the whole grid.  ``init`` = (TinitBottom, Tinittop, rhinittop,
    rhinitbottom, hydrotime, Maccreated).  conv_flux is the last ADVECT3D's
    thermal advection flux (ppadv.f:857-865) and DISFN the last FORGAS3D's
    viscous dissipation (zero without viscosity, as pp.f:454 starts it)."""
    nx, ny, nz = p.nx, p.ny, p.nz
    inner = ((1, 1, 1), (nx, ny, nz))
    with open(path, "wb") as f:
        _record(f, _i4(nx, ny, nz))
        _record(f, _i4(nx, ny, nz, 1))
        _record(f, _i4(0, 0, 0, 0))
        _record(f, _i4(1, 1, 1))
        _record(f, struct.pack("<ddi", zeit, delt, num_iter))
        for m in ("xxa", "xxb", "dxa", "xya", "xyb", "dya", "xza", "xzb", "dza"):
            _record(f, _f8(src.metric(_M[m])))
        for name in ("V", "G", "H", "RH", "T", "ER", "PG"):  # their whole declared boxes
            _record(f, _f8(src.download(name)))
        _record(f, _f8(_sub(src, "CS", nx, ny, nz, (0, 0, 0), (nx + 1, ny + 1, nz + 1))))
        ub = src.download("upperbnd")[1 + 1:1 + 1 + nz, 1 + 1:1 + 1 + ny]  # (1:ny, 1:nz) of (-1:ny+2, -1:nz+2)
        _record(f, np.ascontiguousarray(ub, dtype="<i4").tobytes())
        _record(f, struct.pack("<6d", *init))
        _record(f, _i4(2))  # NCOSYS
        _record(f, _f8(_sub(src, "xkapR", nx, ny, nz, *inner)))
        _record(f, _f8(_sub(src, "FLIMX", nx, ny, nz, *inner)))
        for c in (1, 2, 3):  # conv_flux(c, 1:nx, 1:ny, 1:nz)
            _record(f, _f8(src.download(f"conv_flux{c}")))
        _record(f, _f8(_sub(src, "xkapA", nx, ny, nz, *inner)))
        _record(f, _f8(_sub(src, "STELLARINPUT", nx, ny, nz, *inner)))
        _record(f, _f8(_sub(src, "DISFN", nx, ny, nz, *inner)))
        _record(f, _f8(_sub(src, "xkapP", nx, ny, nz, *inner)))


class _Reader:
    def __init__(self, path):
        with open(path, "rb") as f:
            self.b = f.read()
        self.o = 0

    def record(self) -> bytes:
        if self.o + 4 > len(self.b):
            raise ValueError("truncated file: no record left")
        (n,) = struct.unpack_from("<i", self.b, self.o)
        if n < 0 or self.o + 8 + n > len(self.b):
            raise ValueError(f"bad record length {n} at byte {self.o}")
        (m,) = struct.unpack_from("<i", self.b, self.o + 4 + n)
        if m != n:
            raise ValueError(f"record at byte {self.o}: trailing length {m} != {n}")
        out = self.b[self.o + 4:self.o + 4 + n]
        self.o += 8 + n
        return out

    def f8(self, shape):
        r = self.record()
        a = np.frombuffer(r, dtype="<f8")
        if a.size != int(np.prod(shape)):
            raise ValueError(f"record holds {a.size} values, {int(np.prod(shape))} expected")
        return a.reshape(shape).copy()

    def ints(self, n):
        r = self.record()
        if len(r) != 4 * n:
            raise ValueError(f"record of {len(r)} bytes, {4 * n} expected")
        return struct.unpack(f"<{n}i", r)

    def done(self):
        return self.o == len(self.b)


def read_printout(path):
    """The records of a PRINTOUT(1) file (one rank's view) as a dict: the
    header values, the 1-D grid, and each array over the bounds it was
    written with (numpy axes k, j, i)."""
    r = _Reader(path)
    out = {}
    nx, ny, nz = r.ints(3)
    lnx, lny, lnz, nprocs = r.ints(4)
    out.update(global_n=(nx, ny, nz), local_n=(lnx, lny, lnz), numprocs=nprocs)
    out["myid"], *coords = r.ints(4)
    out["proc_coords"] = tuple(coords)
    out["proc_dims"] = r.ints(3)
    zeit, delt, num_iter = struct.unpack("<ddi", r.record())
    out.update(ZEIT=zeit, DELT=delt, NUM_ITER=num_iter)
    for m, n in (("xxa", lnx + 3), ("xxb", lnx + 4), ("dxa", lnx + 2), ("xya", lny + 3), ("xyb", lny + 4),
                 ("dya", lny + 2), ("xza", lnz + 3), ("xzb", lnz + 4), ("dza", lnz + 2)):
        out[m] = r.f8((n,))
    full = (lnz + 4, lny + 4, lnx + 4)
    out["V"] = r.f8((lnz + 4, lny + 4, lnx + 3))
    out["G"] = r.f8((lnz + 4, lny + 3, lnx + 4))
    out["H"] = r.f8((lnz + 3, lny + 4, lnx + 4))
    for name in ("RH", "T", "ER", "PG"):
        out[name] = r.f8(full)
    out["CS"] = r.f8((lnz + 2, lny + 2, lnx + 2))
    out["upperbnd"] = np.frombuffer(r.record(), dtype="<i4").reshape(lnz, lny).copy()
    out.update(zip(INIT_SCALARS, struct.unpack("<6d", r.record())))
    (out["NCOSYS"],) = r.ints(1)
    inner = (lnz, lny, lnx)
    for name in ("xkapR", "FLIMX", "conv_flux_1", "conv_flux_2", "conv_flux_3", "xkapA", "STELLARINPUT",
                 "DISFN", "xkapP"):
        out[name] = r.f8(inner)
    if not r.done():
        raise ValueError("bytes left after the last PRINTOUT(1) record")
    return out


def write_restart(path, src, p: P.TTParams, zeit: float, delt: float, num_iter: int, init=(0.0,) * 6):
    """The global restart file READRESTART reads (ppinit.f:1563-1598):
    the interior cells of V, G, H, RH, T, ER, PG, CS, upperbnd and the 1-D
    cell centres, for a grid held whole by one rank."""
    nx, ny, nz = p.nx, p.ny, p.nz
    inner = ((1, 1, 1), (nx, ny, nz))
    with open(path, "wb") as f:
        _record(f, _i4(nx))
        _record(f, _i4(ny))
        _record(f, _i4(nz))
        _record(f, _i4(nx, ny, nz, 1))
        _record(f, struct.pack("<ddi", zeit, delt, num_iter))
        for m, n in (("xxb", nx), ("xyb", ny), ("xzb", nz)):
            _record(f, _f8(src.metric(_M[m])[2:2 + n]))  # (1:n) of (-1:n+2)
        for name in ("V", "G", "H", "RH", "T", "ER", "PG", "CS"):
            _record(f, _f8(_sub(src, name, nx, ny, nz, *inner)))
        ub = src.download("upperbnd")[2:2 + nz, 2:2 + ny]
        _record(f, np.ascontiguousarray(ub, dtype="<i4").tobytes())
        _record(f, struct.pack("<6d", *init))


def read_restart(path, nx, ny, nz):
    """READRESTART's reads (ppinit.f:1563-1598) of a global restart file for
    an nx x ny x nz grid.  Arrays have numpy axes k, j, i over the interior;
    the sizes the file declares must match."""
    r = _Reader(path)
    n = (r.ints(1)[0], r.ints(1)[0], r.ints(1)[0])
    if n != (nx, ny, nz):
        raise ValueError(f"restart file is for a {n} grid, not {(nx, ny, nz)}")
    out = {"local_n_and_numprocs": r.ints(4)}
    zeit, delt, num_iter = struct.unpack("<ddi", r.record())
    out.update(ZEIT=zeit, DELT=delt, NUM_ITER=num_iter)
    out["xxb"], out["xyb"], out["xzb"] = r.f8((nx,)), r.f8((ny,)), r.f8((nz,))
    for name in ("V", "G", "H", "RH", "T", "ER", "PG", "CS"):
        out[name] = r.f8((nz, ny, nx))
    out["upperbnd"] = np.frombuffer(r.record(), dtype="<i4").reshape(nz, ny).copy()
    out.update(zip(INIT_SCALARS, struct.unpack("<6d", r.record())))
    if not r.done():
        raise ValueError("bytes left after the last restart record")
    return out


def restart_state(rest: dict, p: P.TTParams, base: dict | None = None):
    """Wire-format arrays for ``Device.load_state`` from ``read_restart``'s
    result: LOCALINIT's copy of the interior cells of T, ER, RH, V, G, H
    (ppinit.f:23-49), for BOUNDS to complete as LOCALINIT does next.

    Everything else comes from ``base`` when given (the run's initial
    model): the fields a restart file does not carry (gravity, mu_star) and
    the ghost cells of the six.  Not all ghost cells are boundary
    conditions: with BCXMIN=12, BOUNDS never sets RH below the bottom (i =
    0, -1; LOCALINIT's fill of them is commented out, ppinit.f:79-87), so
    those keep the initial model's values for the whole run and only the
    initial model can restore them.  Without ``base`` the ghosts are zero,
    as the reference's freshly allocated arrays would be."""
    nx, ny, nz = p.nx, p.ny, p.nz
    st = dict(base or {})
    for name in ("T", "ER", "RH", "V", "G", "H"):
        f = P.FIELD_IDS[name]
        a = np.array(st[name], copy=True) if name in st else P.empty_field(f, nx, ny, nz)
        lo, _, _ = P.field_bounds(f, nx, ny, nz)
        a[1 - lo[2]:1 - lo[2] + nz, 1 - lo[1]:1 - lo[1] + ny, 1 - lo[0]:1 - lo[0] + nx] = rest[name]
        st[name] = a
    return st
