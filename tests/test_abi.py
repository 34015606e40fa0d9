"""The C ABI: every entry point include/ttgpu.h declares is exported by
libttgpu.so; the Python mirrors of the header's structs and field table
agree with it.  No device calls (runs without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

from ttcode_amd import params as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "ttgpu.h")
LIB = os.path.join(ROOT, "ttcode_amd", "libttgpu.so")


def declared():
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\*\s]+?\b(tt_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("tt_init", "tt_finalize", "tt_upload", "tt_download", "tt_bounds", "tt_advect3d",
                 "tt_forgas3d", "tt_coupled_energy", "tt_solve_sor", "tt_schritt", "tt_diffusion",
                 "tt_step", "tt_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build libttgpu.so first (python -m ttcode_amd.build)"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    lib = ctypes.CDLL(LIB)  # loads without a GPU
    for n in declared():
        getattr(lib, n)


def test_struct_sizes_match_header():
    """sizeof of the ctypes mirrors equals the C structs (compiled probe)."""
    probe = r'''
#include <stdio.h>
#include "ttgpu.h"
int main(void){printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(tt_params), sizeof(tt_solver_opts),
 sizeof(tt_sor_stats), sizeof(tt_energy_stats), sizeof(tt_step_stats), sizeof(tt_limits),
 sizeof(tt_viscosity)); return 0;}
'''
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "p.c")
        open(c, "w").write(probe)
        exe = os.path.join(d, "p")
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        sizes = list(map(int, subprocess.check_output([exe]).split()))
    assert sizes == [ctypes.sizeof(P.TTParams), ctypes.sizeof(P.SolverOpts), ctypes.sizeof(P.SorStats),
                     ctypes.sizeof(P.EnergyStats), ctypes.sizeof(P.StepStats), ctypes.sizeof(P.Limits),
                     ctypes.sizeof(P.Viscosity)]


@pytest.mark.parametrize("f", sorted(P.FIELD_NAMES))
def test_field_table_matches_header(f):
    from oracle.ttoracle import Oracle
    o = Oracle(P.make_params(12, 16, 10))
    rc, lo, hi, rank, isint = o.field_info(f)
    assert rc == 0
    plo, phi, dt = P.field_bounds(f, 12, 16, 10)
    assert tuple(lo[:rank]) == tuple(plo) and tuple(hi[:rank]) == tuple(phi)
    assert bool(isint) == (dt == __import__("numpy").int32)


@pytest.mark.parametrize("field,value,msg", [("bcxmin", 8, "BCXMIN"), ("ibdrad", 5, "IBDRAD"),
                                             ("igrav", 2, "IGRAV"),
                                             ("ikaptyp", 15, "IKAPTYP"), ("iflimtyp", 3, "IFLIMTYP"),
                                             ("f_incident", 1, "F_INCIDENT"), ("isigtyp", 2, "ISIGTYP"),
                                             ("nx", 31, "even")])
def test_init_rejects_unsupported_switches(field, value, msg):
    """tt_init refuses switch values the kernels do not implement with a
    positive (fatal) code and a message, before touching a device."""
    lib = ctypes.CDLL(LIB)
    lib.tt_last_error.restype = ctypes.c_char_p
    p = P.make_params(32, 64, 32)
    setattr(p, field, value)
    rc = lib.tt_init(ctypes.byref(p))
    assert rc == 1
    assert msg in lib.tt_last_error().decode()
