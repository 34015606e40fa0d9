"""Times the reference Fortran's own step (oracle/_ref/ttref, flang -O2 as
the reference's makefile builds it, one MPI rank) on this host's CPU, for the
CPU baseline: the seeded workload state, dt from SCHRITT at step 0, steps
III = 2, 3, ... (no ADJUST_SOR_PARAM), each step's wall time from the
harness's system_clock.  Runs where the reference sources are (it needs
oracle/_ref/ttref, built by oracle/Makefile.ref).

    python oracle/time_reference.py [--size 32x64x32] [--steps 20]

Prints one JSON line: the median step time, every step's time, the SOR and
Newton counts of the last step, and the host's CPU model and count.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle.ttoracle import Oracle  # noqa: E402
from tests.golden import make_golden as M  # noqa: E402
from ttcode_amd import params as P  # noqa: E402


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    # the reference's automatic arrays (e.g. ppadv.f:535-543, ppenergy.f:19-26)
    # live on the stack: 36 MB each at 128x256x128
    import resource
    hard = resource.getrlimit(resource.RLIMIT_STACK)[1]
    resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", default="32x64x32", help="nx x ny x nz = n_r x n_phi x n_theta")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--seed", type=int, default=2024)
    a = ap.parse_args()
    nx, ny, nz = (int(v) for v in a.size.split("x"))
    epsmat = 1e-8 if nx * ny * nz <= 32 * 64 * 32 else 1e-10
    # the reference has no power-law case 26; its constant opacity (IKAPTYP 1)
    # at the power law's normalisation kap0 stands in
    p = P.make_params(nx, ny, nz, epsmat=epsmat, ikaptyp=1, kap0=P.KAP0)
    st = P.generate_state(p, seed=a.seed)
    inp = {k: st[k] for k in M.STATE}
    o = Oracle(p, threads=1)
    o.load_state(inp)
    o.bounds()
    o.pressg()
    dt = o.schritt()
    _, sc = M.run_ref("step", p, inp, delt=dt, iii=2, nsteps=a.steps)
    ms = [sc[f"step_ms_{s}"] for s in range(1, a.steps + 1)]
    out = {"size": a.size, "cells": nx * ny * nz, "epsmat": epsmat, "steps": a.steps,
           "median_ms_per_step": float(np.median(ms)), "ms_per_step": ms,
           "last_step_sor_iter": int(sc["SOR_ITER"]), "last_step_newton": int(sc["energy_iter"]),
           "ranks": 1, "threads": 1, "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
           "build": "flang -O2 (oracle/Makefile.ref), MPICH, one rank"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
