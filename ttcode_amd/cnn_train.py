"""Training of the learned SOR initial guess on the library's own FP64 solves.

    python -m ttcode_amd.cnn_train --samples 2000 --out ttcode_amd/data/cnn_guess.npz

Each configuration is the plain-method generator on the n_r x n_phi x
n_theta = 32 x 64 x 32 grid with seeded opacity exponents a in [0, 1],
b in [1, 3] (kappa = kappa0 (rho/rho0)^a (T/T0)^b) and an irradiation flux
amplitude in [0.5, 2] (T_irr scaled by amplitude^(1/4)).  A few steps are
run with collection on: the first Newton iteration of every
COUPLED_ENERGY_EQS call adds one sample -- the network's five features and
the correction delta_ER / (E dt) that the FP64 red-black SOR produced -- to the
training set on the device.  The library then fits the normalisation and
trains with Adam using its own BF16 matrix-core forward and backward kernels.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    __package__ = "ttcode_amd"

from . import params as P  # noqa: E402
from . import ttgpu  # noqa: E402

GRID = (32, 64, 32)   # nx (r), ny (phi), nz (theta)


def sample_params(rng, nx=GRID[0], ny=GRID[1], nz=GRID[2], epsmat=1e-8):
    a = float(rng.uniform(0.0, 1.0))
    b = float(rng.uniform(1.0, 3.0))
    amp = float(rng.uniform(0.5, 2.0))
    p = P.make_params(nx, ny, nz, epsmat=epsmat, kap_a=a, kap_b=b, tirr=P.TIRR * amp ** 0.25)
    return p, (a, b, amp)


def collect(samples, steps_per_config=4, seed=0, grid=GRID, epsmat=1e-8, log=None):
    """Runs configurations until `samples` training samples are on the
    device.  Returns the last Device (the training set lives in the library
    until tt_cnn_release)."""
    rng = np.random.default_rng(seed)
    dev = None
    cfg = 0
    t0 = time.time()
    while True:
        p, knobs = sample_params(rng, *grid, epsmat=epsmat)
        if dev is not None:
            dev.close()
        dev = ttgpu.Device(p)
        if cfg == 0:
            dev.cnn_reserve(samples, *grid)
            dev.cnn_collect(True)
        st = P.generate_state(p, seed=int(rng.integers(1 << 30)))
        dev.load_state(st)
        dev.bounds()
        dev.pressg()
        dt = dev.schritt()
        for s in range(steps_per_config):
            dev.step(dt, 2 + s)
            if dev.cnn_dataset_size() >= samples:
                break
        cfg += 1
        n = dev.cnn_dataset_size()
        if log and cfg % 50 == 0:
            log(f"collected {n}/{samples} samples from {cfg} configurations ({time.time() - t0:.1f} s)")
        if n >= samples:
            break
    dev.cnn_collect(False)
    return dev


def train(dev, epochs=20, batch=8, lr=3e-3, seed=0, init_seed=0, log=None):
    """Fits the normalisation, initialises the weights and trains; returns
    the per-epoch losses."""
    from .cnn_init import init_params
    w, b = init_params(init_seed)
    dev.cnn_set_weights(w, b)
    dev.cnn_fit_norm()
    t0 = time.time()
    # step schedule: half the epochs at lr, a quarter at lr/4, the rest at
    # lr/16 (Adam's moments carry over between the phases)
    e1 = max(1, epochs // 2)
    e2 = max(1, epochs // 4) if epochs > e1 else 0
    e3 = epochs - e1 - e2
    parts = [dev.cnn_train(e1, batch, lr, seed)]
    if e2 > 0:
        parts.append(dev.cnn_train(e2, batch, lr / 4, seed + 1))
    if e3 > 0:
        parts.append(dev.cnn_train(e3, batch, lr / 16, seed + 2))
    losses = np.concatenate(parts)
    if log:
        log(f"trained {epochs} epochs in {time.time() - t0:.1f} s: loss {losses[0]:.4g} -> {losses[-1]:.4g}")
    return losses


def save(dev, path, losses=None):
    w, b = dev.cnn_get_weights()
    norm = dev.cnn_get_norm()
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez_compressed(path, w=w, b=b, norm=norm, losses=np.asarray(losses if losses is not None else []))


def load(dev, path):
    d = np.load(path)
    dev.cnn_set_weights(d["w"], d["b"])
    dev.cnn_set_norm(d["norm"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=2000)
    ap.add_argument("--steps-per-config", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                                                  "cnn_guess.npz"))
    a = ap.parse_args()
    log = lambda s: print(s, flush=True)  # noqa: E731
    dev = collect(a.samples, a.steps_per_config, a.seed, log=log)
    losses = train(dev, a.epochs, a.batch, a.lr, a.seed, log=log)
    save(dev, a.out, losses)
    log(f"wrote {a.out}")
    dev.cnn_release()
    dev.close()


if __name__ == "__main__":
    main()
