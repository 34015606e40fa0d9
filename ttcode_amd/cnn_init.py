"""Initial parameters of the learned-guess network (layout of
include/ttgpu.h: per layer W[tap][ci][co], then the biases)."""
from __future__ import annotations

import numpy as np

CH = [5, 16, 16, 16, 16, 1]


def init_params(seed=0):
    """He-style initialisation, W ~ N(0, 2 / (27 ci)), biases 0; the last
    layer is scaled by 0.1 so that an untrained guess starts near zero."""
    rng = np.random.default_rng(seed)
    ws, bs = [], []
    for l in range(len(CH) - 1):
        ci, co = CH[l], CH[l + 1]
        w = rng.standard_normal((27, ci, co)) * np.sqrt(2.0 / (27 * ci))
        if l == len(CH) - 2:
            w *= 0.1
        ws.append(w.ravel())
        bs.append(np.zeros(co))
    return np.concatenate(ws).astype(np.float32), np.concatenate(bs).astype(np.float32)
