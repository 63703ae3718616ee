"""Host-side helpers of the learned preconditioner: the parameter layout of the library's U-Net (the C layer table in
unet.hip), He initialisation for training from scratch, and weight files (.npz) for trained networks."""
import os

import numpy as np

# (name, cin, cout, taps, relu) in parameter order; per layer W[taps][cout][cin] then bias[cout]
LAYERS = [
    ("E0a", 4, 16, 27, True),
    ("E0b", 16, 16, 27, True),
    ("E1a", 16, 32, 27, True),
    ("E1b", 32, 32, 27, True),
    ("B2a", 32, 64, 27, True),
    ("B2b", 64, 64, 27, True),
    ("D1p", 64, 32, 27, False),
    ("D1b", 32, 32, 27, True),
    ("D0p", 32, 16, 27, False),
    ("D0b", 16, 16, 27, True),
    ("H", 16, 1, 1, False),
]

WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")


def param_count():
    return sum(t * co * ci + co for _, ci, co, t, _ in LAYERS)


def he_init(seed=0, head_scale=1.0):
    """He-normal weights (scaled for the activation that follows), zero biases."""
    rng = np.random.default_rng(seed)
    out = []
    for _, ci, co, t, relu in LAYERS:
        std = np.sqrt((2.0 if relu else 1.0) / (t * ci))
        w = rng.standard_normal(t * co * ci) * std
        out += [w.astype(np.float32), np.zeros(co, np.float32)]
    p = np.concatenate(out)
    if head_scale != 1.0:
        n_head = LAYERS[-1][1]
        p[-1 - n_head:-1] *= head_scale
    return p


def save(path, params, **meta):
    np.savez(path, params=np.asarray(params, np.float32), **{k: np.asarray(v) for k, v in meta.items()})


def load(path):
    with np.load(path, allow_pickle=False) as f:
        return f["params"].astype(np.float32)


def default_weights_path():
    return os.path.join(WEIGHTS_DIR, "unet_slosh.npz")
