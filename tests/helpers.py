"""Shared test helpers."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


def random_fields(shape, seed, b0=None, scale=1e-3):
    rng = np.random.default_rng(seed)
    f = {}
    for n in ("ex", "ey", "ez"):
        f[n] = rng.standard_normal(shape) * scale
    for i, n in enumerate(("bx", "by", "bz")):
        base = 0.0 if b0 is None else b0[i]
        f[n] = base + rng.standard_normal(shape) * scale
    return f


def upload_fields(sim, f):
    for n, a in f.items():
        sim.upload_field(n, a)
