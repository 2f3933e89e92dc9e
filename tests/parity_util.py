"""Helpers shared by the GPU parity tests."""
import numpy as np

REL_TOL = 1e-3     # per-pixel / per-path XYZ or radiance relative tolerance
ABS_FLOOR = 1e-6   # values below this compare absolutely
AGREE_FRACTION = 0.999


def agreement(a, b, rel=REL_TOL, floor=ABS_FLOOR):
    """Fraction of rows (last axis reduced) where |a-b| <= rel*max(|b|, floor)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    ok = np.abs(a - b) <= rel * np.maximum(np.abs(b), floor)
    if ok.ndim > 1:
        ok = ok.reshape(ok.shape[0], -1).all(axis=1)
    return float(ok.mean())


def camera_rays(rng, n, origin, target, spread):
    o = np.broadcast_to(np.asarray(origin, np.float32), (n, 3)).copy()
    d = (np.asarray(target, np.float32) - o) + rng.normal(0, spread, size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


def make_rays(rp, o, d, tmax=np.inf):
    rays = np.zeros(len(o), dtype=rp.RAY_DTYPE)
    rays["origin"] = o
    rays["direction"] = d
    rays["tmax"] = tmax
    return rays
