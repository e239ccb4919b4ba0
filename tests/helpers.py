"""Small shared helpers for the tests (no device access here)."""
from __future__ import annotations

import numpy as np

import oracle
import prestige_amd as pa

STATE_FIELDS = ("x", "y", "z", "vx", "vy", "vz")


def make_scene(scene_id, n_target=0, dx=0.0, seed=0, precision=pa.F64, **overrides):
    d, arrs = pa.scene(scene_id, n_target=n_target, dx=dx, seed=seed)
    d.precision = precision
    for k, v in overrides.items():
        setattr(d, k, v)
    return d, arrs


def upload_all(ctx, arrs):
    for k in ("x", "y", "z", "vx", "vy", "vz", "radius", "mass"):
        if k in arrs:
            ctx.upload(k, arrs[k])


def oracle_state(d, arrs):
    return oracle.State(oracle.params_from_desc(d), arrs)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))


def max_ulp(a, b):
    """Largest distance in units in the last place between two float64 arrays."""
    a = np.asarray(a, np.float64).view(np.int64)
    b = np.asarray(b, np.float64).view(np.int64)
    a = np.where(a < 0, np.int64(-(2**63)) - a, a)
    b = np.where(b < 0, np.int64(-(2**63)) - b, b)
    return int(np.max(np.abs(a - b))) if a.size else 0
