"""Gradients of the calibration losses against central differences (no device)."""
from __future__ import annotations

import numpy as np

import prestige_amd as pa
from prestige_amd import losses


def _fd_check(f, x, g, rng, eps, k=20):
    """Directional central differences of f at x along random directions vs <g, d>."""
    for _ in range(3):
        d = rng.standard_normal(x.shape)
        fd = (f(x + eps * d) - f(x - eps * d)) / (2 * eps)
        an = float((g * d).sum())
        assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)), (fd, an)


def test_heap_loss_gradient():
    rng = np.random.default_rng(0)
    x = rng.random((3, 500))
    t = losses.heap_target(x, seed=3)
    L, gx, gv = losses.heap_loss(x, t)
    assert L > 0 and np.all(gv == 0)
    _fd_check(lambda y: losses.heap_loss(y, t)[0], x, gx, rng, 1e-6)


def test_dam_break_loss_gradient():
    d, a = pa.scene(pa.SCENE_DAM_BREAK, n_target=1000, dx=0.01)
    n, nf = d.n_particles, d.n_moving
    rng = np.random.default_rng(1)
    x = np.stack([a["x"], a["y"], a["z"]]) + rng.uniform(-1e-3, 1e-3, (3, n))
    v = rng.uniform(-0.2, 0.2, (3, n))
    extent = (d.grid_hi[0] + 3 * 0.01 - 6 * 0.01, d.grid_hi[1] - 3 * 0.01)
    target = losses.dam_break_target(x[:, :nf], extent, seed=5)
    L, gx, gv = losses.dam_break_loss(x, v, d.mass, nf, extent, target)
    assert np.all(gx[:, nf:] == 0) and np.all(gv[:, nf:] == 0)
    _fd_check(lambda y: losses.dam_break_loss(y, v, d.mass, nf, extent, target)[0], x, gx, rng, 1e-7)
    _fd_check(lambda w: losses.dam_break_loss(x, w, d.mass, nf, extent, target)[0], v, gv, rng, 1e-6)
