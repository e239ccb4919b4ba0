"""Calibration losses of the differentiable workloads and their gradients with
respect to the final state (the seeds of prestige_adjoint).

heap_loss (DEM pour): mean squared distance of the final positions to a
seeded target heap, L = (1/N) sum_i |x_i - t_i|^2.

dam_break_loss (WCSPH dam break): kinetic energy of the fluid plus the squared
deviation of the free-surface height on a 64 x 32 grid of columns over the
tank floor from a seeded target.  The height of column c is the smooth
maximum H_c = (1/beta) log(1 + sum_i phi_ic exp(beta z_i)) over fluid
particles, phi_ic = (1 - (rho_ic/s)^2)^2 for horizontal distance rho_ic < s
(column spacing s), so H_c is differentiable in every particle position and
an empty column has height 0.

All arrays are (3, n) or (n,) in particle-id order; float64.
"""
from __future__ import annotations

import numpy as np

GRID = (64, 32)


def heap_target(x0: np.ndarray, seed: int, spread: float = 2e-3) -> np.ndarray:
    """Target positions: the initial positions displaced by U(-spread, spread) per axis."""
    rng = np.random.default_rng(seed)
    return x0 + rng.uniform(-spread, spread, x0.shape)


def heap_loss(x: np.ndarray, target: np.ndarray):
    """Returns (L, dL/dx, dL/dv)."""
    n = x.shape[1]
    r = x - target
    return float((r * r).sum() / n), 2.0 * r / n, np.zeros_like(x)


def _columns(xy: np.ndarray, extent, grid=GRID):
    """For each particle, the up-to-9 columns within the support radius and their weights."""
    lx, ly = extent
    nx, ny = grid
    sx, sy = lx / nx, ly / ny
    s = max(sx, sy)
    ci = np.floor(xy[0] / sx).astype(np.int64)
    cj = np.floor(xy[1] / sy).astype(np.int64)
    out = []
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            a, b = ci + di, cj + dj
            ok = (a >= 0) & (a < nx) & (b >= 0) & (b < ny)
            cx, cy = (a + 0.5) * sx, (b + 0.5) * sy
            ddx, ddy = xy[0] - cx, xy[1] - cy
            u2 = (ddx * ddx + ddy * ddy) / (s * s)
            ok &= u2 < 1.0
            w = np.where(ok, (1.0 - u2) ** 2, 0.0)
            # d w / d(x, y) = -4 (1 - u2) (d, d) / s^2
            gx = np.where(ok, -4.0 * (1.0 - u2) * ddx / (s * s), 0.0)
            gy = np.where(ok, -4.0 * (1.0 - u2) * ddy / (s * s), 0.0)
            idx = np.where(ok, a * ny + b, 0)
            out.append((idx, w, gx, gy))
    return out


def surface_heights(x: np.ndarray, extent, beta: float = 50.0, grid=GRID):
    """Smooth-maximum heights H (nx*ny,) of the fluid particles' positions x (3, n_fluid)."""
    cols = _columns(x[:2], extent, grid)
    ez = np.exp(beta * x[2])
    S = np.zeros(grid[0] * grid[1])
    for idx, w, _, _ in cols:
        np.add.at(S, idx, w * ez)
    return np.log1p(S) / beta, S, cols, ez


def dam_break_target(x0_fluid: np.ndarray, extent, seed: int, beta: float = 50.0, spread: float = 0.01):
    H0, _, _, _ = surface_heights(x0_fluid, extent, beta)
    rng = np.random.default_rng(seed)
    return H0 + rng.uniform(-spread, spread, H0.shape)


def dam_break_loss(x: np.ndarray, v: np.ndarray, mass: float, n_fluid: int, extent, target: np.ndarray,
                   beta: float = 50.0):
    """Returns (L, dL/dx, dL/dv) with x, v of shape (3, n) (boundary rows get zero gradient)."""
    xf, vf = x[:, :n_fluid], v[:, :n_fluid]
    ke = 0.5 * mass * float((vf * vf).sum())
    H, S, cols, ez = surface_heights(xf, extent, beta)
    dH = H - target
    L = ke + float((dH * dH).sum())
    gH = 2.0 * dH / (1.0 + S)               # dL/dS_c = dL/dH_c * dH_c/dS_c
    gx = np.zeros_like(x)
    gv = np.zeros_like(v)
    gv[:, :n_fluid] = mass * vf
    for idx, w, wx, wy in cols:
        g = gH[idx]
        gx[0, :n_fluid] += g * wx * ez / beta
        gx[1, :n_fluid] += g * wy * ez / beta
        gx[2, :n_fluid] += g * w * ez
    return L, gx, gv
