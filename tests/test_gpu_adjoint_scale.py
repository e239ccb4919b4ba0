"""Adjoints at larger sizes and over longer runs, against the oracle's complex-step tangent of the same
loss (no differencing noise) and central differences of the device forward.

WCSPH: a 64,000-fluid dam break (196,408 particles) over 200 steps with the dam-break calibration
loss.  The fluid starts from U(-1, 1) mm/s: from rest every pair has v_ij . x_ij = 0 exactly, the kink
of the viscosity switch, where velocity derivatives are one-sided (the adjoint takes the forward's
receding branch, the complex step a branch chosen by the sign of its O(eps^2) real-part term).
DEM: the polydisperse pour over 500 steps with the heap loss."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import prestige_amd as pa
from helpers import make_scene, upload_all
from prestige_amd import losses

pytestmark = pytest.mark.gpu


def _directional(gp, gv0, dp, dv):
    return float(gp @ dp) + float((gv0 * dv).sum())


def test_wcsph_adjoint_200_steps_64k():
    oracle.set_threads(16)
    d, arrs = make_scene(pa.SCENE_DAM_BREAK, n_target=64000, dx=0.0025)
    n, nf = int(d.n_particles), int(d.n_moving)
    rng = np.random.default_rng(7)
    arrs = dict(arrs)
    for k in ("vx", "vy", "vz"):
        v = np.array(arrs[k])
        v[:nf] += rng.uniform(-1e-3, 1e-3, nf)
        arrs[k] = v
    dx = d.h / 1.2
    extent = (d.grid_hi[0] - 3 * dx, d.grid_hi[1] - 3 * dx)
    target = losses.dam_break_target(np.stack([arrs[k] for k in "xyz"])[:, :nf], extent, seed=7)
    steps = 200

    def loss(x, v):
        return losses.dam_break_loss(x, v, d.mass, nf, extent, target)

    def forward(dp, dv, eps):
        dd = d.copy()
        dd.alpha, dd.c0 = d.alpha + eps * dp[0], d.c0 + eps * dp[1]
        with pa.Context(dd) as ctx:
            upload_all(ctx, dict(arrs, vx=arrs["vx"] + eps * dv[0], vy=arrs["vy"] + eps * dv[1],
                                 vz=arrs["vz"] + eps * dv[2]))
            ctx.step(steps)
            x = np.stack([ctx.download(k) for k in "xyz"])
            v = np.stack([ctx.download(k) for k in ("vx", "vy", "vz")])
        return loss(x, v)[0]

    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        ctx.forward_tape(steps, 20)
        x = np.stack([ctx.download(k) for k in "xyz"])
        v = np.stack([ctx.download(k) for k in ("vx", "vy", "vz")])
        _, gx, gv = loss(x, v)
        gp, gv0 = ctx.adjoint(gx, gv)
    p = oracle.params_from_desc(d)
    drng = np.random.default_rng(5)
    for k in range(2):
        dp = drng.standard_normal(2) * np.array([d.alpha, d.c0])
        dv = np.zeros((3, n))
        dv[:, :nf] = drng.standard_normal((3, nf)) * 0.01
        adj = _directional(gp, gv0, dp, dv)
        _, _, tx, tv = oracle.cs_tangent(p, arrs, dp, dv, steps)
        cs = float((gx * tx).sum() + (gv * tv).sum())
        fd = (forward(dp, dv, 1e-7) - forward(dp, dv, -1e-7)) / 2e-7
        print(f"direction {k}: adjoint {adj:.12e} complex step {cs:.12e} fd {fd:.12e}")
        # the same derivative up to summation order: measured 2e-14 to 3e-14
        assert abs(adj - cs) <= 1e-11 * abs(cs)
        # central differences at a 1e-7 relative step: measured 3e-7 to 4e-7
        assert abs(fd - cs) <= 1e-5 * abs(cs)
