"""DEM reverse mode on the device (gradients w.r.t. kn, restitution, mu and the
initial velocities) against the oracle's complex-step forward-mode FP64
tangent and central differences of the device forward.  The scene is the
polydisperse pour with +-0.5 m/s initial velocities, so pair and floor
contacts open, stick and slip inside the run.

The DEM loss is only piecewise smooth: a contact onset or a stick/slip switch
one step earlier changes the derivative discontinuously.  On this scene a
1e-9 m perturbation of the initial positions flips such an event within 300
steps and moves dL/de by 30 % in FP64 itself, so the mixed form (whose
trajectory differs from FP64 by ~3e-7) is compared over 150 steps, where both
trajectories take the same branches."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import prestige_amd as pa
from helpers import make_scene, upload_all

pytestmark = pytest.mark.gpu

NSTEPS = 300
NSTEPS_MIXED = 150


def _scene(precision=pa.F64):
    return make_scene(pa.SCENE_DEM_POUR, n_target=2000, seed=2, precision=precision)


def _seeds(d, seed=1):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((3, d.n_particles)), rng.standard_normal((3, d.n_particles))


def _device(d, arrs, sx, sv, every, nsteps=NSTEPS):
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        ctx.forward_tape(nsteps, every)
        x_before = ctx.download("x")
        gp, gv = ctx.adjoint(sx, sv)
        assert np.array_equal(x_before, ctx.download("x"))
    return gp, gv


def _cs(d, arrs, sx, sv, dparam, dv0, nsteps=NSTEPS):
    d64 = d.copy()
    d64.precision = pa.F64
    _, _, tx, tv = oracle.cs_tangent(oracle.params_from_desc(d64), arrs, dparam, dv0, nsteps)
    return float((sx * tx).sum() + (sv * tv).sum())


def _traj_error(d, arrs, nsteps):
    """Norm-wise relative velocity difference of the mixed and FP64 forwards."""
    out = {}
    for prec in (pa.F64, pa.MIXED):
        dd = d.copy()
        dd.precision = prec
        with pa.Context(dd) as ctx:
            upload_all(ctx, arrs)
            ctx.step(nsteps)
            out[prec] = np.stack([ctx.download(k) for k in ("vx", "vy", "vz")])
    return float(np.linalg.norm(out[pa.MIXED] - out[pa.F64]) / np.linalg.norm(out[pa.F64]))


@pytest.mark.parametrize("precision,nsteps,tol", [(pa.F64, NSTEPS, 1e-7), (pa.MIXED, NSTEPS_MIXED, 1e-4)])
def test_dem_adjoint_matches_complex_step(precision, nsteps, tol):
    d, arrs = _scene(precision)
    st = oracle.State(oracle.params_from_desc(d), arrs)
    assert st.step(nsteps) == 0 and st.contact_count() > 5
    sx, sv = _seeds(d)
    gp, gv = _device(d, arrs, sx, sv, every=50, nsteps=nsteps)
    rng = np.random.default_rng(3)
    dv = rng.standard_normal((3, d.n_particles))
    cases = [("kn", np.array([1.0, 0, 0]), None), ("e", np.array([0, 1.0, 0]), None),
             ("mu", np.array([0, 0, 1.0]), None), ("v0", None, dv),
             ("all", np.array([2e-5, 0.3, -0.2]), dv)]
    if precision == pa.MIXED:
        ev = _traj_error(d, arrs, nsteps)
        print(f"mixed vs FP64 trajectory: relative velocity difference {ev:.3e}")
    res = []
    for name, dp, dv0 in cases:
        ref = _cs(d, arrs, sx, sv, dp, dv0, nsteps)
        pred = (0.0 if dp is None else float(gp @ dp)) + (0.0 if dv0 is None else float((gv * dv0).sum()))
        print(f"precision {precision} {name}: adjoint {pred:.12e} complex-step {ref:.12e} "
              f"rel {abs(pred - ref) / abs(ref):.3e}")
        res.append((name, pred, ref))
    for name, pred, ref in res:
        assert abs(pred - ref) <= tol * abs(ref), (name, pred, ref)


def test_dem_adjoint_matches_central_differences():
    d, arrs = _scene(pa.F64)
    sx, sv = _seeds(d, seed=7)
    gp, _ = _device(d, arrs, sx, sv, every=60)

    def loss(**kw):
        dd = d.copy()
        for k, v in kw.items():
            setattr(dd, k, v)
        with pa.Context(dd) as ctx:
            upload_all(ctx, arrs)
            ctx.step(NSTEPS)
            x = np.stack([ctx.download(k) for k in "xyz"])
            v = np.stack([ctx.download(k) for k in ("vx", "vy", "vz")])
        return float((sx * x).sum() + (sv * v).sum())

    rel = 1e-6
    fd = []
    for name in ("kn", "restitution", "mu"):
        v0 = getattr(d, name)
        e = rel * v0
        kw_p, kw_m = {name: v0 + e}, {name: v0 - e}
        if name == "kn":   # kt = 2/7 kn moves with kn
            kw_p["kt"] = d.kt * (v0 + e) / v0
            kw_m["kt"] = d.kt * (v0 - e) / v0
        fd.append((loss(**kw_p) - loss(**kw_m)) / (2 * e))
    print("adjoint", gp, "fd", fd)
    for a, b in zip(gp, fd):
        assert abs(a - b) <= 1e-4 * abs(b)
