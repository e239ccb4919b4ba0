"""WCSPH reverse mode on the device.

The device adjoint (hand-written, gather form, checkpoint-recompute) is
checked against the oracle's complex-step forward-mode FP64 tangent of the
same trajectory (dot-product form: <seed, J d> = <J^T seed, d>), against
central differences of the device forward, and for independence of the
checkpoint interval."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import prestige_amd as pa
from helpers import make_scene, upload_all

pytestmark = pytest.mark.gpu

NSTEPS = 20


def _scene(precision=pa.F64, seed=0):
    d, arrs = make_scene(pa.SCENE_DAM_BREAK, n_target=1000, dx=0.01, precision=precision)
    rng = np.random.default_rng(seed)
    nf = d.n_moving
    arrs = dict(arrs)
    for k in ("vx", "vy", "vz"):
        v = np.zeros(d.n_particles)
        v[:nf] = rng.uniform(-0.2, 0.2, nf)   # stirred start: viscosity active from step 1
        arrs[k] = v
    return d, arrs


def _seeds(d, seed=1):
    rng = np.random.default_rng(seed)
    n, nf = d.n_particles, d.n_moving
    sx = np.zeros((3, n))
    sv = np.zeros((3, n))
    sx[:, :nf] = rng.standard_normal((3, nf))
    sv[:, :nf] = rng.standard_normal((3, nf))
    return sx, sv


def _device_adjoint(d, arrs, sx, sv, every):
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        ctx.forward_tape(NSTEPS, every)
        xT = np.stack([ctx.download(k) for k in "xyz"])
        gp, gv = ctx.adjoint(sx, sv)
        xT_after = np.stack([ctx.download(k) for k in "xyz"])
    assert np.array_equal(xT, xT_after)   # the final forward state is restored
    return gp, gv, xT


def _cs_directional(d, arrs, sx, sv, dparam, dv0):
    p = oracle.params_from_desc(d)
    _, _, tx, tv = oracle.cs_tangent(p, arrs, dparam, dv0, NSTEPS)
    return float((sx * tx).sum() + (sv * tv).sum())


@pytest.mark.parametrize("precision,tol", [(pa.F64, 1e-9), (pa.MIXED, 1e-4)])
def test_adjoint_matches_complex_step(precision, tol):
    d, arrs = _scene(precision)
    sx, sv = _seeds(d)
    gp, gv, _ = _device_adjoint(d, arrs, sx, sv, every=5)
    rng = np.random.default_rng(2)
    nf = d.n_moving
    d64 = d.copy()
    d64.precision = pa.F64
    # each parameter and the initial-velocity field on its own, then all together
    dirs = []
    dirs.append((np.array([1.0, 0.0]), None))
    dirs.append((np.array([0.0, 1.0]), None))
    dv = np.zeros((3, d.n_particles))
    dv[:, :nf] = rng.standard_normal((3, nf))
    dirs.append((None, dv))
    dirs.append((np.array([0.7, -0.3]), dv))
    for dparam, dv0 in dirs:
        ref = _cs_directional(d64, arrs, sx, sv, dparam, dv0)
        pred = (0.0 if dparam is None else float(gp @ dparam)) + (0.0 if dv0 is None else float((gv * dv0).sum()))
        print(f"precision {precision}: adjoint {pred:.12e} complex-step {ref:.12e}")
        assert abs(pred - ref) <= tol * abs(ref), (pred, ref)
    assert np.all(gv[:, nf:] == 0.0)


def test_checkpoint_interval_does_not_change_gradients():
    d, arrs = _scene(pa.F64)
    sx, sv = _seeds(d)
    g1 = _device_adjoint(d, arrs, sx, sv, every=1)
    g7 = _device_adjoint(d, arrs, sx, sv, every=7)
    g20 = _device_adjoint(d, arrs, sx, sv, every=20)
    for a, b in ((g1, g7), (g1, g20)):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_adjoint_matches_central_differences():
    """Parameter gradients vs central differences of the device forward (FP64)."""
    d, arrs = _scene(pa.F64, seed=4)
    sx, sv = _seeds(d, seed=5)
    gp, _, _ = _device_adjoint(d, arrs, sx, sv, every=10)

    def loss(alpha, c0):
        dd = d.copy()
        dd.alpha, dd.c0 = alpha, c0
        with pa.Context(dd) as ctx:
            upload_all(ctx, arrs)
            ctx.step(NSTEPS)
            x = np.stack([ctx.download(k) for k in "xyz"])
            v = np.stack([ctx.download(k) for k in ("vx", "vy", "vz")])
        return float((sx * x).sum() + (sv * v).sum())

    ea, ec = 1e-5 * d.alpha, 1e-5 * d.c0
    fa = (loss(d.alpha + ea, d.c0) - loss(d.alpha - ea, d.c0)) / (2 * ea)
    fc = (loss(d.alpha, d.c0 + ec) - loss(d.alpha, d.c0 - ec)) / (2 * ec)
    print(f"dalpha adjoint {gp[0]:.10e} fd {fa:.10e}; dc0 adjoint {gp[1]:.10e} fd {fc:.10e}")
    assert abs(gp[0] - fa) <= 1e-4 * abs(fa)
    assert abs(gp[1] - fc) <= 1e-4 * abs(fc)


def test_adjoint_requires_tape():
    d, arrs = _scene(pa.F64)
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        with pytest.raises(pa.PrestigeError) as ei:
            ctx.adjoint()
        assert ei.value.name == "E_STATE"
