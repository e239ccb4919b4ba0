"""DEM on the device vs the FP64 oracle (bit-identical FP64 form; mixed form
within stated norm-wise tolerances)."""
from __future__ import annotations

import numpy as np
import pytest

import prestige_amd as pa
from helpers import make_scene, oracle_state, rel_l2, upload_all

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "wx", "wy", "wz", "ax", "ay", "az")


@pytest.mark.parametrize("scene,kw,nsteps", [
    (pa.SCENE_DEM_LATTICE, dict(n_target=10000, seed=0), 300),
    (pa.SCENE_DEM_POUR, dict(n_target=20000, seed=2), 300),
])
def test_f64_bitexact(scene, kw, nsteps):
    d, arrs = make_scene(scene, **kw)
    st = oracle_state(d, arrs)
    assert st.step(nsteps) == 0
    assert st.contact_count() > 0
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        ctx.step(nsteps)
        got = {k: ctx.download(k) for k in FIELDS}
    for k in FIELDS:
        ref = st.get(k)
        bad = np.flatnonzero(got[k] != ref)
        assert bad.size == 0, f"{k}: {bad.size} differ, e.g. {got[k][bad[:3]]} vs {ref[bad[:3]]}"


def test_mixed_error():
    d, arrs = make_scene(pa.SCENE_DEM_POUR, n_target=20000, seed=2)
    st = oracle_state(d, arrs)
    assert st.step(300) == 0
    dm = d.copy()
    dm.precision = pa.MIXED
    with pa.Context(dm) as ctx:
        upload_all(ctx, arrs)
        ctx.step(300)
        got = {k: ctx.download(k) for k in FIELDS}
    ex = rel_l2(np.stack([got[k] for k in "xyz"]), np.stack([st.get(k) for k in "xyz"]))
    ev = rel_l2(np.stack([got[k] for k in ("vx", "vy", "vz")]), np.stack([st.get(k) for k in ("vx", "vy", "vz")]))
    print(f"mixed DEM: x {ex:.3e} v {ev:.3e}")
    assert ex < 1e-8 and ev < 1e-3


def _events(d, arrs, nsteps, kn_scale=1.0, chunks=1):
    dd = d.copy()
    dd.flags = pa.FLAG_DEM_EVENTS
    dd.kn = d.kn * kn_scale
    dd.kt = d.kt * kn_scale
    with pa.Context(dd) as ctx:
        upload_all(ctx, arrs)
        for _ in range(chunks):
            ctx.step(nsteps // chunks)
        x = ctx.download("x")
        return x, ctx.download("contact_events"), ctx.download("slip_events")


def test_event_fingerprints():
    """Contact and slip fingerprints: identical for identical runs (also split over several calls),
    unchanged by a 1e-13 relative change of kn, changed by a 2x stiffer contact; recording them leaves
    the trajectory bit-identical."""
    d, arrs = make_scene(pa.SCENE_DEM_POUR, n_target=3000, seed=2)
    nsteps = 300
    x0, c0, s0 = _events(d, arrs, nsteps)
    assert c0.sum() > 0 and s0.sum() > 0
    x1, c1, s1 = _events(d, arrs, nsteps, chunks=3)
    assert np.array_equal(x0, x1) and np.array_equal(c0, c1) and np.array_equal(s0, s1)
    _, c2, s2 = _events(d, arrs, nsteps, kn_scale=1.0 + 1e-13)
    assert np.array_equal(c0, c2) and np.array_equal(s0, s2)
    _, c3, s3 = _events(d, arrs, nsteps, kn_scale=2.0)
    assert not (np.array_equal(c0, c3) and np.array_equal(s0, s3))
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        ctx.step(nsteps)
        assert np.array_equal(ctx.download("x"), x0)
