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
