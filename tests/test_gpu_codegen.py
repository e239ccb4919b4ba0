"""Generated equation passes on the device (prestige_run_equations): the
generated eq1 reproduces the hand-written eq1 kernels bit for bit in both loop
forms; defined equations match a numpy restatement of the same sums (exact
counts; sums to 1e-12 relative, as their order differs from numpy's)."""
from __future__ import annotations

import numpy as np
import pytest

import prestige_amd as pa
from helpers import make_scene, upload_all

pytestmark = pytest.mark.gpu


def _pairs_ctx_desc(n, box, cell, seed):
    """PAIRS context over n random points in a box (the equation-probe model)."""
    rng = np.random.default_rng(seed)
    d = pa.desc_init(pa.MODEL_PAIRS)
    d.n_particles = d.n_moving = n
    d.cell_size = cell
    for a in range(3):
        d.grid_lo[a] = 0.0
        d.grid_hi[a] = box
    pts = rng.uniform(0.0, box, (n, 3))
    return d, {"x": pts[:, 0].copy(), "y": pts[:, 1].copy(), "z": pts[:, 2].copy()}


@pytest.mark.parametrize("loop", [pa.LOOP_ALLPAIRS, pa.LOOP_CELLS])
def test_generated_eq1_matches_builtin(loop):
    rng = np.random.default_rng(2)
    d, arrs = make_scene(pa.SCENE_DAM_BREAK, n_target=8000, dx=0.005)
    if loop == pa.LOOP_ALLPAIRS:
        d, arrs = make_scene(pa.SCENE_DEM_LATTICE, n_target=3000, seed=1)
    cutoff = d.cell_size
    force0 = rng.standard_normal(d.n_particles)
    mass = rng.random(d.n_particles)
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        ctx.neighbors_build()
        ctx.upload("force", force0)
        ctx.upload("mass", mass)
        ctx.eval_equation(pa.EQ1, loop, cutoff)
        ref = ctx.download("force")
        ctx.upload("force", force0)
        ctx.run_equations([pa.EQ1], loop, cutoff)
        got = ctx.download("force")
    assert np.array_equal(got, ref)


def _numpy_sums(x, y, z, mass, cutoff):
    """Restatement of the moment / twice / nb bodies below over the pairs
    within cutoff (j == i included); cutoff None means all pairs."""
    P = np.stack([x, y, z], 1)
    n = len(x)

    def pairs(sl):
        if cutoff is None:
            return np.ones((sl.stop - sl.start, n), dtype=bool)
        return ((P[sl, None, :] - P[None, :, :]) ** 2).sum(2) < cutoff * cutoff

    c, s, q = np.zeros(n), np.zeros(n), np.zeros(n)
    for i0 in range(0, n, 512):
        sl = slice(i0, min(n, i0 + 512))
        nb = pairs(sl)
        dx = x[sl, None] - x[None, :]
        c[sl] = nb.sum(1)
        s[sl] = np.where(nb, mass[None, :] * dx * dx, 0.0).sum(1)
    m = 2.0 * s
    for i0 in range(0, n, 512):
        sl = slice(i0, min(n, i0 + 512))
        q[sl] = np.where(pairs(sl), m[None, :] * np.sqrt(np.abs(y[sl, None] - y[None, :])), 0.0).sum(1)
    return c, s, m, q


@pytest.mark.parametrize("loop", [pa.LOOP_ALLPAIRS, pa.LOOP_CELLS])
def test_defined_equations_match_numpy(loop):
    n = 2000 if loop == pa.LOOP_ALLPAIRS else 6000
    d, arrs = _pairs_ctx_desc(n, box=1.0, cell=0.1, seed=7)
    mass = np.random.default_rng(8).random(n)
    eqs = [pa.define_equation("moment", "{ let d = x[i] - x[j]; s[i] += mass[j] * d * d; c[i] += 1.0 }"),
           pa.define_equation("twice", "{ m[i] = 2.0 * s[i]; }"),
           pa.define_equation("nb", "{ q[i] += m[j] * (y[i] - y[j]).abs().sqrt() }")]
    assert pa.fuse_schedule(eqs) == ([0, 0, 1], 2)
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        ctx.upload("mass", mass)
        ctx.run_equations(eqs, loop, d.cell_size)
        got = {k: ctx.download(k) for k in ("c", "s", "m", "q")}
    c, s, m, q = _numpy_sums(arrs["x"], arrs["y"], arrs["z"], mass,
                             None if loop == pa.LOOP_ALLPAIRS else d.cell_size)
    assert np.array_equal(got["c"], c)
    for k, ref in (("s", s), ("m", m), ("q", q)):
        assert np.allclose(got[k], ref, rtol=1e-12, atol=0.0), k
    if loop == pa.LOOP_ALLPAIRS:
        assert np.all(got["c"] == n)


@pytest.mark.parametrize("loop", [pa.LOOP_ALLPAIRS, pa.LOOP_CELLS])
def test_field_read_and_written_in_one_pass_uses_pre_pass_values(loop):
    """a[i] += 0.5 a[j] reads a at [j] and writes it: every j sees the values
    from before the pass (a separate output array), not partial updates; a
    second call starts from the first call's result."""
    n = 1500
    d, arrs = _pairs_ctx_desc(n, box=1.0, cell=0.1, seed=3)
    a0 = np.random.default_rng(4).random(n)
    e = pa.define_equation("smooth", "{ a[i] += 0.5 * a[j] }")
    P = np.stack([arrs["x"], arrs["y"], arrs["z"]], 1)
    if loop == pa.LOOP_CELLS:
        nb = ((P[:, None, :] - P[None, :, :]) ** 2).sum(2) < d.cell_size ** 2
    else:
        nb = np.ones((n, n), dtype=bool)
    step = lambda a: a + 0.5 * np.where(nb, a[None, :], 0.0).sum(1)  # noqa: E731
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        ctx.upload("a", a0)
        ctx.run_equations([e], loop, d.cell_size)
        got1 = ctx.download("a")
        ctx.run_equations([e], loop, d.cell_size)
        got2 = ctx.download("a")
    assert np.allclose(got1, step(a0), rtol=1e-12, atol=0.0)
    assert np.allclose(got2, step(step(a0)), rtol=1e-12, atol=0.0)


def test_missing_input_field_is_reported():
    d, arrs = _pairs_ctx_desc(500, box=1.0, cell=0.1, seed=5)
    e = pa.define_equation("needs_w", "{ r[i] += w[j] }")
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        with pytest.raises(pa.PrestigeError) as ei:
            ctx.run_equations([e], pa.LOOP_CELLS, d.cell_size)
        assert ei.value.name == "E_STATE" and "'w'" in str(ei.value)


def test_generated_density_and_eos_match_the_wcsph_kernel():
    """The built-in density and EOS bodies, generated and run over the cell
    iterator, give the hand-written WCSPH density pass's rho and p (the same
    sums in another association: rho to 1e-13, p to 1e-10 relative, p being
    B ((rho/rho0)^7 - 1) with ~1e-3 of cancellation)."""
    d, arrs = make_scene(pa.SCENE_DAM_BREAK, n_target=8000, dx=0.005)
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        ctx.step(1)          # density and EOS of the uploaded positions, then the move
        rho_ref, p_ref = ctx.download("rho"), ctx.download("p")
    dp = d.copy()
    dp.model = pa.MODEL_PAIRS
    with pa.Context(dp) as ctx:
        upload_all(ctx, arrs)
        ctx.upload("mass", np.full(d.n_particles, d.mass))
        ctx.upload("rho", np.zeros(d.n_particles))
        ctx.run_equations([pa.EQ_SPH_DENSITY, pa.EQ_SPH_EOS], pa.LOOP_CELLS, d.cell_size)
        rho, p = ctx.download("rho"), ctx.download("p")
    assert np.allclose(rho, rho_ref, rtol=1e-13, atol=0.0)
    scale = np.max(np.abs(p_ref))
    assert np.max(np.abs(p - p_ref)) <= 1e-10 * scale


def test_generated_momentum_matches_the_force_pass():
    """The built-in momentum body (tuple-returning helper, generated over the cell iterator) against
    the hand-written WCSPH force pass on the same positions, velocities, rho and p: the acceleration
    minus gravity, norm-wise to 1e-13 (the sums run in another order and association)."""
    d, arrs = make_scene(pa.SCENE_DAM_BREAK, n_target=8000, dx=0.005)
    rng = np.random.default_rng(11)
    arrs = dict(arrs)
    nf = int(d.n_moving)
    for k in ("vx", "vy", "vz"):
        v = np.zeros(d.n_particles)
        v[:nf] = rng.uniform(-0.05, 0.05, nf)
        arrs[k] = v
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        ctx.step(1)
        rho, p = ctx.download("rho"), ctx.download("p")
        a_ref = np.stack([ctx.download(k) for k in ("ax", "ay", "az")])
    dp = d.copy()
    dp.model = pa.MODEL_PAIRS
    with pa.Context(dp) as ctx:
        upload_all(ctx, arrs)
        ctx.upload("mass", np.full(d.n_particles, d.mass))
        ctx.upload("rho", rho)
        ctx.upload("p", p)
        ctx.run_equations([pa.EQ_SPH_MOMENTUM], pa.LOOP_CELLS, d.cell_size)
        a = np.stack([ctx.download(k) for k in ("ax", "ay", "az")])
    g = np.array(d.gravity[:])[:, None]
    want = (a_ref - g)[:, :nf]
    err = np.linalg.norm(a[:, :nf] - want) / np.linalg.norm(want)
    print(f"generated momentum vs force pass: {err:.3e}")
    assert err < 1e-13


def test_generated_contact_matches_the_dem_pass():
    """The built-in contact body (six-tuple helper) against the hand-written DEM pass on the first step
    (fresh tangential history), without walls or gravity: force / m and dt torque / I bit for bit."""
    d, arrs = make_scene(pa.SCENE_DEM_POUR, n_target=3000, seed=2)
    d.wall_mask = 0
    d.gravity[2] = 0.0
    rng = np.random.default_rng(5)
    n = int(d.n_particles)
    arrs = dict(arrs)
    # squeeze the spheres so that many pairs touch, and spin them
    for k in ("x", "y", "z"):
        c = arrs[k].mean()
        arrs[k] = c + 0.93 * (arrs[k] - c)
    w = {k: rng.uniform(-10.0, 10.0, n) for k in ("wx", "wy", "wz")}
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        for k, v in w.items():
            ctx.upload(k, v)
        ctx.run_equations([pa.EQ_DEM_CONTACT], pa.LOOP_CELLS, d.cell_size)
        f = np.stack([ctx.download(k) for k in ("fx", "fy", "fz")])
        t = np.stack([ctx.download(k) for k in ("tx", "ty", "tz")])
        ctx.step(1)
        a_ref = np.stack([ctx.download(k) for k in ("ax", "ay", "az")])
        w_ref = np.stack([ctx.download(k) for k in ("wx", "wy", "wz")])
    m, r = arrs["mass"], arrs["radius"]
    assert np.count_nonzero(f[0]) > 100     # many touching pairs
    assert np.array_equal(f / m, a_ref)
    inertia = 0.4 * m * r * r
    w0 = np.stack([w["wx"], w["wy"], w["wz"]])
    assert np.array_equal(w0 + d.dt * (t / inertia), w_ref)
