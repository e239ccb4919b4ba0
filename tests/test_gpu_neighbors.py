"""Neighbour search on the device vs the FP64 oracle: keys, stable-sort
permutation, slot ids and the cell lower-bound table bit for bit, and the eq1
probe (prestige/src/lib.rs:10) over all pairs and over the cell iterator."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import prestige_amd as pa
from helpers import make_scene, oracle_state, upload_all

pytestmark = pytest.mark.gpu


def _compare_build(ctx, st):
    g = ctx.neighbors_download()
    o = st.neighbors()
    for k in ("keys", "perm", "ids", "lb"):
        assert g[k].shape == o[k].shape, k
        bad = np.flatnonzero(g[k] != o[k])
        assert bad.size == 0, f"{k}: {bad.size} mismatches, first at {bad[:5]}"
    return g


@pytest.mark.parametrize("scene,kw", [
    (pa.SCENE_DAM_BREAK, dict(n_target=8000, dx=0.005)),
    (pa.SCENE_DEM_LATTICE, dict(n_target=10000, seed=3)),
    (pa.SCENE_DEM_POUR, dict(n_target=20000, seed=5)),
])
def test_build_bitexact(scene, kw):
    d, arrs = make_scene(scene, **kw)
    st = oracle_state(d, arrs)
    assert st.build() == 0
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        ctx.neighbors_build()
        g = _compare_build(ctx, st)
        # invariants of the table
        assert np.all(np.diff(g["keys"].astype(np.int64)) >= 0)
        assert g["lb"][0] == 0 and g["lb"][-1] == d.n_particles
        assert np.all(np.diff(g["lb"].astype(np.int64)) >= 0)
        assert np.array_equal(np.sort(g["ids"]), np.arange(d.n_particles, dtype=np.uint32))


def test_rebuild_after_motion_bitexact():
    """Second build starts from the previous sorted order (nearly sorted input)."""
    d, arrs = make_scene(pa.SCENE_DEM_LATTICE, n_target=10000, seed=11)
    st = oracle_state(d, arrs)
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        for _ in range(3):
            assert st.step(25) == 0
            ctx.step(25)
            assert st.build() == 0
            ctx.neighbors_build()
            _compare_build(ctx, st)


@pytest.mark.parametrize("policy,speed,n", [
    ("incremental", 10.0, 20000),    # about a thousand moved slots: bitonic sort in LDS
    ("incremental", 20.0, 100000),   # several thousand: one-block radix sort
    ("incremental", 40.0, 150000),   # more than the LDS holds
    ("radix", 40.0, 150000),
])
def test_sort_policies_over_steps(monkeypatch, policy, speed, n):
    """Each build starts from the previous step's order.  The
    incremental sort (slots that kept their cell stay in order, moved ones are
    sorted and merged) and the full radix sort must both give the oracle's
    stable sort, however many slots moved."""
    monkeypatch.setenv("PRESTIGE_SORT", policy)
    d, arrs = make_scene(pa.SCENE_DEM_POUR, n_target=n, seed=7)
    rng = np.random.default_rng(1)
    for k in ("vx", "vy", "vz"):
        arrs[k] = rng.uniform(-speed, speed, d.n_particles)
    d.dt = 1e-4
    st = oracle_state(d, arrs)
    moved = []
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        prev = None
        for _ in range(4):
            assert st.step(1) == 0
            # spheres this fast outrun the DEM step (dt |v| > r / 4): the step
            # still runs and reports E_CFL; the sort must be exact regardless
            try:
                ctx.step(1)
            except pa.PrestigeError as e:
                assert e.name == "E_CFL" and 3 ** 0.5 * speed * d.dt > 0.25 * arrs["radius"].min(), str(e)
            # rebuild from the step's order with the keys the step computed
            assert st.build() == 0
            ctx.neighbors_build()
            g = _compare_build(ctx, st)
            if prev is not None:
                moved.append(int(np.sum(g["keys"][np.argsort(g["perm"])] != prev)))
            prev = g["keys"].copy()
    print(f"{policy}: slots that changed cell per step {moved}")
    if speed > 20.0:
        assert min(moved) > 16384


@pytest.mark.parametrize("n", [1, 7, 63, 64, 65, 200])
def test_small_and_ragged_counts_over_steps(n):
    """Counts below, at and just past one wavefront: the step kernels' moved
    mask (one word per wavefront, partial last word), the incremental sort
    and the state stay identical to the oracle."""
    d, arrs = make_scene(pa.SCENE_DEM_LATTICE, n_target=n, seed=4)
    rng = np.random.default_rng(n)
    for k in ("vx", "vy", "vz"):
        arrs[k] = rng.uniform(-2.0, 2.0, d.n_particles)
    d.dt = 1e-4
    st = oracle_state(d, arrs)
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        for _ in range(3):
            assert st.step(10) == 0
            ctx.step(10)
            assert st.build() == 0
            ctx.neighbors_build()
            _compare_build(ctx, st)
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            assert np.array_equal(ctx.download(k), st.get(k)), k


def test_eq1_allpairs_bitexact():
    rng = np.random.default_rng(0)
    d, arrs = make_scene(pa.SCENE_DEM_LATTICE, n_target=3000, seed=1)
    force0 = rng.standard_normal(d.n_particles)
    mass = rng.random(d.n_particles)
    expect = oracle.eq1_allpairs(force0, mass)
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        ctx.upload("force", force0)
        ctx.upload("mass", mass)
        ctx.eval_equation(pa.EQ1, pa.LOOP_ALLPAIRS)
        got = ctx.download("force")
    assert np.array_equal(got, expect)
    # closed form for dyadic masses: force0 + sum(mass)
    dy = np.round(mass * 1024) / 1024
    assert np.array_equal(oracle.eq1_allpairs(np.zeros_like(dy), dy), np.full_like(dy, dy.sum()))


@pytest.mark.parametrize("n", [1, 15, 16, 17, 32, 33, 100])
def test_eq1_allpairs_small_counts_bitexact(n):
    """Counts around the kernel's 16-mass batches: below one batch, exactly one or two, and a partial
    last batch after the double-buffered loop."""
    rng = np.random.default_rng(n)
    d, arrs = make_scene(pa.SCENE_DEM_LATTICE, n_target=n, seed=1)
    assert d.n_particles == n
    force0 = rng.standard_normal(n)
    mass = rng.random(n)
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        ctx.upload("force", force0)
        ctx.upload("mass", mass)
        ctx.eval_equation(pa.EQ1, pa.LOOP_ALLPAIRS)
        got = ctx.download("force")
    assert np.array_equal(got, oracle.eq1_allpairs(force0, mass))


@pytest.mark.parametrize("cut_frac", [1.0, 0.5])
def test_eq1_cells_bitexact(cut_frac):
    rng = np.random.default_rng(1)
    d, arrs = make_scene(pa.SCENE_DAM_BREAK, n_target=8000, dx=0.005)
    cutoff = cut_frac * d.cell_size
    force0 = rng.standard_normal(d.n_particles)
    mass = rng.random(d.n_particles)
    st = oracle_state(d, arrs)
    assert st.build() == 0
    expect = st.eq1_cells(cutoff, force0, mass)
    counts = st.eq1_cells(cutoff, np.zeros(d.n_particles), np.ones(d.n_particles))
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        ctx.neighbors_build()
        ctx.upload("force", force0)
        ctx.upload("mass", mass)
        ctx.eval_equation(pa.EQ1, pa.LOOP_CELLS, cutoff)
        got = ctx.download("force")
        ctx.upload("force", np.zeros(d.n_particles))
        ctx.upload("mass", np.ones(d.n_particles))
        ctx.eval_equation(pa.EQ1, pa.LOOP_CELLS, cutoff)
        got_counts = ctx.download("force")
    assert np.array_equal(got, expect)
    assert np.array_equal(got_counts, counts)
    if cut_frac == 1.0:
        side = 20
        inner = np.arange(side**3).reshape(side, side, side)[5:15, 5:15, 5:15].ravel()
        assert np.all(got_counts[inner] == 57.0)


@pytest.mark.slow
def test_full_size_dam_break_build():
    """1,000,000 fluid + 799,320 tank particles: bit-exact build and 57 interior neighbours."""
    d, arrs = make_scene(pa.SCENE_DAM_BREAK, n_target=1_000_000, dx=0.005)
    assert d.n_particles == 1_799_320
    st = oracle_state(d, arrs)
    assert st.build() == 0
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        ctx.neighbors_build()
        g = _compare_build(ctx, st)
        ctx.upload("force", np.zeros(d.n_particles))
        ctx.upload("mass", np.ones(d.n_particles))
        ctx.eval_equation(pa.EQ1, pa.LOOP_CELLS, d.cell_size)
        counts = ctx.download("force")
    side = 100
    inner = np.arange(side**3).reshape(side, side, side)[10:90, 10:90, 10:90].ravel()
    assert np.all(counts[inner] == 57.0)
    assert int((np.diff(g["lb"].astype(np.int64))).sum()) == d.n_particles
