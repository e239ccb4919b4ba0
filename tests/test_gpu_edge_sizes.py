"""Edge sizes of the path: no particles and one particle.

The reference's loop over i in 0..n and j in 0..n (prestige/src/codegen/
simple_cpu.rs:7-8) does nothing for n = 0; here every entry point accepts an
empty context, steps it and returns empty fields.  One WCSPH particle has the
closed form of a lone particle: its only pair is itself (q = 0, W = 1), so
rho = m sigma, p = B ((rho / rho0)^7 - 1), no pair force, and the step is free
fall: v1 = v0 + dt g, x1 = x0 + dt v1.
"""
import math

import numpy as np
import pytest

import prestige_amd as pa
from helpers import STATE_FIELDS, make_scene, oracle_state, upload_all

pytestmark = pytest.mark.gpu

CASES = [(pa.SCENE_DAM_BREAK, pa.F64), (pa.SCENE_DAM_BREAK, pa.MIXED), (pa.SCENE_DAM_BREAK, pa.F32),
         (pa.SCENE_DEM_LATTICE, pa.F64), (pa.SCENE_DEM_LATTICE, pa.MIXED)]


@pytest.mark.parametrize("scene,prec", CASES)
def test_empty_context_steps_and_returns_empty_fields(scene, prec):
    d, _ = make_scene(scene, n_target=1000, dx=0.01, precision=prec)
    d.n_particles = 0
    d.n_moving = 0
    empty = np.zeros(0)
    print(f"empty context: scene {scene} precision {prec}", flush=True)
    with pa.Context(d) as ctx:
        for k in STATE_FIELDS:
            ctx.upload(k, empty)
        if scene == pa.SCENE_DEM_LATTICE:
            ctx.upload("radius", empty)
            ctx.upload("mass", empty)
        ctx.step(20)    # past one read of the device flags (every 16 steps)
        for k in STATE_FIELDS:
            assert ctx.download(k).shape == (0,), k
        ctx.neighbors_build()
        nb = ctx.neighbors_download()
        for k in ("keys", "perm", "ids"):
            assert nb[k].size == 0, k
        # every cell's first slot is the end of the (empty) slot range
        assert nb["lb"].size == int(np.prod(ctx.grid_dims())) + 1
        assert not nb["lb"].any()
        ctx.upload("mass", empty)
        ctx.upload("force", empty)
        for loop in (pa.LOOP_ALLPAIRS, pa.LOOP_CELLS):
            ctx.eval_equation(pa.EQ1, loop, d.cell_size)
            assert ctx.download("force").shape == (0,)


def _lone_particle(prec):
    d, arrs = make_scene(pa.SCENE_DAM_BREAK, n_target=1000, dx=0.01, precision=prec)
    h, rho0 = d.h, d.rho0
    d.mass = 2.0 * rho0 * math.pi * h ** 3    # rho = m sigma = 2 rho0: p > 0
    d.n_particles = 1
    d.n_moving = 1
    lo = np.array(list(d.grid_lo))
    hi = np.array(list(d.grid_hi))
    mid = 0.5 * (lo + hi)
    one = {k: np.array([v]) for k, v in zip(STATE_FIELDS, (mid[0], mid[1], mid[2], 0.3, -0.2, 0.1))}
    return d, one


@pytest.mark.parametrize("prec", [pa.F64, pa.MIXED, pa.F32])
def test_lone_wcsph_particle_closed_form(prec):
    d, one = _lone_particle(prec)
    h = d.h
    sigma = 1.0 / (((math.pi * h) * h) * h)
    rho = d.mass * sigma
    B = ((d.rho0 * d.c0) * d.c0) / 7.0
    p = B * ((rho / d.rho0) ** 7 - 1.0)
    g = np.array(list(d.gravity))
    v0 = np.array([one[k][0] for k in ("vx", "vy", "vz")])
    x0 = np.array([one[k][0] for k in ("x", "y", "z")])
    v1 = v0 + d.dt * g
    x1 = x0 + d.dt * v1
    # FP64: rounding of a few operations; FP32 pair arithmetic and FP32
    # density outputs: one FP32 rounding of rho and of p
    tol = 1e-14 if prec == pa.F64 else 2e-6
    print(f"lone particle: precision {prec}", flush=True)
    with pa.Context(d) as ctx:
        upload_all(ctx, one)
        ctx.step(1)
        assert ctx.download("rho")[0] == pytest.approx(rho, rel=tol)
        assert ctx.download("p")[0] == pytest.approx(p, rel=tol)
        for a, k in enumerate(("vx", "vy", "vz")):
            assert ctx.download(k)[0] == pytest.approx(v1[a], rel=tol, abs=1e-300), k
        # positions: x1 - x0 = dt v1 is resolved to the state's own precision
        xt = 1e-15 if prec != pa.F32 else 3e-7
        for a, k in enumerate(("x", "y", "z")):
            assert ctx.download(k)[0] == pytest.approx(x1[a], rel=xt), k


def test_lone_wcsph_particle_matches_oracle_bit_for_bit():
    d, one = _lone_particle(pa.F64)
    st = oracle_state(d, one)
    with pa.Context(d) as ctx:
        upload_all(ctx, one)
        for _ in range(3):
            assert st.step(5) == 0
            ctx.step(5)
            for k in STATE_FIELDS:
                assert np.array_equal(ctx.download(k), st.get(k)), k
