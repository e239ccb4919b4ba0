"""Failure detection on the device, each condition forced from the host on
valid inputs so that every kernel still runs to its normal end: the kernels'
own bounds and value checks raise the flags, prestige_step reads them every
16 steps and reports NONFINITE before CFL before CAPACITY before DOMAIN."""
from __future__ import annotations

import math

import numpy as np
import pytest

import prestige_amd as pa
from helpers import make_scene, upload_all

pytestmark = pytest.mark.gpu


def _dam(n=8000, precision=pa.F64):
    return make_scene(pa.SCENE_DAM_BREAK, n_target=n, dx=0.005, precision=precision)


def _expect(ctx, nsteps, name):
    with pytest.raises(pa.PrestigeError) as ei:
        ctx.step(nsteps)
    assert ei.value.name == name, str(ei.value)
    return str(ei.value)


@pytest.mark.parametrize("precision", [pa.F64, pa.MIXED])
def test_nonfinite_acceleration(precision):
    """A finite but huge velocity (1e308 m/s) overflows the pair terms to inf."""
    d, arrs = _dam(precision=precision)
    vx = arrs["vx"].copy()
    vx[4000] = 1e308
    with pa.Context(d) as ctx:
        upload_all(ctx, dict(arrs, vx=vx))
        _expect(ctx, 3, "E_NONFINITE")


def test_cfl_runaway_particle_ends_the_call_within_16_steps():
    """One particle at 1.5 c0: the CFL number dt (c0 + |v|) / h exceeds 1/2 at the first step; the call
    returns after step 16 of 100 with the state of 16 steps, bit for bit."""
    d, arrs = _dam()
    vz = arrs["vz"].copy()
    vz[4000] = 1.5 * d.c0
    with pa.Context(d) as ctx:
        upload_all(ctx, dict(arrs, vz=vz))
        msg = _expect(ctx, 100, "E_CFL")
        assert "steps 1-16 of 100" in msg
        got = ctx.download("x")
    with pa.Context(d) as ref:
        upload_all(ref, dict(arrs, vz=vz))
        with pytest.raises(pa.PrestigeError):
            ref.step(16)
        assert np.array_equal(got, ref.download("x"))


def test_cfl_time_step_too_large_at_rest():
    d, arrs = _dam()
    d.dt = 0.6 * d.h / d.c0
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        _expect(ctx, 1, "E_CFL")


def test_cfl_dem_sphere_outruns_step():
    """DEM: a sphere moving more than a quarter radius per step."""
    d, arrs = make_scene(pa.SCENE_DEM_LATTICE, n_target=1000, seed=0)
    vx = arrs["vx"].copy()
    vx[10] = 0.3 * arrs["radius"][10] / d.dt
    with pa.Context(d) as ctx:
        upload_all(ctx, dict(arrs, vx=vx))
        _expect(ctx, 1, "E_CFL")


def test_wcsph_row_capacity():
    """More than 128 particles in one row of three cells: a 12^3 block at spacing h / 3 (about 216
    particles per cell), light enough that p = 0 and nothing moves fast."""
    d = pa.desc_init(pa.MODEL_WCSPH)
    h = 0.01
    g = (np.arange(12) + 0.5) * (h / 3)
    p = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) + 0.1
    n = len(p)
    d.n_particles = d.n_moving = n
    d.h = h
    d.cell_size = 2 * h
    d.mass = 1e-9
    d.c0 = 10.0
    d.dt = 0.25 * h / d.c0
    for a in range(3):
        d.grid_lo[a] = 0.0
        d.grid_hi[a] = 0.3
    z = np.zeros(n)
    with pa.Context(d) as ctx:
        for k, v in zip(("x", "y", "z", "vx", "vy", "vz"), (p[:, 0], p[:, 1], p[:, 2], z, z, z)):
            ctx.upload(k, v)
        _expect(ctx, 1, "E_CAPACITY")


def test_dem_contact_capacity():
    """max_contacts = 2 on a lattice of touching spheres (6 contacts each)."""
    d = pa.desc_init(pa.MODEL_DEM)
    r = 0.01
    g = 0.1 + np.arange(6) * (2 * r * 0.98)
    p = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    n = len(p)
    d.n_particles = d.n_moving = n
    d.max_contacts = 2
    d.cell_size = 2 * r
    d.dt = 1e-5
    d.gravity[2] = 0.0
    for a in range(3):
        d.grid_lo[a] = 0.0
        d.grid_hi[a] = 0.5
    z = np.zeros(n)
    m = np.full(n, 2500.0 * 4.0 / 3.0 * math.pi * r**3)
    with pa.Context(d) as ctx:
        for k, v in zip(("x", "y", "z", "vx", "vy", "vz"), (p[:, 0], p[:, 1], p[:, 2], z, z, z)):
            ctx.upload(k, v)
        ctx.upload("radius", np.full(n, r))
        ctx.upload("mass", m)
        _expect(ctx, 1, "E_CAPACITY")


def test_flags_clear_after_an_error():
    """After an error the flags start from zero again: a fresh upload steps cleanly."""
    d, arrs = _dam()
    vz = arrs["vz"].copy()
    vz[4000] = 1.5 * d.c0
    with pa.Context(d) as ctx:
        upload_all(ctx, dict(arrs, vz=vz))
        _expect(ctx, 2, "E_CFL")
        upload_all(ctx, arrs)
        ctx.step(2)
