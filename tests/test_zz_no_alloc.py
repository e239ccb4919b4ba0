"""prestige_eq1_cells_noalloc (allocation-free eq1 over the cell iterator)
against the context entry point prestige_eval_equation(EQ1, LOOP_CELLS):
bit-identical forces on two calls in a row, the second with a smaller input
and another grid, plus the size limits and the domain check.  The test
allocates its device arrays itself (hipMalloc through ctypes); the entry
point allocates nothing."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import prestige_amd as pa
from helpers import make_scene

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
    return _hip


class DeviceArray:
    """n doubles of device memory, filled from (and read back into) numpy."""

    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host, np.float64)
        self.n = host.size
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), C.c_size_t(max(8, host.nbytes))) == 0
        assert hip().hipMemcpy(self.p, host.ctypes.data_as(C.c_void_p), C.c_size_t(host.nbytes), 1) == 0

    def get(self) -> np.ndarray:
        out = np.empty(self.n, np.float64)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = C.c_void_p()


def _inputs(scene, seed, **kw):
    d, arrs = make_scene(scene, seed=seed, **kw)
    rng = np.random.default_rng(seed)
    n = int(d.n_particles)
    force0 = np.round(rng.standard_normal(n) * 64) / 64
    mass = np.round(rng.random(n) * 256 + 1) / 256
    return d, arrs, force0, mass


def _with_context(d, arrs, force0, mass, cutoff):
    e = pa.desc_init(pa.MODEL_PAIRS)
    e.n_particles = d.n_particles
    e.n_moving = d.n_particles
    e.cell_size = d.cell_size
    for a in range(3):
        e.grid_lo[a] = d.grid_lo[a]
        e.grid_hi[a] = d.grid_hi[a]
    with pa.Context(e) as ctx:
        for k in ("x", "y", "z"):
            ctx.upload(k, arrs[k])
        ctx.upload("force", force0)
        ctx.upload("mass", mass)
        ctx.eval_equation(pa.EQ1, pa.LOOP_CELLS, cutoff)
        return ctx.download("force")


def _noalloc(d, arrs, force0, mass, cutoff, n=None):
    n = int(d.n_particles) if n is None else n
    bufs = [DeviceArray(arrs[k]) for k in ("x", "y", "z")] + [DeviceArray(mass), DeviceArray(force0)]
    lo = (C.c_double * 3)(*d.grid_lo)
    hi = (C.c_double * 3)(*d.grid_hi)
    try:
        st = pa.load_library().prestige_eq1_cells_noalloc(bufs[0].p, bufs[1].p, bufs[2].p, bufs[3].p, bufs[4].p,
                                                          n, lo, hi, d.cell_size, cutoff, None)
        return st, bufs[4].get()
    finally:
        for b in bufs:
            b.free()


def test_noalloc_matches_context_on_two_calls():
    """Two calls in a row, the second smaller and on another grid: each bit-identical to the context path."""
    cases = [(pa.SCENE_DEM_POUR, 1, dict(n_target=60000)),
             (pa.SCENE_DAM_BREAK, 2, dict(n_target=3000, dx=0.01))]
    for scene, seed, kw in cases:
        d, arrs, force0, mass = _inputs(scene, seed, **kw)
        print(f"[noalloc] scene {scene} n {d.n_particles}", flush=True)
        cutoff = 0.9 * d.cell_size
        want = _with_context(d, arrs, force0, mass, cutoff)
        st, got = _noalloc(d, arrs, force0, mass, cutoff)
        assert st == 0, pa.load_library().prestige_last_error(None)
        assert np.array_equal(got, want)
        assert np.any(got != force0)


def test_noalloc_limits_and_domain():
    lib = pa.load_library()
    d, arrs, force0, mass = _inputs(pa.SCENE_DEM_POUR, 3, n_target=2000)
    lo = (C.c_double * 3)(*d.grid_lo)
    hi = (C.c_double * 3)(*d.grid_hi)
    null = C.c_void_p()
    # more particles than the scratch holds: refused before any device work
    st = lib.prestige_eq1_cells_noalloc(null, null, null, null, null, (1 << 20) + 1, lo, hi, d.cell_size,
                                        d.cell_size, None)
    assert st == 4 and b"exceeds" in lib.prestige_last_error(None)
    # a grid with more cells than the scratch holds
    st = lib.prestige_eq1_cells_noalloc(null, null, null, null, null, 10, lo, hi, d.cell_size / 20,
                                        d.cell_size / 20, None)
    assert st == 4 and b"cells" in lib.prestige_last_error(None)
    # cutoff above the cell size
    st = lib.prestige_eq1_cells_noalloc(null, null, null, null, null, 10, lo, hi, d.cell_size,
                                        2 * d.cell_size, None)
    assert st == 1
    # a particle outside the grid: its key is clamped in the kernel and the domain flag reported
    arrs = dict(arrs)
    x = np.array(arrs["x"], copy=True)
    x[7] = d.grid_hi[0] + 1.0
    arrs["x"] = x
    st, _ = _noalloc(d, arrs, force0, mass, d.cell_size)
    assert st == 6
    # and the next call on the same scratch is clean again
    arrs["x"][7] = d.grid_lo[0] + 0.5 * d.cell_size
    st, got = _noalloc(d, arrs, force0, mass, d.cell_size)
    assert st == 0
    assert np.array_equal(got, _with_context(d, arrs, force0, mass, d.cell_size))
