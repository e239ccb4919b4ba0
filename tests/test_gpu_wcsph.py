"""WCSPH on the device vs the FP64 oracle.

FP64 form: bit-identical fields (same operation order, no FMA contraction,
same neighbour visiting order).  FP32 and mixed forms: norm-wise relative
error per field, ||f - f_ref||_2 / ||f_ref||_2 over fluid particles, with
the tolerances written in each test."""
from __future__ import annotations

import numpy as np
import pytest

import prestige_amd as pa
from helpers import make_scene, oracle_state, rel_l2, upload_all

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "rho", "p", "ax", "ay", "az")


def _run_pair(d, arrs, nsteps):
    st = oracle_state(d, arrs)
    assert st.step(nsteps) == 0
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        ctx.step(nsteps)
        got = {k: ctx.download(k) for k in FIELDS}
    ref = {k: st.get(k) for k in FIELDS}
    return got, ref


@pytest.mark.parametrize("nsteps", [1, 30])
def test_f64_bitexact(nsteps):
    d, arrs = make_scene(pa.SCENE_DAM_BREAK, n_target=8000, dx=0.005)
    got, ref = _run_pair(d, arrs, nsteps)
    for k in FIELDS:
        bad = np.flatnonzero(got[k] != ref[k])
        assert bad.size == 0, f"{k}: {bad.size} values differ, e.g. {got[k][bad[:3]]} vs {ref[k][bad[:3]]}"


# The plain FP32 form after 50 steps (no error bound is specified for it; these tolerances only catch
# a broken FP32 path).
@pytest.mark.parametrize("precision,tol", [(pa.F32, {"x": 1e-6, "v": 0.5, "rho": 1e-5, "p": 0.1})])
def test_reduced_precision_error(precision, tol):
    d, arrs = make_scene(pa.SCENE_DAM_BREAK, n_target=8000, dx=0.005)
    st = oracle_state(d, arrs)
    assert st.step(50) == 0
    dm = d.copy()
    dm.precision = precision
    with pa.Context(dm) as ctx:
        upload_all(ctx, arrs)
        ctx.step(50)
        got = {k: ctx.download(k) for k in FIELDS}
    nf = d.n_moving
    ex = rel_l2(np.stack([got[k][:nf] for k in "xyz"]), np.stack([st.get(k)[:nf] for k in "xyz"]))
    ev = rel_l2(np.stack([got[k][:nf] for k in ("vx", "vy", "vz")]),
                np.stack([st.get(k)[:nf] for k in ("vx", "vy", "vz")]))
    er = rel_l2(got["rho"][:nf], st.get("rho")[:nf])
    ep = rel_l2(got["p"][:nf], st.get("p")[:nf])
    print(f"precision {precision}: x {ex:.3e} v {ev:.3e} rho {er:.3e} p {ep:.3e}")
    assert ex < tol["x"] and ev < tol["v"] and er < tol["rho"] and ep < tol["p"]


def test_mixed_with_fp64_density():
    """PRESTIGE_FLAG_DENSITY_F64: the first step's density and pressure are the F64 form's bit for bit
    (same positions, same operations); later steps differ only through the FP32 force."""
    d, arrs = make_scene(pa.SCENE_DAM_BREAK, n_target=8000, dx=0.005)
    dm = d.copy()
    dm.precision = pa.MIXED
    dm.flags = pa.FLAG_DENSITY_F64
    out = {}
    for name, dd in (("f64", d), ("mixed_rho64", dm)):
        with pa.Context(dd) as ctx:
            upload_all(ctx, arrs)
            ctx.step(1)
            out[name] = {k: ctx.download(k) for k in ("rho", "p", "vx")}
            ctx.step(19)
            out[name]["v20"] = np.stack([ctx.download(k) for k in ("vx", "vy", "vz")])
            out[name]["rho20"] = ctx.download("rho")
    assert np.array_equal(out["mixed_rho64"]["rho"], out["f64"]["rho"])
    assert np.array_equal(out["mixed_rho64"]["p"], out["f64"]["p"])
    nf = d.n_moving
    ev = rel_l2(out["mixed_rho64"]["v20"][:, :nf], out["f64"]["v20"][:, :nf])
    er = rel_l2(out["mixed_rho64"]["rho20"][:nf], out["f64"]["rho20"][:nf])
    print(f"mixed with FP64 density after 20 steps: v {ev:.3e} rho {er:.3e}")
    assert ev < 1e-4 and er < 1e-7


def test_domain_error_reported():
    d, arrs = make_scene(pa.SCENE_DAM_BREAK, n_target=8000, dx=0.005)
    x = arrs["x"].copy()
    x[0] = d.grid_hi[0] + 1.0
    with pa.Context(d) as ctx:
        upload_all(ctx, dict(arrs, x=x))
        with pytest.raises(pa.PrestigeError) as ei:
            ctx.step(1)
        assert ei.value.name == "E_DOMAIN"


@pytest.mark.slow
def test_full_size_f64_bitexact_five_steps():
    """The 1.8M-particle workload itself: five steps take the radix sort, the
    incremental sort and, from the third density pass on, the dry-wall reuse
    (about 700k wall particles)."""
    d, arrs = make_scene(pa.SCENE_DAM_BREAK, n_target=1_000_000, dx=0.005)
    got, ref = _run_pair(d, arrs, 5)
    for k in ("x", "vz", "rho", "p", "az"):
        assert np.array_equal(got[k], ref[k]), k


def test_f64_bitexact_far_from_origin():
    """The scan works on FP32 offsets, whose rounding grows with the coordinate
    magnitude; its margin grows with it.  A dam break moved ~1 km from the
    origin must keep exactly the oracle's neighbour sets."""
    d, arrs = make_scene(pa.SCENE_DAM_BREAK, n_target=8000, dx=0.005)
    shift = (1000.0, -500.0, 250.0)
    for a, s in enumerate(shift):
        d.grid_lo[a] += s
        d.grid_hi[a] += s
    arrs = dict(arrs, x=arrs["x"] + shift[0], y=arrs["y"] + shift[1], z=arrs["z"] + shift[2])
    got, ref = _run_pair(d, arrs, 5)
    for k in FIELDS:
        assert np.array_equal(got[k], ref[k]), k


def _layered_scene():
    """A dense slab two cells thick (lattice 3 % tighter than the rest
    spacing, so part of it is under pressure) under a sparse layer (one
    particle per four cells): the wavefronts over the sparse layer span many
    rows, so the union of their row runs below exceeds the LDS stage and those
    runs are scanned from global memory."""
    d = pa.desc_init(pa.MODEL_WCSPH)
    h = 0.012
    dx = h / 1.2
    cs = 2.0 * h
    ncell = (12, 12, 6)
    d.h, d.cell_size, d.rho0, d.c0, d.alpha = h, cs, 1000.0, 20.0, 0.1
    d.mass = d.rho0 * dx**3
    d.dt = 0.25 * h / d.c0
    d.gravity[:] = (0.0, 0.0, -9.81)
    for a in range(3):
        d.grid_lo[a] = 0.0
        d.grid_hi[a] = ncell[a] * cs
    g = np.arange(0.5 * dx, ncell[0] * cs, 0.97 * dx)
    zz = np.arange(cs + 0.5 * dx, 3.0 * cs, 0.97 * dx)
    X, Y, Z = np.meshgrid(g, g, zz, indexing="ij")
    dense = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    cx, cy = np.meshgrid(np.arange(0, ncell[0], 2), np.arange(0, ncell[1], 2), indexing="ij")
    rng = np.random.default_rng(3)
    sparse = np.stack([(cx.ravel() + rng.uniform(0.1, 0.9, cx.size)) * cs,
                       (cy.ravel() + rng.uniform(0.1, 0.9, cx.size)) * cs,
                       np.full(cx.size, 3.4 * cs)], 1)
    pts = np.concatenate([dense, sparse])
    d.n_particles = d.n_moving = len(pts)
    arrs = {"x": pts[:, 0].copy(), "y": pts[:, 1].copy(), "z": pts[:, 2].copy(),
            "vx": np.zeros(len(pts)), "vy": np.zeros(len(pts)), "vz": np.zeros(len(pts))}
    return d, arrs


def test_f64_bitexact_unstaged_runs():
    d, arrs = _layered_scene()
    assert d.n_particles > 4 * 256
    got, ref = _run_pair(d, arrs, 3)
    for k in FIELDS:
        assert np.array_equal(got[k], ref[k]), k


def test_f64_bitexact_dry_walls_over_steps(monkeypatch):
    """Wall particles with no moving particle in their 27 cells at two passes in
    a row keep their density (the wavefronts made only of them skip the scan);
    the 8000-particle column collapses over the tank floor for 150 steps, wetting
    walls that were dry.  FP64 fields stay bit-identical to the oracle, and to a
    run with the reuse switched off, at 60 and 150 steps."""
    d, arrs = make_scene(pa.SCENE_DAM_BREAK, n_target=8000, dx=0.005)
    st = oracle_state(d, arrs)
    runs = {}
    for mode in ("on", "off"):
        monkeypatch.setenv("PRESTIGE_DRY", mode)
        with pa.Context(d) as ctx:
            upload_all(ctx, arrs)
            out = []
            for n in (60, 90):
                ctx.step(n)
                out.append({k: ctx.download(k) for k in FIELDS})
            runs[mode] = out
    for idx, n in enumerate((60, 90)):
        assert st.step(n) == 0
        ref = {k: st.get(k) for k in FIELDS}
        for k in FIELDS:
            assert np.array_equal(runs["on"][idx][k], ref[k]), (idx, k)
            assert np.array_equal(runs["off"][idx][k], ref[k]), (idx, k)


@pytest.mark.parametrize("precision", [pa.MIXED, pa.F32])
def test_dry_wall_reuse_is_exact_in_reduced_precision(monkeypatch, precision):
    """The FP32 and mixed forms keep the outputs of dry wall particles by
    particle id: 150 steps of the collapsing column give the same bits with
    the reuse on and off."""
    d, arrs = make_scene(pa.SCENE_DAM_BREAK, n_target=8000, dx=0.005)
    d.precision = precision
    runs = {}
    for mode in ("on", "off"):
        monkeypatch.setenv("PRESTIGE_DRY", mode)
        with pa.Context(d) as ctx:
            upload_all(ctx, arrs)
            ctx.step(150)
            runs[mode] = {k: ctx.download(k) for k in FIELDS}
    for k in FIELDS:
        assert np.array_equal(runs["on"][k], runs["off"][k]), k


def test_two_contexts_stepped_in_turn():
    """Per-context device state (tile-queue counters, dry-wall passes, the
    incremental sort's moved counts) stays per context: two dam breaks of
    different sizes stepped alternately match the oracle bit for bit."""
    scenes = [make_scene(pa.SCENE_DAM_BREAK, n_target=n, dx=0.005) for n in (8000, 27000)]
    refs = []
    for d, arrs in scenes:
        st = oracle_state(d, arrs)
        assert st.step(40) == 0
        refs.append({k: st.get(k) for k in FIELDS})
    ctxs = [pa.Context(d) for d, _ in scenes]
    try:
        for ctx, (_, arrs) in zip(ctxs, scenes):
            upload_all(ctx, arrs)
        for _ in range(10):
            for ctx in ctxs:
                ctx.step(4)
        for ctx, ref in zip(ctxs, refs):
            for k in FIELDS:
                assert np.array_equal(ctx.download(k), ref[k]), k
    finally:
        for ctx in ctxs:
            ctx.close()


# ---------------------------------------------------------------------------
# Mixed forms after 1000 steps against the FP64 oracle: the specified bound is a relative L2 error of
# at most 1e-5 per field after 1000 steps.  Measured on MI355X for this 8,000-fluid dam break (42,866
# particles with the tank; GPU results are deterministic, so these figures repeat exactly):
#
#   form          rho      p        v        x
#   mixed         3.1e-6   1.9e-4   1.9e-5   6.5e-7
#   mixed_rho64   2.4e-7   1.4e-5   1.3e-6   5.1e-8      (PRESTIGE_FLAG_DENSITY_F64)
#
# A field that meets the bound is tested against the bound itself.  A field that does not is tested
# against its measured limit with 3x headroom (so a lost Kahan carry, an FP32 EOS or a dropped
# compensation still fails), and test_mixed_meets_the_bound marks the gap as a strict expected failure:
# p = B((rho/rho0)^7 - 1) turns a relative density error e into about 7 e / ((rho/rho0)^7 - 1), some
# 500-1000 e at these compressions, so p cannot meet 1e-5 unless rho is near 1e-8; v of the plain mixed
# form carries the FP32 density pair sums through that pressure into the momentum equation.
MIXED_BOUND = 1e-5
MIXED_STEPS = 1000
MIXED_FORMS = {"mixed": 0, "mixed_rho64": pa.FLAG_DENSITY_F64}
MIXED_LIMITS = {
    ("mixed", "rho"): MIXED_BOUND, ("mixed", "x"): MIXED_BOUND,
    ("mixed", "v"): 5.6e-5, ("mixed", "p"): 5.6e-4,
    ("mixed_rho64", "rho"): MIXED_BOUND, ("mixed_rho64", "v"): MIXED_BOUND, ("mixed_rho64", "x"): MIXED_BOUND,
    ("mixed_rho64", "p"): 4.1e-5,
}


@pytest.fixture(scope="module")
def mixed_errors():
    import oracle
    oracle.set_threads(16)
    d, arrs = make_scene(pa.SCENE_DAM_BREAK, n_target=8000, dx=0.005)
    nf = int(d.n_moving)
    st = oracle_state(d, arrs)
    assert st.step(MIXED_STEPS) == 0

    def fields(get):
        return {"x": np.stack([get(k)[:nf] for k in "xyz"]),
                "v": np.stack([get(k)[:nf] for k in ("vx", "vy", "vz")]),
                "rho": np.array(get("rho")[:nf]), "p": np.array(get("p")[:nf])}

    ref = fields(st.get)
    out = {}
    for name, flags in MIXED_FORMS.items():
        dm = d.copy()
        dm.precision = pa.MIXED
        dm.flags = flags
        with pa.Context(dm) as ctx:
            upload_all(ctx, arrs)
            ctx.step(MIXED_STEPS)
            got = fields(ctx.download)
        out[name] = {k: rel_l2(got[k], ref[k]) for k in ref}
        print(name, {k: f"{v:.2e}" for k, v in out[name].items()})
    return out


@pytest.mark.parametrize("form,field", sorted(MIXED_LIMITS))
def test_mixed_error_after_1000_steps(mixed_errors, form, field):
    assert mixed_errors[form][field] <= MIXED_LIMITS[(form, field)], (form, field, mixed_errors[form][field])


@pytest.mark.parametrize("form,field", sorted(k for k, v in MIXED_LIMITS.items() if v > MIXED_BOUND))
@pytest.mark.xfail(strict=True, reason="measured limit above the 1e-5 bound (see MIXED_LIMITS)")
def test_mixed_meets_the_bound(mixed_errors, form, field):
    assert mixed_errors[form][field] <= MIXED_BOUND
