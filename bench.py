#!/usr/bin/env python3
"""Throughput of the particle hot path on one MI355X.

A step is one full WCSPH step of the dam-break workload (neighbour rebuild:
cell keys, stable sort by cell key with the state moved into key order, cell
table; density + Tait EOS gather; momentum + viscosity gather fused with the
symplectic-Euler update).  The headline workload is the largest single-GPU
configuration: 4,019,679 fluid particles (dx = 3.15 mm, a 159^3 block) plus
the 2,006,931 particles of a 3-layer tank, in FP64.  `value` is
particle-updates/s = fluid particles x steps / wall time of the timed steps,
inputs resident in HBM.  Also reported, under "forms": the 1M dam break in
FP64, FP32 and mixed precision, the 4M mixed form with its error against
FP64 after 1, 10, 100 and 1000 steps, the 10,000-sphere DEM lattice in FP64,
and the forward + adjoint cost of the two differentiable workloads.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision f64|f32|mixed] [--workload 4M|1M]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

PEAK_HBM_GBS = 8000.0
PREC = {"f64": 0, "f32": 1, "mixed": 2}
DTYPE = {"f64": "f64", "f32": "f32", "mixed": "f32 pair arithmetic + f64 state/integrator"}
SCENE_1M = dict(n_target=1_000_000, dx=0.005, seed=0)
SCENE_4M = dict(n_target=4_000_000, dx=0.00315, seed=0)
STATE = ("x", "y", "z", "vx", "vy", "vz")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _context(pa, scene_id, scene_args, precision):
    d, arrs = pa.scene(scene_id, **scene_args)
    d.precision = PREC[precision]
    ctx = pa.Context(d)
    for k in STATE + ("radius", "mass"):
        if k in arrs:
            ctx.upload(k, arrs[k])
    return d, arrs, ctx


def run_form(pa, scene_args, precision, steps, warmup, scene_id=None):
    """Throughput over `steps` timed steps with nothing recorded between the
    kernels (a HIP event pair around every launch stretches the stream by
    about 0.15 ms per step).  Then the same steps again from the initial
    state, with HIP events around every kernel on the library's stream: each
    kernel's average duration over the same stretch of the simulation."""
    scene_id = pa.SCENE_DAM_BREAK if scene_id is None else scene_id
    d, arrs, ctx = _context(pa, scene_id, scene_args, precision)
    dem = scene_id != pa.SCENE_DAM_BREAK
    zeros = np.zeros(d.n_particles)

    def reset():
        for k in STATE:
            ctx.upload(k, arrs[k])
        if dem:
            for k in ("wx", "wy", "wz"):
                ctx.upload(k, zeros)

    with ctx:
        ctx.step(warmup)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step(steps)
        ctx.sync()
        t1 = time.perf_counter()
        reset()
        ctx.step(warmup)
        ctx.sync()
        ctx.timing_enable(True)   # per-kernel HIP events on the library's stream
        ctx.step(steps)
        ctx.sync()
        kt = ctx.timing()
        ctx.timing_enable(False)
    sec = t1 - t0
    nf = int(d.n_moving)
    return {"value": nf * steps / sec, "ms_per_step": 1e3 * sec / steps, "n_fluid": nf,
            "n_total": int(d.n_particles), "kernels": kt}


def _profile_table(name, tag):
    """`tag: k=v, ...` line of a committed rocprofv3 summary under profiles/ (empty when absent)."""
    path = ROOT / "profiles" / name
    if not path.exists():
        return {}
    for line in path.read_text().splitlines():
        if line.startswith(tag + ":"):
            pairs = [p.strip().split("=") for p in line.split(":", 1)[1].split(",") if "=" in p]
            return {k: float(v) for k, v in pairs}
    return {}


# committed counter summaries per workload: HBM bytes per launch (FETCH_SIZE / WRITE_SIZE runs) and the
# VALU issue share of each gather (SQ counter runs)
PROFILES = {"4M": ("pmc_traffic_wcsph_4M.md", "pmc_valu_wcsph_4M.md"),
            "1M": ("pmc_traffic_wcsph_1M.md", "pmc_valu_wcsph_1M.md")}


def roofline(kernels, workload=None):
    """Dominant kernel: algorithmic bytes per launch / its HIP-event average duration, vs 8 TB/s.  traffic is
    the rocprofv3-measured HBM bytes per launch of the same kernel on the same workload; `limiter` says
    whether the counters show it bound by VALU instruction issue or by HBM (MFMA plays no part: nothing
    here is a dense contraction, so the roof it is priced against is HBM)."""
    if not kernels:
        return None
    dom = max(kernels, key=lambda k: k["total_ms"])
    avg_s = dom["total_ms"] / dom["launches"] / 1e3
    achieved = dom["bytes_per_launch"] / avg_s / 1e9
    tfile, vfile = PROFILES.get(workload, (None, None))
    traffic = _profile_table(tfile, "traffic_bytes").get(dom["kernel"]) if tfile else None
    valu = _profile_table(vfile, "valu_issue").get(dom["kernel"]) if vfile else None
    out = {"kernel": dom["kernel"], "bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS,
           "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 5),
           "traffic": int(traffic) if traffic else None,
           "avg_us": round(avg_s * 1e6, 3), "bytes_per_launch": dom["bytes_per_launch"]}
    if traffic:
        out["traffic_GB_s"] = round(traffic / avg_s / 1e9, 1)
        out["traffic_frac"] = round(traffic / avg_s / 1e9 / PEAK_HBM_GBS, 4)
    if valu is not None:
        out["valu_issue_frac"] = valu
        out["limiter"] = "valu-issue" if valu > out.get("traffic_frac", 0.0) else "hbm"
    return out


def kernel_table(kernels):
    return [{"kernel": k["kernel"], "avg_us": round(1e3 * k["total_ms"] / k["launches"], 2),
             "launches": k["launches"], "GB_s": round(k["bytes_per_launch"] / (k["total_ms"] / k["launches"]) / 1e6, 1)}
            for k in kernels]


def accuracy(pa, scene_args, marks=(1, 10, 100, 1000)):
    """Norm-wise relative error of the reduced forms against the FP64 form (bit-identical to the FP64
    oracle) over fluid particles, at full size: mixed, mixed with the density pass in FP64
    (PRESTIGE_FLAG_DENSITY_F64), and FP32."""
    forms = {"f64": ("f64", 0), "mixed": ("mixed", 0), "mixed_rho64": ("mixed", 1), "f32": ("f32", 0)}
    ctxs = {}
    for name, (p, flags) in forms.items():
        d, arrs = pa.scene(pa.SCENE_DAM_BREAK, **scene_args)
        d.precision = PREC[p]
        d.flags = flags
        ctxs[name] = pa.Context(d)
        for k in STATE:
            ctxs[name].upload(k, arrs[k])
    nf = int(d.n_moving)
    out = {"steps": list(marks)}
    out.update({name: {"rho": [], "p": [], "v": [], "x": []} for name in forms if name != "f64"})
    done = 0
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))  # noqa: E731
    try:
        for m in marks:
            for c in ctxs.values():
                c.step(m - done)
            done = m
            ref = {k: ctxs["f64"].download(k)[:nf] for k in ("rho", "p", "vx", "vy", "vz", "x", "y", "z")}
            for name in out:
                if name == "steps":
                    continue
                got = {k: ctxs[name].download(k)[:nf] for k in ref}
                out[name]["rho"].append(rel(got["rho"], ref["rho"]))
                out[name]["p"].append(rel(got["p"], ref["p"]))
                out[name]["v"].append(rel(np.stack([got[k] for k in ("vx", "vy", "vz")]),
                                          np.stack([ref[k] for k in ("vx", "vy", "vz")])))
                out[name]["x"].append(rel(np.stack([got[k] for k in "xyz"]), np.stack([ref[k] for k in "xyz"])))
            log(f"[bench]   errors at step {m}: mixed v {out['mixed']['v'][-1]:.2e} p {out['mixed']['p'][-1]:.2e}; "
                f"mixed_rho64 v {out['mixed_rho64']['v'][-1]:.2e} p {out['mixed_rho64']['p'][-1]:.2e}")
    finally:
        for c in ctxs.values():
            c.close()
    return out


def _fwd_adj(pa, scene_id, scene_args, steps, every, seeds_of):
    """Steady-state cost of one calibration iteration on one context: forward
    (steps) vs checkpointed forward tape + adjoint.  One untimed iteration
    first (allocations, first touch); the initial state is re-uploaded before
    each timed run.  Loss evaluation on the host is not timed."""
    d, arrs, ctx = _context(pa, scene_id, scene_args, "mixed")
    keys = [k for k in STATE + ("radius", "mass") if k in arrs]
    dem = scene_id != pa.SCENE_DAM_BREAK
    zeros = np.zeros(d.n_particles)

    def reset():
        for k in keys:
            ctx.upload(k, arrs[k])
        if dem:
            for k in ("wx", "wy", "wz"):
                ctx.upload(k, zeros)
        ctx.sync()

    def final():
        return (np.stack([ctx.download(k) for k in "xyz"]),
                np.stack([ctx.download(k) for k in ("vx", "vy", "vz")]))

    with ctx:
        ctx.step(steps)
        reset()
        ctx.forward_tape(steps, every)
        ctx.adjoint(*seeds_of(d, *final())[1:])
        reset()
        t0 = time.perf_counter()
        ctx.step(steps)
        ctx.sync()
        t_fwd = time.perf_counter() - t0
        reset()
        t0 = time.perf_counter()
        ctx.forward_tape(steps, every)
        ctx.sync()
        t_tape = time.perf_counter() - t0
        L, gx, gv = seeds_of(d, *final())
        t0 = time.perf_counter()
        gp, gv0 = ctx.adjoint(gx, gv)
        t_adj = time.perf_counter() - t0
    return d, t_fwd, t_tape, t_adj, L, gp, gv0


def adjoint_wcsph(pa, steps=200, every=20):
    """Differentiable mixed-precision dam break (1M fluid): forward vs forward + adjoint."""
    from prestige_amd import losses

    d0, arrs = pa.scene(pa.SCENE_DAM_BREAK, **SCENE_1M)
    nf = int(d0.n_moving)
    dx = d0.h / 1.2
    extent = (d0.grid_hi[0] - 3 * dx, d0.grid_hi[1] - 3 * dx)
    target = losses.dam_break_target(np.stack([arrs[k] for k in "xyz"])[:, :nf], extent, seed=7)
    seeds = lambda d, x, v: losses.dam_break_loss(x, v, d.mass, nf, extent, target)  # noqa: E731
    d, t_fwd, t_tape, t_adj, L, gp, gv0 = _fwd_adj(pa, pa.SCENE_DAM_BREAK, SCENE_1M, steps, every, seeds)
    return {"workload": f"WCSPH dam break mixed, {nf} fluid, {steps} steps, checkpoint every {every}",
            "forward_s": round(t_fwd, 4), "tape_s": round(t_tape, 4), "adjoint_s": round(t_adj, 4),
            "fwd_adj_over_fwd": round((t_tape + t_adj) / t_fwd, 3),
            "value": nf * steps / (t_tape + t_adj), "unit": "particle-updates/s (forward + adjoint)",
            "loss": L, "dL_dalpha": float(gp[0]), "dL_dc0": float(gp[1]),
            "dL_dv0_norm": float(np.linalg.norm(gv0))}


def adjoint_dem(pa, steps=500, every=50):
    """Differentiable DEM pour (100k polydisperse spheres, mixed): forward vs forward + adjoint."""
    from prestige_amd import losses

    scene = dict(n_target=100_000, seed=0)
    _, arrs = pa.scene(pa.SCENE_DEM_POUR, **scene)
    target = losses.heap_target(np.stack([arrs[k] for k in "xyz"]), seed=11)
    seeds = lambda d, x, v: losses.heap_loss(x, target)  # noqa: E731
    d, t_fwd, t_tape, t_adj, L, gp, gv0 = _fwd_adj(pa, pa.SCENE_DEM_POUR, scene, steps, every, seeds)
    n = int(d.n_particles)
    return {"workload": f"DEM pour mixed, {n} spheres, {steps} steps, checkpoint every {every}",
            "forward_s": round(t_fwd, 4), "tape_s": round(t_tape, 4), "adjoint_s": round(t_adj, 4),
            "fwd_adj_over_fwd": round((t_tape + t_adj) / t_fwd, 3),
            "value": n * steps / (t_tape + t_adj), "unit": "particle-updates/s (forward + adjoint)",
            "forward_value": n * steps / t_fwd,
            "loss": L, "dL_dkn": float(gp[0]), "dL_de": float(gp[1]), "dL_dmu": float(gp[2]),
            "dL_dv0_norm": float(np.linalg.norm(gv0))}


def _grad_fd_check(pa, scene_id, scene_args, steps, every, loss_of, params, ndir=8, rel_eps=(2e-9, 1e-9, 5e-10),
                   seed=5, v_scale=0.1, events=False, stir=0.0):
    """Directional derivatives of the loss along `ndir` random directions in (parameters, initial
    velocities): the FP64 and mixed adjoints against central differences of the FP64 device forward.
    Each difference uses the largest relative step in rel_eps whose two runs take the same branches as
    the unperturbed run (DEM: identical contact and slip event fingerprints; WCSPH: no events are
    compared, the first step is used), and the spread against the next such step measures its noise.
    stir > 0 adds U(-stir, stir) m/s to the moving particles' initial velocities first: a dam break
    starting at rest has v_i . x_ij = 0 exactly for every pair, the kink of the viscosity switch, where
    the loss has only one-sided derivatives in velocity directions.  Returns per-direction records and
    the worst relative errors over the resolved directions."""
    d0, arrs = pa.scene(scene_id, **scene_args)
    n = int(d0.n_particles)
    if stir > 0.0:
        srng = np.random.default_rng(seed + 1000)
        arrs = dict(arrs)
        nm = int(d0.n_moving)
        for k in ("vx", "vy", "vz"):
            v = np.array(arrs[k], np.float64)
            v[:nm] += srng.uniform(-stir, stir, nm)
            arrs[k] = v
    rng = np.random.default_rng(seed)
    dirs = []
    nm = int(d0.n_moving)
    for _ in range(ndir):
        dp = rng.standard_normal(len(params)) * np.array([abs(getattr(d0, k)) for k in params])
        # initial velocities of the moving particles only: fixed particles keep theirs (the adjoint
        # returns no gradient for them), though their velocity enters the viscous term
        dv = np.zeros((3, n))
        dv[:, :nm] = rng.standard_normal((3, nm)) * v_scale
        dirs.append((dp, dv))
    zeros = np.zeros(n)

    def forward(dp, dv, eps):
        d = d0.copy()
        d.precision = PREC["f64"]
        if events:
            d.flags = pa.FLAG_DEM_EVENTS
        for k, name in enumerate(params):
            setattr(d, name, getattr(d0, name) + eps * dp[k])
            if name == "kn":   # kt = 2/7 kn moves with kn
                d.kt = d0.kt * d.kn / d0.kn
        with pa.Context(d) as ctx:
            for k in STATE + ("radius", "mass"):
                if k in arrs:
                    a = arrs[k]
                    if k in ("vx", "vy", "vz"):
                        a = a + eps * dv["xyz".index(k[1])]
                    ctx.upload(k, a)
            if scene_id != pa.SCENE_DAM_BREAK:
                for k in ("wx", "wy", "wz"):
                    ctx.upload(k, zeros)
            ctx.step(steps)
            x = np.stack([ctx.download(k) for k in "xyz"])
            v = np.stack([ctx.download(k) for k in ("vx", "vy", "vz")])
            ev = (np.stack([ctx.download("contact_events"), ctx.download("slip_events")]) if events else None)
        return loss_of(d, x, v)[0], ev

    grads = {}
    for prec in ("f64", "mixed"):
        d, ar, ctx = _context(pa, scene_id, scene_args, prec)
        with ctx:
            for k in ("vx", "vy", "vz"):
                ctx.upload(k, arrs[k])
            if scene_id != pa.SCENE_DAM_BREAK:
                for k in ("wx", "wy", "wz"):
                    ctx.upload(k, zeros)
            ctx.forward_tape(steps, every)
            x = np.stack([ctx.download(k) for k in "xyz"])
            v = np.stack([ctx.download(k) for k in ("vx", "vy", "vz")])
            _, gx, gv = loss_of(d, x, v)
            grads[prec] = ctx.adjoint(gx, gv)
    _, ev0 = forward(np.zeros(len(params)), np.zeros((3, n)), 0.0)
    recs = []
    for k, (dp, dv) in enumerate(dirs):
        rec = {"direction": k, "same_events": False}
        fds = []
        for eps in rel_eps:
            lp, evp = forward(dp, dv, eps)
            lm, evm = forward(dp, dv, -eps)
            same = (not events) or (np.array_equal(evp, ev0) and np.array_equal(evm, ev0))
            if same:
                fds.append((eps, (lp - lm) / (2 * eps)))
            elif not fds:
                rec.update({"eps": eps, "fd": (lp - lm) / (2 * eps)})
        if fds:
            rec.update({"eps": fds[0][0], "same_events": True, "fd": fds[0][1]})
            # the spread of the difference quotient over the step sizes that keep every branch
            rec["fd_spread"] = max((abs(f - fds[0][1]) for _, f in fds[1:]), default=None)
        for prec, (gp, gv0) in grads.items():
            pred = float(gp @ dp) + float((gv0 * dv).sum())
            rec[f"adjoint_{prec}"] = pred
            rec[f"rel_err_{prec}"] = abs(pred - rec["fd"]) / max(abs(rec["fd"]), 1e-300)
        log(f"[bench]   direction {k}: fd {rec['fd']:.6e} f64 {rec['rel_err_f64']:.2e} "
            f"mixed {rec['rel_err_mixed']:.2e} eps {rec['eps']:.0e} spread {rec.get('fd_spread')} "
            f"same events {rec['same_events']}")
        recs.append(rec)
    ok = [r for r in recs if r["same_events"]]
    # A direction is resolved when its difference quotient stands at least ten times above its own
    # spread over step sizes; the others are reported, not ranked.
    res = [r for r in ok if r.get("fd_spread") is not None and abs(r["fd"]) > 10 * r["fd_spread"]]
    return {"directions": recs, "valid_directions": len(ok), "resolved_directions": len(res),
            "max_rel_err_f64": max((r["rel_err_f64"] for r in res), default=None),
            "max_rel_err_mixed": max((r["rel_err_mixed"] for r in res), default=None)}


def fd_dem(pa, steps=500, every=50):
    """Pour of 100,000 spheres, 500 steps: gradients w.r.t. kn, e, mu and the initial velocities."""
    from prestige_amd import losses

    scene = dict(n_target=100_000, seed=0)
    _, arrs = pa.scene(pa.SCENE_DEM_POUR, **scene)
    target = losses.heap_target(np.stack([arrs[k] for k in "xyz"]), seed=11)
    return _grad_fd_check(pa, pa.SCENE_DEM_POUR, scene, steps, every,
                          lambda d, x, v: losses.heap_loss(x, target), ("kn", "restitution", "mu"),
                          v_scale=0.05, events=True)


def fd_wcsph(pa, steps=200, every=20):
    """Dam break of 1,000,000 fluid particles, 200 steps, the fluid starting with U(-1, 1) mm/s
    velocities: gradients w.r.t. alpha, c0 and the initial velocities (dt fixed)."""
    from prestige_amd import losses

    d0, arrs = pa.scene(pa.SCENE_DAM_BREAK, **SCENE_1M)
    nf = int(d0.n_moving)
    dx = d0.h / 1.2
    extent = (d0.grid_hi[0] - 3 * dx, d0.grid_hi[1] - 3 * dx)
    target = losses.dam_break_target(np.stack([arrs[k] for k in "xyz"])[:, :nf], extent, seed=7)
    return _grad_fd_check(pa, pa.SCENE_DAM_BREAK, SCENE_1M, steps, every,
                          lambda d, x, v: losses.dam_break_loss(x, v, d.mass, nf, extent, target),
                          ("alpha", "c0"), rel_eps=(1e-7, 5e-8), v_scale=0.01, stir=1e-3)


def generated_density(pa, scene_args, reps=20):
    """The built-in density and EOS equation bodies run through the generated path
    (prestige_run_equations: hiprtc-compiled gather pass over the cell iterator) on the
    1M dam break, for comparison with the hand-written density kernel.  avg_us is the
    pass plus the copy of the fields it reads at [j] into slot order."""
    d, a = pa.scene(pa.SCENE_DAM_BREAK, **scene_args)
    d.model = pa.MODEL_PAIRS
    with pa.Context(d) as ctx:
        for k in ("x", "y", "z"):
            ctx.upload(k, a[k])
        ctx.upload("mass", np.full(d.n_particles, d.mass))
        ctx.neighbors_build()
        eqs = [pa.EQ_SPH_DENSITY, pa.EQ_SPH_EOS]
        t0 = time.perf_counter()
        ctx.run_equations(eqs, pa.LOOP_CELLS, d.cell_size)   # first call compiles
        t_first = time.perf_counter() - t0
        ctx.timing_enable(True)
        for _ in range(reps):
            ctx.run_equations(eqs, pa.LOOP_CELLS, d.cell_size)
        tm = {k["kernel"]: k for k in ctx.timing()}
    kt, kc = tm["generated_pass"], tm.get("generated_slot_copy")
    copy_us = 1e3 * kc["total_ms"] / reps if kc else 0.0
    return {"workload": "density + EOS bodies of the equation table, generated, cell iterator, "
                        f"{d.n_particles} particles",
            "avg_us": round(1e3 * kt["total_ms"] / kt["launches"] + copy_us, 1),
            "pass_us": round(1e3 * kt["total_ms"] / kt["launches"], 1), "slot_copy_us": round(copy_us, 1),
            "launches": kt["launches"], "first_call_s": round(t_first, 3)}


def eq1_allpairs(pa, n=10_000, reps=20):
    """The reference's own loop, force[i] += mass[j] over every (i, j) with i outer and j inner, j == i
    included (the generated loop and body of the reference's simple CPU backend), on the 10,000-particle
    lattice: the device kernel (one lane per i, j in order, so each force is the sequential sum) against
    the oracle's literal loop on one host thread, as the reference runs it.  1e8 pairs per call."""
    import oracle

    d, a = pa.scene(pa.SCENE_DEM_LATTICE, n_target=n, seed=0)
    d.model = pa.MODEL_PAIRS
    n = int(d.n_particles)
    mass = np.random.default_rng(3).random(n)
    with pa.Context(d) as ctx:
        for k in ("x", "y", "z"):
            ctx.upload(k, a[k])
        ctx.upload("mass", mass)
        ctx.upload("force", np.zeros(n))
        ctx.eval_equation(pa.EQ1, pa.LOOP_ALLPAIRS)
        ctx.timing_enable(True)
        for _ in range(reps):
            ctx.eval_equation(pa.EQ1, pa.LOOP_ALLPAIRS)
        tm = {k["kernel"]: k for k in ctx.timing()}
    kt = tm["eq1_allpairs"]
    us = 1e3 * kt["total_ms"] / kt["launches"]
    t0 = time.perf_counter()
    oracle.eq1_allpairs(np.zeros(n), mass)
    cpu_s = time.perf_counter() - t0
    return {"workload": f"eq1 force[i] += mass[j], all pairs, {n} particles ({n * n:.0e} pairs per call)",
            "value": n * n / (us * 1e-6), "unit": "pairs/s", "avg_us": round(us, 2), "launches": kt["launches"],
            "cpu": {"value": n * n / cpu_s, "unit": "pairs/s", "cores": 1, "kind": "port",
                    "sample": f"one call of the oracle's literal i-j loop, n = {n}"}}


def eq1_cells(pa, reps=20):
    """eq1 over the cell iterator (j over the 27 cells of i's cell within the cutoff 2h, in row-run order)
    on the 1M dam break: the hand-written kernel and the same body through the generator, against the
    oracle's cell loop on the host cores."""
    import oracle

    d, a = pa.scene(pa.SCENE_DAM_BREAK, **SCENE_1M)
    n = int(d.n_particles)
    cutoff = d.cell_size
    mass = np.random.default_rng(4).random(n)
    out = {"workload": f"eq1 force[i] += mass[j] over j within 2h (cell iterator), {n} particles "
                       "of the 1M dam break"}
    d.model = pa.MODEL_PAIRS
    with pa.Context(d) as ctx:
        for k in ("x", "y", "z"):
            ctx.upload(k, a[k])
        ctx.upload("mass", mass)
        ctx.upload("force", np.zeros(n))
        ctx.neighbors_build()
        ctx.eval_equation(pa.EQ1, pa.LOOP_CELLS, cutoff)
        ctx.run_equations([pa.EQ1], pa.LOOP_CELLS, cutoff)   # compiles
        ctx.timing_enable(True)
        for _ in range(reps):
            ctx.eval_equation(pa.EQ1, pa.LOOP_CELLS, cutoff)
            ctx.run_equations([pa.EQ1], pa.LOOP_CELLS, cutoff)
        tm = {k["kernel"]: k for k in ctx.timing()}
    for name, key in (("hand_written", "eq1_cells"), ("generated", "generated_pass")):
        k = tm[key]
        us = 1e3 * k["total_ms"] / k["launches"]
        out[name] = {"avg_us": round(us, 1), "particles_per_s": n / (us * 1e-6)}
    if "generated_slot_copy" in tm:
        k = tm["generated_slot_copy"]
        out["generated"]["slot_copy_us"] = round(1e3 * k["total_ms"] / k["launches"], 1)
    cores = host_cores()
    prev = oracle.set_threads(cores)
    try:
        st = oracle.State(oracle.params_from_desc(d), a)
        assert st.build() == 0
        st.eq1_cells(cutoff, np.zeros(n), mass)
        t0 = time.perf_counter()
        st.eq1_cells(cutoff, np.zeros(n), mass)
        sec = time.perf_counter() - t0
        st.close()
    finally:
        oracle.set_threads(prev)
    out["cpu"] = {"particles_per_s": n / sec, "cores": cores, "kind": "port",
                  "sample": f"one call of the oracle's cell loop, OpenMP {cores} threads"}
    return out


def copy_probe(nbytes):
    """hipMemcpyDtoD of nbytes (the runtime's copy kernel): reads and writes exactly nbytes of HBM."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpyDtoD(b, a, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"[bench] copy probe: 3 x {nbytes} bytes device to device")


def _cpu_rate(pa, scene_id, scene_args, threads, budget_s, max_steps=50):
    """FP64 oracle steps per second on `threads` host threads, over about budget_s of steps after one
    warm-up step (first touch, thread pool)."""
    import oracle

    d, arrs = pa.scene(scene_id, **scene_args)
    prev = oracle.set_threads(threads)
    try:
        st = oracle.State(oracle.params_from_desc(d), arrs)
        assert st.step(1) == 0
        t0 = time.perf_counter()
        assert st.step(1) == 0
        one = time.perf_counter() - t0
        k = max(1, min(max_steps, int(budget_s / max(one, 1e-6))))
        t0 = time.perf_counter()
        assert st.step(k) == 0
        sec = time.perf_counter() - t0
        st.close()
    finally:
        oracle.set_threads(prev)
    return d, k, sec


def host_cores():
    """Threads the host grants this process (the GPU box exports OMP_NUM_THREADS for its CPU share)."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(pa, scene_args, budget_s=10.0):
    """FP64 oracle (OpenMP over particles) on the host cores, bounded sample of the same workload: all the
    host threads, and one thread."""
    cores = host_cores()
    d, k, sec = _cpu_rate(pa, pa.SCENE_DAM_BREAK, scene_args, cores, budget_s)
    out = {"value": float(d.n_moving) * k / sec, "unit": "particle-updates/s", "cores": cores, "kind": "port",
           "sample": f"{k} FP64 steps of the same {d.n_particles}-particle dam break (after 2 warm-up "
                     f"steps), C oracle, OpenMP {cores} threads"}
    d1, k1, sec1 = _cpu_rate(pa, pa.SCENE_DAM_BREAK, scene_args, 1, budget_s, max_steps=5)
    out["one_core"] = {"value": float(d1.n_moving) * k1 / sec1, "cores": 1,
                       "sample": f"{k1} FP64 steps of the same dam break, 1 thread"}
    return out


def dem_lattice(pa, steps, warmup):
    """The 10,000-sphere FP64 DEM lattice (1000 steps in its configuration) on the device, next to the
    FP64 oracle on the host cores for the same scene."""
    scene = dict(n_target=10_000, seed=0)
    r = run_form(pa, scene, "f64", steps, warmup, scene_id=pa.SCENE_DEM_LATTICE)
    cores = host_cores()
    d, k, sec = _cpu_rate(pa, pa.SCENE_DEM_LATTICE, scene, cores, 5.0, max_steps=1000)
    return {"workload": f"DEM lattice FP64, {r['n_fluid']} spheres r = 0.01 m, kn 1e5, e 0.3, mu 0.5, "
                        f"dt 1e-5, {steps} steps",
            "value": r["value"], "ms_per_step": r["ms_per_step"], "kernels": kernel_table(r["kernels"]),
            "cpu": {"value": float(d.n_moving) * k / sec, "cores": cores, "kind": "port",
                    "sample": f"{k} FP64 steps of the same lattice, C oracle, OpenMP {cores} threads"}}


WORKLOADS = {
    "4M": (SCENE_4M, "WCSPH dam break, 4,019,679 fluid (159^3 block) + 2,006,931 tank particles, dx 3.15 mm, "
                     "h = 1.2 dx, cubic spline, Tait gamma 7, alpha 0.1, dt = 0.25 h / c0"),
    "1M": (SCENE_1M, "WCSPH dam break, 1,000,000 fluid + 799,320 tank particles, dx 5 mm, "
                     "h = 1.2 dx, cubic spline, Tait gamma 7, alpha 0.1, dt = 0.25 h / c0"),
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", choices=list(PREC), default="f64")
    ap.add_argument("--workload", choices=list(WORKLOADS), default="4M")
    ap.add_argument("--no-forms", action="store_true", help="skip the other forms, accuracy and adjoint runs")
    ap.add_argument("--forms", default="all",
                    help="comma-separated subset of the forms to run (f64_1M, f32_1M, mixed_1M, mixed_4M, accuracy, "
                         "dem_lattice_10k, eq1_allpairs_10k, eq1_cells_1M, adjoint_wcsph_1M, adjoint_dem_100k, fd_dem, fd_wcsph, generated)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="first run a plain device-to-device copy of BYTES (a reference kernel of known "
                         "traffic for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only "
                                   "(spatial slabs with halo exchange), so --gpus must be 1"}))
        sys.exit(2)

    import prestige_amd as pa

    pa.load_library()
    if args.calibrate_copy > 0:
        copy_probe(args.calibrate_copy)
    scene_main, workload_text = WORKLOADS[args.workload]
    log(f"[bench] main form {args.precision} ({args.workload}): {args.steps} steps after {args.warmup} warm-up")
    main_res = run_form(pa, scene_main, args.precision, args.steps, args.warmup)
    forms = {}
    want = set(args.forms.split(",")) if args.forms != "all" else None
    run = lambda name: not args.no_forms and (want is None or name in want)  # noqa: E731
    for p in PREC:
        if args.workload == "1M" and p == args.precision:
            continue
        if run(f"{p}_1M"):
            log(f"[bench] form {p} (1M)")
            r = run_form(pa, SCENE_1M, p, min(args.steps, 500), args.warmup)
            forms[f"{p}_1M"] = {"value": r["value"], "ms_per_step": r["ms_per_step"],
                                "roofline": roofline(r["kernels"], "1M"), "kernels": kernel_table(r["kernels"])}
    if run("mixed_4M") and not (args.workload == "4M" and args.precision == "mixed"):
        log("[bench] form mixed (4M, dx = 3.15 mm)")
        r = run_form(pa, SCENE_4M, "mixed", min(args.steps, 300), args.warmup)
        forms["mixed_4M"] = {"value": r["value"], "ms_per_step": r["ms_per_step"], "n_fluid": r["n_fluid"],
                             "n_total": r["n_total"], "roofline": roofline(r["kernels"], "4M"),
                             "kernels": kernel_table(r["kernels"])}
    if run("accuracy"):
        log("[bench] accuracy of the reduced forms vs f64 at 4M over 1000 steps")
        forms.setdefault("mixed_4M", {})["error_vs_f64"] = accuracy(pa, SCENE_4M)
    if run("dem_lattice_10k"):
        log("[bench] DEM lattice 10k FP64, 1000 steps")
        forms["dem_lattice_10k"] = dem_lattice(pa, 1000, 10)
    if run("eq1_allpairs_10k"):
        log("[bench] eq1 all pairs, 10k particles")
        forms["eq1_allpairs_10k"] = eq1_allpairs(pa)
    if run("eq1_cells_1M"):
        log("[bench] eq1 over the cell iterator, 1M dam break")
        forms["eq1_cells_1M"] = eq1_cells(pa)
    if run("adjoint_wcsph_1M"):
        log("[bench] adjoint: WCSPH 1M mixed, 200 steps, K = 20")
        forms["adjoint_wcsph_1M"] = adjoint_wcsph(pa)
    if run("adjoint_dem_100k"):
        log("[bench] adjoint: DEM pour 100k mixed, 500 steps, K = 50")
        forms["adjoint_dem_100k"] = adjoint_dem(pa)
    if run("fd_dem"):
        log("[bench] adjoint vs central differences: DEM pour 100k, 500 steps, 8 directions")
        forms.setdefault("adjoint_dem_100k", {})["fd_check"] = fd_dem(pa)
    if run("fd_wcsph"):
        log("[bench] adjoint vs central differences: WCSPH 1M, 200 steps, 8 directions")
        forms.setdefault("adjoint_wcsph_1M", {})["fd_check"] = fd_wcsph(pa)
    if run("generated"):
        log("[bench] generated density + EOS pass (1M)")
        forms["generated_density_1M"] = generated_density(pa, SCENE_1M)
    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU baseline (oracle, {args.workload})")
        cpu = cpu_baseline(pa, scene_main)
    out = {
        "metric": "particle-updates/s/GPU",
        "value": main_res["value"],
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic (deterministic dam-break lattice)",
        "config": {
            "workload": workload_text,
            "n_fluid": main_res["n_fluid"],
            "n_total": main_res["n_total"],
            "precision": args.precision,
        },
        "roofline": roofline(main_res["kernels"], args.workload),
        "cpu_baseline": cpu,
        "kernels": kernel_table(main_res["kernels"]),
        "forms": forms,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(main_res["value"] / cpu["value"], 2)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
