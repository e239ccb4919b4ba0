"""Physics pins of the oracle's WCSPH terms, each a check that returns a measured error.

Every check is computed from textbook statements written here independently of
the oracle's arrangement of its formulas, and runs on any build of the oracle
(`L`, a library handle from oracle.lib() or oracle.load()), so that
tests/test_oracle_pins.py can run the same checks on builds with a planted
mistake and see them fail.

  * tait_interior     Tait EOS p = B((rho/rho0)^7 - 1), B = rho0 c0^2 / 7, on
                      the interior of a lattice, whose density is the closed
                      form rho/rho0 = 1.000809548 (Monaghan 1994).
  * pressure_gradient The pressure acceleration is the exact gradient of the
                      internal energy V = sum_k m u(rho_k), u' = p / rho^2: the
                      SPH momentum equation follows from that Lagrangian
                      (Monaghan 1992, sec. 3).  The gradient is a complex-step
                      derivative of V computed here from the cubic spline.
  * viscous_pair      Two approaching particles with p clamped to 0: the
                      acceleration is -m Pi_ij grad W_ij with Monaghan's
                      (1992) Pi_ij = -alpha c0 mu_ij / rhobar_ij,
                      mu_ij = h v_ij.r_ij / (r_ij^2 + 0.01 h^2); a receding
                      pair feels nothing.
  * dissipation       A stirred block at p = 0 loses kinetic energy at every
                      step and keeps its momentum.
  * hydrostatic       A settled water column in a tank: the pressure gradient
                      in its interior is rho0 g.
"""
from __future__ import annotations

import math

import numpy as np

import oracle
import prestige_amd as pa

# ---------------------------------------------------------------------------
# textbook kernel and equation of state (Monaghan & Lattanzio 1985; Monaghan 1994)
# ---------------------------------------------------------------------------


def spline_f(q):
    """Cubic spline shape, W = f(q) / (pi h^3); works on real or complex q (branches on the real part)."""
    q = np.asarray(q)
    qr = q.real
    t = 2.0 - q
    return np.where(qr < 1.0, 1.0 - 1.5 * q**2 + 0.75 * q**3, np.where(qr < 2.0, 0.25 * t**3, 0.0))


def spline_df(q):
    q = np.asarray(q)
    t = 2.0 - q
    return np.where(q < 1.0, -3.0 * q + 2.25 * q**2, np.where(q < 2.0, -0.75 * t**2, 0.0))


def tait_B(rho0, c0):
    return rho0 * c0 * c0 / 7.0


def internal_energy(rho, rho0, B):
    """u(rho) = int_rho0^rho p(s) / s^2 ds for p = B((s/rho0)^7 - 1), valid for rho >= rho0 (no clamp)."""
    return B * (rho**6 / (6.0 * rho0**7) + 1.0 / rho) - B * (1.0 / (6.0 * rho0) + 1.0 / rho0)


def brute_density(pos, m, h):
    """rho_i = sum_j m W(|x_i - x_j|, h) over all pairs (j == i included); pos (n, 3), real or complex."""
    d = pos[:, None, :] - pos[None, :, :]
    r = np.sqrt((d * d).sum(-1))
    return m / (math.pi * h**3) * spline_f(r / h).sum(1)


# ---------------------------------------------------------------------------
# scenes
# ---------------------------------------------------------------------------


def _desc(n, h, rho0=1000.0, c0=20.0, alpha=0.1, mass=None, g=0.0, lo=-1.0, hi=1.0):
    d = pa.desc_init(pa.MODEL_WCSPH)
    d.n_particles = n
    d.n_moving = n
    d.h = h
    d.cell_size = 2.0 * h
    d.rho0 = rho0
    d.c0 = c0
    d.alpha = alpha
    d.mass = mass
    d.dt = 0.25 * h / c0
    d.gravity[2] = -g
    for a in range(3):
        d.grid_lo[a] = lo
        d.grid_hi[a] = hi
    return d


def _arrays(pos, vel=None):
    vel = np.zeros_like(pos) if vel is None else vel
    return {"x": pos[:, 0], "y": pos[:, 1], "z": pos[:, 2], "vx": vel[:, 0], "vy": vel[:, 1], "vz": vel[:, 2]}


def _acc(st):
    return np.stack([st.get("ax"), st.get("ay"), st.get("az")], 1)


def _state(L, d, arrs):
    return oracle.State(oracle.params_from_desc(d), arrs, L)


# ---------------------------------------------------------------------------
# checks: each returns the error it measures
# ---------------------------------------------------------------------------


def lattice_ratio():
    """rho / rho0 of an interior site of a cubic lattice at h = 1.2 dx with m = rho0 dx^3."""
    g = np.arange(-3, 4)
    o = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    q = np.linalg.norm(o, axis=1) / 1.2
    return float(spline_f(q).sum() / (math.pi * 1.2**3))


def tait_interior(L):
    """max |p / (B (ratio^7 - 1)) - 1| over the interior of an 8,000-particle block in its tank."""
    d, arrs = pa.scene(pa.SCENE_DAM_BREAK, n_target=8000, dx=0.005)
    st = _state(L, d, arrs)
    assert st.density() == 0
    inner = np.arange(8000).reshape(20, 20, 20)[5:15, 5:15, 5:15].ravel()
    want = tait_B(d.rho0, d.c0) * (lattice_ratio() ** 7 - 1.0)
    p = st.get("p")[inner]
    st.close()
    return float(np.max(np.abs(p / want - 1.0)))


def surface_clamp(L):
    """Largest |p| among the block's top-layer particles, whose density is below rho0 (p clamped at 0)."""
    d, arrs = pa.scene(pa.SCENE_DAM_BREAK, n_target=8000, dx=0.005)
    st = _state(L, d, arrs)
    assert st.density() == 0
    top = np.arange(8000).reshape(20, 20, 20)[19].ravel()   # z-slowest: the last layer is the top
    rho, p = st.get("rho")[top], st.get("p")[top]
    st.close()
    assert np.all(rho < d.rho0)
    return float(np.max(np.abs(p)))


def _cluster(side=4, dx=0.01, jitter=0.15, seed=3):
    rng = np.random.default_rng(seed)
    g = np.arange(side) * dx
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    return pos + rng.uniform(-jitter * dx, jitter * dx, pos.shape)


def pressure_gradient(L):
    """max |a_oracle + grad V / m| / max |a| for a compressed, inviscid 64-particle cluster at rest."""
    dx = 0.01
    h = 1.2 * dx
    pos = _cluster(dx=dx)
    n = len(pos)
    rho0, c0 = 1000.0, 20.0
    # mass such that every particle is compressed (rho >= 1.05 rho0), so p > 0 and u is the smooth branch
    m = 1.05 * rho0 / brute_density(pos, 1.0, h).min()
    d = _desc(n, h, rho0=rho0, c0=c0, alpha=0.0, mass=m)
    st = _state(L, d, _arrays(pos))
    assert st.step(1) == 0
    a = _acc(st)
    st.close()
    B = tait_B(rho0, c0)
    eps = 1e-20 * dx
    grad = np.zeros_like(pos)
    for i in range(n):
        for c in range(3):
            z = pos.astype(np.complex128)
            z[i, c] += 1j * eps
            V = (m * internal_energy(brute_density(z, m, h), rho0, B)).sum()
            grad[i, c] = V.imag / eps
    want = -grad / m
    return float(np.max(np.abs(a - want)) / np.max(np.abs(want)))


def _pair(L, approach: bool, q0=1.3, v0=0.5, alpha=0.5, c0=20.0):
    h = 0.012
    r0 = q0 * h
    pos = np.array([[-0.5 * r0, 0.0, 0.0], [0.5 * r0, 0.0, 0.0]])
    s = 1.0 if approach else -1.0
    vel = np.array([[s * v0, 0.0, 0.0], [-s * v0, 0.0, 0.0]])
    rho0 = 1000.0
    m = 0.5 * rho0 * math.pi * h**3 / (1.0 + float(spline_f(q0)))   # rho = rho0 / 2: p clamped to 0
    d = _desc(2, h, rho0=rho0, c0=c0, alpha=alpha, mass=m)
    st = _state(L, d, _arrays(pos, vel))
    assert st.step(1) == 0
    a, p, rho = _acc(st), st.get("p"), st.get("rho")
    st.close()
    return a, p, rho, dict(h=h, r0=r0, v0=v0, m=m, alpha=alpha, c0=c0, q0=q0, pos=pos, vel=vel)


def viscous_pair(L):
    """Relative error of an approaching pair's acceleration against Monaghan's Pi_ij, plus |a| of a
    receding pair (which must be exactly 0) and |p| of both (clamped to 0)."""
    a, p, rho, k = _pair(L, True)
    h, m = k["h"], k["m"]
    rho_e = brute_density(k["pos"], m, h)
    rij = k["pos"][0] - k["pos"][1]
    vij = k["vel"][0] - k["vel"][1]
    r = np.linalg.norm(rij)
    mu = h * vij.dot(rij) / (r * r + 0.01 * h * h)
    Pi = -k["alpha"] * k["c0"] * mu / (0.5 * (rho_e[0] + rho_e[1]))
    gradW = float(spline_df(r / h)) / (math.pi * h**4) * rij / r       # grad_i W(|x_i - x_j|, h)
    want0 = -m * Pi * gradW
    want = np.stack([want0, -want0])
    err = float(np.max(np.abs(a - want)) / np.max(np.abs(want)))
    a2, p2, _, _ = _pair(L, False)
    return err, float(np.max(np.abs(a2))), float(max(np.max(np.abs(p)), np.max(np.abs(p2))))


def dissipation(L, steps=20):
    """(largest relative kinetic-energy increase over one step, largest relative momentum drift, max |p|)
    for a stirred 216-particle block at p = 0 without gravity."""
    dx = 0.01
    h = 1.2 * dx
    pos = _cluster(side=6, dx=dx, jitter=0.1, seed=5)
    rng = np.random.default_rng(9)
    vel = rng.uniform(-0.5, 0.5, pos.shape)
    rho0 = 1000.0
    m = 0.5 * rho0 / brute_density(pos, 1.0, h).max()            # rho <= rho0 / 2 everywhere: p = 0
    d = _desc(len(pos), h, rho0=rho0, c0=20.0, alpha=0.5, mass=m)
    st = _state(L, d, _arrays(pos, vel))

    def ke_mom():
        v = np.stack([st.get("vx"), st.get("vy"), st.get("vz")], 1)
        return 0.5 * m * float((v * v).sum()), m * v.sum(0)

    ke0, mom0 = ke_mom()
    worst_gain, worst_drift, pmax = -np.inf, 0.0, 0.0
    prev = ke0
    for _ in range(steps):
        assert st.step(1) == 0
        ke, mom = ke_mom()
        worst_gain = max(worst_gain, (ke - prev) / ke0)
        worst_drift = max(worst_drift, float(np.max(np.abs(mom - mom0))) / (m * len(pos) * 0.5))
        pmax = max(pmax, float(np.max(np.abs(st.get("p")))))
        prev = ke
    st.close()
    return worst_gain, worst_drift, pmax


def water_column(nxy=6, nz=20, dx=0.01, layers=3, free=6):
    """A column of nxy x nxy x nz fluid lattice sites in a tank of the same inner section whose floor and
    walls are `layers` lattice layers of fixed particles, walls `free` sites above the water."""
    g = lambda a, b: (np.arange(a, b) + 0.5) * dx  # noqa: E731
    fluid = np.stack(np.meshgrid(g(0, nxy), g(0, nxy), g(0, nz), indexing="ij"), -1).reshape(-1, 3)
    allp = np.stack(np.meshgrid(g(-layers, nxy + layers), g(-layers, nxy + layers), g(-layers, nz + free),
                                indexing="ij"), -1).reshape(-1, 3)
    inside = (allp[:, 0] > 0) & (allp[:, 0] < nxy * dx) & (allp[:, 1] > 0) & (allp[:, 1] < nxy * dx) & (allp[:, 2] > 0)
    wall = allp[~inside]
    return fluid, wall


def hydrostatic(L, settle_s=0.4, samples=10, every=20, alpha=1.0):
    """Relative error of the settled column's interior pressure gradient -dp/dz against rho0 g: after
    settle_s of settling (alpha raised to damp the start-up sloshing; the equilibrium does not depend on
    it), `samples` snapshots `every` steps apart of the fluid particles between 0.2 H and 0.7 H (H the
    water height) pooled into one least-squares slope of p(z)."""
    dx, nz = 0.01, 20
    fluid, wall = water_column(nz=nz, dx=dx)
    pos = np.concatenate([fluid, wall])
    n, nf = len(pos), len(fluid)
    h = 1.2 * dx
    H = nz * dx
    gacc = 9.81
    c0 = 10.0 * math.sqrt(2.0 * gacc * H)
    d = _desc(n, h, rho0=1000.0, c0=c0, alpha=alpha, mass=1000.0 * dx**3, g=gacc, lo=-0.05, hi=0.35)
    d.n_moving = nf
    st = _state(L, d, _arrays(pos))
    assert st.step(int(settle_s / d.dt)) == 0
    zs, ps = [], []
    for _ in range(samples):
        assert st.step(every) == 0
        z, p = st.get("z")[:nf], st.get("p")[:nf]
        sel = (z > 0.2 * H) & (z < 0.7 * H)
        zs.append(z[sel])
        ps.append(p[sel])
    st.close()
    slope = np.polyfit(np.concatenate(zs), np.concatenate(ps), 1)[0]
    return float(abs(-slope / (d.rho0 * gacc) - 1.0))
