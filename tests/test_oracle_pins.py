"""Physics pins of the oracle's WCSPH equation of state, pressure force and
artificial viscosity (no device), and proof that they bite: each planted
mistake below, compiled into a copy of the oracle sources on the CPU, makes
at least the named check fail.

The device's FP64 form is compared with the oracle bit for bit, so these pins
are what fixes the physics of both.
"""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

import oracle
import oracle_pins as P

ORACLE_DIR = Path(oracle.__file__).resolve().parent


def test_tait_interior_closed_form():
    """Interior lattice: p = rho0 c0^2 / 7 ((1.000809548)^7 - 1), to FP64 rounding of the lattice sum."""
    assert P.tait_interior(oracle.lib()) < 1e-8


def test_tait_clamps_surface_pressure():
    """Surface particles are below rho0 and their pressure is clamped to exactly 0."""
    assert P.surface_clamp(oracle.lib()) == 0.0


def test_pressure_force_is_energy_gradient():
    """Inviscid cluster: a_i = -(1/m) dV/dx_i with V = sum m u(rho), u' = p / rho^2."""
    assert P.pressure_gradient(oracle.lib()) < 1e-10


def test_viscous_pair_matches_monaghan():
    err, receding, pmax = P.viscous_pair(oracle.lib())
    assert pmax == 0.0
    assert err < 1e-12
    assert receding == 0.0


def test_viscosity_dissipates_and_conserves_momentum():
    gain, drift, pmax = P.dissipation(oracle.lib())
    assert pmax == 0.0
    assert gain < 0.0           # kinetic energy falls at every step
    assert drift < 1e-12


def test_hydrostatic_column():
    """Settled column: -dp/dz = rho0 g within 6 % (the start-up sloshing is not fully gone)."""
    assert P.hydrostatic(oracle.lib(), settle_s=0.6, samples=10, every=40) < 0.06


# planted mistakes: (name, text in physics_impl.h, replacement, checks that must fail)
MUTATIONS = [
    ("viscosity sign flipped",
     "num = FMA(pri + s->pr[j], drb, -(ach * vdotr)) * dw;",
     "num = FMA(pri + s->pr[j], drb, ach * vdotr) * dw;",
     ("viscous_pair", "dissipation")),
    ("mu without h",
     "const REAL ach = s->alpha_c0 * h;",
     "const REAL ach = s->alpha_c0;",
     ("viscous_pair",)),
    ("eps^2 = 0.1 h^2",
     "s->eps2 = (0.01 * h) * h;",
     "s->eps2 = (0.1 * h) * h;",
     ("viscous_pair",)),
    ("B without /7",
     "s->B = ((P->rho0 * c0) * c0) / 7.0;",
     "s->B = ((P->rho0 * c0) * c0);",
     ("tait_interior", "pressure_gradient")),
    ("Tait exponent 6",
     "REAL r7 = r4 * r2_ * ratio;",
     "REAL r7 = r4 * r2_;",
     ("tait_interior", "pressure_gradient")),
    ("momentum scale m sigma / h doubled",
     "s->msh = (P->mass * s->sigma) * s->inv_h;",
     "s->msh = 2.0 * ((P->mass * s->sigma) * s->inv_h);",
     ("pressure_gradient", "viscous_pair", "hydrostatic")),
    ("pressure force sign flipped",
     "s->ax[i] = P->g[0] - msh * sx;",
     "s->ax[i] = P->g[0] + msh * sx;",
     ("pressure_gradient",)),
]


def _failing(L, name):
    """True when check `name` fails on the oracle build L (a check that cannot run, e.g. a particle thrown
    out of the grid, fails too)."""
    try:
        return _check_fails(L, name)
    except AssertionError:
        return True


def _check_fails(L, name):
    if name == "tait_interior":
        return P.tait_interior(L) >= 1e-8
    if name == "pressure_gradient":
        return P.pressure_gradient(L) >= 1e-10
    if name == "viscous_pair":
        err, receding, pmax = P.viscous_pair(L)
        return not (err < 1e-12 and receding == 0.0 and pmax == 0.0)
    if name == "dissipation":
        gain, drift, pmax = P.dissipation(L)
        return not (gain < 0.0 and drift < 1e-12 and pmax == 0.0)
    if name == "hydrostatic":
        return P.hydrostatic(L, settle_s=0.3, samples=5, every=40) >= 0.06
    raise KeyError(name)


@pytest.mark.parametrize("name,old,new,checks", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_planted_mistake_is_caught(tmp_path, name, old, new, checks):
    for f in ("oracle.c", "oracle.h", "physics_impl.h", "Makefile"):
        shutil.copy(ORACLE_DIR / f, tmp_path / f)
    src = (tmp_path / "physics_impl.h").read_text()
    assert src.count(old) == 1, f"mutation site of '{name}' not found once"
    (tmp_path / "physics_impl.h").write_text(src.replace(old, new))
    subprocess.run(["make", "-s", "-C", str(tmp_path)], check=True)
    L = oracle.load(tmp_path / "liboracle.so")
    for check in checks:
        assert _failing(L, check), f"'{check}' does not catch: {name}"
