"""ctypes binding of oracle/liboracle.so — the CPU FP64 oracle.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
the CPU-baseline leg of bench.py, as the checker; nothing in prestige_amd
imports it.  See oracle.h for what each function restates (reference files
prestige/src/codegen/simple_cpu.rs:7-19, prestige/src/lib.rs:7-12) and for
the complex-step forward-mode derivative.

Parity pins (tests/test_oracle.py): closed forms (spline normalisation,
lattice density 1.000809548 rho0, 57 neighbours on an interior lattice site,
discrete free fall, head-on restitution e, Coulomb sliding), brute-force
all-pairs neighbour sets, momentum conservation, and complex-step vs central
differences.  WCSPH equation of state, pressure force and viscosity
(tests/test_oracle_pins.py): the Tait closed form on the lattice interior,
the pressure force as the exact gradient of the EOS internal energy,
Monaghan's Pi_ij on an approaching pair, dissipation at p = 0 and a
hydrostatic column; planted mistakes compiled into copies of these sources
(load()) show that each check fails when its term is wrong.
"""
from __future__ import annotations

import ctypes as C
import math
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

MODEL_PAIRS, MODEL_WCSPH, MODEL_DEM = 0, 1, 2


class Params(C.Structure):
    _fields_ = [
        ("model", C.c_int32),
        ("max_contacts", C.c_uint32),
        ("n", C.c_uint64),
        ("n_moving", C.c_uint64),
        ("lo", C.c_double * 3),
        ("cs", C.c_double),
        ("dims", C.c_int32 * 3),
        ("pad0", C.c_int32),
        ("g", C.c_double * 3),
        ("dt", C.c_double),
        ("h", C.c_double),
        ("rho0", C.c_double),
        ("c0", C.c_double),
        ("alpha", C.c_double),
        ("mass", C.c_double),
        ("kn", C.c_double),
        ("kt", C.c_double),
        ("e", C.c_double),
        ("mu", C.c_double),
        ("wall_lo", C.c_double * 3),
        ("wall_hi", C.c_double * 3),
        ("wall_mask", C.c_uint32),
        ("pad1", C.c_uint32),
    ]


_lib = None
_D = C.POINTER(C.c_double)
_U32 = C.POINTER(C.c_uint32)


def build() -> Path:
    """Compile liboracle.so with the committed Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def bind(L):
    """Declare the C signatures of an oracle library handle (ctypes.CDLL)."""
    P = C.POINTER(Params)
    L.orc_cell_key.restype = C.c_int64
    L.orc_cell_key.argtypes = [P, C.c_double, C.c_double, C.c_double]
    L.orc_stable_sort.argtypes = [C.c_uint64, _U32, C.c_uint32, _U32]
    L.orc_cell_lb.argtypes = [C.c_uint64, _U32, C.c_uint32, _U32]
    L.orc_spline_f.restype = C.c_double
    L.orc_spline_f.argtypes = [C.c_double]
    L.orc_spline_df.restype = C.c_double
    L.orc_spline_df.argtypes = [C.c_double]
    L.orc_eq1_allpairs.argtypes = [C.c_uint64, _D, _D]
    L.orc_create.restype = C.c_void_p
    L.orc_create.argtypes = [P, _D, _D, _D, _D, _D, _D, _D, _D]
    L.orc_destroy.argtypes = [C.c_void_p]
    L.orc_build.restype = C.c_int
    L.orc_build.argtypes = [C.c_void_p]
    L.orc_neighbors.argtypes = [C.c_void_p, _U32, _U32, _U32, _U32]
    L.orc_eq1_cells.argtypes = [C.c_void_p, C.c_double, _D, _D]
    L.orc_density.restype = C.c_int
    L.orc_density.argtypes = [C.c_void_p]
    L.orc_step.restype = C.c_int
    L.orc_step.argtypes = [C.c_void_p, C.c_uint32]
    L.orc_set_threads.restype = C.c_int
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_get.restype = C.c_int
    L.orc_get.argtypes = [C.c_void_p, C.c_char_p, _D]
    L.orc_contact_count.restype = C.c_uint64
    L.orc_contact_count.argtypes = [C.c_void_p]
    L.orc_cs_tangent.restype = C.c_int
    L.orc_cs_tangent.argtypes = [P, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, C.c_uint32, _D, _D, _D, _D]
    return L


def lib():
    """The oracle built from this directory (compiled on first use)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = bind(C.CDLL(str(LIB_PATH)))
    return _lib


def load(path) -> C.CDLL:
    """Another build of the oracle sources (a test that plants a mistake compiles one), bound but not
    made the default."""
    return bind(C.CDLL(str(path)))


def _d(a):
    return None if a is None else np.ascontiguousarray(a, np.float64).ctypes.data_as(_D)


def params_from_desc(desc) -> Params:
    """Oracle parameters from a prestige_amd.Desc (grid dims as the library computes them)."""
    p = Params()
    p.model = desc.model
    p.max_contacts = desc.max_contacts
    p.n = desc.n_particles
    p.n_moving = desc.n_moving
    p.cs = desc.cell_size
    for a in range(3):
        p.lo[a] = desc.grid_lo[a]
        p.dims[a] = int(math.ceil((desc.grid_hi[a] - desc.grid_lo[a]) / desc.cell_size))
        p.g[a] = desc.gravity[a]
        p.wall_lo[a] = desc.wall_lo[a]
        p.wall_hi[a] = desc.wall_hi[a]
    p.dt = desc.dt
    p.h, p.rho0, p.c0, p.alpha, p.mass = desc.h, desc.rho0, desc.c0, desc.alpha, desc.mass
    p.kn, p.kt, p.e, p.mu = desc.kn, desc.kt, desc.restitution, desc.mu
    p.wall_mask = desc.wall_mask
    return p


class State:
    """FP64 oracle state (orc_state*), on the default oracle or on the library handle L."""

    def __init__(self, params: Params, arrays: dict, L=None):
        L = L or lib()
        self.L = L
        self.p = params
        self.n = int(params.n)
        self._keep = {k: np.ascontiguousarray(v, np.float64) for k, v in arrays.items() if v is not None}
        g = lambda k: _d(self._keep.get(k))  # noqa: E731
        self._h = L.orc_create(C.byref(params), g("x"), g("y"), g("z"), g("vx"), g("vy"), g("vz"),
                               g("radius"), g("mass"))
        if not self._h:
            raise ValueError("orc_create failed")

    def close(self):
        if self._h:
            self.L.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build(self) -> int:
        return self.L.orc_build(self._h)

    def neighbors(self) -> dict:
        nc = int(self.p.dims[0]) * int(self.p.dims[1]) * int(self.p.dims[2])
        keys = np.empty(self.n, np.uint32)
        perm = np.empty(self.n, np.uint32)
        ids = np.empty(self.n, np.uint32)
        lb = np.empty(nc + 1, np.uint32)
        u = lambda a: a.ctypes.data_as(_U32)  # noqa: E731
        self.L.orc_neighbors(self._h, u(keys), u(perm), u(ids), u(lb))
        return {"keys": keys, "perm": perm, "ids": ids, "lb": lb}

    def eq1_cells(self, cutoff: float, force, mass) -> np.ndarray:
        f = np.array(force, np.float64, copy=True)
        m = np.ascontiguousarray(mass, np.float64)
        self.L.orc_eq1_cells(self._h, cutoff, _d(f), _d(m))
        return f

    def density(self) -> int:
        return self.L.orc_density(self._h)

    def step(self, nsteps: int = 1) -> int:
        return self.L.orc_step(self._h, int(nsteps))

    def get(self, name: str) -> np.ndarray:
        out = np.empty(self.n, np.float64)
        st = self.L.orc_get(self._h, name.encode(), _d(out))
        if st:
            raise KeyError(name)
        return out

    def contact_count(self) -> int:
        return int(self.L.orc_contact_count(self._h))


def set_threads(n: int) -> int:
    """OpenMP threads of later oracle calls; returns the previous count."""
    return int(lib().orc_set_threads(int(n)))


def eq1_allpairs(force, mass) -> np.ndarray:
    f = np.array(force, np.float64, copy=True)
    m = np.ascontiguousarray(mass, np.float64)
    lib().orc_eq1_allpairs(len(f), _d(f), _d(m))
    return f


def cs_tangent(params: Params, arrays: dict, dparam, dv0, nsteps: int):
    """Complex-step forward-mode tangent of the final (x, v) for a direction in (params, v0).

    Returns (x_T, v_T, dx_T, dv_T), each (3, n) in id order."""
    n = int(params.n)
    keep = {k: np.ascontiguousarray(v, np.float64) for k, v in arrays.items() if v is not None}
    g = lambda k: _d(keep.get(k))  # noqa: E731
    dp = None if dparam is None else np.ascontiguousarray(dparam, np.float64)
    dv = None if dv0 is None else np.ascontiguousarray(np.asarray(dv0, np.float64).reshape(-1))
    xo, vo, tx, tv = (np.empty(3 * n) for _ in range(4))
    st = lib().orc_cs_tangent(C.byref(params), g("x"), g("y"), g("z"), g("vx"), g("vy"), g("vz"),
                              g("radius"), g("mass"), _d(dp), _d(dv), int(nsteps), _d(xo), _d(vo), _d(tx),
                              _d(tv))
    if st:
        raise RuntimeError(f"orc_cs_tangent status {st}")
    return xo.reshape(3, n), vo.reshape(3, n), tx.reshape(3, n), tv.reshape(3, n)
