"""prestige_amd — MI355X (gfx950) particle hot path behind the prestige equation interface.

Python host binding of libprestige_hip.so (ctypes over the C ABI in
include/prestige_hip.h).  The binding mirrors the reference's interface for
this path: particle data as one float64 slice per named field indexed by
particle id (prestige/src/lib.rs:8), built-in equations described by
name/reads/writes/body (prestige/src/equations/ir.rs:3-14), fusion
(prestige/src/equations/fuse.rs:14-40) and the generated loop text
(prestige/src/codegen/simple_cpu.rs:3-22), which the library executes on the
GPU instead of printing.

There is no CPU fallback: every compute call goes to the HIP library, and
importing the binding fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

__all__ = [
    "Desc", "Context", "PrestigeError", "load_library", "scene", "equation_describe",
    "fuse_schedule", "generate_simple_cpu", "define_equation", "generate_hip", "write_series", "mesh_boundary", "OUT_VTK",
    "OUT_CSV", "MODEL_PAIRS", "MODEL_WCSPH", "MODEL_DEM",
    "F64", "F32", "MIXED", "LOOP_ALLPAIRS", "LOOP_CELLS", "EQ1", "EQ_SPH_DENSITY", "EQ_SPH_EOS",
    "EQ_SPH_MOMENTUM", "EQ_INTEGRATE", "EQ_DEM_CONTACT", "SCENE_DEM_LATTICE", "SCENE_DAM_BREAK",
    "SCENE_DEM_POUR", "ROLE_I", "ROLE_J",
]

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libprestige_hip.so"

ABI_VERSION = 1
MODEL_PAIRS, MODEL_WCSPH, MODEL_DEM = 0, 1, 2
OUT_VTK, OUT_CSV = 0, 1
F64, F32, MIXED = 0, 1, 2
LOOP_ALLPAIRS, LOOP_CELLS = 0, 1
EQ1, EQ_SPH_DENSITY, EQ_SPH_EOS, EQ_SPH_MOMENTUM, EQ_INTEGRATE, EQ_DEM_CONTACT = range(6)
SCENE_DEM_LATTICE, SCENE_DAM_BREAK, SCENE_DEM_POUR = 0, 1, 2
ROLE_I, ROLE_J = 1, 2
FLAG_DENSITY_F64, FLAG_DEM_EVENTS = 1, 2

STATUS = {
    0: "OK", 1: "E_INVALID", 2: "E_HIP", 3: "E_OOM", 4: "E_CAPACITY", 5: "E_NONFINITE",
    6: "E_DOMAIN", 7: "E_UNSUPPORTED", 8: "E_STATE", 9: "E_CFL",
}


class PrestigeError(RuntimeError):
    """A non-zero prestige_status; .status holds the code, .name its symbol."""

    def __init__(self, status: int, message: str):
        self.status = status
        self.name = STATUS.get(status, str(status))
        super().__init__(f"{self.name}: {message}")


class Desc(C.Structure):
    """ctypes mirror of prestige_desc."""

    _fields_ = [
        ("abi_version", C.c_uint32),
        ("model", C.c_int32),
        ("precision", C.c_int32),
        ("max_contacts", C.c_uint32),
        ("n_particles", C.c_uint64),
        ("n_moving", C.c_uint64),
        ("grid_lo", C.c_double * 3),
        ("grid_hi", C.c_double * 3),
        ("cell_size", C.c_double),
        ("gravity", C.c_double * 3),
        ("dt", C.c_double),
        ("h", C.c_double),
        ("rho0", C.c_double),
        ("c0", C.c_double),
        ("alpha", C.c_double),
        ("mass", C.c_double),
        ("kn", C.c_double),
        ("kt", C.c_double),
        ("restitution", C.c_double),
        ("mu", C.c_double),
        ("wall_lo", C.c_double * 3),
        ("wall_hi", C.c_double * 3),
        ("wall_mask", C.c_uint32),
        ("flags", C.c_uint32),
    ]

    def copy(self) -> "Desc":
        d = Desc()
        C.memmove(C.byref(d), C.byref(self), C.sizeof(Desc))
        return d


class SceneParams(C.Structure):
    _fields_ = [("dx", C.c_double), ("n_target", C.c_uint64), ("seed", C.c_uint64)]


_lib = None

_P = C.c_void_p
_D = C.POINTER(C.c_double)
_U32 = C.POINTER(C.c_uint32)


def _bind(lib):
    S = C.c_int
    sig = {
        "prestige_desc_init": (None, [C.POINTER(Desc), C.c_int32]),
        "prestige_create": (S, [C.POINTER(Desc), C.POINTER(_P)]),
        "prestige_destroy": (None, [_P]),
        "prestige_field_upload": (S, [_P, C.c_char_p, _D, C.c_uint64]),
        "prestige_field_download": (S, [_P, C.c_char_p, _D, C.c_uint64]),
        "prestige_neighbors_build": (S, [_P]),
        "prestige_neighbors_download": (S, [_P, _U32, _U32, _U32, _U32]),
        "prestige_grid_dims": (S, [_P, _U32]),
        "prestige_eval_equation": (S, [_P, C.c_int32, C.c_int32, C.c_double]),
        "prestige_eq1_cells_noalloc": (S, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                           _D, _D, C.c_double, C.c_double, C.c_void_p]),
        "prestige_step": (S, [_P, C.c_uint32]),
        "prestige_forward_tape": (S, [_P, C.c_uint32, C.c_uint32]),
        "prestige_adjoint": (S, [_P, _D, _D, _D, C.c_uint32, _D]),
        "prestige_timing_enable": (S, [_P, C.c_int32]),
        "prestige_timing_count": (C.c_uint32, [_P]),
        "prestige_timing_get": (S, [_P, C.c_uint32, C.POINTER(C.c_char_p), _D, C.POINTER(C.c_uint64), _D]),
        "prestige_sync": (S, [_P]),
        "prestige_last_error": (C.c_char_p, [_P]),
        "prestige_equation_describe": (S, [C.c_int32, C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t,
                                           C.c_char_p, C.c_size_t, C.POINTER(C.c_uint8), C.c_size_t,
                                           C.c_char_p, C.c_size_t]),
        "prestige_fuse_schedule": (S, [C.POINTER(C.c_int32), C.c_uint32, _U32, _U32]),
        "prestige_generate_simple_cpu": (S, [C.POINTER(C.c_int32), C.c_uint32, C.c_char_p, C.c_size_t,
                                             C.POINTER(C.c_size_t)]),
        "prestige_equation_define": (S, [C.c_char_p, C.c_char_p, C.POINTER(C.c_int32)]),
        "prestige_generate_hip": (S, [C.POINTER(C.c_int32), C.c_uint32, C.c_int32, C.c_char_p, C.c_size_t,
                                      C.POINTER(C.c_size_t)]),
        "prestige_run_equations": (S, [_P, C.POINTER(C.c_int32), C.c_uint32, C.c_int32, C.c_double]),
        "prestige_write_particles": (S, [_P, C.c_char_p, C.c_int32, C.c_char_p, C.c_double]),
        "prestige_write_series": (S, [C.c_char_p, C.POINTER(C.c_char_p), _D, C.c_uint32]),
        "prestige_mesh_boundary": (S, [C.c_char_p, C.c_double, C.c_uint32, C.c_int32, _D, _D, _D, C.c_uint64,
                                       C.POINTER(C.c_uint64)]),
        "prestige_scene_count": (S, [C.c_int32, C.POINTER(SceneParams), C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_uint64)]),
        "prestige_scene_build": (S, [C.c_int32, C.POINTER(SceneParams), C.POINTER(Desc), _D, _D, _D, _D, _D,
                                     _D, _D, _D]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def load_library(path: str | os.PathLike | None = None):
    """Load libprestige_hip.so (built in-tree by prestige_amd.build)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise ImportError(f"{p} is missing: build it with `python -m prestige_amd.build` (hipcc, gfx950)")
    lib = _bind(C.CDLL(str(p)))
    if path is None:
        _lib = lib
    return lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_D)


def _check(status: int, ctx_handle=None):
    if status != 0:
        msg = load_library().prestige_last_error(ctx_handle)
        raise PrestigeError(status, (msg or b"").decode(errors="replace"))


def desc_init(model: int) -> Desc:
    d = Desc()
    load_library().prestige_desc_init(C.byref(d), model)
    return d


class Context:
    """One device context (prestige_ctx*): owns device buffers and a HIP stream."""

    FIELDS = ("x", "y", "z", "vx", "vy", "vz")

    def __init__(self, desc: Desc):
        lib = load_library()
        self._lib = lib
        self.desc = desc.copy()
        h = _P()
        _check(lib.prestige_create(C.byref(self.desc), C.byref(h)), None)
        self._h = h
        self.n = int(desc.n_particles)

    # -- lifetime ------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.prestige_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, status):
        _check(status, self._h)

    # -- fields --------------------------------------------------------------
    def upload(self, name: str, data) -> None:
        a = np.ascontiguousarray(data, dtype=np.float64)
        if a.shape != (self.n,):
            raise PrestigeError(1, f"field '{name}' must have shape ({self.n},), got {a.shape}")
        self._c(self._lib.prestige_field_upload(self._h, name.encode(), _dptr(a), self.n))

    def download(self, name: str) -> np.ndarray:
        out = np.empty(self.n, dtype=np.float64)
        self._c(self._lib.prestige_field_download(self._h, name.encode(), _dptr(out), self.n))
        return out

    def upload_state(self, arrays: dict) -> None:
        for k, v in arrays.items():
            if v is not None:
                self.upload(k, v)

    # -- neighbour search ----------------------------------------------------
    def neighbors_build(self) -> None:
        self._c(self._lib.prestige_neighbors_build(self._h))

    def grid_dims(self):
        d = (C.c_uint32 * 3)()
        self._c(self._lib.prestige_grid_dims(self._h, d))
        return tuple(int(v) for v in d)

    def neighbors_download(self) -> dict:
        nx, ny, nz = self.grid_dims()
        keys = np.empty(self.n, np.uint32)
        perm = np.empty(self.n, np.uint32)
        ids = np.empty(self.n, np.uint32)
        lb = np.empty(nx * ny * nz + 1, np.uint32)
        u = lambda a: a.ctypes.data_as(_U32)  # noqa: E731
        self._c(self._lib.prestige_neighbors_download(self._h, u(keys), u(perm), u(ids), u(lb)))
        return {"keys": keys, "perm": perm, "ids": ids, "lb": lb}

    # -- compute -------------------------------------------------------------
    def eval_equation(self, eq_id: int, loop: int, cutoff: float = 0.0) -> None:
        self._c(self._lib.prestige_eval_equation(self._h, eq_id, loop, float(cutoff)))

    def write_particles(self, path, fmt: int = 0, fields=(), time: float = 0.0) -> None:
        """Write positions and fields (id order) as legacy binary VTK (fmt OUT_VTK) or CSV (OUT_CSV)."""
        self._c(self._lib.prestige_write_particles(self._h, str(path).encode(), fmt, ",".join(fields).encode(),
                                                   float(time)))

    def run_equations(self, eq_ids, loop: int, cutoff: float = 0.0) -> None:
        """Compile (once per context) and run the generated gather passes of the fused equations."""
        ids = (C.c_int32 * max(1, len(eq_ids)))(*eq_ids)
        self._c(self._lib.prestige_run_equations(self._h, ids, len(eq_ids), loop, float(cutoff)))

    def step(self, nsteps: int = 1) -> None:
        self._c(self._lib.prestige_step(self._h, int(nsteps)))

    def forward_tape(self, nsteps: int, checkpoint_every: int) -> None:
        self._c(self._lib.prestige_forward_tape(self._h, int(nsteps), int(checkpoint_every)))

    def adjoint(self, seed_x=None, seed_v=None, n_params: int | None = None, want_v0: bool = True):
        """Reverse sweep; seeds are (3, n) or 3n arrays in id order.  Returns (dL/dparams, dL/dv0)."""
        if n_params is None:
            n_params = 3 if self.desc.model == MODEL_DEM else 2
        sx = None if seed_x is None else np.ascontiguousarray(np.asarray(seed_x, np.float64).reshape(-1))
        sv = None if seed_v is None else np.ascontiguousarray(np.asarray(seed_v, np.float64).reshape(-1))
        for s in (sx, sv):
            if s is not None and s.shape != (3 * self.n,):
                raise PrestigeError(1, "seeds must hold 3*n values")
        gp = np.zeros(n_params, np.float64)
        gv = np.zeros(3 * self.n, np.float64) if want_v0 else None
        self._c(self._lib.prestige_adjoint(self._h, None if sx is None else _dptr(sx),
                                           None if sv is None else _dptr(sv), _dptr(gp), n_params,
                                           None if gv is None else _dptr(gv)))
        return gp, (None if gv is None else gv.reshape(3, self.n))

    def sync(self) -> None:
        self._c(self._lib.prestige_sync(self._h))

    # -- timing --------------------------------------------------------------
    def timing_enable(self, on: bool = True) -> None:
        self._c(self._lib.prestige_timing_enable(self._h, 1 if on else 0))

    def timing(self) -> list:
        k = self._lib.prestige_timing_count(self._h)
        out = []
        for i in range(k):
            name = C.c_char_p()
            ms = C.c_double()
            launches = C.c_uint64()
            b = C.c_double()
            self._c(self._lib.prestige_timing_get(self._h, i, C.byref(name), C.byref(ms), C.byref(launches),
                                                  C.byref(b)))
            out.append({"kernel": name.value.decode(), "total_ms": ms.value, "launches": launches.value,
                        "bytes_per_launch": b.value})
        return out


# -- scenes --------------------------------------------------------------------
def scene(scene_id: int, n_target: int = 0, dx: float = 0.0, seed: int = 0):
    """Build a benchmark scene: returns (Desc, dict of id-order float64 arrays)."""
    lib = load_library()
    sp = SceneParams(dx, n_target, seed)
    nt = C.c_uint64()
    nm = C.c_uint64()
    _check(lib.prestige_scene_count(scene_id, C.byref(sp), C.byref(nt), C.byref(nm)))
    n = nt.value
    arrs = {k: np.zeros(n, np.float64) for k in ("x", "y", "z", "vx", "vy", "vz", "radius", "mass")}
    d = Desc()
    _check(lib.prestige_scene_build(scene_id, C.byref(sp), C.byref(d), *[_dptr(arrs[k]) for k in
                                    ("x", "y", "z", "vx", "vy", "vz", "radius", "mass")]))
    if d.model != MODEL_DEM:
        arrs.pop("radius")
        arrs.pop("mass")
    return d, arrs


# -- equation interface ----------------------------------------------------------
def equation_describe(eq_id: int) -> dict:
    lib = load_library()
    bufs = [C.create_string_buffer(4096) for _ in range(4)]
    roles = (C.c_uint8 * 64)()
    _check(lib.prestige_equation_describe(eq_id, bufs[0], 4096, bufs[1], 4096, bufs[2], 4096, roles, 64,
                                          bufs[3], 4096))
    reads = [r for r in bufs[1].value.decode().split(",") if r]
    writes = [w for w in bufs[2].value.decode().split(",") if w]
    return {"name": bufs[0].value.decode(), "reads": reads, "writes": writes,
            "roles": {r: int(roles[k]) for k, r in enumerate(reads)}, "body": bufs[3].value.decode()}


def fuse_schedule(eq_ids) -> tuple[list, int]:
    lib = load_library()
    ids = (C.c_int32 * max(1, len(eq_ids)))(*eq_ids)
    out = (C.c_uint32 * max(1, len(eq_ids)))()
    npass = C.c_uint32()
    _check(lib.prestige_fuse_schedule(ids, len(eq_ids), out, C.byref(npass)))
    return [int(out[k]) for k in range(len(eq_ids))], int(npass.value)


def define_equation(name: str, body: str) -> int:
    """Register an equation body (the block #[equation] captures); returns its id."""
    lib = load_library()
    out = C.c_int32()
    _check(lib.prestige_equation_define(name.encode(), body.encode(), C.byref(out)))
    return int(out.value)


def generate_hip(eq_ids, loop: int) -> str:
    """HIP source of the fused equations' gather passes."""
    lib = load_library()
    ids = (C.c_int32 * max(1, len(eq_ids)))(*eq_ids)
    n = C.c_size_t()
    _check(lib.prestige_generate_hip(ids, len(eq_ids), loop, None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    _check(lib.prestige_generate_hip(ids, len(eq_ids), loop, buf, n.value + 1, C.byref(n)))
    return buf.value.decode()


def write_series(path, files, times) -> None:
    """ParaView collection (.pvd) listing files[k] at times[k]."""
    lib = load_library()
    fs = (C.c_char_p * max(1, len(files)))(*[str(f).encode() for f in files])
    ts = np.ascontiguousarray(times, dtype=np.float64)
    _check(lib.prestige_write_series(str(path).encode(), fs, _dptr(ts) if len(ts) else None, len(files)))


def mesh_boundary(path, dx: float, layers: int = 1, side: int = -1) -> np.ndarray:
    """Boundary particles sampled from the surface of a Gmsh .msh file at
    spacing dx: an (n, 3) float64 array, layer 0 first, later layers l*dx
    against (side=-1) or along (side=+1) the face normals."""
    lib = load_library()
    n = C.c_uint64()
    _check(lib.prestige_mesh_boundary(str(path).encode(), dx, layers, side, None, None, None, 0, C.byref(n)))
    xyz = [np.zeros(max(1, n.value), np.float64) for _ in range(3)]
    _check(lib.prestige_mesh_boundary(str(path).encode(), dx, layers, side, *[_dptr(a) for a in xyz], n.value,
                                      C.byref(n)))
    return np.stack([a[:n.value] for a in xyz], 1)


def generate_simple_cpu(eq_ids) -> str:
    lib = load_library()
    ids = (C.c_int32 * max(1, len(eq_ids)))(*eq_ids)
    n = C.c_size_t()
    _check(lib.prestige_generate_simple_cpu(ids, len(eq_ids), None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    _check(lib.prestige_generate_simple_cpu(ids, len(eq_ids), buf, n.value + 1, C.byref(n)))
    return buf.value.decode()
