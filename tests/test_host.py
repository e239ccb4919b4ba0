"""Host-side tests (no device): the C-ABI library loads and exports every symbol
of include/prestige_hip.h, the equation interface mirrors the reference's
(prestige/src/equations/ir.rs:3-14, fuse.rs:14-40, codegen/simple_cpu.rs:3-22),
the scenes are the frozen workloads, and descriptor validation fails loudly."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import prestige_amd as pa

ROOT = Path(__file__).resolve().parents[1]


def _declared_symbols():
    text = (ROOT / "include" / "prestige_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(prestige_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(pa.LIB_PATH))
    syms = _declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_density_f64_flag_is_validated():
    """PRESTIGE_FLAG_DENSITY_F64 belongs to the MIXED WCSPH form; other forms and unknown bits are
    refused before any device work."""
    lib = pa.load_library()
    for model, prec, flags in ((pa.MODEL_WCSPH, pa.F64, 1), (pa.MODEL_DEM, pa.MIXED, 1),
                               (pa.MODEL_WCSPH, pa.MIXED, 2)):
        d, _ = pa.scene(pa.SCENE_DAM_BREAK, n_target=1000, dx=0.01)
        d.model, d.precision, d.flags = model, prec, flags
        h = ctypes.c_void_p()
        assert lib.prestige_create(ctypes.byref(d), ctypes.byref(h)) == 1
        assert b"FLAG" in lib.prestige_last_error(None).upper()


def test_desc_layout_matches_header():
    # offsets of a few fields pin the ctypes mirror to the C struct layout
    assert pa.Desc.n_particles.offset == 16
    assert pa.Desc.grid_lo.offset == 32
    assert pa.Desc.wall_mask.offset == ctypes.sizeof(pa.Desc) - 8
    d = pa.desc_init(pa.MODEL_DEM)
    assert d.abi_version == 1 and d.gravity[2] == -9.81 and abs(d.kt - 2 / 7 * d.kn) < 1e-9


# ---------------------------------------------------------------------------
# the Rust -sys block of INTEGRATION.md against the header
# ---------------------------------------------------------------------------

_C_SCALARS = {"int32_t": "i32", "uint32_t": "u32", "uint64_t": "u64", "uint8_t": "u8", "double": "f64",
              "size_t": "usize", "char": "c_char", "void": "c_void", "prestige_status": "prestige_status",
              "prestige_ctx": "prestige_ctx", "prestige_desc": "prestige_desc",
              "prestige_scene_params": "prestige_scene_params"}


def _c_type_to_rust(decl: str) -> str:
    """Rust spelling of one C parameter or return type (name dropped; T x[3] is a pointer)."""
    decl = " ".join(decl.replace("*", " * ").split())
    array = bool(re.search(r"\[\d*\]$", decl))
    decl = re.sub(r"\s*\[\d*\]$", "", decl)
    toks = decl.split()
    # drop the parameter name: the last identifier when more than the type is present
    if toks and toks[-1] not in ("*",) and toks[-1] not in _C_SCALARS and len(toks) > 1:
        toks = toks[:-1]
    if array:
        toks.append("*")
    # C pointer declarators read right to left: "const char * const *" -> *const *const c_char
    base_const = toks[0] == "const"
    if base_const:
        toks = toks[1:]
    rust = _C_SCALARS[toks[0]]
    const_next = base_const
    for t in toks[1:]:
        if t == "const":
            const_next = True
            continue
        assert t == "*", decl
        rust = ("*const " if const_next else "*mut ") + rust
        const_next = False
    return rust


def _header_functions():
    text = (ROOT / "include" / "prestige_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = []
    for ret, name, params in re.findall(r"\n([A-Za-z_][\w \*]*?)\b(prestige_\w+)\s*\(([^;{]*?)\)\s*;", text, re.S):
        ps = [p.strip() for p in params.split(",") if p.strip() and p.strip() != "void"]
        r = _c_type_to_rust(ret.strip() + " x")
        out.append((name, [_c_type_to_rust(p) for p in ps], None if r == "c_void" else r))
    return out


def _header_struct(name):
    text = (ROOT / "include" / "prestige_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    body = re.search(r"typedef struct " + name + r" \{(.*?)\} " + name + ";", text, re.S).group(1)
    fields = []
    for line in body.split(";"):
        line = " ".join(line.split())
        if not line:
            continue
        m = re.match(r"(\w+) (\w+)(\[(\d+)\])?$", line)
        assert m, line
        t = _C_SCALARS[m.group(1)]
        fields.append((m.group(2), f"[{t}; {m.group(4)}]" if m.group(3) else t))
    return fields


def _rust_block():
    text = (ROOT / "INTEGRATION.md").read_text()
    blocks = re.findall(r"```rust\n(.*?)```", text, re.S)
    return next(b for b in blocks if 'extern "C"' in b)


def _rust_functions(block):
    ext = block[block.index('extern "C" {'):]
    out = []
    for name, params, ret in re.findall(r"pub fn (prestige_\w+)\((.*?)\)\s*(?:->\s*([^;]+))?;", ext, re.S):
        ps = [" ".join(p.split(":", 1)[1].split()) for p in params.split(",") if p.strip()]
        out.append((name, ps, " ".join(ret.split()) if ret else None))
    return out


def _rust_struct(block, name):
    body = re.search(r"pub struct " + name + r" \{(.*?)\}", block, re.S).group(1)
    return [(m.group(1), " ".join(m.group(2).split())) for m in re.finditer(r"pub (\w+): ([^,]+),", body)]


def test_rust_sys_block_matches_header():
    """Every function of the header, in order, with the same parameter and return types, and the
    #[repr(C)] structs with the same fields in the same order and types."""
    block = _rust_block()
    c_fns = _header_functions()
    r_fns = _rust_functions(block)
    assert [f[0] for f in r_fns] == [f[0] for f in c_fns]
    for (cn, cp, cr), (rn, rp, rr) in zip(c_fns, r_fns):
        assert rp == cp, f"{cn}: parameters {rp} != {cp}"
        assert rr == cr, f"{cn}: return {rr} != {cr}"
    for st in ("prestige_desc", "prestige_scene_params"):
        assert _rust_struct(block, st) == _header_struct(st), st


def test_c_to_rust_type_mapping():
    assert _c_type_to_rust("const char* const* files") == "*const *const c_char"
    assert _c_type_to_rust("const char** name") == "*mut *const c_char"
    assert _c_type_to_rust("prestige_ctx** out") == "*mut *mut prestige_ctx"
    assert _c_type_to_rust("const double grid_lo[3]") == "*const f64"
    assert _c_type_to_rust("uint32_t dims[3]") == "*mut u32"
    assert _c_type_to_rust("void* stream") == "*mut c_void"
    assert _c_type_to_rust("uint64_t n") == "u64"


# ---------------------------------------------------------------------------
# equation interface
# ---------------------------------------------------------------------------

def test_eq1_record_matches_macro_inference():
    """#[equation] on eq1 (prestige/src/lib.rs:7-12): reads {force, mass}, writes {force}.

    prestige_macros/src/lib.rs:123-133 records every indexed name as a read,
    the LHS included; :88-120 records compound-assignment targets as writes."""
    e = pa.equation_describe(pa.EQ1)
    assert e["name"] == "eq1"
    assert set(e["reads"]) == {"force", "mass"}
    assert e["writes"] == ["force"]
    assert e["roles"] == {"force": pa.ROLE_I, "mass": pa.ROLE_J}
    assert e["body"] == "{ force [i] += mass [j] ; }"


def test_builtin_records_are_consistent():
    for eq in range(6):
        e = pa.equation_describe(eq)
        assert e["reads"] == sorted(e["reads"])
        # every written slice is also a read (LHS index), as the macro infers
        assert set(e["writes"]) <= set(e["reads"]), e["name"]
        # every name indexed in the body is listed, with the roles the body uses
        idx = set(re.findall(r"\b([a-z_][a-z0-9_]*) \[(i|j)\]", e["body"]))
        names = {n for n, _ in idx}
        assert names == set(e["reads"]), (e["name"], names ^ set(e["reads"]))
        for n, r in idx:
            bit = pa.ROLE_I if r == "i" else pa.ROLE_J
            assert e["roles"][n] & bit, (e["name"], n, r)


def test_generate_simple_cpu_text():
    """Layout of prestige/src/codegen/simple_cpu.rs:7-19."""
    code = pa.generate_simple_cpu([pa.EQ1])
    assert code == "for i in 0..n {\n    for j in 0..n {\n        { force [i] += mass [j] ; }\n    }\n}\n"
    two = pa.generate_simple_cpu([pa.EQ_SPH_DENSITY, pa.EQ_SPH_EOS])
    lines = two.splitlines()
    assert lines[0] == "for i in 0..n {" and lines[1] == "    for j in 0..n {"
    assert lines[2].startswith("        { rho [i] +=") and lines[3].startswith("        { p [i] =")
    assert lines[4:] == ["    }", "}"]
    assert pa.generate_simple_cpu([]) == "for i in 0..n {\n    for j in 0..n {\n    }\n}\n"


def test_fuse_schedule_splits_at_j_hazards():
    # density + EOS fuse (EOS reads rho at [i] only); momentum reads rho[j], p[j] -> new pass;
    # integrate reads only [i] and writes x, v which momentum reads at [j]: double-buffered, same pass
    passes, n = pa.fuse_schedule([pa.EQ_SPH_DENSITY, pa.EQ_SPH_EOS, pa.EQ_SPH_MOMENTUM, pa.EQ_INTEGRATE])
    assert passes == [0, 0, 1, 1] and n == 2
    assert pa.fuse_schedule([pa.EQ1, pa.EQ1]) == ([0, 0], 1)   # force[j] is never read
    assert pa.fuse_schedule([pa.EQ_DEM_CONTACT, pa.EQ_INTEGRATE]) == ([0, 0], 1)
    assert pa.fuse_schedule([pa.EQ_INTEGRATE, pa.EQ_SPH_DENSITY]) == ([0, 1], 2)   # x[j] after x[i] write
    assert pa.fuse_schedule([]) == ([], 0)
    with pytest.raises(pa.PrestigeError):
        pa.fuse_schedule([99])


# ---------------------------------------------------------------------------
# scenes
# ---------------------------------------------------------------------------

def test_dam_break_counts_and_layout():
    d, a = pa.scene(pa.SCENE_DAM_BREAK, n_target=1_000_000, dx=0.005)
    assert d.n_moving == 1_000_000 and d.n_particles == 1_799_320
    assert abs(d.h - 0.006) < 1e-15 and abs(d.cell_size - 0.012) < 1e-15
    assert abs(d.c0 - 10 * np.sqrt(2 * 9.81 * 0.5)) < 1e-12
    assert abs(d.dt - 0.25 * d.h / d.c0) < 1e-18
    nf = 1_000_000
    x, z = a["x"][:nf], a["z"][:nf]
    assert x.min() == 0.0025 and abs(x.max() - 0.4975) < 1e-12 and z.min() == 0.0025
    wall = np.stack([a["x"][nf:], a["y"][nf:], a["z"][nf:]])
    inside = (wall[0] > 0) & (wall[0] < 1.6) & (wall[1] > 0) & (wall[1] < 0.67) & (wall[2] > 0)
    assert not inside.any()
    assert np.all(a["vx"] == 0)
    d4, _ = pa.scene(pa.SCENE_DAM_BREAK, n_target=4_000_000, dx=0.00315)
    assert d4.n_moving == 159**3


def test_dem_scenes_are_deterministic_and_overlap_free():
    d, a = pa.scene(pa.SCENE_DEM_LATTICE, n_target=10_000, seed=0)
    d2, a2 = pa.scene(pa.SCENE_DEM_LATTICE, n_target=10_000, seed=0)
    assert all(np.array_equal(a[k], a2[k]) for k in a)
    _, a3 = pa.scene(pa.SCENE_DEM_LATTICE, n_target=10_000, seed=1)
    assert not np.array_equal(a["x"], a3["x"])
    assert d.n_particles == 10_000 and d.wall_mask == 0x3F
    assert np.all(np.abs(a["vx"]) <= 0.1) and np.all(a["radius"] == 0.01)
    assert abs(a["mass"][0] - 2500 * 4 / 3 * np.pi * 1e-6) < 1e-15
    p = np.stack([a["x"], a["y"], a["z"]], 1)
    # lattice spacing 0.021 with +-0.5 mm jitter: nearest neighbours never overlap
    for s in range(0, 10_000, 2000):
        d2m = ((p[s:s + 2000, None] - p[None]) ** 2).sum(-1)
        np.fill_diagonal(d2m[:, s:s + 2000], 1.0)
        assert d2m.min() >= (0.02) ** 2 - 1e-15
    dp, ap = pa.scene(pa.SCENE_DEM_POUR, n_target=100_000, seed=0)
    assert dp.n_particles == 100_000 and dp.wall_mask == 0x1F
    assert ap["radius"].min() >= 0.008 and ap["radius"].max() <= 0.012
    assert dp.cell_size == 0.024


# ---------------------------------------------------------------------------
# validation and the no-fallback rule
# ---------------------------------------------------------------------------

def _small_desc():
    d, _ = pa.scene(pa.SCENE_DAM_BREAK, n_target=1000, dx=0.01)
    return d


@pytest.mark.parametrize("field,value", [("abi_version", 99), ("model", 7), ("precision", 5),
                                         ("cell_size", 0.0), ("dt", 0.0)])
def test_create_rejects_bad_descriptors(field, value):
    d = _small_desc()
    setattr(d, field, value)
    with pytest.raises(pa.PrestigeError) as ei:
        pa.Context(d)
    assert ei.value.name == "E_INVALID"


def test_create_rejects_slot_counts_beyond_the_index_width():
    # slots are 32-bit indices with one pad value; the mixed WCSPH density walk
    # also keeps a run tag in bit 31
    for prec, n_bad in ((pa.F64, 0xFFFFFFFF), (pa.MIXED, 0x80000000)):
        d = _small_desc()
        d.precision = prec
        d.n_particles = n_bad
        d.n_moving = 0
        with pytest.raises(pa.PrestigeError) as ei:
            pa.Context(d)
        assert ei.value.name == "E_INVALID"


def test_create_rejects_cell_smaller_than_support():
    d = _small_desc()
    d.cell_size = 1.9 * d.h
    with pytest.raises(pa.PrestigeError):
        pa.Context(d)


def test_no_cpu_fallback_without_device():
    """Without a HIP device the library reports an error instead of computing on the CPU."""
    from conftest import gpu_available

    if gpu_available():
        pytest.skip("a device is present")
    with pytest.raises(pa.PrestigeError) as ei:
        pa.Context(_small_desc())
    assert ei.value.name in ("E_HIP", "E_OOM")


def test_write_series_collection(tmp_path):
    """The time-series collection lists each particle file at its time."""
    import xml.etree.ElementTree as ET
    p = tmp_path / "run.pvd"
    pa.write_series(p, ["out_0000.vtk", "out_0001.vtk"], [0.0, 2.5e-4])
    root = ET.parse(p).getroot()
    ds = root.find("Collection").findall("DataSet")
    assert [(d.get("file"), float(d.get("timestep"))) for d in ds] == [("out_0000.vtk", 0.0), ("out_0001.vtk", 2.5e-4)]
