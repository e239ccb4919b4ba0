"""Equation front end and HIP code generator on the host: the macro's
read/write rule and the i/j roles (prestige_macros/src/lib.rs:71-133), the
pass schedule, the generated text, the constructs without a translation, and
that every generated program compiles for gfx950 (hiprtc, no device needed)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import pytest

import prestige_amd as pa

I, J = pa.ROLE_I, pa.ROLE_J


def test_inference_reproduces_the_builtin_sets():
    """Each built-in's read/write sets and roles were written out by hand from
    the macro's rule; inferring them from the body text gives the same sets."""
    for eq in range(6):
        ref = pa.equation_describe(eq)
        new = pa.equation_describe(pa.define_equation(ref["name"] + "_again", ref["body"]))
        assert new["reads"] == ref["reads"], ref["name"]
        assert new["writes"] == ref["writes"], ref["name"]
        assert new["roles"] == ref["roles"], ref["name"]


def test_read_write_rule_and_roles():
    e = pa.equation_describe(pa.define_equation(
        "rule", "{ a[i] = b[j] + c[i]; d[i] -= 1.0; e[i] *= 2.0; if f[i] == g[j] { h2[i] /= 3.0 } }"))
    assert e["reads"] == ["a", "b", "c", "d", "e", "f", "g", "h2"]
    assert e["writes"] == ["a", "d", "e", "h2"]   # == is a comparison, not a write
    assert e["roles"] == {"a": I, "b": J, "c": I, "d": I, "e": I, "f": I, "g": J, "h2": I}


def test_only_indexed_paths_count_and_only_their_first_segment():
    """The macro records an indexed expression only when it is a path, and then
    path.segments[0] (prestige_macros/src/lib.rs:75-79, 105-109, 125-129): a
    field `self.x[i]` is neither read nor written, `crate::m::x[i]` records
    `crate`, and of `x[i][j]` only the inner path counts, as a read."""
    e = pa.equation_describe(pa.define_equation(
        "paths", "{ self.x[i] += m[j]; crate::m::y[i] = w[j]; z[i][j] += 1.0; ::q[i] -= 2.0; }"))
    assert e["reads"] == ["crate", "m", "q", "w", "z"]
    assert e["writes"] == ["crate", "q"]
    with pytest.raises(pa.PrestigeError) as ex:
        pa.generate_hip([pa.define_equation("fieldref", "{ self.x[i] += m[j]; }")], pa.LOOP_CELLS)
    assert ex.value.name == "E_UNSUPPORTED"


def test_schedule():
    # built-in WCSPH list: density + EOS, then momentum + integrate
    sph = [pa.EQ_SPH_DENSITY, pa.EQ_SPH_EOS, pa.EQ_SPH_MOMENTUM, pa.EQ_INTEGRATE]
    assert pa.fuse_schedule(sph) == ([0, 0, 1, 1], 2)
    a = pa.define_equation("acc", "{ s[i] += mass[j]; }")
    b = pa.define_equation("dbl", "{ m[i] = 2.0 * s[i]; }")            # i-only
    c = pa.define_equation("use_j", "{ q[i] += m[j]; }")               # reads m at [j]
    d = pa.define_equation("use_i", "{ r[i] += m[i] * mass[j]; }")     # reads m at [i]
    assert pa.fuse_schedule([a, b]) == ([0, 0], 1)
    assert pa.fuse_schedule([a, b, c]) == ([0, 0, 1], 2)
    # an i-only body runs after the j loop, so a j body reading its output moves on
    assert pa.fuse_schedule([b, d]) == ([0, 1], 2)
    assert pa.fuse_schedule([d, b]) == ([0, 0], 1)


def test_generated_text():
    cells = pa.generate_hip([pa.EQ1], pa.LOOP_CELLS)
    assert 'extern "C" __global__ void __launch_bounds__(64, 8) prestige_pass_0_eq1(' in cells
    assert "force_i += g_mass[js]" in cells and "const double* __restrict__ g_mass" in cells
    assert "< c2" in cells and "lb[row" in cells and "s_q[kk] = q" in cells
    pairs = pa.generate_hip([pa.EQ1], pa.LOOP_ALLPAIRS)
    assert "for (unsigned jd = 0; jd < n; ++jd)" in pairs and "force_i += f_mass[jd]" in pairs
    e = pa.define_equation("methods", "{ let d = (x[i] - x[j]).abs().sqrt(); w[i] += f64::exp(-d) * d.powi(2) }")
    t = pa.generate_hip([e], pa.LOOP_CELLS)
    assert "sqrt ( fabs ( (" in t and "exp ( - d )" in t and "pow ( d , 2 )" in t
    assert "double d =" in t


def test_tuples_and_if_expressions():
    """Tuple values (a, b), their fields t.0, destructuring lets and if expressions (chains included)
    translate: tuples to pt_make / pt_get, if expressions to lambdas returning their branches' values."""
    e = pa.define_equation("tup", "{ let (a, b) = (x[i] - x[j], y[i] - y[j]); "
                                  "let s = if a > 0.0 { let t = a * 2.0; t + b } else if b > 0.0 { b } else { 0.0 }; "
                                  "let p = (s, s * s); w[i] += p.1 - p.0; }")
    t = pa.generate_hip([e], pa.LOOP_CELLS)
    assert "auto __t0 = pt_make (" in t and "double a = pt_get < 0 > ( __t0 )" in t
    assert "( [ & ] ( ) {" in t and "return t + b ;" in t and "else if" in t
    assert "pt_get < 1 > ( p )" in t and "auto p = pt_make ( s , s * s )" in t
    m = pa.generate_hip([pa.EQ_SPH_MOMENTUM], pa.LOOP_CELLS)
    assert "auto f = momentum (" in m and ", prestige_k )" in m and "pt_get < 2 > ( f )" in m


@pytest.mark.parametrize("body,what", [
    ("{ let f = g(x[i]); a[i] += f . 0 ; }", "unknown name"),
    ("{ a[j] += 1.0; }", "through [i] only"),
    ("{ x[i] += 1.0; }", "position"),
    ("{ let t = if x[i] > 0.0 { 1.0 }; a[i] += t; }", "without else"),
    ("{ let t = if x[i] > 0.0 { a[i] += 1.0; } else { 2.0 }; }", "without a value"),
    ("{ for k in 0..3 { a[i] += 1.0; } }", "translation"),
    ("{ a[i] += undefined_name; }", "unknown name"),
    ("{ a[i + 1] += 1.0; }", "other than i or j"),
])
def test_unsupported_constructs(body, what):
    e = pa.define_equation("bad", body)
    with pytest.raises(pa.PrestigeError) as ei:
        pa.generate_hip([e], pa.LOOP_CELLS)
    assert ei.value.name == "E_UNSUPPORTED" and what in str(ei.value)


@pytest.mark.parametrize("name,body", [("9lives", "{ }"), ("ok", "{ a[i] += 1.0; } }"), ("ok", "a[i] += 1.0;")])
def test_define_rejects(name, body):
    with pytest.raises(pa.PrestigeError) as ei:
        pa.define_equation(name, body)
    assert ei.value.name == "E_INVALID"


def _hiprtc():
    path = Path("/opt/rocm/lib/libhiprtc.so")
    if not path.exists():
        pytest.skip("hiprtc is not installed")
    lib = C.CDLL(str(path))
    lib.hiprtcCreateProgram.argtypes = [C.POINTER(C.c_void_p), C.c_char_p, C.c_char_p, C.c_int, C.c_void_p,
                                        C.c_void_p]
    lib.hiprtcCompileProgram.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_char_p)]
    lib.hiprtcGetProgramLogSize.argtypes = [C.c_void_p, C.POINTER(C.c_size_t)]
    lib.hiprtcGetProgramLog.argtypes = [C.c_void_p, C.c_char_p]
    lib.hiprtcDestroyProgram.argtypes = [C.POINTER(C.c_void_p)]
    return lib


def _compile(lib, src: str):
    prog = C.c_void_p()
    assert lib.hiprtcCreateProgram(C.byref(prog), src.encode(), b"gen.hip", 0, None, None) == 0
    opts = [b"--offload-arch=gfx950", b"-O3", b"-std=c++17", b"-ffp-contract=off", b"-Xclang",
            b"-target-feature", b"-Xclang", b"-packed-fp32-ops"]
    arr = (C.c_char_p * len(opts))(*opts)
    r = lib.hiprtcCompileProgram(prog, len(opts), arr)
    n = C.c_size_t()
    lib.hiprtcGetProgramLogSize(prog, C.byref(n))
    log = C.create_string_buffer(n.value + 1)
    lib.hiprtcGetProgramLog(prog, log)
    lib.hiprtcDestroyProgram(C.byref(prog))
    return r, log.value.decode(errors="replace")


def test_generated_programs_compile_for_gfx950():
    lib = _hiprtc()
    m = pa.define_equation("moment", "{ let d = x[i] - x[j]; s[i] += mass[j] * d * d; c[i] += 1.0 }")
    t = pa.define_equation("twice", "{ m[i] = 2.0 * s[i]; }")
    q = pa.define_equation("nb", "{ if i != j { q[i] += m[j] * (y[i] - y[j]).abs().sqrt() } }")
    sm = pa.define_equation("smooth", "{ a[i] += 0.5 * a[j] * kernel_w(x[i] - x[j], y[i] - y[j], z[i] - z[j], h) }")
    tp = pa.define_equation("tup2", "{ let (a, b) = (x[i] - x[j], y[i] - y[j]); "
                                    "let s = if a > 0.0 { let t = a * 2.0; t + b } else if b > 0.0 { b } else { 0.0 }; "
                                    "w[i] += s; }")
    lists = [[pa.EQ1], [pa.EQ_SPH_DENSITY, pa.EQ_SPH_EOS], [m, t, q], [sm], [pa.EQ_SPH_MOMENTUM],
             [pa.EQ_DEM_CONTACT], [tp]]
    for eqs in lists:
        for loop in (pa.LOOP_ALLPAIRS, pa.LOOP_CELLS):
            r, log = _compile(lib, pa.generate_hip(eqs, loop))
            assert r == 0, log
