"""Gmsh surface meshes as boundary particles (host only).

Pinned by closed forms: a box whose sides are multiples of dx samples to
exactly the lattice points of its surface, whatever its triangulation or
file format, and three layers behind an open box tank sampled at the
dam-break spacing give exactly the dam-break scene's wall particles.  For a
polygonal (non-box) tank: every layer-0 point lies on the mesh, no two points
are closer than dx/2, no surface point is farther than dx from a sample, and
layer l lies on the face planes moved l*dx behind the surface."""
from __future__ import annotations

import numpy as np
import pytest

import prestige_amd as pa
from mesh_files import box_surface, prism_tank, write_msh22, write_msh41


def lattice_index(pts, dx, shift):
    f = pts / dx - shift
    idx = np.rint(f)
    assert np.max(np.abs(f - idx)) < 1e-6
    return set(map(tuple, idx.astype(np.int64)))


# -- boxes ------------------------------------------------------------------------
IRREGULAR = [[0.0, 0.37, 0.61, 1.0], [0.0, 0.5, 1.0], [0.0, 0.2, 0.9, 1.0]]


@pytest.mark.parametrize("fmt,breaks,quads", [("2.2", None, True), ("2.2", None, False),
                                              ("4.1", IRREGULAR, False), ("4.1", IRREGULAR, True)])
def test_cube_is_its_surface_lattice(tmp_path, fmt, breaks, quads):
    m, dx = 12, 0.01
    nodes, elems = box_surface((0.0, 0.0, 0.0), (m * dx,) * 3, breaks, quads=quads)
    path = tmp_path / "cube.msh"
    (write_msh22 if fmt == "2.2" else write_msh41)(path, nodes, elems)
    pts = pa.mesh_boundary(path, dx)
    assert len(pts) == 6 * m * m + 2
    g = np.arange(m + 1)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    on = (np.minimum.reduce([X, Y, Z]) == 0) | (np.maximum.reduce([X, Y, Z]) == m)
    want = set(zip(X[on], Y[on], Z[on]))
    assert lattice_index(pts, dx, 0.0) == want


def test_cube_layers_inward(tmp_path):
    """Outward normals, layers against them: nested shells inside the cube."""
    m, dx, nl = 10, 0.02, 3
    nodes, elems = box_surface((0.0, 0.0, 0.0), (m * dx,) * 3)
    path = tmp_path / "cube.msh"
    write_msh22(path, nodes, elems, newline="\r\n")
    pts = pa.mesh_boundary(path, dx, layers=nl, side=-1)
    g = np.arange(m + 1)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    depth = np.minimum.reduce([X, Y, Z, m - X, m - Y, m - Z])
    sel = depth < nl
    want = set(zip(X[sel], Y[sel], Z[sel]))
    assert len(pts) == len(want)
    assert lattice_index(pts, dx, 0.0) == want
    # layer order: the first 6 m^2 + 2 points are the surface itself
    assert np.all(depth[tuple(np.rint(pts[:6 * m * m + 2] / dx).astype(int).T)] == 0)


@pytest.mark.parametrize("breaks", [None, IRREGULAR])
def test_tank_mesh_gives_the_dam_break_walls(tmp_path, breaks):
    d, arrs = pa.scene(pa.SCENE_DAM_BREAK, n_target=1000, dx=0.005)
    dx = 0.005
    walls = np.stack([arrs[k][d.n_moving:] for k in "xyz"], 1)
    hi_idx = np.rint(walls / dx - 0.5).max(0).astype(int)   # (nix + 2, niy + 2, niz - 1)
    nix, niy, niz = hi_idx[0] - 2, hi_idx[1] - 2, hi_idx[2] + 1
    lo = (-0.5 * dx,) * 3
    hi = ((nix + 0.5) * dx, (niy + 0.5) * dx, (niz - 0.5) * dx)
    nodes, elems = box_surface(lo, hi, breaks, open_top=True, inward=True, quads=breaks is None)
    path = tmp_path / "tank.msh"
    (write_msh22 if breaks is None else write_msh41)(path, nodes, elems)
    pts = pa.mesh_boundary(path, dx, layers=3, side=-1)
    assert len(pts) == len(walls)
    assert lattice_index(pts, dx, 0.5) == lattice_index(walls, dx, 0.5)


# -- a polygonal tank ---------------------------------------------------------------
def point_triangle_distance(p, a, b, c):
    """Distance from points p (n, 3) to triangle abc (Ericson, Real-Time
    Collision Detection, 5.1.5: closest point by Voronoi region)."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ap @ ab, ap @ ac
    bp = p - b
    d3, d4 = bp @ ab, bp @ ac
    cp = p - c
    d5, d6 = cp @ ab, cp @ ac
    va = d3 * d6 - d5 * d4
    vb = d5 * d2 - d1 * d6
    vc = d1 * d4 - d3 * d2
    denom = va + vb + vc
    v = np.where(denom != 0, vb / np.where(denom != 0, denom, 1), 0)
    w = np.where(denom != 0, vc / np.where(denom != 0, denom, 1), 0)
    q = a + v[:, None] * ab + w[:, None] * ac                # interior
    # edges and vertices
    def seg(s, t):
        st = t - s
        u = np.clip(((p - s) @ st) / (st @ st), 0, 1)
        return s + u[:, None] * st
    cand = [q, seg(a, b), seg(b, c), seg(c, a)]
    inside = (va >= 0) & (vb >= 0) & (vc >= 0)
    dist = [np.linalg.norm(p - x, axis=1) for x in cand]
    dist[0] = np.where(inside, dist[0], np.inf)
    return np.minimum.reduce(dist)


def mesh_distance(p, nodes, elems):
    out = np.full(len(p), np.inf)
    for _, nd in elems:
        a, b, c = (np.asarray(nodes[t]) for t in nd[:3])
        out = np.minimum(out, point_triangle_distance(p, a, b, c))
    return out


def test_polygonal_tank(tmp_path):
    dx, nseg, radius, height = 0.01, 12, 0.2, 0.15
    nodes, elems = prism_tank(nseg, radius, height)
    path = tmp_path / "prism.msh"
    write_msh41(path, nodes, elems)
    n0 = len(pa.mesh_boundary(path, dx))
    pts = pa.mesh_boundary(path, dx, layers=3, side=-1)
    surf = pts[:n0]
    assert np.max(mesh_distance(surf, nodes, elems)) < 1e-12
    # no duplicates or near-duplicates
    from scipy.spatial import cKDTree
    dmin, _ = cKDTree(surf).query(surf, k=2)
    assert dmin[:, 1].min() >= 0.5 * dx * (1 - 1e-9)
    # coverage: random surface points (area-weighted triangles) have a sample within dx
    rng = np.random.default_rng(0)
    tri = np.array([[nodes[t] for t in nd] for _, nd in elems])
    area = 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1)
    pick = rng.choice(len(tri), 4000, p=area / area.sum())
    r1, r2 = rng.random(4000), rng.random(4000)
    s = np.sqrt(r1)
    q = (1 - s)[:, None] * tri[pick, 0] + (s * (1 - r2))[:, None] * tri[pick, 1] + (s * r2)[:, None] * tri[pick, 2]
    gap, _ = cKDTree(surf).query(q)
    assert gap.max() <= dx
    # point density close to one per dx^2
    assert abs(n0 / (area.sum() / dx**2) - 1.0) < 0.15
    # layer l lies on the tank's face planes moved l*dx outwards, and no
    # closer than l*dx to the tank
    tri_n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    tri_n /= np.linalg.norm(tri_n, axis=1)[:, None]
    n1 = len(pa.mesh_boundary(path, dx, layers=2, side=-1))
    for lay, sub in ((1, pts[n0:n1]), (2, pts[n1:])):
        behind = -np.einsum("fk,pfk->pf", tri_n, sub[:, None, :] - tri[None, :, 0])
        assert np.min(np.abs(behind - lay * dx), axis=1).max() < 1e-12
        assert mesh_distance(sub, nodes, elems).min() >= lay * dx * (1 - 1e-9)


# -- errors -----------------------------------------------------------------------
def _status(fn):
    with pytest.raises(pa.PrestigeError) as ei:
        fn()
    return ei.value.name


def test_errors(tmp_path):
    nodes, elems = box_surface((0, 0, 0), (1, 1, 1))
    good = tmp_path / "good.msh"
    write_msh22(good, nodes, elems)
    assert _status(lambda: pa.mesh_boundary(tmp_path / "missing.msh", 0.1)) == "E_INVALID"
    assert _status(lambda: pa.mesh_boundary(good, 0.0)) == "E_INVALID"
    assert _status(lambda: pa.mesh_boundary(good, 0.1, layers=0)) == "E_INVALID"
    assert _status(lambda: pa.mesh_boundary(good, 0.1, side=0)) == "E_INVALID"
    binary = tmp_path / "bin.msh"
    binary.write_text("$MeshFormat\n4.1 1 8\n$EndMeshFormat\n")
    assert _status(lambda: pa.mesh_boundary(binary, 0.1)) == "E_UNSUPPORTED"
    v3 = tmp_path / "v3.msh"
    v3.write_text("$MeshFormat\n3.0 0 8\n$EndMeshFormat\n")
    assert _status(lambda: pa.mesh_boundary(v3, 0.1)) == "E_UNSUPPORTED"
    dangling = tmp_path / "dangling.msh"
    write_msh22(dangling, nodes, elems + [(2, [1, 2, 99])])
    assert _status(lambda: pa.mesh_boundary(dangling, 0.1)) == "E_INVALID"
    lines_only = tmp_path / "lines.msh"
    write_msh22(lines_only, {1: (0, 0, 0), 2: (1, 0, 0)}, [(1, [1, 2]), (15, [1])])
    assert _status(lambda: pa.mesh_boundary(lines_only, 0.1)) == "E_INVALID"
    # too small a caller array
    import ctypes as C
    lib = pa.load_library()
    n = C.c_uint64()
    buf = (C.c_double * 4)()
    assert lib.prestige_mesh_boundary(str(good).encode(), 0.5, 1, -1, None, None, None, 0, C.byref(n)) == 0
    assert n.value == 6 * 4 + 2
    assert lib.prestige_mesh_boundary(str(good).encode(), 0.5, 1, -1, buf, buf, buf, 4, C.byref(n)) == 1
    assert b"smaller" in lib.prestige_last_error(None)


def test_second_order_and_mixed_elements(tmp_path):
    """6-node triangles and 9-node quadrangles use their corner nodes; line,
    point and tetrahedron elements are skipped."""
    m, dx = 4, 0.25
    nodes, elems = box_surface((0, 0, 0), (1, 1, 1), quads=True)
    mid = max(nodes) + 1
    nodes[mid] = (0.5, 0.5, 0.5)   # stands in for the higher-order nodes
    full = []
    for k, (_, q) in enumerate(elems):
        if k % 2:
            full += [(9, [q[0], q[1], q[2]] + [mid] * 3), (9, [q[0], q[2], q[3]] + [mid] * 3)]
        else:
            full.append((10, q + [mid] * 5))
    full += [(1, [1, 2]), (15, [3]), (4, [1, 2, 3, mid])]
    path = tmp_path / "mixed.msh"
    write_msh41(path, nodes, full)
    pts = pa.mesh_boundary(path, dx)
    assert len(pts) == 6 * m * m + 2
    assert lattice_index(pts, dx, 0.0) == lattice_index(pa.mesh_boundary(_cube(tmp_path), dx), dx, 0.0)


def _cube(tmp_path):
    nodes, elems = box_surface((0, 0, 0), (1, 1, 1))
    path = tmp_path / "unit.msh"
    write_msh22(path, nodes, elems)
    return path
