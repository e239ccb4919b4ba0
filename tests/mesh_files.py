"""MSH 2.2 / 4.1 writers and surface meshes of boxes and polygonal tanks for
the boundary-import tests (written here: no Gmsh in the test environment)."""
from __future__ import annotations

import numpy as np


def write_msh22(path, nodes, elems, newline="\n"):
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$PhysicalNames", "1", '2 1 "wall"',
             "$EndPhysicalNames", "$Nodes", str(len(nodes))]
    lines += [f"{t} {float(x)!r} {float(y)!r} {float(z)!r}" for t, (x, y, z) in nodes.items()]
    lines += ["$EndNodes", "$Elements", str(len(elems))]
    lines += [f"{k + 1} {typ} 2 1 1 " + " ".join(map(str, nd)) for k, (typ, nd) in enumerate(elems)]
    lines += ["$EndElements"]
    path.write_text(newline.join(lines) + newline)


def write_msh41(path, nodes, elems):
    tags = list(nodes)
    lines = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat", "$Entities", "0 0 1 0",
             "1 0 0 0 1 1 1 0 0", "$EndEntities", "$Nodes", f"1 {len(tags)} {min(tags)} {max(tags)}",
             f"2 1 0 {len(tags)}"]
    lines += [str(t) for t in tags]
    lines += [" ".join(repr(float(v)) for v in nodes[t]) for t in tags]
    lines += ["$EndNodes"]
    blocks = {}
    for typ, nd in elems:
        blocks.setdefault(typ, []).append(nd)
    lines += ["$Elements", f"{len(blocks)} {len(elems)} 1 {len(elems)}"]
    k = 1
    for typ, lst in blocks.items():
        lines.append(f"2 1 {typ} {len(lst)}")
        for nd in lst:
            lines.append(f"{k} " + " ".join(map(str, nd)))
            k += 1
    lines += ["$EndElements"]
    path.write_text("\n".join(lines) + "\n")


def box_surface(lo, hi, breaks=None, open_top=False, inward=False, quads=True):
    """Surface of the box [lo, hi] meshed on the tensor grid of breaks (per
    axis, fractions of the side, default the two ends); face normals outward
    (inward=True flips them); quadrangles or triangles with alternating
    diagonals.  Corner nodes get the smallest tags."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    br = breaks or [[0.0, 1.0]] * 3
    coord = [lo[a] + np.asarray(br[a]) * (hi[a] - lo[a]) for a in range(3)]
    nodes, tag = {}, {}

    def node(idx):
        if idx not in tag:
            tag[idx] = None
        return idx

    faces = []
    for a in range(3):
        b, c = (a + 1) % 3, (a + 2) % 3
        for end in (0, len(coord[a]) - 1):
            if open_top and a == 2 and end > 0:
                continue
            out = 1 if end > 0 else -1
            for ib in range(len(coord[b]) - 1):
                for ic in range(len(coord[c]) - 1):
                    quad = []
                    for db, dc in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        idx = [0, 0, 0]
                        idx[a], idx[b], idx[c] = end, ib + db, ic + dc
                        quad.append(node(tuple(idx)))
                    if (out < 0) != inward:
                        quad = quad[::-1]
                    faces.append((quad, (ib + ic) % 2))
    # corners first, then every other node, in a fixed order
    last = [len(coord[a]) - 1 for a in range(3)]
    keys = sorted(tag, key=lambda i: (not all(i[a] in (0, last[a]) for a in range(3)), i))
    for t, k in enumerate(keys, 1):
        tag[k] = t
        nodes[t] = tuple(float(coord[a][k[a]]) for a in range(3))
    elems = []
    for quad, alt in faces:
        q = [tag[k] for k in quad]
        if quads:
            elems.append((3, q))
        elif alt:
            elems += [(2, [q[1], q[2], q[3]]), (2, [q[1], q[3], q[0]])]
        else:
            elems += [(2, [q[0], q[1], q[2]]), (2, [q[0], q[2], q[3]])]
    return nodes, elems


def prism_tank(nseg, radius, height):
    """Open-top regular nseg-gon tank, normals into the tank, bottom fanned
    from its centre, side quads split into triangles."""
    nodes = {1: (0.0, 0.0, 0.0)}
    ang = 2 * np.pi * np.arange(nseg) / nseg
    for k in range(nseg):
        nodes[2 + k] = (radius * np.cos(ang[k]), radius * np.sin(ang[k]), 0.0)
        nodes[2 + nseg + k] = (radius * np.cos(ang[k]), radius * np.sin(ang[k]), height)
    elems = []
    for k in range(nseg):
        a, b = 2 + k, 2 + (k + 1) % nseg
        elems.append((2, [1, a, b]))                         # bottom: normal +z (into the tank)
        at, bt = a + nseg, b + nseg
        elems += [(2, [a, at, bt]), (2, [a, bt, b])]         # side: normal towards the axis
    return nodes, elems


