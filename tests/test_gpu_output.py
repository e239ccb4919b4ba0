"""Particle output: the legacy VTK and CSV files hold exactly the downloaded
fields (id order, FP64 round trip)."""
from __future__ import annotations

import numpy as np
import pytest

import prestige_amd as pa
from helpers import make_scene, upload_all

pytestmark = pytest.mark.gpu


def _read_vtk(path):
    """Minimal reader of the legacy binary POLYDATA the library writes."""
    raw = path.read_bytes()
    pos = 0

    def line():
        nonlocal pos
        e = raw.index(b"\n", pos)
        s = raw[pos:e].decode()
        pos = e + 1
        return s

    def block(nbytes, dt):
        nonlocal pos
        a = np.frombuffer(raw[pos:pos + nbytes], dtype=dt).copy()
        pos += nbytes
        return a

    assert line().startswith("# vtk DataFile")
    title = line()
    assert line() == "BINARY" and line() == "DATASET POLYDATA"
    assert line() == "FIELD FieldData 1" and line() == "TIME 1 1 double"
    time = block(8, ">f8")[0]
    line()
    n = int(line().split()[1])
    pts = block(24 * n, ">f8").reshape(n, 3)
    line()
    assert line().split()[0] == "VERTICES"
    cells = block(8 * n, ">i4").reshape(n, 2)
    line()
    assert line() == f"POINT_DATA {n}"
    fields = {}
    while pos < len(raw):
        name = line().split()[1]
        assert line() == "LOOKUP_TABLE default"
        fields[name] = block(8 * n, ">f8")
        line()
    return title, time, pts, cells, fields


def test_vtk_and_csv_round_trip(tmp_path):
    d, arrs = make_scene(pa.SCENE_DAM_BREAK, n_target=2000, dx=0.01)
    names = ("vx", "vz", "rho", "p")
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        ctx.step(3)
        ref = {k: ctx.download(k) for k in ("x", "y", "z") + names}
        ctx.write_particles(tmp_path / "a.vtk", pa.OUT_VTK, names, time=3 * d.dt)
        ctx.write_particles(tmp_path / "a.csv", pa.OUT_CSV, names, time=3 * d.dt)
    title, t, pts, cells, fields = _read_vtk(tmp_path / "a.vtk")
    assert t == 3 * d.dt
    assert np.array_equal(pts, np.stack([ref["x"], ref["y"], ref["z"]], 1))
    assert np.array_equal(cells[:, 0], np.ones(len(pts))) and np.array_equal(cells[:, 1], np.arange(len(pts)))
    for k in names:
        assert np.array_equal(fields[k], ref[k]), k
    text = (tmp_path / "a.csv").read_text().splitlines()
    assert float(text[0].split()[2]) == 3 * d.dt
    assert text[1] == "id,x,y,z," + ",".join(names)
    tab = np.loadtxt(tmp_path / "a.csv", delimiter=",", skiprows=2)
    assert np.array_equal(tab[:, 0], np.arange(len(pts)))
    for c, k in enumerate(("x", "y", "z") + names):
        assert np.array_equal(tab[:, c + 1], ref[k]), k
