"""Golden vectors (tests/golden/oracle_vectors.npz): oracle regression on the CPU,
bit-exact device parity on the GPU."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

import prestige_amd as pa
from helpers import upload_all

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
import make_golden  # noqa: E402

GOLD = np.load(Path(__file__).resolve().parent / "golden" / "oracle_vectors.npz")


def test_oracle_reproduces_golden_vectors():
    fresh = make_golden.generate()
    assert set(fresh) == set(GOLD.files)
    for k in GOLD.files:
        assert np.array_equal(fresh[k], GOLD[k]), k


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["dam", "dem"])
def test_device_reproduces_golden_fields(name):
    scene, kw, steps, fields = make_golden.CASES[name]
    d, arrs = pa.scene(scene, **kw)
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        ctx.step(steps)
        for f in fields:
            assert np.array_equal(ctx.download(f), GOLD[f"{name}/{f}"]), f


@pytest.mark.gpu
def test_device_reproduces_golden_build():
    scene, kw, _, _ = make_golden.CASES["pour_build"]
    d, arrs = pa.scene(scene, **kw)
    with pa.Context(d) as ctx:
        upload_all(ctx, arrs)
        ctx.neighbors_build()
        nb = ctx.neighbors_download()
    assert np.array_equal(nb["keys"], GOLD["pour_build/keys"])
    assert np.array_equal(nb["perm"], GOLD["pour_build/perm"])
