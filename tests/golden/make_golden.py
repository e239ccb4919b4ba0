"""Write tests/golden/oracle_vectors.npz: small FP64 golden vectors produced by
the CPU oracle (oracle/), which the reference cannot produce (it has no
physics and cannot be built here).  Inputs are the deterministic scenes of
prestige_amd.scene; outputs are oracle fields in particle-id order.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
import prestige_amd as pa  # noqa: E402

OUT = Path(__file__).resolve().parent / "oracle_vectors.npz"

CASES = {
    # name: (scene, scene kwargs, steps, fields)
    "dam": (pa.SCENE_DAM_BREAK, dict(n_target=1000, dx=0.01), 3, ("rho", "p", "vz", "z", "az")),
    "dem": (pa.SCENE_DEM_LATTICE, dict(n_target=2000, seed=0), 50, ("x", "vx", "vz", "wy")),
    "pour_build": (pa.SCENE_DEM_POUR, dict(n_target=4000, seed=5), 0, ()),
}


def generate() -> dict:
    out = {}
    for name, (scene, kw, steps, fields) in CASES.items():
        d, arrs = pa.scene(scene, **kw)
        st = oracle.State(oracle.params_from_desc(d), arrs)
        if steps:
            assert st.step(steps) == 0
            for f in fields:
                out[f"{name}/{f}"] = st.get(f)
        else:
            assert st.build() == 0
            nb = st.neighbors()
            out[f"{name}/keys"] = nb["keys"]
            out[f"{name}/perm"] = nb["perm"]
    return out


def main():
    data = generate()
    np.savez_compressed(OUT, **data)
    h = hashlib.sha256(OUT.read_bytes()).hexdigest()
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes) sha256 {h}")


if __name__ == "__main__":
    main()
