"""Writes tests/golden/mesh_vectors.npz from the reference's own test
generators (tests/utils.py:3-97 of the reference), so that the project's
generators (cholespy_amd.workloads) are pinned against them without the
reference being present at test time. Run in a container where the
reference is mounted:

    python tests/golden/make_mesh_vectors.py /root/reference

Only the generated arrays are stored (faces and COO arrays), no code.
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]


def main(ref_root: str):
    sys.path.insert(0, str(Path(ref_root) / "tests"))
    import utils  # the reference's test generators

    out = {}
    n, faces = utils.get_cube()
    out["cube_n"] = n
    out["cube_faces"] = faces
    v, idx = utils.get_coo_arrays(n, faces, 2.0)
    out["cube_coo_values"], out["cube_coo_idx"] = v, idx
    for level in range(4):
        n, faces = utils.get_icosphere(level)
        out[f"ico{level}_n"] = n
        out[f"ico{level}_faces"] = faces
    n, faces = utils.get_icosphere(2)
    v, idx = utils.get_coo_arrays(n, faces, 2.0)
    out["ico2_coo_values"], out["ico2_coo_idx"] = v, idx
    np.savez_compressed(ROOT / "tests" / "golden" / "mesh_vectors.npz", **out)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
