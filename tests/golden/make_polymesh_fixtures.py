"""Writes the polyMesh fixtures of the real-mesh tests: two meshes shipped
with the reference's tutorials, read with foamhip.polymesh (the reader under
test) and stored as compressed numpy arrays, so the tests need no access to
the reference tree.

    python tests/golden/make_polymesh_fixtures.py [/root/reference]

incompressibleFluid/airFoil2D (2-D, quadrilateral faces, an empty
front-and-back patch) and incompressibleDriftFlux/tank3D (3-D, mixed
triangle and quadrilateral faces, gzipped files, a neighbour list carrying
-1 for the boundary faces).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "openfoam-dev_amd"))
from foamhip.polymesh import PolyMesh  # noqa: E402

CASES = {"airfoil2d": "tutorials/incompressibleFluid/airFoil2D",
         "tank3d": "tutorials/incompressibleDriftFlux/tank3D"}


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    for name, rel in CASES.items():
        m = PolyMesh(os.path.join(ref, rel))
        out = os.path.join(HERE, f"polymesh_{name}.npz")
        np.savez_compressed(
            out, points=m.points, faceStart=m.faceStart, facePoints=m.facePoints,
            owner=m.owner, neighbour=m.neighbour,
            patchName=np.array([p[0] for p in m.patches]),
            patchType=np.array([p[1] for p in m.patches]),
            patchStart=np.array([p[2] for p in m.patches], dtype=np.int32),
            patchSize=np.array([p[3] for p in m.patches], dtype=np.int32))
        print(f"{out}: {m.nCells} cells, {m.nFaces} faces ({m.nInternalFaces} internal), "
              f"{len(m.patches)} patches, {os.path.getsize(out)} bytes")


if __name__ == "__main__":
    main()
