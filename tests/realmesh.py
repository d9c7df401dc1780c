"""The real meshes of the reference's tutorials (fixtures written by
tests/golden/make_polymesh_fixtures.py) and a Laplacian system on them:
fixedValue 0 on the outlet patches, zeroGradient elsewhere (empty patches
add nothing to the assembly either), a smooth diffusivity of the cell
centres and a seeded right-hand side."""
import os

import numpy as np

from foamhip.polymesh import PolyMesh

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MESHES = ("airfoil2d", "tank3d")


def load(name):
    z = np.load(os.path.join(GOLDEN, f"polymesh_{name}.npz"), allow_pickle=False)
    patches = list(zip(z["patchName"].tolist(), z["patchType"].tolist(),
                       z["patchStart"].tolist(), z["patchSize"].tolist()))
    return PolyMesh.from_arrays(z["points"], z["faceStart"], z["facePoints"], z["owner"],
                                z["neighbour"], patches)


def boundary_types(m):
    """Per boundary face: 1 (fixedValue 0) on patches named outlet / OUTL*,
    0 (zeroGradient) elsewhere; values 0."""
    fixed = np.array([1 if n.lower().startswith(("outlet", "outl")) else 0
                      for n, _, _, _ in m.patches], dtype=np.int32)
    bType = fixed[m.patch_of_boundary_faces()]
    return bType, np.zeros(bType.size)


def gamma_of(C):
    """A smooth positive diffusivity of the cell centres C [N, 3]."""
    x, y, z = C[:, 0], C[:, 1], C[:, 2]
    return 1.0 + 0.5*np.sin(0.7*x)*np.cos(0.9*y) + 0.25*np.sin(1.3*z)


def rhs(n, seed):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n)
