"""The polyMesh reader (foamhip/polymesh.py) and the oracle on the real
meshes of the reference's tutorials (airFoil2D, tank3D; CPU)."""
import gzip
import os

import numpy as np
import pytest

import oracle as O
from foamhip.polymesh import PolyMesh
from foamhip.cases import GradedBoxCase
from realmesh import MESHES, boundary_types, gamma_of, load, rhs

HEADER = """FoamFile
{{
    format      ascii;
    class       {cls};
    object      {obj};
}}
// comment line
"""


def _write(d, obj, cls, body, gz=False):
    text = HEADER.format(cls=cls, obj=obj) + body + "\n/* trailing */\n"
    if gz:
        with gzip.open(os.path.join(d, obj + ".gz"), "wt") as f:
            f.write(text)
    else:
        with open(os.path.join(d, obj), "w") as f:
            f.write(text)


def _write_mesh(d, c, gz, neighbour_all, uniform_owner_patch):
    nF = c.faceStart.size - 1
    P = c.points.reshape(-1, 3)
    pts = "\n".join(f"({float(x)!r} {float(y)!r} {float(z)!r})" for x, y, z in P)
    _write(d, "points", "vectorField", f"{len(P)}\n(\n{pts}\n)", gz)
    faces = "\n".join(f"{e - s}(" + " ".join(str(p) for p in c.facePoints[s:e]) + ")"
                      for s, e in zip(c.faceStart[:-1], c.faceStart[1:]))
    _write(d, "faces", "faceList", f"{nF}\n(\n{faces}\n)", gz)
    owner = np.concatenate([c.lower, c.bFaceCells])
    _write(d, "owner", "labelList", f"{nF}\n(\n" + "\n".join(map(str, owner)) + "\n)", gz)
    nb = c.upper if not neighbour_all else np.concatenate([c.upper, -np.ones(c.nB, int)])
    _write(d, "neighbour", "labelList",
           f"{len(nb)}\n(\n" + "\n".join(map(str, nb)) + "\n)", gz)
    _write(d, "boundary", "polyBoundaryMesh",
           f"1\n(\n    walls\n    {{\n        type wall;\n        inGroups 1(wall);\n"
           f"        nFaces {c.nB};\n        startFace {c.F};\n    }}\n)")
    return owner


@pytest.mark.parametrize("gz,neighbour_all", [(False, False), (True, True)])
def test_reader_round_trip(tmp_path, gz, neighbour_all):
    """A mesh written in the reference's ASCII list formats (plain and
    gzipped, neighbour with and without -1 for the boundary faces) reads
    back exactly."""
    c = GradedBoxCase(5, 4, 3, zRatio=3.0, jitter=0.1, seed=2)
    owner = _write_mesh(str(tmp_path), c, gz, neighbour_all, False)
    m = PolyMesh(str(tmp_path))
    np.testing.assert_array_equal(m.points, c.points.reshape(-1, 3))
    np.testing.assert_array_equal(m.faceStart, c.faceStart)
    np.testing.assert_array_equal(m.facePoints, c.facePoints)
    np.testing.assert_array_equal(m.owner, owner)
    np.testing.assert_array_equal(m.lower, c.lower)
    np.testing.assert_array_equal(m.upper, c.upper)
    assert m.nCells == c.N and m.nInternalFaces == c.F
    assert m.patches == [("walls", "wall", c.F, c.nB)]


def test_reader_refuses_binary(tmp_path):
    (tmp_path / "points").write_text(
        "FoamFile\n{\n    format binary;\n    class vectorField;\n}\n3\n(...)")
    with pytest.raises(ValueError, match="binary"):
        PolyMesh(str(tmp_path))


@pytest.mark.parametrize("name", MESHES)
def test_real_mesh_geometry_invariants(name):
    """The oracle's geometry on a real mesh: every cell closed (its
    outward face area vectors sum to zero), positive volumes equal to a
    third of the sum of Sf.Cf over its faces (the pyramid decomposition
    with a closed surface), weights in (0, 1)."""
    m = load(name)
    assert np.all(m.lower < m.upper) and np.all(np.diff(m.lower) >= 0)
    g = O.mesh_geometry(m.nCells, m.lower, m.upper, m.bFaceCells, m.faceStart,
                        m.facePoints, m.points)
    Sf = np.concatenate([g["Sf"].reshape(-1, 3), g["bSf"].reshape(-1, 3)])
    Cf = np.concatenate([g["Cf"].reshape(-1, 3), g["bCf"].reshape(-1, 3)])
    own = m.owner
    nb = np.concatenate([m.neighbour, -np.ones(m.nFaces - m.nInternalFaces, int)])
    close = np.zeros((m.nCells, 3))
    np.add.at(close, own, Sf)
    np.subtract.at(close, nb[:m.nInternalFaces], Sf[:m.nInternalFaces])
    scale = np.abs(Sf).max()
    assert np.abs(close).max() < 1e-12*scale*10
    V = g["V"]
    assert np.all(V > 0)
    sv = np.zeros(m.nCells)
    np.add.at(sv, own, (Sf*Cf).sum(1))
    np.subtract.at(sv, nb[:m.nInternalFaces], (Sf*Cf).sum(1)[:m.nInternalFaces])
    np.testing.assert_allclose(sv/3.0, V, rtol=1e-8, atol=1e-12*np.abs(V).max())
    w = g["weights"]
    assert np.all((w > 0) & (w < 1))


def test_real_mesh_oracle_solve_vs_sparse_direct():
    """PCG+DIC of the oracle on the airFoil2D Laplacian (corrected scheme's
    implicit part, fixedValue outlet) converges to scipy's sparse direct
    solution of the same assembled matrix."""
    sp = pytest.importorskip("scipy.sparse")
    spl = pytest.importorskip("scipy.sparse.linalg")
    m = load("airfoil2d")
    g = O.mesh_geometry(m.nCells, m.lower, m.upper, m.bFaceCells, m.faceStart,
                        m.facePoints, m.points)
    bType, bValue = boundary_types(m)
    a = O.Addr(m.nCells, m.lower, m.upper)
    gamma = gamma_of(g["C"].reshape(-1, 3))
    upper, diag, ic, bc = O.laplacian_assemble(a, gamma, g["weights"], g["magSf"],
                                               g["nonOrthDeltaCoeffs"], m.bFaceCells,
                                               g["bMagSf"], g["bDeltaCoeffs"], bType, bValue)
    b = rhs(m.nCells, 4)
    diag, b = O.add_boundary(m.bFaceCells, ic, bc, diag, b)
    om = O.Matrix(a, diag, upper[:m.nInternalFaces])
    psi, perf = om.solve({"solver": "PCG", "preconditioner": "DIC", "tolerance": 1e-13,
                          "relTol": 0, "maxIter": 5000}, np.zeros(m.nCells), b)
    assert perf["converged"]
    A = sp.coo_matrix((np.concatenate([diag, upper[:m.nInternalFaces],
                                       upper[:m.nInternalFaces]]),
                       (np.concatenate([np.arange(m.nCells), m.lower, m.upper]),
                        np.concatenate([np.arange(m.nCells), m.upper, m.lower]))),
                      shape=(m.nCells, m.nCells)).tocsc()
    x = spl.spsolve(A, b)
    assert np.linalg.norm(psi - x) <= 1e-9*np.linalg.norm(x)
