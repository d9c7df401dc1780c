"""Real meshes of the reference's tutorials end to end on the device:
airFoil2D (2-D, an empty front-and-back patch) and tank3D (3-D, triangles
and quadrilaterals), read by foamhip.polymesh from fixtures.  The device
geometry and the Laplacian assembly (Gauss linear corrected: the implicit
part with nonOrthDeltaCoeffs) are the oracle's bitwise; the pressure-type
solves -- PCG+DIC, the case's GAMG (GaussSeidel smoother, airFoil2D's p
entry) and smoothSolver -- match the FP64 oracle (1e-10, +-2 iterations);
mixed PCG meets the FP64 residual test."""
import ctypes

import numpy as np
import pytest

import foamhip as fh
import oracle as O
from realmesh import MESHES, boundary_types, gamma_of, load, rhs
from test_gpu_geometry import P, dev, device_geometry, host
from test_gpu_solvers import ITER_TOL, SOLUTION_TOL, rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

CONTROLS = [
    {"solver": "PCG", "preconditioner": "DIC", "tolerance": 1e-10, "relTol": 0,
     "maxIter": 5000},
    # airFoil2D/system/fvSolution: the p entry and the U entry
    {"solver": "GAMG", "smoother": "GaussSeidel", "tolerance": 1e-6, "relTol": 0.1,
     "maxIter": 1000},
    {"solver": "smoothSolver", "smoother": "GaussSeidel", "nSweeps": 2,
     "tolerance": 1e-8, "relTol": 0.1, "maxIter": 1000},
    # the p entry's GAMG solved to its tolerance (airFoil2D: about 200 cycles)
    {"solver": "GAMG", "smoother": "GaussSeidel", "tolerance": 1e-6, "relTol": 0,
     "maxIter": 1000},
]


def system(name):
    m = load(name)
    mesh, g = device_geometry(m.nCells, m.lower, m.upper, m.bFaceCells, m.faceStart,
                              m.facePoints, m.points)
    ref = O.mesh_geometry(m.nCells, m.lower, m.upper, m.bFaceCells, m.faceStart,
                          m.facePoints, m.points)
    return m, mesh, g, ref


@pytest.mark.parametrize("name", MESHES)
def test_real_mesh_geometry_and_assembly_bitwise(name):
    m, mesh, g, ref = system(name)
    assert not bool(mesh.info()["tiled"])
    for k in ref:
        np.testing.assert_array_equal(g[k], ref[k], err_msg=k)
    bType, bValue = boundary_types(m)
    gamma = gamma_of(g["C"].reshape(-1, 3))
    a = O.Addr(m.nCells, m.lower, m.upper)
    upper, diag, ic, bc = O.laplacian_assemble(a, gamma, g["weights"], g["magSf"],
                                               g["nonOrthDeltaCoeffs"], m.bFaceCells,
                                               g["bMagSf"], g["bDeltaCoeffs"], bType, bValue)
    b = rhs(m.nCells, 7)
    diag_b, b_b = O.add_boundary(m.bFaceCells, ic, bc, diag, b)
    bnd = fh.Boundary(mesh, m.bFaceCells)
    mat = fh.lduMatrix(mesh)
    nB = m.nFaces - m.nInternalFaces
    d_ic = torch.empty(nB, dtype=torch.float64, device="cuda")
    d_bc = torch.empty_like(d_ic)
    d_src = dev(b)
    ins = [dev(gamma), dev(g["weights"]), dev(g["magSf"]), dev(g["nonOrthDeltaCoeffs"]),
           dev(g["bMagSf"]), dev(g["bDeltaCoeffs"]), dev(bType, np.int32), dev(bValue)]
    torch.cuda.synchronize()
    fh._check(fh.lib().fh_assemble_laplacian(mat.handle, bnd.handle, *[P(t) for t in ins],
                                             P(d_ic), P(d_bc), P(d_src), 1))
    d, u, _ = mat.coeffs()
    np.testing.assert_array_equal(d, diag_b)
    np.testing.assert_array_equal(u, upper[:m.nInternalFaces])
    np.testing.assert_array_equal(host(d_src), b_b)


@pytest.mark.parametrize("name", MESHES)
@pytest.mark.parametrize("ctl", CONTROLS, ids=["PCG-DIC", "GAMG-GS", "smooth-GS", "GAMG-GS-full"])
def test_real_mesh_solves_vs_oracle(name, ctl):
    m, mesh, g, ref = system(name)
    bType, bValue = boundary_types(m)
    gamma = gamma_of(ref["C"].reshape(-1, 3))
    a = O.Addr(m.nCells, m.lower, m.upper)
    upper, diag, ic, bc = O.laplacian_assemble(a, gamma, ref["weights"], ref["magSf"],
                                               ref["nonOrthDeltaCoeffs"], m.bFaceCells,
                                               ref["bMagSf"], ref["bDeltaCoeffs"], bType, bValue)
    b = rhs(m.nCells, 11)
    diag, b = O.add_boundary(m.bFaceCells, ic, bc, diag, b)
    up = upper[:m.nInternalFaces]
    om = O.Matrix(a, diag, up)
    O.reset_direction()
    opsi, operf = om.solve(ctl, np.zeros(m.nCells), b)
    fh.lib().fh_gamg_reset_direction()
    mat = fh.lduMatrix(mesh, diag, up)
    psi = np.zeros(m.nCells)
    perf = fh.solver_New("p", mat, ctl).solve(psi, b)
    print(f"{name} {ctl['solver']}: device {perf.nIterations}, oracle {operf['nIterations']}, "
          f"rel L2 {rel_l2(psi, opsi):.2e}")
    assert perf.converged and operf["converged"]
    assert abs(perf.nIterations - operf["nIterations"]) <= ITER_TOL
    assert perf.normFactor == pytest.approx(operf["normFactor"], rel=1e-12)
    assert rel_l2(psi, opsi) <= SOLUTION_TOL


@pytest.mark.parametrize("name", MESHES)
def test_real_mesh_mixed_pcg(name):
    """Mixed PCG (FP32 inner, the FP32 global DIC on a general mesh) meets
    the FP64 residual test recomputed on the host, and at tolerance 1e-12
    matches the FP64 oracle to 1e-8."""
    m, mesh, g, ref = system(name)
    bType, bValue = boundary_types(m)
    gamma = gamma_of(ref["C"].reshape(-1, 3))
    a = O.Addr(m.nCells, m.lower, m.upper)
    upper, diag, ic, bc = O.laplacian_assemble(a, gamma, ref["weights"], ref["magSf"],
                                               ref["nonOrthDeltaCoeffs"], m.bFaceCells,
                                               ref["bMagSf"], ref["bDeltaCoeffs"], bType, bValue)
    b = rhs(m.nCells, 13)
    diag, b = O.add_boundary(m.bFaceCells, ic, bc, diag, b)
    up = upper[:m.nInternalFaces]
    om = O.Matrix(a, diag, up)
    mat = fh.lduMatrix(mesh, diag, up)
    ctl = {"solver": "PCG", "preconditioner": "DIC", "tolerance": 1e-8, "relTol": 0,
           "maxIter": 5000, "precision": "mixed"}
    psi = np.zeros(m.nCells)
    perf = fh.solver_New("p", mat, ctl).solve(psi, b)
    assert perf.converged
    assert np.abs(b - om.Amul(psi)).sum() / perf.normFactor < 1e-8
    psi2 = np.zeros(m.nCells)
    assert fh.solver_New("p", mat, dict(ctl, tolerance=1e-12)).solve(psi2, b).converged
    opsi, _ = om.solve(dict(ctl, precision="double", tolerance=1e-13), np.zeros(m.nCells), b)
    assert rel_l2(psi2, opsi) <= 1e-8
