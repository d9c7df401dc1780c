"""Coupled (cyclic) interfaces on the device against the oracle, through
the C-ABI (fh_mesh_create_coupled, fh_matrix_set_interfaces): a box whose
x-min and x-max patches form a cyclic pair.  Amul, Tmul, residual and sumA
add the interfaces after the face loop (lduMatrixATmul.C:34-280,
lduMatrixUpdateMatrixInterfaces.C:97-116, cyclicFvPatchField.C:150-174) and
are compared bitwise; the solvers (PCG+DIC, PBiCGStab+DILU, PBiCG+DILU,
smoothSolver with the Gauss-Seidel smoothers, GaussSeidelSmoother.C:117-143)
against the oracle's solves to the FP64 bar (1e-10, +-2 iterations)."""
import numpy as np
import pytest

import foamhip as fh
import oracle as O
from meshes import cyclic_system
from test_gpu_solvers import ITER_TOL, SOLUTION_TOL, rel_l2

pytestmark = pytest.mark.gpu

SHAPE = (24, 16, 12)


def systems(kind, seed=3):
    n, lower, upper, diag, up, lo, ifc, _, _ = cyclic_system(kind, seed, SHAPE, with_dense=False)
    om = O.Matrix(O.Addr(n, lower, upper), diag, up, lo, interfaces=ifc)
    mesh = fh.lduAddressing(n, lower, upper, interfaces=ifc)
    m = fh.lduMatrix(mesh, diag, up, lo)
    m.set_interfaces(ifc["bouCoeffs"], ifc["intCoeffs"])
    return n, om, m, mesh


@pytest.mark.parametrize("kind", ["sym", "asym"])
def test_interface_kernels_bitwise(kind):
    n, om, m, mesh = systems(kind)
    rng = np.random.default_rng(1)
    psi, b = rng.standard_normal(n), rng.standard_normal(n)
    np.testing.assert_array_equal(m.Amul(psi), om.Amul(psi))
    np.testing.assert_array_equal(m.Tmul(psi), om.Tmul(psi))
    np.testing.assert_array_equal(m.residual(psi, b), om.residual(psi, b))
    np.testing.assert_array_equal(m.sumA(), om.sumA())


@pytest.mark.parametrize("ctl,kind", [
    ({"solver": "PCG", "preconditioner": "DIC"}, "sym"),
    ({"solver": "PCG", "preconditioner": "diagonal"}, "sym"),
    ({"solver": "PBiCGStab", "preconditioner": "DILU"}, "asym"),
    ({"solver": "PBiCG", "preconditioner": "DILU"}, "asym"),
    ({"solver": "smoothSolver", "smoother": "symGaussSeidel", "nSweeps": 1}, "sym"),
    ({"solver": "smoothSolver", "smoother": "GaussSeidel", "nSweeps": 2}, "asym"),
    ({"solver": "smoothSolver", "smoother": "DIC", "nSweeps": 1}, "sym"),
])
def test_interface_solvers_vs_oracle(ctl, kind):
    n, om, m, mesh = systems(kind, seed=5)
    b = np.random.default_rng(2).standard_normal(n)
    ctl = dict(ctl, tolerance=1e-8, relTol=0, maxIter=3000)
    opsi, operf = om.solve(ctl, np.zeros(n), b)
    psi = np.zeros(n)
    perf = fh.solver_New("p", m, ctl).solve(psi, b)
    assert perf.converged == operf["converged"]
    assert abs(perf.nIterations - operf["nIterations"]) <= ITER_TOL, (perf, operf)
    assert perf.normFactor == pytest.approx(operf["normFactor"], rel=1e-12)
    assert perf.initialResidual == pytest.approx(operf["initialResidual"], rel=1e-12)
    assert rel_l2(psi, opsi) <= SOLUTION_TOL, rel_l2(psi, opsi)


def test_interface_mixed_pcg():
    """The mixed-precision PCG on a coupled mesh: FP32 products with the
    interface entries, the FP64 residual with them; the FP64 residual test,
    recomputed by the oracle, holds."""
    n, om, m, mesh = systems("sym", seed=7)
    b = np.random.default_rng(3).standard_normal(n)
    ctl = {"solver": "PCG", "preconditioner": "DIC", "tolerance": 1e-9, "relTol": 0,
           "maxIter": 3000, "precision": "mixed"}
    psi = np.zeros(n)
    perf = fh.solver_New("p", m, ctl).solve(psi, b)
    assert perf.converged
    assert np.abs(b - om.Amul(psi)).sum() / perf.normFactor < 1e-9
    opsi, _ = om.solve(dict(ctl, precision="double", tolerance=1e-13), np.zeros(n), b)
    assert rel_l2(psi, opsi) <= 1e-7


def test_interface_errors():
    n, lower, upper, diag, up, lo, ifc, _, _ = cyclic_system("sym", 3, SHAPE, with_dense=False)
    mesh = fh.lduAddressing(n, lower, upper, interfaces=ifc)
    m = fh.lduMatrix(mesh, diag, up)
    b = np.ones(n)
    # the interface coefficients are required on a coupled mesh
    with pytest.raises(fh.FoamHipError, match="interface"):
        fh.solver_New("p", m, {"solver": "PCG", "preconditioner": "DIC"}).solve(np.zeros(n), b)
    m.set_interfaces(ifc["bouCoeffs"])
    with pytest.raises(fh.FoamHipError) as e:
        fh.solver_New("p", m, {"solver": "GAMG", "smoother": "GaussSeidel"}).solve(
            np.zeros(n), b)
    assert e.value.code == 4
    # an interface face out of range is refused at mesh creation
    bad = dict(ifc, nbrCells=np.where(np.arange(len(ifc["nbrCells"])) == 3, n, ifc["nbrCells"]))
    with pytest.raises(fh.FoamHipError, match="out of range"):
        fh.lduAddressing(n, lower, upper, interfaces=bad)
