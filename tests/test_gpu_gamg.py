"""GAMG (solver and preconditioner), PBiCG and the tile-local DIC on the
device against the FP64 oracle.

The device agglomerates once per mesh, the oracle on every solve; both use
the process-global pairing direction, reset before each comparison, and each
comparison uses a fresh mesh so both see the same weights.  Every sum except
the global reductions is taken in the reference order, so the trajectories
agree to rounding: solution relative L2 <= 1e-10, iterations within +-2."""
import numpy as np
import pytest

import foamhip as fh
import oracle as O
from meshes import asym_coeffs, box, random_graph, spd_coeffs
from test_gpu_solvers import ITER_TOL, SOLUTION_TOL, convdiff_case, laplacian_case, rel_l2

pytestmark = pytest.mark.gpu


def run_pair(c_N, lower, upper, diag, up, lo, b, ctl, tol=SOLUTION_TOL, iter_tol=ITER_TOL):
    a = O.Addr(c_N, lower, upper)
    om = O.Matrix(a, diag, up, lo)
    O.reset_direction()
    opsi, operf = om.solve(ctl, np.zeros(c_N), b)
    mesh = fh.lduAddressing(c_N, lower, upper)
    m = fh.lduMatrix(mesh, diag, up, lo)
    fh.lib().fh_gamg_reset_direction()
    psi = np.zeros(c_N)
    perf = fh.solver_New("p", m, ctl).solve(psi, b)
    print(ctl, "device", perf.nIterations, perf.finalResidual, "oracle", operf["nIterations"],
          operf["finalResidual"], "diff", rel_l2(psi, opsi))
    assert perf.converged == operf["converged"]
    assert abs(perf.nIterations - operf["nIterations"]) <= iter_tol
    assert perf.normFactor == pytest.approx(operf["normFactor"], rel=1e-12)
    assert perf.initialResidual == pytest.approx(operf["initialResidual"], rel=1e-12)
    assert rel_l2(psi, opsi) <= tol
    return perf, operf, psi, opsi


GAMG = {"solver": "GAMG", "tolerance": 1e-8, "relTol": 0}


@pytest.mark.parametrize("smoother", ["symGaussSeidel", "GaussSeidel", "DIC", "FDIC",
                                      "DICGaussSeidel"])
def test_gamg_symmetric_smoothers(smoother):
    c, a, diag, upper, b = laplacian_case(16, sigma=1.0)
    ctl = dict(GAMG, smoother=smoother)
    perf, *_ = run_pair(c.N, c.lower, c.upper, diag, upper, None, b, ctl)
    assert perf.solverName == "GAMG"


def test_gamg_fixed_cycles_trajectory():
    """tolerance 0: exactly maxIter V-cycles on both sides."""
    c, a, diag, upper, b = laplacian_case(20, sigma=1.0)
    ctl = dict(GAMG, smoother="symGaussSeidel", tolerance=0, maxIter=6)
    perf, operf, psi, opsi = run_pair(c.N, c.lower, c.upper, diag, upper, None, b, ctl,
                                      iter_tol=0)
    assert perf.nIterations == 6
    assert perf.finalResidual == pytest.approx(operf["finalResidual"], rel=1e-9)


@pytest.mark.parametrize("smoother", ["symGaussSeidel", "DICGaussSeidel"])
def test_gamg_face_area_pair(smoother):
    """faceAreaPair (faceAreaPairGAMGAgglomeration.C:55-80): the hierarchy
    from the face areas, trajectories against the oracle's."""
    c, a, diag, upper, b = laplacian_case(16, sigma=1.0)
    ctl = dict(GAMG, smoother=smoother, agglomerator="faceAreaPair")
    om = O.Matrix(a, diag, upper)
    O.reset_direction()
    opsi, operf = om.solve(dict(ctl, Sf=c.Sf, magSf=c.magSf), np.zeros(c.N), b)
    mesh = fh.lduAddressing(c.N, c.lower, c.upper)
    mesh.set_face_areas(c.Sf, c.magSf)
    m = fh.lduMatrix(mesh, diag, upper)
    fh.lib().fh_gamg_reset_direction()
    psi = np.zeros(c.N)
    perf = fh.solver_New("p", m, ctl).solve(psi, b)
    print("faceAreaPair", smoother, perf.nIterations, operf["nIterations"], rel_l2(psi, opsi))
    assert perf.converged and operf["converged"]
    assert abs(perf.nIterations - operf["nIterations"]) <= ITER_TOL
    assert rel_l2(psi, opsi) <= SOLUTION_TOL
    # without the face areas the agglomerator is refused
    m2 = fh.lduMatrix(fh.lduAddressing(c.N, c.lower, c.upper), diag, upper)
    with pytest.raises(fh.FoamHipError, match="face areas"):
        fh.solver_New("p", m2, ctl).solve(np.zeros(c.N), b)


@pytest.mark.parametrize("opts", [
    {"nPreSweeps": 2},
    {"scaleCorrection": False},
    {"mergeLevels": 2},
    {"nCellsInCoarsestLevel": 40, "nFinestSweeps": 1, "nPostSweeps": 1},
    {"directSolveCoarsest": True},
    {"directSolveCoarsest": True, "nCellsInCoarsestLevel": 300},
])
def test_gamg_controls(opts):
    c, a, diag, upper, b = laplacian_case(16, sigma=0.5)
    ctl = dict(GAMG, smoother="GaussSeidel", **opts)
    run_pair(c.N, c.lower, c.upper, diag, upper, None, b, ctl)


@pytest.mark.parametrize("smoother", ["GaussSeidel", "DILU", "DILUGaussSeidel",
                                      "symGaussSeidel"])
def test_gamg_asymmetric(smoother):
    # diffusivity 1e-2: the coarse Galerkin operators of the upwind
    # operator at 1e-3 make GAMG diverge in the reference arithmetic too
    c, a, diag, upper, lower, b = convdiff_case(14, D=1e-2)
    ctl = dict(GAMG, smoother=smoother, maxIter=200)
    run_pair(c.N, c.lower, c.upper, diag, upper, lower, b, ctl)


def test_gamg_random_graph():
    rng = np.random.default_rng(7)
    n, lower, upper = random_graph(3000, 6, rng, shuffle_within_owner=True)
    diag, up = spd_coeffs(n, lower, upper, rng)
    b = rng.standard_normal(n)
    ctl = dict(GAMG, smoother="symGaussSeidel", tolerance=1e-10)
    run_pair(n, lower, upper, diag, up, None, b, ctl)


def test_gamg_preconditioned_krylov():
    c, a, diag, upper, b = laplacian_case(16, sigma=1.0)
    ctl = {"solver": "PCG", "preconditioner": {"preconditioner": "GAMG",
                                               "smoother": "GaussSeidel"},
           "tolerance": 1e-9, "relTol": 0}
    perf, *_ = run_pair(c.N, c.lower, c.upper, diag, upper, None, b, ctl)
    assert perf.solverName == "GAMGPCG"
    c, a, diag, upper, lower, b = convdiff_case(12, D=1e-2)
    ctl = {"solver": "PBiCGStab", "preconditioner": {"preconditioner": "GAMG",
                                                     "smoother": "GaussSeidel"},
           "tolerance": 1e-9, "relTol": 0}
    run_pair(c.N, c.lower, c.upper, diag, upper, lower, b, ctl)


def test_gamg_preconditioner_subdict_controls():
    """The GAMG preconditioner takes its smoother, sweeps, agglomeration,
    nVcycles and coarsest-level tolerance from the preconditioner
    sub-dictionary (lduMatrixPreconditioner.C:72-82, GAMGPreconditioner.C:
    44-63), not from the solver's dictionary, which names others here."""
    c, a, diag, upper, b = laplacian_case(20, sigma=1.0)
    sub = {"preconditioner": "GAMG", "smoother": "symGaussSeidel", "nVcycles": 1,
           "nCellsInCoarsestLevel": 50, "directSolveCoarsest": 0, "tolerance": 1e-2,
           "relTol": 0.1, "nFinestSweeps": 1}
    ctl = {"solver": "PCG", "preconditioner": sub, "tolerance": 1e-9, "relTol": 0,
           "smoother": "GaussSeidel", "nVcycles": 3, "nCellsInCoarsestLevel": 10,
           "nFinestSweeps": 4}
    perf, operf, *_ = run_pair(c.N, c.lower, c.upper, diag, upper, None, b, ctl)
    # the parent's GAMG entries would make another preconditioner
    parent_like = dict(ctl, preconditioner=dict(sub, smoother="GaussSeidel", nVcycles=3,
                                                nCellsInCoarsestLevel=10, nFinestSweeps=4))
    O.reset_direction()
    _, other = O.Matrix(a, diag, upper).solve(parent_like, np.zeros(c.N), b)
    assert other["nIterations"] != operf["nIterations"]
    with pytest.raises(fh.FoamHipError, match="smoother"):
        fh.controls_from_dict({"solver": "PCG", "preconditioner": "GAMG",
                               "smoother": "GaussSeidel"})


def test_gamg_mixed_precision():
    """FP32 V-cycles on the residual equation inside FP64 residual updates:
    the FP64 convergence test is met, the cycle count stays close to the FP64
    count (the FP32 cycles smooth with multi-colour Gauss-Seidel, not the
    ordered sweep), and at a tight tolerance the solution matches the FP64
    oracle to 1e-8."""
    c, a, diag, upper, b = laplacian_case(24, sigma=1.0)
    om = O.Matrix(a, diag, upper)
    ctl = dict(GAMG, smoother="GaussSeidel")
    O.reset_direction()
    opsi, operf = om.solve(ctl, np.zeros(c.N), b)
    mesh = fh.lduAddressing(c.N, c.lower, c.upper)
    m = fh.lduMatrix(mesh, diag, upper)
    fh.lib().fh_gamg_reset_direction()
    psi = np.zeros(c.N)
    perf = fh.solver_New("p", m, dict(ctl, precision="mixed")).solve(psi, b)
    assert perf.solverName == "GAMG(mixed)"
    assert perf.converged
    r = b - om.Amul(psi)
    assert np.abs(r).sum() / perf.normFactor < 1e-8
    assert perf.nIterations <= 1.25*operf["nIterations"] + ITER_TOL
    tight = dict(ctl, tolerance=1e-13, maxIter=200)
    O.reset_direction()
    xpsi, _ = om.solve(tight, np.zeros(c.N), b)
    psi2 = np.zeros(c.N)
    perf2 = fh.solver_New("p", m, dict(tight, precision="mixed")).solve(psi2, b)
    assert perf2.converged
    assert rel_l2(psi2, xpsi) <= 1e-8


@pytest.mark.parametrize("smoother", ["GaussSeidel", "symGaussSeidel"])
def test_gamg_mixed_precision_unstructured(smoother):
    """Mixed GAMG on a random graph (irregular coarse levels, many colours in
    the FP32 colour sweeps): the FP64 convergence test holds for the true
    residual, and at a tight tolerance the solution matches the FP64
    oracle."""
    rng = np.random.default_rng(7)
    n = 3000
    _, lower, upper = random_graph(n, 6, rng)
    diag, up = spd_coeffs(n, lower, upper, rng)
    b = rng.uniform(-1, 1, n)
    a = O.Addr(n, lower, upper)
    om = O.Matrix(a, diag, up)
    mesh = fh.lduAddressing(n, lower, upper)
    m = fh.lduMatrix(mesh, diag, up)
    ctl = dict(GAMG, smoother=smoother, maxIter=200)
    fh.lib().fh_gamg_reset_direction()
    psi = np.zeros(n)
    perf = fh.solver_New("p", m, dict(ctl, precision="mixed")).solve(psi, b)
    assert perf.solverName == "GAMG(mixed)" and perf.converged
    assert np.abs(b - om.Amul(psi)).sum() / perf.normFactor < 1e-8
    tight = dict(ctl, tolerance=1e-13)
    O.reset_direction()
    xpsi, _ = om.solve(tight, np.zeros(n), b)
    fh.lib().fh_gamg_reset_direction()
    psi2 = np.zeros(n)
    assert fh.solver_New("p", m, dict(tight, precision="mixed")).solve(psi2, b).converged
    assert rel_l2(psi2, xpsi) <= 1e-8


def test_gamg_no_coarse_levels_is_an_error():
    n, lower, upper = box(2, 2, 2)
    rng = np.random.default_rng(1)
    diag, up = spd_coeffs(n, lower, upper, rng)
    mesh = fh.lduAddressing(n, lower, upper)
    m = fh.lduMatrix(mesh, diag, up)
    with pytest.raises(fh.FoamHipError, match="No coarse levels"):
        fh.solver_New("p", m, dict(GAMG, smoother="GaussSeidel")).solve(
            np.zeros(n), np.ones(n))


@pytest.mark.parametrize("precon", ["DILU", "diagonal", "none"])
def test_pbicg(precon):
    # at diffusivity 1e-3 the reference's own PBiCG iteration count moves by
    # up to 30 under a 1e-15 relative change of b (no preconditioner), so
    # the +-2 bar is applied where the iteration is not chaotic
    c, a, diag, upper, lower, b = convdiff_case(14, D=1e-2)
    ctl = {"solver": "PBiCG", "preconditioner": precon, "tolerance": 1e-9, "relTol": 0,
           "maxIter": 2000}
    perf, *_ = run_pair(c.N, c.lower, c.upper, diag, upper, lower, b, ctl)
    assert perf.solverName == precon + "PBiCG"


def test_pbicg_random_graph_shuffled():
    rng = np.random.default_rng(11)
    n, lower, upper = random_graph(1500, 5, rng, shuffle_within_owner=True)
    diag, up, lo = asym_coeffs(n, lower, upper, rng)
    b = rng.standard_normal(n)
    ctl = {"solver": "PBiCG", "preconditioner": "DILU", "tolerance": 1e-10, "relTol": 0}
    run_pair(n, lower, upper, diag, up, lo, b, ctl)


def _tile_blocks(nx, ny, nz, lower, upper, bz):
    def tile(c):
        i = c % nx
        j = (c // nx) % ny
        k = c // (nx * ny)
        return (i // 16, j // 16, k // bz)
    return np.array([tile(l) != tile(u) for l, u in zip(lower, upper)])


@pytest.mark.parametrize("kind", ["DIC", "DILU"])
def test_tile_local_factor_is_block_jacobi(kind):
    """tileDIC / tileDILU on a box = the reference DIC / DILU of the matrix
    with every face between two 16 x 16 x 8 tiles removed (bitwise)."""
    nx, ny, nz = 37, 20, 19
    n, lower, upper = box(nx, ny, nz)
    rng = np.random.default_rng(5)
    if kind == "DIC":
        diag, up = spd_coeffs(n, lower, upper, rng)
        lo = None
    else:
        diag, up, lo = asym_coeffs(n, lower, upper, rng)
    cut = _tile_blocks(nx, ny, nz, lower, upper, 8)
    assert cut.any() and not cut.all()
    up_b = np.where(cut, 0.0, up)
    lo_b = None if lo is None else np.where(cut, 0.0, lo)
    om = O.Matrix(O.Addr(n, lower, upper), diag, up_b, lo_b)
    mesh = fh.lduAddressing(n, lower, upper)
    assert mesh.info()["tiled"] == 1
    m = fh.lduMatrix(mesh, diag, up, lo)
    r = rng.standard_normal(n)
    np.testing.assert_array_equal(m.reciprocalD("tile" + kind), om.reciprocalD(kind))
    np.testing.assert_array_equal(m.precondition("tile" + kind, r), om.precondition(kind, r))


def _numpy_pcg(om, b, prec, tol):
    """PCG of PCG.C:65-193 in numpy with a given preconditioner application;
    the iteration count of exact arithmetic for a preconditioner the solver
    library has no FP64 form of."""
    x = np.zeros(b.size)
    r = b.copy()
    nf = om.normFactor(x, b)
    wArA, p = 1e20, None
    for it in range(10000):
        w = prec(r)
        old, wArA = wArA, w @ r
        p = w.copy() if it == 0 else w + (wArA / old) * p
        q = om.Amul(p)
        alpha = wArA / (q @ p)
        x += alpha * p
        r -= alpha * q
        if np.abs(r).sum() / nf < tol:
            return it + 1
    return -1


def test_mixed_pcg_tile_dic_inner():
    """Mixed PCG with the tile-local inner preconditioner meets the FP64
    residual test, matches the FP64 oracle at a tight tolerance, and at the
    working tolerance takes about the iterations of PCG with the same
    preconditioner in exact arithmetic: on this 40^3 box (rows of 40 cells,
    run on rows zero-padded to 48) the FP16-rounded 16 x 16 x 16
    block-Jacobi DIC plus the coarse correction over the tiles."""
    c, a, diag, upper, b = laplacian_case(40, sigma=1.0)
    om = O.Matrix(a, diag, upper)
    mesh = fh.lduAddressing(c.N, c.lower, c.upper)
    m = fh.lduMatrix(mesh, diag, upper)
    cut = _tile_blocks(40, 40, 40, c.lower, c.upper, 16)
    up16 = upper.astype(np.float32).astype(np.float16).astype(np.float64)
    om_t = O.Matrix(O.Addr(c.N, c.lower, c.upper), diag, np.where(cut, 0.0, up16))
    tile, nT, Aci = _tile_coarse(om, c.N, 40, 40, 40)

    def two_level(r):
        rc = np.bincount(tile, r, nT)
        return om_t.precondition("DIC", r) + (Aci @ rc)[tile]

    exact = {"DIC": _numpy_pcg(om, b, lambda r: om.precondition("DIC", r), 1e-8),
             "tileDIC": _numpy_pcg(om, b, two_level, 1e-8)}
    for inner in ["tileDIC", "DIC"]:
        ctl = {"solver": "PCG", "preconditioner": "DIC", "tolerance": 1e-12, "relTol": 0,
               "maxIter": 3000, "precision": "mixed", "innerPreconditioner": inner}
        psi = np.zeros(c.N)
        perf = fh.solver_New("p", m, ctl).solve(psi, b)
        assert perf.converged
        r = b - om.Amul(psi)
        assert np.abs(r).sum() / perf.normFactor < 1e-12
        opsi, _ = om.solve(dict(ctl, precision="double", tolerance=1e-13), np.zeros(c.N), b)
        assert rel_l2(psi, opsi) <= 1e-8
        p8 = fh.solver_New("p", m, dict(ctl, tolerance=1e-8)).solve(np.zeros(c.N), b)
        print(inner, "mixed its at 1e-8", p8.nIterations, "exact arithmetic", exact[inner])
        assert abs(p8.nIterations - exact[inner]) <= max(3, 0.05 * exact[inner])


def _tile_coarse(om, n, nx, ny, nz):
    """The coarse correction over 16^3 tiles in numpy: P (cell -> tile
    indicator) and the inverse of Ac = P^T A P, from the FP64 matrix."""
    c = np.arange(n)
    nTI, nTJ = (nx + 15) // 16, (ny + 15) // 16
    tile = (c % nx) // 16 + nTI * (((c // nx) % ny) // 16 + nTJ * ((c // (nx * ny)) // 16))
    nT = tile.max() + 1
    AP = np.stack([om.Amul((tile == t).astype(float)) for t in range(nT)], axis=1)
    Ac = np.stack([AP[tile == t].sum(axis=0) for t in range(nT)])
    return tile, nT, np.linalg.inv(Ac)


def test_mixed_pcg_tile_dic_half():
    """Mixed PCG on a box whose tile-local inner preconditioner stores its
    coefficients in FP16 (nx a multiple of 16) and carries the coarse
    correction over the tiles: the FP64 residual test holds, the solution
    matches the FP64 oracle at a tight tolerance, and at the working
    tolerance it takes about the iterations of PCG with that preconditioner
    (the FP16-rounded 16^3 block-Jacobi DIC plus P Ac^-1 P^T) in exact
    arithmetic -- half the one-level count here."""
    c, a, diag, upper, b = laplacian_case(48, sigma=1.0)
    om = O.Matrix(a, diag, upper)
    mesh = fh.lduAddressing(c.N, c.lower, c.upper)
    m = fh.lduMatrix(mesh, diag, upper)
    cut = _tile_blocks(48, 48, 48, c.lower, c.upper, 16)
    up16 = upper.astype(np.float32).astype(np.float16).astype(np.float64)
    om_t = O.Matrix(O.Addr(c.N, c.lower, c.upper), diag, np.where(cut, 0.0, up16))
    tile, nT, Aci = _tile_coarse(om, c.N, 48, 48, 48)

    def two_level(r):
        rc = np.bincount(tile, r, nT)
        return om_t.precondition("DIC", r) + (Aci @ rc)[tile]

    exact = _numpy_pcg(om, b, two_level, 1e-8)
    one_level = _numpy_pcg(om, b, lambda r: om_t.precondition("DIC", r), 1e-8)
    assert exact < 0.75 * one_level
    ctl = {"solver": "PCG", "preconditioner": "DIC", "tolerance": 1e-12, "relTol": 0,
           "maxIter": 3000, "precision": "mixed"}
    psi = np.zeros(c.N)
    perf = fh.solver_New("p", m, ctl).solve(psi, b)
    assert perf.converged and perf.finalResidual < 1e-12
    opsi, _ = om.solve(dict(ctl, precision="double", tolerance=1e-13), np.zeros(c.N), b)
    assert rel_l2(psi, opsi) <= 1e-8
    p8 = fh.solver_New("p", m, dict(ctl, tolerance=1e-8)).solve(np.zeros(c.N), b)
    print("fp16 tileDIC mixed its at 1e-8", p8.nIterations, "exact arithmetic", exact)
    assert abs(p8.nIterations - exact) <= max(3, 0.05 * exact)


@pytest.mark.parametrize("shape", [(64, 48, 32), (32, 32, 80), (16, 16, 16), (208, 160, 32),
                                   (36, 28, 20), (37, 21, 19)])
def test_mixed_pcg_two_level_boxes(shape):
    """Mixed PCG with the two-level inner preconditioner on non-cubic boxes
    (and a single tile, where the coarse part is off; 208 x 160 x 32 has 130
    tiles per plane, so each plane's inverse is split 2 x 2 at 128; rows of
    36 and 37 cells run the tile-local apply on zero-padded rows, and with 37
    the coarse correction is added to w before p is formed instead of inside
    the fused product, which needs rows of a multiple of 4): the FP64
    residual test holds for the true residual, and at a tight tolerance the
    solution matches the FP64 oracle to 1e-8."""
    from foamhip.cases import BoxCase
    nx, ny, nz = shape
    c = BoxCase(nx, ny, nz)
    gamma = c.lognormal(5, 1.0, 4.0)
    a = O.Addr(c.N, c.lower, c.upper)
    upper, diag, ic, bc = O.laplacian_assemble(a, gamma, c.weights, c.magSf, c.deltaCoeffs,
                                               c.bFaceCells, c.bMagSf, c.bDeltaCoeffs,
                                               c.bType, c.bValue)
    b = c.rhs(3)
    diag, b = O.add_boundary(c.bFaceCells, ic, bc, diag, b)
    om = O.Matrix(a, diag, upper)
    mesh = fh.lduAddressing(c.N, c.lower, c.upper)
    m = fh.lduMatrix(mesh, diag, upper)
    ctl = {"solver": "PCG", "preconditioner": "DIC", "tolerance": 1e-8, "relTol": 0,
           "maxIter": 3000, "precision": "mixed"}
    psi = np.zeros(c.N)
    perf = fh.solver_New("p", m, ctl).solve(psi, b)
    assert perf.converged
    assert np.abs(b - om.Amul(psi)).sum() / perf.normFactor < 1e-8
    psi2 = np.zeros(c.N)
    assert fh.solver_New("p", m, dict(ctl, tolerance=1e-12)).solve(psi2, b).converged
    opsi, _ = om.solve(dict(ctl, precision="double", tolerance=1e-13), np.zeros(c.N), b)
    assert rel_l2(psi2, opsi) <= 1e-8


@pytest.mark.parametrize("kind", ["DIC", "DILU"])
@pytest.mark.parametrize("shape", [(37, 20, 19), (40, 40, 40), (16, 32, 48)])
def test_fp32_preconditioners(kind, shape):
    """The FP32 inner preconditioners against the FP64 oracle: the exact
    factor, and the tile-local one = the oracle factor of the matrix with the
    faces between the FP32 tiles (16 x 16 x 16) removed -- for a symmetric
    box of that matrix with its off-diagonal coefficients rounded to FP16
    (the storage of that path; rows that are not a multiple of 16 cells run
    on zero-padded rows).  FP32 bar: relative L2 difference 1e-5."""
    nx, ny, nz = shape
    n, lower, upper = box(nx, ny, nz)
    rng = np.random.default_rng(9)
    if kind == "DIC":
        diag, up = spd_coeffs(n, lower, upper, rng)
        lo = None
    else:
        diag, up, lo = asym_coeffs(n, lower, upper, rng)
    mesh = fh.lduAddressing(n, lower, upper)
    m = fh.lduMatrix(mesh, diag, up, lo)
    r = rng.standard_normal(n)
    om = O.Matrix(O.Addr(n, lower, upper), diag, up, lo)
    w = m.precondition_f32(kind, r).astype(np.float64)
    ref = om.precondition(kind, r)
    assert rel_l2(w, ref) <= 1e-5, rel_l2(w, ref)
    cut = _tile_blocks(nx, ny, nz, lower, upper, 16)
    up_t = up
    if kind == "DIC":
        up_t = up.astype(np.float32).astype(np.float16).astype(np.float64)
    om_t = O.Matrix(O.Addr(n, lower, upper), diag, np.where(cut, 0.0, up_t),
                    None if lo is None else np.where(cut, 0.0, lo))
    w = m.precondition_f32("tile" + kind, r).astype(np.float64)
    ref = om_t.precondition(kind, r)
    assert rel_l2(w, ref) <= 1e-5, rel_l2(w, ref)


def _half_stored(up32):
    """The FP16 storage of the tile-local DIC: scaled by 2^e so that the
    largest magnitude lies in [2^13, 2^14), rounded to FP16, scaled back;
    FP32 when the smallest non-zero magnitude would fall below 2^-14."""
    a = np.abs(up32[up32 != 0]).astype(np.float64)
    e = 14 - int(np.frexp(a.max())[1])
    if a.min() * 2.0 ** e < 2.0 ** -14:
        return up32.astype(np.float64), None
    h = (up32 * np.float32(2.0 ** e)).astype(np.float16).astype(np.float64) / 2.0 ** e
    return h, e


@pytest.mark.parametrize("scale", [1e6, 1e-7, "span"])
def test_tile_dic_half_scaled(scale):
    """The FP16 coefficient storage of the tile-local DIC holds for any
    coefficient magnitude: scaled by 1e6 (plain FP16 would overflow to inf)
    and by 1e-7 (plain FP16 would flush to subnormals or zero) the apply is
    the block-Jacobi DIC of the coefficients rounded at the storage scale;
    magnitudes spanning more than 2^28 keep FP32 coefficients.  Mixed PCG on
    the scaled systems converges in the unscaled system's iterations."""
    nx = ny = nz = 32
    n, lower, upper = box(nx, ny, nz)
    rng = np.random.default_rng(17)
    diag, up = spd_coeffs(n, lower, upper, rng)
    if scale == "span":
        # one z-plane of faces 1e-9 weaker: the span exceeds 2^28
        weak = (lower // (nx * ny)) == 3
        up = np.where(weak, up * 1e-9, up)
        diag = np.zeros(n)
        np.subtract.at(diag, lower, up)
        np.subtract.at(diag, upper, up)
        diag += rng.uniform(0.01, 0.5, n)
        s = 1.0
    else:
        s = scale
    diag, up = diag * s, up * s
    mesh = fh.lduAddressing(n, lower, upper)
    m = fh.lduMatrix(mesh, diag, up)
    r = rng.standard_normal(n)
    up_t, e = _half_stored(up.astype(np.float32))
    assert (e is None) == (scale == "span")
    cut = _tile_blocks(nx, ny, nz, lower, upper, 16)
    om_t = O.Matrix(O.Addr(n, lower, upper), diag, np.where(cut, 0.0, up_t))
    w = m.precondition_f32("tileDIC", r).astype(np.float64)
    assert np.all(np.isfinite(w))
    ref = om_t.precondition("DIC", r)
    assert rel_l2(w, ref) <= 1e-5, rel_l2(w, ref)
    if scale == "span":
        return
    b = rng.uniform(-1, 1, n) * s
    ctl = {"solver": "PCG", "preconditioner": "DIC", "tolerance": 1e-8, "relTol": 0,
           "maxIter": 3000, "precision": "mixed"}
    psi = np.zeros(n)
    perf = fh.solver_New("p", m, ctl).solve(psi, b)
    m1 = fh.lduMatrix(mesh, diag / s, up / s)
    perf1 = fh.solver_New("p", m1, ctl).solve(np.zeros(n), b / s)
    om = O.Matrix(O.Addr(n, lower, upper), diag, up)
    assert perf.converged
    assert np.abs(b - om.Amul(psi)).sum() / perf.normFactor < 1e-8
    assert abs(perf.nIterations - perf1.nIterations) <= 2


@pytest.mark.parametrize("shape", [(48, 40, 37), (37, 40, 21)])
@pytest.mark.parametrize("apply", ["lds", "quarter"])
def test_tile_dic_apply_variants(apply, shape, monkeypatch):
    """The other kernels of the symmetric tile-local DIC with FP16
    coefficients (FOAMHIP_TILE_APPLY: the tile staged in LDS, the
    quarter-row pipeline) give the row walk's result: full and partial
    tiles (ny and nz not multiples of 16; nx 37 run on zero-padded rows),
    against the oracle's block-Jacobi DIC of the FP16-stored coefficients,
    and in a mixed PCG solve (FP64 residual test)."""
    nx, ny, nz = shape
    n, lower, upper = box(nx, ny, nz)
    rng = np.random.default_rng(23)
    diag, up = spd_coeffs(n, lower, upper, rng)
    mesh = fh.lduAddressing(n, lower, upper)
    m = fh.lduMatrix(mesh, diag, up)
    r = rng.standard_normal(n)
    monkeypatch.setenv("FOAMHIP_TILE_APPLY", "rows")
    w_rows = m.precondition_f32("tileDIC", r).astype(np.float64)
    monkeypatch.setenv("FOAMHIP_TILE_APPLY", apply)
    w = m.precondition_f32("tileDIC", r).astype(np.float64)
    up_t, e = _half_stored(up.astype(np.float32))
    assert e is not None
    cut = _tile_blocks(nx, ny, nz, lower, upper, 16)
    om_t = O.Matrix(O.Addr(n, lower, upper), diag, np.where(cut, 0.0, up_t))
    ref = om_t.precondition("DIC", r)
    assert rel_l2(w, ref) <= 1e-5, rel_l2(w, ref)
    assert rel_l2(w, w_rows) <= 1e-6, rel_l2(w, w_rows)
    b = rng.uniform(-1, 1, n)
    ctl = {"solver": "PCG", "preconditioner": "DIC", "tolerance": 1e-8, "relTol": 0,
           "maxIter": 3000, "precision": "mixed"}
    psi = np.zeros(n)
    perf = fh.solver_New("p", m, ctl).solve(psi, b)
    assert perf.converged
    om = O.Matrix(O.Addr(n, lower, upper), diag, up)
    assert np.abs(b - om.Amul(psi)).sum() / perf.normFactor < 1e-8


AGG_CASES = [
    ("box", lambda r: box(19, 14, 11), "spd", 1),
    ("box_merge2", lambda r: box(24, 20, 9), "spd", 2),
    ("box_merge3", lambda r: box(17, 16, 15), "spd", 3),
    ("rand", lambda r: random_graph(3000, 6, r), "spd", 1),
    ("rand_shuffled_merge2", lambda r: random_graph(2500, 5, r, shuffle_within_owner=True), "spd", 2),
    ("asym", lambda r: box(15, 13, 12), "asym", 1),
    # sparse, no chain: cells without faces (numbered after the pairs) and
    # unpaired cells joining clusters through other unpaired cells
    ("sparse_isolated", lambda r: random_graph(4000, 1.2, r, connect_chain=False), "spd", 1),
    ("sparse_merge2", lambda r: random_graph(3000, 2, r, connect_chain=False), "spd", 2),
]


@pytest.mark.parametrize("name,gen,kind,merge", AGG_CASES)
@pytest.mark.parametrize("start_backward", [False, True])
def test_agglomeration_maps_equal_oracle(name, gen, kind, merge, start_backward):
    """pairGAMGAgglomerate.C:123-301 and GAMGAgglomerateLduAddressing.C:
    32-239, 514-611 on the device: every level's restrictAddressing,
    faceRestrictAddressing, faceFlipMap and coarse addressing equal the
    oracle's sequential walk (the direction toggle starting either way)."""
    rng = np.random.default_rng(31)
    n, lo, up = gen(rng)
    if kind == "spd":
        d, u = spd_coeffs(n, lo, up, rng)
        l = None
    else:
        d, u, l = asym_coeffs(n, lo, up, rng)
    ctl = {"nCellsInCoarsestLevel": 8, "mergeLevels": merge}
    O.reset_direction()
    fh.lib().fh_gamg_reset_direction()
    if start_backward:
        # one pair level of a throwaway two-cell mesh (one coarse cell, below
        # nCellsInCoarsestLevel 2: no level kept) flips both toggles once
        tiny = {"nCellsInCoarsestLevel": 2, "mergeLevels": 1}
        assert O.agglomerate(O.Matrix(O.Addr(*box(2, 1, 1)), np.full(2, 4.0), np.ones(1)),
                             tiny) == []
        tm = fh.lduMatrix(fh.lduAddressing(*box(2, 1, 1)), np.full(2, 4.0), np.ones(1))
        assert tm.agglomeration(tiny) == []
        assert fh.lib().fh_gamg_direction() == 0
    ref = O.agglomerate(O.Matrix(O.Addr(n, lo, up), d, u, l), ctl)
    got = fh.lduMatrix(fh.lduAddressing(n, lo, up), d, u, l).agglomeration(ctl)
    assert len(got) == len(ref) >= 2
    for g, r in zip(got, ref):
        for key in ("nFineCells", "nCoarseCells", "nCoarseFaces"):
            assert g[key] == r[key], key
        for key in ("restrictMap", "faceRestrict", "lower", "upper"):
            np.testing.assert_array_equal(g[key], r[key], err_msg=key)
        # the flip of a face inside a coarse cell is not used (the face goes
        # to the diagonal)
        ext = r["faceRestrict"] >= 0
        np.testing.assert_array_equal(g["faceFlip"][ext], r["faceFlip"][ext].astype(np.uint8))
    assert fh.lib().fh_gamg_direction() == O.get_direction()


def test_agglomeration_face_area_ties():
    """faceAreaPair on a uniform box: every weight ties, so the choice is the
    first face of each cell's list (the cellFaces order)."""
    n, lo, up = box(12, 10, 8)
    rng = np.random.default_rng(32)
    d, u = spd_coeffs(n, lo, up, rng)
    Sf = np.zeros((len(lo), 3))
    Sf[:, 0] = 1.0
    magSf = np.ones(len(lo))
    ctl = {"nCellsInCoarsestLevel": 8, "mergeLevels": 1, "agglomerator": "faceAreaPair"}
    O.reset_direction()
    ref = O.agglomerate(O.Matrix(O.Addr(n, lo, up), d, u), dict(ctl, Sf=Sf, magSf=magSf))
    mesh = fh.lduAddressing(n, lo, up)
    mesh.set_face_areas(Sf, magSf)
    fh.lib().fh_gamg_reset_direction()
    got = fh.lduMatrix(mesh, d, u).agglomeration(ctl)
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        np.testing.assert_array_equal(g["restrictMap"], r["restrictMap"])
        np.testing.assert_array_equal(g["faceRestrict"], r["faceRestrict"])
        np.testing.assert_array_equal(g["lower"], r["lower"])
        np.testing.assert_array_equal(g["upper"], r["upper"])
