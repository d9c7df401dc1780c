"""Pins the oracle (oracle/oracle.c) against what the mathematics fixes:
dense linear algebra on the same matrix, closed forms, invariants and
hand-worked known answers.  The reference ships no golden vectors for this
path, so these checks are what makes the oracle trustworthy; each is chosen
so that a dropped term, a swapped upper/lower or a wrong index fails."""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle as O
import numpy_solvers as NS
from meshes import asym_coeffs, box, dense, random_graph, spd_coeffs

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def mk(n, lo, up, rng, kind="spd"):
    a = O.Addr(n, lo, up)
    if kind == "spd":
        d, u = spd_coeffs(n, lo, up, rng)
        return a, O.Matrix(a, d, u), dense(n, lo, up, d, u)
    d, u, l = asym_coeffs(n, lo, up, rng)
    return a, O.Matrix(a, d, u, l), dense(n, lo, up, d, u, l)


MESHES = [
    ("box", lambda rng: box(4, 3, 5)),
    ("rand", lambda rng: random_graph(40, 5, rng)),
    ("rand_shuffled", lambda rng: random_graph(40, 5, rng, shuffle_within_owner=True)),
]


@pytest.mark.parametrize("name,gen", MESHES)
def test_addressing_definitions(name, gen):
    rng = np.random.default_rng(1)
    n, lo, up = gen(rng)
    a = O.Addr(n, lo, up)
    # losort: faces grouped by neighbour, stable in face index
    exp = np.argsort(up, kind="stable")
    assert np.array_equal(a.losort(), exp)
    own = a.ownerStart()
    assert own[0] == 0 and own[-1] == len(lo)
    for c in range(n):
        assert np.all(lo[own[c]:own[c + 1]] == c)
    ls = a.losortStart()
    for c in range(n):
        assert np.all(up[exp[ls[c]:ls[c + 1]]] == c)


@pytest.mark.parametrize("name,gen", MESHES)
@pytest.mark.parametrize("kind", ["spd", "asym"])
def test_matrix_kernels_vs_dense(name, gen, kind):
    rng = np.random.default_rng(2)
    n, lo, up = gen(rng)
    a, m, A = mk(n, lo, up, rng, kind)
    x = rng.standard_normal(n)
    b = rng.standard_normal(n)
    np.testing.assert_allclose(m.Amul(x), A @ x, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(m.Tmul(x), A.T @ x, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(m.residual(x, b), b - A @ x, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(m.sumA(), A.sum(axis=1), rtol=1e-13, atol=1e-13)


def test_norm_factor_closed_forms():
    rng = np.random.default_rng(3)
    n, lo, up = box(5, 4, 3)
    a, m, A = mk(n, lo, up, rng)
    b = rng.standard_normal(n)
    # psi = 0: normFactor = sum|b| + 1e-20 (lduMatrixSolver.C:174-197)
    assert m.normFactor(np.zeros(n), b) == pytest.approx(np.abs(b).sum() + 1e-20, rel=1e-14)
    # psi = c: A psi = c*sumA, xRef = c -> sum|b - c*sumA| + 1e-20
    c = 0.37
    exp = np.abs(b - c * A.sum(axis=1)).sum() + 1e-20
    assert m.normFactor(np.full(n, c), b) == pytest.approx(exp, rel=1e-12)


def ic_factors(A, rD, kind):
    """M = (D~ + L) D~^-1 (D~ + U) with D~ = 1/rD."""
    Dt = np.diag(1.0 / rD)
    L = np.tril(A, -1)
    U = np.triu(A, 1)
    return (Dt + L) @ np.linalg.inv(Dt) @ (Dt + U)


@pytest.mark.parametrize("name,gen", MESHES)
@pytest.mark.parametrize("precon", ["DIC", "FDIC", "DILU"])
def test_incomplete_factor_preconditioners(name, gen, precon):
    rng = np.random.default_rng(4)
    n, lo, up = gen(rng)
    a, m, A = mk(n, lo, up, rng, "asym" if precon == "DILU" else "spd")
    rD = m.reciprocalD(precon)
    M = ic_factors(A, rD, precon)
    # defining property of the zero-fill factorisation: diag(M) = diag(A)
    np.testing.assert_allclose(np.diag(M), np.diag(A), rtol=1e-12)
    # and M has no entry outside the sparsity of A except the fill
    r = rng.standard_normal(n)
    w = m.precondition(precon, r)
    np.testing.assert_allclose(M @ w, r, rtol=1e-10, atol=1e-10)
    if precon == "DILU":
        wT = m.precondition(precon, r, transpose=True)
        np.testing.assert_allclose(M.T @ wT, r, rtol=1e-10, atol=1e-10)


def test_dic_is_exact_cholesky_on_a_chain():
    rng = np.random.default_rng(5)
    n, lo, up = box(12, 1, 1)[0], np.arange(11, dtype=np.int32), np.arange(1, 12, dtype=np.int32)
    a, m, A = mk(n, lo, up, rng)
    r = rng.standard_normal(n)
    np.testing.assert_allclose(m.precondition("DIC", r), np.linalg.solve(A, r), rtol=1e-12)
    psi, perf = m.solve({"solver": "PCG", "preconditioner": "DIC", "tolerance": 1e-12},
                        np.zeros(n), r)
    assert perf["nIterations"] == 1 and perf["converged"]
    # DILU is exact LU on an asymmetric chain: PBiCGStab converges in one step
    a2, m2, A2 = mk(n, lo, up, rng, "asym")
    psi, perf = m2.solve({"solver": "PBiCGStab", "preconditioner": "DILU",
                          "tolerance": 1e-12}, np.zeros(n), r)
    assert perf["nIterations"] == 1
    np.testing.assert_allclose(psi, np.linalg.solve(A2, r), rtol=1e-10)


def test_diagonal_preconditioner():
    rng = np.random.default_rng(6)
    n, lo, up = box(3, 3, 3)
    a, m, A = mk(n, lo, up, rng)
    r = rng.standard_normal(n)
    np.testing.assert_allclose(m.precondition("diagonal", r), r / np.diag(A), rtol=1e-15)
    np.testing.assert_array_equal(m.precondition("none", r), r)


@pytest.mark.parametrize("name,gen", MESHES)
@pytest.mark.parametrize("kind", ["spd", "asym"])
def test_gauss_seidel_is_a_triangular_solve(name, gen, kind):
    rng = np.random.default_rng(7)
    n, lo, up = gen(rng)
    a, m, A = mk(n, lo, up, rng, kind)
    x0 = rng.standard_normal(n)
    b = rng.standard_normal(n)
    L = np.tril(A)          # D + L
    U = np.triu(A, 1)
    fwd = np.linalg.solve(L, b - U @ x0)
    np.testing.assert_allclose(m.smooth("GaussSeidel", x0, b, 1), fwd, rtol=1e-11, atol=1e-12)
    bwd = np.linalg.solve(np.triu(A), b - np.tril(A, -1) @ fwd)
    np.testing.assert_allclose(m.smooth("symGaussSeidel", x0, b, 1), bwd, rtol=1e-11, atol=1e-12)
    # two sweeps = the one-sweep map applied twice
    two = m.smooth("GaussSeidel", m.smooth("GaussSeidel", x0, b, 1), b, 1)
    np.testing.assert_array_equal(m.smooth("GaussSeidel", x0, b, 2), two)


@pytest.mark.parametrize("smoother,kind", [("DIC", "spd"), ("DILU", "asym"), ("FDIC", "spd")])
def test_ilu_smoothers(smoother, kind):
    rng = np.random.default_rng(8)
    n, lo, up = random_graph(30, 4, rng)
    a, m, A = mk(n, lo, up, rng, kind)
    x0 = rng.standard_normal(n)
    b = rng.standard_normal(n)
    M = ic_factors(A, m.reciprocalD("DILU" if smoother == "DILU" else "DIC"), smoother)
    exp = x0 + np.linalg.solve(M, b - A @ x0)
    np.testing.assert_allclose(m.smooth(smoother, x0, b, 1), exp, rtol=1e-10, atol=1e-11)


SOLVES = [
    ({"solver": "PCG", "preconditioner": "DIC"}, "spd"),
    ({"solver": "PCG", "preconditioner": "FDIC"}, "spd"),
    ({"solver": "PCG", "preconditioner": "diagonal"}, "spd"),
    ({"solver": "PCG", "preconditioner": "none"}, "spd"),
    ({"solver": "PBiCGStab", "preconditioner": "DILU"}, "asym"),
    ({"solver": "PBiCGStab", "preconditioner": "diagonal"}, "asym"),
    ({"solver": "PBiCG", "preconditioner": "DILU"}, "asym"),
    ({"solver": "smoothSolver", "smoother": "symGaussSeidel", "nSweeps": 2}, "spd"),
    ({"solver": "smoothSolver", "smoother": "GaussSeidel"}, "asym"),
    ({"solver": "smoothSolver", "smoother": "DICGaussSeidel"}, "spd"),
    ({"solver": "GAMG", "smoother": "GaussSeidel", "nCellsInCoarsestLevel": 4}, "spd"),
    ({"solver": "GAMG", "smoother": "symGaussSeidel", "nCellsInCoarsestLevel": 8,
      "directSolveCoarsest": 1, "mergeLevels": 2}, "spd"),
    ({"solver": "GAMG", "smoother": "GaussSeidel", "nCellsInCoarsestLevel": 4}, "asym"),
    ({"solver": "PCG", "preconditioner": "GAMG", "smoother": "DICGaussSeidel",
      "nCellsInCoarsestLevel": 4}, "spd"),
]


@pytest.mark.parametrize("ctl,kind", SOLVES)
def test_solvers_converge_to_dense_solution(ctl, kind):
    rng = np.random.default_rng(9)
    n, lo, up = box(6, 5, 4)
    O.reset_direction()
    a, m, A = mk(n, lo, up, rng, kind)
    b = rng.standard_normal(n)
    c = dict(ctl, tolerance=1e-10, relTol=0, maxIter=2000)
    psi, perf = m.solve(c, np.zeros(n), b)
    x = np.linalg.solve(A, b)
    assert perf["converged"], perf
    assert perf["finalResidual"] < 1e-10
    # normalised L1 residual recomputed densely
    assert np.abs(b - A @ psi).sum() / perf["normFactor"] == pytest.approx(
        perf["finalResidual"], rel=1e-6, abs=1e-14)
    np.testing.assert_allclose(psi, x, rtol=1e-6, atol=1e-8 * np.abs(x).max())


def test_pcg_terminates_within_n_iterations_unpreconditioned():
    rng = np.random.default_rng(10)
    n, lo, up = box(2, 2, 2)
    a, m, A = mk(n, lo, up, rng)
    b = rng.standard_normal(n)
    psi, perf = m.solve({"solver": "PCG", "preconditioner": "none", "tolerance": 1e-12},
                        np.zeros(n), b)
    assert perf["nIterations"] <= n + 1


def test_solver_controls_semantics():
    rng = np.random.default_rng(11)
    n, lo, up = box(5, 5, 5)
    a, m, A = mk(n, lo, up, rng)
    b = rng.standard_normal(n)
    # maxIter caps; converged stays false
    psi, p = m.solve({"solver": "PCG", "preconditioner": "DIC", "tolerance": 1e-14,
                      "maxIter": 3}, np.zeros(n), b)
    assert p["nIterations"] == 3 and not p["converged"]
    # minIter forces iterations even when converged initially
    psi, p = m.solve({"solver": "PCG", "preconditioner": "DIC", "tolerance": 10.0,
                      "minIter": 2}, np.zeros(n), b)
    assert p["nIterations"] == 2
    # already converged: no iteration
    psi, p = m.solve({"solver": "PCG", "preconditioner": "DIC", "tolerance": 10.0},
                     np.zeros(n), b)
    assert p["nIterations"] == 0 and p["converged"] and p["initialResidual"] == pytest.approx(1.0)
    # relTol
    psi, p = m.solve({"solver": "PCG", "preconditioner": "DIC", "tolerance": 0,
                      "relTol": 1e-2}, np.zeros(n), b)
    assert p["finalResidual"] < 1e-2 * p["initialResidual"] and p["converged"]
    # singular: zero right-hand side with zero initial guess -> initial
    # residual 0 converges; a zero matrix row is singular
    psi, p = m.solve({"solver": "PCG", "preconditioner": "none", "tolerance": 0,
                      "minIter": 1}, np.zeros(n), np.zeros(n))
    assert p["singular"]


def load_kat():
    with open(os.path.join(GOLDEN, "pair_agglomeration_kat.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("case", range(3))
def test_pair_agglomeration_hand_worked(case):
    """Hand-worked known answers of the greedy pairing
    (pairGAMGAgglomerate.C:123-301); see the fixture for the derivation."""
    kat = load_kat()[case]
    a = O.Addr(kat["nCells"], kat["lower"], kat["upper"])
    O.reset_direction()
    for step in kat["steps"]:
        assert O.get_direction() == (1 if step["forward"] else 0)
        n, cmap = O.pair_agglomerate(a, kat["weights"])
        assert n == step["nCoarse"]
        assert list(cmap) == step["map"]


def test_pair_agglomeration_invariants():
    rng = np.random.default_rng(12)
    n, lo, up = random_graph(200, 5, rng)
    a = O.Addr(n, lo, up)
    w = rng.uniform(0.1, 1, len(lo))
    for _ in range(2):
        nc, cmap = O.pair_agglomerate(a, w)
        assert cmap.min() == 0 and cmap.max() == nc - 1
        assert len(np.unique(cmap)) == nc
        assert nc < n
        # clusters are connected through faces of the fine graph
        for c in range(nc):
            members = np.where(cmap == c)[0]
            if len(members) == 1:
                continue
            inside = (cmap[lo] == c) & (cmap[up] == c)
            adj = {m: set() for m in members}
            for l, u in zip(lo[inside], up[inside]):
                adj[l].add(u); adj[u].add(l)
            seen, stack = {members[0]}, [members[0]]
            while stack:
                v = stack.pop()
                for x in adj[v]:
                    if x not in seen:
                        seen.add(x); stack.append(x)
            assert len(seen) == len(members)


@pytest.mark.parametrize("kind", ["spd", "asym"])
def test_galerkin_coarse_matrix(kind):
    """agglomerateMatrix == P^T A P with piecewise-constant P."""
    rng = np.random.default_rng(13)
    n, lo, up = box(6, 6, 4)
    a, m, A = mk(n, lo, up, rng, kind)
    O.reset_direction()
    levels = O.agglomerate(m, {"nCellsInCoarsestLevel": 4, "mergeLevels": 1})
    assert len(levels) >= 2
    lev = levels[0]
    cd, cu, cl = O.agglomerate_matrix(lev, m.diag, m.upper, None if kind == "spd" else m.lower)
    nc = lev["nCoarseCells"]
    P = np.zeros((n, nc))
    P[np.arange(n), lev["restrictMap"]] = 1
    Ac = P.T @ A @ P
    Cd = dense(nc, lev["lower"], lev["upper"], cd, cu, cl)
    np.testing.assert_allclose(Cd, Ac, rtol=1e-12, atol=1e-12)
    # coarse faces are upper-triangular: owner < neighbour, owners sorted
    assert np.all(lev["lower"] < lev["upper"])
    assert np.all(np.diff(lev["lower"]) >= 0)


def test_merge_levels_composes_restriction():
    rng = np.random.default_rng(14)
    n, lo, up = box(8, 8, 4)
    a, m, A = mk(n, lo, up, rng)
    O.reset_direction()
    one = O.agglomerate(m, {"nCellsInCoarsestLevel": 2, "mergeLevels": 1})
    O.reset_direction()
    two = O.agglomerate(m, {"nCellsInCoarsestLevel": 2, "mergeLevels": 2})
    # level 0 of mergeLevels 2 = level 1 map composed with level 0 map
    comp = one[1]["restrictMap"][one[0]["restrictMap"]]
    np.testing.assert_array_equal(two[0]["restrictMap"], comp)
    assert two[0]["nCoarseCells"] == one[1]["nCoarseCells"]


def _gs(A, psi, b):
    """One forward Gauss-Seidel sweep as a triangular solve."""
    DL = np.tril(A)
    U = np.triu(A, 1)
    return np.linalg.solve(DL, b - U @ psi)


def _scale(A, field, source):
    """GAMGSolverScale.C:31-76 in dense form."""
    Af = A @ field
    sf = (source @ field) / (Af @ field)
    return sf * field + (source - sf * Af) / np.diag(A)


def _dense_vcycle(A, Ps, b, psi, scale, nPost, maxPost, nFinest, coarse_solve=None):
    """The V-cycle of GAMGSolverSolve.C:148-443 (no pre-sweeps, Gauss-Seidel)
    composed from dense Galerkin products; the coarsest level is solved
    exactly unless coarse_solve(Ac, src) is given."""
    As = [A]
    for P in Ps:
        As.append(P.T @ As[-1] @ P)
    r = b - A @ psi
    src = [Ps[0].T @ r]
    for P in Ps[1:]:
        src.append(P.T @ src[-1])
    L = len(Ps)
    corr = [None] * L
    corr[L - 1] = (np.linalg.solve(As[L], src[L - 1]) if coarse_solve is None
                   else coarse_solve(As[L], src[L - 1]))
    for l in range(L - 2, -1, -1):
        c = Ps[l + 1] @ corr[l + 1]
        if scale and l < L - 2:
            c = _scale(As[l + 1], c, src[l])
        for _ in range(min(nPost + l, maxPost)):
            c = _gs(As[l + 1], c, src[l])
        corr[l] = c
    f = Ps[0] @ corr[0]
    if scale:
        f = _scale(A, f, r)
    psi = psi + f
    for _ in range(nFinest):
        psi = _gs(A, psi, b)
    return psi


@pytest.mark.parametrize("minCells,scale,nPost,nFinest", [
    (40, 0, 0, 0),      # two levels: pure coarse-grid correction
    (40, 1, 0, 0),      # + finest correction scaling
    (40, 0, 0, 2),      # + finest smoothing
    (9, 1, 2, 1),       # several levels: intermediate smoothing and scaling
    (9, 0, 3, 2),
])
def test_gamg_vcycle_closed_form(minCells, scale, nPost, nFinest):
    """One V-cycle of the oracle = the dense composition of Galerkin
    restriction, exact coarsest solve, prolongation, scaling and
    Gauss-Seidel sweeps built from the piecewise-constant prolongators."""
    rng = np.random.default_rng(16)
    n, lo, up = box(6, 5, 4)
    a, m, A = mk(n, lo, up, rng)
    b = rng.standard_normal(n)
    O.reset_direction()
    levels = O.agglomerate(m, {"nCellsInCoarsestLevel": minCells, "mergeLevels": 1})
    if minCells == 40:
        assert len(levels) == 1
    else:
        assert len(levels) >= 3
    Ps = []
    for lev in levels:
        P = np.zeros((lev["nFineCells"], lev["nCoarseCells"]))
        P[np.arange(lev["nFineCells"]), lev["restrictMap"]] = 1
        Ps.append(P)
    expect = _dense_vcycle(A, Ps, b, np.zeros(n), scale, nPost, 4, nFinest)
    O.reset_direction()
    ctl = {"solver": "GAMG", "smoother": "GaussSeidel", "tolerance": 0, "relTol": 0,
           "maxIter": 1, "nCellsInCoarsestLevel": minCells, "directSolveCoarsest": 1,
           "scaleCorrection": scale, "nPostSweeps": nPost, "nFinestSweeps": nFinest}
    got, perf = m.solve(ctl, np.zeros(n), b)
    assert perf["nIterations"] == 1
    np.testing.assert_allclose(got, expect, rtol=1e-9, atol=1e-12 * np.abs(expect).max())


def test_gamg_preconditioner_two_cycles():
    """GAMGPreconditioner with nVcycles 2 = a second V-cycle on the
    residual left by the first (GAMGPreconditioner.C:87-148), checked as
    PCG+GAMG reaching the dense solution and against a 1-cycle run."""
    rng = np.random.default_rng(17)
    n, lo, up = box(6, 5, 4)
    a, m, A = mk(n, lo, up, rng)
    b = rng.standard_normal(n)
    O.reset_direction()
    levels = O.agglomerate(m, {"nCellsInCoarsestLevel": 40, "mergeLevels": 1})
    P = np.zeros((n, levels[0]["nCoarseCells"]))
    P[np.arange(n), levels[0]["restrictMap"]] = 1
    # one GAMG solver cycle from psi = 0 and a second from its result
    # compose as the preconditioner's two cycles on r = b
    w1 = _dense_vcycle(A, [P], b, np.zeros(n), 0, 0, 4, 1)
    w2 = _dense_vcycle(A, [P], b, w1, 0, 0, 4, 1)
    O.reset_direction()
    ctl = {"solver": "GAMG", "smoother": "GaussSeidel", "tolerance": 0, "relTol": 0,
           "maxIter": 2, "nCellsInCoarsestLevel": 40, "directSolveCoarsest": 1,
           "scaleCorrection": 0, "nFinestSweeps": 1}
    got, _ = m.solve(ctl, np.zeros(n), b)
    np.testing.assert_allclose(got, w2, rtol=1e-9, atol=1e-12 * np.abs(w2).max())


def _prolongators(levels):
    Ps = []
    for lev in levels:
        P = np.zeros((lev["nFineCells"], lev["nCoarseCells"]))
        P[np.arange(lev["nFineCells"]), lev["restrictMap"]] = 1
        Ps.append(P)
    return Ps


@pytest.mark.parametrize("minCells,tol,relTol", [
    (40, 1e-3, 0.1),     # relTol stops the coarsest PCG
    (40, 1e-2, 0),       # tolerance stops it
    (9, 0, 0.05),        # several levels; relTol only
    (40, 0, 0),          # neither: PCG runs to its singularity test or maxIter
])
def test_gamg_coarsest_iterative_solve(minCells, tol, relTol):
    """directSolveCoarsest no (GAMGSolverSolve.C:520-552, GAMGSolver.C:270-322):
    the coarsest correction is PCG+DIC from zero with the GAMG solver's own
    tolerance and relTol and the default maxIter 1000.  One oracle V-cycle
    equals the dense composition whose coarsest solve is a textbook PCG with
    the textbook DIC factor and those controls; an exact solve, a zero
    initial guess dropped or another tolerance would give a different
    cycle."""
    rng = np.random.default_rng(21)
    n, lo, up = box(6, 5, 4)
    a, m, A = mk(n, lo, up, rng)
    b = rng.standard_normal(n)
    O.reset_direction()
    levels = O.agglomerate(m, {"nCellsInCoarsestLevel": minCells, "mergeLevels": 1})
    Ps = _prolongators(levels)
    nIts = []

    def coarse(Ac, src):
        x, it, _ = NS.pcg(Ac, src, np.zeros(src.size), NS.dic_dense(Ac), tol, relTol)
        nIts.append(it)
        return x

    expect = _dense_vcycle(A, Ps, b, np.zeros(n), 0, 1, 4, 1, coarse_solve=coarse)
    exact = _dense_vcycle(A, Ps, b, np.zeros(n), 0, 1, 4, 1)
    O.reset_direction()
    ctl = {"solver": "GAMG", "smoother": "GaussSeidel", "tolerance": tol, "relTol": relTol,
           "maxIter": 1, "minIter": 1, "nCellsInCoarsestLevel": minCells,
           "directSolveCoarsest": 0, "scaleCorrection": 0, "nPostSweeps": 1,
           "nFinestSweeps": 1}
    got, perf = m.solve(ctl, np.zeros(n), b)
    assert perf["nIterations"] == 1
    np.testing.assert_allclose(got, expect, rtol=1e-9, atol=1e-11 * np.abs(expect).max())
    if tol > 0 or relTol > 0:
        # the coarsest PCG stopped early: the cycle is not the exact one
        assert 0 < nIts[0] < levels[-1]["nCoarseCells"]
        assert np.abs(got - exact).max() > 1e-6 * np.abs(exact).max()


def test_gamg_preconditioner_subdict_controls():
    """A GAMG preconditioner reads its controls from the preconditioner
    sub-dictionary (lduMatrixPreconditioner.C:72-82 hands e.dict() to
    GAMGPreconditioner, GAMGPreconditioner.C:44-63), not from the solver's
    dictionary.  The sub-dictionary's smoother, sweeps, coarsest settings
    and nVcycles differ from the parent's here; the oracle's PCG+GAMG
    residual history equals a textbook PCG whose preconditioner is the
    dense V-cycle of the sub-dictionary's settings, and differs from the
    one the parent's settings would give."""
    rng = np.random.default_rng(22)
    n, lo, up = box(6, 5, 4)
    a, m, A = mk(n, lo, up, rng)
    b = rng.standard_normal(n)
    sub = {"preconditioner": "GAMG", "smoother": "GaussSeidel", "nCellsInCoarsestLevel": 40,
           "directSolveCoarsest": 1, "scaleCorrection": 0, "nPostSweeps": 0,
           "nFinestSweeps": 1, "nVcycles": 1, "tolerance": 1e-3}
    # 30 iterations: a forward Gauss-Seidel V-cycle is not a symmetric
    # preconditioner, so PCG stagnates near 1e-4 instead of converging
    parent = {"solver": "PCG", "preconditioner": sub, "tolerance": 1e-10, "relTol": 0,
              "maxIter": 30, "smoother": "symGaussSeidel", "nCellsInCoarsestLevel": 10,
              "nFinestSweeps": 3, "nVcycles": 2, "directSolveCoarsest": 0}
    O.reset_direction()
    Ps = _prolongators(O.agglomerate(m, {"nCellsInCoarsestLevel": 40, "mergeLevels": 1}))

    def vcycles(r, nV, nFinest):
        w = np.zeros(n)
        for _ in range(nV):
            w = _dense_vcycle(A, Ps, r, w, 0, 0, 4, nFinest)
        return w

    x, its, hist = NS.pcg(A, b, np.zeros(n), lambda r: vcycles(r, 1, 1), 1e-10, 0, maxIter=30)
    _, _, hist_p = NS.pcg(A, b, np.zeros(n), lambda r: vcycles(r, 2, 3), 1e-10, 0, maxIter=30)
    O.reset_direction()
    got, perf, ohist = m.solve(parent, np.zeros(n), b, history=True)
    assert perf["nIterations"] == its == 30
    np.testing.assert_allclose(ohist[:31], hist, rtol=1e-8)
    np.testing.assert_allclose(got, x, rtol=1e-8, atol=1e-10 * np.abs(x).max())
    # the parent's settings give another history
    k = min(len(hist_p), len(hist))
    assert np.abs(hist_p[1:k] / hist[1:k] - 1).max() > 1e-1


def test_lu_matches_numpy():
    rng = np.random.default_rng(15)
    A = rng.standard_normal((30, 30))
    A[0, 0] = 0.0     # forces a pivot
    b = rng.standard_normal(30)
    LU, piv = O.lu_decompose(A)
    x = O.lu_backsubstitute(LU, piv, b)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10)


def uniform_case(n, fixed=(0,)):
    """Uniform unit-box geometry, as the case generator defines it."""
    N, lo, up = box(n, n, n)
    h = 1.0 / n
    F = len(lo)
    dirs = np.where(up - lo == 1, 0, np.where(up - lo == n, 1, 2))
    bfc, bpatch = [], []
    for p in range(6):
        axis, hi = p // 2, p % 2
        for k in range(n):
            for j in range(n):
                for i in range(n):
                    idx = (i, j, k)
                    if idx[axis] == (n - 1 if hi else 0):
                        bfc.append(i + n * (j + n * k)); bpatch.append(p)
    bfc, bpatch = np.array(bfc, np.int32), np.array(bpatch)
    return dict(N=N, lo=lo, up=up, h=h, dirs=dirs, w=np.full(F, 0.5), magSf=np.full(F, h * h),
                dc=np.full(F, 1 / h), bfc=bfc, bpatch=bpatch, bMagSf=np.full(len(bfc), h * h),
                bdc=np.full(len(bfc), 2 / h),
                bType=np.isin(bpatch, fixed).astype(np.int32), bValue=np.zeros(len(bfc)))


def test_laplacian_closed_form_uniform_box():
    c = uniform_case(6)
    a = O.Addr(c["N"], c["lo"], c["up"])
    gamma = np.ones(c["N"])
    upper, diag, ic, bc = O.laplacian_assemble(a, gamma, c["w"], c["magSf"], c["dc"], c["bfc"],
                                               c["bMagSf"], c["bdc"], c["bType"], c["bValue"])
    h = c["h"]
    np.testing.assert_allclose(upper, h, rtol=1e-15)
    # interior rows sum to zero before the boundary terms
    np.testing.assert_allclose(diag + O.Matrix(a, np.zeros(c["N"]), upper).sumA(), 0, atol=1e-15)
    # fixedValue faces add -2 gamma h to the diagonal, zeroGradient 0
    np.testing.assert_allclose(ic[c["bType"] == 1], -2 * h, rtol=1e-15)
    assert np.all(ic[c["bType"] == 0] == 0)
    d2, s2 = O.add_boundary(c["bfc"], ic, bc, diag, np.zeros(c["N"]))
    A = dense(c["N"], c["lo"], c["up"], d2, upper)
    # A*1 is non-zero only in cells touching the fixedValue wall
    row = A @ np.ones(c["N"])
    wall = np.zeros(c["N"], bool)
    wall[c["bfc"][c["bType"] == 1]] = True
    np.testing.assert_allclose(row[~wall], 0, atol=1e-14)
    np.testing.assert_allclose(row[wall], -2 * h, rtol=1e-12)


def test_laplacian_variable_gamma_vs_dense_fv():
    """Two-point flux with linear face interpolation, built independently."""
    rng = np.random.default_rng(16)
    c = uniform_case(5)
    a = O.Addr(c["N"], c["lo"], c["up"])
    gamma = np.exp(rng.standard_normal(c["N"]))
    upper, diag, ic, bc = O.laplacian_assemble(a, gamma, c["w"], c["magSf"], c["dc"], c["bfc"],
                                               c["bMagSf"], c["bdc"], c["bType"], c["bValue"])
    N = c["N"]
    A = np.zeros((N, N))
    for l, u in zip(c["lo"], c["up"]):
        g = 0.5 * (gamma[l] + gamma[u]) * c["h"] ** 2 / c["h"]
        A[l, u] += g; A[u, l] += g; A[l, l] -= g; A[u, u] -= g
    for b, cell in enumerate(c["bfc"]):
        if c["bType"][b]:
            A[cell, cell] -= gamma[cell] * c["h"] ** 2 * 2 / c["h"]
    d2, _ = O.add_boundary(c["bfc"], ic, bc, diag, np.zeros(N))
    np.testing.assert_allclose(dense(N, c["lo"], c["up"], d2, upper), A, rtol=1e-12, atol=1e-15)


def test_upwind_vs_dense_fv():
    rng = np.random.default_rng(17)
    c = uniform_case(4)
    a = O.Addr(c["N"], c["lo"], c["up"])
    phi = rng.standard_normal(len(c["lo"]))
    bPhi = rng.standard_normal(len(c["bfc"]))
    lo_, up_, dg, ic, bc = O.upwind_assemble(a, phi, c["bType"], bPhi, c["bValue"])
    N = c["N"]
    A = np.zeros((N, N))
    for f, (l, u) in enumerate(zip(c["lo"], c["up"])):
        if phi[f] >= 0:      # upwind value from the owner
            A[l, l] += phi[f]; A[u, l] -= phi[f]
        else:                # from the neighbour
            A[l, u] += phi[f]; A[u, u] -= phi[f]
    np.testing.assert_allclose(dense(N, c["lo"], c["up"], dg, up_, lo_), A, atol=1e-14)
    zg = c["bType"] == 0
    np.testing.assert_allclose(ic[zg], bPhi[zg])
    np.testing.assert_allclose(bc[~zg], -bPhi[~zg] * c["bValue"][~zg])


def test_gauss_grad_exact_for_linear_fields():
    c = uniform_case(6)
    a = O.Addr(c["N"], c["lo"], c["up"])
    n, h = 6, c["h"]
    ii = np.arange(c["N"])
    C = np.stack([(ii % n + 0.5) * h, ((ii // n) % n + 0.5) * h, (ii // (n * n) + 0.5) * h], 1)
    g = np.array([1.5, -2.0, 0.25])
    phi = C @ g + 3.0
    F = len(c["lo"])
    Sf = np.zeros((F, 3))
    Sf[np.arange(F), c["dirs"]] = h * h
    bSf = np.zeros((len(c["bfc"]), 3))
    axis = c["bpatch"] // 2
    sign = np.where(c["bpatch"] % 2 == 1, 1.0, -1.0)
    bSf[np.arange(len(c["bfc"])), axis] = sign * h * h
    # boundary values of the linear field at the face centres (fixed value)
    bCf = C[c["bfc"]].copy()
    bCf[np.arange(len(c["bfc"])), axis] += sign * h / 2
    bPhi = bCf @ g + 3.0
    grad, bGrad = O.grad_gauss_linear(a, phi, c["w"], Sf.ravel(), np.full(c["N"], h ** 3),
                                      c["bfc"], bSf.ravel(), c["bMagSf"], bPhi,
                                      (sign * g[axis]))
    np.testing.assert_allclose(grad.reshape(-1, 3), np.tile(g, (c["N"], 1)), rtol=1e-10)
    np.testing.assert_allclose(bGrad.reshape(-1, 3), np.tile(g, (len(c["bfc"]), 1)), rtol=1e-10)


def test_gauss_grad_boundary_correction():
    """gradScheme::correctBoundaryConditions (gradScheme.C:123-148): on a
    non-coupled patch the boundary gradient is the adjacent cell's gradient
    with its normal component replaced by the patch snGrad,
    g_b = g_c + n (snGrad - n.g_c).  With a non-linear cell field and
    snGrad unrelated to n.g_c the correction term is far from zero; the
    normal component must equal snGrad and the tangential components the
    cell's, on every patch (both normal signs)."""
    c = uniform_case(5)
    a = O.Addr(c["N"], c["lo"], c["up"])
    n, h = 5, c["h"]
    rng = np.random.default_rng(23)
    phi = rng.uniform(-1, 1, c["N"])
    F = len(c["lo"])
    Sf = np.zeros((F, 3))
    Sf[np.arange(F), c["dirs"]] = h * h
    nB = len(c["bfc"])
    axis = c["bpatch"] // 2
    sign = np.where(c["bpatch"] % 2 == 1, 1.0, -1.0)
    bSf = np.zeros((nB, 3))
    bSf[np.arange(nB), axis] = sign * h * h
    bPhi = rng.uniform(-1, 1, nB)
    snGrad = rng.uniform(-5, 5, nB)
    grad, bGrad = O.grad_gauss_linear(a, phi, c["w"], Sf.ravel(), np.full(c["N"], h ** 3),
                                      c["bfc"], bSf.ravel(), c["bMagSf"], bPhi, snGrad)
    grad, bGrad = grad.reshape(-1, 3), bGrad.reshape(-1, 3)
    nrm = bSf / c["bMagSf"][:, None]
    gc = grad[c["bfc"]]
    # the correction is not small here
    assert np.abs(snGrad - (nrm * gc).sum(1)).min() > 1e-3
    np.testing.assert_allclose((nrm * bGrad).sum(1), snGrad, rtol=1e-12, atol=1e-12)
    tang = bGrad - nrm * (nrm * bGrad).sum(1)[:, None]
    np.testing.assert_allclose(tang, gc - nrm * (nrm * gc).sum(1)[:, None], rtol=1e-12,
                               atol=1e-12)
    # the cell gradient itself: Gauss theorem with the face values
    # w phi_P + (1 - w) phi_N and the patch values, over the volume
    gexp = np.zeros((c["N"], 3))
    phif = c["w"] * phi[c["lo"]] + (1 - c["w"]) * phi[c["up"]]
    np.add.at(gexp, c["lo"], Sf * phif[:, None])
    np.subtract.at(gexp, c["up"], Sf * phif[:, None])
    np.add.at(gexp, c["bfc"], bSf * bPhi[:, None])
    np.testing.assert_allclose(grad, gexp / h ** 3, rtol=1e-12, atol=1e-9)


# ---- lduMatrix H, H1, faceH, sumMagOffDiag -------------------------------

def test_h_operators_hand_worked():
    """Three cells in a line, faces (0,1) and (1,2), upper != lower so a
    swapped role or sign changes the answer; values worked by hand from
    the face loops of lduMatrixTemplates.C:33-111, lduMatrixATmul.C:298-324
    and lduMatrixOperations.C:77-97."""
    lo = np.array([0, 1], np.int32)
    up = np.array([1, 2], np.int32)
    a = O.Addr(3, lo, up)
    m = O.Matrix(a, [1.0, 1.0, 1.0], [2.0, 3.0], [5.0, 7.0])
    psi = np.array([1.0, 10.0, 100.0])
    # H[0] = -2*10; H[1] = -5*1 - 3*100; H[2] = -7*10
    assert m.H(psi).tolist() == [-20.0, -305.0, -70.0]
    assert m.H1().tolist() == [-2.0, -8.0, -7.0]
    # faceH = upper*psi(neighbour) - lower*psi(owner)
    assert m.faceH(psi).tolist() == [2.0 * 10 - 5.0 * 1, 3.0 * 100 - 7.0 * 10]
    mm = O.Matrix(a, [1.0, 1.0, 1.0], [-2.0, 3.0], [5.0, -7.0])
    assert mm.sumMagOffDiag([0.5, 0.0, 0.0]).tolist() == [2.5, 8.0, 7.0]
    # a vector field: component by component
    v = np.stack([psi, 2 * psi, -psi], axis=1)
    np.testing.assert_array_equal(m.H(v), np.stack([m.H(psi), m.H(2 * psi), m.H(-psi)], 1))
    np.testing.assert_array_equal(m.faceH(v)[:, 2], -m.faceH(psi))
    # diagonal-only: H = 0, H1 = 0, faceH is an error
    d = O.Matrix(a, [1.0, 2.0, 3.0])
    assert d.H(psi).tolist() == [0.0, 0.0, 0.0] and d.H1().tolist() == [0.0, 0.0, 0.0]
    with pytest.raises(RuntimeError):
        d.faceH(psi)


@pytest.mark.parametrize("name,gen", MESHES)
@pytest.mark.parametrize("kind", ["spd", "asym"])
def test_h_operators_vs_dense(name, gen, kind):
    rng = np.random.default_rng(5)
    n, lo, up = gen(rng)
    a, m, A = mk(n, lo, up, rng, kind)
    off = A - np.diag(np.diag(A))
    x = rng.standard_normal(n)
    np.testing.assert_allclose(m.H(x), -off @ x, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(m.H1(), -off.sum(axis=1), rtol=1e-13, atol=1e-13)
    s0 = rng.standard_normal(n)
    np.testing.assert_allclose(m.sumMagOffDiag(s0), s0 + np.abs(off).sum(axis=1),
                               rtol=1e-13, atol=1e-13)
    # faceH from the dense coefficients: A[l, u] is the face's upper one
    np.testing.assert_allclose(m.faceH(x), A[lo, up] * x[up] - A[up, lo] * x[lo],
                               rtol=1e-13, atol=1e-13)
    # H of a unit field is H1, to the bit (coefficient * 1 is exact)
    np.testing.assert_array_equal(m.H(np.ones(n)), m.H1())
    # a three-component field is three scalar fields, to the bit
    X = rng.standard_normal((n, 3))
    np.testing.assert_array_equal(m.H(X), np.stack([m.H(X[:, c]) for c in range(3)], 1))


@pytest.mark.parametrize("name,gen", MESHES)
def test_h_operators_identities(name, gen):
    """Negative off-diagonals (a Laplacian): sumMagOffDiag from zero is H1
    to the bit.  Symmetric: the discrete divergence of faceH (out of the
    owner, into the neighbour) is psi*H1 - H(psi)."""
    rng = np.random.default_rng(6)
    n, lo, up = gen(rng)
    a = O.Addr(n, lo, up)
    d, u = spd_coeffs(n, lo, up, rng)
    assert np.all(u <= 0)
    m = O.Matrix(a, d, u)
    # 0 + |a| + |b| and (0 - a) - b round alike for a, b <= 0
    np.testing.assert_array_equal(m.sumMagOffDiag(np.zeros(n)), m.H1())
    x = rng.standard_normal(n)
    fH = m.faceH(x)
    div = np.zeros(n)
    np.add.at(div, lo, fH)
    np.add.at(div, up, -fH)
    np.testing.assert_allclose(div, x * m.H1() - m.H(x), rtol=1e-12, atol=1e-12)


# ---- fvMatrix::relax --------------------------------------------------------

def test_relax_hand_worked():
    """fvMatrix.C:594-740 on three cells in a line, one boundary face at each
    end; values worked by hand (all exact in binary)."""
    lo = np.array([0, 1], np.int32)
    up = np.array([1, 2], np.int32)
    a = O.Addr(3, lo, up)
    m = O.Matrix(a, [1.5, 2.0, 3.5], [-1.0, -2.0])
    bfc = np.array([0, 2], np.int32)
    psi = np.array([1.0, 2.0, 3.0])
    # sumOff = [1, 3, 2]; D + max|ic| = [5.5, 2, 4]; max(|D|, sumOff) = [5.5, 3, 4];
    # /0.5 = [11, 6, 8]; - min ic = [15, 6, 7.5]; S += (D - D0) psi
    d, s = m.relax(bfc, [-4.0, 0.5], 0.5, psi, [10.0, 20.0, 30.0])
    assert d.tolist() == [15.0, 6.0, 7.5]
    assert s.tolist() == [23.5, 28.0, 42.0]
    # three components: the largest magnitude in, the smallest value out
    ic3 = np.array([[-4.0, 1.0, 2.0], [0.5, -0.25, 0.75]])
    P3 = np.array([[1.0, 0.0, -1.0], [2.0, 1.0, 0.0], [3.0, 0.0, 1.0]])
    d, s = m.relax(bfc, ic3, 0.5, P3, np.zeros((3, 3)))
    assert d.tolist() == [15.0, 6.0, 8.75]
    assert s.tolist() == [[13.5, 0.0, -13.5], [8.0, 4.0, 0.0], [15.75, 0.0, 5.25]]
    # alpha <= 0: nothing changes
    d, s = m.relax(bfc, [-4.0, 0.5], 0.0, psi, [10.0, 20.0, 30.0])
    assert d.tolist() == [1.5, 2.0, 3.5] and s.tolist() == [10.0, 20.0, 30.0]


@pytest.mark.parametrize("name,gen", MESHES)
@pytest.mark.parametrize("kind", ["spd", "asym"])
def test_relax_invariants(name, gen, kind):
    rng = np.random.default_rng(8)
    n, lo, up = gen(rng)
    a, m, A = mk(n, lo, up, rng, kind)
    nB = 2 * n
    bfc = rng.integers(0, n, nB).astype(np.int32)
    ic = rng.standard_normal((nB, 3))
    psi = rng.standard_normal((n, 3))
    S = rng.standard_normal((n, 3))
    d1, s1 = m.relax(bfc, ic, 0.7, psi, S)
    # the relaxed equation has the same residual at psi: A' psi - S' = A psi - S
    A1 = A - np.diag(np.diag(A)) + np.diag(d1)
    np.testing.assert_allclose(A1 @ psi - s1, A @ psi - S, rtol=1e-11, atol=1e-11)
    # (D + sum of the faces' smallest coefficients) * alpha is the dominant
    # diagonal, the same for every alpha
    mn = np.zeros(n)
    np.add.at(mn, bfc, ic.min(axis=1))
    d2, _ = m.relax(bfc, ic, 0.35, psi, S)
    np.testing.assert_allclose((d1 + mn) * 0.7, (d2 + mn) * 0.35, rtol=1e-13, atol=1e-13)
    dom = (d1 + mn) * 0.7
    off = np.abs(A - np.diag(np.diag(A))).sum(axis=1)
    mx = np.zeros(n)
    np.add.at(mx, bfc, np.abs(ic).max(axis=1))
    np.testing.assert_allclose(dom, np.maximum(np.abs(np.diag(A) + mx), off), rtol=1e-13)


# ---- faceAreaPair agglomeration --------------------------------------------

def _box_face_areas(nx, ny, nz, lo, up, area=1.0):
    """Axis-aligned face area vectors of a box mesh: the direction from the
    index step between owner and neighbour."""
    step = up - lo
    Sf = np.zeros((len(lo), 3))
    Sf[step == 1, 0] = area
    Sf[step == nx, 1] = area
    Sf[step == nx * ny, 2] = area
    return Sf, np.full(len(lo), area)


def test_face_area_weights_axis_aligned():
    """faceAreaPairGAMGAgglomeration.C:55-80: an x, y or z face of area A
    weighs sqrt(A) times 1, 1.01 or 1.02."""
    A = 2.25
    Sf = np.array([[A, 0, 0], [0, A, 0], [0, 0, A], [0, 0, -A]])
    w = O.face_area_weights(Sf, np.full(4, A))
    np.testing.assert_allclose(w, np.sqrt(A) * np.array([1.0, 1.01, 1.02, 1.02]), rtol=1e-15)


def test_face_area_pair_agglomerates_along_z():
    """On a uniform box the z faces weigh most, so the first level pairs
    every cell with its z neighbour (nz even); algebraicPair with uniform
    coefficients does not."""
    nx, ny, nz = 4, 3, 6
    n, lo, up = box(nx, ny, nz)
    a = O.Addr(n, lo, up)
    Sf, magSf = _box_face_areas(nx, ny, nz, lo, up)
    m = O.Matrix(a, np.full(n, -6.0), np.ones(len(lo)))
    O.reset_direction()
    lev = O.agglomerate(m, {"nCellsInCoarsestLevel": 2, "agglomerator": "faceAreaPair",
                            "Sf": Sf, "magSf": magSf})[0]
    rmap = lev["restrictMap"]
    assert lev["nCoarseCells"] == n // 2
    for cc in range(n // 2):
        cells = np.flatnonzero(rmap == cc)
        assert len(cells) == 2 and cells[1] - cells[0] == nx * ny
    O.reset_direction()
    lev_alg = O.agglomerate(m, {"nCellsInCoarsestLevel": 2})[0]
    assert not np.array_equal(lev_alg["restrictMap"], rmap)


def test_flux_hand_worked():
    """fvMatrix.C:953-1028 on three cells in a line with a boundary face at
    each end: internal faces upper psi(N) - lower psi(P), boundary faces
    ic psi(cell) - bc, plus the correction."""
    lo = np.array([0, 1], np.int32)
    up = np.array([1, 2], np.int32)
    m = O.Matrix(O.Addr(3, lo, up), [1.0, 1.0, 1.0], [2.0, 3.0], [5.0, 7.0])
    psi = np.array([1.0, 10.0, 100.0])
    fl, bf = m.flux([0, 2], [-4.0, 0.5], [8.0, -1.0], psi)
    assert fl.tolist() == [15.0, 230.0]
    assert bf.tolist() == [-4.0 * 1 - 8.0, 0.5 * 100 + 1.0]
    fl, bf = m.flux([0, 2], [-4.0, 0.5], [8.0, -1.0], psi, [0.5, -0.25], [1.0, 2.0])
    assert fl.tolist() == [15.5, 229.75] and bf.tolist() == [-11.0, 53.0]
    # components: cmptMultiply on the boundary
    P3 = np.stack([psi, -psi, 2 * psi], 1)
    ic3 = np.array([[1.0, 2.0, 3.0], [0.5, 0.25, -1.0]])
    bc3 = np.zeros((2, 3))
    fl, bf = m.flux([0, 2], ic3, bc3, P3)
    assert bf.tolist() == [[1.0, -2.0, 6.0], [50.0, -25.0, -200.0]]
    np.testing.assert_array_equal(fl[:, 1], -m.faceH(psi))


# ---------------------------------------------------------------------------
# coupled (cyclic) interfaces: lduMatrixATmul.C:34-280 with
# interfaceBouCoeffs / interfaceIntCoeffs, lduMatrixUpdateMatrixInterfaces.C,
# cyclicFvPatchField.C:150-174, GaussSeidelSmoother.C:117-143
# ---------------------------------------------------------------------------

from meshes import cyclic_box, cyclic_system, dense_interfaces  # noqa: E402


def _cyclic_system(kind, seed=3, shape=(6, 5, 4)):
    n, lower, upper, diag, up, lo, ifc, A, AT = cyclic_system(kind, seed, shape)
    return n, O.Addr(n, lower, upper), diag, up, lo, ifc, A, AT


@pytest.mark.parametrize("kind", ["sym", "asym"])
def test_interfaces_vs_dense(kind):
    """Amul, Tmul, residual and sumA with coupled interfaces = the dense
    matrix with A[faceCells, nbrCells] -= interface coefficients (Tmul:
    the transposed internal part and the interfaceIntCoeffs)."""
    n, a, diag, up, lo, ifc, A, AT = _cyclic_system(kind)
    m = O.Matrix(a, diag, up, lo, interfaces=ifc)
    rng = np.random.default_rng(9)
    psi, b = rng.standard_normal(n), rng.standard_normal(n)
    np.testing.assert_allclose(m.Amul(psi), A @ psi, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(m.Tmul(psi), AT @ psi, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(m.residual(psi, b), b - A @ psi, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(m.sumA(), A.sum(axis=1), rtol=1e-13, atol=1e-13)
    # a plain matrix is unchanged by the extension
    m0 = O.Matrix(a, diag, up, lo)
    np.testing.assert_allclose(m0.Amul(psi), (A - dense_interfaces(n, ifc, ifc["bouCoeffs"])) @ psi,
                               rtol=1e-13, atol=1e-13)


def test_cyclic_laplacian_is_periodic():
    """fvm::laplacian(1, p) on a uniform box with an x cyclic pair (coupled
    patch: internalCoeffs = boundaryCoeffs = -|Sf|*deltaCoeffs, the other
    patches zeroGradient): A sin(2 pi i/nx) = 2 (cos(2 pi/nx) - 1) h^-2 V
    sin(2 pi i/nx), the periodic second difference, exactly up to
    rounding."""
    nx, ny, nz = 8, 3, 2
    n, lower, upper, ifc = cyclic_box(nx, ny, nz)
    h = 1.0 / nx
    # unit cube cells of size h in every direction: |Sf| deltaCoeffs = h
    up = np.full(len(lower), h)
    diag = np.zeros(n)
    a = O.Addr(n, lower, upper)
    O.lib().or_negSumDiag(ctypes.byref(a.s), O._p(up), O._p(up), O._p(diag))
    bou = np.full(len(ifc["faceCells"]), -h)
    np.add.at(diag, ifc["faceCells"], bou)
    m = O.Matrix(a, diag, up, interfaces=dict(ifc, bouCoeffs=bou))
    i = np.arange(n) % nx
    f = np.sin(2 * np.pi * i / nx)
    lam = 2 * (np.cos(2 * np.pi / nx) - 1) * h
    np.testing.assert_allclose(m.Amul(f), lam * f, atol=1e-14)
    # a constant field is in the null space (all patches non-fixed)
    np.testing.assert_allclose(m.Amul(np.ones(n)), 0.0, atol=1e-14)


@pytest.mark.parametrize("ctl,kind", [
    ({"solver": "PCG", "preconditioner": "DIC"}, "sym"),
    ({"solver": "PCG", "preconditioner": "none"}, "sym"),
    ({"solver": "PBiCGStab", "preconditioner": "DILU"}, "asym"),
    ({"solver": "PBiCG", "preconditioner": "DILU"}, "asym"),
    ({"solver": "smoothSolver", "smoother": "symGaussSeidel", "nSweeps": 2}, "sym"),
    ({"solver": "smoothSolver", "smoother": "GaussSeidel", "nSweeps": 1}, "asym"),
    ({"solver": "smoothSolver", "smoother": "DIC", "nSweeps": 1}, "sym"),
])
def test_solvers_with_interfaces_converge(ctl, kind):
    """The Krylov solvers and the smoothSolver with coupled interfaces
    converge to the solution of the dense coupled matrix; the reported
    residual is the reference's, recomputed with the interfaces."""
    n, a, diag, up, lo, ifc, A, AT = _cyclic_system(kind, seed=5)
    m = O.Matrix(a, diag, up, lo, interfaces=ifc)
    b = np.random.default_rng(2).standard_normal(n)
    ctl = dict(ctl, tolerance=1e-10, relTol=0, maxIter=5000)
    psi, perf = m.solve(ctl, np.zeros(n), b)
    assert perf["converged"]
    x = np.linalg.solve(A, b)
    assert np.linalg.norm(psi - x) / np.linalg.norm(x) < 1e-7
    # (PBiCGStab and PBiCG report their recurrence residual)
    r = b - A @ psi
    assert np.abs(r).sum() / perf["normFactor"] == pytest.approx(perf["finalResidual"],
                                                                rel=1e-2)


@pytest.mark.parametrize("sym", [False, True])
def test_gauss_seidel_with_interfaces_closed_form(sym):
    """One Gauss-Seidel sweep with coupled interfaces treats them as a
    Jacobi term of the field before the sweep: psi1 = (D + L)^-1 (b - U psi1
    ... ) = (D + L)^-1 (b - C psi0 - U psi0) for the forward sweep (C the
    interfaces' part); the symmetric sweep then runs backward with the same
    interface term: psi2 = (D + U)^-1 (b - C psi0 - L psi1)."""
    n, a, diag, up, lo, ifc, A, AT = _cyclic_system("asym", seed=8)
    m = O.Matrix(a, diag, up, lo, interfaces=ifc)
    rng = np.random.default_rng(4)
    psi0, b = rng.standard_normal(n), rng.standard_normal(n)
    C = dense_interfaces(n, ifc, ifc["bouCoeffs"])
    Ai = A - C
    D = np.diag(np.diag(Ai))
    L = np.tril(Ai, -1)
    U = np.triu(Ai, 1)
    psi1 = np.linalg.solve(D + L, b - C @ psi0 - U @ psi0)
    ref = psi1 if not sym else np.linalg.solve(D + U, b - C @ psi0 - L @ psi1)
    got = m.smooth("symGaussSeidel" if sym else "GaussSeidel", psi0, b, 1)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_gamg_with_interfaces_is_refused():
    n, a, diag, up, lo, ifc, A, AT = _cyclic_system("sym")
    m = O.Matrix(a, diag, up, lo, interfaces=ifc)
    with pytest.raises(RuntimeError):
        m.solve({"solver": "GAMG", "smoother": "GaussSeidel", "tolerance": 1e-8},
                np.zeros(n), np.ones(n))


# ---------------------------------------------------------------------------
# mesh geometry: face::areaAndCentre, primitiveMesh::makeCellCentresAndVols,
# surfaceInterpolation::makeWeights / makeDeltaCoeffs / makeNonOrthDeltaCoeffs
# ---------------------------------------------------------------------------

def _geometry(c):
    return O.mesh_geometry(c.N, c.lower, c.upper, c.bFaceCells, c.faceStart, c.facePoints,
                           c.points)


def test_geometry_uniform_box_closed_form():
    """Without grading or jitter the geometry is the uniform box's closed
    form (BoxCase): face areas and centres, outward boundary normals, cell
    centres and volumes, weights 1/2, delta coefficients 1/h (2/h at the
    boundary)."""
    from foamhip.cases import BoxCase, GradedBoxCase
    c = GradedBoxCase(5, 4, 3, L=(1.0, 2.0, 0.5), zRatio=1.0, jitter=0.0)
    g = _geometry(c)
    b = BoxCase(5, 4, 3, L=(1.0, 2.0, 0.5))
    for k in ("weights", "magSf", "Sf", "deltaCoeffs", "V", "C", "Cf", "bMagSf", "bSf",
              "bDeltaCoeffs", "bCf"):
        np.testing.assert_allclose(g[k], getattr(b, k), rtol=1e-13, atol=1e-14, err_msg=k)
    np.testing.assert_allclose(g["nonOrthDeltaCoeffs"], g["deltaCoeffs"], rtol=1e-13)


def test_geometry_invariants_graded_jittered():
    """On a z-graded (ratio 100), 15%-jittered box: the cell volumes add to
    the box volume; each cell's outward face-area vectors add to zero (the
    vector area of a polygon is half the sum of its edge cross products, so
    shared faces cancel); each volume is a third of the sum of its outward
    Sf . Cf; the weights lie in (0, 1); the grading of the unjittered mesh
    is the expansion ratio."""
    from foamhip.cases import GradedBoxCase
    c = GradedBoxCase(6, 5, 8, L=(1.0, 1.0, 2.0), zRatio=100.0, jitter=0.15, seed=4)
    g = _geometry(c)
    assert g["V"].sum() == pytest.approx(2.0, rel=1e-13)
    assert np.all(g["V"] > 0)
    Sf, Cf = g["Sf"].reshape(-1, 3), g["Cf"].reshape(-1, 3)
    bSf, bCf = g["bSf"].reshape(-1, 3), g["bCf"].reshape(-1, 3)
    tot = np.zeros((c.N, 3))
    np.add.at(tot, c.lower, Sf)
    np.add.at(tot, c.upper, -Sf)
    np.add.at(tot, c.bFaceCells, bSf)
    assert np.abs(tot).max() < 1e-15
    sc = np.zeros(c.N)
    np.add.at(sc, c.lower, (Sf * Cf).sum(1))
    np.add.at(sc, c.upper, -(Sf * Cf).sum(1))
    np.add.at(sc, c.bFaceCells, (bSf * bCf).sum(1))
    np.testing.assert_allclose(g["V"], sc / 3.0, rtol=1e-10)
    assert np.all((g["weights"] > 0) & (g["weights"] < 1))
    u = GradedBoxCase(2, 2, 8, zRatio=100.0, jitter=0.0)
    z = u.points.reshape(-1, 3)[:, 2][::9]       # the corner line i = j = 0
    dz = np.diff(z)
    assert dz[-1] / dz[0] == pytest.approx(100.0, rel=1e-12)
