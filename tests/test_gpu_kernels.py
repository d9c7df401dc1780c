"""GPU parity of the single operations against the oracle, through the
C-ABI.  The reference loops are reproduced in the same per-row order, so
everything except the tree-reduced normFactor is compared bitwise."""
import os

import numpy as np
import pytest

import foamhip as fh
import oracle as O
from meshes import asym_coeffs, box, hub_graph, random_graph, spd_coeffs

pytestmark = pytest.mark.gpu


class env:
    """Temporarily set environment variables read at mesh setup."""

    def __init__(self, **kw):
        self.kw = {k: str(v) for k, v in kw.items() if v is not None}

    def __enter__(self):
        self.old = {k: os.environ.get(k) for k in self.kw}
        os.environ.update(self.kw)

    def __exit__(self, *a):
        for k, v in self.old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


def make(case, rng, kind, column=None):
    """column=None: default schedule (3-D tiles on boxes, the sync-free row
    sweep elsewhere); column=T: the level-partitioned sweep with T x T
    columns; column="parts": the one-part level sweep instead of the row
    sweep."""
    n, lo, up = case
    if column == "parts":
        e = env(FOAMHIP_SWEEP="parts")
    else:
        e = env(FOAMHIP_COLUMN=column, FOAMHIP_SWEEP="levels" if column else None)
    with e:
        mesh = fh.lduAddressing(n, lo, up)
        a = O.Addr(n, lo, up)
        if kind == "spd":
            d, u = spd_coeffs(n, lo, up, rng)
            return mesh, fh.lduMatrix(mesh, d, u), O.Matrix(a, d, u)
        d, u, l = asym_coeffs(n, lo, up, rng)
        return mesh, fh.lduMatrix(mesh, d, u, l), O.Matrix(a, d, u, l)


CASES = [
    ("box8", lambda r: box(8, 8, 8), None),
    ("box_tiles", lambda r: box(37, 20, 19), None),
    ("box_tiles_flat", lambda r: box(16, 48, 3), None),
    ("box_parts", lambda r: box(20, 18, 6), 4),
    ("box_parts_wide", lambda r: box(33, 17, 9), 8),
    ("rand", lambda r: random_graph(500, 6, r), None),
    ("rand_shuffled", lambda r: random_graph(300, 5, r, shuffle_within_owner=True), None),
    ("rand_parts", lambda r: random_graph(500, 6, r), "parts"),
    ("rand_many_chunks", lambda r: random_graph(140000, 6, r, shuffle_within_owner=True,
                                                connect_chain=False), None),
    ("rand_hubs", lambda r: hub_graph(20000, 5, 12, 40, r), None),
    ("gamg_coarse", lambda r: coarse_box_level(40, 1), None),
    ("chain", lambda r: box(70, 1, 1), None),
]


def coarse_box_level(n, level):
    """The addressing of a GAMG coarse level of an n^3 box (pair
    agglomeration with algebraic weights of a Laplacian-like matrix)."""
    N, lo, up = box(n, n, n)
    rng = np.random.default_rng(3)
    d, u = spd_coeffs(N, lo, up, rng)
    O.reset_direction()
    levels = O.agglomerate(O.Matrix(O.Addr(N, lo, up), d, u),
                           {"nCellsInCoarsestLevel": 10, "mergeLevels": 1})
    lev = levels[level - 1]
    return lev["nCoarseCells"], lev["lower"], lev["upper"]


def test_box_schedule_detected():
    mesh = fh.lduAddressing(*box(20, 18, 6))
    info = mesh.info()
    assert (info["boxNx"], info["boxNy"], info["boxNz"]) == (20, 18, 6)
    assert info["nLevels"] == 20 + 18 + 6 - 2
    assert info["tiled"] == 1
    with env(FOAMHIP_COLUMN=4, FOAMHIP_SWEEP="levels"):
        info = fh.lduAddressing(*box(20, 18, 6)).info()
    assert info["nParts"] == 5 * 5 and info["tiled"] == 0
    # not a box: one part, level schedule
    info = fh.lduAddressing(*random_graph(50, 4, np.random.default_rng(0))).info()
    assert info["tiled"] == 0 and info["boxNx"] == 0


@pytest.mark.parametrize("name,gen,col", CASES)
@pytest.mark.parametrize("kind", ["spd", "asym"])
def test_matrix_kernels_bitwise(name, gen, col, kind):
    rng = np.random.default_rng(20)
    mesh, m, om = make(gen(rng), rng, kind, col)
    n = mesh.nCells
    x = rng.standard_normal(n)
    b = rng.standard_normal(n)
    np.testing.assert_array_equal(m.Amul(x), om.Amul(x))
    np.testing.assert_array_equal(m.Tmul(x), om.Tmul(x))
    np.testing.assert_array_equal(m.residual(x, b), om.residual(x, b))
    np.testing.assert_array_equal(m.sumA(), om.sumA())
    # a tree-ordered global sum: equal to rounding
    assert m.normFactor(x, b) == pytest.approx(om.normFactor(x, b), rel=1e-13)


@pytest.mark.parametrize("name,gen,col", CASES)
@pytest.mark.parametrize("kind", ["spd", "asym"])
def test_h_operators_bitwise(name, gen, col, kind):
    """lduMatrix::H (scalar, vector and six-component fields), H1, faceH and
    sumMagOffDiag against the oracle's face loops."""
    rng = np.random.default_rng(21)
    mesh, m, om = make(gen(rng), rng, kind, col)
    n = mesh.nCells
    x = rng.standard_normal(n)
    np.testing.assert_array_equal(m.H(x), om.H(x))
    for nc in (3, 6):
        X = rng.standard_normal((n, nc))
        np.testing.assert_array_equal(m.H(X), om.H(X))
        np.testing.assert_array_equal(m.faceH(X), om.faceH(X))
    np.testing.assert_array_equal(m.H1(), om.H1())
    np.testing.assert_array_equal(m.faceH(x), om.faceH(x))
    s0 = rng.standard_normal(n)
    np.testing.assert_array_equal(m.sumMagOffDiag(s0), om.sumMagOffDiag(s0))


def test_h_operators_edge_cases():
    rng = np.random.default_rng(22)
    n, lo, up = box(5, 4, 3)
    mesh = fh.lduAddressing(n, lo, up)
    x = rng.standard_normal(n)
    # diagonal-only matrix: H = 0, H1 = 0, sumOff unchanged, faceH refused
    d = fh.lduMatrix(mesh, rng.uniform(1, 2, n))
    assert not d.H(x).any() and not d.H1().any()
    s0 = rng.standard_normal(n)
    np.testing.assert_array_equal(d.sumMagOffDiag(s0), s0)
    with pytest.raises(fh.FoamHipError, match="faceH"):
        d.faceH(x)
    # nCmpt out of range
    m = fh.lduMatrix(mesh, *spd_coeffs(n, lo, up, rng))
    with pytest.raises(fh.FoamHipError):
        m.H(rng.standard_normal((n, 10)))


@pytest.mark.parametrize("name,gen,col", CASES)
def test_preconditioners_bitwise(name, gen, col):
    rng = np.random.default_rng(21)
    for kind, precons in [("spd", ["DIC", "FDIC", "diagonal", "none"]),
                          ("asym", ["DILU", "diagonal", "none"])]:
        mesh, m, om = make(gen(rng), rng, kind, col)
        r = rng.standard_normal(mesh.nCells)
        for p in precons:
            if p not in ("none",):
                np.testing.assert_array_equal(m.reciprocalD(p), om.reciprocalD(p), err_msg=p)
            np.testing.assert_array_equal(m.precondition(p, r), om.precondition(p, r), err_msg=p)
        if kind == "asym":
            np.testing.assert_array_equal(m.precondition("DILU", r, transpose=True),
                                          om.precondition("DILU", r, transpose=True))


@pytest.mark.parametrize("name,gen,col", CASES)
def test_smoothers_bitwise(name, gen, col):
    rng = np.random.default_rng(22)
    for kind, smoothers in [("spd", ["GaussSeidel", "symGaussSeidel", "DIC", "FDIC",
                                     "DICGaussSeidel"]),
                            ("asym", ["GaussSeidel", "symGaussSeidel", "DILU",
                                      "DILUGaussSeidel"])]:
        mesh, m, om = make(gen(rng), rng, kind, col)
        x0 = rng.standard_normal(mesh.nCells)
        b = rng.standard_normal(mesh.nCells)
        for s in smoothers:
            for nS in (1, 2):
                np.testing.assert_array_equal(m.smooth(s, x0, b, nS), om.smooth(s, x0, b, nS),
                                              err_msg=f"{s} x{nS}")


@pytest.mark.parametrize("column", [None, 4])
def test_pipelined_sweep_repeatable(column):
    """The tiled and the partitioned sweeps are deterministic: the same bits
    on every run, whatever the order tiles are scheduled in."""
    rng = np.random.default_rng(23)
    mesh, m, om = make(box(40, 40, 10), rng, "spd", column)
    if column:
        assert mesh.info()["nParts"] == 100
    r = rng.standard_normal(mesh.nCells)
    ref = om.precondition("DIC", r)
    for _ in range(5):
        np.testing.assert_array_equal(m.precondition("DIC", r), ref)
    x0 = rng.standard_normal(mesh.nCells)
    ref = om.smooth("symGaussSeidel", x0, r, 2)
    for _ in range(3):
        np.testing.assert_array_equal(m.smooth("symGaussSeidel", x0, r, 2), ref)


def test_row_sweep_repeatable():
    """The sync-free chunk sweep (persistent wavefronts, per-chunk flags)
    gives the same bits on every run."""
    rng = np.random.default_rng(24)
    case = random_graph(200000, 6, rng, connect_chain=False)
    mesh, m, om = make(case, rng, "asym")
    assert mesh.info()["nParts"] == 1
    r = rng.standard_normal(mesh.nCells)
    ref = om.precondition("DILU", r)
    for _ in range(4):
        np.testing.assert_array_equal(m.precondition("DILU", r), ref)
    x0 = rng.standard_normal(mesh.nCells)
    ref = om.smooth("symGaussSeidel", x0, r, 2)
    for _ in range(3):
        np.testing.assert_array_equal(m.smooth("symGaussSeidel", x0, r, 2), ref)
