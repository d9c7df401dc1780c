"""Mesh geometry on the device against the oracle, bitwise (fh_face_geometry,
fh_cell_geometry, fh_surface_interpolation: faceTemplates.C:79-149,
primitiveMeshCellCentresAndVols.C:72-154, surfaceInterpolation.C:135-330,
fvPatch.C:138-164), on the measured GAMG configuration's mesh -- a box graded
along z with expansion ratio 100 and 15% node jitter -- in blockMesh order
(the box path) and renumbered (the general rows); and the GAMG solve of that
configuration (algebraicPair agglomeration, mergeLevels 2, direct coarsest
solve of ~1000 rows, symGaussSeidel) at 40^3 against the oracle, FP64 and
mixed precision."""
import ctypes

import numpy as np
import pytest

import foamhip as fh
import oracle as O
from foamhip.cases import GradedBoxCase
from meshes import renumber_cells
from test_gpu_solvers import ITER_TOL, SOLUTION_TOL, rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(x, dtype=None):
    return torch.from_numpy(np.ascontiguousarray(x if dtype is None else x.astype(dtype))).to("cuda")


def P(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def device_geometry(N, lower, upper, bFaceCells, faceStart, facePoints, points):
    mesh = fh.lduAddressing(N, lower, upper)
    bnd = fh.Boundary(mesh, bFaceCells)
    F, B = len(lower), len(bFaceCells)
    nAll = F + B
    d = {"fs": dev(faceStart, np.int32), "fp": dev(facePoints, np.int32), "pts": dev(points)}
    z = lambda n: torch.empty(n, dtype=torch.float64, device="cuda")
    Cf, Sf, magSf = z(3 * nAll), z(3 * nAll), z(nAll)
    torch.cuda.synchronize()
    fh._check(fh.lib().fh_face_geometry(nAll, P(d["fs"]), P(d["fp"]), P(d["pts"]), P(Cf), P(Sf),
                                        P(magSf)))
    g = {"Cf": Cf[:3 * F], "Sf": Sf[:3 * F], "magSf": magSf[:F], "bCf": Cf[3 * F:],
         "bSf": Sf[3 * F:], "bMagSf": magSf[F:]}
    g = {k: v.contiguous() for k, v in g.items()}
    C, V = z(3 * N), z(N)
    fh._check(fh.lib().fh_cell_geometry(mesh.handle, bnd.handle, P(g["Cf"]), P(g["Sf"]),
                                        P(g["bCf"]), P(g["bSf"]), P(C), P(V)))
    g["C"], g["V"] = C, V
    for k, n in (("weights", F), ("deltaCoeffs", F), ("nonOrthDeltaCoeffs", F),
                 ("bDeltaCoeffs", B), ("bNonOrthDeltaCoeffs", B)):
        g[k] = z(n)
    fh._check(fh.lib().fh_surface_interpolation(
        mesh.handle, bnd.handle, P(C), P(g["Cf"]), P(g["Sf"]), P(g["magSf"]), P(g["bCf"]),
        P(g["bSf"]), P(g["bMagSf"]), P(g["weights"]), P(g["deltaCoeffs"]),
        P(g["nonOrthDeltaCoeffs"]), P(g["bDeltaCoeffs"]), P(g["bNonOrthDeltaCoeffs"])))
    return mesh, {k: host(v) for k, v in g.items()}


def check_bitwise(got, ref):
    for k in ref:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)


@pytest.mark.parametrize("renumber", [False, True])
def test_geometry_bitwise(renumber):
    c = GradedBoxCase(24, 20, 16, zRatio=100.0, jitter=0.15, seed=3)
    N, lower, upper, fs, fp = c.N, c.lower, c.upper, c.faceStart, c.facePoints
    bfc = c.bFaceCells
    if renumber:
        q = np.random.default_rng(5).permutation(N)
        lower, upper, fs, fp, _ = renumber_cells(N, lower, upper, fs, fp, q)
        bfc = q[bfc].astype(np.int32)
    ref = O.mesh_geometry(N, lower, upper, bfc, fs, fp, c.points)
    mesh, got = device_geometry(N, lower, upper, bfc, fs, fp, c.points)
    assert bool(mesh.info()["tiled"]) != renumber
    check_bitwise(got, ref)


CONFIG4 = {"solver": "GAMG", "smoother": "symGaussSeidel", "agglomerator": "algebraicPair",
           "mergeLevels": 2, "nCellsInCoarsestLevel": 1000, "directSolveCoarsest": True,
           "tolerance": 1e-8, "relTol": 0, "maxIter": 500}


def test_config4_gamg_40():
    """The GAMG configuration at 40^3 on its graded, jittered mesh:
    fvm::laplacian with a log-normal diffusivity assembled on the device
    from the device geometry (nonOrthDeltaCoeffs, the corrected scheme's
    implicit part), then the FP64 GAMG against the oracle's (1e-10, +-2
    cycles) and the mixed GAMG's FP64 residual."""
    c = GradedBoxCase(40, zRatio=100.0, jitter=0.15, seed=3)
    mesh, g = device_geometry(c.N, c.lower, c.upper, c.bFaceCells, c.faceStart, c.facePoints,
                              c.points)
    gamma = c.lognormal(9, 1.0, 4.0)
    a = O.Addr(c.N, c.lower, c.upper)
    upper, diag, ic, bc = O.laplacian_assemble(a, gamma, g["weights"], g["magSf"],
                                               g["nonOrthDeltaCoeffs"], c.bFaceCells, g["bMagSf"],
                                               g["bDeltaCoeffs"], c.bType, c.bValue)
    b = c.rhs(2)
    diag, b = O.add_boundary(c.bFaceCells, ic, bc, diag, b)
    # the device assembly of the same matrix is the oracle's, bitwise
    bnd = fh.Boundary(mesh, c.bFaceCells)
    m = fh.lduMatrix(mesh)
    d_ic = torch.empty(c.nB, dtype=torch.float64, device="cuda")
    d_bc = torch.empty_like(d_ic)
    d_src = dev(c.rhs(2))
    ins = [dev(gamma), dev(g["weights"]), dev(g["magSf"]), dev(g["nonOrthDeltaCoeffs"]),
           dev(g["bMagSf"]), dev(g["bDeltaCoeffs"]), dev(c.bType, np.int32), dev(c.bValue)]
    torch.cuda.synchronize()
    fh._check(fh.lib().fh_assemble_laplacian(m.handle, bnd.handle, *[P(t) for t in ins], P(d_ic),
                                             P(d_bc), P(d_src), 1))
    d, u, _ = m.coeffs()
    np.testing.assert_array_equal(d, diag)
    np.testing.assert_array_equal(u, upper[:c.F])
    np.testing.assert_array_equal(host(d_src), b)
    om = O.Matrix(a, diag, upper[:c.F])
    O.reset_direction()
    opsi, operf = om.solve(CONFIG4, np.zeros(c.N), b)
    fh.lib().fh_gamg_reset_direction()
    psi = np.zeros(c.N)
    perf = fh.solver_New("p", m, CONFIG4).solve(psi, b)
    print(f"GAMG 40^3 graded: device {perf.nIterations} cycles, oracle {operf['nIterations']}, "
          f"rel L2 {rel_l2(psi, opsi):.2e}")
    assert perf.converged and operf["converged"]
    assert abs(perf.nIterations - operf["nIterations"]) <= ITER_TOL
    assert perf.normFactor == pytest.approx(operf["normFactor"], rel=1e-12)
    assert rel_l2(psi, opsi) <= SOLUTION_TOL
    pm = fh.solver_New("p", m, dict(CONFIG4, precision="mixed")).solve(psi2 := np.zeros(c.N), b)
    res = np.abs(b - om.Amul(psi2)).sum() / pm.normFactor
    print(f"mixed GAMG: {pm.nIterations} cycles, FP64 residual {res:.2e}, vs oracle "
          f"{rel_l2(psi2, opsi):.2e}")
    assert pm.converged and res < 1e-8
