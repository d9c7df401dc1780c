"""CPU checks of the drop-in boundary: the libraries load, libfoamhip.so
exports every function include/foamhip.h declares, and the host-side
selection logic behaves like the reference run-time tables.  No call here
touches a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

import foamhip
from foamhip import _lib
from foamhip.cases import BoxCase

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "foamhip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(fh_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_header_declares_what_the_binding_types():
    decl = declared_functions()
    assert len(decl) >= 30
    assert set(decl) == set(_lib.FH_SYMBOLS), set(decl) ^ set(_lib.FH_SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(_lib.LIB_DIR, "libfoamhip.so"))
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert foamhip.lib().fh_version() >= 10000


def test_name_lookup_matches_reference_tables():
    L = foamhip.lib()
    v = ctypes.c_int32()
    for name, want in [("PCG", 0), ("hipPCG", 0), ("PBiCGStab", 1), ("hipPBiCGStab", 1),
                       ("GAMG", 4), ("smoothSolver", 3)]:
        assert L.fh_solver_lookup(name.encode(), ctypes.byref(v)) == 0 and v.value == want
    for name, want in [("DIC", 2), ("hipDIC", 2), ("DILU", 4), ("FDIC", 3), ("diagonal", 1),
                       ("none", 0), ("GAMG", 5)]:
        assert L.fh_preconditioner_lookup(name.encode(), ctypes.byref(v)) == 0 and v.value == want
    assert L.fh_smoother_lookup(b"nonBlockingGaussSeidel", ctypes.byref(v)) == 0 and v.value == 0
    # unknown names are errors listing the valid names (lduMatrixSolver.C:72-78)
    assert L.fh_solver_lookup(b"ICCG", ctypes.byref(v)) == 1
    msg = L.fh_last_error().decode()
    assert "Unknown solver ICCG" in msg and "PCG" in msg and "GAMG" in msg


def test_controls_from_dict_defaults_and_errors():
    c = foamhip.controls_from_dict({"solver": "hipPCG", "preconditioner": "hipDIC"})
    assert (c.maxIter, c.minIter, c.tolerance, c.relTol) == (1000, 0, 1e-6, 0.0)
    assert c.nPostSweeps == 2 and c.nFinestSweeps == 2 and c.maxPostSweeps == 4
    c = foamhip.controls_from_dict({"solver": "PCG", "preconditioner": {"preconditioner": "DIC"},
                                    "tolerance": 1e-8, "relTol": 0.01, "precision": "mixed"})
    assert c.preconditioner == 2 and c.tolerance == 1e-8 and c.relTol == 0.01 and c.precision == 1
    c = foamhip.controls_from_dict({"solver": "GAMG", "smoother": "symGaussSeidel",
                                    "directSolveCoarsest": "yes", "mergeLevels": 2,
                                    "nCellsInCoarsestLevel": 1000})
    assert c.directSolveCoarsest == 1 and c.mergeLevels == 2 and c.nCellsInCoarsestLevel == 1000
    with pytest.raises(foamhip.FoamHipError):
        foamhip.controls_from_dict({"solver": "Nope"})
    with pytest.raises(foamhip.FoamHipError):
        foamhip.controls_from_dict({"preconditioner": "DIC"})
    with pytest.raises(foamhip.FoamHipError):
        foamhip.controls_from_dict({"solver": "PCG", "precision": "half"})


def test_box_case_generator():
    c = BoxCase(4, 3, 5)
    assert c.N == 60 and c.F == 3 * 3 * 5 + 4 * 2 * 5 + 4 * 3 * 4
    assert np.all(c.lower < c.upper) and np.all(np.diff(c.lower) >= 0)
    # per owner the neighbours ascend: +1, +nx, +nx*ny
    assert c.upper[0] == 1 and c.upper[1] == 4 and c.upper[2] == 12
    assert c.nB == 2 * (3 * 5 + 4 * 5 + 4 * 3)
    assert np.all(c.bType[c.bPatch == 0] == 1) and np.all(c.bType[c.bPatch != 0] == 0)
    np.testing.assert_allclose(c.V.sum(), 1.0)
    # closed surface: sum of outward boundary areas is zero per direction
    np.testing.assert_allclose(c.bSf.reshape(-1, 3).sum(0), 0, atol=1e-14)
    b = c.rhs(1)
    # U[-1,1] minus its sample mean
    assert abs(b.mean()) < 1e-15 and np.all(np.abs(b) <= 1.0 + 0.2)
    np.testing.assert_array_equal(b, c.rhs(1))            # deterministic
    g = c.lognormal(7, 1.0, 2.0)
    lg = np.log(g)
    assert abs(lg.mean()) < 1e-10 and abs(lg.std() - 1.0) < 1e-10
    phi, bPhi = c.fourier_flux(3)
    assert phi.shape == (c.F,) and bPhi.shape == (c.nB,)


def test_fourier_velocity_is_nearly_divergence_free():
    c = BoxCase(24)
    phi, bPhi = c.fourier_flux(3)
    div = np.zeros(c.N)
    np.add.at(div, c.lower, phi)
    np.subtract.at(div, c.upper, phi)
    np.add.at(div, c.bFaceCells, bPhi)
    # second-order face-centre quadrature of an exactly solenoidal field
    scale = np.abs(phi).mean()
    assert np.abs(div).mean() < 0.05 * scale


def test_server_lifecycle_without_solves():
    """A server starts its workers, keeps a clock and shuts down with no
    device work; bad requests are refused at submission."""
    from foamhip.serve import Server
    srv = Server(maxConcurrent=3)
    t0 = srv.now()
    assert t0 >= 0.0 and srv.pending() == 0
    assert srv.wait(0.01) is None
    import ctypes
    r = _lib.FhRequest()
    assert foamhip.lib().fh_server_submit(srv.handle, ctypes.byref(r)) == 1
    assert srv.now() >= t0
    srv.close()
    cfg = _lib.FhServerConfig()
    foamhip.lib().fh_server_config_default(ctypes.byref(cfg))
    cfg.maxConcurrent = 0
    h = ctypes.c_void_p()
    assert foamhip.lib().fh_server_create(ctypes.byref(cfg), ctypes.byref(h)) == 1
