"""foamhip: Python view of the MI355X lduMatrix solver library.

It mirrors the run-time-selectable solver interface of OpenFOAM's
lduMatrix (src/OpenFOAM/matrices/lduMatrix/lduMatrix/lduMatrix.H:93-509):

    mesh   = lduAddressing(nCells, lowerAddr, upperAddr)
    matrix = lduMatrix(mesh, diag, upper, lower=None)      # lower None: symmetric
    perf   = solver_New("p", matrix, {"solver": "PCG",
                                      "preconditioner": "DIC",
                                      "tolerance": 1e-8}).solve(psi, source)

Every call goes through the C-ABI of libfoamhip.so (include/foamhip.h); the
library must be built, there is no host fallback.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import FhControls, FhMeshInfo, FhPerf

__all__ = [
    "FoamHipError", "lduAddressing", "lduMatrix", "Boundary", "SolverPerformance",
    "solver_New", "controls_from_dict", "lib", "FhControls", "FhPerf",
]

FH_SOLVER = {"PCG": 0, "PBiCGStab": 1, "PBiCG": 2, "smoothSolver": 3, "GAMG": 4, "diagonal": 5}
FH_PRECON = {"none": 0, "diagonal": 1, "DIC": 2, "FDIC": 3, "DILU": 4, "GAMG": 5,
             "tileDIC": 6, "tileDILU": 7}
FH_SMOOTH = {"GaussSeidel": 0, "symGaussSeidel": 1, "DIC": 2, "DILU": 3,
             "DICGaussSeidel": 4, "DILUGaussSeidel": 5, "FDIC": 6, "none": 7}


class FoamHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_lib.STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code
        self.msg = msg


def lib():
    return _lib.fh()


def _check(status):
    if status != _lib.FH_OK:
        raise FoamHipError(status, lib().fh_last_error().decode(errors="replace"))


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class lduAddressing:
    """lduAddressing (lduAddressing.H:112-207) resident on the device."""

    def __init__(self, nCells, lowerAddr, upperAddr, interfaces=None):
        """interfaces: None, or the coupled (cyclic) interfaces as a dict of
        start [nInterfaces+1], faceCells and nbrCells (faces concatenated in
        interface order; fh_mesh_create_coupled)."""
        self.lower = _i32(lowerAddr)
        self.upper = _i32(upperAddr)
        self.nCells = int(nCells)
        self.nFaces = int(self.lower.size)
        h = ctypes.c_void_p()
        self.nIfFaces = 0
        if interfaces is None:
            _check(lib().fh_mesh_create(self.nCells, self.nFaces, _ptr(self.lower),
                                        _ptr(self.upper), ctypes.byref(h)))
        else:
            self._ifs = [_i32(interfaces[k]) for k in ("start", "faceCells", "nbrCells")]
            self.nIfFaces = int(self._ifs[0][-1])
            _check(lib().fh_mesh_create_coupled(self.nCells, self.nFaces, _ptr(self.lower),
                                                _ptr(self.upper), len(self._ifs[0]) - 1,
                                                *[_ptr(a) for a in self._ifs],
                                                ctypes.byref(h)))
        self.handle = h

    def set_face_areas(self, Sf, magSf):
        """fvMesh::Sf() (nFaces x 3) and magSf() of the internal faces: the
        weights of the faceAreaPair agglomeration."""
        self._Sf = _f64(Sf).reshape(-1)
        self._magSf = _f64(magSf)
        _check(lib().fh_mesh_set_face_areas(self.handle, _ptr(self._Sf), _ptr(self._magSf)))

    def info(self):
        inf = FhMeshInfo()
        _check(lib().fh_mesh_info_get(self.handle, ctypes.byref(inf)))
        return {k: getattr(inf, k) for k, _ in FhMeshInfo._fields_}

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                lib().fh_mesh_destroy(h)
            except Exception:
                pass
            self.handle = None


class Boundary:
    """Flattened non-coupled patch faces of an fvBoundaryMesh."""

    def __init__(self, mesh, bFaceCells):
        self.mesh = mesh
        self.faceCells = _i32(bFaceCells)
        h = ctypes.c_void_p()
        _check(lib().fh_boundary_create(mesh.handle, int(self.faceCells.size),
                                        _ptr(self.faceCells), ctypes.byref(h)))
        self.handle = h
        self.nB = int(self.faceCells.size)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                lib().fh_boundary_destroy(h)
            except Exception:
                pass
            self.handle = None


class lduMatrix:
    """lduMatrix coefficients (lduMatrix.H:79-87) resident on the device."""

    def __init__(self, mesh, diag=None, upper=None, lower=None):
        self.mesh = mesh
        h = ctypes.c_void_p()
        _check(lib().fh_matrix_create(mesh.handle, ctypes.byref(h)))
        self.handle = h
        self.symmetric = lower is None
        if diag is not None:
            self.set(diag, upper, lower)

    def set(self, diag, upper=None, lower=None):
        self._diag = _f64(diag)
        self._upper = _f64(upper) if upper is not None else None
        self._lower = _f64(lower) if lower is not None else None
        self.symmetric = lower is None
        _check(lib().fh_matrix_set(self.handle, _ptr(self._diag), _ptr(self._upper),
                                   _ptr(self._lower)))

    def set_interfaces(self, bouCoeffs, intCoeffs=None):
        """interfaceBouCoeffs (and interfaceIntCoeffs, default the same) of
        a coupled mesh, one value per interface face."""
        self._ifBou = _f64(bouCoeffs)
        self._ifInt = _f64(intCoeffs) if intCoeffs is not None else None
        _check(lib().fh_matrix_set_interfaces(self.handle, _ptr(self._ifBou),
                                              _ptr(self._ifInt)))

    def set_device(self, d_diag, d_upper=None, d_lower=None):
        """Device pointers (ints), e.g. torch tensor data_ptr()."""
        self.symmetric = d_lower is None
        _check(lib().fh_matrix_set_device(self.handle, ctypes.c_void_p(d_diag),
                                          ctypes.c_void_p(d_upper) if d_upper else None,
                                          ctypes.c_void_p(d_lower) if d_lower else None))

    # ---- lduMatrix members ----
    def _unary(self, fn, x, *extra):
        x = _f64(x)
        out = np.empty(self.mesh.nCells)
        _check(fn(self.handle, _ptr(x), *extra, _ptr(out)))
        return out

    def Amul(self, psi):
        return self._unary(lib().fh_amul, psi)

    def Tmul(self, psi):
        return self._unary(lib().fh_tmul, psi)

    def residual(self, psi, source):
        psi, source = _f64(psi), _f64(source)
        out = np.empty(self.mesh.nCells)
        _check(lib().fh_residual(self.handle, _ptr(psi), _ptr(source), _ptr(out)))
        return out

    def sumA(self):
        out = np.empty(self.mesh.nCells)
        _check(lib().fh_sum_a(self.handle, _ptr(out)))
        return out

    def coeffs(self):
        """(diag, upper, lower) read back from the device, natural order
        (lower is upper for a symmetric matrix)."""
        n, f = self.mesh.nCells, self.mesh.nFaces
        d, u, lo = np.empty(n), np.empty(f), np.empty(f)
        _check(lib().fh_matrix_get(self.handle, _ptr(d), _ptr(u), _ptr(lo)))
        return d, u, lo

    def H(self, psi):
        """lduMatrix::H (lduMatrixTemplates.C:33-64): minus the off-diagonal
        part of the matrix times psi; psi of shape (nCells,) or
        (nCells, nCmpt)."""
        psi = _f64(psi)
        nc = 1 if psi.ndim == 1 else psi.shape[1]
        out = np.empty_like(psi)
        _check(lib().fh_h(self.handle, nc, _ptr(psi), _ptr(out)))
        return out

    def H1(self):
        """lduMatrix::H1 (lduMatrixATmul.C:298-324)."""
        out = np.empty(self.mesh.nCells)
        _check(lib().fh_h1(self.handle, _ptr(out)))
        return out

    def faceH(self, psi):
        """lduMatrix::faceH (lduMatrixTemplates.C:77-111): one value (or
        nCmpt values) per face."""
        psi = _f64(psi)
        nc = 1 if psi.ndim == 1 else psi.shape[1]
        nF = self.mesh.nFaces
        out = np.empty((nF,) if psi.ndim == 1 else (nF, nc))
        _check(lib().fh_face_h(self.handle, nc, _ptr(psi), _ptr(out)))
        return out

    def sumMagOffDiag(self, sumOff):
        """lduMatrix::sumMagOffDiag (lduMatrixOperations.C:77-97): returns
        sumOff plus each row's sum of off-diagonal magnitudes."""
        out = _f64(sumOff).copy()
        _check(lib().fh_sum_mag_off_diag(self.handle, _ptr(out)))
        return out

    def normFactor(self, psi, source):
        psi, source = _f64(psi), _f64(source)
        v = ctypes.c_double()
        _check(lib().fh_norm_factor(self.handle, _ptr(psi), _ptr(source), ctypes.byref(v)))
        return v.value

    def reciprocalD(self, preconditioner):
        out = np.empty(self.mesh.nCells)
        _check(lib().fh_reciprocal_d(self.handle, FH_PRECON[preconditioner], _ptr(out)))
        return out

    def precondition(self, preconditioner, rA, transpose=False):
        rA = _f64(rA)
        out = np.empty(self.mesh.nCells)
        fn = lib().fh_precondition_t if transpose else lib().fh_precondition
        _check(fn(self.handle, FH_PRECON[preconditioner], _ptr(rA), _ptr(out)))
        return out

    def precondition_f32(self, preconditioner, rA):
        """The FP32 application used by the mixed-precision inner iterations."""
        r = np.ascontiguousarray(rA, dtype=np.float32)
        out = np.empty(self.mesh.nCells, np.float32)
        _check(lib().fh_precondition_f32(self.handle, FH_PRECON[preconditioner], _ptr(r),
                                         _ptr(out)))
        return out

    def agglomeration(self, controls):
        """The pair-agglomeration hierarchy a GAMG solve with these controls
        ("agglomerator", "nCellsInCoarsestLevel", "mergeLevels") uses, one
        dict per level with GAMGAgglomeration's maps (GAMGAgglomeration.H:
        325-367): restrictMap, faceRestrict, faceFlip and the coarse
        lower/upper addressing."""
        c = controls_from_dict(dict({"smoother": "GaussSeidel"}, **controls, solver="GAMG"))
        n = ctypes.c_int32()
        _check(lib().fh_gamg_agglomeration_levels(self.handle, ctypes.byref(c), ctypes.byref(n)))
        out = []
        for lev in range(n.value):
            sz = np.zeros(4, np.int32)
            _check(lib().fh_gamg_agglomeration_level(self.handle, ctypes.byref(c), lev,
                                                     _ptr(sz), None, None, None, None, None))
            nF, nFF, nC, nCF = (int(v) for v in sz)
            rm, fr = np.empty(nF, np.int32), np.empty(nFF, np.int32)
            ff = np.empty(nFF, np.uint8)
            lo, up = np.empty(nCF, np.int32), np.empty(nCF, np.int32)
            _check(lib().fh_gamg_agglomeration_level(self.handle, ctypes.byref(c), lev,
                                                     _ptr(sz), _ptr(rm), _ptr(fr), _ptr(ff),
                                                     _ptr(lo), _ptr(up)))
            out.append({"nFineCells": nF, "nFineFaces": nFF, "nCoarseCells": nC,
                        "nCoarseFaces": nCF, "restrictMap": rm, "faceRestrict": fr,
                        "faceFlip": ff, "lower": lo, "upper": up})
        return out

    def smooth(self, smoother, psi, source, nSweeps):
        psi = _f64(psi).copy()
        source = _f64(source)
        _check(lib().fh_smooth(self.handle, FH_SMOOTH[smoother], _ptr(psi), _ptr(source),
                               int(nSweeps)))
        return psi

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                lib().fh_matrix_destroy(h)
            except Exception:
                pass
            self.handle = None


class SolverPerformance:
    """SolverPerformance<scalar> (SolverPerformance.H:78-268)."""

    def __init__(self, perf, fieldName):
        self.solverName = perf.solverName.decode()
        self.fieldName = fieldName
        self.initialResidual = perf.initialResidual
        self.finalResidual = perf.finalResidual
        self.nIterations = perf.nIterations
        self.converged = bool(perf.converged)
        self.singular = bool(perf.singular)
        self.nRefinements = perf.nRefinements
        self.normFactor = perf.normFactor
        self.solveMilliseconds = perf.solveMilliseconds

    def line(self):
        """The log line of SolverPerformance::print (SolverPerformance.C:95-126)."""
        if self.singular:
            return f"{self.solverName}:  Solving for {self.fieldName}:  solution singularity"
        return (f"{self.solverName}:  Solving for {self.fieldName}, Initial residual = "
                f"{self.initialResidual:.6g}, Final residual = {self.finalResidual:.6g}, "
                f"No Iterations {self.nIterations}")

    def __repr__(self):
        return self.line()


def _lookup(fn, name):
    v = ctypes.c_int32()
    _check(fn(name.encode(), ctypes.byref(v)))
    return v.value


def controls_from_dict(d):
    """Map an fvSolution solver dictionary to fh_controls.

    Keys and defaults follow lduMatrix::solver::readControls
    (lduMatrixSolver.C:158-164), smoothSolver::readControls and
    GAMGSolver::readControls; the preconditioner may be a word or a
    sub-dictionary with a "preconditioner" entry
    (lduMatrixPreconditioner.C:40-60).  A GAMG preconditioner is a GAMG
    solver built from that sub-dictionary alone (lduMatrixPreconditioner.C:
    72-82, GAMGPreconditioner.C:44-63): its smoother, sweeps, agglomeration,
    nVcycles and coarsest-level tolerances come from there, and the word form
    leaves it without the required smoother entry, as in the reference.
    "precision mixed" selects the FP32 inner / FP64 outer path."""
    if "solver" not in d:
        raise FoamHipError(1, "keyword solver is undefined in the solver dictionary")
    c = _fill_controls(d)
    c.solver = _lookup(lib().fh_solver_lookup, d["solver"])
    pre = d.get("preconditioner", "none")
    sub = pre if isinstance(pre, dict) else None
    name = sub.get("preconditioner", "none") if sub is not None else pre
    c.preconditioner = _lookup(lib().fh_preconditioner_lookup, name)
    if c.preconditioner == FH_PRECON["GAMG"]:
        if sub is None or "smoother" not in sub:
            raise FoamHipError(1, "GAMG preconditioner: keyword smoother is undefined in "
                               "the preconditioner sub-dictionary (GAMGPreconditioner reads "
                               "its controls from it)")
        gc = _fill_controls(sub)
        gc.solver = FH_SOLVER["GAMG"]
        c._precondControls = gc          # kept alive with c
        c.preconditionerControls = ctypes.addressof(gc)
    return c


def _fill_controls(d):
    c = FhControls()
    lib().fh_controls_default(ctypes.byref(c))
    if "smoother" in d:
        c.smoother = _lookup(lib().fh_smoother_lookup, d["smoother"])
    if "innerPreconditioner" in d:
        c.innerPreconditioner = _lookup(lib().fh_preconditioner_lookup,
                                        d["innerPreconditioner"])
    ints = ["maxIter", "minIter", "nSweeps", "nPreSweeps", "preSweepsLevelMultiplier",
            "maxPreSweeps", "nPostSweeps", "postSweepsLevelMultiplier", "maxPostSweeps",
            "nFinestSweeps", "nCellsInCoarsestLevel", "mergeLevels", "nVcycles",
            "maxRefinements"]
    for k in ints:
        if k in d:
            setattr(c, k, int(d[k]))
    if "minCellsPerProcessor" in d:
        c.nCellsInCoarsestLevel = int(d["minCellsPerProcessor"])
    for k in ["tolerance", "relTol", "innerRelTol"]:
        if k in d:
            setattr(c, k, float(d[k]))
    for k in ["scaleCorrection", "directSolveCoarsest"]:
        if k in d:
            v = d[k]
            if isinstance(v, str):
                v = v.lower() in ("yes", "on", "true", "1")
            setattr(c, k, int(bool(v)))
    if "agglomerator" in d:
        if d["agglomerator"] not in ("algebraicPair", "faceAreaPair"):
            raise FoamHipError(1, f"Unknown GAMGAgglomeration type {d['agglomerator']}")
        c.agglomerator = 1 if d["agglomerator"] == "faceAreaPair" else 0
    prec = d.get("precision", "double")
    if prec not in ("double", "mixed"):
        raise FoamHipError(1, f"Unknown precision {prec}; valid: double mixed")
    c.precision = 1 if prec == "mixed" else 0
    return c


class _Solver:
    def __init__(self, fieldName, matrix, controls):
        self.fieldName = fieldName
        self.matrix = matrix
        self.controls = controls

    def solve(self, psi, source):
        """lduMatrix::solver::solve: psi (numpy, in place) holds the initial
        guess and receives the solution."""
        if psi.dtype != np.float64 or not psi.flags.c_contiguous:
            raise FoamHipError(1, "psi must be a contiguous float64 array (solved in place)")
        source = _f64(source)
        perf = FhPerf()
        status = lib().fh_solve(self.matrix.handle, ctypes.byref(self.controls), _ptr(psi),
                                _ptr(source), ctypes.byref(perf))
        _check(status)
        return SolverPerformance(perf, self.fieldName)

    def solve_device(self, d_psi, d_source):
        """Device-resident psi/source (pointers as ints)."""
        perf = FhPerf()
        _check(lib().fh_solve_device(self.matrix.handle, ctypes.byref(self.controls),
                                     ctypes.c_void_p(d_psi), ctypes.c_void_p(d_source),
                                     ctypes.byref(perf)))
        return SolverPerformance(perf, self.fieldName)


def solver_New(fieldName, matrix, solverControls):
    """lduMatrix::solver::New (lduMatrixSolver.C:40-130)."""
    c = solverControls if isinstance(solverControls, FhControls) else \
        controls_from_dict(solverControls)
    return _Solver(fieldName, matrix, c)
