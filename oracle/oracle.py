"""ctypes binding of the oracle (liboracle.so built from oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, never by the product.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_i32, _dbl, _vp = ctypes.c_int32, ctypes.c_double, ctypes.c_void_p

SOLVER = {"PCG": 0, "PBiCGStab": 1, "PBiCG": 2, "smoothSolver": 3, "GAMG": 4, "diagonal": 5}
PRECON = {"none": 0, "diagonal": 1, "DIC": 2, "FDIC": 3, "DILU": 4, "GAMG": 5}
SMOOTH = {"GaussSeidel": 0, "symGaussSeidel": 1, "DIC": 2, "DILU": 3, "DICGaussSeidel": 4,
          "DILUGaussSeidel": 5, "FDIC": 6, "none": 7}


class OrAddr(ctypes.Structure):
    _fields_ = [("nCells", _i32), ("nFaces", _i32), ("lower", _vp), ("upper", _vp),
                ("losort", ctypes.POINTER(_i32)), ("ownerStart", ctypes.POINTER(_i32)),
                ("losortStart", ctypes.POINTER(_i32))]


class OrInterfaces(ctypes.Structure):
    _fields_ = [("nInterfaces", _i32), ("start", _vp), ("faceCells", _vp), ("nbrCells", _vp),
                ("bouCoeffs", _vp), ("intCoeffs", _vp)]


class OrMatrix(ctypes.Structure):
    _fields_ = [("addr", ctypes.POINTER(OrAddr)), ("diag", _vp), ("upper", _vp), ("lower", _vp),
                ("interfaces", ctypes.POINTER(OrInterfaces))]


class OrControls(ctypes.Structure):
    _fields_ = [("solver", _i32), ("preconditioner", _i32), ("smoother", _i32),
                ("maxIter", _i32), ("minIter", _i32), ("tolerance", _dbl), ("relTol", _dbl),
                ("nSweeps", _i32), ("nPreSweeps", _i32), ("preSweepsLevelMultiplier", _i32),
                ("maxPreSweeps", _i32), ("nPostSweeps", _i32),
                ("postSweepsLevelMultiplier", _i32), ("maxPostSweeps", _i32),
                ("nFinestSweeps", _i32), ("scaleCorrection", _i32),
                ("directSolveCoarsest", _i32), ("nCellsInCoarsestLevel", _i32),
                ("mergeLevels", _i32), ("nVcycles", _i32), ("faceWeights", _vp),
                ("precondControls", _vp)]


class OrPerf(ctypes.Structure):
    _fields_ = [("initialResidual", _dbl), ("finalResidual", _dbl), ("nIterations", _i32),
                ("converged", _i32), ("singular", _i32), ("normFactor", _dbl)]


class OrHierarchy(ctypes.Structure):
    _fields_ = [("nLevels", _i32), ("nCells", ctypes.POINTER(_i32)),
                ("nFaces", ctypes.POINTER(_i32)),
                ("lower", ctypes.POINTER(ctypes.POINTER(_i32))),
                ("upper", ctypes.POINTER(ctypes.POINTER(_i32))),
                ("restrictMap", ctypes.POINTER(ctypes.POINTER(_i32))),
                ("faceRestrict", ctypes.POINTER(ctypes.POINTER(_i32))),
                ("faceFlip", ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8)))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C {HERE}`")
        _lib = ctypes.CDLL(path)
        L = _lib
        L.or_addr_init.argtypes = [ctypes.POINTER(OrAddr), _i32, _i32, _vp, _vp]
        L.or_addr_free.argtypes = [ctypes.POINTER(OrAddr)]
        for name in ["or_sum", "or_sumMag", "or_sumSqr", "or_average"]:
            getattr(L, name).argtypes = [_i32, _vp]
            getattr(L, name).restype = _dbl
        L.or_sumProd.argtypes = [_i32, _vp, _vp]
        L.or_sumProd.restype = _dbl
        L.or_face_geometry.argtypes = [_i32, _vp, _vp, _vp, _vp, _vp, _vp]
        L.or_cell_geometry.argtypes = [_i32, _i32, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp,
                                       _vp, _vp]
        L.or_surface_interpolation.argtypes = [_i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                               _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp]
        L.or_set_sum_order.argtypes = [ctypes.c_int]
        M = ctypes.POINTER(OrMatrix)
        L.or_amul.argtypes = [M, _vp, _vp]
        L.or_tmul.argtypes = [M, _vp, _vp]
        L.or_sumA.argtypes = [M, _vp]
        L.or_residual.argtypes = [M, _vp, _vp, _vp]
        L.or_normFactor.argtypes = [M, _vp, _vp, _vp, _vp]
        L.or_normFactor.restype = _dbl
        L.or_negSumDiag.argtypes = [ctypes.POINTER(OrAddr), _vp, _vp, _vp]
        L.or_H.argtypes = [M, _i32, _vp, _vp]
        L.or_H1.argtypes = [M, _vp]
        L.or_faceH.argtypes = [M, _i32, _vp, _vp]
        L.or_faceH.restype = ctypes.c_int
        L.or_sumMagOffDiag.argtypes = [M, _vp]
        L.or_relax.argtypes = [M, _i32, _vp, _vp, _i32, _dbl, _vp, _vp, _vp]
        L.or_face_area_weights.argtypes = [_i32, _vp, _vp, _vp]
        L.or_flux.argtypes = [M, _i32, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        L.or_flux.restype = ctypes.c_int
        L.or_calcReciprocalD_DIC.argtypes = [M, _vp]
        L.or_calcReciprocalD_DILU.argtypes = [M, _vp]
        L.or_precondition.argtypes = [M, ctypes.c_int, _vp, _vp, _vp]
        L.or_preconditionT.argtypes = [M, ctypes.c_int, _vp, _vp, _vp]
        L.or_smooth.argtypes = [M, ctypes.c_int, _vp, _vp, _vp, _i32]
        L.or_controls_default.argtypes = [ctypes.POINTER(OrControls)]
        L.or_solve.argtypes = [M, ctypes.POINTER(OrControls), _vp, _vp,
                               ctypes.POINTER(OrPerf), _vp]
        L.or_solve.restype = ctypes.c_int
        L.or_gamg_reset_direction.argtypes = []
        L.or_gamg_get_direction.restype = ctypes.c_int
        L.or_pair_agglomerate.argtypes = [ctypes.POINTER(OrAddr), _vp, _vp]
        L.or_pair_agglomerate.restype = _i32
        L.or_gamg_agglomerate.argtypes = [M, ctypes.POINTER(OrControls),
                                          ctypes.POINTER(OrHierarchy)]
        L.or_hierarchy_free.argtypes = [ctypes.POINTER(OrHierarchy)]
        L.or_agglomerate_matrix.argtypes = [_i32, _i32, _i32, _vp, _i32, _vp, _vp, _vp, _vp,
                                            _vp, _vp, _vp, _vp]
        L.or_LUDecompose.argtypes = [_i32, _vp, _vp]
        L.or_LUDecompose.restype = ctypes.c_int
        L.or_LUBacksubstitute.argtypes = [_i32, _vp, _vp, _vp]
        L.or_laplacian_assemble.argtypes = [ctypes.POINTER(OrAddr)] + [_vp] * 4 + \
            [_i32] + [_vp] * 9
        L.or_upwind_assemble.argtypes = [ctypes.POINTER(OrAddr), _vp, _i32] + [_vp] * 8
        L.or_add_boundary.argtypes = [_i32, _vp, _vp, _vp, _vp, _vp]
        L.or_grad_gauss_linear.argtypes = [ctypes.POINTER(OrAddr)] + [_vp] * 4 + \
            [_i32] + [_vp] * 7
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _controls(d, keep=None):
    """An fvSolution solver dictionary as OrControls (lduMatrixSolver.C:158-164
    defaults).  A preconditioner sub-dictionary (lduMatrixPreconditioner.C:
    40-60) becomes the controls of the GAMG preconditioner
    (GAMGPreconditioner.C:44-63).  Returns (controls, arrays to keep alive)."""
    keep = [] if keep is None else keep
    c = OrControls()
    lib().or_controls_default(ctypes.byref(c))
    c.solver = SOLVER[d["solver"]] if "solver" in d else SOLVER["GAMG"]
    pre = d.get("preconditioner", "none")
    if isinstance(pre, dict):
        sub, _ = _controls(dict(pre, solver="GAMG", preconditioner="none"), keep)
        keep.append(sub)
        c.precondControls = ctypes.cast(ctypes.pointer(sub), _vp)
        pre = pre["preconditioner"]
    c.preconditioner = PRECON[pre]
    if "smoother" in d:
        c.smoother = SMOOTH[d["smoother"]]
    for k in ["maxIter", "minIter", "nSweeps", "nPreSweeps", "nPostSweeps", "maxPostSweeps",
              "nFinestSweeps", "nCellsInCoarsestLevel", "mergeLevels", "nVcycles",
              "preSweepsLevelMultiplier", "maxPreSweeps", "postSweepsLevelMultiplier"]:
        if k in d:
            setattr(c, k, int(d[k]))
    for k in ["tolerance", "relTol"]:
        if k in d:
            setattr(c, k, float(d[k]))
    for k in ["scaleCorrection", "directSolveCoarsest"]:
        if k in d:
            setattr(c, k, int(bool(d[k])))
    fw = _face_weights(d)
    keep.append(fw)
    c.faceWeights = _p(fw)
    return c, keep


class Addr:
    def __init__(self, nCells, lower, upper):
        self.lower = np.ascontiguousarray(lower, np.int32)
        self.upper = np.ascontiguousarray(upper, np.int32)
        self.nCells = int(nCells)
        self.nFaces = int(self.lower.size)
        self.s = OrAddr()
        lib().or_addr_init(ctypes.byref(self.s), self.nCells, self.nFaces, _p(self.lower),
                           _p(self.upper))

    def losort(self):
        return np.ctypeslib.as_array(self.s.losort, (max(self.nFaces, 1),))[:self.nFaces].copy()

    def ownerStart(self):
        return np.ctypeslib.as_array(self.s.ownerStart, (self.nCells + 1,)).copy()

    def losortStart(self):
        return np.ctypeslib.as_array(self.s.losortStart, (self.nCells + 1,)).copy()

    def __del__(self):
        try:
            lib().or_addr_free(ctypes.byref(self.s))
        except Exception:
            pass


class Matrix:
    """Oracle lduMatrix; lower None -> symmetric (lower aliases upper),
    upper None -> diagonal."""

    def __init__(self, addr, diag, upper=None, lower=None, interfaces=None):
        """interfaces: None, or the coupled interfaces as a dict of arrays
        start [nInterfaces+1], faceCells, nbrCells, bouCoeffs and (for Tmul)
        intCoeffs (defaults to bouCoeffs), faces concatenated in interface
        order."""
        self.addr = addr
        self.diag = _f(diag)
        self.upper = _f(upper) if upper is not None else None
        self.lower = _f(lower) if lower is not None else self.upper
        self.symmetric = lower is None
        self.s = OrMatrix(ctypes.pointer(addr.s), _p(self.diag), _p(self.upper), _p(self.lower))
        self._if = None
        if interfaces is not None:
            st = np.ascontiguousarray(interfaces["start"], np.int32)
            fc = np.ascontiguousarray(interfaces["faceCells"], np.int32)
            nb = np.ascontiguousarray(interfaces["nbrCells"], np.int32)
            bou = _f(interfaces["bouCoeffs"])
            itc = _f(interfaces.get("intCoeffs", bou))
            self._if = (st, fc, nb, bou, itc)
            self._ifs = OrInterfaces(len(st) - 1, _p(st), _p(fc), _p(nb), _p(bou), _p(itc))
            self.s.interfaces = ctypes.pointer(self._ifs)
        self.ref = ctypes.byref(self.s)

    def _out(self):
        return np.empty(self.addr.nCells)

    def Amul(self, psi):
        psi = _f(psi); o = self._out(); lib().or_amul(self.ref, _p(psi), _p(o)); return o

    def Tmul(self, psi):
        psi = _f(psi); o = self._out(); lib().or_tmul(self.ref, _p(psi), _p(o)); return o

    def sumA(self):
        o = self._out(); lib().or_sumA(self.ref, _p(o)); return o

    def residual(self, psi, source):
        psi, source = _f(psi), _f(source)
        o = self._out(); lib().or_residual(self.ref, _p(psi), _p(source), _p(o)); return o

    def H(self, psi):
        """lduMatrix::H; psi of shape (nCells,) or (nCells, nCmpt)."""
        psi = _f(psi)
        nc = 1 if psi.ndim == 1 else psi.shape[1]
        o = np.empty_like(psi)
        lib().or_H(self.ref, nc, _p(psi), _p(o))
        return o

    def H1(self):
        o = self._out(); lib().or_H1(self.ref, _p(o)); return o

    def faceH(self, psi):
        """lduMatrix::faceH: one value (or nCmpt) per face."""
        psi = _f(psi)
        nc = 1 if psi.ndim == 1 else psi.shape[1]
        shape = (self.addr.nFaces,) if psi.ndim == 1 else (self.addr.nFaces, nc)
        o = np.empty(shape)
        if lib().or_faceH(self.ref, nc, _p(psi), _p(o)) != 0:
            raise RuntimeError("Cannot calculate faceH: the matrix does not have any "
                               "off-diagonal coefficients.")
        return o

    def sumMagOffDiag(self, sumOff):
        o = _f(sumOff).copy(); lib().or_sumMagOffDiag(self.ref, _p(o)); return o

    def flux(self, bFaceCells, ic, bc, psi, corr=None, bCorr=None):
        """fvMatrix::flux(), non-coupled patches: (internal, boundary) face
        fluxes; arrays (n,) or (n, nCmpt)."""
        psi, ic, bc = _f(psi), _f(ic), _f(bc)
        nc = 1 if psi.ndim == 1 else psi.shape[1]
        nF, nB = self.addr.nFaces, len(bFaceCells)
        fl = np.empty((nF,) if nc == 1 else (nF, nc))
        bf = np.empty((nB,) if nc == 1 else (nB, nc))
        bfc = np.ascontiguousarray(bFaceCells, np.int32)
        corr = None if corr is None else _f(corr)
        bCorr = None if bCorr is None else _f(bCorr)
        if lib().or_flux(self.ref, nc, _p(psi), nB, _p(bfc), _p(ic), _p(bc), _p(corr),
                         _p(bCorr), _p(fl), _p(bf)) != 0:
            raise RuntimeError("Cannot calculate faceH")
        return fl, bf

    def relax(self, bFaceCells, ic, alpha, psi, source):
        """fvMatrix::relax(alpha), non-coupled patches: returns the relaxed
        (diag, source); ic of shape (nB,) or (nB, nCmpt), psi and source
        (nCells,) or (nCells, nCmpt)."""
        ic, psi = _f(ic), _f(psi)
        nc = 1 if psi.ndim == 1 else psi.shape[1]
        d = self.diag.copy(); s = _f(source).copy()
        bfc = np.ascontiguousarray(bFaceCells, np.int32)
        lib().or_relax(self.ref, len(bfc), _p(bfc), _p(ic), nc, alpha, _p(psi), _p(d), _p(s))
        return d, s

    def normFactor(self, psi, source):
        psi, source = _f(psi), _f(source)
        Apsi = self.Amul(psi)
        tmp = self._out()
        return lib().or_normFactor(self.ref, _p(psi), _p(source), _p(Apsi), _p(tmp))

    def reciprocalD(self, precon):
        o = self._out()
        if precon in ("DIC", "FDIC"):
            lib().or_calcReciprocalD_DIC(self.ref, _p(o))
        elif precon == "DILU":
            lib().or_calcReciprocalD_DILU(self.ref, _p(o))
        elif precon == "diagonal":
            o[:] = 1.0 / self.diag
        else:
            raise ValueError(precon)
        return o

    def precondition(self, precon, r, transpose=False):
        r = _f(r)
        rD = self.reciprocalD(precon) if precon != "none" else self._out()
        o = self._out()
        fn = lib().or_preconditionT if transpose else lib().or_precondition
        fn(self.ref, PRECON[precon], _p(rD), _p(r), _p(o))
        return o

    def smooth(self, smoother, psi, source, nSweeps):
        psi = _f(psi).copy(); source = _f(source)
        rD = self._out()
        if smoother in ("DIC", "FDIC", "DICGaussSeidel"):
            lib().or_calcReciprocalD_DIC(self.ref, _p(rD))
        elif smoother in ("DILU", "DILUGaussSeidel"):
            lib().or_calcReciprocalD_DILU(self.ref, _p(rD))
        lib().or_smooth(self.ref, SMOOTH[smoother], _p(rD), _p(psi), _p(source), int(nSweeps))
        return psi

    def solve(self, controls, psi, source, history=False):
        """controls: dict of fvSolution entries.  Returns (psi, perf dict[, hist])."""
        c, keep = _controls(controls)
        psi = _f(psi).copy(); source = _f(source)
        perf = OrPerf()
        hist = np.zeros(c.maxIter + 2) if history else None
        rc = lib().or_solve(self.ref, ctypes.byref(c), _p(psi), _p(source), ctypes.byref(perf),
                            _p(hist))
        if rc != 0:
            raise RuntimeError(f"oracle solve failed ({rc})")
        pd = {"initialResidual": perf.initialResidual, "finalResidual": perf.finalResidual,
              "nIterations": perf.nIterations, "converged": bool(perf.converged),
              "singular": bool(perf.singular), "normFactor": perf.normFactor}
        if history:
            return psi, pd, hist
        return psi, pd


def reset_direction():
    lib().or_gamg_reset_direction()


def get_direction():
    return lib().or_gamg_get_direction()


def pair_agglomerate(addr, weights):
    w = _f(weights)
    m = np.empty(addr.nCells, np.int32)
    n = lib().or_pair_agglomerate(ctypes.byref(addr.s), _p(w), _p(m))
    return int(n), m


def face_area_weights(Sf, magSf):
    """faceAreaPair weights of the internal faces; Sf of shape (nFaces, 3)
    or flat."""
    Sf, magSf = _f(Sf), _f(magSf)
    w = np.empty(magSf.size)
    lib().or_face_area_weights(int(magSf.size), _p(Sf), _p(magSf), _p(w))
    return w


def _face_weights(controls):
    """The faceAreaPair weights named by the controls ("agglomerator":
    "faceAreaPair" with "Sf" and "magSf"), or None (algebraicPair)."""
    if controls.get("agglomerator", "algebraicPair") != "faceAreaPair":
        return None
    return face_area_weights(controls["Sf"], controls["magSf"])


def agglomerate(matrix, controls):
    """Full pair-agglomeration hierarchy; returns a list of level dicts."""
    c = OrControls()
    lib().or_controls_default(ctypes.byref(c))
    fw = _face_weights(controls)
    c.faceWeights = _p(fw)
    for k, v in controls.items():
        if k in ("agglomerator", "Sf", "magSf"):
            continue
        setattr(c, k, int(v))
    h = OrHierarchy()
    lib().or_gamg_agglomerate(matrix.ref, ctypes.byref(c), ctypes.byref(h))
    levels = []
    for l in range(h.nLevels):
        nF, nFF = h.nCells[l], h.nFaces[l]
        nc, nf = h.nCells[l + 1], h.nFaces[l + 1]
        levels.append({
            "nFineCells": nF, "nFineFaces": nFF, "nCoarseCells": nc, "nCoarseFaces": nf,
            "restrictMap": np.ctypeslib.as_array(h.restrictMap[l], (nF,)).copy(),
            "faceRestrict": np.ctypeslib.as_array(h.faceRestrict[l], (max(nFF, 1),))[:nFF].copy(),
            "faceFlip": np.ctypeslib.as_array(h.faceFlip[l], (max(nFF, 1),))[:nFF].copy(),
            "lower": np.ctypeslib.as_array(h.lower[l + 1], (max(nf, 1),))[:nf].copy(),
            "upper": np.ctypeslib.as_array(h.upper[l + 1], (max(nf, 1),))[:nf].copy(),
        })
    lib().or_hierarchy_free(ctypes.byref(h))
    return levels


def agglomerate_matrix(level, fineDiag, fineUpper, fineLower=None):
    nc, nf = level["nCoarseCells"], level["nCoarseFaces"]
    cd = np.empty(nc); cu = np.empty(max(nf, 1)); cl = np.empty(max(nf, 1)) if fineLower is not None else None
    rm = np.ascontiguousarray(level["restrictMap"], np.int32)
    fr = np.ascontiguousarray(level["faceRestrict"], np.int32)
    ff = np.ascontiguousarray(level["faceFlip"], np.uint8)
    lib().or_agglomerate_matrix(level["nFineFaces"], nc, nf, _p(rm), level["nFineCells"],
                                _p(fr), _p(ff), _p(_f(fineDiag)), _p(_f(fineUpper)),
                                _p(_f(fineLower)) if fineLower is not None else None,
                                _p(cd), _p(cu), _p(cl) if cl is not None else None)
    return cd, cu[:nf], (cl[:nf] if cl is not None else None)


def lu_decompose(A):
    A = np.ascontiguousarray(A, np.float64).copy()
    n = A.shape[0]
    piv = np.empty(n, np.int32)
    rc = lib().or_LUDecompose(n, _p(A), _p(piv))
    if rc != 0:
        raise RuntimeError("singular matrix")
    return A, piv


def lu_backsubstitute(LU, piv, b):
    x = _f(b).copy()
    lib().or_LUBacksubstitute(LU.shape[0], _p(LU), _p(np.ascontiguousarray(piv, np.int32)), _p(x))
    return x


def laplacian_assemble(addr, gamma, weights, magSf, deltaCoeffs, bFaceCells, bMagSf,
                       bDeltaCoeffs, bType, bValue):
    F, N, nB = addr.nFaces, addr.nCells, len(bFaceCells)
    upper = np.empty(max(F, 1)); diag = np.empty(N); ic = np.empty(max(nB, 1)); bc = np.empty(max(nB, 1))
    lib().or_laplacian_assemble(ctypes.byref(addr.s), _p(_f(gamma)), _p(_f(weights)),
                                _p(_f(magSf)), _p(_f(deltaCoeffs)), nB,
                                _p(np.ascontiguousarray(bFaceCells, np.int32)), _p(_f(bMagSf)),
                                _p(_f(bDeltaCoeffs)), _p(np.ascontiguousarray(bType, np.int32)),
                                _p(_f(bValue)), _p(upper), _p(diag), _p(ic), _p(bc))
    return upper[:F], diag, ic[:nB], bc[:nB]


def upwind_assemble(addr, phi, bType, bPhi, bValue):
    F, N, nB = addr.nFaces, addr.nCells, len(bType)
    lower = np.empty(max(F, 1)); upper = np.empty(max(F, 1)); diag = np.empty(N)
    ic = np.empty(max(nB, 1)); bc = np.empty(max(nB, 1))
    lib().or_upwind_assemble(ctypes.byref(addr.s), _p(_f(phi)), nB,
                             _p(np.ascontiguousarray(bType, np.int32)), _p(_f(bPhi)),
                             _p(_f(bValue)), _p(lower), _p(upper), _p(diag), _p(ic), _p(bc))
    return lower[:F], upper[:F], diag, ic[:nB], bc[:nB]


def add_boundary(bFaceCells, ic, bc, diag, source):
    diag = _f(diag).copy(); source = _f(source).copy()
    lib().or_add_boundary(len(bFaceCells), _p(np.ascontiguousarray(bFaceCells, np.int32)),
                          _p(_f(ic)), _p(_f(bc)), _p(diag), _p(source))
    return diag, source


def grad_gauss_linear(addr, phi, weights, Sf, V, bFaceCells, bSf, bMagSf, bPhi, bSnGrad):
    N, nB = addr.nCells, len(bFaceCells)
    grad = np.empty(3 * N); bGrad = np.empty(3 * max(nB, 1))
    lib().or_grad_gauss_linear(ctypes.byref(addr.s), _p(_f(phi)), _p(_f(weights)), _p(_f(Sf)),
                               _p(_f(V)), nB, _p(np.ascontiguousarray(bFaceCells, np.int32)),
                               _p(_f(bSf)), _p(_f(bMagSf)), _p(_f(bPhi)), _p(_f(bSnGrad)),
                               _p(grad), _p(bGrad))
    return grad, bGrad[:3 * nB]


def convection_diffusion(addr, case, phi, bPhi, D):
    """fvm::div(phi,T) upwind - fvm::laplacian(D,T): lduMatrix::operator-=
    on the two assembled matrices (diag, upper, lower, internal and
    boundary coefficients)."""
    lo_c, up_c, dg_c, ic_c, bc_c = upwind_assemble(addr, phi, case.bType, bPhi, case.bValue)
    up_l, dg_l, ic_l, bc_l = laplacian_assemble(
        addr, np.full(addr.nCells, D), case.weights, case.magSf, case.deltaCoeffs,
        case.bFaceCells, case.bMagSf, case.bDeltaCoeffs, case.bType, case.bValue)
    # gamma_f of a uniform field: w*(D - D) + D = D, as the kernel computes
    return dg_c - dg_l, up_c - up_l, lo_c - up_l, ic_c - ic_l, bc_c - bc_l


def mesh_geometry(nCells, lower, upper, bFaceCells, faceStart, facePoints, points):
    """primitiveMesh face and cell geometry and the surfaceInterpolation
    coefficients of a polyhedral mesh (faces: internal then boundary).
    Returns a dict of Cf, Sf, magSf, bCf, bSf, bMagSf, C, V, weights,
    deltaCoeffs, nonOrthDeltaCoeffs, bDeltaCoeffs, bNonOrthDeltaCoeffs."""
    lower = np.ascontiguousarray(lower, np.int32)
    upper = np.ascontiguousarray(upper, np.int32)
    bfc = np.ascontiguousarray(bFaceCells, np.int32)
    fs = np.ascontiguousarray(faceStart, np.int32)
    fp = np.ascontiguousarray(facePoints, np.int32)
    pts = _f(points)
    F, B = len(lower), len(bfc)
    nAll = F + B
    Cf, Sf, magSf = np.empty(3 * nAll), np.empty(3 * nAll), np.empty(nAll)
    lib().or_face_geometry(nAll, _p(fs), _p(fp), _p(pts), _p(Cf), _p(Sf), _p(magSf))
    g = {"Cf": Cf[:3 * F].copy(), "Sf": Sf[:3 * F].copy(), "magSf": magSf[:F].copy(),
         "bCf": Cf[3 * F:].copy(), "bSf": Sf[3 * F:].copy(), "bMagSf": magSf[F:].copy()}
    C, V = np.empty(3 * nCells), np.empty(nCells)
    lib().or_cell_geometry(nCells, F, _p(lower), _p(upper), B, _p(bfc), _p(g["Cf"]),
                           _p(g["Sf"]), _p(g["bCf"]), _p(g["bSf"]), _p(C), _p(V))
    g["C"], g["V"] = C, V
    for k in ("weights", "deltaCoeffs", "nonOrthDeltaCoeffs"):
        g[k] = np.empty(F)
    for k in ("bDeltaCoeffs", "bNonOrthDeltaCoeffs"):
        g[k] = np.empty(B)
    lib().or_surface_interpolation(F, _p(lower), _p(upper), _p(C), _p(g["Cf"]), _p(g["Sf"]),
                                   _p(g["magSf"]), _p(g["weights"]), _p(g["deltaCoeffs"]),
                                   _p(g["nonOrthDeltaCoeffs"]), B, _p(bfc), _p(g["bCf"]),
                                   _p(g["bSf"]), _p(g["bMagSf"]), _p(g["bDeltaCoeffs"]),
                                   _p(g["bNonOrthDeltaCoeffs"]))
    return g


def set_sum_order(order):
    """0: the reference's serial reductions; 1: blocks of 256 summed, then
    the block sums (a different order of the same additions)."""
    lib().or_set_sum_order(int(order))
