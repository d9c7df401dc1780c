"""ctypes loading of the in-tree native libraries.

libfoamhip.so is the product (HIP kernels + C-ABI, include/foamhip.h);
libfoamhip_case.so builds the synthetic hex-box cases.  Both are built
in-tree by ``make -C openfoam-dev_amd``; a missing library is an error,
there is no fallback path.
"""
import ctypes
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_ROOT, "lib")

FH_OK = 0
STATUS_NAMES = {
    1: "FH_ERR_INVALID",
    2: "FH_ERR_HIP",
    3: "FH_ERR_NOMEM",
    4: "FH_ERR_UNSUPPORTED",
    5: "FH_ERR_NUMERIC",
    6: "FH_ERR_SCHEDULE",
}

_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_dbl = ctypes.c_double
_vp = ctypes.c_void_p
_cp = ctypes.c_char_p


class FhControls(ctypes.Structure):
    _fields_ = [
        ("solver", _i32), ("preconditioner", _i32), ("smoother", _i32),
        ("precision", _i32), ("maxIter", _i32), ("minIter", _i32),
        ("tolerance", _dbl), ("relTol", _dbl), ("nSweeps", _i32),
        ("nPreSweeps", _i32), ("preSweepsLevelMultiplier", _i32),
        ("maxPreSweeps", _i32), ("nPostSweeps", _i32),
        ("postSweepsLevelMultiplier", _i32), ("maxPostSweeps", _i32),
        ("nFinestSweeps", _i32), ("scaleCorrection", _i32),
        ("directSolveCoarsest", _i32), ("nCellsInCoarsestLevel", _i32),
        ("mergeLevels", _i32), ("nVcycles", _i32), ("innerRelTol", _dbl),
        ("maxRefinements", _i32), ("innerPreconditioner", _i32), ("agglomerator", _i32),
        ("preconditionerControls", ctypes.c_void_p),
    ]


class FhPerf(ctypes.Structure):
    _fields_ = [
        ("solverName", ctypes.c_char * 64), ("initialResidual", _dbl),
        ("finalResidual", _dbl), ("nIterations", _i32), ("converged", _i32),
        ("singular", _i32), ("nRefinements", _i32), ("normFactor", _dbl),
        ("solveMilliseconds", _dbl),
    ]


class FhMeshInfo(ctypes.Structure):
    _fields_ = [
        ("nCells", _i32), ("nFaces", _i32), ("nLevels", _i32), ("nParts", _i32),
        ("boxNx", _i32), ("boxNy", _i32), ("boxNz", _i32), ("tiled", _i32),
        ("deviceBytes", _i64),
    ]


class FhServerConfig(ctypes.Structure):
    _fields_ = [("maxConcurrent", _i32), ("borrowInputs", _i32), ("memoryBudget", _i64)]


class FhRequest(ctypes.Structure):
    _fields_ = [
        ("id", _i64), ("mesh", _vp), ("diag", _vp), ("upper", _vp), ("lower", _vp),
        ("source", _vp), ("psi0", _vp), ("controls", FhControls),
    ]


class FhResult(ctypes.Structure):
    _fields_ = [
        ("id", _i64), ("status", _i32), ("nCells", _i32), ("perf", FhPerf),
        ("submitTime", _dbl), ("startTime", _dbl), ("endTime", _dbl),
        ("message", ctypes.c_char * 160),
    ]


# every symbol declared in include/foamhip.h: name -> (restype, argtypes)
FH_SYMBOLS = {
    "fh_version": (_i32, []),
    "fh_last_error": (_cp, []),
    "fh_set_device": (_i32, [_i32]),
    "fh_controls_default": (None, [ctypes.POINTER(FhControls)]),
    "fh_solver_lookup": (_i32, [_cp, ctypes.POINTER(_i32)]),
    "fh_preconditioner_lookup": (_i32, [_cp, ctypes.POINTER(_i32)]),
    "fh_smoother_lookup": (_i32, [_cp, ctypes.POINTER(_i32)]),
    "fh_gamg_direction": (_i32, []),
    "fh_gamg_reset_direction": (None, []),
    "fh_gamg_agglomeration_levels": (_i32, [_vp, ctypes.POINTER(FhControls),
                                            ctypes.POINTER(_i32)]),
    "fh_gamg_agglomeration_level": (_i32, [_vp, ctypes.POINTER(FhControls), _i32, _vp, _vp,
                                           _vp, _vp, _vp, _vp]),
    "fh_mesh_create": (_i32, [_i32, _i32, _vp, _vp, ctypes.POINTER(_vp)]),
    "fh_mesh_create_coupled": (_i32, [_i32, _i32, _vp, _vp, _i32, _vp, _vp, _vp,
                                      ctypes.POINTER(_vp)]),
    "fh_matrix_set_interfaces": (_i32, [_vp, _vp, _vp]),
    "fh_face_geometry": (_i32, [_i32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fh_cell_geometry": (_i32, [_vp] * 8),
    "fh_surface_interpolation": (_i32, [_vp] * 14),
    "fh_norm_factor_device_noalloc": (_i32, [_vp, _vp, _vp, _vp]),
    "fh_mesh_info_get": (_i32, [_vp, ctypes.POINTER(FhMeshInfo)]),
    "fh_mesh_destroy": (_i32, [_vp]),
    "fh_matrix_create": (_i32, [_vp, ctypes.POINTER(_vp)]),
    "fh_matrix_set": (_i32, [_vp, _vp, _vp, _vp]),
    "fh_matrix_set_device": (_i32, [_vp, _vp, _vp, _vp]),
    "fh_matrix_destroy": (_i32, [_vp]),
    "fh_solve": (_i32, [_vp, ctypes.POINTER(FhControls), _vp, _vp, ctypes.POINTER(FhPerf)]),
    "fh_solve_device": (_i32, [_vp, ctypes.POINTER(FhControls), _vp, _vp, ctypes.POINTER(FhPerf)]),
    "fh_amul": (_i32, [_vp, _vp, _vp]),
    "fh_tmul": (_i32, [_vp, _vp, _vp]),
    "fh_residual": (_i32, [_vp, _vp, _vp, _vp]),
    "fh_sum_a": (_i32, [_vp, _vp]),
    "fh_matrix_get": (_i32, [_vp, _vp, _vp, _vp]),
    "fh_mesh_set_face_areas": (_i32, [_vp, _vp, _vp]),
    "fh_h": (_i32, [_vp, _i32, _vp, _vp]),
    "fh_h1": (_i32, [_vp, _vp]),
    "fh_face_h": (_i32, [_vp, _i32, _vp, _vp]),
    "fh_sum_mag_off_diag": (_i32, [_vp, _vp]),
    "fh_h_device": (_i32, [_vp, _i32, _vp, _vp]),
    "fh_face_h_device": (_i32, [_vp, _i32, _vp, _vp]),
    "fh_norm_factor": (_i32, [_vp, _vp, _vp, ctypes.POINTER(_dbl)]),
    "fh_reciprocal_d": (_i32, [_vp, _i32, _vp]),
    "fh_precondition": (_i32, [_vp, _i32, _vp, _vp]),
    "fh_precondition_t": (_i32, [_vp, _i32, _vp, _vp]),
    "fh_smooth": (_i32, [_vp, _i32, _vp, _vp, _i32]),
    "fh_precondition_f32": (_i32, [_vp, _i32, _vp, _vp]),
    "fh_profile_kernel": (_i32, [_vp, _i32, _i32, _i32, _i32, ctypes.POINTER(_dbl),
                                 ctypes.POINTER(_dbl)]),
    "fh_kernel_timing": (_i32, [_vp, ctypes.c_uint32]),
    "fh_kernel_timing_get": (_i32, [_vp, _i32, ctypes.POINTER(_dbl), ctypes.POINTER(_i64)]),
    "fh_profile_stream": (_i32, [_i32, _i64, ctypes.POINTER(_dbl)]),
    "fh_boundary_create": (_i32, [_vp, _i32, _vp, ctypes.POINTER(_vp)]),
    "fh_boundary_destroy": (_i32, [_vp]),
    "fh_assemble_laplacian": (_i32, [_vp, _vp] + [_vp] * 10 + [_vp, _i32]),
    "fh_assemble_convection_diffusion": (_i32, [_vp, _vp, _vp, _dbl] + [_vp] * 9 + [_vp, _i32]),
    "fh_grad_gauss_linear": (_i32, [_vp, _vp] + [_vp] * 10),
    "fh_relax": (_i32, [_vp, _vp, _dbl, _i32, _vp, _vp, _vp, _vp]),
    "fh_flux": (_i32, [_vp, _vp, _i32] + [_vp] * 7),
    "fh_server_config_default": (None, [ctypes.POINTER(FhServerConfig)]),
    "fh_server_create": (_i32, [ctypes.POINTER(FhServerConfig), ctypes.POINTER(_vp)]),
    "fh_server_submit": (_i32, [_vp, ctypes.POINTER(FhRequest)]),
    "fh_server_wait": (_i32, [_vp, _dbl, ctypes.POINTER(FhResult), ctypes.POINTER(_i32)]),
    "fh_server_fetch": (_i32, [_vp, _i64, _vp]),
    "fh_server_now": (_i32, [_vp, ctypes.POINTER(_dbl)]),
    "fh_server_pending": (_i32, [_vp, ctypes.POINTER(_i64)]),
    "fh_server_destroy": (_i32, [_vp]),
}

CASE_SYMBOLS = {
    "fhc_box_nfaces": (_i64, [_i32, _i32, _i32]),
    "fhc_box_nbfaces": (_i64, [_i32, _i32, _i32]),
    "fhc_box_faces": (None, [_i32, _i32, _i32, _vp, _vp, _vp]),
    "fhc_box_bfaces": (None, [_i32, _i32, _i32, _vp, _vp]),
    "fhc_box_geometry": (None, [_i32, _i32, _i32, _dbl, _dbl, _dbl] + [_vp] * 11),
    "fhc_uniform": (None, [_u64, _i64, _dbl, _dbl, _vp]),
    "fhc_rhs": (None, [_u64, _i64, _vp]),
    "fhc_lognormal": (None, [_i32, _i32, _i32, _u64, _dbl, _dbl, _vp]),
    "fhc_fourier_flux": (None, [_i32, _i32, _i32, _u64] + [_vp] * 7),
    "fhc_graded_box_points": (None, [_i32, _i32, _i32, _dbl, _dbl, _dbl, _dbl, _dbl, _u64, _vp]),
    "fhc_box_face_points": (None, [_i32, _i32, _i32, _vp]),
}


def _load(name, symbols):
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `make -C {PKG_ROOT}` "
            "(or __graft_entry__.build()); there is no fallback implementation")
    if name == "libfoamhip.so":
        # One HIP runtime per process: PyTorch (used for device buffers by
        # callers and tests) ships its own libamdhip64 under the same soname.
        # Loaded first, it is the one libfoamhip binds to as well; loaded
        # after libfoamhip, torch would run on the other runtime and find no
        # device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
    lib = ctypes.CDLL(path)
    for sym, (res, args) in symbols.items():
        fn = getattr(lib, sym)
        fn.restype = res
        fn.argtypes = args
    return lib


_fh = None
_case = None


def fh():
    """The product library (raises if it is not built)."""
    global _fh
    if _fh is None:
        _fh = _load("libfoamhip.so", FH_SYMBOLS)
    return _fh


def case():
    global _case
    if _case is None:
        _case = _load("libfoamhip_case.so", CASE_SYMBOLS)
    return _case
