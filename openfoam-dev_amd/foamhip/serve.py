"""Python view of the request loop (fh_server_*, include/foamhip.h).

    srv = Server(maxConcurrent=8, memoryBudget=32 << 30)
    srv.submit(7, matrix_of_request, source, controls)   # copies, returns at once
    res = srv.wait(timeout=1.0)       # next completed result (dict) or None
    psi = srv.fetch(res["id"], nCells)

Requests are admitted in arrival order as solve slots and device memory
free up; solves of different requests overlap on the device, and a result
is the same bits whether it ran alone or beside others.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import FhRequest, FhResult, FhServerConfig
from . import FoamHipError, controls_from_dict, lib


def _check(status):
    if status != _lib.FH_OK:
        raise FoamHipError(status, lib().fh_last_error().decode(errors="replace"))


class Server:
    def __init__(self, maxConcurrent=8, memoryBudget=32 << 30, borrowInputs=False):
        """borrowInputs: the server reads the caller's arrays in place; they
        must stay alive and unchanged until the result is returned."""
        cfg = FhServerConfig()
        lib().fh_server_config_default(ctypes.byref(cfg))
        cfg.maxConcurrent = int(maxConcurrent)
        cfg.borrowInputs = 1 if borrowInputs else 0
        cfg.memoryBudget = int(memoryBudget)
        h = ctypes.c_void_p()
        _check(lib().fh_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self.max_concurrent = int(maxConcurrent)
        self._sizes = {}

    def now(self):
        t = ctypes.c_double()
        _check(lib().fh_server_now(self.handle, ctypes.byref(t)))
        return t.value

    def submit(self, rid, mesh, diag, upper, source, controls, lower=None, psi0=None):
        """mesh: foamhip.lduAddressing; arrays are copied before return."""
        c = controls if isinstance(controls, _lib.FhControls) else controls_from_dict(controls)
        arrs = [np.ascontiguousarray(a, np.float64) if a is not None else None
                for a in (diag, upper, lower, source, psi0)]
        ptr = lambda a: ctypes.c_void_p(a.ctypes.data) if a is not None else None
        r = FhRequest(int(rid), mesh.handle, ptr(arrs[0]), ptr(arrs[1]), ptr(arrs[2]),
                      ptr(arrs[3]), ptr(arrs[4]), c)
        _check(lib().fh_server_submit(self.handle, ctypes.byref(r)))
        self._sizes[int(rid)] = mesh.nCells
        # borrowed arrays (converted copies included) live until fetch
        self._keep = getattr(self, "_keep", {})
        self._keep[int(rid)] = arrs

    def wait(self, timeout=1.0):
        res = FhResult()
        got = ctypes.c_int32()
        _check(lib().fh_server_wait(self.handle, float(timeout), ctypes.byref(res),
                                    ctypes.byref(got)))
        if not got.value:
            return None
        p = res.perf
        return {"id": res.id, "status": res.status, "nCells": res.nCells,
                "submit": res.submitTime, "start": res.startTime, "end": res.endTime,
                "nIterations": p.nIterations, "finalResidual": p.finalResidual,
                "converged": bool(p.converged), "solverName": p.solverName.decode(),
                "message": res.message.decode(errors="replace")}

    def fetch(self, rid):
        n = self._sizes.pop(int(rid))
        getattr(self, "_keep", {}).pop(int(rid), None)
        psi = np.empty(n)
        _check(lib().fh_server_fetch(self.handle, int(rid), ctypes.c_void_p(psi.ctypes.data)))
        return psi

    def pending(self):
        c = ctypes.c_int64()
        _check(lib().fh_server_pending(self.handle, ctypes.byref(c)))
        return c.value

    def close(self):
        if getattr(self, "handle", None):
            lib().fh_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
