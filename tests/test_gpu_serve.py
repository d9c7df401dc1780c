"""The request loop: every result of a mixed stream of concurrent requests
is bitwise the result of the same request solved alone, admission by the
memory budget still completes every request, and the timestamps order."""
import numpy as np
import pytest

import foamhip as fh
from foamhip.cases import BoxCase
from foamhip.serve import Server

pytestmark = pytest.mark.gpu


def systems():
    out = []
    for i, (n, ctl) in enumerate([
            (8, {"solver": "PCG", "preconditioner": "DIC", "precision": "mixed"}),
            (12, {"solver": "PCG", "preconditioner": "DIC"}),
            (16, {"solver": "PCG", "preconditioner": "DIC", "precision": "mixed"}),
            (20, {"solver": "GAMG", "smoother": "GaussSeidel"}),
            (24, {"solver": "PCG", "preconditioner": "DIC", "precision": "mixed"}),
            (10, {"solver": "PBiCGStab", "preconditioner": "diagonal"})]):
        c = BoxCase(n)
        mesh = fh.lduAddressing(c.N, c.lower, c.upper)
        gamma = c.lognormal(40 + i, 0.8, 4.0)
        diag, upper, source = c.laplacian_system(gamma, c.rhs(70 + i))
        out.append({"mesh": mesh, "diag": diag, "upper": upper, "source": source,
                    "controls": dict(ctl, tolerance=1e-8, relTol=0, maxIter=3000)})
    return out


def alone(s, source):
    m = fh.lduMatrix(s["mesh"], s["diag"], s["upper"])
    psi = np.zeros(s["mesh"].nCells)
    perf = fh.solver_New("p", m, s["controls"]).solve(psi, source)
    return psi, perf


@pytest.mark.parametrize("budget", [32 << 30, 1])
def test_mix_equals_alone(budget):
    S = systems()
    reqs = []
    for k in range(30):
        s = S[(k * 7) % len(S)]
        src = s["source"] * (1.0 + 0.01 * k)
        reqs.append((s, src))
    ref = [alone(s, src) for s, src in reqs]
    srv = Server(maxConcurrent=6, memoryBudget=budget)
    for k, (s, src) in enumerate(reqs):
        srv.submit(k, s["mesh"], s["diag"], s["upper"], src, s["controls"])
    got = {}
    while len(got) < len(reqs):
        r = srv.wait(60.0)
        assert r is not None, "server stalled"
        assert r["status"] == 0, r["message"]
        assert r["submit"] <= r["start"] <= r["end"]
        got[r["id"]] = (srv.fetch(r["id"]), r)
    assert srv.pending() == 0
    srv.close()
    for k, (psi_ref, perf_ref) in enumerate(ref):
        psi, r = got[k]
        np.testing.assert_array_equal(psi, psi_ref)
        assert r["nIterations"] == perf_ref.nIterations
