"""Serving benchmark: independent Poisson solve requests on one MI355X.

Requests: hex boxes n^3 with n = round(exp(U[ln 16, ln 128])), each with a
seeded log-normal diffusivity (sigma ~ U[0.5, 1.5], correlation length 8
cells) and a seeded right-hand side, fvm::laplacian with the x-min wall
fixed; mixed precision PCG (FP32 inner, FP64 residual) to 1e-8.  The
matrices come from a pool of distinct seeded systems (``pool`` of them,
reused round robin); every request carries its own seeded right-hand side
(generated before the clock starts, so generation stays off it).  First
the saturated throughput is measured (all requests queued at once), then
Poisson arrivals at fractions of it; latency is completion minus
submission on the server clock.

    python -m foamhip.serve_bench --requests 4096 --rates 0.5,0.8,1.0 --concurrent 16
"""
import argparse
import json
import math
import os
import sys
import threading
import time

import numpy as np

from . import lduAddressing
from .cases import BoxCase
from .serve import Server

CONTROLS = {"solver": "PCG", "preconditioner": "DIC", "tolerance": 1e-8, "relTol": 0,
            "precision": "mixed", "maxIter": 5000}


def build_pool(n_systems, seed=2024, nmin=16, nmax=128):
    rng = np.random.default_rng(seed)
    meshes = {}
    pool = []
    for i in range(n_systems):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        sigma = float(rng.uniform(0.5, 1.5))
        if n not in meshes:
            c = BoxCase(n)
            meshes[n] = (c, lduAddressing(c.N, c.lower, c.upper))
        c, mesh = meshes[n]
        gamma = c.lognormal(1000 + i, sigma, 8.0)
        diag, upper, source = c.laplacian_system(gamma, c.rhs(5000 + i))
        pool.append({"n": n, "sigma": sigma, "mesh": mesh, "diag": diag, "upper": upper,
                     "source": source, "case": c})
    return pool


def request_sources(pool, n_requests, seed):
    """Per request k: the right-hand side U[-1,1] minus its mean, seeded by
    (seed, k), on the mesh of pool system k mod len(pool) (the x-min wall's
    value is 0, so the boundary adds nothing to the source)."""
    out = []
    for k in range(n_requests):
        s = pool[k % len(pool)]
        out.append(s["case"].rhs(seed * 1000003 + k))
    return out


def run_stream(srv, pool, n_requests, rate, rng, sources=None):
    """Submit n_requests (Poisson arrivals at `rate`/s, or all at once when
    rate is None) and collect every result; returns the result records."""
    results = []
    done = threading.Event()

    def collect():
        while len(results) < n_requests:
            r = srv.wait(1.0)
            if r is None:
                continue
            if r["status"] != 0:
                raise RuntimeError(f"request {r['id']} failed: {r['message']}")
            srv.fetch(r["id"])
            results.append(r)
        done.set()

    th = threading.Thread(target=collect, daemon=True)
    th.start()
    t0 = srv.now()
    t_next = t0
    window = 4 * srv.max_concurrent
    for k in range(n_requests):
        if rate is None:
            # saturated: keep a bounded window of queued requests
            while k - len(results) >= window:
                time.sleep(0.0005)
        else:
            t_next += rng.exponential(1.0 / rate)
            while True:
                dt = t_next - srv.now()
                if dt <= 0:
                    break
                time.sleep(min(dt, 0.01))
        s = pool[k % len(pool)]
        src = s["source"] if sources is None else sources[k]
        srv.submit(k, s["mesh"], s["diag"], s["upper"], src, CONTROLS)
    th.join()
    return results, t0


def summarize(results, t0):
    lat = np.array([r["end"] - r["submit"] for r in results])
    t_end = max(r["end"] for r in results)
    return {"requests": len(results), "solves_per_s": len(results) / (t_end - t0),
            "p50_ms": 1e3 * float(np.percentile(lat, 50)),
            "p99_ms": 1e3 * float(np.percentile(lat, 99)),
            "mean_ms": 1e3 * float(lat.mean()),
            "mean_iterations": float(np.mean([r["nIterations"] for r in results]))}


def benchmark(n_requests=4096, rates=(0.5, 0.8, 1.0), pool_size=48, max_concurrent=8,
              seed=7, log=print):
    pool = build_pool(pool_size)
    cells = np.mean([p["n"] ** 3 for p in pool])
    log(f"pool of {pool_size} systems, mean {cells:.0f} cells")
    srv = Server(maxConcurrent=max_concurrent, borrowInputs=True)
    rng = np.random.default_rng(seed)
    # warm up: every pooled system once (resident matrices and kernels)
    res, t0 = run_stream(srv, pool, len(pool), None, rng)
    sat_res, t0 = run_stream(srv, pool, n_requests, None, rng,
                             request_sources(pool, n_requests, 1))
    sat = summarize(sat_res, t0)
    log(f"saturated: {sat}")
    out = {"saturated": sat, "rates": {}}
    for i, f in enumerate(rates):
        lam = f * sat["solves_per_s"]
        r, t0 = run_stream(srv, pool, n_requests, lam, rng,
                           request_sources(pool, n_requests, 2 + i))
        s = summarize(r, t0)
        s["arrival_rate_per_s"] = lam
        out["rates"][str(f)] = s
        log(f"rate {f} x saturated ({lam:.1f}/s): {s}")
    srv.close()
    out["pool"] = pool_size
    out["mean_cells"] = float(cells)
    out["max_concurrent"] = max_concurrent
    return out


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--hw-queues", type=int, default=16,
                   help="GPU_MAX_HW_QUEUES for this process (0: keep the environment's)")
    p.add_argument("--requests", type=int, default=4096)
    p.add_argument("--rates", default="0.5,0.8,1.0")
    p.add_argument("--pool", type=int, default=48)
    p.add_argument("--concurrent", type=int, default=16)
    a = p.parse_args()
    # hardware queues for the concurrent solves' streams (HIP's default is
    # 4); set before the runtime starts, over an exported value
    if a.hw_queues > 0:
        os.environ["GPU_MAX_HW_QUEUES"] = str(a.hw_queues)
    out = benchmark(a.requests, tuple(float(x) for x in a.rates.split(",")), a.pool,
                    a.concurrent, log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
