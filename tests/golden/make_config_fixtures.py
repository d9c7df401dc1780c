#!/usr/bin/env python3
"""Writes the oracle's full-size solutions of the measured configurations
as compact summaries (tests/config_cases.py: norm, sum, 32 seeded Gaussian
projections, 2048 seeded entries) for the GPU tests, which cannot afford
the oracle's minutes per solve at these sizes:

  c2_pbicgstab_dilu_128.json  PBiCGStab+DILU, 128^3 upwind convection-
                              diffusion (D 1e-3), tolerance 1e-8, and the
                              same solve with b perturbed by 1e-15 relative
                              and with its global sums in another order (the
                              reference's own spread)
  c3_pcg_dic_256.json         PCG+DIC, 256^3 log-normal Poisson, tolerance
                              1e-8 and 1e-12 (the tight solution the error
                              statements are measured against), the 1e-8
                              solve with b perturbed by 1e-15 relative, and
                              with its global sums in another order

Run from the repository root after building the oracle:
    python tests/golden/make_config_fixtures.py [c2] [c3] [c3-perturbed] [c2-reordered]
        [c3-reordered]
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "openfoam-dev_amd"), os.path.join(ROOT, "oracle"),
                os.path.dirname(HERE)]

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
import config_cases as CC  # noqa: E402


def perf_dict(p):
    return {k: (bool(v) if isinstance(v, (bool, np.bool_)) else
                int(v) if isinstance(v, (int, np.integer)) else float(v))
            for k, v in p.items()}


def solve(om, ctl, N, b):
    t0 = time.time()
    psi, perf = om.solve(ctl, np.zeros(N), b)
    print(f"  {ctl}: {perf['nIterations']} iterations, {time.time() - t0:.0f} s", flush=True)
    return {"controls": ctl, "perf": perf_dict(perf), "psi": CC.summary(psi)}


def c2():
    c, a, diag, upper, lower, b = CC.c2_convection_diffusion()
    om = O.Matrix(a, diag, upper, lower)
    out = {"case": "c2_convection_diffusion(n=128, seed=2, D=1e-3)",
           "solve": solve(om, CC.C2_CONTROLS, c.N, b)}
    out["perturbed"] = perturbed(om, CC.C2_CONTROLS, c.N, b, out["solve"]["psi"])
    with open(os.path.join(HERE, "c2_pbicgstab_dilu_128.json"), "w") as f:
        json.dump(out, f)


def perturbed(om, ctl, N, b, ref_summary):
    """The reference's own sensitivity: the same solve with b perturbed by a
    relative 1e-15 (seeded), i.e. at the level of a different summation
    order."""
    rng = np.random.default_rng(5)
    bp = b * (1.0 + 1e-15 * rng.standard_normal(N))
    psi_p, perf_p = om.solve(ctl, np.zeros(N), bp)
    out = {"perf": perf_dict(perf_p), "rel_l2_vs_solve": CC.rel_diff_estimate(psi_p, ref_summary)}
    print("  perturbed:", out, flush=True)
    return out


def reordered(om, ctl, N, b, ref_summary):
    """The reference's sensitivity to the order of its global sums: the same
    solve with the reductions summed in blocks of 256 (oracle.set_sum_order),
    the only arithmetic in which the device solve differs from the oracle."""
    O.set_sum_order(1)
    try:
        psi_r, perf_r = om.solve(ctl, np.zeros(N), b)
    finally:
        O.set_sum_order(0)
    out = {"perf": perf_dict(perf_r), "rel_l2_vs_solve": CC.rel_diff_estimate(psi_r, ref_summary)}
    print("  reordered sums:", out, flush=True)
    return out


def add_reordered(name, build):
    """Adds the reordered-sum spread to an existing fixture."""
    path = os.path.join(HERE, name)
    with open(path) as f:
        out = json.load(f)
    om, ctl, N, b = build()
    out["reordered"] = reordered(om, ctl, N, b, out["solve"]["psi"])
    with open(path, "w") as f:
        json.dump(out, f)


def c2_system():
    c, a, diag, upper, lower, b = CC.c2_convection_diffusion()
    return O.Matrix(a, diag, upper, lower), CC.C2_CONTROLS, c.N, b


def c3_system():
    c, a, diag, upper, b = CC.c3_poisson_lognormal()
    return O.Matrix(a, diag, upper), CC.C3_CONTROLS, c.N, b


def c3(only_perturbed=False):
    c, a, diag, upper, b = CC.c3_poisson_lognormal()
    om = O.Matrix(a, diag, upper)
    path = os.path.join(HERE, "c3_pcg_dic_256.json")
    if only_perturbed:
        with open(path) as f:
            out = json.load(f)
    else:
        out = {"case": "c3_poisson_lognormal(n=256)",
               "solve": solve(om, CC.C3_CONTROLS, c.N, b),
               "tight": solve(om, dict(CC.C3_CONTROLS, tolerance=CC.C3_TIGHT), c.N, b)}
    out["perturbed"] = perturbed(om, CC.C3_CONTROLS, c.N, b, out["solve"]["psi"])
    with open(path, "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    which = sys.argv[1:] or ["c2", "c3"]
    for w in which:
        print(w, flush=True)
        {"c2": c2, "c3": c3, "c3-perturbed": lambda: c3(only_perturbed=True),
         "c2-reordered": lambda: add_reordered("c2_pbicgstab_dilu_128.json", c2_system),
         "c3-reordered": lambda: add_reordered("c3_pcg_dic_256.json", c3_system)}[w]()
