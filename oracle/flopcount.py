"""Per-unit operation counts of the reference algorithm - TEST INFRASTRUCTURE (oracle/).

Runs the oracle compiled with an operation-counting double (flopcount.cpp, built by the
Makefile as build/liboracle_flops.so) on samples of the benchmark workloads and prints the counts
that bench.py keeps as constants for its roofline figures:

  eval     Phi + grad Phi at points drawn from the triaxial Dehnen density (bench --mode eval)
  hess     Phi, grad Phi and the Hessian at the same points (bench --mode hess)
  dop853   orbits of the 1-300-period workload, DOP853 at accuracy 1e-8, final state and max dE/E:
           operations per step attempt, everything included (12 evaluations, the stage sums,
           error norm and step control, energy, the orbit's start) (bench --mode fp64)
  dprkn8   the same orbits with DPRKN8 (bench --mode fp64 --method dprkn8)

One operation: + - * / on doubles, sqrt and every libm function (flopcount.cpp).  The oracle is
the reference's algorithm restated (agama_oracle.cpp), so these are the reference's operation
counts for the same work, not those of the GPU kernels (whose per-cell tables need fewer).

Usage: python oracle/flopcount.py [--orbits 10000] [--points 20000] [--threads 8]
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle import pyoracle  # noqa: E402
from agama_amd import workloads  # noqa: E402

COEF = os.path.join(os.path.dirname(HERE), "tests", "golden", "dehnen_triax_lmax8.coef")


def count(fn, threads):
    lib = pyoracle.flops()
    lib.flc_reset(threads)
    t0 = time.perf_counter()
    r = fn()
    return lib.flc_ops(threads), r, time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--orbits", type=int, default=10000)
    ap.add_argument("--points", type=int, default=20000)
    ap.add_argument("--threads", type=int, default=min(8, os.cpu_count() or 1))
    args = ap.parse_args()
    th = args.threads
    pot = pyoracle.OraclePotential(COEF, library=pyoracle.flops())
    plain = pyoracle.OraclePotential(COEF)

    pts = workloads.sample_dehnen_positions(args.points, workloads.rng(7))
    ops, r, dt = count(lambda: pot.eval(pts, nthreads=th), th)
    same = all(np.array_equal(a, b) for a, b in zip(r, plain.eval(pts, nthreads=th)))
    print(f"eval: {ops / len(pts):.1f} operations per point ({len(pts)} density points, {dt:.1f} s; "
          f"values bitwise equal to the strict oracle: {same})")
    ops, r, dt = count(lambda: pot.eval_hess(pts, nthreads=th), th)
    print(f"hess: {ops / len(pts):.1f} operations per point")

    radii, Phi, dPhi = workloads.read_coef_file(COEF)
    ic, T = workloads.triaxial_dehnen_orbits(args.orbits, lambda p: plain.eval(p, nthreads=th)[0],
                                             (radii, Phi, dPhi), seed=42, uneven=True)
    for method in ("dop853", "dprkn8"):
        ops, r, dt = count(lambda: pot.orbits(ic, T, trajsize=1, accuracy=1e-8, nthreads=th, method=method), th)
        att, steps, evals = int(r["nattempts"].sum()), int(r["nsteps"].sum()), int(r["nevals"].sum())
        print(f"{method}: {ops / att:.1f} operations per attempt, {ops / steps:.1f} per accepted step, "
              f"{ops / evals:.1f} per evaluation ({args.orbits} orbits of the 1-300-period workload, "
              f"{steps} steps, {att} attempts, {evals} evaluations, {dt:.1f} s on {th} threads)")


if __name__ == "__main__":
    main()
