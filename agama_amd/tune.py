"""Autotuner for the persistent orbit kernel.

The kernel's launch configuration has two run-time knobs that trade registers, LDS and
memory-pipeline use against each other, and their best setting depends on the expansion
(number of harmonics, table size) and on the batch:
  fetch  - how a lane reads the 2-D splines: the per-cell polynomial tables ("cell", one
           contiguous run per point), or the node records with per-lane 16-byte loads
           ("vec2"), per-lane 8-byte loads ("scalar") or the wavefront-cooperative gather
           staged through 20 KB of LDS per wavefront ("coop");
  waves  - the wavefronts per SIMD the kernel is compiled for: 1 (up to 512 registers per
           lane, no spills) or 2 (256 registers, twice the wavefronts to hide latency).
`tune()` integrates a sample of the workload under every setting, checks that all settings
give the same orbits (bitwise within the record modes, to rounding between the record modes and
the per-cell tables), and returns (and optionally saves) the fastest; `apply()` sets it for the
process.

    import agama_amd, agama_amd.tune as tune
    pot = agama_amd.Potential(file="halo.coef")
    best = tune.tune(pot, n=65536, out="orbit_tuning.json")
    tune.apply(best)            # or tune.apply("orbit_tuning.json") in a later process
"""
import json
import time

import numpy as np

from ._native import lib, check

FETCH = {"scalar": 0, "vec2": 1, "coop": 2, "cell": 3}


def current():
    vals = [np.zeros(1, np.int32) for _ in range(4)]
    check(lib.agx_get_orbit_config(*[v.ctypes.data for v in vals]), "get_orbit_config")
    inv = {v: k for k, v in FETCH.items()}
    return {"num_blocks": int(vals[0][0]), "fetch_eval": inv[int(vals[1][0])],
            "fetch_orbit": inv[int(vals[2][0])], "waves": int(vals[3][0])}


def apply(cfg):
    """set a configuration (a dict from tune() or the path of a saved one)"""
    if isinstance(cfg, str):
        with open(cfg) as f:
            cfg = json.load(f)
    cur = current()
    check(lib.agx_set_fetch(FETCH[cfg.get("fetch_eval", cur["fetch_eval"])],
                            FETCH[cfg.get("fetch_orbit", cur["fetch_orbit"])]), "set_fetch")
    check(lib.agx_set_orbit_waves(int(cfg.get("waves", cur["waves"]))), "set_orbit_waves")
    check(lib.agx_set_orbit_launch(int(cfg.get("num_blocks", 0)), 0), "set_orbit_launch")
    return current()


def tune(potential, n=65536, ic=None, time_total=None, accuracy=1e-8, fetches=("cell", "vec2", "coop", "scalar"),
         waves=(1, 2), repeats=2, out=None, verbose=False):
    """time every (fetch, waves) setting on a sample of orbits and return the fastest.
    By default the sample is n triaxial-Dehnen-like orbits (positions and velocities drawn as
    in the batching workload) with integration times of 1-30 circular periods."""
    import agama_amd
    from . import workloads
    if ic is None:
        ic, T = workloads.triaxial_dehnen_orbits(n, lambda p: potential.potential(p), potential.coefs(),
                                                 seed=4242, uneven=True)
        time_total = T / 10
    saved = current()
    results, ref = [], None
    try:
        for fe in fetches:
            for w in waves:
                apply({"fetch_eval": saved["fetch_eval"], "fetch_orbit": fe, "waves": w, "num_blocks": 0})
                best_t, info = np.inf, {}
                for _ in range(repeats):
                    info = {}
                    t0 = time.perf_counter()
                    agama_amd.orbit(potential=potential, ic=ic, time=time_total, accuracy=accuracy, info=info)
                    best_t = min(best_t, time.perf_counter() - t0)
                steps = int(info["nsteps"].sum())
                if ref is None:
                    ref = info["final"]
                    same = True
                else:
                    same = bool(np.array_equal(ref, info["final"]))
                    # rounding-level differences grow along chaotic orbits: nearly all orbits
                    # agree to 1e-8 and none is lost
                    d = np.linalg.norm(ref[:, :3] - info["final"][:, :3], axis=1) / \
                        np.linalg.norm(ref[:, :3], axis=1)
                    if not (np.mean(d <= 1e-8) >= 0.99 and np.all(info["status"] == 0)):
                        raise RuntimeError(f"fetch={fe} waves={w} changes the orbits")
                rec = {"fetch_orbit": fe, "waves": w, "seconds": best_t, "steps_per_s": steps / best_t,
                       "identical_to_first": same}
                results.append(rec)
                if verbose:
                    print(f"tune: fetch={fe:6s} waves={w}: {steps / best_t / 1e6:8.2f} Msteps/s", flush=True)
    finally:
        apply(saved)
    best = max(results, key=lambda r: r["steps_per_s"])
    cfg = {"fetch_eval": saved["fetch_eval"], "fetch_orbit": best["fetch_orbit"], "waves": best["waves"],
           "num_blocks": 0, "sample_orbits": len(ic), "candidates": results}
    if out:
        with open(out, "w") as f:
            json.dump(cfg, f, indent=1)
    return cfg
