"""Autotuner for the persistent orbit kernel.

The kernel's launch configuration has two run-time knobs that trade registers, LDS and
memory-pipeline use against each other, and their best setting depends on the expansion
(number of harmonics, table size) and on the batch (stage 1):
  fetch  - how a lane reads the 2-D splines: the per-cell polynomial tables ("cell", one
           contiguous run per point), or the node records with per-lane 16-byte loads
           ("vec2"), per-lane 8-byte loads ("scalar") or the wavefront-cooperative gather
           staged through 20 KB of LDS per wavefront ("coop");
  waves  - the wavefronts per SIMD the kernel is compiled for: 1 (up to 512 registers per
           lane, no spills) or 2 (256 registers, twice the wavefronts to hide latency).
and the persistent schedule has two more that decide how the batch's work lands on the
wavefronts (stage 2, with the stage-1 choice fixed):
  groups - one wavefront in `stride` integrates the longest expected orbits with 8 lanes per
           orbit (-1: the library's automatic choice, 0: off);
  probe  - steps of the work-order probe that ranks the orbits before the launch (0: the static
           cost key).
`tune()` integrates a sample of the workload under every setting, checks that all settings
give the same orbits (bitwise within the record modes, to rounding between the record modes,
the per-cell tables and group mode), and returns (and optionally saves) the fastest; `apply()`
sets it for the process.  The workgroup size (256 threads: one wavefront per SIMD of a CU with
one LDS copy of the staged tables) and the LDS staging of the grids, power-law arrays and bucket
table are fixed: 256 is the only size at which a CU's four SIMDs share one copy of the ~120 KB of
staged data, and reading those arrays from global memory measured slower in every case.

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


def current_schedule():
    """group mode and probe settings (agx_set_orbit_groups / agx_set_orbit_probe)"""
    gs, pr = np.zeros(1, np.int32), np.zeros(1, np.int32)
    go, gm, pm = (np.zeros(1, np.int64) for _ in range(3))
    check(lib.agx_get_orbit_schedule(gs.ctypes.data, go.ctypes.data, gm.ctypes.data, pr.ctypes.data,
                                     pm.ctypes.data), "get_orbit_schedule")
    return {"group_stride": int(gs[0]), "group_orbits": int(go[0]), "group_min_batch": int(gm[0]),
            "probe_steps": int(pr[0]), "probe_min_batch": int(pm[0])}


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
    sch = current_schedule()
    sch.update({k: cfg[k] for k in sch if k in cfg})
    check(lib.agx_set_orbit_groups(int(sch["group_stride"]), int(sch["group_orbits"]),
                                   int(sch["group_min_batch"])), "set_orbit_groups")
    check(lib.agx_set_orbit_probe(int(sch["probe_steps"]), int(sch["probe_min_batch"])), "set_orbit_probe")
    return dict(current(), **current_schedule())


def tune(potential, n=65536, ic=None, time_total=None, accuracy=1e-8, fetches=("cell", "vec2", "coop", "scalar"),
         waves=(1, 2), groups=(-1, 0, 8, 16, 32), probes=(64, 0), repeats=2, out=None, verbose=False):
    """time every (fetch, waves) setting on a sample of orbits, then every (groups, probe)
    setting with the fastest of those, and return the fastest overall.  The sample counts as a
    batch large enough for group mode and the probe (their minimum batch sizes are lowered to
    it while tuning and saved with the result).  By default the sample is n
    triaxial-Dehnen-like orbits (positions and velocities drawn as in the batching workload)
    with integration times of 1-30 circular periods."""
    import agama_amd
    from . import workloads
    if ic is None:
        ic, T = workloads.triaxial_dehnen_orbits(n, lambda p: potential.potential(p), potential.coefs(),
                                                 seed=4242, uneven=True)
        time_total = T / 10
    saved = dict(current(), **current_schedule())
    sample = len(ic)
    min_batch = {"group_min_batch": min(sample, saved["group_min_batch"]),
                 "probe_min_batch": min(sample, saved["probe_min_batch"])}
    results, ref = [], None

    def run(c):
        nonlocal ref
        apply(dict(c, fetch_eval=saved["fetch_eval"], num_blocks=0, group_orbits=-1, **min_batch))
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
            # rounding-level differences grow along chaotic orbits: nearly all orbits agree to
            # 1e-8 and none is lost
            d = np.linalg.norm(ref[:, :3] - info["final"][:, :3], axis=1) / np.linalg.norm(ref[:, :3], axis=1)
            if not (np.mean(d <= 1e-8) >= 0.99 and np.all(info["status"] == 0)):
                raise RuntimeError(f"{c} changes the orbits")
        results.append(dict(c, seconds=best_t, steps_per_s=steps / best_t, identical_to_first=same))
        if verbose:
            print(f"tune: fetch={c['fetch_orbit']:6s} waves={c['waves']} groups={c['group_stride']:3d} "
                  f"probe={c['probe_steps']:3d}: {steps / best_t / 1e6:8.2f} Msteps/s", flush=True)

    try:
        # stage 1: (fetch, waves) under the current schedule
        for fe in fetches:
            for w in waves:
                run(dict(fetch_orbit=fe, waves=w, group_stride=saved["group_stride"],
                         probe_steps=saved["probe_steps"]))
        # stage 2: (groups, probe) with the fastest of stage 1
        b = max(results, key=lambda r: r["steps_per_s"])
        for g in groups:
            for pr in probes:
                if g != saved["group_stride"] or pr != saved["probe_steps"]:
                    run(dict(fetch_orbit=b["fetch_orbit"], waves=b["waves"], group_stride=g, probe_steps=pr))
    finally:
        apply(saved)
    best = max(results, key=lambda r: r["steps_per_s"])
    cfg = {"fetch_eval": saved["fetch_eval"], "fetch_orbit": best["fetch_orbit"], "waves": best["waves"],
           "num_blocks": 0, "group_stride": best["group_stride"], "group_orbits": -1,
           "probe_steps": best["probe_steps"], **min_batch, "sample_orbits": sample, "candidates": results}
    if out:
        with open(out, "w") as f:
            json.dump(cfg, f, indent=1)
    return cfg
