"""GPU parity of the DOP853 kernel's group mode: the longest orbits of a batch integrated by
groups of 8 lanes, each lane evaluating a share of the Multipole's azimuthal harmonics and the
partial Fourier sums added across the group (agx_set_orbit_groups).  Bars as for the one-lane
kernel: regular orbits agree with the oracle to 1e-7 relative in position after 100 circular
periods at accuracy 1e-8, energy drift no worse than the oracle's; every orbit of a mixed batch
(groups for the first orbits of the order, one lane for the rest) completes with the oracle's
step count; trajectory rows sampled in group mode match the oracle's.  The group results differ
from the one-lane ones (the order of additions in the Fourier sum), which shows the group path
ran."""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

COEF = os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")


@pytest.fixture(scope="module")
def setup():
    import agama_amd
    from oracle.pyoracle import OraclePotential
    return agama_amd.Potential(file=COEF), OraclePotential(COEF)


@pytest.fixture(params=[1, 3], ids=["records", "cells"])
def groups(request):
    """group mode over the node records (orbit fetch mode 1) and over the per-cell tables (3)"""
    import ctypes
    from agama_amd._native import lib
    old = (ctypes.c_int(), ctypes.c_int())
    lib.agx_get_orbit_config(None, ctypes.byref(old[0]), ctypes.byref(old[1]), None)
    assert lib.agx_set_fetch(old[0].value, request.param) == 0
    print(f"orbit fetch mode {request.param}", flush=True)
    yield lib
    assert lib.agx_set_orbit_groups(-1, -1, 65536) == 0   # the library's defaults
    assert lib.agx_set_fetch(old[0].value, old[1].value) == 0


def ics(setup, n, seed, periods=100.0, uneven=False):
    from agama_amd import workloads
    gpu, orc = setup
    return workloads.triaxial_dehnen_orbits(n, lambda p: orc.eval(p)[0], gpu.coefs(), seed=seed,
                                            periods=periods, uneven=uneven)


def test_group_mode_100_periods_match_oracle(setup, groups):
    import agama_amd
    gpu, orc = setup
    ic, T = ics(setup, 256, seed=21)
    print("case: group mode for every orbit, 256 orbits x 100 periods", flush=True)
    assert groups.agx_set_orbit_groups(1, 256, 0) == 0     # every wavefront, every orbit
    info = {}
    agama_amd.orbit(potential=gpu, ic=ic, time=T, accuracy=1e-8, info=info)
    ref = orc.orbits(ic, T, trajsize=1, accuracy=1e-8)
    assert np.all(info["status"] == 0)
    err = np.linalg.norm(info["final"][:, :3] - ref["final"][:, :3], axis=1) / np.linalg.norm(ref["final"][:, :3], axis=1)
    same = info["nsteps"] == ref["nsteps"]
    print(f"position rel err median {np.median(err):.2e}, 90% {np.quantile(err, 0.9):.2e}; "
          f"identical step counts {same.mean():.3f}")
    assert np.median(err) <= 1e-7 and np.quantile(err[same], 0.9) <= 1e-7
    assert np.median(info["maxde"]) <= 1.5 * np.median(ref["maxde"]) + 1e-13
    assert np.max(info["maxde"]) < 1e-5
    # the one-lane kernel on the same batch: close, but not the same bits
    assert groups.agx_set_orbit_groups(0, -1, 0) == 0
    one = {}
    agama_amd.orbit(potential=gpu, ic=ic, time=T, accuracy=1e-8, info=one)
    d = np.linalg.norm(info["final"][:, :3] - one["final"][:, :3], axis=1) / np.linalg.norm(one["final"][:, :3], axis=1)
    print(f"group vs one-lane: median {np.median(d):.2e}, identical {np.mean(d == 0):.3f}")
    assert not np.array_equal(info["final"], one["final"])
    assert np.median(d) <= 1e-7


def test_mixed_batch_and_sampling(setup, groups):
    import agama_amd
    gpu, orc = setup
    ic, T = ics(setup, 4096, seed=26, periods=3, uneven=True)
    T = T / 10
    print("case: automatic stride from 0 orbits up, 4096 uneven orbits", flush=True)
    assert groups.agx_set_orbit_groups(-1, -1, 0) == 0
    info = {}
    agama_amd.orbit(potential=gpu, ic=ic, time=T, info=info)
    ref = orc.orbits(ic, T)
    assert np.all(info["status"] == 0) and np.all(info["nsteps"] > 0)
    err = np.linalg.norm(info["final"][:, :3] - ref["final"][:, :3], axis=1) / np.linalg.norm(ref["final"][:, :3], axis=1)
    print(f"steps identical {np.mean(info['nsteps'] == ref['nsteps']):.3f}, pos err median {np.median(err):.2e}")
    assert np.mean(info["nsteps"] == ref["nsteps"]) > 0.9 and np.median(err) <= 1e-9
    # trajectory rows in a rotating frame written by group lanes
    print("case: group mode with trajsize=11, Omega=0.3, 64 orbits", flush=True)
    assert groups.agx_set_orbit_groups(2, 32, 0) == 0
    ic2, T2 = ics(setup, 64, seed=24, periods=10)
    t, x = agama_amd.orbit(potential=gpu, ic=ic2, time=T2, trajsize=11, Omega=0.3, timestart=1.5,
                           dtype="float64", separateTime=True)
    r = orc.orbits(ic2, T2, trajsize=11, Omega=0.3, tstart=1.5)
    rx = r["traj"].reshape(64, 11, 6)
    assert np.max(np.abs(t - r["time"].reshape(64, 11))) <= 1e-14 * np.max(np.abs(t))
    e = np.linalg.norm(x[:, -1, :3] - rx[:, -1, :3], axis=1) / np.linalg.norm(rx[:, -1, :3], axis=1)
    assert np.median(e) < 1e-7 and np.allclose(x[:, 0], ic2, rtol=1e-14, atol=1e-15)
    with pytest.raises(Exception):
        from agama_amd._native import check
        check(groups.agx_set_orbit_groups(-2, -1, 0), "groups")


def test_stride_one_runs_every_orbit(setup, groups):
    """stride 1 makes every wavefront a group wavefront; a group-orbit count below the batch
    size must not leave the other orbits unintegrated (they go to the groups as well)"""
    import agama_amd
    gpu, orc = setup
    ic, T = ics(setup, 4096, seed=27, periods=2)
    print("case: agx_set_orbit_groups(1, 64, 0) on 4096 orbits x 2 periods", flush=True)
    assert groups.agx_set_orbit_groups(1, 64, 0) == 0
    info = {}
    agama_amd.orbit(potential=gpu, ic=ic, time=T, accuracy=1e-8, info=info)
    ref = orc.orbits(ic, T, trajsize=1, accuracy=1e-8)
    assert np.all(info["status"] == 0) and np.all(info["nsteps"] > 0)
    assert np.all(np.isfinite(info["final"]))
    err = np.linalg.norm(info["final"][:, :3] - ref["final"][:, :3], axis=1) / np.linalg.norm(ref["final"][:, :3], axis=1)
    print(f"position rel err median {np.median(err):.2e}, max {np.max(err):.2e}; "
          f"identical step counts {np.mean(info['nsteps'] == ref['nsteps']):.3f}")
    assert np.median(err) <= 1e-9 and np.quantile(err, 0.99) <= 1e-7


def test_record_second_pass_with_group_defaults(setup, groups):
    """every-step recording whose second pass integrates orbits again: with group mode allowed
    for every batch size (min_batch 0) both passes still run the one-lane kernel, so the orbits
    integrated again take the same steps and their rows equal a one-pass recording bit for bit"""
    import agama_amd
    from agama_amd._native import lib
    gpu, orc = setup
    ic, T = ics(setup, 48, seed=43, periods=3)
    print("case: trajsize=0, 48 orbits, 512-row budget, agx_set_orbit_groups(-1, -1, 0)", flush=True)
    t1, x1 = agama_amd.orbit(potential=gpu, ic=ic, time=T, trajsize=0, dtype="float64", separateTime=True)
    assert groups.agx_set_orbit_groups(-1, -1, 0) == 0
    assert lib.agx_set_record_rows(16 * 32) == 0
    try:
        t2, x2 = agama_amd.orbit(potential=gpu, ic=ic, time=T, trajsize=0, dtype="float64", separateTime=True)
    finally:
        lib.agx_set_record_rows(0)
    for i in range(48):
        assert np.array_equal(t1[i], t2[i]) and np.array_equal(x1[i], x2[i]), i
    ref = orc.orbits_steps(ic, T)
    rows = np.array([len(a) for a in t2])
    print(f"identical step counts with the oracle {np.mean(rows == ref['rows']):.3f}")
    assert np.mean(rows == ref["rows"]) >= 0.9


def test_dprkn8_group_mode(setup):
    """DPRKN8 with group mode (per-cell tables): every orbit in groups agrees with the oracle's
    DPRKN8 at the one-lane kernel's bars and differs from the one-lane results only by rounding;
    a mixed batch (one wavefront in 8 in group mode) completes with the oracle's step counts"""
    import ctypes
    import agama_amd
    from agama_amd._native import lib
    gpu, orc = setup
    old = (ctypes.c_int(), ctypes.c_int())
    lib.agx_get_orbit_config(None, ctypes.byref(old[0]), ctypes.byref(old[1]), None)
    assert lib.agx_set_fetch(old[0].value, 3) == 0
    try:
        ic, T = ics(setup, 256, seed=29, periods=20)
        print("case: DPRKN8, group mode for every orbit, 256 orbits x 20 periods", flush=True)
        assert lib.agx_set_orbit_groups(1, 256, 0) == 0
        info = {}
        agama_amd.orbit(potential=gpu, ic=ic, time=T, accuracy=1e-8, method="dprkn8", info=info)
        ref = orc.orbits(ic, T, trajsize=1, accuracy=1e-8, method="dprkn8")
        assert np.all(info["status"] == 0)
        err = np.linalg.norm(info["final"][:, :3] - ref["final"][:, :3], axis=1) / \
            np.linalg.norm(ref["final"][:, :3], axis=1)
        same = info["nsteps"] == ref["nsteps"]
        print(f"position rel err median {np.median(err):.2e}; identical step counts {same.mean():.3f}")
        assert np.median(err) <= 1e-8 and np.quantile(err[same], 0.9) <= 1e-7
        assert np.max(info["maxde"]) < 1e-5
        assert lib.agx_set_orbit_groups(0, -1, 0) == 0
        one = {}
        agama_amd.orbit(potential=gpu, ic=ic, time=T, accuracy=1e-8, method="dprkn8", info=one)
        d = np.linalg.norm(info["final"][:, :3] - one["final"][:, :3], axis=1) / np.linalg.norm(one["final"][:, :3], axis=1)
        print(f"group vs one-lane: median {np.median(d):.2e}")
        assert not np.array_equal(info["final"], one["final"]) and np.median(d) <= 1e-8
        ic2, T2 = ics(setup, 4096, seed=30, periods=3, uneven=True)
        T2 = T2 / 10
        print("case: DPRKN8, one wavefront in 8 in group mode, 4096 uneven orbits", flush=True)
        assert lib.agx_set_orbit_groups(8, -1, 0) == 0
        info2 = {}
        agama_amd.orbit(potential=gpu, ic=ic2, time=T2, method="dprkn8", info=info2)
        ref2 = orc.orbits(ic2, T2, method="dprkn8")
        assert np.all(info2["status"] == 0) and np.all(info2["nsteps"] > 0)
        err2 = np.linalg.norm(info2["final"][:, :3] - ref2["final"][:, :3], axis=1) / \
            np.linalg.norm(ref2["final"][:, :3], axis=1)
        print(f"steps identical {np.mean(info2['nsteps'] == ref2['nsteps']):.3f}, pos err median {np.median(err2):.2e}")
        assert np.mean(info2["nsteps"] == ref2["nsteps"]) > 0.9 and np.median(err2) <= 1e-9
    finally:
        assert lib.agx_set_orbit_groups(-1, -1, 65536) == 0
        assert lib.agx_set_fetch(old[0].value, old[1].value) == 0


@pytest.mark.parametrize("method", ["dop853", "dprkn8"])
def test_group_mode_in_the_power_law_regions(setup, method):
    """group mode on orbits that stay beyond r_max (outer power-law extrapolation, split over the
    group's lanes) or inside r_min (inner extrapolation): the same orbits as the oracle's"""
    import ctypes
    import agama_amd
    from agama_amd._native import lib
    gpu, orc = setup
    rng = np.random.default_rng(33)
    cases = []
    for lo, hi in ((1500.0, 6000.0), (1e-4, 6e-4)):
        n = 64
        r = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
        u = rng.normal(size=(n, 3))
        u /= np.linalg.norm(u, axis=1)[:, None]
        pos = u * r[:, None]
        grad = orc.eval(pos)[1]
        vc = np.sqrt(np.abs(np.sum(grad * pos, axis=1)))       # sqrt(r |dPhi/dr|)
        w = np.cross(u, rng.normal(size=(n, 3)))
        w /= np.linalg.norm(w, axis=1)[:, None]
        vel = w * (vc * rng.uniform(0.9, 1.0, n))[:, None]      # near-circular: stays in its region
        cases.append((np.hstack([pos, vel]), 3 * 2 * np.pi * r / vc))
    old = (ctypes.c_int(), ctypes.c_int())
    lib.agx_get_orbit_config(None, ctypes.byref(old[0]), ctypes.byref(old[1]), None)
    assert lib.agx_set_fetch(old[0].value, 3) == 0
    try:
        assert lib.agx_set_orbit_groups(1, -1, 0) == 0          # every orbit in groups
        for (ic, T), region in zip(cases, ("beyond r_max", "inside r_min")):
            print(f"case: {method}, group mode, 64 orbits {region}", flush=True)
            info = {}
            agama_amd.orbit(potential=gpu, ic=ic, time=T, accuracy=1e-8, method=method, info=info)
            ref = orc.orbits(ic, T, trajsize=1, accuracy=1e-8, method=method)
            assert np.all(info["status"] == 0)
            rf = np.linalg.norm(ref["final"][:, :3], axis=1)
            err = np.linalg.norm(info["final"][:, :3] - ref["final"][:, :3], axis=1) / rf
            print(f"  final radii {rf.min():.3g}..{rf.max():.3g}; position rel err median {np.median(err):.2e}, "
                  f"max {np.max(err):.2e}; identical step counts {np.mean(info['nsteps'] == ref['nsteps']):.3f}")
            if region == "beyond r_max":
                assert rf.min() > 1000
            else:
                assert rf.max() < 1e-3
            assert np.median(err) <= 1e-9 and np.max(err) <= 1e-6
    finally:
        assert lib.agx_set_orbit_groups(-1, -1, 65536) == 0
        assert lib.agx_set_fetch(old[0].value, old[1].value) == 0
