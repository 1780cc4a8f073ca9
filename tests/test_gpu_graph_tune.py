"""HIP-graph capture of an orbit batch and the launch-configuration autotuner.
The graph is replayed several times with other GPU work and synchronisations in between, and
every replay must reproduce the eager integration exactly; every autotuner setting must give
the same orbits."""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

COEF = os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")


@pytest.fixture(scope="module")
def pot():
    import agama_amd
    return agama_amd.Potential(file=COEF)


@pytest.fixture(params=["static", "probe+groups"])
def schedule(request):
    """the batch's work order: the static cost key, or the probe with group mode at one wavefront
    in 8 (the thresholds lowered to this batch; the probe's order is dealt over the groups)"""
    from agama_amd import lib, tune
    saved = tune.current_schedule()
    if request.param == "probe+groups":
        assert lib.agx_set_orbit_probe(64, 2048) == 0
        assert lib.agx_set_orbit_groups(8, -1, 2048) == 0
    yield request.param
    tune.apply(saved)


def test_orbit_graph_replays_match_eager(pot, schedule):
    import torch
    import agama_amd
    from agama_amd import lib, workloads
    print("schedule:", schedule, flush=True)
    n = 4096
    ic, T = workloads.triaxial_dehnen_orbits(n, lambda p: pot.potential(p), pot.coefs(), seed=51, uneven=True)
    T = T / 20
    dev = torch.device("cuda:0")
    d_ic = torch.from_numpy(ic).to(dev)
    d_T = torch.from_numpy(T).to(dev)
    d_sz = torch.ones(n, dtype=torch.int32, device=dev)
    d_off = torch.arange(n, dtype=torch.int64, device=dev)

    def outputs():
        return dict(fin=torch.empty((n, 6), dtype=torch.float64, device=dev),
                    ns=torch.empty(n, dtype=torch.int64, device=dev),
                    de=torch.empty(n, dtype=torch.float64, device=dev),
                    st=torch.empty(n, dtype=torch.int32, device=dev))
    eager, gout = outputs(), outputs()
    stream = torch.cuda.Stream()
    rc = lib.agx_orbit_batch_device(pot.handle, n, d_ic.data_ptr(), None, d_T.data_ptr(), d_sz.data_ptr(),
                                    d_off.data_ptr(), None, 1e-8, 0.0, int(1e8), None, None,
                                    eager["fin"].data_ptr(), eager["ns"].data_ptr(), None, None,
                                    eager["de"].data_ptr(), eager["st"].data_ptr(), 0, stream.cuda_stream)
    assert rc == 0
    torch.cuda.synchronize()
    g = lib.agx_orbit_graph_create(pot.handle, n, d_ic.data_ptr(), None, d_T.data_ptr(), d_sz.data_ptr(),
                                   d_off.data_ptr(), 1e-8, 0.0, int(1e8), None, None, gout["fin"].data_ptr(),
                                   gout["ns"].data_ptr(), None, None, gout["de"].data_ptr(),
                                   gout["st"].data_ptr(), 0)
    assert g, agama_amd._native.last_error()
    try:
        pts = np.random.default_rng(1).normal(size=(1000, 3))
        for rep in range(3):
            for v in gout.values():
                v.fill_(-1)
            torch.cuda.synchronize()
            assert lib.agx_orbit_graph_launch(g, stream.cuda_stream) == 0
            torch.cuda.synchronize()
            for k in eager:
                assert torch.equal(eager[k], gout[k]), (rep, k)
            pot.potential(pts)          # other GPU work between replays
    finally:
        lib.agx_orbit_graph_destroy(g)


def test_autotuner_settings_agree_and_restore(pot):
    from agama_amd import tune
    before, before_sched = tune.current(), tune.current_schedule()
    cfg = tune.tune(pot, n=4096, repeats=1, groups=(-1, 0, 8), probes=(64, 0), verbose=True)
    assert tune.current() == before and tune.current_schedule() == before_sched
    # stage 1: 4 fetch modes x 2 wave counts; stage 2: 3 x 2 schedules but the current one
    assert len(cfg["candidates"]) == 8 + 5
    assert cfg["fetch_orbit"] in tune.FETCH and cfg["waves"] in (1, 2)
    assert cfg["group_stride"] in (-1, 0, 8) and cfg["probe_steps"] in (64, 0)
    print("tuned:", {k: cfg[k] for k in ("fetch_orbit", "waves", "group_stride", "probe_steps")},
          "identical:", all(c["identical_to_first"] for c in cfg["candidates"]))
    applied = tune.apply(cfg)
    assert applied["fetch_orbit"] == cfg["fetch_orbit"] and applied["waves"] == cfg["waves"]
    assert applied["group_stride"] == cfg["group_stride"] and applied["probe_steps"] == cfg["probe_steps"]
    assert applied["group_min_batch"] == applied["probe_min_batch"] == 4096
    tune.apply(dict(before, **before_sched))
    assert tune.current_schedule() == before_sched
