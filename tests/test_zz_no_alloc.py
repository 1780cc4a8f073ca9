"""GPU test of agx_orbit_batch_device_noalloc, the orbit entry point that allocates nothing (the
longest-first schedule's workspace lives in the library's static device array): called twice
in a row through the C ABI, the second time with a smaller batch (the array then holds the
first call's leftovers), its results equal agx_orbit_batch_device's on the same inputs bit for
bit (the same kernels and the same work order).  Sizes beyond the array's limit are refused
before anything runs."""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

COEF = os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")


def integrate(entry, pot, ic, T, trajsize):
    import torch
    from agama_amd._native import last_error
    dev = torch.device("cuda:0")
    n = len(ic)
    d_ic = torch.from_numpy(ic).to(dev)
    d_T = torch.from_numpy(T).to(dev)
    d_sz = torch.full((n,), trajsize, dtype=torch.int32, device=dev)
    d_off = torch.arange(n, dtype=torch.int64, device=dev) * trajsize
    out = dict(time=torch.full((n * trajsize,), -1.0, dtype=torch.float64, device=dev),
               traj=torch.full((n * trajsize, 6), -1.0, dtype=torch.float64, device=dev),
               final=torch.full((n, 6), -1.0, dtype=torch.float64, device=dev),
               nsteps=torch.full((n,), -1, dtype=torch.int64, device=dev),
               nattempts=torch.full((n,), -1, dtype=torch.int64, device=dev),
               nevals=torch.full((n,), -1, dtype=torch.int64, device=dev),
               maxde=torch.full((n,), -1.0, dtype=torch.float64, device=dev),
               status=torch.full((n,), -1, dtype=torch.int32, device=dev))
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        rc = entry(pot.handle, n, d_ic.data_ptr(), None, d_T.data_ptr(), d_sz.data_ptr(), d_off.data_ptr(), None,
                   1e-8, 0.0, int(1e8), out["time"].data_ptr(), out["traj"].data_ptr(), out["final"].data_ptr(),
                   out["nsteps"].data_ptr(), out["nattempts"].data_ptr(), out["nevals"].data_ptr(),
                   out["maxde"].data_ptr(), out["status"].data_ptr(), 0, stream.cuda_stream)
        assert rc == 0, last_error()
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


def test_noalloc_equals_allocating_entry_point():
    import agama_amd
    from agama_amd import lib, workloads
    from oracle.pyoracle import OraclePotential
    pot = agama_amd.Potential(file=COEF)
    orc = OraclePotential(COEF)
    ic, T = workloads.triaxial_dehnen_orbits(20000, lambda p: orc.eval(p)[0], pot.coefs(), seed=61, periods=3,
                                             uneven=True)
    T = T / 20
    cases = [(ic, T, 1), (ic[:3000].copy(), T[:3000].copy(), 5)]
    got = []
    for i, (c_ic, c_T, sz) in enumerate(cases):   # two no-alloc calls in a row, the second smaller
        print(f"case: agx_orbit_batch_device_noalloc call {i + 1}, {len(c_ic)} orbits, trajsize {sz}", flush=True)
        got.append(integrate(lib.agx_orbit_batch_device_noalloc, pot, c_ic, c_T, sz))
    for i, (c_ic, c_T, sz) in enumerate(cases):
        print(f"case: agx_orbit_batch_device, {len(c_ic)} orbits, trajsize {sz}", flush=True)
        ref = integrate(lib.agx_orbit_batch_device, pot, c_ic, c_T, sz)
        assert np.all(ref["status"] == 0) and np.all(ref["nsteps"] > 0)
        for k in ref:
            assert np.array_equal(got[i][k], ref[k]), (i, k)
    # the rows of the second call are the trajectory (first row = initial condition)
    assert np.array_equal(got[1]["traj"].reshape(3000, 5, 6)[:, 0], cases[1][0])


def test_noalloc_refuses_batches_beyond_its_array():
    import agama_amd
    from agama_amd import lib
    from agama_amd._native import last_error
    pot = agama_amd.Potential(file=COEF)
    print("case: agx_orbit_batch_device_noalloc with 1048577 orbits (refused before any launch)", flush=True)
    rc = lib.agx_orbit_batch_device_noalloc(pot.handle, (1 << 20) + 1, None, None, None, None, None, None, 1e-8,
                                            0.0, int(1e8), None, None, None, None, None, None, None, None, 0, None)
    assert rc == 1    # AGX_EINVAL
    assert "at most 1048576 orbits" in last_error()
