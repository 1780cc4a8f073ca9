"""GPU parity of the batched Multipole evaluation against the CPU oracle.
Tolerance: 1e-12 relative for potential and gradient (FP64 kernel); the FP32-table form is
checked against its own stated error bound."""
import ctypes
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pots():
    import agama_amd
    from oracle.pyoracle import OraclePotential
    out = {}
    for name in ("dehnen_triax_lmax8.coef", "disk_bulge_halo_lmax12.coef"):
        path = os.path.join(GOLDEN, name)
        out[name] = (agama_amd.Potential(file=path), OraclePotential(path))
    return out


def rel_errors(gpu, orc, pts):
    phi, grad = gpu._eval(pts, True, True)
    ophi, ograd = orc.eval(pts)
    ephi = np.abs(phi - ophi) / np.abs(ophi)
    egrad = np.linalg.norm(grad - ograd, axis=1) / np.maximum(np.linalg.norm(ograd, axis=1), 1e-300)
    return ephi, egrad, phi, grad


def test_eval_1m_points_from_density(pots):
    from agama_amd import workloads
    gpu, orc = pots["dehnen_triax_lmax8.coef"]
    pts = workloads.density_points(1_000_000, seed=11)
    ephi, egrad, _, _ = rel_errors(gpu, orc, pts)
    print(f"1M points: max rel err phi {ephi.max():.3e} grad {egrad.max():.3e}")
    assert ephi.max() <= 1e-12 and egrad.max() <= 1e-12


def test_eval_edge_points(pots):
    """axis (R=0), equatorial plane, exact grid radii, both extrapolation branches, origin"""
    gpu, orc = pots["dehnen_triax_lmax8.coef"]
    radii = orc.tables(0)
    rg = np.exp(radii)
    pts = [[0, 0, 1.0], [0, 0, -2.0], [1.0, 0, 0], [0, 1.5, 0], [1e-4, 2e-4, -3e-4], [2e3, 10, 5],
           [1e5, -3e4, 2e4], [0, 0, 5e-4], [0.9999e-3, 0, 0], [1.0001e3, 0, 0]]
    pts += [[r, 0, 0] for r in rg[1:-1]] + [[0, 0, r] for r in rg[1:-1]] + [[r / 2, r / 2, r / np.sqrt(2)] for r in rg[2:-2]]
    pts = np.array(pts, dtype=np.float64)
    ephi, egrad, phi, grad = rel_errors(gpu, orc, pts)
    assert np.all(np.isfinite(phi))
    assert ephi.max() <= 1e-12 and egrad.max() <= 1e-12, (ephi.argmax(), egrad.argmax())


def test_eval_symmetry_and_origin(pots):
    gpu, orc = pots["dehnen_triax_lmax8.coef"]
    rng = np.random.default_rng(3)
    p = rng.normal(size=(1000, 3))
    ph = [gpu.potential(p * s) for s in ([1, 1, 1], [-1, 1, 1], [1, -1, 1], [1, 1, -1], [-1, -1, -1])]
    for q in ph[1:]:
        assert np.max(np.abs(q / ph[0] - 1)) < 1e-12
    p0 = gpu.potential([0.0, 0.0, 0.0])
    assert np.isfinite(p0) and abs(p0 / orc.eval(np.zeros((1, 3)))[0][0] - 1) < 1e-12


def test_eval_axisymmetric_lmax12(pots):
    gpu, orc = pots["disk_bulge_halo_lmax12.coef"]
    rng = np.random.default_rng(5)
    R = 10 ** rng.uniform(-1.5, 2.5, 200_000)
    z = rng.normal(size=R.size) * 0.3 * (1 + R / 5)
    ph = rng.uniform(0, 2 * np.pi, R.size)
    pts = np.stack([R * np.cos(ph), R * np.sin(ph), z], 1)
    ephi, egrad, _, _ = rel_errors(gpu, orc, pts)
    assert ephi.max() <= 1e-12 and egrad.max() <= 1e-12


def test_fp32_tables_error_bound(pots):
    """tables rounded to float: force error vs FP64 oracle stays below 1e-5 (median 1e-7)"""
    import agama_amd
    from agama_amd import workloads
    gpu, orc = pots["dehnen_triax_lmax8.coef"]
    pts = workloads.density_points(200_000, seed=12)
    phi32, g32 = gpu._eval(pts, True, True, precision=agama_amd.PRECISION_FP32)
    ophi, og = orc.eval(pts)
    e = np.linalg.norm(g32 - og, axis=1) / np.linalg.norm(og, axis=1)
    print(f"fp32 tables: force err median {np.median(e):.2e} max {e.max():.2e}")
    assert np.median(e) < 1e-6 and e.max() < 1e-4
    assert np.max(np.abs(phi32 / ophi - 1)) < 1e-5


def test_public_api_shapes_and_c_abi(pots):
    import ctypes
    from agama_amd import lib
    gpu, orc = pots["dehnen_triax_lmax8.coef"]
    p = np.array([0.3, -0.2, 0.5])
    assert np.shape(gpu.potential(p)) == () and gpu.force(p).shape == (3,)
    pot, acc = gpu.eval(np.tile(p, (4, 1)), pot=True, acc=True)
    assert pot.shape == (4,) and acc.shape == (4, 3)
    d = np.empty(3)
    h = lib.agama_createPotential(("type=Multipole, file=" + os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")).encode())
    assert h
    val = lib.agama_evalPotential(h, p.ctypes.data, 0.0, d.ctypes.data, None)
    lib.agama_deletePotential(h)
    ophi, og = orc.eval(p[None])
    assert abs(val / ophi[0] - 1) < 1e-12 and np.max(np.abs(d - og[0])) < 1e-12 * np.linalg.norm(og)
    assert not lib.agama_createPotential(b"type=Multipole, file=/nonexistent")
    assert b"cannot open" in lib.agama_getError()


def test_host_entry_points_share_a_handle_across_threads(pots):
    """agx_eval_many / agx_eval_hess_many with host buffers stage through the handle's own stream
    and a buffer kept between calls: calls of growing and shrinking sizes from four threads at
    once on one handle give the device entry points' results bit for bit, and the calling
    thread's current device is left as it was"""
    import threading
    import torch
    from agama_amd import lib, workloads
    gpu, _ = pots["dehnen_triax_lmax8.coef"]
    pts = workloads.density_points(300_000, seed=14)
    dev = torch.device("cuda:0")
    d_pts = torch.from_numpy(pts).to(dev)
    d_phi = torch.empty(len(pts), dtype=torch.float64, device=dev)
    d_g = torch.empty((len(pts), 3), dtype=torch.float64, device=dev)
    d_h = torch.empty((len(pts), 6), dtype=torch.float64, device=dev)
    assert lib.agx_eval_hess_many_device(gpu.handle, len(pts), d_pts.data_ptr(), 0.0, d_phi.data_ptr(),
                                         d_g.data_ptr(), d_h.data_ptr(), 0, None) == 0
    torch.cuda.synchronize()
    ref_phi, ref_g, ref_h = d_phi.cpu().numpy(), d_g.cpu().numpy(), d_h.cpu().numpy()
    assert lib.agx_eval_many_device(gpu.handle, len(pts), d_pts.data_ptr(), 0.0, d_phi.data_ptr(), d_g.data_ptr(),
                                    0, None) == 0
    torch.cuda.synchronize()
    ref_phi1, ref_g1 = d_phi.cpu().numpy(), d_g.cpu().numpy()
    sizes = [1, 300_000, 17, 120_000, 5_000, 299_999, 64]
    errors = []

    def worker(k):
        try:
            for j, n in enumerate(sizes[k:] + sizes[:k]):
                lo = (k * 7919 + j * 104729) % (len(pts) - n + 1)
                phi, g, h = np.empty(n), np.empty((n, 3)), np.empty((n, 6))
                sl = slice(lo, lo + n)
                p = np.ascontiguousarray(pts[sl])
                if (k + j) % 2:
                    rc = lib.agx_eval_hess_many(gpu.handle, n, p.ctypes.data, 0.0, phi.ctypes.data, g.ctypes.data,
                                                h.ctypes.data, 0)
                    ok = np.array_equal(h, ref_h[sl]) and np.array_equal(phi, ref_phi[sl]) and \
                        np.array_equal(g, ref_g[sl])
                else:
                    rc = lib.agx_eval_many(gpu.handle, n, p.ctypes.data, 0.0, phi.ctypes.data, g.ctypes.data, 0)
                    ok = np.array_equal(phi, ref_phi1[sl]) and np.array_equal(g, ref_g1[sl])
                if rc or not ok:
                    errors.append((k, j, n, rc))
        except Exception as e:   # pragma: no cover - reported below
            errors.append(repr(e))
    print("case: host-buffer evaluations from 4 threads on one handle", flush=True)
    threads = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    assert torch.cuda.current_device() == 0


@pytest.mark.parametrize("fetch", [0, 1, 2, 3])
@pytest.mark.parametrize("name", ["dehnen_triax_lmax8.coef", "disk_bulge_halo_lmax12.coef",
                                  "ferrers_bar_lmax16.coef"])
def test_eval_every_fetch_mode(name, fetch):
    """every evaluation form of the FP64 kernel (8- and 16-byte record gathers, the cooperative
    gather, the per-cell polynomial tables) against the oracle at 200k density points and at
    the grid nodes, cell edges and both extrapolation branches; bar 1e-12 relative"""
    import agama_amd
    from agama_amd import lib, workloads
    from oracle.pyoracle import OraclePotential
    path = os.path.join(GOLDEN, name)
    gpu, orc = agama_amd.Potential(file=path), OraclePotential(path)
    pts = workloads.density_points(200_000, seed=21) * (3.0 if "ferrers" in name else 1.0)
    rg, tg = np.exp(orc.tables(0)), orc.tables(1)
    # points on the (ln r, tau) node lattice: R = r (1 - tau^2) / (1 + tau^2), z = 2 r tau / (1 + tau^2)
    rr, tt = np.meshgrid(rg[::3], tg[::2], indexing="ij")
    R, z = rr * (1 - tt ** 2) / (1 + tt ** 2), 2 * rr * tt / (1 + tt ** 2)
    phi0 = 0.7
    lat = np.stack([R.ravel() * np.cos(phi0), R.ravel() * np.sin(phi0), z.ravel()], axis=1)
    extra = np.array([[0, 0, 1.0], [1e-4, 2e-4, -3e-4], [2e3, 10, 5], [0.9999e-3, 0, 0], [1.0001e3, 0, 0]])
    # both power-law regions in every direction, including the poles: r log-uniform in
    # [rmin/100, rmin] and [rmax, 1e4 rmax]
    rng = np.random.default_rng(22)
    d = rng.normal(size=(4000, 3))
    d[:20, :2] = 0
    d /= np.linalg.norm(d, axis=1)[:, None]
    rr = np.concatenate([rg[0] * np.exp(rng.uniform(np.log(0.01), 0, 2000)),
                         rg[-1] * np.exp(rng.uniform(0, np.log(1e4), 2000))])
    pts = np.concatenate([pts, lat, extra, d * rr[:, None]])
    old = (ctypes.c_int(), ctypes.c_int())
    lib.agx_get_orbit_config(None, ctypes.byref(old[0]), ctypes.byref(old[1]), None)
    try:
        assert lib.agx_set_fetch(fetch, old[1].value) == 0
        ephi, egrad, phi, _ = rel_errors(gpu, orc, pts)
    finally:
        lib.agx_set_fetch(old[0].value, old[1].value)
    print(f"{name} fetch {fetch}: max rel err phi {ephi.max():.3e} grad {egrad.max():.3e}")
    assert np.all(np.isfinite(phi))
    assert ephi.max() <= 1e-12 and egrad.max() <= 1e-12
