"""GPU checks of the MLP surrogate against its CPU oracle (oracle/mlp_oracle.py):
the training GEMM in every operand layout, the fused MFMA forward + input-gradient backward,
one training step's loss and parameter gradient, the potential/force assembly, and a short
training run on the triaxial Dehnen Multipole judged by the FP64 Multipole.
Tolerances: the oracle emulates the device's bf16 roundings in float64, so the remaining
differences come from fp32 accumulation order (and the rare bf16 rounding it flips)."""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

COEF = os.path.join(GOLDEN, "dehnen_triax_lmax8.coef")


def bf16_bits(a):
    from oracle import mlp_oracle as M
    r = M.bf16_round(a).astype(np.float32)
    return (r.view(np.uint32) >> 16).astype(np.uint16)


def bits_to_f(b):
    return (b.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def pot():
    import agama_amd
    return agama_amd.Potential(file=COEF)


def random_net(seed):
    from oracle import mlp_oracle as M
    th = M.unpack(M.init_params(seed))
    g = np.random.Generator(np.random.Philox(seed + 100))
    for k in ("b1", "b2", "b3", "b4", "b5"):
        th[k] = g.normal(size=th[k].shape) * 0.1
    return M.pack(th)


@pytest.mark.parametrize("ak,bk,splits", [(1, 1, 1), (1, 0, 1), (0, 1, 1), (0, 0, 1), (0, 0, 4)])
def test_gemm_layouts(ak, bk, splits):
    from agama_amd.surrogate import debug_gemm
    g = np.random.Generator(np.random.Philox(11))
    M_, N_, K_ = 128, 192, 512
    A = g.normal(size=(M_, K_)); B = g.normal(size=(K_, N_))
    Ab, Bb = bf16_bits(A), bf16_bits(B)
    ref = bits_to_f(Ab) @ bits_to_f(Bb)
    C = debug_gemm(Ab if ak else Ab.T.copy(), Bb.T.copy() if bk else Bb, ak, bk, splits)
    err = np.abs(C - ref).max() / np.abs(ref).max()
    assert err < 1e-5, err


def test_network_matches_oracle(pot):
    import agama_amd
    from oracle import mlp_oracle as M
    th = random_net(3)
    sur = agama_amd.Surrogate.from_params(pot, th, 2, np.log(100), 1e-3, 1e-2, 1e2)
    g = np.random.Generator(np.random.Philox(12))
    n = 1000
    feat = np.empty((n, 4), dtype=np.float32)
    d = np.abs(g.normal(size=(n, 3))); d /= np.linalg.norm(d, axis=1)[:, None]
    feat[:, :3] = d; feat[:, 3] = g.uniform(-1, 1, n)
    out, dfeat = sur.network(feat)
    o_ref, cache = M.forward(th.astype(np.float64), feat.astype(np.float64), bf16=True)
    d_ref, _ = M.backward(cache, np.ones(n), want_params=False)
    eo = np.abs(out - o_ref) / np.sqrt(np.mean(o_ref ** 2))
    ed = np.abs(dfeat - d_ref) / np.sqrt(np.mean(d_ref ** 2))
    print(f"network vs oracle: out err median {np.median(eo):.2e} max {eo.max():.2e}; "
          f"dout/dfeat err median {np.median(ed):.2e} max {ed.max():.2e}")
    # a relative perturbation of 3e-7 in the pre-activations (fp32 accumulation order) flips
    # a bf16 rounding now and then, which moves single outputs by up to ~1e-2 of the RMS in the
    # oracle itself; a layout or indexing error moves every output by O(1)
    assert np.median(eo) < 1e-5 and eo.max() < 3e-2
    assert np.median(ed) < 1e-5 and ed.max() < 3e-2


def test_training_gradient_matches_oracle(pot):
    import agama_amd
    from oracle import mlp_oracle as M
    th = random_net(4)
    sur = agama_amd.Surrogate.from_params(pot, th, 2, np.log(100), 1e-3, 1e-2, 1e2)
    g = np.random.Generator(np.random.Philox(13))
    n = 1024
    feat = np.empty((n, 4), dtype=np.float32)
    d = np.abs(g.normal(size=(n, 3))); d /= np.linalg.norm(d, axis=1)[:, None]
    feat[:, :3] = d; feat[:, 3] = g.uniform(-1, 1, n)
    target = g.normal(size=n).astype(np.float32)
    loss, grad = sur.train_grad(feat, target)
    lref, gref = M.loss_and_grad(th.astype(np.float64), feat.astype(np.float64), target.astype(np.float64), bf16=True)
    assert abs(loss / lref - 1) < 1e-4
    o = 0
    for name, shp in M.SHAPES:
        k = int(np.prod(shp))
        a, b = grad[o:o + k], gref[o:o + k]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"grad {name}: rel err {rel:.2e}")
        assert rel < 2e-3, name
        o += k


def test_surrogate_assembly_matches_oracle(pot):
    """Phi_NN, grad Phi_NN from the GPU vs the oracle's assembly (features, chain rule and
    scaling) with the FP64 l<=2 base taken from a zero network on the GPU"""
    import agama_amd
    from agama_amd import workloads
    from oracle import mlp_oracle as M
    th = random_net(5)
    zero = np.zeros_like(th)
    lns, k = np.log(100), 2e-3
    sur = agama_amd.Surrogate.from_params(pot, th, 2, lns, k, 1e-2, 1e2)
    base = agama_amd.Surrogate.from_params(pot, zero, 2, lns, k, 1e-2, 1e2)
    pts = workloads.density_points(2000, seed=14)
    pb, gb = base._eval(pts, True, True)
    phi, grad = sur._eval(pts, True, True)
    pref, gref = M.surrogate_eval(th.astype(np.float64), pts, pb, gb, lns, k, bf16=True)
    ep = np.abs(phi / pref - 1)
    eg = np.linalg.norm(grad - gref, axis=1) / np.linalg.norm(gref, axis=1)
    print(f"assembly vs oracle: phi err max {ep.max():.2e}; grad err median {np.median(eg):.2e} max {eg.max():.2e}")
    # medians at rounding level; the maxima are single bf16 rounding flips scaled by out_scale
    assert np.median(ep) < 1e-8 and ep.max() < 2e-5
    assert np.median(eg) < 1e-7 and eg.max() < 2e-4
    # the l <= 2 base against the full expansion: close but not equal (l >= 4 terms)
    ophi = pot._eval(pts, True, False)[0]
    assert np.max(np.abs(pb / ophi - 1)) < 0.02


def test_short_training_run(pot, tmp_path):
    """a short GPU training run: the loss falls and the surrogate beats its l<=2 base"""
    import agama_amd
    from agama_amd import workloads
    sur = agama_amd.Surrogate.train(pot, samples=2e6, steps=1500, batch=16384, seed=2)
    hist = sur.loss_history()
    assert len(hist) == 1500 and np.all(np.isfinite(hist))
    first, last = hist[:20].mean(), hist[-50:].mean()
    pts = workloads.density_points(100_000, seed=15)
    r = np.linalg.norm(pts, axis=1)
    pts = pts[(r > 0.02) & (r < 50)]
    ophi, og = pot._eval(pts, True, True)
    phi, g = sur._eval(pts, True, True)
    theta, meta = sur.params()
    base = agama_amd.Surrogate.from_params(pot, np.zeros_like(theta), int(meta[0]), meta[1], meta[2], meta[3], meta[4])
    pb, gb = base._eval(pts, True, True)
    ef = np.linalg.norm(g - og, axis=1) / np.linalg.norm(og, axis=1)
    efb = np.linalg.norm(gb - og, axis=1) / np.linalg.norm(og, axis=1)
    ep = np.abs(phi / ophi - 1)
    print(f"training loss {first:.3e} -> {last:.3e}; out_scale {meta[2]:.3e}; force err median "
          f"{np.median(ef):.2e} (base alone {np.median(efb):.2e}), 99% {np.quantile(ef, .99):.2e}; "
          f"potential err median {np.median(ep):.2e}")
    assert last < 0.05 * first
    assert np.median(ef) < 0.2 * np.median(efb)
    path = str(tmp_path / "s.npz")
    sur.save(path)
    sur2 = agama_amd.Surrogate.load(pot, path)
    phi2, g2 = sur2._eval(pts[:1000], True, True)
    assert np.array_equal(phi2, phi[:1000]) and np.array_equal(g2, g[:1000])


def test_higher_base_order(pot):
    """a base expansion of higher order (l <= 4 instead of 2) leaves the network a smaller residual
    to learn: the same short training gives a smaller potential and force error, and the saved
    parameters carry the base order"""
    import agama_amd
    from agama_amd import workloads
    pts = workloads.density_points(100_000, seed=16)
    r = np.linalg.norm(pts, axis=1)
    pts = pts[(r > 0.02) & (r < 50)]
    ophi, og = pot._eval(pts, True, True)
    err = {}
    for b in (2, 4):
        sur = agama_amd.Surrogate.train(pot, samples=2e6, steps=1500, batch=16384, seed=2, baselmax=b)
        phi, g = sur._eval(pts, True, True)
        err[b] = (np.median(np.abs(phi / ophi - 1)), np.median(np.linalg.norm(g - og, axis=1) / np.linalg.norm(og, axis=1)))
        assert int(sur.params()[1][0]) == b
    print("median potential / force error by base order:", err)
    assert err[4][0] < 0.5 * err[2][0] and err[4][1] < 0.5 * err[2][1]


@pytest.fixture(scope="module")
def trained(pot):
    import agama_amd
    return agama_amd.Surrogate.train(pot, samples=4e6, steps=3000, batch=16384, seed=3)


def test_surrogate_orbits_and_fallback(pot, trained):
    """every orbit either keeps its FP64-judged energy error within max_rel_de with the
    surrogate force, or is integrated again with the FP64 Multipole - and then equals the
    plain FP64 integration exactly"""
    import agama_amd
    from agama_amd import workloads
    from oracle.pyoracle import OraclePotential
    orc = OraclePotential(COEF)
    ic, T = workloads.triaxial_dehnen_orbits(512, lambda p: orc.eval(p)[0], pot.coefs(), seed=31, periods=10)
    ic[0, :3] = [150.0, 0.0, 0.0]          # outside the training radii: must fall back
    info_s, info_p = {}, {}
    agama_amd.orbit(potential=trained, ic=ic, time=T, accuracy=1e-8, info=info_s, max_rel_de=1e-6)
    agama_amd.orbit(potential=pot, ic=ic, time=T, accuracy=1e-8, info=info_p)
    st = info_s["status"]
    fb = (st & 8) != 0
    assert fb[0]
    assert info_s["n_fallback"] == fb.sum()
    assert np.all((st & ~8) == 0) and np.all(info_p["status"] == 0)
    assert np.array_equal(info_s["final"][fb], info_p["final"][fb])
    assert np.array_equal(info_s["nsteps"][fb], info_p["nsteps"][fb])
    ok = ~fb
    assert np.all(info_s["maxde"][ok] <= 1e-6)
    pos = np.linalg.norm(info_s["final"][ok, :3] - info_p["final"][ok, :3], axis=1) / \
        np.linalg.norm(info_p["final"][ok, :3], axis=1)
    steps_ratio = info_s["nsteps"][ok].sum() / max(info_p["nsteps"][ok].sum(), 1)
    print(f"surrogate orbits: {ok.sum()} of {len(ok)} kept the surrogate (dE/E <= 1e-6 over 10 "
          f"periods); their final position vs FP64 median {np.median(pos) if ok.any() else np.nan:.2e}; "
          f"steps surrogate/FP64 {steps_ratio:.3f}; max dE/E surrogate {info_s['maxde'][ok].max() if ok.any() else np.nan:.2e}")


def test_surrogate_orbits_loose_threshold(pot, trained):
    """with a loose energy threshold the surrogate integrates in-range orbits itself until the
    step size collapses (the bf16 rounding in the network makes its force rough on small
    scales, which the error control at accuracy 1e-8 answers with ever shorter steps); the
    orbits that keep the surrogate, their step counts and their FP64-judged energy error are
    reported, and every result is either a surrogate orbit or the exact FP64 integration"""
    import agama_amd
    from agama_amd import workloads
    from oracle.pyoracle import OraclePotential
    orc = OraclePotential(COEF)
    ic, T = workloads.triaxial_dehnen_orbits(256, lambda p: orc.eval(p)[0], pot.coefs(), seed=32, periods=2)
    info, ref = {}, {}
    agama_amd.orbit(potential=trained, ic=ic, time=T, accuracy=1e-8, info=info, max_rel_de=1.0)
    agama_amd.orbit(potential=pot, ic=ic, time=T, accuracy=1e-8, info=ref)
    st = info["status"]
    fb = (st & 8) != 0
    assert np.all((st & ~8) == 0)
    assert np.array_equal(info["final"][fb], ref["final"][fb])
    ok = ~fb
    msg = f"loose threshold, 2 periods: {ok.sum()} of {len(ok)} orbits kept the surrogate"
    if ok.any():
        mde = info["maxde"][ok]
        msg += (f"; their steps surrogate/FP64 {info['nsteps'][ok].sum() / ref['nsteps'][ok].sum():.2f}, "
                f"dE/E median {np.median(mde):.2e} max {mde.max():.2e}")
        assert np.median(mde) < 1e-2
    print(msg)


def test_pilot_routes_a_batch_the_surrogate_cannot_hold(pot, trained):
    """8192 orbits at max_rel_de=1e-6: the pilot (every 64th orbit, 200 steps) finds that the
    surrogate cannot keep the energy error, so every orbit is integrated with the FP64
    Multipole directly - marked as a fallback and equal to the plain FP64 integration"""
    import agama_amd
    from agama_amd import workloads
    from oracle.pyoracle import OraclePotential
    orc = OraclePotential(COEF)
    ic, T = workloads.triaxial_dehnen_orbits(8192, lambda p: orc.eval(p)[0], pot.coefs(), seed=33, periods=2)
    info_s, info_p = {}, {}
    agama_amd.orbit(potential=trained, ic=ic, time=T, accuracy=1e-8, info=info_s, max_rel_de=1e-6)
    agama_amd.orbit(potential=pot, ic=ic, time=T, accuracy=1e-8, info=info_p)
    np_, pf = info_s["pilot"]
    print(f"pilot: {pf} of {np_} orbits fell back; batch fallbacks {info_s['n_fallback']}")
    assert np_ == 1024 and 2 * pf >= np_
    assert info_s["n_fallback"] == 8192 and np.all(info_s["status"] == 8)
    assert np.array_equal(info_s["final"], info_p["final"])
    assert np.array_equal(info_s["nsteps"], info_p["nsteps"])
