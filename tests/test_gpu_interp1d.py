"""GPU parity of Multipole expansions with lmax <= 2, which the reference evaluates with 1-d
quintic splines in ln r per (l, m) term (MultipoleInterp1d, src/potential_multipole.cpp:
1691-1815, chosen at :1509-1511) and the lmax = 2 fast path of the synthesis (:722-772): the
spherical Plummer expansion (l = 0 only), the triaxial Dehnen lmax = 2 expansion (fast path) and
synthetic coefficients with every (l, m) term of lmax = 2 (odd l, sine terms: general path).
Bars as for the 2-d form: potential, gradient and Hessian within 1e-12 of the oracle (relative
to |Phi| and the largest component) at points inside, on and beyond the grid; orbits over 20
circular periods at accuracy 1e-8 with both integrators agree with the oracle to 1e-7 for the
median orbit, with identical step counts for at least 90% of them or for as many as a recompile of
the reference (the oracle with FMA contraction) keeps, less 0.05."""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

CASES = ["plummer_lmax4.coef", "dehnen_triax_lmax2.coef", "general_lmax2"]


def load_coefs(case):
    from agama_amd.workloads import read_coef_file
    from tests.test_oracle_multipole import synthetic_general_coefs
    return synthetic_general_coefs() if case == "general_lmax2" else read_coef_file(os.path.join(GOLDEN, case))


def load(case):
    import agama_amd
    from oracle.pyoracle import OraclePotential
    coefs = load_coefs(case)
    return agama_amd.Potential(coefs=coefs), OraclePotential(coefs=coefs), coefs


def points(radii, n, seed):
    g = np.random.Generator(np.random.Philox(seed))
    r = np.exp(g.uniform(np.log(radii[0] * 0.1), np.log(radii[-1] * 10), n))
    d = g.normal(size=(n, 3))
    pts = d / np.linalg.norm(d, axis=1)[:, None] * r[:, None]
    k = np.arange(len(radii))
    pts[:len(radii)] = np.stack([radii[k] * 0.6, radii[k] * 0.0, radii[k] * 0.8], 1)   # on the nodes
    pts[len(radii):len(radii) + 4] = [[0, 0, 1.0], [0, 0, -0.3], [1e-5, 0, 0], [0.5, 0.5, 0]]
    return pts


@pytest.mark.parametrize("case", CASES)
def test_eval_and_hessian_match_oracle(case):
    gpu, orc, (radii, _, _) = load(case)
    print(f"case: eval + Hessian, 100k points, {case}", flush=True)
    pts = points(radii, 100_000, 71)
    phi, grad = gpu._eval(pts, True, True)
    ophi, og = orc.eval(pts)
    ep = np.max(np.abs(phi / ophi - 1))
    eg = np.max(np.linalg.norm(grad - og, axis=1) / np.linalg.norm(og, axis=1))
    p2, g2, h2 = gpu.eval(pts, pot=True, acc=True, der=True)
    hp, hg, hh = orc.eval_hess(pts)
    eh = np.max(np.abs(-h2 - hh).max(axis=1) / np.abs(hh).max(axis=1))
    print(f"{case}: potential {ep:.1e}, gradient {eg:.1e}, Hessian {eh:.1e}")
    assert ep <= 1e-12 and eg <= 1e-12 and eh <= 1e-12
    assert np.max(np.abs(p2 / hp - 1)) <= 1e-12 and np.max(np.linalg.norm(g2 + hg, axis=1) / np.linalg.norm(hg, axis=1)) <= 1e-12


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("method", ["dop853", "dprkn8"])
def test_orbits_match_oracle(case, method):
    import agama_amd
    gpu, orc, (radii, _, _) = load(case)
    g = np.random.Generator(np.random.Philox(72))
    n = 256
    r = np.exp(g.uniform(np.log(radii[0] * 3), np.log(radii[-1] / 3), n))
    d = g.normal(size=(n, 3))
    pos = d / np.linalg.norm(d, axis=1)[:, None] * r[:, None]
    _, gr = orc.eval(pos)
    vc = np.sqrt(np.linalg.norm(gr, axis=1) * r)
    vel = g.normal(size=(n, 3)) * (0.5 * vc / np.sqrt(3))[:, None]
    ic = np.hstack([pos, vel])
    T = 20 * 2 * np.pi * r / vc
    print(f"case: 256 orbits x 20 periods, {case}, {method}", flush=True)
    info = {}
    agama_amd.orbit(potential=gpu, ic=ic, time=T, accuracy=1e-8, method=method, info=info)
    ref = orc.orbits(ic, T, trajsize=1, accuracy=1e-8, method=method)
    assert np.all(info["status"] == 0) and np.all(ref["status"] == 0)
    err = np.linalg.norm(info["final"][:, :3] - ref["final"][:, :3], axis=1) / np.linalg.norm(ref["final"][:, :3], axis=1)
    same = np.mean(info["nsteps"] == ref["nsteps"])
    # a recompile of the reference (the oracle with FMA contraction) on the same orbits
    from oracle import pyoracle
    from oracle.pyoracle import OraclePotential
    fast = OraclePotential(coefs=load_coefs(case), library=pyoracle.fast())
    re = fast.orbits(ic, T, trajsize=1, accuracy=1e-8, method=method)
    same_re = np.mean(re["nsteps"] == ref["nsteps"])
    print(f"{case} {method}: position err median {np.median(err):.1e}, identical steps {same:.3f} (recompiled "
          f"oracle {same_re:.3f}), max dE/E GPU {info['maxde'].max():.1e} oracle {ref['maxde'].max():.1e}")
    assert np.median(err) <= 1e-7 and same >= min(0.9, same_re - 0.05)
    assert np.median(info["maxde"]) <= 1.5 * np.median(ref["maxde"]) + 1e-13
