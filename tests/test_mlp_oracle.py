"""Pins of the MLP surrogate's CPU oracle (oracle/mlp_oracle.py): its hand-written backward
against central finite differences of its own float64 forward (input gradient, weight
gradients), the feature chain rule against finite differences of the assembled potential,
and bfloat16 rounding against exactly representable cases."""
import numpy as np

from oracle import mlp_oracle as M


def small_net(seed=3, scale=1.0):
    th = M.init_params(seed).astype(np.float64)
    g = np.random.Generator(np.random.Philox(seed + 1))
    # non-zero biases so that every term of the backward is exercised
    p = M.unpack(th)
    for k in ("b1", "b2", "b3", "b4", "b5"):
        p[k] = g.normal(size=p[k].shape) * 0.1
    return M.pack(p).astype(np.float64) * scale


def test_param_count():
    assert M.NPARAM == 198_913


def test_bf16_rounding_known_values():
    v = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7, -2.5, 3.0e38, 1e-40])
    r = M.bf16_round(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                       # tie -> even mantissa
    assert r[2] == 1 + 2 ** -6               # tie -> even mantissa (upwards)
    assert r[3] == 1 + 2 ** -7               # representable
    assert r[4] == -2.5
    assert abs(r[5] / 3.0e38 - 1) < 2 ** -8
    rr = M.bf16_round(np.linspace(-3, 3, 1001))
    assert np.all(np.abs(rr - np.linspace(-3, 3, 1001)) <= 2 ** -8 * np.abs(np.linspace(-3, 3, 1001)) + 1e-300)


def test_input_gradient_matches_finite_differences():
    th = small_net()
    g = np.random.Generator(np.random.Philox(5))
    f = g.uniform(-1, 1, size=(6, 4))
    out, cache = M.forward(th, f)
    dfeat, _ = M.backward(cache, np.ones(len(f)), want_params=False)
    eps = 1e-6
    for i in range(4):
        d = np.zeros(4); d[i] = eps
        fd = (M.forward(th, f + d)[0] - M.forward(th, f - d)[0]) / (2 * eps)
        assert np.allclose(dfeat[:, i], fd, rtol=1e-7, atol=1e-9)


def test_parameter_gradient_matches_finite_differences():
    th = small_net()
    g = np.random.Generator(np.random.Philox(6))
    f = g.uniform(-1, 1, size=(8, 4))
    t = g.normal(size=8)
    loss, grad = M.loss_and_grad(th, f, t, bf16=False)
    # indices across every tensor: first/last element of each plus random ones
    idx, o = [], 0
    for _, shp in M.SHAPES:
        n = int(np.prod(shp))
        idx += [o, o + n - 1, o + int(g.integers(n))]
        o += n
    eps = 1e-6
    for k in idx:
        d = np.zeros_like(th); d[k] = eps
        lp = M.loss_and_grad(th + d, f, t, bf16=False)[0]
        lm = M.loss_and_grad(th - d, f, t, bf16=False)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(grad[k] - fd) <= 1e-6 * (abs(fd) + 1e-3), (k, grad[k], fd)


def test_one_neuron_closed_form():
    """only hidden neuron 0 is connected: out = w5_0 * silu(W4_00 silu(W3_00 silu(W2_00 silu(W1_0.f))))"""
    p = {k: np.zeros(s) for k, s in M.SHAPES}
    p["W1"][0] = [0.3, -0.2, 0.5, 0.7]
    p["W2"][0, 0] = 1.3; p["W3"][0, 0] = -0.8; p["W4"][0, 0] = 1.1; p["w5"][0] = 2.0; p["b5"][0] = 0.25
    th = M.pack(p).astype(np.float64)
    f = np.array([[0.1, 0.2, 0.3, 0.4]])
    q = M.unpack(th)    # the parameters as stored (float32)
    silu = lambda z: z / (1 + np.exp(-z))
    a = silu(q["W1"][0] @ f[0])
    expect = q["w5"][0] * silu(q["W4"][0, 0] * silu(q["W3"][0, 0] * silu(q["W2"][0, 0] * a))) + q["b5"][0]
    # hidden neurons other than 0 have zero input and contribute silu(0) = 0
    assert abs(M.forward(th, f)[0][0] - expect) < 1e-15


def test_surrogate_chain_rule_matches_finite_differences():
    th = small_net(scale=0.5)
    g = np.random.Generator(np.random.Philox(7))
    pos = g.normal(size=(5, 3)) * 2

    def base(x):   # Plummer baseline: closed-form potential and gradient
        r2 = np.sum(x * x, 1)
        return -1 / np.sqrt(1 + r2), x / (1 + r2)[:, None] ** 1.5

    pb, gb = base(pos)
    phi, grad = M.surrogate_eval(th, pos, pb, gb, np.log(100), 0.01, bf16=False)
    eps = 1e-6
    for c in range(3):
        d = np.zeros(3); d[c] = eps
        pp = M.surrogate_eval(th, pos + d, *base(pos + d), np.log(100), 0.01, bf16=False)[0]
        pm = M.surrogate_eval(th, pos - d, *base(pos - d), np.log(100), 0.01, bf16=False)[0]
        assert np.allclose(grad[:, c], (pp - pm) / (2 * eps), rtol=1e-7, atol=1e-10)
