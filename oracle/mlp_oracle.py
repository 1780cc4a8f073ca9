"""CPU oracle for the MLP potential surrogate (test infrastructure only: imported by tests/,
__graft_entry__.smoke() and bench.py's CPU baseline, never by the product path).

The network has no counterpart in the reference; what it approximates is the Multipole
potential (src/potential_multipole.cpp:1969-2080).  It is restated here in float64 numpy:

    features  f = (|x|/r, |y|/r, |z|/r, ln(r) / ln_scale)
    h1 = silu(W1 f + b1), h_l = silu(W_l h_{l-1} + b_l)  (l = 2..4, width 256)
    out = w5 . h4 + b5
    Phi_NN(x) = Phi_base(x) * (1 + out_scale * out)

with Phi_base the FP64 Multipole truncated to l <= 2 (evaluated by the caller).  The
parameter vector is flat float32 in the order W1[256x4] b1 W2[256x256] b2 W3 b3 W4 b4 w5[256]
b5[1] (198,913 values), weights row-major [out][in].

`bf16=True` emulates the device arithmetic: weights and the activations passed between layers
(and to the output layer) are rounded to bfloat16 (round to nearest even); the SiLU
derivatives used by the backward pass are rounded the same way; sums are exact (float64),
which differs from the device's float32 accumulation only by accumulation rounding.

Parity of this oracle is pinned in tests/test_mlp_oracle.py by central finite differences of
its own float64 forward (input gradient and every weight gradient) and by a closed-form
one-neuron network.
"""
import numpy as np

WIDTH = 256
NIN = 4
SHAPES = [("W1", (WIDTH, NIN)), ("b1", (WIDTH,)), ("W2", (WIDTH, WIDTH)), ("b2", (WIDTH,)),
          ("W3", (WIDTH, WIDTH)), ("b3", (WIDTH,)), ("W4", (WIDTH, WIDTH)), ("b4", (WIDTH,)),
          ("w5", (WIDTH,)), ("b5", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(theta):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(theta[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], dtype=np.float32).ravel() for name, _ in SHAPES])


def init_params(seed=1):
    """He-style initial weights (normal, std sqrt(2/fan_in) for hidden layers), zero biases"""
    g = np.random.Generator(np.random.Philox(seed))
    p = {}
    for name, shp in SHAPES:
        if name.startswith("b"):
            p[name] = np.zeros(shp)
        elif name == "w5":
            p[name] = g.normal(size=shp) * np.sqrt(1.0 / WIDTH)
        else:
            p[name] = g.normal(size=shp) * np.sqrt(2.0 / shp[1])
    return pack(p)


def bf16_round(a):
    """round float64/float32 values to the nearest bfloat16 (ties to even), returned as float64"""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def features(pos, ln_scale):
    """f[n,4] and df/dx [n,4,3] from Cartesian positions"""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    r = np.linalg.norm(pos, axis=1)
    a = np.abs(pos)
    f = np.empty((len(pos), 4))
    f[:, :3] = a / r[:, None]
    f[:, 3] = np.log(r) / ln_scale
    J = np.zeros((len(pos), 4, 3))
    for i in range(3):
        J[:, i, i] = np.sign(pos[:, i]) / r
        J[:, i, :] -= (a[:, i] / r ** 3)[:, None] * pos
    J[:, 3, :] = pos / (ln_scale * r ** 2)[:, None]
    return f, J


def _silu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return z * s, s * (1.0 + z * (1.0 - s))


def forward(theta, f, bf16=False):
    """network output out[n] and the cache for backward"""
    p = unpack(theta)
    rnd = bf16_round if bf16 else (lambda v: v)
    Ws = [rnd(p["W1"]), rnd(p["W2"]), rnd(p["W3"]), rnd(p["W4"])]
    bs = [p["b1"], p["b2"], p["b3"], p["b4"]]
    w5 = p["w5"]
    h = np.asarray(f, dtype=np.float64)
    hs, ss = [h], []
    for l in range(4):
        # layer 1 multiplies the float32 features by float32 weights on the device
        W = p["W1"] if l == 0 else Ws[l]
        z = h @ W.T + bs[l]
        hv, sv = _silu(z)
        h = rnd(hv)
        hs.append(h)
        ss.append(rnd(sv))
    out = h @ w5 + p["b5"][0]
    return out, dict(p=p, Ws=Ws, hs=hs, ss=ss, rnd=rnd)


def backward(cache, dout, want_params=True):
    """given dL/dout[n]: dL/dfeatures [n,4] and (optionally) the flat parameter gradient.
    With bf16=True the gradients passed between layers are rounded like the device's."""
    p, Ws, hs, ss, rnd = cache["p"], cache["Ws"], cache["hs"], cache["ss"], cache["rnd"]
    dout = np.asarray(dout, dtype=np.float64)
    grads = {}
    grads["w5"] = hs[4].T @ dout
    grads["b5"] = np.array([dout.sum()])
    g = dout[:, None] * p["w5"][None, :] * ss[3]            # dL/dz4
    for l in (3, 2, 1):
        grads[f"W{l + 1}"] = rnd(g).T @ hs[l]
        grads[f"b{l + 1}"] = g.sum(0)
        g = (rnd(g) @ Ws[l]) * ss[l - 1]                      # dL/dz_l
    grads["W1"] = g.T @ hs[0]
    grads["b1"] = g.sum(0)
    dfeat = g @ p["W1"]
    return dfeat, (pack(grads) if want_params else None)


def surrogate_eval(theta, pos, phi_base, grad_base, ln_scale, out_scale, bf16=True):
    """Phi_NN and grad Phi_NN at positions, given the baseline potential and gradient there"""
    f, J = features(pos, ln_scale)
    out, cache = forward(theta, f, bf16)
    dfeat, _ = backward(cache, np.ones(len(f)), want_params=False)
    dout_dx = np.einsum("nf,nfc->nc", dfeat, J)
    phi = phi_base * (1 + out_scale * out)
    grad = grad_base * (1 + out_scale * out)[:, None] + (phi_base * out_scale)[:, None] * dout_dx
    return phi, grad


def loss_and_grad(theta, f, target, bf16=True):
    """mean squared error of out against target and its parameter gradient"""
    out, cache = forward(theta, f, bf16)
    e = out - target
    loss = np.mean(e * e)
    _, g = backward(cache, 2 * e / len(e))
    return loss, g
