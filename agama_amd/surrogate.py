"""MLP potential surrogate: training on the GPU, evaluation, and the raw network (ctypes over
the C ABI of libagama_amd; see the agx_surrogate_* declarations in include/agama_amd.h).

    pot = agama_amd.Potential(file="halo.coef")
    sur = agama_amd.Surrogate.train(pot, samples=20e6, steps=6000)
    phi, acc = sur.eval(points, pot=True, acc=True)
    sur.save("halo.mlp.npz"); sur = agama_amd.Surrogate.load(pot, "halo.mlp.npz")
"""
import ctypes

import numpy as np

from ._native import lib, check, last_error, AgamaAmdError

NPARAM = 198_913
_f32 = np.float32


def _ptr(a):
    return None if a is None else a.ctypes.data


class Surrogate:
    """Phi_NN(x) = Phi_base(x) (1 + out_scale * MLP(f(x))), Phi_base = the Multipole truncated
    to l <= base_lmax; force by the network's hand-written backward (bf16 MFMA, fp32 sums)"""

    def __init__(self, potential, handle):
        self._pot = potential    # keeps the potential alive: the surrogate refers to it
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.agx_surrogate_destroy(h)
            self._h = None

    # ---- construction ----
    @classmethod
    def train(cls, potential, **params):
        text = ", ".join(f"{k}={v}" for k, v in params.items())
        h = lib.agx_surrogate_train(potential._h, text.encode())
        if not h:
            raise AgamaAmdError(f"surrogate training failed: {last_error()}")
        return cls(potential, h)

    @classmethod
    def from_params(cls, potential, theta, base_lmax, ln_scale, out_scale, rmin, rmax):
        th = np.ascontiguousarray(theta, dtype=_f32)
        if th.size != NPARAM:
            raise ValueError(f"theta must have {NPARAM} values")
        h = lib.agx_surrogate_create(potential._h, _ptr(th), int(base_lmax), float(ln_scale),
                                     float(out_scale), float(rmin), float(rmax))
        if not h:
            raise AgamaAmdError(f"surrogate creation failed: {last_error()}")
        return cls(potential, h)

    @classmethod
    def load(cls, potential, path):
        with np.load(path, allow_pickle=False) as z:
            m = z["meta"]
            return cls.from_params(potential, z["theta"], int(m[0]), m[1], m[2], m[3], m[4])

    def save(self, path):
        theta, meta = self.params()
        np.savez(path, theta=theta, meta=meta, loss=self.loss_history())

    # ---- inspection ----
    def params(self):
        theta = np.empty(NPARAM, dtype=_f32)
        meta = np.empty(8)
        check(lib.agx_surrogate_info(self._h, _ptr(theta), _ptr(meta)), "surrogate info")
        return theta, meta

    @property
    def meta(self):
        m = self.params()[1]
        return dict(base_lmax=int(m[0]), ln_scale=float(m[1]), out_scale=float(m[2]),
                    rmin=float(m[3]), rmax=float(m[4]), steps=int(m[5]), last_loss=float(m[6]))

    def loss_history(self):
        n = lib.agx_surrogate_loss_history(self._h, None, 0)
        out = np.empty(max(n, 0), dtype=_f32)
        if n > 0:
            lib.agx_surrogate_loss_history(self._h, _ptr(out), n)
        return out

    # ---- evaluation ----
    def _eval(self, points, want_pot, want_grad):
        pts = np.ascontiguousarray(points, dtype=np.float64)
        p2 = pts.reshape(-1, 3)
        n = len(p2)
        phi = np.empty(n) if want_pot else None
        grad = np.empty((n, 3)) if want_grad else None
        check(lib.agx_surrogate_eval_many(self._h, n, _ptr(p2), _ptr(phi), _ptr(grad)), "surrogate eval")
        return phi, grad

    def eval(self, points, pot=False, acc=False):
        """like Potential.eval: Phi and/or acceleration -grad Phi"""
        phi, grad = self._eval(points, pot, acc)
        res = []
        if pot:
            res.append(phi)
        if acc:
            res.append(-grad)
        return res[0] if len(res) == 1 else tuple(res)

    def network(self, feat):
        """the raw network on features [n,4] (float32): out[n], d out / d feat [n,4]"""
        f = np.ascontiguousarray(feat, dtype=_f32).reshape(-1, 4)
        out = np.empty(len(f), dtype=_f32)
        d = np.empty((len(f), 4), dtype=_f32)
        check(lib.agx_surrogate_network(self._h, len(f), _ptr(f), _ptr(out), _ptr(d)), "surrogate network")
        return out, d

    def train_grad(self, feat, target):
        """loss and parameter gradient of one batch through the training kernels (no update)"""
        f = np.ascontiguousarray(feat, dtype=_f32).reshape(-1, 4)
        t = np.ascontiguousarray(target, dtype=_f32)
        loss = np.empty(1, dtype=_f32)
        g = np.empty(NPARAM, dtype=_f32)
        check(lib.agx_surrogate_train_grad(self._h, len(f), _ptr(f), _ptr(t), _ptr(loss), _ptr(g)),
              "surrogate train_grad")
        return float(loss[0]), g


def debug_gemm(A, B, a_kcontig, b_kcontig, splits=1):
    """the training GEMM on bf16 bit patterns (uint16 arrays): returns float32 C[M][N]"""
    A = np.ascontiguousarray(A, dtype=np.uint16)
    B = np.ascontiguousarray(B, dtype=np.uint16)
    M, K = A.shape if a_kcontig else A.shape[::-1]
    N = B.shape[0] if b_kcontig else B.shape[1]
    C = np.empty((M, N), dtype=_f32)
    check(lib.agx_debug_gemm(M, N, K, splits, int(a_kcontig), int(b_kcontig), _ptr(A), _ptr(B), _ptr(C)),
          "gemm")
    return C
