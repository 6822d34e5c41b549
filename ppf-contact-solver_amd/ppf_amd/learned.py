"""Learned preconditioner (ctypes over libppf_mi355x.so): a message-passing
network over a Newton system's 3x3-block graph that predicts a block FSAI
factor, M^-1 = G^T G. Block Jacobi (the reference's preconditioner,
solver.cu:357-400) is its zero-output-head starting point and the fallback.

Weights are one fp32 array (row-major [in][out]): W0 16x64, c0; per layer
We 80x64, ce, Wn 128x64, cn; Wo 64x6, co; Wg 144x9, cg.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import PcgResult, check

H, FN, FE, LAYERS = 64, 16, 16, 3
ORDER = (["W0", "c0"] + [k for l in range(LAYERS) for k in (f"We{l}", f"ce{l}", f"Wn{l}", f"cn{l}")]
         + ["Wo", "co", "Wg", "cg"])
SHAPES = {"W0": (FN, H), "c0": (H,), "Wo": (H, 6), "co": (6,), "Wg": (2 * H + FE, 9), "cg": (9,)}
for _l in range(LAYERS):
    SHAPES.update({f"We{_l}": (H + FE, H), f"ce{_l}": (H,), f"Wn{_l}": (2 * H, H),
                   f"cn{_l}": (H,)})


def flatten(w):
    return np.concatenate([np.asarray(w[k], np.float32).reshape(-1) for k in ORDER])


def unflatten(flat):
    out, o = {}, 0
    for k in ORDER:
        size = int(np.prod(SHAPES[k]))
        out[k] = np.asarray(flat[o:o + size], np.float64).reshape(SHAPES[k])
        o += size
    return out


def init_weights(seed=0):
    """He-style weights with zero output heads (M^-1 = block Jacobi)."""
    rng = np.random.default_rng(seed)
    w = {}
    for k in ORDER:
        shape = SHAPES[k]
        if k.startswith("W") and k not in ("Wo", "Wg"):
            w[k] = rng.normal(0.0, np.sqrt(2.0 / shape[0]), shape)
        else:
            w[k] = np.zeros(shape)
    return flatten(w)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class LearnedPreconditioner:
    def __init__(self, weights=None):
        self.lib = _lib.load()
        self.count = self.lib.ppf_lpc_weight_count()
        h = ctypes.c_void_p()
        check(self.lib.ppf_lpc_create(ctypes.byref(h)), "ppf_lpc_create")
        self._h = h
        self.set_weights(init_weights() if weights is None else weights)

    def close(self):
        if getattr(self, "_h", None):
            self.lib.ppf_lpc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_weights(self, flat):
        flat = np.ascontiguousarray(flat, np.float32)
        if flat.size != self.count:
            raise ValueError(f"{flat.size} weights, {self.count} expected")
        check(self.lib.ppf_lpc_set_weights(self._h, flat.ctypes.data_as(ctypes.c_void_p),
                                           self.count, _stream()), "ppf_lpc_set_weights")

    def weights(self):
        out = np.zeros(self.count, np.float32)
        check(self.lib.ppf_lpc_get_weights(self._h, out.ctypes.data_as(ctypes.c_void_p),
                                           self.count, _stream()), "ppf_lpc_get_weights")
        return out

    def factor(self, plan, value, cdiag, b):
        """Network and factor for the system on `plan` (a BlockJacobiPCG)."""
        check(self.lib.ppf_lpc_factor(self._h, plan._h, _ptr(value), _ptr(cdiag), _ptr(b),
                                      _stream()), "ppf_lpc_factor")

    def outputs(self, n, nnz, dev):
        t = torch.empty(6 * n, dtype=torch.float32, device=dev)
        u = torch.empty(9 * nnz, dtype=torch.float32, device=dev)
        gd = torch.empty(9 * n, dtype=torch.float32, device=dev)
        gl = torch.empty(9 * nnz, dtype=torch.float32, device=dev)
        check(self.lib.ppf_lpc_outputs(self._h, _ptr(t), _ptr(u), _ptr(gd), _ptr(gl),
                                       _stream()), "ppf_lpc_outputs")
        return t, u, gd, gl

    def gradient(self, plan, value, cdiag, probes):
        """Loss and weight gradient (numpy) for probes (k, n, 3) on the device."""
        loss = ctypes.c_float()
        g = np.zeros(self.count, np.float32)
        k = probes.shape[0]
        check(self.lib.ppf_lpc_gradient(self._h, plan._h, _ptr(value), _ptr(cdiag),
                                        _ptr(probes), k, ctypes.byref(loss),
                                        g.ctypes.data_as(ctypes.c_void_p), _stream()),
              "ppf_lpc_gradient")
        return loss.value, g

    def adam(self, lr, step):
        check(self.lib.ppf_lpc_adam(self._h, lr, step, _stream()), "ppf_lpc_adam")

    def solve(self, plan, value, cdiag, b, x, tol=1e-3, max_iter=10000):
        res = PcgResult()
        check(self.lib.ppf_pcg_solve_lpc(plan._h, self._h, _ptr(value), _ptr(cdiag), _ptr(b),
                                         _ptr(x), tol, max_iter, ctypes.byref(res), _stream()),
              "ppf_pcg_solve_lpc")
        return res
