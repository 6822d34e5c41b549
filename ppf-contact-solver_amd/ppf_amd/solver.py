"""Block-Jacobi PCG over the reference's split block operator.

Mirrors solver::solve (crates/ppf-cts-solver/src/cpp/solver/solver.cu:2180-2321)
for the block-Jacobi preconditioner: M = B (upper-triangle 3x3 block CSR, applied
both ways) + C (per-row diagonal), solved to |r|_1/|r0|_1 < tol. All compute is
in libppf_mi355x.so; tensors are torch device buffers.
"""
import ctypes

import torch

from . import _lib
from ._lib import PcgResult, check

STATUS = {0: "converged", 1: "truncated", 2: "iteration cap", 3: "non-finite",
          4: "indefinite (pAp)", 5: "non-SPD preconditioner"}


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _need(t, dtype, numel, name):
    if not t.is_cuda or t.dtype != dtype or not t.is_contiguous() or t.numel() != numel:
        raise ValueError(f"{name}: expected contiguous cuda {dtype} with {numel} "
                         f"elements, got {t.dtype} {tuple(t.shape)} on {t.device}")


class BlockJacobiPCG:
    """A solve plan for one sparsity pattern (row_offset: n+1, col_index: nnz, int32)."""

    def __init__(self, row_offset, col_index):
        self.lib = _lib.load()
        self.n = row_offset.numel() - 1
        self.nnz = col_index.numel()
        _need(row_offset, torch.int32, self.n + 1, "row_offset")
        _need(col_index, torch.int32, self.nnz, "col_index")
        self._keep = (row_offset, col_index)
        h = ctypes.c_void_p()
        check(self.lib.ppf_pcg_plan_create(self.n, self.nnz, _ptr(row_offset),
                                           _ptr(col_index), _stream(),
                                           ctypes.byref(h)), "ppf_pcg_plan_create")
        self._h = h

    def repattern(self, row_offset, col_index):
        """Switch to another pattern with the same number of rows."""
        nnz = col_index.numel()
        _need(row_offset, torch.int32, self.n + 1, "row_offset")
        _need(col_index, torch.int32, nnz, "col_index")
        check(self.lib.ppf_pcg_plan_repattern(self._h, nnz, _ptr(row_offset),
                                              _ptr(col_index), _stream()),
              "ppf_pcg_plan_repattern")
        self.nnz = nnz
        self._keep = (row_offset, col_index)

    def close(self):
        if getattr(self, "_h", None):
            self.lib.ppf_pcg_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check_operands(self, value, cdiag):
        _need(value, torch.float32, 9 * self.nnz, "value")
        _need(cdiag, torch.float32, 9 * self.n, "cdiag")

    def solve(self, value, cdiag, b, x, tol=1e-3, max_iter=10000, err_hist=None):
        """Solve in place (x holds x0 on entry). Returns a PcgResult."""
        self._check_operands(value, cdiag)
        _need(b, torch.float32, 3 * self.n, "b")
        _need(x, torch.float32, 3 * self.n, "x")
        res = PcgResult()
        if err_hist is None:
            rc = self.lib.ppf_pcg_solve(self._h, _ptr(value), _ptr(cdiag), _ptr(b),
                                        _ptr(x), tol, max_iter, ctypes.byref(res),
                                        _stream())
        else:
            _need(err_hist, torch.float32, max_iter, "err_hist")
            rc = self.lib.ppf_pcg_solve_hist(self._h, _ptr(value), _ptr(cdiag),
                                             _ptr(b), _ptr(x), tol, max_iter,
                                             _ptr(err_hist), ctypes.byref(res),
                                             _stream())
        check(rc, "ppf_pcg_solve")
        return res

    def spmv(self, value, cdiag, x):
        self._check_operands(value, cdiag)
        _need(x, torch.float32, 3 * self.n, "x")
        y = torch.empty_like(x)
        check(self.lib.ppf_pcg_spmv(self._h, _ptr(value), _ptr(cdiag), _ptr(x),
                                    _ptr(y), _stream()), "ppf_pcg_spmv")
        return y

    def time_spmv(self, value, cdiag, x, reps=50):
        """Mean duration (ms) of one SpMV+dot kernel launch, HIP events."""
        self._check_operands(value, cdiag)
        _need(x, torch.float32, 3 * self.n, "x")
        y = torch.empty_like(x)
        ms = ctypes.c_float()
        check(self.lib.ppf_pcg_time_spmv(self._h, _ptr(value), _ptr(cdiag), _ptr(x),
                                         _ptr(y), reps, ctypes.byref(ms), _stream()),
              "ppf_pcg_time_spmv")
        return ms.value


def bj_invert(blocks):
    """Block-Jacobi inverse of (n, 9) column-major blocks (solver.cu:357-400)."""
    lib = _lib.load()
    blocks = blocks.contiguous()
    n = blocks.numel() // 9
    out = torch.empty_like(blocks)
    bad = torch.zeros(1, dtype=torch.int32, device=blocks.device)
    check(lib.ppf_bj_invert(n, _ptr(blocks), _ptr(out), _ptr(bad), _stream()),
          "ppf_bj_invert")
    return out, int(bad.item())


def symm3x3(blocks):
    """Eigen-decomposition of (n, 9) symmetric blocks: (values asc, vectors)."""
    lib = _lib.load()
    blocks = blocks.contiguous()
    n = blocks.numel() // 9
    val = torch.empty(n * 3, dtype=torch.float32, device=blocks.device)
    vec = torch.empty(n * 9, dtype=torch.float32, device=blocks.device)
    check(lib.ppf_symm3x3(n, _ptr(blocks), _ptr(val), _ptr(vec), _stream()),
          "ppf_symm3x3")
    return val, vec
