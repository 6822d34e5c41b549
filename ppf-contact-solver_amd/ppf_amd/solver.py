"""Block-Jacobi PCG over the reference's split block operator.

Mirrors solver::solve (crates/ppf-cts-solver/src/cpp/solver/solver.cu:2180-2321)
for the block-Jacobi preconditioner: M = B (upper-triangle 3x3 block CSR, applied
both ways) + C (per-row diagonal), solved to |r|_1/|r0|_1 < tol. All compute is
in libppf_mi355x.so; tensors are torch device buffers.
"""
import ctypes

import numpy as np
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

    def set_order(self, order):
        """Solve in another row order (new row k = original row order[k], an
        int32 device tensor kept alive here); takes effect at the next
        repattern. None: the original order."""
        if order is not None:
            _need(order, torch.int32, self.n, "order")
        self._order = order
        check(self.lib.ppf_pcg_plan_set_order(self._h, _ptr(order) if order is not None
                                              else None, _stream()), "ppf_pcg_plan_set_order")

    FORMS = {"auto": 0, "ell": 1, "split": 2, "sym": 3}

    def set_form(self, form):
        """Fix the SpMV form ('auto', 'ell', 'split', 'sym'); returns the form
        the plan now runs."""
        used = ctypes.c_int()
        check(self.lib.ppf_pcg_plan_set_spmv_form(self._h, self.FORMS[form], ctypes.byref(used),
                                                  _stream()), "ppf_pcg_plan_set_spmv_form")
        return {v: k for k, v in self.FORMS.items()}[used.value]

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


class BatchPCG(BlockJacobiPCG):
    """One plan over many independent systems stacked block-diagonally:
    scene c owns rows scene_row[c] .. scene_row[c + 1]. Every kernel launch
    serves all scenes; each scene has its own scalars and result."""

    def __init__(self, row_offset, col_index, scene_row):
        self.lib = _lib.load()
        self.n = row_offset.numel() - 1
        self.nnz = col_index.numel()
        _need(row_offset, torch.int32, self.n + 1, "row_offset")
        _need(col_index, torch.int32, self.nnz, "col_index")
        rows = np.ascontiguousarray(np.asarray(scene_row, np.uint32))
        self.nscene = len(rows) - 1
        self._keep = (row_offset, col_index)
        h = ctypes.c_void_p()
        check(self.lib.ppf_pcg_plan_create_batch(self.n, self.nnz, _ptr(row_offset),
                                                 _ptr(col_index), self.nscene,
                                                 rows.ctypes.data_as(ctypes.c_void_p),
                                                 _stream(), ctypes.byref(h)),
              "ppf_pcg_plan_create_batch")
        self._h = h

    def set_tiles(self, spmv_rows, vector_rows):
        check(self.lib.ppf_pcg_plan_set_tiles(self._h, spmv_rows, vector_rows, _stream()),
              "ppf_pcg_plan_set_tiles")

    def autotune(self, value, cdiag, b, tol=1e-3):
        """Measure candidate tile sizes on this system, keep the fastest:
        returns (spmv rows, vector rows, ms of the capped timing solve)."""
        self._check_operands(value, cdiag)
        _need(b, torch.float32, 3 * self.n, "b")
        rows = np.zeros(2, np.uint32)
        ms = ctypes.c_float()
        check(self.lib.ppf_pcg_plan_autotune(self._h, _ptr(value), _ptr(cdiag), _ptr(b), tol,
                                             rows.ctypes.data_as(ctypes.c_void_p),
                                             ctypes.byref(ms), _stream()),
              "ppf_pcg_plan_autotune")
        return int(rows[0]), int(rows[1]), ms.value

    def solve(self, value, cdiag, b, x, tol=1e-3, max_iter=10000):
        """Solve every scene in place; returns one PcgResult per scene."""
        self._check_operands(value, cdiag)
        _need(b, torch.float32, 3 * self.n, "b")
        _need(x, torch.float32, 3 * self.n, "x")
        res = (PcgResult * self.nscene)()
        check(self.lib.ppf_pcg_solve_batch(self._h, _ptr(value), _ptr(cdiag), _ptr(b), _ptr(x),
                                           tol, max_iter, ctypes.cast(res, ctypes.c_void_p),
                                           _stream()), "ppf_pcg_solve_batch")
        return list(res)


def stack_systems(systems):
    """Block-diagonal union of systems (dicts of row_offset, col_index,
    value, cdiag, b; numpy or torch): the stacked arrays and the first row of
    every scene."""
    def host(a):
        return a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    rp, ci, val, cd, b, rows = [np.zeros(1, np.int64)], [], [], [], [], [0]
    base = nz = 0
    for s in systems:
        r = host(s["row_offset"]).astype(np.int64)
        n = len(r) - 1
        rp.append(r[1:] + nz)
        ci.append(host(s["col_index"]).astype(np.int64) + base)
        val.append(host(s["value"]).reshape(-1))
        cd.append(host(s["cdiag"]).reshape(-1))
        b.append(host(s["b"]).reshape(-1))
        base += n
        nz += int(r[-1])
        rows.append(base)
    return dict(row_offset=np.concatenate(rp).astype(np.int32),
                col_index=np.concatenate(ci).astype(np.int32),
                value=np.concatenate(val).astype(np.float32),
                cdiag=np.concatenate(cd).astype(np.float32),
                b=np.concatenate(b).astype(np.float32)), np.array(rows, np.uint32)


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
