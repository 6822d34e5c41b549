"""The allocation-free PCG solve (ppf_pcg_solve_noalloc) against a plan's.

ppf_pcg_solve_noalloc builds the plan of ppf_pcg_plan_create and runs the solve
of ppf_pcg_solve on a fixed device region of the library, for callers that
cannot allocate. It launches the same kernels on the same tiles, so its
iterates, iteration count, status and residual must equal the plan's bit for
bit. Two calls in a row, the second on a smaller system, share the region:
the second must not see the first's contents. No device or pinned allocation
may happen inside the calls, and a system larger than the region is refused
with an error before any device work.
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from ppf_amd import _lib, scenes
from ppf_amd._lib import PcgResult
from ppf_amd.solver import BlockJacobiPCG
from ppf_amd.stepper import Stepper, scene_to_device

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _grid(dev, res, seed):
    s = scenes.grid_spd_system(res, seed_matrix=seed, seed_rhs=seed + 1)
    t = {k: torch.from_numpy(np.ascontiguousarray(s[k])).to(dev).reshape(-1)
         for k in ("row_offset", "col_index", "value", "cdiag", "b")}
    return s, t


def _drape_system(dev):
    sc = scenes.drape_scene(res=20, height=0.002)
    st = Stepper(scene_to_device(sc, dev))
    p = scenes.step_params()
    for _ in range(3):
        r = st.advance(p)
        assert r["status"] == 0, r
        p["prev_dt"] = r["dt"]
    s = st.system()
    st.close()
    return {k: s[k].reshape(-1).contiguous() for k in ("row_offset", "col_index", "value",
                                                      "cdiag", "b")}


def _noalloc(t, tol, max_iter=10000):
    lib = _lib.load()
    n = t["row_offset"].numel() - 1
    nnz = t["col_index"].numel()
    x = torch.zeros_like(t["b"])
    res = PcgResult()
    torch.cuda.synchronize()
    a0 = _lib.alloc_counts()
    rc = lib.ppf_pcg_solve_noalloc(n, nnz, _ptr(t["row_offset"]), _ptr(t["col_index"]),
                                   _ptr(t["value"]), _ptr(t["cdiag"]), _ptr(t["b"]), _ptr(x),
                                   tol, max_iter, ctypes.byref(res), _stream())
    a1 = _lib.alloc_counts()
    _lib.check(rc, "ppf_pcg_solve_noalloc")
    return x, res, (a1[0] - a0[0], a1[1] - a0[1])


def _planned(t, tol, max_iter=10000):
    plan = BlockJacobiPCG(t["row_offset"], t["col_index"])
    x = torch.zeros_like(t["b"])
    res = plan.solve(t["value"], t["cdiag"], t["b"], x, tol=tol, max_iter=max_iter)
    torch.cuda.synchronize()
    plan.close()
    return x, res


def _same(a, b):
    xa, ra = a
    xb, rb = b
    assert (ra.iter, ra.status, ra.break_cause) == (rb.iter, rb.status, rb.break_cause)
    assert np.float32(ra.resid).view(np.uint32) == np.float32(rb.resid).view(np.uint32)
    assert torch.equal(xa.view(torch.int32), xb.view(torch.int32))


def test_noalloc_equals_plan_solve_two_calls(dev):
    big_s, big = _grid(dev, 60, 11)          # 3,600 rows
    small = _drape_system(dev)               # drape with sphere contact, fewer rows
    tiny_s, tiny = _grid(dev, 24, 13)        # 576 rows
    n_big = big["row_offset"].numel() - 1
    n_small = small["row_offset"].numel() - 1
    assert n_small < n_big
    want = [_planned(big, 1e-3), _planned(small, 1e-3), _planned(tiny, 1e-5)]
    got = []
    for t, tol in ((big, 1e-3), (small, 1e-3), (tiny, 1e-5)):
        x, res, (na, nf) = _noalloc(t, tol)
        print(f"noalloc: n {t['row_offset'].numel() - 1} nnz {t['col_index'].numel()} "
              f"iter {res.iter} status {res.status} resid {res.resid:.3e} "
              f"allocs {na} frees {nf}", flush=True)
        assert na == 0 and nf == 0
        got.append((x, res))
    for g, w in zip(got, want):
        assert g[1].status == 0
        _same(g, w)
    # and the solve itself against the oracle (iteration schedule, solution)
    ref = oracle.pcg(big_s, tol=1e-3)
    assert abs(got[0][1].iter - ref["iter"]) <= 4
    xg = got[0][0].cpu().numpy()
    assert np.abs(xg - ref["x"]).max() / np.abs(ref["x"]).max() < 1e-2


def test_noalloc_repeat_is_bitwise(dev):
    _, t = _grid(dev, 40, 21)
    a = _noalloc(t, 1e-3)
    _, u = _grid(dev, 52, 22)
    _noalloc(u, 1e-3)                        # another system in between
    b = _noalloc(t, 1e-3)
    _same(a[:2], b[:2])


def test_noalloc_refuses_oversize(dev):
    lib = _lib.load()
    d = torch.zeros(64, dtype=torch.int32, device=dev)
    f = torch.zeros(64, dtype=torch.float32, device=dev)
    res = PcgResult()
    rc = lib.ppf_pcg_solve_noalloc(1 << 20, 1 << 23, _ptr(d), _ptr(d), _ptr(f), _ptr(f), _ptr(f),
                                   _ptr(f), 1e-3, 100, ctypes.byref(res), _stream())
    assert rc == 1
    assert "beyond the fixed" in lib.ppf_last_error().decode()
