"""Batched block-Jacobi PCG (one plan over many uneven systems stacked
block-diagonally) through the C ABI.

Each scene keeps its own scalars, so:
  * a scene's solution is bitwise the same whatever else is in the batch and
    wherever it sits (same tile sizes), and identical scenes give identical
    solutions;
  * against a solve of the scene on its own, iteration counts agree within
    the check stride (the partial sums are grouped by other tiles, so the
    last bits differ), and every scene meets the stopping rule: |b - M x|_1 /
    |b|_1 < tol in fp64 on the host (x0 = 0, so r0 = b), with 1.5x slack for
    the fp32 residual the device measured.
The plain per-scene solve is itself pinned to the oracle's PCG in
test_gpu_pcg.py.
"""
import numpy as np
import pytest
import torch

from ppf_amd import scenes
from ppf_amd._lib import load
from ppf_amd.solver import BatchPCG, BlockJacobiPCG, stack_systems

pytestmark = pytest.mark.gpu


def cloth(res, seed):
    import oracle
    sc = scenes.drape_scene(res=res, height=0.05)
    sh = dict(sc["shell"])
    x = sc["mesh"]["x"].copy()
    rng = np.random.default_rng(seed)
    sh["x"] = x + rng.normal(0, 2e-3, x.shape).astype(np.float32)
    sh["x0"] = x
    sh["target"] = x
    rp, ci = scenes.upper_pattern(sh["nv"], sh["face"], sh["hinge"])
    f, cd, val, _ = oracle.elastic_assemble(sh, rp, ci, 1.0 / 60.0)
    return dict(row_offset=rp.astype(np.int32), col_index=ci.astype(np.int32),
                value=val.reshape(-1).astype(np.float32), cdiag=cd.astype(np.float32),
                b=f.astype(np.float32))


def systems():
    out = [cloth(9, 0), cloth(21, 1), cloth(14, 2)]
    g = scenes.grid_spd_system(12)
    out.append({k: g[k] for k in ("row_offset", "col_index", "value", "cdiag", "b")})
    lr = scenes.long_row_spd_system(n=300, seed=4)
    out.append({k: lr[k] for k in ("row_offset", "col_index", "value", "cdiag", "b")})
    out.append(cloth(33, 3))
    return out


def to_dev(s, dev):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev).reshape(-1)
            for k, v in s.items()}


def solve_batch(dev, syss, tiles=None, autotune=False):
    st, rows = stack_systems(syss)
    t = to_dev(st, dev)
    plan = BatchPCG(t["row_offset"], t["col_index"], rows)
    tuned = None
    if tiles:
        plan.set_tiles(*tiles)
    if autotune:
        tuned = plan.autotune(t["value"], t["cdiag"], t["b"])
    x = torch.zeros_like(t["b"])
    res = plan.solve(t["value"], t["cdiag"], t["b"], x)
    plan.close()
    xs = x.cpu().numpy()
    return [xs[3 * rows[c]:3 * rows[c + 1]] for c in range(len(syss))], res, tuned


def rel_resid(s, x):
    n = len(s["row_offset"]) - 1
    M = scenes.sparse_from_upper(n, s["row_offset"], s["col_index"], s["value"], s["cdiag"])
    b = s["b"].astype(np.float64)
    return np.abs(b - M @ x.astype(np.float64)).sum() / np.abs(b).sum()


def test_batch_matches_single_solves(dev):
    syss = systems()
    xs, res, _ = solve_batch(dev, syss)
    for c, s in enumerate(syss):
        t = to_dev(s, dev)
        plan = BlockJacobiPCG(t["row_offset"], t["col_index"])
        x = torch.zeros_like(t["b"])
        one = plan.solve(t["value"], t["cdiag"], t["b"], x)
        plan.close()
        rr = rel_resid(s, xs[c])
        print(f"scene {c}: n {len(s['row_offset']) - 1} batch {res[c].iter} it "
              f"(status {res[c].status}), alone {one.iter} it; fp64 residual {rr:.2e}")
        assert res[c].status == 0 and one.status == 0
        assert abs(res[c].iter - one.iter) <= 4
        assert rr < 1.5e-3


def test_scene_results_do_not_depend_on_the_batch(dev):
    syss = systems()
    tiles = (256, 512)
    a, ra, _ = solve_batch(dev, [syss[1], syss[0], syss[4]], tiles)
    b, rb, _ = solve_batch(dev, [syss[5], syss[3], syss[1]], tiles)
    assert np.array_equal(a[0], b[2]) and ra[0].iter == rb[2].iter
    # identical scenes, identical solutions
    c, rc, _ = solve_batch(dev, [syss[2]] * 5, tiles)
    for k in range(1, 5):
        assert np.array_equal(c[0], c[k]) and rc[k].iter == rc[0].iter


def test_autotune_keeps_a_candidate_and_still_solves(dev):
    syss = systems()
    xs, res, tuned = solve_batch(dev, syss, autotune=True)
    r1, r2, ms = tuned
    print(f"autotuned tiles: SpMV {r1} rows, vector kernels {r2} rows ({ms:.3f} ms)")
    assert r1 in (128, 256, 512, 1024, 2048) and r2 in (256, 512, 1024, 2048)
    for c, s in enumerate(syss):
        assert res[c].status == 0
        assert rel_resid(s, xs[c]) < 1.5e-3


def test_batch_argument_errors(dev):
    import ctypes
    syss = systems()[:2]
    st, rows = stack_systems(syss)
    t = to_dev(st, dev)
    bad = rows.copy()
    bad[1] += 1  # the first scene's last row now couples into the second
    with pytest.raises(RuntimeError):
        BatchPCG(t["row_offset"], t["col_index"], bad)
    plan = BatchPCG(t["row_offset"], t["col_index"], rows)
    x = torch.zeros_like(t["b"])
    # the one-result entry point refuses a many-scene plan
    lib = load()
    from ppf_amd._lib import PcgResult
    r = PcgResult()
    rc = lib.ppf_pcg_solve(plan._h, ctypes.c_void_p(t["value"].data_ptr()),
                           ctypes.c_void_p(t["cdiag"].data_ptr()),
                           ctypes.c_void_p(t["b"].data_ptr()), ctypes.c_void_p(x.data_ptr()),
                           1e-3, 100, ctypes.byref(r),
                           ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc != 0 and b"ppf_pcg_solve_batch" in lib.ppf_last_error()
    plan.close()
