"""GPU parity of the block-Jacobi PCG path (through the C ABI) against the CPU
oracle, fp64 references and size-independent properties.

Tolerances (fp32 work, different reduction order than the oracle):
  * eigensolve / inverse: element-wise within 1e-5 of the block scale (the
    device acosf/cosf may differ from the host's by an ulp, which the
    trigonometric form amplifies near repeated eigenvalues);
  * SpMV: within 2e-6 of max|y| against the oracle, 1e-5 against fp64;
  * PCG: same status, reported iteration within 4 (the check schedule's
    stride), residual curves within 1e-3 relative over the first 20
    iterations, solution within 1e-2 of max|x|, and the true fp64 L1 residual
    below 1.5 * tol.
"""
import numpy as np
import pytest
import torch

import oracle
from ppf_amd import scenes
from ppf_amd.solver import BlockJacobiPCG, bj_invert, symm3x3

pytestmark = pytest.mark.gpu


def to_dev(s, dev):
    return {k: torch.from_numpy(np.ascontiguousarray(s[k])).to(dev).reshape(-1)
            for k in ("row_offset", "col_index", "value", "cdiag", "b")}


def solve_gpu(s, dev, tol=1e-3, max_iter=10000, x0=None, hist=True):
    t = to_dev(s, dev)
    plan = BlockJacobiPCG(t["row_offset"], t["col_index"])
    x = torch.zeros_like(t["b"]) if x0 is None else torch.from_numpy(x0).to(dev)
    h = torch.zeros(max_iter, dtype=torch.float32, device=dev) if hist else None
    res = plan.solve(t["value"], t["cdiag"], t["b"], x, tol=tol, max_iter=max_iter,
                     err_hist=h)
    torch.cuda.synchronize()
    out = dict(iter=res.iter, resid=res.resid, status=res.status, err0=res.err0,
               x=x.cpu().numpy())
    if hist:
        out["err_hist"] = h.cpu().numpy()[:res.iter]
    plan.close()
    return out


def sym_blocks(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        lam = np.exp(rng.uniform(-6, 6, 3))
        m = q @ np.diag(lam) @ q.T
        out.append(np.ascontiguousarray((0.5 * (m + m.T)).T.reshape(9)))
    out += [np.eye(3).reshape(9), np.diag([1.0, 1, 2]).reshape(9),
            (-5 * np.eye(3)).reshape(9), np.zeros(9)]
    return np.stack(out).astype(np.float32)


def test_symm3x3_matches_oracle(dev):
    a = sym_blocks(4000, 0)
    vg, qg = symm3x3(torch.from_numpy(a).to(dev))
    vg = vg.cpu().numpy().reshape(-1, 3)
    qg = qg.cpu().numpy().reshape(-1, 9)
    vo, qo = oracle.symm3x3(a)
    scale = np.maximum(np.abs(a).max(1), 1e-30)
    assert (np.abs(vg - vo).max(1) / scale).max() < 1e-5
    exact = np.mean(np.all(vg.view(np.uint32) == vo.view(np.uint32), 1))
    print(f"eigenvalues bit-identical for {exact:.1%} of blocks")
    # Eigenvectors are compared through the reconstruction they produce; an
    # ulp of acosf/cosf moves a near-repeated pair's vectors by up to the
    # method's own fp32 resolution (3.4e-4 of scale, measured on the oracle).
    rec_g = np.einsum("kij,kj,klj->kil", qg.reshape(-1, 3, 3), vg, qg.reshape(-1, 3, 3))
    rec_o = np.einsum("kij,kj,klj->kil", qo.reshape(-1, 3, 3), vo, qo.reshape(-1, 3, 3))
    assert (np.abs(rec_g - rec_o).max((1, 2)) / scale).max() < 5e-4


def test_bj_invert_matches_oracle(dev):
    a = sym_blocks(4000, 1)[:-2]  # drop the non-SPD probes
    ig, bad = bj_invert(torch.from_numpy(a).to(dev))
    ig = ig.cpu().numpy()
    io, bad_o = oracle.bj_invert(a)
    assert bad == bad_o == 0
    scale = np.abs(io).max(1)
    err = np.abs(ig - io).max(1) / scale
    assert np.median(err) < 1e-6
    # Against the exact fp64 inverse the device lands in the oracle's own
    # error envelope: a near-repeated eigenvalue pair turns an ulp of
    # acosf/cosf into a rotation of its eigenvectors, which the inverse
    # weights by the eigenvalue gap (blocks here span condition 1e5).
    ex = np.linalg.inv(a.reshape(-1, 3, 3).transpose(0, 2, 1).astype(np.float64))
    ex = ex.transpose(0, 2, 1).reshape(-1, 9)
    exs = np.abs(ex).max(1)
    eg = np.abs(ig - ex).max(1) / exs
    eo = np.abs(io - ex).max(1) / exs
    print(f"vs fp64: gpu p99 {np.percentile(eg, 99):.2e} max {eg.max():.2e}; "
          f"oracle p99 {np.percentile(eo, 99):.2e} max {eo.max():.2e}")
    assert np.percentile(eg, 99) <= 2 * np.percentile(eo, 99) + 1e-6
    assert eg.max() <= 2 * eo.max() + 1e-6
    _, bad2 = bj_invert(torch.from_numpy(sym_blocks(3, 2)[-2:]).to(dev))
    assert bad2 == 2  # -5 I and 0 trap in the reference


@pytest.mark.parametrize("kind", ["grid", "element", "longrows"])
def test_spmv_matches_oracle_and_fp64(dev, kind):
    s = {"grid": lambda: scenes.grid_spd_system(64, seed_matrix=2),
         "element": lambda: scenes.element_spd_system(12, seed=3),
         "longrows": lambda: scenes.long_row_spd_system(seed=4)}[kind]()
    t = to_dev(s, dev)
    plan = BlockJacobiPCG(t["row_offset"], t["col_index"])
    x = np.random.default_rng(5).standard_normal(3 * s["n"]).astype(np.float32)
    y = plan.spmv(t["value"], t["cdiag"], torch.from_numpy(x).to(dev)).cpu().numpy()
    yo = oracle.spmv(s, x)
    M = scenes.sparse_from_upper(s["n"], s["row_offset"], s["col_index"],
                                 s["value"], s["cdiag"])
    y64 = M @ x.astype(np.float64)
    assert np.abs(y - yo).max() <= 2e-6 * np.abs(yo).max()
    assert np.abs(y - y64).max() <= 1e-5 * np.abs(y64).max()
    plan.close()


CASES = [("grid16", lambda: scenes.grid_spd_system(16, seed_matrix=20, seed_rhs=21)),
         ("grid96", lambda: scenes.grid_spd_system(96, seed_matrix=22, seed_rhs=23)),
         ("grid256", lambda: scenes.grid_spd_system(256, seed_matrix=24, seed_rhs=25)),
         ("element14", lambda: scenes.element_spd_system(14, seed=26)),
         # Rows up to ~50 blocks: exercises the CSR walk instead of the ELL form.
         ("longrows", lambda: scenes.long_row_spd_system(seed=27))]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_pcg_matches_oracle(dev, name, make):
    s = make()
    g = solve_gpu(s, dev)
    o = oracle.pcg(s, hist=True)
    print(f"{name}: gpu iter {g['iter']} resid {g['resid']:.4e} | "
          f"oracle iter {o['iter']} resid {o['resid']:.4e}")
    assert g["status"] == o["status"] == 0
    assert abs(g["iter"] - o["iter"]) <= 4
    assert abs(g["err0"] - o["err0"]) <= 1e-5 * o["err0"]
    k = min(20, len(g["err_hist"]), len(o["err_hist"]))
    rel = np.abs(g["err_hist"][:k] - o["err_hist"][:k]) / o["err_hist"][:k]
    assert rel.max() < 1e-3, rel
    assert np.abs(g["x"] - o["x"]).max() < 1e-2 * np.abs(o["x"]).max()
    M = scenes.sparse_from_upper(s["n"], s["row_offset"], s["col_index"],
                                 s["value"], s["cdiag"])
    b = s["b"].astype(np.float64)
    true = np.abs(b - M @ g["x"].astype(np.float64)).sum() / np.abs(b).sum()
    assert true < 1.5e-3


def test_pcg_is_deterministic(dev):
    s = scenes.grid_spd_system(128, seed_matrix=30, seed_rhs=31)
    a = solve_gpu(s, dev)
    b = solve_gpu(s, dev)
    assert a["iter"] == b["iter"]
    assert np.array_equal(a["x"].view(np.uint32), b["x"].view(np.uint32))


def test_pcg_edge_cases(dev):
    s = scenes.grid_spd_system(12, seed_matrix=40)
    # Zero right-hand side: |r0|_1 = 0 returns at once (solver.cu:1029-1031).
    z = dict(s, b=np.zeros_like(s["b"]))
    g = solve_gpu(z, dev)
    assert g["status"] == 0 and g["iter"] == 1 and g["err0"] == 0.0
    assert not np.any(g["x"])
    # Iteration cap.
    g = solve_gpu(s, dev, max_iter=3)
    o = oracle.pcg(s, max_iter=3)
    assert g["status"] == o["status"] == 2 and g["iter"] == o["iter"] == 3
    # Warm start: the rule is relative to |b - A x0|_1.
    x0 = oracle.pcg(s, tol=1e-2)["x"]
    g = solve_gpu(s, dev, x0=x0)
    o = oracle.pcg(s, x0=x0)
    assert g["status"] == o["status"] == 0
    assert abs(g["err0"] - o["err0"]) <= 1e-4 * o["err0"]
    assert abs(g["iter"] - o["iter"]) <= 4
    # One row.
    one = scenes.grid_spd_system(2)
    g = solve_gpu(one, dev)
    assert g["status"] == 0
    # Indefinite operator: reported, not converged.
    bad = dict(s, cdiag=(-50.0 * s["cdiag"]).astype(np.float32))
    g = solve_gpu(bad, dev)
    o = oracle.pcg(bad)
    assert g["status"] in (4, 5) and o["status"] in (4, 5)


def test_plan_repattern_matches_fresh_plan(dev):
    """One plan switched between patterns of the same n -- a chain (short
    rows), medium fans (a wider ELL form), long fans (CSR form, a larger
    pattern) and back -- solves each bit-identically to a plan built for it."""
    order = [scenes.long_row_spd_system(hubs=0, seed=60),
             scenes.long_row_spd_system(hubs=3, fan=10, seed=61),
             scenes.long_row_spd_system(hubs=6, fan=60, seed=62),
             scenes.long_row_spd_system(hubs=0, seed=63)]
    t0 = to_dev(order[0], dev)
    plan = BlockJacobiPCG(t0["row_offset"], t0["col_index"])
    for k, s in enumerate(order):
        t = to_dev(s, dev)
        if k:
            plan.repattern(t["row_offset"], t["col_index"])
        x = torch.zeros_like(t["b"])
        r = plan.solve(t["value"], t["cdiag"], t["b"], x)
        fresh = solve_gpu(s, dev, hist=False)
        print(f"pattern {k}: nnz {len(s['col_index'])} iter {r.iter}/{fresh['iter']}")
        assert r.status == fresh["status"] == 0
        assert r.iter == fresh["iter"]
        assert np.array_equal(x.cpu().numpy().view(np.uint32), fresh["x"].view(np.uint32))
    plan.close()


def permuted_upper(s, order):
    """P A P^T of an upper-stored system on the host: new row k = old row
    order[k]; a block whose rows swap order moves to the other row,
    transposed; columns ascending per row."""
    rp = np.asarray(s["row_offset"], np.int64)
    ci = np.asarray(s["col_index"], np.int64)
    n = len(rp) - 1
    inv = np.empty(n, np.int64)
    inv[order] = np.arange(n)
    rows = np.repeat(np.arange(n), np.diff(rp))
    a, b = inv[rows], inv[ci]
    r, c = np.minimum(a, b), np.maximum(a, b)
    blk = np.asarray(s["value"], np.float32).reshape(-1, 3, 3)
    blk = np.where((b < a)[:, None, None], blk.transpose(0, 2, 1), blk)
    key = np.lexsort((c, r))
    prp = np.zeros(n + 1, np.int64)
    np.add.at(prp, r + 1, 1)
    prp = np.cumsum(prp)
    cd = np.asarray(s["cdiag"], np.float32).reshape(-1, 9)[order]
    bb = np.asarray(s["b"], np.float32).reshape(-1, 3)[order]
    return dict(row_offset=prp.astype(np.int32), col_index=c[key].astype(np.int32),
                value=np.ascontiguousarray(blk[key]).reshape(-1), cdiag=cd.reshape(-1),
                b=bb.reshape(-1))


@pytest.mark.parametrize("kind", ["grid", "longrows"])
def test_plan_row_order(dev, kind):
    """A plan solving in another row order gives, in the original order,
    bitwise the solution of the host-permuted system solved by a plain plan;
    the identity order changes nothing; the order takes effect at the next
    repattern (solving before it is refused)."""
    from ppf_amd._lib import PPFError
    s = (scenes.grid_spd_system(20, seed_matrix=3, seed_rhs=4) if kind == "grid"
         else scenes.long_row_spd_system(n=300, seed=9))
    n = len(s["row_offset"]) - 1
    t = to_dev(s, dev)
    order = np.random.default_rng(5).permutation(n).astype(np.int32)
    plan = BlockJacobiPCG(t["row_offset"], t["col_index"])
    od = torch.from_numpy(order).to(dev)
    plan.set_order(od)
    x = torch.zeros_like(t["b"])
    with pytest.raises(PPFError):
        plan.solve(t["value"], t["cdiag"], t["b"], x)
    plan.repattern(t["row_offset"], t["col_index"])
    r = plan.solve(t["value"], t["cdiag"], t["b"], x)
    y = plan.spmv(t["value"], t["cdiag"], t["b"])
    ps = permuted_upper(s, order)
    pt = to_dev(ps, dev)
    pplan = BlockJacobiPCG(pt["row_offset"], pt["col_index"])
    px = torch.zeros_like(pt["b"])
    pr = pplan.solve(pt["value"], pt["cdiag"], pt["b"], px)
    pplan.close()
    print(f"{kind}: ordered {r.iter} it, host-permuted {pr.iter} it")
    assert r.status == pr.status == 0 and r.iter == pr.iter
    xo = x.cpu().numpy().reshape(-1, 3)
    xp = px.cpu().numpy().reshape(-1, 3)
    assert np.array_equal(xo[order].view(np.uint32), xp.view(np.uint32))
    # the SpMV in the original order: y(b) = M b, the host-permuted plan's
    # product of the permuted b, scattered back
    bp = np.asarray(s["b"], np.float32).reshape(-1, 3)[order].reshape(-1)
    pplan = BlockJacobiPCG(pt["row_offset"], pt["col_index"])
    yp = pplan.spmv(pt["value"], pt["cdiag"], torch.from_numpy(bp).to(dev)).cpu().numpy()
    pplan.close()
    assert np.array_equal(y.cpu().numpy().reshape(-1, 3)[order].view(np.uint32),
                          yp.reshape(-1, 3).view(np.uint32))
    # identity order: bitwise the unordered plan
    plan.set_order(torch.arange(n, dtype=torch.int32, device=dev))
    plan.repattern(t["row_offset"], t["col_index"])
    xi = torch.zeros_like(t["b"])
    ri = plan.solve(t["value"], t["cdiag"], t["b"], xi)
    fresh = solve_gpu(s, dev, hist=False)
    assert ri.iter == fresh["iter"]
    assert np.array_equal(xi.cpu().numpy().view(np.uint32), fresh["x"].view(np.uint32))
    plan.close()


def test_plan_rejects_order_that_is_not_a_permutation(dev):
    """An order with an entry >= n or a repeated entry is refused
    (PPF_ERR_ARG) before the plan uses it; the plan keeps its original order
    and its solve is bitwise the fresh plan's."""
    from ppf_amd._lib import PPFError
    s = scenes.grid_spd_system(12, seed_matrix=3, seed_rhs=4)
    n = len(s["row_offset"]) - 1
    t = to_dev(s, dev)
    plan = BlockJacobiPCG(t["row_offset"], t["col_index"])
    good = np.random.default_rng(2).permutation(n).astype(np.int32)
    dup = good.copy()
    dup[7] = dup[3]
    big = good.copy()
    big[5] = n + 11
    for bad in (dup, big):
        with pytest.raises(PPFError):
            plan.set_order(torch.from_numpy(bad).to(dev))
    plan.repattern(t["row_offset"], t["col_index"])
    x = torch.zeros_like(t["b"])
    r = plan.solve(t["value"], t["cdiag"], t["b"], x)
    fresh = solve_gpu(s, dev, hist=False)
    assert r.status == 0 and r.iter == fresh["iter"]
    assert np.array_equal(x.cpu().numpy().view(np.uint32), fresh["x"].view(np.uint32))
    plan.close()


def test_plan_rejects_malformed_pattern(dev):
    from ppf_amd._lib import PPFError
    ro = torch.tensor([0, 2, 3], dtype=torch.int32, device=dev)
    ci = torch.tensor([1, 0, 1], dtype=torch.int32, device=dev)  # row 0 unsorted
    with pytest.raises(PPFError):
        BlockJacobiPCG(ro, ci)


def test_pcg_full_size_grid_converges(dev):
    """1024x1024 grid (1,048,576 rows): converged, true fp64 residual, and
    deterministic across two solves."""
    s = scenes.grid_spd_system(1024)
    a = solve_gpu(s, dev, hist=False)
    b = solve_gpu(s, dev, hist=False)
    print(f"grid1024: iter {a['iter']} resid {a['resid']:.4e}")
    assert a["status"] == 0 and a["resid"] < 1e-3
    assert np.array_equal(a["x"].view(np.uint32), b["x"].view(np.uint32))
    M = scenes.sparse_from_upper(s["n"], s["row_offset"], s["col_index"],
                                 s["value"], s["cdiag"])
    bb = s["b"].astype(np.float64)
    true = np.abs(bb - M @ a["x"].astype(np.float64)).sum() / np.abs(bb).sum()
    assert true < 1.5e-3


def _contact_system(dev):
    """The last Newton system of a draped sheet in contact (stepper), and the
    stepper's locality row order."""
    from ppf_amd.stepper import Stepper, scene_to_device
    sc = scenes.drape_scene(res=40, height=0.002)
    st = Stepper(scene_to_device(sc, dev))
    p = scenes.step_params()
    for _ in range(4):
        r = st.advance(p)
        assert r["status"] == 0
        p["prev_dt"] = r["dt"]
    s = {k: v.cpu().numpy() for k, v in st.system().items()}
    order = st.solve_order().cpu().numpy().astype(np.int32)
    st.close()
    return s, order


@pytest.mark.parametrize("kind", ["contact", "contact_ordered", "longrows", "grid_split"])
@pytest.mark.parametrize("tiles", [None, 64, 8192])
def test_sym_form_bitwise_equals_split(dev, kind, tiles):
    """The symmetric SpMV form (in-tile mirrors handed over through LDS) is
    bit-identical to the split form: the product, and every iterate of a
    whole solve. Tiles of 64 rows put most mirrors outside their tile; tiles
    of 8,192 rows overflow the LDS places, so the tables' cap is crossed."""
    if kind.startswith("contact"):
        s, order = _contact_system(dev)
    elif kind == "longrows":
        s, order = scenes.long_row_spd_system(n=3000, hubs=20, fan=40, seed=71), None
    else:
        s, order = scenes.grid_spd_system(64, seed_matrix=5, seed_rhs=6), None
    t = to_dev(s, dev)
    plan = BlockJacobiPCG(t["row_offset"], t["col_index"])
    if kind == "contact_ordered":
        plan.set_order(torch.from_numpy(order).to(dev))
        plan.repattern(t["row_offset"], t["col_index"])
    if tiles:
        from ppf_amd.solver import _stream
        assert plan.lib.ppf_pcg_plan_set_tiles(plan._h, tiles, tiles, _stream()) == 0
    out = {}
    for form in ("split", "sym"):
        assert plan.set_form(form) == form
        y = plan.spmv(t["value"], t["cdiag"], t["b"]).cpu().numpy()
        x = torch.zeros_like(t["b"])
        r = plan.solve(t["value"], t["cdiag"], t["b"], x)
        assert r.status == 0
        out[form] = (y, r.iter, r.resid, x.cpu().numpy())
    assert plan.set_form("auto") in ("ell", "sym")
    plan.close()
    (ya, ia, ra, xa), (yb, ib, rb, xb) = out["split"], out["sym"]
    print(f"{kind} tiles {tiles}: n {len(s['row_offset']) - 1} nnz {len(s['col_index'])} "
          f"iterations {ia}/{ib}")
    assert np.array_equal(ya.view(np.uint32), yb.view(np.uint32))
    assert ia == ib and ra == rb
    assert np.array_equal(xa.view(np.uint32), xb.view(np.uint32))
