"""GPU parity of the learned preconditioner (through the C ABI) against the
fp64 oracle (oracle/learned.py), and what training gains over block Jacobi.

Tolerances: the dense layers run in bf16 (8-bit mantissa) with fp32
accumulation, features are rounded to bf16, so the network outputs and
gradients agree with the fp64 oracle to a few percent: outputs within 4% of
their largest magnitude, the G blocks within 2% (the heads' tanh is
bounded), gradients with cosine similarity above 0.98 and norms within 10%.
The factor apply inside PCG is fp32: with zero output heads the solve
matches block Jacobi's iteration count within the check stride.
"""
import numpy as np
import pytest
import torch

import learned as OL
import oracle
from ppf_amd import scenes
from ppf_amd.learned import LearnedPreconditioner, flatten, init_weights
from ppf_amd.solver import BlockJacobiPCG

pytestmark = pytest.mark.gpu


def cloth_system(res, seed, dt=1.0 / 60.0):
    sc = scenes.drape_scene(res=res, height=0.05)
    sh = dict(sc["shell"])
    x = sc["mesh"]["x"].copy()
    rng = np.random.default_rng(seed)
    sh["x"] = x + rng.normal(0, 2e-3, x.shape).astype(np.float32)
    sh["x0"] = x
    sh["target"] = x
    rp, ci = scenes.upper_pattern(sh["nv"], sh["face"], sh["hinge"])
    f, cd, val, _ = oracle.elastic_assemble(sh, rp, ci, dt)
    return dict(n=int(sh["nv"]), row_offset=rp.astype(np.int32), col_index=ci.astype(np.int32),
                value=val.reshape(-1).astype(np.float32), cdiag=cd.astype(np.float32),
                b=f.astype(np.float32))


def to_dev(s, dev):
    return {k: torch.from_numpy(np.ascontiguousarray(s[k])).to(dev).reshape(-1)
            for k in ("row_offset", "col_index", "value", "cdiag", "b")}


def oracle_weights(seed):
    w = OL.init_weights(seed, zero_head=False)
    rng = np.random.default_rng(seed + 100)
    w["Wo"] = rng.normal(0, 0.3, w["Wo"].shape)
    w["Wg"] = rng.normal(0, 0.3, w["Wg"].shape)
    w["co"] = rng.normal(0, 0.1, 6)
    w["cg"] = rng.normal(0, 0.1, 9)
    return w


def slot_to_lower(s, g):
    """For every stored off-diagonal slot, the oracle's lower-edge index."""
    ro, ci = s["row_offset"].astype(np.int64), s["col_index"].astype(np.int64)
    n = len(ro) - 1
    rows = np.repeat(np.arange(n), np.diff(ro))
    low = g["lower"]
    key = {(int(g["dst"][e]), int(g["src"][e])): k for k, e in enumerate(low)}
    out = np.full(len(ci), -1)
    for sl in range(len(ci)):
        if ci[sl] != rows[sl]:
            out[sl] = key[(int(ci[sl]), int(rows[sl]))]
    return out


@pytest.mark.parametrize("kind", ["cloth", "longrows"])
def test_forward_and_factor_match_oracle(dev, kind):
    s = cloth_system(10, 1) if kind == "cloth" else scenes.long_row_spd_system(n=150, seed=3)
    s = {k: (np.asarray(v, np.int32) if k in ("row_offset", "col_index") else v)
         for k, v in s.items()}
    t = to_dev(s, dev)
    w = oracle_weights(5)
    plan = BlockJacobiPCG(t["row_offset"], t["col_index"])
    lpc = LearnedPreconditioner(flatten(w))
    lpc.factor(plan, t["value"], t["cdiag"], t["b"])
    n, nnz = s["n"], len(s["col_index"])
    td, ud, gd, gl = (a.cpu().numpy() for a in lpc.outputs(n, nnz, dev))
    g = OL.system_graph(s)
    x, f = OL.features(g, s["b"])
    to, uo, _ = OL.forward(w, g, x, f)
    m = slot_to_lower(s, g)
    off = m >= 0
    ud = ud.reshape(-1, 9)
    terr = np.abs(td.reshape(-1, 6) - to).max() / np.abs(to).max()
    uerr = np.abs(ud[off] - uo[m[off]]).max() / np.abs(uo).max()
    print(f"{kind}: t rel err {terr:.3e}, u rel err {uerr:.3e}")
    assert terr < 0.04 and uerr < 0.04
    assert np.all(ud[~off] == 0.0)
    Gd, Ge = OL.factor(to, uo)
    Gdiag, Glow = OL.preconditioner(g, Gd, Ge)
    gd = gd.reshape(-1, 3, 3).transpose(0, 2, 1)
    gl = gl.reshape(-1, 3, 3).transpose(0, 2, 1)
    derr = np.abs(gd - Gdiag).max() / np.abs(Gdiag).max()
    lerr = np.abs(gl[off] - Glow[m[off]]).max() / np.abs(Gdiag).max()
    print(f"{kind}: Gd rel err {derr:.3e}, Gl rel err {lerr:.3e}")
    assert derr < 0.02 and lerr < 0.02
    lpc.close()
    plan.close()


def test_zero_heads_solve_like_block_jacobi(dev):
    s = cloth_system(16, 2)
    t = to_dev(s, dev)
    plan = BlockJacobiPCG(t["row_offset"], t["col_index"])
    x = torch.zeros_like(t["b"])
    bj = plan.solve(t["value"], t["cdiag"], t["b"], x)
    lpc = LearnedPreconditioner(init_weights(3))
    lpc.factor(plan, t["value"], t["cdiag"], t["b"])
    x2 = torch.zeros_like(t["b"])
    r = lpc.solve(plan, t["value"], t["cdiag"], t["b"], x2)
    print(f"block Jacobi {bj.iter} iterations, zero-head factor {r.iter}")
    assert r.status == 0 and bj.status == 0
    assert abs(r.iter - bj.iter) <= 4
    # the factor path then goes back to block Jacobi on the same plan
    x3 = torch.zeros_like(t["b"])
    bj2 = plan.solve(t["value"], t["cdiag"], t["b"], x3)
    assert bj2.iter == bj.iter
    assert torch.equal(x3, x)
    lpc.close()
    plan.close()


def test_gradient_matches_oracle(dev):
    s = cloth_system(10, 4)
    s = {k: (np.asarray(v, np.int32) if k in ("row_offset", "col_index") else v)
         for k, v in s.items()}
    t = to_dev(s, dev)
    w = oracle_weights(7)
    plan = BlockJacobiPCG(t["row_offset"], t["col_index"])
    lpc = LearnedPreconditioner(flatten(w))
    lpc.factor(plan, t["value"], t["cdiag"], t["b"])
    z = OL.probe_vectors(s["n"], 4, seed=11)
    zt = torch.from_numpy(z.astype(np.float32)).to(dev).contiguous()
    loss, grad = lpc.gradient(plan, t["value"], t["cdiag"], zt)
    g = OL.system_graph(s)
    x, f = OL.features(g, s["b"])
    lo, go = OL.backward(w, g, x, f, z)
    go = flatten(go).astype(np.float64)
    cos = grad @ go / (np.linalg.norm(grad) * np.linalg.norm(go))
    ratio = np.linalg.norm(grad) / np.linalg.norm(go)
    print(f"loss gpu {loss:.6f} oracle {lo:.6f}; gradient cosine {cos:.5f}, norm ratio {ratio:.4f}")
    assert abs(loss - lo) < 0.02 * abs(lo) + 1e-3
    assert cos > 0.98 and abs(ratio - 1.0) < 0.1
    # per block of weights, too (heads and the first layer included)
    from ppf_amd.learned import ORDER, SHAPES
    o = 0
    for k in ORDER:
        size = int(np.prod(SHAPES[k]))
        a, b = grad[o:o + size], go[o:o + size]
        if np.linalg.norm(b) > 1e-6 * np.linalg.norm(go):
            c = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
            assert c > 0.95, (k, c)
        o += size
    lpc.close()
    plan.close()


def test_training_beats_block_jacobi_on_unseen_systems(dev):
    """Adam on the device's own forward/backward kernels over a few cloth
    Newton systems; on unseen systems (another size and seed) the trained
    factor needs fewer PCG iterations than block Jacobi to 1e-3."""
    train = [to_dev(cloth_system(16, k), dev) for k in range(4)]
    plans = [BlockJacobiPCG(t["row_offset"], t["col_index"]) for t in train]
    lpc = LearnedPreconditioner(init_weights(0))
    gen = torch.Generator(device=dev).manual_seed(0)
    first = last = None
    for step in range(1, 301):
        k = step % len(train)
        t, plan = train[k], plans[k]
        n = t["b"].numel() // 3
        lpc.factor(plan, t["value"], t["cdiag"], t["b"])
        z = (torch.randint(0, 2, (4, n, 3), generator=gen, device=dev).float() * 2 - 1).contiguous()
        loss, _ = lpc.gradient(plan, t["value"], t["cdiag"], z)
        lpc.adam(2e-3, step)
        first = loss if first is None else first
        last = loss
    print(f"training loss {first:.4f} -> {last:.4f}")
    assert last < first
    for seed in (100, 101):
        t = to_dev(cloth_system(24, seed), dev)
        plan = BlockJacobiPCG(t["row_offset"], t["col_index"])
        x = torch.zeros_like(t["b"])
        bj = plan.solve(t["value"], t["cdiag"], t["b"], x)
        lpc.factor(plan, t["value"], t["cdiag"], t["b"])
        x2 = torch.zeros_like(t["b"])
        r = lpc.solve(plan, t["value"], t["cdiag"], t["b"], x2)
        print(f"unseen system {seed}: block Jacobi {bj.iter} iterations, learned {r.iter}")
        assert r.status == 0 and bj.status == 0
        assert r.iter <= 0.8 * bj.iter
        plan.close()
    lpc.close()
    for p in plans:
        p.close()


def contact_system(dev, res=12, layers=2, steps=4, seed=3):
    """The last Newton system of a small stacked drape in contact (sheets on
    the spinning cylinder, Dirichlet-eliminated pins): contact blocks beside
    the elastic ones, as on the stacked-layer scene."""
    from ppf_amd.stepper import Stepper, scene_to_device
    sc = scenes.stacked_drape_scene(res=res, layers=layers, seed=seed, n_around=24, n_along=8,
                                    clearance=1.2e-3)
    ns = layers * res * res
    t = scene_to_device(sc, dev)
    t["fix_delta"] = torch.zeros_like(t["mesh"]["x"])
    st = Stepper(t)
    p = scenes.step_params()
    clock, pairs = 0.0, 0
    for _ in range(steps):
        now = scenes.spin_fix_positions(sc, clock)[ns:].copy()
        aim = scenes.spin_fix_positions(sc, clock + p["dt"])[ns:].copy()
        t["fix_pos"].view(-1, 3)[ns:] = torch.from_numpy(aim).to(dev)
        t["fix_delta"].view(-1, 3)[ns:] = torch.from_numpy(aim - now).to(dev)
        r = st.advance(p)
        assert r["status"] == 0, r
        pairs = r["pairs_max"]
        p["prev_dt"] = r["dt"]
        clock += r["dt"]
    s = st.system()
    st.close()
    h = {k: v.cpu().numpy() for k, v in s.items()}
    h["n"] = len(h["row_offset"]) - 1
    h["cdiag"] = h["cdiag"].reshape(-1, 9)
    return h, pairs


def test_forward_and_gradient_on_contact_system(dev):
    """Forward outputs, factor and loss gradient on a stacked-drape contact
    system (contact and Dirichlet rows beside the elastic ones) against the
    fp64 oracle, to the same bars as on the cloth systems."""
    s, pairs = contact_system(dev)
    assert pairs > 0
    s = {k: (np.asarray(v, np.int32) if k in ("row_offset", "col_index") else v)
         for k, v in s.items()}
    t = to_dev(s, dev)
    w = oracle_weights(9)
    plan = BlockJacobiPCG(t["row_offset"], t["col_index"])
    lpc = LearnedPreconditioner(flatten(w))
    lpc.factor(plan, t["value"], t["cdiag"], t["b"])
    n, nnz = s["n"], len(s["col_index"])
    td, ud, gd, gl = (a.cpu().numpy() for a in lpc.outputs(n, nnz, dev))
    g = OL.system_graph(s)
    x, f = OL.features(g, s["b"])
    to, uo, _ = OL.forward(w, g, x, f)
    m = slot_to_lower(s, g)
    off = m >= 0
    ud = ud.reshape(-1, 9)
    terr = np.abs(td.reshape(-1, 6) - to).max() / np.abs(to).max()
    uerr = np.abs(ud[off] - uo[m[off]]).max() / np.abs(uo).max()
    Gd, Ge = OL.factor(to, uo)
    Gdiag, Glow = OL.preconditioner(g, Gd, Ge)
    gd = gd.reshape(-1, 3, 3).transpose(0, 2, 1)
    gl = gl.reshape(-1, 3, 3).transpose(0, 2, 1)
    derr = np.abs(gd - Gdiag).max() / np.abs(Gdiag).max()
    lerr = np.abs(gl[off] - Glow[m[off]]).max() / np.abs(Gdiag).max()
    z = OL.probe_vectors(n, 4, seed=12)
    zt = torch.from_numpy(z.astype(np.float32)).to(dev).contiguous()
    loss, grad = lpc.gradient(plan, t["value"], t["cdiag"], zt)
    lo, go = OL.backward(w, g, x, f, z)
    go = flatten(go).astype(np.float64)
    cos = grad @ go / (np.linalg.norm(grad) * np.linalg.norm(go))
    ratio = np.linalg.norm(grad) / np.linalg.norm(go)
    print(f"contact system: n {n}, {nnz} blocks, {pairs} pairs; t {terr:.2e} u {uerr:.2e} "
          f"Gd {derr:.2e} Gl {lerr:.2e}; loss {loss:.5f}/{lo:.5f} grad cos {cos:.5f} "
          f"norm ratio {ratio:.4f}")
    assert terr < 0.04 and uerr < 0.04
    assert derr < 0.02 and lerr < 0.02
    assert abs(loss - lo) < 0.02 * abs(lo) + 1e-3
    assert cos > 0.98 and abs(ratio - 1.0) < 0.1
    lpc.close()
    plan.close()


@pytest.mark.timeout(600)
def test_training_on_stacked_contact_systems_unseen_seeds(dev):
    """Trained on the Newton systems the block-Jacobi solver produces on the
    stacked drape (8 sheets of 64 x 64 on the spinning cylinder, seeds 0-2),
    the learned factor needs fewer PCG iterations than block Jacobi on the
    contact systems of unseen seeds (4 and 5)."""
    from ppf_amd import lpc_train
    train = lpc_train.collect_drape_systems(dev, res=64, layers=8, seeds=(0, 1, 2), steps=12)
    lpc, hist = lpc_train.train(train, steps=800, lr=2e-3)
    del train
    print(f"training loss {np.mean(hist[:20]):.4f} -> {np.mean(hist[-20:]):.4f}")
    assert np.mean(hist[-20:]) < np.mean(hist[:20])
    test = lpc_train.collect_drape_systems(dev, res=64, layers=8, seeds=(4, 5), steps=12,
                                           every=3)
    rows = lpc_train.compare(lpc, test)
    for r in rows:
        print(f"unseen: block Jacobi {r['bj_iter']} iterations, learned {r['lpc_iter']}")
        assert r["bj_status"] == 0 and r["lpc_status"] == 0
    ratio = sum(r["lpc_iter"] for r in rows) / sum(r["bj_iter"] for r in rows)
    print(f"learned / block-Jacobi iterations on unseen seeds: {ratio:.3f}")
    assert ratio < 0.8
    lpc.close()
