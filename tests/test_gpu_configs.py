"""The configurations as written, on the GPU through the C ABI, against the
oracle and against the properties the mathematics fixes.

- Plain-method drape (64 x 64 cloth, edge 1/63 m, 1,280-triangle icosphere of
  radius 0.3, gap 1e-3, friction 0.5, dt 1/60, 60 steps, block-Jacobi PCG to
  1e-3): every step from the GPU's state against the oracle's step from the
  same state -- status, Newton count and zero intersections on both, dt and
  contact count, positions. Beside it, on every step in contact, the fp32
  solve of that step's last Newton system against an fp64 direct solve of
  the same system, for the GPU's solver and the oracle's: the two fp32
  solvers stop at different points of the 1e-3 tolerance ball, and the
  GPU/oracle position gap is of the size of that ball.
- Stacked-layer drape at full size (8 sheets of 256 x 256, 527,360 vertices,
  cylinder spinning at 2 rad/s): two time steps in contact against the
  oracle, and 24 time steps of properties (status, zero intersections by the
  intersection check, no cloth vertex inside the cylinder's inscribed radius,
  finite positions, contact pairs present).
- 64 uneven draped sheets (sides log-uniform in 32..384, sphere radius
  0.1..0.4 m, seed 5), their last Newton systems of time step 3 solved in one
  batched plan against one plan per scene, whose solves are pinned to the
  oracle on the smallest scenes.
"""
import copy

import numpy as np
import pytest
import torch

import oracle
from ppf_amd import scenes
from ppf_amd.solver import BatchPCG, BlockJacobiPCG, stack_systems
from ppf_amd.stepper import Stepper, scene_to_device
from test_gpu_step import (SOLVE_BALL_FACTOR, solve_error_vs_fp64, ulp_nearest,
                           within_ulp_copies)

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(900)
def test_plain_drape_60_steps(dev):
    sc = scenes.drape_scene()  # 64 x 64 at 0.05 m above the sphere
    sco = copy.deepcopy(sc)
    t = scene_to_device(sc, dev)
    gpu = Stepper(t)
    orc = oracle.OracleStepper(sco)
    p = scenes.step_params()
    worst = worst_ratio = worst_solve = 0.0
    contact_steps = within = 0
    for k in range(60):
        xg = t["mesh"]["x"].cpu().numpy().reshape(-1, 3)
        pg = t["prev"].cpu().numpy().reshape(-1, 3)
        sco["mesh"]["x"][:] = xg
        sco["prev"][:] = pg
        rg = gpu.advance(p)
        ro = orc.advance(p)
        xn = t["mesh"]["x"].cpu().numpy().reshape(-1, 3)
        xo = sco["mesh"]["x"]
        scale = np.abs(xo).max()
        err = np.abs(xn - xo).max() / scale
        disp = np.abs(xo - xg).max()
        eg, eo = solve_error_vs_fp64(gpu.system(), p["cg_tol"], scale)
        bar = max(1e-5, SOLVE_BALL_FACTOR * max(eg, eo))
        # a step on a discontinuity of the algorithm: compared with the
        # oracle's steps from one-ulp copies of the start state
        near = ulp_nearest(orc, sco, xg, pg, p, xn, scale) if err >= bar else (err, 0.0)
        line = (f"step {k}: status {rg['status']}/{ro['status']} newton {rg['newton_steps']}/"
                f"{ro['newton_steps']} pcg {rg['pcg_iter_total']}/{ro['pcg_iter_total']} "
                f"pairs {rg['pairs_max']}/{ro['pairs_max']} isect {rg['intersections']}/"
                f"{ro['intersections']} dt {rg['dt']:.6g}/{ro['dt']:.6g} pos err {err:.2e} "
                f"(disp {disp / scale:.2e}); last Newton solve vs fp64: gpu {eg:.2e} oracle "
                f"{eo:.2e}; bar {bar:.2e}; nearest one-ulp copy {near[0]:.2e} (spread {near[1]:.2e})")
        print(line, flush=True)
        assert rg["status"] == ro["status"] == 0, line
        assert rg["intersections"] == ro["intersections"] == 0, line
        assert rg["newton_steps"] == ro["newton_steps"], line
        assert abs(rg["dt"] - ro["dt"]) <= 1e-3 * ro["dt"], line
        assert abs(rg["pairs_max"] - ro["pairs_max"]) <= 0.02 * ro["pairs_max"], line
        # The two fp32 solves are equally far from the exact solution: the
        # GPU's solver is no less accurate than the oracle's.
        assert eg < 3.0 * eo + 1e-7, line
        assert err < bar or within_ulp_copies(err, near, bar), line
        contact_steps += rg["pairs_max"] > 0
        within += err < 1e-5
        worst = max(worst, min(err, near[0]))
        worst_ratio = max(worst_ratio, min(err, near[0]) / bar)
        worst_solve = max(worst_solve, eg, eo)
        p["prev_dt"] = rg["dt"]
    gpu.close()
    print(f"worst position gap {worst:.2e} ({worst_ratio:.2f} of its bar); {within} of 60 "
          f"steps within 1e-5; worst fp32-vs-fp64 solve distance {worst_solve:.2e}; "
          f"{contact_steps} steps in contact")
    assert contact_steps >= 40


def _spin(sc, t, ns, clock, dt, dev, sco=None, orc=None):
    now = scenes.spin_fix_positions(sc, clock)[ns:].copy()
    aim = scenes.spin_fix_positions(sc, clock + dt)[ns:].copy()
    t["fix_pos"].view(-1, 3)[ns:] = torch.from_numpy(aim).to(dev)
    t["fix_delta"].view(-1, 3)[ns:] = torch.from_numpy(aim - now).to(dev)
    if sco is not None:
        sco["fix_pos"][ns:] = aim
        sco["fix_delta"][ns:] = aim - now
        orc.delta[:] = sco["fix_delta"].reshape(-1)


@pytest.mark.timeout(1200)
def test_stacked_drape_full_size(dev):
    res, layers = 256, 8
    sc = scenes.stacked_drape_scene(res=res, layers=layers)
    ns = layers * res * res
    sc["fix_delta"] = np.zeros_like(sc["fix_pos"])
    t = scene_to_device(sc, dev)
    gpu = Stepper(t)
    p = scenes.step_params()
    radius = 0.2
    inscribed = radius * np.cos(np.pi / 96)  # the cylinder is a 96-gon
    clock = 0.0
    compare_at = (6, 7)
    sco = orc = None
    pairs_seen = 0
    for k in range(24):
        if k in compare_at:
            if orc is None:
                sco = copy.deepcopy(sc)
                orc = oracle.OracleStepper(sco)
            sco["mesh"]["x"][:] = t["mesh"]["x"].cpu().numpy().reshape(-1, 3)
            sco["prev"][:] = t["prev"].cpu().numpy().reshape(-1, 3)
            _spin(sc, t, ns, clock, p["dt"], dev, sco, orc)
        else:
            _spin(sc, t, ns, clock, p["dt"], dev)
        xg0 = t["mesh"]["x"].cpu().numpy().reshape(-1, 3)
        rg = gpu.advance(p)
        xn = t["mesh"]["x"].cpu().numpy().reshape(-1, 3)
        line = (f"step {k}: status {rg['status']} newton {rg['newton_steps']} pcg "
                f"{rg['pcg_iter_total']} pairs {rg['pairs_max']} isect {rg['intersections']} "
                f"toi {rg['toi_advanced']:.4f} ms {rg['ms_total']:.1f}")
        assert rg["status"] == 0 and rg["intersections"] == 0, line
        assert np.all(np.isfinite(xn)), line
        rad = np.hypot(xn[:ns, 0], xn[:ns, 1])
        assert rad.min() >= inscribed - 1e-6, (line, rad.min())
        pairs_seen = max(pairs_seen, rg["pairs_max"])
        if k in compare_at:
            ro = orc.advance(p)
            xo = sco["mesh"]["x"]
            err = np.abs(xn - xo).max() / np.abs(xo).max()
            line += (f"; oracle newton {ro['newton_steps']} pcg {ro['pcg_iter_total']} pairs "
                     f"{ro['pairs_max']} isect {ro['intersections']} toi "
                     f"{ro['toi_advanced']:.4f}; pos err {err:.2e} (disp "
                     f"{np.abs(xo - xg0).max() / np.abs(xo).max():.2e})")
            print(line, flush=True)
            assert ro["status"] == 0 and ro["intersections"] == 0, line
            assert rg["newton_steps"] == ro["newton_steps"], line
            assert abs(rg["dt"] - ro["dt"]) <= 1e-2 * ro["dt"], line
            assert abs(rg["pairs_max"] - ro["pairs_max"]) <= 0.01 * ro["pairs_max"], line
            assert err < 5e-4, line
        else:
            print(line, flush=True)
        p["prev_dt"] = rg["dt"]
        clock += rg["dt"]
    gpu.close()
    assert pairs_seen > 1_000_000


@pytest.mark.timeout(900)
def test_batch_of_64_uneven_scenes(dev):
    rng = np.random.default_rng(5)
    count = 64
    sides = np.exp(rng.uniform(np.log(32), np.log(384), count)).astype(int)
    radii = rng.uniform(0.1, 0.4, count)
    systems = []
    for side, rad in zip(sides, radii):
        sc = scenes.drape_scene(res=int(side), radius=float(rad), height=0.002)
        st = Stepper(scene_to_device(sc, dev))
        p = scenes.step_params()
        for _ in range(3):
            r = st.advance(p)
            assert r["status"] == 0, (side, r)
            p["prev_dt"] = r["dt"]
        systems.append(st.system())
        st.close()
    stacked, rows = stack_systems(systems)
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in stacked.items()}
    plan = BatchPCG(t["row_offset"], t["col_index"], rows)
    r1, r2, _ = plan.autotune(t["value"], t["cdiag"], t["b"])
    x = torch.zeros_like(t["b"])
    bat = plan.solve(t["value"], t["cdiag"], t["b"], x)
    plan.close()
    xb = x.cpu().numpy()
    print(f"64 scenes, {int(rows[-1])} rows, tiles {r1}/{r2}, batched iterations max "
          f"{max(r.iter for r in bat)} sum {sum(r.iter for r in bat)}")
    order = np.argsort(sides)
    worst = 0.0
    for c, s in enumerate(systems):
        pl = BlockJacobiPCG(s["row_offset"], s["col_index"])
        xs = torch.zeros_like(s["b"])
        one = pl.solve(s["value"], s["cdiag"], s["b"], xs)
        pl.close()
        assert bat[c].status == 0 and one.status == 0
        # the batch's tiles sum in another order than the single plan's: over
        # ~1,800 fp32 iterations the stopping check can land a few checks apart
        assert abs(bat[c].iter - one.iter) <= max(4, 0.005 * one.iter), (c, bat[c].iter,
                                                                          one.iter)
        h = {k: v.cpu().numpy() for k, v in s.items()}
        n = len(h["row_offset"]) - 1
        M = scenes.sparse_from_upper(n, h["row_offset"], h["col_index"], h["value"], h["cdiag"])
        xc = xb[3 * int(rows[c]):3 * int(rows[c + 1])].astype(np.float64)
        b = h["b"].astype(np.float64)
        # True (fp64) L1 residuals of the batched and the single solution.
        # The solver stops on its recurrence residual; over thousands of fp32
        # iterations the true residual drifts from it alike in both solves.
        resid = np.abs(b - M @ xc).sum() / np.abs(b).sum()
        rs = np.abs(b - M @ xs.cpu().numpy().astype(np.float64)).sum() / np.abs(b).sum()
        worst = max(worst, resid / max(rs, 1e-3))
        assert resid < max(1.5e-3, 1.5 * rs), (c, resid, rs, one.iter)
        if c in order[:3]:  # the smallest scenes: the single solve against the oracle
            ob = oracle.pcg({**h, "n": n}, tol=1e-3)
            xo = ob["x"]
            rel = np.abs(xs.cpu().numpy() - xo).max() / np.abs(xo).max()
            print(f"scene {c} (side {sides[c]}): gpu {one.iter} oracle {ob['iter']} "
                  f"iterations, max rel dx {rel:.2e}")
            assert abs(one.iter - ob["iter"]) <= 4 and rel < 1e-2
    print(f"worst batched / single true-residual ratio {worst:.3f}")
