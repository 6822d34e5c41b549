"""GPU time step (through the C ABI) against the CPU oracle's step, started
from the same state every step.

The stages are bit-exact or ulp-close individually, but the PCG sums in a
different order than the oracle's, so its iterates (and everything after
them) agree only to the solver tolerance. The bar is the one stated for
positions: per step, the GPU's positions agree with the oracle's to 1e-5 of
the coordinate scale (max |x|); the error relative to the step's own
displacement is printed as well. The fp32 solves' distance from an fp64
re-solve of the step's Newton system widens it on a stiff step
(SOLVE_BALL_FACTOR). A step that amplifies ulp-level differences (one-ulp
changes of its start state move its outcome 100x further, continuously or
across a discontinuity) is compared with the oracle's steps from one-ulp
copies of the start state (ulp_nearest, within_ulp_copies).
The stacked sheets on the spinning cylinder
(400+ contacts per step) are held to 5e-4, dt and toi to 1%: there the fp32
solve's own tolerance ball is wider than 1e-5 and can tip a contact at the
edge of the gap (the test states the measured spread).
Newton iteration counts and the status must agree, the taken dt to 1e-3
relative (it is dt0 times the accumulated CCD toi, whose probe loop
quantizes).
"""
import copy

import numpy as np
import pytest
import torch

import oracle
from ppf_amd import scenes
from ppf_amd.stepper import Stepper, scene_to_device

pytestmark = pytest.mark.gpu


def solve_error_vs_fp64(sysd, tol, scale):
    """|x32 - x64|_inf / scale of the GPU's fp32 PCG solve and of the oracle's
    fp32 PCG solve of one Newton system (x64: an fp64 direct solve of the same
    system): how far each fp32 solver's answer lies from the exact one."""
    import scipy.sparse.linalg as sla
    from ppf_amd.solver import BlockJacobiPCG
    h = {k: v.cpu().numpy() for k, v in sysd.items()}
    n = len(h["row_offset"]) - 1
    M = scenes.sparse_from_upper(n, h["row_offset"], h["col_index"], h["value"],
                                 h["cdiag"]).tocsc()
    exact = sla.spsolve(M, h["b"].astype(np.float64))
    plan = BlockJacobiPCG(sysd["row_offset"], sysd["col_index"])
    x = torch.zeros_like(sysd["b"])
    r = plan.solve(sysd["value"], sysd["cdiag"], sysd["b"], x, tol=tol)
    plan.close()
    assert r.status == 0
    ob = oracle.pcg({**h, "n": n}, tol=tol)
    eg = np.abs(x.cpu().numpy().astype(np.float64) - exact).max() / scale
    eo = np.abs(ob["x"].astype(np.float64) - exact).max() / scale
    return eg, eo


# Per-step position bar: the two fp32 solvers (different summation orders)
# stop at different points of the tolerance ball of each Newton system. Each
# point lies within e of the exact (fp64) solution, e measured per step on the
# step's last Newton system; two of them within 2e, over up to three Newton
# iterations 6e. The bar is max(floor, 6 e).
SOLVE_BALL_FACTOR = 6.0
# A step can sit on a discontinuity of the algorithm itself: measured on the
# oracle alone, 2 of 11 copies of the drape's step 7 whose start coordinates
# moved by one ulp each (random direction) put vertex 63 4.2e-5 of the
# coordinate scale away from where the other 9 put it (same contact count,
# PCG counts and step sizes). The two implementations differ at the ulp level
# (summation orders, device math), so the GPU may land on either side. On
# C1's step 3 the copies spread continuously instead, up to 3.2e-5. Such a
# step passes if the GPU's positions are within the bar of the oracle's step
# from SOME start state within one ulp of the same state, or no further from
# the oracle's positions than the farthest of those steps (ULP_TRIALS seeded
# copies): the GPU's gap is then one that one-ulp input changes make in the
# oracle itself.
ULP_TRIALS = 16


def within_ulp_copies(err, near, bar):
    """The verdict of ulp_nearest's (nearest, spread) for a step whose gap
    `err` exceeds `bar`."""
    return near[0] < bar or err <= near[1]


def ulp_nearest(orc, sco, x_start, prev_start, params, x_gpu, scale, fix_start=None,
                trials=ULP_TRIALS, seed=0):
    """Step the oracle from `trials` copies of the start state whose every
    coordinate moved by one ulp in a random direction. Returns (the smallest
    |x_gpu - x'|_inf / scale over the copies, the largest |x' - x|_inf / scale,
    x the oracle's unperturbed result). The unperturbed result, post-step prev
    and pins are restored in sco afterwards."""
    rng = np.random.default_rng(seed)
    x_ref = sco["mesh"]["x"].copy()
    prev_after = sco["prev"].copy()
    fix_after = sco["fix_pos"].copy() if sco.get("fix_pos") is not None else None
    nearest, spread = np.inf, 0.0
    for _ in range(trials):
        toward = np.where(rng.random(x_start.shape) < 0.5, np.inf, -np.inf).astype(np.float32)
        sco["mesh"]["x"][:] = np.nextafter(np.asarray(x_start, np.float32), toward)
        sco["prev"][:] = prev_start
        if fix_start is not None:
            sco["fix_pos"][:] = fix_start
        orc.advance(dict(params))
        xp = sco["mesh"]["x"]
        nearest = min(nearest, float(np.abs(x_gpu - xp).max() / scale))
        spread = max(spread, float(np.abs(xp - x_ref).max() / scale))
    sco["mesh"]["x"][:] = x_ref
    sco["prev"][:] = prev_after
    if fix_after is not None:
        sco["fix_pos"][:] = fix_after
    return nearest, spread


def _compare(dev, sc, steps, params, bar=1e-5):
    """Every step from the GPU's state against the oracle's step from the
    same state. `bar` is the floor of the per-step position bar (relative to
    the coordinate scale); the bar widens to 6x the fp32 solves' measured
    distance from an fp64 re-solve of that step's last Newton system; a step
    beyond it is compared with the oracle's steps from one-ulp copies of the
    start state (ulp_nearest, within_ulp_copies)."""
    sco = copy.deepcopy(sc)
    t = scene_to_device(sc, dev)
    gpu = Stepper(t)
    orc = oracle.OracleStepper(sco)
    p = dict(params)
    worst = worst_ratio = 0.0
    for k in range(steps):
        xg = t["mesh"]["x"].cpu().numpy().reshape(-1, 3)
        pg = t["prev"].cpu().numpy().reshape(-1, 3)
        sco["mesh"]["x"][:] = xg
        sco["prev"][:] = pg
        fix0 = sco["fix_pos"].copy() if sco.get("fix_pos") is not None else None
        rg = gpu.advance(p)
        ro = orc.advance(p)
        xn = t["mesh"]["x"].cpu().numpy().reshape(-1, 3)
        xo = sco["mesh"]["x"]
        scale = np.abs(xo).max()
        err = np.abs(xn - xo).max() / scale
        disp = np.abs(xo - xg).max()
        rel_disp = np.abs(xn - xo).max() / max(disp, 1e-30)
        eg, eo = solve_error_vs_fp64(gpu.system(), p["cg_tol"], scale)
        step_bar = max(bar, SOLVE_BALL_FACTOR * max(eg, eo))
        near = None
        if err >= step_bar:
            near = ulp_nearest(orc, sco, xg, pg, p, xn, scale, fix0)
        print(f"step {k}: status {rg['status']}/{ro['status']} newton "
              f"{rg['newton_steps']}/{ro['newton_steps']} pcg {rg['pcg_iter_total']}/"
              f"{ro['pcg_iter_total']} pairs {rg['pairs_max']}/{ro['pairs_max']} dt "
              f"{rg['dt']:.6g}/{ro['dt']:.6g} pos err {err:.2e} (of disp {rel_disp:.2e}); "
              f"fp32 solves vs fp64 {eg:.2e}/{eo:.2e}, bar {step_bar:.2e}"
              + ("" if near is None else f"; one-ulp copies: nearest {near[0]:.2e}, "
                 f"spread {near[1]:.2e}") + f" ms {rg['ms_total']:.2f}", flush=True)
        assert rg["status"] == ro["status"] == 0, (rg, ro)
        assert rg["newton_steps"] == ro["newton_steps"]
        # dt = dt0 * accumulated toi: the ACCD stops at quantized probe times,
        # so ulp-level input differences can move it by a few 1e-5 relative
        assert abs(rg["dt"] - ro["dt"]) <= 1e-3 * ro["dt"]
        assert abs(rg["pairs_max"] - ro["pairs_max"]) <= 0.02 * ro["pairs_max"]
        if near is not None:
            # beyond the bar: within what one-ulp copies of the start state
            # make of the oracle's own step
            assert within_ulp_copies(err, near, step_bar), (err, near)
            err = min(err, near[0])
        worst = max(worst, err)
        worst_ratio = max(worst_ratio, err / step_bar)
        p["prev_dt"] = rg["dt"]
    gpu.close()
    print(f"worst position gap {worst:.2e}, {worst_ratio:.2f} of its bar")
    return worst


@pytest.mark.parametrize("cg_tol,bar", [(1e-3, 1e-5), (1e-5, 1e-5)])
def test_drape_small(dev, cg_tol, bar):
    # Floor 1e-5 of the coordinate scale; a step whose Newton systems are
    # stiff enough that the fp32 solves sit further than that from the fp64
    # answer gets 6x that distance. At cg_tol 1e-5 step 7 lands 2.2e-5 from
    # the oracle with both solves within 3.1e-7 of fp64: that step sits on a
    # discontinuity (ulp_nearest).
    sc = scenes.drape_scene(res=16, height=0.003)
    _compare(dev, sc, 14, scenes.step_params(cg_tol=cg_tol), bar)


def test_drape_pinned_corners(dev):
    res = 12
    sc = scenes.drape_scene(res=res, height=0.004, pins=[0, res - 1, res * res - 1])
    fix = sc["fix_pos"].copy()
    _compare(dev, sc, 10, scenes.step_params())
    t = scene_to_device(sc, dev)
    gpu = Stepper(t)
    p = scenes.step_params()
    for _ in range(5):
        r = gpu.advance(p)
        assert r["status"] == 0
        p["prev_dt"] = r["dt"]
    x = t["mesh"]["x"].cpu().numpy().reshape(-1, 3)
    pins = [0, res - 1, res * res - 1]
    assert np.array_equal(x[pins], fix[pins])
    gpu.close()


def test_drape_c1_resolution(dev):
    # The plain-method drape at its full 64 x 64 resolution, close to the
    # sphere so contact starts within the first steps; 500-900 PCG iterations
    # per contact system. Bar: 1e-5 of the coordinate scale, or 6x the fp32
    # solves' distance from the fp64 answer on the stiffer steps.
    sc = scenes.drape_scene(res=64, height=0.002)
    _compare(dev, sc, 6, scenes.step_params(cg_tol=1e-5), bar=1e-5)


def test_newton_solve_error_vs_fp64(dev):
    """The fp32 PCG solution of a contact Newton system against an fp64
    direct solve of the same system: stated, and bounded by what the L1
    residual rule allows (error <= condition * tolerance)."""
    import scipy.sparse.linalg as sla
    from ppf_amd.solver import BlockJacobiPCG
    sc = scenes.drape_scene(res=32, height=0.002)
    t = scene_to_device(sc, dev)
    gpu = Stepper(t)
    p = scenes.step_params()
    for _ in range(4):
        r = gpu.advance(p)
        assert r["status"] == 0
        p["prev_dt"] = r["dt"]
    s = gpu.system()
    n = s["row_offset"].numel() - 1
    host = {k: v.cpu().numpy() for k, v in s.items()}
    M = scenes.sparse_from_upper(n, host["row_offset"], host["col_index"], host["value"],
                                 host["cdiag"]).tocsc()
    b = host["b"].astype(np.float64)
    exact = sla.spsolve(M, b)
    out = []
    for tol in (1e-3, 1e-5):
        plan = BlockJacobiPCG(s["row_offset"], s["col_index"])
        x = torch.zeros_like(s["b"])
        res = plan.solve(s["value"], s["cdiag"], s["b"], x, tol=tol)
        plan.close()
        xg = x.cpu().numpy().astype(np.float64)
        err = np.linalg.norm(xg - exact) / np.linalg.norm(exact)
        ob = oracle.pcg({**host, "n": n}, tol=tol)
        oerr = np.linalg.norm(ob["x"] - exact) / np.linalg.norm(exact)
        resid = np.abs(b - M @ xg).sum() / np.abs(b).sum()
        print(f"tol {tol:g}: gpu iter {res.iter} rel err vs fp64 {err:.3e} (L1 resid "
              f"{resid:.2e}); oracle iter {ob['iter']} rel err {oerr:.3e}")
        assert res.status == 0 and resid < 1.5 * tol
        out.append(err)
    # a tighter tolerance moves the fp32 solution towards the fp64 one
    assert out[1] < out[0] and out[1] < 1e-3
    gpu.close()


def test_last_newton_system_is_a_solvable_dirichlet_system(dev):
    from ppf_amd.solver import BlockJacobiPCG
    res = 12
    pins = [0, res - 1]
    sc = scenes.drape_scene(res=res, height=0.004, pins=pins)
    t = scene_to_device(sc, dev)
    gpu = Stepper(t)
    p = scenes.step_params()
    for _ in range(3):
        r = gpu.advance(p)
        assert r["status"] == 0
        p["prev_dt"] = r["dt"]
    s = gpu.system()
    ro = s["row_offset"].cpu().numpy()
    ci = s["col_index"].cpu().numpy()
    n = res * res
    assert len(ro) == n + 1 and ro[0] == 0 and np.all(np.diff(ro) >= 0) and ro[-1] == len(ci)
    # upper storage: columns ascending and >= the row
    rows = np.repeat(np.arange(n), np.diff(ro))
    assert np.all(ci >= rows)
    val = s["value"].cpu().numpy().reshape(-1, 9)
    cd = s["cdiag"].cpu().numpy().reshape(-1, 9)
    # pinned rows and columns are eliminated: no stored block touches a pinned
    # vertex (the solve reads none), and a pinned row's diagonal is the identity
    for v in pins:
        touch = (rows == v) | (ci == v)
        assert not np.any(touch)
        assert np.array_equal(cd[v], np.eye(3, dtype=np.float32).reshape(9))
    plan = BlockJacobiPCG(s["row_offset"], s["col_index"])
    x = torch.zeros_like(s["b"])
    res_ = plan.solve(s["value"], s["cdiag"], s["b"], x)
    assert res_.status == 0
    plan.close()
    gpu.close()


def test_stacked_sheets_on_spinning_cylinder(dev):
    # the stacked-layer scene at 2 x 16 x 16, cylinder pinned and driven
    # kinematically (positions aimed at each step's end, step deltas for the
    # rewind of shortened steps): per step against the oracle
    sc = scenes.stacked_drape_scene(res=16, layers=2, n_around=24, n_along=8)
    ns = 2 * 16 * 16
    sc["fix_delta"] = np.zeros_like(sc["fix_pos"])
    sco = copy.deepcopy(sc)
    t = scene_to_device(sc, dev)
    gpu = Stepper(t)
    orc = oracle.OracleStepper(sco)
    # 200-1200 PCG iterations per step with 400+ active contacts. Even at
    # cg_tol 1e-5 the two fp32 solvers (different summation orders) stop at
    # different points of the tolerance ball of these stiff systems: measured
    # per-step differences up to 2.9e-5 of the coordinate scale (step 7, 360
    # vs 362 iterations), and 1.85e-4 (step 5, 1,158 vs 1,148 iterations, 2%
    # of that step's displacement) with another summation order of the same
    # SpMV. With the solve in another row order, step 7 (468 contacts) ends
    # with one contact more and a 0.4% longer CCD step (toi 0.3098 vs 0.3085):
    # a contact at the edge of the gap decided by the solve's last bits. The
    # bounds: status and Newton count exact; dt, toi and the contact count
    # within 1%; positions within 5e-4 of the coordinate scale.
    p = scenes.step_params(cg_tol=1e-5)
    clock = 0.0
    for k in range(8):
        now = scenes.spin_fix_positions(sc, clock)[ns:].copy()
        aim = scenes.spin_fix_positions(sc, clock + p["dt"])[ns:].copy()
        t["fix_pos"].view(-1, 3)[ns:] = torch.from_numpy(aim).to(dev)
        t["fix_delta"].view(-1, 3)[ns:] = torch.from_numpy(aim - now).to(dev)
        sco["fix_pos"][ns:] = aim
        sco["fix_delta"][ns:] = aim - now
        orc.delta[:] = sco["fix_delta"].reshape(-1)
        sco["mesh"]["x"][:] = t["mesh"]["x"].cpu().numpy().reshape(-1, 3)
        sco["prev"][:] = t["prev"].cpu().numpy().reshape(-1, 3)
        xg0 = sco["mesh"]["x"].copy()
        rg = gpu.advance(p)
        ro = orc.advance(p)
        xn = t["mesh"]["x"].cpu().numpy().reshape(-1, 3)
        xo = sco["mesh"]["x"]
        err = np.abs(xn - xo).max() / np.abs(xo).max()
        print(f"step {k}: newton {rg['newton_steps']}/{ro['newton_steps']} pcg "
              f"{rg['pcg_iter_total']}/{ro['pcg_iter_total']} pairs {rg['pairs_max']}/"
              f"{ro['pairs_max']} toi {rg['toi_advanced']:.4f}/{ro['toi_advanced']:.4f} "
              f"pos err {err:.2e} (disp {np.abs(xo - xg0).max():.2e})")
        assert rg["status"] == ro["status"] == 0
        assert rg["newton_steps"] == ro["newton_steps"]
        assert abs(rg["dt"] - ro["dt"]) <= 1e-2 * ro["dt"]
        assert abs(rg["toi_advanced"] - ro["toi_advanced"]) <= 1e-2 * ro["toi_advanced"]
        assert abs(rg["pairs_max"] - ro["pairs_max"]) <= 0.01 * ro["pairs_max"]
        assert err < 5e-4
        # the rewound pin targets (fix - (1 - toi) delta) differ only by what
        # the two toi values give, and the cylinder follows the oracle's
        fx = t["fix_pos"].cpu().numpy().reshape(-1, 3)[ns:]
        dtoi = abs(rg["toi_advanced"] - ro["toi_advanced"])
        bound = dtoi * np.abs(aim - now).max() + 1e-6 * np.abs(fx).max()
        assert np.abs(fx - sco["fix_pos"][ns:]).max() <= bound
        assert np.abs(xn[ns:] - xo[ns:]).max() <= bound
        p["prev_dt"] = rg["dt"]
        clock += rg["dt"]
    gpu.close()


def test_dirichlet_elimination_bitwise_vs_oracle(dev):
    """One Newton iteration of a drape with moving pins (pinned vertices 0 and
    70 have free neighbours with higher indices -- their couplings are lifted
    by transpose --, 11 and 143 free neighbours with lower ones). The right-hand
    side after the lifts (main.cu:954-1049) equals the oracle's bit for bit:
    both sum each free row's own lifts in entry order, then its mirror lifts
    in source-row order. The kept blocks equal the oracle's non-eliminated
    blocks exactly, and a second run gives the same bits."""
    res = 12
    pins = [0, res - 1, 70, res * res - 1]
    sc = scenes.drape_scene(res=res, height=0.004, pins=pins)
    sc["fix_pos"][pins] += np.array([1e-3, -2e-3, 1.5e-3], np.float32)
    p = scenes.step_params(max_newton_steps=2, min_newton_steps=10)
    sco = copy.deepcopy(sc)
    orc = oracle.OracleStepper(sco)
    ro = orc.advance(p)
    o = orc.last_system()
    runs = []
    for _ in range(2):
        t = scene_to_device(copy.deepcopy(sc), dev)
        gpu = Stepper(t)
        rg = gpu.advance(p)
        assert rg["status"] == ro["status"] == 5 and rg["newton_steps"] == 1
        runs.append({k: v.cpu().numpy() for k, v in gpu.system().items()})
        gpu.close()
    g = runs[0]
    assert np.array_equal(g["b"].view(np.uint32), runs[1]["b"].view(np.uint32))
    assert np.array_equal(g["b"].view(np.uint32), o["b"].view(np.uint32))
    n = res * res
    assert np.array_equal(g["cdiag"].reshape(-1, 9).view(np.uint32), o["cdiag"].view(np.uint32))
    grow = np.repeat(np.arange(n), np.diff(g["row_offset"]))
    gb = {(int(r), int(c)): v for r, c, v in zip(grow, g["col_index"],
                                                  g["value"].reshape(-1, 9))}
    orow = np.repeat(np.arange(n), np.diff(o["row_offset"].astype(np.int64)))
    pinned = np.zeros(n, bool)
    pinned[pins] = True
    kept = 0
    for r, c, v in zip(orow, o["col_index"], o["value"]):
        if pinned[r] or pinned[c]:
            assert not np.any(v) and (int(r), int(c)) not in gb
        else:
            assert np.array_equal(gb[(int(r), int(c))].view(np.uint32), v.view(np.uint32))
            kept += 1
    assert kept == len(gb)
    # With the pins held in place the lifts vanish (p = 0), and nothing else
    # in a free row changes: the difference is the lifts alone, non-zero on
    # the pins' neighbours on both sides of them in index order, zero elsewhere.
    sc0 = scenes.drape_scene(res=res, height=0.004, pins=pins)
    o0 = oracle.OracleStepper(sc0)
    o0.advance(p)
    b0 = o0.last_system()["b"].reshape(-1, 3)
    d = np.abs(g["b"].reshape(-1, 3) - b0).max(1)
    d[pins] = 0
    nbr = np.zeros(n, bool)
    for r, c in zip(orow, o["col_index"]):
        if pinned[r] != pinned[c]:
            nbr[c if pinned[r] else r] = True
    assert np.all(d[~nbr] == 0)
    assert np.all(d[[1, 12, 71, 82]] > 0.1)  # mirror lifts (pins 0 and 70)
    assert np.all(d[[10, 69, 130, 142]] > 0.1)  # own lifts (pins 11, 70, 143)


def _floor_scene(res=12, height=0.004):
    sc = scenes.drape_scene(res=res, height=height)
    sc["static"] = None
    y = float(sc["mesh"]["x"][:, 1].min())
    return sc, y


def test_cloth_on_analytic_floor_vs_oracle(dev):
    # analytic floor (data.constraint.floor; contact.cu:1154-1382, 2318-2470)
    # with friction: per step against the oracle from the same state
    sc, y = _floor_scene()
    sc["floors"] = [dict(ground=(0.0, y - 0.003, 0.0), ghat=1e-3, friction=0.5)]
    _compare(dev, sc, 14, scenes.step_params(), bar=1e-5)


def test_cloth_on_analytic_spheres_vs_oracle(dev):
    # a static analytic sphere with friction, a kinematic one beside it and a
    # tilted floor below: per step against the oracle
    sc = scenes.drape_scene(res=14, height=0.003, radius=0.3)
    sc["static"] = None
    sc["spheres"] = [dict(center=(0.0, 0.0, 0.0), ghat=1e-3, friction=0.5, radius=0.3),
                     dict(center=(0.45, 0.05, 0.0), ghat=1e-3, friction=0.0, radius=0.2,
                          kinematic=True)]
    up = np.array([0.1, 1.0, 0.0]) / np.linalg.norm([0.1, 1.0, 0.0])
    sc["floors"] = [dict(ground=(0.0, -0.2, 0.0), ghat=1e-3, friction=0.3, up=tuple(up))]
    _compare(dev, sc, 12, scenes.step_params(), bar=1e-5)


def test_cloth_in_analytic_bowl_vs_oracle(dev):
    # a reversed bowl (a container whose upper half is a cylinder) holding the
    # cloth, its rim far above
    # every vertex starts inside (corner radius sqrt(2 * 0.15^2 + 0.2^2) =
    # 0.29 < 0.3 - ghat); the corners reach the wall after about 1 cm of fall
    sc = scenes.drape_scene(res=10, size=0.3, height=0.0, radius=0.0)
    sc["static"] = None
    x = sc["mesh"]["x"]
    x[:, 1] = -0.2
    sc["prev"][:] = x
    sc["spheres"] = [dict(center=(0.0, 0.0, 0.0), ghat=1e-3, friction=0.2, radius=0.3,
                          bowl=True, reverse=True)]
    _compare(dev, sc, 10, scenes.step_params(), bar=1e-5)


def test_pin_driven_through_analytic_floor_fails_like_oracle(dev):
    sc, y = _floor_scene(res=8)
    pins = [0]
    sc["mesh"]["vflag"][pins] |= 1
    sc["shell"]["vflag"] = sc["mesh"]["vflag"].copy()
    sc["fix_pos"] = sc["mesh"]["x"].copy()
    sc["floors"] = [dict(ground=(0.0, y - 0.001, 0.0), ghat=1e-3, friction=0.0)]
    sc["fix_pos"][0, 1] = y - 0.01
    sco = copy.deepcopy(sc)
    t = scene_to_device(sc, dev)
    gpu = Stepper(t)
    rg = gpu.advance(scenes.step_params())
    ro = oracle.OracleStepper(sco).advance(scenes.step_params())
    gpu.close()
    assert rg["status"] == ro["status"] == 6
