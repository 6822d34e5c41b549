"""Many uneven scenes stepped together (ppf_stepper_advance_scenes) through
the C ABI.

- Every scene of a batch, every step, against the oracle's step of that scene
  alone from the same state: status, Newton count, dt, contact count, zero
  intersections, positions within the per-step bar of the one-scene tests
  (1e-5 of the coordinate scale, or 6x the fp32 solves' distance from an fp64
  re-solve of the scene's Newton system on stiff steps).
- Isolation: a scene's trajectory does not depend on the other scenes of its
  batch -- the same scene beside different companions steps bit for bit alike
  (per-scene solver scalars, line-search minima and Newton control).
- The first step's contact pairs of each scene equal those of a one-scene
  stepper from the same state (the trees never pair elements of different
  scenes, though the scenes overlap in space).
"""
import copy

import numpy as np
import pytest
import torch

import oracle
from ppf_amd import scenes
from ppf_amd.stepper import Stepper, scene_to_device

pytestmark = pytest.mark.gpu


def _scenes():
    # uneven sizes, radii and heights; all centred on the origin (they
    # overlap in space)
    return [scenes.drape_scene(res=12, radius=0.3, height=0.002),
            scenes.drape_scene(res=18, radius=0.15, height=0.003),
            scenes.drape_scene(res=9, radius=0.25, height=0.0015),
            scenes.drape_scene(res=15, radius=0.4, height=0.0025)]


def _slice(a, rows, c):
    return a.reshape(-1, 3)[rows[c]:rows[c + 1]]


def test_batch_steps_match_oracle_per_scene(dev):
    from test_gpu_step import (SOLVE_BALL_FACTOR, solve_error_vs_fp64, ulp_nearest,
                                within_ulp_copies)
    scs = _scenes()
    packed, rows = scenes.pack_scenes(copy.deepcopy(scs))
    t = scene_to_device(packed, dev)
    st = Stepper(t)
    p = scenes.step_params()
    ns = len(scs)
    prev_dt = [p["dt"]] * ns
    worst = 0.0
    for k in range(8):
        x0 = t["mesh"]["x"].cpu().numpy().reshape(-1, 3).copy()
        pv0 = t["prev"].cpu().numpy().reshape(-1, 3).copy()
        res = st.advance_scenes(p, prev_dt)
        x1 = t["mesh"]["x"].cpu().numpy().reshape(-1, 3)
        for c in range(ns):
            sco = copy.deepcopy(scs[c])
            sco["mesh"]["x"][:] = _slice(x0, rows, c)
            sco["prev"][:] = _slice(pv0, rows, c)
            orc = oracle.OracleStepper(sco)
            po = dict(p)
            po["prev_dt"] = prev_dt[c]
            ro = orc.advance(po)
            rg = res[c]
            xo = sco["mesh"]["x"]
            scale = np.abs(xo).max()
            err = np.abs(_slice(x1, rows, c) - xo).max() / scale
            bar = 1e-5
            if err >= bar:
                # a stiff step: the bar widens to the fp32 solves' distance from
                # an fp64 re-solve of this scene's Newton system (stepped alone
                # from the same state)
                one = copy.deepcopy(scs[c])
                one["mesh"]["x"][:] = _slice(x0, rows, c)
                one["prev"][:] = _slice(pv0, rows, c)
                single = Stepper(scene_to_device(one, dev))
                single.advance(po)
                eg, eo = solve_error_vs_fp64(single.system(), p["cg_tol"], scale)
                single.close()
                bar = max(bar, SOLVE_BALL_FACTOR * max(eg, eo))
            near = (err, 0.0)
            if err >= bar:
                # a step on a discontinuity of the algorithm: the oracle's
                # steps from one-ulp copies of the scene's start state
                near = ulp_nearest(orc, sco, _slice(x0, rows, c), _slice(pv0, rows, c), po,
                                   _slice(x1, rows, c), scale)
            line = (f"step {k} scene {c}: status {rg['status']}/{ro['status']} newton "
                    f"{rg['newton_steps']}/{ro['newton_steps']} pcg {rg['pcg_iter_total']}/"
                    f"{ro['pcg_iter_total']} pairs {rg['pairs_max']}/{ro['pairs_max']} dt "
                    f"{rg['dt']:.6g}/{ro['dt']:.6g} isect {rg['intersections']} pos err "
                    f"{err:.2e} bar {bar:.2e} nearest one-ulp copy {near[0]:.2e}")
            print(line, flush=True)
            assert rg["status"] == ro["status"] == 0, line
            assert rg["intersections"] == 0, line
            assert rg["newton_steps"] == ro["newton_steps"], line
            assert abs(rg["dt"] - ro["dt"]) <= 1e-3 * ro["dt"], line
            assert abs(rg["pairs_max"] - ro["pairs_max"]) <= 0.02 * max(ro["pairs_max"], 1), line
            assert err < bar or within_ulp_copies(err, near, bar), line
            worst = max(worst, min(err, near[0]))
            prev_dt[c] = rg["dt"]
    st.close()
    print(f"worst position gap over the batch {worst:.2e}")


def test_scene_isolated_from_its_companions(dev):
    scs = _scenes()
    runs = []
    for companions in ([scs[1], scs[2]], [scs[3]]):
        group = [copy.deepcopy(scs[0])] + [copy.deepcopy(s) for s in companions]
        packed, rows = scenes.pack_scenes(group)
        t = scene_to_device(packed, dev)
        st = Stepper(t)
        p = scenes.step_params()
        prev_dt = [p["dt"]] * len(group)
        out = []
        for _ in range(6):
            res = st.advance_scenes(p, prev_dt)
            prev_dt = [r["dt"] for r in res]
            out.append((res[0]["newton_steps"], res[0]["pcg_iter_total"], res[0]["dt"]))
        x = t["mesh"]["x"].cpu().numpy().reshape(-1, 3)[rows[0]:rows[1]].copy()
        st.close()
        runs.append((out, x))
        print("scene 0 beside", len(companions), "companions:", out)
    assert runs[0][0] == runs[1][0]
    assert np.array_equal(runs[0][1].view(np.uint32), runs[1][1].view(np.uint32))


def test_first_step_pairs_equal_one_scene_steppers(dev):
    scs = _scenes()
    packed, rows = scenes.pack_scenes(copy.deepcopy(scs))
    t = scene_to_device(packed, dev)
    st = Stepper(t)
    p = scenes.step_params()
    # step until the scenes are in contact, then compare one more step
    prev_dt = [p["dt"]] * len(scs)
    for _ in range(2):
        res = st.advance_scenes(p, prev_dt)
        prev_dt = [r["dt"] for r in res]
    x = t["mesh"]["x"].cpu().numpy().reshape(-1, 3).copy()
    pv = t["prev"].cpu().numpy().reshape(-1, 3).copy()
    res = st.advance_scenes(p, prev_dt)
    st.close()
    for c, sc in enumerate(scs):
        one = copy.deepcopy(sc)
        one["mesh"]["x"][:] = _slice(x, rows, c)
        one["prev"][:] = _slice(pv, rows, c)
        s1 = Stepper(scene_to_device(one, dev))
        q = dict(p)
        q["prev_dt"] = prev_dt[c]
        r1 = s1.advance(q)
        s1.close()
        print(f"scene {c}: batched pairs {res[c]['pairs_max']} newton {res[c]['newton_steps']}; "
              f"alone {r1['pairs_max']} newton {r1['newton_steps']}")
        assert res[c]["pairs_max"] > 0
        # the first Newton iteration's pairs depend on the start state only
        assert res[c]["newton_steps"] == r1["newton_steps"]
        assert abs(res[c]["pairs_max"] - r1["pairs_max"]) <= 0.02 * r1["pairs_max"]
