"""The solver-backend boundary on the GPU: the C++ mock host fills the host's
DataSet / ParamSet from a scene and drives libsimbackend_mi355x.so in
Backend::run's call order; its result is compared with the stage library's
stepper on the same scene (the same kernels in the same order: bit-identical)
and with the CPU oracle (positions within 1e-5 of max |x| per step, the
stepper's parity bar).

The mock host runs as a child process of the test.
"""
import copy
import os
import subprocess

import numpy as np
import pytest
import torch

from ppf_amd import scenes
from ppf_amd.datapack import read_pack, scene_arrays, write_pack
from ppf_amd.stepper import Stepper, scene_to_device

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "ppf-contact-solver_amd", "mock_host")


def run_host(tmp_path, arrays, log=False):
    scene = tmp_path / "scene.pack"
    out = tmp_path / "out.pack"
    write_pack(scene, arrays)
    cmd = [HOST, str(scene), str(out)] + ([str(tmp_path)] if log else [])
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    print(p.stdout[-2000:], p.stderr[-2000:])
    return p.returncode, (read_pack(out) if out.exists() else None), p.stderr


def stepper_run(dev, sc, params, steps, schedule=None, deltas=None, vflag_or=None,
                eflag_or=None):
    t = scene_to_device(sc, dev)
    if vflag_or is not None:
        t["mesh"]["vflag"] |= torch.from_numpy(vflag_or.view(np.int32)).to(dev)
    if eflag_or is not None:
        t["mesh"]["eflag"] |= torch.from_numpy(eflag_or.view(np.int32)).to(dev)
    if deltas is not None and t.get("fix_delta") is None:
        t["fix_delta"] = torch.zeros_like(t["mesh"]["x"])
    st = Stepper(t)
    p = dict(params)
    pins = np.nonzero(sc["mesh"]["vflag"] & 1)[0]
    dts = []
    for k in range(steps):
        if schedule is not None:
            fp = t["fix_pos"].view(-1, 3)
            fp[pins] = torch.from_numpy(schedule[k]).to(dev)
            t["fix_delta"].view(-1, 3)[pins] = torch.from_numpy(deltas[k]).to(dev)
        r = st.advance(p)
        assert r["status"] == 0, r
        p["prev_dt"] = r["dt"]
        dts.append(r["dt"])
    x = t["mesh"]["x"].cpu().numpy().reshape(-1, 3)
    prev = t["prev"].cpu().numpy().reshape(-1, 3)
    st.close()
    return x, prev, np.array(dts, np.float32)


def test_mock_host_drape_matches_stepper_and_oracle(dev, tmp_path):
    sc = scenes.drape_scene(res=16, height=0.003)
    params = scenes.step_params()
    steps = 12
    rc, out, _ = run_host(tmp_path, scene_arrays(sc, params, steps), log=True)
    assert rc == 0
    flags = out["flags"].reshape(-1, 6)
    assert flags.shape[0] == steps and flags.all()
    x, prev, dts = stepper_run(dev, copy.deepcopy(sc), params, steps)
    xh = out["curr"].reshape(-1, 3)
    assert np.array_equal(xh.view(np.uint32), x.view(np.uint32))
    assert np.array_equal(out["prev"].reshape(-1, 3).view(np.uint32), prev.view(np.uint32))
    assert np.array_equal(out["dt"], dts)
    assert abs(out["time"][-1] - float(np.sum(dts.astype(np.float64)))) < 1e-6
    # The host's run is the stepper's, bit for bit; on this trajectory every
    # step of the stepper matches the oracle's step from the same state (to
    # 1e-5 of the coordinate scale, or 6x the fp32 solves' distance from an
    # fp64 re-solve on stiff steps). A free-running comparison would measure
    # the chaos of contact events, not the step.
    from test_gpu_step import _compare
    _compare(dev, copy.deepcopy(sc), steps, params, bar=1e-5)
    assert os.path.exists(tmp_path / "advance.newton_steps.out")
    assert out["dyn_counts"][1] == sc["mesh"]["nv"] + 1


def test_mock_host_strain_limited_drape_matches_stepper(dev, tmp_path):
    # FaceParam.strainlimit through the host's DataSet: the backend's run is
    # the stepper's with the same per-face limits, bit for bit.
    sc = scenes.drape_scene(res=12, height=0.003, strain_limit=0.01)
    params = scenes.step_params()
    steps = 6
    rc, out, _ = run_host(tmp_path, scene_arrays(sc, params, steps))
    assert rc == 0
    x, prev, dts = stepper_run(dev, copy.deepcopy(sc), params, steps)
    assert np.array_equal(out["curr"].reshape(-1, 3).view(np.uint32), x.view(np.uint32))
    assert np.array_equal(out["dt"], dts)
    sc2 = scenes.drape_scene(res=12, height=0.003)
    x2, _, _ = stepper_run(dev, sc2, params, steps)
    assert not np.array_equal(x2, x)      # the limit changed the run


def test_mock_host_plasticity_matches_stepper(dev, tmp_path):
    # FaceParam / HingeParam plasticity through the host's DataSet: positions
    # and the rest shape the host fetches afterwards (fetch_inv_rest,
    # fetch_rest_angles) equal the stepper's run bit for bit.
    sc = scenes.drape_scene(res=12, height=0.003, plasticity=(5.0, 0.001),
                            bend_plasticity=(5.0, 0.01))
    params = scenes.step_params()
    steps = 6
    rc, out, _ = run_host(tmp_path, scene_arrays(sc, params, steps))
    assert rc == 0
    sc2 = copy.deepcopy(sc)
    t = scene_to_device(sc2, dev)
    st = Stepper(t)
    p = dict(params)
    for _ in range(steps):
        r = st.advance(p)
        assert r["status"] == 0, r
        p["prev_dt"] = r["dt"]
    x = t["mesh"]["x"].cpu().numpy()
    inv = t["shell"]["inv_rest"].cpu().numpy()
    hr = t["shell"]["h_rest"].cpu().numpy()
    st.close()
    assert np.array_equal(out["curr"].view(np.uint32), x.view(np.uint32))
    assert np.array_equal(out["inv_rest"].view(np.uint32), inv.view(np.uint32))
    assert np.array_equal(out["rest_angle"].view(np.uint32), hr.view(np.uint32))
    assert not np.array_equal(inv, np.asarray(sc["shell"]["inv_rest"], np.float32).reshape(-1))
    assert not np.array_equal(hr, np.asarray(sc["shell"]["h_rest"], np.float32))


def test_mock_host_moving_pins(dev, tmp_path):
    res = 12
    pins = [0, res - 1]
    sc = scenes.drape_scene(res=res, height=0.004, pins=pins)
    params = scenes.step_params()
    steps = 8
    base = sc["fix_pos"][pins].astype(np.float64)
    # pins slide along x at 0.05 m/s: position at the step's end, delta over it
    sched = np.stack([base + [0.05 * params["dt"] * (k + 1), 0, 0] for k in range(steps)])
    deltas = np.broadcast_to(np.array([0.05 * params["dt"], 0, 0]), sched.shape)
    sched, deltas = sched.astype(np.float32), np.ascontiguousarray(deltas, np.float32)
    rc, out, _ = run_host(tmp_path, scene_arrays(sc, params, steps, sched, deltas))
    assert rc == 0
    x, prev, dts = stepper_run(dev, copy.deepcopy(sc), params, steps, sched, deltas)
    xh = out["curr"].reshape(-1, 3)
    assert np.array_equal(xh.view(np.uint32), x.view(np.uint32))
    assert np.array_equal(out["dt"], dts)
    # pinned vertices end where the (possibly rewound) schedule put them
    last = sched[-1] - (1.0 - dts[-1] / params["dt"]) * deltas[-1]
    assert np.abs(xh[pins] - last).max() < 1e-6


def test_mock_host_collision_windows(dev, tmp_path):
    # every sheet vertex in a group whose window has not opened yet: the
    # sheet's own queries find nothing; compared with the stepper given the
    # same query-inactive flags
    sc = scenes.drape_scene(res=12, height=0.002)
    params = scenes.step_params()
    steps = 4
    nv = sc["mesh"]["nv"]
    dmap = np.zeros(nv, np.uint32)
    windows = np.zeros((1, 8, 2), np.float32)
    windows[0, 0] = (10.0, 20.0)
    rc, out, _ = run_host(tmp_path, scene_arrays(sc, params, steps,
                                                  windows=(dmap, windows, np.array([1]))))
    assert rc == 0
    x, _, _ = stepper_run(dev, copy.deepcopy(sc), params, steps,
                          vflag_or=np.full(nv, 4, np.uint32),
                          eflag_or=np.full(sc["mesh"]["ne"], 4, np.uint32))
    assert np.array_equal(out["curr"].reshape(-1, 3).view(np.uint32), x.view(np.uint32))


def test_mock_host_refuses_intersecting_start(dev, tmp_path):
    sc = scenes.drape_scene(res=8, height=0.002)
    x = sc["mesh"]["x"].reshape(-1, 3)
    x[27, 1] -= 0.5   # pull one vertex into the sphere: its edges cross it
    sc["prev"] = x.copy()
    rc, out, err = run_host(tmp_path, scene_arrays(sc, scenes.step_params(), 1))
    assert rc == 3
    assert "code 1" in err


def test_mock_host_refuses_unsupported_content(dev, tmp_path):
    sc = scenes.drape_scene(res=8, height=0.002)
    a = scene_arrays(sc, scenes.step_params(), 1)
    a["stitch"] = np.array([1], np.uint32)
    rc, out, err = run_host(tmp_path, a)
    assert rc == 3
    assert "code 4" in err and "stitch" in err


def test_mock_host_analytic_floor_matches_stepper(dev, tmp_path):
    """Analytic colliders through the solver-backend boundary
    (DataSet.constraint.floor / sphere, ParamSet.constraint_tol): the mock host's
    run equals the stage ABI stepper's bit for bit."""
    sc = scenes.drape_scene(res=12, height=0.004)
    sc["static"] = None
    y = float(sc["mesh"]["x"][:, 1].min())
    sc["floors"] = [dict(ground=(0.0, y - 0.003, 0.0), ghat=1e-3, friction=0.5)]
    sc["spheres"] = [dict(center=(0.7, y, 0.0), ghat=1e-3, friction=0.2, radius=0.25,
                          kinematic=True)]
    params = scenes.step_params()
    steps = 8
    rc, out, _ = run_host(tmp_path, scene_arrays(sc, params, steps))
    assert rc == 0
    x, prev, dts = stepper_run(dev, copy.deepcopy(sc), params, steps)
    assert np.array_equal(out["curr"].reshape(-1, 3).view(np.uint32), x.view(np.uint32))
    assert np.array_equal(out["dt"], dts)
    assert float(x[:, 1].min()) > y - 0.003  # above the floor
