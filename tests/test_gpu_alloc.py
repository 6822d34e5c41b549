"""Steady-state allocation and the carried contact pattern.

The reference counts device allocations and frees per step and states the
goal: once the pools have warmed up, a step performs zero dynamic device
allocation (main/main.cu:280-286, 1635-1646; buffer/buffer.hpp). Here every
device and pinned-host allocation of the library is counted
(ppf_alloc_counts), buffers grow to a high-water mark and stay:
- the stepper: the same time steps run twice from the same start state --
  the second pass needs exactly the sizes of the first, so it allocates and
  frees nothing, and it reproduces the first pass bit for bit;
- the solver-backend boundary: a host that checkpoints every step
  (fetch_dyn_counts / fetch_dyn, gather_current_positions) allocates nothing
  once the drape's contact set stops growing, and the per-step
  device-alloc / device-free log channels read zero.
update_dyn (resume, csrmat.cu:136-153) carries a pattern that fetch_dyn
reports until the next step replaces it.
"""
import copy
import os

import numpy as np
import pytest
import torch

from ppf_amd import _lib, scenes
from ppf_amd.datapack import scene_arrays
from ppf_amd.stepper import Stepper, scene_to_device

from test_gpu_backend import run_host

pytestmark = pytest.mark.gpu


def test_stepper_second_pass_allocates_nothing(dev):
    sc = scenes.drape_scene(res=24, height=0.002)
    t = scene_to_device(sc, dev)
    x0, prev0 = t["mesh"]["x"].clone(), t["prev"].clone()
    st = Stepper(t)
    passes = []
    for _ in range(2):
        t["mesh"]["x"].copy_(x0)
        t["prev"].copy_(prev0)
        p = scenes.step_params()
        a0 = _lib.alloc_counts()
        per_step = []
        for _ in range(6):
            r = st.advance(p)
            assert r["status"] == 0, r
            per_step.append((r["device_alloc"], r["device_free"], r["pairs_max"]))
            p["prev_dt"] = r["dt"]
        a1 = _lib.alloc_counts()
        passes.append((a1[0] - a0[0], a1[1] - a0[1], per_step, t["mesh"]["x"].cpu().numpy()))
        print("pass: allocs", passes[-1][0], "frees", passes[-1][1], "per step", per_step)
    st.close()
    assert passes[0][2][-1][2] > 0  # contact reached within the steps
    assert passes[1][0] == 0 and passes[1][1] == 0
    assert all(a == 0 and f == 0 for a, f, _ in passes[1][2])
    assert np.array_equal(passes[0][3].view(np.uint32), passes[1][3].view(np.uint32))


def test_backend_checkpointing_host_reaches_zero_allocation(dev, tmp_path):
    sc = scenes.drape_scene(res=16, height=0.002)
    steps = 24
    rc, out, _ = run_host(tmp_path, scene_arrays(sc, scenes.step_params(), steps), log=True)
    assert rc == 0
    allocs, frees = out["allocs"].astype(np.int64), out["frees"].astype(np.int64)
    print("cumulative allocs per step", allocs.tolist())
    # the last third of the steps allocates and frees nothing
    assert np.all(allocs[-8:] == allocs[-8]) and np.all(frees[-8:] == frees[-8])
    lines = open(tmp_path / "advance.device-alloc.out").read().split()
    per_step = [int(v) for v in lines[1::2]]
    assert len(per_step) == steps and all(v == 0 for v in per_step[-8:])


def test_update_dyn_carries_the_pattern(dev, tmp_path):
    sc = scenes.drape_scene(res=12, height=0.002)
    nv = sc["mesh"]["nv"]
    rng = np.random.default_rng(4)
    rows = []
    for i in range(nv):
        cols = np.unique(rng.integers(i, nv, size=rng.integers(0, 4)))
        rows.append(cols)
    off = np.concatenate([[0], np.cumsum([len(c) for c in rows])]).astype(np.uint32)
    idx = np.concatenate(rows).astype(np.uint32)
    a = scene_arrays(sc, scenes.step_params(), 2)
    a["carry_dyn_offset"] = off
    a["carry_dyn_index"] = idx
    rc, out, _ = run_host(tmp_path, a)
    assert rc == 0
    assert np.array_equal(out["dyn0_offset"], off) and np.array_equal(out["dyn0_index"], idx)
    # The steps do not depend on the carried pattern, and afterwards
    # fetch_dyn reports the steps' own: the run equals one without it.
    rc2, plain, _ = run_host(tmp_path, scene_arrays(sc, scenes.step_params(), 2))
    assert rc2 == 0
    assert np.array_equal(out["curr"].view(np.uint32), plain["curr"].view(np.uint32))
    assert np.array_equal(out["dyn_counts"], plain["dyn_counts"])
    assert out["dyn_counts"][1] == nv + 1
