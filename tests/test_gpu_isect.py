"""GPU parity of the post-step intersection check (through the C ABI)
against the CPU oracle: the number of intersecting pairs and the set of
(type, elem0, elem1) records are identical, and so is every record's payload
(the same fp32 positions). Record order is the order threads claimed slots
and is not compared."""
import numpy as np
import pytest
import torch

import oracle
from ppf_amd import scenes
from ppf_amd.ccd import static_to_device
from ppf_amd.narrow import NarrowPhase, mesh_to_device

pytestmark = pytest.mark.gpu


def _keyed(rec):
    return {(int(r["type"]), int(r["elem0"]), int(r["elem1"])): r for r in rec}


def _masses(m, dev, zero=()):
    out = []
    for k in ("nv", "ne", "nf"):
        a = np.ones(m[k], np.float32)
        if k in zero:
            a[:] = 0
        out.append(a)
    return out, [torch.from_numpy(a).to(dev) for a in out]


def _check(dev, m, sm=None, cap=100000, surface=None):
    t = mesh_to_device(m, dev)
    nph = NarrowPhase()
    nph.build(t)
    host, devm = _masses(m, dev)
    st = static_to_device(sm, dev) if sm is not None else None
    n, rec = nph.intersection_check(t, *devm, static=st, surface=surface, cap=cap)
    on, orec = oracle.intersection_check(m, *host, static=sm, surface=surface, cap=cap)
    assert n == on
    if n <= cap:
        g, o = _keyed(rec), _keyed(orec)
        assert set(g) == set(o)
        for k in g:
            assert np.array_equal(g[k]["positions"].view(np.uint32),
                                  o[k]["positions"].view(np.uint32)), k
            assert g[k]["num_verts0"] == o[k]["num_verts0"]
            assert g[k]["num_verts1"] == o[k]["num_verts1"]
    nph.close()
    return n, rec


def _crossing_sheets(res, seed):
    x, f = scenes.stacked_sheets(res, layers=2, gap=0.0, noise=1e-4, seed=seed,
                                 shifts=[(0, 0), (0.37, 0.21)])
    n = res * res
    x = x.astype(np.float64)
    x[n:, 2] += (x[n:, 0] - 0.05) * 0.4
    return x.astype(np.float32), f


def test_face_edge_crossings(dev):
    x, f = _crossing_sheets(48, 0)
    m = scenes.contact_mesh(x, f, ghat=1e-3)
    n, _ = _check(dev, m)
    assert n > 100


def test_offsets_edge_edge_point_point(dev):
    x, f = scenes.stacked_sheets(40, layers=2, gap=2e-4, noise=3e-5, seed=3,
                                 shifts=[(0, 0), (0.008, 0.006)])
    m = scenes.contact_mesh(x, f, ghat=1e-3, offset=1.5e-4)
    n, rec = _check(dev, m)
    types = set(int(r["type"]) for r in rec)
    assert n > 100 and {1, 3} <= types
    # surface_vert_count limits the point-point pass to the first vertices
    n2, rec2 = _check(dev, m, surface=m["nv"] // 2)
    assert n2 < n


def test_collision_mesh_and_cap(dev):
    m, x1, vm, em, fm, sm = scenes.static_feature_scene("floor", res=32)
    y = x1.astype(np.float64)
    y[:, 2] = -1e-3 + y[:, 0] * 0.05
    m2 = dict(m)
    m2["x"] = y.astype(np.float32)
    n, rec = _check(dev, m2, sm=sm)
    assert n > 10 and all(int(r["type"]) == 2 for r in rec)
    # A small cap still counts every pair.
    n3, rec3 = _check(dev, m2, sm=sm, cap=5)
    assert n3 == n and len(rec3) == 5


def test_clear_scene(dev):
    m, x1, vm, em, fm, sm = scenes.falling_scene(res=48, layers=3, seed=4,
                                                  ground_gap=2e-4)
    n, rec = _check(dev, m, sm=sm)
    assert n == 0 and len(rec) == 0
