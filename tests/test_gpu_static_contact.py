"""GPU parity of contact against the static collision mesh (pair kinds 4-6,
through the C ABI) against the CPU oracle.

* Pair sets: the GPU's collision-mesh pairs equal the oracle's (integer:
  exact; compared as sets, the order within a query follows the traversal).
* Assembly: on the same pair list (self-contact pairs followed by
  collision-mesh pairs) the gradient and every block agree with the oracle to
  1e-6 of the largest entry, and the bit-identical fraction is reported; the
  expected result is bit-exact (same fp32 expressions and order).
"""
import numpy as np
import pytest
import torch

import oracle
from ppf_amd import scenes
from ppf_amd.ccd import static_to_device
from ppf_amd.narrow import ContactAssembler, NarrowPhase, mesh_to_device, pairs_to_numpy

pytestmark = pytest.mark.gpu


def near_scene(feature, res, height=4e-4, tilt=0.0, seed=1):
    m, x1, vm, em, fm, sm = scenes.static_feature_scene(feature, res=res)
    x = m["x"].astype(np.float64)
    x[:, 2] -= 1e-3 - height
    x[:, 2] += tilt * x[:, 0]
    rng = np.random.default_rng(seed)
    x[:, 2] += rng.normal(0, 5e-5, len(x))
    m = dict(m)
    m["x"] = x.astype(np.float32)
    sm = dict(sm)
    for k, n in (("v", sm["nv"]), ("e", sm["ne"]), ("f", sm["nf"])):
        sm[k + "_fric"] = rng.uniform(0.1, 0.6, n).astype(np.float32)
    vm = vm.copy()
    vm[::17] = 0.0
    return m, vm, em, fm, sm


SCENES = [("floor", 48, 0.0), ("floor", 64, 0.003), ("spike", 48, 0.0), ("ridge", 64, 0.0)]


def _setup(dev, feature, res, tilt):
    m, vm, em, fm, sm = near_scene(feature, res, tilt=tilt)
    t = mesh_to_device(m, dev)
    st = static_to_device(sm, dev)
    md = [torch.from_numpy(a).to(dev) for a in (vm, em, fm)]
    nph = NarrowPhase()
    nph.build(t)
    return m, vm, em, fm, sm, t, st, md, nph


@pytest.mark.parametrize("feature,res,tilt", SCENES)
def test_static_pairs_match_oracle(dev, feature, res, tilt):
    m, vm, em, fm, sm, t, st, md, nph = _setup(dev, feature, res, tilt)
    g = pairs_to_numpy(nph.static_pairs(t, st, *md))
    o = oracle.static_proximity(m, sm, vm, em, fm)
    gs = {(int(p["kind"]), int(p["a"]), int(p["b"]), int(p["count"])) for p in g}
    os_ = {(int(p["kind"]), int(p["a"]), int(p["b"]), int(p["count"])) for p in o}
    print(f"{feature} {res}: {len(g)} pairs, kinds {np.bincount(g['kind'], minlength=7)[4:]}")
    assert len(gs) == len(g)
    assert gs == os_ and len(gs) > 0
    nph.close()


@pytest.mark.parametrize("feature,res,tilt", SCENES)
def test_static_assembly_matches_oracle(dev, feature, res, tilt):
    m, vm, em, fm, sm, t, st, md, nph = _setup(dev, feature, res, tilt)
    pairs = torch.cat([nph.proximity(t), nph.static_pairs(t, st, *md)])
    rng = np.random.default_rng(7)
    x0 = (m["x"] + rng.normal(0, 2e-5, m["x"].shape)).astype(np.float32)
    mass = vm.copy()
    fric = {"v": rng.uniform(0, 0.5, m["nv"]).astype(np.float32),
            "e": rng.uniform(0, 0.5, m["ne"]).astype(np.float32),
            "f": rng.uniform(0, 0.5, m["nf"]).astype(np.float32)}
    elastic = scenes.face_element_bsr(m["nv"], m["face"], seed=3, scale=1.0)
    el_t = tuple(torch.from_numpy(np.ascontiguousarray(a)).reshape(-1).to(dev) for a in elastic)
    asm = ContactAssembler()
    g, rp, rows, cols, vals = asm.assemble(
        t, torch.from_numpy(x0).reshape(-1).to(dev), torch.from_numpy(mass).to(dev),
        {k: torch.from_numpy(v).to(dev) for k, v in fric.items()}, pairs,
        elastic=el_t, static=st)
    pn = pairs_to_numpy(pairs)
    og, orow, ocol, oval = oracle.contact_assemble(m, x0, mass, fric, pn, elastic=elastic,
                                                   static=sm)
    g = g.cpu().numpy()
    rows, cols = rows.cpu().numpy(), cols.cpu().numpy()
    vals = vals.cpu().numpy().reshape(-1, 9)
    nst = int(np.sum((pn["kind"] & 255) >= 4))
    print(f"{feature} {res}: {len(pn)} pairs ({nst} collision-mesh), {len(rows)} blocks; "
          f"gradient bit-exact {np.mean(g.view(np.uint32) == og.view(np.uint32)):.4f}, "
          f"blocks {np.mean(vals.view(np.uint32) == oval.view(np.uint32)):.4f}")
    assert nst > 0
    assert np.array_equal(rows.astype(np.uint32), orow)
    assert np.array_equal(cols.astype(np.uint32), ocol)
    assert np.abs(g - og).max() <= 1e-6 * np.abs(og).max()
    assert np.abs(vals - oval).max() <= 1e-6 * np.abs(oval).max()
    asm.close()
    nph.close()
