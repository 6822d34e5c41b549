"""GPU parity of the elastic + inertia assembly (through the C ABI) against
the CPU oracle.

* The fixed pattern equals the numpy enumeration of face / hinge vertex
  pairs plus the diagonal (integer: exact).
* Inertia, air drag and Baraff-Witkin terms use only +, -, *, / and sqrt
  (correctly rounded on both sides) in the same order: force, diagonal and
  blocks are bit-exact when bending is off.
* Bending goes through atan2f, whose device and host implementations may
  differ in the last place; with bending on, every value agrees to 1e-5 of the
  largest entry of its array (tolerance stated here, measured worst case
  printed).
"""
import numpy as np
import pytest
import torch

import oracle
from ppf_amd import scenes
from ppf_amd.elastic import ElasticAssembler, shell_to_device

pytestmark = pytest.mark.gpu


def _run(dev, sh):
    t = shell_to_device(sh, dev)
    asm = ElasticAssembler(t)
    rp, ci = asm.pattern()
    force, cdiag, value = asm.assemble(t, sh["dt"])
    rp = rp.cpu().numpy().astype(np.uint32)
    ci = ci.cpu().numpy().astype(np.uint32)
    out = (force.cpu().numpy(), cdiag.cpu().numpy(), value.cpu().numpy()[:9 * len(ci)])
    asm.close()
    return rp, ci, out


@pytest.mark.parametrize("res", [8, 64])
def test_pattern(dev, res):
    sh = scenes.shell_scene(res, seed=1)
    rp, ci, _ = _run(dev, sh)
    wr, wc = scenes.upper_pattern(sh["nv"], sh["face"], sh["hinge"])
    assert np.array_equal(rp, wr) and np.array_equal(ci, wc)


@pytest.mark.parametrize("res,pins", [(16, 0), (64, 7)])
def test_membrane_inertia_bit_exact(dev, res, pins):
    sh = scenes.shell_scene(res, seed=2, pins=pins, stretch=0.05)
    sh["h_bend"][:] = 0.0                 # bending off: no atan2f on the path
    rp, ci, (f, cd, v) = _run(dev, sh)
    of, ocd, ov, miss = oracle.elastic_assemble(sh, rp, ci, sh["dt"])
    assert miss == 0
    assert np.array_equal(f.view(np.uint32), of.view(np.uint32))
    assert np.array_equal(cd.view(np.uint32), ocd.view(np.uint32))
    assert np.array_equal(v.view(np.uint32), ov.reshape(-1).view(np.uint32))


@pytest.mark.parametrize("res,seed", [(16, 3), (64, 4), (128, 5)])
def test_with_bending(dev, res, seed):
    sh = scenes.shell_scene(res, seed=seed, stretch=0.05, bump=0.1)
    sh["fflag"][::17] = 1                 # some excluded faces
    sh["hflag"][::13] = 1                 # some excluded hinges
    rp, ci, (f, cd, v) = _run(dev, sh)
    of, ocd, ov, miss = oracle.elastic_assemble(sh, rp, ci, sh["dt"])
    ov = ov.reshape(-1)
    assert miss == 0
    assert np.array_equal(cd.view(np.uint32), ocd.view(np.uint32))
    for name, a, b in (("force", f, of), ("blocks", v, ov)):
        err = np.abs(a - b).max() / np.abs(b).max()
        same = np.mean(a.view(np.uint32) == b.view(np.uint32))
        print(f"{name}: max rel err {err:.2e}, bit-identical {same:.4f}")
        assert err < 1e-5, name
