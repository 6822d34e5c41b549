"""GPU parity of the contact force / Hessian assembly (through the C ABI)
against the CPU oracle on identical pair lists. Both sum every block and every
gradient row in pair order with the same fp32 expressions, so the expected
agreement is bit-exact; the test reports the bit-exact fraction and also
bounds the difference (1e-6 of the largest entry) so a single-ulp device/host
difference in a library function would be named rather than hidden."""
import numpy as np
import pytest
import torch

import oracle
from ppf_amd import scenes
from ppf_amd.narrow import ContactAssembler, NarrowPhase, mesh_to_device, pairs_to_numpy

pytestmark = pytest.mark.gpu


def run(res, seed, dev, elastic_on=True, mode=0):
    m, x0, mass, fric, elastic = scenes.contact_scenario(res=res, seed=seed)
    t = mesh_to_device(m, dev)
    nph = NarrowPhase()
    nph.build(t)
    pairs = nph.proximity(t)
    asm = ContactAssembler()
    el_t = tuple(torch.from_numpy(np.ascontiguousarray(a)).reshape(-1).to(dev)
                 for a in elastic) if elastic_on else None
    g, rp, rows, cols, vals = asm.assemble(
        t, torch.from_numpy(x0).reshape(-1).to(dev), torch.from_numpy(mass).to(dev),
        {k: torch.from_numpy(v).to(dev) for k, v in fric.items()}, pairs,
        friction_mode=mode, elastic=el_t)
    torch.cuda.synchronize()
    pn = pairs_to_numpy(pairs)
    og, orow, ocol, oval = oracle.contact_assemble(
        m, x0, mass, fric, pn, friction_mode=mode,
        elastic=elastic if elastic_on else None)
    return (g.cpu().numpy(), rp.cpu().numpy(), rows.cpu().numpy(), cols.cpu().numpy(),
            vals.cpu().numpy().reshape(-1, 9), og, orow, ocol, oval, m, len(pn))


@pytest.mark.parametrize("res,seed,elastic_on,mode",
                         [(24, 0, True, 0), (32, 1, False, 1), (40, 2, True, 2)])
def test_assembly_matches_oracle(dev, res, seed, elastic_on, mode):
    g, rp, rows, cols, vals, og, orow, ocol, oval, m, npairs = run(
        res, seed, dev, elastic_on, mode)
    print(f"res {res}: {npairs} pairs, {len(rows)} blocks; gradient bit-exact "
          f"{np.mean(g.view(np.uint32) == og.view(np.uint32)):.4f}, blocks "
          f"{np.mean(vals.view(np.uint32) == oval.view(np.uint32)):.4f}")
    assert np.array_equal(rows.astype(np.uint32), orow)
    assert np.array_equal(cols.astype(np.uint32), ocol)
    assert np.abs(g - og).max() <= 1e-6 * np.abs(og).max()
    assert np.abs(vals - oval).max() <= 1e-6 * np.abs(oval).max()
    # Row offsets of the block CSR.
    expect = np.searchsorted(orow, np.arange(m["nv"] + 1))
    assert np.array_equal(rp.astype(np.int64), expect)


def test_assembly_is_deterministic(dev):
    a = run(24, 3, dev)
    b = run(24, 3, dev)
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
    assert np.array_equal(a[4].view(np.uint32), b[4].view(np.uint32))
