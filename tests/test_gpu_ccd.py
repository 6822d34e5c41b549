"""GPU parity of the ACCD kinds and the contact line search (through the C
ABI) against the CPU oracle. Both evaluate the same fp32 expressions in the
same order with no contraction, and sqrt / division are correctly rounded on
both sides, so every toi and overlap flag must be bit-exact. The line search
is a minimum over the pairs whose swept boxes overlap; that set does not
depend on the tree, so the result is bit-exact too."""
import numpy as np
import pytest
import torch

import oracle
from ppf_amd import scenes
from ppf_amd.ccd import LineSearch, ccd_batch, static_to_device
from ppf_amd.narrow import mesh_to_device

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_ccd_batch_bit_exact(dev, kind):
    n = 3000
    x0, x1 = scenes.ccd_pairs(kind, n, seed=40 + kind)
    rng = np.random.default_rng(kind)
    off = rng.choice([0.0, 1e-5, 1e-4], n).astype(np.float32)
    gh = rng.uniform(5e-4, 2e-3, n).astype(np.float32)
    toi, ov = ccd_batch(kind, torch.from_numpy(x0).to(dev), torch.from_numpy(x1).to(dev),
                        torch.from_numpy(off).to(dev), torch.from_numpy(gh).to(dev))
    toi, ov = toi.cpu().numpy(), ov.cpu().numpy()
    ref = np.array([oracle.ccd(kind, x0[i], x1[i], float(off[i]), float(gh[i]))
                    for i in range(n)])
    rt = ref[:, 0].astype(np.float32)
    ro = ref[:, 1].astype(np.int32)
    hits = int(np.sum(rt < 1.25))
    print(f"kind {kind}: {hits} of {n} pairs stop early, {int(ro.sum())} overlaps")
    assert hits > n // 10
    bad = np.nonzero(toi.view(np.uint32) != rt.view(np.uint32))[0]
    assert len(bad) == 0, (bad[:5], toi[bad[:5]], rt[bad[:5]])
    assert np.array_equal(ov, ro)


def test_ccd_batch_degenerate_inputs(dev):
    # Zero motion, coincident columns and starting overlaps.
    x0 = np.zeros((4, 4, 3), np.float32)
    x1 = np.zeros((4, 4, 3), np.float32)
    x0[1] = [[0, 0, 1e-3], [0, 0, 0], [1e-2, 0, 0], [0, 1e-2, 0]]
    x1[1] = x0[1]
    x0[2] = [[2e-3, 2e-3, 5e-5], [0, 0, 0], [1e-2, 0, 0], [0, 1e-2, 0]]
    x1[2] = x0[2]
    x1[2, 0, 2] -= 1e-3
    x0[3] = [[2e-3, 2e-3, 5e-3], [0, 0, 0], [1e-2, 0, 0], [0, 1e-2, 0]]
    x1[3] = x0[3]
    x1[3, 0, 2] -= 1e-2
    off = np.full(4, 1e-4, np.float32)
    gh = np.full(4, 1e-3, np.float32)
    toi, ov = ccd_batch(0, torch.from_numpy(x0).to(dev), torch.from_numpy(x1).to(dev),
                        torch.from_numpy(off).to(dev), torch.from_numpy(gh).to(dev))
    toi, ov = toi.cpu().numpy(), ov.cpu().numpy()
    for i in range(4):
        rt, ro = oracle.ccd(0, x0[i], x1[i], float(off[i]), float(gh[i]))
        assert np.float32(rt).view(np.uint32) == toi[i].view(np.uint32), i
        assert bool(ov[i]) == ro, i
    assert ov[0] == 1 and toi[0] == 0.0      # collapsed frame
    assert toi[1] == 1.25                    # no relative motion
    assert ov[2] == 1 and toi[2] == 0.0      # starts inside the offset
    assert 0.0 < toi[3] < 1.25 and ov[3] == 0


SCENES = {
    "sheets2": dict(res=24, layers=2, seed=0),
    "sheets3_flags": dict(res=32, layers=3, seed=1),
    "sheets3_static": dict(res=24, layers=3, seed=2, static=True, ground_gap=2e-4),
    "big_static": dict(res=96, layers=2, seed=3, static=True, drop=2e-3,
                       ground_gap=2e-4),
}


def _scene(spec):
    spec = dict(spec)
    static = spec.pop("static", False)
    m, x1, vm, em, fm, sm = scenes.falling_scene(**spec)
    return m, x1, vm, em, fm, (sm if static else None)


def _to(dev, *arrs):
    return [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in arrs]


@pytest.mark.parametrize("name", list(SCENES))
def test_line_search_bit_exact(dev, name):
    m, x1, vm, em, fm, sm = _scene(SCENES[name])
    t = mesh_to_device(m, dev)
    st = static_to_device(sm, dev) if sm is not None else None
    x1d, vmd, emd, fmd = _to(dev, x1.reshape(-1), vm, em, fm)
    ls = LineSearch()
    toi, ov = ls(t, x1d, vmd, emd, fmd, static=st)
    rt, ro = oracle.line_search(m, x1, vm, em, fm, static=sm)
    print(f"{name}: nv {m['nv']} toi gpu {toi!r} oracle {rt!r} overlap {ov}/{ro}")
    assert np.float32(toi).view(np.uint32) == np.float32(rt).view(np.uint32)
    assert ov == ro
    assert 0.0 < toi < 1.0
    ls.close()


def test_line_search_refit_and_no_hit(dev):
    # A second call with rebuild off refits the swept trees; a zero step
    # certifies the whole sweep (1.0).
    m, x1, vm, em, fm, sm = _scene(SCENES["sheets3_static"])
    t = mesh_to_device(m, dev)
    st = static_to_device(sm, dev)
    vmd, emd, fmd = _to(dev, vm, em, fm)
    ls = LineSearch()
    toi0, _ = ls(t, *_to(dev, x1.reshape(-1)), vmd, emd, fmd, static=st)
    rng = np.random.default_rng(77)
    x2 = (m["x"] + 0.5 * (x1 - m["x"]) +
          rng.normal(0, 1e-4, x1.shape)).astype(np.float32)
    toi1, ov1 = ls(t, *_to(dev, x2.reshape(-1)), vmd, emd, fmd, static=st, rebuild=False)
    rt1, ro1 = oracle.line_search(m, x2, vm, em, fm, static=sm)
    assert np.float32(toi1).view(np.uint32) == np.float32(rt1).view(np.uint32)
    assert ov1 == ro1
    toi2, ov2 = ls(t, *_to(dev, m["x"].reshape(-1)), vmd, emd, fmd, static=st,
                   rebuild=False)
    assert toi2 == 1.0 and not ov2
    ls.close()


def test_line_search_overlap_flag(dev):
    # Two sheets closer than their offsets: toi 0 and the overlap flag.
    x, f = scenes.stacked_sheets(12, layers=2, gap=8e-5, noise=0.0, seed=5)
    m = scenes.contact_mesh(x, f, ghat=1e-3, offset=5e-5)
    x1 = x.copy()
    x1[:, 2] += np.where(np.arange(len(x)) < len(x) // 2, -1e-3, 1e-3).astype(np.float32)
    nv, ne, nf = m["nv"], m["ne"], m["nf"]
    ones = [np.ones(k, np.float32) for k in (nv, ne, nf)]
    ls = LineSearch()
    toi, ov = ls(mesh_to_device(m, dev), *_to(dev, x1.reshape(-1), *ones))
    rt, ro = oracle.line_search(m, x1, *ones)
    assert ro and ov
    assert toi == rt == 0.0
    ls.close()


@pytest.mark.parametrize("feature,mass", [("spike", "fmass"), ("ridge", "emass"),
                                          ("floor", "vmass")])
def test_line_search_collision_mesh_kinds(dev, feature, mass):
    # Each static feature is decided by one collision-mesh pair kind; without
    # the feature (or with the masses that kind tests set to zero) the step
    # goes further.
    m, x1, vm, em, fm, sm = scenes.static_feature_scene(feature)
    t = mesh_to_device(m, dev)
    st = static_to_device(sm, dev)
    ls = LineSearch()
    masses = {"vmass": vm, "emass": em, "fmass": fm}
    toi, ov = ls(t, *_to(dev, x1.reshape(-1), vm, em, fm), static=st)
    rt, ro = oracle.line_search(m, x1, vm, em, fm, static=sm)
    assert np.float32(toi).view(np.uint32) == np.float32(rt).view(np.uint32)
    assert not ov and not ro
    free, _ = oracle.line_search(m, x1, vm, em, fm)
    assert rt < free
    masses[mass] = np.zeros_like(masses[mass])
    toi0, _ = ls(t, *_to(dev, x1.reshape(-1), masses["vmass"], masses["emass"],
                         masses["fmass"]), static=st)
    rt0, _ = oracle.line_search(m, x1, masses["vmass"], masses["emass"],
                                masses["fmass"], static=sm)
    assert np.float32(toi0).view(np.uint32) == np.float32(rt0).view(np.uint32)
    assert rt0 > rt
    ls.close()


def _comb_mesh(cluster=1000, seed=0):
    """Tiny triangles whose centroids quantize to the Morton codes
    2^30 - 2^b (b = 0..29) -- a comb: every split peels one of them off on
    the left and continues on the right -- plus a run of equal codes at the
    top, and one triangle that sweeps the whole box."""
    pts = []
    for b in range(30):
        q = [1024 - 2 ** int(np.ceil(max(b - a, 0) / 3)) for a in range(3)]
        pts.append([min(v, 1023) + 0.5 for v in q])
    pts.append([0.25, 0.25, 0.25])
    rng = np.random.default_rng(seed)
    pts += list(1023.0 + 0.9 * rng.random((cluster, 3)))
    tri = np.array([[0, 0, 0], [0.01, 0, 0], [0, 0.01, 0]])
    x = (np.asarray(pts)[:, None, :] + tri[None]).reshape(-1, 3)
    # a target in the run of equal codes, across the diagonal the mover's
    # first vertex travels
    c = 1023.3
    target = np.array([[c - 0.01, c, c], [c + 0.01, c - 0.01, c], [c, c + 0.01, c]])
    mover = np.array([[0.5, 0.5, 0.5], [0.51, 0.5, 0.5], [0.5, 0.51, 0.5]])
    x = np.concatenate([x, target, mover]).astype(np.float32)
    faces = np.arange(len(x)).reshape(-1, 3).astype(np.uint32)
    m = scenes.contact_mesh(x, faces, ghat=1e-3, offset=0.0)
    x1 = x.copy()
    x1[-3:] += np.float32(1030.0)
    return m, x1


def _walk_stack_need(nodes, root):
    need, st = 1, [root]
    while st:
        v = st.pop()
        if nodes[v][1]:
            need = max(need, len(st) + 2)
            st += [nodes[v][0] - 1, nodes[v][1] - 1]
    return need


def test_line_search_deep_trees_take_the_long_stacks(dev):
    # The comb's trees need more than the short first pass's 36 stack
    # entries for a query that overlaps everything, and at most the 48 of
    # the second pass: the first pass flags the overflow, the second runs
    # every sweep again, and the result is the oracle's bit for bit.
    m, x1 = _comb_mesh()
    needs = [_walk_stack_need(*oracle.bvh_build(m, k)[::2]) for k in (0, 1, 2)]
    print(f"walk stack entries needed (face, edge, vertex trees): {needs}")
    assert max(needs) > 36 and max(needs) <= 48
    ones = [np.ones(m[k], np.float32) for k in ("nv", "ne", "nf")]
    t = mesh_to_device(m, dev)
    ls = LineSearch()
    toi, ov = ls(t, *_to(dev, x1.reshape(-1), *ones))
    rt, ro = oracle.line_search(m, x1, *ones)
    ls.close()
    print(f"deep trees: toi gpu {toi!r} oracle {rt!r} overlap {ov}/{ro}")
    assert np.float32(toi).view(np.uint32) == np.float32(rt).view(np.uint32)
    assert ov == ro
    assert 0.0 < toi < 1.0
