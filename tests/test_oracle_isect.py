"""CPU oracle: the post-step intersection check.

Pinned by geometry computed independently in fp64:
  * face-edge records are exactly the (face, edge) pairs, sharing no vertex,
    whose segment strictly crosses the triangle (signed-volume test);
  * edge-edge records are the edge pairs closer than the sum of their offsets
    (exact segment distance), point-point records the vertex pairs closer
    than theirs;
  * collision-mesh records are the dynamic edges crossing a static triangle;
  * every filter (fixed, zero mass, same rigid body, both colliders) removes
    exactly its pairs, and separated meshes are clear.
"""
import numpy as np

import oracle
from ppf_amd import scenes


def orient(a, b, c, d):
    """Signed volume of (a, b, c, d) in fp64 (vectorised over leading axes)."""
    return np.einsum("...i,...i->...", np.cross(b - a, c - a), d - a)


def crossing_pairs(x, edges, faces):
    """(face, edge) pairs, sharing no vertex, whose segment strictly crosses
    the triangle (fp64)."""
    x = x.astype(np.float64)
    T = x[faces]                                   # (F, 3, 3)
    out = set()
    for ei, (a, b) in enumerate(edges):
        p, q = x[a], x[b]
        s1 = orient(T[:, 0], T[:, 1], T[:, 2], p)
        s2 = orient(T[:, 0], T[:, 1], T[:, 2], q)
        o0 = orient(p, q, T[:, 0], T[:, 1])
        o1 = orient(p, q, T[:, 1], T[:, 2])
        o2 = orient(p, q, T[:, 2], T[:, 0])
        hit = (s1 * s2 < 0) & (((o0 > 0) & (o1 > 0) & (o2 > 0)) |
                               ((o0 < 0) & (o1 < 0) & (o2 < 0)))
        share = np.isin(faces, [a, b]).any(1)
        for fi in np.nonzero(hit & ~share)[0]:
            out.add((int(fi), ei))
    return out


def seg_dist(p1, q1, p2, q2):
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, e, f = d1 @ d1, d2 @ d2, d2 @ r
    c = d1 @ r
    b = d1 @ d2
    den = a * e - b * b
    s = np.clip((b * f - c * e) / den, 0, 1) if den > 1e-30 else 0.0
    t = (b * s + f) / e
    if t < 0:
        t, s = 0.0, np.clip(-c / a, 0, 1)
    elif t > 1:
        t, s = 1.0, np.clip((b - c) / a, 0, 1)
    return np.linalg.norm((p1 + d1 * s) - (p2 + d2 * t))


def crossing_sheets(res=10, seed=0):
    """A flat sheet and a second sheet tilted through it."""
    x, f = scenes.stacked_sheets(res, layers=2, gap=0.0, noise=1e-4, seed=seed,
                                 shifts=[(0, 0), (0.37, 0.21)])
    n = res * res
    x = x.astype(np.float64)
    x[n:, 2] += (x[n:, 0] - 0.05) * 0.4          # tilt the second sheet
    return x.astype(np.float32), f


def ones(m):
    return [np.ones(m[k], np.float32) for k in ("nv", "ne", "nf")]


def test_face_edge_records_are_the_crossings():
    x, f = crossing_sheets()
    m = scenes.contact_mesh(x, f, ghat=1e-3)
    n, rec = oracle.intersection_check(m, *ones(m), cap=100000)
    got = {(int(r["elem0"]), int(r["elem1"])) for r in rec if r["type"] == 0}
    want = crossing_pairs(x, m["edge"], m["face"])
    assert n == len(rec) and len(want) > 10
    assert got == want
    assert all(r["type"] == 0 for r in rec)      # offset 0: no distance records
    # Record payload: the face's vertices then the edge's.
    r = rec[0]
    fv = m["face"][r["elem0"]]
    ev = m["edge"][r["elem1"]]
    assert r["num_verts0"] == 3 and r["num_verts1"] == 2
    assert np.array_equal(r["positions"].reshape(5, 3), x[np.r_[fv, ev]])


def test_offset_records_edge_edge_and_point_point():
    # Two parallel sheets 2e-4 apart with offsets 1.5e-4 per element: edge
    # pairs and vertex pairs closer than 3e-4 are recorded.
    x, f = scenes.stacked_sheets(8, layers=2, gap=2e-4, noise=3e-5, seed=3,
                                 shifts=[(0, 0), (0.008, 0.006)])
    m = scenes.contact_mesh(x, f, ghat=1e-3, offset=1.5e-4)
    n, rec = oracle.intersection_check(m, *ones(m), cap=100000)
    assert n == len(rec)
    ee = {(int(r["elem0"]), int(r["elem1"])) for r in rec if r["type"] == 1}
    pp = {(int(r["elem0"]), int(r["elem1"])) for r in rec if r["type"] == 3}
    X = x.astype(np.float64)
    E = m["edge"]
    want_ee = set()
    for a in range(len(E)):
        for b in range(a):
            ea, eb = E[a], E[b]
            if set(ea.tolist()) & set(eb.tolist()):
                continue                       # shared vertex: separate rule
            if seg_dist(X[ea[0]], X[ea[1]], X[eb[0]], X[eb[1]]) < 3e-4:
                want_ee.add((a, b))
    ee_free = {p for p in ee if not set(E[p[0]].tolist()) & set(E[p[1]].tolist())}
    assert ee_free == want_ee and len(want_ee) > 0
    d = np.linalg.norm(X[:, None] - X[None], axis=2)
    i, j = np.nonzero(np.tril(d < 3e-4, -1))
    assert pp == set(zip(i.tolist(), j.tolist())) and len(pp) > 0


def test_filters_and_clear():
    x, f = crossing_sheets(seed=1)
    base = scenes.contact_mesh(x, f, ghat=1e-3)
    nv, ne, nf = base["nv"], base["ne"], base["nf"]
    n0, _ = oracle.intersection_check(base, *ones(base))
    assert n0 > 0

    def run(masses=None, **kw):
        m = dict(base)
        m.update(kw)
        return oracle.intersection_check(m, *(masses or ones(base)))[0]

    assert run(vbody=np.full(nv, 3, np.uint32)) == 0
    assert run(vflag=np.full(nv, 2, np.uint32)) == 0
    assert run(eflag=np.ones(ne, np.uint32), fflag=np.ones(nf, np.uint32)) == 0
    assert run(masses=[np.zeros(nv, np.float32), np.zeros(ne, np.float32),
                       np.zeros(nf, np.float32)]) == 0
    # One sheet colliders only: cross-sheet pairs still count.
    vf = np.zeros(nv, np.uint32)
    vf[: nv // 2] = 2
    assert run(vflag=vf) == n0
    # Separated sheets are clear.
    y = x.copy()
    y[nv // 2:, 2] += 0.2
    assert run(x=y) == 0


def test_collision_mesh_crossings():
    m, x1, vm, em, fm, sm = scenes.static_feature_scene("floor")
    # The sheet after falling through the floor, tilted so edges cross it.
    y = x1.astype(np.float64)
    y[:, 2] = -1e-3 + (y[:, 0] - 0.0) * 0.05
    m2 = dict(m)
    m2["x"] = y.astype(np.float32)
    n, rec = oracle.intersection_check(m2, vm, em, fm, static=sm, cap=100000)
    got = {int(r["elem1"]) for r in rec if r["type"] == 2}
    s0, s1, s2 = sm["x"].astype(np.float64)[sm["face"][0]]
    yy = m2["x"].astype(np.float64)
    want = set()
    for ei, (a, b) in enumerate(m["edge"]):
        v1, v2 = orient(s0, s1, s2, yy[a]), orient(s0, s1, s2, yy[b])
        if v1 * v2 < 0:
            want.add(ei)
    assert got == want and len(want) > 0
    # Massless edges do not test against the collision mesh.
    n0, rec0 = oracle.intersection_check(m2, vm, np.zeros_like(em), fm, static=sm,
                                         cap=100000)
    assert not any(r["type"] == 2 for r in rec0)
