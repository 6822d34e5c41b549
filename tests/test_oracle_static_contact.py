"""CPU oracle: contact against the static collision mesh (pair kinds 4-6).

Pinned by:
  * brute force: the enumerated pairs are exactly the (vertex, static face),
    (static vertex, face) and (edge, static edge) pairs closer than
    offset + ghat (fp64 distances; pairs within 1e-6 relative of the threshold
    are not asserted), with the pinned / massless / fixed filters;
  * finite differences of the fp64 push energy -k d^3 / (3 ghat),
    d = dist - offset - ghat, at frozen stiffness: with friction off the
    assembled force is its gradient for every kind;
  * closed forms: the stiffness of a vertex with no elastic term is m / gap^2;
    the kinetic friction force is tangential with magnitude mu |f_n|.
"""
import numpy as np
import pytest

import oracle
from ppf_amd import scenes


def seg_dist(p1, q1, p2, q2):
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, e, f = d1 @ d1, d2 @ d2, d2 @ r
    c, b = d1 @ r, d1 @ d2
    den = a * e - b * b
    s = np.clip((b * f - c * e) / den, 0, 1) if den > 1e-30 else 0.0
    t = (b * s + f) / e
    if t < 0:
        t, s = 0.0, np.clip(-c / a, 0, 1)
    elif t > 1:
        t, s = 1.0, np.clip((b - c) / a, 0, 1)
    return np.linalg.norm((p1 + d1 * s) - (p2 + d2 * t))


def point_seg(p, a, b):
    d = b - a
    t = np.clip((p - a) @ d / (d @ d), 0, 1)
    return np.linalg.norm(p - (a + t * d))


def point_tri(p, t0, t1, t2):
    n = np.cross(t1 - t0, t2 - t0)
    nn = n @ n
    u = np.cross(p - t0, t2 - t0) @ n / nn
    v = np.cross(t1 - t0, p - t0) @ n / nn
    if u >= 0 and v >= 0 and u + v <= 1:
        return abs((p - t0) @ n) / np.sqrt(nn)
    return min(point_seg(p, t0, t1), point_seg(p, t1, t2), point_seg(p, t2, t0))


def near_scene(feature, height=4e-4, res=8, tilt=0.0):
    """A sheet hovering `height` above the static feature (inside the
    1e-3 contact gap), optionally tilted so distances vary."""
    m, x1, vm, em, fm, sm = scenes.static_feature_scene(feature, res=res)
    x = m["x"].astype(np.float64)
    # the features' tops sit 1e-3 below the sheet: lower the sheet to `height`
    x[:, 2] -= 1e-3 - height
    x[:, 2] += tilt * x[:, 0]
    m = dict(m)
    m["x"] = x.astype(np.float32)
    rng = np.random.default_rng(1)
    sm = dict(sm)
    for k, n in (("v", sm["nv"]), ("e", sm["ne"]), ("f", sm["nf"])):
        sm[k + "_fric"] = rng.uniform(0.1, 0.6, n).astype(np.float32)
    return m, vm, em, fm, sm


def brute_pairs(m, sm, vm, em, fm):
    X, S = m["x"].astype(np.float64), sm["x"].astype(np.float64)
    want, near = set(), set()

    def test(key, d, r):
        if abs(d - r) < 1e-6 * r:
            near.add(key)
        elif d < r:
            want.add(key)

    for i in range(m["nv"]):
        if (m["vflag"][i] & 1) or vm[i] <= 0:
            continue
        for fi, f in enumerate(sm["face"]):
            r = (m["v_off"][i] + sm["f_off"][fi]) + 0.5 * (m["v_ghat"][i] + sm["f_ghat"][fi])
            test((4, i, fi), point_tri(X[i], *S[f]), r)
    for sv in range(sm["nv"]):
        for fi, f in enumerate(m["face"]):
            if (m["fflag"][fi] & 1) or fm[fi] <= 0:
                continue
            r = (m["f_off"][fi] + sm["v_off"][sv]) + 0.5 * (m["f_ghat"][fi] + sm["v_ghat"][sv])
            test((5, sv, fi), point_tri(S[sv], *X[f]), r)
    for ei, e in enumerate(m["edge"]):
        if (m["eflag"][ei] & 1) or em[ei] <= 0:
            continue
        for sj, s in enumerate(sm["edge"]):
            r = (m["e_off"][ei] + sm["e_off"][sj]) + 0.5 * (m["e_ghat"][ei] + sm["e_ghat"][sj])
            test((6, ei, sj), seg_dist(X[e[0]], X[e[1]], S[s[0]], S[s[1]]), r)
    return want, near


@pytest.mark.parametrize("feature,tilt", [("floor", 0.0), ("floor", 0.01), ("spike", 0.0),
                                          ("ridge", 0.0)])
def test_static_pairs_brute_force(feature, tilt):
    m, vm, em, fm, sm = near_scene(feature, tilt=tilt)
    # a few filtered elements
    vm = vm.copy()
    vm[::11] = 0.0
    m["vflag"] = m["vflag"].copy()
    m["vflag"][5::13] |= 1
    got = oracle.static_proximity(m, sm, vm, em, fm)
    gs = {(int(p["kind"]), int(p["a"]), int(p["b"])) for p in got}
    assert len(gs) == len(got) and all(p["count"] == 1 for p in got)
    want, near = brute_pairs(m, sm, vm, em, fm)
    assert want - gs == set()
    assert gs - want - near == set()
    kinds = {k[0] for k in gs}
    print(feature, len(gs), sorted(kinds))
    assert len(want) > 0


def push_energy(d, k, gh):
    return -k * d ** 3 / (3 * gh) if d < 0 else 0.0


def test_push_force_is_energy_gradient():
    # One pair of each kind in isolation, friction off, no elastic term.
    rng = np.random.default_rng(3)
    gh, off = 1e-3, 5e-5
    # kind 4: vertex over a static triangle
    tri = np.array([[-0.01, -0.01, 0], [0.01, -0.01, 0], [0, 0.01, 0]], np.float32)
    for kind in (4, 5, 6):
        for trial in range(5):
            if kind == 4:
                p = np.array([rng.uniform(-2e-3, 2e-3), rng.uniform(-2e-3, 2e-3),
                              rng.uniform(2e-4, 9e-4)], np.float32)
                x = p[None]
                faces = np.zeros((0, 3), np.uint32)
                mesh = scenes.contact_mesh(x, faces, ghat=gh, offset=off)
                smesh = scenes.static_mesh(tri, np.array([[0, 1, 2]], np.uint32), ghat=gh,
                                           offset=off)
                pair = (4, 0, 0)
            elif kind == 5:
                x = tri.copy()
                x[:, 2] = rng.uniform(2e-4, 9e-4) + rng.uniform(-5e-5, 5e-5, 3)
                mesh = scenes.contact_mesh(x, np.array([[0, 1, 2]], np.uint32), ghat=gh,
                                           offset=off)
                smesh = scenes.static_mesh(np.array([[1e-3, 5e-4, 0.0]], np.float32),
                                           np.zeros((0, 3), np.uint32), ghat=gh, offset=off)
                pair = (5, 0, 0)
            else:
                x = np.array([[-0.01, 0.0, 0.0], [0.01, 0.0, 0.0]], np.float32)
                x[:, 2] = rng.uniform(2e-4, 9e-4) + rng.uniform(-5e-5, 5e-5, 2)
                x[:, 1] += rng.uniform(-1e-3, 1e-3, 2)
                mesh = scenes.contact_mesh(x, np.zeros((0, 3), np.uint32), ghat=gh, offset=off)
                mesh["edge"] = np.array([[0, 1]], np.uint32)
                mesh["ne"] = 1
                for k in ("e_ghat", "e_off"):
                    mesh[k] = np.full(1, gh if "ghat" in k else off, np.float32)
                mesh["eflag"] = np.zeros(1, np.uint32)
                mesh["e2f_off"] = np.zeros(2, np.uint32)
                smesh = scenes.static_mesh(np.array([[0.0, -0.01, 0.0], [0.0, 0.01, 0.0]],
                                                    np.float32),
                                           np.zeros((0, 3), np.uint32), ghat=gh, offset=off)
                smesh["edge"] = np.array([[0, 1]], np.uint32)
                smesh["ne"] = 1
                smesh["e_ghat"] = np.full(1, gh, np.float32)
                smesh["e_off"] = np.full(1, off, np.float32)
                pair = (6, 0, 0)
            for k, n in (("v", smesh["nv"]), ("e", smesh["ne"]), ("f", smesh["nf"])):
                smesh[k + "_fric"] = np.zeros(max(n, 1), np.float32)
            nv = mesh["nv"]
            mass = np.full(nv, 2e-4, np.float32)
            fric = {"v": np.zeros(nv, np.float32), "e": np.zeros(max(1, mesh["ne"]), np.float32),
                    "f": np.zeros(max(1, mesh["nf"]), np.float32)}
            pr = np.zeros(1, oracle.PAIR_DTYPE)
            pr["kind"], pr["a"], pr["b"], pr["count"] = pair[0], pair[1], pair[2], 1
            grad, rows, cols, vals = oracle.contact_assemble(mesh, x, mass, fric, pr,
                                                             static=smesh)
            X = x.astype(np.float64)
            S = smesh["x"].astype(np.float64)

            def dist(Y):
                if kind == 4:
                    return point_tri(Y[0], *S[[0, 1, 2]])
                if kind == 5:
                    return point_tri(S[0], *Y[[0, 1, 2]])
                return seg_dist(Y[0], Y[1], S[0], S[1])

            d0 = dist(X)
            gap = d0 - 2 * off
            # stiffness with no elastic term: sum c_i^2 m_i / (gap^2 sum c_i^2) =
            # m / gap^2 for equal masses
            kst = 2e-4 / gap ** 2
            r = 2 * off + gh
            g = np.zeros_like(X)
            h = 1e-8
            for i in range(len(X)):
                for dd in range(3):
                    Yp, Ym = X.copy(), X.copy()
                    Yp[i, dd] += h
                    Ym[i, dd] -= h
                    g[i, dd] = (push_energy(dist(Yp) - r, kst, gh) -
                                push_energy(dist(Ym) - r, kst, gh)) / (2 * h)
            err = np.abs(grad.reshape(-1, 3) - g).max() / np.abs(g).max()
            assert err < 2e-3, (kind, trial, err)


def test_kinetic_friction_closed_form():
    # A vertex sliding over a static floor: the friction force is tangential,
    # opposes the slip and has magnitude mu * |normal force|.
    gh, off, mu = 1e-3, 0.0, 0.4
    tri = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32)
    x = np.array([[0.0, 0.0, 3e-4]], np.float32)
    x0 = np.array([[-1e-3, 0.0, 3e-4]], np.float32)     # slip +x by 1e-3
    mesh = scenes.contact_mesh(x, np.zeros((0, 3), np.uint32), ghat=gh, offset=off)
    smesh = scenes.static_mesh(tri, np.array([[0, 1, 2]], np.uint32), ghat=gh, offset=off)
    smesh["v_fric"] = np.full(3, mu, np.float32)
    smesh["e_fric"] = np.full(3, mu, np.float32)
    smesh["f_fric"] = np.full(1, mu, np.float32)
    pr = np.zeros(1, oracle.PAIR_DTYPE)
    pr["kind"], pr["count"] = 4, 1
    fric = {"v": np.full(1, mu, np.float32), "e": np.zeros(1, np.float32),
            "f": np.zeros(1, np.float32)}
    grad, _, _, _ = oracle.contact_assemble(mesh, x0, np.full(1, 1e-4, np.float32), fric, pr,
                                            friction_eps=1e-5, static=smesh)
    fric0 = dict(fric, v=np.zeros(1, np.float32))
    g0, _, _, _ = oracle.contact_assemble(mesh, x0, np.full(1, 1e-4, np.float32), fric0, pr,
                                          friction_eps=1e-5, static=smesh)
    fn = g0.reshape(3)
    ft = grad.reshape(3) - fn
    assert abs(fn[0]) < 1e-12 and abs(fn[1]) < 1e-12 and fn[2] < 0
    assert ft[2] == 0 and abs(ft[1]) < 1e-12 and ft[0] > 0     # gradient opposes -x drag
    assert np.isclose(np.linalg.norm(ft), mu * abs(fn[2]), rtol=1e-5)
