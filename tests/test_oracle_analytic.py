"""The oracle's analytic sphere / floor colliders (contact.cu:1154-1382,
2318-2470) against what the mathematics fixes.

- Barrier: for a vertex inside the contact gap ghat of a floor, with no
  elastic block, the stiffness is m / gap^2 and the push barrier's gradient
  and Hessian are k (-d^2 / ghat) up and k (-2 d / ghat) up up^T, d the depth
  below the shifted plane (ground + ghat up); with the stiffness held fixed
  the gradient is the derivative of the energy k (-d^3 / (3 ghat)) (central
  differences in fp64). The same along the radial normal of a sphere, for a
  reversed sphere (inside), a bowl (the upper half a cylinder: the centre
  moves to the vertex's height) and a kinematic sphere (stiffness m / ghat^2).
  Outside the gap and past the thickness nothing acts.
- Friction: a vertex sliding along the floor feels lam u with lam =
  mu |f_push| / max(eps, |u|) (kinetic) against its tangential displacement.
- Line search: a vertex swept straight through a floor or radially through a
  sphere reaches it at the closed-form fraction h0 / (h0 - h1),
  (r0 - r) / (r0 - r1) of the max_t-extended sweep; kinematic colliders are
  not swept; a pinned vertex driven from outside to inside a static collider
  is flagged (the smallest index), one already inside or staying outside is
  not.
"""
import numpy as np

import oracle
from ppf_amd import scenes


def _one_vertex(x, x0=None, mass=0.01, fric=0.0):
    x = np.asarray(x, np.float32).reshape(1, 3)
    m = scenes.contact_mesh(x, np.zeros((0, 3), np.uint32))
    x0 = x if x0 is None else np.asarray(x0, np.float32).reshape(1, 3)
    return m, x0, np.array([mass], np.float32), np.array([fric], np.float32)


def _push(k, d, ghat, n):
    n = np.asarray(n, np.float64)
    g = k * (-(d * d) / ghat) * n
    H = k * (-2.0 * d / ghat) * np.outer(n, n)
    return g, H


def test_floor_barrier_closed_form_and_energy():
    ghat, mass = 1e-3, 0.01
    up = np.array([0.0, 1.0, 0.0])
    for h in (2e-4, 5e-4, 9e-4):
        m, x0, ms, fr = _one_vertex([0.1, h, -0.2], mass=mass)
        floor = dict(ground=(0, 0, 0), ghat=ghat, friction=0.0, up=tuple(up))
        g, H, nc = oracle.analytic_assemble(m, x0, ms, fr, floors=[floor])
        k = mass / (h * h)
        d = h - ghat
        gw, Hw = _push(k, d, ghat, up)
        assert nc == 1
        assert np.allclose(g, gw, rtol=2e-6, atol=0)
        assert np.allclose(H[0].reshape(3, 3).T, Hw, rtol=2e-6, atol=1e-12)
        # gradient of k (-d^3 / (3 ghat)) at fixed k, central differences
        eps = 1e-7
        E = lambda y: k * (-(y - ghat) ** 3 / (3 * ghat))
        fd = (E(h + eps) - E(h - eps)) / (2 * eps)
        assert abs(g[1] - fd) <= 1e-4 * abs(fd)
    # outside the gap: nothing
    m, x0, ms, fr = _one_vertex([0.0, 1.5e-3, 0.0])
    g, H, nc = oracle.analytic_assemble(m, x0, ms, fr, floors=[dict(ground=(0, 0, 0), ghat=ghat,
                                                                   friction=0.0)])
    assert nc == 0 and not g.any() and not H.any()
    # deeper than the thickness: passes through
    m, x0, ms, fr = _one_vertex([0.0, -0.02, 0.0])
    g, H, nc = oracle.analytic_assemble(m, x0, ms, fr, floors=[dict(
        ground=(0, 0, 0), ghat=ghat, friction=0.0, thickness=0.01)])
    assert nc == 0 and not g.any()


def test_tilted_floor_and_kinematic_gap_floor():
    ghat, mass = 2e-3, 0.02
    up = np.array([1.0, 2.0, -2.0]) / 3.0
    ground = np.array([0.3, -0.1, 0.2])
    h = 7e-4
    x = ground + h * up + np.array([0.05, 0.0, 0.0]) - 0.05 * up[0] * up  # still h above
    h = float((x - ground) @ up)
    m, x0, ms, fr = _one_vertex(x, mass=mass)
    g, H, _ = oracle.analytic_assemble(m, x0, ms, fr, floors=[dict(
        ground=tuple(ground), ghat=ghat, friction=0.0, up=tuple(up))])
    gw, Hw = _push(mass / (h * h), h - ghat, ghat, up)
    assert np.allclose(g, gw, rtol=5e-5)
    assert np.allclose(H[0].reshape(3, 3).T, Hw, rtol=5e-5, atol=1e-9)
    # a kinematic floor clamps the gap from below at constraint_tol * ghat
    m, x0, ms, fr = _one_vertex([0.0, 1e-6, 0.0], mass=mass)
    g, _, _ = oracle.analytic_assemble(m, x0, ms, fr, floors=[dict(
        ground=(0, 0, 0), ghat=ghat, friction=0.0, kinematic=True)], constraint_tol=0.01)
    k = mass / (0.01 * ghat) ** 2
    gw, _ = _push(k, 1e-6 - ghat, ghat, (0, 1, 0))
    assert np.allclose(g, gw, rtol=5e-5)


def _radial(c, x):
    """Distance and unit direction of the fp32-rounded vertex from the fp32
    centre (in fp64: the closed forms see the oracle's actual inputs)."""
    d = np.asarray(x, np.float32).astype(np.float64) - np.asarray(c, np.float32).astype(np.float64)
    r = np.linalg.norm(d)
    return r, d / r


def test_sphere_barriers_closed_form():
    # A small sphere: the fp32 gap |x - c| - R loses about log2(R / gap) bits
    # to cancellation, so R = 0.02 keeps the closed forms tight (at R = 0.3 a
    # 4e-4 gap is good to ~1e-4 only, on any fp32 implementation).
    ghat, mass, R = 1e-3, 0.01, 0.02
    c = np.array([0.01, -0.02, 0.005])
    n0 = np.array([2.0, -1.0, 2.0]) / 3.0
    # solid sphere: vertex just outside the surface
    x = c + (R + 4e-4) * n0
    r, n = _radial(c, x)
    m, x0, ms, fr = _one_vertex(x, mass=mass)
    sphere = dict(center=tuple(c), ghat=ghat, friction=0.0, radius=R)
    g, H, nc = oracle.analytic_assemble(m, x0, ms, fr, spheres=[sphere])
    gw, Hw = _push(mass / (r - R) ** 2, r - (R + ghat), ghat, n)
    assert nc == 1 and np.allclose(g, gw, rtol=2e-5)
    assert np.allclose(H[0].reshape(3, 3).T, Hw, rtol=2e-5, atol=1e-9)
    # reversed (a container): vertex just inside, pushed inwards
    x = c + (R - 4e-4) * n0
    r, n = _radial(c, x)
    m, x0, ms, fr = _one_vertex(x, mass=mass)
    g, _, _ = oracle.analytic_assemble(m, x0, ms, fr, spheres=[dict(sphere, reverse=True)])
    # o - proj = (r - (R - ghat)) n, normal -n: d = -(r - R + ghat) < 0
    gw, _ = _push(mass / (R - r) ** 2, -(r - (R - ghat)), ghat, -n)
    assert np.allclose(g, gw, rtol=2e-5)
    # bowl above the centre: a vertical cylinder wall (centre at the vertex's height)
    xb = c + np.array([R - 3e-4, 0.05, 0.0])
    cb = np.array([c[0], np.float32(xb[1]), c[2]])
    rb, _ = _radial(cb, xb)
    m, x0, ms, fr = _one_vertex(xb, mass=mass)
    g, _, nc = oracle.analytic_assemble(m, x0, ms, fr, spheres=[dict(sphere, bowl=True,
                                                                     reverse=True)])
    gw, _ = _push(mass / (R - rb) ** 2, -(rb - (R - ghat)), ghat, (-1.0, 0.0, 0.0))
    assert nc == 1 and np.allclose(g, gw, rtol=2e-5, atol=1e-9)
    # kinematic: stiffness m / ghat^2, target on the shifted surface
    x = c + (R + 6e-4) * n0
    r, n = _radial(c, x)
    m, x0, ms, fr = _one_vertex(x, mass=mass)
    g, _, nc = oracle.analytic_assemble(m, x0, ms, fr, spheres=[dict(sphere, kinematic=True)])
    gw, _ = _push(mass / ghat ** 2, r - (R + ghat), ghat, n)
    assert nc == 1 and np.allclose(g, gw, rtol=5e-4)
    # pinned or massless vertices are skipped
    m, x0, ms, fr = _one_vertex(c + (R + 4e-4) * n0, mass=0.0)
    g, _, nc = oracle.analytic_assemble(m, x0, ms, fr, spheres=[sphere])
    assert nc == 0 and not g.any()


def test_floor_friction_kinetic():
    ghat, mass, mu = 1e-3, 0.01, 0.5
    h = 5e-4
    slide = np.array([3e-3, 0.0, -4e-3])  # |u| = 5e-3 > eps: kinetic
    x = np.array([0.0, h, 0.0])
    m, x0, ms, fr = _one_vertex(x, x0=x - slide, mass=mass, fric=mu)
    g, H, _ = oracle.analytic_assemble(m, x0, ms, fr, floors=[dict(
        ground=(0, 0, 0), ghat=ghat, friction=mu)], friction_eps=1e-5)
    k = mass / (h * h)
    fp, Hp = _push(k, h - ghat, ghat, (0, 1, 0))
    lam = mu * np.linalg.norm(fp) / np.linalg.norm(slide)
    assert np.allclose(g, fp + lam * slide, rtol=5e-5)
    s = slide / np.linalg.norm(slide)
    w = np.cross([0, 1, 0], s)
    assert np.allclose(H[0].reshape(3, 3).T, Hp + lam * np.outer(w, w), rtol=5e-5, atol=1e-9)


def test_line_search_closed_forms_and_pins():
    max_t = 1.25
    floor = dict(ground=(0, 0, 0), ghat=1e-3, friction=0.0)
    x0 = np.array([[0.2, 0.01, 0.0]], np.float32)
    x1 = np.array([[0.2, -0.006, 0.0]], np.float32)
    ext = x0 + max_t * (x1 - x0)
    toi, bad = oracle.analytic_toi(x0, x1, [0], [0.01], floors=[floor], max_t=max_t)
    t = 0.01 / (0.01 - ext[0, 1])
    assert bad is None and abs(toi - max_t * t) <= 1e-6
    # kinematic floors are not swept; a vertex above stays unclamped
    toi, _ = oracle.analytic_toi(x0, x1, [0], [0.01], floors=[dict(floor, kinematic=True)],
                                 max_t=max_t)
    assert toi == np.float32(max_t)
    # radially through a sphere
    c = np.array([0.0, 0.0, 0.0])
    n = np.array([0.0, 0.6, 0.8])
    sp = dict(center=tuple(c), ghat=1e-3, friction=0.0, radius=0.3)
    a0 = (c + 0.34 * n)[None].astype(np.float32)
    a1 = (c + 0.30 * n)[None].astype(np.float32)
    toi, bad = oracle.analytic_toi(a0, a1, [0], [0.01], spheres=[sp], max_t=max_t)
    r1 = 0.34 + max_t * (0.30 - 0.34)
    assert bad is None and abs(toi - max_t * (0.34 - 0.3) / (0.34 - r1)) <= 2e-6
    # pins: vertex 1 driven through the floor, vertex 0 stays above, vertex 2
    # starts below (already inside: not a crossing)
    p0 = np.array([[0, 0.01, 0], [0, 0.002, 0], [0, -0.01, 0]], np.float32)
    p1 = np.array([[0, 0.009, 0], [0, -0.002, 0], [0, -0.02, 0]], np.float32)
    _, bad = oracle.analytic_toi(p0, p1, [1, 1, 1], [0, 0, 0], floors=[floor], max_t=max_t)
    assert bad == 1
    _, bad = oracle.analytic_toi(p0, p1, [1, 1, 1], [0, 0, 0],
                                 floors=[dict(floor, kinematic=True)], max_t=max_t)
    assert bad is None


def test_cloth_on_floor_stays_above_it():
    """The oracle's time steps of a cloth dropped onto an analytic floor
    (no sphere mesh): every vertex stays on the feasible side of the floor,
    the cloth comes to rest on it."""
    sc = scenes.drape_scene(res=10, height=0.004)
    sc["static"] = None
    y_floor = float(sc["mesh"]["x"][:, 1].min()) - 0.003
    sc["floors"] = [dict(ground=(0.0, y_floor, 0.0), ghat=1e-3, friction=0.5)]
    st = oracle.OracleStepper(sc)
    p = scenes.step_params()
    lowest = []
    for _ in range(16):
        r = st.advance(p)
        assert r["status"] == 0, r
        p["prev_dt"] = r["dt"]
        lowest.append(float(sc["mesh"]["x"][:, 1].min()) - y_floor)
    assert min(lowest) > 0.0
    assert lowest[-1] < 1e-3  # resting within the contact gap


def test_pinned_vertex_through_floor_fails_the_step():
    res = 8
    sc = scenes.drape_scene(res=res, height=0.004, pins=[0])
    sc["static"] = None
    y = float(sc["mesh"]["x"][0, 1])
    sc["floors"] = [dict(ground=(0.0, y - 0.001, 0.0), ghat=1e-3, friction=0.0)]
    sc["fix_pos"][0, 1] = y - 0.01  # driven below the floor in one step
    st = oracle.OracleStepper(sc)
    r = st.advance(scenes.step_params())
    assert r["status"] == 6
