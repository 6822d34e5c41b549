"""CPU oracle: contact barrier + friction gradient and Hessian assembly.

Pins: fp64 closed forms of a single point-face contact (cubic barrier force and
curvature, the elasticity-free stiffness m/g^2, both friction branches),
finite differences of the barrier energy, and the PSD structure of the
assembled contact Hessian.
"""
import numpy as np

import oracle
from ppf_amd import scenes


def single_pf(h, slip, mass=2e-4, ghat=1e-3, mu=0.3, eps=1e-5):
    tri = np.array([[0, 0, 0], [0.01, 0, 0], [0, 0.01, 0]], np.float32)
    x = np.concatenate([tri, [[0.003, 0.002, h]]]).astype(np.float32)
    m = scenes.contact_mesh(x, np.array([[0, 1, 2]], np.uint32), ghat=ghat)
    x0 = x.copy()
    x0[3] -= np.array(slip, np.float32)
    fric = {"v": np.full(4, mu, np.float32), "e": np.full(3, mu, np.float32),
            "f": np.full(1, mu, np.float32)}
    massv = np.full(4, mass, np.float32)
    pairs = np.array([(0, 3, 0, 1)], dtype=oracle.PAIR_DTYPE)
    return m, x0, massv, fric, pairs, eps


def test_single_point_face_closed_form():
    ghat, mass, h = 1e-3, 2e-4, 4e-4
    for slip, kinetic in (((3e-6, 1e-6, 0.0), False), ((4e-5, -2e-5, 0.0), True)):
        m, x0, massv, fric, pairs, eps = single_pf(h, slip, mass=mass, ghat=ghat)
        grad, rows, cols, vals = oracle.contact_assemble(m, x0, massv, fric, pairs,
                                                         friction_eps=eps)
        w = np.array([1.0, -0.5, -0.3, -0.2])  # barycentric of (0.003, 0.002)
        g = h
        k = mass / g ** 2                        # equal masses, no elastic term
        y = g - ghat
        fn = k * (-2 * y * y / ghat)             # d/dg of -2 y^3 / (3 ghat)
        n = np.array([0, 0, 1.0])
        f = fn * n
        H = k * 4 * (1 - g / ghat) * np.outer(n, n)
        contact = -n @ f
        u = np.array(slip)
        u = u - n * (n @ u)
        un = np.linalg.norm(u)
        lam = 0.3 * contact / max(eps, un)
        f = f + lam * u
        if kinetic:
            s = u / un
            wd = np.cross(n, s)
            H = H + lam * np.outer(wd, wd)
        else:
            H = H + lam * (np.eye(3) - np.outer(n, n))
        idx = [3, 0, 1, 2]
        for ii, v in enumerate(idx):
            assert np.allclose(grad[3 * v:3 * v + 3], w[ii] * f, rtol=2e-4,
                               atol=1e-9 * abs(fn)), (v, grad[3 * v:3 * v + 3], w[ii] * f)
        blocks = {(int(r), int(c)): vals[k2].reshape(3, 3).T for k2, (r, c) in
                  enumerate(zip(rows, cols))}
        assert len(blocks) == 10
        for ii in range(4):
            for jj in range(4):
                i, j = idx[ii], idx[jj]
                if i <= j:
                    assert np.allclose(blocks[(i, j)], w[ii] * w[jj] * H, rtol=2e-4,
                                       atol=1e-6 * np.abs(H).max())


def test_barrier_gradient_matches_finite_difference():
    # Without friction (slip 0 leaves u = 0) the gradient on the query vertex
    # is k * dE/dz with the stiffness held fixed, E = -2 (g - ghat)^3 / (3 ghat).
    ghat, mass = 1e-3, 2e-4
    for h in (1e-4, 5e-4, 9e-4):
        m, x0, massv, fric, pairs, eps = single_pf(h, (0, 0, 0), mass=mass, ghat=ghat)
        grad, *_ = oracle.contact_assemble(m, x0, massv, fric, pairs)
        k = mass / h ** 2
        E = lambda z: -2 * min(z - ghat, 0.0) ** 3 / (3 * ghat)
        d = 1e-9
        fd = k * (E(h + d) - E(h - d)) / (2 * d)
        assert abs(grad[3 * 3 + 2] - fd) <= 1e-3 * abs(fd)
        # Outside the gap the barrier is flat.
    m, x0, massv, fric, pairs, eps = single_pf(1.2e-3, (0, 0, 0), ghat=ghat)
    grad, rows, cols, vals = oracle.contact_assemble(m, x0, massv, fric, pairs)
    assert not np.any(grad) and not np.any(vals)


def test_assembled_contact_hessian_is_psd_and_symmetric():
    m, x0, mass, fric, elastic = scenes.contact_scenario(res=12)
    pairs = oracle.proximity(m)
    assert len(pairs) > 50
    grad, rows, cols, vals = oracle.contact_assemble(m, x0, mass, fric, pairs,
                                                     elastic=elastic)
    assert np.all(np.isfinite(grad)) and np.all(np.isfinite(vals))
    nv = m["nv"]
    M = np.zeros((3 * nv, 3 * nv))
    for (r, c), b in zip(zip(rows, cols), vals):
        blk = b.reshape(3, 3).T.astype(np.float64)
        M[3 * r:3 * r + 3, 3 * c:3 * c + 3] += blk
        if r != c:
            M[3 * c:3 * c + 3, 3 * r:3 * r + 3] += blk.T
    assert np.allclose(M, M.T, atol=1e-6 * np.abs(M).max())
    ev = np.linalg.eigvalsh(M)
    assert ev.min() >= -1e-5 * ev.max()
    # Rows sorted, columns ascending within a row, upper triangle only.
    key = rows.astype(np.int64) * nv + cols
    assert np.all(np.diff(key) > 0) and np.all(rows <= cols)


def _elastic_upper(K, nv):
    """Upper block CSR (diagonal included) of a dense (3nv, 3nv) matrix, the
    layout of the elastic snapshot the stiffness reads (blocks column-major)."""
    rp, col, val = [0], [], []
    for i in range(nv):
        for j in range(i, nv):
            col.append(j)
            val.append(K[3 * i:3 * i + 3, 3 * j:3 * j + 3].T.reshape(-1))
        rp.append(len(col))
    return (np.array(rp, np.uint32), np.array(col, np.uint32),
            np.array(val, np.float32).reshape(-1))


def _spd(rng, d):
    a = rng.standard_normal((d, d))
    return (a @ a.T + d * np.eye(d)) * 50.0


def _stiffness_from_force(grad, v, w_v, n, g, ghat):
    # barrier force on vertex v = w_v * k * cubic'(g) * n (no friction: mu = 0)
    gp = -2.0 * (g - ghat) ** 2 / ghat
    comp = int(np.argmax(np.abs(n)))
    return grad[3 * v + comp] / (w_v * gp * n[comp])


def test_elastic_inclusive_stiffness_closed_form():
    """barrier.cu:49-86: k = w^T (H_el + diag(m_i / g^2)) w with w the
    normalised stack of w_i * n over the contact's vertices and a zero mass
    replaced by the contact's largest mass. Worked here in fp64 from the
    elastic blocks themselves: every block K_ij of the contact's vertices
    (distinct values per block, off-diagonal blocks not symmetric) enters with
    weight w_i w_j, so a block read from the wrong row or column, a dropped
    inertia term or a missing zero-mass substitution changes k."""
    rng = np.random.default_rng(11)
    ghat, h = 1e-3, 4e-4
    nofric = {"v": np.zeros(4, np.float32), "e": np.zeros(3, np.float32),
              "f": np.zeros(1, np.float32)}
    # Point-face: query vertex 3 over face (0, 1, 2), n = +z, weights
    # (1, -0.5, -0.3, -0.2) on (3, 0, 1, 2).
    tri = np.array([[0, 0, 0], [0.01, 0, 0], [0, 0.01, 0]], np.float32)
    x = np.concatenate([tri, [[0.003, 0.002, h]]]).astype(np.float32)
    m = scenes.contact_mesh(x, np.array([[0, 1, 2]], np.uint32), ghat=ghat)
    K = _spd(rng, 12)
    K[0:3, 3:6] += rng.standard_normal((3, 3)) * 40.0  # a non-symmetric block pair
    K[3:6, 0:3] = K[0:3, 3:6].T
    el = _elastic_upper(K.astype(np.float32).astype(np.float64), 4)
    idx = [3, 0, 1, 2]
    w = np.array([1.0, -0.5, -0.3, -0.2])
    n = np.array([0.0, 0.0, 1.0])
    pairs = np.array([(0, 3, 0, 1)], dtype=oracle.PAIR_DTYPE)
    for mass in (np.full(4, 2e-4), np.array([3e-4, 1e-4, 0.0, 2e-4])):
        grad, *_ = oracle.contact_assemble(m, x.copy(), mass.astype(np.float32), nofric,
                                           pairs, elastic=el)
        mmax = mass[idx].max()
        quad = sum(w[a] * w[b] * (n @ K[3 * idx[a]:3 * idx[a] + 3,
                                       3 * idx[b]:3 * idx[b] + 3] @ n)
                   for a in range(4) for b in range(4))
        inert = sum(w[a] ** 2 * (mass[idx[a]] if mass[idx[a]] > 0 else mmax) / h ** 2
                    for a in range(4))
        k_exp = (quad + inert) / np.sum(w ** 2)
        k_got = _stiffness_from_force(grad, 3, 1.0, n, h, ghat)
        assert abs(k_got - k_exp) <= 2e-4 * abs(k_exp), (mass, k_got, k_exp)
        # the elasticity-free value differs: the elastic term is not negligible
        assert abs(inert / np.sum(w ** 2) - k_exp) > 0.05 * k_exp
    # Point-point along a skew direction, vertices 0 and 3 of two triangles.
    e = np.array([1.0, 2.0, 2.0]) / 3.0
    x2 = np.array([[0, 0, 0], [0.01, 0, 0], [0, 0.01, 0],
                   [0, 0, 0], [0.01, 0.01, 0.02], [0.02, 0, 0.01]], np.float64)
    x2[3] = h * e
    x2[4:] += x2[3]
    x2 = x2.astype(np.float32)
    m2 = scenes.contact_mesh(x2, np.array([[0, 1, 2], [3, 4, 5]], np.uint32), ghat=ghat)
    g = float(np.linalg.norm(x2[3].astype(np.float64) - x2[0]))
    e = (x2[0].astype(np.float64) - x2[3]) / g       # ex = x_0 - x_3 (w = (1, -1))
    K2 = _spd(rng, 18)
    K2[0:3, 9:12] += rng.standard_normal((3, 3)) * 60.0
    K2[9:12, 0:3] = K2[0:3, 9:12].T
    el2 = _elastic_upper(K2.astype(np.float32).astype(np.float64), 6)
    nf2 = {"v": np.zeros(6, np.float32), "e": np.zeros(6, np.float32),
           "f": np.zeros(2, np.float32)}
    pp = np.array([(2, 0, 3, 1)], dtype=oracle.PAIR_DTYPE)
    for mass in (np.full(6, 2e-4), np.array([0.0, 1e-4, 1e-4, 5e-4, 1e-4, 1e-4])):
        grad, *_ = oracle.contact_assemble(m2, x2.copy(), mass.astype(np.float32), nf2, pp,
                                           elastic=el2)
        B = lambda i, j: K2[3 * i:3 * i + 3, 3 * j:3 * j + 3]
        quad = e @ (B(0, 0) + B(3, 3) - B(0, 3) - B(3, 0)) @ e
        mm = [mass[0], mass[3]]
        mmax = max(mm)
        inert = sum((q if q > 0 else mmax) for q in mm) / g ** 2
        k_exp = (quad + inert) / 2.0
        k_got = _stiffness_from_force(grad, 0, 1.0, e, g, ghat)
        assert abs(k_got - k_exp) <= 2e-4 * abs(k_exp), (mass, k_got, k_exp)
        # the force on vertex 3 is the opposite one
        assert np.allclose(grad[9:12], -grad[0:3], rtol=1e-5, atol=0)
