"""The scene data the step reads, against the reference scene builder's
definitions (crates/ppf-cts-solver/src/builder.rs, scene.rs), by closed forms:

- inv_rest2x2 of a shell face (builder.rs:157-228): the rest edges projected
  into the face's own frame (first edge along the first axis, the in-plane
  normal of it second), inverted. A right triangle with legs a, b gives
  diag(1/a, 1/b); an equilateral one of side s gives [[1/s, -1/(s sqrt 3)],
  [0, 2/(s sqrt 3)]]; any rigid motion of a face leaves it unchanged, and the
  projected edges keep the rest edges' Gram matrix.
- lumped masses (scene.rs:1791-1799, builder.rs:571-605): face mass = density
  * area, a vertex gets a third of each incident face's area and mass. On the
  square cloth every triangle has area h^2 / 2, so a vertex's mass is density
  * h^2 / 6 times its incident face count, and the total is density * size^2.
- Lame parameters (builder.rs:1739-1743).
- the substep clamp (builder.rs:268-275): dt = min(dt, 0.9999 / fps) in f32.
"""
import numpy as np

from ppf_amd import scenes


def _inv_rest(tri):
    faces = np.array([[0, 1, 2]], np.uint32)
    return scenes.shell_rest(np.asarray(tri, np.float64), faces)["inv_rest"][0]


def _mat(cm):
    """Column-major 2x2 (as stored) -> matrix."""
    return np.array([[cm[0], cm[2]], [cm[1], cm[3]]], np.float64)


def test_inv_rest_right_and_equilateral_triangles():
    a, b = 0.3, 0.7
    got = _mat(_inv_rest([[0, 0, 0], [a, 0, 0], [0, 0, b]]))
    assert np.allclose(got, np.diag([1 / a, 1 / b]), rtol=1e-6, atol=0)
    s = 0.25
    h = s * np.sqrt(3) / 2
    got = _mat(_inv_rest([[0, 0, 0], [s, 0, 0], [s / 2, h, 0]]))
    want = np.array([[1 / s, -1 / (s * np.sqrt(3))], [0, 2 / (s * np.sqrt(3))]])
    assert np.allclose(got, want, rtol=1e-6, atol=1e-7)


def test_inv_rest_is_invariant_under_rigid_motion_and_keeps_the_metric():
    rng = np.random.default_rng(0)
    for _ in range(50):
        tri = rng.normal(size=(3, 3))
        q, r = np.linalg.qr(rng.normal(size=(3, 3)))
        q = q * np.sign(np.diag(r))
        if np.linalg.det(q) < 0:
            q[:, 0] = -q[:, 0]
        moved = tri @ q.T + rng.normal(size=3)
        A, B = _mat(_inv_rest(tri)), _mat(_inv_rest(moved))
        assert np.allclose(A, B, rtol=1e-4, atol=1e-5 * np.abs(A).max())
        D = np.linalg.inv(A)  # the projected edges
        E = np.stack([tri[1] - tri[0], tri[2] - tri[0]], 1)
        assert np.allclose(D.T @ D, E.T @ E, rtol=1e-5, atol=1e-6 * np.abs(E).max() ** 2)
        # first edge on the first axis, positive orientation
        assert abs(D[1, 0]) <= 1e-6 * np.abs(D).max() and D[0, 0] > 0 and D[1, 1] > 0


def test_lumped_masses_of_the_square_cloth():
    res, size, density = 9, 1.0, 0.2
    sc = scenes.drape_scene(res=res, size=size, density=density)
    sh = sc["shell"]
    h = size / (res - 1)
    faces = sh["face"].astype(np.int64)
    count = np.bincount(faces.reshape(-1), minlength=res * res)
    assert np.allclose(sh["vmass"], density * h * h / 6 * count, rtol=1e-6)
    assert np.allclose(sh["varea"], h * h / 6 * count, rtol=1e-6)
    assert np.allclose(sh["fmass"], density * h * h / 2, rtol=1e-6)
    assert np.isclose(sh["vmass"].astype(np.float64).sum(), density * size * size, rtol=1e-5)
    # interior vertices of the alternating-diagonal triangulation touch 8 and 4
    # faces in a checkerboard (datamodel/mesh.rs:181-211)
    interior = count.reshape(res, res)[1:-1, 1:-1]
    i, j = np.meshgrid(np.arange(1, res - 1), np.arange(1, res - 1), indexing="ij")
    assert np.array_equal(interior, np.where((i + j) % 2 == 0, 8, 4))


def test_lame_parameters():
    mu, lam = scenes.convert_prop(1e4, 0.25)
    assert np.isclose(mu, 1e4 / 2.5, rtol=1e-7) and np.isclose(lam, 1e4 * 0.25 / (1.25 * 0.5),
                                                                 rtol=1e-7)
    sh = scenes.drape_scene(res=4)["shell"]
    assert np.all(sh["f_mu"] == np.float32(mu)) and np.all(sh["f_lambda"] == np.float32(lam))


def test_substep_clamp():
    p = scenes.step_params()
    assert np.float32(p["dt"]) == np.float32(0.9999) / np.float32(60.0)
    assert abs(p["dt"] - 0.016665) < 1e-9
    assert p["prev_dt"] == p["dt"]
    # a dt already under a frame is kept
    assert scenes.step_params(dt=0.01)["dt"] == float(np.float32(0.01))
    assert scenes.clamp_substep_dt(0.5, 30.0) == float(np.float32(0.9999) / np.float32(30.0))
