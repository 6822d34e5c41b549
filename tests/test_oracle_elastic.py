"""CPU oracle: inertia, air drag, Baraff-Witkin membrane and dihedral bending.

Pinned by:
  * the reference's own BaraffWitkin and face_compute_force_hessian, built
    from their headers into oracle/_ref: element gradients and Hessians agree
    bit for bit on random inputs;
  * finite differences of independently written fp64 energies (inertia, air
    drag with the normal held fixed, stretch, shear, bending): the assembled
    force is their gradient;
  * the assembled blocks equal the sum of element Hessians placed by vertex
    index, and the assembled matrix is symmetric positive semi-definite.
"""
import numpy as np
import pytest

import oracle
from ppf_amd import scenes

REF = oracle.load_ref()


@pytest.mark.skipif(REF is None, reason="oracle/_ref not built (needs the reference tree)")
def test_element_terms_match_reference_build():
    rng = np.random.default_rng(0)
    for _ in range(500):
        F = (np.eye(3, 2) + rng.normal(0, 0.3, (3, 2))).astype(np.float32)
        for kind in (0, 1):
            g, h = oracle.elastic_local(kind, F, 1.7)
            rg, rh = oracle.ref_elastic_local(kind, F, 1.7)
            assert np.array_equal(g.view(np.uint32), rg.view(np.uint32))
            assert np.array_equal(h.view(np.uint32), rh.view(np.uint32))
        p = rng.normal(0, 1, (4, 3)).astype(np.float32)
        rest = float(rng.uniform(-1, 1))
        g, h = oracle.elastic_local(2, p[[2, 1, 0, 3]], rest)
        rg, rh = oracle.ref_elastic_local(2, p, rest)
        assert np.array_equal(g.view(np.uint32), rg.view(np.uint32))
        assert np.array_equal(h.view(np.uint32), rh.view(np.uint32))


def dihedral64(p0, p1, p2, p3):
    n1 = np.cross(p1 - p0, p2 - p0)
    n2 = np.cross(p2 - p3, p1 - p3)
    e = p1 - p2
    return np.arctan2(np.cross(n2, n1) @ e / np.linalg.norm(e), n1 @ n2)


def energy64(sh, x, normals, air_density, air_friction):
    """Total fp64 energy of the shell at positions x (the gradient the
    assembly must produce); vertex normals held fixed."""
    dt = sh["dt"]
    X0, T = sh["x0"].astype(np.float64), sh["target"].astype(np.float64)
    E = 0.0
    free = (sh["vflag"] & 1) == 0
    d = x - T
    E += 0.5 * np.sum(sh["vmass"][free] * np.sum(d[free] ** 2, 1)) / dt ** 2
    z = x - X0
    nz = np.any(np.abs(normals) > 1e-5, 1) & free
    nn = np.einsum("ij,ij->i", normals, z)
    tz = z - normals * nn[:, None]
    E += 0.5 * np.sum((sh["varea"] * air_density)[nz] *
                      (nn[nz] ** 2 + air_friction * np.sum(tz[nz] ** 2, 1))) / dt ** 2
    f = sh["face"].astype(np.int64)
    inv = sh["inv_rest"].astype(np.float64).reshape(-1, 2, 2).transpose(0, 2, 1)
    dX = np.stack([x[f[:, 1]] - x[f[:, 0]], x[f[:, 2]] - x[f[:, 0]]], 2)   # (nf,3,2)
    F = dX @ inv
    nu, nv_ = np.linalg.norm(F[:, :, 0], axis=1), np.linalg.norm(F[:, :, 1], axis=1)
    i6 = np.einsum("ij,ij->i", F[:, :, 0], F[:, :, 1])
    E += np.sum(sh["fmass"] * (0.5 * sh["f_mu"] * ((nu - 1) ** 2 + (nv_ - 1) ** 2) +
                               0.5 * sh["f_lambda"] * i6 ** 2))
    H = sh["hinge"].astype(np.int64)
    vm, va = sh["vmass"].astype(np.float64), sh["varea"].astype(np.float64)
    for k, h in enumerate(H):
        dens = [vm[q] / va[q] for q in h if va[q] > 0]
        areal = np.mean(dens) if dens else 0.0
        stiff = 1.28e-5 * sh["h_bend"][k] * sh["h_length"][k] ** 2 / sh["h_area"][k] * areal
        th = dihedral64(x[h[2]], x[h[1]], x[h[0]], x[h[3]])
        E += 0.5 * stiff * (th - sh["h_rest"][k]) ** 2
    return E


def vertex_normals(sh, x):
    n = np.zeros((sh["nv"], 3))
    for i in range(sh["nv"]):
        for k in range(sh["v2f_off"][i], sh["v2f_off"][i + 1]):
            a, b, c = sh["face"][sh["v2f"][k]]
            n[i] += np.cross(x[b] - x[a], x[c] - x[a])
    ln = np.linalg.norm(n, axis=1)
    n[ln > 0] /= ln[ln > 0, None]
    return n


@pytest.mark.parametrize("pins", [0, 5])
def test_force_is_energy_gradient(pins):
    sh = scenes.shell_scene(7, seed=3, pins=pins, stretch=0.05)
    rp, ci = scenes.upper_pattern(sh["nv"], sh["face"], sh["hinge"])
    ad, af = 1e-3, 0.2
    force, cdiag, val, miss = oracle.elastic_assemble(sh, rp, ci, sh["dt"], ad, af)
    assert miss == 0
    x = sh["x"].astype(np.float64)
    nrm = vertex_normals(sh, sh["x"].astype(np.float64))
    g = np.zeros_like(x)
    eps = 1e-6
    for i in range(sh["nv"]):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, d] += eps
            xm[i, d] -= eps
            g[i, d] = (energy64(sh, xp, nrm, ad, af) - energy64(sh, xm, nrm, ad, af)) / (2 * eps)
    f = force.reshape(-1, 3)
    err = np.abs(f - g).max() / np.abs(g).max()
    assert err < 2e-3, err
    if pins:
        # pinned vertices carry no inertia / air term (only element forces)
        assert np.all(cdiag.reshape(-1, 9)[:pins] == 0)


def test_blocks_are_placed_element_hessians():
    sh = scenes.shell_scene(6, seed=5, stretch=0.05)
    rp, ci = scenes.upper_pattern(sh["nv"], sh["face"], sh["hinge"])
    force, cdiag, val, miss = oracle.elastic_assemble(sh, rp, ci, sh["dt"])
    nv = sh["nv"]
    dense = np.zeros((3 * nv, 3 * nv))
    for r in range(nv):
        for k in range(rp[r], rp[r + 1]):
            c = ci[k]
            B = val[k].reshape(3, 3, order="F")
            dense[3 * r:3 * r + 3, 3 * c:3 * c + 3] += B
            if c != r:
                dense[3 * c:3 * c + 3, 3 * r:3 * r + 3] += B.T
    want = np.zeros_like(dense)
    x = sh["x"]
    for k, tri in enumerate(sh["face"].astype(np.int64)):
        inv = sh["inv_rest"][k].reshape(2, 2, order="F").astype(np.float64)
        F = (np.stack([x[tri[1]] - x[tri[0]], x[tri[2]] - x[tri[0]]], 1) @ inv)
        h = (oracle.elastic_local(0, F, sh["f_mu"][k])[1].astype(np.float64) +
             oracle.elastic_local(1, F, sh["f_lambda"][k])[1])
        g = [-inv[0] - inv[1], inv[0], inv[1]]
        for a in range(3):
            for b in range(3):
                blk = sum(g[a][d] * g[b][e] * h[3 * d:3 * d + 3, 3 * e:3 * e + 3]
                          for d in range(2) for e in range(2))
                want[3 * tri[a]:3 * tri[a] + 3, 3 * tri[b]:3 * tri[b] + 3] += sh["fmass"][k] * blk
    vm, va = sh["vmass"], sh["varea"]
    for k, h in enumerate(sh["hinge"].astype(np.int64)):
        dens = [vm[q] / va[q] for q in h if va[q] > 0]
        stiff = 1.28e-5 * sh["h_bend"][k] * sh["h_length"][k] ** 2 / sh["h_area"][k] * np.mean(dens)
        r = [h[2], h[1], h[0], h[3]]
        _, H = oracle.elastic_local(2, x[r], sh["h_rest"][k])
        for a in range(4):
            for b in range(4):
                want[3 * r[a]:3 * r[a] + 3, 3 * r[b]:3 * r[b] + 3] += stiff * H[3 * a:3 * a + 3, 3 * b:3 * b + 3]
    # upper blocks come from the element block with the lower vertex first
    assert np.allclose(dense, want, rtol=1e-5, atol=1e-5 * np.abs(want).max())
    full = dense + np.kron(np.eye(nv), np.ones((3, 3))) * 0
    for i in range(nv):
        full[3 * i:3 * i + 3, 3 * i:3 * i + 3] += cdiag[9 * i:9 * i + 9].reshape(3, 3, order="F")
    assert np.allclose(full, full.T, atol=1e-4 * np.abs(full).max())
    ev = np.linalg.eigvalsh(0.5 * (full + full.T))
    assert ev.min() > -1e-5 * ev.max()
