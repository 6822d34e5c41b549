"""CPU oracle: 3x3 symmetric eigensolve and block-Jacobi inverse.

Pins: bit-exact against the reference's own eigsolve.hpp compiled host-side
(oracle/_ref, built from the reference tree), fp64 numpy.linalg.eigh, the exact
fp64 inverse, and the eigenvalue floor rule of solver.cu:357-400.
"""
import numpy as np
import pytest

import oracle


def colmajor(m):
    return np.ascontiguousarray(np.asarray(m, np.float32).T.reshape(9))


def from_colmajor(v):
    return np.asarray(v, np.float64).reshape(3, 3).T


def sym_cases(seed=0, count=400):
    rng = np.random.default_rng(seed)
    cases = []
    for _ in range(count):
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        lam = np.exp(rng.uniform(-8, 8, 3)) * rng.choice([1, 1, 1, -1], 3)
        cases.append(q @ np.diag(lam) @ q.T)
    # Degenerate and special shapes named by eigsolve.hpp's comments.
    cases += [np.zeros((3, 3)), np.eye(3), -5 * np.eye(3), np.diag([2.0, 3, 5]),
              np.diag([1.0, 1, 2]), np.diag([1.0, 2, 2]), np.diag([1e20, 1, 1e-20]),
              np.full((3, 3), 1.0), np.diag([1.0, 1.0 + 1e-6, 1.0 - 1e-6]),
              np.array([[4.0, 1e-7, 0], [1e-7, 4.0, 0], [0, 0, 4.0]])]
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    cases.append(q @ np.diag([1.0, 1.0, 1e-3]) @ q.T)  # clustered pair
    cases.append(q @ np.diag([1e9, 1.0, 1e-2]) @ q.T)  # contact-like condition
    cases.append(np.outer([1.0, 2, 3], [1.0, 2, 3]))     # rank one
    return [0.5 * (c + c.T) for c in cases]


def test_symm3x3_bit_exact_vs_reference_build():
    if oracle.load_ref() is None:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    a = np.stack([colmajor(c) for c in sym_cases()])
    v0, q0 = oracle.symm3x3(a)
    v1, q1 = oracle.ref_symm3x3(a)
    assert np.array_equal(v0.view(np.uint32), v1.view(np.uint32))
    assert np.array_equal(q0.view(np.uint32), q1.view(np.uint32))


def test_symm3x3_vs_eigh():
    cases = sym_cases(seed=1)
    a = np.stack([colmajor(c) for c in cases])
    val, vec = oracle.symm3x3(a)
    for k, c in enumerate(cases):
        c32 = from_colmajor(a[k])
        ref = np.linalg.eigvalsh(c32)
        s = max(np.abs(c32).max(), 1e-30)
        # Ascending up to the rounding of the middle value (3q - e1 - e3).
        assert np.all(np.diff(val[k]) >= -1e-5 * s), "ascending order"
        # The trigonometric closed form resolves eigenvalues to about
        # sqrt(eps) of the max-abs scale in fp32 (measured max 3.4e-4 over
        # 5000 random log-uniform spectra); the bit-exact test above pins the
        # reference's own rounding.
        assert np.allclose(val[k], ref, atol=5e-4 * s, rtol=0), (k, val[k], ref)
        Q = from_colmajor(vec[k])
        assert np.allclose(Q.T @ Q, np.eye(3), atol=1e-4), k
        assert np.abs(c32 @ Q - Q * val[k]).max() <= 1e-3 * s, k


def spd_blocks(rng, count, lo, hi):
    out = []
    for _ in range(count):
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        lam = np.exp(rng.uniform(np.log(lo), np.log(hi), 3))
        out.append(q @ np.diag(lam) @ q.T)
    return out


def test_bj_invert_matches_inverse_when_well_conditioned():
    rng = np.random.default_rng(2)
    blocks = spd_blocks(rng, 1000, 1.0, 1e2)
    a = np.stack([colmajor(b) for b in blocks])
    inv, bad = oracle.bj_invert(a)
    assert bad == 0
    errs = []
    for k, b in enumerate(blocks):
        b32 = from_colmajor(a[k])
        ref = np.linalg.inv(b32)
        got = from_colmajor(inv[k])
        errs.append(np.abs(got - ref).max() / np.abs(ref).max())
        assert np.array_equal(got, got.T)
    # The cross-product eigenvectors lose digits on a near-repeated pair, so
    # a rare block is off by up to a few percent (a preconditioner, never the
    # solution); typical blocks are at fp32 level.
    errs = np.array(errs)
    assert np.median(errs) < 1e-5
    assert np.percentile(errs, 99) < 1e-3
    assert errs.max() < 5e-2


def test_bj_invert_floors_unresolvable_eigenvalues():
    rng = np.random.default_rng(3)
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    lam = np.array([1e9, 1e2, 1e-2])  # 1e-2 < 1e-6 * 1e9: floored to 1e3
    b = q @ np.diag(lam) @ q.T
    inv, bad = oracle.bj_invert(colmajor(b)[None])
    assert bad == 0
    got = np.linalg.eigvalsh(from_colmajor(inv[0]))
    expect = np.sort(1.0 / np.maximum(lam, 1e-6 * lam.max()))
    assert np.allclose(got, expect, rtol=1e-3), (got, expect)


def test_bj_invert_symmetrizes_input():
    rng = np.random.default_rng(4)
    b = spd_blocks(rng, 1, 1, 10)[0]
    skew = np.array([[0, 1e-3, 0], [-1e-3, 0, 0], [0, 0, 0]])
    i0, _ = oracle.bj_invert(colmajor(b)[None])
    i1, _ = oracle.bj_invert(colmajor(b + skew)[None])
    assert np.array_equal(i0, i1)


def test_bj_invert_flags_non_spd_blocks():
    a = np.stack([colmajor(-np.eye(3)), colmajor(np.zeros((3, 3))),
                  colmajor(np.diag([np.nan, 1, 1])), colmajor(np.eye(3))])
    _, bad = oracle.bj_invert(a)
    assert bad == 3
