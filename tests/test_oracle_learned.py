"""CPU oracle of the learned preconditioner (oracle/learned.py: a
message-passing network predicting a block FSAI factor Ghat, M^-1 = G^T G,
G = Ghat L^-1).

Pinned by what the mathematics fixes:
  * the directed edge graph reproduces A x exactly and pairs every edge with
    its reverse; Ahat has unit diagonal blocks by construction;
  * with the output heads at zero, M^-1 is exactly the block-Jacobi inverse;
  * M^-1 is symmetric positive definite for arbitrary weights (saturated);
  * the probe loss converges to the exact log Kaporin number computed from
    the dense matrices;
  * every gradient matches central finite differences of the loss (fp64);
  * the exact FSAI factor of the pattern (closed form: per-row local solves)
    has a lower Kaporin number than block Jacobi and no pattern factor beats it.
"""
import numpy as np
import pytest

import learned as L
from ppf_amd import scenes


def small_system(seed=0, n=60):
    return scenes.long_row_spd_system(n=n, hubs=2, fan=8, seed=seed)


def dense(s):
    return scenes.sparse_from_upper(s["n"], s["row_offset"], s["col_index"], s["value"],
                                    s["cdiag"]).toarray()


def test_graph_reproduces_matvec():
    s = small_system(1)
    g = L.system_graph(s)
    M = dense(s)
    x = np.random.default_rng(2).normal(size=(s["n"], 3))
    y = np.einsum("nab,nb->na", g["diag"], x)
    np.add.at(y, g["dst"], np.einsum("eab,eb->ea", g["blk"], x[g["src"]]))
    assert np.allclose(y.reshape(-1), M @ x.reshape(-1), rtol=1e-12, atol=1e-9)
    e = np.arange(len(g["dst"]))
    assert np.array_equal(g["rev"][g["rev"]], e)
    assert np.array_equal(g["dst"][g["rev"]], g["src"])
    # Ahat = Lbd^-1 A Lbd^-T has identity diagonal blocks
    Lbd = np.zeros_like(M)
    for i in range(s["n"]):
        Lbd[3 * i:3 * i + 3, 3 * i:3 * i + 3] = g["L"][i]
    Ah = np.linalg.solve(Lbd, np.linalg.solve(Lbd, M).T).T
    for i in range(s["n"]):
        assert np.allclose(Ah[3 * i:3 * i + 3, 3 * i:3 * i + 3], np.eye(3), atol=1e-10)


def dense_precond(g, Gd, Ge):
    n = g["n"]
    Gdiag, Glow = L.preconditioner(g, Gd, Ge)
    G = np.zeros((3 * n, 3 * n))
    for j in range(n):
        G[3 * j:3 * j + 3, 3 * j:3 * j + 3] = Gdiag[j]
    low = g["lower"]
    for k, e in enumerate(low):
        j, i = g["dst"][e], g["src"][e]
        G[3 * j:3 * j + 3, 3 * i:3 * i + 3] = Glow[k]
    return G


def test_zero_heads_give_block_jacobi_and_any_weights_spd():
    s = small_system(3)
    g = L.system_graph(s)
    x, f = L.features(g, s["b"])
    w = L.init_weights(0)
    t, u, _ = L.forward(w, g, x, f)
    assert np.all(t == 0.0) and np.all(u == 0.0)
    G = dense_precond(g, *L.factor(t, u))
    Minv = G.T @ G
    n = s["n"]
    for i in range(n):
        assert np.allclose(Minv[3 * i:3 * i + 3, 3 * i:3 * i + 3], np.linalg.inv(g["diag"][i]),
                           rtol=1e-10, atol=1e-12)
    off = Minv.copy()
    for i in range(n):
        off[3 * i:3 * i + 3, 3 * i:3 * i + 3] = 0
    assert np.abs(off).max() == 0.0
    w = L.init_weights(1, zero_head=False)
    w["Wo"] *= 80.0
    w["Wg"] *= 80.0
    t, u, _ = L.forward(w, g, x, f)
    G = dense_precond(g, *L.factor(t, u))
    Minv = G.T @ G
    assert np.allclose(Minv, Minv.T)
    assert np.linalg.eigvalsh(Minv).min() > 0.0
    # the apply routine computes the same G^T G r
    r = np.random.default_rng(4).normal(size=(n, 3))
    z = L.apply_preconditioner(g, *L.preconditioner(g, *L.factor(t, u)), r)
    assert np.allclose(z.reshape(-1), Minv @ r.reshape(-1))


def log_kaporin(g, Gd, Ge, s):
    n = s["n"]
    G = dense_precond(g, Gd, Ge)
    K = G @ dense(s) @ G.T
    sign, logdet = np.linalg.slogdet(K)
    return np.log(np.trace(K) / (3 * n)) - logdet / (3 * n)


def test_probe_loss_is_the_log_kaporin_number():
    s = small_system(4)
    g = L.system_graph(s)
    x, f = L.features(g, s["b"])
    w = L.init_weights(2, zero_head=False)
    t, u, _ = L.forward(w, g, x, f)
    Gd, Ge = L.factor(t, u)
    z = L.probe_vectors(s["n"], 3000, seed=5)
    est = L.loss_parts(g, Gd, Ge, z)[0]
    # the loss drops the constant log det Ahat / 3n
    Lbd_logdet = sum(np.linalg.slogdet(g["diag"][i])[1] for i in range(s["n"]))
    const = (np.linalg.slogdet(dense(s))[1] - Lbd_logdet) / (3 * s["n"])
    exact = log_kaporin(g, Gd, Ge, s) + const
    assert abs(est - exact) < 0.02 * abs(exact) + 2e-3


@pytest.mark.parametrize("seed", [0, 1])
def test_gradients_match_finite_differences(seed):
    s = small_system(10 + seed, n=40)
    g = L.system_graph(s)
    x, f = L.features(g, s["b"])
    z = L.probe_vectors(s["n"], 3, seed=seed)
    w = L.init_weights(seed, zero_head=False)
    for k in w:
        w[k] = w[k] + np.random.default_rng(seed + 7).normal(0, 0.05, w[k].shape)
    _, grads = L.backward(w, g, x, f, z)
    rng = np.random.default_rng(seed + 3)
    eps = 1e-6
    for name in sorted(w):
        flat = w[name].reshape(-1)
        for idx in rng.choice(flat.size, size=min(5, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            lp = L.loss(w, g, x, f, z)
            flat[idx] = old - eps
            lm = L.loss(w, g, x, f, z)
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            an = grads[name].reshape(-1)[idx]
            assert abs(fd - an) <= 1e-7 + 1e-4 * abs(fd), (name, idx, fd, an)


def test_exact_fsai_beats_block_jacobi():
    # classical FSAI on the pattern: Ghat row j solves Ahat[S,S] g = e_j
    # (scaled), minimising the Kaporin number over the pattern
    s = small_system(6, n=40)
    g = L.system_graph(s)
    n = s["n"]
    M = dense(s)
    Lbd = np.zeros_like(M)
    for i in range(n):
        Lbd[3 * i:3 * i + 3, 3 * i:3 * i + 3] = g["L"][i]
    Ah = np.linalg.solve(Lbd, np.linalg.solve(Lbd, M).T).T
    Gh = np.zeros_like(M)
    low = g["lower"]
    for j in range(n):
        cols = [j] + sorted(int(g["src"][e]) for e in low if g["dst"][e] == j)
        idx = np.concatenate([np.arange(3 * c, 3 * c + 3) for c in cols])
        for d in range(3):
            row = 3 * j + d
            S = np.array([k for k in idx if k // 3 < j or k <= row])
            e = (S == row).astype(float)
            sol = np.linalg.solve(Ah[np.ix_(S, S)], e)
            Gh[row, S] = sol / np.sqrt(sol[S == row][0])
    K = Gh @ Ah @ Gh.T
    kap = np.log(np.trace(K) / (3 * n)) - np.linalg.slogdet(K)[1] / (3 * n)
    kap_bj = np.log(np.trace(Ah) / (3 * n)) - np.linalg.slogdet(Ah)[1] / (3 * n)
    assert kap < kap_bj
    # a random pattern factor is never better than the FSAI minimiser
    rng = np.random.default_rng(1)
    for _ in range(3):
        R = Gh + np.where(Gh != 0, rng.normal(0, 0.05, Gh.shape), 0)
        KR = R @ Ah @ R.T
        assert np.log(np.trace(KR) / (3 * n)) - np.linalg.slogdet(KR)[1] / (3 * n) >= kap - 1e-9
