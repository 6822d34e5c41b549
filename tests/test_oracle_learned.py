"""CPU oracle of the learned block preconditioner (oracle/learned.py).

Pinned by what the mathematics fixes:
  * zero output weights give exactly the block-Jacobi inverse;
  * every P_i is symmetric positive definite for arbitrary weights;
  * the loss equals |P A - I|_F^2 / n in expectation: with many probes it
    approaches the exact Frobenius norm computed from the dense matrix;
  * every gradient matches central finite differences of the loss (fp64);
  * the graph is the symmetric closure of the stored upper blocks (it
    reproduces A x exactly) and rev pairs every edge with its reverse.
"""
import numpy as np
import pytest

import learned as L
from ppf_amd import scenes


def small_system(seed=0):
    return scenes.long_row_spd_system(n=60, hubs=2, fan=8, seed=seed)


def test_graph_reproduces_matvec():
    s = small_system(1)
    g = L.system_graph(s)
    M = scenes.sparse_from_upper(s["n"], s["row_offset"], s["col_index"], s["value"],
                                 s["cdiag"]).toarray()
    x = np.random.default_rng(2).normal(size=(s["n"], 3))
    y = np.einsum("nab,nb->na", g["diag"], x)
    np.add.at(y, g["dst"], np.einsum("eab,eb->ea", g["blk"], x[g["src"]]))
    assert np.allclose(y.reshape(-1), M @ x.reshape(-1), rtol=1e-12, atol=1e-9)
    e = np.arange(len(g["dst"]))
    assert np.array_equal(g["rev"][g["rev"]], e)
    assert np.array_equal(g["dst"][g["rev"]], g["src"])


def test_zero_head_is_block_jacobi_and_spd():
    s = small_system(3)
    g = L.system_graph(s)
    x, f = L.features(g, s["b"])
    C = np.linalg.cholesky(L.bj_inverse(g["diag"]))
    w = L.init_weights(0)
    t, _ = L.forward(w, g, x, f)
    assert np.all(t == 0.0)
    P = L.precond_blocks(t, C)
    assert np.allclose(P, np.linalg.inv(g["diag"]), rtol=1e-10, atol=1e-12)
    w = L.init_weights(1, zero_head=False)
    w["Wo"] *= 50.0     # saturate the tanh: T at its bounds
    t, _ = L.forward(w, g, x, f)
    P = L.precond_blocks(t, C)
    assert np.allclose(P, P.transpose(0, 2, 1))
    assert np.linalg.eigvalsh(P).min() > 0.0


def test_probe_loss_estimates_frobenius_norm():
    s = small_system(4)
    g = L.system_graph(s)
    n = s["n"]
    C = np.linalg.cholesky(L.bj_inverse(g["diag"]))
    w = L.init_weights(2, zero_head=False)
    x, f = L.features(g, s["b"])
    t, _ = L.forward(w, g, x, f)
    P = L.precond_blocks(t, C)
    M = scenes.sparse_from_upper(n, s["row_offset"], s["col_index"], s["value"],
                                 s["cdiag"]).toarray()
    Pd = np.zeros((3 * n, 3 * n))
    for i in range(n):
        Pd[3 * i:3 * i + 3, 3 * i:3 * i + 3] = P[i]
    exact = np.linalg.norm(Pd @ M - np.eye(3 * n)) ** 2 / n
    z, u = L.probe_data(g, nprobe=4000, seed=5)
    est, _ = L.loss_and_dP(P, z, u)
    assert abs(est - exact) < 0.03 * exact


@pytest.mark.parametrize("seed", [0, 1])
def test_gradients_match_finite_differences(seed):
    s = small_system(10 + seed)
    g = L.system_graph(s)
    x, f = L.features(g, s["b"])
    C = np.linalg.cholesky(L.bj_inverse(g["diag"]))
    z, u = L.probe_data(g, nprobe=3, seed=seed)
    w = L.init_weights(seed, zero_head=False)
    for k in w:
        w[k] = w[k] + np.random.default_rng(seed + 7).normal(0, 0.05, w[k].shape)
    _, grads = L.backward(w, g, x, f, C, z, u)
    rng = np.random.default_rng(seed + 3)
    eps = 1e-6
    for name in sorted(w):
        flat = w[name].reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            lp = L.loss(w, g, x, f, C, z, u)
            flat[idx] = old - eps
            lm = L.loss(w, g, x, f, C, z, u)
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            an = grads[name].reshape(-1)[idx]
            assert abs(fd - an) <= 1e-6 + 1e-4 * abs(fd), (name, idx, fd, an)
