"""Oracle (test infrastructure only) of the learned block preconditioner:
a message-passing network over the 3x3-block Hessian graph that maps each
row's diagonal block, its neighbour blocks and the right-hand side to a
correction T_i of the row's block-Jacobi inverse,

    P_i = C_i (I + T_i)(I + T_i)^T C_i^T,   C_i C_i^T = (A_ii)^-1 (block Jacobi),
    T_i = lower triangle of 0.5 tanh(t_i)  (6 outputs; |T_jj| < 0.5 keeps I + T
          non-singular, so P_i is SPD for every weight value),

trained to minimise the probe loss L = mean_k sum_i |P_i (A z_k)_i - z_k,i|^2
over Rademacher probes z_k (a stochastic estimate of |P A - I|_F^2). With all
weights zero, T = 0 and P is exactly block Jacobi. The reference has no
learned preconditioner; its block-Jacobi preconditioner (solver.cu:357-400)
is the fallback and the baseline.

Everything here is fp64 numpy; the device path computes the same network in
bf16 (MFMA, fp32 accumulate) with fp32 elementwise work, and is compared with
this file to stated tolerances. The gradients are pinned by finite
differences of `loss` (tests/test_oracle_learned.py).

Network (hidden H = 64):
    h0 = relu(x W0 + c0)                                 x: node features (16)
    for l in 1..3:
        m_e = relu([h_src(e), f_e] We_l + ce_l)          f_e: edge features (16)
        M_i = sum over edges e into i of m_e / max(1, deg_i)
        h_l = relu([h_{l-1}, M] Wn_l + cn_l)
    t = h3 Wo + co                                       (6 outputs per row)
"""
import numpy as np

H = 64
FN = 16   # node feature width
FE = 16   # edge feature width
LAYERS = 3
LOWER = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]  # (row, col) of t's 6 entries


def blocks_of(value):
    """Column-major 9-float blocks -> (k, 3, 3) row-major matrices."""
    return np.asarray(value, np.float64).reshape(-1, 3, 3).transpose(0, 2, 1)


def system_graph(sys):
    """Symmetric directed edge list of a stored-upper system: for every
    stored off-diagonal block (i, j), edges j->i (block A_ij) and i->j (block
    A_ji = A_ij^T), grouped by destination row in ascending source order.
    Returns dst, src, the block of each edge (as applied to x_src), the
    diagonal blocks A_ii (stored diagonal + cdiag), and rev (index of the
    reverse edge)."""
    n = int(sys["n"]) if "n" in sys else len(sys["row_offset"]) - 1
    ro = np.asarray(sys["row_offset"], np.int64)
    ci = np.asarray(sys["col_index"], np.int64)
    rows = np.repeat(np.arange(n), np.diff(ro))
    B = blocks_of(sys["value"])
    diag = blocks_of(sys["cdiag"]).copy()
    on = ci == rows
    np.add.at(diag, rows[on], B[on])
    off = ~on
    i, j, b = rows[off], ci[off], B[off]
    dst = np.concatenate([i, j])
    src = np.concatenate([j, i])
    blk = np.concatenate([b, b.transpose(0, 2, 1)])
    order = np.lexsort((src, dst))
    dst, src, blk = dst[order], src[order], blk[order]
    key = dst * n + src
    rkey = src * n + dst
    rev = np.searchsorted(key, rkey)
    return dict(n=n, dst=dst, src=src, blk=blk, diag=diag, rev=rev)


def bj_inverse(diag):
    """Block-Jacobi inverse blocks (exact inverse of the SPD diagonal block)."""
    return np.linalg.inv(diag)


def features(g, b):
    """Node features (n, 16) and edge features (E, 16), scale-free:
    Dhat_i = A_ii / s_i^2 with s_i^2 = tr(A_ii)/3; edge block
    A_ij / (s_i s_j) and its Frobenius norm; per row the sum of those norms,
    the degree / 16 and b_i / s_i normalised by its RMS over rows."""
    n, dst, src, blk, diag = g["n"], g["dst"], g["src"], g["blk"], g["diag"]
    s2 = np.trace(diag, axis1=1, axis2=2) / 3.0
    s = np.sqrt(s2)
    x = np.zeros((n, FN))
    x[:, 0:9] = (diag / s2[:, None, None]).reshape(n, 9)
    eb = blk / (s[dst] * s[src])[:, None, None]
    en = np.sqrt((eb ** 2).sum((1, 2)))
    f = np.zeros((len(dst), FE))
    f[:, 0:9] = eb.reshape(-1, 9)
    f[:, 9] = en
    rowsum = np.zeros(n)
    np.add.at(rowsum, dst, en)
    deg = np.bincount(dst, minlength=n).astype(np.float64)
    x[:, 9] = rowsum
    x[:, 10] = deg / 16.0
    bh = np.asarray(b, np.float64).reshape(n, 3) / s[:, None]
    rms = np.sqrt((bh ** 2).sum() / max(1, n)) or 1.0
    x[:, 11:14] = bh / rms
    return x, f


def init_weights(seed=0, scale=1.0, zero_head=True):
    """He-style initial weights; the output head starts at zero (P = BJ)."""
    rng = np.random.default_rng(seed)
    w = {"W0": rng.normal(0, scale * np.sqrt(2.0 / FN), (FN, H)), "c0": np.zeros(H)}
    for l in range(LAYERS):
        w[f"We{l}"] = rng.normal(0, scale * np.sqrt(2.0 / (H + FE)), (H + FE, H))
        w[f"ce{l}"] = np.zeros(H)
        w[f"Wn{l}"] = rng.normal(0, scale * np.sqrt(2.0 / (2 * H)), (2 * H, H))
        w[f"cn{l}"] = np.zeros(H)
    w["Wo"] = np.zeros((H, 6)) if zero_head else rng.normal(0, 0.05, (H, 6))
    w["co"] = np.zeros(6)
    return w


def relu(a):
    return np.maximum(a, 0.0)


def forward(w, g, x, f):
    """Returns t (n, 6) and the cache the backward pass needs."""
    n, dst, src = g["n"], g["dst"], g["src"]
    deg = np.maximum(np.bincount(dst, minlength=n), 1).astype(np.float64)
    h = relu(x @ w["W0"] + w["c0"])
    cache = {"h": [h], "m": [], "M": []}
    for l in range(LAYERS):
        m = relu(np.concatenate([h[src], f], 1) @ w[f"We{l}"] + w[f"ce{l}"])
        M = np.zeros((n, H))
        np.add.at(M, dst, m)
        M /= deg[:, None]
        h = relu(np.concatenate([h, M], 1) @ w[f"Wn{l}"] + w[f"cn{l}"])
        cache["m"].append(m)
        cache["M"].append(M)
        cache["h"].append(h)
    t = h @ w["Wo"] + w["co"]
    cache["deg"] = deg
    return t, cache


def lower_of(t):
    T = np.zeros((len(t), 3, 3))
    th = 0.5 * np.tanh(t)
    for k, (r, c) in enumerate(LOWER):
        T[:, r, c] = th[:, k]
    return T


def precond_blocks(t, C):
    """P_i = C_i (I + T_i)(I + T_i)^T C_i^T."""
    R = np.eye(3)[None] + lower_of(t)
    CR = C @ R
    return CR @ CR.transpose(0, 2, 1)


def probe_data(g, nprobe=4, seed=0):
    """Rademacher probes z (k, n, 3) and u = A z."""
    rng = np.random.default_rng(seed)
    n = g["n"]
    z = rng.choice([-1.0, 1.0], size=(nprobe, n, 3))
    u = np.einsum("nab,knb->kna", g["diag"], z)
    for k in range(nprobe):
        np.add.at(u[k], g["dst"], np.einsum("eab,eb->ea", g["blk"], z[k][g["src"]]))
    return z, u


def loss_and_dP(P, z, u):
    """L = mean_k sum_i |P_i u_ki - z_ki|^2 / n and dL/dP_i."""
    k, n = z.shape[0], z.shape[1]
    r = np.einsum("nab,knb->kna", P, u) - z
    L = (r ** 2).sum() / (k * n)
    dP = 2.0 * np.einsum("kna,knb->nab", r, u) / (k * n)
    return L, dP


def loss(w, g, x, f, C, z, u):
    t, _ = forward(w, g, x, f)
    return loss_and_dP(precond_blocks(t, C), z, u)[0]


def backward(w, g, x, f, C, z, u):
    """Loss and gradients of every weight."""
    dst, src = g["dst"], g["src"]
    t, cache = forward(w, g, x, f)
    R = np.eye(3)[None] + lower_of(t)
    CR = C @ R
    P = CR @ CR.transpose(0, 2, 1)
    L, dP = loss_and_dP(P, z, u)
    # P = C Q C^T, Q = R R^T: dQ = C^T dP C; dR = (dQ + dQ^T) R
    dQ = C.transpose(0, 2, 1) @ dP @ C
    dR = (dQ + dQ.transpose(0, 2, 1)) @ R
    th = np.tanh(t)
    dt = np.zeros_like(t)
    for k, (r_, c_) in enumerate(LOWER):
        dt[:, k] = dR[:, r_, c_] * 0.5 * (1.0 - th[:, k] ** 2)
    grads = {}
    h3 = cache["h"][LAYERS]
    grads["Wo"] = h3.T @ dt
    grads["co"] = dt.sum(0)
    dh = dt @ w["Wo"].T
    deg = cache["deg"]
    for l in reversed(range(LAYERS)):
        h_prev, M, h = cache["h"][l], cache["M"][l], cache["h"][l + 1]
        dZ = dh * (h > 0)
        X = np.concatenate([h_prev, M], 1)
        grads[f"Wn{l}"] = X.T @ dZ
        grads[f"cn{l}"] = dZ.sum(0)
        dX = dZ @ w[f"Wn{l}"].T
        dh_prev = dX[:, :H].copy()
        dM = dX[:, H:]
        m = cache["m"][l]
        dm = dM[dst] / deg[dst][:, None]
        dY = dm * (m > 0)
        Xe = np.concatenate([h_prev[src], f], 1)
        grads[f"We{l}"] = Xe.T @ dY
        grads[f"ce{l}"] = dY.sum(0)
        ge = (dY @ w[f"We{l}"].T)[:, :H]
        np.add.at(dh_prev, src, ge)   # back to the source rows
        dh = dh_prev
    dZ0 = dh * (cache["h"][0] > 0)
    grads["W0"] = x.T @ dZ0
    grads["c0"] = dZ0.sum(0)
    return L, grads
