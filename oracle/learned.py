"""Oracle (test infrastructure only) of the learned preconditioner: a
message-passing network over the 3x3-block Hessian graph that predicts a
sparse approximate inverse factor (FSAI form).

For the system A (stored upper blocks + per-row diagonal), with A_ii = L_i L_i^T
(Cholesky of the diagonal blocks) and the normalised blocks
Ahat_ij = L_i^-1 A_ij L_j^-T, the network outputs a block lower-triangular
Ghat on the pattern of A:

    Ghat_jj = I + lower(0.5 tanh t_j)     (6 outputs per row; its diagonal lies
                                          in (0.5, 1.5), so Ghat is non-singular)
    Ghat_ji = 0.5 tanh(u_e)               (9 outputs per stored block, i < j)

and the preconditioner is M^-1 = G^T G with G = Ghat L^-1, symmetric positive
definite for every weight value. With the output heads at zero, Ghat = I and
M^-1 is exactly block Jacobi (the reference's preconditioner,
solver.cu:357-400, which stays the fallback and the baseline). Training
minimises the log Kaporin number of Ghat Ahat Ghat^T,

    L = log(tr(Ghat Ahat Ghat^T) / 3n) - (2 / 3n) sum_j log det Ghat_jj,

whose minimiser over the pattern is the classical FSAI factor; the trace is
estimated with Rademacher probes z (tr = E[y^T Ahat y], y = Ghat^T z).

Everything here is fp64 numpy; the device computes the same network in bf16
(MFMA, fp32 accumulate) and is compared with this file to stated tolerances.
Gradients are pinned by finite differences of `loss`.

Network (hidden H = 64):
    h0 = relu(x W0 + c0)                              x: node features (16)
    for l in 1..3:
        m_e = relu([h_src(e), f_e] We_l + ce_l)       f_e: edge features (16)
        M_i = sum over edges e into i of m_e / max(1, deg_i)
        h_l = relu([h_{l-1}, M] Wn_l + cn_l)
    t_j = h3_j Wo + co                                (6 per row)
    u_e = [h3_dst, h3_src, f_e] Wg + cg               (9 per lower edge)
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
    """Directed edges of a stored-upper system: for every stored off-diagonal
    block (i, j), edge j->i (block A_ij) and edge i->j (block A_ji = A_ij^T),
    grouped by destination row in ascending source order. Also the diagonal
    blocks A_ii (stored diagonal + cdiag), their Cholesky factors, the
    normalised edge blocks and the reverse of every edge."""
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
    rev = np.searchsorted(key, src * n + dst)
    L = np.linalg.cholesky(diag)
    Li = np.linalg.inv(L)
    ahat = Li[dst] @ blk @ Li[src].transpose(0, 2, 1)
    return dict(n=n, dst=dst, src=src, blk=blk, diag=diag, rev=rev, L=L, Li=Li, ahat=ahat,
                lower=np.nonzero(src < dst)[0])


def features(g, b):
    """Node (n, 16) and edge (E, 16) features of the normalised system:
    edge: Ahat_e (9), |Ahat_e|_F, 1 if the edge is a lower entry; node: the
    sum and max of |Ahat_e|_F over its edges, degree / 16, and L^-1 b
    normalised by its RMS over rows (3)."""
    n, dst = g["n"], g["dst"]
    en = np.sqrt((g["ahat"] ** 2).sum((1, 2)))
    f = np.zeros((len(dst), FE))
    f[:, 0:9] = g["ahat"].reshape(-1, 9)
    f[:, 9] = en
    f[:, 10] = (g["src"] < dst).astype(np.float64)
    x = np.zeros((n, FN))
    np.add.at(x[:, 0], dst, en)
    np.maximum.at(x[:, 1], dst, en)
    x[:, 2] = np.bincount(dst, minlength=n) / 16.0
    bh = np.einsum("nab,nb->na", g["Li"], np.asarray(b, np.float64).reshape(n, 3))
    rms = np.sqrt((bh ** 2).sum() / max(1, n)) or 1.0
    x[:, 3:6] = bh / rms
    return x, f


def init_weights(seed=0, scale=1.0, zero_head=True):
    """He-style initial weights; the output heads start at zero (M^-1 = BJ)."""
    rng = np.random.default_rng(seed)
    w = {"W0": rng.normal(0, scale * np.sqrt(2.0 / FN), (FN, H)), "c0": np.zeros(H)}
    for l in range(LAYERS):
        w[f"We{l}"] = rng.normal(0, scale * np.sqrt(2.0 / (H + FE)), (H + FE, H))
        w[f"ce{l}"] = np.zeros(H)
        w[f"Wn{l}"] = rng.normal(0, scale * np.sqrt(2.0 / (2 * H)), (2 * H, H))
        w[f"cn{l}"] = np.zeros(H)
    if zero_head:
        w["Wo"], w["Wg"] = np.zeros((H, 6)), np.zeros((2 * H + FE, 9))
    else:
        w["Wo"] = rng.normal(0, 0.05, (H, 6))
        w["Wg"] = rng.normal(0, 0.05, (2 * H + FE, 9))
    w["co"], w["cg"] = np.zeros(6), np.zeros(9)
    return w


def relu(a):
    return np.maximum(a, 0.0)


def forward(w, g, x, f):
    """Head outputs t (n, 6), u (lower edges, 9) and the backward cache."""
    n, dst, src, low = g["n"], g["dst"], g["src"], g["lower"]
    deg = np.maximum(np.bincount(dst, minlength=n), 1).astype(np.float64)
    h = relu(x @ w["W0"] + w["c0"])
    cache = {"h": [h], "m": [], "M": [], "deg": deg}
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
    Xg = np.concatenate([h[dst[low]], h[src[low]], f[low]], 1)
    u = Xg @ w["Wg"] + w["cg"]
    cache["Xg"] = Xg
    return t, u, cache


def factor(t, u):
    """Ghat diagonal blocks (n, 3, 3) and lower edge blocks (lower edges, 3, 3);
    edge outputs are read column-major, like every block here."""
    Gd = np.repeat(np.eye(3)[None], len(t), 0)
    th = 0.5 * np.tanh(t)
    for k, (r, c) in enumerate(LOWER):
        Gd[:, r, c] += th[:, k]
    Ge = (0.5 * np.tanh(u)).reshape(-1, 3, 3).transpose(0, 2, 1)
    return Gd, Ge


def probe_vectors(n, nprobe=4, seed=0):
    return np.random.default_rng(seed).choice([-1.0, 1.0], size=(nprobe, n, 3))


def ahat_apply(g, y):
    """Ahat y for y (n, 3) (unit diagonal blocks)."""
    out = y.copy()
    np.add.at(out, g["dst"], np.einsum("eab,eb->ea", g["ahat"], y[g["src"]]))
    return out


def ghat_t_apply(g, Gd, Ge, z):
    """y = Ghat^T z: y_a = sum over row entries (j, a) of Ghat_ja^T z_j."""
    low = g["lower"]
    y = np.einsum("nba,nb->na", Gd, z)
    np.add.at(y, g["src"][low], np.einsum("eba,eb->ea", Ge, z[g["dst"][low]]))
    return y


def loss_parts(g, Gd, Ge, z):
    """Loss and dL/dGd, dL/dGe for probes z (k, n, 3)."""
    n = g["n"]
    low = g["lower"]
    K = len(z)
    q = 0.0
    dGd = np.zeros_like(Gd)
    dGe = np.zeros_like(Ge)
    ws = []
    for k in range(K):
        y = ghat_t_apply(g, Gd, Ge, z[k])
        w_ = 2.0 * ahat_apply(g, y)
        q += (y * ahat_apply(g, y)).sum() / K
        ws.append(w_)
    diag = np.stack([Gd[:, d, d] for d in range(3)], 1)
    L = np.log(q / (3 * n)) - (2.0 / (3 * n)) * np.log(diag).sum()
    for k in range(K):
        dGd += np.einsum("na,nb->nab", z[k], ws[k]) / (K * q)
        dGe += np.einsum("ea,eb->eab", z[k][g["dst"][low]], ws[k][g["src"][low]]) / (K * q)
    for d in range(3):
        dGd[:, d, d] -= (2.0 / (3 * n)) / diag[:, d]
    return L, dGd, dGe


def loss(w, g, x, f, z):
    t, u, _ = forward(w, g, x, f)
    Gd, Ge = factor(t, u)
    return loss_parts(g, Gd, Ge, z)[0]


def backward(w, g, x, f, z):
    """Loss and the gradient of every weight."""
    dst, src, low = g["dst"], g["src"], g["lower"]
    t, u, cache = forward(w, g, x, f)
    Gd, Ge = factor(t, u)
    L, dGd, dGe = loss_parts(g, Gd, Ge, z)
    th = np.tanh(t)
    dt = np.zeros_like(t)
    for k, (r, c) in enumerate(LOWER):
        dt[:, k] = dGd[:, r, c] * 0.5 * (1.0 - th[:, k] ** 2)
    du = dGe.transpose(0, 2, 1).reshape(-1, 9) * 0.5 * (1.0 - np.tanh(u) ** 2)
    grads = {"Wo": cache["h"][LAYERS].T @ dt, "co": dt.sum(0),
             "Wg": cache["Xg"].T @ du, "cg": du.sum(0)}
    dh = dt @ w["Wo"].T
    dXg = du @ w["Wg"].T
    np.add.at(dh, dst[low], dXg[:, :H])
    np.add.at(dh, src[low], dXg[:, H:2 * H])
    deg = cache["deg"]
    for l in reversed(range(LAYERS)):
        h_prev, M, h = cache["h"][l], cache["M"][l], cache["h"][l + 1]
        dZ = dh * (h > 0)
        X = np.concatenate([h_prev, M], 1)
        grads[f"Wn{l}"] = X.T @ dZ
        grads[f"cn{l}"] = dZ.sum(0)
        dX = dZ @ w[f"Wn{l}"].T
        dh_prev = dX[:, :H].copy()
        m = cache["m"][l]
        dY = (dX[:, H:][dst] / deg[dst][:, None]) * (m > 0)
        Xe = np.concatenate([h_prev[src], f], 1)
        grads[f"We{l}"] = Xe.T @ dY
        grads[f"ce{l}"] = dY.sum(0)
        np.add.at(dh_prev, src, (dY @ w[f"We{l}"].T)[:, :H])
        dh = dh_prev
    dZ0 = dh * (cache["h"][0] > 0)
    grads["W0"] = x.T @ dZ0
    grads["c0"] = dZ0.sum(0)
    return L, grads


def preconditioner(g, Gd, Ge):
    """G = Ghat L^-1 as (diagonal blocks (n,3,3), lower edge blocks)."""
    low = g["lower"]
    return Gd @ g["Li"], Ge @ g["Li"][g["src"][low]]


def apply_preconditioner(g, Gdiag, Glow, r):
    """z = G^T G r for r (n, 3)."""
    low = g["lower"]
    y = np.einsum("nab,nb->na", Gdiag, r)
    np.add.at(y, g["dst"][low], np.einsum("eab,eb->ea", Glow, r[g["src"][low]]))
    z = np.einsum("nba,nb->na", Gdiag, y)
    np.add.at(z, g["src"][low], np.einsum("eba,eb->ea", Glow, y[g["dst"][low]]))
    return z
