"""Synthetic inputs of the hot-path workloads (numpy, host side).

Grid triangulation: generate_rect_faces
(crates/ppf-cts-core/src/datamodel/mesh.rs:181-211), alternating diagonals.
"""
import numpy as np


def rect_faces(res_x, res_y):
    """(2*(res_x-1)*(res_y-1), 3) faces of a res_x x res_y vertex grid, vertex
    index i*res_y + j, diagonal alternating with (i % 2 == j % 2)."""
    i, j = np.meshgrid(np.arange(res_x - 1), np.arange(res_y - 1), indexing="xy")
    # mesh.rs iterates j outer, i inner.
    i = i.reshape(-1).astype(np.int64)
    j = j.reshape(-1).astype(np.int64)
    v0 = i * res_y + j
    v1 = v0 + 1
    v2 = v0 + res_y
    v3 = v2 + 1
    same = (i % 2) == (j % 2)
    fa = np.where(same[:, None], np.stack([v0, v1, v3], 1), np.stack([v0, v1, v2], 1))
    fb = np.where(same[:, None], np.stack([v0, v3, v2], 1), np.stack([v1, v3, v2], 1))
    faces = np.empty((2 * len(v0), 3), dtype=np.int64)
    faces[0::2] = fa
    faces[1::2] = fb
    return faces.astype(np.uint32)


def unique_edges(faces):
    """Sorted unique undirected edges (E, 2) with e[0] < e[1]."""
    f = np.asarray(faces).astype(np.int64).reshape(-1, 3)
    if f.size == 0:
        return np.zeros((0, 2), np.uint32)
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    e.sort(axis=1)
    key = np.unique(e[:, 0] * (int(f.max()) + 1) + e[:, 1])
    m = int(f.max()) + 1
    return np.stack([key // m, key % m], 1).astype(np.uint32)


def _colmajor(blocks):
    """(k, 3, 3) row-major numpy blocks -> (k, 9) column-major floats."""
    return np.ascontiguousarray(blocks.transpose(0, 2, 1).reshape(-1, 9))


def upper_bsr(n, pairs, pair_blocks, diag_blocks):
    """Assemble the reference's upper-triangle block CSR from strictly-upper
    (i, j) pairs (i < j, unique) and per-row diagonal blocks, rows sorted by j.
    Returns row_offset (n+1), col_index, value (nnz, 9) column-major."""
    pairs = np.asarray(pairs, dtype=np.int64)
    rows = np.concatenate([np.arange(n, dtype=np.int64), pairs[:, 0]])
    cols = np.concatenate([np.arange(n, dtype=np.int64), pairs[:, 1]])
    blocks = np.concatenate([diag_blocks, pair_blocks]).astype(np.float32)
    order = np.lexsort((cols, rows))
    rows, cols, blocks = rows[order], cols[order], blocks[order]
    row_offset = np.zeros(n + 1, dtype=np.int64)
    np.add.at(row_offset, rows + 1, 1)
    row_offset = np.cumsum(row_offset)
    return (row_offset.astype(np.int32), cols.astype(np.int32),
            _colmajor(blocks).astype(np.float32))


def haar_rotations(rng, k):
    g = rng.standard_normal((k, 3, 3))
    q, r = np.linalg.qr(g)
    d = np.sign(np.diagonal(r, axis1=1, axis2=2))
    d[d == 0] = 1.0
    return q * d[:, None, :]


def grid_spd_system(res, seed_matrix=0, seed_rhs=1, mass=1.0, h=1.0 / 60.0,
                    lam_lo=1e2, lam_hi=1e6):
    """Kernel-isolation PCG system on a res x res triangulated grid.

    A = mass*I + h^2 * sum_edges [[K_e, -K_e], [-K_e, K_e]], K_e = Q^T diag(l) Q,
    l log-uniform in [lam_lo, lam_hi], Q Haar-random (seed_matrix); b ~ N(0,1)
    (seed_rhs); x0 = 0. SPD by construction. Returns a dict of numpy arrays:
    row_offset, col_index, value (upper blocks incl. diagonal stiffness sums),
    cdiag (mass*I per row), b, edges.
    """
    n = res * res
    edges = unique_edges(rect_faces(res, res)).astype(np.int64)
    rng = np.random.default_rng(seed_matrix)
    ne = len(edges)
    lam = np.exp(rng.uniform(np.log(lam_lo), np.log(lam_hi), size=(ne, 3)))
    q = haar_rotations(rng, ne)
    K = np.einsum("eki,ek,ekj->eij", q, lam, q) * (h * h)
    K = 0.5 * (K + K.transpose(0, 2, 1))
    diag = np.zeros((n, 3, 3))
    for c in range(2):
        for a in range(3):
            for b in range(3):
                diag[:, a, b] += np.bincount(edges[:, c], weights=K[:, a, b],
                                             minlength=n)
    ro, ci, val = upper_bsr(n, edges, -K, diag)
    cdiag = _colmajor(np.broadcast_to(mass * np.eye(3), (n, 3, 3))).astype(np.float32)
    b = np.random.default_rng(seed_rhs).standard_normal(3 * n).astype(np.float32)
    return dict(n=n, row_offset=ro, col_index=ci, value=val, cdiag=cdiag, b=b,
                edges=edges.astype(np.uint32))


def element_spd_system(res, seed=0, mass=1.0, scale=10.0):
    """SPD test system whose off-diagonal blocks are NOT symmetric: per face a
    random 9x9 G^T G element matrix (so the (i,j) and (j,i) blocks of every pair
    are transposes, not copies), plus mass*I on the diagonal."""
    n = res * res
    faces = rect_faces(res, res).astype(np.int64)
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((len(faces), 6, 9))
    E = np.einsum("fki,fkj->fij", g, g) * (scale / 9.0)
    blocks = {}
    for f in range(len(faces)):
        for a in range(3):
            for c in range(3):
                i, j = int(faces[f, a]), int(faces[f, c])
                if i <= j:
                    blk = E[f, 3 * a:3 * a + 3, 3 * c:3 * c + 3]
                    blocks[(i, j)] = blocks.get((i, j), 0) + blk
    diag = np.zeros((n, 3, 3))
    pairs, pblk = [], []
    for (i, j), blk in blocks.items():
        if i == j:
            diag[i] = blk
        else:
            pairs.append((i, j))
            pblk.append(blk)
    ro, ci, val = upper_bsr(n, np.array(pairs), np.array(pblk), diag)
    cdiag = _colmajor(np.broadcast_to(mass * np.eye(3), (n, 3, 3))).astype(np.float32)
    b = rng.standard_normal(3 * n).astype(np.float32)
    return dict(n=n, row_offset=ro, col_index=ci, value=val, cdiag=cdiag, b=b)


def dense_from_upper(n, row_offset, col_index, value, cdiag):
    """fp64 dense matrix of the split operator (small n only)."""
    M = np.zeros((3 * n, 3 * n))
    for i in range(n):
        for k in range(row_offset[i], row_offset[i + 1]):
            j = int(col_index[k])
            blk = value[k].reshape(3, 3).T.astype(np.float64)  # column-major
            M[3 * i:3 * i + 3, 3 * j:3 * j + 3] += blk
            if j != i:
                M[3 * j:3 * j + 3, 3 * i:3 * i + 3] += blk.T
        M[3 * i:3 * i + 3, 3 * i:3 * i + 3] += cdiag[i].reshape(3, 3).T
    return M


def sparse_from_upper(n, row_offset, col_index, value, cdiag):
    """fp64 scipy CSR of the split operator."""
    import scipy.sparse as sp
    rows = np.repeat(np.arange(n), np.diff(row_offset))
    cols = col_index.astype(np.int64)
    blk = value.reshape(-1, 3, 3).transpose(0, 2, 1).astype(np.float64)  # [r, c]
    r_list, c_list, v_list = [], [], []
    a, b = np.meshgrid(np.arange(3), np.arange(3), indexing="ij")
    a, b = a.reshape(-1), b.reshape(-1)
    r_list.append((3 * rows[:, None] + a).reshape(-1))
    c_list.append((3 * cols[:, None] + b).reshape(-1))
    v_list.append(blk[:, a, b].reshape(-1))
    off = rows != cols
    r_list.append((3 * cols[off][:, None] + b).reshape(-1))
    c_list.append((3 * rows[off][:, None] + a).reshape(-1))
    v_list.append(blk[off][:, a, b].reshape(-1))
    cd = cdiag.reshape(-1, 3, 3).transpose(0, 2, 1).astype(np.float64)
    ii = np.arange(n)
    r_list.append((3 * ii[:, None] + a).reshape(-1))
    c_list.append((3 * ii[:, None] + b).reshape(-1))
    v_list.append(cd[:, a, b].reshape(-1))
    return sp.csr_matrix((np.concatenate(v_list), (np.concatenate(r_list),
                                                    np.concatenate(c_list))),
                         shape=(3 * n, 3 * n))


def long_row_spd_system(n=400, hubs=4, fan=48, seed=0, mass=1.0):
    """SPD system whose hub rows couple to `fan` other rows (rows far longer
    than a mesh's), built from random rank-3 pair terms [[K, -K], [-K, K]]
    plus a chain, so both the short-row and long-row SpMV forms are used."""
    rng = np.random.default_rng(seed)
    pairs = set((i, i + 1) for i in range(n - 1))
    for h in range(hubs):
        for j in rng.choice(n, size=fan, replace=False):
            if j != h:
                pairs.add((min(h, int(j)), max(h, int(j))))
    pairs = np.array(sorted(pairs), dtype=np.int64)
    q = haar_rotations(rng, len(pairs))
    lam = np.exp(rng.uniform(0, 3, size=(len(pairs), 3)))
    K = np.einsum("eki,ek,ekj->eij", q, lam, q)
    K = 0.5 * (K + K.transpose(0, 2, 1))
    diag = np.zeros((n, 3, 3))
    for c in range(2):
        np.add.at(diag, pairs[:, c], K)
    ro, ci, val = upper_bsr(n, pairs, -K, diag)
    cdiag = _colmajor(np.broadcast_to(mass * np.eye(3), (n, 3, 3))).astype(np.float32)
    b = rng.standard_normal(3 * n).astype(np.float32)
    return dict(n=n, row_offset=ro, col_index=ci, value=val, cdiag=cdiag, b=b)


# ---- contact meshes --------------------------------------------------------

def mesh_topology(nv, faces):
    """Edges (sorted unique, e[0] < e[1]) and the CSR neighbor tables the
    narrow phase reads: vertex->edges, vertex->faces, edge->faces (the
    reference's data.mesh.neighbor.vertex.edge / .face and .edge.face)."""
    faces = np.asarray(faces, dtype=np.int64)
    e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
    e.sort(axis=1)
    key = e[:, 0] * nv + e[:, 1]
    ukey, inv = np.unique(key, return_inverse=True)
    edges = np.stack([ukey // nv, ukey % nv], 1)
    ne, nf = len(edges), len(faces)

    def csr(rows, cols, nrow):
        order = np.lexsort((cols, rows))
        off = np.zeros(nrow + 1, np.int64)
        np.add.at(off, rows + 1, 1)
        return np.cumsum(off).astype(np.uint32), cols[order].astype(np.uint32)

    v2e_off, v2e = csr(np.concatenate([edges[:, 0], edges[:, 1]]),
                       np.concatenate([np.arange(ne), np.arange(ne)]), nv)
    v2f_off, v2f = csr(faces.reshape(-1), np.repeat(np.arange(nf), 3), nv)
    fe = inv.reshape(3, nf).T  # the three edges of every face
    e2f_off, e2f = csr(fe.reshape(-1), np.repeat(np.arange(nf), 3), ne)
    return dict(edge=edges.astype(np.uint32), v2e_off=v2e_off, v2e=v2e,
                v2f_off=v2f_off, v2f=v2f, e2f_off=e2f_off, e2f=e2f)


def contact_mesh(x, faces, ghat=1e-3, offset=0.0):
    """Contact mesh dict for the narrow phase (uniform contact parameters)."""
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
    faces = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
    nv = len(x)
    topo = mesh_topology(nv, faces)
    ne, nf = len(topo["edge"]), len(faces)
    m = dict(nv=nv, ne=ne, nf=nf, x=x, face=faces, **topo)
    for k, n in (("v", nv), ("e", ne), ("f", nf)):
        m[k + "_ghat"] = np.full(n, ghat, np.float32)
        m[k + "_off"] = np.full(n, offset, np.float32)
    m["vflag"] = np.zeros(nv, np.uint32)
    m["vbody"] = np.zeros(nv, np.uint32)
    m["eflag"] = np.zeros(ne, np.uint32)
    m["fflag"] = np.zeros(nf, np.uint32)
    return m


def random_triangles(nf, seed=2, jitter=5e-3):
    """Independent triangles: centroid ~ U[0,1]^3, every vertex = centroid +
    U[-jitter, jitter]^3 (the broad/narrow-phase kernel-isolation shape)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 1.0, size=(nf, 1, 3))
    x = (c + rng.uniform(-jitter, jitter, size=(nf, 3, 3))).reshape(-1, 3)
    faces = np.arange(3 * nf, dtype=np.uint32).reshape(nf, 3)
    return x.astype(np.float32), faces


def stacked_sheets(res=32, layers=2, gap=6e-4, noise=2e-4, shift=(0.31, 0.17),
                   seed=3, size=0.1, shifts=None):
    """`layers` square sheets (res x res vertices, side `size`) stacked `gap`
    apart in z with N(0, noise) vertex noise and odd layers shifted by `shift`
    cells (or layer k by shifts[k]): dense self-proximity with shared mesh
    topology, so the multiplicity rules of the point-edge / point-point kinds
    are exercised (nearly aligned vertices give high point-point counts)."""
    rng = np.random.default_rng(seed)
    f0 = rect_faces(res, res).astype(np.int64)
    h = size / (res - 1)
    i, j = np.meshgrid(np.arange(res), np.arange(res), indexing="ij")
    xs, faces = [], []
    for layer in range(layers):
        sx, sy = (shifts[layer] if shifts is not None
                  else (shift if layer % 2 else (0.0, 0.0)))
        p = np.stack([(i.reshape(-1) + sx) * h, (j.reshape(-1) + sy) * h,
                      np.full(res * res, layer * gap)], 1)
        p[:, 2] += rng.normal(0.0, noise, res * res)
        xs.append(p)
        faces.append(f0 + layer * res * res)
    return (np.concatenate(xs).astype(np.float32),
            np.concatenate(faces).astype(np.uint32))


def face_element_bsr(nv, faces, seed=0, scale=1.0):
    """Random PSD per-face 9x9 element matrices G^T G assembled into the
    reference's upper block CSR (stand-in for an elastic Hessian: its
    off-diagonal blocks are unsymmetric, so a lookup that forgets to transpose
    (i > j) is caught)."""
    faces = np.asarray(faces, dtype=np.int64)
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((len(faces), 5, 9))
    E = np.einsum("fki,fkj->fij", g, g) * (scale / 9.0)
    keys, blks = [], []
    for a in range(3):
        for c in range(3):
            i, j = faces[:, a], faces[:, c]
            sel = i <= j
            keys.append(np.stack([i[sel], j[sel]], 1))
            blks.append(E[sel, 3 * a:3 * a + 3, 3 * c:3 * c + 3])
    keys = np.concatenate(keys)
    blks = np.concatenate(blks)
    key = keys[:, 0] * nv + keys[:, 1]
    uk, inv = np.unique(key, return_inverse=True)
    acc = np.zeros((len(uk), 3, 3))
    np.add.at(acc, inv, blks)
    rows, cols = uk // nv, uk % nv
    diag = np.zeros((nv, 3, 3))
    on = rows == cols
    diag[rows[on]] = acc[on]
    ro, ci, val = upper_bsr(nv, np.stack([rows[~on], cols[~on]], 1), acc[~on], diag)
    return ro, ci, val


def contact_scenario(res=24, seed=0, slip=2e-5):
    """Stacked-sheet contact mesh plus everything the contact assembly reads:
    start positions x0 = x - slip noise (both friction branches occur),
    masses (a few zero, as pinned vertices carry), friction per element, and a
    random elastic upper block CSR on the face pattern."""
    x, f = stacked_sheets(res, layers=3, gap=4e-4, noise=5e-5,
                          shifts=[(0, 0), (0.31, 0.17), (0.02, 0.03)], seed=seed)
    m = contact_mesh(x, f, ghat=1e-3)
    rng = np.random.default_rng(seed + 100)
    x0 = (x + rng.normal(0.0, slip, x.shape)).astype(np.float32)
    mass = rng.uniform(1e-5, 1e-4, m["nv"]).astype(np.float32)
    mass[rng.random(m["nv"]) < 0.05] = 0.0
    fric = {"v": rng.uniform(0.0, 0.5, m["nv"]).astype(np.float32),
            "e": rng.uniform(0.0, 0.5, m["ne"]).astype(np.float32),
            "f": rng.uniform(0.0, 0.5, m["nf"]).astype(np.float32)}
    elastic = face_element_bsr(m["nv"], f, seed=seed + 7, scale=1.0)
    return m, x0, mass, fric, elastic


def static_mesh(x, faces, ghat=1e-3, offset=0.0):
    """Static collision mesh dict (data.constraint.mesh layout): positions,
    unique edges, faces and uniform per-element contact gap / offset."""
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
    faces = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
    edge = unique_edges(faces).astype(np.uint32)
    nv, ne, nf = len(x), len(edge), len(faces)
    sm = dict(nv=nv, ne=ne, nf=nf, x=x, edge=edge, face=faces)
    for k, n in (("v", nv), ("e", ne), ("f", nf)):
        sm[k + "_ghat"] = np.full(n, ghat, np.float32)
        sm[k + "_off"] = np.full(n, offset, np.float32)
    return sm


def falling_scene(res=24, layers=2, seed=0, drop=3e-3, ground_res=None,
                  flags=True, offset=5e-5, ground_gap=1e-3):
    """Line-search scene: `layers` noisy sheets stacked 1.5e-3 apart whose
    target positions x1 fall by `drop` (upper layers faster, so sheets pass
    through each other) with per-vertex jitter, above a static ground grid
    placed `ground_gap` below the lowest sheet. With `flags`, a few vertices are
    pinned, a few marked collider and one sheet is a rigid body, so every
    pair filter is exercised. Returns (mesh, x1, vmass, emass, fmass,
    static mesh)."""
    rng = np.random.default_rng(seed)
    x, f = stacked_sheets(res, layers=layers, gap=1.5e-3, noise=1e-4,
                          shifts=[(0.13 * k, 0.29 * k) for k in range(layers)],
                          seed=seed + 1)
    m = contact_mesh(x, f, ghat=1e-3, offset=offset)
    nv, ne, nf = m["nv"], m["ne"], m["nf"]
    per = res * res
    x1 = x.copy()
    for k in range(layers):
        x1[k * per:(k + 1) * per, 2] -= drop * (1.0 + 0.6 * k)
    x1 += rng.normal(0.0, 2e-4, x1.shape).astype(np.float32)
    if flags:
        pin = rng.random(nv) < 0.03
        m["vflag"][pin] |= 1
        m["vflag"][rng.random(nv) < 0.03] |= 2
        if layers > 2:
            m["vbody"][2 * per:3 * per] = 1
        x1[pin] = x[pin]
        e = m["edge"]
        m["eflag"] = ((m["vflag"][e[:, 0]] & 1) & (m["vflag"][e[:, 1]] & 1)).astype(np.uint32)
        fc = m["face"]
        m["fflag"] = ((m["vflag"][fc[:, 0]] & 1) & (m["vflag"][fc[:, 1]] & 1)
                      & (m["vflag"][fc[:, 2]] & 1)).astype(np.uint32)
    vmass = rng.uniform(1e-5, 1e-4, nv).astype(np.float32)
    vmass[rng.random(nv) < 0.02] = 0.0
    emass = rng.uniform(1e-5, 1e-4, ne).astype(np.float32)
    fmass = rng.uniform(1e-5, 1e-4, nf).astype(np.float32)
    gr = ground_res or max(4, res // 2)
    gx, gf = stacked_sheets(gr, layers=1, gap=0.0, noise=0.0, seed=seed + 2,
                            size=0.12)
    gx = gx + np.float32([-0.01, -0.01, float(x[:, 2].min()) - ground_gap])
    sm = static_mesh(gx, gf, ghat=1e-3, offset=offset)
    return m, x1.astype(np.float32), vmass, emass, fmass, sm


def ccd_pairs(kind, n, seed=0, scale=0.01):
    """n random primitive pairs of one CCD kind (0 point-triangle, 1
    point-edge, 2 point-point, 3 edge-edge) whose first primitive is driven
    toward (and often through) the second, so many sweeps end in contact.
    Returns x0, x1 of shape (n, columns, 3) float32."""
    cols = {0: 4, 1: 3, 2: 2, 3: 4}[kind]
    split = 2 if kind == 3 else 1
    rng = np.random.default_rng(seed)
    x0 = rng.uniform(-1, 1, (n, cols, 3)) * scale
    v = rng.uniform(-1, 1, (n, cols, 3)) * (scale * (2e-3 if kind in (1, 2) else 0.4))
    pull = x0[:, split:].mean(1) - x0[:, :split].mean(1)
    v[:, :split] += rng.uniform(0.5, 2.0, (n, 1, 1)) * pull[:, None, :]
    return x0.astype(np.float32), (x0 + v).astype(np.float32)


def static_feature_scene(feature, res=8, offset=5e-5):
    """A flat sheet (rotated 30 degrees about z, so no edge is axis aligned)
    falling 5e-3 at uniform speed onto one static feature that decides the
    step through one collision-mesh pair kind:
      'spike' -- a tetrahedron tip 1e-3 below the sheet inside a face
                 (static vertex vs dynamic face);
      'ridge' -- a steep roof whose ridge edge runs 1e-3 below the sheet
                 across its edges (dynamic edge vs static edge);
      'floor' -- one large triangle 1e-3 below (dynamic vertex vs static
                 face).
    Returns (mesh, x1, vmass, emass, fmass, static mesh)."""
    x, f = stacked_sheets(res, layers=1, gap=0.0, noise=0.0, seed=0, size=0.1)
    c, s_ = np.cos(np.pi / 6), np.sin(np.pi / 6)
    rot = np.array([[c, -s_, 0], [s_, c, 0], [0, 0, 1]])
    x = (x.astype(np.float64) - [0.05, 0.05, 0]) @ rot.T
    x = x.astype(np.float32)
    m = contact_mesh(x, f, ghat=1e-3, offset=offset)
    x1 = (x - np.float32([0, 0, 5e-3])).astype(np.float32)
    if feature == "spike":
        # tip under the centroid of a face near the middle
        fc = x[f].mean(1)
        k = int(np.argmin(np.linalg.norm(fc[:, :2] - [0.003, 0.002], axis=1)))
        tip = [fc[k, 0], fc[k, 1], -1e-3]
        base = [[tip[0] - 0.02, tip[1] - 0.02, -3e-2], [tip[0] + 0.02, tip[1] - 0.02, -3e-2],
                [tip[0], tip[1] + 0.02, -3e-2]]
        sx = np.array([tip] + base, np.float32)
        sf = np.array([[0, 1, 2], [0, 2, 3], [0, 3, 1]], np.uint32)
    elif feature == "ridge":
        sx = np.array([[-0.08, 0.0013, -1e-3], [0.08, 0.0013, -1e-3],
                       [-0.08, -0.0037, -2e-2], [0.08, -0.0037, -2e-2],
                       [-0.08, 0.0063, -2e-2], [0.08, 0.0063, -2e-2]], np.float32)
        sf = np.array([[0, 1, 2], [1, 3, 2], [0, 4, 1], [1, 4, 5]], np.uint32)
    elif feature == "floor":
        # one large triangle: its edges and vertices are far from the sheet
        sx = np.array([[-1.0, -1.0, -1e-3], [1.0, -1.0, -1e-3], [0.0, 1.0, -1e-3]],
                      np.float32)
        sf = np.array([[0, 1, 2]], np.uint32)
    else:
        raise ValueError(feature)
    sm = static_mesh(sx, sf, ghat=1e-3, offset=offset)
    ones = lambda n: np.ones(n, np.float32)
    return m, x1, ones(m["nv"]), ones(m["ne"]), ones(m["nf"]), sm


def shell_hinges(faces, with_faces=False):
    """Interior edges shared by two faces as hinges (e0, e1, flap0, flap1): the
    stored layout whose remap (h2, h1, h0, h3) puts the shared edge in the
    middle. Sorted by edge; flap0 belongs to the lower-index face. With
    `with_faces`, also the two face indices of every hinge."""
    f = np.asarray(faces, np.int64).reshape(-1, 3)
    nf = len(f)
    if nf == 0:
        h = np.zeros((0, 4), np.uint32)
        return (h, np.zeros((0, 2), np.int64)) if with_faces else h
    a = f.reshape(-1)
    b = f[:, [1, 2, 0]].reshape(-1)
    c = f[:, [2, 0, 1]].reshape(-1)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    n = int(f.max()) + 1
    key = lo * n + hi
    order = np.argsort(key, kind="stable")
    ks = key[order]
    dup = np.nonzero(ks[1:] == ks[:-1])[0]
    i0, i1 = order[dup], order[dup + 1]
    h = np.stack([lo[i0], hi[i0], c[i0], c[i1]], 1).astype(np.uint32)
    if with_faces:
        return h, np.stack([i0 // 3, i1 // 3], 1)
    return h


def vertex_face_csr(nv, faces):
    f = np.asarray(faces, np.int64)
    cnt = np.bincount(f.reshape(-1), minlength=nv)
    off = np.concatenate([[0], np.cumsum(cnt)]).astype(np.uint32)
    order = np.argsort(f.reshape(-1), kind="stable")
    v2f = (order // 3).astype(np.uint32)
    return off, v2f


def _dihedral64(p0, p1, p2, p3):
    """Signed dihedral angle of the remapped hinge (X0..X3) in fp64 (the
    atan2 form of dihedral_angle.hpp:86-108)."""
    n1 = np.cross(p1 - p0, p2 - p0)
    n2 = np.cross(p2 - p3, p1 - p3)
    e = p1 - p2
    s = np.einsum("...i,...i->...", np.cross(n2, n1), e) / np.linalg.norm(e, axis=-1)
    c = np.einsum("...i,...i->...", n1, n2)
    return np.arctan2(s, c)


def convert_prop(young, poisson):
    """Lame parameters of a Young's modulus and Poisson ratio, in f32
    (builder.rs:1739-1743)."""
    e, nu = np.float32(young), np.float32(poisson)
    one, two = np.float32(1.0), np.float32(2.0)
    return float(e / (two * (one + nu))), float(e * nu / ((one + nu) * (one - two * nu)))


def shell_rest(rest, faces, density=0.2, young=1e4, poisson=0.25, bend=10.0):
    """Rest-state shell data the elastic assembly reads (builder.rs:157-272,
    339-644, mesh.rs hinges): lumped vertex mass / area, face mass,
    inv_rest2x2 (column-major) in each face's own 2D frame, hinges with rest
    dihedral angle / edge length / combined area, Lame parameters from young /
    poisson (builder.rs:1739-1743)."""
    rest = np.asarray(rest, np.float64).reshape(-1, 3)
    faces = np.asarray(faces, np.uint32).reshape(-1, 3)
    nv, nf = len(rest), len(faces)
    P = rest[faces.astype(np.int64)]
    e1, e2 = P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]
    cr = np.cross(e1, e2)
    area = 0.5 * np.linalg.norm(cr, axis=1)
    l1 = np.linalg.norm(e1, axis=1)
    u = e1 / l1[:, None]
    w = np.cross(cr, e1)
    w /= np.linalg.norm(w, axis=1)[:, None]
    D = np.stack([np.stack([l1, np.einsum("ij,ij->i", e2, u)], 1),
                  np.stack([np.zeros(nf), np.einsum("ij,ij->i", e2, w)], 1)], 1)
    inv = np.linalg.inv(D)                              # (nf, 2, 2) row-major
    inv_rest = inv.transpose(0, 2, 1).reshape(nf, 4)    # column-major
    # FaceProp area and mass = density * area (scene.rs:1791-1799); a vertex
    # gets a third of each incident face's area and mass (builder.rs:571-605)
    area = area.astype(np.float32)
    fmass = np.float32(density) * area
    fi = faces.reshape(-1).astype(np.int64)
    varea = np.zeros(nv, np.float32)
    vmass = np.zeros(nv, np.float32)
    np.add.at(varea, fi, np.repeat(area / np.float32(3.0), 3))
    np.add.at(vmass, fi, np.repeat(fmass / np.float32(3.0), 3))
    hinge, hf = shell_hinges(faces, with_faces=True)
    H = hinge.astype(np.int64)
    hl = np.linalg.norm(rest[H[:, 1]] - rest[H[:, 0]], axis=1)
    harea = area[hf[:, 0]] + area[hf[:, 1]] if len(H) else np.zeros(0)
    mu, lam = convert_prop(young, poisson)
    hrest = (_dihedral64(rest[H[:, 2]], rest[H[:, 1]], rest[H[:, 0]], rest[H[:, 3]])
             if len(H) else np.zeros(0))
    v2f_off, v2f = vertex_face_csr(nv, faces)
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    return dict(nv=nv, nf=nf, nh=len(hinge), face=faces, hinge=hinge, v2f_off=v2f_off,
                v2f=v2f, vmass=f32(vmass), varea=f32(varea), vflag=np.zeros(nv, np.uint32),
                fmass=f32(fmass), inv_rest=f32(inv_rest), f_mu=f32(np.full(nf, mu)),
                f_lambda=f32(np.full(nf, lam)), fflag=np.zeros(nf, np.uint32),
                h_rest=f32(hrest), h_length=f32(hl), h_area=f32(harea),
                h_bend=f32(np.full(len(hinge), bend)), hflag=np.zeros(len(hinge), np.uint32))


def shell_scene(res=16, size=1.0, density=0.2, young=1e4, poisson=0.25, bend=10.0,
                dt=1.0 / 60.0, seed=0, bump=0.05, stretch=0.02, pins=0, gravity=-9.8):
    """A square cloth (res x res vertices, side `size`, the alternating
    triangulation of rect_faces) with everything the elastic assembly reads
    (shell_rest), the current state x0 = rest + a smooth bump, a perturbed
    Newton iterate x, and the inertial target x0 + dt^2 g. `pins` vertices
    (the first ones) are pinned."""
    rng = np.random.default_rng(seed)
    faces = rect_faces(res, res).astype(np.uint32)
    i, j = np.meshgrid(np.arange(res), np.arange(res), indexing="ij")
    h = size / (res - 1)
    rest = np.stack([i.reshape(-1) * h, j.reshape(-1) * h, np.zeros(res * res)], 1)
    sh = shell_rest(rest, faces, density, young, poisson, bend)
    x0 = rest.copy()
    x0[:, 2] = bump * np.sin(np.pi * rest[:, 0] / size) * np.sin(np.pi * rest[:, 1] / size)
    x = x0 + rng.normal(0.0, stretch * h, x0.shape)
    target = x0 + np.array([0.0, 0.0, gravity]) * dt * dt
    sh["vflag"][:pins] = 1
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    sh.update(x=f32(x), x0=f32(x0), target=f32(target), dt=dt)
    return sh


def upper_pattern(nv, faces, hinges):
    """Upper-triangle block pattern (row offsets, columns) of every face and
    hinge vertex pair plus every diagonal (numpy, for checking)."""
    pairs = [np.stack([np.arange(nv), np.arange(nv)], 1)]
    f = np.asarray(faces, np.int64)
    for a in range(3):
        for b in range(3):
            pairs.append(np.stack([f[:, a], f[:, b]], 1))
    hh = np.asarray(hinges, np.int64).reshape(-1, 4)
    for a in range(4):
        for b in range(4):
            pairs.append(np.stack([hh[:, a], hh[:, b]], 1))
    p = np.concatenate(pairs)
    p = p[p[:, 0] <= p[:, 1]]
    key = np.unique(p[:, 0] * nv + p[:, 1])
    rows, cols = key // nv, key % nv
    ro = np.searchsorted(rows, np.arange(nv + 1)).astype(np.uint32)
    return ro, cols.astype(np.uint32)


def icosphere(radius=0.3, subdiv=3):
    """Icosphere: the icosahedron with every face split in four `subdiv`
    times and the new vertices projected to the sphere (subdiv 3: 642
    vertices, 1,280 faces)."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
         (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = list(f)
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]
        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return (np.array(verts) * radius).astype(np.float32), np.array(faces, np.uint32)


def drape_scene(res=64, size=1.0, height=0.05, radius=0.3, subdiv=3, ghat=1e-3, offset=0.0,
                friction=0.5, density=0.2, young=1e4, poisson=0.25, bend=10.0, pins=None,
                strain_limit=0.0, plasticity=(0.0, 0.0), bend_plasticity=(0.0, 0.0)):
    """Square cloth (res x res vertices, side `size`, in the xz plane, y up)
    centred `height` above a static icosphere of `radius` at the origin: the
    plain-method drape configuration (64 x 64, edge 1/63 m, 1,280-triangle
    sphere, gap 1e-3, friction 0.5, density 0.2, Young 1e4, Poisson 0.25, bend
    10). `pins`: indices of pinned vertices (held in place). `strain_limit`
    > 0 limits every face's stretch (FaceParam.strainlimit; the frontend's
    strain-limit, examples/headless.py uses 0.05). `plasticity` and
    `bend_plasticity`: (rate, threshold) of every face's and hinge's rest-shape
    creep (FaceParam.plasticity / HingeParam.plasticity; rate 0 = off).
    Returns a dict of numpy arrays: mesh (contact mesh, x = current
    positions), shell, static, prev, emass, fric, fix_pos."""
    faces = rect_faces(res, res).astype(np.uint32)
    i, j = np.meshgrid(np.arange(res), np.arange(res), indexing="ij")
    h = size / (res - 1)
    rest = np.stack([i.reshape(-1) * h - size / 2, np.zeros(res * res),
                     j.reshape(-1) * h - size / 2], 1)
    x = rest + np.array([0.0, radius + height, 0.0])
    shell = shell_rest(rest, faces, density, young, poisson, bend)
    mesh = contact_mesh(x.astype(np.float32), faces, ghat=ghat, offset=offset)
    sx, sf = icosphere(radius, subdiv)
    static = static_mesh(sx, sf, ghat=ghat, offset=offset)
    for k, n in (("v", static["nv"]), ("e", static["ne"]), ("f", static["nf"])):
        static[k + "_fric"] = np.full(n, friction, np.float32)
    e = mesh["edge"].astype(np.int64)
    emass = (0.5 * (shell["vmass"][e[:, 0]] + shell["vmass"][e[:, 1]])).astype(np.float32)
    fric = {"v": np.full(mesh["nv"], friction, np.float32),
            "e": np.full(mesh["ne"], friction, np.float32),
            "f": np.full(mesh["nf"], friction, np.float32)}
    fix_pos = None
    if pins is not None and len(pins):
        mesh["vflag"][np.asarray(pins)] |= 1
        shell["vflag"] = mesh["vflag"].copy()
        fix_pos = mesh["x"].copy()
    if strain_limit > 0.0:
        shell["f_strainlimit"] = np.full(len(faces), strain_limit, np.float32)
    if plasticity[0] > 0.0:
        shell["f_plasticity"] = np.full(len(faces), plasticity[0], np.float32)
        shell["f_plastic_threshold"] = np.full(len(faces), plasticity[1], np.float32)
    if bend_plasticity[0] > 0.0:
        nh = len(shell["hinge"])
        shell["h_plasticity"] = np.full(nh, bend_plasticity[0], np.float32)
        shell["h_plastic_threshold"] = np.full(nh, bend_plasticity[1], np.float32)
    return dict(mesh=mesh, shell=shell, static=static, prev=mesh["x"].copy(), emass=emass,
                fric=fric, fix_pos=fix_pos)


MAX_SUBSTEP_FRAME_FRACTION = 0.9999  # builder.rs:268


def clamp_substep_dt(dt, fps):
    """A requested substep dt kept strictly under one frame, in f32
    (builder.rs:270-275): min(dt, 0.9999 / fps)."""
    return float(min(np.float32(dt), np.float32(MAX_SUBSTEP_FRAME_FRACTION) / np.float32(fps)))


def step_params(dt=1.0 / 60.0, fps=60.0, **kw):
    """ParamSet defaults the step reads (crates/ppf-cts-core/src/datamodel/
    params.rs:215-280), with gravity (0, -9.8, 0). dt is clamped under one
    frame of `fps` as the scene builder does (builder.rs:348): the configs'
    dt = 1/60 at 60 fps steps 0.016665 s."""
    dt = clamp_substep_dt(dt, fps)
    p = dict(dt=dt, prev_dt=dt, gravity=(0.0, -9.8, 0.0), cg_tol=1e-3, cg_max_iter=10000,
             target_toi=0.25, line_search_max_t=1.25, ccd_eps=1e-7, max_dx=1.0,
             friction_eps=1e-5, friction_mode=0, min_newton_steps=0, max_newton_steps=2048,
             air_density=1e-3, air_friction=0.2, wind=(0.0, 0.0, 0.0), disable_contact=0,
             check_intersection=1, constraint_tol=0.01, eiganalysis_eps=1e-2)
    p.update(kw)
    return p


def cylinder(radius=0.2, length=1.2, n_around=64, n_along=24):
    """Closed-side cylinder surface along z (no caps), centred at the origin:
    n_along rings of n_around vertices, alternating diagonals."""
    th = 2 * np.pi * np.arange(n_around) / n_around
    z = np.linspace(-length / 2, length / 2, n_along)
    x = np.stack([np.tile(radius * np.cos(th), n_along), np.tile(radius * np.sin(th), n_along),
                  np.repeat(z, n_around)], 1)
    faces = []
    for j in range(n_along - 1):
        for i in range(n_around):
            a = j * n_around + i
            b = j * n_around + (i + 1) % n_around
            c = a + n_around
            d = b + n_around
            faces += [(a, b, d), (a, d, c)] if (i + j) % 2 else [(a, b, c), (b, d, c)]
    return x.astype(np.float32), np.array(faces, np.uint32)


def stacked_drape_scene(res=256, layers=8, gap=2e-3, noise=1e-4, size=1.0, radius=0.2,
                        clearance=5e-3, seed=3, ghat=1e-3, friction=0.5, density=0.2,
                        young=1e4, poisson=0.25, bend=10.0, n_around=96, n_along=32):
    """Stacked-layer drape: `layers` square sheets (res x res, side `size`,
    xz plane) stacked `gap` apart with N(0, noise) vertex noise, the lowest
    `clearance` above a horizontal cylinder of `radius` along z. The cylinder
    is part of the dynamic mesh with every vertex pinned (Dirichlet), its
    faces and hinges excluded from the elastic terms, so it can be driven
    kinematically (rotation: `spin_fix_positions`). Same dict layout as
    drape_scene, plus 'cyl' (indices of the cylinder vertices) and 'cyl_rest'."""
    rng = np.random.default_rng(seed)
    f0 = rect_faces(res, res).astype(np.int64)
    i, j = np.meshgrid(np.arange(res), np.arange(res), indexing="ij")
    h = size / (res - 1)
    xs, fs, rests = [], [], []
    for k in range(layers):
        p = np.stack([i.reshape(-1) * h - size / 2, np.zeros(res * res),
                      j.reshape(-1) * h - size / 2], 1)
        rests.append(p.copy())
        p[:, 1] = radius + clearance + k * gap + rng.normal(0.0, noise, res * res)
        xs.append(p)
        fs.append(f0 + k * res * res)
    cx, cf = cylinder(radius, 1.2 * size, n_around, n_along)
    ns = layers * res * res
    x = np.concatenate(xs + [cx.astype(np.float64)])
    rest = np.concatenate(rests + [cx.astype(np.float64)])
    faces = np.concatenate(fs + [cf.astype(np.int64) + ns]).astype(np.uint32)
    nfs = layers * len(f0)
    shell = shell_rest(rest, faces, density, young, poisson, bend)
    mesh = contact_mesh(x.astype(np.float32), faces, ghat=ghat, offset=0.0)
    cyl = np.arange(ns, len(x))
    mesh["vflag"][cyl] |= 1
    shell["vflag"] = mesh["vflag"].copy()
    shell["fflag"][nfs:] = 1
    hv = shell["hinge"].astype(np.int64)
    shell["hflag"][(hv >= ns).any(1)] = 1
    mesh["fflag"][nfs:] = 1
    e = mesh["edge"].astype(np.int64)
    mesh["eflag"][(e >= ns).all(1)] = 1
    emass = (0.5 * (shell["vmass"][e[:, 0]] + shell["vmass"][e[:, 1]])).astype(np.float32)
    fric = {"v": np.full(mesh["nv"], friction, np.float32),
            "e": np.full(mesh["ne"], friction, np.float32),
            "f": np.full(mesh["nf"], friction, np.float32)}
    static = None
    return dict(mesh=mesh, shell=shell, static=static, prev=mesh["x"].copy(), emass=emass,
                fric=fric, fix_pos=mesh["x"].copy(), cyl=cyl, cyl_rest=cx)


def spin_fix_positions(sc, t, omega=2.0):
    """Fix positions of the cylinder rotated by omega * t about z (all other
    entries keep their values)."""
    c, s = np.cos(omega * t), np.sin(omega * t)
    r = sc["cyl_rest"].astype(np.float64)
    fix = sc["fix_pos"]
    fix[sc["cyl"], 0] = c * r[:, 0] - s * r[:, 1]
    fix[sc["cyl"], 1] = s * r[:, 0] + c * r[:, 1]
    fix[sc["cyl"], 2] = r[:, 2]
    return fix


def pack_scenes(scs):
    """Many independent scenes (drape_scene layout, each with a static mesh
    or none) as one scene dict for a many-scene stepper: vertices, edges,
    faces and hinges of scene c follow those of scene c - 1, every index is
    shifted accordingly, and mesh['vscene'] / static['vscene'] give each
    vertex's scene. Scene c's vertices are rows[c] .. rows[c + 1] of the
    result (returned beside it)."""
    ns = len(scs)
    nvs = [int(sc["mesh"]["nv"]) for sc in scs]
    nes = [int(sc["mesh"]["ne"]) for sc in scs]
    nfs = [int(sc["mesh"]["nf"]) for sc in scs]
    rows = np.concatenate([[0], np.cumsum(nvs)]).astype(np.int64)
    eoff = np.concatenate([[0], np.cumsum(nes)]).astype(np.int64)
    foff = np.concatenate([[0], np.cumsum(nfs)]).astype(np.int64)
    cat = lambda parts, dt: np.ascontiguousarray(np.concatenate(parts), dt)

    def csr(key_off, vals, row_base, val_base):
        offs, cols = [], []
        base = 0
        for c, sc in enumerate(scs):
            o = sc["mesh"][key_off].astype(np.int64)
            offs.append(o[:-1] + base if c < ns - 1 else o + base)
            cols.append(sc["mesh"][vals].astype(np.int64) + val_base[c])
            base += int(o[-1])
        return cat(offs, np.uint32), cat(cols, np.uint32)

    m = dict(nv=int(rows[-1]), ne=int(eoff[-1]), nf=int(foff[-1]), nscene=ns)
    m["x"] = cat([sc["mesh"]["x"].reshape(-1, 3) for sc in scs], np.float32)
    m["edge"] = cat([sc["mesh"]["edge"].astype(np.int64) + rows[c] for c, sc in enumerate(scs)],
                    np.uint32)
    m["face"] = cat([sc["mesh"]["face"].astype(np.int64) + rows[c] for c, sc in enumerate(scs)],
                    np.uint32)
    m["v2e_off"], m["v2e"] = csr("v2e_off", "v2e", rows, eoff)
    m["v2f_off"], m["v2f"] = csr("v2f_off", "v2f", rows, foff)
    m["e2f_off"], m["e2f"] = csr("e2f_off", "e2f", eoff, foff)
    for k in ("v_ghat", "v_off", "e_ghat", "e_off", "f_ghat", "f_off"):
        m[k] = cat([sc["mesh"][k] for sc in scs], np.float32)
    for k in ("vflag", "vbody", "eflag", "fflag"):
        m[k] = cat([sc["mesh"][k] for sc in scs], np.uint32)
    m["vscene"] = np.repeat(np.arange(ns), nvs).astype(np.uint32)

    sh = dict(nv=m["nv"], nf=m["nf"])
    nhs = [int(sc["shell"]["nh"]) for sc in scs]
    sh["nh"] = int(sum(nhs))
    sh["face"] = m["face"]
    sh["hinge"] = cat([sc["shell"]["hinge"].astype(np.int64).reshape(-1, 4) + rows[c]
                       for c, sc in enumerate(scs)], np.uint32)
    offs, cols, base = [], [], 0
    for c, sc in enumerate(scs):
        o = sc["shell"]["v2f_off"].astype(np.int64)
        offs.append(o[:-1] + base if c < ns - 1 else o + base)
        cols.append(sc["shell"]["v2f"].astype(np.int64) + foff[c])
        base += int(o[-1])
    sh["v2f_off"], sh["v2f"] = cat(offs, np.uint32), cat(cols, np.uint32)
    for k in ("vmass", "varea", "fmass", "inv_rest", "f_mu", "f_lambda", "h_rest", "h_length",
              "h_area", "h_bend"):
        sh[k] = cat([np.asarray(sc["shell"][k], np.float32).reshape(
            (-1, 4) if k == "inv_rest" else (-1,)) for sc in scs], np.float32)
    for k in ("vflag", "fflag", "hflag"):
        sh[k] = cat([sc["shell"][k] for sc in scs], np.uint32)

    statics = [sc.get("static") for sc in scs]
    static = None
    if any(s is not None for s in statics):
        snv = [int(s["nv"]) if s is not None else 0 for s in statics]
        srow = np.concatenate([[0], np.cumsum(snv)]).astype(np.int64)
        parts = [s for s in statics if s is not None]
        static = dict(nv=int(srow[-1]), ne=int(sum(int(s["ne"]) for s in parts)),
                      nf=int(sum(int(s["nf"]) for s in parts)))
        static["x"] = cat([s["x"].reshape(-1, 3) for s in parts], np.float32)
        static["edge"] = cat([s["edge"].astype(np.int64) + srow[c]
                              for c, s in enumerate(statics) if s is not None], np.uint32)
        static["face"] = cat([s["face"].astype(np.int64) + srow[c]
                              for c, s in enumerate(statics) if s is not None], np.uint32)
        for k in ("v_ghat", "v_off", "e_ghat", "e_off", "f_ghat", "f_off", "v_fric", "e_fric",
                  "f_fric"):
            if all(k in s for s in parts):
                static[k] = cat([s[k] for s in parts], np.float32)
        static["vscene"] = np.repeat(np.arange(ns), snv).astype(np.uint32)

    fric = {k: cat([sc["fric"][k] for sc in scs], np.float32) for k in ("v", "e", "f")}
    fix = None
    if any(sc.get("fix_pos") is not None for sc in scs):
        fix = cat([sc["fix_pos"] if sc.get("fix_pos") is not None else sc["mesh"]["x"]
                   for sc in scs], np.float32)
    packed = dict(mesh=m, shell=sh, static=static,
                  prev=cat([sc["prev"].reshape(-1, 3) for sc in scs], np.float32),
                  emass=cat([sc["emass"] for sc in scs], np.float32), fric=fric, fix_pos=fix)
    return packed, rows
