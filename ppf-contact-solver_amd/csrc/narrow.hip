// narrow.hip -- contact proximity narrow phase on MI355X.
//
// Replaces the stage-0 traversal of contact::embed_contact_force_hessian
// (contact/contact.cu:1622-1745) and the acceptance rules of its four pair
// kinds (contact.cu:432-870): point-face, point-edge (multiplicity from the
// edge's faces), point-point (index < vertex, multiplicity from the incident
// edges and faces) and edge-edge (edge < index, no shared vertex). Queries run
// in BVH leaf (Morton) order, one thread each, with the traversal stack in LDS.
// Children are tested before they are pushed. Two passes -- count, exclusive
// scan, fill -- give every query a contiguous, deterministic output range. The
// count pass keeps each query's first pairs; the fill pass copies those and
// walks the trees again only for queries with more.
#include "bvh_query.hpp"
#include "ccd.hpp"
#include "distance.hpp"
#include "narrow.hpp"
#include "scan.hpp"

namespace ppf {

namespace {

constexpr int kQT = kQueryThreads;
constexpr int kStack = kQueryStack;
using Q = QBox;

// Pairs the count pass keeps per query. The fill pass then copies them
// instead of walking the tree again; only queries with more pairs walk again.
constexpr uint32_t kStash = 8;

struct Ctx {
    ppf_contact_mesh m;
    BvhView fb, eb, vb;
    uint32_t *cnt;      // per query (vertex leaves, then edge leaves)
    const uint32_t *ofs;
    ppf_pair *out;
    uint64_t cap;
    uint32_t *overflow; // [0] stack overflow, [1] output overflow
    ppf_pair *stash;    // kStash per query, filled by the count pass (may be null)
    uint32_t *redo;     // fill pass: only these leaf positions (null: all)
    uint32_t nredo;
    unsigned long long *stats; // walk statistics of the count pass (null: off)
};

__device__ inline f3 X(const ppf_contact_mesh &m, uint32_t v) {
    return load3(m.x + 3 * (size_t)v);
}
__device__ inline bool vdyn(const ppf_contact_mesh &m, uint32_t v) {
    return (m.vflag[v] & 1u) == 0;
}
__device__ inline bool vcol(const ppf_contact_mesh &m, uint32_t v) {
    return (m.vflag[v] & 2u) != 0;
}

// The count pass (FILL false) counts the query's pairs in `slot` and keeps
// its first kStash in the stash; the fill pass writes them from the query's
// offset on.
template <bool FILL>
__device__ inline void emit(const Ctx &c, uint64_t &slot, uint32_t qpos, uint32_t kind,
                            uint32_t a, uint32_t b, uint32_t count) {
    if (FILL) {
        if (slot < c.cap) {
            c.out[slot] = ppf_pair{kind, a, b, count};
        } else {
            atomicOr(c.overflow + 1, 1u);
        }
    } else if (c.stash && slot < kStash) {
        c.stash[(size_t)qpos * kStash + slot] = ppf_pair{kind, a, b, count};
    }
    ++slot;
}

// Leaf position this thread queries: one per leaf, or (a fill pass after a
// stashed count) the listed ones. False: none. Consecutive leaves share an
// XCD (and its L2), as do their tree walks.
template <bool FILL>
__device__ inline bool query_pos(const Ctx &c, uint32_t n, uint32_t &leaf_pos) {
    const uint32_t t = xcd_logical_block() * kQT + threadIdx.x;
    if (!FILL || !c.redo) {
        leaf_pos = t;
        return t < n;
    }
    if (t >= c.nredo) {
        return false;
    }
    leaf_pos = c.redo[t];
    return true;
}

template <bool FILL>
__global__ __launch_bounds__(kQT) void k_vertex_queries(Ctx c) {
    __shared__ uint32_t stk_all[kStack * kQT];
    uint32_t leaf_pos;
    const ppf_contact_mesh &m = c.m;
    const bool has = query_pos<FILL>(c, c.vb.n, leaf_pos);
    WalkCount wc;
    [&] {
    if (!has) {
        return;
    }
    uint32_t *stk = stk_all + threadIdx.x;
    const uint32_t i = c.vb.node[2 * (size_t)leaf_pos] - 1;
    const f3 p = X(m, i);
    const float qm = 0.5f * m.v_ghat[i] + m.v_off[i];
    Q q = (m.vflag[i] & 4u) ? empty_box()
                            : Q{{p.x - qm, p.y - qm, p.z - qm}, {p.x + qm, p.y + qm, p.z + qm}};
    q.scene = vertex_scene(m.vscene, m.nscene, i);
    uint64_t slot = FILL ? (uint64_t)c.ofs[leaf_pos] : 0;
    uint32_t total = 0;
    // point-face (contact.cu:717-785)
    total += traverse(c.fb, q, stk, c.overflow, [&](uint32_t fi) {
        const uint32_t f0 = m.face[3 * (size_t)fi], f1 = m.face[3 * (size_t)fi + 1],
                       f2 = m.face[3 * (size_t)fi + 2];
        const bool either_dyn = vdyn(m, i) || (m.fflag[fi] & 1u) == 0;
        const bool same_body = m.vbody[i] != 0 && m.vbody[i] == m.vbody[f0];
        const bool both_col = vcol(m, i) && vcol(m, f0);
        if (f0 == i || f1 == i || f2 == i || !either_dyn || same_body || both_col) {
            return false;
        }
        const f3 t0 = X(m, f0), t1 = X(m, f1), t2 = X(m, f2);
        float w[3];
        pt_coeff(p, t0, t1, t2, w);
        if (!interior<3>(w)) {
            return false;
        }
        const f3 e = (w[0] * (p - t0) + w[1] * (p - t1)) + w[2] * (p - t2);
        const float off = m.v_off[i] + m.f_off[fi];
        const float gh = 0.5f * (m.v_ghat[i] + m.f_ghat[fi]);
        const float r = gh + off;
        if (dot3(e, e) < r * r) {
            emit<FILL>(c, slot, leaf_pos, 0, i, fi, 1);
            return true;
        }
        return false;
    }, &wc);
    // point-edge (contact.cu:622-715)
    total += traverse(c.eb, q, stk, c.overflow, [&](uint32_t ei) {
        const uint32_t e0 = m.edge[2 * (size_t)ei], e1 = m.edge[2 * (size_t)ei + 1];
        const bool either_dyn = vdyn(m, i) || (m.eflag[ei] & 1u) == 0;
        const bool same_body = m.vbody[i] != 0 && m.vbody[i] == m.vbody[e0];
        const bool both_col = vcol(m, i) && vcol(m, e0);
        if (e0 == i || e1 == i || !either_dyn || same_body || both_col) {
            return false;
        }
        const f3 a = X(m, e0), b = X(m, e1);
        float w[2];
        pe_coeff(p, a, b, w);
        if (!interior<2>(w)) {
            return false;
        }
        const f3 e = w[0] * (p - a) + w[1] * (p - b);
        const float off = m.v_off[i] + m.e_off[ei];
        const float gh = 0.5f * (m.v_ghat[i] + m.e_ghat[ei]);
        const float r = gh + off;
        if (!(dot3(e, e) < r * r)) {
            return false;
        }
        uint32_t count = 0, nofric = 0;
        const uint32_t k0 = m.e2f_off[ei], k1 = m.e2f_off[ei + 1];
        if (k1 > k0) {
            for (uint32_t k = k0; k < k1; ++k) {
                const uint32_t f = m.e2f[k];
                const uint32_t g0 = m.face[3 * (size_t)f], g1 = m.face[3 * (size_t)f + 1],
                               g2 = m.face[3 * (size_t)f + 2];
                if (g0 == i || g1 == i || g2 == i) {
                    ++count;
                    nofric = 1;
                } else {
                    float cc[3];
                    pt_coeff(p, X(m, g0), X(m, g1), X(m, g2), cc);
                    if (!interior<3>(cc)) {
                        ++count;
                    }
                }
            }
        } else {
            count = 1;
        }
        if (count) {
            emit<FILL>(c, slot, leaf_pos, 1u | (nofric << 8), i, ei, count);
            return true;
        }
        return false;
    }, &wc);
    // point-point (contact.cu:432-620)
    total += traverse(c.vb, q, stk, c.overflow, [&](uint32_t j) {
        const bool either_dyn = vdyn(m, i) || vdyn(m, j);
        const bool same_body = m.vbody[i] != 0 && m.vbody[i] == m.vbody[j];
        const bool both_col = vcol(m, i) && vcol(m, j);
        if (!(j < i) || !either_dyn || same_body || both_col) {
            return false;
        }
        const f3 e = p - X(m, j);
        const float off = m.v_off[i] + m.v_off[j];
        const float gh = 0.5f * (m.v_ghat[i] + m.v_ghat[j]);
        const float r = gh + off;
        if (!(dot3(e, e) < r * r)) {
            return false;
        }
        uint32_t count = 0, nofric = 0;
        const uint32_t k0 = m.v2e_off[j], k1 = m.v2e_off[j + 1];
        if (k1 > k0) {
            for (uint32_t k = k0; k < k1; ++k) {
                const uint32_t ed = m.v2e[k];
                const uint32_t a0 = m.edge[2 * (size_t)ed], a1 = m.edge[2 * (size_t)ed + 1];
                if (a0 == i || a1 == i) {
                    ++count;
                    nofric = 1;
                } else {
                    float cc[2];
                    pe_coeff(p, X(m, a0), X(m, a1), cc);
                    if (!interior<2>(cc)) {
                        ++count;
                    }
                }
            }
            for (uint32_t k = m.v2f_off[j]; k < m.v2f_off[j + 1]; ++k) {
                const uint32_t f = m.v2f[k];
                const uint32_t g0 = m.face[3 * (size_t)f], g1 = m.face[3 * (size_t)f + 1],
                               g2 = m.face[3 * (size_t)f + 2];
                if (g0 == i || g1 == i || g2 == i) {
                    ++count;
                    nofric = 1;
                } else {
                    float cc[3];
                    pt_coeff(p, X(m, g0), X(m, g1), X(m, g2), cc);
                    if (!interior<3>(cc)) {
                        ++count;
                    }
                }
            }
        } else {
            count = 1;
        }
        if (count) {
            emit<FILL>(c, slot, leaf_pos, 2u | (nofric << 8), i, j, count);
            return true;
        }
        return false;
    }, &wc);
    if (!FILL) {
        c.cnt[leaf_pos] = total;
    }
    }();
    if (!FILL) {
        walk_flush(c.stats, kWalkProxVertex, wc, has);
    }
}

// edge-edge (contact.cu:787-870), queries in edge-BVH leaf order.
template <bool FILL>
__global__ __launch_bounds__(kQT) void k_edge_queries(Ctx c) {
    __shared__ uint32_t stk_all[kStack * kQT];
    uint32_t leaf_pos;
    const ppf_contact_mesh &m = c.m;
    const bool has = query_pos<FILL>(c, c.eb.n, leaf_pos);
    WalkCount wc;
    [&] {
    if (!has) {
        return;
    }
    uint32_t *stk = stk_all + threadIdx.x;
    const uint32_t ei = c.eb.node[2 * (size_t)leaf_pos] - 1;
    const uint32_t a0 = m.edge[2 * (size_t)ei], a1 = m.edge[2 * (size_t)ei + 1];
    const f3 pa = X(m, a0), pb = X(m, a1);
    const float qm = 0.5f * m.e_ghat[ei] + m.e_off[ei];
    Q q{{fminf(pa.x, pb.x), fminf(pa.y, pb.y), fminf(pa.z, pb.z)},
        {fmaxf(pa.x, pb.x), fmaxf(pa.y, pb.y), fmaxf(pa.z, pb.z)}};
    if (qm != 0.0f) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            q.mn[d] -= qm;
            q.mx[d] += qm;
        }
    }
    if (m.eflag[ei] & 4u) {
        q = empty_box();
    }
    q.scene = vertex_scene(m.vscene, m.nscene, a0);
    const uint32_t qpos = c.vb.n + leaf_pos;
    uint64_t slot = FILL ? (uint64_t)c.ofs[qpos] : 0;
    const uint32_t total = traverse(c.eb, q, stk, c.overflow, [&](uint32_t ej) {
        const uint32_t b0 = m.edge[2 * (size_t)ej], b1 = m.edge[2 * (size_t)ej + 1];
        const bool either_dyn = (m.eflag[ei] & 1u) == 0 || (m.eflag[ej] & 1u) == 0;
        const bool same_body = m.vbody[a0] != 0 && m.vbody[a0] == m.vbody[b0];
        const bool both_col = vcol(m, a0) && vcol(m, b0);
        const bool shared = a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1;
        if (!(ei < ej) || shared || !either_dyn || same_body || both_col) {
            return false;
        }
        const f3 qa = X(m, b0), qb = X(m, b1);
        float w[4];
        ee_coeff(pa, pb, qa, qb, w);
        if (!interior<4>(w)) {
            return false;
        }
        const f3 x0 = w[0] * pa + w[1] * pb;
        const f3 x1 = w[2] * qa + w[3] * qb;
        const f3 e = x0 - x1;
        const float off = m.e_off[ei] + m.e_off[ej];
        const float gh = 0.5f * (m.e_ghat[ei] + m.e_ghat[ej]);
        const float r = gh + off;
        if (dot3(e, e) < r * r) {
            emit<FILL>(c, slot, qpos, 3, ei, ej, 1);
            return true;
        }
        return false;
    }, &wc);
    if (!FILL) {
        c.cnt[qpos] = total;
    }
    }();
    if (!FILL) {
        walk_flush(c.stats, kWalkProxEdge, wc, has);
    }
}

// Copies the stashed pairs of every query with at most kStash pairs to its
// output range (one thread per stash entry) and lists the other queries,
// vertex leaves from redo[0] on, edge leaves from redo[nv] on, counted in
// overflow[2] and overflow[3]. The list order varies from run to run; the
// output does not (every query writes its own range).
__global__ __launch_bounds__(256) void k_unstash(Ctx c) {
    const uint32_t nq = c.vb.n + c.eb.n;
    const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint32_t q = (uint32_t)(t / kStash), k = (uint32_t)(t % kStash);
    uint32_t n = 0;
    if (q < nq) {
        n = c.cnt[q];
        if (k < n && n <= kStash) {
            const uint64_t slot = (uint64_t)c.ofs[q] + k;
            if (slot < c.cap) {
                c.out[slot] = c.stash[t];
            } else {
                atomicOr(c.overflow + 1, 1u);
            }
        }
    }
    // one append per wave and kind
    const bool redo = q < nq && k == 0 && n > kStash;
    const bool edge = q >= c.vb.n;
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (int kind = 0; kind < 2; ++kind) {
        const bool mine = redo && edge == (kind == 1);
        const uint64_t m = __ballot(mine);
        if (m == 0) {
            continue;
        }
        const uint32_t first = (uint32_t)__builtin_ctzll(m);
        uint32_t at = 0;
        if (lane == first) {
            at = atomicAdd(c.overflow + 2 + kind, (uint32_t)__popcll(m));
        }
        at = __shfl(at, (int)first);
        if (mine) {
            uint32_t *list = kind ? c.redo + c.vb.n : c.redo;
            list[at + lanes_below(m)] = kind ? q - c.vb.n : q;
        }
    }
}

// ---- collision-mesh pairs (contact.cu:899-1152, 1789-1954) -------------------

struct SCtx {
    ppf_contact_mesh m;
    ppf_static_mesh sm;
    const float *vmass, *emass, *fmass;
    BvhView fb, eb, vb, sfb, seb;
    uint32_t *cnt; // per query: vertex leaves, static vertices, edge leaves
    const uint32_t *ofs;
    ppf_pair *out;
    uint64_t cap;
    uint32_t *overflow;
};

template <bool FILL>
__device__ inline void emit_s(const SCtx &c, uint64_t &slot, uint32_t kind, uint32_t a,
                              uint32_t b) {
    if (FILL) {
        if (slot < c.cap) {
            c.out[slot] = ppf_pair{kind, a, b, 1};
        } else {
            atomicOr(c.overflow + 1, 1u);
        }
        ++slot;
    }
}

__device__ inline f3 SX(const ppf_static_mesh &sm, uint32_t v) {
    return load3(sm.x + 3 * (size_t)v);
}

// Dynamic vertex vs static face (kind 4), one thread per vertex leaf.
template <bool FILL> __global__ __launch_bounds__(kQT) void k_m2c(SCtx c) {
    __shared__ uint32_t stk_all[kStack * kQT];
    const uint32_t qpos = blockIdx.x * kQT + threadIdx.x;
    if (qpos >= c.vb.n) {
        return;
    }
    uint32_t *stk = stk_all + threadIdx.x;
    const ppf_contact_mesh &m = c.m;
    const uint32_t i = c.vb.node[2 * (size_t)qpos] - 1;
    uint64_t slot = FILL ? (uint64_t)c.ofs[qpos] : 0;
    uint32_t total = 0;
    if ((m.vflag[i] & 5u) == 0 && c.vmass[i] > 0.0f) {
        const f3 p = X(m, i);
        QBox q = point_box(p, 0.5f * m.v_ghat[i] + m.v_off[i]);
        q.scene = vertex_scene(m.vscene, m.nscene, i);
        total = traverse(c.sfb, q, stk, c.overflow, [&](uint32_t fi) {
            const uint32_t f0 = c.sm.face[3 * (size_t)fi], f1 = c.sm.face[3 * (size_t)fi + 1],
                           f2 = c.sm.face[3 * (size_t)fi + 2];
            const f3 t0 = SX(c.sm, f0) - p, t1 = SX(c.sm, f1) - p, t2 = SX(c.sm, f2) - p;
            float w[3];
            pt_coeff_unclassified(mk3(0.0f, 0.0f, 0.0f), t0, t1, t2, w);
            const f3 y = (w[0] * t0 + w[1] * t1) + w[2] * t2;
            const float off = m.v_off[i] + c.sm.f_off[fi];
            const float gh = 0.5f * (m.v_ghat[i] + c.sm.f_ghat[fi]);
            const float r = off + gh;
            if (dot3(y, y) < r * r) {
                emit_s<FILL>(c, slot, 4, i, fi);
                return true;
            }
            return false;
        });
    }
    if (!FILL) {
        c.cnt[qpos] = total;
    }
}

// Static vertex vs dynamic face (kind 5), one thread per static vertex.
template <bool FILL> __global__ __launch_bounds__(kQT) void k_c2m(SCtx c) {
    __shared__ uint32_t stk_all[kStack * kQT];
    const uint32_t sv = blockIdx.x * kQT + threadIdx.x;
    if (sv >= c.sm.nv) {
        return;
    }
    uint32_t *stk = stk_all + threadIdx.x;
    const ppf_contact_mesh &m = c.m;
    const uint32_t qpos = c.vb.n + sv;
    uint64_t slot = FILL ? (uint64_t)c.ofs[qpos] : 0;
    const f3 y = SX(c.sm, sv);
    QBox q = point_box(y, 0.5f * c.sm.v_ghat[sv] + c.sm.v_off[sv]);
    q.scene = vertex_scene(c.sm.vscene, m.nscene, sv);
    const uint32_t total = traverse(c.fb, q, stk, c.overflow, [&](uint32_t fi) {
        if ((m.fflag[fi] & 1u) != 0 || !(c.fmass[fi] > 0.0f)) {
            return false;
        }
        const uint32_t f0 = m.face[3 * (size_t)fi], f1 = m.face[3 * (size_t)fi + 1],
                       f2 = m.face[3 * (size_t)fi + 2];
        const f3 t0 = X(m, f0) - y, t1 = X(m, f1) - y, t2 = X(m, f2) - y;
        float w[3];
        pt_coeff_unclassified(mk3(0.0f, 0.0f, 0.0f), t0, t1, t2, w);
        const f3 e = (w[0] * t0 + w[1] * t1) + w[2] * t2;
        const float off = m.f_off[fi] + c.sm.v_off[sv];
        const float gh = 0.5f * (m.f_ghat[fi] + c.sm.v_ghat[sv]);
        const float r = off + gh;
        if (dot3(e, e) < r * r) {
            emit_s<FILL>(c, slot, 5, sv, fi);
            return true;
        }
        return false;
    });
    if (!FILL) {
        c.cnt[qpos] = total;
    }
}

// Dynamic edge vs static edge (kind 6), one thread per edge leaf.
template <bool FILL> __global__ __launch_bounds__(kQT) void k_cm_ee(SCtx c) {
    __shared__ uint32_t stk_all[kStack * kQT];
    const uint32_t lp = blockIdx.x * kQT + threadIdx.x;
    if (lp >= c.eb.n) {
        return;
    }
    uint32_t *stk = stk_all + threadIdx.x;
    const ppf_contact_mesh &m = c.m;
    const uint32_t qpos = c.vb.n + c.sm.nv + lp;
    const uint32_t ei = c.eb.node[2 * (size_t)lp] - 1;
    uint64_t slot = FILL ? (uint64_t)c.ofs[qpos] : 0;
    uint32_t total = 0;
    if ((m.eflag[ei] & 5u) == 0 && c.emass[ei] > 0.0f) {
        const uint32_t a0 = m.edge[2 * (size_t)ei], a1 = m.edge[2 * (size_t)ei + 1];
        const f3 p0 = X(m, a0), p1 = X(m, a1);
        QBox q = seg_box(p0, p1, 0.5f * m.e_ghat[ei] + m.e_off[ei]);
        q.scene = vertex_scene(m.vscene, m.nscene, a0);
        total = traverse(c.seb, q, stk, c.overflow, [&](uint32_t sj) {
            const uint32_t b0 = c.sm.edge[2 * (size_t)sj], b1 = c.sm.edge[2 * (size_t)sj + 1];
            const f3 q0 = SX(c.sm, b0), q1 = SX(c.sm, b1);
            float w[4];
            ee_coeff(p0, p1, q0, q1, w);
            const f3 e = (w[0] * p0 + w[1] * p1) - (w[2] * q0 + w[3] * q1);
            const float off = m.e_off[ei] + c.sm.e_off[sj];
            const float gh = 0.5f * (m.e_ghat[ei] + c.sm.e_ghat[sj]);
            const float r = off + gh;
            if (dot3(e, e) < r * r) {
                emit_s<FILL>(c, slot, 6, ei, sj);
                return true;
            }
            return false;
        });
    }
    if (!FILL) {
        c.cnt[qpos] = total;
    }
}

uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

} // namespace

Narrow::~Narrow() {
    if (cnt_) (void)dev_free(cnt_);
    if (ofs_) (void)dev_free(ofs_);
    if (scratch_) (void)dev_free(scratch_);
    if (stash_) (void)dev_free(stash_);
    if (redo_) (void)dev_free(redo_);
    if (flags_) (void)dev_free(flags_);
    if (h_flags_) (void)host_free(h_flags_);
}

int Narrow::reserve(uint32_t queries) {
    if (queries <= qcap_) {
        return PPF_OK;
    }
    if (cnt_) (void)dev_free(cnt_);
    if (ofs_) (void)dev_free(ofs_);
    if (scratch_) (void)dev_free(scratch_);
    if (stash_) (void)dev_free(stash_);
    if (redo_) (void)dev_free(redo_);
    cnt_ = ofs_ = scratch_ = redo_ = nullptr;
    stash_ = nullptr;
    qcap_ = 0;
    PPF_HIP(dev_alloc(&cnt_, (queries + 1) * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&ofs_, (queries + 1) * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&scratch_, scan_scratch_words(queries + 1) * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&stash_, (size_t)queries * kStash * sizeof(ppf_pair)));
    PPF_HIP(dev_alloc(&redo_, (size_t)queries * sizeof(uint32_t)));
    if (!flags_) {
        PPF_HIP(dev_alloc(&flags_, 4 * sizeof(uint32_t)));
        // four flag words, then the pair total
        PPF_HIP(host_alloc(reinterpret_cast<void **>(&h_flags_),
                              8 * sizeof(uint32_t)));
    }
    qcap_ = queries;
    return PPF_OK;
}

int Narrow::build_trees(const ppf_contact_mesh &m, hipStream_t s) {
    PrimSet f, e, v;
    f.n = m.nf, f.arity = 3, f.elem = m.face, f.ghat = m.f_ghat, f.offset = m.f_off;
    e.n = m.ne, e.arity = 2, e.elem = m.edge, e.ghat = m.e_ghat, e.offset = m.e_off;
    v.n = m.nv, v.arity = 1, v.elem = nullptr, v.ghat = m.v_ghat, v.offset = m.v_off;
    f.vscene = e.vscene = v.vscene = m.vscene;
    f.nscene = e.nscene = v.nscene = m.nscene;
    int rc = face_.build(f, m.x, nullptr, 0.0f, s);
    if (rc == PPF_OK) {
        rc = edge_.build(e, m.x, nullptr, 0.0f, s);
    }
    if (rc == PPF_OK) {
        rc = vert_.build(v, m.x, nullptr, 0.0f, s);
    }
    return rc;
}

int Narrow::refit_trees(const ppf_contact_mesh &m, hipStream_t s) {
    if (face_.view().n != m.nf || edge_.view().n != m.ne || vert_.view().n != m.nv) {
        return build_trees(m, s);
    }
    PrimSet f, e, v;
    f.n = m.nf, f.arity = 3, f.elem = m.face, f.ghat = m.f_ghat, f.offset = m.f_off;
    e.n = m.ne, e.arity = 2, e.elem = m.edge, e.ghat = m.e_ghat, e.offset = m.e_off;
    v.n = m.nv, v.arity = 1, v.elem = nullptr, v.ghat = m.v_ghat, v.offset = m.v_off;
    f.vscene = e.vscene = v.vscene = m.vscene;
    f.nscene = e.nscene = v.nscene = m.nscene;
    int rc = face_.refit(f, m.x, nullptr, 0.0f, s);
    if (rc == PPF_OK) {
        rc = edge_.refit(e, m.x, nullptr, 0.0f, s);
    }
    if (rc == PPF_OK) {
        rc = vert_.refit(v, m.x, nullptr, 0.0f, s);
    }
    return rc;
}

int Narrow::proximity(const ppf_contact_mesh &m, ppf_pair *out, uint64_t cap,
                      uint64_t *count, hipStream_t s) {
    const uint32_t nq = m.nv + m.ne;
    int rc = reserve(nq);
    if (rc != PPF_OK) {
        return rc;
    }
    PPF_HIP(hipMemsetAsync(flags_, 0, 4 * sizeof(uint32_t), s));
    const bool fill = out && cap;
    Ctx c;
    c.m = m;
    c.fb = face_.view();
    c.eb = edge_.view();
    c.vb = vert_.view();
    c.cnt = cnt_;
    c.ofs = ofs_;
    c.out = out;
    c.cap = cap;
    c.overflow = flags_;
    c.stash = fill ? stash_ : nullptr;
    c.redo = nullptr;
    c.nredo = 0;
    c.stats = walk_stats_device();
    if (m.nv) {
        WalkClock clk(kWalkProxVertex, s);
        k_vertex_queries<false><<<xcd_grid(cdiv(m.nv, kQT)), kQT, 0, s>>>(c);
        PPF_HIP(hipGetLastError());
        clk.stop();
    }
    if (m.ne) {
        WalkClock clk(kWalkProxEdge, s);
        k_edge_queries<false><<<xcd_grid(cdiv(m.ne, kQT)), kQT, 0, s>>>(c);
        PPF_HIP(hipGetLastError());
        clk.stop();
    }
    PPF_HIP(hipMemsetAsync(cnt_ + nq, 0, sizeof(uint32_t), s));
    PPF_HIP(exclusive_scan_u32(cnt_, ofs_, nq + 1, scratch_, s));
    if (fill && nq) {
        c.redo = redo_;
        const uint64_t entries = (uint64_t)nq * kStash;
        k_unstash<<<(uint32_t)((entries + 255) / 256), 256, 0, s>>>(c);
        PPF_HIP(hipGetLastError());
    }
    // total and the redo counts in one readback
    PPF_HIP(hipMemcpyAsync(h_flags_ + 4, ofs_ + nq, sizeof(uint32_t),
                           hipMemcpyDeviceToHost, s));
    PPF_HIP(hipMemcpyAsync(h_flags_, flags_, 4 * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, s));
    PPF_HIP(hipStreamSynchronize(s));
    *count = h_flags_[4];
    if (fill) {
        const uint32_t nv_redo = h_flags_[2], ne_redo = h_flags_[3];
        if (nv_redo) {
            Ctx cv = c;
            cv.nredo = nv_redo;
            k_vertex_queries<true><<<xcd_grid(cdiv(nv_redo, kQT)), kQT, 0, s>>>(cv);
            PPF_HIP(hipGetLastError());
        }
        if (ne_redo) {
            Ctx ce = c;
            ce.redo = redo_ + m.nv;
            ce.nredo = ne_redo;
            k_edge_queries<true><<<xcd_grid(cdiv(ne_redo, kQT)), kQT, 0, s>>>(ce);
            PPF_HIP(hipGetLastError());
        }
        PPF_HIP(hipMemcpyAsync(h_flags_, flags_, 4 * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, s));
        PPF_HIP(hipStreamSynchronize(s));
    }
    if (h_flags_[0]) {
        set_error("proximity: a BVH traversal exceeded its %d-entry stack", kStack);
        return PPF_ERR_CAPACITY;
    }
    return PPF_OK;
}

int Narrow::build_static(const ppf_static_mesh &sm, uint32_t nscene, hipStream_t s) {
    if (sm.nv == 0) {
        return PPF_OK;
    }
    if (static_x_ == sm.x && sface_.view().n == sm.nf && sedge_.view().n == sm.ne &&
        static_nv_ == sm.nv) {
        return PPF_OK;
    }
    PrimSet f, e;
    f.n = sm.nf, f.arity = 3, f.elem = sm.face, f.ghat = sm.f_ghat, f.offset = sm.f_off;
    e.n = sm.ne, e.arity = 2, e.elem = sm.edge, e.ghat = sm.e_ghat, e.offset = sm.e_off;
    f.vscene = e.vscene = sm.vscene;
    f.nscene = e.nscene = nscene;
    int rc = sface_.build(f, sm.x, nullptr, 0.0f, s);
    if (rc == PPF_OK) {
        rc = sedge_.build(e, sm.x, nullptr, 0.0f, s);
    }
    if (rc == PPF_OK) {
        static_x_ = sm.x;
        static_nv_ = sm.nv;
    }
    return rc;
}

int Narrow::static_proximity(const ppf_contact_mesh &m, const ppf_static_mesh &sm,
                             const float *vmass, const float *emass, const float *fmass,
                             ppf_pair *out, uint64_t cap, uint64_t *count, hipStream_t s) {
    *count = 0;
    if (sm.nv == 0) {
        return PPF_OK;
    }
    int rc = build_static(sm, m.nscene, s);
    if (rc != PPF_OK) {
        return rc;
    }
    const uint32_t nq = m.nv + sm.nv + m.ne;
    rc = reserve(nq);
    if (rc != PPF_OK) {
        return rc;
    }
    PPF_HIP(hipMemsetAsync(flags_, 0, 4 * sizeof(uint32_t), s));
    SCtx c;
    c.m = m;
    c.sm = sm;
    c.vmass = vmass;
    c.emass = emass;
    c.fmass = fmass;
    c.fb = face_.view();
    c.eb = edge_.view();
    c.vb = vert_.view();
    c.sfb = sface_.view();
    c.seb = sedge_.view();
    c.cnt = cnt_;
    c.ofs = ofs_;
    c.out = out;
    c.cap = cap;
    c.overflow = flags_;
    auto launch = [&](bool fill) -> int {
        if (m.nv) {
            if (fill) k_m2c<true><<<cdiv(m.nv, kQT), kQT, 0, s>>>(c);
            else k_m2c<false><<<cdiv(m.nv, kQT), kQT, 0, s>>>(c);
            PPF_HIP(hipGetLastError());
        }
        if (fill) k_c2m<true><<<cdiv(sm.nv, kQT), kQT, 0, s>>>(c);
        else k_c2m<false><<<cdiv(sm.nv, kQT), kQT, 0, s>>>(c);
        PPF_HIP(hipGetLastError());
        if (m.ne) {
            if (fill) k_cm_ee<true><<<cdiv(m.ne, kQT), kQT, 0, s>>>(c);
            else k_cm_ee<false><<<cdiv(m.ne, kQT), kQT, 0, s>>>(c);
            PPF_HIP(hipGetLastError());
        }
        return PPF_OK;
    };
    rc = launch(false);
    if (rc != PPF_OK) {
        return rc;
    }
    PPF_HIP(hipMemsetAsync(cnt_ + nq, 0, sizeof(uint32_t), s));
    PPF_HIP(exclusive_scan_u32(cnt_, ofs_, nq + 1, scratch_, s));
    uint32_t total = 0;
    PPF_HIP(hipMemcpyAsync(&total, ofs_ + nq, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PPF_HIP(hipStreamSynchronize(s));
    *count = total;
    if (out && cap) {
        rc = launch(true);
        if (rc != PPF_OK) {
            return rc;
        }
    }
    PPF_HIP(hipMemcpyAsync(h_flags_, flags_, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PPF_HIP(hipStreamSynchronize(s));
    if (h_flags_[0]) {
        set_error("static proximity: a BVH traversal exceeded its %d-entry stack", kStack);
        return PPF_ERR_CAPACITY;
    }
    return PPF_OK;
}

} // namespace ppf
