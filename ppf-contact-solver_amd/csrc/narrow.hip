// narrow.hip -- contact proximity narrow phase on MI355X.
//
// Replaces the stage-0 traversal of contact::embed_contact_force_hessian
// (contact/contact.cu:1622-1745) and the acceptance rules of its four pair
// kinds (contact.cu:432-870): point-face, point-edge (multiplicity from the
// edge's faces), point-point (index < vertex, multiplicity from the incident
// edges and faces) and edge-edge (edge < index, no shared vertex). Queries run
// in BVH leaf (Morton) order, one thread each, with the traversal stack in LDS.
// Children are tested before they are pushed. Two passes -- count, exclusive
// scan, fill -- give every query a contiguous, deterministic output range.
#include "bvh_query.hpp"
#include "distance.hpp"
#include "narrow.hpp"
#include "scan.hpp"

namespace ppf {

namespace {

constexpr int kQT = kQueryThreads;
constexpr int kStack = kQueryStack;
using Q = QBox;

struct Ctx {
    ppf_contact_mesh m;
    BvhView fb, eb, vb;
    uint32_t *cnt;      // per query (vertex leaves, then edge leaves)
    const uint32_t *ofs;
    ppf_pair *out;
    uint64_t cap;
    uint32_t *overflow; // [0] stack overflow, [1] output overflow
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

template <bool FILL>
__device__ inline void emit(const Ctx &c, uint64_t &slot, uint32_t kind,
                            uint32_t a, uint32_t b, uint32_t count) {
    if (FILL) {
        if (slot < c.cap) {
            c.out[slot] = ppf_pair{kind, a, b, count};
        } else {
            atomicOr(c.overflow + 1, 1u);
        }
        ++slot;
    }
}

template <bool FILL>
__global__ __launch_bounds__(kQT) void k_vertex_queries(Ctx c) {
    __shared__ uint32_t stk_all[kStack * kQT];
    const uint32_t leaf_pos = blockIdx.x * kQT + threadIdx.x;
    const ppf_contact_mesh &m = c.m;
    if (leaf_pos >= c.vb.n) {
        return;
    }
    uint32_t *stk = stk_all + threadIdx.x;
    const uint32_t i = c.vb.node[2 * (size_t)leaf_pos] - 1;
    const f3 p = X(m, i);
    const float qm = 0.5f * m.v_ghat[i] + m.v_off[i];
    const Q q{{p.x - qm, p.y - qm, p.z - qm}, {p.x + qm, p.y + qm, p.z + qm}};
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
            emit<FILL>(c, slot, 0, i, fi, 1);
            return true;
        }
        return false;
    });
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
            emit<FILL>(c, slot, 1u | (nofric << 8), i, ei, count);
            return true;
        }
        return false;
    });
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
            emit<FILL>(c, slot, 2u | (nofric << 8), i, j, count);
            return true;
        }
        return false;
    });
    if (!FILL) {
        c.cnt[leaf_pos] = total;
    }
}

// edge-edge (contact.cu:787-870), queries in edge-BVH leaf order.
template <bool FILL>
__global__ __launch_bounds__(kQT) void k_edge_queries(Ctx c) {
    __shared__ uint32_t stk_all[kStack * kQT];
    const uint32_t leaf_pos = blockIdx.x * kQT + threadIdx.x;
    const ppf_contact_mesh &m = c.m;
    if (leaf_pos >= c.eb.n) {
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
            emit<FILL>(c, slot, 3, ei, ej, 1);
            return true;
        }
        return false;
    });
    if (!FILL) {
        c.cnt[qpos] = total;
    }
}

uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

} // namespace

Narrow::~Narrow() {
    if (cnt_) (void)hipFree(cnt_);
    if (ofs_) (void)hipFree(ofs_);
    if (scratch_) (void)hipFree(scratch_);
    if (flags_) (void)hipFree(flags_);
    if (h_flags_) (void)hipHostFree(h_flags_);
}

int Narrow::reserve(uint32_t queries) {
    if (queries <= qcap_) {
        return PPF_OK;
    }
    if (cnt_) (void)hipFree(cnt_);
    if (ofs_) (void)hipFree(ofs_);
    if (scratch_) (void)hipFree(scratch_);
    cnt_ = ofs_ = scratch_ = nullptr;
    PPF_HIP(hipMalloc(&cnt_, (queries + 1) * sizeof(uint32_t)));
    PPF_HIP(hipMalloc(&ofs_, (queries + 1) * sizeof(uint32_t)));
    PPF_HIP(hipMalloc(&scratch_, scan_scratch_words(queries + 1) * sizeof(uint32_t)));
    if (!flags_) {
        PPF_HIP(hipMalloc(&flags_, 4 * sizeof(uint32_t)));
        PPF_HIP(hipHostMalloc(reinterpret_cast<void **>(&h_flags_),
                              4 * sizeof(uint32_t), hipHostMallocDefault));
    }
    qcap_ = queries;
    return PPF_OK;
}

int Narrow::build_trees(const ppf_contact_mesh &m, hipStream_t s) {
    PrimSet f, e, v;
    f.n = m.nf, f.arity = 3, f.elem = m.face, f.ghat = m.f_ghat, f.offset = m.f_off;
    e.n = m.ne, e.arity = 2, e.elem = m.edge, e.ghat = m.e_ghat, e.offset = m.e_off;
    v.n = m.nv, v.arity = 1, v.elem = nullptr, v.ghat = m.v_ghat, v.offset = m.v_off;
    int rc = face_.build(f, m.x, nullptr, 0.0f, s);
    if (rc == PPF_OK) {
        rc = edge_.build(e, m.x, nullptr, 0.0f, s);
    }
    if (rc == PPF_OK) {
        rc = vert_.build(v, m.x, nullptr, 0.0f, s);
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
    if (m.nv) {
        k_vertex_queries<false><<<cdiv(m.nv, kQT), kQT, 0, s>>>(c);
        PPF_HIP(hipGetLastError());
    }
    if (m.ne) {
        k_edge_queries<false><<<cdiv(m.ne, kQT), kQT, 0, s>>>(c);
        PPF_HIP(hipGetLastError());
    }
    PPF_HIP(hipMemsetAsync(cnt_ + nq, 0, sizeof(uint32_t), s));
    PPF_HIP(exclusive_scan_u32(cnt_, ofs_, nq + 1, scratch_, s));
    uint32_t total = 0;
    PPF_HIP(hipMemcpyAsync(&total, ofs_ + nq, sizeof(uint32_t),
                           hipMemcpyDeviceToHost, s));
    PPF_HIP(hipStreamSynchronize(s));
    *count = total;
    if (out && cap) {
        if (m.nv) {
            k_vertex_queries<true><<<cdiv(m.nv, kQT), kQT, 0, s>>>(c);
            PPF_HIP(hipGetLastError());
        }
        if (m.ne) {
            k_edge_queries<true><<<cdiv(m.ne, kQT), kQT, 0, s>>>(c);
            PPF_HIP(hipGetLastError());
        }
    }
    PPF_HIP(hipMemcpyAsync(h_flags_, flags_, 4 * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, s));
    PPF_HIP(hipStreamSynchronize(s));
    if (h_flags_[0]) {
        set_error("proximity: a BVH traversal exceeded its %d-entry stack", kStack);
        return PPF_ERR_CAPACITY;
    }
    return PPF_OK;
}

} // namespace ppf
