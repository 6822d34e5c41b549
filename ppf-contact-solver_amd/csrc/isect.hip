// isect.hip -- post-step intersection check on MI355X.
//
// Replaces contact::check_intersection (contact/contact.cu:3056-3175) with its
// testers: face-edge (segment strictly crossing a triangle, sharing no
// vertex; contact.cu:2879-2946 over ppf_isect::edge_triangle_intersect,
// contact/intersect_core.hpp:33-81), edge-edge (closer than the summed
// offsets; contact.cu:2762-2877), collision-mesh face vs edge
// (contact.cu:2948-2980) and point-point among surface vertices
// (contact.cu:2982-3054). One thread per edge (per vertex) in BVH leaf
// order; every accepted pair takes a record slot from one counter, so the
// count is exact and the first `cap` records are kept.
#include "bvh_query.hpp"
#include "ccd.hpp"
#include "isect.hpp"

namespace ppf {

namespace {

constexpr int kQT = kQueryThreads;
constexpr int kStack = kQueryStack;

struct IsCtx {
    ppf_contact_mesh m;
    ppf_static_mesh sm;
    BvhView fb, eb, vb, cfb;
    const float *vmass, *emass, *fmass;
    uint32_t nsurf;
    ppf_isect_record *rec;
    uint32_t cap;
    uint32_t *ctr; // [0] pairs found, [1] stack overflow, [2 + c] found in scene c
    uint32_t ns;   // scenes counted separately (0: one scene)
};

__device__ inline f3 X(const float *x, uint32_t v) { return load3(x + 3 * (size_t)v); }
__device__ inline bool vcol(const ppf_contact_mesh &m, uint32_t v) {
    return (m.vflag[v] & 2u) != 0;
}

__device__ inline float dot3a(const float *a, const float *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// p inside triangle (0, d1, d2) by the 2x2 Gram solve; zero area -> false.
__device__ bool point_triangle_inside(const float *p, const float *d1, const float *d2) {
    const float a00 = dot3a(d1, d1), a01 = dot3a(d1, d2), a11 = dot3a(d2, d2);
    const float b0 = dot3a(d1, p), b1 = dot3a(d2, p);
    const float det = a00 * a11 - a01 * a01;
    if (det == 0.0f) {
        return false;
    }
    const float w0 = (a11 * b0 - a01 * b1) / det;
    const float w1 = (a00 * b1 - a01 * b0) / det;
    const float w2 = 1.0f - w0 - w1;
    float lo = w0 < w1 ? w0 : w1;
    lo = lo < w2 ? lo : w2;
    float hi = w0 > w1 ? w0 : w1;
    hi = hi > w2 ? hi : w2;
    return lo >= 0.0f && hi <= 1.0f;
}

// Segment (a0, a1) strictly crossing triangle (b0, b1, b2), in coordinates
// relative to b0.
__device__ bool edge_triangle(f3 a0, f3 a1, f3 b0, f3 b1, f3 b2) {
    const f3 d1 = b1 - b0, d2 = b2 - b0, e0 = a0 - b0, e1 = a1 - b0;
    const float n[3] = {d1.y * d2.z - d1.z * d2.y, d1.z * d2.x - d1.x * d2.z,
                        d1.x * d2.y - d1.y * d2.x};
    const float s1 = e0.x * n[0] + e0.y * n[1] + e0.z * n[2];
    const float s2 = e1.x * n[0] + e1.y * n[1] + e1.z * n[2];
    if (s1 * s2 < 0.0f) {
        const float t = s1 / (s1 - s2);
        const float r[3] = {(1.0f - t) * e0.x + t * e1.x, (1.0f - t) * e0.y + t * e1.y,
                            (1.0f - t) * e0.z + t * e1.z};
        const float u[3] = {d1.x, d1.y, d1.z}, v[3] = {d2.x, d2.y, d2.z};
        return point_triangle_inside(r, u, v);
    }
    return false;
}

__device__ void record(const IsCtx &c, uint32_t type, uint32_t e0, uint32_t e1,
                       const f3 *a, uint32_t na, const f3 *b, uint32_t nb, uint32_t scene) {
    const uint32_t slot = atomicAdd(c.ctr, 1u);
    if (c.ns) {
        atomicAdd(c.ctr + 2 + scene, 1u);
    }
    if (slot < c.cap) {
        ppf_isect_record r;
        r.type = type, r.elem0 = e0, r.elem1 = e1, r.num_verts0 = na, r.num_verts1 = nb;
        uint32_t k = 0;
        for (uint32_t i = 0; i < na; ++i) {
            r.positions[k++] = a[i].x, r.positions[k++] = a[i].y, r.positions[k++] = a[i].z;
        }
        for (uint32_t i = 0; i < nb; ++i) {
            r.positions[k++] = b[i].x, r.positions[k++] = b[i].y, r.positions[k++] = b[i].z;
        }
        for (; k < 15; ++k) {
            r.positions[k] = 0.0f;
        }
        c.rec[slot] = r;
    }
}

__global__ __launch_bounds__(kQT) void k_isect_edges(IsCtx c) {
    __shared__ uint32_t stk_all[kStack * kQT];
    const uint32_t leaf = xcd_logical_block() * kQT + threadIdx.x;
    if (leaf >= c.eb.n) {
        return;
    }
    uint32_t *stk = stk_all + threadIdx.x;
    const ppf_contact_mesh &m = c.m;
    const uint32_t ei = c.eb.node[2 * (size_t)leaf] - 1;
    const uint32_t a0 = m.edge[2 * (size_t)ei], a1 = m.edge[2 * (size_t)ei + 1];
    const f3 y[2] = {X(m.x, a0), X(m.x, a1)};
    QBox q = seg_box(y[0], y[1], 0.0f);
    q.scene = vertex_scene(m.vscene, m.nscene, a0);
    const bool efix = (m.eflag[ei] & 1u) != 0;
    const bool emass_pos = c.emass[ei] > 0.0f;
    // face-edge (contact.cu:2879-2946)
    traverse(c.fb, q, stk, c.ctr + 1, [&](uint32_t fi) {
        const uint32_t f0 = m.face[3 * (size_t)fi], f1 = m.face[3 * (size_t)fi + 1],
                       f2 = m.face[3 * (size_t)fi + 2];
        const bool either_dyn = (m.fflag[fi] & 1u) == 0 || !efix;
        const bool either_nonzero = c.fmass[fi] > 0.0f || emass_pos;
        const bool same = m.vbody[a0] != 0 && m.vbody[a0] == m.vbody[f0];
        const bool both = vcol(m, a0) && vcol(m, f0);
        if (!either_dyn || !either_nonzero || same || both) {
            return false;
        }
        if (a0 == f0 || a0 == f1 || a0 == f2 || a1 == f0 || a1 == f1 || a1 == f2) {
            return false;
        }
        const f3 fv[3] = {X(m.x, f0), X(m.x, f1), X(m.x, f2)};
        if (edge_triangle(y[0], y[1], fv[0], fv[1], fv[2])) {
            record(c, 0, fi, ei, fv, 3, y, 2, q.scene);
            return true;
        }
        return false;
    });
    // edge-edge (contact.cu:2762-2877)
    traverse(c.eb, q, stk, c.ctr + 1, [&](uint32_t ej) {
        if (!(ej < ei)) {
            return false;
        }
        const uint32_t b0 = m.edge[2 * (size_t)ej], b1 = m.edge[2 * (size_t)ej + 1];
        const bool either_dyn = !efix || (m.eflag[ej] & 1u) == 0;
        const bool either_nonzero = emass_pos || c.emass[ej] > 0.0f;
        const bool same = m.vbody[a0] != 0 && m.vbody[a0] == m.vbody[b0];
        const bool both = vcol(m, a0) && vcol(m, b0);
        if (!either_dyn || !either_nonzero || same || both) {
            return false;
        }
        const float off = m.e_off[ei] + m.e_off[ej];
        const f3 ev[2] = {X(m.x, b0), X(m.x, b1)};
        const bool shared = a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1;
        if (!shared) {
            float w[4];
            ee_coeff(y[0], y[1], ev[0], ev[1], w);
            const f3 p = w[0] * y[0] + w[1] * y[1];
            const f3 r = w[2] * ev[0] + w[3] * ev[1];
            const f3 e = p - r;
            if (dot3(e, e) < off * off) {
                record(c, 1, ei, ej, y, 2, ev, 2, q.scene);
                return true;
            }
            return false;
        }
        const uint32_t ea[2] = {a0, a1}, eb[2] = {b0, b1};
        const int ij[4][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}};
        for (int k = 0; k < 4; ++k) {
            const int i = ij[k][0], j = ij[k][1];
            if (ea[i] == eb[j]) {
                const f3 q0 = X(m.x, ea[i]), q1 = X(m.x, ea[1 - i]), q2 = X(m.x, eb[1 - j]);
                float c0[2], c1[2];
                pe_coeff_unclassified(q1, q0, q2, c0);
                pe_coeff_unclassified(q2, q0, q1, c1);
                const f3 e0 = (c0[0] * q0 + c0[1] * q2) - q1;
                const f3 e1 = (c1[0] * q0 + c1[1] * q1) - q2;
                if (fminf(dot3(e0, e0), dot3(e1, e1)) < off * off) {
                    record(c, 1, ei, ej, y, 2, ev, 2, q.scene);
                    return true;
                }
            }
        }
        return false;
    });
    // collision-mesh face vs edge (contact.cu:2948-2980, 3120-3130)
    if (c.sm.nv && emass_pos) {
        traverse(c.cfb, q, stk, c.ctr + 1, [&](uint32_t fi) {
            const uint32_t f0 = c.sm.face[3 * (size_t)fi], f1 = c.sm.face[3 * (size_t)fi + 1],
                           f2 = c.sm.face[3 * (size_t)fi + 2];
            const f3 fv[3] = {X(c.sm.x, f0), X(c.sm.x, f1), X(c.sm.x, f2)};
            if (edge_triangle(y[0], y[1], fv[0], fv[1], fv[2])) {
                record(c, 2, fi, ei, fv, 3, y, 2, q.scene);
                return true;
            }
            return false;
        });
    }
}

// point-point among surface vertices (contact.cu:2982-3054, 3135-3155)
__global__ __launch_bounds__(kQT) void k_isect_verts(IsCtx c) {
    __shared__ uint32_t stk_all[kStack * kQT];
    const uint32_t leaf = xcd_logical_block() * kQT + threadIdx.x;
    if (leaf >= c.vb.n) {
        return;
    }
    const ppf_contact_mesh &m = c.m;
    const uint32_t i = c.vb.node[2 * (size_t)leaf] - 1;
    if (i >= c.nsurf) {
        return;
    }
    uint32_t *stk = stk_all + threadIdx.x;
    const f3 p = X(m.x, i);
    QBox q = point_box(p, m.v_off[i]);
    q.scene = vertex_scene(m.vscene, m.nscene, i);
    traverse(c.vb, q, stk, c.ctr + 1, [&](uint32_t j) {
        if (!(j < i)) {
            return false;
        }
        const bool either_dyn = (m.vflag[i] & 1u) == 0 || (m.vflag[j] & 1u) == 0;
        const bool either_nonzero = c.vmass[i] > 0.0f || c.vmass[j] > 0.0f;
        const bool same = m.vbody[i] != 0 && m.vbody[i] == m.vbody[j];
        const bool both = vcol(m, i) && vcol(m, j);
        if (!either_dyn || !either_nonzero || same || both) {
            return false;
        }
        const float off = m.v_off[i] + m.v_off[j];
        const f3 qj = X(m.x, j);
        const f3 e = p - qj;
        if (dot3(e, e) < off * off) {
            record(c, 3, i, j, &p, 1, &qj, 1, q.scene);
            return true;
        }
        return false;
    });
}

uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

} // namespace

IntersectionCheck::~IntersectionCheck() {
    if (rec_) (void)dev_free(rec_);
    if (ctr_) (void)dev_free(ctr_);
    if (h_ctr_) (void)host_free(h_ctr_);
}

int IntersectionCheck::run(const BvhView &face, const BvhView &edge, const BvhView &vert,
                           const ppf_contact_mesh &m, const ppf_static_mesh *sm,
                           const float *vmass, const float *emass, const float *fmass,
                           uint32_t surface_vert_count, ppf_isect_record *records,
                           uint32_t cap, uint32_t *count, hipStream_t s,
                           uint32_t *scene_count) {
    const uint32_t ns = (m.vscene && m.nscene > 1) ? m.nscene : 0u;
    if (2 + ns > ctr_cap_) {
        if (ctr_) {
            PPF_HIP(dev_free(ctr_));
            ctr_ = nullptr;
        }
        if (h_ctr_) {
            PPF_HIP(host_free(h_ctr_));
            h_ctr_ = nullptr;
        }
        PPF_HIP(dev_alloc(&ctr_, (2 + ns) * sizeof(uint32_t)));
        PPF_HIP(host_alloc(reinterpret_cast<void **>(&h_ctr_), (2 + ns) * sizeof(uint32_t)));
        ctr_cap_ = 2 + ns;
    }
    const uint32_t need = cap < 256u ? 256u : cap;
    if (need > rec_cap_) {
        if (rec_) {
            PPF_HIP(dev_free(rec_));
            rec_ = nullptr;
        }
        PPF_HIP(dev_alloc(&rec_, (size_t)need * sizeof(ppf_isect_record)));
        rec_cap_ = need;
    }
    ppf_static_mesh none{};
    const ppf_static_mesh &S = sm ? *sm : none;
    if (S.nv && (cm_x_ != S.x || cm_nf_ != S.nf || cm_face_.view().n != S.nf)) {
        PrimSet sf;
        sf.n = S.nf, sf.arity = 3, sf.elem = S.face, sf.ghat = S.f_ghat, sf.offset = S.f_off;
        sf.vscene = S.vscene;
        sf.nscene = m.nscene;
        int rc = cm_face_.build(sf, S.x, nullptr, 0.0f, s);
        if (rc != PPF_OK) {
            return rc;
        }
        cm_x_ = S.x;
        cm_nf_ = S.nf;
    }
    PPF_HIP(hipMemsetAsync(ctr_, 0, (2 + ns) * sizeof(uint32_t), s));
    IsCtx c;
    c.m = m;
    c.sm = S;
    c.fb = face;
    c.eb = edge;
    c.vb = vert;
    c.cfb = S.nv ? cm_face_.view() : BvhView{};
    c.vmass = vmass;
    c.emass = emass;
    c.fmass = fmass;
    c.nsurf = surface_vert_count < m.nv ? surface_vert_count : m.nv;
    c.rec = rec_;
    c.cap = rec_cap_;
    c.ctr = ctr_;
    c.ns = ns;
    if (m.ne) {
        k_isect_edges<<<xcd_grid(cdiv(m.ne, kQT)), kQT, 0, s>>>(c);
        PPF_HIP(hipGetLastError());
    }
    if (c.nsurf) {
        k_isect_verts<<<xcd_grid(cdiv(m.nv, kQT)), kQT, 0, s>>>(c);
        PPF_HIP(hipGetLastError());
    }
    PPF_HIP(hipMemcpyAsync(h_ctr_, ctr_, (2 + ns) * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           s));
    PPF_HIP(hipStreamSynchronize(s));
    if (h_ctr_[1]) {
        set_error("intersection check: a BVH traversal exceeded its %d-entry stack", kStack);
        return PPF_ERR_CAPACITY;
    }
    *count = h_ctr_[0];
    for (uint32_t k = 0; scene_count && k < ns; ++k) {
        scene_count[k] = h_ctr_[2 + k];
    }
    const uint32_t keep = h_ctr_[0] < cap ? h_ctr_[0] : cap;
    if (records && keep) {
        PPF_HIP(hipMemcpy(records, rec_, (size_t)keep * sizeof(ppf_isect_record),
                          hipMemcpyDeviceToHost));
    }
    return PPF_OK;
}

} // namespace ppf
