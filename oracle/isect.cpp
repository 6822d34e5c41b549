// isect.cpp -- CPU oracle: the post-step intersection check.
//
// TEST INFRASTRUCTURE ONLY (the checker of the HIP path, never shipped).
//
// Restates in fp32 (-ffp-contract=off):
//   * ppf_isect::point_triangle_inside / edge_triangle_intersect, called with
//     coordinates relative to the triangle's first vertex as the device
//     wrapper does                                contact/intersect_core.hpp:33-81,
//                                                 contact/contact.cu:2742-2760
//   * EdgeEdgeIntersectTester (offset test, shared-vertex point-edge form)
//                                                 contact/contact.cu:2762-2877
//   * FaceEdgeIntersectTester                     contact/contact.cu:2879-2946
//   * CollisionMeshFaceEdgeIntersectTester        contact/contact.cu:2948-2980
//   * PointPointIntersectTester                   contact/contact.cu:2982-3054
//   * check_intersection (per-edge queries with the unpadded edge box, per
//     surface-vertex queries with the offset box)  contact/contact.cu:3056-3175
// Trees are the narrow-phase trees at the checked positions (margin
// 0.5*ghat + offset) and the collision-mesh face tree (same margin).
#include "oracle_bvh.hpp"
#include "oracle_distance.hpp"
#include <algorithm>
#include <cstdint>
#include <vector>

namespace {

using orc::Box;
using orc::Bvh;
using orc::V3;
using orc::v3;

bool point_triangle_inside(const float *p, const float *d1, const float *d2) {
    auto dot = [](const float *a, const float *b) {
        return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    };
    float a00 = dot(d1, d1), a01 = dot(d1, d2), a11 = dot(d2, d2);
    float b0 = dot(d1, p), b1 = dot(d2, p);
    float det = a00 * a11 - a01 * a01;
    if (det == 0.0f) {
        return false;
    }
    float w0 = (a11 * b0 - a01 * b1) / det;
    float w1 = (a00 * b1 - a01 * b0) / det;
    float w2 = 1.0f - w0 - w1;
    // Selection order of the reference's ternaries (matters only for NaN).
    float lo = w0 < w1 ? w0 : w1;
    lo = lo < w2 ? lo : w2;
    float hi = w0 > w1 ? w0 : w1;
    hi = hi > w2 ? hi : w2;
    return lo >= 0.0f && hi <= 1.0f;
}

// Segment (a0, a1) strictly crossing triangle (b0, b1, b2).
bool edge_triangle(const V3 &a0, const V3 &a1, const V3 &b0, const V3 &b1, const V3 &b2) {
    const V3 d1 = orc::sub(b1, b0), d2 = orc::sub(b2, b0);
    const V3 e0 = orc::sub(a0, b0), e1 = orc::sub(a1, b0);
    // The core subtracts the (zero) first vertex again: exact.
    const float n[3] = {d1[1] * d2[2] - d1[2] * d2[1], d1[2] * d2[0] - d1[0] * d2[2],
                        d1[0] * d2[1] - d1[1] * d2[0]};
    const float s1 = e0[0] * n[0] + e0[1] * n[1] + e0[2] * n[2];
    const float s2 = e1[0] * n[0] + e1[1] * n[1] + e1[2] * n[2];
    if (s1 * s2 < 0.0f) {
        const float t = s1 / (s1 - s2);
        const float r[3] = {(1.0f - t) * e0[0] + t * e1[0], (1.0f - t) * e0[1] + t * e1[1],
                            (1.0f - t) * e0[2] + t * e1[2]};
        return point_triangle_inside(r, d1.x, d2.x);
    }
    return false;
}

} // namespace

extern "C" {

// Same layouts as ccd.cpp / narrow.cpp.
struct oracle_isect_mesh {
    uint32_t nv, ne, nf;
    const float *x;
    const uint32_t *edge, *face;
    const uint32_t *v2e_off, *v2e, *v2f_off, *v2f, *e2f_off, *e2f;
    const float *v_ghat, *v_off, *e_ghat, *e_off, *f_ghat, *f_off;
    const uint32_t *vflag, *vbody, *eflag, *fflag;
};

struct oracle_isect_static {
    uint32_t nv, ne, nf;
    const float *x;
    const uint32_t *edge, *face;
    const float *v_ghat, *v_off, *e_ghat, *e_off, *f_ghat, *f_off;
};

// IntersectionRecord (data.hpp:741-748).
struct oracle_isect_record {
    uint32_t type, elem0, elem1, num_verts0, num_verts1;
    float positions[15];
};

// Number of intersecting pairs (0 = clear); the first `cap` records in the
// order found (edges ascending, then surface vertices ascending).
uint32_t oracle_intersection_check(const oracle_isect_mesh *m, const oracle_isect_static *sm,
                                   const float *vmass, const float *emass,
                                   const float *fmass, uint32_t surface_vert_count,
                                   oracle_isect_record *rec, uint32_t cap) {
    const uint32_t nv = m->nv, ne = m->ne, nf = m->nf;
    std::vector<V3> x(nv);
    for (uint32_t i = 0; i < nv; ++i) {
        x[i] = v3(m->x[3 * i], m->x[3 * i + 1], m->x[3 * i + 2]);
    }
    auto prim_tree = [&](Bvh &t, uint32_t n, int k, const uint32_t *el, const float *gh,
                         const float *off, const std::vector<V3> &pos) {
        std::vector<V3> c(n);
        std::vector<Box> pb(n);
        for (uint32_t i = 0; i < n; ++i) {
            V3 p[3];
            for (int a = 0; a < k; ++a) {
                p[a] = pos[el ? el[(size_t)k * i + a] : i];
            }
            c[i] = p[0];
            pb[i] = orc::leaf_box(p, k, 0.5f * gh[i] + off[i]);
        }
        orc::build_bvh(t, n, c, pb);
    };
    Bvh fb, eb, vb, cfb;
    prim_tree(fb, nf, 3, m->face, m->f_ghat, m->f_off, x);
    prim_tree(eb, ne, 2, m->edge, m->e_ghat, m->e_off, x);
    prim_tree(vb, nv, 1, nullptr, m->v_ghat, m->v_off, x);
    std::vector<V3> sx;
    if (sm && sm->nv) {
        sx.resize(sm->nv);
        for (uint32_t i = 0; i < sm->nv; ++i) {
            sx[i] = v3(sm->x[3 * i], sm->x[3 * i + 1], sm->x[3 * i + 2]);
        }
        prim_tree(cfb, sm->nf, 3, sm->face, sm->f_ghat, sm->f_off, sx);
    }
    uint32_t count = 0;
    auto record = [&](uint32_t type, uint32_t e0, uint32_t e1, const V3 *a, uint32_t na,
                      const V3 *b, uint32_t nb) {
        if (count < cap) {
            oracle_isect_record &r = rec[count];
            r = oracle_isect_record{};
            r.type = type, r.elem0 = e0, r.elem1 = e1, r.num_verts0 = na, r.num_verts1 = nb;
            uint32_t k = 0;
            for (uint32_t i = 0; i < na; ++i)
                for (int d = 0; d < 3; ++d) r.positions[k++] = a[i][d];
            for (uint32_t i = 0; i < nb; ++i)
                for (int d = 0; d < 3; ++d) r.positions[k++] = b[i][d];
        }
        ++count;
    };
    auto vcol = [&](uint32_t v) { return (m->vflag[v] & 2u) != 0; };
    for (uint32_t ei = 0; ei < ne; ++ei) {
        const uint32_t a0 = m->edge[2 * ei], a1 = m->edge[2 * ei + 1];
        const V3 y[2] = {x[a0], x[a1]};
        const Box q = orc::make_box(y, 2, 0.0f);
        // face-edge
        orc::query(fb, q, [&](uint32_t fi) {
            const uint32_t *f = m->face + 3 * fi;
            bool either_dyn = (m->fflag[fi] & 1u) == 0 || (m->eflag[ei] & 1u) == 0;
            bool either_nonzero = fmass[fi] > 0.0f || emass[ei] > 0.0f;
            bool same = m->vbody[a0] != 0 && m->vbody[a0] == m->vbody[f[0]];
            bool both = vcol(a0) && vcol(f[0]);
            if (!either_dyn || !either_nonzero || same || both) {
                return false;
            }
            if (a0 == f[0] || a0 == f[1] || a0 == f[2] || a1 == f[0] || a1 == f[1] ||
                a1 == f[2]) {
                return false;
            }
            const V3 fv[3] = {x[f[0]], x[f[1]], x[f[2]]};
            if (edge_triangle(y[0], y[1], fv[0], fv[1], fv[2])) {
                record(0, fi, ei, fv, 3, y, 2);
                return true;
            }
            return false;
        });
        // edge-edge
        orc::query(eb, q, [&](uint32_t ej) {
            if (!(ej < ei)) {
                return false;
            }
            const uint32_t b0 = m->edge[2 * ej], b1 = m->edge[2 * ej + 1];
            bool either_dyn = (m->eflag[ei] & 1u) == 0 || (m->eflag[ej] & 1u) == 0;
            bool either_nonzero = emass[ei] > 0.0f || emass[ej] > 0.0f;
            bool same = m->vbody[a0] != 0 && m->vbody[a0] == m->vbody[b0];
            bool both = vcol(a0) && vcol(b0);
            if (!either_dyn || !either_nonzero || same || both) {
                return false;
            }
            const float off = m->e_off[ei] + m->e_off[ej];
            const V3 ev1[2] = {x[b0], x[b1]};
            const bool shared = a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1;
            if (!shared) {
                float c[4];
                orc::ee_coeff(y[0], y[1], ev1[0], ev1[1], c);
                V3 p = orc::add(orc::scale(c[0], y[0]), orc::scale(c[1], y[1]));
                V3 r = orc::add(orc::scale(c[2], ev1[0]), orc::scale(c[3], ev1[1]));
                V3 e = orc::sub(p, r);
                if (orc::dot(e, e) < off * off) {
                    record(1, ei, ej, y, 2, ev1, 2);
                    return true;
                }
                return false;
            }
            const uint32_t ea[2] = {a0, a1}, eb2[2] = {b0, b1};
            const int ij[4][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}};
            for (int k = 0; k < 4; ++k) {
                const int i = ij[k][0], j = ij[k][1];
                if (ea[i] == eb2[j]) {
                    const V3 q0 = x[ea[i]], q1 = x[ea[1 - i]], q2 = x[eb2[1 - j]];
                    float c0[2], c1[2];
                    orc::pe_coeff_unclassified(q1, q0, q2, c0);
                    orc::pe_coeff_unclassified(q2, q0, q1, c1);
                    V3 e0 = orc::sub(orc::add(orc::scale(c0[0], q0), orc::scale(c0[1], q2)), q1);
                    V3 e1 = orc::sub(orc::add(orc::scale(c1[0], q0), orc::scale(c1[1], q1)), q2);
                    if (std::min(orc::dot(e0, e0), orc::dot(e1, e1)) < off * off) {
                        record(1, ei, ej, y, 2, ev1, 2);
                        return true;
                    }
                }
            }
            return false;
        });
        // collision-mesh face vs edge
        if (sm && sm->nv && emass[ei] > 0.0f) {
            orc::query(cfb, q, [&](uint32_t fi) {
                const uint32_t *f = sm->face + 3 * fi;
                const V3 fv[3] = {sx[f[0]], sx[f[1]], sx[f[2]]};
                if (edge_triangle(y[0], y[1], fv[0], fv[1], fv[2])) {
                    record(2, fi, ei, fv, 3, y, 2);
                    return true;
                }
                return false;
            });
        }
    }
    const uint32_t ns = std::min(surface_vert_count, nv);
    for (uint32_t i = 0; i < ns; ++i) {
        const Box q = orc::point_box(x[i], m->v_off[i]);
        orc::query(vb, q, [&](uint32_t j) {
            if (!(j < i)) {
                return false;
            }
            bool either_dyn = (m->vflag[i] & 1u) == 0 || (m->vflag[j] & 1u) == 0;
            bool either_nonzero = vmass[i] > 0.0f || vmass[j] > 0.0f;
            bool same = m->vbody[i] != 0 && m->vbody[i] == m->vbody[j];
            bool both = vcol(i) && vcol(j);
            if (!either_dyn || !either_nonzero || same || both) {
                return false;
            }
            const float off = m->v_off[i] + m->v_off[j];
            V3 e = orc::sub(x[i], x[j]);
            if (orc::dot(e, e) < off * off) {
                record(3, i, j, &x[i], 1, &x[j], 1);
                return true;
            }
            return false;
        });
    }
    return count;
}
}
