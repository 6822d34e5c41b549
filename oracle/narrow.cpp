// narrow.cpp -- CPU oracle: LBVH build and the contact proximity narrow phase.
//
// TEST INFRASTRUCTURE ONLY (the checker of the HIP path, never shipped).
//
// Restates in fp32 (-ffp-contract=off):
//   * distance::point_edge_distance_coeff        contact/distance.hpp:14-26
//   * distance::point_triangle_distance_coeff    contact/distance.hpp:46-81
//   * distance::edge_edge_distance_coeff         contact/distance.hpp:112-192
//   * LBVH: centroids, scene bounds, 10-bit/axis Morton codes, stable sort,
//     Karras radix tree with index tie-break, leaf boxes inflated by
//     0.5*ghat + offset, bottom-up box union       lbvh/lbvh.cu:108-385, 623-857
//   * aabb::make / overlap / query (stack DFS)    contact/aabb.hpp:15-107
//   * the stage-0 acceptance rules of the four contact kinds
//       point-point  (index < vertex, multiplicity from incident edges/faces)
//                                                  contact/contact.cu:432-620
//       point-edge   (interior projection, multiplicity from edge faces)
//                                                  contact/contact.cu:622-715
//       point-face   (interior projection)         contact/contact.cu:717-785
//       edge-edge    (edge < index, no shared vertex, interior closest points)
//                                                  contact/contact.cu:787-870
//     with the shared filters either_dyn / same_pdrd_body / both_collider and
//     the query order of contact.cu:1622-1745 (queries in Morton order).
#include "oracle_bvh.hpp"
#include "oracle_distance.hpp"
#include "oracle_math.hpp"
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace {

using orc::V3;
using orc::v3;
using orc::Box;
using orc::Bvh;
using orc::build_bvh;
using orc::leaf_box;
using orc::make_box;
using orc::point_box;
using orc::query;
using orc::clamp_unit;
using orc::ee_coeff;
using orc::pe_coeff;
using orc::pt_coeff;

inline bool interior(const float *c, int n) {
    float mx = c[0], mn = c[0];
    for (int i = 1; i < n; ++i) {
        mx = std::max(mx, c[i]);
        mn = std::min(mn, c[i]);
    }
    return mx < 1.0f && mn > 0.0f;
}

} // namespace

extern "C" {

// Contact mesh for the narrow phase. Neighbor tables are CSR:
// v2e / v2f (edges / faces incident to a vertex), e2f (faces incident to an
// edge). vflag bit0 = pinned (fix_index != 0), bit1 = collider; eflag/fflag
// bit0 = fixed. vbody = PD rigid body id (0 = none).
struct oracle_contact_mesh {
    uint32_t nv, ne, nf;
    const float *x;
    const uint32_t *edge, *face;
    const uint32_t *v2e_off, *v2e, *v2f_off, *v2f, *e2f_off, *e2f;
    const float *v_ghat, *v_off, *e_ghat, *e_off, *f_ghat, *f_off;
    const uint32_t *vflag, *vbody, *eflag, *fflag;
};

// One accepted pair: kind 0 point-face, 1 point-edge, 2 point-point,
// 3 edge-edge, plus bit 8 set when friction is excluded (a neighbor element of
// the candidate contains the query vertex, contact.cu:498-531, 685-695);
// a = query element (vertex, or edge for kind 3), b = candidate element,
// count = multiplicity.
struct oracle_pair {
    uint32_t kind, a, b, count;
};

// Build the face (kind 0), edge (kind 1) or vertex (kind 2) BVH of a static
// mesh: writes 2*(2n-1) node words, (2n-1) boxes as 6 floats, and the root.
void oracle_bvh_build(const oracle_contact_mesh *m, int kind, uint32_t *nodes,
                      float *boxes, uint32_t *root) {
    std::vector<V3> x(m->nv);
    for (uint32_t i = 0; i < m->nv; ++i) {
        x[i] = v3(m->x[3 * i], m->x[3 * i + 1], m->x[3 * i + 2]);
    }
    uint32_t n = kind == 0 ? m->nf : (kind == 1 ? m->ne : m->nv);
    std::vector<V3> cent(n);
    std::vector<Box> pb(n);
    for (uint32_t i = 0; i < n; ++i) {
        if (kind == 0) {
            const uint32_t *f = m->face + 3 * i;
            V3 p[3] = {x[f[0]], x[f[1]], x[f[2]]};
            cent[i] = v3((p[0][0] + p[1][0] + p[2][0]) / 3.0f,
                         (p[0][1] + p[1][1] + p[2][1]) / 3.0f,
                         (p[0][2] + p[1][2] + p[2][2]) / 3.0f);
            pb[i] = leaf_box(p, 3, 0.5f * m->f_ghat[i] + m->f_off[i]);
        } else if (kind == 1) {
            const uint32_t *e = m->edge + 2 * i;
            V3 p[2] = {x[e[0]], x[e[1]]};
            cent[i] = v3((p[0][0] + p[1][0]) / 2.0f, (p[0][1] + p[1][1]) / 2.0f,
                         (p[0][2] + p[1][2]) / 2.0f);
            pb[i] = leaf_box(p, 2, 0.5f * m->e_ghat[i] + m->e_off[i]);
        } else {
            cent[i] = x[i];
            pb[i] = leaf_box(&x[i], 1, 0.5f * m->v_ghat[i] + m->v_off[i]);
        }
    }
    Bvh b;
    build_bvh(b, n, cent, pb);
    size_t nn = n ? 2 * (size_t)n - 1 : 0;
    for (size_t k = 0; k < nn; ++k) {
        nodes[2 * k] = b.node[2 * k];
        nodes[2 * k + 1] = b.node[2 * k + 1];
        for (int d = 0; d < 3; ++d) {
            boxes[6 * k + d] = b.box[k].mn[d];
            boxes[6 * k + 3 + d] = b.box[k].mx[d];
        }
    }
    *root = b.root;
}

// Stage-0 proximity pairs of all four kinds, in the reference's query order
// (vertex queries in vertex-BVH leaf order, PF then PE then PP; then edge
// queries in edge-BVH leaf order). Returns the pair count; writes at most
// `cap` pairs.
uint64_t oracle_proximity(const oracle_contact_mesh *m, oracle_pair *out,
                          uint64_t cap) {
    const uint32_t nv = m->nv, ne = m->ne, nf = m->nf;
    std::vector<V3> x(nv);
    for (uint32_t i = 0; i < nv; ++i) {
        x[i] = v3(m->x[3 * i], m->x[3 * i + 1], m->x[3 * i + 2]);
    }
    auto vdyn = [&](uint32_t v) { return (m->vflag[v] & 1u) == 0; };
    auto vcol = [&](uint32_t v) { return (m->vflag[v] & 2u) != 0; };
    Bvh fb, eb, vb;
    {
        std::vector<V3> c(nf);
        std::vector<Box> pb(nf);
        for (uint32_t i = 0; i < nf; ++i) {
            const uint32_t *f = m->face + 3 * i;
            V3 p[3] = {x[f[0]], x[f[1]], x[f[2]]};
            c[i] = v3((p[0][0] + p[1][0] + p[2][0]) / 3.0f,
                      (p[0][1] + p[1][1] + p[2][1]) / 3.0f,
                      (p[0][2] + p[1][2] + p[2][2]) / 3.0f);
            pb[i] = leaf_box(p, 3, 0.5f * m->f_ghat[i] + m->f_off[i]);
        }
        build_bvh(fb, nf, c, pb);
    }
    {
        std::vector<V3> c(ne);
        std::vector<Box> pb(ne);
        for (uint32_t i = 0; i < ne; ++i) {
            const uint32_t *e = m->edge + 2 * i;
            V3 p[2] = {x[e[0]], x[e[1]]};
            c[i] = v3((p[0][0] + p[1][0]) / 2.0f, (p[0][1] + p[1][1]) / 2.0f,
                      (p[0][2] + p[1][2]) / 2.0f);
            pb[i] = leaf_box(p, 2, 0.5f * m->e_ghat[i] + m->e_off[i]);
        }
        build_bvh(eb, ne, c, pb);
    }
    {
        std::vector<Box> pb(nv);
        for (uint32_t i = 0; i < nv; ++i) {
            pb[i] = leaf_box(&x[i], 1, 0.5f * m->v_ghat[i] + m->v_off[i]);
        }
        build_bvh(vb, nv, x, pb);
    }
    // Queries in fixed chunks, each chunk's pairs in query order; chunks are
    // concatenated in order, so the output is independent of the thread count.
    const uint32_t kChunk = 4096;
    const uint32_t ncv = (nv + kChunk - 1) / kChunk, nce = (ne + kChunk - 1) / kChunk;
    std::vector<std::vector<oracle_pair>> res(ncv + nce);
#pragma omp parallel for schedule(dynamic, 1)
    for (long ch = 0; ch < (long)ncv; ++ch) {
    std::vector<oracle_pair> &loc = res[ch];
    auto emit = [&](uint32_t kind, uint32_t a, uint32_t b, uint32_t count) {
        loc.push_back(oracle_pair{kind, a, b, count});
    };
    const uint32_t l0 = (uint32_t)ch * kChunk, l1 = std::min(nv, l0 + kChunk);
    for (uint32_t leaf = l0; leaf < l1; ++leaf) {
        const uint32_t i = vb.node[2 * leaf] - 1;
        const V3 &p = x[i];
        Box q = point_box(p, 0.5f * m->v_ghat[i] + m->v_off[i]);
        q.active = (m->vflag[i] & 4u) == 0; // collision window (contact.cu:1637)
        // point-face
        query(fb, q, [&](uint32_t fi) {
            const uint32_t *f = m->face + 3 * fi;
            bool either_dyn = vdyn(i) || (m->fflag[fi] & 1u) == 0;
            bool same_body = m->vbody[i] != 0 && m->vbody[i] == m->vbody[f[0]];
            bool both_col = vcol(i) && vcol(f[0]);
            if (f[0] == i || f[1] == i || f[2] == i || !either_dyn || same_body ||
                both_col) {
                return false;
            }
            float c[3];
            pt_coeff(p, x[f[0]], x[f[1]], x[f[2]], c);
            if (!interior(c, 3)) {
                return false;
            }
            V3 e = orc::add(orc::add(orc::scale(c[0], orc::sub(p, x[f[0]])),
                                     orc::scale(c[1], orc::sub(p, x[f[1]]))),
                            orc::scale(c[2], orc::sub(p, x[f[2]])));
            float off = m->v_off[i] + m->f_off[fi];
            float gh = 0.5f * (m->v_ghat[i] + m->f_ghat[fi]);
            float r = gh + off;
            if (orc::sqnorm(e) < r * r) {
                emit(0, i, fi, 1);
                return true;
            }
            return false;
        });
        // point-edge
        query(eb, q, [&](uint32_t ei) {
            const uint32_t *ed = m->edge + 2 * ei;
            bool either_dyn = vdyn(i) || (m->eflag[ei] & 1u) == 0;
            bool same_body = m->vbody[i] != 0 && m->vbody[i] == m->vbody[ed[0]];
            bool both_col = vcol(i) && vcol(ed[0]);
            if (ed[0] == i || ed[1] == i || !either_dyn || same_body || both_col) {
                return false;
            }
            float c[2];
            pe_coeff(p, x[ed[0]], x[ed[1]], c);
            if (!interior(c, 2)) {
                return false;
            }
            V3 e = orc::add(orc::scale(c[0], orc::sub(p, x[ed[0]])),
                            orc::scale(c[1], orc::sub(p, x[ed[1]])));
            float off = m->v_off[i] + m->e_off[ei];
            float gh = 0.5f * (m->v_ghat[i] + m->e_ghat[ei]);
            float r = gh + off;
            if (!(orc::sqnorm(e) < r * r)) {
                return false;
            }
            uint32_t count = 0, nofric = 0;
            uint32_t f0 = m->e2f_off[ei], f1 = m->e2f_off[ei + 1];
            if (f1 > f0) {
                for (uint32_t k = f0; k < f1; ++k) {
                    const uint32_t *f = m->face + 3 * m->e2f[k];
                    if (f[0] == i || f[1] == i || f[2] == i) {
                        ++count;
                        nofric = 1;
                    } else {
                        float cc[3];
                        pt_coeff(p, x[f[0]], x[f[1]], x[f[2]], cc);
                        if (!interior(cc, 3)) {
                            ++count;
                        }
                    }
                }
            } else {
                count = 1;
            }
            if (count) {
                emit(1 | (nofric << 8), i, ei, count);
                return true;
            }
            return false;
        });
        // point-point
        query(vb, q, [&](uint32_t j) {
            bool either_dyn = vdyn(i) || vdyn(j);
            bool same_body = m->vbody[i] != 0 && m->vbody[i] == m->vbody[j];
            bool both_col = vcol(i) && vcol(j);
            if (!(j < i) || !either_dyn || same_body || both_col) {
                return false;
            }
            V3 e = orc::sub(p, x[j]);
            float off = m->v_off[i] + m->v_off[j];
            float gh = 0.5f * (m->v_ghat[i] + m->v_ghat[j]);
            float r = gh + off;
            if (!(orc::sqnorm(e) < r * r)) {
                return false;
            }
            uint32_t count = 0, nofric = 0;
            uint32_t e0 = m->v2e_off[j], e1 = m->v2e_off[j + 1];
            if (e1 > e0) {
                for (uint32_t k = e0; k < e1; ++k) {
                    const uint32_t *ed = m->edge + 2 * m->v2e[k];
                    if (ed[0] == i || ed[1] == i) {
                        ++count;
                        nofric = 1;
                    } else {
                        float cc[2];
                        pe_coeff(p, x[ed[0]], x[ed[1]], cc);
                        if (!interior(cc, 2)) {
                            ++count;
                        }
                    }
                }
                for (uint32_t k = m->v2f_off[j]; k < m->v2f_off[j + 1]; ++k) {
                    const uint32_t *f = m->face + 3 * m->v2f[k];
                    if (f[0] == i || f[1] == i || f[2] == i) {
                        ++count;
                        nofric = 1;
                    } else {
                        float cc[3];
                        pt_coeff(p, x[f[0]], x[f[1]], x[f[2]], cc);
                        if (!interior(cc, 3)) {
                            ++count;
                        }
                    }
                }
            } else {
                count = 1;
            }
            if (count) {
                emit(2 | (nofric << 8), i, j, count);
                return true;
            }
            return false;
        });
    }
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (long ch = 0; ch < (long)nce; ++ch) {
    std::vector<oracle_pair> &loc = res[ncv + ch];
    auto emit = [&](uint32_t kind, uint32_t a, uint32_t b, uint32_t count) {
        loc.push_back(oracle_pair{kind, a, b, count});
    };
    const uint32_t l0 = (uint32_t)ch * kChunk, l1 = std::min(ne, l0 + kChunk);
    for (uint32_t leaf = l0; leaf < l1; ++leaf) {
        const uint32_t ei = eb.node[2 * leaf] - 1;
        const uint32_t *ea = m->edge + 2 * ei;
        V3 pts[2] = {x[ea[0]], x[ea[1]]};
        Box q = make_box(pts, 2, 0.5f * m->e_ghat[ei] + m->e_off[ei]);
        q.active = (m->eflag[ei] & 4u) == 0; // collision window (contact.cu:1728)
        query(eb, q, [&](uint32_t ej) {
            const uint32_t *eb2 = m->edge + 2 * ej;
            bool either_dyn = (m->eflag[ei] & 1u) == 0 || (m->eflag[ej] & 1u) == 0;
            bool same_body = m->vbody[ea[0]] != 0 && m->vbody[ea[0]] == m->vbody[eb2[0]];
            bool both_col = vcol(ea[0]) && vcol(eb2[0]);
            bool shared = ea[0] == eb2[0] || ea[0] == eb2[1] || ea[1] == eb2[0] ||
                          ea[1] == eb2[1];
            if (!(ei < ej) || shared || !either_dyn || same_body || both_col) {
                return false;
            }
            float c[4];
            ee_coeff(x[ea[0]], x[ea[1]], x[eb2[0]], x[eb2[1]], c);
            if (!interior(c, 4)) {
                return false;
            }
            V3 x0 = orc::add(orc::scale(c[0], x[ea[0]]), orc::scale(c[1], x[ea[1]]));
            V3 x1 = orc::add(orc::scale(c[2], x[eb2[0]]), orc::scale(c[3], x[eb2[1]]));
            V3 e = orc::sub(x0, x1);
            float off = m->e_off[ei] + m->e_off[ej];
            float gh = 0.5f * (m->e_ghat[ei] + m->e_ghat[ej]);
            float r = gh + off;
            if (orc::sqnorm(e) < r * r) {
                emit(3, ei, ej, 1);
                return true;
            }
            return false;
        });
    }
    }
    uint64_t total = 0;
    for (const auto &r : res) {
        for (const oracle_pair &q : r) {
            if (total < cap) {
                out[total] = q;
            }
            ++total;
        }
    }
    return total;
}

// Distance coefficients, exported for the pinning tests.
void oracle_pe_coeff(const float *p, const float *e0, const float *e1, float *w) {
    pe_coeff(v3(p[0], p[1], p[2]), v3(e0[0], e0[1], e0[2]),
             v3(e1[0], e1[1], e1[2]), w);
}
void oracle_pt_coeff(const float *p, const float *t0, const float *t1,
                     const float *t2, float *w) {
    pt_coeff(v3(p[0], p[1], p[2]), v3(t0[0], t0[1], t0[2]),
             v3(t1[0], t1[1], t1[2]), v3(t2[0], t2[1], t2[2]), w);
}
void oracle_ee_coeff(const float *a0, const float *a1, const float *b0,
                     const float *b1, float *w) {
    ee_coeff(v3(a0[0], a0[1], a0[2]), v3(a1[0], a1[1], a1[2]),
             v3(b0[0], b0[1], b0[2]), v3(b1[0], b1[1], b1[2]), w);
}
}

extern "C" {

// Collision mesh for the static pairs (prefix layout of ppf_static_mesh).
struct oracle_static_view_n {
    uint32_t nv, ne, nf;
    const float *x;
    const uint32_t *edge, *face;
    const float *v_ghat, *v_off, *e_ghat, *e_off, *f_ghat, *f_off;
};

// Collision-mesh proximity pairs (contact.cu:899-1152, 1789-1954): kind 4 per
// unpinned vertex of positive mass vs static faces, kind 5 per static vertex
// vs faces that are not fixed and have mass, kind 6 per edge that is not
// fixed and has mass vs static edges; candidates from boxes of margin
// 0.5 * ghat + offset on both sides, accepted when the squared distance is
// below (offset + ghat)^2. Queries in index order. Returns the pair count;
// writes at most `cap`.
uint64_t oracle_static_proximity(const oracle_contact_mesh *m, const oracle_static_view_n *sm,
                                 const float *vmass, const float *emass, const float *fmass,
                                 oracle_pair *out, uint64_t cap) {
    const uint32_t nv = m->nv, ne = m->ne, nf = m->nf;
    std::vector<V3> x(nv), sx(sm->nv);
    for (uint32_t i = 0; i < nv; ++i) x[i] = v3(m->x[3 * i], m->x[3 * i + 1], m->x[3 * i + 2]);
    for (uint32_t i = 0; i < sm->nv; ++i)
        sx[i] = v3(sm->x[3 * i], sm->x[3 * i + 1], sm->x[3 * i + 2]);
    auto tree = [](Bvh &t, uint32_t n, int k, const uint32_t *el, const float *gh,
                   const float *off, const std::vector<V3> &pos) {
        std::vector<V3> c(n);
        std::vector<Box> pb(n);
        for (uint32_t i = 0; i < n; ++i) {
            V3 p[3];
            for (int a = 0; a < k; ++a) p[a] = pos[el[(size_t)k * i + a]];
            c[i] = p[0];
            pb[i] = leaf_box(p, k, 0.5f * gh[i] + off[i]);
        }
        build_bvh(t, n, c, pb);
    };
    Bvh fb, sfb, seb;
    tree(fb, nf, 3, m->face, m->f_ghat, m->f_off, x);
    tree(sfb, sm->nf, 3, sm->face, sm->f_ghat, sm->f_off, sx);
    tree(seb, sm->ne, 2, sm->edge, sm->e_ghat, sm->e_off, sx);
    uint64_t n = 0;
    auto emit = [&](uint32_t kind, uint32_t a, uint32_t b) {
        if (n < cap) out[n] = oracle_pair{kind, a, b, 1};
        ++n;
    };
    const V3 zero = v3(0, 0, 0);
    for (uint32_t i = 0; i < nv; ++i) {
        if ((m->vflag[i] & 5u) != 0 || !(vmass[i] > 0.0f)) continue;
        const V3 p = x[i];
        query(sfb, point_box(p, 0.5f * m->v_ghat[i] + m->v_off[i]), [&](uint32_t fi) {
            const uint32_t *f = sm->face + 3 * (size_t)fi;
            const V3 t0 = orc::sub(sx[f[0]], p), t1 = orc::sub(sx[f[1]], p),
                     t2 = orc::sub(sx[f[2]], p);
            float c[3];
            orc::pt_coeff_unclassified(zero, t0, t1, t2, c);
            const V3 y = orc::add(orc::add(orc::scale(c[0], t0), orc::scale(c[1], t1)),
                                  orc::scale(c[2], t2));
            const float r = (m->v_off[i] + sm->f_off[fi]) + 0.5f * (m->v_ghat[i] + sm->f_ghat[fi]);
            if (orc::dot(y, y) < r * r) {
                emit(4, i, fi);
                return true;
            }
            return false;
        });
    }
    for (uint32_t sv = 0; sv < sm->nv; ++sv) {
        const V3 y = sx[sv];
        query(fb, point_box(y, 0.5f * sm->v_ghat[sv] + sm->v_off[sv]), [&](uint32_t fi) {
            if ((m->fflag[fi] & 1u) != 0 || !(fmass[fi] > 0.0f)) return false;
            const uint32_t *f = m->face + 3 * (size_t)fi;
            const V3 t0 = orc::sub(x[f[0]], y), t1 = orc::sub(x[f[1]], y),
                     t2 = orc::sub(x[f[2]], y);
            float c[3];
            orc::pt_coeff_unclassified(zero, t0, t1, t2, c);
            const V3 e = orc::add(orc::add(orc::scale(c[0], t0), orc::scale(c[1], t1)),
                                  orc::scale(c[2], t2));
            const float r = (m->f_off[fi] + sm->v_off[sv]) + 0.5f * (m->f_ghat[fi] + sm->v_ghat[sv]);
            if (orc::dot(e, e) < r * r) {
                emit(5, sv, fi);
                return true;
            }
            return false;
        });
    }
    for (uint32_t ei = 0; ei < ne; ++ei) {
        if ((m->eflag[ei] & 5u) != 0 || !(emass[ei] > 0.0f)) continue;
        const V3 p0 = x[m->edge[2 * ei]], p1 = x[m->edge[2 * ei + 1]];
        const V3 seg[2] = {p0, p1};
        query(seb, orc::make_box(seg, 2, 0.5f * m->e_ghat[ei] + m->e_off[ei]), [&](uint32_t sj) {
            const V3 q0 = sx[sm->edge[2 * sj]], q1 = sx[sm->edge[2 * sj + 1]];
            float c[4];
            orc::ee_coeff(p0, p1, q0, q1, c);
            const V3 e = orc::sub(orc::add(orc::scale(c[0], p0), orc::scale(c[1], p1)),
                                  orc::add(orc::scale(c[2], q0), orc::scale(c[3], q1)));
            const float r = (m->e_off[ei] + sm->e_off[sj]) + 0.5f * (m->e_ghat[ei] + sm->e_ghat[sj]);
            if (orc::dot(e, e) < r * r) {
                emit(6, ei, sj);
                return true;
            }
            return false;
        });
    }
    return n;
}
}
