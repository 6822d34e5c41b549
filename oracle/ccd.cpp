// ccd.cpp -- CPU oracle: additive CCD (ACCD) and the contact line search.
//
// TEST INFRASTRUCTURE ONLY (the checker of the HIP path, never shipped).
//
// Restates in fp32 (-ffp-contract=off):
//   * accd::centerize / rescale / max_relative_u / directional_advance /
//     ccd_helper / park_floor                   contact/accd.hpp:64-471
//   * the squared-distance functors (point-triangle and point-edge through the
//     unclassified closest-point coefficients)  contact/accd.hpp:473-545,
//                                               contact/distance.hpp:223-299
//   * point_triangle_ccd, point_edge_ccd, point_point_ccd, edge_edge_ccd
//                                               contact/accd.hpp:547-696
//   * the line search over the dynamic mesh (point-face with the own-face
//     point-edge case, point-point, edge-edge with the shared-vertex
//     point-edge pair) and a static collision mesh (M2C, C2M, CM-EE),
//     vertex-face minimum seeding the edge-edge sweep, result / max_t
//                                               contact/contact.cu:1956-2316,
//                                               2472-2740
// The candidate pairs come from swept boxes (the same conservative boxes as
// the reference); the minimum does not depend on how candidates are found.
#include "oracle_bvh.hpp"
#include "oracle_distance.hpp"
#include "oracle_math.hpp"
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

namespace {

using orc::V3;
using orc::v3;
using orc::clamp_unit;
using orc::ee_coeff;
using orc::pe_coeff;
using orc::pe_coeff_unclassified;
using orc::pt_coeff;
using orc::pt_coeff_unclassified;

const float kFltEps = 1.19209290e-7f;

// ---- ACCD -----------------------------------------------------------------

// Column-major 3 x C frame.
template <int C> struct Frame {
    V3 c[C];
};

template <int C> void centerize(Frame<C> &x) {
    const float s = (float)(1.0 / C);
    V3 mov = v3(0, 0, 0);
    for (int k = 0; k < C; ++k) {
        mov = orc::add(mov, orc::scale(s, x.c[k]));
    }
    for (int k = 0; k < C; ++k) {
        x.c[k] = orc::sub(x.c[k], mov);
    }
}

template <int C> float rescale(Frame<C> &x, Frame<C> &dx, float max_t, bool &coincident) {
    float m0 = 0.0f, m1 = 0.0f;
    bool first = true;
    for (int k = 0; k < C; ++k) {
        for (int r = 0; r < 3; ++r) {
            float a = std::fabs(x.c[k][r]);
            float b = std::fabs(x.c[k][r] + max_t * dx.c[k][r]);
            m0 = first ? a : std::max(m0, a);
            m1 = first ? b : std::max(m1, b);
            first = false;
        }
    }
    float me = std::max(m0, m1);
    float scale = 0.99f / me;
    coincident = false;
    if (!(scale > 0.0f) || !std::isfinite(scale)) {
        coincident = true;
        return 0.0f;
    }
    float scaled = 1.0f;
    const float s = 8.0f;
    while (true) {
        if (scale > s) {
            for (int k = 0; k < C; ++k) {
                x.c[k] = orc::scale(s, x.c[k]);
                dx.c[k] = orc::scale(s, dx.c[k]);
            }
            scaled *= s;
            scale /= s;
        } else {
            for (int k = 0; k < C; ++k) {
                x.c[k] = orc::scale(scale, x.c[k]);
                dx.c[k] = orc::scale(scale, dx.c[k]);
            }
            scaled *= scale;
            break;
        }
    }
    return scaled;
}

template <int C, int SPLIT> float max_relative_u(const Frame<C> &u) {
    float mu = 0.0f;
    for (int i = 0; i < SPLIT; ++i) {
        for (int j = SPLIT; j < C; ++j) {
            mu = std::max(mu, orc::sqnorm(orc::sub(u.c[i], u.c[j])));
        }
    }
    return std::sqrt(mu);
}

inline V3 vabs(const V3 &a) {
    return v3(std::fabs(a[0]), std::fabs(a[1]), std::fabs(a[2]));
}

template <int C, int SPLIT>
float directional_advance(const Frame<C> &x, const Frame<C> &dx, const V3 &w,
                          float w_norm, float dip, float park, float max_t) {
    const float round_slack = 2.0f * kFltEps;
    const float coord_slack = 6.2e-7f;
    const V3 wa = vabs(w);
    const float dip_w = dip * w_norm;
    const float park_w = park * w_norm;
    float step = std::numeric_limits<float>::max();
    for (int i = 0; i < SPLIT; ++i) {
        for (int j = SPLIT; j < C; ++j) {
            const V3 gap = orc::sub(x.c[j], x.c[i]);
            const V3 rel = orc::sub(dx.c[i], dx.c[j]);
            const float err =
                round_slack * (orc::dot(wa, vabs(gap)) + max_t * orc::dot(wa, vabs(rel)) +
                               park_w) +
                coord_slack * w_norm;
            const float projected = orc::dot(w, gap) - err;
            if (!(projected - dip_w > 0.0f)) {
                return 0.0f;
            }
            const float rate = orc::dot(w, rel);
            if (rate > 0.0f) {
                step = std::fmin(step, std::fmax(0.0f, projected - park_w) / rate);
            }
        }
    }
    return step;
}

// Squared distance of the frame's two primitives and the closest-point
// direction (from the first primitive to the second).
template <int C, int SPLIT> float sqdist(const Frame<C> &x, V3 &dir) {
    if (C == 4 && SPLIT == 1) { // point-triangle
        float c[3];
        pt_coeff_unclassified(x.c[0], x.c[1], x.c[2], x.c[3], c);
        V3 y = orc::add(orc::add(orc::scale(c[0], orc::sub(x.c[1], x.c[0])),
                                 orc::scale(c[1], orc::sub(x.c[2], x.c[0]))),
                        orc::scale(c[2], orc::sub(x.c[3], x.c[0])));
        dir = y;
    } else if (C == 4) { // edge-edge
        float c[4];
        ee_coeff(x.c[0], x.c[1], x.c[2], x.c[3], c);
        V3 a = orc::add(orc::scale(c[0], x.c[0]), orc::scale(c[1], x.c[1]));
        V3 b = orc::add(orc::scale(c[2], x.c[2]), orc::scale(c[3], x.c[3]));
        dir = orc::sub(b, a);
    } else if (C == 3) { // point-edge
        float c[2];
        pe_coeff_unclassified(x.c[0], x.c[1], x.c[2], c);
        V3 q = orc::add(orc::scale(c[0], x.c[1]), orc::scale(c[1], x.c[2]));
        dir = orc::sub(q, x.c[0]);
    } else { // point-point
        dir = orc::sub(x.c[1], x.c[0]);
    }
    return orc::dot(dir, dir);
}

template <int C> Frame<C> advance(const Frame<C> &x0, const Frame<C> &dx, float t) {
    Frame<C> x;
    for (int k = 0; k < C; ++k) {
        x.c[k] = orc::add(x0.c[k], orc::scale(t, dx.c[k]));
    }
    return x;
}

struct CcdOut {
    bool overlap = false;
};

template <int C, int SPLIT>
float ccd_helper(const Frame<C> &x0, const Frame<C> &dx, float u_max, float offset,
                 float eps, float floor_gap, float max_t, CcdOut &out) {
    Frame<C> x = x0;
    V3 w;
    float lower_t = 0.0f, toi = 0.0f;
    float d2 = sqdist<C, SPLIT>(x0, w);
    const float target_squared = offset * offset;
    if (!(d2 > target_squared)) {
        out.overlap = true;
        return 0.0f;
    }
    const float clearance = std::sqrt(d2) - offset;
    const float floor_clear = std::fmax(2.0f * eps, floor_gap);
    const float park = offset + std::fmin(floor_clear,
                                          std::fmax(clearance - 0.5f * eps,
                                                    0.5f * clearance));
    const float overshoot = std::fmax(0.0f, (park - offset) - 0.5f * floor_clear);
    const float dip = park - overshoot;
    const float park_squared = std::fmin(park * park, std::nextafter(d2, 0.0f));
    const float inv_u_max = 1.0f / u_max;
    const unsigned max_probes = 4096u;
    unsigned probes = 0;
    while (true) {
        if (!(d2 > park_squared)) {
            break;
        }
        lower_t = toi;
        const float dist = std::sqrt(d2);
        float step = (dist - dip) * inv_u_max;
        if (toi + step <= max_t) {
            step = std::fmax(step, directional_advance<C, SPLIT>(
                                       x, dx, w, dist * (1.0f + 2.0f * kFltEps), dip,
                                       park, max_t));
        }
        toi += step;
        if (toi > max_t) {
            return max_t;
        } else if (++probes > max_probes) {
            return lower_t;
        }
        x = advance(x0, dx, toi);
        d2 = sqdist<C, SPLIT>(x, w);
    }
    float upper_t = toi;
    float window = upper_t - lower_t;
    while (true) {
        toi = 0.5f * (upper_t + lower_t);
        x = advance(x0, dx, toi);
        V3 unused;
        d2 = sqdist<C, SPLIT>(x, unused);
        if (d2 > park_squared) {
            lower_t = toi;
        } else {
            upper_t = toi;
        }
        const float nw = upper_t - lower_t;
        if (nw == window) {
            break;
        }
        window = nw;
    }
    if (!(lower_t > 0.0f)) {
        out.overlap = true;
        return 0.0f;
    }
    return lower_t;
}

// A CCD kind over C columns: x0 / x1 of every column, SPLIT columns on the
// first primitive.
template <int C, int SPLIT>
float ccd(const V3 *p0, const V3 *p1, float offset, float ghat, float max_t,
          float ccd_eps, CcdOut &out) {
    Frame<C> x0, dx;
    for (int k = 0; k < C; ++k) {
        x0.c[k] = p0[k];
        dx.c[k] = orc::sub(p1[k], p0[k]);
    }
    centerize(x0);
    centerize(dx);
    bool coincident = false;
    const float scale = rescale(x0, dx, max_t, coincident);
    if (coincident) {
        out.overlap = true;
    }
    if (!(scale > 0.0f)) {
        return 0.0f;
    }
    const float u_max = max_relative_u<C, SPLIT>(dx);
    if (u_max) {
        return ccd_helper<C, SPLIT>(x0, dx, u_max, scale * offset, scale * ccd_eps,
                                    scale * (1e-2f * ghat), max_t, out);
    }
    return max_t;
}

} // namespace

extern "C" {

// Single-pair CCD, exported for the pinning tests. kind: 0 point-triangle
// (p, t0, t1, t2), 1 point-edge (p, e0, e1), 2 point-point, 3 edge-edge
// (a0, a1, b0, b1). x0 / x1: 3 * columns floats. Returns the toi; *overlap
// receives the reference's overlap flag.
float oracle_ccd(int kind, const float *x0, const float *x1, float offset,
                 float ghat, float max_t, float ccd_eps, int *overlap) {
    V3 a[4], b[4];
    const int cols = kind == 0 || kind == 3 ? 4 : (kind == 1 ? 3 : 2);
    for (int k = 0; k < cols; ++k) {
        a[k] = v3(x0[3 * k], x0[3 * k + 1], x0[3 * k + 2]);
        b[k] = v3(x1[3 * k], x1[3 * k + 1], x1[3 * k + 2]);
    }
    CcdOut o;
    float r;
    if (kind == 0) {
        r = ccd<4, 1>(a, b, offset, ghat, max_t, ccd_eps, o);
    } else if (kind == 1) {
        r = ccd<3, 1>(a, b, offset, ghat, max_t, ccd_eps, o);
    } else if (kind == 2) {
        r = ccd<2, 1>(a, b, offset, ghat, max_t, ccd_eps, o);
    } else {
        r = ccd<4, 2>(a, b, offset, ghat, max_t, ccd_eps, o);
    }
    *overlap = o.overlap ? 1 : 0;
    return r;
}

// Dynamic mesh for the line search: the same layout as oracle_contact_mesh
// (x = the step's start positions x0), plus the target positions x1 and the
// element masses the collision-mesh functors test.
struct oracle_ls_mesh {
    uint32_t nv, ne, nf;
    const float *x;
    const uint32_t *edge, *face;
    const uint32_t *v2e_off, *v2e, *v2f_off, *v2f, *e2f_off, *e2f;
    const float *v_ghat, *v_off, *e_ghat, *e_off, *f_ghat, *f_off;
    const uint32_t *vflag, *vbody, *eflag, *fflag;
};

// Static collision mesh (data.constraint.mesh): positions, elements and the
// per-element contact gap / offset. nv == 0 disables it.
struct oracle_static_mesh {
    uint32_t nv, ne, nf;
    const float *x;
    const uint32_t *edge, *face;
    const float *v_ghat, *v_off, *e_ghat, *e_off, *f_ghat, *f_off;
};

struct oracle_ls_params {
    const float *x1;     // target positions (3 nv)
    const float *vmass, *emass, *fmass; // VertexProp / EdgeProp / FaceProp mass
    float max_t;         // ParamSet.line_search_max_t
    float ccd_eps;       // ParamSet.ccd_eps
};

// Largest certified step fraction in [0, 1] (contact::line_search returns
// min toi / max_t); *overlap set when any pair started inside its offset.
float oracle_line_search(const oracle_ls_mesh *m, const oracle_static_mesh *sm,
                         const oracle_ls_params *lp, int *overlap) {
    using orc::Box;
    using orc::Bvh;
    const uint32_t nv = m->nv, ne = m->ne, nf = m->nf;
    const float max_t = lp->max_t;
    std::vector<V3> x0(nv), x1(nv), xs(nv);
    for (uint32_t i = 0; i < nv; ++i) {
        x0[i] = v3(m->x[3 * i], m->x[3 * i + 1], m->x[3 * i + 2]);
        x1[i] = v3(lp->x1[3 * i], lp->x1[3 * i + 1], lp->x1[3 * i + 2]);
        // extrapolated sweep end, lbvh.cu:644-646
        xs[i] = orc::add(orc::scale(max_t, orc::sub(x1[i], x0[i])), x0[i]);
    }
    auto vdyn = [&](uint32_t v) { return (m->vflag[v] & 1u) == 0; };
    auto vcol = [&](uint32_t v) { return (m->vflag[v] & 2u) != 0; };
    auto swept_box = [&](const uint32_t *idx, int k, float margin) {
        V3 pts[6];
        for (int a = 0; a < k; ++a) {
            pts[a] = x0[idx[a]];
            pts[k + a] = xs[idx[a]];
        }
        return orc::leaf_box(pts, 2 * k, margin);
    };
    Bvh fb, eb, vb;
    {
        std::vector<V3> c(nf);
        std::vector<Box> pb(nf);
        for (uint32_t i = 0; i < nf; ++i) {
            const uint32_t *f = m->face + 3 * i;
            c[i] = x0[f[0]];
            pb[i] = swept_box(f, 3, 0.5f * m->f_ghat[i] + m->f_off[i]);
        }
        orc::build_bvh(fb, nf, c, pb);
        std::vector<V3> ce(ne);
        std::vector<Box> pe(ne);
        for (uint32_t i = 0; i < ne; ++i) {
            const uint32_t *e = m->edge + 2 * i;
            ce[i] = x0[e[0]];
            pe[i] = swept_box(e, 2, 0.5f * m->e_ghat[i] + m->e_off[i]);
        }
        orc::build_bvh(eb, ne, ce, pe);
        std::vector<Box> pv(nv);
        for (uint32_t i = 0; i < nv; ++i) {
            pv[i] = swept_box(&i, 1, 0.5f * m->v_ghat[i] + m->v_off[i]);
        }
        orc::build_bvh(vb, nv, x0, pv);
    }
    Bvh sfb, seb;
    std::vector<V3> sx(sm ? sm->nv : 0);
    if (sm && sm->nv) {
        for (uint32_t i = 0; i < sm->nv; ++i) {
            sx[i] = v3(sm->x[3 * i], sm->x[3 * i + 1], sm->x[3 * i + 2]);
        }
        std::vector<V3> c(sm->nf);
        std::vector<Box> pb(sm->nf);
        for (uint32_t i = 0; i < sm->nf; ++i) {
            const uint32_t *f = sm->face + 3 * i;
            V3 p[3] = {sx[f[0]], sx[f[1]], sx[f[2]]};
            c[i] = p[0];
            pb[i] = orc::leaf_box(p, 3, 0.5f * sm->f_ghat[i] + sm->f_off[i]);
        }
        orc::build_bvh(sfb, sm->nf, c, pb);
        std::vector<V3> ce(sm->ne);
        std::vector<Box> pe(sm->ne);
        for (uint32_t i = 0; i < sm->ne; ++i) {
            const uint32_t *e = sm->edge + 2 * i;
            V3 p[2] = {sx[e[0]], sx[e[1]]};
            ce[i] = p[0];
            pe[i] = orc::leaf_box(p, 2, 0.5f * sm->e_ghat[i] + sm->e_off[i]);
        }
        orc::build_bvh(seb, sm->ne, ce, pe);
    }
    bool ovl = false;
    auto run = [&](int kind, const V3 *a, const V3 *b, float off, float gh) {
        CcdOut o;
        float r;
        if (kind == 0) {
            r = ccd<4, 1>(a, b, off, gh, max_t, lp->ccd_eps, o);
        } else if (kind == 1) {
            r = ccd<3, 1>(a, b, off, gh, max_t, lp->ccd_eps, o);
        } else if (kind == 2) {
            r = ccd<2, 1>(a, b, off, gh, max_t, lp->ccd_eps, o);
        } else {
            r = ccd<4, 2>(a, b, off, gh, max_t, lp->ccd_eps, o);
        }
        if (o.overlap) {
            ovl = true;
        }
        return r;
    };
    float t_vf = max_t;
    for (uint32_t i = 0; i < nv; ++i) {
        const float ext = 0.5f * m->v_ghat[i] + m->v_off[i];
        V3 seg[2] = {x0[i], xs[i]};
        Box q = orc::make_box(seg, 2, ext);
        q.active = (m->vflag[i] & 4u) == 0; // collision window (contact.cu:2563)
        // point-face (contact.cu:2110-2164)
        orc::query(fb, q, [&](uint32_t fi) {
            const uint32_t *f = m->face + 3 * fi;
            bool either_dyn = vdyn(i) || (m->fflag[fi] & 1u) == 0;
            bool same_body = m->vbody[i] != 0 && m->vbody[i] == m->vbody[f[0]];
            bool both_col = vcol(i) && vcol(f[0]);
            if (!either_dyn || same_body || both_col) {
                return false;
            }
            int dup = -1;
            for (int k = 0; k < 3; ++k) {
                if (f[k] == i) {
                    dup = k;
                    break;
                }
            }
            float r;
            if (dup < 0) {
                V3 a[4] = {x0[i], x0[f[0]], x0[f[1]], x0[f[2]]};
                V3 b[4] = {x1[i], x1[f[0]], x1[f[1]], x1[f[2]]};
                r = run(0, a, b, m->v_off[i] + m->f_off[fi],
                        0.5f * (m->v_ghat[i] + m->f_ghat[fi]));
            } else {
                const uint32_t p = f[dup], j = f[(dup + 1) % 3], k = f[(dup + 2) % 3];
                V3 a[3] = {x0[p], x0[j], x0[k]};
                V3 b[3] = {x1[p], x1[j], x1[k]};
                r = run(1, a, b, 2.0f * m->v_off[i], m->v_ghat[i]);
            }
            if (r < max_t) {
                t_vf = std::fmin(t_vf, r);
                return true;
            }
            return false;
        });
        // point-point (contact.cu:2266-2316)
        orc::query(vb, q, [&](uint32_t j) {
            bool either_dyn = vdyn(i) || vdyn(j);
            bool same_body = m->vbody[i] != 0 && m->vbody[i] == m->vbody[j];
            bool both_col = vcol(i) && vcol(j);
            if (!(j < i) || !either_dyn || same_body || both_col) {
                return false;
            }
            V3 a[2] = {x0[i], x0[j]}, b[2] = {x1[i], x1[j]};
            float r = run(2, a, b, m->v_off[i] + m->v_off[j],
                          0.5f * (m->v_ghat[i] + m->v_ghat[j]));
            if (r < max_t) {
                t_vf = std::fmin(t_vf, r);
                return true;
            }
            return false;
        });
        // dynamic vertex vs static face (contact.cu:1956-1994)
        if (sm && sm->nf && (m->vflag[i] & 1u) == 0) {
            orc::query(sfb, q, [&](uint32_t fi) {
                if (lp->vmass[i] == 0.0f) {
                    return false;
                }
                const uint32_t *f = sm->face + 3 * fi;
                V3 a[4] = {x0[i], sx[f[0]], sx[f[1]], sx[f[2]]};
                V3 b[4] = {x1[i], sx[f[0]], sx[f[1]], sx[f[2]]};
                float r = run(0, a, b, m->v_off[i] + sm->f_off[fi],
                              0.5f * (m->v_ghat[i] + sm->f_ghat[fi]));
                if (r < max_t) {
                    t_vf = std::fmin(t_vf, r);
                    return true;
                }
                return false;
            });
        }
    }
    // static vertex vs dynamic face (contact.cu:1996-2034)
    if (sm) {
        for (uint32_t i = 0; i < sm->nv; ++i) {
            const Box q = orc::point_box(sx[i], 0.5f * sm->v_ghat[i] + sm->v_off[i]);
            orc::query(fb, q, [&](uint32_t fi) {
                if ((m->fflag[fi] & 1u) != 0 || !(lp->fmass[fi] > 0.0f)) {
                    return false;
                }
                const uint32_t *f = m->face + 3 * fi;
                V3 a[4] = {sx[i], x0[f[0]], x0[f[1]], x0[f[2]]};
                V3 b[4] = {sx[i], x1[f[0]], x1[f[1]], x1[f[2]]};
                float r = run(0, a, b, m->f_off[fi] + sm->v_off[i],
                              0.5f * (m->f_ghat[fi] + sm->v_ghat[i]));
                if (r < max_t) {
                    t_vf = std::fmin(t_vf, r);
                    return true;
                }
                return false;
            });
        }
    }
    const float seed = t_vf;
    float t_ee = max_t;
    for (uint32_t ei = 0; ei < ne; ++ei) {
        const uint32_t *ea = m->edge + 2 * ei;
        const float ext = 0.5f * m->e_ghat[ei] + m->e_off[ei];
        V3 s0[2] = {x0[ea[0]], x0[ea[1]]};
        V3 s1[2] = {orc::add(orc::scale(seed, orc::sub(x1[ea[0]], x0[ea[0]])), x0[ea[0]]),
                    orc::add(orc::scale(seed, orc::sub(x1[ea[1]], x0[ea[1]])), x0[ea[1]])};
        Box b0 = orc::make_box(s0, 2, ext), b1 = orc::make_box(s1, 2, ext), q;
        for (int d = 0; d < 3; ++d) {
            q.mn[d] = std::min(b0.mn[d], b1.mn[d]);
            q.mx[d] = std::max(b0.mx[d], b1.mx[d]);
        }
        q.active = (m->eflag[ei] & 4u) == 0; // collision window (contact.cu:2688)
        float toi = seed;
        // edge-edge (contact.cu:2166-2264)
        orc::query(eb, q, [&](uint32_t ej) {
            const uint32_t *eb2 = m->edge + 2 * ej;
            bool either_dyn = (m->eflag[ei] & 1u) == 0 || (m->eflag[ej] & 1u) == 0;
            bool same_body = m->vbody[ea[0]] != 0 && m->vbody[ea[0]] == m->vbody[eb2[0]];
            bool both_col = vcol(ea[0]) && vcol(eb2[0]);
            if (!(ei < ej) || !either_dyn || same_body || both_col) {
                return false;
            }
            const float off = m->e_off[ei] + m->e_off[ej];
            const float gh = 0.5f * (m->e_ghat[ei] + m->e_ghat[ej]);
            float r = max_t;
            bool shared = ea[0] == eb2[0] || ea[0] == eb2[1] || ea[1] == eb2[0] ||
                          ea[1] == eb2[1];
            if (!shared) {
                V3 a[4] = {x0[ea[0]], x0[ea[1]], x0[eb2[0]], x0[eb2[1]]};
                V3 b[4] = {x1[ea[0]], x1[ea[1]], x1[eb2[0]], x1[eb2[1]]};
                r = run(3, a, b, off, gh);
            } else {
                const int ij[4][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}};
                for (int k = 0; k < 4; ++k) {
                    const int ii = ij[k][0], jj = ij[k][1];
                    if (ea[ii] == eb2[jj]) {
                        const uint32_t i0 = ea[ii], i1 = ea[1 - ii], i2 = eb2[1 - jj];
                        V3 a0[3] = {x0[i1], x0[i0], x0[i2]}, b0[3] = {x1[i1], x1[i0], x1[i2]};
                        V3 a1[3] = {x0[i2], x0[i0], x0[i1]}, b1[3] = {x1[i2], x1[i0], x1[i1]};
                        const float r0 = run(1, a0, b0, off, gh);
                        const float r1 = run(1, a1, b1, off, gh);
                        r = std::fmin(r0, r1);
                        break;
                    }
                }
            }
            if (r < max_t) {
                toi = std::fmin(toi, r);
                return true;
            }
            return false;
        });
        t_ee = std::fmin(t_ee, toi);
        // dynamic edge vs static edge (contact.cu:2036-2076)
        if (sm && sm->ne) {
            float toi2 = seed;
            orc::query(seb, q, [&](uint32_t sj) {
                if ((m->eflag[ei] & 1u) != 0 || !(lp->emass[ei] > 0.0f)) {
                    return false;
                }
                const uint32_t *se = sm->edge + 2 * sj;
                V3 a[4] = {x0[ea[0]], x0[ea[1]], sx[se[0]], sx[se[1]]};
                V3 b[4] = {x1[ea[0]], x1[ea[1]], sx[se[0]], sx[se[1]]};
                float r = run(3, a, b, m->e_off[ei] + sm->e_off[sj],
                              0.5f * (m->e_ghat[ei] + sm->e_ghat[sj]));
                if (r < max_t) {
                    toi2 = std::fmin(toi2, r);
                    return true;
                }
                return false;
            });
            t_ee = std::fmin(t_ee, toi2);
        }
    }
    *overlap = ovl ? 1 : 0;
    return std::fmin(t_vf, t_ee) / max_t;
}
}
