// oracle_distance.hpp -- CPU oracle: closest-point coefficients.
//
// TEST INFRASTRUCTURE ONLY (the checker of the HIP path, never shipped).
//
// Restates in fp32 (-ffp-contract=off):
//   * distance::point_edge_distance_coeff        contact/distance.hpp:14-26
//   * distance::point_triangle_distance_coeff    contact/distance.hpp:46-81
//   * distance::edge_edge_distance_coeff         contact/distance.hpp:112-192
//   * point_triangle_distance_coeff_unclassified contact/distance.hpp:223-283
//   * point_edge_distance_coeff_unclassified     contact/distance.hpp:285-299
#pragma once
#include "oracle_math.hpp"
#include <algorithm>

namespace orc {

inline void pe_coeff(const V3 &p, const V3 &e0, const V3 &e1, float w[2]) {
    V3 r = sub(e1, e0);
    float d = dot(r, r);
    if (d > 0.0f) {
        float t = dot(r, sub(p, e0)) / d;
        w[0] = 1.0f - t;
        w[1] = t;
    } else {
        w[0] = w[1] = 1.0f / 2.0f;
    }
}

inline void pt_coeff(const V3 &p, const V3 &t0, const V3 &t1, const V3 &t2, float w[3]) {
    V3 r0 = sub(t1, t0), r1 = sub(t2, t0), d = sub(p, t0);
    V3 n = cross(r0, r1);
    float nn = dot(n, n);
    if (nn > 0.0f) {
        float inv = 1.0f / nn;
        float u = dot(cross(d, r1), n) * inv;
        float v = dot(cross(r0, d), n) * inv;
        w[0] = 1.0f - u - v;
        w[1] = u;
        w[2] = v;
        return;
    }
    V3 e1 = sub(t2, t1), e2 = sub(t0, t2);
    float l0 = dot(r0, r0), l1 = dot(e1, e1), l2 = dot(e2, e2);
    float c[2];
    if (l0 >= l1 && l0 >= l2) {
        pe_coeff(p, t0, t1, c);
        w[0] = c[0], w[1] = c[1], w[2] = 0.0f;
    } else if (l1 >= l2) {
        pe_coeff(p, t1, t2, c);
        w[0] = 0.0f, w[1] = c[0], w[2] = c[1];
    } else {
        pe_coeff(p, t2, t0, c);
        w[0] = c[1], w[1] = 0.0f, w[2] = c[0];
    }
}

inline float clamp_unit(float v) { return v > 0.0f ? (v < 1.0f ? v : 1.0f) : 0.0f; }

inline void ee_coeff(const V3 &a0, const V3 &a1, const V3 &b0, const V3 &b1, float w[4]) {
    V3 r0 = sub(a1, a0), r1 = sub(b1, b0), d = sub(a0, b0);
    float a = dot(r0, r0), e = dot(r1, r1), b = dot(r0, r1);
    float c = dot(r0, d), f = dot(r1, d);
    float inv_a = a > 0.0f ? 1.0f / a : 0.0f;
    float inv_e = e > 0.0f ? 1.0f / e : 0.0f;
    V3 n = cross(r0, r1);
    float nn = dot(n, n);
    float inv_nn = nn > 0.0f ? 1.0f / nn : 0.0f;
    float s_stat = clamp_unit(dot(n, cross(r1, d)) * inv_nn);
    float t_stat = clamp_unit((b * s_stat + f) * inv_e);
    const float cs[5] = {s_stat, 0.0f, 1.0f, clamp_unit(-c * inv_a),
                         clamp_unit((b - c) * inv_a)};
    const float ct[5] = {t_stat, clamp_unit(f * inv_e), clamp_unit((f + b) * inv_e),
                         0.0f, 1.0f};
    float bs = 0.0f, bt = 0.0f, best = 3.40282347e38f;
    for (int i = 0; i < 5; ++i) {
        V3 v = sub(add(d, scale(cs[i], r0)), scale(ct[i], r1));
        float dist = dot(v, v);
        if (dist < best) {
            best = dist;
            bs = cs[i];
            bt = ct[i];
        }
    }
    w[0] = 1.0f - bs;
    w[1] = bs;
    w[2] = 1.0f - bt;
    w[3] = bt;
}

// Closest point of the triangle (4 candidates, selected by realized distance).
inline void pt_coeff_unclassified(const V3 &p, const V3 &t0, const V3 &t1, const V3 &t2,
                           float best_w[3]) {
    V3 r0 = sub(t1, t0), r1 = sub(t2, t0), d = sub(p, t0);
    float cand[4][3];
    bool adm[4] = {false, true, true, true};
    pt_coeff(p, t0, t1, t2, cand[0]);
    adm[0] = std::min(cand[0][0], std::min(cand[0][1], cand[0][2])) >= 0.0f;
    float c[2];
    pe_coeff(p, t0, t1, c);
    float s01 = clamp_unit(c[1]);
    pe_coeff(p, t1, t2, c);
    float s12 = clamp_unit(c[1]);
    pe_coeff(p, t2, t0, c);
    float s20 = clamp_unit(c[1]);
    cand[1][0] = 1.0f - s01, cand[1][1] = s01, cand[1][2] = 0.0f;
    cand[2][0] = 0.0f, cand[2][1] = 1.0f - s12, cand[2][2] = s12;
    cand[3][0] = s20, cand[3][1] = 0.0f, cand[3][2] = 1.0f - s20;
    best_w[0] = cand[1][0], best_w[1] = cand[1][1], best_w[2] = cand[1][2];
    float best = 3.40282347e38f;
    for (int i = 0; i < 4; ++i) {
        if (!adm[i]) {
            continue;
        }
        V3 v = sub(sub(d, scale(cand[i][1], r0)),
                        scale(cand[i][2], r1));
        float dist = dot(v, v);
        if (dist < best) {
            best = dist;
            best_w[0] = cand[i][0], best_w[1] = cand[i][1], best_w[2] = cand[i][2];
        }
    }
}

inline void pe_coeff_unclassified(const V3 &p, const V3 &e0, const V3 &e1, float w[2]) {
    pe_coeff(p, e0, e1, w);
    if (!(w[0] >= 0.0f && w[0] <= 1.0f)) {
        if (w[0] > 1.0f) {
            w[0] = 1.0f, w[1] = 0.0f;
        } else {
            w[0] = 0.0f, w[1] = 1.0f;
        }
    }
}

} // namespace orc
