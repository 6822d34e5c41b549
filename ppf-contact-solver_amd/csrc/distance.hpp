// distance.hpp -- closest-point barycentric coefficients (device).
//
// The same formulas as the reference's contact/distance.hpp: point-edge
// projection (14-26); point-triangle projection through the cross product with
// the longest-edge fallback for a zero-area triangle (46-81); segment-segment
// closest points as the best of five clamped candidates, the interior one
// taken through the cross product (112-192). Kept op for op: the contact set
// is decided by comparing these coefficients against 0 and 1.
#pragma once
#include "common.hpp"

namespace ppf {

PPF_HD void pe_coeff(f3 p, f3 e0, f3 e1, float w[2]) {
    const f3 r = e1 - e0;
    const float d = dot3(r, r);
    if (d > 0.0f) {
        const float t = dot3(r, p - e0) / d;
        w[0] = 1.0f - t;
        w[1] = t;
    } else {
        w[0] = w[1] = 1.0f / 2.0f;
    }
}

PPF_HD void pt_coeff(f3 p, f3 t0, f3 t1, f3 t2, float w[3]) {
    const f3 r0 = t1 - t0, r1 = t2 - t0, d = p - t0;
    const f3 n = cross3(r0, r1);
    const float nn = dot3(n, n);
    if (nn > 0.0f) {
        const float inv = 1.0f / nn;
        const float u = dot3(cross3(d, r1), n) * inv;
        const float v = dot3(cross3(r0, d), n) * inv;
        w[0] = 1.0f - u - v;
        w[1] = u;
        w[2] = v;
        return;
    }
    const f3 e1 = t2 - t1, e2 = t0 - t2;
    const float l0 = dot3(r0, r0), l1 = dot3(e1, e1), l2 = dot3(e2, e2);
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

PPF_HD float clamp_unit(float v) { return v > 0.0f ? (v < 1.0f ? v : 1.0f) : 0.0f; }

PPF_HD void ee_coeff(f3 a0, f3 a1, f3 b0, f3 b1, float w[4]) {
    const f3 r0 = a1 - a0, r1 = b1 - b0, d = a0 - b0;
    const float a = dot3(r0, r0), e = dot3(r1, r1), b = dot3(r0, r1);
    const float c = dot3(r0, d), f = dot3(r1, d);
    const float inv_a = a > 0.0f ? 1.0f / a : 0.0f;
    const float inv_e = e > 0.0f ? 1.0f / e : 0.0f;
    const f3 n = cross3(r0, r1);
    const float nn = dot3(n, n);
    const float inv_nn = nn > 0.0f ? 1.0f / nn : 0.0f;
    const float s_stat = clamp_unit(dot3(n, cross3(r1, d)) * inv_nn);
    const float t_stat = clamp_unit((b * s_stat + f) * inv_e);
    const float cs[5] = {s_stat, 0.0f, 1.0f, clamp_unit(-c * inv_a),
                         clamp_unit((b - c) * inv_a)};
    const float ct[5] = {t_stat, clamp_unit(f * inv_e), clamp_unit((f + b) * inv_e),
                         0.0f, 1.0f};
    float bs = 0.0f, bt = 0.0f, best = 3.40282347e38f;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const f3 v = (d + cs[i] * r0) - ct[i] * r1;
        const float dist = dot3(v, v);
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

template <int N> PPF_HD bool interior(const float *c) {
    float mx = c[0], mn = c[0];
#pragma unroll
    for (int i = 1; i < N; ++i) {
        mx = mx < c[i] ? c[i] : mx;
        mn = c[i] < mn ? c[i] : mn;
    }
    return mx < 1.0f && mn > 0.0f;
}

} // namespace ppf
