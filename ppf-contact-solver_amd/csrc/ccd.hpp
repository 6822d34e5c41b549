// ccd.hpp -- additive continuous collision detection (device).
//
// The reference's ACCD (contact/accd.hpp:64-696): centre and rescale the
// sweep frame, bound the closing speed by the largest inter-primitive
// relative displacement, conservatively advance (with the directional
// certificate along the closest-point direction) until the pair reaches its
// parking clearance, then bisect to the last certified time. Returns the
// largest certified toi in [0, max_t]; 0 with `overlap` set when the pair
// starts inside its offset or the frame collapses. Kept op for op: the toi
// decides the accepted step.
#pragma once
#include "distance.hpp"

namespace ppf {

constexpr float kCcdFltEps = 1.19209290e-7f;

template <int C> struct Frame {
    f3 c[C];
};

PPF_HD f3 fabs3(f3 a) { return mk3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }

PPF_HD void pt_coeff_unclassified(f3 p, f3 t0, f3 t1, f3 t2, float bw[3]) {
    const f3 r0 = t1 - t0, r1 = t2 - t0, d = p - t0;
    float cand[4][3];
    bool adm[4] = {false, true, true, true};
    pt_coeff(p, t0, t1, t2, cand[0]);
    adm[0] = fminf(cand[0][0], fminf(cand[0][1], cand[0][2])) >= 0.0f;
    float c[2];
    pe_coeff(p, t0, t1, c);
    const float s01 = clamp_unit(c[1]);
    pe_coeff(p, t1, t2, c);
    const float s12 = clamp_unit(c[1]);
    pe_coeff(p, t2, t0, c);
    const float s20 = clamp_unit(c[1]);
    cand[1][0] = 1.0f - s01, cand[1][1] = s01, cand[1][2] = 0.0f;
    cand[2][0] = 0.0f, cand[2][1] = 1.0f - s12, cand[2][2] = s12;
    cand[3][0] = s20, cand[3][1] = 0.0f, cand[3][2] = 1.0f - s20;
    bw[0] = cand[1][0], bw[1] = cand[1][1], bw[2] = cand[1][2];
    float best = 3.40282347e38f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (!adm[i]) {
            continue;
        }
        const f3 v = (d - cand[i][1] * r0) - cand[i][2] * r1;
        const float dist = dot3(v, v);
        if (dist < best) {
            best = dist;
            bw[0] = cand[i][0], bw[1] = cand[i][1], bw[2] = cand[i][2];
        }
    }
}

PPF_HD void pe_coeff_unclassified(f3 p, f3 e0, f3 e1, float w[2]) {
    pe_coeff(p, e0, e1, w);
    if (!(w[0] >= 0.0f && w[0] <= 1.0f)) {
        if (w[0] > 1.0f) {
            w[0] = 1.0f, w[1] = 0.0f;
        } else {
            w[0] = 0.0f, w[1] = 1.0f;
        }
    }
}

template <int C> PPF_HD void centerize(Frame<C> &x) {
    const float s = (float)(1.0 / C);
    f3 mov = mk3(0, 0, 0);
#pragma unroll
    for (int k = 0; k < C; ++k) {
        mov = mov + s * x.c[k];
    }
#pragma unroll
    for (int k = 0; k < C; ++k) {
        x.c[k] = x.c[k] - mov;
    }
}

// Returns the scale applied, 0 (and coincident) for a collapsed frame.
template <int C>
PPF_HD float rescale(Frame<C> &x, Frame<C> &dx, float max_t, bool &coincident) {
    float m0 = 0.0f, m1 = 0.0f;
#pragma unroll
    for (int k = 0; k < C; ++k) {
        const float a[3] = {x.c[k].x, x.c[k].y, x.c[k].z};
        const float v[3] = {dx.c[k].x, dx.c[k].y, dx.c[k].z};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const float p = fabsf(a[r]);
            const float q = fabsf(a[r] + max_t * v[r]);
            m0 = (k == 0 && r == 0) ? p : fmaxf(m0, p);
            m1 = (k == 0 && r == 0) ? q : fmaxf(m1, q);
        }
    }
    float scale = 0.99f / fmaxf(m0, m1);
    coincident = false;
    if (!(scale > 0.0f) || !isfinite(scale)) {
        coincident = true;
        return 0.0f;
    }
    float scaled = 1.0f;
    const float s = 8.0f;
    while (true) {
        if (scale > s) {
#pragma unroll
            for (int k = 0; k < C; ++k) {
                x.c[k] = s * x.c[k];
                dx.c[k] = s * dx.c[k];
            }
            scaled *= s;
            scale /= s;
        } else {
#pragma unroll
            for (int k = 0; k < C; ++k) {
                x.c[k] = scale * x.c[k];
                dx.c[k] = scale * dx.c[k];
            }
            scaled *= scale;
            break;
        }
    }
    return scaled;
}

template <int C, int SPLIT> PPF_HD float max_relative_u(const Frame<C> &u) {
    float mu = 0.0f;
#pragma unroll
    for (int i = 0; i < SPLIT; ++i) {
#pragma unroll
        for (int j = SPLIT; j < C; ++j) {
            const f3 d = u.c[i] - u.c[j];
            mu = fmaxf(mu, dot3(d, d));
        }
    }
    return sqrtf(mu);
}

template <int C, int SPLIT>
PPF_HD float directional_advance(const Frame<C> &x, const Frame<C> &dx, f3 w,
                                 float w_norm, float dip, float park, float max_t) {
    const float round_slack = 2.0f * kCcdFltEps;
    const float coord_slack = 6.2e-7f;
    const f3 wa = fabs3(w);
    const float dip_w = dip * w_norm;
    const float park_w = park * w_norm;
    float step = 3.40282347e38f;
#pragma unroll
    for (int i = 0; i < SPLIT; ++i) {
#pragma unroll
        for (int j = SPLIT; j < C; ++j) {
            const f3 gap = x.c[j] - x.c[i];
            const f3 rel = dx.c[i] - dx.c[j];
            const float err =
                round_slack * (dot3(wa, fabs3(gap)) + max_t * dot3(wa, fabs3(rel)) +
                               park_w) +
                coord_slack * w_norm;
            const float projected = dot3(w, gap) - err;
            if (!(projected - dip_w > 0.0f)) {
                return 0.0f;
            }
            const float rate = dot3(w, rel);
            if (rate > 0.0f) {
                step = fminf(step, fmaxf(0.0f, projected - park_w) / rate);
            }
        }
    }
    return step;
}

// Squared distance and closest-point direction (first primitive -> second).
template <int C, int SPLIT> PPF_HD float sqdist(const Frame<C> &x, f3 &dir) {
    if (C == 4 && SPLIT == 1) {
        float c[3];
        pt_coeff_unclassified(x.c[0], x.c[1], x.c[2], x.c[3], c);
        dir = (c[0] * (x.c[1] - x.c[0]) + c[1] * (x.c[2] - x.c[0])) +
              c[2] * (x.c[3] - x.c[0]);
    } else if (C == 4) {
        float c[4];
        ee_coeff(x.c[0], x.c[1], x.c[2], x.c[3], c);
        const f3 a = c[0] * x.c[0] + c[1] * x.c[1];
        const f3 b = c[2] * x.c[2] + c[3] * x.c[3];
        dir = b - a;
    } else if (C == 3) {
        float c[2];
        pe_coeff_unclassified(x.c[0], x.c[1], x.c[2], c);
        dir = (c[0] * x.c[1] + c[1] * x.c[2]) - x.c[0];
    } else {
        dir = x.c[1] - x.c[0];
    }
    return dot3(dir, dir);
}

template <int C> PPF_HD Frame<C> sweep_at(const Frame<C> &x0, const Frame<C> &dx, float t) {
    Frame<C> x;
#pragma unroll
    for (int k = 0; k < C; ++k) {
        x.c[k] = x0.c[k] + t * dx.c[k];
    }
    return x;
}

template <int C, int SPLIT>
PPF_HD float ccd_helper(const Frame<C> &x0, const Frame<C> &dx, float u_max,
                        float offset, float eps, float floor_gap, float max_t,
                        bool &overlap) {
    Frame<C> x = x0;
    f3 w;
    float lower_t = 0.0f, toi = 0.0f;
    float d2 = sqdist<C, SPLIT>(x0, w);
    if (!(d2 > offset * offset)) {
        overlap = true;
        return 0.0f;
    }
    const float clearance = sqrtf(d2) - offset;
    const float floor_clear = fmaxf(2.0f * eps, floor_gap);
    const float park =
        offset + fminf(floor_clear, fmaxf(clearance - 0.5f * eps, 0.5f * clearance));
    const float overshoot = fmaxf(0.0f, (park - offset) - 0.5f * floor_clear);
    const float dip = park - overshoot;
    const float park_squared = fminf(park * park, nextafterf(d2, 0.0f));
    const float inv_u_max = 1.0f / u_max;
    unsigned probes = 0;
    while (true) {
        if (!(d2 > park_squared)) {
            break;
        }
        lower_t = toi;
        const float dist = sqrtf(d2);
        float step = (dist - dip) * inv_u_max;
        if (toi + step <= max_t) {
            step = fmaxf(step, directional_advance<C, SPLIT>(
                                   x, dx, w, dist * (1.0f + 2.0f * kCcdFltEps), dip,
                                   park, max_t));
        }
        toi += step;
        if (toi > max_t) {
            return max_t;
        } else if (++probes > 4096u) {
            return lower_t;
        }
        x = sweep_at(x0, dx, toi);
        d2 = sqdist<C, SPLIT>(x, w);
    }
    float upper_t = toi;
    float window = upper_t - lower_t;
    while (true) {
        toi = 0.5f * (upper_t + lower_t);
        x = sweep_at(x0, dx, toi);
        f3 unused;
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
        overlap = true;
        return 0.0f;
    }
    return lower_t;
}

// One pair kind over C columns (SPLIT on the first primitive); p0 / p1 are
// the columns at the start and at the end of the unit step.
template <int C, int SPLIT>
PPF_HD float ccd(const f3 *p0, const f3 *p1, float offset, float ghat, float max_t,
                 float ccd_eps, bool &overlap) {
    Frame<C> x0, dx;
#pragma unroll
    for (int k = 0; k < C; ++k) {
        x0.c[k] = p0[k];
        dx.c[k] = p1[k] - p0[k];
    }
    centerize(x0);
    centerize(dx);
    bool coincident = false;
    const float scale = rescale(x0, dx, max_t, coincident);
    if (coincident) {
        overlap = true;
    }
    if (!(scale > 0.0f)) {
        return 0.0f;
    }
    const float u_max = max_relative_u<C, SPLIT>(dx);
    if (u_max != 0.0f) {
        return ccd_helper<C, SPLIT>(x0, dx, u_max, scale * offset, scale * ccd_eps,
                                    scale * (1e-2f * ghat), max_t, overlap);
    }
    return max_t;
}

} // namespace ppf
