// plasticity.cpp -- CPU oracle: rest-shape creep of shell faces and hinges
// after a committed step.
//
// TEST INFRASTRUCTURE ONLY (the checker of the HIP path, never shipped).
//
// Restates plasticity::update_face_plasticity (plasticity/plasticity.cu:21-78)
// and plasticity::update_hinge_plasticity (plasticity.cu:145-185), called at
// main.cu:1615-1624 with ParamSet.dt. fp32, in the device code's expression
// order; alpha = 1 - exp(-rate dt) uses the C library's expf where the device
// takes its hardware exponential (a few ulps apart). Pinned in
// tests/test_oracle_plasticity.py: the new rest shape of a face reproduces
// U diag(s') V^T in its own frame (numpy SVD), rigid motions and the dead zone
// leave it unchanged, the singular values creep by the closed form of the
// dead-zone rule, and the hinge rule is checked against its closed form on
// known dihedral angles.
#include <cmath>
#include <cstdint>
#include <cstring>

extern "C" {
void oracle_svd32(uint32_t n, const float *F, float *U, float *S, float *V);
float oracle_dihedral_angle(const float *p);
}

namespace {

struct P3 {
    float x, y, z;
};

P3 sub(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
float dot(P3 a, P3 b) {
    float acc = 0.0f;
    acc = acc + a.x * b.x;
    acc = acc + a.y * b.y;
    acc = acc + a.z * b.z;
    return acc;
}
P3 cross(P3 a, P3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
P3 unit(P3 v) {
    const float z = dot(v, v);
    if (!(z > 0.0f)) return v;
    const float n = std::sqrt(z);
    return {v.x / n, v.y / n, v.z / n};
}
P3 at(const float *x, uint32_t i) { return {x[3 * (size_t)i], x[3 * (size_t)i + 1], x[3 * (size_t)i + 2]}; }

} // namespace

extern "C" {

// Creep of every face with f_rate > 0 and every non-excluded hinge
// (hflag bit0 clear) with h_rate > 0, at positions x; inv_rest (4 nf) and
// h_rest (nh) are updated in place. Null rate arrays: none.
void oracle_plasticity(uint32_t nf, uint32_t nh, const uint32_t *face, const uint32_t *hinge,
                       const uint32_t *hflag, const float *x, const float *f_rate,
                       const float *f_thr, const float *h_rate, const float *h_thr, float dt,
                       float *inv_rest, float *h_rest) {
    for (uint32_t k = 0; f_rate && k < nf; ++k) {
        const float rate = f_rate[k];
        if (!(rate > 0.0f)) continue;
        const float thr = f_thr[k];
        const float alpha = 1.0f - std::exp(-rate * dt);
        const uint32_t *v = face + 3 * (size_t)k;
        float *inv = inv_rest + 4 * (size_t)k;
        const P3 x0 = at(x, v[0]);
        const P3 d0 = sub(at(x, v[1]), x0), d1 = sub(at(x, v[2]), x0);
        float F[6];
        F[0] = (0.0f + d0.x * inv[0]) + d1.x * inv[1];
        F[1] = (0.0f + d0.y * inv[0]) + d1.y * inv[1];
        F[2] = (0.0f + d0.z * inv[0]) + d1.z * inv[1];
        F[3] = (0.0f + d0.x * inv[2]) + d1.x * inv[3];
        F[4] = (0.0f + d0.y * inv[2]) + d1.y * inv[3];
        F[5] = (0.0f + d0.z * inv[2]) + d1.z * inv[3];
        float U[6], S[2], V[4];
        oracle_svd32(1, F, U, S, V);
        float s[2] = {S[0], S[1]};
        bool changed = false;
        for (int i = 0; i < 2; ++i) {
            if (std::fabs(S[i] - 1.0f) > thr) {
                const float target = S[i] < 1.0f ? 1.0f - thr : 1.0f + thr;
                s[i] = S[i] + alpha * (target - S[i]);
                changed = true;
            }
        }
        if (!changed) continue;
        float Fn[6];
        for (int c = 0; c < 2; ++c)
            for (int r = 0; r < 3; ++r)
                Fn[r + 3 * c] = (0.0f + (U[r] * s[0]) * V[c]) + (U[3 + r] * s[1]) * V[c + 2];
        const P3 n = unit(cross(d0, d1));
        const P3 t = unit(d0);
        const P3 b = unit(cross(n, t));
        const float a00 = dot(t, d0), a01 = dot(t, d1), a10 = dot(b, d0), a11 = dot(b, d1);
        const P3 fc0 = {Fn[0], Fn[1], Fn[2]}, fc1 = {Fn[3], Fn[4], Fn[5]};
        const float p00 = dot(t, fc0), p01 = dot(t, fc1), p10 = dot(b, fc0), p11 = dot(b, fc1);
        const float det = a00 * a11 - a01 * a10;
        const float id = 1.0f / det;
        const float i00 = a11 * id, i01 = -a01 * id, i10 = -a10 * id, i11 = a00 * id;
        inv[0] = i00 * p00 + i01 * p10;
        inv[1] = i10 * p00 + i11 * p10;
        inv[2] = i00 * p01 + i01 * p11;
        inv[3] = i10 * p01 + i11 * p11;
    }
    for (uint32_t k = 0; h_rate && k < nh; ++k) {
        if (hflag[k] & 1u) continue;
        const float rate = h_rate[k];
        if (!(rate > 0.0f)) continue;
        const float thr = h_thr[k];
        const float alpha = 1.0f - std::exp(-rate * dt);
        const uint32_t *h = hinge + 4 * (size_t)k;
        // the bending order (h2, h1, h0, h3) (dihedral_angle.hpp:15-17)
        const uint32_t order[4] = {h[2], h[1], h[0], h[3]};
        float p[12];
        for (int a = 0; a < 4; ++a)
            for (int d = 0; d < 3; ++d) p[3 * a + d] = x[3 * (size_t)order[a] + d];
        const float theta = oracle_dihedral_angle(p);
        const float rest = h_rest[k];
        const float delta = theta - rest;
        if (std::fabs(delta) > thr) {
            const float edge = delta > 0.0f ? thr : -thr;
            h_rest[k] = rest + alpha * (delta - edge);
        }
    }
}

} // extern "C"
