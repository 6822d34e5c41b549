// elastic.cpp -- CPU oracle: inertia, air drag, membrane and bending terms.
//
// TEST INFRASTRUCTURE ONLY (the checker of the HIP path, never shipped).
//
// Restates in fp32 (-ffp-contract=off), op for op:
//   * embed_vertex_force_hessian: air damper on the area-weighted vertex
//     normal, momentum m (y - target) / dt^2, m I / dt^2 into the per-vertex
//     diagonal, only for unpinned vertices      energy/energy.cu:182-270, 700-713
//     (air_damper.hpp:13-42, momentum.hpp:14-22, utility.cu:15-32)
//   * embed_face_force_hessian, Baraff-Witkin branch: F = dX inv_rest,
//     stretch + shear gradient / PSD Hessian, mass-scaled, mapped to the three
//     vertices                                   energy/energy.cu:304-393,
//     energy/model/baraffwitkin.hpp:27-226, utility/utility.cu:185-282
//   * embed_hinge_force_hessian: areal-density-normalized Discrete Shells
//     stiffness, dihedral angle (atan2 form), its gradient and the Wu-Kim
//     projected Hessian (F-part eigenpairs + PSD-projected W-part)
//                                                energy/energy.cu:448-553,
//     energy/model/dihedral_angle.hpp:15-355, linalg/eigsolve.hpp:245-316
//   * FixedCSRMat::push (upper triangle, zero entries skipped)
//                                                csrmat/csrmat.cu:544-594
// Contributions are summed in a fixed order: per stored block, faces by
// index then hinges by index (local block order within an element); per
// vertex, the vertex term, then faces, then hinges.
#include "oracle_math.hpp"
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

using orc::V3;
using orc::v3;

// ---- small dense helpers (column-major, accumulate from zero) ---------------

template <int R, int C> struct Mat {
    float m[R * C];
    float &operator()(int r, int c) { return m[r + R * c]; }
    float operator()(int r, int c) const { return m[r + R * c]; }
    static Mat zero() {
        Mat z;
        std::memset(z.m, 0, sizeof(z.m));
        return z;
    }
};

template <int N> struct Vec {
    float v[N];
    float &operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
};

template <int N> float vdot(const Vec<N> &a, const Vec<N> &b) {
    float acc = 0.0f;
    for (int i = 0; i < N; ++i) {
        acc = acc + a[i] * b[i];
    }
    return acc;
}

template <int N> Vec<N> vscale(float s, const Vec<N> &a) {
    Vec<N> r;
    for (int i = 0; i < N; ++i) r[i] = s * a[i];
    return r;
}
template <int N> Vec<N> vadd(const Vec<N> &a, const Vec<N> &b) {
    Vec<N> r;
    for (int i = 0; i < N; ++i) r[i] = a[i] + b[i];
    return r;
}
template <int N> Vec<N> vsub(const Vec<N> &a, const Vec<N> &b) {
    Vec<N> r;
    for (int i = 0; i < N; ++i) r[i] = a[i] - b[i];
    return r;
}
template <int N> void vnormalize(Vec<N> &a) {
    float n = std::sqrt(vdot(a, a));
    if (n > 0.0f) {
        for (int i = 0; i < N; ++i) a[i] = a[i] / n;
    }
}

V3 vdiv(const V3 &a, float s) { return v3(a[0] / s, a[1] / s, a[2] / s); }
V3 vmul(const V3 &a, float s) { return v3(a[0] * s, a[1] * s, a[2] * s); }
float vnorm(const V3 &a) { return std::sqrt(orc::dot(a, a)); }
void normalize3(V3 &a) {
    float n = vnorm(a);
    if (n > 0.0f) {
        a = vdiv(a, n);
    }
}

// 3x3 outer product a b^T (each entry 0 + a_r b_c).
orc::M3 outer(const V3 &a, const V3 &b) {
    orc::M3 r;
    for (int c = 0; c < 3; ++c)
        for (int rr = 0; rr < 3; ++rr) r(rr, c) = 0.0f + a[rr] * b[c];
    return r;
}
orc::M3 mscale(float s, const orc::M3 &a) {
    orc::M3 r;
    for (int i = 0; i < 9; ++i) r.m[i] = s * a.m[i];
    return r;
}
orc::M3 mdiv(const orc::M3 &a, float s) {
    orc::M3 r;
    for (int i = 0; i < 9; ++i) r.m[i] = a.m[i] / s;
    return r;
}
orc::M3 madd(const orc::M3 &a, const orc::M3 &b) {
    orc::M3 r;
    for (int i = 0; i < 9; ++i) r.m[i] = a.m[i] + b.m[i];
    return r;
}
orc::M3 msub(const orc::M3 &a, const orc::M3 &b) {
    orc::M3 r;
    for (int i = 0; i < 9; ++i) r.m[i] = a.m[i] - b.m[i];
    return r;
}
orc::M3 mzero() {
    orc::M3 r;
    for (int i = 0; i < 9; ++i) r.m[i] = 0.0f;
    return r;
}
orc::M3 meye() {
    orc::M3 r = mzero();
    r(0, 0) = r(1, 1) = r(2, 2) = 1.0f;
    return r;
}
orc::M3 mtrans(const orc::M3 &a) {
    orc::M3 r;
    for (int c = 0; c < 3; ++c)
        for (int rr = 0; rr < 3; ++rr) r(rr, c) = a(c, rr);
    return r;
}
// matrix * vector, accumulated over k from zero (smat.hpp operator*)
V3 mv(const orc::M3 &a, const V3 &x) {
    V3 r = v3(0, 0, 0);
    for (int k = 0; k < 3; ++k)
        for (int rr = 0; rr < 3; ++rr) r[rr] = r[rr] + a(rr, k) * x[k];
    return r;
}

// ---- inertia + air drag ------------------------------------------------------

struct Params {
    float dt, air_density, air_friction;
    float wind[3];
};

void vertex_term(const V3 &y, const V3 &x, const V3 &target, const V3 &normal, float mass,
                 float area, const Params &p, V3 &f_out, orc::M3 &h_out) {
    V3 f = v3(0, 0, 0);
    orc::M3 H = mzero();
    const float dt2 = p.dt * p.dt;
    bool nz = false;
    for (int d = 0; d < 3; ++d) {
        if (!(std::fabs(normal[d]) <= 1e-5f)) nz = true;
    }
    if (nz && p.air_density != 0.0f) {
        const V3 wind = v3(p.wind[0], p.wind[1], p.wind[2]);
        const V3 z = orc::sub(orc::sub(y, x), orc::scale(p.dt, wind));
        const orc::M3 P = msub(meye(), outer(normal, normal));
        // normal * n.z / dt^2 + ((air_friction * P) * z) / dt^2
        const V3 g = orc::add(vdiv(vmul(normal, orc::dot(normal, z)), dt2),
                              vdiv(mv(mscale(p.air_friction, P), z), dt2));
        const orc::M3 h = madd(mdiv(outer(normal, normal), dt2),
                               mdiv(mscale(p.air_friction, P), dt2));
        const float s = area * p.air_density;
        f = orc::add(f, orc::scale(s, g));
        H = madd(H, mscale(s, h));
    }
    f = orc::add(f, orc::scale(mass, vdiv(orc::sub(y, target), dt2)));
    H = madd(H, mscale(mass, mdiv(meye(), dt2)));
    f_out = f;
    h_out = H;
}

// ---- Baraff-Witkin -----------------------------------------------------------

using M32 = Mat<3, 2>;
using M66 = Mat<6, 6>;

V3 col(const M32 &F, int c) { return v3(F(0, c), F(1, c), F(2, c)); }

M32 stretch_gradient(const M32 &F, float mu) {
    const V3 fu = col(F, 0), fv = col(F, 1);
    const float nu = vnorm(fu), nv = vnorm(fv);
    const V3 a = vdiv(vmul(fu, nu - 1.0f), nu);
    const V3 b = vdiv(vmul(fv, nv - 1.0f), nv);
    M32 r;
    for (int d = 0; d < 3; ++d) {
        r(d, 0) = mu * a[d];
        r(d, 1) = mu * b[d];
    }
    return r;
}

// One column's 3x3 block of the stretch Hessian.
orc::M3 stretch_block(const V3 &q, float norm_q) {
    V3 qa = orc::cross(v3(q[2], 0.0f, -q[0]), q);
    V3 qb = orc::cross(v3(0.0f, q[2], -q[1]), q);
    V3 qc = orc::cross(v3(q[1], -q[0], 0.0f), q);
    const float s[3] = {orc::dot(qa, qa), orc::dot(qb, qb), orc::dot(qc, qc)};
    V3 q2;
    if (s[0] > s[1] && s[0] > s[2]) {
        q2 = qa;
    } else if (s[1] > s[0] && s[1] > s[2]) {
        q2 = qb;
    } else {
        q2 = qc;
    }
    V3 q1 = q;
    V3 q3 = orc::cross(q1, q2);
    normalize3(q1);
    normalize3(q2);
    normalize3(q3);
    const float l1 = 1.0f, l2 = 1.0f - 1.0f / norm_q, l3 = 1.0f - 1.0f / norm_q;
    // max(0, l) * q * q^T parses as ((max(0, l) * q) * q^T).
    orc::M3 H = mzero();
    H = madd(H, outer(orc::scale(std::max(0.0f, l1), q1), q1));
    H = madd(H, outer(orc::scale(std::max(0.0f, l2), q2), q2));
    H = madd(H, outer(orc::scale(std::max(0.0f, l3), q3), q3));
    return H;
}

M66 stretch_hessian(const M32 &F, float mu) {
    const orc::M3 H1 = stretch_block(col(F, 0), vnorm(col(F, 0)));
    const orc::M3 H2 = stretch_block(col(F, 1), vnorm(col(F, 1)));
    M66 H = M66::zero();
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
            H(r, c) = H1(r, c);
            H(3 + r, 3 + c) = H2(r, c);
        }
    M66 out;
    for (int i = 0; i < 36; ++i) out.m[i] = mu * H.m[i];
    return out;
}

M32 shear_gradient(const M32 &F, float lmd) {
    const float w = orc::dot(col(F, 0), col(F, 1));
    M32 r;
    for (int d = 0; d < 3; ++d) {
        r(d, 0) = lmd * (w * F(d, 1));
        r(d, 1) = lmd * (w * F(d, 0));
    }
    return r;
}

M66 shear_hessian(const M32 &F, float lmd) {
    const float I6 = orc::dot(col(F, 0), col(F, 1));
    float I2 = 0.0f;
    for (int i = 0; i < 6; ++i) I2 = I2 + F.m[i] * F.m[i];
    Vec<6> g, Hg;
    for (int d = 0; d < 3; ++d) {
        g[d] = F(d, 1);
        g[3 + d] = F(d, 0);
        Hg[d] = F(d, 0);
        Hg[3 + d] = F(d, 1);
    }
    const float lmd0 = I2 + std::sqrt(I2 * I2 + 12.0f * I6 * I6);
    Vec<6> q0 = vadd(vscale(2.0f * I6, Hg), vscale(lmd0, g));
    Vec<6> z, w1{}, w2{}, w3{};
    float lmd1, lmd2;
    if (I6 > 0) {
        z = q0;
        lmd1 = lmd2 = 2.0f * I6;
        w1[0] = 1, w1[1] = 1;
        w2[2] = 1, w2[3] = 1;
        w3[4] = 1, w3[5] = 1;
    } else {
        lmd1 = lmd2 = -2.0f * I6;
        const float tmp = I2 - std::sqrt(I2 * I2 + 12.0f * I6 * I6);
        z = vadd(vscale(2.0f * I6, Hg), vscale(tmp, g));
        w1[0] = 1, w1[1] = -1;
        w2[2] = 1, w2[3] = -1;
        w3[4] = 1, w3[5] = -1;
    }
    const float a1 = vdot(z, w1), a2 = vdot(z, w2), a3 = vdot(z, w3);
    const Vec<6> q1_1 = vsub(vadd(vscale(a3, w1), vscale(a3, w2)), vscale(a1 - a2, w3));
    const Vec<6> q1_2 = vadd(vsub(vscale(a2, w1), vscale(a1 - a3, w2)), vscale(a2, w3));
    const Vec<6> q1_3 = vadd(vadd(vscale(-(a2 - a3), w1), vscale(a1, w2)), vscale(a1, w3));
    const float n1[3] = {vdot(q1_1, q1_1), vdot(q1_2, q1_2), vdot(q1_3, q1_3)};
    Vec<6> q1;
    if (n1[0] > n1[1] && n1[0] > n1[2]) {
        q1 = q1_1;
    } else if (n1[1] > n1[0] && n1[1] > n1[2]) {
        q1 = q1_2;
    } else {
        q1 = q1_3;
    }
    const float b1 = vdot(q1, w1), b2 = vdot(q1, w2), b3 = vdot(q1, w3);
    // 2x2 systems [[p, q], [r, s]] (row-major) with their adjugates.
    auto det2 = [](float p, float q, float r, float s) { return p * s - q * r; };
    auto adj_mul = [](float p, float q, float r, float s, float x, float y, float out[2]) {
        // adj = [[s, -q], [-r, p]] (smat column-major product, from zero)
        out[0] = (0.0f + s * x) + (-q) * y;
        out[1] = (0.0f + (-r) * x) + p * y;
    };
    const float d1 = det2(a2, a3, b2, b3), d2 = det2(a3, a1, b3, b1), d3 = det2(a1, a2, b1, b2);
    float c45[2], c56[2], c64[2];
    adj_mul(a1, a2, b1, b2, -a3, -b3, c45);
    adj_mul(a2, a3, b2, b3, -a1, -b1, c56);
    adj_mul(a3, a1, b3, b1, -a2, -b2, c64);
    const Vec<6> q2_1 = vadd(vadd(vscale(c45[0], w1), vscale(c45[1], w2)), vscale(d3, w3));
    const Vec<6> q2_2 = vadd(vadd(vscale(d1, w1), vscale(c56[0], w2)), vscale(c56[1], w3));
    const Vec<6> q2_3 = vadd(vadd(vscale(c64[1], w1), vscale(d2, w2)), vscale(c64[0], w3));
    const float n2[3] = {vdot(q2_1, q2_1), vdot(q2_2, q2_2), vdot(q2_3, q2_3)};
    Vec<6> q2;
    if (n2[0] > n2[1] && n2[0] > n2[2]) {
        q2 = q2_1;
    } else if (n2[1] > n2[0] && n2[1] > n2[2]) {
        q2 = q2_2;
    } else {
        q2 = q2_3;
    }
    vnormalize(q0);
    vnormalize(q1);
    vnormalize(q2);
    M66 res = M66::zero();
    auto acc = [&](float l, const Vec<6> &q) {
        const float s = std::max(0.0f, l);
        for (int c = 0; c < 6; ++c)
            for (int r = 0; r < 6; ++r) res(r, c) = res(r, c) + s * (0.0f + q[r] * q[c]);
    };
    acc(lmd0, q0);
    acc(lmd1, q1);
    acc(lmd2, q2);
    M66 out;
    const float h = 0.5f * lmd;
    for (int i = 0; i < 36; ++i) out.m[i] = h * res.m[i];
    return out;
}

// inv_rest: column-major 2x2 (m[r + 2c]).
void face_term(const V3 &x0, const V3 &x1, const V3 &x2, const float *inv, float mass,
               float mu, float lmd, float f_out[9], float h_out[81]) {
    M32 dx;
    for (int d = 0; d < 3; ++d) {
        dx(d, 0) = x1[d] - x0[d];
        dx(d, 1) = x2[d] - x0[d];
    }
    M32 F = M32::zero();
    for (int c = 0; c < 2; ++c)
        for (int k = 0; k < 2; ++k) {
            const float b = inv[k + 2 * c];
            for (int r = 0; r < 3; ++r) F(r, c) = F(r, c) + dx(r, k) * b;
        }
    const M32 g0 = stretch_gradient(F, mu), g1 = shear_gradient(F, lmd);
    const M66 h0 = stretch_hessian(F, mu), h1 = shear_hessian(F, lmd);
    M32 dedF;
    for (int i = 0; i < 6; ++i) dedF.m[i] = g0.m[i] + g1.m[i];
    M66 d2;
    for (int i = 0; i < 36; ++i) d2.m[i] = h0.m[i] + h1.m[i];
    // g_a rows of the deformation map
    const float gg[3][2] = {{-inv[0] - inv[1], -inv[2] - inv[3]},
                            {inv[0], inv[2]},
                            {inv[1], inv[3]}};
    for (int a = 0; a < 3; ++a)
        for (int d = 0; d < 3; ++d) {
            const float fe = (0.0f + gg[a][0] * dedF(d, 0)) + gg[a][1] * dedF(d, 1);
            f_out[3 * a + d] = 0.0f + mass * fe;
        }
    // 9x9 column-major
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            float blk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int d = 0; d < 2; ++d)
                for (int e = 0; e < 2; ++e) {
                    const float s = gg[a][d] * gg[b][e];
                    for (int c = 0; c < 3; ++c)
                        for (int r = 0; r < 3; ++r)
                            blk[r + 3 * c] = blk[r + 3 * c] + s * d2(3 * d + r, 3 * e + c);
                }
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r)
                    h_out[(3 * a + r) + 9 * (3 * b + c)] = 0.0f + mass * blk[r + 3 * c];
        }
}

// ---- dihedral bending (Wu-Kim projected Hessian) ----------------------------

float dihedral_angle(const V3 &v0, const V3 &v1, const V3 &v2, const V3 &v3_) {
    const V3 n1 = orc::cross(orc::sub(v1, v0), orc::sub(v2, v0));
    const V3 n2 = orc::cross(orc::sub(v2, v3_), orc::sub(v1, v3_));
    const V3 e = orc::sub(v1, v2);
    const float en = vnorm(e);
    if (en <= 0.0f) {
        return 0.0f;
    }
    const float s = orc::dot(orc::cross(n2, n1), e) / en;
    const float c = orc::dot(n1, n2);
    return std::atan2(s, c);
}

// Gradient of the angle in argument order (X0, X1, X2, X3); the reference's
// parameter names are (v2, v0, v1, v3).
void dihedral_grad(const V3 &X0, const V3 &X1, const V3 &X2, const V3 &X3, V3 g[4]) {
    const V3 &v2 = X0, &v0 = X1, &v1 = X2, &v3_ = X3;
    const V3 e0 = orc::sub(v1, v0), e1 = orc::sub(v2, v0), e2 = orc::sub(v3_, v0);
    const V3 e3 = orc::sub(v2, v1), e4 = orc::sub(v3_, v1);
    const V3 n1 = orc::cross(e0, e1), n2 = orc::cross(e2, e0);
    const float n1s = orc::dot(n1, n1), n2s = orc::dot(n2, n2);
    const float en = vnorm(e0);
    g[0] = orc::scale(-en / n1s, n1);
    g[1] = orc::sub(orc::scale(-orc::dot(e0, e3) / (en * n1s), n1),
                    orc::scale(orc::dot(e0, e4) / (en * n2s), n2));
    g[2] = orc::add(orc::scale(orc::dot(e0, e1) / (en * n1s), n1),
                    orc::scale(orc::dot(e0, e2) / (en * n2s), n2));
    g[3] = orc::scale(-en / n2s, n2);
}

void psd_project4(float A[4][4]) {
    float V[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    float scale = 0.0f;
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i) scale = std::fmax(scale, std::fabs(A[i][j]));
    if (scale == 0.0f) {
        return;
    }
    const float tiny = 1e-7f * scale;
    for (int sweep = 0; sweep < 24; ++sweep) {
        float off = 0.0f;
        for (int p = 0; p < 4; ++p)
            for (int q = p + 1; q < 4; ++q) off = std::fmax(off, std::fabs(A[p][q]));
        if (off <= tiny) {
            break;
        }
        for (int p = 0; p < 4; ++p)
            for (int q = p + 1; q < 4; ++q) {
                const float apq = A[p][q];
                if (std::fabs(apq) <= tiny) {
                    continue;
                }
                const float app = A[p][p], aqq = A[q][q];
                const float tau = (aqq - app) / (2.0f * apq);
                const float t = (tau >= 0.0f ? 1.0f : -1.0f) /
                                (std::fabs(tau) + std::sqrt(1.0f + tau * tau));
                const float c = 1.0f / std::sqrt(1.0f + t * t);
                const float s = t * c;
                A[p][p] = c * c * app - 2.0f * s * c * apq + s * s * aqq;
                A[q][q] = s * s * app + 2.0f * s * c * apq + c * c * aqq;
                A[p][q] = 0.0f;
                A[q][p] = 0.0f;
                for (int k = 0; k < 4; ++k) {
                    if (k == p || k == q) continue;
                    const float akp = A[k][p], akq = A[k][q];
                    const float np = c * akp - s * akq, nq = s * akp + c * akq;
                    A[k][p] = np;
                    A[p][k] = np;
                    A[k][q] = nq;
                    A[q][k] = nq;
                }
                for (int k = 0; k < 4; ++k) {
                    const float vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    float lam[4];
    for (int i = 0; i < 4; ++i) lam[i] = std::fmax(A[i][i], 0.0f);
    for (int i = 0; i < 4; ++i)
        for (int j = i; j < 4; ++j) {
            float sum = 0.0f;
            for (int k = 0; k < 4; ++k) sum += V[i][k] * lam[k] * V[j][k];
            A[i][j] = sum;
            A[j][i] = sum;
        }
}

// Force (12, per unit stiffness, X0..X3 order) and 12x12 Hessian (row-major
// H[i][j], X0..X3 order) of E = 1/2 (theta - theta0)^2 at X0..X3 (the
// remapped hinge).
void hinge_unit(const V3 &X0, const V3 &X1, const V3 &X2, const V3 &X3, float rest,
                float f12[12], float H[12][12]) {
    const float angle = dihedral_angle(X0, X1, X2, X3);
    V3 g[4];
    dihedral_grad(X0, X1, X2, X3, g);
    const float gdiff = angle - rest;
    for (int k = 0; k < 4; ++k)
        for (int d = 0; d < 3; ++d) f12[3 * k + d] = gdiff * g[k][d];
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) H[i][j] = 0.0f;
    const V3 y1 = orc::sub(X1, X2), y2 = orc::sub(X0, X2), y3 = orc::sub(X3, X2);
    const float ly1 = vnorm(y1);
    if (ly1 <= 1e-9f) {
        return;
    }
    const V3 t1 = vdiv(y1, ly1);
    const orc::M3 sigma = msub(meye(), outer(t1, t1));
    const V3 z0 = orc::sub(y2, vmul(t1, orc::dot(t1, y2)));
    const V3 z1 = orc::sub(y3, vmul(t1, orc::dot(t1, y3)));
    const float l0 = vnorm(z0), l1 = vnorm(z1);
    const V3 bcr = orc::cross(z0, z1);
    const float lb = vnorm(bcr);
    if (l0 <= 1e-9f || l1 <= 1e-9f || lb <= 1e-9f * l0 * l1) {
        return;
    }
    const V3 tb = vdiv(bcr, lb);
    const float Sgeom = (orc::dot(tb, t1) >= 0.0f) ? 1.0f : -1.0f;
    float Hp[12][12];
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) Hp[i][j] = 0.0f;
    const V3 e0p = orc::cross(z0, tb), e1p = orc::cross(z1, tb);
    const float S3 = -Sgeom;
    const float w = S3 * gdiff;
    const float gamma = (l1 * l1) / (l0 * l0);
    const float gm1 = gamma - 1.0f, gp1 = gamma + 1.0f;
    const float r = std::sqrt(4.0f * w * w * (gm1 / gp1) * (gm1 / gp1) + 1.0f);
    const float Rm = std::sqrt(std::fmax(0.0f, 2.0f * (2.0f * w * w + 1.0f - r)));
    const float Rp = std::sqrt(std::fmax(0.0f, 2.0f * (2.0f * w * w + 1.0f + r)));
    const float Lh = 1.0f / (l0 * l0) + 1.0f / (l1 * l1);
    const float cc = 4.0f * w / gp1;
    // eta = (-1/ly1) * (t1.y * sigma + t1 z^T)
    auto eta = [&](const V3 &y, const V3 &z) {
        return mscale(-1.0f / ly1, madd(mscale(orc::dot(t1, y), sigma), outer(t1, z)));
    };
    const orc::M3 eta2 = eta(y2, z0), eta3 = eta(y3, z1);
    const orc::M3 e2T = mtrans(eta2), e3T = mtrans(eta3);
    const orc::M3 m0 = msub(mscale(-1.0f, e2T), sigma);
    const orc::M3 m1 = msub(mscale(-1.0f, e3T), sigma);
    auto accF = [&](float lam, const V3 &qa, const V3 &qb) {
        if (lam <= 0.0f) {
            return;
        }
        const float n2 = orc::dot(qa, qa) + orc::dot(qb, qb);
        if (n2 <= 1e-20f) {
            return;
        }
        const V3 b0 = orc::add(mv(m0, qa), mv(m1, qb));
        const V3 b1 = orc::add(mv(e2T, qa), mv(e3T, qb));
        const V3 b2 = mv(sigma, qa);
        const V3 b3 = mv(sigma, qb);
        float p12[12];
        for (int c = 0; c < 3; ++c) {
            p12[c] = b0[c];
            p12[3 + c] = b1[c];
            p12[6 + c] = b2[c];
            p12[9 + c] = b3[c];
        }
        const float s = lam / n2;
        for (int i = 0; i < 12; ++i)
            for (int j = 0; j < 12; ++j) Hp[i][j] += s * p12[i] * p12[j];
    };
    const int sr[4] = {-1, -1, 1, 1};
    const float RR[4] = {Rm, Rm, Rp, Rp};
    const int sR[4] = {-1, 1, -1, 1};
    for (int mm = 0; mm < 4; ++mm) {
        const float R = RR[mm];
        const float a = w * (gm1 * gm1 / gp1 + sr[mm] * gp1 * r + sR[mm] * gm1 * R);
        const float b = -0.5f * gamma * R * R + 2.0f * w * w -
                        sR[mm] * 0.5f * (gm1 + sr[mm] * gp1 * r) * R;
        const float d = 1.0f + sr[mm] * r + sR[mm] * R;
        accF(0.25f * Lh * d, orc::add(orc::scale(a, z0), orc::scale(b, e0p)),
             orc::add(orc::scale(cc, z1), orc::scale(d, e1p)));
    }
    const float cos_t = orc::dot(z0, z1) / (l0 * l1);
    const float sin_t = lb / (l0 * l1);
    const float beta = (l1 / l0 - l0 / l1) * cos_t;
    const float alpha0 = 0.5f * (-beta + std::sqrt(beta * beta + 4.0f));
    const float ginv = w / sin_t;
    accF(ginv * (cos_t / (l1 * l1) - alpha0 / (l0 * l1)), orc::scale(alpha0, tb), tb);
    accF(ginv * (cos_t / (l1 * l1) + 1.0f / (alpha0 * l0 * l1)), tb,
         orc::scale(-alpha0, tb));

    const V3 tau0 = vdiv(z0, l0), tau1 = vdiv(z1, l1);
    const V3 tau1p = orc::cross(tau1, tb);
    const float a0 = orc::dot(t1, y2) / (ly1 * l0);
    const float a1 = orc::dot(t1, y3) / (ly1 * l1);
    const float c0 = 1.0f / l0, c1 = 1.0f / l1;
    const float k0 = -4.0f * orc::dot(tau0, tau1p) / (ly1 * ly1);
    const float k1 = 2.0f * orc::dot(tau0, tau1p) / ly1;
    const V3 s01 = orc::add(tau0, tau1), d01 = orc::sub(tau0, tau1);
    float pc[4][12];
    for (int c = 0; c < 3; ++c) {
        pc[0][c] = s01[c], pc[0][3 + c] = -s01[c], pc[0][6 + c] = 0.0f, pc[0][9 + c] = 0.0f;
        pc[1][c] = d01[c], pc[1][3 + c] = -d01[c], pc[1][6 + c] = 0.0f, pc[1][9 + c] = 0.0f;
        pc[2][c] = (-a1 - a0 + c1 + c0) * t1[c];
        pc[2][3 + c] = (a1 + a0) * t1[c];
        pc[2][6 + c] = -c0 * t1[c];
        pc[2][9 + c] = -c1 * t1[c];
        pc[3][c] = (-a1 + a0 + c1 - c0) * t1[c];
        pc[3][3 + c] = (a1 - a0) * t1[c];
        pc[3][6 + c] = c0 * t1[c];
        pc[3][9 + c] = -c1 * t1[c];
    }
    float gg[4], g23 = 0.0f;
    for (int i = 0; i < 4; ++i) {
        float s = 0.0f;
        for (int m = 0; m < 12; ++m) s += pc[i][m] * pc[i][m];
        gg[i] = s;
    }
    for (int m = 0; m < 12; ++m) g23 += pc[2][m] * pc[3][m];
    const float det23 = gg[2] * gg[3] - g23 * g23;
    if (gg[0] > 1e-5f && gg[1] > 1e-5f && det23 > 1e-6f * gg[2] * gg[3]) {
        const float we = -gdiff * Sgeom;
        float E[4][4] = {{0.0f}};
        E[0][0] = we * (-k0 * gg[0]);
        E[1][1] = we * (k0 * gg[1]);
        E[0][2] = we * (k1 * gg[2]);
        E[2][0] = E[0][2];
        E[0][3] = we * (k1 * g23);
        E[3][0] = E[0][3];
        E[1][2] = we * (k1 * g23);
        E[2][1] = E[1][2];
        E[1][3] = we * (k1 * gg[3]);
        E[3][1] = E[1][3];
        float Li[4][4] = {{0.0f}};
        Li[0][0] = 1.0f / std::sqrt(gg[0]);
        Li[1][1] = 1.0f / std::sqrt(gg[1]);
        const float L22 = std::sqrt(gg[2]);
        const float L32 = g23 / L22;
        const float L33 = std::sqrt(std::fmax(1e-30f, gg[3] - L32 * L32));
        Li[2][2] = 1.0f / L22;
        Li[3][3] = 1.0f / L33;
        Li[3][2] = -L32 / (L22 * L33);
        float LiE[4][4], Ss[4][4], LtS[4][4], Core[4][4];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                float s = 0.0f;
                for (int k = 0; k < 4; ++k) s += Li[i][k] * E[k][j];
                LiE[i][j] = s;
            }
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                float s = 0.0f;
                for (int k = 0; k < 4; ++k) s += LiE[i][k] * Li[j][k];
                Ss[i][j] = s;
            }
        psd_project4(Ss);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                float s = 0.0f;
                for (int k = 0; k < 4; ++k) s += Li[k][i] * Ss[k][j];
                LtS[i][j] = s;
            }
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                float s = 0.0f;
                for (int k = 0; k < 4; ++k) s += LtS[i][k] * Li[k][j];
                Core[i][j] = s;
            }
        float PC[12][4];
        for (int m = 0; m < 12; ++m)
            for (int j = 0; j < 4; ++j) {
                float s = 0.0f;
                for (int i = 0; i < 4; ++i) s += pc[i][m] * Core[i][j];
                PC[m][j] = s;
            }
        for (int m = 0; m < 12; ++m)
            for (int n = 0; n < 12; ++n) {
                float s = 0.0f;
                for (int j = 0; j < 4; ++j) s += PC[m][j] * pc[j][n];
                Hp[m][n] += s;
            }
    }
    const int perm[4] = {2, 1, 0, 3};
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    H[3 * perm[a] + i][3 * perm[b] + j] = Hp[3 * a + i][3 * b + j];
}

} // namespace

extern "C" {

// Shell dataset of the elastic assembly. Positions: x (eval_x, the Newton
// iterate), x0 (vertex.curr), target (inertial target y). Faces / hinges as
// data.mesh.mesh (hinge = 4 vertices in the stored order; the bending term
// remaps it to (h2, h1, h0, h3)). v2f_off/v2f: vertex->face CSR in the order
// the vertex normal sums. Per vertex: mass, area, flag bit0 = pinned. Per
// face: mass, inv_rest2x2 (4 floats, column-major), mu, lambda, flag bit0 =
// excluded (fixed / collider / rest-excluded). Per hinge: rest angle, edge
// length, rest area of the two faces, bend (resolved, directional weight
// included), flag bit0 = excluded.
struct oracle_shell {
    uint32_t nv, nf, nh;
    const float *x, *x0, *target;
    const uint32_t *face, *hinge, *v2f_off, *v2f;
    const float *vmass, *varea;
    const uint32_t *vflag;
    const float *fmass, *inv_rest, *f_mu, *f_lambda;
    const uint32_t *fflag;
    const float *h_rest, *h_length, *h_area, *h_bend;
    const uint32_t *hflag;
    const float *f_strainlimit, *f_shrink;
    const float *f_plasticity, *f_plastic_threshold;
    const float *h_plasticity, *h_plastic_threshold;
};

struct oracle_elastic_params {
    float dt, air_density, air_friction;
    float wind[3];
};

// Element-local terms, for pinning against the reference's own code:
// kind 0 stretch gradient (6) + Hessian (36) of F (6, column-major) with
// s0 = mu; kind 1 shear gradient + Hessian with s0 = lambda; kind 2 the
// unit-stiffness hinge force (12) + Hessian (144 row-major) at the 4 points
// (12 floats, already remapped) with s0 = rest angle.
void oracle_elastic_local(int kind, const float *in, float s0, float *grad, float *hess) {
    if (kind == 0 || kind == 1) {
        M32 F;
        for (int i = 0; i < 6; ++i) F.m[i] = in[i];
        const M32 g = kind == 0 ? stretch_gradient(F, s0) : shear_gradient(F, s0);
        const M66 h = kind == 0 ? stretch_hessian(F, s0) : shear_hessian(F, s0);
        std::memcpy(grad, g.m, sizeof(g.m));
        std::memcpy(hess, h.m, sizeof(h.m));
        return;
    }
    V3 X[4];
    for (int k = 0; k < 4; ++k) X[k] = v3(in[3 * k], in[3 * k + 1], in[3 * k + 2]);
    float H[12][12];
    hinge_unit(X[0], X[1], X[2], X[3], s0, grad, H);
    std::memcpy(hess, H, sizeof(H));
}

float oracle_dihedral_angle(const float *p) {
    return dihedral_angle(v3(p[0], p[1], p[2]), v3(p[3], p[4], p[5]), v3(p[6], p[7], p[8]),
                          v3(p[9], p[10], p[11]));
}

// Elastic + inertia assembly. Writes force (3 nv), the per-vertex diagonal
// cdiag (9 nv, column-major) and the upper-block values `val` of the given
// pattern (row offsets rp, columns ci; nnz blocks). Returns the number of
// element blocks that found no slot in the pattern (0 for a complete one).
uint32_t oracle_elastic_assemble(const oracle_shell *s, const oracle_elastic_params *p,
                                 const uint32_t *rp, const uint32_t *ci, float *force,
                                 float *cdiag, float *val) {
    const uint32_t nv = s->nv;
    auto X = [&](const float *a, uint32_t i) { return v3(a[3 * i], a[3 * i + 1], a[3 * i + 2]); };
    Params pr{p->dt, p->air_density, p->air_friction, {p->wind[0], p->wind[1], p->wind[2]}};
    const size_t nnz = rp[nv];
    std::memset(val, 0, nnz * 9 * sizeof(float));
    std::vector<float> f(3 * (size_t)nv, 0.0f);
    for (uint32_t i = 0; i < nv; ++i) {
        for (int e = 0; e < 9; ++e) cdiag[9 * (size_t)i + e] = 0.0f;
        if (s->vflag[i] & 1u) {
            continue;
        }
        V3 n = v3(0, 0, 0);
        for (uint32_t k = s->v2f_off[i]; k < s->v2f_off[i + 1]; ++k) {
            const uint32_t *fc = s->face + 3 * (size_t)s->v2f[k];
            const V3 z0 = X(s->x, fc[0]), z1 = X(s->x, fc[1]), z2 = X(s->x, fc[2]);
            n = orc::add(n, orc::cross(orc::sub(z1, z0), orc::sub(z2, z0)));
        }
        if (orc::dot(n, n) != 0.0f) {
            normalize3(n);
        }
        V3 fv;
        orc::M3 hv;
        vertex_term(X(s->x, i), X(s->x0, i), X(s->target, i), n, s->vmass[i], s->varea[i], pr,
                    fv, hv);
        for (int d = 0; d < 3; ++d) f[3 * (size_t)i + d] = 0.0f + fv[d];
        for (int e = 0; e < 9; ++e) cdiag[9 * (size_t)i + e] = 0.0f + hv.m[e];
    }
    uint32_t missing = 0;
    auto push = [&](uint32_t r, uint32_t c, const float blk[9]) {
        if (r > c) {
            return;
        }
        for (uint32_t k = rp[r]; k < rp[r + 1]; ++k) {
            if (ci[k] == c) {
                for (int e = 0; e < 9; ++e) {
                    if (blk[e] != 0.0f) {
                        val[9 * (size_t)k + e] += blk[e];
                    }
                }
                return;
            }
        }
        ++missing;
    };
    // Element contributions are staged so that every block and every force
    // entry sums faces first, then hinges, each by element index.
    std::vector<float> fe(9 * (size_t)s->nf), he(81 * (size_t)s->nf);
    std::vector<char> fon(s->nf, 0);
    for (uint32_t k = 0; k < s->nf; ++k) {
        if (s->fflag[k] & 1u) continue;
        const uint32_t *fc = s->face + 3 * (size_t)k;
        if (!(s->f_mu[k] > 0.0f)) continue;
        face_term(X(s->x, fc[0]), X(s->x, fc[1]), X(s->x, fc[2]), s->inv_rest + 4 * (size_t)k,
                  s->fmass[k], s->f_mu[k], s->f_lambda[k], &fe[9 * (size_t)k],
                  &he[81 * (size_t)k]);
        fon[k] = 1;
    }
    std::vector<float> hf(12 * (size_t)s->nh);
    std::vector<float> hh(144 * (size_t)s->nh);
    std::vector<char> hon(s->nh, 0);
    for (uint32_t k = 0; k < s->nh; ++k) {
        if (s->hflag[k] & 1u) continue;
        const uint32_t *h = s->hinge + 4 * (size_t)k;
        float areal = 0.0f;
        int cnt = 0;
        for (int q = 0; q < 4; ++q) {
            const float a = s->varea[h[q]];
            if (a > 0.0f) {
                areal += s->vmass[h[q]] / a;
                ++cnt;
            }
        }
        if (cnt > 0) areal /= cnt;
        const float length = s->h_length[k], area = s->h_area[k];
        const float stiff = (area > 1e-12f)
                                ? 1.28e-5f * s->h_bend[k] * (length * length / area) * areal
                                : 0.0f;
        if (!(stiff > 0.0f)) continue;
        const uint32_t r[4] = {h[2], h[1], h[0], h[3]};
        float g12[12], H[12][12];
        hinge_unit(X(s->x, r[0]), X(s->x, r[1]), X(s->x, r[2]), X(s->x, r[3]), s->h_rest[k],
                   g12, H);
        for (int i = 0; i < 12; ++i) hf[12 * (size_t)k + i] = g12[i] * stiff;
        for (int i = 0; i < 12; ++i)
            for (int j = 0; j < 12; ++j) hh[144 * (size_t)k + 12 * i + j] = H[i][j] * stiff;
        hon[k] = 1;
    }
    for (uint32_t k = 0; k < s->nf; ++k) {
        if (!fon[k]) continue;
        const uint32_t *fc = s->face + 3 * (size_t)k;
        for (int a = 0; a < 3; ++a)
            for (int d = 0; d < 3; ++d) f[3 * (size_t)fc[a] + d] += fe[9 * (size_t)k + 3 * a + d];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                float blk[9];
                for (int c = 0; c < 3; ++c)
                    for (int rr = 0; rr < 3; ++rr)
                        blk[rr + 3 * c] = he[81 * (size_t)k + (3 * a + rr) + 9 * (3 * b + c)];
                push(fc[a], fc[b], blk);
            }
    }
    for (uint32_t k = 0; k < s->nh; ++k) {
        if (!hon[k]) continue;
        const uint32_t *h = s->hinge + 4 * (size_t)k;
        const uint32_t r[4] = {h[2], h[1], h[0], h[3]};
        for (int a = 0; a < 4; ++a)
            for (int d = 0; d < 3; ++d) f[3 * (size_t)r[a] + d] += hf[12 * (size_t)k + 3 * a + d];
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                float blk[9];
                for (int c = 0; c < 3; ++c)
                    for (int rr = 0; rr < 3; ++rr)
                        blk[rr + 3 * c] = hh[144 * (size_t)k + 12 * (3 * a + rr) + (3 * b + c)];
                push(r[a], r[b], blk);
            }
    }
    std::memcpy(force, f.data(), f.size() * sizeof(float));
    return missing;
}
}
