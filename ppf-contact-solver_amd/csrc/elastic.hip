// elastic.hip -- inertia, air drag, membrane and bending assembly on MI355X.
//
// Replaces energy::embed_momentum_force_hessian and the shell part of
// energy::embed_elastic_force_hessian (energy/energy.cu:182-270, 304-393,
// 448-553, 700-794):
//   * per unpinned vertex: air damper on the area-weighted vertex normal
//     (air_damper.hpp:13-42, utility.cu:15-32) and momentum m (y - target) /
//     dt^2 with m I / dt^2 into the per-vertex diagonal (momentum.hpp:14-22);
//   * per face: Baraff-Witkin stretch + shear on F = dX inv_rest, mass
//     scaled and mapped to the vertices (baraffwitkin.hpp:27-226,
//     utility.cu:185-282);
//   * per hinge: areal-density-normalized Discrete Shells stiffness times the
//     dihedral-angle force and the Wu-Kim projected Hessian
//     (dihedral_angle.hpp:15-355, eigsolve.hpp:245-316).
// Elements write their upper blocks to staging; one gather per stored block
// and per vertex then sums the contributions in a fixed order (faces by
// index, then hinges by index; zero entries skipped as FixedCSRMat::push,
// csrmat.cu:544-594). The sums are therefore deterministic and reproduce a
// sequential element loop exactly; the reference folds them with float
// atomics in arbitrary order.
#include "elastic.hpp"
#include "scan.hpp"
#include "sort.hpp"

namespace ppf {

namespace {

constexpr int kNT = 256;
constexpr uint32_t kFacePairs = 6;   // (0,0) (1,1) (2,2) (0,1) (0,2) (1,2)
constexpr uint32_t kHingePairs = 10; // 4 diagonal + 6 off-diagonal

__device__ __constant__ uint8_t kFP[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
__device__ __constant__ uint8_t kHP[10][2] = {{0, 0}, {1, 1}, {2, 2}, {3, 3}, {0, 1},
                                              {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};

uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// ---- small dense helpers: column-major, products accumulated from zero -------

struct M3 {
    float m[9];
};

__device__ inline M3 m_zero() {
    M3 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.m[i] = 0.0f;
    return r;
}
__device__ inline M3 m_eye() {
    M3 r = m_zero();
    r.m[0] = r.m[4] = r.m[8] = 1.0f;
    return r;
}
// a b^T: every entry 0 + a_r b_c (exact: a single product).
__device__ inline M3 m_outer(f3 a, f3 b) {
    const float av[3] = {a.x, a.y, a.z}, bv[3] = {b.x, b.y, b.z};
    M3 r;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int rr = 0; rr < 3; ++rr) r.m[rr + 3 * c] = av[rr] * bv[c];
    return r;
}
__device__ inline M3 m_scale(float s, const M3 &a) {
    M3 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.m[i] = s * a.m[i];
    return r;
}
__device__ inline M3 m_div(const M3 &a, float s) {
    M3 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.m[i] = a.m[i] / s;
    return r;
}
__device__ inline M3 m_add(const M3 &a, const M3 &b) {
    M3 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.m[i] = a.m[i] + b.m[i];
    return r;
}
__device__ inline M3 m_sub(const M3 &a, const M3 &b) {
    M3 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.m[i] = a.m[i] - b.m[i];
    return r;
}
__device__ inline M3 m_trans(const M3 &a) {
    M3 r;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int rr = 0; rr < 3; ++rr) r.m[rr + 3 * c] = a.m[c + 3 * rr];
    return r;
}
__device__ inline f3 m_mul(const M3 &a, f3 x) {
    float r[3] = {0.0f, 0.0f, 0.0f};
    const float xv[3] = {x.x, x.y, x.z};
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int rr = 0; rr < 3; ++rr) r[rr] = r[rr] + a.m[rr + 3 * k] * xv[k];
    return mk3(r[0], r[1], r[2]);
}
__device__ inline f3 v_div(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
__device__ inline f3 v_mul(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ inline float v_norm(f3 a) { return sqrtf(dot3(a, a)); }
__device__ inline void v_normalize(f3 &a) {
    const float n = v_norm(a);
    if (n > 0.0f) {
        a = v_div(a, n);
    }
}

__device__ inline f3 X(const float *p, uint32_t i) { return load3(p + 3 * (size_t)i); }

// ---- per-vertex inertia + air drag ------------------------------------------

__global__ __launch_bounds__(kNT) void k_vertex(ppf_shell sh, ppf_elastic_params p,
                                                float *__restrict__ vf,
                                                float *__restrict__ cdiag) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i >= sh.nv) {
        return;
    }
    float *fo = vf + 3 * (size_t)i;
    float *ho = cdiag + 9 * (size_t)i;
    if (sh.vflag[i] & 1u) {
        fo[0] = fo[1] = fo[2] = 0.0f;
        for (int e = 0; e < 9; ++e) ho[e] = 0.0f;
        return;
    }
    f3 n = mk3(0.0f, 0.0f, 0.0f);
    for (uint32_t k = sh.v2f_off[i]; k < sh.v2f_off[i + 1]; ++k) {
        const uint32_t fi = sh.v2f[k];
        const f3 z0 = X(sh.x, sh.face[3 * (size_t)fi]);
        const f3 z1 = X(sh.x, sh.face[3 * (size_t)fi + 1]);
        const f3 z2 = X(sh.x, sh.face[3 * (size_t)fi + 2]);
        n = n + cross3(z1 - z0, z2 - z0);
    }
    if (dot3(n, n) != 0.0f) {
        v_normalize(n);
    }
    const f3 y = X(sh.x, i), x = X(sh.x0, i), tg = X(sh.target, i);
    const float dt = (p.scene_dt && p.vscene) ? p.scene_dt[p.vscene[i]] : p.dt;
    const float dt2 = dt * dt;
    f3 f = mk3(0.0f, 0.0f, 0.0f);
    M3 H = m_zero();
    const bool nz = !(fabsf(n.x) <= 1e-5f) || !(fabsf(n.y) <= 1e-5f) || !(fabsf(n.z) <= 1e-5f);
    if (nz && p.air_density != 0.0f) {
        const f3 wind = mk3(p.wind[0], p.wind[1], p.wind[2]);
        const f3 z = (y - x) - dt * wind;
        const M3 P = m_sub(m_eye(), m_outer(n, n));
        const f3 g = v_div(v_mul(n, dot3(n, z)), dt2) + v_div(m_mul(m_scale(p.air_friction, P), z), dt2);
        const M3 h = m_add(m_div(m_outer(n, n), dt2), m_div(m_scale(p.air_friction, P), dt2));
        const float s = sh.varea[i] * p.air_density;
        f = f + s * g;
        H = m_add(H, m_scale(s, h));
    }
    const float mass = sh.vmass[i];
    f = f + mass * v_div(y - tg, dt2);
    H = m_add(H, m_scale(mass, m_div(m_eye(), dt2)));
    fo[0] = 0.0f + f.x, fo[1] = 0.0f + f.y, fo[2] = 0.0f + f.z;
    for (int e = 0; e < 9; ++e) ho[e] = 0.0f + H.m[e];
}

// ---- Baraff-Witkin ------------------------------------------------------------

// max(0, l) * q * q^T parses as ((max(0, l) * q) * q^T).
__device__ M3 stretch_block(f3 q, float nq) {
    const f3 qa = cross3(mk3(q.z, 0.0f, -q.x), q);
    const f3 qb = cross3(mk3(0.0f, q.z, -q.y), q);
    const f3 qc = cross3(mk3(q.y, -q.x, 0.0f), q);
    const float s0 = dot3(qa, qa), s1 = dot3(qb, qb), s2 = dot3(qc, qc);
    f3 q2;
    if (s0 > s1 && s0 > s2) {
        q2 = qa;
    } else if (s1 > s0 && s1 > s2) {
        q2 = qb;
    } else {
        q2 = qc;
    }
    f3 q1 = q;
    f3 q3 = cross3(q1, q2);
    v_normalize(q1);
    v_normalize(q2);
    v_normalize(q3);
    const float l2 = 1.0f - 1.0f / nq;
    M3 H = m_zero();
    H = m_add(H, m_outer(fmaxf(0.0f, 1.0f) * q1, q1));
    H = m_add(H, m_outer(fmaxf(0.0f, l2) * q2, q2));
    H = m_add(H, m_outer(fmaxf(0.0f, l2) * q3, q3));
    return H;
}

struct V6 {
    float v[6];
};
__device__ inline float d6(const V6 &a, const V6 &b) {
    float acc = 0.0f;
#pragma unroll
    for (int i = 0; i < 6; ++i) acc = acc + a.v[i] * b.v[i];
    return acc;
}
__device__ inline V6 s6(float s, const V6 &a) {
    V6 r;
#pragma unroll
    for (int i = 0; i < 6; ++i) r.v[i] = s * a.v[i];
    return r;
}
__device__ inline V6 a6(const V6 &a, const V6 &b) {
    V6 r;
#pragma unroll
    for (int i = 0; i < 6; ++i) r.v[i] = a.v[i] + b.v[i];
    return r;
}
__device__ inline V6 m6(const V6 &a, const V6 &b) {
    V6 r;
#pragma unroll
    for (int i = 0; i < 6; ++i) r.v[i] = a.v[i] - b.v[i];
    return r;
}
__device__ inline void n6(V6 &a) {
    const float n = sqrtf(d6(a, a));
    if (n > 0.0f) {
#pragma unroll
        for (int i = 0; i < 6; ++i) a.v[i] = a.v[i] / n;
    }
}

// F: column-major 3x2 (F[r + 3c]). Writes dE/dF (6) and d2E/dF2 (36,
// column-major) of stretch(mu) + shear(lambda).
__device__ void bw_terms(const float F[6], float mu, float lmd, float g[6], float h[36]) {
    const f3 fu = mk3(F[0], F[1], F[2]), fv = mk3(F[3], F[4], F[5]);
    // stretch gradient
    const float nu = v_norm(fu), nv = v_norm(fv);
    const f3 ga = v_div(v_mul(fu, nu - 1.0f), nu), gb = v_div(v_mul(fv, nv - 1.0f), nv);
    // shear gradient
    const float w = dot3(fu, fv);
    const float sg[6] = {lmd * (w * F[3]), lmd * (w * F[4]), lmd * (w * F[5]),
                         lmd * (w * F[0]), lmd * (w * F[1]), lmd * (w * F[2])};
    const float st[6] = {mu * ga.x, mu * ga.y, mu * ga.z, mu * gb.x, mu * gb.y, mu * gb.z};
#pragma unroll
    for (int i = 0; i < 6; ++i) g[i] = st[i] + sg[i];
    // stretch Hessian (block diagonal, scaled by mu)
    const M3 H1 = stretch_block(fu, nu), H2 = stretch_block(fv, nv);
    float hs[36];
#pragma unroll
    for (int i = 0; i < 36; ++i) hs[i] = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            hs[r + 6 * c] = H1.m[r + 3 * c];
            hs[(3 + r) + 6 * (3 + c)] = H2.m[r + 3 * c];
        }
#pragma unroll
    for (int i = 0; i < 36; ++i) hs[i] = mu * hs[i];
    // shear Hessian
    const float I6 = w;
    float I2 = 0.0f;
#pragma unroll
    for (int i = 0; i < 6; ++i) I2 = I2 + F[i] * F[i];
    V6 gv, Hg;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        gv.v[d] = F[3 + d];
        gv.v[3 + d] = F[d];
        Hg.v[d] = F[d];
        Hg.v[3 + d] = F[3 + d];
    }
    const float root = sqrtf(I2 * I2 + 12.0f * I6 * I6);
    const float lmd0 = I2 + root;
    V6 q0 = a6(s6(2.0f * I6, Hg), s6(lmd0, gv));
    V6 z, w1, w2, w3;
#pragma unroll
    for (int i = 0; i < 6; ++i) w1.v[i] = w2.v[i] = w3.v[i] = 0.0f;
    float lmd1;
    if (I6 > 0) {
        z = q0;
        lmd1 = 2.0f * I6;
        w1.v[0] = 1, w1.v[1] = 1;
        w2.v[2] = 1, w2.v[3] = 1;
        w3.v[4] = 1, w3.v[5] = 1;
    } else {
        lmd1 = -2.0f * I6;
        const float tmp = I2 - root;
        z = a6(s6(2.0f * I6, Hg), s6(tmp, gv));
        w1.v[0] = 1, w1.v[1] = -1;
        w2.v[2] = 1, w2.v[3] = -1;
        w3.v[4] = 1, w3.v[5] = -1;
    }
    const float a1 = d6(z, w1), a2 = d6(z, w2), a3 = d6(z, w3);
    const V6 q1_1 = m6(a6(s6(a3, w1), s6(a3, w2)), s6(a1 - a2, w3));
    const V6 q1_2 = a6(m6(s6(a2, w1), s6(a1 - a3, w2)), s6(a2, w3));
    const V6 q1_3 = a6(a6(s6(-(a2 - a3), w1), s6(a1, w2)), s6(a1, w3));
    const float n10 = d6(q1_1, q1_1), n11 = d6(q1_2, q1_2), n12 = d6(q1_3, q1_3);
    V6 q1;
    if (n10 > n11 && n10 > n12) {
        q1 = q1_1;
    } else if (n11 > n10 && n11 > n12) {
        q1 = q1_2;
    } else {
        q1 = q1_3;
    }
    const float b1 = d6(q1, w1), b2 = d6(q1, w2), b3 = d6(q1, w3);
    // [[p, q], [r, s]]: det p s - q r; adj(A) x from zero.
    const float det1 = a2 * b3 - a3 * b2, det2 = a3 * b1 - a1 * b3, det3 = a1 * b2 - a2 * b1;
    const float c45_0 = (0.0f + b2 * (-a3)) + (-a2) * (-b3);
    const float c45_1 = (0.0f + (-b1) * (-a3)) + a1 * (-b3);
    const float c56_0 = (0.0f + b3 * (-a1)) + (-a3) * (-b1);
    const float c56_1 = (0.0f + (-b2) * (-a1)) + a2 * (-b1);
    const float c64_0 = (0.0f + b1 * (-a2)) + (-a1) * (-b2);
    const float c64_1 = (0.0f + (-b3) * (-a2)) + a3 * (-b2);
    const V6 q2_1 = a6(a6(s6(c45_0, w1), s6(c45_1, w2)), s6(det3, w3));
    const V6 q2_2 = a6(a6(s6(det1, w1), s6(c56_0, w2)), s6(c56_1, w3));
    const V6 q2_3 = a6(a6(s6(c64_1, w1), s6(det2, w2)), s6(c64_0, w3));
    const float n20 = d6(q2_1, q2_1), n21 = d6(q2_2, q2_2), n22 = d6(q2_3, q2_3);
    V6 q2;
    if (n20 > n21 && n20 > n22) {
        q2 = q2_1;
    } else if (n21 > n20 && n21 > n22) {
        q2 = q2_2;
    } else {
        q2 = q2_3;
    }
    n6(q0);
    n6(q1);
    n6(q2);
    float res[36];
#pragma unroll
    for (int i = 0; i < 36; ++i) res[i] = 0.0f;
    const float l0p = fmaxf(0.0f, lmd0), l1p = fmaxf(0.0f, lmd1);
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            res[r + 6 * c] = res[r + 6 * c] + l0p * (q0.v[r] * q0.v[c]);
            res[r + 6 * c] = res[r + 6 * c] + l1p * (q1.v[r] * q1.v[c]);
            res[r + 6 * c] = res[r + 6 * c] + l1p * (q2.v[r] * q2.v[c]);
        }
    const float half = 0.5f * lmd;
#pragma unroll
    for (int i = 0; i < 36; ++i) h[i] = hs[i] + half * res[i];
}

__global__ __launch_bounds__(kNT) void k_face(ppf_shell sh, float *__restrict__ fblk,
                                              float *__restrict__ ff,
                                              uint8_t *__restrict__ active) {
    const uint32_t k = blockIdx.x * kNT + threadIdx.x;
    if (k >= sh.nf) {
        return;
    }
    const bool on = !(sh.fflag[k] & 1u) && sh.f_mu[k] > 0.0f;
    active[k] = on ? 1 : 0;
    if (!on) {
        return;
    }
    const uint32_t v[3] = {sh.face[3 * (size_t)k], sh.face[3 * (size_t)k + 1],
                           sh.face[3 * (size_t)k + 2]};
    const f3 x0 = X(sh.x, v[0]), x1 = X(sh.x, v[1]), x2 = X(sh.x, v[2]);
    const float *inv = sh.inv_rest + 4 * (size_t)k;
    const float i0 = inv[0], i1 = inv[1], i2 = inv[2], i3 = inv[3];
    const f3 d0 = x1 - x0, d1 = x2 - x0;
    // F = dX inv_rest (3x2 * 2x2, accumulated from zero)
    const float F[6] = {(0.0f + d0.x * i0) + d1.x * i1, (0.0f + d0.y * i0) + d1.y * i1,
                        (0.0f + d0.z * i0) + d1.z * i1, (0.0f + d0.x * i2) + d1.x * i3,
                        (0.0f + d0.y * i2) + d1.y * i3, (0.0f + d0.z * i2) + d1.z * i3};
    float g[6], h[36];
    bw_terms(F, sh.f_mu[k], sh.f_lambda[k], g, h);
    const float mass = sh.fmass[k];
    const float gg[3][2] = {{-i0 - i1, -i2 - i3}, {i0, i2}, {i1, i3}};
    float *fo = ff + 9 * (size_t)k;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float fe = (0.0f + gg[a][0] * g[d]) + gg[a][1] * g[3 + d];
            fo[3 * a + d] = 0.0f + mass * fe;
        }
    float *bo = fblk + 9 * kFacePairs * (size_t)k;
    for (uint32_t p = 0; p < kFacePairs; ++p) {
        // block (a, b) for the stored (lower vertex index, higher) order
        uint32_t a = kFP[p][0], b = kFP[p][1];
        if (v[a] > v[b]) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        float blk[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) blk[e] = 0.0f;
#pragma unroll
        for (int d = 0; d < 2; ++d)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const float s = gg[a][d] * gg[b][e];
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int r = 0; r < 3; ++r)
                        blk[r + 3 * c] = blk[r + 3 * c] + s * h[(3 * d + r) + 6 * (3 * e + c)];
            }
#pragma unroll
        for (int e = 0; e < 9; ++e) bo[9 * p + e] = 0.0f + mass * blk[e];
    }
}

// ---- strain limiting --------------------------------------------------------------
//
// strainlimiting.cu:17-100 (the barrier on the stretch beyond rest of each
// singular value of F, scaled by an elastic-inclusive stiffness) and 243-305
// (its line search), with the one-rotation Jacobi SVD of utility.cu:43-130,
// the cubic barrier of cubic.hpp:13-40 and the 3x2 eigen-analysis of
// eigenanalysis.cu:18-78. The strain-limit curvature table is diagonal, so
// its two scaling modes are the singular directions themselves.

// F = U diag(S) V^T of a 3x2 F (column-major), singular values ascending.
struct Svd32 {
    float U[6]; // 3x2, column-major
    float S[2];
    float V[4]; // 2x2, column-major
};

__device__ inline Svd32 svd32(const float F[6]) {
    float A[3][2] = {{F[0], F[3]}, {F[1], F[4]}, {F[2], F[5]}};
    float W[2][2] = {{1.0f, 0.0f}, {0.0f, 1.0f}};
    float a = 0.0f, b = 0.0f, g = 0.0f;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        a += A[r][0] * A[r][0];
        b += A[r][1] * A[r][1];
        g += A[r][0] * A[r][1];
    }
    // the rotation that makes the two columns orthogonal (identity if they are)
    const float zeta = (b - a) / (2.0f * g);
    const float t = (g == 0.0f) ? 0.0f
                                : (zeta >= 0.0f ? 1.0f : -1.0f) /
                                      (fabsf(zeta) + sqrtf(1.0f + zeta * zeta));
    const float c = 1.0f / sqrtf(1.0f + t * t), sn = c * t;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const float ap = A[r][0], aq = A[r][1];
        A[r][0] = c * ap - sn * aq;
        A[r][1] = sn * ap + c * aq;
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const float vp = W[r][0], vq = W[r][1];
        W[r][0] = c * vp - sn * vq;
        W[r][1] = sn * vp + c * vq;
    }
    float sv[2], u[3][2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        float n2 = 0.0f;
#pragma unroll
        for (int r = 0; r < 3; ++r) n2 += A[r][k] * A[r][k];
        const float nrm = sqrtf(n2);
        sv[k] = nrm;
        const float inv = nrm > 0.0f ? 1.0f / nrm : 0.0f;
#pragma unroll
        for (int r = 0; r < 3; ++r) u[r][k] = A[r][k] * inv;
    }
    const int i0 = sv[1] < sv[0] ? 1 : 0, i1 = 1 - i0;
    Svd32 out;
    out.S[0] = sv[i0], out.S[1] = sv[i1];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        out.U[r] = u[r][i0];
        out.U[3 + r] = u[r][i1];
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        out.V[r] = W[r][i0];
        out.V[2 + r] = W[r][i1];
    }
    return out;
}

__device__ inline float sl_limit(const ppf_shell &sh, uint32_t k) {
    const float shrink = sh.f_shrink ? sh.f_shrink[k] : 1.0f;
    return (sh.f_strainlimit[k] + 1.0f) / shrink - 1.0f;
}

// F = dX inv_rest of face k at positions x (column-major, as k_face).
__device__ inline void face_F(const float *x, const uint32_t v[3], const float *inv, float F[6]) {
    const f3 x0 = X(x, v[0]), x1 = X(x, v[1]), x2 = X(x, v[2]);
    const float i0 = inv[0], i1 = inv[1], i2 = inv[2], i3 = inv[3];
    const f3 d0 = x1 - x0, d1 = x2 - x0;
    F[0] = (0.0f + d0.x * i0) + d1.x * i1, F[1] = (0.0f + d0.y * i0) + d1.y * i1;
    F[2] = (0.0f + d0.z * i0) + d1.z * i1, F[3] = (0.0f + d0.x * i2) + d1.x * i3;
    F[4] = (0.0f + d0.y * i2) + d1.y * i3, F[5] = (0.0f + d0.z * i2) + d1.z * i3;
}

// Block (i, j) of the upper-stored elastic pattern (zero when absent),
// transposed when i > j (FixedCSRMat's full-matrix view).
__device__ inline void pattern_block(const uint32_t *rp, const uint32_t *ci, const float *val,
                                     uint32_t i, uint32_t j, float blk[9]) {
    const uint32_t r = i < j ? i : j, c = i < j ? j : i;
#pragma unroll
    for (int e = 0; e < 9; ++e) blk[e] = 0.0f;
    for (uint32_t k = rp[r]; k < rp[r + 1]; ++k) {
        if (ci[k] == c) {
            const float *b = val + 9 * (size_t)k;
            if (i <= j) {
#pragma unroll
                for (int e = 0; e < 9; ++e) blk[e] = b[e];
            } else {
#pragma unroll
                for (int rr = 0; rr < 3; ++rr)
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) blk[rr + 3 * cc] = b[cc + 3 * rr];
            }
            break;
        }
    }
}

// M = U3 Q V^T (3x2, column-major) of a 3x2 mode Q (column-major) with
// U3 = [U0, U1, U0 x U1].
__device__ inline void sl_mode(const float U3[9], const float Q[6], const float V[4], float M[6]) {
    float T[6];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < 3; ++k) acc = acc + U3[r + 3 * k] * Q[k + 3 * c];
            T[r + 3 * c] = acc;
        }
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            float acc = 0.0f;
#pragma unroll
            for (int j = 0; j < 2; ++j) acc = acc + T[r + 3 * j] * V[c + 2 * j];
            M[r + 3 * c] = acc;
        }
}

// Strain-limit terms of every limited face, staged like k_face's (forces per
// face vertex, the six upper blocks per face) with an active flag.
__global__ __launch_bounds__(kNT) void k_sl_face(ppf_shell sh, const uint32_t *__restrict__ rp,
                                                 const uint32_t *__restrict__ ci,
                                                 const float *__restrict__ el, float eig_eps,
                                                 float *__restrict__ slblk,
                                                 float *__restrict__ slf,
                                                 uint8_t *__restrict__ active) {
    const uint32_t k = blockIdx.x * kNT + threadIdx.x;
    if (k >= sh.nf) {
        return;
    }
    active[k] = 0;
    const float sl = sh.f_strainlimit[k];
    if ((sh.fflag[k] & 1u) || !(sl > 0.0f)) {
        return;
    }
    const uint32_t v[3] = {sh.face[3 * (size_t)k], sh.face[3 * (size_t)k + 1],
                           sh.face[3 * (size_t)k + 2]};
    const float *inv = sh.inv_rest + 4 * (size_t)k;
    float F[6];
    face_F(sh.x, v, inv, F);
    const Svd32 svd = svd32(F);
    float gs[2] = {svd.S[0] - 1.0f, svd.S[1] - 1.0f}; // stretch beyond rest
    if (!(fmaxf(gs[0], gs[1]) > 0.0f)) {
        return;
    }
    active[k] = 1;
    // barrier on sl - stretch with gap sl (cubic.hpp:22-40, offset 0)
    float deda[2] = {0.0f, 0.0f}, h[2] = {0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        if (gs[i] > 0.0f) {
            const float y = (sl - gs[i]) - 0.0f;
            const float yy = y - sl;
            deda[i] = -(yy < 0.0f ? -2.0f * yy * yy / sl : 0.0f);
            h[i] = y - sl < 0.0f ? 4.0f * (1.0f - y / sl) : 0.0f;
        }
    }
    // elastic-inclusive stiffness at the step start: w^T H_el w + m / d^2
    float stiff = 0.0f;
    {
        const f3 y0 = X(sh.x0, v[0]), y1 = X(sh.x0, v[1]), y2 = X(sh.x0, v[2]);
        const float s3 = 1.0f / 3.0f;
        const f3 cen = (s3 * y0 + s3 * y1) + s3 * y2;
        const f3 wv[3] = {y0 - cen, y1 - cen, y2 - cen};
        for (int ii = 0; ii < 3; ++ii) {
            float hv[3] = {0.0f, 0.0f, 0.0f};
            for (int jj = 0; jj < 3; ++jj) {
                float blk[9];
                pattern_block(rp, ci, el, v[ii], v[jj], blk);
                const float w[3] = {wv[jj].x, wv[jj].y, wv[jj].z};
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c) hv[r] = hv[r] + blk[r + 3 * c] * w[c];
            }
            stiff = stiff + ((wv[ii].x * hv[0] + wv[ii].y * hv[1]) + wv[ii].z * hv[2]);
        }
        const float d = sl_limit(sh, k) - fmaxf(gs[0], gs[1]);
        stiff = stiff + sh.fmass[k] / (d * d);
    }
    const float a0 = gs[0] + 1.0f, a1 = gs[1] + 1.0f;
    // dE/dF = U diag(deda) V^T
    float dF[6];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r)
            dF[r + 3 * c] = (0.0f + (svd.U[r] * deda[0]) * svd.V[c]) +
                            (svd.U[3 + r] * deda[1]) * svd.V[c + 2];
    // projected d2E/dF2: six modes (twist, flip, two out-of-plane, two scaling)
    float U3[9];
    {
        const f3 u0 = mk3(svd.U[0], svd.U[1], svd.U[2]), u1 = mk3(svd.U[3], svd.U[4], svd.U[5]);
        const f3 u2 = cross3(u0, u1);
        U3[0] = u0.x, U3[1] = u0.y, U3[2] = u0.z, U3[3] = u1.x, U3[4] = u1.y, U3[5] = u1.z;
        U3[6] = u2.x, U3[7] = u2.y, U3[8] = u2.z;
    }
    const float is2 = 0.7071067811865475f;
    const float denom = a0 - a1;
    float lam[6];
    lam[0] = fmaxf(0.0f, (deda[0] + deda[1]) / (a0 + a1));
    lam[1] = fmaxf(0.0f, fabsf(denom) > eig_eps ? (deda[0] - deda[1]) / denom
                                                : 0.5f * (h[0] + h[1]) - 0.5f * (0.0f + 0.0f));
    lam[2] = fmaxf(0.0f, deda[0] / a0);
    lam[3] = fmaxf(0.0f, deda[1] / a1);
    // scaling modes in ascending curvature order
    const int lo = h[1] < h[0] ? 1 : 0;
    lam[4] = fmaxf(0.0f, h[lo]);
    lam[5] = fmaxf(0.0f, h[1 - lo]);
    float H6[36];
#pragma unroll
    for (int e = 0; e < 36; ++e) H6[e] = 0.0f;
    for (int i = 0; i < 6; ++i) {
        if (lam[i] == 0.0f) {
            continue;
        }
        float Q[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (i == 0) {
            Q[3] = is2, Q[1] = -is2; // (0,1) = 1/sqrt2, (1,0) = -1/sqrt2
        } else if (i == 1) {
            Q[3] = is2, Q[1] = is2;
        } else if (i == 2) {
            Q[2] = 1.0f;             // (2,0)
        } else if (i == 3) {
            Q[5] = 1.0f;             // (2,1)
        } else {
            const int ax = i == 4 ? lo : 1 - lo;
            Q[ax + 3 * ax] = 1.0f;   // (ax, ax)
        }
        float M[6];
        sl_mode(U3, Q, svd.V, M);
        for (int n = 0; n < 6; ++n)
#pragma unroll
            for (int m = 0; m < 6; ++m) H6[m + 6 * n] = H6[m + 6 * n] + (lam[i] * M[m]) * M[n];
    }
    const float i0 = inv[0], i1 = inv[1], i2 = inv[2], i3 = inv[3];
    const float gg[3][2] = {{-i0 - i1, -i2 - i3}, {i0, i2}, {i1, i3}};
    float *fo = slf + 9 * (size_t)k;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float fe = (0.0f + gg[a][0] * dF[d]) + gg[a][1] * dF[3 + d];
            fo[3 * a + d] = 0.0f + stiff * fe;
        }
    float *bo = slblk + 9 * kFacePairs * (size_t)k;
    for (uint32_t p = 0; p < kFacePairs; ++p) {
        uint32_t a = kFP[p][0], b = kFP[p][1];
        if (v[a] > v[b]) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        float blk[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) blk[e] = 0.0f;
#pragma unroll
        for (int d = 0; d < 2; ++d)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const float sc = gg[a][d] * gg[b][e];
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int r = 0; r < 3; ++r)
                        blk[r + 3 * c] = blk[r + 3 * c] + sc * H6[(3 * d + r) + 6 * (3 * e + c)];
            }
#pragma unroll
        for (int e = 0; e < 9; ++e) bo[9 * p + e] = 0.0f + stiff * blk[e];
    }
}

// Adds the staged strain-limit blocks to the stored blocks, per block in
// face order (the faces' share of the elastic slot table).
__global__ __launch_bounds__(kNT) void k_sl_gather_blocks(
    uint32_t nnz, uint32_t nf, const uint32_t *__restrict__ slot_off,
    const uint32_t *__restrict__ src, const float *__restrict__ slblk,
    const uint8_t *__restrict__ active, float *__restrict__ value) {
    const uint32_t s = blockIdx.x * kNT + threadIdx.x;
    if (s >= nnz) {
        return;
    }
    const uint32_t nfp = kFacePairs * nf;
    float *o = value + 9 * (size_t)s;
    float acc[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) acc[e] = o[e];
    for (uint32_t c = slot_off[s]; c < slot_off[s + 1]; ++c) {
        const uint32_t q = src[c];
        if (q >= nfp || !active[q / kFacePairs]) {
            continue;
        }
        const float *blk = slblk + 9 * (size_t)q;
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            const float y = blk[e];
            if (y != 0.0f) {
                acc[e] += y;
            }
        }
    }
#pragma unroll
    for (int e = 0; e < 9; ++e) o[e] = acc[e];
}

// Adds the staged strain-limit forces per vertex in face order.
__global__ __launch_bounds__(kNT) void k_sl_gather_force(
    uint32_t nv, uint32_t nf, const uint32_t *__restrict__ off,
    const uint32_t *__restrict__ src, const float *__restrict__ slf,
    const uint8_t *__restrict__ active, float *__restrict__ force) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i >= nv) {
        return;
    }
    float f0 = force[3 * (size_t)i], f1 = force[3 * (size_t)i + 1], f2 = force[3 * (size_t)i + 2];
    for (uint32_t c = off[i]; c < off[i + 1]; ++c) {
        const uint32_t q = src[c];
        if (q >= 3 * nf || !active[q / 3]) {
            continue;
        }
        const float *g = slf + 9 * (size_t)(q / 3) + 3 * (q % 3);
        f0 += g[0];
        f1 += g[1];
        f2 += g[2];
    }
    force[3 * (size_t)i] = f0;
    force[3 * (size_t)i + 1] = f1;
    force[3 * (size_t)i + 2] = f2;
}

// Largest stretch beyond rest of the two singular values of F.
__device__ inline float max_stretch(const float F[6]) {
    const Svd32 s = svd32(F);
    return fmaxf(s.S[0] - 1.0f, s.S[1] - 1.0f);
}

// Strain-limit line search of one face per thread (strainlimiting.cu:243-305):
// the sweep F0 + t (F1 - F0), t in [0, max_t]; a face whose stretch reaches
// its limit by max_t is bisected to the last t below it.
__global__ __launch_bounds__(kNT) void k_sl_toi(ppf_shell sh, const float *__restrict__ x0,
                                                const float *__restrict__ x1, float max_t,
                                                const uint32_t *__restrict__ vscene,
                                                uint32_t *__restrict__ toi_bits) {
    const uint32_t k = blockIdx.x * kNT + threadIdx.x;
    if (k >= sh.nf) {
        return;
    }
    const float sl = sh.f_strainlimit[k];
    if ((sh.fflag[k] & 1u) || !(sl > 0.0f)) {
        return;
    }
    const uint32_t v[3] = {sh.face[3 * (size_t)k], sh.face[3 * (size_t)k + 1],
                           sh.face[3 * (size_t)k + 2]};
    const float *inv = sh.inv_rest + 4 * (size_t)k;
    const float limit = sl_limit(sh, k);
    float F0[6], F1[6], dF[6], Ft[6];
    face_F(x0, v, inv, F0);
    face_F(x1, v, inv, F1);
#pragma unroll
    for (int e = 0; e < 6; ++e) dF[e] = F1[e] - F0[e];
    float t = max_t;
#pragma unroll
    for (int e = 0; e < 6; ++e) Ft[e] = F0[e] + t * dF[e];
    if (max_stretch(Ft) >= limit) {
        float upper = t, lower = 0.0f;
        float window = upper - lower;
        while (true) {
            t = 0.5f * (upper + lower);
#pragma unroll
            for (int e = 0; e < 6; ++e) Ft[e] = F0[e] + t * dF[e];
            if (max_stretch(Ft) - limit < 0.0f) {
                lower = t;
            } else {
                upper = t;
            }
            const float nw = upper - lower;
            if (nw == window) {
                break;
            }
            window = nw;
        }
        t = lower;
    }
    if (t < max_t) {
        atomicMin(toi_bits + (vscene ? vscene[v[0]] : 0u), __float_as_uint(t));
    }
}

// ---- plasticity -------------------------------------------------------------------
//
// Rest-shape creep after a committed step (plasticity.cu:21-78 faces, 145-185
// hinges): the singular values of F (faces) and the dihedral angle (hinges)
// outside the dead zone |s - 1| <= threshold, |angle - rest| <= threshold are
// pulled toward its edge by alpha = 1 - exp(-rate dt) (the reference takes the
// hardware exponential there; __expf is gfx950's).

__device__ inline f3 unit3(f3 v) {
    const float z = dot3(v, v);
    if (!(z > 0.0f)) {
        return v;
    }
    const float n = sqrtf(z);
    return mk3(v.x / n, v.y / n, v.z / n);
}

// One face per thread; faces of scenes that did not commit (ok[scene] == 0)
// are skipped.
__global__ __launch_bounds__(kNT) void k_plastic_face(ppf_shell sh, float dt,
                                                      const uint32_t *__restrict__ vscene,
                                                      const uint32_t *__restrict__ ok,
                                                      float *__restrict__ inv_rest) {
    const uint32_t k = blockIdx.x * kNT + threadIdx.x;
    if (k >= sh.nf) {
        return;
    }
    const float rate = sh.f_plasticity[k];
    if (!(rate > 0.0f)) {
        return;
    }
    const uint32_t v[3] = {sh.face[3 * (size_t)k], sh.face[3 * (size_t)k + 1],
                           sh.face[3 * (size_t)k + 2]};
    if (ok && !ok[vscene ? vscene[v[0]] : 0u]) {
        return;
    }
    const float thr = sh.f_plastic_threshold[k];
    const float alpha = 1.0f - __expf(-rate * dt);
    float *inv = inv_rest + 4 * (size_t)k;
    float F[6];
    face_F(sh.x, v, inv, F);
    const Svd32 svd = svd32(F);
    float s[2] = {svd.S[0], svd.S[1]};
    bool changed = false;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        if (fabsf(svd.S[i] - 1.0f) > thr) {
            const float target = svd.S[i] < 1.0f ? 1.0f - thr : 1.0f + thr;
            s[i] = svd.S[i] + alpha * (target - svd.S[i]);
            changed = true;
        }
    }
    if (!changed) {
        return;
    }
    // F' = U diag(s) V^T
    float Fn[6];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r)
            Fn[r + 3 * c] = (0.0f + (svd.U[r] * s[0]) * svd.V[c]) + (svd.U[3 + r] * s[1]) * svd.V[c + 2];
    // the face's in-plane frame (t along the first edge, b = n x t) maps dX
    // and F' to 2x2; the new rest inverse is (P dX)^-1 (P F')
    const f3 x0 = X(sh.x, v[0]);
    const f3 d0 = X(sh.x, v[1]) - x0, d1 = X(sh.x, v[2]) - x0;
    const f3 n = unit3(cross3(d0, d1));
    const f3 t = unit3(d0);
    const f3 b = unit3(cross3(n, t));
    const float a00 = dot3(t, d0), a01 = dot3(t, d1), a10 = dot3(b, d0), a11 = dot3(b, d1);
    const f3 fc0 = mk3(Fn[0], Fn[1], Fn[2]), fc1 = mk3(Fn[3], Fn[4], Fn[5]);
    const float p00 = dot3(t, fc0), p01 = dot3(t, fc1), p10 = dot3(b, fc0), p11 = dot3(b, fc1);
    const float det = a00 * a11 - a01 * a10;
    const float id = 1.0f / det;
    const float i00 = a11 * id, i01 = -a01 * id, i10 = -a10 * id, i11 = a00 * id;
    inv[0] = i00 * p00 + i01 * p10;
    inv[1] = i10 * p00 + i11 * p10;
    inv[2] = i00 * p01 + i01 * p11;
    inv[3] = i10 * p01 + i11 * p11;
}

__device__ float dihedral_angle(f3 v0, f3 v1, f3 v2, f3 v3);

// One hinge per thread.
__global__ __launch_bounds__(kNT) void k_plastic_hinge(ppf_shell sh, float dt,
                                                       const uint32_t *__restrict__ vscene,
                                                       const uint32_t *__restrict__ ok,
                                                       float *__restrict__ h_rest) {
    const uint32_t k = blockIdx.x * kNT + threadIdx.x;
    if (k >= sh.nh || (sh.hflag[k] & 1u)) {
        return;
    }
    const float rate = sh.h_plasticity[k];
    if (!(rate > 0.0f)) {
        return;
    }
    const uint32_t *h = sh.hinge + 4 * (size_t)k;
    if (ok && !ok[vscene ? vscene[h[0]] : 0u]) {
        return;
    }
    const float thr = sh.h_plastic_threshold[k];
    const float alpha = 1.0f - __expf(-rate * dt);
    const float theta = dihedral_angle(X(sh.x, h[2]), X(sh.x, h[1]), X(sh.x, h[0]), X(sh.x, h[3]));
    const float rest = h_rest[k];
    const float delta = theta - rest;
    if (fabsf(delta) > thr) {
        const float edge = delta > 0.0f ? thr : -thr;
        h_rest[k] = rest + alpha * (delta - edge);
    }
}

// ---- dihedral bending -----------------------------------------------------------

__device__ float dihedral_angle(f3 v0, f3 v1, f3 v2, f3 v3) {
    const f3 n1 = cross3(v1 - v0, v2 - v0);
    const f3 n2 = cross3(v2 - v3, v1 - v3);
    const f3 e = v1 - v2;
    const float en = v_norm(e);
    if (en <= 0.0f) {
        return 0.0f;
    }
    const float s = dot3(cross3(n2, n1), e) / en;
    const float c = dot3(n1, n2);
    return atan2f(s, c);
}

__device__ void psd_project4(float A[4][4]) {
    float V[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) V[i][j] = i == j ? 1.0f : 0.0f;
    float scale = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) scale = fmaxf(scale, fabsf(A[i][j]));
    if (scale == 0.0f) {
        return;
    }
    const float tiny = 1e-7f * scale;
    for (int sweep = 0; sweep < 24; ++sweep) {
        float off = 0.0f;
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int q = p + 1; q < 4; ++q) off = fmaxf(off, fabsf(A[p][q]));
        if (off <= tiny) {
            break;
        }
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int q = p + 1; q < 4; ++q) {
                const float apq = A[p][q];
                if (fabsf(apq) <= tiny) {
                    continue;
                }
                const float app = A[p][p], aqq = A[q][q];
                const float tau = (aqq - app) / (2.0f * apq);
                const float t = (tau >= 0.0f ? 1.0f : -1.0f) / (fabsf(tau) + sqrtf(1.0f + tau * tau));
                const float c = 1.0f / sqrtf(1.0f + t * t);
                const float s = t * c;
                A[p][p] = c * c * app - 2.0f * s * c * apq + s * s * aqq;
                A[q][q] = s * s * app + 2.0f * s * c * apq + c * c * aqq;
                A[p][q] = 0.0f;
                A[q][p] = 0.0f;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (k == p || k == q) continue;
                    const float akp = A[k][p], akq = A[k][q];
                    const float np = c * akp - s * akq, nq = s * akp + c * akq;
                    A[k][p] = np;
                    A[p][k] = np;
                    A[k][q] = nq;
                    A[q][k] = nq;
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const float vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    float lam[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) lam[i] = fmaxf(A[i][i], 0.0f);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = i; j < 4; ++j) {
            float sum = 0.0f;
#pragma unroll
            for (int k = 0; k < 4; ++k) sum += V[i][k] * lam[k] * V[j][k];
            A[i][j] = sum;
            A[j][i] = sum;
        }
}

// Unit-stiffness hinge force (12, X0..X3 order) and the projected Hessian in
// the Wu-Kim P-order (Hp, 12x12 row-major; P = (X2, X1, X0, X3)).
__device__ void hinge_unit(f3 X0, f3 X1, f3 X2, f3 X3, float rest, float f12[12],
                           float Hp[12][12]) {
    const float angle = dihedral_angle(X0, X1, X2, X3);
    {
        // gradient with the reference's parameter order (v2, v0, v1, v3)
        const f3 v2 = X0, v0 = X1, v1 = X2, v3 = X3;
        const f3 e0 = v1 - v0, e1 = v2 - v0, e2 = v3 - v0, e3 = v2 - v1, e4 = v3 - v1;
        const f3 n1 = cross3(e0, e1), n2 = cross3(e2, e0);
        const float n1s = dot3(n1, n1), n2s = dot3(n2, n2);
        const float en = v_norm(e0);
        const f3 g0 = (-en / n1s) * n1;
        const f3 g1 = (-dot3(e0, e3) / (en * n1s)) * n1 - (dot3(e0, e4) / (en * n2s)) * n2;
        const f3 g2 = (dot3(e0, e1) / (en * n1s)) * n1 + (dot3(e0, e2) / (en * n2s)) * n2;
        const f3 g3 = (-en / n2s) * n2;
        const float gd = angle - rest;
        const f3 gs[4] = {g0, g1, g2, g3};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            f12[3 * k] = gd * gs[k].x;
            f12[3 * k + 1] = gd * gs[k].y;
            f12[3 * k + 2] = gd * gs[k].z;
        }
    }
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) Hp[i][j] = 0.0f;
    const f3 y1 = X1 - X2, y2 = X0 - X2, y3 = X3 - X2;
    const float ly1 = v_norm(y1);
    if (ly1 <= 1e-9f) {
        return;
    }
    const f3 t1 = v_div(y1, ly1);
    const M3 sigma = m_sub(m_eye(), m_outer(t1, t1));
    const f3 z0 = y2 - v_mul(t1, dot3(t1, y2));
    const f3 z1 = y3 - v_mul(t1, dot3(t1, y3));
    const float l0 = v_norm(z0), l1 = v_norm(z1);
    const f3 bcr = cross3(z0, z1);
    const float lb = v_norm(bcr);
    if (l0 <= 1e-9f || l1 <= 1e-9f || lb <= 1e-9f * l0 * l1) {
        return;
    }
    const f3 tb = v_div(bcr, lb);
    const float Sgeom = (dot3(tb, t1) >= 0.0f) ? 1.0f : -1.0f;
    const float gdiff = angle - rest;
    const f3 e0p = cross3(z0, tb), e1p = cross3(z1, tb);
    const float w = -Sgeom * gdiff;
    const float gamma = (l1 * l1) / (l0 * l0);
    const float gm1 = gamma - 1.0f, gp1 = gamma + 1.0f;
    const float r = sqrtf(4.0f * w * w * (gm1 / gp1) * (gm1 / gp1) + 1.0f);
    const float Rm = sqrtf(fmaxf(0.0f, 2.0f * (2.0f * w * w + 1.0f - r)));
    const float Rp = sqrtf(fmaxf(0.0f, 2.0f * (2.0f * w * w + 1.0f + r)));
    const float Lh = 1.0f / (l0 * l0) + 1.0f / (l1 * l1);
    const float cc = 4.0f * w / gp1;
    const M3 eta2 = m_scale(-1.0f / ly1, m_add(m_scale(dot3(t1, y2), sigma), m_outer(t1, z0)));
    const M3 eta3 = m_scale(-1.0f / ly1, m_add(m_scale(dot3(t1, y3), sigma), m_outer(t1, z1)));
    const M3 e2T = m_trans(eta2), e3T = m_trans(eta3);
    const M3 m0 = m_sub(m_scale(-1.0f, e2T), sigma);
    const M3 m1 = m_sub(m_scale(-1.0f, e3T), sigma);
    auto accF = [&](float lam, f3 qa, f3 qb) {
        if (lam <= 0.0f) {
            return;
        }
        const float n2 = dot3(qa, qa) + dot3(qb, qb);
        if (n2 <= 1e-20f) {
            return;
        }
        const f3 b0 = m_mul(m0, qa) + m_mul(m1, qb);
        const f3 b1 = m_mul(e2T, qa) + m_mul(e3T, qb);
        const f3 b2 = m_mul(sigma, qa);
        const f3 b3 = m_mul(sigma, qb);
        const float p12[12] = {b0.x, b0.y, b0.z, b1.x, b1.y, b1.z,
                               b2.x, b2.y, b2.z, b3.x, b3.y, b3.z};
        const float s = lam / n2;
        for (int i = 0; i < 12; ++i) {
            const float si = s * p12[i];
            for (int j = 0; j < 12; ++j) Hp[i][j] += si * p12[j];
        }
    };
    const float sr[4] = {-1.0f, -1.0f, 1.0f, 1.0f};
    const float RR[4] = {Rm, Rm, Rp, Rp};
    const float sR[4] = {-1.0f, 1.0f, -1.0f, 1.0f};
    for (int mm = 0; mm < 4; ++mm) {
        const float R = RR[mm];
        const float a = w * (gm1 * gm1 / gp1 + sr[mm] * gp1 * r + sR[mm] * gm1 * R);
        const float b = -0.5f * gamma * R * R + 2.0f * w * w -
                        sR[mm] * 0.5f * (gm1 + sr[mm] * gp1 * r) * R;
        const float d = 1.0f + sr[mm] * r + sR[mm] * R;
        accF(0.25f * Lh * d, a * z0 + b * e0p, cc * z1 + d * e1p);
    }
    const float cos_t = dot3(z0, z1) / (l0 * l1);
    const float sin_t = lb / (l0 * l1);
    const float beta = (l1 / l0 - l0 / l1) * cos_t;
    const float alpha0 = 0.5f * (-beta + sqrtf(beta * beta + 4.0f));
    const float ginv = w / sin_t;
    accF(ginv * (cos_t / (l1 * l1) - alpha0 / (l0 * l1)), alpha0 * tb, tb);
    accF(ginv * (cos_t / (l1 * l1) + 1.0f / (alpha0 * l0 * l1)), tb, (-alpha0) * tb);

    const f3 tau0 = v_div(z0, l0), tau1 = v_div(z1, l1);
    const f3 tau1p = cross3(tau1, tb);
    const float a0 = dot3(t1, y2) / (ly1 * l0);
    const float a1 = dot3(t1, y3) / (ly1 * l1);
    const float c0 = 1.0f / l0, c1 = 1.0f / l1;
    const float k0 = -4.0f * dot3(tau0, tau1p) / (ly1 * ly1);
    const float k1 = 2.0f * dot3(tau0, tau1p) / ly1;
    const f3 s01 = tau0 + tau1, d01 = tau0 - tau1;
    const float tv[3] = {t1.x, t1.y, t1.z};
    const float sv[3] = {s01.x, s01.y, s01.z}, dv[3] = {d01.x, d01.y, d01.z};
    float pc[4][12];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        pc[0][c] = sv[c], pc[0][3 + c] = -sv[c], pc[0][6 + c] = 0.0f, pc[0][9 + c] = 0.0f;
        pc[1][c] = dv[c], pc[1][3 + c] = -dv[c], pc[1][6 + c] = 0.0f, pc[1][9 + c] = 0.0f;
        pc[2][c] = (-a1 - a0 + c1 + c0) * tv[c];
        pc[2][3 + c] = (a1 + a0) * tv[c];
        pc[2][6 + c] = -c0 * tv[c];
        pc[2][9 + c] = -c1 * tv[c];
        pc[3][c] = (-a1 + a0 + c1 - c0) * tv[c];
        pc[3][3 + c] = (a1 - a0) * tv[c];
        pc[3][6 + c] = c0 * tv[c];
        pc[3][9 + c] = -c1 * tv[c];
    }
    float gg[4], g23 = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float s = 0.0f;
#pragma unroll
        for (int m = 0; m < 12; ++m) s += pc[i][m] * pc[i][m];
        gg[i] = s;
    }
#pragma unroll
    for (int m = 0; m < 12; ++m) g23 += pc[2][m] * pc[3][m];
    const float det23 = gg[2] * gg[3] - g23 * g23;
    if (gg[0] > 1e-5f && gg[1] > 1e-5f && det23 > 1e-6f * gg[2] * gg[3]) {
        const float we = -gdiff * Sgeom;
        float E[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) E[i][j] = 0.0f;
        E[0][0] = we * (-k0 * gg[0]);
        E[1][1] = we * (k0 * gg[1]);
        E[0][2] = E[2][0] = we * (k1 * gg[2]);
        E[0][3] = E[3][0] = we * (k1 * g23);
        E[1][2] = E[2][1] = we * (k1 * g23);
        E[1][3] = E[3][1] = we * (k1 * gg[3]);
        float Li[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) Li[i][j] = 0.0f;
        Li[0][0] = 1.0f / sqrtf(gg[0]);
        Li[1][1] = 1.0f / sqrtf(gg[1]);
        const float L22 = sqrtf(gg[2]);
        const float L32 = g23 / L22;
        const float L33 = sqrtf(fmaxf(1e-30f, gg[3] - L32 * L32));
        Li[2][2] = 1.0f / L22;
        Li[3][3] = 1.0f / L33;
        Li[3][2] = -L32 / (L22 * L33);
        float LiE[4][4], Ss[4][4], Core[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float s = 0.0f;
#pragma unroll
                for (int k = 0; k < 4; ++k) s += Li[i][k] * E[k][j];
                LiE[i][j] = s;
            }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float s = 0.0f;
#pragma unroll
                for (int k = 0; k < 4; ++k) s += LiE[i][k] * Li[j][k];
                Ss[i][j] = s;
            }
        psd_project4(Ss);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float s = 0.0f;
#pragma unroll
                for (int k = 0; k < 4; ++k) s += Li[k][i] * Ss[k][j];
                LiE[i][j] = s; // LtS
            }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float s = 0.0f;
#pragma unroll
                for (int k = 0; k < 4; ++k) s += LiE[i][k] * Li[k][j];
                Core[i][j] = s;
            }
        for (int m = 0; m < 12; ++m) {
            float PC[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float s = 0.0f;
#pragma unroll
                for (int i = 0; i < 4; ++i) s += pc[i][m] * Core[i][j];
                PC[j] = s;
            }
            for (int n = 0; n < 12; ++n) {
                float s = 0.0f;
#pragma unroll
                for (int j = 0; j < 4; ++j) s += PC[j] * pc[j][n];
                Hp[m][n] += s;
            }
        }
    }
}

__global__ __launch_bounds__(64) void k_hinge(ppf_shell sh, float *__restrict__ hblk,
                                              float *__restrict__ hf,
                                              uint8_t *__restrict__ active) {
    const uint32_t k = blockIdx.x * 64 + threadIdx.x;
    if (k >= sh.nh) {
        return;
    }
    const uint32_t *h = sh.hinge + 4 * (size_t)k;
    const uint32_t hv[4] = {h[0], h[1], h[2], h[3]};
    float stiff = 0.0f;
    if (!(sh.hflag[k] & 1u)) {
        float areal = 0.0f;
        int cnt = 0;
        for (int q = 0; q < 4; ++q) {
            const float a = sh.varea[hv[q]];
            if (a > 0.0f) {
                areal += sh.vmass[hv[q]] / a;
                ++cnt;
            }
        }
        if (cnt > 0) {
            areal /= (float)cnt;
        }
        const float length = sh.h_length[k], area = sh.h_area[k];
        stiff = (area > 1e-12f) ? 1.28e-5f * sh.h_bend[k] * (length * length / area) * areal
                                : 0.0f;
    }
    const bool on = stiff > 0.0f;
    active[k] = on ? 1 : 0;
    if (!on) {
        return;
    }
    const uint32_t rv[4] = {hv[2], hv[1], hv[0], hv[3]}; // remapped order X0..X3
    float f12[12];
    float Hp[12][12];
    hinge_unit(X(sh.x, rv[0]), X(sh.x, rv[1]), X(sh.x, rv[2]), X(sh.x, rv[3]), sh.h_rest[k],
               f12, Hp);
    float *fo = hf + 12 * (size_t)k;
#pragma unroll
    for (int i = 0; i < 12; ++i) fo[i] = f12[i] * stiff;
    // codebase block (A, B) = P-order block (perm[A], perm[B]), perm = (2, 1, 0, 3)
    const int perm[4] = {2, 1, 0, 3};
    float *bo = hblk + 9 * kHingePairs * (size_t)k;
    for (uint32_t p = 0; p < kHingePairs; ++p) {
        uint32_t A = kHP[p][0], B = kHP[p][1];
        if (rv[A] > rv[B]) {
            const uint32_t t = A;
            A = B;
            B = t;
        }
        const int pa = perm[A], pb = perm[B];
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) bo[9 * p + r + 3 * c] = Hp[3 * pa + r][3 * pb + c] * stiff;
    }
}

// ---- pattern + contribution tables --------------------------------------------

struct Topo {
    uint32_t nv, nf, nh;
    const uint32_t *face, *hinge;
};

// Stored (row, col) of contribution `src`: face pairs, then hinge pairs, then
// one diagonal entry per vertex (so every row has its diagonal slot).
__device__ inline void pair_rc(const Topo &t, uint32_t src, uint32_t &row, uint32_t &col) {
    const uint32_t nfp = kFacePairs * t.nf, nhp = kHingePairs * t.nh;
    uint32_t a, b;
    if (src < nfp) {
        const uint32_t k = src / kFacePairs, p = src % kFacePairs;
        a = t.face[3 * (size_t)k + kFP[p][0]];
        b = t.face[3 * (size_t)k + kFP[p][1]];
    } else if (src < nfp + nhp) {
        const uint32_t k = (src - nfp) / kHingePairs, p = (src - nfp) % kHingePairs;
        const uint32_t *h = t.hinge + 4 * (size_t)k;
        const uint32_t rv[4] = {h[2], h[1], h[0], h[3]};
        a = rv[kHP[p][0]];
        b = rv[kHP[p][1]];
    } else {
        a = b = src - nfp - nhp;
    }
    row = a < b ? a : b;
    col = a < b ? b : a;
}

__global__ void k_pair_init(Topo t, uint32_t n, uint32_t *__restrict__ key,
                            uint32_t *__restrict__ val) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i >= n) {
        return;
    }
    uint32_t r, c;
    pair_rc(t, i, r, c);
    key[i] = c;
    val[i] = i;
}

__global__ void k_pair_rowkey(Topo t, uint32_t n, const uint32_t *__restrict__ val,
                              uint32_t *__restrict__ key) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i >= n) {
        return;
    }
    uint32_t r, c;
    pair_rc(t, val[i], r, c);
    key[i] = r;
}

__global__ void k_pair_heads(Topo t, uint32_t n, const uint32_t *__restrict__ val,
                             uint32_t *__restrict__ head) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i >= n) {
        return;
    }
    uint32_t r, c;
    pair_rc(t, val[i], r, c);
    uint32_t h = 1;
    if (i > 0) {
        uint32_t rp, cp;
        pair_rc(t, val[i - 1], rp, cp);
        h = (rp != r || cp != c) ? 1u : 0u;
    }
    head[i] = h;
}

// Slot of every head: column, start of its contribution range, row counts.
__global__ void k_pair_slots(Topo t, uint32_t n, const uint32_t *__restrict__ val,
                             const uint32_t *__restrict__ head,
                             const uint32_t *__restrict__ sid, uint32_t *__restrict__ ci,
                             uint32_t *__restrict__ slot_off, uint32_t *__restrict__ rcount) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i >= n || !head[i]) {
        return;
    }
    uint32_t r, c;
    pair_rc(t, val[i], r, c);
    const uint32_t s = sid[i];
    ci[s] = c;
    slot_off[s] = i;
    atomicAdd(rcount + r, 1u);
}

// Force contributions: face corner k*3 + a, then hinge corner 3 nf + 4 k + a.
__device__ inline uint32_t corner_vertex(const Topo &t, uint32_t src) {
    if (src < 3 * t.nf) {
        return t.face[src];
    }
    const uint32_t q = src - 3 * t.nf, k = q / 4, a = q % 4;
    const uint32_t *h = t.hinge + 4 * (size_t)k;
    const uint32_t rv[4] = {h[2], h[1], h[0], h[3]};
    return rv[a];
}

__global__ void k_corner_init(Topo t, uint32_t n, uint32_t *__restrict__ key,
                              uint32_t *__restrict__ val) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i >= n) {
        return;
    }
    key[i] = corner_vertex(t, i);
    val[i] = i;
}

__global__ void k_corner_count(Topo t, uint32_t n, uint32_t *__restrict__ cnt) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i >= n) {
        return;
    }
    atomicAdd(cnt + corner_vertex(t, i), 1u);
}

// ---- per-assembly gathers ---------------------------------------------------------

__global__ __launch_bounds__(kNT) void k_gather_blocks(
    uint32_t nnz, uint32_t nf, uint32_t nh, const uint32_t *__restrict__ slot_off,
    const uint32_t *__restrict__ src, const float *__restrict__ fblk,
    const float *__restrict__ hblk, const uint8_t *__restrict__ factive,
    const uint8_t *__restrict__ hactive, float *__restrict__ value) {
    const uint32_t s = blockIdx.x * kNT + threadIdx.x;
    if (s >= nnz) {
        return;
    }
    const uint32_t nfp = kFacePairs * nf, nhp = kHingePairs * nh;
    float acc[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) acc[e] = 0.0f;
    for (uint32_t c = slot_off[s]; c < slot_off[s + 1]; ++c) {
        const uint32_t q = src[c];
        const float *blk;
        if (q < nfp) {
            if (!factive[q / kFacePairs]) continue;
            blk = fblk + 9 * (size_t)q;
        } else if (q < nfp + nhp) {
            if (!hactive[(q - nfp) / kHingePairs]) continue;
            blk = hblk + 9 * (size_t)(q - nfp);
        } else {
            continue; // diagonal placeholder
        }
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            const float y = blk[e];
            if (y != 0.0f) {
                acc[e] += y;
            }
        }
    }
    float *o = value + 9 * (size_t)s;
#pragma unroll
    for (int e = 0; e < 9; ++e) o[e] = acc[e];
}

__global__ __launch_bounds__(kNT) void k_gather_force(
    uint32_t nv, uint32_t nf, const uint32_t *__restrict__ off,
    const uint32_t *__restrict__ src, const float *__restrict__ vf,
    const float *__restrict__ ff, const float *__restrict__ hf,
    const uint8_t *__restrict__ factive, const uint8_t *__restrict__ hactive,
    float *__restrict__ force) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i >= nv) {
        return;
    }
    float f0 = vf[3 * (size_t)i], f1 = vf[3 * (size_t)i + 1], f2 = vf[3 * (size_t)i + 2];
    for (uint32_t c = off[i]; c < off[i + 1]; ++c) {
        const uint32_t q = src[c];
        const float *g;
        if (q < 3 * nf) {
            if (!factive[q / 3]) continue;
            g = ff + 9 * (size_t)(q / 3) + 3 * (q % 3);
        } else {
            const uint32_t h = (q - 3 * nf) / 4, a = (q - 3 * nf) % 4;
            if (!hactive[h]) continue;
            g = hf + 12 * (size_t)h + 3 * a;
        }
        f0 += g[0];
        f1 += g[1];
        f2 += g[2];
    }
    force[3 * (size_t)i] = f0;
    force[3 * (size_t)i + 1] = f1;
    force[3 * (size_t)i + 2] = f2;
}

__global__ void k_row_offsets(uint32_t nv, const uint32_t *__restrict__ cnt,
                              uint32_t *__restrict__ rp, uint32_t total) {
    // rp = exclusive scan of cnt is written by the scan; the tail entry here.
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        rp[nv] = total;
    }
}

} // namespace

ElasticAssembler::~ElasticAssembler() { release(); }

void ElasticAssembler::release() {
    for (void *p : {(void *)rp_, (void *)ci_, (void *)slot_off_, (void *)slot_src_,
                    (void *)vtx_off_, (void *)vtx_src_, (void *)fblk_, (void *)hblk_,
                    (void *)ff_, (void *)hf_, (void *)vf_, (void *)slblk_, (void *)slf_,
                    (void *)toi_word_}) {
        if (p) {
            (void)dev_free(p);
        }
    }
    rp_ = ci_ = slot_off_ = slot_src_ = vtx_off_ = vtx_src_ = toi_word_ = nullptr;
    fblk_ = hblk_ = ff_ = hf_ = vf_ = slblk_ = slf_ = nullptr;
    nnz_ = 0;
}

int ElasticAssembler::create(const ppf_shell &topo, hipStream_t s) {
    release();
    nv_ = topo.nv, nf_ = topo.nf, nh_ = topo.nh;
    Topo t{topo.nv, topo.nf, topo.nh, topo.face, topo.hinge};
    const uint64_t np64 = (uint64_t)kFacePairs * nf_ + (uint64_t)kHingePairs * nh_ + nv_;
    const uint64_t nc64 = 3ull * nf_ + 4ull * nh_;
    if (np64 >= (1ull << 32) || nc64 >= (1ull << 32)) {
        set_error("elastic: mesh too large for 32-bit contribution tables");
        return PPF_ERR_CAPACITY;
    }
    const uint32_t np = (uint32_t)np64, nc = (uint32_t)nc64;
    uint32_t kb = 1;
    while (kb < 32 && (1u << kb) <= nv_) {
        ++kb;
    }
    const size_t nmax = np > nc ? np : nc;
    uint32_t *key = nullptr, *head = nullptr, *sid = nullptr, *scratch = nullptr;
    const size_t sw = radix_sort_scratch_words(nmax) > scan_scratch_words(nmax + 1)
                          ? radix_sort_scratch_words(nmax)
                          : scan_scratch_words(nmax + 1);
    PPF_HIP(dev_alloc(&key, nmax * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&head, (nmax + 1) * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&sid, (nmax + 1) * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&scratch, sw * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&slot_src_, (size_t)np * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&vtx_src_, (size_t)(nc ? nc : 1) * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&rp_, ((size_t)nv_ + 1) * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&vtx_off_, ((size_t)nv_ + 1) * sizeof(uint32_t)));
    // Blocks: stable sort by column, then stable sort by row.
    k_pair_init<<<cdiv(np, kNT), kNT, 0, s>>>(t, np, key, slot_src_);
    PPF_HIP(hipGetLastError());
    PPF_HIP(radix_sort_pairs(key, slot_src_, np, kb, scratch, s));
    k_pair_rowkey<<<cdiv(np, kNT), kNT, 0, s>>>(t, np, slot_src_, key);
    PPF_HIP(hipGetLastError());
    PPF_HIP(radix_sort_pairs(key, slot_src_, np, kb, scratch, s));
    k_pair_heads<<<cdiv(np, kNT), kNT, 0, s>>>(t, np, slot_src_, head);
    PPF_HIP(hipGetLastError());
    PPF_HIP(exclusive_scan_u32(head, sid, np, scratch, s));
    uint32_t last_sid = 0, last_head = 0;
    PPF_HIP(hipMemcpyAsync(&last_sid, sid + np - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PPF_HIP(hipMemcpyAsync(&last_head, head + np - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PPF_HIP(hipStreamSynchronize(s));
    nnz_ = last_sid + last_head;
    PPF_HIP(dev_alloc(&ci_, (size_t)nnz_ * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&slot_off_, ((size_t)nnz_ + 1) * sizeof(uint32_t)));
    PPF_HIP(hipMemsetAsync(head + np, 0, sizeof(uint32_t), s));
    uint32_t *rcount = key; // reuse: per-row slot counts
    PPF_HIP(hipMemsetAsync(rcount, 0, (size_t)nv_ * sizeof(uint32_t), s));
    k_pair_slots<<<cdiv(np, kNT), kNT, 0, s>>>(t, np, slot_src_, head, sid, ci_, slot_off_,
                                               rcount);
    PPF_HIP(hipGetLastError());
    PPF_HIP(hipMemcpyAsync(slot_off_ + nnz_, &np, sizeof(uint32_t), hipMemcpyHostToDevice, s));
    PPF_HIP(exclusive_scan_u32(rcount, rp_, nv_, scratch, s));
    k_row_offsets<<<1, 1, 0, s>>>(nv_, rcount, rp_, nnz_);
    PPF_HIP(hipGetLastError());
    // Force contributions: stable sort of corners by vertex.
    if (nc) {
        k_corner_init<<<cdiv(nc, kNT), kNT, 0, s>>>(t, nc, key, vtx_src_);
        PPF_HIP(hipGetLastError());
        PPF_HIP(radix_sort_pairs(key, vtx_src_, nc, kb, scratch, s));
        PPF_HIP(hipMemsetAsync(head, 0, (size_t)nv_ * sizeof(uint32_t), s));
        k_corner_count<<<cdiv(nc, kNT), kNT, 0, s>>>(t, nc, head);
        PPF_HIP(hipGetLastError());
        PPF_HIP(exclusive_scan_u32(head, vtx_off_, nv_, scratch, s));
    } else {
        PPF_HIP(hipMemsetAsync(vtx_off_, 0, (size_t)nv_ * sizeof(uint32_t), s));
    }
    k_row_offsets<<<1, 1, 0, s>>>(nv_, nullptr, vtx_off_, nc);
    PPF_HIP(hipGetLastError());
    // Staging: blocks, element forces, activity bytes (after the forces).
    PPF_HIP(dev_alloc(&fblk_, (9ull * kFacePairs * nf_ + 1) * sizeof(float)));
    PPF_HIP(dev_alloc(&hblk_, (9ull * kHingePairs * nh_ + 1) * sizeof(float)));
    PPF_HIP(dev_alloc(&ff_, (9ull * nf_ + 1) * sizeof(float) + nf_ + nh_ + 8));
    PPF_HIP(dev_alloc(&hf_, (12ull * nh_ + 1) * sizeof(float)));
    PPF_HIP(dev_alloc(&vf_, (3ull * nv_ + 1) * sizeof(float)));
    PPF_HIP(hipStreamSynchronize(s));
    (void)dev_free(key);
    (void)dev_free(head);
    (void)dev_free(sid);
    (void)dev_free(scratch);
    return PPF_OK;
}

int ElasticAssembler::strain_limit(const ppf_shell &sh, float eig_eps, float *force,
                                   float *value, hipStream_t s) {
    if (!sh.f_strainlimit || !nf_) {
        return PPF_OK;
    }
    if (!slblk_) {
        PPF_HIP(dev_alloc(&slblk_, 9 * (size_t)kFacePairs * nf_ * sizeof(float)));
        PPF_HIP(dev_alloc(&slf_, (9 * (size_t)nf_ + nf_ / 4 + 1) * sizeof(float)));
    }
    uint8_t *active = reinterpret_cast<uint8_t *>(slf_ + 9 * (size_t)nf_);
    k_sl_face<<<cdiv(nf_, kNT), kNT, 0, s>>>(sh, rp_, ci_, value, eig_eps, slblk_, slf_, active);
    PPF_HIP(hipGetLastError());
    if (nnz_) {
        k_sl_gather_blocks<<<cdiv(nnz_, kNT), kNT, 0, s>>>(nnz_, nf_, slot_off_, slot_src_,
                                                          slblk_, active, value);
        PPF_HIP(hipGetLastError());
    }
    if (nv_) {
        k_sl_gather_force<<<cdiv(nv_, kNT), kNT, 0, s>>>(nv_, nf_, vtx_off_, vtx_src_, slf_,
                                                         active, force);
        PPF_HIP(hipGetLastError());
    }
    return PPF_OK;
}

int ElasticAssembler::plasticity(const ppf_shell &sh, float dt, const uint32_t *vscene,
                                 const uint32_t *ok, float *inv_rest, float *h_rest,
                                 hipStream_t s) {
    if ((sh.f_plasticity && !sh.f_plastic_threshold) ||
        (sh.h_plasticity && !sh.h_plastic_threshold)) {
        set_error("elastic: plasticity rates without thresholds");
        return PPF_ERR_ARG;
    }
    if (sh.f_plasticity && nf_) {
        k_plastic_face<<<cdiv(nf_, kNT), kNT, 0, s>>>(sh, dt, vscene, ok, inv_rest);
        PPF_HIP(hipGetLastError());
    }
    if (sh.h_plasticity && nh_) {
        k_plastic_hinge<<<cdiv(nh_, kNT), kNT, 0, s>>>(sh, dt, vscene, ok, h_rest);
        PPF_HIP(hipGetLastError());
    }
    return PPF_OK;
}

int ElasticAssembler::toi_word(uint32_t **out) {
    if (!toi_word_) {
        PPF_HIP(dev_alloc(&toi_word_, sizeof(uint32_t)));
    }
    *out = toi_word_;
    return PPF_OK;
}

int ElasticAssembler::strain_limit_toi(const ppf_shell &sh, const float *x0, const float *x1,
                                       float max_t, const uint32_t *vscene,
                                       uint32_t *toi_bits, hipStream_t s) {
    if (!sh.f_strainlimit || !nf_) {
        return PPF_OK;
    }
    k_sl_toi<<<cdiv(nf_, kNT), kNT, 0, s>>>(sh, x0, x1, max_t, vscene, toi_bits);
    PPF_HIP(hipGetLastError());
    return PPF_OK;
}

int ElasticAssembler::assemble(const ppf_shell &sh, const ppf_elastic_params &p, float *force,
                               float *cdiag, float *value, hipStream_t s) {
    uint8_t *factive = reinterpret_cast<uint8_t *>(ff_ + 9ull * nf_ + 1);
    uint8_t *hactive = factive + nf_;
    if (nv_) {
        k_vertex<<<cdiv(nv_, kNT), kNT, 0, s>>>(sh, p, vf_, cdiag);
        PPF_HIP(hipGetLastError());
    }
    if (nf_) {
        k_face<<<cdiv(nf_, kNT), kNT, 0, s>>>(sh, fblk_, ff_, factive);
        PPF_HIP(hipGetLastError());
    }
    if (nh_) {
        k_hinge<<<cdiv(nh_, 64), 64, 0, s>>>(sh, hblk_, hf_, hactive);
        PPF_HIP(hipGetLastError());
    }
    if (nnz_) {
        k_gather_blocks<<<cdiv(nnz_, kNT), kNT, 0, s>>>(nnz_, nf_, nh_, slot_off_, slot_src_,
                                                       fblk_, hblk_, factive, hactive, value);
        PPF_HIP(hipGetLastError());
    }
    if (nv_) {
        k_gather_force<<<cdiv(nv_, kNT), kNT, 0, s>>>(nv_, nf_, vtx_off_, vtx_src_, vf_, ff_,
                                                      hf_, factive, hactive, force);
        PPF_HIP(hipGetLastError());
    }
    return PPF_OK;
}

} // namespace ppf
