// eig3.hpp -- closed-form symmetric 3x3 eigensolve and the block-Jacobi
// inverse, on the device.
//
// The algorithm is the reference's (linalg::eig::symm3x3,
// crates/ppf-cts-solver/src/cpp/linalg/eigsolve.hpp:97-232, and
// solver::invert, crates/ppf-cts-solver/src/cpp/solver/solver.cu:357-400):
// max-abs scaling, trigonometric (Cardano) eigenvalues with the acos argument
// clamped, cross-product eigenvectors with a Gram-Schmidt repair for a
// clustered pair, identity for a degenerate triad, ascending output order, and
// eigenvalues floored at 1e-6 * lambda_max in the inverse. It is kept op for op
// because the preconditioner it produces decides PCG iteration counts.
#pragma once
#include "common.hpp"

namespace ppf {

struct m3 {
    float m[9];
    PPF_HD float &operator()(int r, int c) { return m[r + 3 * c]; }
    PPF_HD float operator()(int r, int c) const { return m[r + 3 * c]; }
};

PPF_HD float sqn3(f3 a) { return dot3(a, a); }
PPF_HD f3 nrm3(f3 a) {
    float n = sqrtf(sqn3(a));
    if (n > 0.0f) {
        return mk3(a.x / n, a.y / n, a.z / n);
    }
    return a;
}
PPF_HD f3 col3(const m3 &a, int c) { return mk3(a(0, c), a(1, c), a(2, c)); }
PPF_HD m3 eye3() {
    m3 r;
    for (int i = 0; i < 9; ++i) {
        r.m[i] = 0.0f;
    }
    r.m[0] = r.m[4] = r.m[8] = 1.0f;
    return r;
}

namespace eig_detail {

constexpr float kPi = 3.14159265358979323846f;

PPF_HD f3 values_desc(const m3 &A) {
    float p1 = A(0, 1) * A(0, 1) + A(0, 2) * A(0, 2) + A(1, 2) * A(1, 2);
    float tr = 0.0f;
    tr = tr + A(0, 0);
    tr = tr + A(1, 1);
    tr = tr + A(2, 2);
    float q = tr / 3.0f;
    float d0 = A(0, 0) - q, d1 = A(1, 1) - q, d2 = A(2, 2) - q;
    float p2 = d0 * d0 + d1 * d1 + d2 * d2 + 2.0f * p1;
    float p = sqrtf(p2 / 6.0f);
    if (fabsf(p) < 1e-8f) {
        return mk3(0.0f, 0.0f, 0.0f);
    }
    float ip = 1.0f / p;
    m3 B;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            float qi = (r == c) ? q * 1.0f : q * 0.0f;
            B(r, c) = ip * (A(r, c) - qi);
        }
    }
    const float *m = B.m;
    float det = m[0] * (m[4] * m[8] - m[7] * m[5]) -
                m[3] * (m[1] * m[8] - m[7] * m[2]) +
                m[6] * (m[1] * m[5] - m[4] * m[2]);
    float r = det / 2.0f;
    r = r < -1.0f ? -1.0f : (r > 1.0f ? 1.0f : r);
    float phi = acosf(r) / 3.0f;
    float e1 = q + 2.0f * p * cosf(phi);
    float e3 = q + 2.0f * p * cosf(phi + 2.0f * kPi / 3.0f);
    float e2 = 3.0f * q - e1 - e3;
    return mk3(e1, e2, e3);
}

PPF_HD f3 largest(f3 a, f3 b, f3 c) {
    float an = sqn3(a), bn = sqn3(b), cn = sqn3(c);
    if (an > bn) {
        return (an > cn) ? a : c;
    }
    return (bn > cn) ? b : c;
}

PPF_HD void ortho(const m3 &A, f3 &v1, f3 &v2) {
    const float eps = 1e-7f;
    f3 u = col3(A, 0), v = col3(A, 1), w = col3(A, 2);
    f3 q = largest(cross3(u, v), cross3(v, w), cross3(w, u));
    if (sqn3(q) < eps) {
        f3 pp = largest(u, v, w);
        f3 x = cross3(pp, mk3(1.0f, 0.0f, 0.0f));
        if (sqn3(x) < eps) {
            x = cross3(pp, mk3(0.0f, 1.0f, 0.0f));
        }
        f3 y = cross3(pp, x);
        v1 = nrm3(x);
        v2 = nrm3(y);
    } else {
        v1 = nrm3(q);
        v2 = mk3(0.0f, 0.0f, 0.0f);
    }
}

PPF_HD m3 shift(const m3 &A, float l) {
    m3 S = A;
    S(0, 0) = A(0, 0) - l * 1.0f;
    S(1, 1) = A(1, 1) - l * 1.0f;
    S(2, 2) = A(2, 2) - l * 1.0f;
    return S;
}

PPF_HD m3 vectors(const m3 &A, f3 lmd) {
    f3 v1, v2;
    ortho(shift(A, lmd.x), v1, v2);
    if (sqn3(v2) == 0.0f) {
        f3 t1, t2;
        ortho(shift(A, lmd.y), t1, t2);
        f3 g = t1 - dot3(v1, t1) * v1;
        v2 = nrm3(g);
    }
    f3 w = nrm3(cross3(v1, v2));
    v2 = cross3(w, v1);
    if (sqn3(v1) < 0.5f || sqn3(w) < 0.5f) {
        return eye3();
    }
    m3 R;
    R(0, 0) = v1.x, R(1, 0) = v1.y, R(2, 0) = v1.z;
    R(0, 1) = v2.x, R(1, 1) = v2.y, R(2, 1) = v2.z;
    R(0, 2) = w.x, R(1, 2) = w.y, R(2, 2) = w.z;
    return R;
}

} // namespace eig_detail

// Eigenvalues ascending (val), eigenvectors as the columns of vec.
PPF_HD void symm3x3(const m3 &A, f3 &val, m3 &vec) {
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        s = fmaxf(s, fabsf(A.m[i]));
    }
    if (s <= 0.0f) {
        val = mk3(0.0f, 0.0f, 0.0f);
        vec = eye3();
        return;
    }
    float is = 1.0f / s;
    m3 B;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        B.m[i] = is * A.m[i];
    }
    f3 lmd = eig_detail::values_desc(B);
    m3 ev;
    if (sqn3(lmd) > 0.0f) {
        ev = eig_detail::vectors(B, lmd);
    } else {
        ev = eye3();
        float tr = 0.0f;
        tr = tr + B(0, 0);
        tr = tr + B(1, 1);
        tr = tr + B(2, 2);
        float q = tr / 3.0f;
        lmd = mk3(q, q, q);
    }
    val = mk3(s * lmd.z, s * lmd.y, s * lmd.x);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        vec(r, 0) = ev(r, 2);
        vec(r, 1) = ev(r, 1);
        vec(r, 2) = ev(r, 0);
    }
}

// Block-Jacobi inverse of a (numerically) symmetric block. Returns false
// where the reference's asserts would trap (lambda_max non-finite or <= 0).
PPF_HD bool bj_invert(const float *in, float *out) {
    m3 sym;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        sym(a, a) = in[a + 3 * a];
#pragma unroll
        for (int b = a + 1; b < 3; ++b) {
            float s = 0.5f * (in[a + 3 * b] + in[b + 3 * a]);
            sym(a, b) = s;
            sym(b, a) = s;
        }
    }
    f3 lam;
    m3 Q;
    symm3x3(sym, lam, Q);
    float lmax = fmaxf(lam.x, fmaxf(lam.y, lam.z));
    bool ok = isfinite(lmax) && lmax > 0.0f;
    const float fl = lmax * 1.0e-6f;
    float l[3] = {lam.x, lam.y, lam.z};
    float m00 = 0, m01 = 0, m02 = 0, m11 = 0, m12 = 0, m22 = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float il = 1.0f / fmaxf(l[k], fl);
        float q0 = Q(0, k), q1 = Q(1, k), q2 = Q(2, k);
        m00 += il * q0 * q0;
        m01 += il * q0 * q1;
        m02 += il * q0 * q2;
        m11 += il * q1 * q1;
        m12 += il * q1 * q2;
        m22 += il * q2 * q2;
    }
    out[0] = m00, out[1] = m01, out[2] = m02;
    out[3] = m01, out[4] = m11, out[5] = m12;
    out[6] = m02, out[7] = m12, out[8] = m22;
    return ok;
}

} // namespace ppf
