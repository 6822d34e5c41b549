// linalg.cpp -- CPU oracle: closed-form symmetric 3x3 eigensolve and the
// block-Jacobi diagonal-block inverse.
//
// TEST INFRASTRUCTURE ONLY (the checker of the HIP path, never shipped).
//
// Restates, op for op in fp32:
//   * linalg::eig::symm3x3            crates/ppf-cts-solver/src/cpp/linalg/eigsolve.hpp:195-232
//     with eigvalues3                 eigsolve.hpp:97-120
//          pick_largest / find_ortho3x3  eigsolve.hpp:122-155
//          eigvectors3x3              eigsolve.hpp:157-193
//   * solver::invert                  crates/ppf-cts-solver/src/cpp/solver/solver.cu:357-400
// Host cos is the library cosf, as the reference's own host build
// (float_math.hpp:55-62 takes std::cos off the device).
#include "oracle_math.hpp"
#include <cstdint>

namespace orc {

static const float kPi = 3.14159265358979323846f;

// Eigenvalues of a max-abs-scaled symmetric matrix, DESCENDING.
static V3 eigvalues3(const M3 &A) {
    float p1 = A(0, 1) * A(0, 1) + A(0, 2) * A(0, 2) + A(1, 2) * A(1, 2);
    float tr = 0.0f;
    tr = tr + A(0, 0);
    tr = tr + A(1, 1);
    tr = tr + A(2, 2);
    float q = tr / 3.0f;
    float d0 = A(0, 0) - q, d1 = A(1, 1) - q, d2 = A(2, 2) - q;
    float p2 = d0 * d0 + d1 * d1 + d2 * d2 + 2.0f * p1;
    float p = std::sqrt(p2 / 6.0f);
    if (std::fabs(p) < 1e-8f) {
        return v3(0.0f, 0.0f, 0.0f);
    }
    float ip = 1.0f / p;
    M3 B;
    for (int c = 0; c < 3; ++c) {
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
    if (r < -1.0f) {
        r = -1.0f;
    } else if (r > 1.0f) {
        r = 1.0f;
    }
    float phi = std::acos(r) / 3.0f;
    float e1 = q + 2.0f * p * std::cos(phi);
    float e3 = q + 2.0f * p * std::cos(phi + 2.0f * kPi / 3.0f);
    float e2 = 3.0f * q - e1 - e3;
    return v3(e1, e2, e3);
}

static V3 pick_largest(const V3 &a, const V3 &b, const V3 &c) {
    float an = sqnorm(a), bn = sqnorm(b), cn = sqnorm(c);
    if (an > bn) {
        return (an > cn) ? a : c;
    }
    return (bn > cn) ? b : c;
}

struct Ortho {
    V3 v1, v2;
};

static Ortho find_ortho(const M3 &A) {
    const float eps = 1e-7f;
    V3 u = col(A, 0), v = col(A, 1), w = col(A, 2);
    V3 q = pick_largest(cross(u, v), cross(v, w), cross(w, u));
    Ortho o;
    if (sqnorm(q) < eps) {
        V3 pp = pick_largest(u, v, w);
        V3 x = cross(pp, v3(1.0f, 0.0f, 0.0f));
        if (sqnorm(x) < eps) {
            x = cross(pp, v3(0.0f, 1.0f, 0.0f));
        }
        V3 y = cross(pp, x);
        o.v1 = normalized(x);
        o.v2 = normalized(y);
    } else {
        o.v1 = normalized(q);
        o.v2 = v3(0.0f, 0.0f, 0.0f);
    }
    return o;
}

static M3 shifted(const M3 &A, float l) {
    M3 S = A;
    for (int k = 0; k < 3; ++k) {
        S(k, k) = A(k, k) - l * 1.0f;
    }
    return S;
}

static M3 eigvectors3(const M3 &A, const V3 &lmd) {
    Ortho uv = find_ortho(shifted(A, lmd[0]));
    if (sqnorm(uv.v2) == 0.0f) {
        Ortho t = find_ortho(shifted(A, lmd[1]));
        V3 g = sub(t.v1, scale(dot(uv.v1, t.v1), uv.v1));
        uv.v2 = normalized(g);
    }
    V3 w = normalized(cross(uv.v1, uv.v2));
    uv.v2 = cross(w, uv.v1);
    if (sqnorm(uv.v1) < 0.5f || sqnorm(w) < 0.5f) {
        return identity();
    }
    M3 R;
    for (int r = 0; r < 3; ++r) {
        R(r, 0) = uv.v1[r];
        R(r, 1) = uv.v2[r];
        R(r, 2) = w[r];
    }
    return R;
}

// Eigenvalues ASCENDING, eigenvectors as columns.
void symm3x3(const M3 &A, V3 &val, M3 &vec) {
    float s = 0.0f;
    for (int i = 0; i < 9; ++i) {
        s = std::fmax(s, std::fabs(A.m[i]));
    }
    if (s <= 0.0f) {
        val = v3(0.0f, 0.0f, 0.0f);
        vec = identity();
        return;
    }
    float is = 1.0f / s;
    M3 B;
    for (int i = 0; i < 9; ++i) {
        B.m[i] = is * A.m[i];
    }
    V3 lmd = eigvalues3(B);
    M3 ev;
    if (sqnorm(lmd) > 0.0f) {
        ev = eigvectors3(B, lmd);
    } else {
        ev = identity();
        float tr = 0.0f;
        tr = tr + B(0, 0);
        tr = tr + B(1, 1);
        tr = tr + B(2, 2);
        float q = tr / 3.0f;
        lmd = v3(q, q, q);
    }
    val = v3(s * lmd[2], s * lmd[1], s * lmd[0]);
    for (int r = 0; r < 3; ++r) {
        vec(r, 0) = ev(r, 2);
        vec(r, 1) = ev(r, 1);
        vec(r, 2) = ev(r, 0);
    }
}

// Block-Jacobi inverse of one symmetric diagonal block, eigenvalues floored at
// 1e-6 * lambda_max. Returns false where the reference traps (non-finite or
// non-positive lambda_max).
bool bj_invert(const float *in, float *out) {
    M3 sym;
    for (int a = 0; a < 3; ++a) {
        sym(a, a) = in[a + 3 * a];
        for (int b = a + 1; b < 3; ++b) {
            float s = 0.5f * (in[a + 3 * b] + in[b + 3 * a]);
            sym(a, b) = s;
            sym(b, a) = s;
        }
    }
    V3 lam;
    M3 Q;
    symm3x3(sym, lam, Q);
    float lmax = std::fmax(lam[0], std::fmax(lam[1], lam[2]));
    bool ok = std::isfinite(lmax) && lmax > 0.0f;
    float mi[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const float floor_ = lmax * 1.0e-6f;
    for (int k = 0; k < 3; ++k) {
        float il = 1.0f / std::fmax(lam[k], floor_);
        float q0 = Q(0, k), q1 = Q(1, k), q2 = Q(2, k);
        mi[0] += il * q0 * q0;
        mi[3] += il * q0 * q1;
        mi[6] += il * q0 * q2;
        mi[4] += il * q1 * q1;
        mi[7] += il * q1 * q2;
        mi[8] += il * q2 * q2;
    }
    mi[1] = mi[3];
    mi[2] = mi[6];
    mi[5] = mi[7];
    for (int i = 0; i < 9; ++i) {
        out[i] = mi[i];
    }
    return ok;
}

} // namespace orc

extern "C" {

void oracle_symm3x3(const float *a9, float *val3, float *vec9) {
    orc::M3 A;
    for (int i = 0; i < 9; ++i) {
        A.m[i] = a9[i];
    }
    orc::V3 v;
    orc::M3 Q;
    orc::symm3x3(A, v, Q);
    for (int i = 0; i < 3; ++i) {
        val3[i] = v[i];
    }
    for (int i = 0; i < 9; ++i) {
        vec9[i] = Q.m[i];
    }
}

// Returns the number of blocks whose inverse would trap in the reference.
uint32_t oracle_bj_invert(uint32_t n, const float *in9, float *out9) {
    uint32_t bad = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (!orc::bj_invert(in9 + 9 * (size_t)i, out9 + 9 * (size_t)i)) {
            ++bad;
        }
    }
    return bad;
}
}
