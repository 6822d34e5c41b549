// oracle_math.hpp -- fp32 vector / 3x3 helpers for the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY. The oracle is the checker of the HIP path; the
// product never links it. Every routine is plain single-precision C++ compiled
// with -ffp-contract=off, so each expression rounds exactly where it is written.
//
// Matrices are column-major 3x3 (m[r + 3*c]), the layout of the reference's
// Mat3x3f (crates/ppf-cts-solver/src/cpp/linalg/smat.hpp).
#pragma once
#include <cmath>
#include <cstdint>

namespace orc {

struct V3 {
    float x[3];
    float &operator[](int i) { return x[i]; }
    float operator[](int i) const { return x[i]; }
};
struct M3 {
    float m[9];
    float &operator()(int r, int c) { return m[r + 3 * c]; }
    float operator()(int r, int c) const { return m[r + 3 * c]; }
};

inline V3 v3(float a, float b, float c) { return V3{{a, b, c}}; }
inline V3 sub(const V3 &a, const V3 &b) {
    return v3(a[0] - b[0], a[1] - b[1], a[2] - b[2]);
}
inline V3 add(const V3 &a, const V3 &b) {
    return v3(a[0] + b[0], a[1] + b[1], a[2] + b[2]);
}
inline V3 scale(float s, const V3 &a) { return v3(s * a[0], s * a[1], s * a[2]); }
// Left-to-right accumulation from zero, as smat.hpp's dot/squaredNorm.
inline float dot(const V3 &a, const V3 &b) {
    float acc = 0.0f;
    acc = acc + a[0] * b[0];
    acc = acc + a[1] * b[1];
    acc = acc + a[2] * b[2];
    return acc;
}
inline float sqnorm(const V3 &a) { return dot(a, a); }
inline V3 cross(const V3 &a, const V3 &b) {
    return v3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
              a[0] * b[1] - a[1] * b[0]);
}
// Zero stays zero (no 0/0), smat.hpp:119-134.
inline V3 normalized(const V3 &a) {
    float n = std::sqrt(sqnorm(a));
    V3 r = a;
    if (n > 0.0f) {
        r[0] = a[0] / n;
        r[1] = a[1] / n;
        r[2] = a[2] / n;
    }
    return r;
}
inline V3 col(const M3 &a, int c) { return v3(a(0, c), a(1, c), a(2, c)); }
inline M3 identity() {
    M3 r{};
    r(0, 0) = r(1, 1) = r(2, 2) = 1.0f;
    return r;
}
// y = A x with the reference's UnrolledMat3x3f term order (solver.cu:196-203).
inline V3 matvec(const float *a, const float *x) {
    return v3(a[0] * x[0] + a[3] * x[1] + a[6] * x[2],
              a[1] * x[0] + a[4] * x[1] + a[7] * x[2],
              a[2] * x[0] + a[5] * x[1] + a[8] * x[2]);
}
// y = A^T x (solver.cu:204-210).
inline V3 matvec_t(const float *a, const float *x) {
    return v3(a[0] * x[0] + a[1] * x[1] + a[2] * x[2],
              a[3] * x[0] + a[4] * x[1] + a[5] * x[2],
              a[6] * x[0] + a[7] * x[1] + a[8] * x[2]);
}

} // namespace orc
