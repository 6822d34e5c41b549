// ref_driver.cpp -- extern "C" wrappers around the reference's own
// host-compilable headers, compiled from where they lie under the reference
// tree (see ../Makefile, target `ref`). This file holds only wrappers; the
// arithmetic is the reference's. TEST INFRASTRUCTURE ONLY: it pins the oracle.
#include "contact/distance.hpp"
#include "linalg/eigsolve.hpp"

// common.hpp declares the host logger the Rust side normally provides; the
// wrapped functions never log.
extern "C" void print_rust(const char *) {}

namespace {
Vec3f v3(const float *p) { return Vec3f(p[0], p[1], p[2]); }
} // namespace

extern "C" {

// linalg::eig::symm3x3 (linalg/eigsolve.hpp:195-232).
void ref_symm3x3(const float *a9, float *val3, float *vec9) {
    linalg::eig::M3 A;
    for (int i = 0; i < 9; ++i) {
        A.m[i] = a9[i];
    }
    linalg::eig::V3 v;
    linalg::eig::M3 Q;
    linalg::eig::symm3x3(A, v, Q);
    for (int i = 0; i < 3; ++i) {
        val3[i] = v[i];
    }
    for (int i = 0; i < 9; ++i) {
        vec9[i] = Q.m[i];
    }
}

// distance::point_edge_distance_coeff (contact/distance.hpp:14-26).
void ref_pe_coeff(const float *p, const float *e0, const float *e1, float *w) {
    Vec2f c = distance::point_edge_distance_coeff<float, float>(v3(p), v3(e0), v3(e1));
    w[0] = c[0];
    w[1] = c[1];
}

// distance::point_triangle_distance_coeff (contact/distance.hpp:46-81).
void ref_pt_coeff(const float *p, const float *t0, const float *t1,
                  const float *t2, float *w) {
    Vec3f c = distance::point_triangle_distance_coeff<float, float>(
        v3(p), v3(t0), v3(t1), v3(t2));
    w[0] = c[0];
    w[1] = c[1];
    w[2] = c[2];
}

// distance::edge_edge_distance_coeff (contact/distance.hpp:112-192).
void ref_ee_coeff(const float *a0, const float *a1, const float *b0,
                  const float *b1, float *w) {
    Vec4f c = distance::edge_edge_distance_coeff<float, float>(v3(a0), v3(a1),
                                                               v3(b0), v3(b1));
    for (int i = 0; i < 4; ++i) {
        w[i] = c[i];
    }
}
}
