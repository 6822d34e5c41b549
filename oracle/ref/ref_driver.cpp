// ref_driver.cpp -- extern "C" wrappers around the reference's own
// host-compilable headers, compiled from where they lie under the reference
// tree (see ../Makefile, target `ref`). This file holds only wrappers; the
// arithmetic is the reference's. TEST INFRASTRUCTURE ONLY: it pins the oracle.
#include "contact/distance.hpp"
#include "energy/model/baraffwitkin.hpp"
#include "energy/model/dihedral_angle.hpp"
#include "linalg/eigsolve.hpp"

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

// BaraffWitkin::stretch_gradient / stretch_hessian (kind 0) and
// shear_gradient / shear_hessian (kind 1), energy/model/baraffwitkin.hpp:27-226.
// F: 6 floats column-major; grad 6, hess 36 column-major.
void ref_bw(int kind, const float *f, float s, float *grad, float *hess) {
    Mat3x2f F;
    for (int c = 0; c < 2; ++c)
        for (int r = 0; r < 3; ++r) F(r, c) = f[r + 3 * c];
    Mat3x2f g = kind == 0 ? BaraffWitkin::stretch_gradient(F, s)
                          : BaraffWitkin::shear_gradient(F, s);
    Mat6x6f h = kind == 0 ? BaraffWitkin::stretch_hessian(F, s)
                          : BaraffWitkin::shear_hessian(F, s);
    for (int c = 0; c < 2; ++c)
        for (int r = 0; r < 3; ++r) grad[r + 3 * c] = g(r, c);
    for (int c = 0; c < 6; ++c)
        for (int r = 0; r < 6; ++r) hess[r + 6 * c] = h(r, c);
}

// dihedral_angle::face_compute_force_hessian (energy/model/dihedral_angle.hpp:
// 130-355) on the hinge (0, 1, 2, 3) over four points p (12 floats, stored
// order). Writes the force (12, remapped vertex order) and the 12x12 Hessian
// row-major.
void ref_hinge(const float *p, float rest, float *force, float *hess) {
    Vec3f pts[4];
    for (int k = 0; k < 4; ++k) pts[k] = v3(p + 3 * k);
    Vec<Vec3f> vertex;
    vertex.data = pts;
    vertex.size = 4;
    vertex.allocated = 4;
    Vec4u hinge(0, 1, 2, 3);
    Mat3x4f f;
    Mat12x12f H;
    dihedral_angle::face_compute_force_hessian(vertex, hinge, rest, f, H);
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 3; ++r) force[3 * c + r] = f(r, c);
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j < 12; ++j) hess[12 * i + j] = H(i, j);
}
}
