// ref_driver.cpp -- extern "C" wrappers around the reference's own
// host-compilable headers, compiled from where they lie under the reference
// tree (see ../Makefile, target `ref`). This file holds only wrappers; the
// arithmetic is the reference's. TEST INFRASTRUCTURE ONLY: it pins the oracle.
#include "linalg/eigsolve.hpp"

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
}
