// assemble.hpp -- contact force / Hessian assembly (internal interface).
#pragma once
#include "common.hpp"
#include "ppf_mi355x.h"

namespace ppf {

struct PairEval;

// Owns the per-pair evaluations, the sort records and the assembled
// upper-triangle contact blocks (valid until the next assemble).
class ContactAssembler {
  public:
    ~ContactAssembler();
    int assemble(const ppf_contact_mesh &m, const ppf_contact_params &cp,
                 const ppf_pair *pairs, uint64_t npairs, float *grad, hipStream_t s);
    uint64_t nnz() const { return nnz_; }
    const uint32_t *row_offset() const { return rp_; }
    const uint32_t *rows() const { return out_row_; }
    const uint32_t *cols() const { return out_col_; }
    const float *values() const { return out_val_; }

  private:
    int reserve(uint64_t npairs, uint32_t nv);
    void release();
    PairEval *pe_ = nullptr;
    uint32_t *buf_ = nullptr, *out_row_ = nullptr, *out_col_ = nullptr, *rp_ = nullptr;
    float *out_val_ = nullptr;
    uint64_t pcap_ = 0, nnz_ = 0;
    uint32_t vcap_ = 0;
};

// Analytic sphere / floor terms of every vertex added to grad (3 nv) and the
// per-vertex diagonal cdiag (9 nv); spheres and floors are device arrays.
int analytic_assemble(const ppf_contact_mesh &m, const ppf_contact_params &cp,
                      const ppf_sphere *spheres, uint32_t nsphere, const ppf_floor *floors,
                      uint32_t nfloor, float constraint_tol, float *grad, float *cdiag,
                      hipStream_t s);

} // namespace ppf
