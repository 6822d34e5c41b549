// elastic.hpp -- inertia, air drag, membrane and bending assembly (internal).
#pragma once
#include "common.hpp"
#include "ppf_mi355x.h"

namespace ppf {

// Owns the fixed upper-block pattern of a shell (every face and hinge vertex
// pair plus the diagonal), the contribution tables that sum element blocks
// into it in a fixed order, and the per-element staging buffers.
class ElasticAssembler {
  public:
    ~ElasticAssembler();
    int create(const ppf_shell &topo, hipStream_t s);
    int assemble(const ppf_shell &sh, const ppf_elastic_params &p, float *force,
                 float *cdiag, float *value, hipStream_t s);
    uint32_t nv() const { return nv_; }
    uint32_t nf() const { return nf_; }
    uint32_t nh() const { return nh_; }
    uint32_t nnz() const { return nnz_; }
    const uint32_t *row_offset() const { return rp_; }
    const uint32_t *col_index() const { return ci_; }

  private:
    void release();
    uint32_t nv_ = 0, nf_ = 0, nh_ = 0, nnz_ = 0;
    uint32_t *rp_ = nullptr, *ci_ = nullptr;
    uint32_t *slot_off_ = nullptr, *slot_src_ = nullptr; // block contributions
    uint32_t *vtx_off_ = nullptr, *vtx_src_ = nullptr;   // force contributions
    float *fblk_ = nullptr, *hblk_ = nullptr;             // staged element blocks
    float *ff_ = nullptr, *hf_ = nullptr;                 // staged element forces
    float *vf_ = nullptr;                                 // vertex-term forces
};

} // namespace ppf
