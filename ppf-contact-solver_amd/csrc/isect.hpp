// isect.hpp -- post-step intersection check (internal interface).
#pragma once
#include "lbvh.hpp"
#include "ppf_mi355x.h"

namespace ppf {

// Owns the collision-mesh face tree and the record buffer. The dynamic
// trees are the narrow phase's, built at the checked positions.
class IntersectionCheck {
  public:
    ~IntersectionCheck();
    int run(const BvhView &face, const BvhView &edge, const BvhView &vert,
            const ppf_contact_mesh &m, const ppf_static_mesh *sm, const float *vmass,
            const float *emass, const float *fmass, uint32_t surface_vert_count,
            ppf_isect_record *records, uint32_t cap, uint32_t *count, hipStream_t s,
            uint32_t *scene_count = nullptr);

  private:
    Lbvh cm_face_;
    const float *cm_x_ = nullptr;
    uint32_t cm_nf_ = 0;
    ppf_isect_record *rec_ = nullptr;
    uint32_t rec_cap_ = 0;
    uint32_t *ctr_ = nullptr, *h_ctr_ = nullptr;
    uint32_t ctr_cap_ = 0;
};

} // namespace ppf
