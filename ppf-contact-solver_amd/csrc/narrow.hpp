// narrow.hpp -- contact proximity narrow phase (internal interface).
#pragma once
#include "lbvh.hpp"
#include "ppf_mi355x.h"

namespace ppf {

// Face, edge and vertex trees of one contact mesh plus the per-query count /
// offset arrays of the two-pass pair output.
class Narrow {
  public:
    ~Narrow();
    int build_trees(const ppf_contact_mesh &m, hipStream_t s);
    // Refit the boxes of the last build at new positions (same topology); a
    // size change rebuilds.
    int refit_trees(const ppf_contact_mesh &m, hipStream_t s);
    // All stage-0 pairs; writes min(count, cap) pairs to `out` (may be null to
    // count only). Returns PPF_ERR_CAPACITY if a traversal stack overflowed.
    int proximity(const ppf_contact_mesh &m, ppf_pair *out, uint64_t cap,
                  uint64_t *count, hipStream_t s);
    Lbvh &tree(int kind) { return kind == 0 ? face_ : (kind == 1 ? edge_ : vert_); }
    // Collision-mesh pairs (kinds 4, 5, 6) at the positions of the last
    // build_trees; the static trees are rebuilt when the static mesh changes.
    int static_proximity(const ppf_contact_mesh &m, const ppf_static_mesh &sm,
                         const float *vmass, const float *emass, const float *fmass,
                         ppf_pair *out, uint64_t cap, uint64_t *count, hipStream_t s);

  private:
    int reserve(uint32_t queries);
    // Trees of the collision mesh (nscene: the dynamic mesh's scene count).
    int build_static(const ppf_static_mesh &sm, uint32_t nscene, hipStream_t s);
    Lbvh face_, edge_, vert_, sface_, sedge_;
    const float *static_x_ = nullptr;
    uint32_t static_nv_ = 0;
    uint32_t qcap_ = 0;
    uint32_t *cnt_ = nullptr, *ofs_ = nullptr, *scratch_ = nullptr;
    ppf_pair *stash_ = nullptr;  // the count pass's first pairs of every query
    uint32_t *redo_ = nullptr;   // queries with more pairs than the stash holds
    uint32_t *flags_ = nullptr, *h_flags_ = nullptr;
};

} // namespace ppf
