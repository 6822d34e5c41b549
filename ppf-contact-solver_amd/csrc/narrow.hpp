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
    // All stage-0 pairs; writes min(count, cap) pairs to `out` (may be null to
    // count only). Returns PPF_ERR_CAPACITY if a traversal stack overflowed.
    int proximity(const ppf_contact_mesh &m, ppf_pair *out, uint64_t cap,
                  uint64_t *count, hipStream_t s);
    Lbvh &tree(int kind) { return kind == 0 ? face_ : (kind == 1 ? edge_ : vert_); }

  private:
    int reserve(uint32_t queries);
    Lbvh face_, edge_, vert_;
    uint32_t qcap_ = 0;
    uint32_t *cnt_ = nullptr, *ofs_ = nullptr, *scratch_ = nullptr;
    uint32_t *flags_ = nullptr, *h_flags_ = nullptr;
};

} // namespace ppf
