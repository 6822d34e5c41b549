// linesearch.hpp -- contact CCD line search (internal interface).
#pragma once
#include "lbvh.hpp"
#include "ppf_mi355x.h"

namespace ppf {

// Owns the swept trees of the dynamic mesh (faces, edges, vertices), the
// trees of the static collision mesh and the result words.
class LineSearch {
  public:
    ~LineSearch();
    // A mesh of many scenes also returns every scene's toi and overlap flag
    // (host arrays of m.nscene; the pair sweeps never cross scenes).
    int run(const ppf_contact_mesh &m, const ppf_static_mesh *sm, const ppf_ls_params &p,
            bool rebuild, float *toi, int *overlap, hipStream_t s, float *scene_toi = nullptr,
            int *scene_overlap = nullptr, uint32_t *pin_bad = nullptr);

  private:
    Lbvh face_, edge_, vert_, sface_, sedge_;
    const float *static_x_ = nullptr;
    uint32_t *out_ = nullptr, *h_out_ = nullptr;
    uint32_t out_cap_ = 0;
};

// Batched single-pair ACCD (kind 0 point-triangle, 1 point-edge,
// 2 point-point, 3 edge-edge).
int ccd_batch(int kind, uint32_t n, const float *x0, const float *x1, const float *offset,
              const float *ghat, float max_t, float eps, float *toi, int32_t *overlap,
              hipStream_t s);

} // namespace ppf
