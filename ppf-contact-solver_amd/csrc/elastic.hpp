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
    // Strain-limit barrier of the faces with a strain limit at sh.x
    // (stiffness from the step-start positions sh.x0 and the elastic blocks
    // already in `value`): force and the face blocks of `value` gain its
    // terms, in face order after the elastic ones. No-op when sh has none.
    int strain_limit(const ppf_shell &sh, float eig_eps, float *force, float *value,
                     hipStream_t s);
    // Strain-limit line search from x0 to x1 (largest certified t <= max_t
    // of every limited face): the minimum over each scene's faces into
    // toi_bits[scene] (float bits, atomically; preset by the caller), the
    // scene of a face being its first vertex's (vscene null: one scene).
    int strain_limit_toi(const ppf_shell &sh, const float *x0, const float *x1, float max_t,
                         const uint32_t *vscene, uint32_t *toi_bits, hipStream_t s);
    // Rest-shape creep after a committed step at sh.x (inv_rest and h_rest
    // hold the current rest data and receive the new one; they may be
    // sh.inv_rest / sh.h_rest). ok: per-scene commit flags (null: every
    // face and hinge), the scene of an element being its first vertex's
    // (vscene null: one scene). No-op without plasticity rates.
    int plasticity(const ppf_shell &sh, float dt, const uint32_t *vscene, const uint32_t *ok,
                   float *inv_rest, float *h_rest, hipStream_t s);
    // One device word owned by the assembler (the stage entry's line-search
    // minimum).
    int toi_word(uint32_t **out);

  private:
    void release();
    uint32_t nv_ = 0, nf_ = 0, nh_ = 0, nnz_ = 0;
    uint32_t *rp_ = nullptr, *ci_ = nullptr;
    uint32_t *slot_off_ = nullptr, *slot_src_ = nullptr; // block contributions
    uint32_t *vtx_off_ = nullptr, *vtx_src_ = nullptr;   // force contributions
    float *fblk_ = nullptr, *hblk_ = nullptr;             // staged element blocks
    float *ff_ = nullptr, *hf_ = nullptr;                 // staged element forces
    float *vf_ = nullptr;                                 // vertex-term forces
    float *slblk_ = nullptr, *slf_ = nullptr;             // staged strain-limit terms
    uint32_t *toi_word_ = nullptr;
};

} // namespace ppf
