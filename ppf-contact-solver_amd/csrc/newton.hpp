// newton.hpp -- one time step of the contact solver (internal interface).
#pragma once
#include <vector>

#include "assemble.hpp"
#include "elastic.hpp"
#include "isect.hpp"
#include "learned.hpp"
#include "linesearch.hpp"
#include "narrow.hpp"
#include "pcg.hpp"

namespace ppf {

// Owns every per-step buffer and stage of the Newton loop of one scene, or of
// many independent scenes stepped together (one launch per stage over all of
// them); the scene's arrays stay with the caller.
class Stepper {
  public:
    ~Stepper();
    int create(const ppf_scene &sc, hipStream_t s);
    // Re-read the pin set (mesh.vflag bit0, fix_pos) after it changed.
    int refresh_pins(const ppf_scene &sc, hipStream_t s);
    int advance(const ppf_scene &sc, const ppf_step_params &p, ppf_step_result *out,
                hipStream_t s);
    // One time step of every scene of a many-scene mesh (sc.mesh.vscene):
    // prev_dt and out hold one entry per scene (p.prev_dt is not read).
    int advance_scenes(const ppf_scene &sc, const ppf_step_params &p, const float *prev_dt,
                       ppf_step_result *out, hipStream_t s);
    uint32_t scenes() const { return ns_; }
    const float *eval_x() const { return ex_; }
    uint32_t nv() const { return nv_; }
    // Learned preconditioner for the Newton solves (null: block Jacobi).
    void set_lpc(Lpc *lpc) { lpc_ = lpc; }
    static constexpr uint32_t kMaxRecords = 256;
    // Records of the last post-step intersection check.
    uint32_t records(ppf_isect_record *out, uint32_t cap) const;
    const ContactAssembler &contact() const { return contact_; }
    // The last Newton system solved (merged upper blocks after Dirichlet
    // elimination -- no block touches a pinned vertex --, per-row diagonal,
    // right-hand side).
    void system(uint32_t *nnz, const uint32_t **rp, const uint32_t **ci, const float **val,
                const float **cdiag, const float **rhs) const {
        *nnz = sys_nnz_, *rp = mrp_, *ci = mci_, *val = mval_, *cdiag = cdiag_, *rhs = force_;
    }
    // Row order of the solve (new row k = vertex order[k]), null before create.
    const uint32_t *solve_order() const { return order_; }

  private:
    int newton_iteration(const ppf_scene &sc, const ppf_step_params &p, float dt,
                         bool first_iteration, ppf_step_result *out, float *toi_out,
                         float *recale_out, hipStream_t s);
    int reserve_pairs(uint64_t n);
    int reserve_merged(uint64_t nnz);
    int reserve_lifts(uint64_t nl);
    void release();
    uint32_t nv_ = 0, ne_ = 0, nf_ = 0, nh_ = 0;
    // scenes: count, the caller's per-vertex scene ids, the first vertex of
    // each scene (host), per-scene step scalars (device and pinned host: 5
    // floats and 3 words per scene) and the scenes still in their Newton loop
    uint32_t ns_ = 1;
    const uint32_t *vscene_ = nullptr;
    std::vector<uint32_t> scene_row_;
    float *sf_ = nullptr, *h_sf_ = nullptr;
    uint32_t *su_ = nullptr, *h_su_ = nullptr;
    std::vector<uint8_t> active_;
    // analytic colliders of the step (device copies of the scene's arrays)
    ppf_sphere *sph_ = nullptr;
    ppf_floor *flr_ = nullptr;
    uint32_t sph_cap_ = 0, flr_cap_ = 0, nsph_ = 0, nflr_ = 0;
    ElasticAssembler elastic_;
    Narrow narrow_;
    ContactAssembler contact_;
    LineSearch ls_;
    IntersectionCheck isect_;
    float *ex_ = nullptr, *target_ = nullptr, *dx_ = nullptr, *force_ = nullptr;
    float *cdiag_ = nullptr, *eval_ = nullptr, *cgrad_ = nullptr;
    uint32_t *mask_ = nullptr, *scal_ = nullptr, *h_scal_ = nullptr;
    // row order of the solve (mesh pieces, each in Morton order), set once
    uint32_t *order_ = nullptr;
    bool order_changed_ = false;
    ppf_pair *pairs_ = nullptr;
    uint64_t pair_cap_ = 0;
    uint32_t *mrp_ = nullptr, *mci_ = nullptr, *mcnt_ = nullptr, *scan_ = nullptr;
    float *mval_ = nullptr;
    uint64_t mcap_ = 0;
    // Dirichlet mirror lifts: per-row counts / offsets, records (grown)
    uint32_t *lcnt_ = nullptr, *lrp_ = nullptr, *lbuf_ = nullptr;
    uint64_t lcap_ = 0;
    uint32_t npinned_ = 0;
    PcgPlanImpl *plan_ = nullptr;
    Lpc *lpc_ = nullptr;
    uint32_t sys_nnz_ = 0;
    ppf_isect_record records_[kMaxRecords] = {};
    uint32_t records_count_ = 0;
    hipEvent_t ev_[8] = {};
};

} // namespace ppf
