// pcg.hpp -- internal interface of the block-Jacobi PCG (pcg.hip).
#pragma once
#include "common.hpp"
#include "ppf_mi355x.h"

namespace ppf {

struct PcgPlanImpl;

int pcg_plan_create(uint32_t n, uint32_t nnz, const uint32_t *rp,
                    const uint32_t *col, hipStream_t st, PcgPlanImpl **out);
// A plan over the block-diagonal union of nscene systems: scene c owns rows
// scene_row[c] .. scene_row[c + 1] (host array, nscene + 1 entries) and its
// entries couple only those rows. One launch per kernel serves all scenes;
// each scene keeps its own scalars, stopping rule and result.
int pcg_plan_create_scenes(uint32_t n, uint32_t nnz, const uint32_t *rp, const uint32_t *col,
                           uint32_t nscene, const uint32_t *scene_row, hipStream_t st,
                           PcgPlanImpl **out);
uint32_t pcg_plan_scenes(const PcgPlanImpl *pl);
// Solve in another row order: new row k is original row order[k] (device
// array of n, owned by the caller and left untouched until the order changes).
// Takes effect at the next pattern (pcg_plan_repattern); values, right-hand
// side and iterate stay in the original order at every entry point.
int pcg_plan_set_order(PcgPlanImpl *pl, const uint32_t *order, hipStream_t st);
bool pcg_plan_ordered(const PcgPlanImpl *pl);
// The system in the plan's row order (the caller's arrays when the plan keeps
// the original order): what a factor computed on pcg_pattern() must see.
// b may be null.
int pcg_order_system(PcgPlanImpl *pl, const float *val, const float *cdiag, const float *b,
                     const float **oval, const float **ocdiag, const float **ob,
                     hipStream_t st);
// Rows per tile of the SpMV and of the vector kernels (every scene cut into
// tiles; 0, 0: one resident round over a one-scene plan).
int pcg_plan_set_tiles(PcgPlanImpl *pl, uint32_t r1, uint32_t r2, hipStream_t st);
// Fix the SpMV form (PPF_SPMV_*; AUTO: the plan's own choice per pattern) and
// re-apply it to the current pattern; the form in use.
int pcg_plan_set_spmv_form(PcgPlanImpl *pl, int form, hipStream_t st);
int pcg_plan_spmv_form(const PcgPlanImpl *pl);
// Measure candidate tile sizes on this system and keep the fastest.
int pcg_plan_autotune(PcgPlanImpl *pl, const float *val, const float *cdiag, const float *b,
                      float tol, uint32_t *best_r1, uint32_t *best_r2, float *best_ms,
                      hipStream_t st);
void pcg_plan_destroy(PcgPlanImpl *pl);
// Point the plan at a new pattern of the same n (buffers grow as needed; the
// iteration graph is re-captured only when the SpMV form changes).
int pcg_plan_repattern(PcgPlanImpl *pl, uint32_t nnz, const uint32_t *rp, const uint32_t *col,
                       hipStream_t st);
// The factor preconditioner M^-1 = G^T G (null: block Jacobi), bf16
// column-major blocks in 10-entry records: Gd the diagonal blocks (n), Gl the
// lower blocks by slot (Gl[s] = G_ji for the stored slot s = (i, j), i < j),
// Gr the rows of G in order (row j at j + tp[j]: G_jj, then G_ji for the
// row's mirrored entries in transpose-table order).
struct PcgFactor {
    const uint16_t *Gd, *Gl, *Gr;
};
// res: one result per scene. Many-scene plans take no factor and no history.
int pcg_solve(PcgPlanImpl *pl, const float *val, const float *cdiag,
              const float *b, float *x, float tol, uint32_t max_iter,
              float *err_hist, ppf_pcg_result *res, hipStream_t st,
              const PcgFactor *fac = nullptr);
// Pattern arrays of the plan (device): row offsets, columns, transpose table
// (per row the mirrored entries: source row and slot), diagonal slot per row.
struct PcgPattern {
    uint32_t n, nnz;
    const uint32_t *rp, *col, *tp, *tcol, *tslot, *dslot;
};
PcgPattern pcg_pattern(const PcgPlanImpl *pl);
int pcg_time_spmv(PcgPlanImpl *pl, const float *val, const float *cdiag,
                  const float *x, float *y, uint32_t reps, float *avg_ms,
                  hipStream_t st);
int pcg_spmv(PcgPlanImpl *pl, const float *val, const float *cdiag,
             const float *x, float *y, hipStream_t st);
// Bytes of the fixed device region of the allocation-free solve.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
// Plan and solve in one call on that region (no device or pinned allocation,
// no graph): the same kernels, tiles and results as a plan's solve.
int pcg_solve_noalloc(uint32_t n, uint32_t nnz, const uint32_t *rp, const uint32_t *col,
                      const float *val, const float *cdiag, const float *b, float *x,
                      float tol, uint32_t max_iter, ppf_pcg_result *res, hipStream_t st);

} // namespace ppf
