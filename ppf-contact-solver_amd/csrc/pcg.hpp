// pcg.hpp -- internal interface of the block-Jacobi PCG (pcg.hip).
#pragma once
#include "common.hpp"
#include "ppf_mi355x.h"

namespace ppf {

struct PcgPlanImpl;

int pcg_plan_create(uint32_t n, uint32_t nnz, const uint32_t *rp,
                    const uint32_t *col, hipStream_t st, PcgPlanImpl **out);
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

} // namespace ppf
