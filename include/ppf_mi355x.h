/* ppf_mi355x.h -- C ABI of the MI355X-native per-Newton-step hot path.
 *
 * All pointers named d_* are device (HBM) pointers; sizes are element counts.
 * Every function returns 0 (PPF_OK) on success or a PPF_ERR_* code, with a
 * message in ppf_last_error(). `stream` is a hipStream_t (NULL = null stream).
 * 3x3 blocks are 9 floats, column-major, as the reference's Mat3x3f
 * (crates/ppf-cts-solver/src/cpp/linalg/smat.hpp).
 *
 * Block-sparse matrices follow the reference's FixedCSRMat semantics
 * (crates/ppf-cts-solver/src/cpp/csrmat/csrmat.hpp:71-88): each row i stores
 * the blocks (i, j) with j >= i in ascending j, the operator applies every
 * stored off-diagonal block both as (i, j) and transposed as (j, i), and a
 * separate per-row diagonal C_i is added (solver.cu:213-259).
 */
#ifndef PPF_MI355X_H
#define PPF_MI355X_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PPF_OK = 0,
    PPF_ERR_ARG = 1,      /* invalid argument or shape */
    PPF_ERR_HIP = 2,      /* HIP runtime error */
    PPF_ERR_NODEV = 3,    /* no gfx950 device visible */
    PPF_ERR_CAPACITY = 4, /* a fixed-capacity buffer would overflow */
    PPF_ERR_SOLVER = 5    /* solver invariant violated (reference aborts) */
};

/* Last error message of the calling thread ("" if none). */
const char *ppf_last_error(void);

/* Device and pinned-host allocations / frees the library has made since it
 * was loaded (replaces the reference's g_device_alloc_count /
 * g_device_free_count, main/main.cu:280-286, 1635-1646, which it logs per
 * step as device-alloc / device-free). Buffers grow to a high-water mark and
 * are kept, so once warmed up a time step adds nothing to either count. */
void ppf_alloc_counts(uint64_t *allocs, uint64_t *frees);

/* Statistics of the BVH walks, a measurement mode (off by default; no
 * reference counterpart). While on, the walk kernels add up, per kind:
 * launches, queries (elements that walk), nodes taken off the walk stack
 * (an internal one reads its 64-byte walk record: both children's boxes,
 * scene tags and codes), leaves among them, and sweeps (CCD pairs run by the line
 * search), and every launch of those kernels is timed with HIP events (the
 * host waits for it). Kinds: 0 line-search vertex sweeps (point-face,
 * point-point, vertex vs collision face), 1 line-search edge sweeps
 * (edge-edge, edge vs collision edge), 2 proximity vertex queries, 3
 * proximity edge queries (the count passes). Enabling zeroes the counters;
 * disabling keeps them. Not for concurrent use from several threads. */
typedef struct ppf_walk_stats {
    uint64_t launches, queries, nodes, leaves, sweeps;
    double ms;
} ppf_walk_stats;
int ppf_walk_stats_enable(int on);
int ppf_walk_stats_read(ppf_walk_stats *out /* [4] */);

/* Device allocation and release through the same counters, for code layered
 * on this library (the solver-backend boundary; the reference's
 * mem::malloc_device / mem::free_device under buffer/buffer.hpp). */
int ppf_device_alloc(void **ptr, size_t bytes);
void ppf_device_free(void *ptr);

/* Library / device identity: writes the gfx arch name of device 0. */
int ppf_device_info(char *arch, uint32_t arch_len, uint32_t *cu_count);

/* ------------------------------------------------------------------------
 * Block-Jacobi PCG (replaces solver::solve, solver.cu:2180-2321, for the
 * block-Jacobi preconditioner; the recurrence is solver::cg_device,
 * solver.cu:921-1383).
 * ---------------------------------------------------------------------- */

/* Outcome of one solve. status: 0 converged, 1 truncated (p^T A p inside its
 * own fp32 round-off; success, solver.cu:1287-1303), 2 iteration cap,
 * 3 non-finite residual, 4 p^T A p negative beyond the round-off bound,
 * 5 r^T M^-1 r non-positive or a non-invertible diagonal block (the reference
 * aborts with PPF_FATAL_SOLVER_INVARIANT on 4 and 5). */
typedef struct ppf_pcg_result {
    uint32_t iter;     /* reported iteration count (reference check schedule) */
    float resid;       /* |r|_1 / |r0|_1 at the reported iteration */
    int32_t status;
    int32_t break_cause; /* 0 none, 1 pAp, 2 rz, 3 noise (solver.cu:46-57) */
    int32_t break_iter;
    float err0;        /* |b - A x0|_1 */
} ppf_pcg_result;

typedef struct ppf_pcg_plan ppf_pcg_plan;

/* Build a solve plan for a fixed sparsity pattern: the transpose table
 * (reference: builder.rs:1336-1345 for the fixed matrix, csrmat.cu:302-433 for
 * the dynamic one) and all per-solve scratch. d_row_offset has n+1 entries,
 * d_col_index nnz entries; both must stay valid for the plan's lifetime. */
int ppf_pcg_plan_create(uint32_t n, uint32_t nnz, const uint32_t *d_row_offset,
                        const uint32_t *d_col_index, void *stream,
                        ppf_pcg_plan **out);
void ppf_pcg_plan_destroy(ppf_pcg_plan *plan);

/* Point an existing plan at another pattern with the same n (the contact
 * matrix changes pattern every Newton iteration; reference: the per-iteration
 * rebuild of the dynamic matrix in csrmat.cu:302-433). Scratch grows when the
 * new pattern is larger; the previous pattern's arrays may be freed after. */
int ppf_pcg_plan_repattern(ppf_pcg_plan *plan, uint32_t nnz, const uint32_t *d_row_offset,
                           const uint32_t *d_col_index, void *stream);

/* Solve M x = b to |r|_1/|r0|_1 < tol (x holds x0 on entry). d_value: nnz
 * blocks; d_cdiag: n blocks; d_b, d_x: 3n floats. Mirrors solver::solve +
 * solver::cg (solver.cu:1631-1675). */
int ppf_pcg_solve(ppf_pcg_plan *plan, const float *d_value,
                  const float *d_cdiag, const float *d_b, float *d_x, float tol,
                  uint32_t max_iter, ppf_pcg_result *out, void *stream);

/* ------------------------------------------------------------------------
 * Learned preconditioner. No counterpart in the reference, whose block
 * Jacobi (solver.cu:357-400) is the starting point (zero output heads) and
 * the fallback. A message-passing network over the system's 3x3-block graph
 * (node features from the diagonal and right-hand side, edge features from
 * the normalised blocks L_i^-1 A_ij L_j^-T, A_ii = L_i L_i^T; 3 layers of
 * width 64 on bf16 MFMA with fp32 accumulation) predicts a block FSAI factor
 * Ghat on A's pattern; M^-1 = G^T G with G = Ghat L^-1 is symmetric positive
 * definite for any weights. Trained on the log Kaporin number of
 * Ghat Ahat Ghat^T with Rademacher probes.
 * ---------------------------------------------------------------------- */

typedef struct ppf_lpc ppf_lpc;

/* Number of fp32 weights (row-major [in][out]: W0 16x64, c0; per layer
 * We 80x64, ce, Wn 128x64, cn; Wo 64x6, co; Wg 144x9, cg). */
uint32_t ppf_lpc_weight_count(void);
int ppf_lpc_create(ppf_lpc **out);
void ppf_lpc_destroy(ppf_lpc *h);
int ppf_lpc_set_weights(ppf_lpc *h, const float *host_weights, uint32_t count, void *stream);
int ppf_lpc_get_weights(ppf_lpc *h, float *host_weights, uint32_t count, void *stream);

/* Graph, features, network and factor for the system (plan's pattern,
 * d_value, d_cdiag, right-hand side d_b; in the plan's row order when it has
 * one, ppf_pcg_plan_set_order). */
int ppf_lpc_factor(ppf_lpc *h, ppf_pcg_plan *plan, const float *d_value,
                   const float *d_cdiag, const float *d_b, void *stream);

/* Device copies of the last factor's head outputs t (n x 6), u (nnz x 9, by
 * slot; zero on diagonal slots) and of G: diagonal blocks (n x 9) and lower
 * blocks by slot (nnz x 9, G_ji for the slot (i, j)). Any may be null. */
int ppf_lpc_outputs(ppf_lpc *h, float *d_t, float *d_u, float *d_gd, float *d_gl,
                    void *stream);

/* Loss and weight gradient (host, ppf_lpc_weight_count floats) for the
 * probes d_probes (nprobe x n x 3), after ppf_lpc_factor on the same system.
 * Plans in the original row order only. */
int ppf_lpc_gradient(ppf_lpc *h, ppf_pcg_plan *plan, const float *d_value,
                     const float *d_cdiag, const float *d_probes, uint32_t nprobe,
                     float *loss, float *host_grad, void *stream);

/* One Adam step (beta 0.9 / 0.999) on the last gradient. */
int ppf_lpc_adam(ppf_lpc *h, float lr, uint32_t step, void *stream);

/* PCG with M^-1 = G^T G from the last ppf_lpc_factor on this system; same
 * stopping rule and result as ppf_pcg_solve. */
int ppf_pcg_solve_lpc(ppf_pcg_plan *plan, ppf_lpc *h, const float *d_value,
                      const float *d_cdiag, const float *d_b, float *d_x, float tol,
                      uint32_t max_iter, ppf_pcg_result *out, void *stream);

/* Same solve, also writing |r|_1 after every iteration into d_err_hist
 * (max_iter floats) for residual-curve comparisons. */
int ppf_pcg_solve_hist(ppf_pcg_plan *plan, const float *d_value,
                       const float *d_cdiag, const float *d_b, float *d_x,
                       float tol, uint32_t max_iter, float *d_err_hist,
                       ppf_pcg_result *out, void *stream);

/* y = M x for the plan's pattern (solver::apply, solver.cu:213-259). */
int ppf_pcg_spmv(ppf_pcg_plan *plan, const float *d_value, const float *d_cdiag,
                 const float *d_x, float *d_y, void *stream);

/* Solve in another row order: new row k is original row d_order[k] (device
 * array of n, kept by the caller until the next call). The upper pattern is
 * rebuilt in that order at the next ppf_pcg_plan_repattern (blocks whose rows
 * swap order move to the other row, transposed); every entry point keeps
 * taking and returning values, right-hand sides and iterates in the original
 * order. A locality order (e.g. the vertices' Morton order) keeps the
 * transposed half's block reads in L2. d_order null: the original order. One-
 * scene plans only. No reference counterpart (the reference solves in the
 * mesh's order). */
int ppf_pcg_plan_set_order(ppf_pcg_plan *plan, const uint32_t *d_order, void *stream);

/* Batched PCG: one plan over many independent systems ("scenes") stacked
 * block-diagonally: scene c owns rows h_scene_row[c] .. h_scene_row[c + 1]
 * (host array of nscene + 1 rising row indices, 0 .. n) and its blocks couple
 * only its own rows (checked). Each kernel launch serves every scene; each
 * scene keeps its own alpha, beta, stopping rule and result, so a scene's
 * iterates do not depend on the others'. Work is cut into tiles of whole
 * rows of one scene. No reference counterpart: the reference solves one
 * scene per process (solver::solve, solver.cu:2180-2321, once per scene). */
int ppf_pcg_plan_create_batch(uint32_t n, uint32_t nnz, const uint32_t *d_row_offset,
                              const uint32_t *d_col_index, uint32_t nscene,
                              const uint32_t *h_scene_row, void *stream, ppf_pcg_plan **out);

/* Rows per tile of the SpMV (spmv_rows) and of the vector kernels
 * (vector_rows). 0, 0 on a one-scene plan: one resident round of workgroups
 * over the whole system (the default). */
int ppf_pcg_plan_set_tiles(ppf_pcg_plan *plan, uint32_t spmv_rows, uint32_t vector_rows,
                           void *stream);

/* SpMV forms of a plan. ELL: a padded per-row entry table (rows of at most 32
 * stored + mirrored blocks). SPLIT: the pattern and its transpose table read
 * directly. SYM (one-scene plans): the split form, except that a stored
 * block whose mirror row lies in the same workgroup tile is read once and
 * its transposed product handed over through LDS; results are bit-identical
 * to SPLIT. AUTO (the default) picks per pattern: ELL when the widest row is
 * at most twice the average, else SYM on a one-scene plan and SPLIT on a
 * batch. */
enum { PPF_SPMV_AUTO = 0, PPF_SPMV_ELL = 1, PPF_SPMV_SPLIT = 2, PPF_SPMV_SYM = 3 };

/* Fix the SpMV form of a plan (PPF_SPMV_*) and apply it to the current
 * pattern; *form_in_use (may be null) receives the form the plan now runs.
 * Measurement entry point, no reference counterpart (the reference has one
 * operator form, solver.cu:213-259). */
int ppf_pcg_plan_set_spmv_form(ppf_pcg_plan *plan, int form, int *form_in_use, void *stream);

/* Autotuner: time the same capped solve (24 iterations) of this system
 * under candidate tile sizes (SpMV 128..2048 rows, vector kernels 256..2048)
 * and keep the fastest. h_tile_rows[2] receives the chosen sizes and *h_ms
 * (may be null) the chosen candidate's time. */
int ppf_pcg_plan_autotune(ppf_pcg_plan *plan, const float *d_value, const float *d_cdiag,
                          const float *d_b, float tol, uint32_t *h_tile_rows, float *h_ms,
                          void *stream);

/* Solve every scene of a batch plan: h_out receives nscene results. The
 * stopping rule and statuses are ppf_pcg_solve's, per scene. */
int ppf_pcg_solve_batch(ppf_pcg_plan *plan, const float *d_value, const float *d_cdiag,
                        const float *d_b, float *d_x, float tol, uint32_t max_iter,
                        ppf_pcg_result *h_out, void *stream);

/* Time `reps` back-to-back launches of the PCG's SpMV kernel (y = M x with
 * the x.Mx partials, the kernel of one CG iteration) with HIP events on
 * `stream`; *avg_ms receives the mean duration of one launch. Measurement
 * entry point, no reference counterpart. */
int ppf_pcg_time_spmv(ppf_pcg_plan *plan, const float *d_value,
                      const float *d_cdiag, const float *d_x, float *d_y,
                      uint32_t reps, float *avg_ms, void *stream);

/* ppf_pcg_plan_create + ppf_pcg_solve in one call that allocates nothing
 * (no device, pinned or graph allocation; the caller passes inputs and
 * outputs). No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture of their own, tight serving loops,
 * harnesses that forbid hipMalloc). The plan's scratch lives in a fixed
 * 32 MiB device region of the library, one per GPU, held until the process
 * ends: it fits about 104 n + 256 n (the widest ELL table) + 24 nnz + 0.5 MiB
 * bytes, e.g. 32,768 rows with 262,144 stored blocks; larger systems return
 * PPF_ERR_ARG. Same kernels, tiles and iterates as a plan's solve (results
 * equal bit for bit). Every call on a GPU shares that GPU's region: calls
 * that overlap on one GPU (several host threads or streams) race. Blocks
 * until the solve ends. */
int ppf_pcg_solve_noalloc(uint32_t n, uint32_t nnz, const uint32_t *d_row_offset,
                          const uint32_t *d_col_index, const float *d_value,
                          const float *d_cdiag, const float *d_b, float *d_x, float tol,
                          uint32_t max_iter, ppf_pcg_result *out, void *stream);

/* Block-Jacobi inverse of n symmetric 3x3 blocks (solver::invert,
 * solver.cu:357-400). *d_bad (device uint32) counts blocks the reference
 * would trap on. */
int ppf_bj_invert(uint32_t n, const float *d_in, float *d_out, uint32_t *d_bad,
                  void *stream);

/* Symmetric 3x3 eigendecomposition of n blocks, eigenvalues ascending
 * (linalg::eig::symm3x3, linalg/eigsolve.hpp:195-232). */
int ppf_symm3x3(uint32_t n, const float *d_in, float *d_val, float *d_vec,
                void *stream);

/* ------------------------------------------------------------------------
 * LBVH and the contact proximity narrow phase.
 * ---------------------------------------------------------------------- */

/* A contact mesh (all device pointers): positions x (3 nv floats), edges
 * (2 ne), faces (3 nf), the reference's neighbor tables as CSR
 * (data.mesh.neighbor: vertex->edge, vertex->face, edge->face; data.hpp:85-98),
 * per-element contact gap and offset (VertexParam/EdgeParam/FaceParam ghat and
 * offset, resolved per element), vflag bit0 = pinned (fix_index != 0),
 * bit1 = collider, vbody = PD rigid body id (0 none), eflag/fflag bit0 =
 * fixed. vflag / eflag bit2 = outside the element's collision window: the
 * queries the vertex (or edge) issues find nothing, while it can still be
 * found by others (the inactive query boxes of contact.cu:1637, 1728,
 * 1868, 1940, 2563, 2591, 2688). */
typedef struct ppf_contact_mesh {
    uint32_t nv, ne, nf;
    const float *x;
    const uint32_t *edge, *face;
    const uint32_t *v2e_off, *v2e, *v2f_off, *v2f, *e2f_off, *e2f;
    const float *v_ghat, *v_off, *e_ghat, *e_off, *f_ghat, *f_off;
    const uint32_t *vflag, *vbody, *eflag, *fflag;
    /* Many independent scenes in one mesh (no counterpart in the reference,
     * which steps one scene per process): the scene id of every vertex, the
     * vertices, edges and faces of a scene contiguous and scene ids rising
     * with them; elements of different scenes never meet. nscene scenes;
     * vscene null (or nscene <= 1) = one scene. */
    const uint32_t *vscene;
    uint32_t nscene;
} ppf_contact_mesh;

/* One accepted stage-0 pair: kind (low byte) 0 point-face, 1 point-edge,
 * 2 point-point, 3 edge-edge, with bit 8 set when the pair carries no friction
 * (a neighbor element of the candidate contains the query vertex,
 * contact.cu:498-531, 685-695); a = query element (vertex; edge for kind 3),
 * b = candidate element (face, edge, vertex, edge); count = multiplicity.
 * Collision-mesh pairs (contact.cu:899-1152): kind 4 dynamic vertex a vs
 * static face b, 5 static vertex a vs dynamic face b, 6 dynamic edge a vs
 * static edge b; count 1. */
typedef struct ppf_pair {
    uint32_t kind, a, b, count;
} ppf_pair;

typedef struct ppf_narrow ppf_narrow;

int ppf_narrow_create(ppf_narrow **out);
void ppf_narrow_destroy(ppf_narrow *h);

/* Build the face, edge and vertex LBVHs of the mesh at its positions
 * (replaces lbvh::build_face_bvh / build_edge_bvh / build_vertex_bvh,
 * lbvh/lbvh.cu:863-1026). */
int ppf_narrow_build(ppf_narrow *h, const ppf_contact_mesh *m, void *stream);

/* Copy one tree (kind 0 face, 1 edge, 2 vertex): 2*(2n-1) node words and
 * (2n-1) boxes of 8 floats (min xyz, active, max xyz, pad). Either output may
 * be null; *n and *root are always written. */
int ppf_narrow_tree(ppf_narrow *h, int kind, uint32_t *d_nodes, float *d_boxes,
                    uint32_t *n, uint32_t *root, void *stream);

/* All stage-0 proximity pairs of the mesh (the detection half of
 * contact::embed_contact_force_hessian, contact/contact.cu:1398-1787, with the
 * acceptance rules of contact.cu:432-870), grouped by query in BVH leaf order.
 * Writes min(*count, cap) pairs to d_out (may be null to count only). Needs
 * ppf_narrow_build on the same positions first. */
int ppf_narrow_proximity(ppf_narrow *h, const ppf_contact_mesh *m,
                         ppf_pair *d_out, uint64_t cap, uint64_t *count,
                         void *stream);

struct ppf_static_mesh;

/* All collision-mesh proximity pairs (kinds 4, 5, 6; the detection of
 * contact::embed_constraint_force_hessian, contact/contact.cu:1789-1954, with
 * the acceptance tests of contact.cu:899-1152): per unpinned vertex of
 * positive mass against the static faces, per static vertex against the
 * faces that are not fixed and have mass, per edge that is not fixed and has
 * mass against the static edges; grouped by query. Element masses are
 * VertexProp / EdgeProp / FaceProp mass. Needs ppf_narrow_build at m->x. */
int ppf_narrow_static(ppf_narrow *h, const ppf_contact_mesh *m,
                      const struct ppf_static_mesh *sm, const float *d_vmass,
                      const float *d_emass, const float *d_fmass, ppf_pair *d_out,
                      uint64_t cap, uint64_t *count, void *stream);

/* One intersecting pair (IntersectionRecord, data.hpp:741-748): type 0
 * face-edge, 1 edge-edge, 2 collision-mesh face vs edge, 3 point-point;
 * elem0 / elem1 as the reference records them; positions packs elem0's
 * vertices then elem1's. */
typedef struct ppf_isect_record {
    uint32_t type, elem0, elem1, num_verts0, num_verts1;
    float positions[15];
} ppf_isect_record;

struct ppf_static_mesh;

/* Post-step intersection check at m->x (replaces contact::check_intersection,
 * contact/contact.cu:3056-3175): edges crossing faces or the collision mesh
 * (sm may be null), edge pairs and surface-vertex pairs (the first
 * surface_vert_count vertices) closer than their summed offsets. Element
 * masses are VertexProp / EdgeProp / FaceProp mass. *count receives the number
 * of intersecting pairs (0 = clear); the first min(*count, cap) records are
 * copied to the host array `records` (may be null). Needs ppf_narrow_build at
 * the same positions first. */
int ppf_intersection_check(ppf_narrow *h, const ppf_contact_mesh *m,
                           const struct ppf_static_mesh *sm, const float *d_vmass,
                           const float *d_emass, const float *d_fmass,
                           uint32_t surface_vert_count, ppf_isect_record *records,
                           uint32_t cap, uint32_t *count, void *stream);

/* ------------------------------------------------------------------------
 * Contact barrier + friction force and Hessian assembly.
 * ---------------------------------------------------------------------- */

/* Per-solve contact material: the step's start positions x0 (vertex.curr,
 * 3 nv), per-vertex mass (VertexProp.mass), per-element friction coefficients,
 * FrictionMode (0 min, 1 max, 2 mean; data.hpp:79), ParamSet.friction_eps, and
 * the elastic-only upper block CSR the stiffness reads its local Hessian from
 * (the tmp_fixed snapshot of main.cu:806; may be null = no elastic term). */
typedef struct ppf_contact_params {
    const float *x0;
    const float *mass;
    const float *v_fric, *e_fric, *f_fric;
    int32_t friction_mode;
    float friction_eps;
    const uint32_t *el_rp, *el_col;
    const float *el_val;
    /* the collision mesh the kind 4-6 pairs refer to (host pointer to a struct
     * of device pointers; may be null when there are no such pairs) */
    const struct ppf_static_mesh *sm;
} ppf_contact_params;

typedef struct ppf_contact ppf_contact;

int ppf_contact_create(ppf_contact **out);
void ppf_contact_destroy(ppf_contact *h);

/* Assemble the contact energy gradient (d_grad, 3 nv floats, overwritten) and
 * the upper-triangle 3x3 contact blocks of the pairs (replaces the stage-1
 * fill of contact::embed_contact_force_hessian, contact/contact.cu:262-354 and
 * 1398-1787, with barrier::compute_stiffness, barrier/barrier.cu:49-86, the
 * cubic barrier and Friction). Blocks are summed per (row, col) in pair order;
 * *nnz receives their count. m->x holds the Newton iterate (eval_x). */
int ppf_contact_assemble(ppf_contact *h, const ppf_contact_mesh *m,
                         const ppf_contact_params *p, const ppf_pair *d_pairs,
                         uint64_t npairs, float *d_grad, uint64_t *nnz,
                         void *stream);

/* The blocks of the last assemble, owned by h: row offsets (nv+1), row and
 * column per block (sorted by row, then column), 9 floats per block. */
int ppf_contact_blocks(ppf_contact *h, const uint32_t **d_row_offset,
                       const uint32_t **d_rows, const uint32_t **d_cols,
                       const float **d_vals, uint64_t *nnz);

/* ------------------------------------------------------------------------
 * CCD step filter (contact line search).
 * ---------------------------------------------------------------------- */

/* Static collision mesh (data.constraint.mesh, data.hpp:225-240): positions,
 * edges, faces and per-element contact gap / offset, all device pointers.
 * nv == 0 (or a null pointer to it) means no collision mesh. */
typedef struct ppf_static_mesh {
    uint32_t nv, ne, nf;
    const float *x;
    const uint32_t *edge, *face;
    const float *v_ghat, *v_off, *e_ghat, *e_off, *f_ghat, *f_off;
    /* per-element friction (the collision mesh's Vertex/Edge/FaceParam
     * friction); read by the contact assembly only */
    const float *v_fric, *e_fric, *f_fric;
    /* scene id of every collision-mesh vertex (as ppf_contact_mesh.vscene;
     * null = one scene) */
    const uint32_t *vscene;
} ppf_static_mesh;

/* Analytic colliders (Sphere and Floor of data.constraint, data.hpp:449-467;
 * the same layout as data.rs:689-709). A sphere pushes vertices out of it
 * (reverse: keeps them inside; bowl: a sphere whose upper half is a
 * cylinder), a floor keeps them above the plane through `ground` normal to
 * `up`; thickness > 0 lets vertices deeper than it pass through; kinematic
 * colliders move and are held by the barrier alone (no line-search clamp). */
typedef struct ppf_sphere {
    float center[3];
    float ghat, friction, radius, thickness;
    bool bowl, reverse, kinematic;
} ppf_sphere;

typedef struct ppf_floor {
    float ground[3];
    float ghat, friction, thickness;
    float up[3];
    bool kinematic;
} ppf_floor;

/* Line-search inputs besides the mesh (whose x is the step's start x0): the
 * target positions x1 (3 nv), VertexProp / EdgeProp / FaceProp mass (the
 * collision-mesh sweeps skip massless elements), ParamSet.line_search_max_t
 * and ParamSet.ccd_eps. */
typedef struct ppf_ls_params {
    const float *x1;
    const float *vmass, *emass, *fmass;
    float max_t;
    float ccd_eps;
    /* analytic colliders (device arrays; vertex_constraint_line_search,
     * contact.cu:2318-2470); the mesh's vflag bit0 marks pinned vertices */
    const ppf_sphere *spheres;
    uint32_t nsphere;
    const ppf_floor *floors;
    uint32_t nfloor;
} ppf_ls_params;

typedef struct ppf_ccd ppf_ccd;

int ppf_ccd_create(ppf_ccd **out);
void ppf_ccd_destroy(ppf_ccd *h);

/* Largest certified fraction of the step x0 -> x1 (replaces
 * contact::line_search, contact/contact.cu:2472-2740, with the ACCD of
 * contact/accd.hpp:64-696): the minimum time of impact over all point-face,
 * point-point, edge-edge and collision-mesh pairs whose swept boxes overlap,
 * divided by max_t; *toi = 1 when nothing is hit. *overlap is set when a pair
 * starts inside its offset. rebuild = 0 refits the swept trees of the last
 * call on the same mesh sizes instead of rebuilding them. */
int ppf_line_search(ppf_ccd *h, const ppf_contact_mesh *m, const ppf_static_mesh *sm,
                    const ppf_ls_params *p, int rebuild, float *toi, int *overlap,
                    void *stream);

/* ACCD of n independent pairs of one kind (accd::point_triangle_ccd,
 * point_edge_ccd, point_point_ccd, edge_edge_ccd; contact/accd.hpp:547-696):
 * kind 0 point-triangle (p, t0, t1, t2), 1 point-edge (p, e0, e1),
 * 2 point-point, 3 edge-edge (a0, a1, b0, b1). d_x0 / d_x1 hold 3 * columns
 * floats per pair, d_offset / d_ghat one float per pair. Writes the toi per
 * pair to d_toi and the overlap flag (0/1) to d_overlap (may be null). */
int ppf_ccd_batch(int kind, uint32_t n, const float *d_x0, const float *d_x1,
                  const float *d_offset, const float *d_ghat, float max_t,
                  float ccd_eps, float *d_toi, int32_t *d_overlap, void *stream);

/* ------------------------------------------------------------------------
 * Inertia, air drag, membrane and bending (elastic) assembly.
 * ---------------------------------------------------------------------- */

/* A shell (all device pointers). Positions: x = the Newton iterate (eval_x),
 * x0 = vertex.curr, target = the inertial target y. Faces and hinges as
 * data.mesh.mesh (hinge = 4 vertices in stored order; the bending term works
 * on (h2, h1, h0, h3), dihedral_angle.hpp:15-17). v2f_off / v2f: vertex->face
 * CSR in the order the vertex normal sums (data.mesh.neighbor.vertex.face).
 * Per vertex: VertexProp mass and area, vflag bit0 = pinned (fix_index != 0).
 * Per face: FaceProp mass, inv_rest2x2 (4 floats, column-major), FaceParam mu
 * and lambda, fflag bit0 = excluded (fixed, collider or rest-excluded). Per
 * hinge: HingeProp rest angle, length and area, the resolved bend
 * (hinge_bend_directional applied), hflag bit0 = excluded. */
typedef struct ppf_shell {
    uint32_t nv, nf, nh;
    const float *x, *x0, *target;
    const uint32_t *face, *hinge, *v2f_off, *v2f;
    const float *vmass, *varea;
    const uint32_t *vflag;
    const float *fmass, *inv_rest, *f_mu, *f_lambda;
    const uint32_t *fflag;
    const float *h_rest, *h_length, *h_area, *h_bend;
    const uint32_t *hflag;
    /* Strain limiting (strainlimiting.cu:46-100, 243-305): per face the
     * FaceParam strainlimit (0 = off) and min(shrink_x, shrink_y) (null =
     * 1); the barrier keeps the larger singular value of F below
     * 1 + (strainlimit + 1) / shrink - 1. Null f_strainlimit: no face is
     * limited. */
    const float *f_strainlimit;
    const float *f_shrink;
    /* Plasticity (plasticity.cu:21-78, 145-185): per face the FaceParam
     * plasticity rate and threshold, per hinge the HingeParam ones (null
     * rates: none). After every committed step the rest shape creeps toward
     * the dead zone |s - 1| <= threshold of each singular value (faces) and
     * |angle - rest| <= threshold (hinges), by 1 - exp(-rate dt): inv_rest
     * and h_rest are then rewritten in place, so they must point to writable
     * device memory. */
    const float *f_plasticity, *f_plastic_threshold;
    const float *h_plasticity, *h_plastic_threshold;
} ppf_shell;

/* dt and the ParamSet air terms (air_density, air_friction, wind). */
typedef struct ppf_elastic_params {
    float dt, air_density, air_friction;
    float wind[3];
    /* many scenes in one shell: the time step of vertex i is
     * scene_dt[vscene[i]] (both device arrays; null: dt for every vertex) */
    const float *scene_dt;
    const uint32_t *vscene;
} ppf_elastic_params;

typedef struct ppf_elastic ppf_elastic;

/* Build the fixed upper-block pattern of the shell (every face and hinge
 * vertex pair plus every diagonal; the shell stencils of builder.rs:1244-1355)
 * and the tables that sum element blocks into it. Only nv/nf/nh, face and
 * hinge are read. */
int ppf_elastic_create(const ppf_shell *topology, void *stream, ppf_elastic **out);
void ppf_elastic_destroy(ppf_elastic *h);

/* The pattern, owned by h: row offsets (nv+1) and column indices (nnz). */
int ppf_elastic_pattern(ppf_elastic *h, const uint32_t **d_row_offset,
                        const uint32_t **d_col_index, uint32_t *nnz);

/* Elastic + inertia force and Hessian at sh->x (replaces
 * energy::embed_momentum_force_hessian and the shell part of
 * energy::embed_elastic_force_hessian, energy/energy.cu:182-553, 700-794):
 * d_force (3 nv) and d_cdiag (9 nv, the per-vertex diagonal) are overwritten,
 * d_value (nnz blocks of the pattern) receives the face and hinge blocks.
 * Contributions are summed per block in a fixed order (faces, then hinges, by
 * index). sh must have the topology of create. */
int ppf_elastic_assemble(ppf_elastic *h, const ppf_shell *sh, const ppf_elastic_params *p,
                         float *d_force, float *d_cdiag, float *d_value, void *stream);

/* Strain-limit force and Hessian at sh->x of every face with
 * sh->f_strainlimit > 0 (replaces
 * strainlimiting::embed_strainlimiting_force_hessian,
 * strainlimiting/strainlimiting.cu:46-100): the cubic barrier on
 * strainlimit - (s - 1) of each singular value s > 1 of F, projected by the
 * 3x2 eigen-analysis with eiganalysis_eps (ParamSet.eiganalysis_eps),
 * scaled by w^T H_el w + fmass / d^2 (w: the face's step-start vertices
 * sh->x0 about their centroid; H_el: the face's 9x9 blocks of d_value;
 * d: the limit minus the larger stretch). d_value must hold the blocks
 * ppf_elastic_assemble wrote (the reference's tmp_fixed); the terms are added
 * to d_force and d_value per vertex and per block in face order. A shell with
 * a null f_strainlimit is left as it is. */
int ppf_elastic_strain_limit(ppf_elastic *h, const ppf_shell *sh, float eiganalysis_eps,
                             float *d_force, float *d_value, void *stream);

/* Strain-limit line search of the step d_x0 -> d_x1 (replaces
 * strainlimiting::line_search, strainlimiting/strainlimiting.cu:243-305):
 * per limited face the largest t <= line_search_max_t, by bisection, that
 * keeps the larger singular value of F(x0) + t (F(x1) - F(x0)) below
 * 1 + its limit; *toi = the minimum over the faces / line_search_max_t
 * (1 when none binds). */
int ppf_elastic_strain_limit_toi(ppf_elastic *h, const ppf_shell *sh, const float *d_x0,
                                 const float *d_x1, float line_search_max_t, float *toi,
                                 void *stream);

/* Rest-shape creep after a committed step at positions sh->x (replaces
 * plasticity::update_face_plasticity and update_hinge_plasticity,
 * plasticity/plasticity.cu:21-78, 145-185, called at main.cu:1615-1624
 * with ParamSet.dt): per face with f_plasticity > 0, each singular value s
 * of F = dX inv_rest with |s - 1| > threshold moves by
 * alpha (1 -/+ threshold - s), alpha = 1 - exp(-rate dt), and the new
 * inv_rest maps the face's current in-plane frame to U diag(s') V^T; per
 * non-excluded hinge with h_plasticity > 0, the rest angle moves by
 * alpha (delta - sign(delta) threshold) when |delta| = |angle - rest| >
 * threshold. d_inv_rest (4 nf) and d_h_rest (nh) hold the current rest
 * data and receive the new one (they may be sh->inv_rest and sh->h_rest
 * themselves); entries that do not creep are left as they are. */
int ppf_elastic_plasticity(ppf_elastic *h, const ppf_shell *sh, float dt, float *d_inv_rest,
                           float *d_h_rest, void *stream);

/* ------------------------------------------------------------------------
 * One time step: the Newton loop of main_helper::advance.
 * ---------------------------------------------------------------------- */

/* A scene (device pointers). mesh.x is vertex.curr and prev is vertex.prev:
 * both are updated in place by a successful advance. The shell carries the
 * elastic data (its x, x0 and target fields are ignored: the stepper owns the
 * iterate); shell.vmass is also the contact mass and the pinned flag is
 * mesh.vflag bit0. collision.nv == 0 means no collision mesh. fix_pos holds
 * the target position of every pinned vertex (3 nv; may be null when no
 * vertex is pinned); emass is EdgeProp mass; *_fric the per-element
 * friction. fix_delta (3 nv, may be null) is each pin's displacement over the
 * step (FixPair::step_delta: zero for a static pin): when the Newton loop
 * shortens the step to toi_advanced < 1, fix_pos is rewound in place to
 * fix_pos - (1 - toi_advanced) * fix_delta before the final iteration
 * (rewind_kinematic_fix, main.cu:513-545). */
typedef struct ppf_scene {
    ppf_contact_mesh mesh;
    ppf_shell shell;
    ppf_static_mesh collision;
    float *prev;
    const float *emass;
    const float *v_fric, *e_fric, *f_fric;
    float *fix_pos;
    const float *fix_delta;
    /* analytic colliders (host arrays, read at every advance) */
    const ppf_sphere *spheres;
    uint32_t nsphere;
    const ppf_floor *floors;
    uint32_t nfloor;
} ppf_scene;

/* The ParamSet fields the step reads (data.hpp:515-568). */
typedef struct ppf_step_params {
    float dt, prev_dt;
    float gravity[3];
    float cg_tol;
    uint32_t cg_max_iter;
    float target_toi, line_search_max_t, ccd_eps, max_dx;
    float friction_eps;
    int32_t friction_mode;
    uint32_t min_newton_steps, max_newton_steps;
    float air_density, air_friction;
    float wind[3];
    int32_t disable_contact;
    int32_t check_intersection;
    /* ParamSet.constraint_tol: the smallest gap a kinematic floor's barrier
     * sees, as a fraction of its ghat (default 0.01) */
    float constraint_tol;
    /* ParamSet.eiganalysis_eps: below this gap of the two singular values the
     * strain-limit Hessian's flip mode takes the mean curvature (default
     * 1e-2) */
    float eiganalysis_eps;
} ppf_step_params;

/* Outcome of one step (StepResult, data.hpp:570-595). status: 0 success,
 * 1 PCG failed, 2 CCD failed (toi <= FLT_EPSILON), 3 a contact pair starts
 * inside its offset, 4 intersection after the step, 5 Newton stalled
 * (max_newton_steps), 6 a pinned vertex is driven through a static analytic
 * collider (StepResult.pin_feasible = false, main.cu:1240-1271). The step is
 * committed (prev <- curr, curr <- the new positions) on 0 and on 4, as
 * main.cu:1515-1629 commits before the host sees the intersection flag. Times
 * are HIP-event milliseconds on the stream. */
typedef struct ppf_step_result {
    int32_t status;
    uint32_t newton_steps;
    uint32_t pcg_iter_total, pcg_iter_max;
    uint32_t pairs_max;
    uint32_t intersections;
    float toi_advanced;
    float dt;
    float ms_elastic, ms_contact, ms_solve, ms_line_search, ms_total;
    uint32_t lpc_fallbacks; /* solves the learned factor handed to block Jacobi */
    uint32_t device_alloc, device_free; /* allocations / frees during the step */
} ppf_step_result;

typedef struct ppf_stepper ppf_stepper;

int ppf_stepper_create(const ppf_scene *scene, void *stream, ppf_stepper **out);
void ppf_stepper_destroy(ppf_stepper *h);

/* Advance the scene by one step (replaces main_helper::advance,
 * main/main.cu:267-1649, for shells with fix pins, self contact and a static
 * collision mesh): inertial target, Newton iterations (elastic + contact
 * assembly, Dirichlet elimination of pinned vertices, block-Jacobi PCG, CCD
 * line search, toi accumulation to target_toi, then one error-reduction
 * iteration at the shortened dt), post-step intersection check, commit of
 * prev / curr. */
int ppf_stepper_advance(ppf_stepper *h, const ppf_scene *scene, const ppf_step_params *p,
                        ppf_step_result *out, void *stream);

/* One time step of every scene of a stepper created over many scenes
 * (scene->mesh.vscene / nscene, scene->collision.vscene). No counterpart in
 * the reference, which steps one scene per process (main_helper::advance per
 * scene): each scene runs its own Newton loop -- dt, toi bookkeeping, Newton
 * count, status, commit --, and every stage (trees, proximity, assembly,
 * block-Jacobi PCG with per-scene scalars, line search with per-scene minima,
 * intersection check) is one launch over all scenes; elements of different
 * scenes never meet. prev_dt: host array, one entry per scene (the step
 * before's dt; p->prev_dt is not read); out: one result per scene. */
int ppf_stepper_advance_scenes(ppf_stepper *h, const ppf_scene *scene,
                               const ppf_step_params *p, const float *prev_dt,
                               ppf_step_result *out, void *stream);

/* Precondition the stepper's Newton solves with the learned factor of h
 * (recomputed for every Newton system; ms_solve includes it), falling back
 * to block Jacobi for a solve that does not reach the tolerance. NULL
 * restores block Jacobi. The handle must outlive its use. */
int ppf_stepper_set_preconditioner(ppf_stepper *s, ppf_lpc *h);

/* Re-read the pin set after the scene's mesh.vflag bit0 or fix_pos changed
 * (update_constraint, main.cu:1988-2077). */
int ppf_stepper_update_pins(ppf_stepper *h, const ppf_scene *scene, void *stream);

/* The last Newton system the stepper solved: n rows, nnz stored upper blocks
 * of the merged matrix after Dirichlet elimination (main.cu:954-1049), the
 * per-row diagonal and the right-hand side -- the (Hessian, residual) pair a
 * linear-system dump records (dump_linsys.hpp:6-16). Writes *n and *nnz and
 * copies into whichever of d_row_offset (n + 1), d_col_index (nnz), d_value
 * (9 nnz), d_cdiag (9 n), d_rhs (3 n) is non-null. */
int ppf_stepper_system(ppf_stepper *h, uint32_t *n, uint32_t *nnz, uint32_t *d_row_offset,
                       uint32_t *d_col_index, float *d_value, float *d_cdiag, float *d_rhs,
                       void *stream);

/* The row order the stepper's solves run in (new row k = vertex
 * d_order[k]; n entries copied into d_order): the mesh pieces one after the
 * other, each in the Morton order of its initial positions (see
 * ppf_pcg_plan_set_order). */
int ppf_stepper_solve_order(ppf_stepper *h, uint32_t *d_order, void *stream);

/* Records of the stepper's last post-step intersection check (the first
 * min(cap, 256) of them; contact::check_intersection, contact.cu:3056-3175).
 * Returns the number of intersections that check found. */
uint32_t ppf_stepper_intersection_records(ppf_stepper *h, ppf_isect_record *out,
                                          uint32_t cap);

/* Upper 3x3 blocks of the last contact matrix the stepper assembled (rows =
 * vertices; self contact and collision mesh, before Dirichlet elimination).
 * Writes *nnz; fills whichever of d_row_offset (nv + 1), d_col_index (nnz),
 * d_value (9 nnz) is non-null. */
int ppf_stepper_contact_pattern(ppf_stepper *h, uint32_t *nnz, uint32_t *d_row_offset,
                                uint32_t *d_col_index, float *d_value, void *stream);

#ifdef __cplusplus
}
#endif

#endif
