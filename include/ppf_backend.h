/* ppf_backend.h -- the solver-backend C ABI the Rust host links against.
 *
 * libsimbackend_mi355x.so exports exactly the functions the host imports:
 * the 16 of crates/ppf-cts-solver/src/backend.rs:28-62, set_log_path
 * (main.rs:43-45) and ppf_fatal_code / ppf_fatal_detail
 * (status_writer.rs:39-46), with the behaviour of their implementations in
 * crates/ppf-cts-solver/src/cpp/main/main.cu:1673-2350. The structs below are
 * the host's #[repr(C)] types of data.rs, cvec.rs and cvecvec.rs, field for
 * field (names as in data.rs, a cts_ prefix on the type names); their byte
 * layout is checked against the layout the C rules give data.rs.
 *
 * Ownership follows the reference: the host allocates every array;
 * initialize() copies what the solver needs to the device and keeps the
 * DataSet by value (so fetch() writes into the host's own curr/prev buffers)
 * and the ParamSet by pointer (advance() reads it every step and writes
 * prev_dt and time back, main.cu:1572-1573). Calls are sequential from one
 * thread; one scene per process, as in the reference.
 *
 * Supported scene content: shell faces (Baraff-Witkin membrane) with hinges,
 * fix pins, one static collision mesh, self contact and collision-mesh
 * contact. initialize() returns false (fatal code 4, a detail naming the
 * feature) for content outside it: tets, rods, stitches, pulls, torques,
 * analytic spheres and floors, PDRD bodies, grains, translation locks,
 * strain limits, plasticity, non-cubic barriers.
 */
#ifndef PPF_BACKEND_H
#define PPF_BACKEND_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* nalgebra aliases of data.rs:11-18: column-major arrays, 4-byte aligned. */
typedef struct cts_Vec3f { float v[3]; } cts_Vec3f;
typedef struct cts_Vec2u { uint32_t v[2]; } cts_Vec2u;
typedef struct cts_Vec3u { uint32_t v[3]; } cts_Vec3u;
typedef struct cts_Vec4u { uint32_t v[4]; } cts_Vec4u;
typedef struct cts_Vec6u { uint32_t v[6]; } cts_Vec6u;
typedef struct cts_Vec6f { float v[6]; } cts_Vec6f;
typedef struct cts_Mat2x2f { float m[4]; } cts_Mat2x2f;
typedef struct cts_Mat3x3f { float m[9]; } cts_Mat3x3f;

/* CVec<T> (cvec.rs:10-14) and CVecVec<T> (cvecvec.rs:10-17). */
#define CTS_VEC(T) struct { T *data; uint32_t size; uint32_t allocated; }
#define CTS_VECVEC(T)                                                          \
    struct {                                                                   \
        T *data;                                                               \
        uint32_t *offset;                                                      \
        uint32_t size, nnz, nnz_allocated, offset_allocated;                   \
    }
typedef CTS_VEC(void) cts_CVec;
typedef CTS_VECVEC(void) cts_CVecVec;

/* Enums are 4-byte ints (data.rs:530-626). */
enum { CTS_MODEL_ARAP = 0, CTS_MODEL_STVK, CTS_MODEL_BARAFF_WITKIN, CTS_MODEL_SNHK,
       CTS_MODEL_PDRD };
enum { CTS_BARRIER_CUBIC = 0, CTS_BARRIER_QUAD, CTS_BARRIER_LOG };
enum { CTS_FRICTION_MIN = 0, CTS_FRICTION_MAX, CTS_FRICTION_MEAN };
enum { CTS_PRECOND_BLOCK_JACOBI = 0, CTS_PRECOND_SCHWARZ };

typedef struct cts_FfiVertexNeighbor {
    CTS_VECVEC(uint32_t) face, hinge, edge, rod;
} cts_FfiVertexNeighbor;
typedef struct cts_FfiHingeNeighbor { CTS_VECVEC(uint32_t) face; } cts_FfiHingeNeighbor;
typedef struct cts_FfiEdgeNeighbor { CTS_VECVEC(uint32_t) face; } cts_FfiEdgeNeighbor;

typedef struct cts_FfiMesh {
    CTS_VEC(cts_Vec3u) face;
    CTS_VEC(cts_Vec4u) hinge;
    CTS_VEC(cts_Vec2u) edge;
    CTS_VEC(cts_Vec4u) tet;
} cts_FfiMesh;

typedef struct cts_FfiNeighbor {
    cts_FfiVertexNeighbor vertex;
    cts_FfiHingeNeighbor hinge;
    cts_FfiEdgeNeighbor edge;
} cts_FfiNeighbor;

typedef struct cts_Type {
    CTS_VEC(uint8_t) face, vertex, hinge;
} cts_Type;

typedef struct cts_FfiMeshInfo {
    cts_FfiMesh mesh;
    cts_FfiNeighbor neighbor;
    cts_Type ttype;
} cts_FfiMeshInfo;

typedef struct cts_VertexParam {
    float ghat, offset, friction;
} cts_VertexParam;

typedef struct cts_EdgeParam {
    float stiffness, bend, ghat, offset, friction, strainlimit, plasticity,
        plasticity_threshold;
    bool bend_rest_from_geometry;
    float deform_damping, bend_damping;
} cts_EdgeParam;

typedef struct cts_FaceParam {
    uint32_t model;
    float mu, lambda, friction, ghat, offset, bend, strainlimit, shrink_x, shrink_y,
        pressure, plasticity, plasticity_threshold, bend_plasticity,
        bend_plasticity_threshold;
    bool bend_rest_from_geometry;
    float deform_damping, bend_damping, bend_warp, bend_weft;
} cts_FaceParam;

typedef struct cts_HingeParam {
    float bend, ghat, offset, plasticity, plasticity_threshold, bend_damping, bend_warp,
        bend_weft;
} cts_HingeParam;

typedef struct cts_TetParam {
    uint32_t model;
    float mu, lambda, shrink, plasticity, plasticity_threshold, deform_damping;
} cts_TetParam;

typedef struct cts_VertexProp {
    float area, volume, mass, rest_bend_angle;
    uint32_t fix_index, pull_index, param_index, pdrd_body_index;
    bool collider;
} cts_VertexProp;

typedef struct cts_EdgeProp {
    float length, initial_length, mass;
    bool fixed;
    uint32_t param_index;
} cts_EdgeProp;

typedef struct cts_FaceProp {
    float area, mass;
    bool fixed, rest_excluded, collider;
    uint32_t param_index;
} cts_FaceProp;

typedef struct cts_HingeProp {
    float length, area, rest_angle, uv_edge_sin2;
    bool fixed, collider;
    uint32_t param_index;
} cts_HingeProp;

typedef struct cts_TetProp {
    float mass, volume;
    bool fixed, rest_excluded;
    uint32_t param_index;
} cts_TetProp;

typedef struct cts_PdrdBodyProp {
    cts_Vec3f rest_centroid;
    cts_Mat3x3f rest_gram_inv;
    float volume;
    uint32_t vertex_start, vertex_count;
    float mass_per_vertex;
    uint32_t joint_mode;
    cts_Vec3f joint_axis, joint_pin;
} cts_PdrdBodyProp;

typedef struct cts_PropSet {
    CTS_VEC(cts_VertexProp) vertex;
    CTS_VEC(cts_EdgeProp) edge;
    CTS_VEC(cts_FaceProp) face;
    CTS_VEC(cts_HingeProp) hinge;
    CTS_VEC(cts_TetProp) tet;
    CTS_VEC(cts_PdrdBodyProp) pdrd_body;
} cts_PropSet;

typedef struct cts_ParamArrays {
    CTS_VEC(cts_VertexParam) vertex;
    CTS_VEC(cts_EdgeParam) edge;
    CTS_VEC(cts_FaceParam) face;
    CTS_VEC(cts_HingeParam) hinge;
    CTS_VEC(cts_TetParam) tet;
} cts_ParamArrays;

typedef struct cts_FixPair {
    cts_Vec3f position, step_delta;
    float ghat;
    uint32_t index;
    bool kinematic;
} cts_FixPair;

typedef struct cts_PullPair {
    cts_Vec3f position;
    float weight;
    uint32_t index;
} cts_PullPair;

typedef struct cts_TorqueGroup {
    uint32_t axis_component, vertex_start, vertex_count, hint_vertex;
} cts_TorqueGroup;

typedef struct cts_TorqueVertex {
    float magnitude;
    uint32_t index, group_id;
} cts_TorqueVertex;

typedef struct cts_Stitch {
    cts_Vec6u index;
    cts_Vec6f weight;
    float stiffness;
} cts_Stitch;

typedef struct cts_Sphere {
    cts_Vec3f center;
    float ghat, friction, radius, thickness;
    bool bowl, reverse, kinematic;
} cts_Sphere;

typedef struct cts_Floor {
    cts_Vec3f ground;
    float ghat, friction, thickness;
    cts_Vec3f up;
    bool kinematic;
} cts_Floor;

typedef struct cts_CollisionMeshPropSet {
    CTS_VEC(cts_VertexProp) vertex;
    CTS_VEC(cts_FaceProp) face;
    CTS_VEC(cts_EdgeProp) edge;
} cts_CollisionMeshPropSet;

typedef struct cts_CollisionMeshParamArrays {
    CTS_VEC(cts_VertexParam) vertex;
    CTS_VEC(cts_FaceParam) face;
    CTS_VEC(cts_EdgeParam) edge;
} cts_CollisionMeshParamArrays;

typedef struct cts_CollisionMesh {
    CTS_VEC(cts_Vec3f) vertex;
    CTS_VEC(cts_Vec3u) face;
    CTS_VEC(cts_Vec2u) edge;
    cts_CollisionMeshPropSet prop;
    cts_CollisionMeshParamArrays param_arrays;
    cts_FfiNeighbor neighbor;
} cts_CollisionMesh;

typedef struct cts_Constraint {
    CTS_VEC(cts_FixPair) fix;
    CTS_VEC(cts_PullPair) pull;
    CTS_VEC(cts_TorqueGroup) torque_groups;
    CTS_VEC(cts_TorqueVertex) torque_vertices;
    CTS_VEC(cts_Sphere) sphere;
    CTS_VEC(cts_Floor) floor;
    CTS_VEC(cts_Stitch) stitch;
    cts_CollisionMesh mesh;
} cts_Constraint;

typedef struct cts_RestShapeUpdate {
    CTS_VEC(cts_Mat2x2f) inv_rest2x2;
    CTS_VEC(cts_Mat3x3f) inv_rest3x3;
    CTS_VEC(uint8_t) exclude_face;
    CTS_VEC(uint8_t) exclude_tet;
} cts_RestShapeUpdate;

typedef struct cts_ParamSet {
    double time;
    float air_friction, air_density, constraint_tol, prev_dt, dt, playback;
    uint32_t min_newton_steps;
    float target_toi, stitch_length_factor;
    uint32_t cg_max_iter;
    float cg_tol, line_search_max_t, ccd_eps, max_dx, eiganalysis_eps, friction_eps,
        isotropic_air_friction;
    cts_Vec3f gravity, wind;
    uint32_t barrier, friction_mode, csrmat_max_nnz;
    float fix_xz;
    bool disable_contact, inactive_momentum;
    uint32_t precond, schwarz_levels, max_newton_steps;
    bool disable_pin_dof_removal;
} cts_ParamSet;

typedef struct cts_StepResult {
    double time;
    bool ccd_success, pcg_success, intersection_free, newton_progress, pin_feasible,
        contact_separated;
} cts_StepResult;

typedef struct cts_IntersectionRecord {
    uint32_t itype; /* 0 face-edge, 1 edge-edge, 2 collision mesh, 3 point-point */
    uint32_t elem0, elem1, num_verts0, num_verts1;
    float positions[15];
} cts_IntersectionRecord;

typedef struct cts_VertexSet {
    CTS_VEC(cts_Vec3f) prev;
    CTS_VEC(cts_Vec3f) curr;
} cts_VertexSet;

typedef struct cts_TranslationLock {
    cts_Vec3f axis;
    float total_mass;
    uint32_t pdrd_body_index, dmap_index;
    cts_Vec3f rotation_axis;
    uint32_t rotation_mode;
    cts_Vec3f anchor;
} cts_TranslationLock;

typedef struct cts_DataSet {
    cts_VertexSet vertex;
    cts_FfiMeshInfo mesh;
    cts_PropSet prop;
    cts_ParamArrays param_arrays;
    CTS_VEC(cts_Mat2x2f) inv_rest2x2;
    CTS_VEC(cts_Mat3x3f) inv_rest3x3;
    cts_Constraint constraint;
    CTS_VECVEC(uint32_t) fixed_index_table;
    CTS_VECVEC(cts_Vec2u) transpose_table;
    uint32_t rod_count, shell_face_count, surface_vert_count;
    CTS_VEC(uint32_t) pdrd_vert_list;
    CTS_VEC(cts_Vec3f) pdrd_rest_centered;
    CTS_VEC(cts_Vec3f) grain_omega;
    CTS_VEC(float) grain_inv_inertia;
    CTS_VEC(cts_Vec3f) grain_torque;
    CTS_VEC(float) grain_ang_stiff;
    CTS_VEC(cts_Vec3f) grain_contact_normal;
    CTS_VEC(float) grain_inv_inertia_center;
    CTS_VEC(cts_Vec3f) grain_omega_prev;
    CTS_VEC(cts_Mat3x3f) grain_a;
    CTS_VEC(cts_Mat3x3f) grain_b;
    CTS_VEC(cts_Vec3f) grain_grot;
    CTS_VEC(uint32_t) tet_hess_slots, face_hess_slots, edge_hess_slots, hinge_hess_slots,
        rod_bend_hess_slots, stitch_hess_slots;
    CTS_VEC(cts_TranslationLock) translation_lock;
    CTS_VEC(uint32_t) translation_lock_index;
    CTS_VEC(cts_Vec3f) translation_lock_initial;
    CTS_VEC(uint32_t) statistics_object_index, statistics_static_object_index,
        statistics_contact_count;
} cts_DataSet;

/* Fatal codes (fatal.hpp:49-54). */
enum { CTS_FATAL_NONE = 0, CTS_FATAL_INIT_INTERSECTION = 1, CTS_FATAL_OOM = 2,
       CTS_FATAL_DRIVER = 3, CTS_FATAL_SOLVER = 4 };

/* ---- the 19 entry points ------------------------------------------------- */

/* Copy the scene to the device, build the contact trees and check that the
 * scene does not intersect itself at prev or curr (main.cu:126-265,
 * 1877-1920). false: g_ppf_fatal_code / detail say why (1 intersecting scene,
 * 3 no device or HIP error, 4 unsupported content). */
bool initialize(const cts_DataSet *data, cts_ParamSet *param);

/* One time step of dt = param->dt * param->playback (main.cu:267-1649):
 * Newton loop, CCD line search, intersection check; on success prev <- curr,
 * curr <- the new positions, param->prev_dt = dt and param->time advance. */
void advance(cts_StepResult *result);

/* Device curr / prev into the host DataSet's own buffers (main.cu:1926-1937);
 * statistics_contact_count (when present) is zero-filled. */
void fetch(void);
/* Inverse rest matrices and rest angles back into the host arrays
 * (main.cu:1939-1968); without plasticity they are the values given. */
void fetch_inv_rest(void);
void fetch_rest_angles(void);

/* Pattern of the last assembled contact matrix (upper 3x3 blocks, rows =
 * vertices; main.cu:1970-1986): counts, then index / offset and optionally
 * the blocks (value may be null). update_dyn accepts a carried-over pattern;
 * the pattern is rebuilt every Newton iteration here, so it is validated and
 * otherwise has no effect on the result. */
void fetch_dyn_counts(uint32_t *n_value, uint32_t *n_offset);
void fetch_dyn(uint32_t *index, float *value, uint32_t *offset);
void update_dyn(const uint32_t *index, const uint32_t *offset);

/* Replace the pin set (positions, step deltas) and re-derive the fixed flags
 * of elements whose vertices are all pinned (main.cu:1988-2077). */
void update_constraint(const cts_Constraint *constraint);
/* Streamed rest shape: new inv_rest2x2 and per-face exclusion (2079-2117). */
void update_rest_shape(const cts_RestShapeUpdate *update);

/* prev = curr - v dt for the listed vertices (main.cu:2119-2153). */
void override_velocity(const uint32_t *indices, uint32_t count, float vx, float vy, float vz,
                       float dt);
/* out[3k..3k+2] = curr[indices[k]] (main.cu:2155-2191). */
void gather_current_positions(const uint32_t *indices, uint32_t count, float *out);
/* prev -= (w x (curr - c)) dt for the listed vertices (main.cu:2193-2228). */
void override_angular_velocity(const uint32_t *indices, uint32_t count, float wx, float wy,
                               float wz, float cx, float cy, float cz, float dt);

/* Collision windows (main.cu:2230-2320): vertex i takes part in contact at
 * time t when its group vert_dmap[i] has no window or t lies in one of its
 * [start, end) windows (windows: n_groups x 8 pairs of floats); faces and
 * edges are active when any of their vertices is. */
void init_collision_windows(const uint32_t *vert_dmap, uint32_t vert_count,
                            const float *windows, const uint32_t *window_counts,
                            uint32_t n_groups);
void refresh_collision_active(float time);

/* Records of the last intersection check (at most max_count, at most 256). */
uint32_t fetch_intersection_records(cts_IntersectionRecord *out, uint32_t max_count);

/* Directory for the per-step log channels (main.cu:1681-1683). */
void set_log_path(const char *data_dir);

unsigned char ppf_fatal_code(void);
const char *ppf_fatal_detail(void);

extern unsigned char g_ppf_fatal_code;
extern char g_ppf_fatal_detail[512];

#ifdef __cplusplus
}
#endif

#endif /* PPF_BACKEND_H */
