// backend.hip -- the 19 solver-backend entry points the Rust host links
// (crates/ppf-cts-solver/src/backend.rs:28-62, main.rs:43-45,
// status_writer.rs:39-46), over the stage library of include/ppf_mi355x.h.
//
// initialize() turns the host's DataSet into the per-element device arrays
// the stepper reads (resolved contact gaps, offsets, friction and material
// per element, flag words, pin targets), builds the contact trees and checks
// the start state for intersections; advance() runs one step of the stepper
// and maps its outcome onto StepResult. The reference keeps the host DataSet
// by value and writes back into the host's arrays (fetch, update_constraint);
// so does this file.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ppf_backend.h"
#include "../../include/ppf_mi355x.h"

extern "C" {
unsigned char g_ppf_fatal_code = 0;
char g_ppf_fatal_detail[512] = {0};
}

namespace {

constexpr uint32_t kMaxWindows = 8;      // MAX_COLLISION_WINDOWS (main.cu:47)
constexpr uint32_t kMaxRecords = 256;    // MAX_INTERSECTION_RECORDS (data.hpp:737)
constexpr unsigned kThreads = 256;

void set_detail(unsigned char code, const char *fmt, ...) {
    g_ppf_fatal_code = code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_ppf_fatal_detail, sizeof(g_ppf_fatal_detail), fmt, ap);
    va_end(ap);
    fprintf(stderr, "ppf backend: %s\n", g_ppf_fatal_detail);
    fflush(stderr);
}

// An unrecoverable failure inside a step: report and end the process, as
// ppf_fatal does (fatal.hpp:60-80); the host's exit hook reads the code.
[[noreturn]] void fatal(unsigned char code, const char *fmt, ...) {
    g_ppf_fatal_code = code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_ppf_fatal_detail, sizeof(g_ppf_fatal_detail), fmt, ap);
    va_end(ap);
    fprintf(stderr, "ppf backend: fatal: %s\n", g_ppf_fatal_detail);
    fflush(stderr);
    exit(1);
}

void hip_or_die(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        fatal(CTS_FATAL_DRIVER, "%s: %s", what, hipGetErrorString(e));
    }
}

void ppf_or_die(int rc, const char *what) {
    if (rc != 0) {
        fatal(CTS_FATAL_DRIVER, "%s: %s", what, ppf_last_error());
    }
}

__global__ void k_override_velocity(const uint32_t *idx, uint32_t n, const float *curr,
                                    float *prev, float dx, float dy, float dz) {
    const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= n) {
        return;
    }
    const size_t v = 3 * (size_t)idx[k];
    prev[v] = curr[v] - dx;
    prev[v + 1] = curr[v + 1] - dy;
    prev[v + 2] = curr[v + 2] - dz;
}

__global__ void k_gather(const uint32_t *idx, uint32_t n, const float *curr, float *out) {
    const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= n) {
        return;
    }
    const size_t v = 3 * (size_t)idx[k];
    out[3 * (size_t)k] = curr[v];
    out[3 * (size_t)k + 1] = curr[v + 1];
    out[3 * (size_t)k + 2] = curr[v + 2];
}

__global__ void k_override_angular(const uint32_t *idx, uint32_t n, const float *curr,
                                   float *prev, float wx, float wy, float wz, float cx,
                                   float cy, float cz, float dt) {
    const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= n) {
        return;
    }
    const size_t v = 3 * (size_t)idx[k];
    const float rx = curr[v] - cx, ry = curr[v + 1] - cy, rz = curr[v + 2] - cz;
    const float vx = wy * rz - wz * ry, vy = wz * rx - wx * rz, vz = wx * ry - wy * rx;
    prev[v] = prev[v] - vx * dt;
    prev[v + 1] = prev[v + 1] - vy * dt;
    prev[v + 2] = prev[v + 2] - vz * dt;
}

struct Backend {
    bool ready = false;
    cts_DataSet host{};
    cts_ParamSet *param = nullptr;
    hipStream_t s = nullptr;
    std::vector<void *> allocs;
    uint32_t nv = 0, ne = 0, nf = 0, nh = 0;
    // device arrays the stepper reads through its scene
    float *curr = nullptr, *prev = nullptr, *fix_pos = nullptr, *fix_delta = nullptr;
    float *inv_rest = nullptr;
    float *h_rest = nullptr; // HingeProp.rest_angle, rewritten by plasticity
    uint32_t *vflag = nullptr, *eflag = nullptr, *fflag = nullptr;
    uint32_t *sh_vflag = nullptr, *sh_fflag = nullptr, *hflag = nullptr;
    ppf_scene sc{};
    ppf_stepper *stepper = nullptr;
    // analytic colliders of the current constraint (copied: the host's
    // arrays need not outlive the call that handed them over)
    std::vector<ppf_sphere> spheres;
    std::vector<ppf_floor> floors;
    // collision windows
    bool cw = false;
    std::vector<uint32_t> cw_dmap, cw_counts;
    std::vector<float> cw_windows;
    std::vector<uint8_t> v_active, e_active;
    // last intersection check
    std::vector<cts_IntersectionRecord> records;
    uint32_t record_count = 0;
    std::string log_dir;
    // Scratch of the host calls between steps (index lists, gathered
    // positions, the contact pattern): grown to a high-water mark and kept,
    // so a host that calls them every step allocates nothing once warmed up
    // (the reference's pooled buffers, buffer/buffer.hpp).
    struct Scratch {
        void *p = nullptr;
        size_t bytes = 0;
    };
    Scratch scr_idx, scr_out, scr_dyn;
    void *grow(Scratch &sc, size_t bytes) {
        if (bytes > sc.bytes) {
            ppf_device_free(sc.p);
            sc = Scratch{};
            const size_t cap = bytes + bytes / 4 + 256;
            ppf_or_die(ppf_device_alloc(&sc.p, cap), "ppf_device_alloc");
            sc.bytes = cap;
        }
        return sc.p;
    }
    // Contact pattern carried in by update_dyn (a resumed run), reported by
    // fetch_dyn until the next step replaces it (csrmat.cu:136-153).
    bool carried = false;
    std::vector<uint32_t> carried_rp, carried_ci;

    template <class T> T *put(const T *h, size_t n) {
        void *p = nullptr;
        const size_t bytes = (n ? n : 1) * sizeof(T);
        ppf_or_die(ppf_device_alloc(&p, bytes), "ppf_device_alloc");
        allocs.push_back(p);
        if (n) {
            hip_or_die(hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
        }
        return static_cast<T *>(p);
    }
    template <class T> T *put(const std::vector<T> &h) { return put(h.data(), h.size()); }

    void release() {
        if (stepper) {
            ppf_stepper_destroy(stepper);
            stepper = nullptr;
        }
        for (void *p : allocs) {
            ppf_device_free(p);
        }
        allocs.clear();
        for (Scratch *sc : {&scr_idx, &scr_out, &scr_dyn}) {
            ppf_device_free(sc->p);
            *sc = Scratch{};
        }
        carried = false;
        carried_rp.clear();
        carried_ci.clear();
        if (s) {
            (void)hipStreamDestroy(s);
            s = nullptr;
        }
        ready = false;
        cw = false;
        records.clear();
        record_count = 0;
    }

    // Flag words from the host props, the pin set and the collision windows.
    void upload_flags() {
        const cts_VertexProp *vp = host.prop.vertex.data;
        std::vector<uint32_t> vf(nv), svf(nv), ef(ne), ff(nf), sff(nf), hf(nh);
        for (uint32_t i = 0; i < nv; ++i) {
            const bool pinned = vp[i].fix_index != 0;
            const bool active = !cw || i >= v_active.size() || v_active[i];
            vf[i] = (pinned ? 1u : 0u) | (vp[i].collider ? 2u : 0u) | (active ? 0u : 4u);
            svf[i] = pinned ? 1u : 0u;
        }
        for (uint32_t e = 0; e < ne; ++e) {
            const bool active = !cw || e_active[e];
            ef[e] = (host.prop.edge.data[e].fixed ? 1u : 0u) | (active ? 0u : 4u);
        }
        for (uint32_t f = 0; f < nf; ++f) {
            const cts_FaceProp &fp = host.prop.face.data[f];
            ff[f] = fp.fixed ? 1u : 0u;
            sff[f] = (fp.fixed || fp.rest_excluded || fp.collider) ? 1u : 0u;
        }
        for (uint32_t h = 0; h < nh; ++h) {
            const cts_HingeProp &hp = host.prop.hinge.data[h];
            hf[h] = (hp.fixed || hp.collider) ? 1u : 0u;
        }
        auto up = [&](uint32_t *d, const std::vector<uint32_t> &v) {
            if (!v.empty()) {
                hip_or_die(hipMemcpyAsync(d, v.data(), v.size() * sizeof(uint32_t),
                                          hipMemcpyHostToDevice, s),
                           "hipMemcpyAsync");
            }
        };
        up(vflag, vf);
        up(sh_vflag, svf);
        up(eflag, ef);
        up(fflag, ff);
        up(sh_fflag, sff);
        up(hflag, hf);
        hip_or_die(hipStreamSynchronize(s), "hipStreamSynchronize");
    }

    // Pin targets and step deltas of the current constraint.
    void upload_pins() {
        std::vector<float> pos(3 * (size_t)nv, 0.0f), del(3 * (size_t)nv, 0.0f);
        const cts_Constraint &c = host.constraint;
        for (uint32_t k = 0; k < c.fix.size; ++k) {
            const cts_FixPair &fx = c.fix.data[k];
            for (int d = 0; d < 3; ++d) {
                pos[3 * (size_t)fx.index + d] = fx.position.v[d];
                del[3 * (size_t)fx.index + d] = fx.kinematic ? fx.step_delta.v[d] : 0.0f;
            }
        }
        hip_or_die(hipMemcpyAsync(fix_pos, pos.data(), pos.size() * sizeof(float),
                                  hipMemcpyHostToDevice, s),
                   "hipMemcpyAsync");
        hip_or_die(hipMemcpyAsync(fix_delta, del.data(), del.size() * sizeof(float),
                                  hipMemcpyHostToDevice, s),
                   "hipMemcpyAsync");
        hip_or_die(hipStreamSynchronize(s), "hipStreamSynchronize");
    }

    // The constraint's analytic colliders (cts_Sphere / cts_Floor have the
    // layout of ppf_sphere / ppf_floor: data.rs:689-709).
    void copy_colliders() {
        const cts_Constraint &c = host.constraint;
        spheres.resize(c.sphere.size);
        floors.resize(c.floor.size);
        if (c.sphere.size) {
            std::memcpy(spheres.data(), c.sphere.data, c.sphere.size * sizeof(ppf_sphere));
        }
        if (c.floor.size) {
            std::memcpy(floors.data(), c.floor.data, c.floor.size * sizeof(ppf_floor));
        }
        sc.spheres = spheres.empty() ? nullptr : spheres.data();
        sc.nsphere = (uint32_t)spheres.size();
        sc.floors = floors.empty() ? nullptr : floors.data();
        sc.nfloor = (uint32_t)floors.size();
    }

    // fix_index / pull_index and the all-pinned element flags from the
    // constraint, written into the host's prop arrays (main.cu:2018-2076).
    void derive_fixed() {
        cts_VertexProp *vp = host.prop.vertex.data;
        for (uint32_t i = 0; i < nv; ++i) {
            vp[i].fix_index = 0;
            vp[i].pull_index = 0;
        }
        for (uint32_t k = 0; k < host.constraint.fix.size; ++k) {
            vp[host.constraint.fix.data[k].index].fix_index = k + 1;
        }
        auto pinned = [&](uint32_t v) { return vp[v].fix_index > 0; };
        for (uint32_t f = 0; f < nf; ++f) {
            const uint32_t *fv = host.mesh.mesh.face.data[f].v;
            host.prop.face.data[f].fixed = pinned(fv[0]) && pinned(fv[1]) && pinned(fv[2]);
        }
        for (uint32_t e = 0; e < ne; ++e) {
            const uint32_t *ev = host.mesh.mesh.edge.data[e].v;
            host.prop.edge.data[e].fixed = pinned(ev[0]) && pinned(ev[1]);
        }
        for (uint32_t h = 0; h < nh; ++h) {
            const uint32_t *hv = host.mesh.mesh.hinge.data[h].v;
            host.prop.hinge.data[h].fixed =
                pinned(hv[0]) && pinned(hv[1]) && pinned(hv[2]) && pinned(hv[3]);
        }
    }

    void refresh_active(float time) {
        v_active.assign(nv, 1);
        const uint32_t n = std::min<uint32_t>(nv, (uint32_t)cw_dmap.size());
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t g = cw_dmap[i];
            const uint32_t cnt = g < cw_counts.size() ? cw_counts[g] : 0;
            bool active = cnt == 0;
            for (uint32_t w = 0; w < cnt && w < kMaxWindows; ++w) {
                const float ts = cw_windows[(size_t)g * kMaxWindows * 2 + 2 * w];
                const float te = cw_windows[(size_t)g * kMaxWindows * 2 + 2 * w + 1];
                if (time >= ts && time < te) {
                    active = true;
                    break;
                }
            }
            v_active[i] = active ? 1 : 0;
        }
        e_active.assign(ne, 1);
        for (uint32_t e = 0; e < ne; ++e) {
            const uint32_t *ev = host.mesh.mesh.edge.data[e].v;
            e_active[e] = (v_active[ev[0]] || v_active[ev[1]]) ? 1 : 0;
        }
    }

    void log_step(const ppf_step_result &r) {
        if (log_dir.empty()) {
            return;
        }
        const double t = param->time;
        auto put_line = [&](const char *name, double v, bool integer) {
            const std::string path = log_dir + "/advance" + (name[0] ? "." : "") + name + ".out";
            FILE *fp = fopen(path.c_str(), "a");
            if (!fp) {
                return;
            }
            if (integer) {
                fprintf(fp, "%f %lu\n", t, (unsigned long)v);
            } else {
                fprintf(fp, "%f %e\n", t, v);
            }
            fclose(fp);
        };
        put_line("", r.ms_total, true);
        put_line("newton_steps", r.newton_steps, true);
        put_line("iter", r.pcg_iter_total, true);
        put_line("num_contact", r.pairs_max, true);
        put_line("toi_advanced", r.toi_advanced, false);
        put_line("final_dt", r.dt, false);
        put_line("matrix_assembly", r.ms_elastic + r.ms_contact, true);
        put_line("linsolve", r.ms_solve, true);
        put_line("line_search", r.ms_line_search, true);
        put_line("device-alloc", r.device_alloc, true);
        put_line("device-free", r.device_free, true);
    }
};

Backend g;

static_assert(sizeof(cts_Sphere) == sizeof(ppf_sphere) && sizeof(cts_Floor) == sizeof(ppf_floor),
              "analytic collider layouts of the two C ABIs differ");

// Content this backend does not integrate: initialize() refuses it by name.
const char *unsupported(const cts_DataSet &d, const cts_ParamSet &p) {
    if (d.mesh.mesh.tet.size) return "tetrahedral solids";
    if (d.rod_count) return "rods";
    if (d.shell_face_count != d.mesh.mesh.face.size) return "faces that are not shell faces";
    if (d.surface_vert_count != d.vertex.curr.size) return "interior (non-surface) vertices";
    if (d.prop.pdrd_body.size || d.pdrd_vert_list.size) return "PDRD rigid bodies";
    if (d.translation_lock.size) return "translation locks";
    for (uint32_t i = 0; i < d.grain_inv_inertia.size; ++i)
        if (d.grain_inv_inertia.data[i] > 0.0f) return "grains";
    const cts_Constraint &c = d.constraint;
    if (c.pull.size) return "pull constraints";
    if (c.torque_groups.size || c.torque_vertices.size) return "torques";
    if (c.stitch.size) return "stitches";
    if (p.barrier != CTS_BARRIER_CUBIC) return "non-cubic barriers";
    if (p.precond != CTS_PRECOND_BLOCK_JACOBI) return "the Schwarz preconditioner";
    if (p.fix_xz != 0.0f) return "fix_xz";
    if (p.inactive_momentum) return "inactive momentum";
    if (p.disable_pin_dof_removal) return "penalty pins (disable_pin_dof_removal)";
    if (p.isotropic_air_friction != 0.0f) return "isotropic air friction";
    for (uint32_t i = 0; i < d.vertex.curr.size; ++i)
        if (d.prop.vertex.data[i].pdrd_body_index) return "PDRD rigid bodies";
    for (uint32_t f = 0; f < d.prop.face.size; ++f) {
        const uint32_t k = d.prop.face.data[f].param_index;
        if (k >= d.param_arrays.face.size) return "a face parameter index out of range";
        const cts_FaceParam &fp = d.param_arrays.face.data[k];
        if (fp.model != CTS_MODEL_BARAFF_WITKIN) return "shell models other than Baraff-Witkin";
        if (fp.pressure != 0.0f) return "pressure";
        if (fp.shrink_x != 1.0f || fp.shrink_y != 1.0f) return "rest-shape shrink";
        if (fp.deform_damping != 0.0f || fp.bend_damping != 0.0f) return "Rayleigh damping";
    }
    for (uint32_t h = 0; h < d.prop.hinge.size; ++h) {
        const uint32_t k = d.prop.hinge.data[h].param_index;
        if (k >= d.param_arrays.hinge.size) return "a hinge parameter index out of range";
        const cts_HingeParam &hp = d.param_arrays.hinge.data[k];
        if (hp.bend_damping != 0.0f) return "Rayleigh damping";
    }
    for (uint32_t i = 0; i < d.prop.vertex.size; ++i)
        if (d.prop.vertex.data[i].param_index >= d.param_arrays.vertex.size)
            return "a vertex parameter index out of range";
    for (uint32_t e = 0; e < d.prop.edge.size; ++e)
        if (d.prop.edge.data[e].param_index >= d.param_arrays.edge.size)
            return "an edge parameter index out of range";
    return nullptr;
}

const char *bad_sizes(const cts_DataSet &d) {
    const uint32_t nv = d.vertex.curr.size, ne = d.mesh.mesh.edge.size,
                   nf = d.mesh.mesh.face.size, nh = d.mesh.mesh.hinge.size;
    if (d.vertex.prev.size != nv) return "vertex.prev and vertex.curr differ in size";
    if (d.prop.vertex.size != nv || d.prop.edge.size != ne || d.prop.face.size != nf ||
        d.prop.hinge.size != nh)
        return "prop arrays do not match the mesh";
    if (d.inv_rest2x2.size != nf) return "inv_rest2x2 does not have one entry per face";
    const auto &nb = d.mesh.neighbor;
    if (nb.vertex.face.size != nv || nb.vertex.edge.size != nv || nb.edge.face.size != ne)
        return "neighbor tables do not match the mesh";
    for (uint32_t k = 0; k < d.constraint.fix.size; ++k)
        if (d.constraint.fix.data[k].index >= nv) return "a fix pin index out of range";
    const cts_CollisionMesh &cm = d.constraint.mesh;
    if (cm.prop.vertex.size != cm.vertex.size || cm.prop.face.size != cm.face.size ||
        cm.prop.edge.size != cm.edge.size)
        return "collision-mesh props do not match its mesh";
    return nullptr;
}

// The collision mesh as resolved per-element arrays.
void upload_collision(Backend &b) {
    const cts_CollisionMesh &cm = b.host.constraint.mesh;
    ppf_static_mesh &sm = b.sc.collision;
    sm = ppf_static_mesh{};
    sm.nv = cm.vertex.size;
    sm.ne = cm.edge.size;
    sm.nf = cm.face.size;
    if (!sm.nv) {
        return;
    }
    std::vector<float> vg(sm.nv), vo(sm.nv), vfr(sm.nv), eg(sm.ne), eo(sm.ne), efr(sm.ne),
        fg(sm.nf), fo(sm.nf), ffr(sm.nf);
    auto vparam = [&](uint32_t k) -> const cts_VertexParam & {
        if (k >= cm.param_arrays.vertex.size)
            fatal(CTS_FATAL_SOLVER, "collision-mesh vertex parameter index out of range");
        return cm.param_arrays.vertex.data[k];
    };
    auto eparam = [&](uint32_t k) -> const cts_EdgeParam & {
        if (k >= cm.param_arrays.edge.size)
            fatal(CTS_FATAL_SOLVER, "collision-mesh edge parameter index out of range");
        return cm.param_arrays.edge.data[k];
    };
    auto fparam = [&](uint32_t k) -> const cts_FaceParam & {
        if (k >= cm.param_arrays.face.size)
            fatal(CTS_FATAL_SOLVER, "collision-mesh face parameter index out of range");
        return cm.param_arrays.face.data[k];
    };
    for (uint32_t i = 0; i < sm.nv; ++i) {
        const cts_VertexParam &p = vparam(cm.prop.vertex.data[i].param_index);
        vg[i] = p.ghat, vo[i] = p.offset, vfr[i] = p.friction;
    }
    for (uint32_t e = 0; e < sm.ne; ++e) {
        const cts_EdgeParam &p = eparam(cm.prop.edge.data[e].param_index);
        eg[e] = p.ghat, eo[e] = p.offset, efr[e] = p.friction;
    }
    for (uint32_t f = 0; f < sm.nf; ++f) {
        const cts_FaceParam &p = fparam(cm.prop.face.data[f].param_index);
        fg[f] = p.ghat, fo[f] = p.offset, ffr[f] = p.friction;
    }
    sm.x = b.put(&cm.vertex.data[0].v[0], 3 * (size_t)sm.nv);
    sm.edge = b.put(&cm.edge.data[0].v[0], 2 * (size_t)sm.ne);
    sm.face = b.put(&cm.face.data[0].v[0], 3 * (size_t)sm.nf);
    sm.v_ghat = b.put(vg), sm.v_off = b.put(vo), sm.v_fric = b.put(vfr);
    sm.e_ghat = b.put(eg), sm.e_off = b.put(eo), sm.e_fric = b.put(efr);
    sm.f_ghat = b.put(fg), sm.f_off = b.put(fo), sm.f_fric = b.put(ffr);
}

// Every per-element array of the stepper's scene.
void upload_scene(Backend &b) {
    const cts_DataSet &d = b.host;
    const uint32_t nv = b.nv, ne = b.ne, nf = b.nf, nh = b.nh;
    b.curr = b.put(&d.vertex.curr.data[0].v[0], 3 * (size_t)nv);
    b.prev = b.put(&d.vertex.prev.data[0].v[0], 3 * (size_t)nv);
    const uint32_t *edge = ne ? &d.mesh.mesh.edge.data[0].v[0] : nullptr;
    const uint32_t *face = nf ? &d.mesh.mesh.face.data[0].v[0] : nullptr;
    const uint32_t *hinge = nh ? &d.mesh.mesh.hinge.data[0].v[0] : nullptr;
    uint32_t *d_edge = b.put(edge, 2 * (size_t)ne);
    uint32_t *d_face = b.put(face, 3 * (size_t)nf);
    uint32_t *d_hinge = b.put(hinge, 4 * (size_t)nh);
    const auto &nb = d.mesh.neighbor;
    uint32_t *v2e_off = b.put(nb.vertex.edge.offset, (size_t)nv + 1);
    uint32_t *v2e = b.put(nb.vertex.edge.data, nb.vertex.edge.offset[nv]);
    uint32_t *v2f_off = b.put(nb.vertex.face.offset, (size_t)nv + 1);
    uint32_t *v2f = b.put(nb.vertex.face.data, nb.vertex.face.offset[nv]);
    uint32_t *e2f_off = b.put(nb.edge.face.offset, (size_t)ne + 1);
    uint32_t *e2f = b.put(nb.edge.face.data, nb.edge.face.offset[ne]);

    std::vector<float> vg(nv), vo(nv), vfr(nv), vmass(nv), varea(nv);
    std::vector<float> eg(ne), eo(ne), efr(ne), emass(ne);
    std::vector<float> fg(nf), fo(nf), ffr(nf), fmass(nf), mu(nf), lam(nf), fsl(nf);
    std::vector<float> fpl(nf), fplt(nf);
    bool any_sl = false, any_fpl = false, any_hpl = false;
    std::vector<float> hrest(nh), hlen(nh), harea(nh), hbend(nh), hpl(nh), hplt(nh);
    std::vector<uint32_t> vbody(nv, 0);
    for (uint32_t i = 0; i < nv; ++i) {
        const cts_VertexProp &vp = d.prop.vertex.data[i];
        const cts_VertexParam &p = d.param_arrays.vertex.data[vp.param_index];
        vg[i] = p.ghat, vo[i] = p.offset, vfr[i] = p.friction;
        vmass[i] = vp.mass, varea[i] = vp.area;
    }
    for (uint32_t e = 0; e < ne; ++e) {
        const cts_EdgeProp &ep = d.prop.edge.data[e];
        const cts_EdgeParam &p = d.param_arrays.edge.data[ep.param_index];
        eg[e] = p.ghat, eo[e] = p.offset, efr[e] = p.friction, emass[e] = ep.mass;
    }
    for (uint32_t f = 0; f < nf; ++f) {
        const cts_FaceProp &fp = d.prop.face.data[f];
        const cts_FaceParam &p = d.param_arrays.face.data[fp.param_index];
        fg[f] = p.ghat, fo[f] = p.offset, ffr[f] = p.friction, fmass[f] = fp.mass;
        mu[f] = p.mu, lam[f] = p.lambda;
        fsl[f] = p.strainlimit;
        any_sl = any_sl || p.strainlimit > 0.0f;
        fpl[f] = p.plasticity, fplt[f] = p.plasticity_threshold;
        any_fpl = any_fpl || p.plasticity > 0.0f;
    }
    for (uint32_t h = 0; h < nh; ++h) {
        const cts_HingeProp &hp = d.prop.hinge.data[h];
        const cts_HingeParam &p = d.param_arrays.hinge.data[hp.param_index];
        hrest[h] = hp.rest_angle, hlen[h] = hp.length, harea[h] = hp.area;
        hpl[h] = p.plasticity, hplt[h] = p.plasticity_threshold;
        any_hpl = any_hpl || p.plasticity > 0.0f;
        // hinge_bend_directional (data.hpp:319-326)
        hbend[h] = hp.uv_edge_sin2 < 0.0f
                       ? p.bend
                       : p.bend + p.bend_warp * hp.uv_edge_sin2 +
                             p.bend_weft * (1.0f - hp.uv_edge_sin2);
    }
    b.vflag = b.put(std::vector<uint32_t>(nv, 0));
    b.sh_vflag = b.put(std::vector<uint32_t>(nv, 0));
    b.eflag = b.put(std::vector<uint32_t>(ne, 0));
    b.fflag = b.put(std::vector<uint32_t>(nf, 0));
    b.sh_fflag = b.put(std::vector<uint32_t>(nf, 0));
    b.hflag = b.put(std::vector<uint32_t>(nh, 0));
    b.fix_pos = b.put(std::vector<float>(3 * (size_t)nv, 0.0f));
    b.fix_delta = b.put(std::vector<float>(3 * (size_t)nv, 0.0f));
    b.inv_rest = b.put(nf ? &d.inv_rest2x2.data[0].m[0] : nullptr, 4 * (size_t)nf);

    ppf_contact_mesh &m = b.sc.mesh;
    m = ppf_contact_mesh{};
    m.nv = nv, m.ne = ne, m.nf = nf;
    m.x = b.curr;
    m.edge = d_edge, m.face = d_face;
    m.v2e_off = v2e_off, m.v2e = v2e, m.v2f_off = v2f_off, m.v2f = v2f;
    m.e2f_off = e2f_off, m.e2f = e2f;
    m.v_ghat = b.put(vg), m.v_off = b.put(vo), m.e_ghat = b.put(eg), m.e_off = b.put(eo);
    m.f_ghat = b.put(fg), m.f_off = b.put(fo);
    m.vflag = b.vflag, m.vbody = b.put(vbody), m.eflag = b.eflag, m.fflag = b.fflag;

    ppf_shell &sh = b.sc.shell;
    sh = ppf_shell{};
    sh.nv = nv, sh.nf = nf, sh.nh = nh;
    sh.face = d_face, sh.hinge = d_hinge, sh.v2f_off = v2f_off, sh.v2f = v2f;
    sh.vmass = b.put(vmass), sh.varea = b.put(varea), sh.vflag = b.sh_vflag;
    sh.fmass = b.put(fmass), sh.inv_rest = b.inv_rest, sh.f_mu = b.put(mu),
    sh.f_lambda = b.put(lam), sh.fflag = b.sh_fflag;
    b.h_rest = b.put(hrest);
    sh.h_rest = b.h_rest, sh.h_length = b.put(hlen), sh.h_area = b.put(harea),
    sh.h_bend = b.put(hbend), sh.hflag = b.hflag;
    // strain limits (shrink is refused, so every face's shrink is 1)
    sh.f_strainlimit = any_sl ? b.put(fsl) : nullptr;
    sh.f_shrink = nullptr;
    // plasticity (plasticity.cu:21-78, 145-185): the stepper rewrites
    // inv_rest and h_rest in place after each committed step
    if (any_fpl) {
        sh.f_plasticity = b.put(fpl), sh.f_plastic_threshold = b.put(fplt);
    }
    if (any_hpl) {
        sh.h_plasticity = b.put(hpl), sh.h_plastic_threshold = b.put(hplt);
    }

    b.sc.prev = b.prev;
    b.sc.emass = b.put(emass);
    b.sc.v_fric = b.put(vfr), b.sc.e_fric = b.put(efr), b.sc.f_fric = b.put(ffr);
    b.sc.fix_pos = b.fix_pos;
    b.sc.fix_delta = b.fix_delta;
    upload_collision(b);
}

// Intersection check of the mesh at positions d_x (3 nv floats); keeps the
// records for fetch_intersection_records.
uint32_t check_state(Backend &b, const float *d_x) {
    ppf_narrow *nw = nullptr;
    ppf_or_die(ppf_narrow_create(&nw), "ppf_narrow_create");
    ppf_contact_mesh m = b.sc.mesh;
    m.x = d_x;
    ppf_or_die(ppf_narrow_build(nw, &m, b.s), "ppf_narrow_build");
    std::vector<ppf_isect_record> rec(kMaxRecords);
    uint32_t count = 0;
    const float *ve = b.sc.emass;
    ppf_or_die(ppf_intersection_check(nw, &m, b.sc.collision.nv ? &b.sc.collision : nullptr,
                                      b.sc.shell.vmass, ve, b.sc.shell.fmass, b.nv, rec.data(),
                                      kMaxRecords, &count, b.s),
               "ppf_intersection_check");
    ppf_narrow_destroy(nw);
    b.record_count = count;
    b.records.assign(std::min(count, kMaxRecords), cts_IntersectionRecord{});
    static_assert(sizeof(ppf_isect_record) == sizeof(cts_IntersectionRecord),
                  "intersection record layouts differ");
    if (!b.records.empty()) {
        std::memcpy(b.records.data(), rec.data(), b.records.size() * sizeof(ppf_isect_record));
    }
    return count;
}

// The host's vertex indices, checked, in the backend's index scratch.
uint32_t *device_indices(Backend &b, const uint32_t *idx, uint32_t count) {
    for (uint32_t k = 0; k < count; ++k) {
        if (idx[k] >= b.nv) {
            fatal(CTS_FATAL_SOLVER, "vertex index %u out of range (%u vertices)", idx[k], b.nv);
        }
    }
    auto *d = static_cast<uint32_t *>(b.grow(b.scr_idx, count * sizeof(uint32_t)));
    hip_or_die(hipMemcpyAsync(d, idx, count * sizeof(uint32_t), hipMemcpyHostToDevice, b.s),
               "hipMemcpyAsync");
    return d;
}

void require_ready(const char *what) {
    if (!g.ready) {
        fatal(CTS_FATAL_SOLVER, "%s called before a successful initialize", what);
    }
}

} // namespace

extern "C" {

unsigned char ppf_fatal_code(void) { return g_ppf_fatal_code; }

const char *ppf_fatal_detail(void) { return g_ppf_fatal_detail; }

void set_log_path(const char *data_dir) { g.log_dir = data_dir ? data_dir : ""; }

bool initialize(const cts_DataSet *data, cts_ParamSet *param) {
    g.release();
    g_ppf_fatal_code = 0;
    g_ppf_fatal_detail[0] = 0;
    if (!data || !param) {
        set_detail(CTS_FATAL_SOLVER, "initialize: null DataSet or ParamSet");
        return false;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_detail(CTS_FATAL_DRIVER, "no HIP device is visible to this process; this build "
                                     "runs on gfx950 (MI355X)");
        return false;
    }
    if (const char *why = bad_sizes(*data)) {
        set_detail(CTS_FATAL_SOLVER, "initialize: %s", why);
        return false;
    }
    if (const char *what = unsupported(*data, *param)) {
        set_detail(CTS_FATAL_SOLVER, "initialize: this backend does not support %s", what);
        return false;
    }
    g.host = *data;
    g.param = param;
    g.nv = data->vertex.curr.size;
    g.ne = data->mesh.mesh.edge.size;
    g.nf = data->mesh.mesh.face.size;
    g.nh = data->mesh.mesh.hinge.size;
    hip_or_die(hipStreamCreateWithFlags(&g.s, hipStreamNonBlocking), "hipStreamCreate");
    g.derive_fixed();
    upload_scene(g);
    g.copy_colliders();
    g.upload_flags();
    g.upload_pins();
    ppf_or_die(ppf_stepper_create(&g.sc, g.s, &g.stepper), "ppf_stepper_create");
    g.ready = true;
    if (!param->disable_contact && g.nv) {
        // the start state must be intersection free at prev and curr
        // (main.cu:230-262)
        if (check_state(g, g.prev) || check_state(g, g.curr)) {
            set_detail(CTS_FATAL_INIT_INTERSECTION,
                       "the scene is already self-intersecting at t=0; the solver advances "
                       "from a separated configuration and cannot untangle one");
            return false;
        }
    }
    return true;
}

void advance(cts_StepResult *result) {
    require_ready("advance");
    cts_ParamSet &p = *g.param;
    ppf_step_params sp{};
    sp.dt = p.dt * p.playback;
    sp.prev_dt = p.prev_dt;
    for (int d = 0; d < 3; ++d) {
        sp.gravity[d] = p.gravity.v[d];
        sp.wind[d] = p.wind.v[d];
    }
    sp.cg_tol = p.cg_tol;
    sp.cg_max_iter = p.cg_max_iter;
    sp.target_toi = p.target_toi;
    sp.line_search_max_t = p.line_search_max_t;
    sp.ccd_eps = p.ccd_eps;
    sp.max_dx = p.max_dx;
    sp.friction_eps = p.friction_eps;
    sp.friction_mode = (int32_t)p.friction_mode;
    sp.min_newton_steps = p.min_newton_steps;
    sp.max_newton_steps = p.max_newton_steps;
    sp.air_density = p.air_density;
    sp.air_friction = p.air_friction;
    sp.disable_contact = p.disable_contact ? 1 : 0;
    sp.check_intersection = p.disable_contact ? 0 : 1;
    sp.constraint_tol = p.constraint_tol;
    sp.eiganalysis_eps = p.eiganalysis_eps;
    ppf_step_result r{};
    ppf_or_die(ppf_stepper_advance(g.stepper, &g.sc, &sp, &r, g.s), "advance");
    g.carried = false; // the step's own contact pattern replaces a carried one
    cts_StepResult out{};
    out.ccd_success = r.status != 2;
    out.pcg_success = r.status != 1;
    out.intersection_free = r.status != 4;
    out.newton_progress = r.status != 5;
    out.pin_feasible = r.status != 6;
    out.contact_separated = r.status != 3;
    if (r.status == 0 || r.status == 4) {
        // the step is committed (main.cu:1572-1629)
        p.prev_dt = r.dt;
        p.time += (double)(p.prev_dt / p.playback);
    }
    out.time = p.time;
    g.record_count = ppf_stepper_intersection_records(g.stepper, nullptr, 0);
    if (g.record_count) {
        std::vector<ppf_isect_record> rec(kMaxRecords);
        ppf_stepper_intersection_records(g.stepper, rec.data(), kMaxRecords);
        g.records.assign(std::min(g.record_count, kMaxRecords), cts_IntersectionRecord{});
        std::memcpy(g.records.data(), rec.data(), g.records.size() * sizeof(ppf_isect_record));
    } else {
        g.records.clear();
    }
    g.log_step(r);
    if (result) {
        *result = out;
    }
}

void fetch(void) {
    require_ready("fetch");
    hip_or_die(hipMemcpy(&g.host.vertex.curr.data[0].v[0], g.curr, 3 * (size_t)g.nv * 4,
                         hipMemcpyDeviceToHost),
               "fetch");
    hip_or_die(hipMemcpy(&g.host.vertex.prev.data[0].v[0], g.prev, 3 * (size_t)g.nv * 4,
                         hipMemcpyDeviceToHost),
               "fetch");
    if (g.host.statistics_contact_count.size) {
        std::memset(g.host.statistics_contact_count.data, 0,
                    g.host.statistics_contact_count.size * sizeof(uint32_t));
    }
}

void fetch_inv_rest(void) {
    require_ready("fetch_inv_rest");
    if (g.nf) {
        hip_or_die(hipMemcpy(&g.host.inv_rest2x2.data[0].m[0], g.inv_rest,
                             4 * (size_t)g.nf * sizeof(float), hipMemcpyDeviceToHost),
                   "fetch_inv_rest");
    }
}

void fetch_rest_angles(void) {
    // HingeProp.rest_angle as bend plasticity left it (main.cu:1950-1968);
    // rods (VertexProp.rest_bend_angle) are refused, so only hinges change
    require_ready("fetch_rest_angles");
    if (g.nh) {
        std::vector<float> rest(g.nh);
        hip_or_die(hipMemcpy(rest.data(), g.h_rest, g.nh * sizeof(float),
                             hipMemcpyDeviceToHost),
                   "fetch_rest_angles");
        for (uint32_t h = 0; h < g.nh; ++h) {
            g.host.prop.hinge.data[h].rest_angle = rest[h];
        }
    }
}

void fetch_dyn_counts(uint32_t *n_value, uint32_t *n_offset) {
    require_ready("fetch_dyn_counts");
    uint32_t nnz = 0;
    if (g.carried) {
        nnz = g.carried_rp[g.nv];
    } else {
        ppf_or_die(ppf_stepper_contact_pattern(g.stepper, &nnz, nullptr, nullptr, nullptr, g.s),
                   "fetch_dyn_counts");
    }
    if (n_value) *n_value = nnz;
    if (n_offset) *n_offset = g.nv + 1;
}

// The contact pattern the next step starts from: the last step's contact
// blocks (upper triangle, rows = vertices), or the pattern update_dyn carried
// in if no step has run since (main.cu:1970-1986, csrmat.cu:136-153). Values
// are written when `value` is non-null (the carried pattern has none: zeros).
void fetch_dyn(uint32_t *index, float *value, uint32_t *offset) {
    require_ready("fetch_dyn");
    const size_t rows = (size_t)g.nv + 1;
    if (g.carried) {
        const uint32_t nnz = g.carried_rp[g.nv];
        if (offset) std::memcpy(offset, g.carried_rp.data(), rows * sizeof(uint32_t));
        if (index && nnz) std::memcpy(index, g.carried_ci.data(), nnz * sizeof(uint32_t));
        if (value && nnz) std::memset(value, 0, 9 * (size_t)nnz * sizeof(float));
        return;
    }
    uint32_t nnz = 0;
    ppf_or_die(ppf_stepper_contact_pattern(g.stepper, &nnz, nullptr, nullptr, nullptr, g.s),
               "fetch_dyn");
    const size_t off_b = rows * 4, idx_b = (size_t)nnz * 4, val_b = value ? 36 * (size_t)nnz : 0;
    auto *base = static_cast<char *>(g.grow(g.scr_dyn, off_b + idx_b + val_b + 16));
    auto *d_off = reinterpret_cast<uint32_t *>(base);
    auto *d_idx = reinterpret_cast<uint32_t *>(base + off_b);
    auto *d_val = value ? reinterpret_cast<float *>(base + off_b + idx_b) : nullptr;
    ppf_or_die(ppf_stepper_contact_pattern(g.stepper, &nnz, d_off, d_idx, d_val, g.s),
               "fetch_dyn");
    if (offset) {
        hip_or_die(hipMemcpyAsync(offset, d_off, off_b, hipMemcpyDeviceToHost, g.s), "fetch_dyn");
    }
    if (index && nnz) {
        hip_or_die(hipMemcpyAsync(index, d_idx, idx_b, hipMemcpyDeviceToHost, g.s), "fetch_dyn");
    }
    if (value && nnz) {
        hip_or_die(hipMemcpyAsync(value, d_val, val_b, hipMemcpyDeviceToHost, g.s), "fetch_dyn");
    }
    hip_or_die(hipStreamSynchronize(g.s), "fetch_dyn");
}

// A carried-over contact pattern (resume, csrmat.cu:136-153): kept as the
// pattern fetch_dyn reports until the next step. The stepper rebuilds its
// contact blocks from the step's own pairs every Newton iteration, so the
// pattern does not change what a step computes. Checked for well-formedness.
void update_dyn(const uint32_t *index, const uint32_t *offset) {
    require_ready("update_dyn");
    if (!index || !offset) {
        return;
    }
    if (offset[0] != 0) {
        fatal(CTS_FATAL_SOLVER, "update_dyn: offset[0] != 0");
    }
    for (uint32_t i = 0; i < g.nv; ++i) {
        if (offset[i + 1] < offset[i]) {
            fatal(CTS_FATAL_SOLVER, "update_dyn: offsets decrease at row %u", i);
        }
        for (uint32_t k = offset[i]; k < offset[i + 1]; ++k) {
            if (index[k] >= g.nv) {
                fatal(CTS_FATAL_SOLVER, "update_dyn: column %u out of range", index[k]);
            }
        }
    }
    g.carried_rp.assign(offset, offset + g.nv + 1);
    g.carried_ci.assign(index, index + offset[g.nv]);
    g.carried = true;
}

void update_constraint(const cts_Constraint *constraint) {
    require_ready("update_constraint");
    if (!constraint) {
        return;
    }
    const cts_Constraint &c = *constraint;
    if (c.pull.size || c.torque_groups.size || c.torque_vertices.size || c.stitch.size) {
        fatal(CTS_FATAL_SOLVER, "update_constraint: pulls, torques and stitches are not "
                                "supported by this backend");
    }
    for (uint32_t k = 0; k < c.fix.size; ++k) {
        if (c.fix.data[k].index >= g.nv) {
            fatal(CTS_FATAL_SOLVER, "update_constraint: pin index %u out of range",
                  c.fix.data[k].index);
        }
    }
    g.host.constraint = c;
    g.derive_fixed();
    g.copy_colliders();
    g.upload_flags();
    g.upload_pins();
    ppf_or_die(ppf_stepper_update_pins(g.stepper, &g.sc, g.s), "update_constraint");
}

void update_rest_shape(const cts_RestShapeUpdate *update) {
    require_ready("update_rest_shape");
    if (!update) {
        return;
    }
    if (update->inv_rest2x2.size) {
        if (update->inv_rest2x2.size != g.nf) {
            fatal(CTS_FATAL_SOLVER, "update_rest_shape: %u inverse rest matrices for %u faces",
                  update->inv_rest2x2.size, g.nf);
        }
        hip_or_die(hipMemcpy(g.inv_rest, &update->inv_rest2x2.data[0].m[0],
                             4 * (size_t)g.nf * sizeof(float), hipMemcpyHostToDevice),
                   "update_rest_shape");
    }
    if (update->exclude_face.size) {
        for (uint32_t i = 0; i < update->exclude_face.size && i < g.nf; ++i) {
            g.host.prop.face.data[i].rest_excluded = update->exclude_face.data[i] != 0;
        }
        g.upload_flags();
    }
}

void override_velocity(const uint32_t *indices, uint32_t count, float vx, float vy, float vz,
                       float dt) {
    require_ready("override_velocity");
    if (count == 0 || dt <= 0.0f) {
        return;
    }
    uint32_t *d = device_indices(g, indices, count);
    k_override_velocity<<<(count + kThreads - 1) / kThreads, kThreads, 0, g.s>>>(
        d, count, g.curr, g.prev, vx * dt, vy * dt, vz * dt);
    hip_or_die(hipGetLastError(), "override_velocity");
    hip_or_die(hipStreamSynchronize(g.s), "override_velocity");
}

void gather_current_positions(const uint32_t *indices, uint32_t count, float *out) {
    require_ready("gather_current_positions");
    if (count == 0) {
        return;
    }
    uint32_t *d = device_indices(g, indices, count);
    auto *d_out = static_cast<float *>(g.grow(g.scr_out, 3 * (size_t)count * sizeof(float)));
    k_gather<<<(count + kThreads - 1) / kThreads, kThreads, 0, g.s>>>(d, count, g.curr, d_out);
    hip_or_die(hipGetLastError(), "gather_current_positions");
    hip_or_die(hipMemcpyAsync(out, d_out, 3 * (size_t)count * sizeof(float),
                              hipMemcpyDeviceToHost, g.s),
               "gather_current_positions");
    hip_or_die(hipStreamSynchronize(g.s), "gather_current_positions");
}

void override_angular_velocity(const uint32_t *indices, uint32_t count, float wx, float wy,
                               float wz, float cx, float cy, float cz, float dt) {
    require_ready("override_angular_velocity");
    if (count == 0 || dt <= 0.0f) {
        return;
    }
    uint32_t *d = device_indices(g, indices, count);
    k_override_angular<<<(count + kThreads - 1) / kThreads, kThreads, 0, g.s>>>(
        d, count, g.curr, g.prev, wx, wy, wz, cx, cy, cz, dt);
    hip_or_die(hipGetLastError(), "override_angular_velocity");
    hip_or_die(hipStreamSynchronize(g.s), "override_angular_velocity");
}

void init_collision_windows(const uint32_t *vert_dmap, uint32_t vert_count,
                            const float *windows, const uint32_t *window_counts,
                            uint32_t n_groups) {
    require_ready("init_collision_windows");
    g.cw_dmap.assign(vert_dmap, vert_dmap + vert_count);
    g.cw_counts.assign(window_counts, window_counts + n_groups);
    g.cw_windows.assign(windows, windows + (size_t)n_groups * kMaxWindows * 2);
    for (uint32_t i = 0; i < vert_count; ++i) {
        if (g.cw_dmap[i] >= n_groups) {
            fatal(CTS_FATAL_SOLVER, "init_collision_windows: group %u out of range",
                  g.cw_dmap[i]);
        }
    }
    g.cw = true;
    g.v_active.assign(g.nv, 1);
    g.e_active.assign(g.ne, 1);
    g.upload_flags();
}

void refresh_collision_active(float time) {
    require_ready("refresh_collision_active");
    if (!g.cw) {
        return;
    }
    g.refresh_active(time);
    g.upload_flags();
}

uint32_t fetch_intersection_records(cts_IntersectionRecord *out, uint32_t max_count) {
    const uint32_t n = std::min<uint32_t>((uint32_t)g.records.size(), max_count);
    for (uint32_t k = 0; out && k < n; ++k) {
        out[k] = g.records[k];
    }
    return n;
}

} // extern "C"
