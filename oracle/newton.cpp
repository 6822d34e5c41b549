// newton.cpp -- CPU oracle: one time step (the Newton loop).
//
// TEST INFRASTRUCTURE ONLY (the checker of the HIP path, never shipped).
//
// Restates main_helper::advance (main/main.cu:267-1649) for shells with fix
// pins, self contact and a static collision mesh, on the oracle's own
// stages: inertial target (main.cu:491-511), elastic + inertia assembly,
// contact pairs and their barrier / friction terms with the elastic blocks as
// stiffness reference, the merged upper matrix, Dirichlet elimination of
// pinned vertices (main.cu:954-1049), block-Jacobi PCG, the max_dx clamp,
// the CCD line search and the toi bookkeeping of main.cu:1177-1513, then the
// intersection check and the prev / curr commit (main.cu:1515-1600).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// Layouts shared with the other oracle files (C linkage: only the layout
// matters).
struct MeshView {
    uint32_t nv, ne, nf;
    const float *x;
    const uint32_t *edge, *face;
    const uint32_t *v2e_off, *v2e, *v2f_off, *v2f, *e2f_off, *e2f;
    const float *v_ghat, *v_off, *e_ghat, *e_off, *f_ghat, *f_off;
    const uint32_t *vflag, *vbody, *eflag, *fflag;
};
struct ShellView {
    uint32_t nv, nf, nh;
    const float *x, *x0, *target;
    const uint32_t *face, *hinge, *v2f_off, *v2f;
    const float *vmass, *varea;
    const uint32_t *vflag;
    const float *fmass, *inv_rest, *f_mu, *f_lambda;
    const uint32_t *fflag;
    const float *h_rest, *h_length, *h_area, *h_bend;
    const uint32_t *hflag;
    const float *f_strainlimit, *f_shrink; // null: no face is strain-limited / no shrink
    const float *f_plasticity, *f_plastic_threshold; // null: no plasticity
    const float *h_plasticity, *h_plastic_threshold;
};
struct StaticView {
    uint32_t nv, ne, nf;
    const float *x;
    const uint32_t *edge, *face;
    const float *v_ghat, *v_off, *e_ghat, *e_off, *f_ghat, *f_off;
    const float *v_fric, *e_fric, *f_fric;
};
struct ElasticParams {
    float dt, air_density, air_friction;
    float wind[3];
};
struct Pair {
    uint32_t kind, a, b, count;
};
struct ContactParams {
    const float *x0;
    const float *mass;
    const float *v_fric, *e_fric, *f_fric;
    int friction_mode;
    float friction_eps;
    const uint32_t *el_rp, *el_col;
    const float *el_val;
    const StaticView *sm;
};
struct LsParams {
    const float *x1;
    const float *vmass, *emass, *fmass;
    float max_t, ccd_eps;
};
struct PcgResult {
    uint32_t iter;
    float resid;
    int32_t status, break_cause, break_iter;
    float err0;
};
struct IsectRecord {
    uint32_t type, elem0, elem1, num_verts0, num_verts1;
    float positions[15];
};

} // namespace

extern "C" {
uint32_t oracle_elastic_assemble(const ShellView *s, const ElasticParams *p,
                                 const uint32_t *rp, const uint32_t *ci, float *force,
                                 float *cdiag, float *val);
uint64_t oracle_proximity(const MeshView *m, Pair *out, uint64_t cap);
uint64_t oracle_static_proximity(const MeshView *m, const StaticView *sm, const float *vmass,
                                 const float *emass, const float *fmass, Pair *out,
                                 uint64_t cap);
uint64_t oracle_contact_assemble(const MeshView *m, const ContactParams *cp, const Pair *pairs,
                                 uint64_t npairs, float *grad, uint32_t *rows, uint32_t *cols,
                                 float *vals, uint64_t cap);
int oracle_pcg(uint32_t n, const uint32_t *rp, const uint32_t *col, const float *val,
               const float *cdiag, const float *b, float *x, float tol, uint32_t max_iter,
               float *err_hist, PcgResult *out);
float oracle_line_search(const MeshView *m, const StaticView *sm, const LsParams *lp,
                         int *overlap);
uint32_t oracle_intersection_check(const MeshView *m, const StaticView *sm, const float *vmass,
                                   const float *emass, const float *fmass,
                                   uint32_t surface_vert_count, IsectRecord *rec, uint32_t cap);
struct oracle_sphere {
    float center[3];
    float ghat, friction, radius, thickness;
    bool bowl, reverse, kinematic;
};
struct oracle_floor {
    float ground[3];
    float ghat, friction, thickness;
    float up[3];
    bool kinematic;
};
uint32_t oracle_analytic_assemble(const MeshView *m, const ContactParams *cp,
                                  const oracle_sphere *sp, uint32_t nsp,
                                  const oracle_floor *fl, uint32_t nfl,
                                  float constraint_tol, float *grad, float *cdiag);
void oracle_strainlimit_assemble(uint32_t nv, uint32_t nf, const uint32_t *face,
                                 const float *inv_rest, const uint32_t *fflag,
                                 const float *fmass, const float *sl, const float *shrink,
                                 const float *x, const float *x0, const uint32_t *rp,
                                 const uint32_t *ci, float eig_eps, float *force, float *value);
void oracle_plasticity(uint32_t nf, uint32_t nh, const uint32_t *face, const uint32_t *hinge,
                       const uint32_t *hflag, const float *x, const float *f_rate,
                       const float *f_thr, const float *h_rate, const float *h_thr, float dt,
                       float *inv_rest, float *h_rest);
float oracle_strainlimit_toi(uint32_t nf, const uint32_t *face, const float *inv_rest,
                             const uint32_t *fflag, const float *sl, const float *shrink,
                             const float *x0, const float *x1, float max_t);
void oracle_analytic_toi(uint32_t nv, const float *x0a, const float *x1a, const uint32_t *vflag,
                         const float *vmass, const oracle_sphere *sp, uint32_t nsp,
                         const oracle_floor *fl, uint32_t nfl, float max_t, float *toi,
                         uint32_t *pin_bad);

struct oracle_scene {
    MeshView mesh;
    ShellView shell;
    StaticView collision;
    float *prev;
    const float *emass;
    const float *v_fric, *e_fric, *f_fric;
    float *fix_pos;
    const float *fix_delta;
    const oracle_sphere *spheres;
    uint32_t nsphere;
    const oracle_floor *floors;
    uint32_t nfloor;
};

struct oracle_step_params {
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
    float constraint_tol;
    float eiganalysis_eps;
};

struct oracle_step_result {
    int32_t status;
    uint32_t newton_steps;
    uint32_t pcg_iter_total, pcg_iter_max;
    uint32_t pairs_max;
    uint32_t intersections;
    float toi_advanced;
    float dt;
};

// The last Newton system the oracle solved (merged upper blocks with the
// eliminated ones zeroed, diagonal, right-hand side), for the tests.
static struct {
    uint32_t n = 0;
    std::vector<uint32_t> rp, ci;
    std::vector<float> val, cdiag, rhs;
} g_last;

uint32_t oracle_last_system(uint32_t *n, uint32_t *rp, uint32_t *ci, float *val, float *cdiag,
                            float *rhs) {
    if (n) *n = g_last.n;
    if (rp) std::copy(g_last.rp.begin(), g_last.rp.end(), rp);
    if (ci) std::copy(g_last.ci.begin(), g_last.ci.end(), ci);
    if (val) std::copy(g_last.val.begin(), g_last.val.end(), val);
    if (cdiag) std::copy(g_last.cdiag.begin(), g_last.cdiag.end(), cdiag);
    if (rhs) std::copy(g_last.rhs.begin(), g_last.rhs.end(), rhs);
    return (uint32_t)g_last.ci.size();
}

// Per Newton iteration of the last step: PCG iterations, step scale of the
// max_dx clamp, CCD toi, largest |dx| (4 floats each), for the tests.
static std::vector<float> g_trace;

uint32_t oracle_step_trace(float *out, uint32_t cap) {
    const uint32_t k = (uint32_t)(g_trace.size() / 4);
    for (uint32_t i = 0; i < std::min(k, cap); ++i)
        std::copy(g_trace.begin() + 4 * i, g_trace.begin() + 4 * i + 4, out + 4 * i);
    return k;
}

// One step; mesh.x (curr) and prev are updated on success. Returns 0.
int oracle_step(oracle_scene *sc, const oracle_step_params *p, oracle_step_result *out) {
    g_trace.clear();
    *out = oracle_step_result{};
    const uint32_t n = sc->mesh.nv;
    const size_t n3 = 3 * (size_t)n;
    float *curr = const_cast<float *>(sc->mesh.x);
    // elastic pattern: face / hinge vertex pairs + diagonal, upper triangle
    std::vector<uint64_t> keys;
    for (uint32_t i = 0; i < n; ++i) keys.push_back((uint64_t)i * n + i);
    for (uint32_t k = 0; k < sc->shell.nf; ++k)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                uint32_t r = sc->shell.face[3 * k + a], c = sc->shell.face[3 * k + b];
                if (r <= c) keys.push_back((uint64_t)r * n + c);
            }
    for (uint32_t k = 0; k < sc->shell.nh; ++k)
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                uint32_t r = sc->shell.hinge[4 * k + a], c = sc->shell.hinge[4 * k + b];
                if (r <= c) keys.push_back((uint64_t)r * n + c);
            }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    std::vector<uint32_t> erp(n + 1, 0), eci(keys.size());
    for (size_t k = 0; k < keys.size(); ++k) {
        eci[k] = (uint32_t)(keys[k] % n);
        erp[keys[k] / n + 1]++;
    }
    for (uint32_t i = 0; i < n; ++i) erp[i + 1] += erp[i];
    std::vector<uint32_t> mask(n, 0);
    uint32_t npinned = 0;
    if (sc->fix_pos)
        for (uint32_t i = 0; i < n; ++i)
            if (sc->mesh.vflag[i] & 1u) mask[i] = 1, ++npinned;

    std::vector<float> ex(curr, curr + n3), target(n3), dx(n3), force(n3), cgrad(n3),
        cdiag(9 * (size_t)n), eval(9 * keys.size());
    float dt = p->dt;
    const float prev_dt = p->prev_dt > 0.0f ? p->prev_dt : p->dt;
    auto compute_target = [&](float h) {
        for (uint32_t i = 0; i < n; ++i) {
            if ((sc->mesh.vflag[i] & 1u) && sc->fix_pos) {
                for (int d = 0; d < 3; ++d) target[3 * i + d] = sc->fix_pos[3 * i + d];
                continue;
            }
            const float tr = h / prev_dt, h2 = h * h;
            for (int d = 0; d < 3; ++d) {
                const float x1 = curr[3 * i + d], x0 = sc->prev[3 * i + d];
                const float y = (x1 - x0) * tr + h2 * p->gravity[d];
                target[3 * i + d] = x1 + y;
            }
        }
    };
    compute_target(dt);
    double toi_advanced = 0.0;
    uint32_t step = 1;
    bool final_step = false;
    std::vector<Pair> pairs;
    while (true) {
        if (!final_step && p->max_newton_steps > 0 && step >= p->max_newton_steps) {
            out->status = 5;
            break;
        }
        if (final_step) {
            // rewind_kinematic_fix (main.cu:531-545): pins back to the
            // integrated fraction of their step, in place
            if (sc->fix_pos && sc->fix_delta && toi_advanced < 1.0) {
                const float back = 1.0f - (float)toi_advanced;
                for (uint32_t i = 0; i < n; ++i)
                    if (sc->mesh.vflag[i] & 1u)
                        for (int d = 0; d < 3; ++d)
                            sc->fix_pos[3 * i + d] -= back * sc->fix_delta[3 * i + d];
            }
            dt *= (float)toi_advanced;
            compute_target(dt);
        }
        // inertia + elastic
        for (uint32_t i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d)
                dx[3 * i + d] = mask[i] ? ex[3 * i + d] - target[3 * i + d] : 0.0f;
        ShellView sh = sc->shell;
        sh.x = ex.data();
        sh.x0 = curr;
        sh.target = target.data();
        ElasticParams ep{dt, p->air_density, p->air_friction, {p->wind[0], p->wind[1], p->wind[2]}};
        oracle_elastic_assemble(&sh, &ep, erp.data(), eci.data(), force.data(), cdiag.data(),
                                eval.data());
        // contact
        MeshView m = sc->mesh;
        m.x = ex.data();
        pairs.clear();
        if (!p->disable_contact) {
            uint64_t c = oracle_proximity(&m, nullptr, 0);
            pairs.resize(c);
            oracle_proximity(&m, pairs.data(), c);
            if (sc->collision.nv) {
                uint64_t c2 = oracle_static_proximity(&m, &sc->collision, sc->shell.vmass,
                                                      sc->emass, sc->shell.fmass, nullptr, 0);
                pairs.resize(c + c2);
                oracle_static_proximity(&m, &sc->collision, sc->shell.vmass, sc->emass,
                                        sc->shell.fmass, pairs.data() + c, c2);
            }
        }
        out->pairs_max = std::max<uint32_t>(out->pairs_max, (uint32_t)pairs.size());
        ContactParams cp{curr, sc->shell.vmass, sc->v_fric, sc->e_fric, sc->f_fric,
                         p->friction_mode, p->friction_eps, erp.data(), eci.data(),
                         eval.data(), sc->collision.nv ? &sc->collision : nullptr};
        const uint64_t cap = 16 * std::max<size_t>(1, pairs.size());
        std::vector<uint32_t> crow(cap), ccol(cap);
        std::vector<float> cval(9 * cap);
        const uint64_t nb = oracle_contact_assemble(&m, &cp, pairs.data(), pairs.size(),
                                                    cgrad.data(), crow.data(), ccol.data(),
                                                    cval.data(), cap);
        // analytic spheres and floors (contact.cu:1154-1382) into the contact
        // gradient and the per-vertex diagonal
        oracle_analytic_assemble(&m, &cp, sc->spheres, sc->nsphere, sc->floors, sc->nfloor,
                                 p->constraint_tol, cgrad.data(), cdiag.data());
        // strain limiting (main.cu:809-822), its stiffness from the elastic
        // blocks alone, its terms then joining force and blocks
        oracle_strainlimit_assemble(n, sh.nf, sh.face, sh.inv_rest, sh.fflag, sh.fmass,
                                    sh.f_strainlimit, sh.f_shrink, ex.data(), curr, erp.data(),
                                    eci.data(), p->eiganalysis_eps, force.data(), eval.data());
        for (size_t k = 0; k < n3; ++k) force[k] = force[k] + cgrad[k];
        // merged upper matrix: elastic pattern + contact blocks (A + B)
        std::vector<uint32_t> mrp(n + 1, 0), mci;
        std::vector<float> mval;
        {
            std::vector<uint32_t> crp(n + 1, 0);
            for (uint64_t k = 0; k < nb; ++k) crp[crow[k] + 1]++;
            for (uint32_t i = 0; i < n; ++i) crp[i + 1] += crp[i];
            for (uint32_t i = 0; i < n; ++i) {
                uint32_t a = erp[i], ae = erp[i + 1], b = crp[i], be = crp[i + 1];
                while (a < ae || b < be) {
                    if (b >= be || (a < ae && eci[a] < ccol[b])) {
                        mci.push_back(eci[a]);
                        for (int e = 0; e < 9; ++e) mval.push_back(eval[9 * (size_t)a + e]);
                        ++a;
                    } else if (a >= ae || ccol[b] < eci[a]) {
                        mci.push_back(ccol[b]);
                        for (int e = 0; e < 9; ++e) mval.push_back(cval[9 * (size_t)b + e]);
                        ++b;
                    } else {
                        mci.push_back(eci[a]);
                        for (int e = 0; e < 9; ++e)
                            mval.push_back(eval[9 * (size_t)a + e] + cval[9 * (size_t)b + e]);
                        ++a;
                        ++b;
                    }
                }
                mrp[i + 1] = (uint32_t)mci.size();
            }
        }
        if (npinned) {
            // Lifts of the couplings to pinned vertices into the free rows
            // (main.cu:954-1049; the reference scatters them atomically, in no
            // fixed order). Here each free row takes its own lifts (row i
            // free, column j pinned: f_i -= B_ij p_j) in entry order, then the
            // mirror lifts (row j pinned, column i free: f_i -= B_ji^T p_j) in
            // (source row, entry) order. Blocks touching a pinned vertex are
            // zeroed.
            auto lift = [&](uint32_t row, const float *b, uint32_t pv, bool tr) {
                float pr[3];
                for (int d = 0; d < 3; ++d) pr[d] = ex[3 * pv + d] - target[3 * pv + d];
                for (int r = 0; r < 3; ++r) {
                    float c = !tr ? (b[r] * pr[0] + b[r + 3] * pr[1]) + b[r + 6] * pr[2]
                                  : (b[3 * r] * pr[0] + b[3 * r + 1] * pr[1]) + b[3 * r + 2] * pr[2];
                    force[3 * row + r] += -c;
                }
            };
            for (uint32_t i = 0; i < n; ++i) {
                if (mask[i]) continue;
                for (uint32_t k = mrp[i]; k < mrp[i + 1]; ++k)
                    if (mask[mci[k]]) lift(i, &mval[9 * (size_t)k], mci[k], false);
            }
            for (uint32_t i = 0; i < n; ++i) {
                if (!mask[i]) continue;
                for (uint32_t k = mrp[i]; k < mrp[i + 1]; ++k)
                    if (!mask[mci[k]]) lift(mci[k], &mval[9 * (size_t)k], i, true);
            }
            for (uint32_t i = 0; i < n; ++i)
                for (uint32_t k = mrp[i]; k < mrp[i + 1]; ++k)
                    if (mask[i] || mask[mci[k]]) std::fill(&mval[9 * (size_t)k], &mval[9 * (size_t)k] + 9, 0.0f);
            for (uint32_t i = 0; i < n; ++i)
                if (mask[i]) {
                    for (int e = 0; e < 9; ++e) cdiag[9 * (size_t)i + e] = (e % 4 == 0) ? 1.0f : 0.0f;
                    for (int d = 0; d < 3; ++d) force[3 * i + d] = ex[3 * i + d] - target[3 * i + d];
                }
        }
        g_last.n = n;
        g_last.rp = mrp;
        g_last.ci = mci;
        g_last.val = mval;
        g_last.cdiag = cdiag;
        g_last.rhs = force;
        PcgResult pr{};
        oracle_pcg(n, mrp.data(), mci.data(), mval.data(), cdiag.data(), force.data(), dx.data(),
                   p->cg_tol, p->cg_max_iter, nullptr, &pr);
        out->newton_steps += 1;
        out->pcg_iter_total += pr.iter;
        out->pcg_iter_max = std::max(out->pcg_iter_max, pr.iter);
        if (pr.status != 0 && pr.status != 1) {
            out->status = 1;
            break;
        }
        float max_dx = 0.0f;
        for (uint32_t i = 0; i < n; ++i) {
            const float a = dx[3 * i], b = dx[3 * i + 1], c = dx[3 * i + 2];
            max_dx = std::max(max_dx, std::sqrt((a * a + b * b) + c * c));
        }
        const float recale = std::fmin(1.0f, p->max_dx / max_dx);
        for (size_t k = 0; k < n3; ++k) {
            target[k] = ex[k];
            ex[k] = ex[k] - recale * dx[k];
        }
        float toi = 1.0f;
        if (!p->disable_contact) {
            MeshView lm = sc->mesh;
            lm.x = target.data();
            LsParams lp{ex.data(), sc->shell.vmass, sc->emass, sc->shell.fmass,
                        p->line_search_max_t, p->ccd_eps};
            int overlap = 0;
            float t = oracle_line_search(&lm, sc->collision.nv ? &sc->collision : nullptr,
                                         &lp, &overlap);
            // the analytic colliders' entry times join the minimum (in max_t
            // units); pins driven into a static collider fail the step
            // (main.cu:1240-1271)
            if (sc->nsphere || sc->nfloor) {
                float ta = p->line_search_max_t;
                uint32_t pin_bad = 0xffffffffu;
                oracle_analytic_toi(n, target.data(), ex.data(), sc->mesh.vflag,
                                    sc->shell.vmass, sc->spheres, sc->nsphere, sc->floors,
                                    sc->nfloor, p->line_search_max_t, &ta, &pin_bad);
                if (pin_bad != 0xffffffffu) {
                    out->status = 6;
                    break;
                }
                t = std::fmin(t, ta / p->line_search_max_t);
            }
            toi = std::fmin(toi, t);
            if (overlap) {
                out->status = 3;
                break;
            }
        }
        if (sc->shell.f_strainlimit) {
            // the strain-limit line search joins the minimum (main.cu:1321-1331)
            const float t = oracle_strainlimit_toi(
                sc->shell.nf, sc->shell.face, sc->shell.inv_rest, sc->shell.fflag,
                sc->shell.f_strainlimit, sc->shell.f_shrink, target.data(), ex.data(),
                p->line_search_max_t);
            toi = std::fmin(toi, t / p->line_search_max_t);
        }
        g_trace.insert(g_trace.end(), {(float)pr.iter, recale, toi, max_dx});
        if (toi <= FLT_EPSILON) {
            out->status = 2;
            break;
        }
        if (!final_step) {
            toi_advanced += std::max(0.0, 1.0 - toi_advanced) * (double)(recale * toi);
        }
        for (size_t k = 0; k < n3; ++k) ex[k] = target[k] + toi * (ex[k] - target[k]);
        if (final_step) {
            break;
        } else if (toi_advanced >= p->target_toi && step >= p->min_newton_steps) {
            final_step = true;
        } else {
            ++step;
            compute_target(dt);
        }
    }
    out->toi_advanced = (float)toi_advanced;
    out->dt = dt;
    if (out->status == 0 && p->check_intersection) {
        MeshView m = sc->mesh;
        m.x = ex.data();
        out->intersections = oracle_intersection_check(
            &m, sc->collision.nv ? &sc->collision : nullptr, sc->shell.vmass, sc->emass,
            sc->shell.fmass, n, nullptr, 0);
        if (out->intersections) out->status = 4;
    }
    // committed also when the check finds intersections (main.cu:1515-1629)
    if (out->status == 0 || out->status == 4) {
        std::memcpy(sc->prev, curr, n3 * sizeof(float));
        std::memcpy(curr, ex.data(), n3 * sizeof(float));
        // rest-shape creep at the committed positions with the requested dt
        // (main.cu:1615-1624)
        const ShellView &s = sc->shell;
        oracle_plasticity(s.nf, s.nh, s.face, s.hinge, s.hflag, curr, s.f_plasticity,
                          s.f_plastic_threshold, s.h_plasticity, s.h_plastic_threshold, p->dt,
                          const_cast<float *>(s.inv_rest), const_cast<float *>(s.h_rest));
    }
    return 0;
}
}
