// assemble.cpp -- CPU oracle: contact barrier + friction force and 3x3-block
// Hessian assembly for a list of accepted proximity pairs.
//
// TEST INFRASTRUCTURE ONLY (the checker of the HIP path, never shipped).
//
// Restates in fp32 (-ffp-contract=off):
//   * Proximity<N> of each kind (index tuple and weights summing to zero)
//                                            contact/contact.cu:432-870
//   * embed_contact_force_hess<N>            contact/contact.cu:262-354
//       anchored ex / ex0, normal, stiffness, barrier force and Hessian,
//       friction, extension to 3N with w_i w_j, multiplicity `count`
//   * barrier::compute_stiffness<N>          barrier/barrier.cu:49-86
//       (elastic Hessian blocks looked up in the elastic-only upper block CSR
//        as FixedCSRMat::operator(), csrmat/csrmat.cu:511-542)
//   * compute_edge_gradient / _hessian       barrier/barrier.cu:103-114
//   * cubic::gradient / curvature            barrier/cubic.hpp:22-40
//   * Friction                               energy/model/friction.hpp:19-81
//   * combine_friction                       contact/contact.cu:64-74
// Accumulation order (the oracle's own, matched by the HIP kernels): blocks
// and force rows summed in pair order, then in the (ii, jj) / ii loop order of
// atomic_embed_hessian / atomic_embed_force (contact.cu:175-190,
// utility/utility.hpp:63-72), starting from zero.
#include "oracle_distance.hpp"
#include "oracle_math.hpp"
#include <cmath>
#include <cstdint>
#include <map>
#include <utility>
#include <vector>

namespace {

using orc::V3;
using orc::v3;
using orc::ee_coeff;
using orc::pe_coeff;
using orc::pt_coeff;

// Elastic block (i, j) of an upper block CSR, transposed when i > j.
void lookup(uint32_t i, uint32_t j, const uint32_t *rp, const uint32_t *col,
            const float *val, float out[9]) {
    for (int e = 0; e < 9; ++e) {
        out[e] = 0.0f;
    }
    if (!rp) {
        return;
    }
    bool tr = false;
    if (i > j) {
        std::swap(i, j);
        tr = true;
    }
    for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
        if (col[k] == j) {
            for (int e = 0; e < 9; ++e) {
                out[e] = 0.0f + val[9 * (size_t)k + e];
            }
            break;
        } else if (col[k] > j) {
            break;
        }
    }
    if (tr) {
        std::swap(out[1], out[3]);
        std::swap(out[2], out[6]);
        std::swap(out[5], out[7]);
    }
}

float cubic_gradient(float g, float ghat, float offset) {
    g -= offset;
    float y = g - ghat;
    return y < 0.0f ? -2.0f * y * y / ghat : 0.0f;
}

float cubic_curvature(float g, float ghat, float offset) {
    g -= offset;
    return (g - ghat < 0.0f) ? 4.0f * (1.0f - g / ghat) : 0.0f;
}

float combine(float a, float b, int mode) {
    if (mode == 1) {
        return std::fmax(a, b);
    }
    if (mode == 2) {
        return 0.5f * (a + b);
    }
    return std::fmin(a, b);
}


// Friction (energy/model/friction.hpp:19-81) added to f and H.
void add_friction(const V3 &nrm, const V3 &dx, float mu, float eps, V3 &f, float H[9]) {
    const float contact = -orc::dot(nrm, f);
    const float ndx = orc::dot(nrm, dx);
    const V3 u = orc::sub(dx, v3(nrm[0] * ndx, nrm[1] * ndx, nrm[2] * ndx));
    const float un = std::sqrt(orc::sqnorm(u));
    const bool kinetic = un > eps;
    const V3 sd = kinetic ? v3(u[0] / un, u[1] / un, u[2] / un) : v3(0, 0, 0);
    const float lam = mu > 0.0f ? mu * contact / std::fmax(eps, un) : 0.0f;
    f = orc::add(f, orc::scale(lam, u));
    float fh[9];
    if (kinetic) {
        const V3 wd = orc::cross(nrm, sd);
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) fh[r + 3 * c] = lam * (wd[r] * wd[c]);
    } else {
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) {
                const float id = r == c ? 1.0f : 0.0f;
                fh[r + 3 * c] = lam * (id - nrm[r] * nrm[c]);
            }
    }
    for (int k = 0; k < 9; ++k) {
        H[k] = H[k] + fh[k];
    }
}

// push::gradient / push::hessian scaled by k (energy/model/push.hpp:26-46).
void push_terms(const V3 &e, const V3 &n, float eps, float k, V3 &f, float H[9]) {
    const float d = orc::dot(e, n);
    V3 g = v3(0, 0, 0);
    float cv = 0.0f;
    if (d < 0.0f) {
        g = orc::scale(-(d * d) / eps, n);
        cv = -2.0f * d / eps;
    }
    f = orc::scale(k, g);
    const V3 cn = orc::scale(cv, n);
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) H[r + 3 * c] = k * (cn[r] * n[c]);
}

// (K ne).ne / |ne|^2 over N vertices, K = elastic blocks + mass / gap^2 on
// the diagonal blocks, ne = c_i n (contact.cu:1010-1028, 1092-1106).
float ext_stiffness(int N, const uint32_t *idx, const float *c, const V3 &n, float gap2,
                    const float *mass, const uint32_t *rp, const uint32_t *col,
                    const float *val) {
    float ne[12], Kn[12];
    for (int j = 0; j < 3 * N; ++j) {
        ne[j] = c[j / 3] * n[j % 3];
        Kn[j] = 0.0f;
    }
    // row-major 3N x 3N product accumulated over columns in order
    for (int col_ = 0; col_ < 3 * N; ++col_) {
        const int jj = col_ / 3, cc = col_ % 3;
        for (int row = 0; row < 3 * N; ++row) {
            const int ii = row / 3, r = row % 3;
            float blk[9];
            lookup(idx[ii], idx[jj], rp, col, val, blk);
            float v = blk[r + 3 * cc];
            if (ii == jj) {
                v = v + (r == cc ? mass[idx[ii]] / gap2 : 0.0f);
            } else {
                v = v + 0.0f;
            }
            Kn[row] = Kn[row] + v * ne[col_];
        }
    }
    float num = 0.0f, den = 0.0f;
    for (int j = 0; j < 3 * N; ++j) num = num + Kn[j] * ne[j];
    for (int j = 0; j < 3 * N; ++j) den = den + ne[j] * ne[j];
    return num / den;
}
} // namespace

extern "C" {

struct oracle_contact_mesh; // narrow.cpp
struct oracle_mesh_view {   // the same layout as oracle_contact_mesh
    uint32_t nv, ne, nf;
    const float *x;
    const uint32_t *edge, *face;
    const uint32_t *v2e_off, *v2e, *v2f_off, *v2f, *e2f_off, *e2f;
    const float *v_ghat, *v_off, *e_ghat, *e_off, *f_ghat, *f_off;
    const uint32_t *vflag, *vbody, *eflag, *fflag;
};

struct oracle_pair_in {
    uint32_t kind, a, b, count;
};

// Contact material and solver constants for the assembly.
// Collision mesh (the same layout as oracle_static_mesh of ccd.cpp, with the
// per-element friction appended).
struct oracle_static_view {
    uint32_t nv, ne, nf;
    const float *x;
    const uint32_t *edge, *face;
    const float *v_ghat, *v_off, *e_ghat, *e_off, *f_ghat, *f_off;
    const float *v_fric, *e_fric, *f_fric;
};

struct oracle_contact_params {
    const float *x0;     // vertex.curr, 3 nv (the step's start positions)
    const float *mass;   // per vertex
    const float *v_fric, *e_fric, *f_fric;
    int friction_mode;   // 0 min, 1 max, 2 mean
    float friction_eps;
    const uint32_t *el_rp, *el_col; // elastic-only upper block CSR (may be null)
    const float *el_val;
    const oracle_static_view *sm;   // collision mesh of kind 4-6 pairs (may be null)
};

// Assemble the contact gradient (grad, 3 nv, written) and the upper-triangle
// contact blocks, returned sorted by (row, col) with duplicates summed. Returns
// the block count; writes at most `cap` blocks.
uint64_t oracle_contact_assemble(const oracle_mesh_view *m,
                                 const oracle_contact_params *cp,
                                 const oracle_pair_in *pairs, uint64_t npairs,
                                 float *grad, uint32_t *rows, uint32_t *cols,
                                 float *vals, uint64_t cap) {
    const uint32_t nv = m->nv;
    auto X = [&](const float *a, uint32_t v) {
        return v3(a[3 * (size_t)v], a[3 * (size_t)v + 1], a[3 * (size_t)v + 2]);
    };
    for (size_t k = 0; k < 3 * (size_t)nv; ++k) {
        grad[k] = 0.0f;
    }
    std::map<std::pair<uint32_t, uint32_t>, std::vector<float>> blocks;
    for (uint64_t pi = 0; pi < npairs; ++pi) {
        const uint32_t kind = pairs[pi].kind & 0xffu;
        const bool no_fric = (pairs[pi].kind >> 8) & 1u;
        const uint32_t a = pairs[pi].a, b = pairs[pi].b;
        uint32_t idx[4];
        float w[4];
        int N;
        if (kind >= 4) {
            const oracle_static_view &sm = *cp->sm;
            auto SX = [&](uint32_t v) { return X(sm.x, v); };
            V3 f;
            float H[9];
            if (kind == 4) {
                const V3 p = X(m->x, a);
                const V3 q = orc::sub(X(cp->x0, a), p);
                const uint32_t *fc = sm.face + 3 * (size_t)b;
                const V3 t0 = orc::sub(SX(fc[0]), p), t1 = orc::sub(SX(fc[1]), p),
                         t2 = orc::sub(SX(fc[2]), p);
                float c[3];
                orc::pt_coeff_unclassified(v3(0, 0, 0), t0, t1, t2, c);
                const V3 y = orc::add(orc::add(orc::scale(c[0], t0), orc::scale(c[1], t1)),
                                      orc::scale(c[2], t2));
                const V3 e = v3(-y[0], -y[1], -y[2]);
                const float off = m->v_off[a] + sm.f_off[b];
                const float gh = 0.5f * (m->v_ghat[a] + sm.f_ghat[b]);
                const float mu = combine(cp->v_fric[a], sm.f_fric[b], cp->friction_mode);
                const V3 n = orc::normalized(e);
                const V3 proj = orc::add(y, orc::scale(off + gh, n));
                const float g = std::sqrt(orc::sqnorm(e)) - off;
                const float gap2 = g * g;
                float blk[9];
                lookup(a, a, cp->el_rp, cp->el_col, cp->el_val, blk);
                const V3 Hn = v3((blk[0] * n[0] + blk[3] * n[1]) + blk[6] * n[2],
                                 (blk[1] * n[0] + blk[4] * n[1]) + blk[7] * n[2],
                                 (blk[2] * n[0] + blk[5] * n[1]) + blk[8] * n[2]);
                const float k = orc::dot(n, Hn) + cp->mass[a] / gap2;
                push_terms(v3(-proj[0], -proj[1], -proj[2]), n, gh, k, f, H);
                add_friction(n, v3(-q[0], -q[1], -q[2]), mu, cp->friction_eps, f, H);
                N = 1;
                idx[0] = a;
                w[0] = 1.0f;
            } else if (kind == 5) {
                const V3 yv = SX(a);
                const uint32_t *fc = m->face + 3 * (size_t)b;
                const V3 t0 = orc::sub(X(m->x, fc[0]), yv), t1 = orc::sub(X(m->x, fc[1]), yv),
                         t2 = orc::sub(X(m->x, fc[2]), yv);
                const V3 s0 = orc::sub(X(cp->x0, fc[0]), yv),
                         s1 = orc::sub(X(cp->x0, fc[1]), yv),
                         s2 = orc::sub(X(cp->x0, fc[2]), yv);
                float c[3];
                orc::pt_coeff_unclassified(v3(0, 0, 0), t0, t1, t2, c);
                const V3 p = orc::add(orc::add(orc::scale(c[0], t0), orc::scale(c[1], t1)),
                                      orc::scale(c[2], t2));
                const V3 q = orc::add(orc::add(orc::scale(c[0], s0), orc::scale(c[1], s1)),
                                      orc::scale(c[2], s2));
                const float off = m->f_off[b] + sm.v_off[a];
                const float gh = 0.5f * (m->f_ghat[b] + sm.v_ghat[a]);
                const float mu = combine(cp->f_fric[b], sm.v_fric[a], cp->friction_mode);
                const V3 n = orc::normalized(p);
                const float g = std::sqrt(orc::sqnorm(p)) - off;
                const float gap2 = g * g;
                const uint32_t fi[3] = {fc[0], fc[1], fc[2]};
                const float k = ext_stiffness(3, fi, c, n, gap2, cp->mass, cp->el_rp,
                                              cp->el_col, cp->el_val);
                const V3 proj = orc::scale(off + gh, n);
                push_terms(orc::sub(p, proj), n, gh, k, f, H);
                add_friction(n, orc::sub(p, q), mu, cp->friction_eps, f, H);
                N = 3;
                idx[0] = fc[0], idx[1] = fc[1], idx[2] = fc[2];
                w[0] = c[0], w[1] = c[1], w[2] = c[2];
            } else {
                const uint32_t *ea = m->edge + 2 * (size_t)a;
                const uint32_t *eb = sm.edge + 2 * (size_t)b;
                const V3 p0 = X(m->x, ea[0]), p1 = X(m->x, ea[1]);
                const V3 q0 = SX(eb[0]), q1 = SX(eb[1]);
                const V3 b0 = X(cp->x0, ea[0]), b1 = X(cp->x0, ea[1]);
                float c[4];
                ee_coeff(p0, p1, q0, q1, c);
                const V3 x = orc::add(orc::scale(c[0], p0), orc::scale(c[1], p1));
                const V3 y = orc::add(orc::scale(c[2], q0), orc::scale(c[3], q1));
                const V3 z = orc::add(orc::scale(c[0], b0), orc::scale(c[1], b1));
                const V3 e = orc::sub(x, y);
                const float off = m->e_off[a] + sm.e_off[b];
                const float gh = 0.5f * (m->e_ghat[a] + sm.e_ghat[b]);
                const float mu = combine(cp->e_fric[a], sm.e_fric[b], cp->friction_mode);
                const V3 n = orc::normalized(e);
                const V3 proj = orc::scale(off + gh, n);
                const float g = std::sqrt(orc::sqnorm(e)) - off;
                const float gap2 = g * g;
                const uint32_t ei[2] = {ea[0], ea[1]};
                const float k = ext_stiffness(2, ei, c, n, gap2, cp->mass, cp->el_rp,
                                              cp->el_col, cp->el_val);
                push_terms(orc::sub(e, proj), n, gh, k, f, H);
                add_friction(n, orc::sub(x, z), mu, cp->friction_eps, f, H);
                N = 2;
                idx[0] = ea[0], idx[1] = ea[1];
                w[0] = c[0], w[1] = c[1];
            }
            const float cnt = (float)pairs[pi].count;
            for (int ii = 0; ii < N; ++ii)
                for (int cc = 0; cc < 3; ++cc) grad[3 * (size_t)idx[ii] + cc] += cnt * (w[ii] * f[cc]);
            for (int ii = 0; ii < N; ++ii)
                for (int jj = 0; jj < N; ++jj) {
                    const uint32_t i = idx[ii], j = idx[jj];
                    if (i <= j) {
                        std::vector<float> &acc = blocks[{i, j}];
                        if (acc.empty()) acc.assign(9, 0.0f);
                        const float s = w[ii] * w[jj];
                        for (int k = 0; k < 9; ++k) acc[k] = acc[k] + cnt * (s * H[k]);
                    }
                }
            continue;
        }
        float gh, off, mu;
        if (kind == 0) {
            const uint32_t *f = m->face + 3 * (size_t)b;
            float c[3];
            pt_coeff(X(m->x, a), X(m->x, f[0]), X(m->x, f[1]), X(m->x, f[2]), c);
            N = 4;
            idx[0] = a, idx[1] = f[0], idx[2] = f[1], idx[3] = f[2];
            w[0] = 1.0f, w[1] = -c[0], w[2] = -c[1], w[3] = -c[2];
            off = m->v_off[a] + m->f_off[b];
            gh = 0.5f * (m->v_ghat[a] + m->f_ghat[b]);
            mu = combine(cp->v_fric[a], cp->f_fric[b], cp->friction_mode);
        } else if (kind == 1) {
            const uint32_t *e = m->edge + 2 * (size_t)b;
            float c[2];
            pe_coeff(X(m->x, a), X(m->x, e[0]), X(m->x, e[1]), c);
            N = 3;
            idx[0] = a, idx[1] = e[0], idx[2] = e[1];
            w[0] = 1.0f, w[1] = -c[0], w[2] = -c[1];
            off = m->v_off[a] + m->e_off[b];
            gh = 0.5f * (m->v_ghat[a] + m->e_ghat[b]);
            mu = combine(cp->v_fric[a], cp->e_fric[b], cp->friction_mode);
        } else if (kind == 2) {
            N = 2;
            idx[0] = a, idx[1] = b;
            w[0] = 1.0f, w[1] = -1.0f;
            off = m->v_off[a] + m->v_off[b];
            gh = 0.5f * (m->v_ghat[a] + m->v_ghat[b]);
            mu = combine(cp->v_fric[a], cp->v_fric[b], cp->friction_mode);
        } else {
            const uint32_t *ea = m->edge + 2 * (size_t)a, *eb = m->edge + 2 * (size_t)b;
            float c[4];
            ee_coeff(X(m->x, ea[0]), X(m->x, ea[1]), X(m->x, eb[0]), X(m->x, eb[1]), c);
            N = 4;
            idx[0] = ea[0], idx[1] = ea[1], idx[2] = eb[0], idx[3] = eb[1];
            w[0] = c[0], w[1] = c[1], w[2] = -c[2], w[3] = -c[3];
            off = m->e_off[a] + m->e_off[b];
            gh = 0.5f * (m->e_ghat[a] + m->e_ghat[b]);
            mu = combine(cp->e_fric[a], cp->e_fric[b], cp->friction_mode);
        }
        // Anchored separations (contact.cu:286-307).
        V3 ex0 = v3(0, 0, 0), ex = v3(0, 0, 0);
        const V3 an0 = X(cp->x0, idx[0]), an = X(m->x, idx[0]);
        float mass[4];
        for (int ii = 0; ii < N; ++ii) {
            mass[ii] = cp->mass[idx[ii]];
            ex0 = orc::add(ex0, orc::scale(w[ii], orc::sub(X(cp->x0, idx[ii]), an0)));
            ex = orc::add(ex, orc::scale(w[ii], orc::sub(X(m->x, idx[ii]), an)));
        }
        const V3 dx = orc::sub(ex, ex0);
        const V3 nrm = orc::normalized(ex);
        // Stiffness (barrier.cu:49-86).
        const float norm = std::sqrt(orc::sqnorm(ex));
        const float g = norm - off;
        const float sqr_x = g * g;
        float m_max = 0.0f;
        for (int ii = 0; ii < N; ++ii) {
            if (mass[ii] > m_max) {
                m_max = mass[ii];
            }
        }
        float wv[12];
        for (int ii = 0; ii < N; ++ii) {
            for (int c = 0; c < 3; ++c) {
                wv[3 * ii + c] = w[ii] * ex[c];
            }
        }
        {
            float sq = 0.0f;
            for (int k = 0; k < 3 * N; ++k) {
                sq = sq + wv[k] * wv[k];
            }
            float ln = std::sqrt(sq);
            if (ln > 0.0f) {
                for (int k = 0; k < 3 * N; ++k) {
                    wv[k] = wv[k] / ln;
                }
            }
        }
        float stiff = 0.0f;
        for (int ii = 0; ii < N; ++ii) {
            const float m_ii = mass[ii] > 0.0f ? mass[ii] : m_max;
            float hv[3] = {0.0f, 0.0f, 0.0f}; // rows 3ii..3ii+2 of (local_hess w)
            for (int jj = 0; jj < N; ++jj) {
                float blk[9];
                lookup(idx[ii], idx[jj], cp->el_rp, cp->el_col, cp->el_val, blk);
                if (ii == jj) {
                    const float s = m_ii / sqr_x;
                    blk[0] = blk[0] + s;
                    blk[4] = blk[4] + s;
                    blk[8] = blk[8] + s;
                }
                for (int r = 0; r < 3; ++r) {
                    for (int c = 0; c < 3; ++c) {
                        hv[r] = hv[r] + blk[r + 3 * c] * wv[3 * jj + c];
                    }
                }
            }
            for (int r = 0; r < 3; ++r) {
                stiff = stiff + hv[r] * wv[3 * ii + r];
            }
        }
        // Barrier force and Hessian (barrier.cu:103-114).
        const float gl = std::sqrt(orc::sqnorm(ex));
        const V3 nn = v3(ex[0] / gl, ex[1] / gl, ex[2] / gl);
        const float gg = cubic_gradient(gl, gh, off);
        V3 f = orc::scale(stiff, v3(nn[0] * gg, nn[1] * gg, nn[2] * gg));
        const float cv = cubic_curvature(std::sqrt(orc::sqnorm(ex)), gh, off);
        const float e2 = orc::sqnorm(ex);
        float H[9];
        for (int c = 0; c < 3; ++c) {
            for (int r = 0; r < 3; ++r) {
                H[r + 3 * c] = stiff * ((cv * ex[r]) * ex[c] / e2);
            }
        }
        if (!no_fric) {
            add_friction(nrm, dx, mu, cp->friction_eps, f, H);
        }
        const float cnt = (float)pairs[pi].count;
        for (int ii = 0; ii < N; ++ii) {
            for (int c = 0; c < 3; ++c) {
                grad[3 * (size_t)idx[ii] + c] += cnt * (w[ii] * f[c]);
            }
        }
        for (int ii = 0; ii < N; ++ii) {
            for (int jj = 0; jj < N; ++jj) {
                const uint32_t i = idx[ii], j = idx[jj];
                if (i <= j) {
                    std::vector<float> &acc = blocks[{i, j}];
                    if (acc.empty()) {
                        acc.assign(9, 0.0f);
                    }
                    const float s = w[ii] * w[jj];
                    for (int k = 0; k < 9; ++k) {
                        acc[k] = acc[k] + cnt * (s * H[k]);
                    }
                }
            }
        }
    }
    uint64_t nb = 0;
    for (const auto &kv : blocks) {
        if (nb < cap) {
            rows[nb] = kv.first.first;
            cols[nb] = kv.first.second;
            for (int k = 0; k < 9; ++k) {
                vals[9 * nb + k] = kv.second[k];
            }
        }
        ++nb;
    }
    return nb;
}
}

// ---- analytic colliders ----------------------------------------------------
//
// Per vertex against analytic spheres and floors (data.constraint.sphere /
// floor; contact/contact.cu:1154-1382, embed_vertex_constraint_force_hessian):
// push barrier with stiffness n.(K_ii n) + m / gap^2 (K_ii the vertex's
// elastic diagonal block) plus friction for static colliders, the mass-only
// stiffness m / ghat^2 and no friction for kinematic spheres. Pinned and
// massless vertices are skipped (Dirichlet rows; static solids). The terms of
// each collider are added to the vertex's force row and diagonal block in
// collider order (push, then friction), onto what is already there.

namespace {

// The friction terms alone (energy/model/friction.hpp:19-81) for the push
// force fp: gradient g, Hessian Hf.
void friction_terms(const V3 &nrm, const V3 &dx, float mu, float eps, const V3 &fp, V3 &g,
                    float Hf[9]) {
    const float contact = -orc::dot(nrm, fp);
    const float ndx = orc::dot(nrm, dx);
    const V3 u = orc::sub(dx, v3(nrm[0] * ndx, nrm[1] * ndx, nrm[2] * ndx));
    const float un = std::sqrt(orc::sqnorm(u));
    const bool kinetic = un > eps;
    const V3 sd = kinetic ? v3(u[0] / un, u[1] / un, u[2] / un) : v3(0, 0, 0);
    const float lam = mu > 0.0f ? mu * contact / std::fmax(eps, un) : 0.0f;
    g = orc::scale(lam, u);
    if (kinetic) {
        const V3 wd = orc::cross(nrm, sd);
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) Hf[r + 3 * c] = lam * (wd[r] * wd[c]);
    } else {
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) {
                const float id = r == c ? 1.0f : 0.0f;
                Hf[r + 3 * c] = lam * (id - nrm[r] * nrm[c]);
            }
    }
}

} // namespace

extern "C" {

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

// Adds every vertex's analytic terms to grad (3 nv) and cdiag (9 nv). Returns
// the number of (vertex, collider) contacts.
uint32_t oracle_analytic_assemble(const oracle_mesh_view *m, const oracle_contact_params *cp,
                                  const oracle_sphere *sp, uint32_t nsp,
                                  const oracle_floor *fl, uint32_t nfl,
                                  float constraint_tol, float *grad, float *cdiag) {
    uint32_t contacts = 0;
    for (uint32_t i = 0; i < m->nv; ++i) {
        const float mass = cp->mass[i];
        if ((m->vflag[i] & 1u) || !(mass > 0.0f)) {
            continue;
        }
        const V3 x = v3(m->x[3 * i], m->x[3 * i + 1], m->x[3 * i + 2]);
        const V3 x0 = v3(cp->x0[3 * i], cp->x0[3 * i + 1], cp->x0[3 * i + 2]);
        const V3 dx = orc::sub(x, x0);
        float lh[9];
        lookup(i, i, cp->el_rp, cp->el_col, cp->el_val, lh);
        V3 F = v3(grad[3 * i], grad[3 * i + 1], grad[3 * i + 2]);
        float *H = cdiag + 9 * (size_t)i;
        for (uint32_t j = 0; j < nsp; ++j) {
            const oracle_sphere &s = sp[j];
            const float ghat = s.ghat;
            const float mu = combine(s.friction, cp->v_fric[i], cp->friction_mode);
            bool reverse = s.reverse;
            float radius = s.radius;
            V3 center = v3(s.center[0], s.center[1], s.center[2]);
            if (s.bowl && x[1] > center[1]) {
                center = v3(center[0], x[1], center[2]);
            }
            const float d2 = orc::sqnorm(orc::sub(x, center));
            if (s.kinematic) {
                if (d2 == 0.0f) {
                    continue;
                }
                const V3 normal = orc::normalized(orc::sub(x, center));
                const float eff = reverse ? (radius - ghat) : (radius + ghat);
                const V3 target = orc::scale(eff, normal);
                const V3 o = orc::sub(x, center);
                if (s.bowl) {
                    reverse = true;
                }
                const float k = mass / (ghat * ghat);
                const float r2 = eff * eff;
                const float dist = std::sqrt(d2);
                const float depth = reverse ? (dist - s.radius) : (s.radius - dist);
                if (s.thickness > 0.0f && depth > s.thickness) {
                    continue;
                }
                const bool hit = reverse ? d2 > r2 : d2 < r2;
                if (hit) {
                    ++contacts;
                    const V3 nn = reverse ? v3(-normal[0], -normal[1], -normal[2]) : normal;
                    V3 fp;
                    float Hp[9];
                    push_terms(orc::sub(o, target), nn, ghat, k, fp, Hp);
                    F = orc::add(F, fp);
                    for (int e = 0; e < 9; ++e) H[e] = H[e] + Hp[e];
                }
                continue;
            }
            radius = reverse ? radius - ghat : radius + ghat;
            const float r2 = radius * radius;
            const bool hit = reverse ? d2 > r2 : d2 < r2;
            if (!hit) {
                continue;
            }
            const float dist = std::sqrt(d2);
            const float depth = reverse ? (dist - s.radius) : (s.radius - dist);
            if (s.thickness > 0.0f && depth > s.thickness) {
                continue;
            }
            ++contacts;
            V3 normal = orc::normalized(orc::sub(x, center));
            const V3 proj = orc::scale(radius, normal);
            const V3 o = orc::sub(x, center);
            if (reverse) {
                normal = v3(-normal[0], -normal[1], -normal[2]);
            }
            const float gap = reverse ? s.radius - std::sqrt(d2) : std::sqrt(d2) - s.radius;
            const float k = orc::dot(normal, orc::matvec(lh, normal.x)) + mass / (gap * gap);
            V3 fp, fg;
            float Hp[9], Hf[9];
            push_terms(orc::sub(o, proj), normal, ghat, k, fp, Hp);
            F = orc::add(F, fp);
            for (int e = 0; e < 9; ++e) H[e] = H[e] + Hp[e];
            friction_terms(normal, dx, mu, cp->friction_eps, fp, fg, Hf);
            F = orc::add(F, fg);
            for (int e = 0; e < 9; ++e) H[e] = H[e] + Hf[e];
        }
        for (uint32_t j = 0; j < nfl; ++j) {
            const oracle_floor &f = fl[j];
            const float ghat = f.ghat;
            const float mu = combine(f.friction, cp->v_fric[i], cp->friction_mode);
            const V3 up = v3(f.up[0], f.up[1], f.up[2]);
            const V3 ground = orc::add(v3(f.ground[0], f.ground[1], f.ground[2]),
                                       orc::scale(ghat, up));
            const V3 e = orc::sub(x, ground);
            const float eu = orc::dot(e, up);
            if (!(eu < 0.0f)) {
                continue;
            }
            if (f.thickness > 0.0f && -eu > f.thickness) {
                continue;
            }
            ++contacts;
            const V3 proj = orc::scale(-eu, up);
            float gap = orc::dot(orc::sub(x, v3(f.ground[0], f.ground[1], f.ground[2])), up);
            if (f.kinematic) {
                gap = std::fmax(gap, constraint_tol * ghat);
            }
            const float k = orc::dot(up, orc::matvec(lh, up.x)) + mass / (gap * gap);
            V3 fp, fg;
            float Hp[9], Hf[9];
            push_terms(v3(-proj[0], -proj[1], -proj[2]), up, ghat, k, fp, Hp);
            F = orc::add(F, fp);
            for (int q = 0; q < 9; ++q) H[q] = H[q] + Hp[q];
            friction_terms(up, dx, mu, cp->friction_eps, fp, fg, Hf);
            F = orc::add(F, fg);
            for (int q = 0; q < 9; ++q) H[q] = H[q] + Hf[q];
        }
        grad[3 * i] = F[0], grad[3 * i + 1] = F[1], grad[3 * i + 2] = F[2];
    }
    return contacts;
}

// The analytic part of the line search (contact.cu:2318-2470): per vertex
// the fraction of the swept step x0 -> x0 + max_t (x1 - x0) at which it
// reaches a static sphere or floor, times max_t, folded into *toi (a minimum
// in [0, max_t]); a pinned vertex driven from outside to inside a static
// collider sets *pin_bad to the smallest such vertex index. Kinematic
// colliders are not swept.
void oracle_analytic_toi(uint32_t nv, const float *x0a, const float *x1a, const uint32_t *vflag,
                         const float *vmass, const oracle_sphere *sp, uint32_t nsp,
                         const oracle_floor *fl, uint32_t nfl, float max_t, float *toi,
                         uint32_t *pin_bad) {
    for (uint32_t i = 0; i < nv; ++i) {
        const V3 y0 = v3(x0a[3 * i], x0a[3 * i + 1], x0a[3 * i + 2]);
        const V3 y1 = v3(x1a[3 * i], x1a[3 * i + 1], x1a[3 * i + 2]);
        const V3 x1 = orc::add(orc::scale(max_t, orc::sub(y1, y0)), y0);
        const V3 &x0 = y0;
        auto centers = [&](const oracle_sphere &s, V3 &c0, V3 &c1) {
            const V3 c = v3(s.center[0], s.center[1], s.center[2]);
            c0 = (s.bowl && x0[1] > c[1]) ? v3(c[0], x0[1], c[2]) : c;
            c1 = (s.bowl && x1[1] > c[1]) ? v3(c[0], x1[1], c[2]) : c;
        };
        if (vflag[i] & 1u) {
            for (uint32_t j = 0; j < nsp; ++j) {
                const oracle_sphere &s = sp[j];
                if (s.kinematic) continue;
                V3 c0, c1;
                centers(s, c0, c1);
                const float r = s.radius;
                const float r0 = std::sqrt(orc::sqnorm(orc::sub(x0, c0)));
                const float r1 = std::sqrt(orc::sqnorm(orc::sub(x1, c1)));
                const float d0 = s.reverse ? (r - r0) : (r0 - r);
                const float d1 = s.reverse ? (r - r1) : (r1 - r);
                if (d0 >= 0.0f && d1 < 0.0f) *pin_bad = std::min(*pin_bad, i);
            }
            for (uint32_t j = 0; j < nfl; ++j) {
                const oracle_floor &f = fl[j];
                if (f.kinematic) continue;
                const V3 up = v3(f.up[0], f.up[1], f.up[2]);
                const V3 g = v3(f.ground[0], f.ground[1], f.ground[2]);
                const float h0 = orc::dot(up, orc::sub(x0, g));
                const float h1 = orc::dot(up, orc::sub(x1, g));
                if (h0 >= 0.0f && h1 < 0.0f) *pin_bad = std::min(*pin_bad, i);
            }
        } else if (vmass[i] > 0.0f) {
            for (uint32_t j = 0; j < nsp; ++j) {
                const oracle_sphere &s = sp[j];
                if (s.kinematic) continue;
                V3 c0, c1;
                centers(s, c0, c1);
                const float r = s.radius;
                const float r0 = std::sqrt(orc::sqnorm(orc::sub(x0, c0)));
                const float r1 = std::sqrt(orc::sqnorm(orc::sub(x1, c1)));
                const float depth0 = s.reverse ? (r0 - r) : (r - r0);
                if (s.thickness > 0.0f && depth0 > s.thickness) continue;
                if ((r0 - r) * (r1 - r) <= 0.0f) {
                    const float t = (r - r0) / (r1 - r0);
                    *toi = std::fmin(*toi, max_t * t);
                }
            }
            for (uint32_t j = 0; j < nfl; ++j) {
                const oracle_floor &f = fl[j];
                if (f.kinematic) continue;
                const V3 up = v3(f.up[0], f.up[1], f.up[2]);
                const V3 g = v3(f.ground[0], f.ground[1], f.ground[2]);
                const float h0 = orc::dot(up, orc::sub(x0, g));
                const float h1 = orc::dot(up, orc::sub(x1, g));
                if (f.thickness > 0.0f && -h0 > f.thickness) continue;
                if (h1 < 0.0f) {
                    const float t = -h0 / (h1 - h0);
                    *toi = std::fmin(*toi, max_t * t);
                }
            }
        }
    }
}

} // extern "C"
