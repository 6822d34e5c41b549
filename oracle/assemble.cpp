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
#include "oracle_math.hpp"
#include <cmath>
#include <cstdint>
#include <map>
#include <utility>
#include <vector>

namespace {

using orc::V3;
using orc::v3;

void pe_coeff(const V3 &p, const V3 &e0, const V3 &e1, float w[2]) {
    V3 r = orc::sub(e1, e0);
    float d = orc::dot(r, r);
    if (d > 0.0f) {
        float t = orc::dot(r, orc::sub(p, e0)) / d;
        w[0] = 1.0f - t;
        w[1] = t;
    } else {
        w[0] = w[1] = 1.0f / 2.0f;
    }
}

void pt_coeff(const V3 &p, const V3 &t0, const V3 &t1, const V3 &t2, float w[3]) {
    V3 r0 = orc::sub(t1, t0), r1 = orc::sub(t2, t0), d = orc::sub(p, t0);
    V3 n = orc::cross(r0, r1);
    float nn = orc::dot(n, n);
    if (nn > 0.0f) {
        float inv = 1.0f / nn;
        float u = orc::dot(orc::cross(d, r1), n) * inv;
        float v = orc::dot(orc::cross(r0, d), n) * inv;
        w[0] = 1.0f - u - v;
        w[1] = u;
        w[2] = v;
        return;
    }
    V3 e1 = orc::sub(t2, t1), e2 = orc::sub(t0, t2);
    float l0 = orc::dot(r0, r0), l1 = orc::dot(e1, e1), l2 = orc::dot(e2, e2);
    float c[2];
    if (l0 >= l1 && l0 >= l2) {
        pe_coeff(p, t0, t1, c);
        w[0] = c[0], w[1] = c[1], w[2] = 0.0f;
    } else if (l1 >= l2) {
        pe_coeff(p, t1, t2, c);
        w[0] = 0.0f, w[1] = c[0], w[2] = c[1];
    } else {
        pe_coeff(p, t2, t0, c);
        w[0] = c[1], w[1] = 0.0f, w[2] = c[0];
    }
}

float clamp_unit(float v) { return v > 0.0f ? (v < 1.0f ? v : 1.0f) : 0.0f; }

void ee_coeff(const V3 &a0, const V3 &a1, const V3 &b0, const V3 &b1, float w[4]) {
    V3 r0 = orc::sub(a1, a0), r1 = orc::sub(b1, b0), d = orc::sub(a0, b0);
    float a = orc::dot(r0, r0), e = orc::dot(r1, r1), b = orc::dot(r0, r1);
    float c = orc::dot(r0, d), f = orc::dot(r1, d);
    float inv_a = a > 0.0f ? 1.0f / a : 0.0f;
    float inv_e = e > 0.0f ? 1.0f / e : 0.0f;
    V3 n = orc::cross(r0, r1);
    float nn = orc::dot(n, n);
    float inv_nn = nn > 0.0f ? 1.0f / nn : 0.0f;
    float s_stat = clamp_unit(orc::dot(n, orc::cross(r1, d)) * inv_nn);
    float t_stat = clamp_unit((b * s_stat + f) * inv_e);
    const float cs[5] = {s_stat, 0.0f, 1.0f, clamp_unit(-c * inv_a),
                         clamp_unit((b - c) * inv_a)};
    const float ct[5] = {t_stat, clamp_unit(f * inv_e), clamp_unit((f + b) * inv_e),
                         0.0f, 1.0f};
    float bs = 0.0f, bt = 0.0f, best = 3.40282347e38f;
    for (int i = 0; i < 5; ++i) {
        V3 v = orc::sub(orc::add(d, orc::scale(cs[i], r0)), orc::scale(ct[i], r1));
        float dist = orc::dot(v, v);
        if (dist < best) {
            best = dist;
            bs = cs[i];
            bt = ct[i];
        }
    }
    w[0] = 1.0f - bs;
    w[1] = bs;
    w[2] = 1.0f - bt;
    w[3] = bt;
}

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
struct oracle_contact_params {
    const float *x0;     // vertex.curr, 3 nv (the step's start positions)
    const float *mass;   // per vertex
    const float *v_fric, *e_fric, *f_fric;
    int friction_mode;   // 0 min, 1 max, 2 mean
    float friction_eps;
    const uint32_t *el_rp, *el_col; // elastic-only upper block CSR (may be null)
    const float *el_val;
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
            // Friction (friction.hpp:19-81).
            const float contact = -orc::dot(nrm, f);
            const float ndx = orc::dot(nrm, dx);
            const V3 u = orc::sub(dx, v3(nrm[0] * ndx, nrm[1] * ndx, nrm[2] * ndx));
            const float un = std::sqrt(orc::sqnorm(u));
            const bool kinetic = un > cp->friction_eps;
            const V3 sd = kinetic ? v3(u[0] / un, u[1] / un, u[2] / un) : v3(0, 0, 0);
            const float lam = mu > 0.0f ? mu * contact / std::fmax(cp->friction_eps, un)
                                        : 0.0f;
            f = orc::add(f, orc::scale(lam, u));
            float fh[9];
            if (kinetic) {
                const V3 wd = orc::cross(nrm, sd);
                for (int c = 0; c < 3; ++c) {
                    for (int r = 0; r < 3; ++r) {
                        fh[r + 3 * c] = lam * (wd[r] * wd[c]);
                    }
                }
            } else {
                for (int c = 0; c < 3; ++c) {
                    for (int r = 0; r < 3; ++r) {
                        const float id = r == c ? 1.0f : 0.0f;
                        fh[r + 3 * c] = lam * (id - nrm[r] * nrm[c]);
                    }
                }
            }
            for (int k = 0; k < 9; ++k) {
                H[k] = H[k] + fh[k];
            }
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
