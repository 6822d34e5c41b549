// assemble.hip -- contact barrier + friction force and 3x3-block Hessian
// assembly on MI355X.
//
// Replaces the stage-1 half of contact::embed_contact_force_hessian
// (contact/contact.cu:262-354 with barrier/barrier.cu:49-114,
// barrier/cubic.hpp:22-40, energy/model/friction.hpp:19-81). The reference
// scatters with float atomics into a fixed and a dynamic CSR; here every pair
// is evaluated once (one thread), and its contributions become records keyed
// by vertex (force) or by (row, col) (upper-triangle blocks). Stable radix
// sorts group the records, and each group is summed sequentially in record
// order (pair order, then the reference's (ii, jj) loop order), so the result
// is deterministic and independent of the scheduling. Every block of a pair is
// count * w_i * w_j * H with one symmetric 3x3 H, so a record stores only the
// pair and the two positions.
#include "assemble.hpp"
#include "ccd.hpp"
#include "distance.hpp"
#include "scan.hpp"
#include "sort.hpp"

namespace ppf {

// One evaluated pair: vertex tuple (unused slots kInvalid), weights, the
// pair's force vector and 3x3 Hessian (before the w_i / w_i w_j extension),
// and its multiplicity.
struct PairEval {
    uint32_t idx[4];
    float w[4];
    float f[3];
    float H[9];
    float cnt;
    uint32_t n;
};

namespace {

constexpr int kNT = 256;
constexpr uint32_t kInvalid = 0xffffffffu;

__device__ inline f3 X(const float *a, uint32_t v) { return load3(a + 3 * (size_t)v); }

__device__ inline float combine(float a, float b, int mode) {
    if (mode == 1) {
        return fmaxf(a, b);
    }
    if (mode == 2) {
        return 0.5f * (a + b);
    }
    return fminf(a, b);
}

// Elastic block (i, j) of the upper block CSR, transposed when i > j
// (FixedCSRMat::operator(), csrmat/csrmat.cu:511-542).
__device__ inline void lookup(uint32_t i, uint32_t j, const uint32_t *rp,
                              const uint32_t *col, const float *val, float out[9]) {
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        out[e] = 0.0f;
    }
    if (!rp) {
        return;
    }
    bool tr = false;
    if (i > j) {
        const uint32_t t = i;
        i = j;
        j = t;
        tr = true;
    }
    for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
        const uint32_t c = col[k];
        if (c == j) {
#pragma unroll
            for (int e = 0; e < 9; ++e) {
                out[e] = 0.0f + val[9 * (size_t)k + e];
            }
            break;
        } else if (c > j) {
            break;
        }
    }
    if (tr) {
        float t;
        t = out[1], out[1] = out[3], out[3] = t;
        t = out[2], out[2] = out[6], out[6] = t;
        t = out[5], out[5] = out[7], out[7] = t;
    }
}


// Friction (energy/model/friction.hpp:19-81) of the push force fp: gradient
// g and Hessian Hf alone (normal force -n.fp, tangential displacement of dx,
// kinetic above eps).
__device__ inline void friction_terms(f3 nrm, f3 dx, float mu, float eps, f3 fp, f3 &g,
                                      float Hf[9]) {
    const float contact = -dot3(nrm, fp);
    const float ndx = dot3(nrm, dx);
    const f3 u = dx - mk3(nrm.x * ndx, nrm.y * ndx, nrm.z * ndx);
    const float un = sqrtf(dot3(u, u));
    const bool kinetic = un > eps;
    const f3 sd = kinetic ? mk3(u.x / un, u.y / un, u.z / un) : mk3(0, 0, 0);
    const float lam = mu > 0.0f ? mu * contact / fmaxf(eps, un) : 0.0f;
    g = lam * u;
    const float na[3] = {nrm.x, nrm.y, nrm.z};
    if (kinetic) {
        const f3 wd = cross3(nrm, sd);
        const float wa[3] = {wd.x, wd.y, wd.z};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                Hf[r + 3 * c] = lam * (wa[r] * wa[c]);
            }
        }
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const float id = r == c ? 1.0f : 0.0f;
                Hf[r + 3 * c] = lam * (id - na[r] * na[c]);
            }
        }
    }
}

// Friction (energy/model/friction.hpp:19-81) on the pair force f and Hessian
// H: normal force -n.f, tangential displacement of dx, kinetic above eps.
__device__ inline void add_friction(f3 nrm, f3 dx, float mu, float eps, f3 &f, float H[9]) {
    const float contact = -dot3(nrm, f);
    const float ndx = dot3(nrm, dx);
    const f3 u = dx - mk3(nrm.x * ndx, nrm.y * ndx, nrm.z * ndx);
    const float un = sqrtf(dot3(u, u));
    const bool kinetic = un > eps;
    const f3 sd = kinetic ? mk3(u.x / un, u.y / un, u.z / un) : mk3(0, 0, 0);
    const float lam = mu > 0.0f ? mu * contact / fmaxf(eps, un) : 0.0f;
    f = f + lam * u;
    const float na[3] = {nrm.x, nrm.y, nrm.z};
    if (kinetic) {
        const f3 wd = cross3(nrm, sd);
        const float wa[3] = {wd.x, wd.y, wd.z};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                H[r + 3 * c] = H[r + 3 * c] + lam * (wa[r] * wa[c]);
            }
        }
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const float id = r == c ? 1.0f : 0.0f;
                H[r + 3 * c] = H[r + 3 * c] + lam * (id - na[r] * na[c]);
            }
        }
    }
}

// Push barrier of a collision-mesh pair (energy/model/push.hpp:26-46):
// f = k * (-d^2 / eps) n and H = k * ((-2 d / eps) n) n^T for d = e.n < 0.
__device__ inline void push_terms(f3 e, f3 n, float eps, float k, f3 &f, float H[9]) {
    const float d = dot3(e, n);
    f3 g = mk3(0.0f, 0.0f, 0.0f);
    float cv = 0.0f;
    if (d < 0.0f) {
        g = (-(d * d) / eps) * n;
        cv = -2.0f * d / eps;
    }
    f = k * g;
    const f3 cn = cv * n;
    const float ca[3] = {cn.x, cn.y, cn.z}, na[3] = {n.x, n.y, n.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            H[r + 3 * c] = k * (ca[r] * na[c]);
        }
    }
}

__device__ inline f3 normalized(f3 e) {
    const float n = sqrtf(dot3(e, e));
    return n > 0.0f ? mk3(e.x / n, e.y / n, e.z / n) : e;
}

// Stiffness (K ne).ne / |ne|^2 of the N vertices idx with weights c along n,
// K = elastic blocks + mass / gap^2 on the diagonal blocks.
template <int N>
__device__ float ext_stiffness(const ppf_contact_params &cp, const uint32_t *idx,
                               const float *c, f3 n, float gap2) {
    const float na[3] = {n.x, n.y, n.z};
    float ne[3 * N];
#pragma unroll
    for (int j = 0; j < 3 * N; ++j) ne[j] = c[j / 3] * na[j % 3];
    float Kn[3 * N];
#pragma unroll
    for (int j = 0; j < 3 * N; ++j) Kn[j] = 0.0f;
    for (int jj = 0; jj < N; ++jj) {
        for (int ii = 0; ii < N; ++ii) {
            float blk[9];
            lookup(idx[ii], idx[jj], cp.el_rp, cp.el_col, cp.el_val, blk);
            const float sd = ii == jj ? cp.mass[idx[ii]] / gap2 : 0.0f;
#pragma unroll
            for (int e = 0; e < 9; ++e) {
                blk[e] = blk[e] + ((e % 4 == 0) ? sd : 0.0f);
            }
            // block (ii, jj) occupies rows 3ii.., cols 3jj..; K ne accumulates
            // over the 3N columns in order
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
#pragma unroll
                for (int r = 0; r < 3; ++r) Kn[3 * ii + r] = Kn[3 * ii + r] + blk[r + 3 * cc] * ne[3 * jj + cc];
        }
    }
    float num = 0.0f, den = 0.0f;
#pragma unroll
    for (int j = 0; j < 3 * N; ++j) num = num + Kn[j] * ne[j];
#pragma unroll
    for (int j = 0; j < 3 * N; ++j) den = den + ne[j] * ne[j];
    return num / den;
}

// Collision-mesh pairs (contact.cu:899-1152): kind 4 dynamic vertex a vs
// static face b, 5 static vertex a vs dynamic face b, 6 dynamic edge a vs
// static edge b.
__device__ void eval_static(const ppf_contact_mesh &m, const ppf_contact_params &cp,
                            const ppf_static_mesh &sm, uint32_t kind, uint32_t a, uint32_t b,
                            uint32_t idx[4], float w[4], int &N, f3 &f, float H[9]) {
    const f3 zero = mk3(0.0f, 0.0f, 0.0f);
    if (kind == 4) {
        const f3 p = X(m.x, a);
        const f3 q = X(cp.x0, a) - p;
        const uint32_t f0 = sm.face[3 * (size_t)b], f1 = sm.face[3 * (size_t)b + 1],
                       f2 = sm.face[3 * (size_t)b + 2];
        const f3 t0 = X(sm.x, f0) - p, t1 = X(sm.x, f1) - p, t2 = X(sm.x, f2) - p;
        float c[3];
        pt_coeff_unclassified(zero, t0, t1, t2, c);
        const f3 y = (c[0] * t0 + c[1] * t1) + c[2] * t2;
        const f3 e = mk3(-y.x, -y.y, -y.z);
        const float off = m.v_off[a] + sm.f_off[b];
        const float gh = 0.5f * (m.v_ghat[a] + sm.f_ghat[b]);
        const float mu = combine(cp.v_fric[a], sm.f_fric[b], cp.friction_mode);
        const f3 n = normalized(e);
        const f3 proj = y + (off + gh) * n;
        const float g = sqrtf(dot3(e, e)) - off;
        const float gap2 = g * g;
        float blk[9];
        lookup(a, a, cp.el_rp, cp.el_col, cp.el_val, blk);
        const float k = dot3(n, mv(blk, n)) + cp.mass[a] / gap2;
        push_terms(mk3(-proj.x, -proj.y, -proj.z), n, gh, k, f, H);
        add_friction(n, mk3(-q.x, -q.y, -q.z), mu, cp.friction_eps, f, H);
        N = 1;
        idx[0] = a;
        w[0] = 1.0f;
    } else if (kind == 5) {
        const f3 yv = X(sm.x, a);
        const uint32_t fc[3] = {m.face[3 * (size_t)b], m.face[3 * (size_t)b + 1],
                                m.face[3 * (size_t)b + 2]};
        const f3 t0 = X(m.x, fc[0]) - yv, t1 = X(m.x, fc[1]) - yv, t2 = X(m.x, fc[2]) - yv;
        const f3 s0 = X(cp.x0, fc[0]) - yv, s1 = X(cp.x0, fc[1]) - yv,
                 s2 = X(cp.x0, fc[2]) - yv;
        float c[3];
        pt_coeff_unclassified(zero, t0, t1, t2, c);
        const f3 p = (c[0] * t0 + c[1] * t1) + c[2] * t2;
        const f3 q = (c[0] * s0 + c[1] * s1) + c[2] * s2;
        const float off = m.f_off[b] + sm.v_off[a];
        const float gh = 0.5f * (m.f_ghat[b] + sm.v_ghat[a]);
        const float mu = combine(cp.f_fric[b], sm.v_fric[a], cp.friction_mode);
        const f3 n = normalized(p);
        const float g = sqrtf(dot3(p, p)) - off;
        const float gap2 = g * g;
        const float k = ext_stiffness<3>(cp, fc, c, n, gap2);
        const f3 proj = (off + gh) * n;
        push_terms(p - proj, n, gh, k, f, H);
        add_friction(n, p - q, mu, cp.friction_eps, f, H);
        N = 3;
        idx[0] = fc[0], idx[1] = fc[1], idx[2] = fc[2];
        w[0] = c[0], w[1] = c[1], w[2] = c[2];
    } else {
        const uint32_t e2[2] = {m.edge[2 * (size_t)a], m.edge[2 * (size_t)a + 1]};
        const f3 p0 = X(m.x, e2[0]), p1 = X(m.x, e2[1]);
        const f3 q0 = X(sm.x, sm.edge[2 * (size_t)b]), q1 = X(sm.x, sm.edge[2 * (size_t)b + 1]);
        const f3 b0 = X(cp.x0, e2[0]), b1 = X(cp.x0, e2[1]);
        float c[4];
        ee_coeff(p0, p1, q0, q1, c);
        const f3 x = c[0] * p0 + c[1] * p1;
        const f3 y = c[2] * q0 + c[3] * q1;
        const f3 z = c[0] * b0 + c[1] * b1;
        const f3 e = x - y;
        const float off = m.e_off[a] + sm.e_off[b];
        const float gh = 0.5f * (m.e_ghat[a] + sm.e_ghat[b]);
        const float mu = combine(cp.e_fric[a], sm.e_fric[b], cp.friction_mode);
        const f3 n = normalized(e);
        const f3 proj = (off + gh) * n;
        const float g = sqrtf(dot3(e, e)) - off;
        const float gap2 = g * g;
        const float k = ext_stiffness<2>(cp, e2, c, n, gap2);
        push_terms(e - proj, n, gh, k, f, H);
        add_friction(n, x - z, mu, cp.friction_eps, f, H);
        N = 2;
        idx[0] = e2[0], idx[1] = e2[1];
        w[0] = c[0], w[1] = c[1];
    }
}

__global__ __launch_bounds__(kNT) void k_pair_eval(ppf_contact_mesh m,
                                                   ppf_contact_params cp,
                                                   ppf_static_mesh sm,
                                                   const ppf_pair *__restrict__ pairs,
                                                   uint64_t npairs,
                                                   PairEval *__restrict__ out) {
    const uint64_t pi = (uint64_t)blockIdx.x * kNT + threadIdx.x;
    if (pi >= npairs) {
        return;
    }
    const ppf_pair pr = pairs[pi];
    const uint32_t kind = pr.kind & 0xffu;
    const bool no_fric = (pr.kind >> 8) & 1u;
    const uint32_t a = pr.a, b = pr.b;
    uint32_t idx[4] = {0, 0, 0, 0};
    float w[4] = {0, 0, 0, 0};
    int N;
    if (kind >= 4) {
        f3 f;
        float H[9];
        eval_static(m, cp, sm, kind, a, b, idx, w, N, f, H);
        PairEval pe;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            pe.idx[k] = k < N ? idx[k] : kInvalid;
            pe.w[k] = w[k];
        }
        pe.f[0] = f.x, pe.f[1] = f.y, pe.f[2] = f.z;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            pe.H[k] = H[k];
        }
        pe.cnt = (float)pr.count;
        pe.n = (uint32_t)N;
        out[pi] = pe;
        return;
    }
    float gh, off, mu;
    if (kind == 0) {
        const uint32_t f0 = m.face[3 * (size_t)b], f1 = m.face[3 * (size_t)b + 1],
                       f2 = m.face[3 * (size_t)b + 2];
        float c[3];
        pt_coeff(X(m.x, a), X(m.x, f0), X(m.x, f1), X(m.x, f2), c);
        N = 4;
        idx[0] = a, idx[1] = f0, idx[2] = f1, idx[3] = f2;
        w[0] = 1.0f, w[1] = -c[0], w[2] = -c[1], w[3] = -c[2];
        off = m.v_off[a] + m.f_off[b];
        gh = 0.5f * (m.v_ghat[a] + m.f_ghat[b]);
        mu = combine(cp.v_fric[a], cp.f_fric[b], cp.friction_mode);
    } else if (kind == 1) {
        const uint32_t e0 = m.edge[2 * (size_t)b], e1 = m.edge[2 * (size_t)b + 1];
        float c[2];
        pe_coeff(X(m.x, a), X(m.x, e0), X(m.x, e1), c);
        N = 3;
        idx[0] = a, idx[1] = e0, idx[2] = e1;
        w[0] = 1.0f, w[1] = -c[0], w[2] = -c[1];
        off = m.v_off[a] + m.e_off[b];
        gh = 0.5f * (m.v_ghat[a] + m.e_ghat[b]);
        mu = combine(cp.v_fric[a], cp.e_fric[b], cp.friction_mode);
    } else if (kind == 2) {
        N = 2;
        idx[0] = a, idx[1] = b;
        w[0] = 1.0f, w[1] = -1.0f;
        off = m.v_off[a] + m.v_off[b];
        gh = 0.5f * (m.v_ghat[a] + m.v_ghat[b]);
        mu = combine(cp.v_fric[a], cp.v_fric[b], cp.friction_mode);
    } else {
        const uint32_t a0 = m.edge[2 * (size_t)a], a1 = m.edge[2 * (size_t)a + 1];
        const uint32_t b0 = m.edge[2 * (size_t)b], b1 = m.edge[2 * (size_t)b + 1];
        float c[4];
        ee_coeff(X(m.x, a0), X(m.x, a1), X(m.x, b0), X(m.x, b1), c);
        N = 4;
        idx[0] = a0, idx[1] = a1, idx[2] = b0, idx[3] = b1;
        w[0] = c[0], w[1] = c[1], w[2] = -c[2], w[3] = -c[3];
        off = m.e_off[a] + m.e_off[b];
        gh = 0.5f * (m.e_ghat[a] + m.e_ghat[b]);
        mu = combine(cp.e_fric[a], cp.e_fric[b], cp.friction_mode);
    }
    // Anchored separations (contact.cu:286-307).
    f3 ex0 = mk3(0, 0, 0), ex = mk3(0, 0, 0);
    const f3 an0 = X(cp.x0, idx[0]), an = X(m.x, idx[0]);
    float mass[4];
    for (int ii = 0; ii < N; ++ii) {
        mass[ii] = cp.mass[idx[ii]];
        ex0 = ex0 + w[ii] * (X(cp.x0, idx[ii]) - an0);
        ex = ex + w[ii] * (X(m.x, idx[ii]) - an);
    }
    const f3 dx = ex - ex0;
    f3 nrm = ex;
    {
        const float nl = sqrtf(dot3(ex, ex));
        if (nl > 0.0f) {
            nrm = mk3(ex.x / nl, ex.y / nl, ex.z / nl);
        }
    }
    // Elasticity-inclusive stiffness (barrier.cu:49-86).
    const float g = sqrtf(dot3(ex, ex)) - off;
    const float sqr_x = g * g;
    float m_max = 0.0f;
    for (int ii = 0; ii < N; ++ii) {
        if (mass[ii] > m_max) {
            m_max = mass[ii];
        }
    }
    float wv[12];
    const float exa[3] = {ex.x, ex.y, ex.z};
    for (int ii = 0; ii < N; ++ii) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            wv[3 * ii + c] = w[ii] * exa[c];
        }
    }
    {
        float sq = 0.0f;
        for (int k = 0; k < 3 * N; ++k) {
            sq = sq + wv[k] * wv[k];
        }
        const float ln = sqrtf(sq);
        if (ln > 0.0f) {
            for (int k = 0; k < 3 * N; ++k) {
                wv[k] = wv[k] / ln;
            }
        }
    }
    float stiff = 0.0f;
    for (int ii = 0; ii < N; ++ii) {
        const float m_ii = mass[ii] > 0.0f ? mass[ii] : m_max;
        float hv[3] = {0.0f, 0.0f, 0.0f};
        for (int jj = 0; jj < N; ++jj) {
            float blk[9];
            lookup(idx[ii], idx[jj], cp.el_rp, cp.el_col, cp.el_val, blk);
            if (ii == jj) {
                const float s = m_ii / sqr_x;
                blk[0] = blk[0] + s;
                blk[4] = blk[4] + s;
                blk[8] = blk[8] + s;
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    hv[r] = hv[r] + blk[r + 3 * c] * wv[3 * jj + c];
                }
            }
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            stiff = stiff + hv[r] * wv[3 * ii + r];
        }
    }
    // Cubic barrier force and Hessian (barrier.cu:103-114, cubic.hpp:22-40).
    const float gl = sqrtf(dot3(ex, ex));
    const f3 nn = mk3(ex.x / gl, ex.y / gl, ex.z / gl);
    float gg;
    {
        const float gs = gl - off;
        const float y = gs - gh;
        gg = y < 0.0f ? -2.0f * y * y / gh : 0.0f;
    }
    f3 f = stiff * mk3(nn.x * gg, nn.y * gg, nn.z * gg);
    float cv;
    {
        const float gs = sqrtf(dot3(ex, ex)) - off;
        cv = (gs - gh < 0.0f) ? 4.0f * (1.0f - gs / gh) : 0.0f;
    }
    const float e2 = dot3(ex, ex);
    float H[9];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            H[r + 3 * c] = stiff * ((cv * exa[r]) * exa[c] / e2);
        }
    }
    if (!no_fric) {
        add_friction(nrm, dx, mu, cp.friction_eps, f, H);
    }
    PairEval pe;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        pe.idx[k] = k < N ? idx[k] : kInvalid;
        pe.w[k] = w[k];
    }
    pe.f[0] = f.x, pe.f[1] = f.y, pe.f[2] = f.z;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        pe.H[k] = H[k];
    }
    pe.cnt = (float)pr.count;
    pe.n = (uint32_t)N;
    out[pi] = pe;
}

// Force records: 4 slots per pair, key = vertex (kInvalid for unused slots),
// value = 4 * pair + position. Block records: 16 slots per pair for the
// ordered position pairs (ii, jj), valid when idx[ii] <= idx[jj]; key pair
// (row, col) = (idx[ii], idx[jj]), value = 16 * pair + 4 * ii + jj.
__global__ __launch_bounds__(kNT) void k_records(const PairEval *__restrict__ pe,
                                                 uint64_t npairs,
                                                 uint32_t *__restrict__ fkey,
                                                 uint32_t *__restrict__ fval,
                                                 uint32_t *__restrict__ hrow,
                                                 uint32_t *__restrict__ hcol,
                                                 uint32_t *__restrict__ hval) {
    const uint64_t pi = (uint64_t)blockIdx.x * kNT + threadIdx.x;
    if (pi >= npairs) {
        return;
    }
    const PairEval &p = pe[pi];
#pragma unroll
    for (int ii = 0; ii < 4; ++ii) {
        fkey[4 * pi + ii] = p.idx[ii];
        fval[4 * pi + ii] = (uint32_t)(4 * pi + ii);
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            const uint32_t i = p.idx[ii], j = p.idx[jj];
            const bool ok = i != kInvalid && j != kInvalid && i <= j;
            const uint64_t r = 16 * pi + 4 * ii + jj;
            hrow[r] = ok ? i : kInvalid;
            hcol[r] = ok ? j : kInvalid;
            hval[r] = (uint32_t)r;
        }
    }
}

// Row (which == 0) or column (which == 1) key of block records given by id.
__global__ void k_block_keys(const PairEval *__restrict__ pe,
                             const uint32_t *__restrict__ ids, uint64_t n, int which,
                             uint32_t *__restrict__ dst) {
    const uint64_t i = (uint64_t)blockIdx.x * kNT + threadIdx.x;
    if (i >= n) {
        return;
    }
    const uint32_t id = ids[i];
    const PairEval &p = pe[id >> 4];
    const uint32_t a = p.idx[(id >> 2) & 3], b = p.idx[id & 3];
    const bool ok = a != kInvalid && b != kInvalid && a <= b;
    dst[i] = ok ? (which ? b : a) : kInvalid;
}

// Gradient: one thread per vertex walks its sorted run of force records.
__global__ __launch_bounds__(kNT) void k_force_reduce(
    const uint32_t *__restrict__ keys, const uint32_t *__restrict__ vals,
    uint64_t nrec, const uint32_t *__restrict__ vstart, uint32_t nv,
    const PairEval *__restrict__ pe, float *__restrict__ grad) {
    const uint32_t v = blockIdx.x * kNT + threadIdx.x;
    if (v >= nv) {
        return;
    }
    float g[3] = {0.0f, 0.0f, 0.0f};
    for (uint64_t r = vstart[v]; r < nrec && keys[r] == v; ++r) {
        const uint32_t rec = vals[r];
        const PairEval &p = pe[rec >> 2];
        const int ii = rec & 3;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            g[c] += p.cnt * (p.w[ii] * p.f[c]);
        }
    }
    store3(grad + 3 * (size_t)v, mk3(g[0], g[1], g[2]));
}

// First record of every vertex in the sorted force keys (lower bound).
__global__ void k_lower_bound(const uint32_t *__restrict__ keys, uint64_t n,
                              uint32_t nq, uint32_t *__restrict__ out) {
    const uint32_t v = blockIdx.x * kNT + threadIdx.x;
    if (v > nq) {
        return;
    }
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if (keys[mid] < v) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    out[v] = (uint32_t)lo;
}

// Segment heads of the sorted (row, col) block keys.
__global__ void k_heads(const uint32_t *__restrict__ row, const uint32_t *__restrict__ col,
                        uint64_t n, uint32_t *__restrict__ head) {
    const uint64_t i = (uint64_t)blockIdx.x * kNT + threadIdx.x;
    if (i >= n) {
        return;
    }
    const bool valid = row[i] != kInvalid;
    head[i] = valid && (i == 0 || row[i] != row[i - 1] || col[i] != col[i - 1]) ? 1u : 0u;
}

// One thread per unique block: sum its run of records in order.
__global__ __launch_bounds__(kNT) void k_block_reduce(
    const uint32_t *__restrict__ row, const uint32_t *__restrict__ col,
    const uint32_t *__restrict__ rec, const uint32_t *__restrict__ head,
    const uint32_t *__restrict__ pos, uint64_t n, const PairEval *__restrict__ pe,
    uint32_t *__restrict__ orow, uint32_t *__restrict__ ocol,
    float *__restrict__ oval) {
    const uint64_t i = (uint64_t)blockIdx.x * kNT + threadIdx.x;
    if (i >= n || !head[i]) {
        return;
    }
    const uint32_t r0 = row[i], c0 = col[i];
    float acc[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        acc[k] = 0.0f;
    }
    for (uint64_t q = i; q < n && row[q] == r0 && col[q] == c0; ++q) {
        const uint32_t id = rec[q];
        const PairEval &p = pe[id >> 4];
        const int ii = (id >> 2) & 3, jj = id & 3;
        const float s = p.w[ii] * p.w[jj];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            acc[k] = acc[k] + p.cnt * (s * p.H[k]);
        }
    }
    const uint32_t o = pos[i];
    orow[o] = r0;
    ocol[o] = c0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        oval[9 * (size_t)o + k] = acc[k];
    }
}

uint32_t cdiv64(uint64_t a, uint32_t b) { return (uint32_t)((a + b - 1) / b); }

// Analytic spheres and floors per vertex (data.constraint.sphere / floor;
// embed_vertex_constraint_force_hessian, contact/contact.cu:1154-1382): the
// push barrier with stiffness n.(K_ii n) + m / gap^2 (K_ii the vertex's
// elastic diagonal block) and friction for static colliders, stiffness
// m / ghat^2 and no friction for kinematic spheres. Pinned vertices (their
// rows are eliminated) and massless ones (static solids) are skipped. Each
// collider's push, then its friction, is added to the vertex's gradient row
// and diagonal block, in collider order (one thread per vertex: no atomics).
__global__ __launch_bounds__(kNT) void k_analytic(ppf_contact_mesh m, ppf_contact_params cp,
                                                  const ppf_sphere *__restrict__ sp,
                                                  uint32_t nsp, const ppf_floor *__restrict__ fl,
                                                  uint32_t nfl, float constraint_tol,
                                                  float *__restrict__ grad,
                                                  float *__restrict__ cdiag) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i >= m.nv) {
        return;
    }
    const float mass = cp.mass[i];
    if ((m.vflag[i] & 1u) || !(mass > 0.0f)) {
        return;
    }
    const f3 x = X(m.x, i);
    const f3 dx = x - X(cp.x0, i);
    float lh[9];
    lookup(i, i, cp.el_rp, cp.el_col, cp.el_val, lh);
    f3 F = X(grad, i);
    float *Hg = cdiag + 9 * (size_t)i;
    float H[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        H[e] = Hg[e];
    }
    f3 fp, fg;
    float Hp[9], Hf[9];
    for (uint32_t j = 0; j < nsp; ++j) {
        const ppf_sphere s = sp[j];
        const float ghat = s.ghat;
        const float mu = combine(s.friction, cp.v_fric[i], cp.friction_mode);
        bool reverse = s.reverse;
        float radius = s.radius;
        f3 center = mk3(s.center[0], s.center[1], s.center[2]);
        if (s.bowl && x.y > center.y) {
            center = mk3(center.x, x.y, center.z);
        }
        const f3 o = x - center;
        const float d2 = dot3(o, o);
        if (s.kinematic) {
            if (d2 == 0.0f) {
                continue;
            }
            const f3 normal = normalized(o);
            const float eff = reverse ? (radius - ghat) : (radius + ghat);
            const f3 target = eff * normal;
            if (s.bowl) {
                reverse = true;
            }
            const float k = mass / (ghat * ghat);
            const float r2 = eff * eff;
            const float dist = sqrtf(d2);
            const float depth = reverse ? (dist - s.radius) : (s.radius - dist);
            if (s.thickness > 0.0f && depth > s.thickness) {
                continue;
            }
            if (reverse ? d2 > r2 : d2 < r2) {
                const f3 nn = reverse ? mk3(-normal.x, -normal.y, -normal.z) : normal;
                push_terms(o - target, nn, ghat, k, fp, Hp);
                F = F + fp;
#pragma unroll
                for (int e = 0; e < 9; ++e) {
                    H[e] = H[e] + Hp[e];
                }
            }
            continue;
        }
        radius = reverse ? radius - ghat : radius + ghat;
        const float r2 = radius * radius;
        if (!(reverse ? d2 > r2 : d2 < r2)) {
            continue;
        }
        const float dist = sqrtf(d2);
        const float depth = reverse ? (dist - s.radius) : (s.radius - dist);
        if (s.thickness > 0.0f && depth > s.thickness) {
            continue;
        }
        f3 normal = normalized(o);
        const f3 proj = radius * normal;
        if (reverse) {
            normal = mk3(-normal.x, -normal.y, -normal.z);
        }
        const float gap = reverse ? s.radius - sqrtf(d2) : sqrtf(d2) - s.radius;
        const float k = dot3(normal, mv(lh, normal)) + mass / (gap * gap);
        push_terms(o - proj, normal, ghat, k, fp, Hp);
        F = F + fp;
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            H[e] = H[e] + Hp[e];
        }
        friction_terms(normal, dx, mu, cp.friction_eps, fp, fg, Hf);
        F = F + fg;
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            H[e] = H[e] + Hf[e];
        }
    }
    for (uint32_t j = 0; j < nfl; ++j) {
        const ppf_floor f = fl[j];
        const float ghat = f.ghat;
        const float mu = combine(f.friction, cp.v_fric[i], cp.friction_mode);
        const f3 up = mk3(f.up[0], f.up[1], f.up[2]);
        const f3 g0 = mk3(f.ground[0], f.ground[1], f.ground[2]);
        const f3 e = x - (g0 + ghat * up);
        const float eu = dot3(e, up);
        if (!(eu < 0.0f)) {
            continue;
        }
        if (f.thickness > 0.0f && -eu > f.thickness) {
            continue;
        }
        const f3 proj = (-eu) * up;
        float gap = dot3(x - g0, up);
        if (f.kinematic) {
            gap = fmaxf(gap, constraint_tol * ghat);
        }
        const float k = dot3(up, mv(lh, up)) + mass / (gap * gap);
        push_terms(mk3(-proj.x, -proj.y, -proj.z), up, ghat, k, fp, Hp);
        F = F + fp;
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            H[q] = H[q] + Hp[q];
        }
        friction_terms(up, dx, mu, cp.friction_eps, fp, fg, Hf);
        F = F + fg;
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            H[q] = H[q] + Hf[q];
        }
    }
    store3(grad + 3 * (size_t)i, F);
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        Hg[e] = H[e];
    }
}

} // namespace

ContactAssembler::~ContactAssembler() { release(); }

void ContactAssembler::release() {
    for (void *p : {(void *)pe_, (void *)buf_, (void *)out_row_, (void *)out_col_,
                    (void *)out_val_, (void *)rp_}) {
        if (p) {
            (void)dev_free(p);
        }
    }
    pe_ = nullptr;
    buf_ = out_row_ = out_col_ = rp_ = nullptr;
    out_val_ = nullptr;
    pcap_ = 0;
    vcap_ = 0;
}

int ContactAssembler::reserve(uint64_t npairs, uint32_t nv) {
    if (npairs <= pcap_ && nv <= vcap_) {
        return PPF_OK;
    }
    release();
    npairs = grow_cap(npairs, 256);
    const uint64_t nh = 16 * npairs, nf = 4 * npairs;
    const size_t sortw = radix_sort_scratch_words(nh);
    const size_t scanw = scan_scratch_words(nh + 1);
    // words: fkey fval (nf each), hrow hcol hval head pos (nh each),
    // vstart (nv + 1), sort scratch, scan scratch.
    const size_t words = 2 * nf + 5 * nh + (nv + 1) + sortw + scanw;
    PPF_HIP(dev_alloc(&pe_, npairs * sizeof(PairEval)));
    PPF_HIP(dev_alloc(&buf_, words * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&out_row_, nh * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&out_col_, nh * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&out_val_, 9 * nh * sizeof(float)));
    PPF_HIP(dev_alloc(&rp_, (nv + 1) * sizeof(uint32_t)));
    pcap_ = npairs;
    vcap_ = nv;
    return PPF_OK;
}

int ContactAssembler::assemble(const ppf_contact_mesh &m, const ppf_contact_params &cp,
                               const ppf_pair *pairs, uint64_t npairs, float *grad,
                               hipStream_t s) {
    const uint32_t nv = m.nv;
    int rc = reserve(npairs, nv);
    if (rc != PPF_OK) {
        return rc;
    }
    nnz_ = 0;
    if (npairs == 0) {
        PPF_HIP(hipMemsetAsync(grad, 0, 3 * (size_t)nv * sizeof(float), s));
        PPF_HIP(hipMemsetAsync(rp_, 0, (nv + 1) * sizeof(uint32_t), s));
        return PPF_OK;
    }
    if (16 * npairs >= kInvalid) {
        set_error("contact assemble: %llu pairs exceed the record index range",
                  (unsigned long long)npairs);
        return PPF_ERR_CAPACITY;
    }
    const uint64_t nh = 16 * npairs, nf = 4 * npairs;
    uint32_t *fkey = buf_, *fval = fkey + nf;
    uint32_t *hrow = fval + nf, *hcol = hrow + nh, *hval = hcol + nh;
    uint32_t *head = hval + nh, *pos = head + nh;
    uint32_t *vstart = pos + nh;
    uint32_t *sortw = vstart + (nv + 1);
    uint32_t *scanw = sortw + radix_sort_scratch_words(nh);

    ppf_static_mesh sm{};
    if (cp.sm) {
        sm = *cp.sm;
    }
    k_pair_eval<<<cdiv64(npairs, kNT), kNT, 0, s>>>(m, cp, sm, pairs, npairs, pe_);
    PPF_HIP(hipGetLastError());
    k_records<<<cdiv64(npairs, kNT), kNT, 0, s>>>(pe_, npairs, fkey, fval, hrow,
                                                  hcol, hval);
    PPF_HIP(hipGetLastError());
    // Gradient: group force records by vertex, sum runs in record order.
    PPF_HIP(radix_sort_pairs(fkey, fval, nf, 32, sortw, s));
    k_lower_bound<<<cdiv64(nv + 1, kNT), kNT, 0, s>>>(fkey, nf, nv, vstart);
    PPF_HIP(hipGetLastError());
    k_force_reduce<<<cdiv64(nv, kNT), kNT, 0, s>>>(fkey, fval, nf, vstart, nv, pe_,
                                                   grad);
    PPF_HIP(hipGetLastError());
    // Blocks: stable sort by column, then stable sort by row: records end up
    // ordered by (row, col), equal keys in record order.
    PPF_HIP(radix_sort_pairs(hcol, hval, nh, 32, sortw, s));
    k_block_keys<<<cdiv64(nh, kNT), kNT, 0, s>>>(pe_, hval, nh, 0, hrow);
    PPF_HIP(hipGetLastError());
    PPF_HIP(radix_sort_pairs(hrow, hval, nh, 32, sortw, s));
    k_block_keys<<<cdiv64(nh, kNT), kNT, 0, s>>>(pe_, hval, nh, 1, hcol);
    PPF_HIP(hipGetLastError());
    k_heads<<<cdiv64(nh, kNT), kNT, 0, s>>>(hrow, hcol, nh, head);
    PPF_HIP(hipGetLastError());
    PPF_HIP(exclusive_scan_u32(head, pos, nh, scanw, s));
    uint32_t last_pos = 0, last_head = 0;
    PPF_HIP(hipMemcpyAsync(&last_pos, pos + nh - 1, sizeof(uint32_t),
                           hipMemcpyDeviceToHost, s));
    PPF_HIP(hipMemcpyAsync(&last_head, head + nh - 1, sizeof(uint32_t),
                           hipMemcpyDeviceToHost, s));
    PPF_HIP(hipStreamSynchronize(s));
    nnz_ = (uint64_t)last_pos + last_head;
    k_block_reduce<<<cdiv64(nh, kNT), kNT, 0, s>>>(hrow, hcol, hval, head, pos,
                                                   nh, pe_, out_row_, out_col_,
                                                   out_val_);
    PPF_HIP(hipGetLastError());
    k_lower_bound<<<cdiv64(nv + 1, kNT), kNT, 0, s>>>(out_row_, nnz_, nv, rp_);
    PPF_HIP(hipGetLastError());
    return PPF_OK;
}

int analytic_assemble(const ppf_contact_mesh &m, const ppf_contact_params &cp,
                      const ppf_sphere *spheres, uint32_t nsphere, const ppf_floor *floors,
                      uint32_t nfloor, float constraint_tol, float *grad, float *cdiag,
                      hipStream_t s) {
    if (m.nv == 0 || (nsphere == 0 && nfloor == 0)) {
        return PPF_OK;
    }
    k_analytic<<<cdiv64(m.nv, kNT), kNT, 0, s>>>(m, cp, spheres, nsphere, floors, nfloor,
                                                  constraint_tol, grad, cdiag);
    PPF_HIP(hipGetLastError());
    return PPF_OK;
}

} // namespace ppf
