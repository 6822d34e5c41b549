// linesearch.hip -- contact CCD line search on MI355X.
//
// Replaces contact::line_search (contact/contact.cu:2472-2740) with the CCD
// functors of contact.cu:1956-2316: per dynamic vertex the point-face
// (own faces: the vertex against the opposite edge), point-point and
// vertex-vs-static-face sweeps; per static vertex the face sweeps; the
// vertex-face minimum then seeds every edge's sweep box for the edge-edge
// (shared vertex: two point-edge sweeps) and edge-vs-static-edge sweeps.
// Boxes are the reference's swept boxes; the minima are combined with
// integer atomics on the float bits (all values are non-negative), which is
// exact and order independent.
#include "bvh_query.hpp"
#include "ccd.hpp"
#include "linesearch.hpp"
#include <type_traits>
#include <vector>

namespace ppf {

namespace {

constexpr int kStack = kQueryStack;
// The sweeps first run with shorter stacks: 36 entries (10 KB of LDS per
// one-wave workgroup) let 16 workgroups share a CU, 4 waves per SIMD, against
// 12 with 48. A walk deeper than 35 levels flags an overflow, and the pass
// runs again with 48-entry stacks.
constexpr int kShortStack = 36;
constexpr int kLsWaves = 4;
using Box6 = QBox;

struct LsCtx {
    ppf_contact_mesh m;
    ppf_static_mesh sm;
    ppf_ls_params lp;
    BvhView fb, eb, vb, sfb, seb;
    uint32_t *out; // [0] t_vf bits, [1] t_ee bits, [2] overlap, [3] stack overflow,
                   // [4] smallest pinned vertex driven into an analytic collider
    // many scenes: per scene t_vf bits [0, ns), t_ee bits [ns, 2 ns), overlap
    // [2 ns, 3 ns); null for one scene
    uint32_t *sout;
    uint32_t ns;
    unsigned long long *stats; // walk statistics (null: off)
};

__device__ inline f3 X(const float *a, uint32_t v) { return load3(a + 3 * (size_t)v); }

__device__ inline f3 XS(const LsCtx &c, uint32_t v) {
    const f3 a = X(c.m.x, v), b = X(c.lp.x1, v);
    const float t = c.lp.max_t;
    return mk3(t * (b.x - a.x) + a.x, t * (b.y - a.y) + a.y, t * (b.z - a.z) + a.z);
}

// One sweep's result; `scene` is the query's scene and `ee` selects the
// edge-edge minimum (many scenes keep their minima per scene).
__device__ inline void note(const LsCtx &c, bool ovl, float r, float &best, uint32_t scene = 0,
                            bool ee = false) {
    if (ovl) {
        atomicOr(c.out + 2, 1u);
        if (c.sout) {
            atomicOr(c.sout + 2 * c.ns + scene, 1u);
        }
    }
    if (r < c.lp.max_t) {
        best = fminf(best, r);
        if (c.sout) {
            atomicMin(c.sout + (ee ? c.ns : 0u) + scene, __float_as_uint(r));
        }
    }
}

constexpr int kWQ = kWaveQueries;

// Smallest value over the wave, then one atomic (values are non-negative
// floats, so their bits order like unsigned integers).
__device__ inline void wave_min_out(uint32_t *out, float best, float above) {
    float v = best;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v = fminf(v, __shfl_xor(v, o));
    }
    if (threadIdx.x == 0 && v < above) {
        atomicMin(out, __float_as_uint(v));
    }
}

// Per dynamic vertex: point-face (contact.cu:2110-2164), point-point
// (contact.cu:2266-2316) and vertex vs static face (contact.cu:1956-1994).
// The filters run on the query's lane while it walks; the sweeps run from the
// wave queue.
template <int STACK>
__global__ __launch_bounds__(kWQ, kLsWaves) void k_ls_vertex(LsCtx c) {
    __shared__ WaveScratchT<STACK> ws;
    const uint32_t leaf = xcd_logical_block() * kWQ + threadIdx.x;
    const bool active = leaf < c.vb.n;
    const ppf_contact_mesh &m = c.m;
    const float max_t = c.lp.max_t, eps = c.lp.ccd_eps;
    const uint32_t i = active ? c.vb.node[2 * (size_t)leaf] - 1 : 0;
    const uint32_t vf = active ? m.vflag[i] : 4u;
    Box6 q = (vf & 4u) ? empty_box()
                       : seg_box(X(m.x, i), XS(c, i), 0.5f * m.v_ghat[i] + m.v_off[i]);
    q.scene = vertex_scene(m.vscene, m.nscene, i);
    const bool idyn = (vf & 1u) == 0, icol = (vf & 2u) != 0;
    const uint32_t ibody = active ? m.vbody[i] : 0u;
    float best = max_t;
    WalkCount wc;
    traverse_wave(
        c.fb, q, active, i, ws, c.out + 3,
        [&](uint32_t fi) {
            const uint32_t f0 = m.face[3 * (size_t)fi];
            const bool either = idyn || (m.fflag[fi] & 1u) == 0;
            const bool same = ibody != 0 && ibody == m.vbody[f0];
            const bool both = icol && (m.vflag[f0] & 2u) != 0;
            return either && !same && !both;
        },
        [&](uint32_t v, uint32_t fi) {
            const uint32_t f[3] = {m.face[3 * (size_t)fi], m.face[3 * (size_t)fi + 1],
                                   m.face[3 * (size_t)fi + 2]};
            int dup = -1;
            for (int k = 0; k < 3; ++k) {
                if (f[k] == v) {
                    dup = k;
                    break;
                }
            }
            bool ovl = false;
            float r;
            if (dup < 0) {
                const f3 a[4] = {X(m.x, v), X(m.x, f[0]), X(m.x, f[1]), X(m.x, f[2])};
                const f3 b[4] = {X(c.lp.x1, v), X(c.lp.x1, f[0]), X(c.lp.x1, f[1]),
                                 X(c.lp.x1, f[2])};
                r = ccd<4, 1>(a, b, m.v_off[v] + m.f_off[fi],
                              0.5f * (m.v_ghat[v] + m.f_ghat[fi]), max_t, eps, ovl);
            } else {
                const uint32_t p = f[dup], j = f[(dup + 1) % 3], k = f[(dup + 2) % 3];
                const f3 a[3] = {X(m.x, p), X(m.x, j), X(m.x, k)};
                const f3 b[3] = {X(c.lp.x1, p), X(c.lp.x1, j), X(c.lp.x1, k)};
                r = ccd<3, 1>(a, b, 2.0f * m.v_off[v], m.v_ghat[v], max_t, eps, ovl);
            }
            note(c, ovl, r, best, vertex_scene(m.vscene, m.nscene, v));
        }, &wc);
    traverse_wave(
        c.vb, q, active, i, ws, c.out + 3,
        [&](uint32_t j) {
            const bool either = idyn || (m.vflag[j] & 1u) == 0;
            const bool same = ibody != 0 && ibody == m.vbody[j];
            const bool both = icol && (m.vflag[j] & 2u) != 0;
            return j < i && either && !same && !both;
        },
        [&](uint32_t v, uint32_t j) {
            const f3 a[2] = {X(m.x, v), X(m.x, j)};
            const f3 b[2] = {X(c.lp.x1, v), X(c.lp.x1, j)};
            bool ovl = false;
            const float r = ccd<2, 1>(a, b, m.v_off[v] + m.v_off[j],
                                      0.5f * (m.v_ghat[v] + m.v_ghat[j]), max_t, eps, ovl);
            note(c, ovl, r, best, vertex_scene(m.vscene, m.nscene, v));
        }, &wc);
    if (c.sm.nf) {
        const bool go = active && idyn && c.lp.vmass[i] != 0.0f;
        traverse_wave(
            c.sfb, q, go, i, ws, c.out + 3, [&](uint32_t) { return true; },
            [&](uint32_t v, uint32_t fi) {
                const uint32_t f0 = c.sm.face[3 * (size_t)fi], f1 = c.sm.face[3 * (size_t)fi + 1],
                               f2 = c.sm.face[3 * (size_t)fi + 2];
                const f3 t0 = X(c.sm.x, f0), t1 = X(c.sm.x, f1), t2 = X(c.sm.x, f2);
                const f3 a[4] = {X(m.x, v), t0, t1, t2};
                const f3 b[4] = {X(c.lp.x1, v), t0, t1, t2};
                bool ovl = false;
                const float r = ccd<4, 1>(a, b, m.v_off[v] + c.sm.f_off[fi],
                                          0.5f * (m.v_ghat[v] + c.sm.f_ghat[fi]), max_t, eps,
                                          ovl);
                note(c, ovl, r, best, vertex_scene(m.vscene, m.nscene, v));
            }, &wc);
    }
    wave_min_out(c.out + 0, best, max_t);
    walk_flush(c.stats, kWalkLsVertex, wc, active);
}

// static vertex vs dynamic face (contact.cu:1996-2034)
template <int STACK>
__global__ __launch_bounds__(kWQ, kLsWaves) void k_ls_static_vertex(LsCtx c) {
    __shared__ WaveScratchT<STACK> ws;
    const uint32_t i = xcd_logical_block() * kWQ + threadIdx.x;
    const bool active = i < c.sm.nv;
    const ppf_contact_mesh &m = c.m;
    const float max_t = c.lp.max_t, eps = c.lp.ccd_eps;
    const uint32_t ii = active ? i : 0;
    const f3 p = X(c.sm.x, ii);
    const float qm = 0.5f * c.sm.v_ghat[ii] + c.sm.v_off[ii];
    const Box6 q{{p.x - qm, p.y - qm, p.z - qm}, {p.x + qm, p.y + qm, p.z + qm},
                 vertex_scene(c.sm.vscene, m.nscene, ii)};
    float best = max_t;
    traverse_wave(
        c.fb, q, active, ii, ws, c.out + 3,
        [&](uint32_t fi) { return (m.fflag[fi] & 1u) == 0 && c.lp.fmass[fi] > 0.0f; },
        [&](uint32_t v, uint32_t fi) {
            const f3 pv = X(c.sm.x, v);
            const uint32_t f0 = m.face[3 * (size_t)fi], f1 = m.face[3 * (size_t)fi + 1],
                           f2 = m.face[3 * (size_t)fi + 2];
            const f3 a[4] = {pv, X(m.x, f0), X(m.x, f1), X(m.x, f2)};
            const f3 b[4] = {pv, X(c.lp.x1, f0), X(c.lp.x1, f1), X(c.lp.x1, f2)};
            bool ovl = false;
            const float r = ccd<4, 1>(a, b, m.f_off[fi] + c.sm.v_off[v],
                                      0.5f * (m.f_ghat[fi] + c.sm.v_ghat[v]), max_t, eps,
                                      ovl);
            note(c, ovl, r, best, vertex_scene(c.sm.vscene, m.nscene, v));
        });
    wave_min_out(c.out + 0, best, max_t);
}

// Analytic spheres and floors per vertex (vertex_constraint_line_search,
// contact.cu:2318-2470): the swept step x0 -> x0 + max_t (x1 - x0) against
// each static collider; the entry time (times max_t) joins the vertex-face
// minimum. A pinned vertex driven from outside into a static collider cannot
// yield: the smallest such vertex index is flagged. Kinematic colliders are
// held by the barrier alone and not swept.
__global__ __launch_bounds__(256) void k_ls_analytic(LsCtx c) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const ppf_contact_mesh &m = c.m;
    const float max_t = c.lp.max_t;
    float toi = max_t;
    if (i < m.nv) {
        const f3 y0 = X(m.x, i), y1 = X(c.lp.x1, i);
        const f3 x0 = y0;
        const f3 x1 = max_t * (y1 - y0) + y0;
        const bool pinned = (m.vflag[i] & 1u) != 0;
        if (pinned || c.lp.vmass[i] > 0.0f) {
            for (uint32_t j = 0; j < c.lp.nsphere; ++j) {
                const ppf_sphere sp = c.lp.spheres[j];
                if (sp.kinematic) {
                    continue;
                }
                const f3 cc = mk3(sp.center[0], sp.center[1], sp.center[2]);
                const f3 c0 = (sp.bowl && x0.y > cc.y) ? mk3(cc.x, x0.y, cc.z) : cc;
                const f3 c1 = (sp.bowl && x1.y > cc.y) ? mk3(cc.x, x1.y, cc.z) : cc;
                const float r = sp.radius;
                const f3 a0 = x0 - c0, a1 = x1 - c1;
                const float r0 = sqrtf(dot3(a0, a0)), r1 = sqrtf(dot3(a1, a1));
                if (pinned) {
                    const float d0 = sp.reverse ? (r - r0) : (r0 - r);
                    const float d1 = sp.reverse ? (r - r1) : (r1 - r);
                    if (d0 >= 0.0f && d1 < 0.0f) {
                        atomicMin(c.out + 4, i);
                    }
                    continue;
                }
                const float depth0 = sp.reverse ? (r0 - r) : (r - r0);
                if (sp.thickness > 0.0f && depth0 > sp.thickness) {
                    continue;
                }
                if ((r0 - r) * (r1 - r) <= 0.0f) {
                    const float t = (r - r0) / (r1 - r0);
                    toi = fminf(toi, max_t * t);
                }
            }
            for (uint32_t j = 0; j < c.lp.nfloor; ++j) {
                const ppf_floor fl = c.lp.floors[j];
                if (fl.kinematic) {
                    continue;
                }
                const f3 up = mk3(fl.up[0], fl.up[1], fl.up[2]);
                const f3 g = mk3(fl.ground[0], fl.ground[1], fl.ground[2]);
                const float h0 = dot3(up, x0 - g), h1 = dot3(up, x1 - g);
                if (pinned) {
                    if (h0 >= 0.0f && h1 < 0.0f) {
                        atomicMin(c.out + 4, i);
                    }
                    continue;
                }
                if (fl.thickness > 0.0f && -h0 > fl.thickness) {
                    continue;
                }
                if (h1 < 0.0f) {
                    const float t = -h0 / (h1 - h0);
                    toi = fminf(toi, max_t * t);
                }
            }
        }
    }
    float v = toi;
    for (int o = 32; o > 0; o >>= 1) {
        v = fminf(v, __shfl_xor(v, o));
    }
    if ((threadIdx.x & 63) == 0 && v < max_t) {
        atomicMin(c.out + 0, __float_as_uint(v));
    }
}

// edge-edge and edge vs static edge, seeded by the vertex-face minimum.
template <int STACK>
__global__ __launch_bounds__(kWQ, kLsWaves) void k_ls_edge(LsCtx c) {
    __shared__ WaveScratchT<STACK> ws;
    const uint32_t leaf = xcd_logical_block() * kWQ + threadIdx.x;
    const bool active = leaf < c.eb.n;
    const ppf_contact_mesh &m = c.m;
    const float max_t = c.lp.max_t, eps = c.lp.ccd_eps;
    const uint32_t ei = active ? c.eb.node[2 * (size_t)leaf] - 1 : 0;
    const uint32_t a0 = m.edge[2 * (size_t)ei], a1 = m.edge[2 * (size_t)ei + 1];
    // the vertex-face minimum of the edge's own scene bounds its sweep box
    const float seed = __uint_as_float(
        c.sout ? c.sout[vertex_scene(m.vscene, m.nscene, a0)] : c.out[0]);
    const float ext = 0.5f * m.e_ghat[ei] + m.e_off[ei];
    const f3 p0 = X(m.x, a0), p1 = X(m.x, a1), q0 = X(c.lp.x1, a0), q1 = X(c.lp.x1, a1);
    const f3 s0 = mk3(seed * (q0.x - p0.x) + p0.x, seed * (q0.y - p0.y) + p0.y,
                      seed * (q0.z - p0.z) + p0.z);
    const f3 s1 = mk3(seed * (q1.x - p1.x) + p1.x, seed * (q1.y - p1.y) + p1.y,
                      seed * (q1.z - p1.z) + p1.z);
    const Box6 b0 = seg_box(p0, p1, ext), b1 = seg_box(s0, s1, ext);
    Box6 q;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        q.mn[d] = fminf(b0.mn[d], b1.mn[d]);
        q.mx[d] = fmaxf(b0.mx[d], b1.mx[d]);
    }
    const uint32_t ef = active ? m.eflag[ei] : 4u;
    if (ef & 4u) {
        q = empty_box();
    }
    q.scene = vertex_scene(m.vscene, m.nscene, a0);
    const bool edyn = (ef & 1u) == 0;
    const bool acol = (m.vflag[a0] & 2u) != 0;
    const uint32_t abody = m.vbody[a0];
    float best = seed;
    WalkCount wc;
    traverse_wave(
        c.eb, q, active, ei, ws, c.out + 3,
        [&](uint32_t ej) {
            const uint32_t b0i = m.edge[2 * (size_t)ej];
            const bool either = edyn || (m.eflag[ej] & 1u) == 0;
            const bool same = abody != 0 && abody == m.vbody[b0i];
            const bool both = acol && (m.vflag[b0i] & 2u) != 0;
            return ei < ej && either && !same && !both;
        },
        [&](uint32_t e, uint32_t ej) {
            const uint32_t e0 = m.edge[2 * (size_t)e], e1 = m.edge[2 * (size_t)e + 1];
            const uint32_t b0i = m.edge[2 * (size_t)ej], b1i = m.edge[2 * (size_t)ej + 1];
            const float off = m.e_off[e] + m.e_off[ej];
            const float gh = 0.5f * (m.e_ghat[e] + m.e_ghat[ej]);
            const uint32_t ea[2] = {e0, e1}, eb[2] = {b0i, b1i};
            float r = max_t;
            bool ovl = false;
            const bool shared = e0 == b0i || e0 == b1i || e1 == b0i || e1 == b1i;
            if (!shared) {
                const f3 a[4] = {X(m.x, e0), X(m.x, e1), X(m.x, b0i), X(m.x, b1i)};
                const f3 b[4] = {X(c.lp.x1, e0), X(c.lp.x1, e1), X(c.lp.x1, b0i),
                                 X(c.lp.x1, b1i)};
                r = ccd<4, 2>(a, b, off, gh, max_t, eps, ovl);
            } else {
                const int ij[4][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}};
                for (int k = 0; k < 4; ++k) {
                    const int ii = ij[k][0], jj = ij[k][1];
                    if (ea[ii] == eb[jj]) {
                        const uint32_t i0 = ea[ii], i1 = ea[1 - ii], i2 = eb[1 - jj];
                        const f3 x0a[3] = {X(m.x, i1), X(m.x, i0), X(m.x, i2)};
                        const f3 x1a[3] = {X(c.lp.x1, i1), X(c.lp.x1, i0), X(c.lp.x1, i2)};
                        const f3 x0b[3] = {X(m.x, i2), X(m.x, i0), X(m.x, i1)};
                        const f3 x1b[3] = {X(c.lp.x1, i2), X(c.lp.x1, i0), X(c.lp.x1, i1)};
                        const float r0 = ccd<3, 1>(x0a, x1a, off, gh, max_t, eps, ovl);
                        const float r1 = ccd<3, 1>(x0b, x1b, off, gh, max_t, eps, ovl);
                        r = fminf(r0, r1);
                        break;
                    }
                }
            }
            note(c, ovl, r, best, vertex_scene(m.vscene, m.nscene, e0), true);
        }, &wc);
    if (c.sm.ne) {
        const bool go = active && edyn && c.lp.emass[ei] > 0.0f;
        traverse_wave(
            c.seb, q, go, ei, ws, c.out + 3, [&](uint32_t) { return true; },
            [&](uint32_t e, uint32_t sj) {
                const uint32_t e0 = m.edge[2 * (size_t)e], e1 = m.edge[2 * (size_t)e + 1];
                const uint32_t s0i = c.sm.edge[2 * (size_t)sj],
                               s1i = c.sm.edge[2 * (size_t)sj + 1];
                const f3 w0 = X(c.sm.x, s0i), w1 = X(c.sm.x, s1i);
                const f3 a[4] = {X(m.x, e0), X(m.x, e1), w0, w1};
                const f3 b[4] = {X(c.lp.x1, e0), X(c.lp.x1, e1), w0, w1};
                bool ovl = false;
                const float r = ccd<4, 2>(a, b, m.e_off[e] + c.sm.e_off[sj],
                                          0.5f * (m.e_ghat[e] + c.sm.e_ghat[sj]), max_t,
                                          eps, ovl);
                note(c, ovl, r, best, vertex_scene(m.vscene, m.nscene, e0), true);
            }, &wc);
    }
    wave_min_out(c.out + 1, best, seed);
    walk_flush(c.stats, kWalkLsEdge, wc, active);
}

template <int C, int SPLIT>
__global__ __launch_bounds__(256) void k_ccd_batch(uint32_t n, const float *__restrict__ x0,
                                                   const float *__restrict__ x1,
                                                   const float *__restrict__ offset,
                                                   const float *__restrict__ ghat,
                                                   float max_t, float eps,
                                                   float *__restrict__ toi,
                                                   int32_t *__restrict__ overlap) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) {
        return;
    }
    f3 a[C], b[C];
#pragma unroll
    for (int k = 0; k < C; ++k) {
        a[k] = load3(x0 + 3 * ((size_t)i * C + k));
        b[k] = load3(x1 + 3 * ((size_t)i * C + k));
    }
    bool ovl = false;
    toi[i] = ccd<C, SPLIT>(a, b, offset[i], ghat[i], max_t, eps, ovl);
    if (overlap) {
        overlap[i] = ovl ? 1 : 0;
    }
}

uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

} // namespace

int ccd_batch(int kind, uint32_t n, const float *x0, const float *x1, const float *offset,
              const float *ghat, float max_t, float eps, float *toi, int32_t *overlap,
              hipStream_t s) {
    if (n == 0) {
        return PPF_OK;
    }
    const dim3 g(cdiv(n, 256)), b(256);
    switch (kind) {
    case 0:
        k_ccd_batch<4, 1><<<g, b, 0, s>>>(n, x0, x1, offset, ghat, max_t, eps, toi, overlap);
        break;
    case 1:
        k_ccd_batch<3, 1><<<g, b, 0, s>>>(n, x0, x1, offset, ghat, max_t, eps, toi, overlap);
        break;
    case 2:
        k_ccd_batch<2, 1><<<g, b, 0, s>>>(n, x0, x1, offset, ghat, max_t, eps, toi, overlap);
        break;
    default:
        k_ccd_batch<4, 2><<<g, b, 0, s>>>(n, x0, x1, offset, ghat, max_t, eps, toi, overlap);
        break;
    }
    PPF_HIP(hipGetLastError());
    return PPF_OK;
}

LineSearch::~LineSearch() {
    if (out_) (void)dev_free(out_);
    if (h_out_) (void)host_free(h_out_);
}

int LineSearch::run(const ppf_contact_mesh &m, const ppf_static_mesh *sm,
                    const ppf_ls_params &lp, bool rebuild, float *toi, int *overlap,
                    hipStream_t s, float *scene_toi, int *scene_overlap, uint32_t *pin_bad) {
    const uint32_t ns = (m.vscene && m.nscene > 1) ? m.nscene : 0u;
    const uint32_t words = 5 + 3 * ns;
    if (words > out_cap_) {
        if (out_) {
            PPF_HIP(dev_free(out_));
            out_ = nullptr;
        }
        if (h_out_) {
            PPF_HIP(host_free(h_out_));
            h_out_ = nullptr;
        }
        PPF_HIP(dev_alloc(&out_, words * sizeof(uint32_t)));
        PPF_HIP(host_alloc(reinterpret_cast<void **>(&h_out_), words * sizeof(uint32_t)));
        out_cap_ = words;
    }
    ppf_static_mesh none{};
    const ppf_static_mesh &S = sm ? *sm : none;
    PrimSet f, e, v;
    f.n = m.nf, f.arity = 3, f.elem = m.face, f.ghat = m.f_ghat, f.offset = m.f_off;
    e.n = m.ne, e.arity = 2, e.elem = m.edge, e.ghat = m.e_ghat, e.offset = m.e_off;
    v.n = m.nv, v.arity = 1, v.elem = nullptr, v.ghat = m.v_ghat, v.offset = m.v_off;
    f.vscene = e.vscene = v.vscene = m.vscene;
    f.nscene = e.nscene = v.nscene = m.nscene;
    const bool fresh = rebuild || face_.view().n != m.nf || edge_.view().n != m.ne ||
                       vert_.view().n != m.nv;
    int rc;
    if (fresh) {
        rc = face_.build(f, m.x, lp.x1, lp.max_t, s);
        if (rc == PPF_OK) rc = edge_.build(e, m.x, lp.x1, lp.max_t, s);
        if (rc == PPF_OK) rc = vert_.build(v, m.x, lp.x1, lp.max_t, s);
    } else {
        rc = face_.refit(f, m.x, lp.x1, lp.max_t, s);
        if (rc == PPF_OK) rc = edge_.refit(e, m.x, lp.x1, lp.max_t, s);
        if (rc == PPF_OK) rc = vert_.refit(v, m.x, lp.x1, lp.max_t, s);
    }
    if (rc != PPF_OK) {
        return rc;
    }
    if (S.nv && (rebuild || sface_.view().n != S.nf || sedge_.view().n != S.ne ||
                 static_x_ != S.x)) {
        PrimSet sf, se;
        sf.n = S.nf, sf.arity = 3, sf.elem = S.face, sf.ghat = S.f_ghat, sf.offset = S.f_off;
        se.n = S.ne, se.arity = 2, se.elem = S.edge, se.ghat = S.e_ghat, se.offset = S.e_off;
        sf.vscene = se.vscene = S.vscene;
        sf.nscene = se.nscene = m.nscene;
        rc = sface_.build(sf, S.x, nullptr, 0.0f, s);
        if (rc == PPF_OK) rc = sedge_.build(se, S.x, nullptr, 0.0f, s);
        if (rc != PPF_OK) {
            return rc;
        }
        static_x_ = S.x;
    }
    // minima start at max_t (per scene too), flags at zero
    const uint32_t tmax = __builtin_bit_cast(uint32_t, lp.max_t);
    std::vector<uint32_t> init(words, 0u);
    for (uint32_t k = 0; k < words; ++k) {
        init[k] = (k < 2 || (k >= 5 && k < 5 + 2 * ns)) ? tmax : 0u;
    }
    init[4] = 0xffffffffu;
    PPF_HIP(hipMemcpyAsync(out_, init.data(), words * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    LsCtx c;
    c.m = m;
    c.sm = S;
    c.lp = lp;
    c.fb = face_.view();
    c.eb = edge_.view();
    c.vb = vert_.view();
    c.sfb = S.nv ? sface_.view() : BvhView{};
    c.seb = S.nv ? sedge_.view() : BvhView{};
    c.out = out_;
    c.sout = ns ? out_ + 5 : nullptr;
    c.ns = ns;
    c.stats = walk_stats_device();
    auto sweeps = [&](auto stack) -> int {
        constexpr int D = decltype(stack)::value;
        if (m.nv) {
            WalkClock clk(kWalkLsVertex, s);
            k_ls_vertex<D><<<xcd_grid(cdiv(m.nv, kWQ)), kWQ, 0, s>>>(c);
            PPF_HIP(hipGetLastError());
            clk.stop();
        }
        if (S.nv && m.nf) {
            k_ls_static_vertex<D><<<xcd_grid(cdiv(S.nv, kWQ)), kWQ, 0, s>>>(c);
            PPF_HIP(hipGetLastError());
        }
        if (m.nv && (lp.nsphere || lp.nfloor)) { // before the edges: it seeds them
            k_ls_analytic<<<cdiv(m.nv, 256), 256, 0, s>>>(c);
            PPF_HIP(hipGetLastError());
        }
        if (m.ne) {
            WalkClock clk(kWalkLsEdge, s);
            k_ls_edge<D><<<xcd_grid(cdiv(m.ne, kWQ)), kWQ, 0, s>>>(c);
            PPF_HIP(hipGetLastError());
            clk.stop();
        }
        PPF_HIP(hipMemcpyAsync(h_out_, out_, words * sizeof(uint32_t), hipMemcpyDeviceToHost,
                               s));
        PPF_HIP(hipStreamSynchronize(s));
        return PPF_OK;
    };
    rc = sweeps(std::integral_constant<int, kShortStack>{});
    if (rc == PPF_OK && h_out_[3]) {
        // a walk overflowed the short stacks: every sweep again from the start
        PPF_HIP(hipMemcpyAsync(out_, init.data(), words * sizeof(uint32_t),
                               hipMemcpyHostToDevice, s));
        rc = sweeps(std::integral_constant<int, kStack>{});
    }
    if (rc != PPF_OK) {
        return rc;
    }
    if (h_out_[3]) {
        set_error("line search: a BVH traversal exceeded its %d-entry stack", kStack);
        return PPF_ERR_CAPACITY;
    }
    const float tvf = __builtin_bit_cast(float, h_out_[0]);
    const float tee = __builtin_bit_cast(float, h_out_[1]);
    *toi = fminf(tvf, tee) / lp.max_t;
    *overlap = h_out_[2] ? 1 : 0;
    for (uint32_t c = 0; c < ns; ++c) {
        const float svf = __builtin_bit_cast(float, h_out_[5 + c]);
        const float see = __builtin_bit_cast(float, h_out_[5 + ns + c]);
        if (scene_toi) scene_toi[c] = fminf(svf, see) / lp.max_t;
        if (scene_overlap) scene_overlap[c] = h_out_[5 + 2 * ns + c] ? 1 : 0;
    }
    if (pin_bad) {
        *pin_bad = h_out_[4];
    }
    return PPF_OK;
}

} // namespace ppf
