// newton.hip -- one time step of the contact solver on MI355X.
//
// The Newton loop of main_helper::advance (main/main.cu:267-1649) for
// shells with fix pins, self contact and a static collision mesh:
//   target  = x + (x - x_prev) dt / dt_prev + dt^2 g (fix position if pinned)
//                                                        main.cu:491-511
//   loop: inertia + elastic assembly; contact pairs at the iterate, contact
//         and collision-mesh barrier + friction with the elastic blocks as
//         the stiffness reference; Dirichlet elimination of pinned vertices
//         (main.cu:954-1049); block-Jacobi PCG on the merged upper matrix;
//         step clamped by max_dx; CCD line search from the previous iterate
//         (line_search_max_t sweep); toi accumulated until target_toi, then
//         one error-reduction iteration at dt * toi_advanced
//                                                        main.cu:683-1514
//   commit: intersection check, prev <- curr, curr <- iterate
//                                                        main.cu:1515-1600
// Every stage is a kernel sequence of this library; the host only reads the
// scalars that steer the loop (pair counts, toi, PCG status).
#include "newton.hpp"
#include "scan.hpp"
#include "sort.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

namespace ppf {

namespace {

constexpr int kNT = 256;

uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

__device__ inline f3 X(const float *p, uint32_t i) { return load3(p + 3 * (size_t)i); }

// Scene of vertex i (0 for a one-scene mesh).
__device__ inline uint32_t vsc(const uint32_t *vs, uint32_t i) { return vs ? vs[i] : 0u; }

// Per-scene step scalars on the device, ns of each: dt, prev_dt, recale (the
// max_dx clamp), toi (of the line search) and back (the pin rewind), as
// consecutive float arrays; max |dx| bits, active flags and pair counts as
// consecutive word arrays. A one-scene step uses the same arrays with ns = 1.
enum { kDt = 0, kPrevDt = 1, kRecale = 2, kToi = 3, kBack = 4, kSceneFloats = 5 };
enum { kMaxNorm = 0, kActive = 1, kPairs = 2, kSlToi = 3, kSceneWords = 4 };

// Inertial target (main.cu:491-511), with each scene's dt and prev_dt.
__global__ void k_target(uint32_t n, const float *__restrict__ curr,
                         const float *__restrict__ prev, const uint32_t *__restrict__ vflag,
                         const float *__restrict__ fix, const uint32_t *__restrict__ vs,
                         const float *__restrict__ sf, uint32_t ns, f3 g,
                         float *__restrict__ target) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i >= n) {
        return;
    }
    if ((vflag[i] & 1u) && fix) {
        store3(target + 3 * (size_t)i, X(fix, i));
        return;
    }
    const uint32_t c = vsc(vs, i);
    const float dt = sf[kDt * ns + c], prev_dt = sf[kPrevDt * ns + c];
    const f3 x1 = X(curr, i), x0 = X(prev, i);
    const float tr = dt / prev_dt, h2 = dt * dt;
    const f3 d = x1 - x0;
    const f3 y = mk3(d.x * tr, d.y * tr, d.z * tr) + h2 * g;
    store3(target + 3 * (size_t)i, x1 + y);
}

// Walk pins back to the fraction of their step the loop integrated
// (main.cu:531-545): fix -= back * delta (a static pin's delta is zero), for
// the scenes whose back is non-zero.
__global__ void k_rewind(uint32_t n, const uint32_t *__restrict__ vflag, float *__restrict__ fix,
                         const float *__restrict__ delta, const uint32_t *__restrict__ vs,
                         const float *__restrict__ sf, uint32_t ns) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i >= n || !(vflag[i] & 1u)) {
        return;
    }
    const float back = sf[kBack * ns + vsc(vs, i)];
    if (back == 0.0f) {
        return;
    }
    const f3 d = X(delta, i);
    store3(fix + 3 * (size_t)i, X(fix, i) - mk3(back * d.x, back * d.y, back * d.z));
}

// dx = eval_x - target on pinned vertices, 0 elsewhere (main.cu:788-801).
__global__ void k_dx_init(uint32_t n, const uint32_t *__restrict__ mask,
                          const float *__restrict__ ex, const float *__restrict__ target,
                          float *__restrict__ dx) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i >= n) {
        return;
    }
    const f3 d = mask[i] ? X(ex, i) - X(target, i) : mk3(0.0f, 0.0f, 0.0f);
    store3(dx + 3 * (size_t)i, d);
}

__global__ void k_mask(uint32_t n, const uint32_t *__restrict__ vflag, const float *fix,
                       uint32_t *__restrict__ mask) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i < n) {
        mask[i] = (fix && (vflag[i] & 1u)) ? 1u : 0u;
    }
}

__global__ void k_add(uint64_t n, const float *__restrict__ a, const float *__restrict__ b,
                      float *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * kNT + threadIdx.x;
    if (i < n) {
        out[i] = a[i] + b[i];
    }
}

// Merged upper block CSR of the Newton system, A = elastic pattern, B =
// contact blocks (A + B where both hold a column), with the Dirichlet
// elimination of pinned vertices (main.cu:954-1049) folded in: a block that
// touches a pinned vertex is not kept (the reference zeroes it in place; the
// solve then reads no all-zero blocks and a pinned row is its diagonal only),
// and its coupling is lifted into the free row's right-hand side:
//   row i free, column j pinned:  f_i -= B_ij p_j       (own lift)
//   row i pinned, column j free:  f_j -= B_ij^T p_i     (mirror lift)
// with p = eval_x - target on the pinned vertex. The reference scatters the
// lifts with float atomics; here a free row takes its own lifts in entry
// order, then its mirror lifts in (source row, entry) order, so every run
// sums in the same order. mask == null: no pins, nothing is eliminated.
__device__ inline bool pinned(const uint32_t *mask, uint32_t v) {
    return mask && mask[v] != 0u;
}

__global__ void k_merge_count(uint32_t n, const uint32_t *__restrict__ rpa,
                              const uint32_t *__restrict__ cia,
                              const uint32_t *__restrict__ rpb,
                              const uint32_t *__restrict__ cib,
                              const uint32_t *__restrict__ mask, uint32_t *__restrict__ cnt,
                              uint32_t *__restrict__ mcnt) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i >= n) {
        return;
    }
    const bool mi = pinned(mask, i);
    uint32_t a = rpa[i], ae = rpa[i + 1], b = rpb[i], be = rpb[i + 1], c = 0, m = 0;
    while (a < ae || b < be) {
        uint32_t j;
        if (b >= be || (a < ae && cia[a] < cib[b])) {
            j = cia[a++];
        } else if (a >= ae || cib[b] < cia[a]) {
            j = cib[b++];
        } else {
            j = cia[a];
            ++a;
            ++b;
        }
        const bool mj = pinned(mask, j);
        c += (!mi && !mj) ? 1u : 0u;
        m += (mi && !mj) ? 1u : 0u;
    }
    cnt[i] = c;
    if (mcnt) {
        mcnt[i] = m;
    }
}

// Kept blocks, own lifts, and one record per mirror lift: the free row it
// goes to (key), and the lifted vector B_ij^T p_i (3 floats at rec[3 l]).
__global__ void k_merge_fill(uint32_t n, const uint32_t *__restrict__ rpa,
                             const uint32_t *__restrict__ cia, const float *__restrict__ va,
                             const uint32_t *__restrict__ rpb, const uint32_t *__restrict__ cib,
                             const float *__restrict__ vb, const uint32_t *__restrict__ mask,
                             const float *__restrict__ ex, const float *__restrict__ target,
                             const uint32_t *__restrict__ rpm, uint32_t *__restrict__ cim,
                             float *__restrict__ vm, const uint32_t *__restrict__ lrp,
                             uint32_t *__restrict__ lkey, uint32_t *__restrict__ lidx,
                             float *__restrict__ lrec, float *__restrict__ force) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i >= n) {
        return;
    }
    const bool mi = pinned(mask, i);
    uint32_t a = rpa[i], ae = rpa[i + 1], b = rpb[i], be = rpb[i + 1], o = rpm[i];
    uint32_t l = mi ? lrp[i] : 0u;
    f3 pi = mk3(0.0f, 0.0f, 0.0f);
    if (mi) {
        pi = X(ex, i) - X(target, i);
    }
    float f[3] = {0.0f, 0.0f, 0.0f};
    bool lifted = false;
    if (mask && !mi) {
        const f3 fi = X(force, i);
        f[0] = fi.x, f[1] = fi.y, f[2] = fi.z;
    }
    while (a < ae || b < be) {
        float blk[9];
        uint32_t j;
        if (b >= be || (a < ae && cia[a] < cib[b])) {
            j = cia[a];
            for (int e = 0; e < 9; ++e) blk[e] = va[9 * (size_t)a + e];
            ++a;
        } else if (a >= ae || cib[b] < cia[a]) {
            j = cib[b];
            for (int e = 0; e < 9; ++e) blk[e] = vb[9 * (size_t)b + e];
            ++b;
        } else {
            j = cia[a];
            for (int e = 0; e < 9; ++e) blk[e] = va[9 * (size_t)a + e] + vb[9 * (size_t)b + e];
            ++a;
            ++b;
        }
        const bool mj = pinned(mask, j);
        if (!mi && !mj) {
            cim[o] = j;
            float *dst = vm + 9 * (size_t)o;
            for (int e = 0; e < 9; ++e) dst[e] = blk[e];
            ++o;
        } else if (!mi) { // free row, pinned column: f_i -= B_ij p_j
            const f3 p = X(ex, j) - X(target, j);
            for (int r = 0; r < 3; ++r) {
                const float c = (blk[r] * p.x + blk[r + 3] * p.y) + blk[r + 6] * p.z;
                f[r] += -c;
            }
            lifted = true;
        } else if (!mj) { // pinned row, free column: f_j -= B_ij^T p_i, later
            float *rec = lrec + 3 * (size_t)l;
            for (int r = 0; r < 3; ++r) {
                rec[r] = (blk[3 * r] * pi.x + blk[3 * r + 1] * pi.y) + blk[3 * r + 2] * pi.z;
            }
            lkey[l] = j;
            lidx[l] = l;
            ++l;
        }
    }
    if (lifted) {
        store3(force + 3 * (size_t)i, mk3(f[0], f[1], f[2]));
    }
}

// Mirror lifts, records sorted (stably) by the free row they go to: the
// first record of each row's run subtracts the run in order.
__global__ void k_mirror_lift(uint32_t nl, const uint32_t *__restrict__ lkey,
                              const uint32_t *__restrict__ lidx,
                              const float *__restrict__ lrec, float *__restrict__ force) {
    const uint32_t t = blockIdx.x * kNT + threadIdx.x;
    if (t >= nl || (t > 0 && lkey[t - 1] == lkey[t])) {
        return;
    }
    const uint32_t row = lkey[t];
    const f3 fi = X(force, row);
    float f[3] = {fi.x, fi.y, fi.z};
    for (uint32_t u = t; u < nl && lkey[u] == row; ++u) {
        const float *rec = lrec + 3 * (size_t)lidx[u];
        for (int r = 0; r < 3; ++r) {
            f[r] += -rec[r];
        }
    }
    store3(force + 3 * (size_t)row, mk3(f[0], f[1], f[2]));
}

__global__ void k_dirichlet_diag(uint32_t n, const uint32_t *__restrict__ mask,
                                 const float *__restrict__ ex,
                                 const float *__restrict__ target, float *__restrict__ cdiag,
                                 float *__restrict__ force) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i >= n || !mask[i]) {
        return;
    }
    float *c = cdiag + 9 * (size_t)i;
    for (int e = 0; e < 9; ++e) c[e] = (e % 4 == 0) ? 1.0f : 0.0f;
    store3(force + 3 * (size_t)i, X(ex, i) - X(target, i));
}

// max_i |dx_i| per scene as float bits (non-negative, so the integer order is
// the float order). A wave whose vertices share one scene adds one atomic.
__global__ void k_max_norm(uint32_t n, const float *__restrict__ dx,
                           const uint32_t *__restrict__ vs, uint32_t *__restrict__ out) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    float v = 0.0f;
    uint32_t c = 0;
    if (i < n) {
        const f3 d = X(dx, i);
        v = sqrtf(dot3(d, d));
        c = vsc(vs, i);
    }
    const uint32_t c0 = __shfl(c, 0);
    if (__ballot(c != c0 && i < n) == 0) {
        for (int o = 32; o > 0; o >>= 1) {
            v = fmaxf(v, __shfl_xor(v, o));
        }
        if ((threadIdx.x & 63) == 0 && v > 0.0f) {
            atomicMax(out + c0, __float_as_uint(v));
        }
    } else if (i < n && v > 0.0f) {
        atomicMax(out + c, __float_as_uint(v));
    }
}

// recale = min(1, max_dx / max |dx|) per scene (main.cu:1177-1180).
__global__ void k_recale(uint32_t ns, float max_dx, const uint32_t *__restrict__ su,
                         float *__restrict__ sf) {
    const uint32_t c = blockIdx.x * kNT + threadIdx.x;
    if (c < ns) {
        sf[kRecale * ns + c] = fminf(1.0f, max_dx / __uint_as_float(su[kMaxNorm * ns + c]));
    }
}

// target <- eval_x; eval_x <- eval_x - recale * dx (main.cu:1177-1187), for
// the scenes still iterating; a finished scene keeps eval_x and gets
// target = eval_x (its sweep is then empty).
__global__ void k_step(uint64_t n3, const uint32_t *__restrict__ vs,
                       const float *__restrict__ sf, const uint32_t *__restrict__ su,
                       uint32_t ns, const float *__restrict__ dx, float *__restrict__ ex,
                       float *__restrict__ target) {
    const uint64_t i = (uint64_t)blockIdx.x * kNT + threadIdx.x;
    if (i < n3) {
        const uint32_t c = vsc(vs, (uint32_t)(i / 3));
        const float x = ex[i];
        target[i] = x;
        if (su[kActive * ns + c]) {
            ex[i] = x - sf[kRecale * ns + c] * dx[i];
        }
    }
}

// eval_x <- target + toi (eval_x - target) (main.cu:1322-1329), per scene
// still iterating.
__global__ void k_shorten(uint64_t n3, const uint32_t *__restrict__ vs,
                          const float *__restrict__ sf, const uint32_t *__restrict__ su,
                          uint32_t ns, const float *__restrict__ target,
                          float *__restrict__ ex) {
    const uint64_t i = (uint64_t)blockIdx.x * kNT + threadIdx.x;
    if (i < n3) {
        const uint32_t c = vsc(vs, (uint32_t)(i / 3));
        if (su[kActive * ns + c]) {
            const float t = target[i];
            ex[i] = t + sf[kToi * ns + c] * (ex[i] - t);
        }
    }
}

// A finished scene takes no Newton step: zero right-hand side and start, so
// its solve ends at once.
__global__ void k_freeze(uint32_t n, const uint32_t *__restrict__ vs,
                         const uint32_t *__restrict__ su, uint32_t ns,
                         float *__restrict__ force, float *__restrict__ dx) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i < n && !su[kActive * ns + vsc(vs, i)]) {
        store3(force + 3 * (size_t)i, mk3(0.0f, 0.0f, 0.0f));
        store3(dx + 3 * (size_t)i, mk3(0.0f, 0.0f, 0.0f));
    }
}

// Contact pairs per scene (the scene of the pair's query element).
__global__ void k_scene_pairs(uint64_t np, const ppf_pair *__restrict__ pairs,
                              const uint32_t *__restrict__ edge,
                              const uint32_t *__restrict__ vs,
                              const uint32_t *__restrict__ svs, uint32_t *__restrict__ cnt) {
    const uint64_t t = (uint64_t)blockIdx.x * kNT + threadIdx.x;
    if (t >= np) {
        return;
    }
    const ppf_pair q = pairs[t];
    const uint32_t kind = q.kind & 255u;
    uint32_t c;
    if (kind == 3 || kind == 6) {
        c = vs[edge[2 * (size_t)q.a]];
    } else if (kind == 5) {
        c = svs ? svs[q.a] : 0u;
    } else {
        c = vs[q.a];
    }
    atomicAdd(cnt + c, 1u);
}

// prev <- curr, curr <- eval_x for the scenes that commit their step.
__global__ void k_commit(uint32_t n, const uint32_t *__restrict__ vs,
                         const uint32_t *__restrict__ ok, float *__restrict__ prev,
                         float *__restrict__ curr, const float *__restrict__ ex) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i < n && ok[vs[i]]) {
        store3(prev + 3 * (size_t)i, X(curr, i));
        store3(curr + 3 * (size_t)i, X(ex, i));
    }
}

} // namespace

Stepper::~Stepper() { release(); }

void Stepper::release() {
    for (void *p : {(void *)ex_, (void *)target_, (void *)dx_, (void *)force_, (void *)cdiag_,
                    (void *)eval_, (void *)cgrad_, (void *)mask_, (void *)scal_,
                    (void *)pairs_, (void *)mrp_, (void *)mci_, (void *)mcnt_,
                    (void *)scan_, (void *)mval_, (void *)order_, (void *)lcnt_,
                    (void *)lrp_, (void *)lbuf_, (void *)sf_, (void *)su_, (void *)sph_,
                    (void *)flr_}) {
        if (p) {
            (void)dev_free(p);
        }
    }
    if (h_scal_) {
        (void)host_free(h_scal_);
    }
    if (h_sf_) {
        (void)host_free(h_sf_);
    }
    if (h_su_) {
        (void)host_free(h_su_);
    }
    sf_ = h_sf_ = nullptr;
    su_ = h_su_ = nullptr;
    sph_ = nullptr;
    flr_ = nullptr;
    sph_cap_ = flr_cap_ = nsph_ = nflr_ = 0;
    ns_ = 1;
    vscene_ = nullptr;
    scene_row_.clear();
    if (plan_) {
        pcg_plan_destroy(plan_);
        plan_ = nullptr;
    }
    for (hipEvent_t &e : ev_) {
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
    }
    ex_ = target_ = dx_ = force_ = cdiag_ = eval_ = cgrad_ = nullptr;
    mask_ = scal_ = h_scal_ = nullptr;
    order_ = nullptr;
    order_changed_ = false;
    pairs_ = nullptr;
    mrp_ = mci_ = mcnt_ = scan_ = nullptr;
    lcnt_ = lrp_ = lbuf_ = nullptr;
    mval_ = nullptr;
    pair_cap_ = mcap_ = lcap_ = 0;
}

namespace {

uint32_t spread10(uint32_t v) { // 10 bits -> every third bit of 30
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

uint32_t spread16(uint32_t v) { // 16 bits -> every other bit of 32
    v &= 0xffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

// Row order of the solve: mesh pieces (connected by faces) one after the
// other, each in the Morton order of its vertices' initial positions within
// the piece's own box, so that a vertex's mesh neighbours are near it in the
// order (one Morton curve over all vertices would interleave stacked
// sheets). An axis along which a piece is thinner than 1% of its largest
// extent is left out (a sheet gets a 2D curve: its noise across the sheet
// would otherwise scramble the order).
void piece_morton_order(uint32_t nv, const std::vector<float> &x,
                        const std::vector<uint32_t> &face,
                        std::vector<uint32_t> &order) {
    std::vector<uint32_t> up(nv);
    for (uint32_t v = 0; v < nv; ++v) {
        up[v] = v;
    }
    auto find = [&](uint32_t v) {
        while (up[v] != v) {
            up[v] = up[up[v]];
            v = up[v];
        }
        return v;
    };
    for (size_t f = 0; f + 2 < face.size(); f += 3) {
        if (face[f] >= nv || face[f + 1] >= nv || face[f + 2] >= nv) {
            continue; // malformed faces are the assembly's to reject
        }
        for (int k = 1; k < 3; ++k) {
            const uint32_t a = find(face[f]), b = find(face[f + k]);
            if (a != b) {
                up[std::max(a, b)] = std::min(a, b); // a piece is named by its lowest vertex
            }
        }
    }
    std::vector<uint32_t> piece(nv);
    std::vector<float> box(6 * (size_t)nv);
    for (uint32_t v = 0; v < nv; ++v) {
        piece[v] = find(v);
        if (piece[v] == v) {
            for (int d = 0; d < 3; ++d) {
                box[6 * (size_t)v + d] = FLT_MAX;
                box[6 * (size_t)v + 3 + d] = -FLT_MAX;
            }
        }
    }
    for (uint32_t v = 0; v < nv; ++v) {
        float *b = &box[6 * (size_t)piece[v]];
        for (int d = 0; d < 3; ++d) {
            b[d] = std::min(b[d], x[3 * (size_t)v + d]);
            b[3 + d] = std::max(b[3 + d], x[3 * (size_t)v + d]);
        }
    }
    std::vector<uint64_t> key(nv);
    for (uint32_t v = 0; v < nv; ++v) {
        const float *b = &box[6 * (size_t)piece[v]];
        float ext[3], big = 0.0f;
        for (int d = 0; d < 3; ++d) {
            ext[d] = b[3 + d] - b[d];
            big = std::max(big, ext[d]);
        }
        int axes[3], na = 0;
        for (int d = 0; d < 3; ++d) {
            if (ext[d] > 0.01f * big) {
                axes[na++] = d;
            }
        }
        uint32_t code = 0;
        const uint32_t cells = na == 3 ? 1024u : 65536u;
        for (int k = 0; k < na; ++k) {
            const int d = axes[k];
            const float t = (x[3 * (size_t)v + d] - b[d]) / ext[d];
            const uint32_t q = (uint32_t)std::min((float)(cells - 1),
                                                  std::max(0.0f, t * (float)cells));
            code |= na == 3 ? spread10(q) << (2 - k)
                            : (na == 2 ? spread16(q) << (1 - k) : q);
        }
        key[v] = ((uint64_t)piece[v] << 32) | code;
    }
    order.resize(nv);
    for (uint32_t v = 0; v < nv; ++v) {
        order[v] = v;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
}

} // namespace

int Stepper::create(const ppf_scene &sc, hipStream_t s) {
    release();
    nv_ = sc.mesh.nv, ne_ = sc.mesh.ne, nf_ = sc.mesh.nf, nh_ = sc.shell.nh;
    if (sc.shell.nv != nv_ || sc.shell.nf != nf_) {
        set_error("stepper: shell and contact mesh sizes differ");
        return PPF_ERR_ARG;
    }
    int rc = elastic_.create(sc.shell, s);
    if (rc != PPF_OK) {
        return rc;
    }
    const size_t n3 = 3 * (size_t)nv_;
    PPF_HIP(dev_alloc(&ex_, n3 * sizeof(float)));
    PPF_HIP(dev_alloc(&target_, n3 * sizeof(float)));
    PPF_HIP(dev_alloc(&dx_, n3 * sizeof(float)));
    PPF_HIP(dev_alloc(&force_, n3 * sizeof(float)));
    PPF_HIP(dev_alloc(&cgrad_, n3 * sizeof(float)));
    PPF_HIP(dev_alloc(&cdiag_, 9 * (size_t)nv_ * sizeof(float)));
    PPF_HIP(dev_alloc(&eval_, 9 * ((size_t)elastic_.nnz() + 1) * sizeof(float)));
    PPF_HIP(dev_alloc(&mask_, ((size_t)nv_ + 1) * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&order_, ((size_t)nv_ + 1) * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&scal_, 4 * sizeof(uint32_t)));
    PPF_HIP(host_alloc(reinterpret_cast<void **>(&h_scal_), 4 * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&mrp_, ((size_t)nv_ + 1) * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&mcnt_, ((size_t)nv_ + 1) * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&lcnt_, ((size_t)nv_ + 1) * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&lrp_, ((size_t)nv_ + 1) * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&scan_, scan_scratch_words(nv_ + 1) * sizeof(uint32_t)));
    for (hipEvent_t &e : ev_) {
        PPF_HIP(hipEventCreate(&e));
    }
    if (sc.mesh.vscene && sc.mesh.nscene > 1) {
        // Many scenes: each scene's vertices contiguous, ids rising from 0.
        ns_ = sc.mesh.nscene;
        vscene_ = sc.mesh.vscene;
        std::vector<uint32_t> hv(nv_);
        PPF_HIP(hipMemcpyAsync(hv.data(), vscene_, (size_t)nv_ * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, s));
        PPF_HIP(hipStreamSynchronize(s));
        for (uint32_t v = 0; v < nv_; ++v) {
            if (hv[v] >= ns_ || (v > 0 && hv[v] < hv[v - 1])) {
                set_error("stepper: vertex scene ids must rise from 0 to %u (vertex %u has %u)",
                          ns_ - 1, v, hv[v]);
                return PPF_ERR_ARG;
            }
        }
        // first vertex of each scene (an empty scene starts where the next does)
        scene_row_.assign(ns_ + 1, nv_);
        for (uint32_t v = nv_; v-- > 0;) {
            scene_row_[hv[v]] = v;
        }
        for (uint32_t c = ns_; c-- > 0;) {
            scene_row_[c] = std::min(scene_row_[c], scene_row_[c + 1]);
        }
    }
    PPF_HIP(dev_alloc(&sf_, 5 * (size_t)ns_ * sizeof(float)));
    PPF_HIP(dev_alloc(&su_, kSceneWords * (size_t)ns_ * sizeof(uint32_t)));
    PPF_HIP(host_alloc(reinterpret_cast<void **>(&h_sf_), 5 * (size_t)ns_ * sizeof(float)));
    PPF_HIP(host_alloc(reinterpret_cast<void **>(&h_su_),
                       kSceneWords * (size_t)ns_ * sizeof(uint32_t)));
    if (nv_ > 1 && ns_ == 1) {
        std::vector<float> hx(n3);
        std::vector<uint32_t> hf(3 * (size_t)nf_), order;
        PPF_HIP(hipMemcpyAsync(hx.data(), sc.mesh.x, n3 * sizeof(float), hipMemcpyDeviceToHost,
                               s));
        if (nf_) {
            PPF_HIP(hipMemcpyAsync(hf.data(), sc.mesh.face, hf.size() * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, s));
        }
        PPF_HIP(hipStreamSynchronize(s));
        piece_morton_order(nv_, hx, hf, order);
        PPF_HIP(hipMemcpyAsync(order_, order.data(), nv_ * sizeof(uint32_t),
                               hipMemcpyHostToDevice, s));
        PPF_HIP(hipStreamSynchronize(s)); // the host order ends here
        order_changed_ = true;
    }
    return refresh_pins(sc, s);
}

int Stepper::refresh_pins(const ppf_scene &sc, hipStream_t s) {
    k_mask<<<cdiv(nv_, kNT), kNT, 0, s>>>(nv_, sc.mesh.vflag, sc.fix_pos, mask_);
    PPF_HIP(hipGetLastError());
    // Count the pinned vertices once per pin set (Dirichlet elimination runs
    // only when there are any).
    npinned_ = 0;
    if (sc.fix_pos && nv_) {
        std::vector<uint32_t> h(nv_);
        PPF_HIP(hipMemcpyAsync(h.data(), mask_, (size_t)nv_ * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, s));
        PPF_HIP(hipStreamSynchronize(s));
        for (uint32_t v : h) {
            npinned_ += v;
        }
    }
    return PPF_OK;
}

int Stepper::reserve_pairs(uint64_t n) {
    if (n <= pair_cap_) {
        return PPF_OK;
    }
    if (pairs_) {
        PPF_HIP(dev_free(pairs_));
        pairs_ = nullptr;
    }
    const uint64_t cap = grow_cap(n);
    PPF_HIP(dev_alloc(&pairs_, cap * sizeof(ppf_pair)));
    pair_cap_ = cap;
    return PPF_OK;
}

int Stepper::reserve_merged(uint64_t nnz) {
    if (nnz <= mcap_) {
        return PPF_OK;
    }
    if (mci_) (void)dev_free(mci_);
    if (mval_) (void)dev_free(mval_);
    mci_ = nullptr;
    mval_ = nullptr;
    const uint64_t cap = grow_cap(nnz);
    PPF_HIP(dev_alloc(&mci_, cap * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&mval_, 9 * cap * sizeof(float)));
    mcap_ = cap;
    return PPF_OK;
}

// Mirror-lift records: keys, indices, 3 floats each, and the sort's scratch.
int Stepper::reserve_lifts(uint64_t nl) {
    if (nl <= lcap_) {
        return PPF_OK;
    }
    if (lbuf_) {
        (void)dev_free(lbuf_);
        lbuf_ = nullptr;
    }
    const uint64_t cap = grow_cap(nl, 256);
    PPF_HIP(dev_alloc(&lbuf_, (5 * cap + radix_sort_scratch_words(cap)) * sizeof(uint32_t)));
    lcap_ = cap;
    return PPF_OK;
}

int Stepper::newton_iteration(const ppf_scene &sc, const ppf_step_params &p, float dt,
                              bool first_iteration, ppf_step_result *out, float *toi_out,
                              float *recale_out, hipStream_t s) {
    const uint32_t n = nv_, ns = ns_;
    const size_t n3 = 3 * (size_t)n;
    const uint32_t *vs = ns > 1 ? vscene_ : nullptr;
    // Inertia + elastic at the iterate.
    PPF_HIP(hipEventRecord(ev_[0], s));
    ppf_shell sh = sc.shell;
    sh.x = ex_;
    sh.x0 = sc.mesh.x;
    sh.target = target_;
    ppf_elastic_params ep;
    ep.dt = dt;
    ep.air_density = p.air_density;
    ep.air_friction = p.air_friction;
    ep.wind[0] = p.wind[0], ep.wind[1] = p.wind[1], ep.wind[2] = p.wind[2];
    ep.scene_dt = vs ? sf_ + kDt * ns : nullptr;
    ep.vscene = vs;
    k_dx_init<<<cdiv(n, kNT), kNT, 0, s>>>(n, mask_, ex_, target_, dx_);
    PPF_HIP(hipGetLastError());
    int rc = elastic_.assemble(sh, ep, force_, cdiag_, eval_, s);
    if (rc != PPF_OK) {
        return rc;
    }
    PPF_HIP(hipEventRecord(ev_[1], s));
    // Contact pairs and their force / blocks (tmp_fixed = elastic blocks).
    ppf_contact_mesh m = sc.mesh;
    m.x = ex_;
    uint64_t npairs = 0;
    if (!p.disable_contact) {
        // Trees are built once per step and refit per iteration (the pair set
        // depends only on the boxes containing the primitives).
        rc = first_iteration ? narrow_.build_trees(m, s) : narrow_.refit_trees(m, s);
        if (rc != PPF_OK) {
            return rc;
        }
        uint64_t nself = 0, nstat = 0;
        rc = narrow_.proximity(m, pairs_, pair_cap_, &nself, s);
        if (rc != PPF_OK) {
            return rc;
        }
        if (nself > pair_cap_) {
            rc = reserve_pairs(nself);
            if (rc != PPF_OK) {
                return rc;
            }
            rc = narrow_.proximity(m, pairs_, pair_cap_, &nself, s);
            if (rc != PPF_OK) {
                return rc;
            }
        }
        if (sc.collision.nv) {
            rc = narrow_.static_proximity(m, sc.collision, sc.shell.vmass, sc.emass,
                                          sc.shell.fmass, pairs_ + nself,
                                          pair_cap_ - nself, &nstat, s);
            if (rc != PPF_OK) {
                return rc;
            }
            if (nself + nstat > pair_cap_) {
                rc = reserve_pairs(nself + nstat);
                if (rc != PPF_OK) {
                    return rc;
                }
                rc = narrow_.proximity(m, pairs_, pair_cap_, &nself, s);
                if (rc == PPF_OK) {
                    rc = narrow_.static_proximity(m, sc.collision, sc.shell.vmass, sc.emass,
                                                  sc.shell.fmass, pairs_ + nself,
                                                  pair_cap_ - nself, &nstat, s);
                }
                if (rc != PPF_OK) {
                    return rc;
                }
            }
        }
        npairs = nself + nstat;
    }
    if (vs) { // pairs per scene, read with the line search's results
        PPF_HIP(hipMemsetAsync(su_ + kPairs * ns, 0, ns * sizeof(uint32_t), s));
        if (npairs) {
            k_scene_pairs<<<cdiv(npairs, kNT), kNT, 0, s>>>(
                npairs, pairs_, sc.mesh.edge, vs, sc.collision.vscene, su_ + kPairs * ns);
            PPF_HIP(hipGetLastError());
        }
    } else if (npairs > out->pairs_max) {
        out->pairs_max = (uint32_t)npairs;
    }
    ppf_contact_params cp;
    cp.x0 = sc.mesh.x;
    cp.mass = sc.shell.vmass;
    cp.v_fric = sc.v_fric;
    cp.e_fric = sc.e_fric;
    cp.f_fric = sc.f_fric;
    cp.friction_mode = p.friction_mode;
    cp.friction_eps = p.friction_eps;
    cp.el_rp = elastic_.row_offset();
    cp.el_col = elastic_.col_index();
    cp.el_val = eval_;
    cp.sm = sc.collision.nv ? &sc.collision : nullptr;
    rc = contact_.assemble(m, cp, pairs_, npairs, cgrad_, s);
    if (rc != PPF_OK) {
        return rc;
    }
    rc = analytic_assemble(m, cp, sph_, nsph_, flr_, nflr_, p.constraint_tol, cgrad_, cdiag_, s);
    if (rc != PPF_OK) {
        return rc;
    }
    // strain limiting (main.cu:809-822): its stiffness reads the elastic
    // blocks alone, as the contact's did (tmp_fixed), then its terms join
    rc = elastic_.strain_limit(sh, p.eiganalysis_eps, force_, eval_, s);
    if (rc != PPF_OK) {
        return rc;
    }
    k_add<<<cdiv(n3, kNT), kNT, 0, s>>>(n3, force_, cgrad_, force_);
    PPF_HIP(hipGetLastError());
    PPF_HIP(hipEventRecord(ev_[2], s));
    // Merged matrix with the Dirichlet elimination folded in, then PCG.
    const uint32_t *pm = npinned_ ? mask_ : nullptr;
    k_merge_count<<<cdiv(n, kNT), kNT, 0, s>>>(n, elastic_.row_offset(), elastic_.col_index(),
                                                contact_.row_offset(), contact_.cols(), pm,
                                                mcnt_, pm ? lcnt_ : nullptr);
    PPF_HIP(hipGetLastError());
    PPF_HIP(hipMemsetAsync(mcnt_ + n, 0, sizeof(uint32_t), s));
    PPF_HIP(exclusive_scan_u32(mcnt_, mrp_, n + 1, scan_, s));
    PPF_HIP(hipMemcpyAsync(h_scal_, mrp_ + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    h_scal_[1] = 0;
    if (pm) {
        PPF_HIP(hipMemsetAsync(lcnt_ + n, 0, sizeof(uint32_t), s));
        PPF_HIP(exclusive_scan_u32(lcnt_, lrp_, n + 1, scan_, s));
        PPF_HIP(hipMemcpyAsync(h_scal_ + 1, lrp_ + n, sizeof(uint32_t), hipMemcpyDeviceToHost,
                               s));
    }
    PPF_HIP(hipStreamSynchronize(s));
    const uint32_t nnz = h_scal_[0], nlift = h_scal_[1];
    rc = reserve_merged(nnz);
    if (rc == PPF_OK) {
        rc = reserve_lifts(nlift);
    }
    if (rc != PPF_OK) {
        return rc;
    }
    sys_nnz_ = nnz;
    uint32_t *lkey = lbuf_, *lidx = lbuf_ + lcap_;
    float *lrec = reinterpret_cast<float *>(lbuf_ + 2 * lcap_);
    uint32_t *lsort = lbuf_ + 5 * lcap_;
    k_merge_fill<<<cdiv(n, kNT), kNT, 0, s>>>(
        n, elastic_.row_offset(), elastic_.col_index(), eval_, contact_.row_offset(),
        contact_.cols(), contact_.values(), pm, ex_, target_, mrp_, mci_, mval_, lrp_, lkey,
        lidx, lrec, force_);
    PPF_HIP(hipGetLastError());
    if (nlift) {
        uint32_t bits = 1;
        while (bits < 32 && (n >> bits) != 0) {
            ++bits;
        }
        PPF_HIP(radix_sort_pairs(lkey, lidx, nlift, bits, lsort, s));
        k_mirror_lift<<<cdiv(nlift, kNT), kNT, 0, s>>>(nlift, lkey, lidx, lrec, force_);
        PPF_HIP(hipGetLastError());
    }
    if (pm) {
        k_dirichlet_diag<<<cdiv(n, kNT), kNT, 0, s>>>(n, mask_, ex_, target_, cdiag_, force_);
        PPF_HIP(hipGetLastError());
    }
    if (vs) {
        k_freeze<<<cdiv(n, kNT), kNT, 0, s>>>(n, vs, su_, ns, force_, dx_);
        PPF_HIP(hipGetLastError());
    }
    // One solver plan: later iterations only swap the pattern in (in the
    // current row order). Many scenes: one plan over their block-diagonal
    // union, each scene with its own scalars and stopping rule.
    if (!plan_) {
        rc = ns > 1 ? pcg_plan_create_scenes(n, nnz, mrp_, mci_, ns, scene_row_.data(), s, &plan_)
                    : pcg_plan_create(n, nnz, mrp_, mci_, s, &plan_);
        if (rc != PPF_OK) {
            return rc;
        }
    }
    if (order_changed_) {
        rc = pcg_plan_set_order(plan_, order_, s);
        if (rc != PPF_OK) {
            return rc;
        }
        order_changed_ = false;
    }
    rc = pcg_plan_repattern(plan_, nnz, mrp_, mci_, s);
    if (rc != PPF_OK) {
        return rc;
    }
    std::vector<ppf_pcg_result> pr(ns);
    bool solved = false;
    if (lpc_ && ns == 1) {
        // learned factor; block Jacobi when it cannot be formed or the solve
        // does not reach the tolerance
        const float *ov = nullptr, *oc = nullptr, *ob = nullptr;
        rc = pcg_order_system(plan_, mval_, cdiag_, force_, &ov, &oc, &ob, s);
        if (rc == PPF_OK) {
            rc = lpc_->prepare(pcg_pattern(plan_), ov, oc, ob, s);
        }
        if (rc == PPF_OK) {
            rc = lpc_->forward(s);
        }
        if (rc == PPF_OK) {
            const PcgFactor fac = lpc_->factor();
            rc = pcg_solve(plan_, mval_, cdiag_, force_, dx_, p.cg_tol, p.cg_max_iter, nullptr,
                           pr.data(), s, &fac);
        }
        solved = rc == PPF_OK && (pr[0].status == 0 || pr[0].status == 1);
        if (!solved) {
            ++out->lpc_fallbacks;
            k_dx_init<<<cdiv(n, kNT), kNT, 0, s>>>(n, mask_, ex_, target_, dx_);
            PPF_HIP(hipGetLastError());
        }
    }
    if (!solved) {
        rc = pcg_solve(plan_, mval_, cdiag_, force_, dx_, p.cg_tol, p.cg_max_iter, nullptr,
                       pr.data(), s);
        if (rc != PPF_OK) {
            return rc;
        }
    }
    bool any_failed = false;
    for (uint32_t c = 0; c < ns; ++c) {
        if (!active_[c]) {
            continue;
        }
        out[c].pcg_iter_total += pr[c].iter;
        if (pr[c].iter > out[c].pcg_iter_max) {
            out[c].pcg_iter_max = pr[c].iter;
        }
        if (pr[c].status != 0 && pr[c].status != 1) {
            out[c].status = 1;
            any_failed = true;
        }
    }
    if (ns == 1 && any_failed) {
        return PPF_OK;
    }
    if (any_failed) { // the failed scenes take no step
        for (uint32_t c = 0; c < ns; ++c) {
            h_su_[kActive * ns + c] = (active_[c] && out[c].status == 0) ? 1u : 0u;
        }
        PPF_HIP(hipMemcpyAsync(su_ + kActive * ns, h_su_ + kActive * ns, ns * sizeof(uint32_t),
                               hipMemcpyHostToDevice, s));
    }
    PPF_HIP(hipEventRecord(ev_[3], s));
    // Clamp the step (max_dx, per scene on the device), move, CCD line
    // search, shorten.
    PPF_HIP(hipMemsetAsync(su_ + kMaxNorm * ns, 0, ns * sizeof(uint32_t), s));
    k_max_norm<<<cdiv(n, kNT), kNT, 0, s>>>(n, dx_, vs, su_ + kMaxNorm * ns);
    PPF_HIP(hipGetLastError());
    k_recale<<<cdiv(ns, kNT), kNT, 0, s>>>(ns, p.max_dx, su_, sf_);
    PPF_HIP(hipGetLastError());
    k_step<<<cdiv(n3, kNT), kNT, 0, s>>>(n3, vs, sf_, su_, ns, dx_, ex_, target_);
    PPF_HIP(hipGetLastError());
    // the recale factors (and pair counts) come back with the line search's
    // synchronization
    PPF_HIP(hipMemcpyAsync(h_sf_ + kRecale * ns, sf_ + kRecale * ns, ns * sizeof(float),
                           hipMemcpyDeviceToHost, s));
    if (vs) {
        PPF_HIP(hipMemcpyAsync(h_su_ + kPairs * ns, su_ + kPairs * ns, ns * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, s));
    }
    for (uint32_t c = 0; c < ns; ++c) {
        toi_out[c] = 1.0f;
    }
    bool overlap_any = false;
    if (!p.disable_contact) {
        ppf_contact_mesh lm = sc.mesh;
        lm.x = target_;
        ppf_ls_params lp;
        lp.x1 = ex_;
        lp.vmass = sc.shell.vmass;
        lp.emass = sc.emass;
        lp.fmass = sc.shell.fmass;
        lp.max_t = p.line_search_max_t;
        lp.ccd_eps = p.ccd_eps;
        lp.spheres = sph_;
        lp.nsphere = nsph_;
        lp.floors = flr_;
        lp.nfloor = nflr_;
        int overlap = 0;
        float t = 1.0f;
        uint32_t pin_bad = 0xffffffffu;
        std::vector<int> sovl(ns, 0);
        rc = ls_.run(lm, sc.collision.nv ? &sc.collision : nullptr, lp, first_iteration, &t,
                     &overlap, s, vs ? toi_out : nullptr, vs ? sovl.data() : nullptr, &pin_bad);
        if (rc != PPF_OK) {
            return rc;
        }
        if (pin_bad != 0xffffffffu) {
            // a pinned vertex is driven into a static analytic collider it
            // cannot yield to: the step fails (main.cu:1240-1271); analytic
            // colliders come with one scene only
            out[0].status = 6;
            return PPF_OK;
        }
        if (!vs) {
            toi_out[0] = fminf(1.0f, t);
            sovl[0] = overlap;
        }
        for (uint32_t c = 0; c < ns; ++c) {
            if (active_[c] && out[c].status == 0 && sovl[c]) {
                out[c].status = 3;
                overlap_any = true;
            }
        }
    }
    const bool sl = sc.shell.f_strainlimit != nullptr;
    if (sl) {
        // the strain-limit line search over the same step joins the minimum
        // (main.cu:1321-1343)
        const uint32_t tb = __builtin_bit_cast(uint32_t, p.line_search_max_t);
        for (uint32_t c = 0; c < ns; ++c) {
            h_su_[kSlToi * ns + c] = tb;
        }
        PPF_HIP(hipMemcpyAsync(su_ + kSlToi * ns, h_su_ + kSlToi * ns, ns * sizeof(uint32_t),
                               hipMemcpyHostToDevice, s));
        rc = elastic_.strain_limit_toi(sc.shell, target_, ex_, p.line_search_max_t, vs,
                                       su_ + kSlToi * ns, s);
        if (rc != PPF_OK) {
            return rc;
        }
        PPF_HIP(hipMemcpyAsync(h_su_ + kSlToi * ns, su_ + kSlToi * ns, ns * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, s));
    }
    PPF_HIP(hipEventRecord(ev_[4], s));
    PPF_HIP(hipEventSynchronize(ev_[4]));
    if (sl) {
        for (uint32_t c = 0; c < ns; ++c) {
            const float t = __builtin_bit_cast(float, h_su_[kSlToi * ns + c]);
            toi_out[c] = fminf(toi_out[c], t / p.line_search_max_t);
        }
    }
    float ms;
    PPF_HIP(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
    const float ms_el = ms;
    PPF_HIP(hipEventElapsedTime(&ms, ev_[1], ev_[2]));
    const float ms_co = ms;
    PPF_HIP(hipEventElapsedTime(&ms, ev_[2], ev_[3]));
    const float ms_so = ms;
    PPF_HIP(hipEventElapsedTime(&ms, ev_[3], ev_[4]));
    const float ms_ls = ms;
    bool any_short = false;
    for (uint32_t c = 0; c < ns; ++c) {
        if (!active_[c]) {
            continue;
        }
        out[c].ms_elastic += ms_el;
        out[c].ms_contact += ms_co;
        out[c].ms_solve += ms_so;
        out[c].ms_line_search += ms_ls;
        if (vs && h_su_[kPairs * ns + c] > out[c].pairs_max) {
            out[c].pairs_max = h_su_[kPairs * ns + c];
        }
        recale_out[c] = h_sf_[kRecale * ns + c];
        if (out[c].status == 0 && toi_out[c] <= FLT_EPSILON) {
            out[c].status = 2;
        }
        if (out[c].status == 0) {
            h_sf_[kToi * ns + c] = toi_out[c];
            any_short = true;
        }
    }
    if (ns == 1 && (overlap_any || out->status != 0)) {
        return PPF_OK;
    }
    if (any_short) {
        // the scenes that failed in this iteration take no shortened step
        for (uint32_t c = 0; c < ns; ++c) {
            h_su_[kActive * ns + c] = (active_[c] && out[c].status == 0) ? 1u : 0u;
        }
        PPF_HIP(hipMemcpyAsync(sf_ + kToi * ns, h_sf_ + kToi * ns, ns * sizeof(float),
                               hipMemcpyHostToDevice, s));
        PPF_HIP(hipMemcpyAsync(su_ + kActive * ns, h_su_ + kActive * ns, ns * sizeof(uint32_t),
                               hipMemcpyHostToDevice, s));
        k_shorten<<<cdiv(n3, kNT), kNT, 0, s>>>(n3, vs, sf_, su_, ns, target_, ex_);
        PPF_HIP(hipGetLastError());
    }
    return PPF_OK;
}

uint32_t Stepper::records(ppf_isect_record *out, uint32_t cap) const {
    const uint32_t keep = std::min(std::min(cap, kMaxRecords), records_count_);
    for (uint32_t k = 0; out && k < keep; ++k) {
        out[k] = records_[k];
    }
    return records_count_;
}

int Stepper::advance(const ppf_scene &sc, const ppf_step_params &p, ppf_step_result *out,
                     hipStream_t s) {
    if (ns_ != 1) {
        set_error("stepper: a stepper of %u scenes advances them together", ns_);
        return PPF_ERR_ARG;
    }
    return advance_scenes(sc, p, &p.prev_dt, out, s);
}

// One time step of every scene (main.cu:267-1649 per scene). The scenes run
// their Newton loops side by side: each keeps its own dt, toi bookkeeping,
// Newton count and status; every stage is one launch over all of them, and a
// scene that has finished its loop takes no further steps while the others
// iterate.
int Stepper::advance_scenes(const ppf_scene &sc, const ppf_step_params &p, const float *prev_dt,
                            ppf_step_result *out, hipStream_t s) {
    const uint32_t ns = ns_;
    for (uint32_t c = 0; c < ns; ++c) {
        out[c] = ppf_step_result{};
    }
    if ((sc.nsphere && !sc.spheres) || (sc.nfloor && !sc.floors)) {
        set_error("stepper: analytic collider counts without arrays");
        return PPF_ERR_ARG;
    }
    if (ns > 1 && (sc.nsphere || sc.nfloor)) {
        set_error("stepper: analytic colliders apply to one scene (a stepper of %u scenes has "
                  "collision meshes per scene instead)",
                  ns);
        return PPF_ERR_ARG;
    }
    // the analytic colliders of this step (kinematic ones move between steps)
    if (sc.nsphere > sph_cap_) {
        if (sph_) {
            PPF_HIP(dev_free(sph_));
            sph_ = nullptr;
        }
        PPF_HIP(dev_alloc(&sph_, sc.nsphere * sizeof(ppf_sphere)));
        sph_cap_ = sc.nsphere;
    }
    if (sc.nfloor > flr_cap_) {
        if (flr_) {
            PPF_HIP(dev_free(flr_));
            flr_ = nullptr;
        }
        PPF_HIP(dev_alloc(&flr_, sc.nfloor * sizeof(ppf_floor)));
        flr_cap_ = sc.nfloor;
    }
    nsph_ = sc.nsphere;
    nflr_ = sc.nfloor;
    if (nsph_) {
        PPF_HIP(hipMemcpyAsync(sph_, sc.spheres, nsph_ * sizeof(ppf_sphere),
                               hipMemcpyHostToDevice, s));
    }
    if (nflr_) {
        PPF_HIP(hipMemcpyAsync(flr_, sc.floors, nflr_ * sizeof(ppf_floor), hipMemcpyHostToDevice,
                               s));
    }
    records_count_ = 0;
    uint64_t alloc0 = 0, free0 = 0;
    alloc_totals(&alloc0, &free0);
    const uint32_t n = nv_;
    const size_t n3 = 3 * (size_t)n;
    const uint32_t *vs = ns > 1 ? vscene_ : nullptr;
    PPF_HIP(hipEventRecord(ev_[5], s));
    const f3 g = mk3(p.gravity[0], p.gravity[1], p.gravity[2]);
    std::vector<float> dt(ns, p.dt);
    std::vector<double> toi_advanced(ns, 0.0);
    std::vector<uint32_t> step(ns, 1u);
    std::vector<uint8_t> final_step(ns, 0u);
    std::vector<float> toi(ns), recale(ns);
    active_.assign(ns, 1u);
    for (uint32_t c = 0; c < ns; ++c) {
        h_sf_[kDt * ns + c] = dt[c];
        h_sf_[kPrevDt * ns + c] = prev_dt[c] > 0.0f ? prev_dt[c] : p.dt;
        h_sf_[kBack * ns + c] = 0.0f;
        h_su_[kActive * ns + c] = 1u;
    }
    PPF_HIP(hipMemcpyAsync(sf_, h_sf_, kSceneFloats * ns * sizeof(float), hipMemcpyHostToDevice,
                           s));
    PPF_HIP(hipMemcpyAsync(su_ + kActive * ns, h_su_ + kActive * ns, ns * sizeof(uint32_t),
                           hipMemcpyHostToDevice, s));
    k_target<<<cdiv(n, kNT), kNT, 0, s>>>(n, sc.mesh.x, sc.prev, sc.mesh.vflag, sc.fix_pos, vs,
                                           sf_, ns, g, target_);
    PPF_HIP(hipGetLastError());
    PPF_HIP(hipMemcpyAsync(ex_, sc.mesh.x, n3 * sizeof(float), hipMemcpyDeviceToDevice, s));
    for (uint32_t iter = 0;; ++iter) {
        uint32_t live = 0;
        for (uint32_t c = 0; c < ns; ++c) {
            if (active_[c] && !final_step[c] && p.max_newton_steps > 0 &&
                step[c] >= p.max_newton_steps) {
                out[c].status = 5;
                active_[c] = 0u;
            }
            live += active_[c];
        }
        if (live == 0) {
            break;
        }
        if (live < ns) { // the scenes that have left the loop take no step
            for (uint32_t c = 0; c < ns; ++c) {
                h_su_[kActive * ns + c] = active_[c];
            }
            PPF_HIP(hipMemcpyAsync(su_ + kActive * ns, h_su_ + kActive * ns,
                                   ns * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        }
        int rc = newton_iteration(sc, p, dt[0], iter == 0, out, toi.data(), recale.data(), s);
        if (rc != PPF_OK) {
            return rc;
        }
        bool moved = false;
        for (uint32_t c = 0; c < ns; ++c) {
            h_sf_[kBack * ns + c] = 0.0f;
            if (!active_[c]) {
                continue;
            }
            out[c].newton_steps += 1;
            if (out[c].status != 0) {
                active_[c] = 0u;
                continue;
            }
            if (final_step[c]) {
                active_[c] = 0u;
                continue;
            }
            toi_advanced[c] +=
                std::max(0.0, 1.0 - toi_advanced[c]) * (double)(recale[c] * toi[c]);
            if (toi_advanced[c] >= p.target_toi && step[c] >= p.min_newton_steps) {
                // the error-reduction iteration at dt * toi_advanced, with the
                // pins walked back to that fraction of their step
                final_step[c] = 1u;
                if (sc.fix_pos && sc.fix_delta && toi_advanced[c] < 1.0) {
                    h_sf_[kBack * ns + c] = 1.0f - (float)toi_advanced[c];
                    moved = true;
                }
                dt[c] *= (float)toi_advanced[c];
            } else {
                ++step[c];
            }
            h_sf_[kDt * ns + c] = dt[c];
        }
        bool any = false;
        for (uint32_t c = 0; c < ns; ++c) {
            any = any || active_[c];
        }
        if (!any) {
            break;
        }
        PPF_HIP(hipMemcpyAsync(sf_, h_sf_, kSceneFloats * ns * sizeof(float),
                               hipMemcpyHostToDevice, s));
        if (moved) {
            k_rewind<<<cdiv(n, kNT), kNT, 0, s>>>(n, sc.mesh.vflag, sc.fix_pos, sc.fix_delta, vs,
                                                  sf_, ns);
            PPF_HIP(hipGetLastError());
        }
        k_target<<<cdiv(n, kNT), kNT, 0, s>>>(n, sc.mesh.x, sc.prev, sc.mesh.vflag, sc.fix_pos,
                                               vs, sf_, ns, g, target_);
        PPF_HIP(hipGetLastError());
        // the host's copy of the scalars is staged again after the next
        // synchronization
        PPF_HIP(hipStreamSynchronize(s));
    }
    bool check = false;
    for (uint32_t c = 0; c < ns; ++c) {
        out[c].toi_advanced = (float)toi_advanced[c];
        out[c].dt = dt[c];
        check = check || out[c].status == 0;
    }
    if (check && p.check_intersection) {
        ppf_contact_mesh m = sc.mesh;
        m.x = ex_;
        int rc = narrow_.build_trees(m, s);
        if (rc != PPF_OK) {
            return rc;
        }
        uint32_t count = 0;
        std::vector<uint32_t> per(ns, 0u);
        rc = isect_.run(narrow_.tree(0).view(), narrow_.tree(1).view(), narrow_.tree(2).view(),
                        m, sc.collision.nv ? &sc.collision : nullptr, sc.shell.vmass,
                        sc.emass, sc.shell.fmass, n, records_, kMaxRecords, &count, s,
                        vs ? per.data() : nullptr);
        if (rc != PPF_OK) {
            return rc;
        }
        records_count_ = count;
        if (!vs) {
            per[0] = count;
        }
        for (uint32_t c = 0; c < ns; ++c) {
            if (out[c].status == 0) {
                out[c].intersections = per[c];
                if (per[c]) {
                    out[c].status = 4;
                }
            }
        }
    }
    // prev <- curr, curr <- iterate for the scenes that commit
    if (!vs) {
        if (out[0].status == 0 || out[0].status == 4) {
            PPF_HIP(hipMemcpyAsync(sc.prev, sc.mesh.x, n3 * sizeof(float),
                                   hipMemcpyDeviceToDevice, s));
            PPF_HIP(hipMemcpyAsync(const_cast<float *>(sc.mesh.x), ex_, n3 * sizeof(float),
                                   hipMemcpyDeviceToDevice, s));
        }
    } else {
        for (uint32_t c = 0; c < ns; ++c) {
            h_su_[kActive * ns + c] = (out[c].status == 0 || out[c].status == 4) ? 1u : 0u;
        }
        PPF_HIP(hipMemcpyAsync(su_ + kActive * ns, h_su_ + kActive * ns, ns * sizeof(uint32_t),
                               hipMemcpyHostToDevice, s));
        k_commit<<<cdiv(n, kNT), kNT, 0, s>>>(n, vs, su_ + kActive * ns, sc.prev,
                                               const_cast<float *>(sc.mesh.x), ex_);
        PPF_HIP(hipGetLastError());
    }
    // rest-shape creep of the scenes that committed, at their new positions,
    // with the step's requested dt (main.cu:1615-1624); the shell's inv_rest
    // and h_rest are rewritten in place
    if ((sc.shell.f_plasticity || sc.shell.h_plasticity) &&
        (vs || out[0].status == 0 || out[0].status == 4)) {
        ppf_shell sh = sc.shell;
        sh.x = sc.mesh.x;
        const int prc = elastic_.plasticity(sh, p.dt, vs, vs ? su_ + kActive * ns : nullptr,
                                            const_cast<float *>(sc.shell.inv_rest),
                                            const_cast<float *>(sc.shell.h_rest), s);
        if (prc != PPF_OK) {
            return prc;
        }
    }
    PPF_HIP(hipEventRecord(ev_[6], s));
    PPF_HIP(hipEventSynchronize(ev_[6]));
    float ms_total = 0.0f;
    PPF_HIP(hipEventElapsedTime(&ms_total, ev_[5], ev_[6]));
    uint64_t alloc1 = 0, free1 = 0;
    alloc_totals(&alloc1, &free1);
    for (uint32_t c = 0; c < ns; ++c) {
        out[c].ms_total = ms_total;
        out[c].device_alloc = (uint32_t)(alloc1 - alloc0);
        out[c].device_free = (uint32_t)(free1 - free0);
    }
    return PPF_OK;
}

} // namespace ppf
