// lbvh.hip -- LBVH build and refit on MI355X.
//
// Tree semantics follow lbvh/lbvh.cu:108-385 (Morton quantization, Karras
// split with the index tie-break) and the swept leaf boxes of
// lbvh/lbvh.cu:623-731; the inner boxes are unions of the children. The
// per-level launches and host round trips of the reference are replaced by a
// single bottom-up kernel with per-node arrival counters: the second child to
// arrive writes the union and continues upward.
#include "lbvh.hpp"
#include <algorithm>
#include "ppf_mi355x.h"
#include "sort.hpp"

namespace ppf {

namespace {

constexpr int kNT = 256;
constexpr uint32_t kNoParent = 0xffffffffu;

__device__ inline uint32_t prim_vertex(const PrimSet &p, uint32_t i, int k) {
    return p.elem ? p.elem[(size_t)i * p.arity + k] : i;
}

__device__ inline uint32_t expand_bits(uint32_t v) {
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Centroids (lbvh.cu:120-167 term order) and per-block min/max partials.
__global__ __launch_bounds__(kNT) void k_centroid(PrimSet p, const float *__restrict__ x,
                                                  float *__restrict__ cent,
                                                  float *__restrict__ part) {
    __shared__ float s[6][kNT];
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    float c[3] = {3.4e38f, 3.4e38f, 3.4e38f}, mx[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    if (i < p.n) {
        const f3 a = load3(x + 3 * (size_t)prim_vertex(p, i, 0));
        f3 m;
        if (p.arity == 3) {
            const f3 b = load3(x + 3 * (size_t)prim_vertex(p, i, 1));
            const f3 d = load3(x + 3 * (size_t)prim_vertex(p, i, 2));
            m = mk3((a.x + b.x + d.x) / 3.0f, (a.y + b.y + d.y) / 3.0f,
                    (a.z + b.z + d.z) / 3.0f);
        } else if (p.arity == 2) {
            const f3 b = load3(x + 3 * (size_t)prim_vertex(p, i, 1));
            m = mk3((a.x + b.x) / 2.0f, (a.y + b.y) / 2.0f, (a.z + b.z) / 2.0f);
        } else {
            m = a;
        }
        store3(cent + 3 * (size_t)i, m);
        c[0] = mx[0] = m.x;
        c[1] = mx[1] = m.y;
        c[2] = mx[2] = m.z;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        s[d][threadIdx.x] = c[d];
        s[3 + d][threadIdx.x] = mx[d];
    }
    __syncthreads();
    for (int w = kNT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                s[d][threadIdx.x] = fminf(s[d][threadIdx.x], s[d][threadIdx.x + w]);
                s[3 + d][threadIdx.x] =
                    fmaxf(s[3 + d][threadIdx.x], s[3 + d][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) {
        part[6 * (size_t)blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
    }
}

__global__ __launch_bounds__(kNT) void k_bounds(uint32_t nb, float *__restrict__ part,
                                                float *__restrict__ bounds) {
    __shared__ float s[6][kNT];
    float v[6] = {3.4e38f, 3.4e38f, 3.4e38f, -3.4e38f, -3.4e38f, -3.4e38f};
    for (uint32_t b = threadIdx.x; b < nb; b += kNT) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            v[d] = fminf(v[d], part[6 * (size_t)b + d]);
            v[3 + d] = fmaxf(v[3 + d], part[6 * (size_t)b + 3 + d]);
        }
    }
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        s[d][threadIdx.x] = v[d];
    }
    __syncthreads();
    for (int w = kNT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                s[d][threadIdx.x] = fminf(s[d][threadIdx.x], s[d][threadIdx.x + w]);
                s[3 + d][threadIdx.x] =
                    fmaxf(s[3 + d][threadIdx.x], s[3 + d][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) {
        bounds[threadIdx.x] = s[threadIdx.x][0];
    }
}

// 30-bit Morton codes (lbvh.cu:253-279).
__global__ __launch_bounds__(kNT) void k_morton(uint32_t n, const float *__restrict__ cent,
                                                const float *__restrict__ bounds,
                                                uint32_t *__restrict__ code,
                                                uint32_t *__restrict__ idx) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i >= n) {
        return;
    }
    uint32_t q[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float lo = bounds[d], hi = bounds[3 + d];
        const float sc = (hi - lo) > 1e-10f ? 1023.0f / (hi - lo) : 0.0f;
        q[d] = (uint32_t)fminf(1023.0f, fmaxf(0.0f, (cent[3 * (size_t)i + d] - lo) * sc));
    }
    code[i] = expand_bits(q[0]) | (expand_bits(q[1]) << 1) | (expand_bits(q[2]) << 2);
    idx[i] = i;
}

__device__ inline uint32_t prim_scene(const PrimSet &p, uint32_t prim) {
    return (p.vscene && p.nscene > 1) ? p.vscene[prim_vertex(p, prim, 0)] : 0u;
}

// Float <-> unsigned with the same order (negative floats reversed), so that
// integer atomics take exact float minima and maxima.
__device__ inline uint32_t ord_bits(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float ord_float(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// Per-scene centroid bounds (6 ordered words per scene: min xyz, max xyz),
// one atomic per wave and component where a wave's primitives share a scene.
__global__ __launch_bounds__(kNT) void k_scene_bounds(PrimSet p, const float *__restrict__ cent,
                                                      uint32_t *__restrict__ sb) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    const bool ok = i < p.n;
    const uint32_t c = ok ? prim_scene(p, i) : 0u;
    float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    if (ok) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo[d] = hi[d] = cent[3 * (size_t)i + d];
        }
    }
    const uint32_t c0 = __shfl(c, 0);
    if (__ballot(ok && c != c0) == 0) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            for (int o = 32; o > 0; o >>= 1) {
                lo[d] = fminf(lo[d], __shfl_xor(lo[d], o));
                hi[d] = fmaxf(hi[d], __shfl_xor(hi[d], o));
            }
        }
        if ((threadIdx.x & 63) == 0 && __ballot(ok) != 0) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                atomicMin(sb + 6 * c0 + d, ord_bits(lo[d]));
                atomicMax(sb + 6 * c0 + 3 + d, ord_bits(hi[d]));
            }
        }
    } else if (ok) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            atomicMin(sb + 6 * c + d, ord_bits(lo[d]));
            atomicMax(sb + 6 * c + 3 + d, ord_bits(hi[d]));
        }
    }
}

// Morton codes quantized in each primitive's own scene box.
__global__ __launch_bounds__(kNT) void k_morton_scenes(PrimSet p, const float *__restrict__ cent,
                                                       const uint32_t *__restrict__ sb,
                                                       uint32_t *__restrict__ code,
                                                       uint32_t *__restrict__ idx) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i >= p.n) {
        return;
    }
    const uint32_t c = prim_scene(p, i);
    uint32_t q[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float lo = ord_float(sb[6 * c + d]), hi = ord_float(sb[6 * c + 3 + d]);
        const float sc = (hi - lo) > 1e-10f ? 1023.0f / (hi - lo) : 0.0f;
        q[d] = (uint32_t)fminf(1023.0f, fmaxf(0.0f, (cent[3 * (size_t)i + d] - lo) * sc));
    }
    code[i] = expand_bits(q[0]) | (expand_bits(q[1]) << 1) | (expand_bits(q[2]) << 2);
    idx[i] = i;
}

// Empty scene boxes (min at the top of the order, max at the bottom).
__global__ void k_scene_bounds_init(uint32_t ns, uint32_t *__restrict__ sb) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < 6 * ns) {
        sb[k] = (k % 6) < 3 ? 0xffffffffu : 0u;
    }
}

// Scene of each sorted leaf (the key of the second, stable sort pass).
__global__ __launch_bounds__(kNT) void k_scene_keys(PrimSet p, const uint32_t *__restrict__ idx,
                                                    uint32_t *__restrict__ key) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i < p.n) {
        key[i] = prim_scene(p, idx[i]);
    }
}

// Morton codes in the final leaf order (from the per-primitive copy).
__global__ __launch_bounds__(kNT) void k_regather(uint32_t n, const uint32_t *__restrict__ idx,
                                                  const uint32_t *__restrict__ byprim,
                                                  uint32_t *__restrict__ code) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i < n) {
        code[i] = byprim[idx[i]];
    }
}

// Common-prefix length of the keys of sorted leaves i and j: the Morton code
// with the index as tie-break (lbvh.cu:289-303), preceded by the scene id
// when the tree holds many scenes (sk: scene per sorted leaf, or null).
__device__ inline int lcp(const uint32_t *__restrict__ k, const uint32_t *__restrict__ sk, int n,
                          int i, int j) {
    if (j < 0 || j >= n) {
        return -1;
    }
    int base = 0;
    if (sk) {
        const uint32_t sa = sk[i], sb = sk[j];
        if (sa != sb) {
            return __clz(sa ^ sb);
        }
        base = 32;
    }
    const uint32_t a = k[i], b = k[j];
    if (a == b) {
        return base + 32 + __clz((uint32_t)(i ^ j));
    }
    return base + __clz(a ^ b);
}

// Karras radix tree over the sorted codes (lbvh.cu:305-385), parents inline.
__global__ __launch_bounds__(kNT) void k_karras(uint32_t n, const uint32_t *__restrict__ k,
                                                const uint32_t *__restrict__ sk,
                                                const uint32_t *__restrict__ idx,
                                                uint32_t *__restrict__ node,
                                                uint32_t *__restrict__ parent) {
    const int i = blockIdx.x * kNT + threadIdx.x;
    const int nn = (int)n;
    if (i < nn) {
        node[2 * (size_t)i] = idx[i] + 1;
        node[2 * (size_t)i + 1] = 0;
    }
    if (i == 0) {
        parent[nn == 1 ? 0 : n] = kNoParent;
    }
    if (i >= nn - 1) {
        return;
    }
    const int d = lcp(k, sk, nn, i, i + 1) > lcp(k, sk, nn, i, i - 1) ? 1 : -1;
    const int lmin = lcp(k, sk, nn, i, i - d);
    int lmax = 2;
    while (lcp(k, sk, nn, i, i + lmax * d) > lmin) {
        lmax *= 2;
    }
    int len = 0;
    for (int t = lmax / 2; t >= 1; t /= 2) {
        if (lcp(k, sk, nn, i, i + (len + t) * d) > lmin) {
            len += t;
        }
    }
    const int j = i + len * d;
    const int l = min(i, j), r = max(i, j);
    const int nodelcp = lcp(k, sk, nn, l, r);
    int s = 0, t = r - l;
    do {
        t = (t + 1) / 2;
        if (lcp(k, sk, nn, l, l + s + t) > nodelcp) {
            s += t;
        }
    } while (t > 1);
    const int split = l + s;
    const uint32_t lc = (l == split) ? (uint32_t)l : n + (uint32_t)split;
    const uint32_t rc = (split + 1 == r) ? (uint32_t)r : n + (uint32_t)split + 1;
    node[2 * ((size_t)n + i)] = lc + 1;
    node[2 * ((size_t)n + i) + 1] = rc + 1;
    parent[lc] = n + i;
    parent[rc] = n + i;
}

// Swept leaf box (lbvh.cu:623-731).
__global__ __launch_bounds__(kNT) void k_leaf_boxes(PrimSet p, const float *__restrict__ x0,
                                                    const float *__restrict__ x1, float t,
                                                    const uint32_t *__restrict__ node,
                                                    Aabb *__restrict__ box) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i >= p.n) {
        return;
    }
    const uint32_t prim = node[2 * (size_t)i] - 1;
    const float margin = 0.5f * p.ghat[prim] + p.offset[prim];
    float lo0[3], hi0[3], lo1[3], hi1[3];
    for (int k = 0; k < p.arity; ++k) {
        const uint32_t v = prim_vertex(p, prim, k);
        const f3 y0 = load3(x0 + 3 * (size_t)v);
        f3 z = y0;
        if (x1) {
            const f3 y1 = load3(x1 + 3 * (size_t)v);
            z = mk3(t * (y1.x - y0.x) + y0.x, t * (y1.y - y0.y) + y0.y,
                    t * (y1.z - y0.z) + y0.z);
        }
        const float a[3] = {y0.x, y0.y, y0.z}, b[3] = {z.x, z.y, z.z};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo0[d] = k ? fminf(lo0[d], a[d]) : a[d];
            hi0[d] = k ? fmaxf(hi0[d], a[d]) : a[d];
            lo1[d] = k ? fminf(lo1[d], b[d]) : b[d];
            hi1[d] = k ? fmaxf(hi1[d], b[d]) : b[d];
        }
    }
    Aabb bx;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        bx.mn[d] = fminf(lo0[d], lo1[d]) - margin;
        bx.mx[d] = fmaxf(hi0[d], hi1[d]) + margin;
    }
    bx.active = 1.0f;
    bx.pad = __uint_as_float(prim_scene(p, prim));
    box[i] = bx;
}

// Depth of every internal node (steps to the root along the parent links),
// the deepest level and a per-depth count.
__global__ __launch_bounds__(kNT) void k_depth(uint32_t n, uint32_t root,
                                               const uint32_t *__restrict__ parent,
                                               uint32_t *__restrict__ depth,
                                               uint32_t *__restrict__ maxd) {
    const uint32_t k = blockIdx.x * kNT + threadIdx.x;
    if (k >= n - 1) {
        return;
    }
    uint32_t cur = n + k, d = 0;
    while (cur != root && d < n) {
        cur = parent[cur];
        ++d;
    }
    depth[k] = d;
    atomicMax(maxd, d);
}

// Per-depth counts. Most nodes of a workgroup share a few depths, so they
// are counted in LDS first (depths below kLdsDepth) and each workgroup adds
// one total per depth: a global atomic per node would serialize on the few
// hot counters.
constexpr uint32_t kLdsDepth = 256;

__global__ __launch_bounds__(kNT) void k_depth_count(uint32_t m, const uint32_t *__restrict__ depth,
                                                     uint32_t *__restrict__ count) {
    __shared__ uint32_t h[kLdsDepth];
    for (uint32_t d = threadIdx.x; d < kLdsDepth; d += kNT) {
        h[d] = 0;
    }
    __syncthreads();
    const uint32_t k = blockIdx.x * kNT + threadIdx.x;
    if (k < m) {
        const uint32_t d = depth[k];
        if (d < kLdsDepth) {
            atomicAdd(h + d, 1u);
        } else {
            atomicAdd(count + d, 1u);
        }
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < kLdsDepth; d += kNT) {
        if (h[d]) {
            atomicAdd(count + d, h[d]);
        }
    }
}

// Internal nodes grouped by depth (order within a depth is irrelevant: a
// node's box depends only on its children's). Each workgroup reserves one
// range per depth and ranks its nodes inside it in LDS.
__global__ __launch_bounds__(kNT) void k_depth_scatter(uint32_t m, uint32_t n,
                                                       const uint32_t *__restrict__ depth,
                                                       uint32_t *__restrict__ fill,
                                                       uint32_t *__restrict__ level) {
    __shared__ uint32_t h[kLdsDepth], base[kLdsDepth];
    for (uint32_t d = threadIdx.x; d < kLdsDepth; d += kNT) {
        h[d] = 0;
    }
    __syncthreads();
    const uint32_t k = blockIdx.x * kNT + threadIdx.x;
    uint32_t d = 0, rank = 0;
    if (k < m) {
        d = depth[k];
        if (d < kLdsDepth) {
            rank = atomicAdd(h + d, 1u);
        } else {
            level[atomicAdd(fill + d, 1u)] = n + k;
        }
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < kLdsDepth; e += kNT) {
        base[e] = h[e] ? atomicAdd(fill + e, h[e]) : 0u;
    }
    __syncthreads();
    if (k < m && d < kLdsDepth) {
        level[base[d] + rank] = n + k;
    }
}

// Boxes of the internal nodes of one depth from their children's (the
// deeper levels are done by the launches before this one).
__global__ __launch_bounds__(kNT) void k_refit_level(const uint32_t *__restrict__ level,
                                                     uint32_t count,
                                                     const uint32_t *__restrict__ node,
                                                     Aabb *__restrict__ box) {
    const uint32_t k = blockIdx.x * kNT + threadIdx.x;
    if (k >= count) {
        return;
    }
    const uint32_t par = level[k];
    const Aabb a = box[node[2 * (size_t)par] - 1];
    const Aabb b = box[node[2 * (size_t)par + 1] - 1];
    Aabb u;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        u.mn[d] = fminf(a.mn[d], b.mn[d]);
        u.mx[d] = fmaxf(a.mx[d], b.mx[d]);
    }
    u.active = (a.active != 0.0f && b.active != 0.0f) ? 1.0f : 0.0f;
    const uint32_t ta = __float_as_uint(a.pad), tb = __float_as_uint(b.pad);
    u.pad = __uint_as_float(ta == tb ? ta : kMixedScene);
    box[par] = u;
}

// The walk records of the internal nodes from the refit boxes.
__global__ __launch_bounds__(kNT) void k_walk_records(uint32_t n,
                                                      const uint32_t *__restrict__ node,
                                                      const Aabb *__restrict__ box,
                                                      WalkRec *__restrict__ rec) {
    const uint32_t i = blockIdx.x * kNT + threadIdx.x;
    if (i + 1 >= n) {
        return;
    }
    const uint32_t v = n + i;
    WalkRec w;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const uint32_t c = node[2 * (size_t)v + k] - 1;
        const bool leaf = node[2 * (size_t)c + 1] == 0;
        w.code[k] = leaf ? (node[2 * (size_t)c] - 1) | kLeafCode : c;
        const Aabb b = box[c];
        w.tag[k] = b.active != 0.0f ? __float_as_uint(b.pad) : kInactiveTag;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            w.box[6 * k + d] = b.mn[d];
            w.box[6 * k + 3 + d] = b.mx[d];
        }
    }
    rec[i] = w;
}

uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

} // namespace

Lbvh::~Lbvh() { release(); }

void Lbvh::release() {
    for (void *p : {(void *)node_, (void *)parent_, (void *)depth_, (void *)level_,
                    (void *)code_,
                    (void *)idx_, (void *)scratch_, (void *)cent_, (void *)part_,
                    (void *)box_, (void *)rec_, (void *)sb_}) {
        if (p) {
            (void)dev_free(p);
        }
    }
    node_ = parent_ = depth_ = level_ = code_ = idx_ = scratch_ = nullptr;
    cent_ = part_ = nullptr;
    box_ = nullptr;
    rec_ = nullptr;
    sb_ = nullptr;
    sb_cap_ = 0;
    cap_ = 0;
}

int Lbvh::reserve(uint32_t n) {
    if (n <= cap_) {
        return PPF_OK;
    }
    release();
    const size_t nodes = 2 * (size_t)n;
    const size_t nb = cdiv(n, kNT);
    PPF_HIP(dev_alloc(&node_, 2 * nodes * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&parent_, nodes * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&depth_, n * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&level_, n * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&code_, n * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&idx_, n * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&scratch_, radix_sort_scratch_words(n) * sizeof(uint32_t)));
    PPF_HIP(dev_alloc(&cent_, 3 * (size_t)n * sizeof(float)));
    PPF_HIP(dev_alloc(&part_, (6 * nb + 6) * sizeof(float)));
    PPF_HIP(dev_alloc(&box_, nodes * sizeof(Aabb)));
    PPF_HIP(dev_alloc(&rec_, std::max<size_t>(n, 1) * sizeof(WalkRec)));
    cap_ = n;
    return PPF_OK;
}

int Lbvh::build(const PrimSet &p, const float *x0, const float *x1, float t,
                hipStream_t s) {
    int rc = reserve(p.n);
    if (rc != PPF_OK) {
        return rc;
    }
    if (p.n >= kLeafCode) {
        set_error("Lbvh: %u primitives exceed the walk records' index range", p.n);
        return PPF_ERR_ARG;
    }
    n_ = p.n;
    root_ = n_ == 1 ? 0 : n_;
    if (n_ == 0) {
        return PPF_OK;
    }
    const uint32_t nb = cdiv(n_, kNT);
    float *bounds = part_ + 6 * (size_t)nb;
    k_centroid<<<nb, kNT, 0, s>>>(p, x0, cent_, part_);
    PPF_HIP(hipGetLastError());
    k_bounds<<<1, kNT, 0, s>>>(nb, part_, bounds);
    PPF_HIP(hipGetLastError());
    const bool scenes = p.vscene && p.nscene > 1;
    if (scenes) {
        // each scene quantized in its own box: a scene's tree does not depend
        // on the other scenes of the mesh
        if (6 * (size_t)p.nscene > sb_cap_) {
            if (sb_) {
                PPF_HIP(dev_free(sb_));
                sb_ = nullptr;
            }
            PPF_HIP(dev_alloc(&sb_, 6 * (size_t)p.nscene * sizeof(uint32_t)));
            sb_cap_ = 6 * (size_t)p.nscene;
        }
        k_scene_bounds_init<<<cdiv(6 * p.nscene, kNT), kNT, 0, s>>>(p.nscene, sb_);
        k_scene_bounds<<<nb, kNT, 0, s>>>(p, cent_, sb_);
        k_morton_scenes<<<nb, kNT, 0, s>>>(p, cent_, sb_, code_, idx_);
    } else {
        k_morton<<<nb, kNT, 0, s>>>(n_, cent_, bounds, code_, idx_);
    }
    PPF_HIP(hipGetLastError());
    if (scenes) { // per-primitive codes, kept for the regather after the scene pass
        PPF_HIP(hipMemcpyAsync(level_, code_, n_ * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                               s));
    }
    PPF_HIP(radix_sort_pairs(code_, idx_, n_, 30, scratch_, s));
    const uint32_t *sk = nullptr;
    if (scenes) {
        // stable second pass by scene: leaves ordered by (scene, code, index)
        uint32_t bits = 1;
        while (bits < 32 && ((p.nscene - 1) >> bits) != 0) {
            ++bits;
        }
        k_scene_keys<<<nb, kNT, 0, s>>>(p, idx_, depth_);
        PPF_HIP(hipGetLastError());
        PPF_HIP(radix_sort_pairs(depth_, idx_, n_, bits, scratch_, s));
        k_regather<<<nb, kNT, 0, s>>>(n_, idx_, level_, code_);
        PPF_HIP(hipGetLastError());
        sk = depth_;
    }
    k_karras<<<nb, kNT, 0, s>>>(n_, code_, sk, idx_, node_, parent_);
    PPF_HIP(hipGetLastError());
    rc = levels(s);
    if (rc != PPF_OK) {
        return rc;
    }
    return refit(p, x0, x1, t, s);
}

// Internal nodes grouped by depth (lbvh.cu:479-616 computes the same levels):
// the order refit visits them in. The sort keys and indices are free by now
// and hold the per-depth counts and fill cursors.
int Lbvh::levels(hipStream_t s) {
    loff_.assign(1, 0);
    if (n_ < 2) {
        return PPF_OK;
    }
    const uint32_t m = n_ - 1;
    uint32_t *maxd = scratch_;
    PPF_HIP(hipMemsetAsync(maxd, 0, sizeof(uint32_t), s));
    k_depth<<<cdiv(m, kNT), kNT, 0, s>>>(n_, root_, parent_, depth_, maxd);
    PPF_HIP(hipGetLastError());
    uint32_t hmax = 0;
    PPF_HIP(hipMemcpyAsync(&hmax, maxd, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PPF_HIP(hipStreamSynchronize(s));
    if (hmax >= n_) {
        set_error("Lbvh: parent links do not reach the root");
        return PPF_ERR_SOLVER;
    }
    const uint32_t levels = hmax + 1;
    uint32_t *count = code_, *fill = idx_;
    PPF_HIP(hipMemsetAsync(count, 0, levels * sizeof(uint32_t), s));
    k_depth_count<<<cdiv(m, kNT), kNT, 0, s>>>(m, depth_, count);
    PPF_HIP(hipGetLastError());
    std::vector<uint32_t> h(levels);
    PPF_HIP(hipMemcpyAsync(h.data(), count, levels * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           s));
    PPF_HIP(hipStreamSynchronize(s));
    loff_.assign(levels + 1, 0);
    for (uint32_t d = 0; d < levels; ++d) {
        loff_[d + 1] = loff_[d] + h[d];
    }
    PPF_HIP(hipMemcpyAsync(fill, loff_.data(), levels * sizeof(uint32_t),
                           hipMemcpyHostToDevice, s));
    k_depth_scatter<<<cdiv(m, kNT), kNT, 0, s>>>(m, n_, depth_, fill, level_);
    PPF_HIP(hipGetLastError());
    // loff_ must stay intact until the copy above has read it
    PPF_HIP(hipStreamSynchronize(s));
    return PPF_OK;
}

int Lbvh::refit(const PrimSet &p, const float *x0, const float *x1, float t,
                hipStream_t s) {
    if (p.n != n_) {
        set_error("Lbvh::refit: primitive count %u differs from the build (%u)",
                  p.n, n_);
        return PPF_ERR_ARG;
    }
    if (n_ == 0) {
        return PPF_OK;
    }
    k_leaf_boxes<<<cdiv(n_, kNT), kNT, 0, s>>>(p, x0, x1, t, node_, box_);
    PPF_HIP(hipGetLastError());
    // Inner boxes level by level from the deepest: a kernel boundary orders
    // each level after its children, with no fences or arrival counters.
    for (size_t d = loff_.size() - 1; d-- > 0;) {
        const uint32_t cnt = loff_[d + 1] - loff_[d];
        if (cnt) {
            k_refit_level<<<cdiv(cnt, kNT), kNT, 0, s>>>(level_ + loff_[d], cnt, node_, box_);
            PPF_HIP(hipGetLastError());
        }
    }
    if (n_ > 1) {
        k_walk_records<<<cdiv(n_ - 1, kNT), kNT, 0, s>>>(n_, node_, box_, rec_);
        PPF_HIP(hipGetLastError());
    }
    return PPF_OK;
}

BvhView Lbvh::view() const {
    BvhView v;
    v.n = n_;
    v.root = root_;
    v.node = node_;
    v.box = box_;
    v.rec = rec_;
    return v;
}

} // namespace ppf
