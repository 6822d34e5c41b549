// bvh_query.hpp -- box queries over an LBVH (device).
//
// aabb::overlap / aabb::query of contact/aabb.hpp:70-107: depth-first from the
// root, inactive boxes never overlap. One thread per query; the per-thread
// stack lives in LDS, strided by the workgroup size so a wave's stack slots
// fall in distinct banks. Children are tested before they are pushed, so the
// stack only ever holds overlapping nodes. A query that would push past the
// stack flags `overflow` and stops (the caller reports a capacity error).
#pragma once
#include "lbvh.hpp"

namespace ppf {

constexpr int kQueryThreads = 128; // threads per query workgroup
// Stack entries per thread. A depth-first walk that tests children before
// pushing them holds at most depth + 1 entries. A Karras tree over 30-bit
// Morton codes is at most 30 levels deep over distinct codes, plus log2 of
// the largest run of equal codes (ties split by index): 48 entries cover runs
// of 2^17. The C3 trees (1-1.5M leaves) are 28-30 levels deep. A deeper walk
// raises the overflow flag (a capacity error), it never writes past the stack.
// 48 rather than 64 entries lets 12 one-wave workgroups share a CU's LDS.
constexpr int kQueryStack = 48;

// A query box and the scene of its query element (0 in a one-scene mesh).
struct QBox {
    float mn[3], mx[3];
    uint32_t scene = 0;
};

// Box overlap (aabb::overlap), restricted to nodes holding the query's scene:
// a node whose leaves all belong to another scene is never entered.
__device__ inline bool overlaps(const Aabb &a, const QBox &q) {
    const uint32_t tag = __float_as_uint(a.pad);
    return a.active != 0.0f && (tag == q.scene || tag == kMixedScene) &&
           (a.mn[0] <= q.mx[0] && a.mx[0] >= q.mn[0]) &&
           (a.mn[1] <= q.mx[1] && a.mx[1] >= q.mn[1]) &&
           (a.mn[2] <= q.mx[2] && a.mx[2] >= q.mn[2]);
}

// Scene id of a vertex (0 when the mesh holds one scene).
__device__ inline uint32_t vertex_scene(const uint32_t *vscene, uint32_t nscene, uint32_t v) {
    return (vscene && nscene > 1) ? vscene[v] : 0u;
}

__device__ inline Aabb load_box(const Aabb *b, uint32_t i) {
    const float4 *p = reinterpret_cast<const float4 *>(b + i);
    const float4 lo = p[0];
    const float4 hi = p[1];
    Aabb r;
    r.mn[0] = lo.x, r.mn[1] = lo.y, r.mn[2] = lo.z, r.active = lo.w;
    r.mx[0] = hi.x, r.mx[1] = hi.y, r.mx[2] = hi.z, r.pad = hi.w;
    return r;
}

// Box of the points a, b (min / max), then -/+ margin when it is non-zero
// (aabb::make, contact/aabb.hpp:45-60).
__device__ inline QBox seg_box(f3 a, f3 b, float margin) {
    QBox q{{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)},
           {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)},
           0u};
    if (margin != 0.0f) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            q.mn[d] -= margin;
            q.mx[d] += margin;
        }
    }
    return q;
}

// A box that overlaps nothing: the query of an element outside its
// collision window (an inactive query box, contact.cu:1637, 1728, 2563).
__device__ inline QBox empty_box() {
    const float inf = __builtin_inff();
    return QBox{{inf, inf, inf}, {-inf, -inf, -inf}, 0u};
}

// Point +/- margin (aabb::make, contact/aabb.hpp:62-69).
__device__ inline QBox point_box(f3 p, float margin) {
    return QBox{{p.x - margin, p.y - margin, p.z - margin},
                {p.x + margin, p.y + margin, p.z + margin},
                0u};
}

// The stack code of the root: its node id, or for a one-leaf tree its
// primitive | kLeafCode.
__device__ inline uint32_t root_code(const BvhView &bvh) {
    return bvh.n == 1 ? (bvh.node[2 * (size_t)bvh.root] - 1) | kLeafCode : bvh.root;
}

// Reads internal node v's walk record and tests both children against q:
// codes and verdicts (aabb::overlap on the stored boxes and scene tags).
__device__ inline void enter(const BvhView &bvh, uint32_t v, const QBox &q, uint32_t code[2],
                             bool ov[2]) {
    const uint4 *r = reinterpret_cast<const uint4 *>(bvh.rec + (v - bvh.n));
    const uint4 a = r[0], b = r[1], c = r[2], d = r[3];
    const float b0[6] = {__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z),
                         __uint_as_float(b.w), __uint_as_float(c.x), __uint_as_float(c.y)};
    const float b1[6] = {__uint_as_float(c.z), __uint_as_float(c.w), __uint_as_float(d.x),
                         __uint_as_float(d.y), __uint_as_float(d.z), __uint_as_float(d.w)};
    code[0] = a.x, code[1] = a.y;
    const uint32_t tag[2] = {a.z, a.w};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const float *bx = k ? b1 : b0;
        ov[k] = (tag[k] == q.scene || tag[k] == kMixedScene) &&
                (bx[0] <= q.mx[0] && bx[3] >= q.mn[0]) &&
                (bx[1] <= q.mx[1] && bx[4] >= q.mn[1]) &&
                (bx[2] <= q.mx[2] && bx[5] >= q.mn[2]);
    }
}

// Work of one lane's tree walks, for the walk statistics: nodes taken off
// the stack (the root included; an internal one costs its 64-byte walk
// record, a leaf nothing more), leaves among them, and candidates swept
// (queued CCD work).
struct WalkCount {
    uint32_t nodes = 0, leaves = 0, work = 0;
};

// Walk statistics of one kind of query, accumulated on the device (u64):
// launches, queries, nodes, leaves, work. Kinds: line-search vertex sweeps,
// line-search edge sweeps, proximity vertex queries, proximity edge queries.
enum WalkKind { kWalkLsVertex = 0, kWalkLsEdge = 1, kWalkProxVertex = 2, kWalkProxEdge = 3 };
constexpr int kWalkKinds = 4, kWalkFields = 5;

// Adds a workgroup's counts to stats[kind] (null: statistics off). Every
// thread calls it; `query` marks a lane that had a query.
__device__ inline void walk_flush(unsigned long long *stats, int kind, const WalkCount &wc,
                                  bool query) {
    if (!stats) {
        return;
    }
    unsigned long long *s = stats + kWalkFields * kind;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    uint32_t v[4] = {query ? 1u : 0u, wc.nodes, wc.leaves, wc.work};
#pragma unroll
    for (int f = 0; f < 4; ++f) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            v[f] += __shfl_xor(v[f], off, kWave);
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            atomicAdd(s + 1 + f, (unsigned long long)v[f]);
        }
    }
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        atomicAdd(s, 1ull);
    }
}

// Visits every leaf whose box overlaps q; `leaf(prim)` returns whether the
// candidate was accepted. Returns the accepted count. `stk` is this thread's
// first LDS slot (stride kQueryThreads).
template <class F>
__device__ uint32_t traverse(const BvhView &bvh, const QBox &q, uint32_t *stk,
                             uint32_t *overflow, F &&leaf, WalkCount *wc = nullptr) {
    if (bvh.n == 0 || !overlaps(load_box(bvh.box, bvh.root), q)) {
        return 0;
    }
    uint32_t count = 0;
    uint32_t sp = 0;
    stk[(sp++) * kQueryThreads] = root_code(bvh);
    while (sp) {
        const uint32_t v = stk[(--sp) * kQueryThreads];
        if (wc) {
            ++wc->nodes;
        }
        if (v & kLeafCode) {
            if (wc) {
                ++wc->leaves;
            }
            if (leaf(v & ~kLeafCode)) {
                ++count;
            }
            continue;
        }
        uint32_t code[2];
        bool ov[2];
        enter(bvh, v, q, code, ov);
        if (sp + 2 > (uint32_t)kQueryStack) {
            atomicOr(overflow, 1u);
            return count;
        }
        if (ov[0]) {
            stk[(sp++) * kQueryThreads] = code[0];
        }
        if (ov[1]) {
            stk[(sp++) * kQueryThreads] = code[1];
        }
    }
    return count;
}

// ---- wave-cooperative traversal --------------------------------------------
//
// For queries whose per-candidate work is long and data dependent (the CCD
// sweeps): a workgroup is one wave of kWaveQueries queries. Every lane walks
// its own query one node per round; a leaf that `accept(prim)` keeps enters a
// queue in LDS as (query id, prim). Whenever the queue holds a wave's worth of
// entries, and when the walks are over, every lane takes one entry and runs
// `work(qid, prim)`, so the long work runs with (nearly) all lanes active
// instead of with the few lanes that happen to sit on a leaf.

constexpr int kWaveQueries = 64;        // one wave per workgroup
constexpr int kWaveQueue = 2 * kWaveQueries;

template <int STACK> struct WaveScratchT {
    uint32_t stk[STACK * kWaveQueries];
    uint32_t qid[kWaveQueue], prim[kWaveQueue];
};
using WaveScratch = WaveScratchT<kQueryStack>;

__device__ inline uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// `active`: this lane has a query (lanes past the end pass false). All lanes
// of the workgroup must call it (the rounds are wave uniform).
template <int STACK, class A, class W>
__device__ void traverse_wave(const BvhView &bvh, const QBox &q, bool active, uint32_t qid,
                              WaveScratchT<STACK> &ws, uint32_t *overflow, A &&accept, W &&work,
                              WalkCount *wc = nullptr) {
    const uint32_t lane = threadIdx.x;
    uint32_t *stk = ws.stk + lane;
    uint32_t sp = 0;
    if (active && bvh.n != 0 && overlaps(load_box(bvh.box, bvh.root), q)) {
        stk[0] = root_code(bvh);
        sp = 1;
    }
    uint32_t queued = 0; // wave uniform
    while (true) {
        uint32_t cand = 0xffffffffu;
        if (sp) {
            const uint32_t v = stk[(--sp) * kWaveQueries];
            if (wc) {
                ++wc->nodes;
            }
            if (v & kLeafCode) {
                const uint32_t prim = v & ~kLeafCode;
                if (wc) {
                    ++wc->leaves;
                }
                if (accept(prim)) {
                    cand = prim;
                }
            } else {
                uint32_t code[2];
                bool ov[2];
                enter(bvh, v, q, code, ov);
                if (sp + 2 > (uint32_t)STACK) {
                    atomicOr(overflow, 1u);
                    sp = 0;
                } else {
                    if (ov[0]) {
                        stk[(sp++) * kWaveQueries] = code[0];
                    }
                    if (ov[1]) {
                        stk[(sp++) * kWaveQueries] = code[1];
                    }
                }
            }
        }
        const uint64_t m = __ballot(cand != 0xffffffffu);
        if (cand != 0xffffffffu) {
            const uint32_t at = queued + lanes_below(m);
            ws.qid[at] = qid;
            ws.prim[at] = cand;
        }
        queued += (uint32_t)__popcll(m);
        const bool walking = __ballot(sp != 0) != 0;
        if (queued >= (uint32_t)kWaveQueries || (!walking && queued)) {
            __syncthreads();
            const uint32_t take = min(queued, (uint32_t)kWaveQueries);
            if (lane < take) {
                const uint32_t at = queued - 1 - lane;
                work(ws.qid[at], ws.prim[at]);
                if (wc) {
                    ++wc->work;
                }
            }
            queued -= take;
            __syncthreads();
        }
        if (!walking && !queued) {
            break;
        }
    }
}

} // namespace ppf
