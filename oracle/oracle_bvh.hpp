// oracle_bvh.hpp -- CPU oracle LBVH (shared by the narrow phase and the
// line search). TEST INFRASTRUCTURE ONLY.
//
// Restates lbvh/lbvh.cu:108-385 (Morton codes, stable sort, Karras radix tree
// with the index tie-break), lbvh.cu:623-857 (leaf boxes, unions) and
// contact/aabb.hpp:15-107 (box make / overlap / stack query).
#pragma once
#include "oracle_math.hpp"
#include <algorithm>
#include <cstdint>
#include <vector>

namespace orc {

struct Box {
    float mn[3], mx[3];
    bool active;
};

// ---- LBVH -----------------------------------------------------------------

inline uint32_t expand_bits(uint32_t v) {
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

inline int clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }

struct Bvh {
    uint32_t n = 0;
    std::vector<uint32_t> node; // 2 per node: leaf (prim+1, 0), inner (l+1, r+1)
    std::vector<Box> box;
    uint32_t root = 0;
};

// lcp of sorted codes at i and j with the index tie-break of lbvh.cu:289-303.
inline int lcp(const std::vector<uint32_t> &k, int n, int i, int j) {
    if (j < 0 || j >= n) {
        return -1;
    }
    if (k[i] == k[j]) {
        return 32 + clz32((uint32_t)(i ^ j));
    }
    return clz32(k[i] ^ k[j]);
}

inline void karras(const std::vector<uint32_t> &k, int n, int i, int &l, int &r,
            int &split) {
    int d = lcp(k, n, i, i + 1) > lcp(k, n, i, i - 1) ? 1 : -1;
    int lmin = lcp(k, n, i, i - d);
    int lmax = 2;
    while (lcp(k, n, i, i + lmax * d) > lmin) {
        lmax *= 2;
    }
    int len = 0;
    for (int t = lmax / 2; t >= 1; t /= 2) {
        if (lcp(k, n, i, i + (len + t) * d) > lmin) {
            len += t;
        }
    }
    int j = i + len * d;
    l = std::min(i, j);
    r = std::max(i, j);
    int node = lcp(k, n, l, r);
    int s = 0, t = r - l;
    do {
        t = (t + 1) / 2;
        if (lcp(k, n, l, l + s + t) > node) {
            s += t;
        }
    } while (t > 1);
    split = l + s;
}

// centroids: npp points per primitive, cent[i] = sum / npp in lbvh.cu's order.
inline void build_bvh(Bvh &bvh, uint32_t n, const std::vector<orc::V3> &cent,
               const std::vector<Box> &prim_box) {
    bvh.n = n;
    bvh.node.assign(2 * (size_t)std::max(1u, 2 * n - 1), 0);
    bvh.box.assign(std::max(1u, 2 * n - 1), Box{});
    if (n == 0) {
        return;
    }
    float mn[3] = {3.4e38f, 3.4e38f, 3.4e38f}, mx[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    for (uint32_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            mn[d] = std::fmin(mn[d], cent[i][d]);
            mx[d] = std::fmax(mx[d], cent[i][d]);
        }
    }
    float sc[3];
    for (int d = 0; d < 3; ++d) {
        sc[d] = (mx[d] - mn[d]) > 1e-10f ? 1023.0f / (mx[d] - mn[d]) : 0.0f;
    }
    std::vector<uint32_t> code(n), idx(n);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t q[3];
        for (int d = 0; d < 3; ++d) {
            q[d] = (uint32_t)std::fmin(1023.0f,
                                       std::fmax(0.0f, (cent[i][d] - mn[d]) * sc[d]));
        }
        code[i] = expand_bits(q[0]) | (expand_bits(q[1]) << 1) |
                  (expand_bits(q[2]) << 2);
        idx[i] = i;
    }
    std::stable_sort(idx.begin(), idx.end(),
                     [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
    std::vector<uint32_t> sk(n);
    for (uint32_t i = 0; i < n; ++i) {
        sk[i] = code[idx[i]];
        bvh.node[2 * i] = idx[i] + 1;
        bvh.node[2 * i + 1] = 0;
        bvh.box[i] = prim_box[idx[i]];
    }
    std::vector<uint32_t> parent(2 * n - 1, 0xffffffffu);
    for (int i = 0; i + 1 < (int)n; ++i) {
        int l, r, s;
        karras(sk, (int)n, i, l, r, s);
        uint32_t lc = (l == s) ? (uint32_t)l : n + (uint32_t)s;
        uint32_t rc = (s + 1 == r) ? (uint32_t)r : n + (uint32_t)s + 1;
        bvh.node[2 * (n + i)] = lc + 1;
        bvh.node[2 * (n + i) + 1] = rc + 1;
        parent[lc] = n + i;
        parent[rc] = n + i;
    }
    bvh.root = n == 1 ? 0 : n;
    // Union of children, children first (post-order by explicit recursion).
    std::vector<uint32_t> order;
    std::vector<uint32_t> st{bvh.root};
    while (!st.empty()) {
        uint32_t v = st.back();
        st.pop_back();
        order.push_back(v);
        if (bvh.node[2 * v + 1]) {
            st.push_back(bvh.node[2 * v] - 1);
            st.push_back(bvh.node[2 * v + 1] - 1);
        }
    }
    for (size_t k = order.size(); k-- > 0;) {
        uint32_t v = order[k];
        if (bvh.node[2 * v + 1]) {
            const Box &a = bvh.box[bvh.node[2 * v] - 1];
            const Box &b = bvh.box[bvh.node[2 * v + 1] - 1];
            Box &o = bvh.box[v];
            for (int d = 0; d < 3; ++d) {
                o.mn[d] = std::fmin(a.mn[d], b.mn[d]);
                o.mx[d] = std::fmax(a.mx[d], b.mx[d]);
            }
            o.active = a.active && b.active;
        }
    }
}

inline bool overlap(const Box &a, const Box &b) {
    if (!a.active || !b.active) {
        return false;
    }
    return (a.mn[0] <= b.mx[0] && a.mx[0] >= b.mn[0]) &&
           (a.mn[1] <= b.mx[1] && a.mx[1] >= b.mn[1]) &&
           (a.mn[2] <= b.mx[2] && a.mx[2] >= b.mn[2]);
}

template <class F> uint32_t query(const Bvh &bvh, const Box &q, F &&leaf) {
    if (bvh.n == 0) {
        return 0;
    }
    std::vector<uint32_t> st;
    st.reserve(128);
    st.push_back(bvh.root);
    uint32_t count = 0;
    while (!st.empty()) {
        uint32_t v = st.back();
        st.pop_back();
        if (!overlap(bvh.box[v], q)) {
            continue;
        }
        if (bvh.node[2 * v + 1] == 0) {
            if (leaf(bvh.node[2 * v] - 1)) {
                ++count;
            }
        } else {
            st.push_back(bvh.node[2 * v] - 1);
            st.push_back(bvh.node[2 * v + 1] - 1);
        }
    }
    return count;
}

inline Box make_box(const orc::V3 *pts, int m, float margin) {
    Box b;
    for (int d = 0; d < 3; ++d) {
        float lo = pts[0][d], hi = pts[0][d];
        for (int k = 1; k < m; ++k) {
            lo = std::fmin(lo, pts[k][d]);
            hi = std::fmax(hi, pts[k][d]);
        }
        b.mn[d] = lo;
        b.mx[d] = hi;
    }
    b.active = true;
    if (margin != 0.0f) {
        for (int d = 0; d < 3; ++d) {
            b.mn[d] -= margin;
            b.mx[d] += margin;
        }
    }
    return b;
}

inline Box point_box(const orc::V3 &x, float margin) {
    Box b;
    for (int d = 0; d < 3; ++d) {
        b.mn[d] = x[d] - margin;
        b.mx[d] = x[d] + margin;
    }
    b.active = true;
    return b;
}

// Leaf box of lbvh.cu:623-731 with x1 == x0: min/max then -/+ margin.
inline Box leaf_box(const orc::V3 *pts, int m, float margin) {
    Box b;
    for (int d = 0; d < 3; ++d) {
        float lo = pts[0][d], hi = pts[0][d];
        for (int k = 1; k < m; ++k) {
            lo = std::fmin(lo, pts[k][d]);
            hi = std::fmax(hi, pts[k][d]);
        }
        b.mn[d] = std::fmin(lo, lo) - margin;
        b.mx[d] = std::fmax(hi, hi) + margin;
    }
    b.active = true;
    return b;
}


} // namespace orc
