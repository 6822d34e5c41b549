// lbvh.hpp -- linear BVH over faces, edges or vertices (device).
//
// Same tree as the reference's LBVH (lbvh/lbvh.cu:108-385): centroids, scene
// bounds of the centroids, 10-bit-per-axis Morton codes, a stable sort of
// (code, index), and the Karras radix tree with the index tie-break on equal
// codes. Node i < n is leaf i (prim + 1, 0); node n + i is internal node i
// (left + 1, right + 1); the root is node n (leaf 0 when n == 1). Boxes are
// refit bottom-up one depth level per launch (levels computed once per build).
#pragma once
#include <vector>

#include "common.hpp"

namespace ppf {

// Box of one node: min xyz, active flag, max xyz, scene tag (32 B, two 16 B
// loads). The tag holds the bits of the scene id every leaf below the node
// belongs to, or kMixedScene when they belong to several (0 throughout for a
// one-scene tree).
struct alignas(32) Aabb {
    float mn[3];
    float active;
    float mx[3];
    float pad;
};
constexpr uint32_t kMixedScene = 0xffffffffu;

// Walk record of internal node n + i (64 B, four 16 B loads): both children's
// codes, scene tags and boxes, so a walk reads one record per internal node
// it enters instead of the node and then its children's boxes. A child code
// is the child's node id, or for a leaf its primitive | kLeafCode. An
// inactive child's tag is kInactiveTag, which matches no query's scene.
struct alignas(64) WalkRec {
    uint32_t code[2];
    uint32_t tag[2];
    float box[12]; // child 0 min xyz, max xyz, then child 1
};
constexpr uint32_t kLeafCode = 0x80000000u;
constexpr uint32_t kInactiveTag = 0xfffffffeu;

struct BvhView {
    uint32_t n = 0;            // leaves
    uint32_t root = 0;
    const uint32_t *node = nullptr; // 2 words per node
    const Aabb *box = nullptr;
    const WalkRec *rec = nullptr;   // n - 1 records (internal nodes)
};

// Primitive set a tree is built over: `arity` vertex indices per primitive
// (3 faces, 2 edges, 1 vertices; index = element for vertices), margin
// 0.5*ghat + offset per primitive.
struct PrimSet {
    uint32_t n = 0;
    int arity = 1;
    const uint32_t *elem = nullptr;  // n * arity vertex indices (null: identity)
    const float *ghat = nullptr;     // per primitive
    const float *offset = nullptr;   // per primitive
    // Many scenes: scene id per vertex (a primitive's scene is its first
    // vertex's) and the scene count. The tree then orders its leaves by
    // (scene, Morton code), so each scene is one subtree, and tags every node
    // with its scene. Null / nscene <= 1: one scene.
    const uint32_t *vscene = nullptr;
    uint32_t nscene = 0;
};

class Lbvh {
  public:
    ~Lbvh();
    // Allocates for up to n primitives (grows when needed).
    int reserve(uint32_t n);
    // Full build from positions x0 (swept to x0 + t (x1 - x0) when x1 is set).
    int build(const PrimSet &p, const float *x0, const float *x1, float t,
              hipStream_t s);
    // Refit the leaf boxes and the inner boxes, same topology.
    int refit(const PrimSet &p, const float *x0, const float *x1, float t,
              hipStream_t s);
    BvhView view() const;
    uint32_t *nodes() { return node_; }
    Aabb *boxes() { return box_; }

  private:
    uint32_t cap_ = 0, n_ = 0, root_ = 0;
    uint32_t *node_ = nullptr, *parent_ = nullptr, *depth_ = nullptr, *level_ = nullptr;
    std::vector<uint32_t> loff_{0}; // level offsets into level_ (host copy)
    uint32_t *code_ = nullptr, *idx_ = nullptr, *scratch_ = nullptr;
    float *cent_ = nullptr, *part_ = nullptr;
    Aabb *box_ = nullptr;
    WalkRec *rec_ = nullptr;
    uint32_t *sb_ = nullptr; // per-scene centroid bounds (many scenes)
    size_t sb_cap_ = 0;
    void release();
    int levels(hipStream_t s);
};

} // namespace ppf
