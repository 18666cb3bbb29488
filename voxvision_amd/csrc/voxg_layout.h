// Flattened device layout of a voxvision octree (shared by host and device).
//
// One 48-byte record per node, breadth-first: the root is record 0, the
// non-empty children of an inner node are contiguous, in subspace order,
// starting at `ref`, and the first levels of the tree form a prefix of the
// array.  Leaf dots are packed xyz floats in canonical voxel-id order, so
// the id of a regular-leaf voxel is its dot index.
#pragma once
#include <stdint.h>

namespace voxg {

enum : uint32_t {
    NT_INNER = 0,
    NT_LEAF = 1,
    NT_DENSE = 2,
    NT_TYPE_MASK = 3,
    NT_OVERFLOW = 4,
};

constexpr uint32_t kEnd = 0xffffffffu;
constexpr uint32_t kMaxLeafDots = (1u << 24) - 1;

struct alignas(16) NodeRec {
    float bmin[3];  // bounding box (tree.c:57-72: tight box of the voxels)
    float bmax[3];
    uint32_t ref;   // inner: first child record; leaf: first dot;
                    // dense: first voxel id
    uint32_t info;  // [1:0] type, [2] overflow; inner: [15:8] child mask;
                    // leaf: [31:8] dot count
    float c[3];     // inner: centre of subdivision (tree.c:31-55);
                    // dense: its dimensions in voxels (3 x uint32)
    uint32_t escape;  // next record in depth-first order after this subtree
};
static_assert(sizeof(NodeRec) == 48, "NodeRec must be 48 bytes");

inline uint32_t node_type(uint32_t info) { return info & NT_TYPE_MASK; }
inline uint32_t child_mask(uint32_t info) { return (info >> 8) & 0xffu; }
inline uint32_t leaf_count(uint32_t info) { return info >> 8; }

}  // namespace voxg
