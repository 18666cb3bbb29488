// Light manager of libvoxgpu: the shadowless light spheres of
// src/voxrnd/lights.c kept in the reference's M-tree (src/voxtrees/mtree.c),
// on the host, and its flattened form for the device.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace voxg {

// Device record of one M-tree node, depth-first preorder: an inner node's
// first child is the next record, `escape` is the record after its subtree.
struct alignas(16) LightNode {
    float c[3];        // bounding sphere centre
    float sqr_radius;  // and squared radius
    uint32_t info;     // bit 0: leaf; bits 8..: sphere count of a leaf
    uint32_t first;    // leaf: index of its first sphere
    uint32_t escape;
    uint32_t pad;
};

// Device record of one light sphere.
struct alignas(16) LightRec {
    float c[3];
    float radius;
    float color[3];
    float sqr_radius;
};

struct LightSphere {
    float center[3];
    float color[3];
    float radius, sqr_radius;
};

// The M-tree of the reference, node for node: insertion descends to the
// child whose bounding-sphere centre is nearest, a node that reaches ten
// entries splits around its two farthest entries, bounding spheres are
// merged entry by entry with the reference's fp32 formula.
class LightTree {
  public:
    LightTree();
    ~LightTree();
    LightTree(const LightTree &) = delete;
    LightTree &operator=(const LightTree &) = delete;

    // 1 inserted; 0 refused as the reference refuses (radius not positive,
    // the sphere already present); -1 refused where the reference is
    // undefined (a centre already present, non-finite values).  The caller
    // checks the colour (check_light).
    int insert(const float center[3], float radius, const float color[3]);
    int remove(const float center[3], float radius);
    void clear();
    size_t count() const;
    // vox_get_light on the host
    void light(const float ambient[3], const float p[3], float out[3]) const;
    // depth-first records for the device
    void flatten(std::vector<LightNode> &nodes,
                 std::vector<LightRec> &spheres) const;

    struct Node;

  private:
    Node *root_;
    Node *alloc(bool leaf);
    void destroy(Node *n);
    Node *split(Node *n);
    Node *add(Node *n, const LightSphere &s);
    Node *find(Node *n, const float c[3], float r) const;
    bool has_center(const Node *n, const float c[3]) const;
    Node *drop_leaf(Node *leaf);
};

}  // namespace voxg
