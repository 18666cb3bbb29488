// Host-side tree construction and flattening (see tree_build.cpp).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/voxgpu.h"
#include "voxg_layout.h"

namespace voxg {

struct HostTree {
    std::vector<NodeRec> nodes;   // breadth-first records
    std::vector<uint32_t> counts; // dots_num of every record (host only)
    std::vector<const void *> src;  // reference node of every record, if any
    std::vector<float> dots;      // 3 floats per regular-leaf voxel
    uint64_t ndots = 0, voxel_ids = 0;
    uint64_t inner = 0, leaves = 0, dense = 0, overflow = 0;
    uint32_t depth = 0;
    float voxel[3] = {1.0f, 1.0f, 1.0f};
};

int build_tree(const float *xyz, size_t n, size_t stride, const float V[3],
               HostTree &out, std::string &err);
int flatten_reference(const struct vox_node *root, const float V[3],
                      HostTree &out, std::string &err);
// vox_make_tree into reference-layout nodes owned by the caller (freed with
// destroy_reference); nodes and leaf dot arrays come from aligned_alloc.
int build_reference(const float *xyz, size_t n, size_t stride, const float V[3],
                    struct vox_node **root, std::string &err);
void destroy_reference(struct vox_node *t);
uint64_t tree_hash(const HostTree &t);
int voxel_coords(const HostTree &t, size_t n, const uint32_t *ids,
                 float *xyz);

}  // namespace voxg

namespace voxg_internal {
// Records msg as the calling thread's voxg_last_error() and returns code
// (defined with the HIP code in voxgpu.hip).
int fail(int code, const std::string &msg);
}  // namespace voxg_internal
