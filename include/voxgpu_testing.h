/*
 * voxgpu_testing.h -- test hooks of libvoxgpu.  No counterpart in the
 * reference; for the library's own tests of its failure handling, which
 * force an error from the host so that no kernel ever fails on purpose.
 */
#ifndef VOXGPU_TESTING_H
#define VOXGPU_TESTING_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Arms a one-shot failure on the calling thread: the checkpoint-th host
 * step of the next render calls (1 = the first traversal launch after the
 * block pre-pass, then the replay and output launches of FAST and ADAPTIVE
 * renders) returns VOXG_EIO instead of launching, as a failed launch
 * would.  0 disarms.  Later calls must give the results of a render that
 * never failed. */
int voxg_test_fail_at(int checkpoint);

/* Caps the traversal stack bound of later launches (process-wide; 0
 * removes the cap).  Below a tree's true bound, rays that need more
 * pending entries lose the extra ones, finish as misses, and the stream's
 * overflow flag is set for voxg_synchronize and the next launch to
 * report.  Every access stays inside the buffers sized for the cap. */
int voxg_test_stack_bound(unsigned bound);

/* While hold is nonzero the server admits no request: submissions queue
 * up, and the first admission after the hold is lifted takes them all
 * as one group (up to the slot and memory limits), so that a test can
 * form a group of a chosen mix. */
struct voxg_server;
int voxg_test_server_hold(struct voxg_server *s, int hold);

/* The render pre-pass's cut of the tree that voxg_tree_build would make
 * from these voxels (host only, no device needed): up to cap boxes as 6
 * floats each (min xyz, max xyz) into out, their number into *count (0:
 * no cut, the tree is too deep or the cut would be the root alone). */
int voxg_test_plan_cut(const float *xyz, size_t n, size_t stride,
                       const float voxel[3], float *out, size_t cap,
                       size_t *count);

#ifdef __cplusplus
}
#endif
#endif
