/*
 * voxgpu.h -- C ABI of libvoxgpu, the MI355X (gfx950) implementation of the
 * voxtrees/voxrnd ray-casting path of shamazmazum/voxvision.
 *
 * Plain C: pointers, sizes and status codes, no C++ or torch types.  Device
 * buffers are HIP device pointers (hipMalloc'd or torch CUDA tensors); all
 * batched calls are asynchronous on the given stream unless stated.  Calls
 * on distinct streams may come from any threads at once; calls on one
 * stream must come from one thread at a time (the library keeps per-stream
 * scratch buffers, grown on demand after the stream's earlier work).
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   voxg_tree_upload / voxg_tree_build   vox_make_tree      src/voxtrees/tree.h:60
 *   voxg_tree_plan(_reference)           vox_make_tree      src/voxtrees/tree.h:60
 *                                        (host-side sizing, no device)
 *   voxg_tree_release                    vox_destroy_tree   src/voxtrees/tree.h:67
 *   voxg_tree_get_info                   vox_voxels_in_tree src/voxtrees/tree.h:72,
 *                                        vox_bounding_box   src/voxtrees/tree.h:77
 *   voxg_ray_batch          vox_ray_tree_intersection src/voxtrees/search.h:24-26
 *   voxg_ball_batch         vox_tree_ball_collidep    src/voxtrees/search.h:36-37
 *   voxg_render             vox_render per-pixel loop src/voxrnd/renderer.h:208,
 *                           with screen2world of the simple camera
 *                           src/voxrnd/simple-camera.c:21-31
 *   voxg_redraw             vox_redraw                src/voxrnd/renderer.h:169
 *   voxg_cd_collide_batch   vox_cd_collide            src/voxrnd/cd.h:47
 *   voxg_read_raw_data      vox_read_raw_data         src/voxtrees/datareader.h:35
 *   voxg_raw_select         its sample loop           src/voxtrees/datareader.c:49-75
 *   voxg_raw_count / _write the same loop in two phases src/voxtrees/datareader.c:49-75
 *   voxg_edit_insert        vox_insert_voxel          src/voxtrees/tree.h:96
 *   voxg_edit_delete        vox_delete_voxel          src/voxtrees/tree.h:114
 *   voxg_edit_rebuild       vox_rebuild_tree          src/voxtrees/tree.h:86
 */
#ifndef VOXGPU_H
#define VOXGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* HIP stream handle (same declaration as hip_runtime_api.h). */
typedef struct ihipStream_t *hipStream_t;

/* Status codes: 0 ok, negative errno-style values on failure. */
#define VOXG_OK 0
#define VOXG_EIO (-5)      /* a HIP call failed */
#define VOXG_ENOMEM (-12)  /* device or host allocation failed */
#define VOXG_ENODEV (-19)  /* no usable gfx950 device */
#define VOXG_EINVAL (-22)  /* bad argument */
#define VOXG_ERANGE (-34)  /* size beyond a documented limit */

/* Result code for a ray that hits nothing. */
#define VOXG_MISS 0xffffffffu

/* ---- reference tree layout (src/voxvision.h.in:45-56,79-83 with
 *      SSE_INTRIN; src/voxtrees/tree.h:15-38) ----
 * Only needed by callers that hand over a tree built by the reference's
 * vox_make_tree; voxg_tree_upload reads it and never writes it. */
#ifndef VOXGPU_NO_REFERENCE_LAYOUT
typedef float vox_dot[4] __attribute__((aligned(16)));
struct vox_box {
    vox_dot min;
    vox_dot max;
};
#define VOX_LEAF 1
#define VOX_DENSE_LEAF 2
#define VOX_LEAF_MASK 3
#define VOX_OVERFLOW 4
struct vox_node {
    struct vox_box bounding_box;
    unsigned int flags;
    unsigned int dots_num;
    union {
        vox_dot *dots;
        struct {
            vox_dot center;
            struct vox_node *children[8];
        } inner;
    } data;
};
#endif

/* Opaque handle of a flattened, device-resident, immutable tree. */
typedef struct voxg_tree voxg_tree;

typedef struct {
    uint64_t nodes, inner, leaves, dense, overflow, dots, voxel_ids;
    uint32_t depth;          /* levels, root = 1 */
    uint32_t stack_entries;  /* traversal stack bound: one entry per
                                inner level, depth-1 (at least 1) */
    uint64_t device_bytes;   /* flattened tree size in HBM */
    float bbox_min[3], bbox_max[3];
    float voxel[3];          /* voxel size snapshot taken at build/upload */
} voxg_tree_info;

/* Select the HIP device used by later calls on this thread (default 0). */
int voxg_init(int device);

/* Build a tree from n voxel min-corners, exactly as vox_make_tree(set, n)
 * would (tree.c:249-338), and upload it.  xyz holds n points of `stride`
 * floats each (3 = packed, 4 = vox_dot layout); it is copied, never
 * permuted.  voxel = vox_voxel at build time (NULL = {1,1,1}).
 * n == 0 yields an empty tree on which every query misses. */
int voxg_tree_build(const float *xyz, size_t n, size_t stride,
                    const float voxel[3], voxg_tree **out);

/* Flatten a tree built by the reference's vox_make_tree (reference layout
 * above) and upload it.  root may be NULL (empty tree).  voxel must be the
 * vox_voxel value the tree was built with (NULL = {1,1,1}). */
int voxg_tree_upload(const struct vox_node *root, const float voxel[3],
                     voxg_tree **out);

/* Host-only planning: build / flatten exactly as voxg_tree_build /
 * voxg_tree_upload would, without touching the device, and report the
 * tree's statistics (device_bytes = HBM the upload will take) and its
 * structure hash.  Either output may be NULL. */
int voxg_tree_plan(const float *xyz, size_t n, size_t stride,
                   const float voxel[3], voxg_tree_info *info,
                   uint64_t *hash);
int voxg_tree_plan_reference(const struct vox_node *root, const float voxel[3],
                             voxg_tree_info *info, uint64_t *hash);

int voxg_tree_release(voxg_tree *tree);
int voxg_tree_get_info(const voxg_tree *tree, voxg_tree_info *out);

/* Structure hash of the tree (identical for identical trees, whatever the
 * origin: build, upload). */
uint64_t voxg_tree_hash(const voxg_tree *tree);

/* Host-side lookups.  Voxel ids: regular-leaf dots are numbered first, in
 * depth-first preorder over children 0..7 and stored order inside a leaf;
 * dense-leaf voxels follow, leaf by leaf in the same preorder, as
 * base + (ix*dy + iy)*dz + iz.  Writes the min corner of each voxel. */
int voxg_tree_voxel_coords(const voxg_tree *tree, size_t n,
                           const uint32_t *ids, float *xyz);
/* Node index (as returned in leaf_out) -> the reference node it came from
 * (trees made by voxg_tree_upload only; NULL otherwise). */
const struct vox_node *voxg_tree_node_ptr(const voxg_tree *tree,
                                          uint32_t node);

/* Batched vox_ray_tree_intersection over n rays (search.c:8-153).
 * Inputs: SoA device arrays ox..dz (origin, direction).  Outputs (each may
 * be NULL): voxel_out = voxel id or VOXG_MISS; leaf_out = node index of the
 * returned leaf or VOXG_MISS; hx,hy,hz = hit point, left untouched on a
 * miss (search.c leaves res untouched).  Results equal the reference's
 * fp32 results bit for bit. */
int voxg_ray_batch(const voxg_tree *tree, size_t n, const float *ox,
                   const float *oy, const float *oz, const float *dx,
                   const float *dy, const float *dz, uint32_t *voxel_out,
                   uint32_t *leaf_out, float *hx, float *hy, float *hz,
                   hipStream_t stream);

/* voxg_ray_batch for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments and results as
 * voxg_ray_batch, bit for bit; the launch's control record, counters and
 * stack spill buffer live in a 32 MiB array that each GPU holds for the
 * life of the process, and no allocating call is made.  The launch runs
 * the VOXG_LAUNCH_RAYS configuration with fewer waves per CU when the
 * array cannot hold the spill entries of all of them; a tree too deep for
 * one wave per CU is rejected with VOXG_ERANGE.  Every call on a GPU shares
 * that GPU's array: calls that overlap on one GPU (other streams, other
 * threads) race and must be serialised by the caller. */
int voxg_ray_batch_noalloc(const voxg_tree *tree, size_t n, const float *ox,
                           const float *oy, const float *oz, const float *dx,
                           const float *dy, const float *dz,
                           uint32_t *voxel_out, uint32_t *leaf_out, float *hx,
                           float *hy, float *hz, hipStream_t stream);

/* Batched vox_tree_ball_collidep (search.c:155-183): out[i] = 1 when the
 * ball (c_i, r_i) touches a voxel (squared distance strictly below r^2). */
int voxg_ball_batch(const voxg_tree *tree, size_t n, const float *cx,
                    const float *cy, const float *cz, const float *r,
                    uint8_t *out, hipStream_t stream);

/* Editable host trees (vox_insert_voxel / vox_delete_voxel,
 * src/voxtrees/tree.h:96,114, tree.c:437-835).  The tree lives in host
 * memory in the reference's own node layout; edits follow the reference's
 * algorithm, so the structure after any edit sequence is the one the
 * reference would have.  voxg_edit_commit flattens and uploads the current
 * state as a new, independent device tree (a version): trees committed
 * earlier stay valid and unchanged until released.  Not thread-safe per
 * handle. */
typedef struct voxg_edit_tree voxg_edit_tree;
/* vox_make_tree over n points (stride 3 or 4 floats; voxel NULL = 1,1,1,
 * else every component > 0, or VOXG_EINVAL). */
int voxg_edit_create(const float *xyz, size_t n, size_t stride,
                     const float voxel[3], voxg_edit_tree **out);
/* vox_insert_voxel: the point is aligned down to the voxel grid first;
 * *inserted (may be NULL) = 1 when it was not in the tree and was added. */
int voxg_edit_insert(voxg_edit_tree *t, const float xyz[3], int *inserted);
/* vox_delete_voxel: *deleted = 1 when the voxel was there and was removed. */
int voxg_edit_delete(voxg_edit_tree *t, const float xyz[3], int *deleted);
/* vox_rebuild_tree (tree.c:422-432): the same voxels, rebuilt from scratch. */
int voxg_edit_rebuild(voxg_edit_tree *t);
/* The current root (NULL when empty), valid until the next edit. */
const struct vox_node *voxg_edit_root(const voxg_edit_tree *t);
size_t voxg_edit_voxels(const voxg_edit_tree *t);
/* Upload the current state as a new device tree (voxg_tree_upload). */
int voxg_edit_commit(const voxg_edit_tree *t, voxg_tree **out);
int voxg_edit_release(voxg_edit_tree *t);

/* vox_cd_collide (src/voxrnd/cd.c:38-55) for n bodies against one tree.
 * Body i is a ball of radius[i] at (px,py,pz)[i] that was last at
 * (lx,ly,lz)[i].  If the two positions differ (vox_dot_equalp,
 * voxvision.h.in:52: x, y, z all compare equal means "not moved") and the
 * ball collides with the tree (vox_tree_ball_collidep), the body goes back
 * to its last position; otherwise its last position becomes its position.
 * Positions are updated in place; collided (may be NULL) gets 1 for a body
 * moved back, else 0.  All pointers are device pointers.  The reference's
 * gravity (vox_cd_gravity, cd.c:33-36) is a stub and has no counterpart. */
int voxg_cd_collide_batch(const voxg_tree *tree, size_t n, float *px,
                          float *py, float *pz, float *lx, float *ly,
                          float *lz, const float *radius, uint8_t *collided,
                          hipStream_t stream);

/* Simple-camera descriptor (fields of struct vox_simple_camera,
 * simple-camera.c:10-17).  mul, xsub, ysub as set by set_window_size
 * (simple-camera.c:120-136); see voxg_camera_set_window_size. */
typedef struct {
    float position[3];
    float rotation[4]; /* quaternion w, i, j, k */
    float mul, xsub, ysub;
    float fov;         /* tan of the half angle; only used by the setter */
} voxg_camera;

/* simple-camera.c:120-136, on the host. */
void voxg_camera_set_window_size(voxg_camera *cam, int w, int h);

/* Quality flags: renderer.h:81-131 */
#define VOXG_QUALITY_BEST 0x0
#define VOXG_QUALITY_FAST 0x1
#define VOXG_QUALITY_ADAPTIVE 0x2
#define VOXG_QUALITY_RAY_MERGE 0x4
#define VOXG_QUALITY_RAY_MERGE_ACCURATE 0x8

/* ---- light manager (src/voxrnd/lights.h, lights.c; src/voxtrees/mtree.c)
 * An ambient light and shadowless light spheres, kept in the reference's
 * M-tree so that the lights containing a point are summed in the
 * reference's order.  Host object; a render that uses it uploads its
 * current state to the device first (synchronously, when it changed since
 * the last upload), so change lights outside graph capture.  Not
 * thread-safe per handle. */
typedef struct voxg_lights voxg_lights;
/* vox_make_light_manager: no lights, ambient light (1, 1, 1). */
int voxg_lights_create(voxg_lights **out);
/* vox_destroy_light_manager (waits for the device). */
int voxg_lights_release(voxg_lights *lights);
/* vox_set_ambient_light: *set = 1 when every component is in [0, 1] (the
 * light changed), else 0; set may be NULL. */
int voxg_lights_set_ambient(voxg_lights *lights, const float color[3],
                            int *set);
/* vox_insert_shadowless_light: *inserted = 1 when added, 0 when the
 * reference refuses it (a colour component outside [0, 1], radius not
 * positive, the same sphere present).  VOXG_EINVAL for what the
 * reference's M-tree cannot hold (a light with this centre present,
 * src/voxtrees/mtree.h) or a non-finite centre or radius. */
int voxg_lights_insert(voxg_lights *lights, const float center[3],
                       float radius, const float color[3], int *inserted);
/* vox_delete_shadowless_light: *deleted = 1 when the sphere was there. */
int voxg_lights_delete(voxg_lights *lights, const float center[3],
                       float radius, int *deleted);
/* vox_delete_shadowless_lights */
int voxg_lights_clear(voxg_lights *lights);
/* vox_shadowless_lights_number */
size_t voxg_lights_count(const voxg_lights *lights);
/* vox_get_light at n points (host): rgb[3i..3i+2] for xyz[3i..3i+2]. */
int voxg_lights_get(const voxg_lights *lights, size_t n, const float *xyz,
                    float *rgb);

typedef struct {
    unsigned quality;
    /* Colour (renderer.c:27-52): device pointer to a 32^3 byte texture
     * indexed x*1024 + y*32 + z; without lights a texel c maps to the
     * 32-bit pixel (c<<rshift)|(c<<gshift)|(c<<bshift)|amask, with lights
     * each channel is min(light, 1) * c truncated to 8 bits (the light of
     * voxg_lights_get at the hit point).  texture may be NULL when squares
     * is NULL. */
    const uint8_t *texture;
    uint32_t rshift, gshift, bshift, amask;
    /* vox_context_set_light_manager (renderer.h:184): NULL = no lights */
    const voxg_lights *lights;
} voxg_render_params;

/* The vox_render per-pixel loop for n_frames camera poses at W x H
 * in any quality mode (BEST, FAST, ADAPTIVE with or without ray merging).
 * (W % 16 == 0, H % 4 == 0, renderer.c:97-100; W, H < 65536 and
 * n_frames * W * H < 2^31, else VOXG_ERANGE).  cams: device array of
 * n_frames descriptors.  Pixel (x, y) of frame f lives at index
 * f*W*H + cs*16 + p with block cs = (y/4)*(W/4) + x/4 and p = (y%4)*4 + x%4
 * (the reference's square layout, renderer.h:14).  Outputs (each may be
 * NULL): squares = colours (pixels that miss are untouched), voxel_px =
 * voxel id or VOXG_MISS, hit_px = 3 planes of f*W*H floats (x, y, z),
 * untouched on a miss.  FAST and ADAPTIVE renders (with or without ray
 * merging) use a device scratch buffer of the stream, 20 bytes per pixel +
 * 4 per 4x4 block, allocated on the first call that needs more and kept
 * until the process ends. */
int voxg_render(const voxg_tree *tree, size_t n_frames,
                const voxg_camera *cams, int W, int H,
                const voxg_render_params *params, uint32_t *squares,
                uint32_t *voxel_px, float *hit_px, hipStream_t stream);

/* Traversal counters of one counting run (node bounding-box tests and
 * voxel-box tests), for the algorithmic-bytes model.  Synchronous. */
typedef struct {
    uint64_t node_visits;
    uint64_t voxel_tests;
    uint64_t rays;            /* searches started at the root */
    uint64_t leaf_traversals; /* FAST-mode searches started at a leaf */
    uint64_t fast_blocks, best_blocks, merge_blocks; /* render modes */
    uint64_t stack_spills;    /* pending pairs moved from the LDS ring to
                                 the HBM spill buffer (launch-config
                                 dependent; not a reference counter) */
} voxg_counts;
int voxg_ray_batch_count(const voxg_tree *tree, size_t n, const float *ox,
                         const float *oy, const float *oz, const float *dx,
                         const float *dy, const float *dz, voxg_counts *out,
                         hipStream_t stream);
int voxg_ball_batch_count(const voxg_tree *tree, size_t n, const float *cx,
                          const float *cy, const float *cz, const float *r,
                          voxg_counts *out, hipStream_t stream);
int voxg_render_count(const voxg_tree *tree, size_t n_frames,
                      const voxg_camera *cams, int W, int H,
                      const voxg_render_params *params, voxg_counts *out,
                      hipStream_t stream);

/* ---- batched form: many trees in one persistent launch ----
 * A batch is a device work queue of items, each a tree and a run of
 * consecutive rays [first, first+count) of the ray arrays passed at launch.
 * voxg_batch_create uploads the queue once and allocates the batch's own
 * launch workspace (call it outside graph capture); voxg_batch_ray is a
 * single kernel launch that needs no memset and uses only buffers owned
 * by the batch, so it can be captured in a HIP graph and replayed for as
 * long as the batch lives, whatever other calls run in between.  Results
 * are those of voxg_ray_batch on each item's tree.  Trees must outlive the
 * batch; the VOXG_LAUNCH_BATCH launch configuration is taken at creation.
 * Launches of one batch must not overlap (one stream, or synchronise);
 * voxg_batch_release waits for the device. */
typedef struct voxg_batch voxg_batch;
typedef struct {
    const voxg_tree *tree;
    size_t first;
    size_t count;
} voxg_batch_item;
int voxg_batch_create(size_t n_items, const voxg_batch_item *items,
                      voxg_batch **out);
int voxg_batch_ray(const voxg_batch *batch, const float *ox, const float *oy,
                   const float *oz, const float *dx, const float *dy,
                   const float *dz, uint32_t *voxel_out, uint32_t *leaf_out,
                   float *hx, float *hy, float *hz, hipStream_t stream);
int voxg_batch_release(voxg_batch *batch);

/* ---- serving: a request loop over resident trees ----
 * A server keeps trees resident, accepts render and ball-query requests
 * from any thread, admits them in arrival order while request slots and
 * its device-memory budget allow, and launches every request admitted
 * together as one group (one multi-job render launch plus one multi-job
 * ball launch) on one of several streams, so groups that arrive while
 * others run overlap on the device.  Results are copied into the caller's
 * host buffers (pass pinned memory, e.g. hipHostMalloc, for asynchronous
 * copies) and equal those of voxg_render / voxg_ball_batch on the same
 * inputs.  Ball coordinates are copied at submission into 64 MiB of
 * pinned host memory the server owns (into pageable memory while that is
 * full).  A request that can never fit the budget is rejected
 * (VOXG_ENOMEM). */
typedef struct voxg_server voxg_server;
typedef struct {
    size_t hbm_budget;      /* device bytes for request buffers;
                               0 = half of the free HBM at creation */
    int max_inflight;       /* requests admitted at once; 0 = 256 */
    int streams;            /* launch streams; 0 = 4 */
    const uint8_t *texture; /* host 32^3 colour texture; NULL = no colours */
    uint32_t rshift, gshift, bshift, amask;
} voxg_server_config;
typedef struct {
    int tree;               /* id from voxg_server_add_tree */
    voxg_camera cam;        /* host copy */
    int W, H;               /* W % 16 == 0, H % 4 == 0 */
    unsigned quality;       /* VOXG_QUALITY_* */
    uint32_t *voxel_out;    /* host, W*H in the square layout, or NULL */
    float *hit_out;         /* host, 3 planes of W*H (unspecified where
                               the pixel missed), or NULL */
    uint32_t *color_out;    /* host, W*H (needs a texture; 0 where the
                               pixel missed, as after vox_redraw), or NULL */
} voxg_render_request;
typedef struct {
    int tree;
    size_t n;
    const float *cx, *cy, *cz, *r; /* host; copied at submission */
    uint8_t *out;                  /* host, n */
} voxg_ball_request;
typedef struct {
    int status;             /* 0 ok, < 0 error code */
    double t_submit, t_admit, t_done; /* seconds on voxg_server_now's clock */
    uint64_t group;         /* launch group the request ran in */
    uint32_t group_size;    /* requests in that group */
} voxg_request_info;
typedef struct {
    uint64_t submitted, completed, rejected, groups;
    uint64_t peak_inflight, peak_groups_inflight;
    uint64_t peak_bytes, budget;
} voxg_server_stats;
int voxg_server_create(const voxg_server_config *cfg, voxg_server **out);
int voxg_server_add_tree(voxg_server *s, const voxg_tree *tree, int *id);
/* Tree versioning: requests submitted after this call use `tree` for id;
 * requests submitted before keep the version current at their submission
 * (e.g. a voxg_edit_commit of an edited scene replaces the previous one).
 * Release a replaced version only after the requests submitted before the
 * replacement have completed (voxg_server_wait). */
int voxg_server_replace_tree(voxg_server *s, int id, const voxg_tree *tree);
int voxg_server_submit_render(voxg_server *s, const voxg_render_request *r,
                              uint64_t *ticket);
int voxg_server_submit_balls(voxg_server *s, const voxg_ball_request *r,
                             uint64_t *ticket);
/* Blocks until the request is done (or timeout_s passes: VOXG_EIO with
 * info->status untouched); a finished ticket is forgotten after wait. */
int voxg_server_wait(voxg_server *s, uint64_t ticket, double timeout_s,
                     voxg_request_info *info);
int voxg_server_get_stats(voxg_server *s, voxg_server_stats *st);
double voxg_server_now(const voxg_server *s);
/* Waits for the work in flight, then frees the server (not the trees). */
int voxg_server_destroy(voxg_server *s);

/* vox_redraw (src/voxrnd/renderer.c:321-327): copy_squares
 * (src/voxrnd/copy-helper.c:5-40) of n_frames frames from the square
 * layout of voxg_render to row-major pixels,
 *   pixels[f*W*H + y*W + x] = squares[f*W*H + ((y/4)*(W/4) + x/4)*16
 *                                     + (y%4)*4 + x%4],
 * then, when clear != 0, zeroes the squares as vox_redraw does, so the
 * next frame starts black.  W % 16 == 0, H % 4 == 0; both buffers are
 * 16-byte aligned device memory.  The SDL surface update is the caller's. */
int voxg_redraw(uint32_t *squares, uint32_t *pixels, size_t n_frames, int W,
                int H, int clear, hipStream_t stream);

/* The sample loop of vox_read_raw_data (src/voxtrees/datareader.c:49-75)
 * on the device: samples holds dim[0] x dim[1] x dim[2] little-endian
 * samples of samplesize (1..4) bytes, i slowest and k fastest; the sample
 * at (i, j, k) is selected when lo <= value < hi (the reference's
 * callback, as used by its callers: `sample > t` is lo = t + 1, hi = 2^32;
 * the Lua ranged reader is lo = min, hi = max).  Selected voxels' min
 * corners (i*V0, j*V1, k*V2) go to xyz_out (packed xyz, device) in the
 * reader's order, at most `capacity` of them; *count_out (device) gets the
 * number selected.  xyz_out may be NULL to count only.  Fewer than 2^32
 * samples.  voxel = vox_voxel (NULL = {1,1,1}). */
int voxg_raw_select(const void *samples, const unsigned dim[3],
                    unsigned samplesize, uint64_t lo, uint64_t hi,
                    const float voxel[3], float *xyz_out, size_t capacity,
                    unsigned long long *count_out, hipStream_t stream);

/* voxg_raw_select in two phases around the caller's allocation of the
 * exact output (the same sample loop, datareader.c:49-75, split where the
 * reference grows its dot array).  voxg_raw_count reads the samples once
 * and leaves the selection in `work` (device, 16-byte aligned,
 * voxg_raw_workspace_bytes(dim) bytes: a bit per sample and an offset per
 * 4096 samples) and the number selected in *count_out (device, or NULL).
 * voxg_raw_write then writes the selected voxels' corners from `work`
 * alone, as voxg_raw_select does (at most `capacity` of them); it does not
 * read the samples, so `work` must come from voxg_raw_count on a volume of
 * the same dim, and stays valid until it is reused.
 * voxg_raw_workspace_bytes returns 0 for a volume of 2^32 samples or more. */
size_t voxg_raw_workspace_bytes(const unsigned dim[3]);
int voxg_raw_count(const void *samples, const unsigned dim[3],
                   unsigned samplesize, uint64_t lo, uint64_t hi, void *work,
                   unsigned long long *count_out, hipStream_t stream);
int voxg_raw_write(const unsigned dim[3], const float voxel[3],
                   const void *work, float *xyz_out, size_t capacity,
                   hipStream_t stream);

/* vox_read_raw_data (datareader.c:11-85) with the test lo <= value < hi:
 * reads the file (its size must be dim[0]*dim[1]*dim[2]*samplesize,
 * "Wrong size of dataset" otherwise), selects the voxels on the device
 * (voxg_raw_select), builds the tree on the host (voxg_tree_build) and
 * uploads it.  Synchronous. */
int voxg_read_raw_data(const char *filename, const unsigned dim[3],
                       unsigned samplesize, uint64_t lo, uint64_t hi,
                       const float voxel[3], voxg_tree **out);

/* Launch configuration of the traversal kernels (autotuned by
 * voxvision_amd.autotune).  Each kind of launch has one of its own, with a
 * default that is the fastest measured for it on MI355X:
 *   VOXG_LAUNCH_RAYS         voxg_ray_batch, voxg_ray_batch_count
 *   VOXG_LAUNCH_RENDER       voxg_render and voxg_render_count in FAST and
 *                            ADAPTIVE modes (one 4x4 block per unit)
 *   VOXG_LAUNCH_RENDER_BEST  the same in BEST mode (one pixel per unit by
 *                            default, handed out to large groups of lanes)
 *   VOXG_LAUNCH_BATCH        voxg_batch_ray and the serving loop's launches
 * Fields:
 *   stack_lds     stack entries per lane kept in LDS (power of two 2..32,
 *                 20 bytes each; at most 160 KiB per workgroup); deeper
 *                 stacks spill to HBM and stay exact
 *   block         threads per workgroup (64, 128, 192 or 256)
 *   waves_per_cu  persistent waves launched per CU (1..64)
 *   refill        minimum idle lanes before a wave fetches new work (1..64)
 *   chunk         consecutive rays a lane traces before it needs new work
 *                 in voxg_ray_batch and the batched form (1..4096; lowered
 *                 for a launch too small to give every persistent lane a
 *                 unit; a render lane takes one 4x4 block, or part of
 *                 one in BEST mode, see split_best)
 *   render_cull   1: voxg_render first finishes, in a coherent pre-pass,
 *                 the 4x4 blocks whose rays all miss the scene box; 0: every
 *                 block goes through the traversal
 *   leaf_batch    lanes of a wave that wait at a leaf before the wave
 *                 searches its voxels (1..64; fewer when no lane of the wave
 *                 can go on traversing)
 *   finish_batch  lanes of a wave whose traversal ended before the wave
 *                 stores their results and starts their next rays (1..64;
 *                 fewer when no other lane is left)
 *   split_best    BEST-mode voxg_render: a lane takes 16 >> split_best
 *                 pixels of a 4x4 block at a time (0..4; the other modes
 *                 always take whole blocks)
 * A kind ignores the fields it has no use for (chunk in renders, split_best
 * outside BEST renders, render_cull outside renders).  None of them changes
 * any result. */
typedef struct {
    int stack_lds;
    int block;
    int waves_per_cu;
    int refill;
    int chunk;
    int render_cull;
    int leaf_batch;
    int finish_batch;
    int split_best;
} voxg_launch_config;
#define VOXG_LAUNCH_RAYS 0
#define VOXG_LAUNCH_RENDER 1
#define VOXG_LAUNCH_RENDER_BEST 2
#define VOXG_LAUNCH_BATCH 3
#define VOXG_LAUNCH_KINDS 4
/* Sets the configuration of every kind of launch. */
int voxg_set_launch_config(const voxg_launch_config *cfg);
/* The configuration of VOXG_LAUNCH_RAYS. */
int voxg_get_launch_config(voxg_launch_config *cfg);
/* One kind's configuration (VOXG_EINVAL for an unknown kind). */
int voxg_set_kind_launch_config(int kind, const voxg_launch_config *cfg);
int voxg_get_kind_launch_config(int kind, voxg_launch_config *cfg);

/* Human-readable text of the last error on this thread. */
const char *voxg_last_error(void);

/* Waits for the work queued on the stream, then reports VOXG_EIO if a
 * traversal launched on it since the last check exceeded its stack bound
 * (one entry per inner level, which the search order never exceeds; the
 * affected rays would be left as misses).  The flag is also
 * checked, without waiting, at the start of every later traversal launch
 * on the stream and by the counting calls.  Replaces the host-side wait a
 * caller of the reference's synchronous searches does not need
 * (src/voxtrees/search.h:24-37 return when done). */
int voxg_synchronize(hipStream_t stream);

/* Frees the buffers the library keeps for the stream (on the current
 * device), after waiting for it.  Call before hipStreamDestroy on a stream
 * the library has used; a later call on the stream allocates them again. */
int voxg_release_stream(hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif
