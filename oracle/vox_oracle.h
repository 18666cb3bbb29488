/*
 * vox_oracle.h -- CPU oracle for the voxtrees/voxrnd ray-casting path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (shamazmazum/voxvision, SSE build) used as the checker for the
 * HIP implementation and as the CPU baseline.  Only tests/, the smoke entry
 * point and the benchmark's cpu_baseline leg may load it.  The product
 * library (libvoxgpu) never links or calls anything in this directory.
 *
 * Every function names the reference file:line it restates.  Floating-point
 * rules followed (the reference's RELEASE build: clang, -march=native,
 * SSE_INTRIN, default -ffp-contract=on):
 *   - fused multiply-add exactly where clang contracts a*b+c inside one
 *     expression; every other operation is rounded separately
 *     (this file is compiled with -ffp-contract=off);
 *   - SSE min/max semantics: min(a,b) = a<b ? a : b, max(a,b) = a>b ? a : b;
 *   - correctly rounded IEEE division, no flush-to-zero.
 *
 * Pinning (how this restatement is checked, see tests/):
 *   - reference known answers: tests/tests.c quaternion/rotation table,
 *     camera tests, tree invariants (check_tree), the commit-676d50c ray;
 *   - closed forms of benchmarks/voxtrees/hit-box.c and ball-intersection.c;
 *   - FP64 brute force over all voxels (hit/miss, hit voxel, hit point);
 *   - brute-force ball test with the same fp32 formula (exact by monotonicity).
 *   Parity unpinned: exact dot order inside leaves and exact hit-point bits
 *   against the compiled reference (the reference cannot be built here);
 *   render colours (depend on vn3d/SDL, taken as inputs).
 */
#ifndef VOX_ORACLE_H
#define VOX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reference data layout: src/voxvision.h.in:45-56 (SSE_INTRIN), :79-83 */
typedef float orc_dot[4] __attribute__((aligned(16)));

struct orc_box {
    orc_dot min;
    orc_dot max;
};

/* Reference node flags and node layout: src/voxtrees/tree.h:15-38 */
#define ORC_LEAF 1
#define ORC_DENSE_LEAF 2
#define ORC_LEAF_MASK 3
#define ORC_OVERFLOW 4

struct orc_node {
    struct orc_box bounding_box;
    unsigned int flags;
    unsigned int dots_num;
    union {
        orc_dot *dots;
        struct {
            orc_dot center;
            struct orc_node *children[8];
        } inner;
    } data;
};

/* Traversal counters (mirror the DTrace probes in search.c:25-149). */
typedef struct {
    uint64_t node_visits;   /* hit_box calls on node bounding boxes */
    uint64_t voxel_tests;   /* hit_box / box_ball_interp calls on leaf voxels */
} orc_counters;

#define ORC_MISS 0xffffffffu

/* Global voxel size: src/voxtrees/tree.c:11, params.h:38 */
void orc_set_voxel(float x, float y, float z);
void orc_get_voxel(float out[3]);

/* ---- geometry primitives (src/voxtrees/geom-sse.c) ---- */
float orc_abs_metric(const orc_dot a, const orc_dot b);
float orc_sqr_metric(const orc_dot a, const orc_dot b);
float orc_sqr_norm(const orc_dot a);
int orc_get_subspace_idx(const orc_dot center, const orc_dot dot);
int orc_get_corrected_subspace_idx(const orc_dot center, const orc_dot dot,
                                   const orc_dot dir);
int orc_hit_box(const struct orc_box *box, const orc_dot origin,
                const orc_dot dir, orc_dot res);
int orc_hit_plane_within_box(const orc_dot origin, const orc_dot dir,
                             const orc_dot planedot, int planenum,
                             orc_dot res, const struct orc_box *box);
int orc_box_ball_interp(const struct orc_box *box, const orc_dot center,
                        float radius);
int orc_dense_set_p(const struct orc_box *box, size_t n);

/* ---- tree construction (src/voxtrees/tree.c:31-122, 212-338) ---- */
struct orc_node *orc_make_tree(orc_dot *set, size_t n);
void orc_destroy_tree(struct orc_node *tree);
size_t orc_voxels_in_tree(const struct orc_node *tree);
/* Flat convenience: build from n x 4 floats (copied, the copy is permuted) */
struct orc_node *orc_make_tree_from(const float *xyzw, size_t n);

/* Tree statistics and canonical voxel numbering.
 * Voxel ids: regular-leaf dots first, in depth-first preorder over
 * children 0..7 and stored order inside a leaf; then dense-leaf voxels,
 * leaf by leaf in the same preorder, as base + (ix*dy + iy)*dz + iz for a
 * dense leaf of dimensions (dx,dy,dz). */
typedef struct {
    uint64_t nodes, inner, leaves, dense, overflow, dots, voxel_ids;
    uint32_t depth;
} orc_tree_stats;
void orc_tree_stats_get(const struct orc_node *tree, orc_tree_stats *st);
/* Write every voxel id's min corner (3 floats per id) and a 64-bit
 * structure hash (layout-independent) of the tree. */
void orc_tree_voxel_table(const struct orc_node *tree, float *xyz_out);
uint64_t orc_tree_hash(const struct orc_node *tree);

/* ---- search (src/voxtrees/search.c) ---- */
const struct orc_node *orc_ray_tree_intersection(const struct orc_node *tree,
                                                 const orc_dot origin,
                                                 const orc_dot dir,
                                                 orc_dot res);
int orc_tree_ball_collidep(const struct orc_node *tree, const orc_dot center,
                           float radius);

/* Batched loops over SoA inputs (single thread unless nthreads > 1).
 * voxel_out: canonical voxel id or ORC_MISS; hit_out: 3 x n floats (SoA),
 * left untouched on a miss.  counters may be NULL. */
void orc_ray_batch(const struct orc_node *tree, size_t n, const float *ox,
                   const float *oy, const float *oz, const float *dx,
                   const float *dy, const float *dz, uint32_t *voxel_out,
                   float *hx, float *hy, float *hz, int nthreads,
                   orc_counters *counters);
void orc_ball_batch(const struct orc_node *tree, size_t n, const float *cx,
                    const float *cy, const float *cz, const float *r,
                    uint8_t *out, int nthreads, orc_counters *counters);

/* ---- camera (src/voxrnd/simple-camera.c, vect-ops-sse.c) ---- */
typedef struct {
    float position[3];
    float rotation[4]; /* w, i, j, k */
    float fov;         /* tan of the half angle (simple-camera.c:151) */
    float mul, xsub, ysub;
} orc_camera;
void orc_camera_set_window_size(orc_camera *cam, int w, int h);
void orc_rotate_vector(const float q[4], const float v[3], float res[3]);
void orc_rotate_vector_q(const float q[4], const float v[4], float res[4]);
void orc_quat_mul(const float q1[4], const float q2[4], float res[4]);
void orc_screen2world(const orc_camera *cam, int sx, int sy, orc_dot ray);

/* ---- renderer per-pixel loop (src/voxrnd/renderer.c:23-52, 197-319) ---- */
#define ORC_QUALITY_BEST 0
#define ORC_QUALITY_FAST 1
#define ORC_QUALITY_ADAPTIVE 2
#define ORC_QUALITY_RAY_MERGE 4
#define ORC_QUALITY_RAY_MERGE_ACCURATE 8

typedef struct orc_lights orc_lights;
typedef struct {
    const uint8_t *texture; /* 32^3 bytes, index x*1024 + y*32 + z */
    uint32_t rshift, gshift, bshift, amask; /* SDL_MapRGB for a 32-bit format */
    const orc_lights *lights; /* NULL: no light manager (renderer.c:38) */
} orc_shading;

typedef struct {
    uint64_t blocks, fast_blocks, best_blocks, merge_blocks;
    uint64_t root_traversals, leaf_traversals;
    orc_counters c;
} orc_render_stats;

/* squares: (W/4)*(H/4) blocks of 16 uint32 (square[p], p = y*4 + x).
 * Pixels that miss are left untouched.  Optional per-pixel outputs in the
 * same square layout (index cs*16 + p): voxel id (ORC_MISS on miss) and
 * hit point (3 planes of W*H floats, untouched on a miss). */
int orc_render(const struct orc_node *scene, const orc_camera *cam, int W,
               int H, unsigned quality, const orc_shading *shade,
               uint32_t *squares, uint32_t *voxel_px, float *hit_px,
               int nthreads, orc_render_stats *stats);

/* orc_render with a trace of the loop's decisions (each output may be
 * NULL): block_flags[cs] = bit 0 block traced in BEST mode, bit 1 every
 * other pixel merged; px_src[i] (square layout) = 0 root search (corners
 * included), 1 answered by a try of the previous pixel's leaf, 2 a merged
 * copy of the previous pixel; try_box = 6 planes of W*H floats (min xyz,
 * max xyz) holding the tried leaf's box where px_src is 1. */
int orc_render_trace(const struct orc_node *scene, const orc_camera *cam,
                     int W, int H, unsigned quality,
                     const orc_shading *shade, uint32_t *squares,
                     uint32_t *voxel_px, float *hit_px, int nthreads,
                     orc_render_stats *stats, uint8_t *block_flags,
                     uint8_t *px_src, float *try_box);

/* Inner nodes in depth-first preorder: 9 floats each (box min xyz, box
 * max xyz, centre xyz) and the voxel count (dots_num, tree.c:270).  Writes
 * at most cap of them; returns how many there are. */
size_t orc_tree_inner_nodes(const struct orc_node *tree, float *out,
                            uint64_t *counts, size_t cap);

/* For each ray, the most inner nodes on the search path of
 * vox_ray_tree_intersection that still had a child behind a dividing plane
 * to visit, over the search's node visits (single thread). */
void orc_ray_max_pending(const struct orc_node *tree, size_t n,
                         const float *ox, const float *oy, const float *oz,
                         const float *dx, const float *dy, const float *dz,
                         uint32_t *out);

/* Ray directions of every pixel of a W x H frame (orc_screen2world), as 3
 * planes of W*H floats in the square layout. */
void orc_camera_rays(const orc_camera *cam, int W, int H, float *dxyz);

/* ---- lights: src/voxrnd/lights.c, src/voxtrees/mtree.c ---- */
typedef struct {
    float center[3];
    float color[3];
    float radius, sqr_radius;
} orc_sphere;
orc_lights *orc_lights_create(void);                 /* ambient 1,1,1 */
void orc_lights_destroy(orc_lights *l);
int orc_lights_set_ambient(orc_lights *l, const float color[3]);
int orc_lights_insert(orc_lights *l, const float center[3], float radius,
                      const float color[3]);
int orc_lights_delete(orc_lights *l, const float center[3], float radius);
void orc_lights_clear(orc_lights *l);
size_t orc_lights_count(const orc_lights *l);
/* vox_get_light: ambient plus, for each sphere containing p in the
 * M-tree's walk order, (1 - min(r, |p - c|) / r) * colour. */
void orc_get_light(const orc_lights *l, const float p[3], float light[3]);
/* Spheres in the M-tree's depth-first order; returns the count. */
size_t orc_lights_spheres(const orc_lights *l, orc_sphere *out, size_t cap);
unsigned orc_lights_height(const orc_lights *l);

/* ---- FP64 brute force (independent of the tree) ---- */
/* For each ray: first voxel hit along o + t d, t >= 0 (closed boxes, an
 * origin inside a box hits at t=0).  voxel_idx: index into dots or -1.
 * tie: 1 when the result is ambiguous at relative tolerance eps (two
 * entries within eps, a grazing voxel, or the hit point within eps of a
 * voxel edge). */
void orc_brute64_rays(const float *dots_xyz, size_t nd, const float voxel[3],
                      size_t n, const float *ox, const float *oy,
                      const float *oz, const float *dx, const float *dy,
                      const float *dz, double eps, int32_t *voxel_idx,
                      double *hit_xyz, uint8_t *tie, int nthreads);
/* orc_brute64_rays for voxels on the grid (dots multiples of voxel), by an
 * FP64 grid walk: the same voxel and hit point; near misses more than two
 * cells past the first hit do not count as ties.  Returns 0, or -22 when a
 * dot is off the grid, -34 / -12 when the grid is too large. */
int orc_dda64_rays(const float *dots_xyz, size_t nd, const float voxel[3],
                   size_t n, const float *ox, const float *oy,
                   const float *oz, const float *dx, const float *dy,
                   const float *dz, double eps, int32_t *voxel_idx,
                   double *hit_xyz, uint8_t *tie, int nthreads);
/* Brute-force ball test with the reference fp32 formula over all voxels. */
void orc_brute_balls(const float *dots_xyz, size_t nd, const float voxel[3],
                     size_t n, const float *cx, const float *cy,
                     const float *cz, const float *r, uint8_t *out,
                     int nthreads);

/* check_tree invariants of tests/tests.c:239-300; 0 = all hold. */
int orc_check_tree(const struct orc_node *tree);

/* Number of online CPUs usable by this process (sched_getaffinity). */
int orc_cpu_count(void);

/* ---- tree edits (tree.c:437-835) ---- */
/* vox_insert_voxel / vox_delete_voxel on *tree (which may become NULL or
 * a new root): the voxel is aligned to the grid first (vox_align); insert
 * returns 1 when the voxel was not in the tree and was added, delete 1
 * when it was there and was removed.  Voxel ids are reassigned after the
 * edit.  The reference's behaviour for an overflow leaf that gains a
 * voxel is undefined (a buffer of VOX_MAX_DOTS+1 dots receives all of
 * them); here that leaf is rebuilt from its dots and the voxel. */
int orc_insert_voxel(struct orc_node **tree, const float xyz[3]);
int orc_delete_voxel(struct orc_node **tree, const float xyz[3]);
/* vox_make_dense_leaf (tree.c:229-240): a dense leaf filling the box. */
struct orc_node *orc_make_dense_leaf_box(const float mn[3], const float mx[3]);
/* flags of the root (ORC_LEAF / ORC_DENSE_LEAF / ORC_OVERFLOW bits, 0 for
 * an inner node), -1 for an empty tree */
int orc_tree_flags(const struct orc_node *tree);
/* vox_rebuild_tree (tree.c:422-432): vox_make_tree over flatten_tree. */
struct orc_node *orc_rebuild_tree(const struct orc_node *tree);

/* ---- around the render loop and the tree ---- */
/* copy_squares (copy-helper.c:5-40) for one frame: 4x4 squares (16
 * uint32 each, ws x hs of them) to a row-major (4 ws) x (4 hs) image. */
void orc_copy_squares(const uint32_t *squares, uint32_t *pixels, unsigned ws,
                      unsigned hs);
/* vox_cd_collide (cd.c:38-55) for n bodies, one after the other: a body
 * whose position differs from its last position (vox_dot_equalp,
 * voxvision.h.in:52) is moved back when a ball of its radius collides with
 * the tree, else its last position becomes the position.  pos and last
 * are n x 3, updated in place; collided[i] = 1 when body i was moved
 * back. */
void orc_cd_collide(const struct orc_node *tree, size_t n, float *pos,
                    float *last, const float *radius, uint8_t *collided);
/* The sample loop of vox_read_raw_data (datareader.c:49-75) over an
 * in-memory volume of dim[0] x dim[1] x dim[2] little-endian samples of
 * samplesize bytes, with the test lo <= sample < hi.  Writes the selected
 * voxels' min corners (i*V0, j*V1, k*V2) in the reader's order into
 * xyz_out (n x 3, room for every sample) and returns their number. */
size_t orc_read_raw(const uint8_t *data, const unsigned dim[3],
                    unsigned samplesize, uint64_t lo, uint64_t hi,
                    float *xyz_out);

#ifdef __cplusplus
}
#endif
#endif
