/*
 * vox_oracle.c -- CPU oracle (restatement of the reference algorithm).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU
 * baseline.  Nothing in the product library links this file.
 * Compiled with -ffp-contract=off; fmaf() marks the sites where the
 * reference's clang build contracts a*b+c into one fused operation.
 * All citations are relative to the reference repository root.
 */
#define _GNU_SOURCE
#include "vox_oracle.h"

#include <limits.h>
#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* src/voxtrees/tree.c:11 -- vox_voxel = {1,1,1} (lane 3 zero). */
static orc_dot orc_voxel = {1.0f, 1.0f, 1.0f, 0.0f};

void orc_set_voxel(float x, float y, float z)
{
    orc_voxel[0] = x;
    orc_voxel[1] = y;
    orc_voxel[2] = z;
    orc_voxel[3] = 0.0f;
}

void orc_get_voxel(float out[3])
{
    out[0] = orc_voxel[0];
    out[1] = orc_voxel[1];
    out[2] = orc_voxel[2];
}

/* MINPS / MAXPS lane semantics (the second operand wins on ties, zeros of
 * either sign and NaN). */
static inline float smin(float a, float b) { return a < b ? a : b; }
static inline float smax(float a, float b) { return a > b ? a : b; }

/* ------------------------------------------------------------------ */
/* Geometry: src/voxtrees/geom-sse.c                                   */
/* ------------------------------------------------------------------ */

/* geom-sse.c:4-10 -- |a-b| per lane, summed as (x + y) + z. */
float orc_abs_metric(const orc_dot a, const orc_dot b)
{
    float x = fabsf(a[0] - b[0]);
    float y = fabsf(a[1] - b[1]);
    float z = fabsf(a[2] - b[2]);
    return (x + y) + z;
}

/* geom-sse.c:12-17 */
float orc_sqr_metric(const orc_dot a, const orc_dot b)
{
    float x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    x = x * x;
    y = y * y;
    z = z * z;
    return (x + y) + z;
}

/* geom-sse.c:19-23 */
float orc_sqr_norm(const orc_dot a)
{
    float x = a[0] * a[0], y = a[1] * a[1], z = a[2] * a[2];
    return (x + y) + z;
}

/* geom-sse.c:64-68 -- bit i set when dot[i] < center[i]. */
int orc_get_subspace_idx(const orc_dot center, const orc_dot dot)
{
    int r = 0, i;
    for (i = 0; i < 3; i++) r |= (dot[i] < center[i]) << i;
    return r;
}

/* geom-sse.c:70-82 -- on the dividing plane the direction decides. */
int orc_get_corrected_subspace_idx(const orc_dot center, const orc_dot dot,
                                   const orc_dot dir)
{
    int r = 0, i;
    for (i = 0; i < 3; i++) {
        int b = (dot[i] == center[i]) ? (dir[i] < 0.0f) : (dot[i] < center[i]);
        r |= b << i;
    }
    return r;
}

/* geom-sse.c:84-116 (with fit_into_box :34-40).
 * An origin inside or on the box is a hit at the origin.  Otherwise the
 * candidate distance per axis is (fit-o)/d (0 where d == 0), the largest
 * one wins (t <= 0 is a miss), lanes whose distance equals t take the
 * clamped coordinate, the others fma(t, d, o); any lane outside the closed
 * box is a miss. */
int orc_hit_box(const struct orc_box *box, const orc_dot origin,
                const orc_dot dir, orc_dot res)
{
    float fit[3], sub[3], dist[3], m[3], hit[3];
    int i;
    for (i = 0; i < 3; i++) {
        fit[i] = smin(smax(origin[i], box->min[i]), box->max[i]);
        sub[i] = fit[i] - origin[i];
    }
    if (sub[0] == 0.0f && sub[1] == 0.0f && sub[2] == 0.0f) {
        res[0] = origin[0];
        res[1] = origin[1];
        res[2] = origin[2];
        res[3] = 0.0f;
        return 1;
    }
    for (i = 0; i < 3; i++)
        dist[i] = (dir[i] == 0.0f) ? 0.0f : sub[i] / dir[i];
    /* The shuffle/max network of :101-105, lane by lane. */
    m[0] = smax(smax(dist[0], dist[2]), smax(dist[0], dist[1]));
    m[1] = smax(smax(dist[1], dist[0]), smax(dist[1], dist[2]));
    m[2] = smax(smax(dist[2], dist[1]), smax(dist[2], dist[0]));
    if (m[0] <= 0.0f) return 0;
    for (i = 0; i < 3; i++) {
        hit[i] = (m[i] == dist[i]) ? fit[i] : fmaf(m[i], dir[i], origin[i]);
        if (hit[i] < box->min[i] || hit[i] > box->max[i]) return 0;
    }
    res[0] = hit[0];
    res[1] = hit[1];
    res[2] = hit[2];
    res[3] = 0.0f;
    return 1;
}

/* geom-sse.c:118-142 */
int orc_hit_plane_within_box(const orc_dot origin, const orc_dot dir,
                             const orc_dot planedot, int planenum,
                             orc_dot res, const struct orc_box *box)
{
    float k = planedot[planenum] - origin[planenum];
    float r[3];
    int i;
    if (k == 0.0f || ((dir[planenum] < 0.0f) != (k < 0.0f))) return 0;
    k = k / dir[planenum];
    for (i = 0; i < 3; i++) {
        r[i] = fmaf(k, dir[i], origin[i]);
        if (r[i] < box->min[i] || r[i] > box->max[i]) return 0;
    }
    res[0] = r[0];
    res[1] = r[1];
    res[2] = r[2];
    res[3] = 0.0f;
    return 1;
}

/* geom-sse.c:144-149 -- squared distance from the centre to the box,
 * computed from the clamped centre, strictly below r*r. */
int orc_box_ball_interp(const struct orc_box *box, const orc_dot center,
                        float radius)
{
    orc_dot fit;
    int i;
    for (i = 0; i < 3; i++)
        fit[i] = smin(smax(center[i], box->min[i]), box->max[i]);
    fit[3] = 0.0f;
    return orc_sqr_metric(fit, center) < radius * radius;
}

/* geom.c:157-167.  clang contracts n*vox_volume - bb_volume. */
int orc_dense_set_p(const struct orc_box *box, size_t n)
{
    float s0 = box->max[0] - box->min[0];
    float s1 = box->max[1] - box->min[1];
    float s2 = box->max[2] - box->min[2];
    float bbv = (s0 * s1) * s2;
    float vv = (orc_voxel[0] * orc_voxel[1]) * orc_voxel[2];
    return fabsf(fmaf((float)n, vv, -bbv)) < vv;
}

/* ------------------------------------------------------------------ */
/* Tree construction: src/voxtrees/tree.c                              */
/* ------------------------------------------------------------------ */

/* Every node carries a hidden 16-byte header (first voxel id, preorder
 * index) in front of the reference layout, so the node itself keeps the
 * exact reference layout (tree.h:28-38). */
typedef struct {
    uint32_t id_base;
    uint32_t preorder;
    uint32_t pad[2];
} orc_hdr;

/* tree.c:212-227 allocates only the union member a node uses; the oracle
 * always allocates the whole node, which changes no field offset. */
static struct orc_node *node_alloc(int flavor)
{
    size_t size = sizeof(orc_hdr) + sizeof(struct orc_node);
    char *mem;
    (void)flavor;
    mem = aligned_alloc(16, (size + 15) & ~(size_t)15);
    memset(mem, 0, size);
    return (struct orc_node *)(mem + sizeof(orc_hdr));
}

static orc_hdr *node_hdr(const struct orc_node *n)
{
    return (orc_hdr *)((char *)n - sizeof(orc_hdr));
}

/* tree.c:57-72 */
static void calc_bounding_box(const orc_dot set[], size_t n,
                              struct orc_box *box)
{
    float mn[3], mx[3];
    size_t i;
    int l;
    for (l = 0; l < 3; l++) mn[l] = mx[l] = set[0][l];
    for (i = 1; i < n; i++)
        for (l = 0; l < 3; l++) {
            mn[l] = smin(mn[l], set[i][l]);
            mx[l] = smax(mx[l], set[i][l]);
        }
    for (l = 0; l < 3; l++) {
        box->min[l] = mn[l];
        box->max[l] = mx[l] + orc_voxel[l];
    }
    box->min[3] = set[0][3];
    box->max[3] = 0.0f;
}

/* tree.c:31-55 -- sequential float sum of (p - min), divided by (float)n,
 * plus min, then snapped up to the voxel grid. */
static void find_center(const orc_dot set[], size_t n,
                        const struct orc_box *box, orc_dot res)
{
    float len = (float)n;
    float sum[3] = {0.0f, 0.0f, 0.0f};
    size_t i;
    int l;
    for (i = 0; i < n; i++)
        for (l = 0; l < 3; l++) sum[l] = sum[l] + (set[i][l] - box->min[l]);
    for (l = 0; l < 3; l++) {
        float s = sum[l] / len;
        s = s + box->min[l];
        res[l] = ceilf(s / orc_voxel[l]) * orc_voxel[l];
    }
    res[3] = 0.0f;
}

/* tree.c:77-81 -- sign bits of dot - center. */
static int subspace_simd(const orc_dot center, const orc_dot dot)
{
    int r = 0, l;
    for (l = 0; l < 3; l++) r |= (signbit(dot[l] - center[l]) ? 1 : 0) << l;
    return r;
}

/* tree.c:83-102 -- in-place partition; matching dots are swapped to the
 * front of [offset, n) in scan order. */
static size_t sort_set(orc_dot set[], size_t n, size_t offset, int subspace,
                       const orc_dot center)
{
    size_t i, counter = offset;
    for (i = offset; i < n; i++) {
        if (subspace_simd(center, set[i]) == subspace) {
            orc_dot tmp;
            memcpy(tmp, set[counter], sizeof(orc_dot));
            memcpy(set[counter], set[i], sizeof(orc_dot));
            memcpy(set[i], tmp, sizeof(orc_dot));
            counter++;
        }
    }
    return counter;
}

/* tree.c:249-338 */
struct orc_node *orc_make_tree(orc_dot *set, size_t n)
{
    struct orc_node *node = NULL;
    if (n > 0) {
        struct orc_box box;
        int densep, leafp;
        calc_bounding_box(set, n, &box);
        densep = orc_dense_set_p(&box, n) ? ORC_DENSE_LEAF : 0;
        leafp = (n <= 7) ? ORC_LEAF : 0;
        node = node_alloc(densep | leafp);
        node->dots_num = (unsigned int)n;
        memcpy(&node->bounding_box, &box, sizeof(box));
        if (densep) {
            node->flags = ORC_DENSE_LEAF;
        } else if (leafp) {
            node->data.dots = aligned_alloc(16, 7 * sizeof(orc_dot));
            memcpy(node->data.dots, set, n * sizeof(orc_dot));
            node->flags = ORC_LEAF;
        } else {
            size_t offset = 0, new_offset;
            int idx;
            node->flags = 0;
            find_center(set, n, &node->bounding_box, node->data.inner.center);
            for (idx = 0; idx < 8; idx++) {
                new_offset = sort_set(set, n, offset, idx,
                                      node->data.inner.center);
                if (offset == 0 && new_offset == n) {
                    /* tree.c:281-305: no split possible -> overflow leaf */
                    struct orc_node *nn = node_alloc(ORC_LEAF);
                    memcpy(&nn->bounding_box, &node->bounding_box,
                           sizeof(struct orc_box));
                    nn->flags = ORC_LEAF | ORC_OVERFLOW;
                    nn->dots_num = (unsigned int)n;
                    nn->data.dots = aligned_alloc(16, n * sizeof(orc_dot));
                    memcpy(nn->data.dots, set, n * sizeof(orc_dot));
                    free(node_hdr(node));
                    node = nn;
                    break;
                }
                node->data.inner.children[idx] =
                    orc_make_tree(set + offset, new_offset - offset);
                offset = new_offset;
            }
        }
    }
    return node;
}

struct orc_node *orc_make_tree_from(const float *xyzw, size_t n)
{
    orc_dot *set;
    struct orc_node *t;
    if (n == 0) return NULL;
    set = aligned_alloc(16, n * sizeof(orc_dot));
    memcpy(set, xyzw, n * sizeof(orc_dot));
    t = orc_make_tree(set, n);
    free(set);
    if (t) {
        /* assign the canonical voxel ids */
        orc_tree_stats st;
        orc_tree_stats_get(t, &st);
    }
    return t;
}

/* tree.c:345-356 */
void orc_destroy_tree(struct orc_node *tree)
{
    int i;
    if (tree) {
        if (tree->flags & ORC_LEAF) free(tree->data.dots);
        else if (!(tree->flags & ORC_LEAF_MASK))
            for (i = 0; i < 8; i++)
                orc_destroy_tree(tree->data.inner.children[i]);
        free(node_hdr(tree));
    }
}

/* tree.c:340-343 */
size_t orc_voxels_in_tree(const struct orc_node *tree)
{
    return tree ? tree->dots_num : 0;
}

/* geom.c:192-196 -- get_dimensions (float division, truncated). */
static void dense_dims(const struct orc_node *t, uint32_t dim[3])
{
    int l;
    for (l = 0; l < 3; l++)
        dim[l] = (uint32_t)(size_t)((t->bounding_box.max[l] -
                                     t->bounding_box.min[l]) /
                                    orc_voxel[l]);
}

static void stats_walk(const struct orc_node *t, orc_tree_stats *st,
                       uint32_t depth)
{
    int i;
    if (!t) return;
    st->nodes++;
    if (depth > st->depth) st->depth = depth;
    if (t->flags & ORC_DENSE_LEAF) {
        uint32_t dim[3];
        dense_dims(t, dim);
        st->dense++;
        st->voxel_ids += (uint64_t)dim[0] * dim[1] * dim[2];
    } else if (t->flags & ORC_LEAF) {
        st->leaves++;
        if (t->flags & ORC_OVERFLOW) st->overflow++;
        st->dots += t->dots_num;
        st->voxel_ids += t->dots_num;
    } else {
        st->inner++;
        for (i = 0; i < 8; i++)
            stats_walk(t->data.inner.children[i], st, depth + 1);
    }
}

/* Canonical voxel ids: regular-leaf dots first (preorder, stored order),
 * then dense-leaf voxels (preorder, (ix*dy + iy)*dz + iz). */
static void id_walk(const struct orc_node *t, uint64_t *dot_next,
                    uint64_t *dense_next, uint64_t *pre)
{
    int i;
    if (!t) return;
    node_hdr(t)->preorder = (uint32_t)(*pre)++;
    if (t->flags & ORC_DENSE_LEAF) {
        uint32_t dim[3];
        dense_dims(t, dim);
        node_hdr(t)->id_base = (uint32_t)*dense_next;
        *dense_next += (uint64_t)dim[0] * dim[1] * dim[2];
    } else if (t->flags & ORC_LEAF) {
        node_hdr(t)->id_base = (uint32_t)*dot_next;
        *dot_next += t->dots_num;
    } else {
        for (i = 0; i < 8; i++)
            id_walk(t->data.inner.children[i], dot_next, dense_next, pre);
    }
}

void orc_tree_stats_get(const struct orc_node *tree, orc_tree_stats *st)
{
    uint64_t dn = 0, dd, pre = 0;
    memset(st, 0, sizeof(*st));
    stats_walk(tree, st, 1);
    dd = st->dots;
    id_walk(tree, &dn, &dd, &pre);
}

static void table_walk(const struct orc_node *t, float *xyz)
{
    int i;
    if (!t) return;
    if (t->flags & ORC_DENSE_LEAF) {
        uint32_t dim[3], a, b, c;
        uint64_t id = node_hdr(t)->id_base;
        dense_dims(t, dim);
        for (a = 0; a < dim[0]; a++)
            for (b = 0; b < dim[1]; b++)
                for (c = 0; c < dim[2]; c++, id++) {
                    xyz[3 * id + 0] = t->bounding_box.min[0] + a * orc_voxel[0];
                    xyz[3 * id + 1] = t->bounding_box.min[1] + b * orc_voxel[1];
                    xyz[3 * id + 2] = t->bounding_box.min[2] + c * orc_voxel[2];
                }
    } else if (t->flags & ORC_LEAF) {
        uint64_t id = node_hdr(t)->id_base;
        unsigned k;
        for (k = 0; k < t->dots_num; k++, id++) {
            xyz[3 * id + 0] = t->data.dots[k][0];
            xyz[3 * id + 1] = t->data.dots[k][1];
            xyz[3 * id + 2] = t->data.dots[k][2];
        }
    } else {
        for (i = 0; i < 8; i++) table_walk(t->data.inner.children[i], xyz);
    }
}

void orc_tree_voxel_table(const struct orc_node *tree, float *xyz_out)
{
    orc_tree_stats st;
    orc_tree_stats_get(tree, &st);
    table_walk(tree, xyz_out);
}

static uint64_t mix64(uint64_t h, uint64_t v)
{
    h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    h *= 0xff51afd7ed558ccdull;
    return h ^ (h >> 33);
}

static uint64_t fbits(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

/* Structure hash: preorder, child slot, flags, counts, box/centre bits and
 * leaf dots in stored order.  The product's flattened tree hashes the
 * same way. */
static uint64_t hash_walk(const struct orc_node *t, uint64_t h, int slot)
{
    int i, l;
    if (!t) return mix64(h, 0xdead0000ull + (uint64_t)slot);
    h = mix64(h, 0x1000ull + (uint64_t)slot);
    h = mix64(h, t->flags & 7u);
    h = mix64(h, t->dots_num);
    for (l = 0; l < 3; l++) {
        h = mix64(h, fbits(t->bounding_box.min[l]));
        h = mix64(h, fbits(t->bounding_box.max[l]));
    }
    if (t->flags & ORC_DENSE_LEAF) return h;
    if (t->flags & ORC_LEAF) {
        unsigned k;
        for (k = 0; k < t->dots_num; k++)
            for (l = 0; l < 3; l++) h = mix64(h, fbits(t->data.dots[k][l]));
        return h;
    }
    for (l = 0; l < 3; l++) h = mix64(h, fbits(t->data.inner.center[l]));
    for (i = 0; i < 8; i++) h = hash_walk(t->data.inner.children[i], h, i);
    return h;
}

uint64_t orc_tree_hash(const struct orc_node *tree)
{
    return hash_walk(tree, 0x12345678ull, 8);
}

/* ------------------------------------------------------------------ */
/* Search: src/voxtrees/search.c                                       */
/* ------------------------------------------------------------------ */

/* search.c:8-153.  Additionally reports the winning dot slot (-1 for a
 * dense leaf) and traversal counters. */
/* Optional per-thread tracker for orc_ray_max_pending: for each inner node
 * on the recursion path, how many of its children behind dividing planes
 * (non-empty, not yet visited) are still pending; `max` records the most
 * nodes with a pending child seen at any node visit.  Test-only
 * instrumentation: the search's results and order are unchanged. */
typedef struct {
    int depth;
    int rem[64];
    int max;
} pend_tracker;
static __thread pend_tracker *g_pend;

static void pend_visit(void)
{
    int k, cur = 0;
    for (k = 0; k < g_pend->depth; k++) cur += g_pend->rem[k] > 0;
    if (cur > g_pend->max) g_pend->max = cur;
}

/* Non-empty children of `tree` reached through its dividing planes in the
 * order of search.c:105-147 (counted from the same primitive calls). */
static int plane_children(const struct orc_node *tree, const orc_dot e,
                          const orc_dot dir, int subspace)
{
    orc_dot pi[3];
    float m[3];
    int idx[3], cnt = 0, i, j, n = 0;
    for (i = 0; i < 3; i++)
        if (orc_hit_plane_within_box(e, dir, tree->data.inner.center, i,
                                     pi[cnt], &tree->bounding_box)) {
            m[cnt] = orc_abs_metric(e, pi[cnt]);
            idx[cnt++] = i;
        }
    for (i = 0; i < cnt; i++) {
        for (j = i + 1; j < cnt; j++)
            if (m[j] < m[i]) {
                float tm = m[j]; int ti = idx[j];
                m[j] = m[i]; idx[j] = idx[i];
                m[i] = tm; idx[i] = ti;
            }
        subspace ^= 1 << idx[i];
        n += tree->data.inner.children[subspace] != NULL;
    }
    return n;
}

static const struct orc_node *rti(const struct orc_node *tree,
                                  const orc_dot origin, const orc_dot dir,
                                  orc_dot res, int *dot_slot,
                                  orc_counters *c)
{
    orc_dot bb_inter;
    orc_dot plane_inter[3];
    int plane_idx[3];
    unsigned i, plane_counter = 0;
    int subspace;
    const struct orc_node *leaf = NULL;

    if (!tree) return NULL;
    c->node_visits++;
    if (g_pend) pend_visit();
    if (!orc_hit_box(&tree->bounding_box, origin, dir, bb_inter)) return NULL;

    /* search.c:31-37 */
    if (tree->flags & ORC_DENSE_LEAF) {
        memcpy(res, bb_inter, sizeof(orc_dot));
        *dot_slot = -1;
        return tree;
    }
    /* search.c:39-81 */
    if (tree->flags & ORC_LEAF) {
        float dist_closest = INFINITY, dist_far;
        struct orc_box voxel;
        orc_dot far_inter;
        for (i = 0; i < tree->dots_num; i++) {
            int l;
            c->voxel_tests++;
            for (l = 0; l < 3; l++) {
                voxel.min[l] = tree->data.dots[i][l];
                voxel.max[l] = tree->data.dots[i][l] + orc_voxel[l];
            }
            if (orc_hit_box(&voxel, bb_inter, dir, far_inter)) {
                dist_far = orc_abs_metric(bb_inter, far_inter);
                if (dist_far == 0.0f) {
                    memcpy(res, far_inter, sizeof(orc_dot));
                    *dot_slot = (int)i;
                    return tree;
                }
                if (dist_far < dist_closest) {
                    dist_closest = dist_far;
                    memcpy(res, far_inter, sizeof(orc_dot));
                    *dot_slot = (int)i;
                    leaf = tree;
                }
            }
        }
        return leaf;
    }

    /* search.c:84-97 */
    subspace = orc_get_corrected_subspace_idx(tree->data.inner.center,
                                              bb_inter, dir);
    if (g_pend) {
        g_pend->rem[g_pend->depth] =
            plane_children(tree, bb_inter, dir, subspace);
        g_pend->depth++;
    }
    leaf = rti(tree->data.inner.children[subspace], bb_inter, dir, res,
               dot_slot, c);
    if (leaf) {
        if (g_pend) g_pend->depth--;
        return leaf;
    }

    /* search.c:105-113 */
    for (i = 0; i < 3; i++) {
        plane_idx[plane_counter] = (int)i;
        if (orc_hit_plane_within_box(bb_inter, dir, tree->data.inner.center,
                                     (int)i, plane_inter[plane_counter],
                                     &tree->bounding_box))
            plane_counter++;
    }
    /* search.c:115-148 -- selection sort by L1 distance, then recurse */
    for (i = 0; i < plane_counter; i++) {
        unsigned j;
        for (j = i + 1; j < plane_counter; j++) {
            if (orc_abs_metric(bb_inter, plane_inter[j]) <
                orc_abs_metric(bb_inter, plane_inter[i])) {
                orc_dot tmp;
                int t2;
                memcpy(tmp, plane_inter[j], sizeof(orc_dot));
                memcpy(plane_inter[j], plane_inter[i], sizeof(orc_dot));
                memcpy(plane_inter[i], tmp, sizeof(orc_dot));
                t2 = plane_idx[j];
                plane_idx[j] = plane_idx[i];
                plane_idx[i] = t2;
            }
        }
        subspace = subspace ^ (1 << plane_idx[i]);
        if (g_pend && tree->data.inner.children[subspace])
            g_pend->rem[g_pend->depth - 1]--;
        if ((leaf = rti(tree->data.inner.children[subspace], plane_inter[i],
                        dir, res, dot_slot, c))) {
            if (g_pend) g_pend->depth--;
            return leaf;
        }
    }
    if (g_pend) g_pend->depth--;
    return NULL;
}

/* For each ray, the most inner nodes on the search path that still had a
 * child behind a dividing plane to visit, over all node visits of the
 * search: the entries a stack holding one entry per such node needs. */
void orc_ray_max_pending(const struct orc_node *tree, size_t n,
                         const float *ox, const float *oy, const float *oz,
                         const float *dx, const float *dy, const float *dz,
                         uint32_t *out)
{
    size_t r;
    pend_tracker t;
    orc_counters c = {0, 0};
    g_pend = &t;
    for (r = 0; r < n; r++) {
        orc_dot o = {ox[r], oy[r], oz[r], 0.0f};
        orc_dot d = {dx[r], dy[r], dz[r], 0.0f};
        orc_dot res;
        int slot;
        memset(&t, 0, sizeof(t));
        (void)rti(tree, o, d, res, &slot, &c);
        out[r] = (uint32_t)t.max;
    }
    g_pend = NULL;
}

const struct orc_node *orc_ray_tree_intersection(const struct orc_node *tree,
                                                 const orc_dot origin,
                                                 const orc_dot dir,
                                                 orc_dot res)
{
    int slot;
    orc_counters c = {0, 0};
    return rti(tree, origin, dir, res, &slot, &c);
}

/* search.c:155-183 */
static int ball(const struct orc_node *tree, const orc_dot center,
                float radius, orc_counters *c)
{
    unsigned i;
    if (!tree) return 0;
    c->node_visits++;
    if (orc_box_ball_interp(&tree->bounding_box, center, radius)) {
        if (tree->flags & ORC_DENSE_LEAF) return 1;
        if (tree->flags & ORC_LEAF) {
            struct orc_box voxel;
            for (i = 0; i < tree->dots_num; i++) {
                int l;
                c->voxel_tests++;
                for (l = 0; l < 3; l++) {
                    voxel.min[l] = tree->data.dots[i][l];
                    voxel.max[l] = tree->data.dots[i][l] + orc_voxel[l];
                }
                if (orc_box_ball_interp(&voxel, center, radius)) return 1;
            }
        } else {
            for (i = 0; i < 8; i++)
                if (ball(tree->data.inner.children[i], center, radius, c))
                    return 1;
        }
    }
    return 0;
}

int orc_tree_ball_collidep(const struct orc_node *tree, const orc_dot center,
                           float radius)
{
    orc_counters c = {0, 0};
    return ball(tree, center, radius, &c);
}

/* Canonical voxel id of a search result: for a dense leaf the voxel that
 * holds the hit point, its index clamped to the leaf's dimensions on the
 * float, before the conversion (NaN goes to 0, +-inf to the last or first
 * voxel: a plain conversion of those is undefined). */
static uint32_t voxel_id_of(const struct orc_node *leaf, const orc_dot res,
                            int slot)
{
    if (!leaf) return ORC_MISS;
    if (leaf->flags & ORC_DENSE_LEAF) {
        uint32_t dim[3], ix[3];
        int l;
        dense_dims(leaf, dim);
        for (l = 0; l < 3; l++) {
            float f = floorf((res[l] - leaf->bounding_box.min[l]) /
                             orc_voxel[l]);
            ix[l] = (uint32_t)(int)fminf(fmaxf(f, 0.0f),
                                         (float)((int)dim[l] - 1));
        }
        return node_hdr(leaf)->id_base + (ix[0] * dim[1] + ix[1]) * dim[2] +
               ix[2];
    }
    return node_hdr(leaf)->id_base + (uint32_t)slot;
}

/* ------------------------------------------------------------------ */
/* Thread pool for the batched loops (the dispatch_apply analogue,      */
/* src/gcd-stubs.c:23-27 / renderer.c:214).                             */
/* ------------------------------------------------------------------ */

int orc_cpu_count(void)
{
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) return CPU_COUNT(&set);
    return 1;
}

typedef void (*orc_task_fn)(void *ctx, size_t begin, size_t end,
                            orc_counters *c);

typedef struct {
    orc_task_fn fn;
    void *ctx;
    size_t n, chunk;
    atomic_size_t next;
    orc_counters *per_thread;
} orc_pool;

typedef struct {
    orc_pool *pool;
    int tid;
} orc_worker_arg;

static void *orc_worker(void *p)
{
    orc_worker_arg *a = p;
    orc_pool *pool = a->pool;
    for (;;) {
        size_t b = atomic_fetch_add(&pool->next, pool->chunk);
        size_t e;
        if (b >= pool->n) break;
        e = b + pool->chunk < pool->n ? b + pool->chunk : pool->n;
        pool->fn(pool->ctx, b, e, &pool->per_thread[a->tid]);
    }
    return NULL;
}

static void orc_parallel(orc_task_fn fn, void *ctx, size_t n, size_t chunk,
                         int nthreads, orc_counters *total)
{
    orc_pool pool;
    pthread_t th[256];
    orc_worker_arg args[256];
    orc_counters local[256];
    int t;
    if (nthreads <= 0) nthreads = orc_cpu_count();
    if (nthreads > 256) nthreads = 256;
    memset(local, 0, sizeof(local));
    pool.fn = fn;
    pool.ctx = ctx;
    pool.n = n;
    pool.chunk = chunk ? chunk : 1;
    atomic_init(&pool.next, 0);
    pool.per_thread = local;
    if (nthreads == 1) {
        fn(ctx, 0, n, &local[0]);
    } else {
        for (t = 0; t < nthreads; t++) {
            args[t].pool = &pool;
            args[t].tid = t;
            pthread_create(&th[t], NULL, orc_worker, &args[t]);
        }
        for (t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    }
    if (total) {
        memset(total, 0, sizeof(*total));
        for (t = 0; t < nthreads; t++) {
            total->node_visits += local[t].node_visits;
            total->voxel_tests += local[t].voxel_tests;
        }
    }
}

typedef struct {
    const struct orc_node *tree;
    const float *ox, *oy, *oz, *dx, *dy, *dz;
    uint32_t *voxel_out;
    float *hx, *hy, *hz;
} ray_ctx;

static void ray_task(void *p, size_t b, size_t e, orc_counters *c)
{
    ray_ctx *x = p;
    size_t i;
    for (i = b; i < e; i++) {
        orc_dot o = {x->ox[i], x->oy[i], x->oz[i], 0.0f};
        orc_dot d = {x->dx[i], x->dy[i], x->dz[i], 0.0f};
        orc_dot res;
        int slot = -1;
        const struct orc_node *leaf = rti(x->tree, o, d, res, &slot, c);
        if (x->voxel_out) x->voxel_out[i] = voxel_id_of(leaf, res, slot);
        if (leaf && x->hx) {
            x->hx[i] = res[0];
            x->hy[i] = res[1];
            x->hz[i] = res[2];
        }
    }
}

void orc_ray_batch(const struct orc_node *tree, size_t n, const float *ox,
                   const float *oy, const float *oz, const float *dx,
                   const float *dy, const float *dz, uint32_t *voxel_out,
                   float *hx, float *hy, float *hz, int nthreads,
                   orc_counters *counters)
{
    ray_ctx x = {tree, ox, oy, oz, dx, dy, dz, voxel_out, hx, hy, hz};
    orc_parallel(ray_task, &x, n, 256, nthreads, counters);
}

typedef struct {
    const struct orc_node *tree;
    const float *cx, *cy, *cz, *r;
    uint8_t *out;
} ball_ctx;

static void ball_task(void *p, size_t b, size_t e, orc_counters *c)
{
    ball_ctx *x = p;
    size_t i;
    for (i = b; i < e; i++) {
        orc_dot ctr = {x->cx[i], x->cy[i], x->cz[i], 0.0f};
        x->out[i] = (uint8_t)ball(x->tree, ctr, x->r[i], c);
    }
}

void orc_ball_batch(const struct orc_node *tree, size_t n, const float *cx,
                    const float *cy, const float *cz, const float *r,
                    uint8_t *out, int nthreads, orc_counters *counters)
{
    ball_ctx x = {tree, cx, cy, cz, r, out};
    orc_parallel(ball_task, &x, n, 1024, nthreads, counters);
}

/* ------------------------------------------------------------------ */
/* Camera: src/voxrnd/simple-camera.c, src/voxrnd/vect-ops-sse.c        */
/* ------------------------------------------------------------------ */

/* simple-camera.c:120-136 */
void orc_camera_set_window_size(orc_camera *cam, int w, int h)
{
    if (w > h) {
        cam->xsub = cam->fov;
        cam->ysub = (cam->fov * (float)h) / (float)w;
        cam->mul = (cam->fov * 2.0f) / (float)w;
    } else {
        cam->ysub = cam->fov;
        cam->xsub = (cam->fov * (float)w) / (float)h;
        cam->mul = (cam->fov * 2.0f) / (float)h;
    }
}

/* vect-ops-sse.c:6-14 on quaternion-layout lanes (w, i, j, k): the
 * vector part is (a_j b_k - b_j a_k, a_k b_i - b_k a_i, a_i b_j - b_i a_j),
 * each product rounded before the subtraction; lane 0 is a_w b_w - b_w a_w. */
static void cross_q(const float a[4], const float b[4], float r[4])
{
    r[0] = a[0] * b[0] - b[0] * a[0];
    r[1] = a[2] * b[3] - b[2] * a[3];
    r[2] = a[3] * b[1] - b[3] * a[1];
    r[3] = a[1] * b[2] - b[1] * a[2];
}

/* vect-ops-sse.c:16-26: t = 2 cross(q, v); r = fma(q_w, t, v) + cross(q, t) */
void orc_rotate_vector_q(const float q[4], const float v[4], float res[4])
{
    float t[4], c2[4], r[4];
    int l;
    cross_q(q, v, t);
    for (l = 0; l < 4; l++) t[l] = t[l] * 2.0f;
    cross_q(q, t, c2);
    for (l = 0; l < 4; l++) r[l] = fmaf(q[0], t[l], v[l]) + c2[l];
    memcpy(res, r, sizeof(r));
}

/* vect-ops-sse.c:58-64 */
void orc_rotate_vector(const float q[4], const float v[3], float res[3])
{
    float vq[4] = {0.0f, v[0], v[1], v[2]}, r[4];
    orc_rotate_vector_q(q, vq, r);
    res[0] = r[1];
    res[1] = r[2];
    res[2] = r[3];
}

/* vect-ops-sse.c:28-45 */
void orc_quat_mul(const float q1[4], const float q2[4], float res[4])
{
    float p[4], r3[4], out[4];
    int l;
    float re;
    for (l = 0; l < 4; l++) p[l] = q1[l] * q2[l];
    re = (p[0] + (-p[1])) + ((-p[2]) + (-p[3]));
    cross_q(q1, q2, r3);
    for (l = 1; l < 4; l++) {
        float r1 = q1[0] * q2[l];
        float r2 = q2[0] * q1[l];
        out[l] = (r1 + r2) + r3[l];
    }
    out[0] = re;
    memcpy(res, out, sizeof(out));
}

/* simple-camera.c:21-31 */
void orc_screen2world(const orc_camera *cam, int sx, int sy, orc_dot ray)
{
    float v[3];
    v[0] = fmaf(cam->mul, (float)sx, -cam->xsub);
    v[1] = 1.0f;
    v[2] = fmaf(-cam->mul, (float)sy, cam->ysub);
    orc_rotate_vector(cam->rotation, v, ray);
    ray[3] = 0.0f;
}

/* ------------------------------------------------------------------ */
/* Renderer per-pixel loop: src/voxrnd/renderer.c                       */
/* ------------------------------------------------------------------ */

/* renderer.c:23-25 */
static const int rendering_order[16] = {0, 1, 4,  5,  2,  3,  6,  7,
                                        8, 9, 12, 13, 10, 11, 14, 15};

/* renderer.c:27-52 without a light manager: texture lookup, then
 * SDL_MapRGB(c, c, c) for a 32-bit format. */
static uint32_t get_color(const orc_shading *sh, const orc_dot inter)
{
    int x, y, z, idx;
    uint32_t c;
    if (!sh || !sh->texture) return 0xffffffffu;
    x = abs((int)(inter[0] / orc_voxel[0])) & 31;
    y = abs((int)(inter[1] / orc_voxel[1])) & 31;
    z = abs((int)(inter[2] / orc_voxel[2])) & 31;
    idx = x * 32 * 32 + y * 32 + z;
    c = sh->texture[idx];
    if (sh->lights) {
        /* renderer.c:40-49: each channel is min(light, 1) * colour,
         * truncated to 8 bits */
        float light[3];
        uint32_t r, g, b;
        orc_get_light(sh->lights, inter, light);
        r = (uint32_t)(uint8_t)(fminf(light[0], 1.0f) * (float)c);
        g = (uint32_t)(uint8_t)(fminf(light[1], 1.0f) * (float)c);
        b = (uint32_t)(uint8_t)(fminf(light[2], 1.0f) * (float)c);
        return (r << sh->rshift) | (g << sh->gshift) | (b << sh->bshift) |
               sh->amask;
    }
    return (c << sh->rshift) | (c << sh->gshift) | (c << sh->bshift) |
           sh->amask;
}

typedef struct {
    const struct orc_node *scene;
    const orc_camera *cam;
    int W, H, ws;
    int rnd_mode, merge_mode;
    const orc_shading *shade;
    uint32_t *squares;
    uint32_t *voxel_px;
    float *hit_px;
    /* optional trace of the loop's decisions (orc_render_trace) */
    uint8_t *block_flags;
    uint8_t *px_src;
    float *try_box;
    atomic_ullong fast_blocks, best_blocks, merge_blocks, root_tr, leaf_tr;
} render_ctx;

/* Trace of pixel (px, py): where its result came from, and for a leaf try
 * that answered it, the box of the leaf that was tried. */
static void px_trace(render_ctx *x, int px, int py, int src,
                     const struct orc_node *tried)
{
    size_t cs, i, np;
    int l;
    if (!x->px_src) return;
    cs = (size_t)(py >> 2) * x->ws + (size_t)(px >> 2);
    i = cs * 16 + (size_t)((py & 3) * 4 + (px & 3));
    np = (size_t)x->W * x->H;
    x->px_src[i] = (uint8_t)src;
    if (x->try_box && tried)
        for (l = 0; l < 3; l++) {
            x->try_box[l * np + i] = tried->bounding_box.min[l];
            x->try_box[(3 + l) * np + i] = tried->bounding_box.max[l];
        }
}

/* Per-pixel outputs use the square layout: index cs*16 + p. */
static void px_out(render_ctx *x, int px, int py, const struct orc_node *leaf,
                   const orc_dot inter, uint32_t vid)
{
    size_t cs = (size_t)(py >> 2) * x->ws + (size_t)(px >> 2);
    size_t i = cs * 16 + (size_t)((py & 3) * 4 + (px & 3));
    if (x->voxel_px) x->voxel_px[i] = leaf ? vid : ORC_MISS;
    if (x->hit_px && leaf) {
        size_t np = (size_t)x->W * x->H;
        x->hit_px[i] = inter[0];
        x->hit_px[np + i] = inter[1];
        x->hit_px[2 * np + i] = inter[2];
    }
}

#define LENGTH_THRESHOLD 1.69 /* renderer.c:197, a double constant */
#define MAX_DIST 150          /* renderer.c:198 */

/* renderer.c:214-318: one 4x4 block */
static void render_task(void *p, size_t b, size_t e, orc_counters *c)
{
    render_ctx *x = p;
    size_t cs;
    for (cs = b; cs < e; cs++) {
        const struct orc_node *corner1, *corner2, *leaf = NULL;
        orc_dot dir1, dir2, inter1, inter2, origin;
        int block_rnd_mode = x->rnd_mode, block_merge_mode = 0;
        int i, xstart, ystart, istart = 0, iend = 16, prev_p = 0, slot = -1;
        uint32_t *out = x->squares ? x->squares + 16 * cs : NULL;
        uint32_t vid = ORC_MISS, vid1 = ORC_MISS;
        unsigned long long nroot = 0, nleaf = 0;

        ystart = (int)(cs / (size_t)x->ws) << 2;
        xstart = (int)(cs % (size_t)x->ws) << 2;
        memcpy(origin, x->cam->position, 3 * sizeof(float));
        origin[3] = 0.0f;

        if (block_rnd_mode == ORC_QUALITY_ADAPTIVE) {
            block_merge_mode = (x->merge_mode == ORC_QUALITY_RAY_MERGE);
            istart = 1;
            orc_screen2world(x->cam, xstart, ystart, dir1);
            corner1 = rti(x->scene, origin, dir1, inter1, &slot, c);
            nroot++;
            block_rnd_mode = ORC_QUALITY_FAST;
            leaf = corner1;
            vid1 = voxel_id_of(corner1, inter1, slot);
            px_out(x, xstart, ystart, corner1, inter1, vid1);
            px_trace(x, xstart, ystart, 0, NULL);
            vid = vid1;
            if (corner1 != NULL) {
                iend = 15;
                if (out) out[0] = get_color(x->shade, inter1);
                orc_screen2world(x->cam, xstart + 3, ystart + 3, dir2);
                corner2 = rti(x->scene, origin, dir2, inter2, &slot, c);
                nroot++;
                px_trace(x, xstart + 3, ystart + 3, 0, NULL);
                px_out(x, xstart + 3, ystart + 3, corner2, inter2,
                       voxel_id_of(corner2, inter2, slot));
                if (corner2 != NULL) {
                    float d1, d2, criteria, dist;
                    if (out) out[15] = get_color(x->shade, inter2);
                    d1 = orc_sqr_metric(inter1, origin);
                    d2 = orc_sqr_norm(dir1);
                    criteria = d1 / d2 * orc_sqr_metric(dir1, dir2);
                    dist = orc_sqr_metric(inter1, inter2);
                    if ((double)(dist / criteria) > LENGTH_THRESHOLD)
                        block_rnd_mode = ORC_QUALITY_BEST;
                    if (d1 > (float)(MAX_DIST * MAX_DIST)) {
                        block_merge_mode =
                            (block_rnd_mode == ORC_QUALITY_FAST &&
                             x->merge_mode == ORC_QUALITY_RAY_MERGE_ACCURATE)
                                ? 1
                                : block_merge_mode;
                    } else {
                        block_merge_mode = 0;
                    }
                }
            }
        }
        if (block_rnd_mode == ORC_QUALITY_FAST) atomic_fetch_add(&x->fast_blocks, 1);
        else atomic_fetch_add(&x->best_blocks, 1);
        if (block_merge_mode) atomic_fetch_add(&x->merge_blocks, 1);
        if (x->block_flags)
            x->block_flags[cs] = (uint8_t)((block_rnd_mode == ORC_QUALITY_BEST) |
                                           (block_merge_mode ? 2 : 0));

        for (i = istart; i < iend; i++) {
            int pp = rendering_order[i];
            int y = pp / 4, xx = pp % 4;
            int merge = block_merge_mode && (i & 1);
            orc_screen2world(x->cam, xx + xstart, y + ystart, dir1);
            if (!merge) {
                const struct orc_node *tried = NULL;
                if (block_rnd_mode == ORC_QUALITY_FAST) {
                    if (leaf != NULL) {
                        tried = leaf;
                        leaf = rti(leaf, origin, dir1, inter1, &slot, c);
                        nleaf++;
                    }
                    if (leaf == NULL) {
                        tried = NULL;
                        leaf = rti(x->scene, origin, dir1, inter1, &slot, c);
                        nroot++;
                    }
                } else {
                    leaf = rti(x->scene, origin, dir1, inter1, &slot, c);
                    nroot++;
                }
                vid = voxel_id_of(leaf, inter1, slot);
                px_trace(x, xx + xstart, y + ystart, tried ? 1 : 0, tried);
            } else {
                px_trace(x, xx + xstart, y + ystart, 2, NULL);
            }
            if (leaf != NULL) {
                uint32_t color = 0;
                if (out) {
                    color = merge ? out[prev_p] : get_color(x->shade, inter1);
                    out[pp] = color;
                }
            }
            px_out(x, xx + xstart, y + ystart, leaf, inter1, vid);
            prev_p = pp;
        }
        atomic_fetch_add(&x->root_tr, nroot);
        atomic_fetch_add(&x->leaf_tr, nleaf);
    }
}

int orc_render(const struct orc_node *scene, const orc_camera *cam, int W,
               int H, unsigned quality, const orc_shading *shade,
               uint32_t *squares, uint32_t *voxel_px, float *hit_px,
               int nthreads, orc_render_stats *stats)
{
    return orc_render_trace(scene, cam, W, H, quality, shade, squares,
                            voxel_px, hit_px, nthreads, stats, NULL, NULL,
                            NULL);
}

int orc_render_trace(const struct orc_node *scene, const orc_camera *cam,
                     int W, int H, unsigned quality,
                     const orc_shading *shade, uint32_t *squares,
                     uint32_t *voxel_px, float *hit_px, int nthreads,
                     orc_render_stats *stats, uint8_t *block_flags,
                     uint8_t *px_src, float *try_box)
{
    render_ctx x;
    orc_counters tot;
    size_t nsq;
    /* renderer.c:97-100 bad_geometry, :179-195 quality checks */
    if ((W & 0xf) || (H & 0x3) || W <= 0 || H <= 0) return -22;
    if ((quality & 0xc) && (quality & 3) != ORC_QUALITY_ADAPTIVE) return -22;
    if ((quality & 3) > 2 || (quality & 0xc) > 8) return -22;
    memset(&x, 0, sizeof(x));
    x.scene = scene;
    x.cam = cam;
    x.W = W;
    x.H = H;
    x.ws = W >> 2;
    x.rnd_mode = (int)(quality & 3);
    x.merge_mode = (int)(quality & 0xc);
    x.shade = shade;
    x.squares = squares;
    x.voxel_px = voxel_px;
    x.hit_px = hit_px;
    x.block_flags = block_flags;
    x.px_src = px_src;
    x.try_box = try_box;
    nsq = (size_t)(W >> 2) * (size_t)(H >> 2);
    orc_parallel(render_task, &x, nsq, 64, nthreads, &tot);
    if (stats) {
        stats->blocks = nsq;
        stats->fast_blocks = x.fast_blocks;
        stats->best_blocks = x.best_blocks;
        stats->merge_blocks = x.merge_blocks;
        stats->root_traversals = x.root_tr;
        stats->leaf_traversals = x.leaf_tr;
        stats->c = tot;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Tree invariants: restatement of check_tree, tests/tests.c:239-300    */
/* ------------------------------------------------------------------ */
static int dot_between(const struct orc_box *b, const float *d)
{
    int i;
    for (i = 0; i < 3; i++)
        if (d[i] < b->min[i] || d[i] > b->max[i]) return 0;
    return 1;
}

/* Returns 0 when every invariant holds, else a code naming the first
 * failure: 1 dense count, 2 leaf voxel outside box, 3 child box outside
 * parent, 4 child in the wrong subspace, 5 voxel count mismatch. */
int orc_check_tree(const struct orc_node *tree)
{
    int i, r;
    float half[3], neg_half[3];
    if (!tree) return 0;
    for (i = 0; i < 3; i++) {
        half[i] = orc_voxel[i] / 2;
        neg_half[i] = -orc_voxel[i] / 2;
    }
    if (tree->flags & ORC_DENSE_LEAF) {
        float s0 = tree->bounding_box.max[0] - tree->bounding_box.min[0];
        float s1 = tree->bounding_box.max[1] - tree->bounding_box.min[1];
        float s2 = tree->bounding_box.max[2] - tree->bounding_box.min[2];
        float bbv = s0 * s1 * s2;
        bbv /= orc_voxel[0] * orc_voxel[1] * orc_voxel[2];
        return orc_voxels_in_tree(tree) == (size_t)bbv ? 0 : 1;
    }
    if (tree->flags & ORC_LEAF) {
        unsigned k;
        for (k = 0; k < tree->dots_num; k++) {
            float tmp[3];
            for (i = 0; i < 3; i++) tmp[i] = tree->data.dots[k][i] + orc_voxel[i];
            if (!dot_between(&tree->bounding_box, tree->data.dots[k])) return 2;
            if (!dot_between(&tree->bounding_box, tmp)) return 2;
        }
        return 0;
    } else {
        size_t sum = 0;
        for (i = 0; i < 8; i++) {
            const struct orc_node *ch = tree->data.inner.children[i];
            orc_dot tmp;
            int l;
            sum += orc_voxels_in_tree(ch);
            if (ch) {
                if (!dot_between(&tree->bounding_box, ch->bounding_box.min) ||
                    !dot_between(&tree->bounding_box, ch->bounding_box.max))
                    return 3;
                for (l = 0; l < 3; l++) tmp[l] = ch->bounding_box.min[l] + half[l];
                tmp[3] = 0.0f;
                if (orc_get_subspace_idx(tree->data.inner.center, tmp) != i) return 4;
                for (l = 0; l < 3; l++) tmp[l] = ch->bounding_box.max[l] + neg_half[l];
                if (orc_get_subspace_idx(tree->data.inner.center, tmp) != i) return 4;
            }
            if ((r = orc_check_tree(ch)) != 0) return r;
        }
        return orc_voxels_in_tree(tree) == sum ? 0 : 5;
    }
}

/* ------------------------------------------------------------------ */
/* src/voxrnd/copy-helper.c:5-40 -- square layout to pixel rows        */
/* ------------------------------------------------------------------ */
void orc_copy_squares(const uint32_t *squares, uint32_t *pixels, unsigned ws,
                      unsigned hs)
{
    /* square (sx, sy) holds pixel rows 4 sy .. 4 sy + 3 and columns
     * 4 sx .. 4 sx + 3, four pixels per row in order */
    unsigned w = ws * 4, sy, sx, line;
    for (sy = 0; sy < hs; sy++)
        for (sx = 0; sx < ws; sx++)
            for (line = 0; line < 4; line++)
                memcpy(pixels + (size_t)(4 * sy + line) * w + 4 * sx,
                       squares + ((size_t)sy * ws + sx) * 16 + 4 * line,
                       4 * sizeof(uint32_t));
}

/* ------------------------------------------------------------------ */
/* src/voxrnd/cd.c:38-55 -- collision of moving bodies with the scene  */
/* ------------------------------------------------------------------ */
void orc_cd_collide(const struct orc_node *tree, size_t n, float *pos,
                    float *last, const float *radius, uint8_t *collided)
{
    size_t i;
    for (i = 0; i < n; i++) {
        float *p = pos + 3 * i, *l = last + 3 * i;
        orc_dot c = {p[0], p[1], p[2], 0.0f};
        collided[i] = 0;
        /* vox_dot_equalp: all of x, y, z compare equal */
        if (p[0] == l[0] && p[1] == l[1] && p[2] == l[2]) continue;
        if (orc_tree_ball_collidep(tree, c, radius[i])) {
            memcpy(p, l, 3 * sizeof(float));
            collided[i] = 1;
        } else {
            memcpy(l, p, 3 * sizeof(float));
        }
    }
}

/* ------------------------------------------------------------------ */
/* src/voxtrees/datareader.c:49-75 -- samples to voxel corners          */
/* ------------------------------------------------------------------ */
size_t orc_read_raw(const uint8_t *data, const unsigned dim[3],
                    unsigned samplesize, uint64_t lo, uint64_t hi,
                    float *xyz_out)
{
    unsigned i, j, k, l;
    size_t pos = 0, counter = 0;
    for (i = 0; i < dim[0]; i++)
        for (j = 0; j < dim[1]; j++)
            for (k = 0; k < dim[2]; k++) {
                uint32_t value = 0;
                for (l = 0; l < samplesize; l++)
                    value |= (uint32_t)data[pos++] << (8 * l);
                if (value >= lo && value < hi) {
                    xyz_out[3 * counter + 0] = i * orc_voxel[0];
                    xyz_out[3 * counter + 1] = j * orc_voxel[1];
                    xyz_out[3 * counter + 2] = k * orc_voxel[2];
                    counter++;
                }
            }
    return counter;
}

/* ------------------------------------------------------------------ */
/* src/voxtrees/tree.c:358-835 -- flattening and edits                 */
/* ------------------------------------------------------------------ */

/* geom.c:192-196 */
static void get_dims(const struct orc_box *b, size_t dim[3])
{
    int l;
    for (l = 0; l < 3; l++)
        dim[l] = (size_t)((b->max[l] - b->min[l]) / orc_voxel[l]);
}

/* geom.c:198-209: a box one voxel thick along two axes; *which = the
 * other axis */
static int stripep(const struct orc_box *b, int *which)
{
    size_t dim[3];
    int l, count = 0;
    get_dims(b, dim);
    for (l = 0; l < 3; l++) {
        if (dim[l] == 1) count++;
        else *which = l;
    }
    return count == 2;
}

/* geom.c:169-181 */
static int divide_box(const struct orc_box *box, const orc_dot center,
                      struct orc_box *res, int idx)
{
    int l;
    memcpy(res, box, sizeof(*res));
    for (l = 0; l < 3; l++) {
        if (idx & (1 << l)) res->max[l] = center[l];
        else res->min[l] = center[l];
        if (res->min[l] == res->max[l]) return 0;
    }
    return 1;
}

/* geom-sse.c:42-49: inside [min, max) on x, y, z */
static int voxel_in_box(const struct orc_box *b, const orc_dot d)
{
    int l;
    for (l = 0; l < 3; l++)
        if (d[l] < b->min[l] || d[l] >= b->max[l]) return 0;
    return 1;
}

/* geom-sse.c:51-62 */
static void closest_vertex(const struct orc_box *b, const orc_dot d,
                           orc_dot res)
{
    int l;
    for (l = 0; l < 3; l++) {
        float d1 = fabsf(b->min[l] - d[l]), d2 = fabsf(b->max[l] - d[l]);
        res[l] = (d1 < d2) ? b->min[l] : b->max[l];
    }
    res[3] = (fabsf(b->min[3] - d[3]) < fabsf(b->max[3] - d[3])) ? b->min[3]
                                                                 : b->max[3];
}

/* voxvision.h.in:52 */
static int dot_equalp(const orc_dot a, const orc_dot b)
{
    return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}

/* tree.c:137-143 (SSE variant) */
static void update_bounding_box(struct orc_box *b, const orc_dot d)
{
    int l;
    for (l = 0; l < 3; l++) {
        float dmax = d[l] + orc_voxel[l];
        b->min[l] = smin(d[l], b->min[l]);
        b->max[l] = smax(dmax, b->max[l]);
    }
}

/* tree.c:128-134 (SSE variant) */
static void vox_align(orc_dot d)
{
    int l;
    for (l = 0; l < 3; l++) d[l] = floorf(d[l] / orc_voxel[l]) * orc_voxel[l];
}

/* tree.c:229-240 */
static struct orc_node *make_dense_leaf(const struct orc_box *box)
{
    struct orc_node *res = node_alloc(ORC_DENSE_LEAF);
    size_t dim[3];
    memcpy(&res->bounding_box, box, sizeof(*box));
    res->flags = ORC_DENSE_LEAF;
    get_dims(box, dim);
    res->dots_num = (unsigned int)(dim[0] * dim[1] * dim[2]);
    return res;
}

/* tree.c:364-420: dots of a subtree in DFS order; a dense leaf's voxels
 * by repeated addition of the voxel size from the box minimum */
static size_t flatten_tree(const struct orc_node *tree, orc_dot *set)
{
    size_t count = 0;
    if (!tree) return 0;
    if (tree->flags & ORC_LEAF) {
        memcpy(set, tree->data.dots, tree->dots_num * sizeof(orc_dot));
        count = tree->dots_num;
    } else if (tree->flags & ORC_DENSE_LEAF) {
        size_t dim[3], i, j, k;
        orc_dot cur = {0.0f, 0.0f, 0.0f, 0.0f};
        get_dims(&tree->bounding_box, dim);
        cur[0] = tree->bounding_box.min[0];
        for (i = 0; i < dim[0]; i++) {
            cur[1] = tree->bounding_box.min[1];
            for (j = 0; j < dim[1]; j++) {
                cur[2] = tree->bounding_box.min[2];
                for (k = 0; k < dim[2]; k++) {
                    memcpy(set[count], cur, sizeof(orc_dot));
                    cur[2] += orc_voxel[2];
                    count++;
                }
                cur[1] += orc_voxel[1];
            }
            cur[0] += orc_voxel[0];
        }
    } else {
        int i;
        for (i = 0; i < 8; i++) {
            size_t sub = flatten_tree(tree->data.inner.children[i], set);
            set += sub;
            count += sub;
        }
    }
    return count;
}

struct orc_node *orc_rebuild_tree(const struct orc_node *tree)
{
    struct orc_node *t = NULL;
    if (tree) {
        orc_dot *dots = aligned_alloc(16, tree->dots_num * sizeof(orc_dot) + 16);
        flatten_tree(tree, dots);
        t = orc_make_tree(dots, tree->dots_num);
        free(dots);
        if (t) {
            orc_tree_stats st;
            orc_tree_stats_get(t, &st);
        }
    }
    return t;
}

/* tree.c:437-458; the descent compares dot < center (geom-sse.c:64-68) */
static int voxel_in_tree(const struct orc_node *tree, const orc_dot v)
{
    unsigned i;
    if (tree && voxel_in_box(&tree->bounding_box, v)) {
        if (tree->flags & ORC_DENSE_LEAF) return 1;
        if (tree->flags & ORC_LEAF) {
            for (i = 0; i < tree->dots_num; i++)
                if (dot_equalp(v, tree->data.dots[i])) return 1;
        } else {
            return voxel_in_tree(
                tree->data.inner.children[orc_get_subspace_idx(
                    tree->data.inner.center, v)],
                v);
        }
    }
    return 0;
}

static void insert_(struct orc_node **tree_ptr, const orc_dot voxel);

/* tree.c:461-501 */
static struct orc_node *insert_in_big_dense(struct orc_node *tree,
                                            const orc_dot voxel)
{
    struct orc_node *node = node_alloc(0);
    orc_dot mid;
    int l, idx1, idx2;
    memcpy(&node->bounding_box, &tree->bounding_box, sizeof(struct orc_box));
    update_bounding_box(&node->bounding_box, voxel);
    node->flags = 0;
    node->dots_num = tree->dots_num + 1;
    closest_vertex(&tree->bounding_box, voxel, node->data.inner.center);
    for (l = 0; l < 4; l++)
        mid[l] = (tree->bounding_box.min[l] + tree->bounding_box.max[l]) / 2.0f;
    idx1 = subspace_simd(node->data.inner.center, mid);
    idx2 = subspace_simd(node->data.inner.center, voxel);
    node->data.inner.children[idx1] = tree;
    insert_(&node->data.inner.children[idx2], voxel);
    return node;
}

/* a leaf (or dense leaf) rebuilt with one more voxel */
static struct orc_node *remake_with(const struct orc_node *tree,
                                    const orc_dot voxel)
{
    orc_dot *dots = aligned_alloc(16, (tree->dots_num + 1) * sizeof(orc_dot) + 16);
    struct orc_node *node;
    flatten_tree(tree, dots);
    memcpy(dots[tree->dots_num], voxel, sizeof(orc_dot));
    node = orc_make_tree(dots, tree->dots_num + 1);
    free(dots);
    return node;
}

/* tree.c:504-605 */
static void insert_(struct orc_node **tree_ptr, const orc_dot voxel)
{
    struct orc_node *tree, *node;
    struct orc_box bb;
again:
    tree = *tree_ptr;
    node = tree;
    if (tree) {
        memcpy(&bb, &tree->bounding_box, sizeof(bb));
        update_bounding_box(&tree->bounding_box, voxel);
        if (orc_dense_set_p(&tree->bounding_box, tree->dots_num + 1)) {
            if (tree->flags & ORC_DENSE_LEAF) {
                tree->dots_num++;
            } else {
                node = make_dense_leaf(&tree->bounding_box);
                orc_destroy_tree(tree);
            }
            goto done;
        }
        if (tree->flags & ORC_DENSE_LEAF)
            memcpy(&tree->bounding_box, &bb, sizeof(bb));
    } else {
        orc_dot one[1];
        memcpy(one[0], voxel, sizeof(orc_dot));
        node = orc_make_tree(one, 1);
        goto done;
    }

    if (tree->flags & ORC_DENSE_LEAF) {
        if (tree->dots_num < 7) {
            node = remake_with(tree, voxel);
            orc_destroy_tree(tree);
        } else {
            node = insert_in_big_dense(tree, voxel);
        }
    } else if (tree->flags & ORC_LEAF) {
        if (tree->dots_num < 7) {
            memcpy(tree->data.dots[tree->dots_num], voxel, sizeof(orc_dot));
            tree->dots_num++;
        } else {
            node = remake_with(tree, voxel);
            orc_destroy_tree(tree);
        }
    } else {
        int idx = subspace_simd(tree->data.inner.center, voxel);
        tree->dots_num++;
        tree_ptr = &tree->data.inner.children[idx];
        goto again;
    }
done:
    *tree_ptr = node;
}

int orc_insert_voxel(struct orc_node **tree, const float xyz[3])
{
    orc_dot v = {xyz[0], xyz[1], xyz[2], 0.0f};
    int res;
    vox_align(v);
    res = !voxel_in_tree(*tree, v);
    if (res) {
        orc_tree_stats st;
        insert_(tree, v);
        orc_tree_stats_get(*tree, &st);
    }
    return res;
}

/* tree.c:628-681 */
static struct orc_node *delete_from_dense_stripe(struct orc_node *tree,
                                                 const orc_dot voxel)
{
    int idx = 0;
    struct orc_node *node, *child;
    struct orc_box bb;
    orc_dot other;
    int l;
    if (!stripep(&tree->bounding_box, &idx)) return NULL;
    if (dot_equalp(tree->bounding_box.min, voxel)) {
        tree->dots_num--;
        tree->bounding_box.min[idx] += orc_voxel[idx];
        return tree;
    }
    for (l = 0; l < 4; l++) other[l] = voxel[l] + orc_voxel[l];
    if (dot_equalp(tree->bounding_box.max, other)) {
        tree->dots_num--;
        tree->bounding_box.max[idx] -= orc_voxel[idx];
        return tree;
    }
    node = node_alloc(0);
    memcpy(&node->bounding_box, &tree->bounding_box, sizeof(struct orc_box));
    node->flags = 0;
    node->dots_num = tree->dots_num - 1;
    memcpy(node->data.inner.center, voxel, sizeof(orc_dot));
    divide_box(&tree->bounding_box, voxel, &bb, 1 << idx);
    node->data.inner.children[1 << idx] = make_dense_leaf(&bb);
    divide_box(&tree->bounding_box, voxel, &bb, 0);
    child = make_dense_leaf(&bb);
    node->data.inner.children[0] = child;
    child->dots_num--;
    child->bounding_box.min[idx] += orc_voxel[idx];
    orc_destroy_tree(tree);
    return node;
}

/* tree.c:683-745 */
static struct orc_node *delete_from_big_dense(struct orc_node *tree,
                                              const orc_dot voxel)
{
    struct orc_node *node;
    struct orc_box bb;
    int i, l;
    node = delete_from_dense_stripe(tree, voxel);
    if (node) return node;
    if (dot_equalp(tree->bounding_box.min, voxel)) {
        orc_dot other;
        for (l = 0; l < 4; l++) other[l] = voxel[l] + orc_voxel[l];
        node = node_alloc(0);
        memcpy(&node->bounding_box, &tree->bounding_box,
               sizeof(struct orc_box));
        node->flags = 0;
        node->dots_num = tree->dots_num - 1;
        memcpy(node->data.inner.center, other, sizeof(orc_dot));
        for (i = 0; i < 7; i++)
            if (divide_box(&tree->bounding_box, other, &bb, i))
                node->data.inner.children[i] = make_dense_leaf(&bb);
        orc_destroy_tree(tree);
        return node;
    }
    node = node_alloc(0);
    memcpy(&node->bounding_box, &tree->bounding_box, sizeof(struct orc_box));
    node->flags = 0;
    node->dots_num = tree->dots_num - 1;
    memcpy(node->data.inner.center, voxel, sizeof(orc_dot));
    for (i = 0; i < 8; i++)
        if (divide_box(&tree->bounding_box, voxel, &bb, i))
            node->data.inner.children[i] = make_dense_leaf(&bb);
    node->data.inner.children[0] =
        delete_from_big_dense(node->data.inner.children[0], voxel);
    orc_destroy_tree(tree);
    return node;
}

/* tree.c:748-818 */
static void delete_(struct orc_node **tree_ptr, const orc_dot voxel)
{
    struct orc_node *tree, *node;
    unsigned i;
again:
    tree = *tree_ptr;
    if (tree->dots_num == 1) {
        orc_destroy_tree(tree);
        *tree_ptr = NULL;
        return;
    }
    if (tree->flags & ORC_LEAF) {
        for (i = 0; i < tree->dots_num; i++)
            if (dot_equalp(tree->data.dots[i], voxel)) break;
        tree->dots_num--;
        memmove(tree->data.dots + i, tree->data.dots + i + 1,
                sizeof(orc_dot) * (tree->dots_num - i));
    } else if (tree->flags & ORC_DENSE_LEAF) {
        if (tree->dots_num <= 7) {
            orc_dot set[7];
            node = delete_from_dense_stripe(tree, voxel);
            if (node) {
                *tree_ptr = node;
                return;
            }
            flatten_tree(tree, set);
            for (i = 0; i < tree->dots_num; i++)
                if (dot_equalp(set[i], voxel)) break;
            memmove(set + i, set + i + 1,
                    sizeof(orc_dot) * (tree->dots_num - i - 1));
            *tree_ptr = orc_make_tree(set, tree->dots_num - 1);
            orc_destroy_tree(tree);
        } else {
            *tree_ptr = delete_from_big_dense(tree, voxel);
        }
    } else {
        int idx = subspace_simd(tree->data.inner.center, voxel);
        node = tree->data.inner.children[idx];
        if (node->dots_num == tree->dots_num) {
            /* the only child: it replaces its parent */
            *tree_ptr = node;
            free(node_hdr(tree));
            goto again;
        }
        tree->dots_num--;
        tree_ptr = &tree->data.inner.children[idx];
        goto again;
    }
}

int orc_delete_voxel(struct orc_node **tree, const float xyz[3])
{
    orc_dot v = {xyz[0], xyz[1], xyz[2], 0.0f};
    int res;
    vox_align(v);
    res = voxel_in_tree(*tree, v);
    if (res) {
        orc_tree_stats st;
        delete_(tree, v);
        if (*tree) orc_tree_stats_get(*tree, &st);
    }
    return res;
}

struct orc_node *orc_make_dense_leaf_box(const float mn[3], const float mx[3])
{
    struct orc_box b;
    struct orc_node *t;
    orc_tree_stats st;
    memset(&b, 0, sizeof(b));
    memcpy(b.min, mn, 3 * sizeof(float));
    memcpy(b.max, mx, 3 * sizeof(float));
    t = make_dense_leaf(&b);
    orc_tree_stats_get(t, &st);
    return t;
}

int orc_tree_flags(const struct orc_node *tree)
{
    return tree ? (int)tree->flags : -1;
}

/* ------------------------------------------------------------------ */
/* Inner nodes for the tests: box, centre and voxel count of each inner */
/* node, in depth-first preorder over children 0..7.                    */
/* ------------------------------------------------------------------ */
static void inner_walk(const struct orc_node *t, float *out, uint64_t *cnt,
                       size_t cap, size_t *k)
{
    int i;
    if (!t || (t->flags & ORC_LEAF_MASK)) return;
    if (*k < cap) {
        for (i = 0; i < 3; i++) {
            out[9 * *k + i] = t->bounding_box.min[i];
            out[9 * *k + 3 + i] = t->bounding_box.max[i];
            out[9 * *k + 6 + i] = t->data.inner.center[i];
        }
        cnt[*k] = t->dots_num;
    }
    (*k)++;
    for (i = 0; i < 8; i++) inner_walk(t->data.inner.children[i], out, cnt,
                                       cap, k);
}

size_t orc_tree_inner_nodes(const struct orc_node *tree, float *out,
                            uint64_t *counts, size_t cap)
{
    size_t k = 0;
    inner_walk(tree, out, counts, cap, &k);
    return k;
}

/* The fp32 ray directions of a frame, in the square layout (index
 * cs*16 + p): orc_screen2world of every pixel (simple-camera.c:21-31). */
void orc_camera_rays(const orc_camera *cam, int W, int H, float *dxyz)
{
    size_t np = (size_t)W * H, cs, ws = (size_t)(W >> 2);
    int p;
    for (cs = 0; cs < np / 16; cs++)
        for (p = 0; p < 16; p++) {
            orc_dot r;
            size_t i = cs * 16 + (size_t)p;
            orc_screen2world(cam, (int)(cs % ws) * 4 + (p & 3),
                             (int)(cs / ws) * 4 + (p >> 2), r);
            dxyz[i] = r[0];
            dxyz[np + i] = r[1];
            dxyz[2 * np + i] = r[2];
        }
}
