/*
 * vox_lights.c -- CPU oracle: M-tree of light spheres and the lit colour.
 *
 * TEST INFRASTRUCTURE ONLY (see vox_oracle.h).  Restates
 *   src/voxtrees/mtree.c:11-489  (M-tree: insert with node splits, bounding
 *                                 spheres, removal, the containing-spheres
 *                                 walk and its callback order)
 *   src/voxrnd/lights.c:1-116    (light manager, vox_get_light)
 * Floating point as the reference's clang build computes it: sqr_metric is
 * (dx*dx + dy*dy) + dz*dz (geom-sse.c:12-17); pow(r, 2) rounds r*r to
 * float once; the bounding-sphere centre c*(a-b)+b and the light sum
 * color + add*light_color are fused multiply-adds (one expression each).
 */
#include "vox_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MT_MAX 9 /* MTREE_MAX_CHILDREN, src/voxtrees/mtree.h */

struct orc_mt_node {
    orc_sphere bs;
    struct orc_mt_node *parent;
    unsigned leaf, num;
    union {
        struct orc_mt_node *children[MT_MAX + 1];
        orc_sphere *spheres;
    } d;
};

struct orc_lights {
    float ambient[3];
    struct orc_mt_node *root;
};

static float sqr_metric3(const float *a, const float *b)
{
    float x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    x = x * x;
    y = y * y;
    z = z * z;
    return (x + y) + z;
}

static float sqr_f(float r) { return r * r; } /* (float)pow(r, 2) */

static struct orc_mt_node *mt_alloc(int leaf)
{
    struct orc_mt_node *n = calloc(1, sizeof(*n));
    n->leaf = (unsigned)leaf;
    if (leaf) n->d.spheres = calloc(MT_MAX + 1, sizeof(orc_sphere));
    return n;
}

static const orc_sphere *item_sphere(const struct orc_mt_node *n, unsigned i)
{
    return n->leaf ? &n->d.spheres[i] : &n->d.children[i]->bs;
}

/* mtree.c:22-56 */
static int bounding_sphere(const orc_sphere *s, orc_sphere *b)
{
    float dist = sqrtf(sqr_metric3(s->center, b->center));
    float crit = (1 + (b->radius - s->radius) / dist) / 2;
    int l;
    if (crit < 0) {
        *b = *s;
    } else if (crit < 1) {
        b->radius = ((s->radius + b->radius) + dist) / 2;
        b->sqr_radius = sqr_f(b->radius);
        for (l = 0; l < 3; l++)
            b->center[l] = fmaf(crit, b->center[l] - s->center[l],
                                s->center[l]);
    } else {
        return 0;
    }
    return 1;
}

/* mtree.c:59-76 */
static int node_bounding_sphere(struct orc_mt_node *n)
{
    unsigned i;
    int changed = 0;
    n->bs = *item_sphere(n, 0);
    for (i = 1; i < n->num; i++) changed += bounding_sphere(item_sphere(n, i),
                                                            &n->bs);
    return !!changed;
}

/* mtree.c:82-89 */
static void propagate(struct orc_mt_node *n)
{
    while (n && node_bounding_sphere(n)) n = n->parent;
}

/* mtree.c:91-115: the pair of items with the largest squared distance
 * (strictly larger, first found wins).  The reference leaves c1, c2 unset
 * when all centres coincide; that cannot happen here, inserts of a centre
 * already present are refused. */
static void farest_centers(const struct orc_mt_node *n, float *c1, float *c2)
{
    unsigned i, j;
    float maxdist = 0;
    memcpy(c1, item_sphere(n, 0)->center, 3 * sizeof(float));
    memcpy(c2, item_sphere(n, 0)->center, 3 * sizeof(float));
    for (i = 0; i < n->num; i++)
        for (j = i + 1; j < n->num; j++) {
            float d = sqr_metric3(item_sphere(n, i)->center,
                                  item_sphere(n, j)->center);
            if (d > maxdist) {
                memcpy(c1, item_sphere(n, i)->center, 3 * sizeof(float));
                memcpy(c2, item_sphere(n, j)->center, 3 * sizeof(float));
                maxdist = d;
            }
        }
}

/* mtree.c:117-189 */
static struct orc_mt_node *split_node(struct orc_mt_node *node)
{
    float c1[3], c2[3];
    unsigned i;
    struct orc_mt_node *parent = node->parent, *res = parent;
    struct orc_mt_node *n1 = mt_alloc((int)node->leaf);
    struct orc_mt_node *n2 = mt_alloc((int)node->leaf);
    n1->parent = n2->parent = parent;
    farest_centers(node, c1, c2);
    for (i = 0; i < node->num; i++) {
        const float *c = item_sphere(node, i)->center;
        float d1 = sqr_metric3(c, c1), d2 = sqr_metric3(c, c2);
        struct orc_mt_node *w = d1 < d2 ? n1 : n2;
        if (node->leaf) {
            w->d.spheres[w->num] = node->d.spheres[i];
        } else {
            w->d.children[w->num] = node->d.children[i];
            node->d.children[i]->parent = w;
        }
        w->num++;
    }
    node_bounding_sphere(n1);
    node_bounding_sphere(n2);
    if (node->leaf) free(node->d.spheres);
    free(node);
    if (parent) {
        for (i = 0; i < parent->num; i++)
            if (parent->d.children[i] == node) parent->d.children[i] = n1;
        parent->d.children[parent->num++] = n2;
        propagate(parent);
        if (parent->num == MT_MAX + 1) res = split_node(parent);
    } else {
        res = mt_alloc(0);
        res->num = 2;
        res->d.children[0] = n1;
        res->d.children[1] = n2;
        node_bounding_sphere(res);
        n1->parent = n2->parent = res;
    }
    return res;
}

/* mtree.c:191-217 */
static struct orc_mt_node *contains(struct orc_mt_node *n, const float *c,
                                    float r)
{
    unsigned i;
    if (!n) return NULL;
    if (!(sqr_metric3(c, n->bs.center) < n->bs.sqr_radius)) return NULL;
    if (n->leaf) {
        for (i = 0; i < n->num; i++)
            if (n->d.spheres[i].center[0] == c[0] &&
                n->d.spheres[i].center[1] == c[1] &&
                n->d.spheres[i].center[2] == c[2] &&
                n->d.spheres[i].radius == r)
                return n;
        return NULL;
    }
    for (i = 0; i < n->num; i++) {
        struct orc_mt_node *f = contains(n->d.children[i], c, r);
        if (f) return f;
    }
    return NULL;
}

/* mtree.c:219-255 */
static struct orc_mt_node *add_helper(struct orc_mt_node *node,
                                      const orc_sphere *s)
{
    struct orc_mt_node *res = node, *child = NULL;
    unsigned i;
    float mindist = INFINITY;
    if (!node) {
        res = mt_alloc(1);
        res->bs = *s;
        res->num = 1;
        res->d.spheres[0] = *s;
    } else if (node->leaf) {
        node->d.spheres[node->num++] = *s;
        propagate(node);
        if (node->num == MT_MAX + 1) res = split_node(node);
    } else {
        for (i = 0; i < node->num; i++) {
            float d = sqr_metric3(s->center, node->d.children[i]->bs.center);
            if (d < mindist) {
                mindist = d;
                child = node->d.children[i];
            }
        }
        res = add_helper(child, s);
    }
    return res;
}

static int center_present(const struct orc_mt_node *n, const float *c)
{
    unsigned i;
    if (!n) return 0;
    if (n->leaf) {
        for (i = 0; i < n->num; i++)
            if (n->d.spheres[i].center[0] == c[0] &&
                n->d.spheres[i].center[1] == c[1] &&
                n->d.spheres[i].center[2] == c[2])
                return 1;
        return 0;
    }
    for (i = 0; i < n->num; i++)
        if (center_present(n->d.children[i], c)) return 1;
    return 0;
}

static void mt_destroy(struct orc_mt_node *n)
{
    unsigned i;
    if (!n) return;
    if (n->leaf) free(n->d.spheres);
    else
        for (i = 0; i < n->num; i++) mt_destroy(n->d.children[i]);
    free(n);
}

/* mtree.c:278-322 */
static struct orc_mt_node *node_delete(struct orc_mt_node *node)
{
    struct orc_mt_node *parent = node->parent;
    unsigned i;
    mt_destroy(node);
    if (!parent) return NULL;
    for (i = 0; i < parent->num; i++)
        if (parent->d.children[i] == node) break;
    if (parent->num == 2) {
        struct orc_mt_node *other = parent->d.children[1 - i];
        struct orc_mt_node *top = parent->parent;
        free(parent);
        other->parent = top;
        if (!top) return other;
        for (i = 0; i < top->num; i++)
            if (top->d.children[i] == parent) break;
        top->d.children[i] = other;
        propagate(top);
        return top;
    }
    parent->num--;
    memmove(&parent->d.children[i], &parent->d.children[i + 1],
            sizeof(parent->d.children[0]) * (parent->num - i));
    parent->d.children[parent->num] = NULL;
    propagate(parent);
    return parent;
}

/* ---- light manager: src/voxrnd/lights.c ---- */
orc_lights *orc_lights_create(void)
{
    orc_lights *l = calloc(1, sizeof(*l));
    l->ambient[0] = l->ambient[1] = l->ambient[2] = 1.0f; /* lights.c:20 */
    return l;
}

void orc_lights_destroy(orc_lights *l)
{
    if (!l) return;
    mt_destroy(l->root);
    free(l);
}

static int check_light(const float *c) /* lights.c:81-89 */
{
    int i;
    for (i = 0; i < 3; i++)
        if (c[i] < 0 || c[i] > 1) return 0;
    return 1;
}

int orc_lights_set_ambient(orc_lights *l, const float color[3])
{
    if (!check_light(color)) return 0;
    memcpy(l->ambient, color, sizeof(l->ambient));
    return 1;
}

/* lights.c:44-57 + mtree.c:257-270.  Returns 1 inserted, 0 refused as the
 * reference refuses (colour outside [0, 1], radius not positive, the same
 * sphere present), -1 refused where the reference is undefined: a centre
 * already present (its M-tree cannot hold two spheres with one centre,
 * src/voxtrees/mtree.h, vox_mtree_add_sphere) or a non-finite centre or
 * radius. */
int orc_lights_insert(orc_lights *l, const float center[3], float radius,
                      const float color[3])
{
    orc_sphere s;
    struct orc_mt_node *n;
    if (!check_light(color)) return 0;
    if (!(radius > 0) || contains(l->root, center, radius)) return 0;
    if (!isfinite(center[0]) || !isfinite(center[1]) ||
        !isfinite(center[2]) || !isfinite(radius))
        return -1; /* undefined in the reference (NaN distances) */
    if (center_present(l->root, center)) return -1;
    memset(&s, 0, sizeof(s));
    memcpy(s.center, center, 3 * sizeof(float));
    memcpy(s.color, color, 3 * sizeof(float));
    s.radius = radius;
    s.sqr_radius = sqr_f(radius);
    n = add_helper(l->root, &s);
    if (!n->parent) l->root = n;
    return 1;
}

/* lights.c:59-67 + mtree.c:325-369 */
int orc_lights_delete(orc_lights *l, const float center[3], float radius)
{
    struct orc_mt_node *leaf = contains(l->root, center, radius), *last;
    unsigned i;
    if (!(radius > 0) || !leaf) return 0;
    if (leaf->num == 1) {
        last = node_delete(leaf);
    } else {
        for (i = 0; i < leaf->num; i++)
            if (leaf->d.spheres[i].center[0] == center[0] &&
                leaf->d.spheres[i].center[1] == center[1] &&
                leaf->d.spheres[i].center[2] == center[2] &&
                leaf->d.spheres[i].radius == radius)
                break;
        memmove(&leaf->d.spheres[i], &leaf->d.spheres[i + 1],
                sizeof(orc_sphere) * (leaf->num - i - 1));
        leaf->num--;
        propagate(leaf);
        last = leaf;
    }
    if (!last || !last->parent) l->root = last;
    return 1;
}

void orc_lights_clear(orc_lights *l)
{
    mt_destroy(l->root);
    l->root = NULL;
}

static size_t mt_count(const struct orc_mt_node *n)
{
    size_t k = 0;
    unsigned i;
    if (!n) return 0;
    if (n->leaf) return n->num;
    for (i = 0; i < n->num; i++) k += mt_count(n->d.children[i]);
    return k;
}

size_t orc_lights_count(const orc_lights *l) { return mt_count(l->root); }

/* lights.c:91-101: the light callback, one fused multiply-add a channel */
static void add_light(const orc_sphere *s, const float *p, float *col)
{
    float dist = fminf(s->radius, sqrtf(sqr_metric3(s->center, p)));
    float add = 1 - dist / s->radius;
    int l;
    for (l = 0; l < 3; l++) col[l] = fmaf(add, s->color[l], col[l]);
}

/* mtree.c:463-491: depth-first over children in order; a one-sphere leaf
 * reports its sphere on the leaf's own bounding-sphere test */
static void containing(const struct orc_mt_node *n, const float *p,
                       float *col)
{
    unsigned i;
    if (!n) return;
    if (!(sqr_metric3(n->bs.center, p) < n->bs.sqr_radius)) return;
    if (n->leaf) {
        if (n->num == 1) {
            add_light(&n->d.spheres[0], p, col);
            return;
        }
        for (i = 0; i < n->num; i++)
            if (sqr_metric3(n->d.spheres[i].center, p) <
                n->d.spheres[i].sqr_radius)
                add_light(&n->d.spheres[i], p, col);
        return;
    }
    for (i = 0; i < n->num; i++) containing(n->d.children[i], p, col);
}

/* lights.c:103-116 */
void orc_get_light(const orc_lights *l, const float p[3], float light[3])
{
    float col[3];
    memcpy(col, l->ambient, sizeof(col));
    containing(l->root, p, col);
    memcpy(light, col, sizeof(col));
}

/* The spheres in the order of a depth-first walk (children in order): the
 * order in which the containing-spheres walk reports them.  Writes at most
 * cap spheres; returns how many there are. */
static void walk_all(const struct orc_mt_node *n, orc_sphere *out, size_t cap,
                     size_t *k)
{
    unsigned i;
    if (!n) return;
    if (n->leaf) {
        for (i = 0; i < n->num; i++) {
            if (*k < cap) out[*k] = n->d.spheres[i];
            (*k)++;
        }
        return;
    }
    for (i = 0; i < n->num; i++) walk_all(n->d.children[i], out, cap, k);
}

size_t orc_lights_spheres(const orc_lights *l, orc_sphere *out, size_t cap)
{
    size_t k = 0;
    walk_all(l->root, out, cap, &k);
    return k;
}

/* Height of the M-tree (0 when empty). */
static unsigned mt_height(const struct orc_mt_node *n)
{
    unsigned h = 0, i;
    if (!n) return 0;
    if (n->leaf) return 1;
    for (i = 0; i < n->num; i++) {
        unsigned c = mt_height(n->d.children[i]);
        if (c > h) h = c;
    }
    return h + 1;
}

unsigned orc_lights_height(const orc_lights *l) { return mt_height(l->root); }
