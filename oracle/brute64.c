/*
 * brute64.c -- tree-free reference answers used to pin the oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing here follows the reference's code: it
 * is the mathematics the reference's search must agree with.
 *   - orc_brute64_rays: FP64 slab test of every voxel box [v, v+V] against
 *     the half-line o + t d (t >= 0, closed boxes, an origin inside a box
 *     hits at t = 0); the first voxel hit wins.  A ray is flagged as a tie
 *     when fp32 rounding could change the answer: two entry distances
 *     within eps, a voxel grazed (entry and exit within eps, or missed by
 *     less than eps), or the hit point within eps of an edge of the hit
 *     voxel.  eps is relative to the scene's coordinate scale.
 *   - orc_brute_balls: the reference fp32 ball/box formula
 *     (src/voxtrees/geom-sse.c:144-149) over every voxel.  Because box
 *     clamping is exact and rounding is monotone, pruning by node boxes
 *     cannot change the answer, so the tree search must match this bit
 *     for bit.
 */
#include "vox_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    const float *dots;
    size_t nd;
    double V[3];
    const float *ox, *oy, *oz, *dx, *dy, *dz;
    double eps;
    int32_t *vidx;
    double *hit;
    uint8_t *tie;
    size_t n;
    atomic_size_t next;
    double scale;
} b64_ctx;

static void b64_one(b64_ctx *x, size_t r)
{
    double o[3] = {x->ox[r], x->oy[r], x->oz[r]};
    double d[3] = {x->dx[r], x->dy[r], x->dz[r]};
    double dn = fmax(fabs(d[0]), fmax(fabs(d[1]), fabs(d[2])));
    double scale = x->scale;
    double eps_len = x->eps * scale;
    double eps_t = eps_len / dn;
    double best = INFINITY, second = INFINITY;
    int32_t bi = -1;
    int tie = 0;
    size_t k;
    int a;
    for (k = 0; k < x->nd; k++) {
        double lo[3], hi[3], tmin = 0.0, tmax = INFINITY;
        int empty = 0, amb = 0;
        for (a = 0; a < 3; a++) {
            lo[a] = x->dots[3 * k + a];
            hi[a] = lo[a] + x->V[a];
            if (d[a] == 0.0) {
                if (o[a] < lo[a] || o[a] > hi[a]) {
                    empty = 1;
                    if (o[a] >= lo[a] - eps_len && o[a] <= hi[a] + eps_len)
                        amb = 1;
                }
                if (fabs(o[a] - lo[a]) <= eps_len ||
                    fabs(o[a] - hi[a]) <= eps_len)
                    amb = 1;
            } else {
                double t1 = (lo[a] - o[a]) / d[a];
                double t2 = (hi[a] - o[a]) / d[a];
                double tn = t1 < t2 ? t1 : t2, tf = t1 < t2 ? t2 : t1;
                if (tn > tmin) tmin = tn;
                if (tf < tmax) tmax = tf;
            }
        }
        if (empty) {
            if (amb && tmin <= tmax + eps_t) tie = 1; /* parallel grazing */
            continue;
        }
        if (tmin > tmax) {
            if (tmin - tmax <= eps_t) tie = 1; /* near miss */
            continue;
        }
        if (tmax - tmin <= eps_t) tie = 1; /* grazing hit */
        if (amb) tie = 1;
        if (tmin < best) {
            second = best;
            best = tmin;
            bi = (int32_t)k;
        } else if (tmin < second) {
            second = tmin;
        }
    }
    if (bi >= 0) {
        double p[3];
        int near = 0;
        if (second - best <= eps_t) tie = 1;
        for (a = 0; a < 3; a++) {
            double lo = x->dots[3 * bi + a], hi = lo + x->V[a];
            p[a] = o[a] + best * d[a];
            if (fabs(p[a] - lo) <= eps_len || fabs(p[a] - hi) <= eps_len)
                near++;
        }
        if (near >= 2) tie = 1;
        if (x->hit) {
            x->hit[3 * r + 0] = p[0];
            x->hit[3 * r + 1] = p[1];
            x->hit[3 * r + 2] = p[2];
        }
    }
    x->vidx[r] = bi;
    if (x->tie) x->tie[r] = (uint8_t)tie;
}

static void *b64_worker(void *p)
{
    b64_ctx *x = p;
    for (;;) {
        size_t b = atomic_fetch_add(&x->next, 16), e, r;
        if (b >= x->n) break;
        e = b + 16 < x->n ? b + 16 : x->n;
        for (r = b; r < e; r++) b64_one(x, r);
    }
    return NULL;
}

void orc_brute64_rays(const float *dots_xyz, size_t nd, const float voxel[3],
                      size_t n, const float *ox, const float *oy,
                      const float *oz, const float *dx, const float *dy,
                      const float *dz, double eps, int32_t *voxel_idx,
                      double *hit_xyz, uint8_t *tie, int nthreads)
{
    b64_ctx x;
    pthread_t th[256];
    size_t k, r;
    int t, a;
    double scale = 1.0;
    memset(&x, 0, sizeof(x));
    x.dots = dots_xyz;
    x.nd = nd;
    for (a = 0; a < 3; a++) x.V[a] = voxel[a];
    x.ox = ox; x.oy = oy; x.oz = oz;
    x.dx = dx; x.dy = dy; x.dz = dz;
    x.eps = eps;
    x.vidx = voxel_idx;
    x.hit = hit_xyz;
    x.tie = tie;
    x.n = n;
    for (k = 0; k < 3 * nd; k++) scale = fmax(scale, fabs(dots_xyz[k]) + 1.0);
    for (r = 0; r < n; r++) {
        scale = fmax(scale, fabs(ox[r]));
        scale = fmax(scale, fabs(oy[r]));
        scale = fmax(scale, fabs(oz[r]));
    }
    x.scale = scale;
    atomic_init(&x.next, 0);
    if (nthreads <= 0) nthreads = orc_cpu_count();
    if (nthreads > 256) nthreads = 256;
    for (t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, b64_worker, &x);
    for (t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}

typedef struct {
    const float *dots;
    size_t nd;
    float V[3];
    const float *cx, *cy, *cz, *r;
    uint8_t *out;
    size_t n;
    atomic_size_t next;
} bb_ctx;

static void *bb_worker(void *p)
{
    bb_ctx *x = p;
    for (;;) {
        size_t b = atomic_fetch_add(&x->next, 64), e, i, k;
        if (b >= x->n) break;
        e = b + 64 < x->n ? b + 64 : x->n;
        for (i = b; i < e; i++) {
            orc_dot c = {x->cx[i], x->cy[i], x->cz[i], 0.0f};
            uint8_t hit = 0;
            for (k = 0; k < x->nd && !hit; k++) {
                struct orc_box vb;
                int a;
                for (a = 0; a < 3; a++) {
                    vb.min[a] = x->dots[3 * k + a];
                    vb.max[a] = vb.min[a] + x->V[a];
                }
                vb.min[3] = vb.max[3] = 0.0f;
                hit = (uint8_t)orc_box_ball_interp(&vb, c, x->r[i]);
            }
            x->out[i] = hit;
        }
    }
    return NULL;
}

void orc_brute_balls(const float *dots_xyz, size_t nd, const float voxel[3],
                     size_t n, const float *cx, const float *cy,
                     const float *cz, const float *r, uint8_t *out,
                     int nthreads)
{
    bb_ctx x;
    pthread_t th[256];
    int t;
    memset(&x, 0, sizeof(x));
    x.dots = dots_xyz;
    x.nd = nd;
    x.V[0] = voxel[0];
    x.V[1] = voxel[1];
    x.V[2] = voxel[2];
    x.cx = cx; x.cy = cy; x.cz = cz; x.r = r;
    x.out = out;
    x.n = n;
    atomic_init(&x.next, 0);
    if (nthreads <= 0) nthreads = orc_cpu_count();
    if (nthreads > 256) nthreads = 256;
    for (t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, bb_worker, &x);
    for (t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}

/*
 * orc_dda64_rays: the answers of orc_brute64_rays for voxels on a grid
 * (every dot a multiple of the voxel size), without testing every voxel:
 * an FP64 grid walk (Amanatides-Woo) visits the cells the half-line passes
 * through in order of entry, and at each one applies the brute force's
 * test (b64_voxel) to the occupied voxels of the 3x3x3 cells around it,
 * each voxel once.  Every voxel the ray hits is a visited cell, and every
 * voxel it misses by less than a voxel is next to one, so the first hit,
 * the second entry and the near misses in front of the hit are seen.  The
 * walk stops two cells past the first hit: near misses further on cannot
 * change the first hit, so unlike the brute force they do not flag a tie
 * (the tie set is a subset of the brute force's, the voxels and hit points
 * are the same).
 */
typedef struct {
    double best, second;
    int32_t bi;
    int tie;
} b64_acc;

/* One voxel box [lo, lo+V] against the ray: the brute force's rules. */
static void b64_voxel(const double o[3], const double d[3], const double lo[3],
                      const double V[3], double eps_len, double eps_t,
                      int32_t k, b64_acc *acc)
{
    double tmin = 0.0, tmax = INFINITY;
    int empty = 0, amb = 0, a;
    for (a = 0; a < 3; a++) {
        double hi = lo[a] + V[a];
        if (d[a] == 0.0) {
            if (o[a] < lo[a] || o[a] > hi) {
                empty = 1;
                if (o[a] >= lo[a] - eps_len && o[a] <= hi + eps_len) amb = 1;
            }
            if (fabs(o[a] - lo[a]) <= eps_len || fabs(o[a] - hi) <= eps_len)
                amb = 1;
        } else {
            double t1 = (lo[a] - o[a]) / d[a];
            double t2 = (hi - o[a]) / d[a];
            double tn = t1 < t2 ? t1 : t2, tf = t1 < t2 ? t2 : t1;
            if (tn > tmin) tmin = tn;
            if (tf < tmax) tmax = tf;
        }
    }
    if (empty) {
        if (amb && tmin <= tmax + eps_t) acc->tie = 1;
        return;
    }
    if (tmin > tmax) {
        if (tmin - tmax <= eps_t) acc->tie = 1;
        return;
    }
    if (tmax - tmin <= eps_t) acc->tie = 1;
    if (amb) acc->tie = 1;
    if (tmin < acc->best) {
        acc->second = acc->best;
        acc->best = tmin;
        acc->bi = k;
    } else if (tmin < acc->second) {
        acc->second = tmin;
    }
}

typedef struct {
    const float *dots;
    double V[3];
    int lo[3], dim[3];
    const int32_t *grid;  /* cell -> dot index or -1 */
    const float *ox, *oy, *oz, *dx, *dy, *dz;
    double eps, scale;
    int32_t *vidx;
    double *hit;
    uint8_t *tie;
    size_t n;
    atomic_size_t next;
} dda_ctx;

#define DDA_SET 8192 /* per-ray set of tested voxels (open addressing) */

static int32_t dda_cell(const dda_ctx *x, long i, long j, long k)
{
    if (i < 0 || j < 0 || k < 0 || i >= x->dim[0] || j >= x->dim[1] ||
        k >= x->dim[2])
        return -1;
    return x->grid[((size_t)i * x->dim[1] + (size_t)j) * x->dim[2] +
                   (size_t)k];
}

static void dda_one(dda_ctx *x, size_t r, int32_t *set, uint32_t *gen,
                    uint32_t stamp)
{
    double o[3] = {x->ox[r], x->oy[r], x->oz[r]};
    double d[3] = {x->dx[r], x->dy[r], x->dz[r]};
    double dn = fmax(fabs(d[0]), fmax(fabs(d[1]), fabs(d[2])));
    double eps_len = x->eps * x->scale, eps_t = eps_len / dn;
    double g0[3], g1[3], t0 = 0.0, t1 = INFINITY, tdelta[3], tnext[3];
    long cell[3], step[3];
    b64_acc acc = {INFINITY, INFINITY, -1, 0};
    int a, ok = 1;
    /* the grid's box grown by one cell on every side (near misses) */
    for (a = 0; a < 3; a++) {
        g0[a] = (x->lo[a] - 1) * x->V[a];
        g1[a] = (x->lo[a] + x->dim[a] + 1) * x->V[a];
        if (d[a] == 0.0) {
            if (o[a] < g0[a] || o[a] > g1[a]) ok = 0;
        } else {
            double ta = (g0[a] - o[a]) / d[a], tb = (g1[a] - o[a]) / d[a];
            if (ta > tb) { double t = ta; ta = tb; tb = t; }
            if (ta > t0) t0 = ta;
            if (tb < t1) t1 = tb;
        }
    }
    if (ok && t0 <= t1) {
        for (a = 0; a < 3; a++) {
            double p = o[a] + t0 * d[a];
            double c = floor(p / x->V[a]);
            /* stay inside the grown box whatever the rounding */
            if (c < x->lo[a] - 1) c = x->lo[a] - 1;
            if (c > x->lo[a] + x->dim[a]) c = x->lo[a] + x->dim[a];
            cell[a] = (long)c;
            if (d[a] > 0) {
                step[a] = 1;
                tnext[a] = ((cell[a] + 1) * x->V[a] - o[a]) / d[a];
                tdelta[a] = x->V[a] / d[a];
            } else if (d[a] < 0) {
                step[a] = -1;
                tnext[a] = (cell[a] * x->V[a] - o[a]) / d[a];
                tdelta[a] = -x->V[a] / d[a];
            } else {
                step[a] = 0;
                tnext[a] = INFINITY;
                tdelta[a] = INFINITY;
            }
        }
        for (;;) {
            long di, dj, dk;
            double tcell;
            for (di = -1; di <= 1; di++)
                for (dj = -1; dj <= 1; dj++)
                    for (dk = -1; dk <= 1; dk++) {
                        long ci = cell[0] + di - x->lo[0];
                        long cj = cell[1] + dj - x->lo[1];
                        long ck = cell[2] + dk - x->lo[2];
                        int32_t k = dda_cell(x, ci, cj, ck);
                        uint32_t h;
                        double lo3[3];
                        if (k < 0) continue;
                        h = ((uint32_t)k * 2654435761u) & (DDA_SET - 1);
                        while (gen[h] == stamp && set[h] != k)
                            h = (h + 1) & (DDA_SET - 1);
                        if (gen[h] == stamp) continue; /* tested already */
                        gen[h] = stamp;
                        set[h] = k;
                        for (a = 0; a < 3; a++) lo3[a] = x->dots[3 * k + a];
                        b64_voxel(o, d, lo3, x->V, eps_len, eps_t, k, &acc);
                    }
            /* next cell: smallest crossing */
            a = tnext[0] < tnext[1] ? (tnext[0] < tnext[2] ? 0 : 2)
                                    : (tnext[1] < tnext[2] ? 1 : 2);
            tcell = tnext[a];
            if (tcell > t1) break;
            /* past the first hit by more than two cells: no voxel that
               could tie with it or precede it is left */
            if (acc.bi >= 0 &&
                tcell > acc.best + 2.0 * fmax(x->V[0], fmax(x->V[1],
                                              x->V[2])) / dn + eps_t)
                break;
            cell[a] += step[a];
            tnext[a] += tdelta[a];
        }
    }
    if (acc.bi >= 0) {
        double p[3];
        int near = 0;
        if (acc.second - acc.best <= eps_t) acc.tie = 1;
        for (a = 0; a < 3; a++) {
            double lo = x->dots[3 * acc.bi + a], hi = lo + x->V[a];
            p[a] = o[a] + acc.best * d[a];
            if (fabs(p[a] - lo) <= eps_len || fabs(p[a] - hi) <= eps_len)
                near++;
        }
        if (near >= 2) acc.tie = 1;
        if (x->hit) {
            x->hit[3 * r + 0] = p[0];
            x->hit[3 * r + 1] = p[1];
            x->hit[3 * r + 2] = p[2];
        }
    }
    x->vidx[r] = acc.bi;
    if (x->tie) x->tie[r] = (uint8_t)acc.tie;
}

static void *dda_worker(void *p)
{
    dda_ctx *x = p;
    int32_t *set = malloc(DDA_SET * sizeof(int32_t));
    uint32_t *gen = calloc(DDA_SET, sizeof(uint32_t));
    uint32_t stamp = 0;
    if (!set || !gen) abort();
    for (;;) {
        size_t b = atomic_fetch_add(&x->next, 16), e, r;
        if (b >= x->n) break;
        e = b + 16 < x->n ? b + 16 : x->n;
        for (r = b; r < e; r++) {
            if (++stamp == 0) {
                memset(gen, 0, DDA_SET * sizeof(uint32_t));
                stamp = 1;
            }
            dda_one(x, r, set, gen, stamp);
        }
    }
    free(set);
    free(gen);
    return NULL;
}

int orc_dda64_rays(const float *dots_xyz, size_t nd, const float voxel[3],
                   size_t n, const float *ox, const float *oy,
                   const float *oz, const float *dx, const float *dy,
                   const float *dz, double eps, int32_t *voxel_idx,
                   double *hit_xyz, uint8_t *tie, int nthreads)
{
    dda_ctx x;
    pthread_t th[256];
    size_t k, r, cells;
    int t, a, hi[3];
    double scale = 1.0;
    int32_t *grid;
    memset(&x, 0, sizeof(x));
    for (a = 0; a < 3; a++) {
        x.V[a] = voxel[a];
        x.lo[a] = 0;
        hi[a] = 0;
    }
    for (k = 0; k < nd; k++)
        for (a = 0; a < 3; a++) {
            double c = dots_xyz[3 * k + a] / x.V[a];
            if (c != floor(c) || fabs(c) > 1e7) return -22; /* off grid */
            if (k == 0 || (int)c < x.lo[a]) x.lo[a] = (int)c;
            if (k == 0 || (int)c > hi[a]) hi[a] = (int)c;
        }
    for (a = 0; a < 3; a++) x.dim[a] = nd ? hi[a] - x.lo[a] + 1 : 1;
    cells = (size_t)x.dim[0] * x.dim[1] * x.dim[2];
    if (cells > ((size_t)1 << 31)) return -34;
    grid = malloc(cells * sizeof(int32_t));
    if (!grid) return -12;
    memset(grid, 0xff, cells * sizeof(int32_t));
    for (k = 0; k < nd; k++) {
        size_t i = (size_t)((long)(dots_xyz[3 * k] / x.V[0]) - x.lo[0]);
        size_t j = (size_t)((long)(dots_xyz[3 * k + 1] / x.V[1]) - x.lo[1]);
        size_t l = (size_t)((long)(dots_xyz[3 * k + 2] / x.V[2]) - x.lo[2]);
        grid[(i * x.dim[1] + j) * x.dim[2] + l] = (int32_t)k;
    }
    x.grid = grid;
    x.dots = dots_xyz;
    x.ox = ox; x.oy = oy; x.oz = oz;
    x.dx = dx; x.dy = dy; x.dz = dz;
    x.eps = eps;
    x.vidx = voxel_idx;
    x.hit = hit_xyz;
    x.tie = tie;
    x.n = n;
    for (k = 0; k < 3 * nd; k++) scale = fmax(scale, fabs(dots_xyz[k]) + 1.0);
    for (r = 0; r < n; r++) {
        scale = fmax(scale, fabs(ox[r]));
        scale = fmax(scale, fabs(oy[r]));
        scale = fmax(scale, fabs(oz[r]));
    }
    x.scale = scale;
    atomic_init(&x.next, 0);
    if (nthreads <= 0) nthreads = orc_cpu_count();
    if (nthreads > 256) nthreads = 256;
    for (t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, dda_worker, &x);
    for (t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(grid);
    return 0;
}
