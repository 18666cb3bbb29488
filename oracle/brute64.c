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
