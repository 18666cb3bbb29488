// Conservative ray / box test of the render pre-pass.
//
// may_enter(lo, hi, o, d, rcp) is false only when the ray o + t d, t >= 0,
// misses the box [lo, hi] enlarged by M = 2^-17 S on every side, where S
// is the largest magnitude among the coordinates of o, lo and hi.  Every
// ray that the exact test of the traversal (hit_box, geom-sse.c:84-116)
// lets in passes within 2^-21 S of the box: its hit point is the correctly
// rounded o + m d (or a clamped coordinate of o) for a quotient m rounded
// once.  This test's slab bounds ((lo - M - o) * rcp(d), two rounded
// subtractions, a reciprocal within 1 ulp, a rounded product) are off by
// less than 2^-20 S / |d_i|, so a ray that the exact test lets in always
// passes here.  Inputs for which those bounds do not hold (non-finite
// values, coordinates beyond 2^40, a direction with all components below
// 2^-30) give true; a component below 2^-60 is treated as zero (the ray is
// then parallel to that slab for any t at which the other coordinates
// could still be inside the box).
//
// rcp(x) must return 1/x within 1 ulp for 2^-60 <= |x| <= 2^128.
#pragma once
#include <math.h>
#include <stdint.h>

// What may_enter needs of the box and the origin, computed once per origin
// (a camera position is the origin of all of its pixels' rays).
struct CullOrigin {
    float a[3], b[3];  // (lo - M) - o and (hi + M) - o per axis
    bool unsure;       // o not finite or too large: every ray may enter
};

// S: the largest magnitude among the coordinates of o, lo and hi
__host__ __device__ inline float cull_scale(const float lo[3],
                                            const float hi[3],
                                            const float o[3]) {
    float S = 0.0f;
    for (int i = 0; i < 3; ++i)
        S = fmaxf(S, fmaxf(fabsf(o[i]), fmaxf(fabsf(lo[i]), fabsf(hi[i]))));
    return S;
}

__host__ __device__ inline CullOrigin cull_origin(const float lo[3],
                                                  const float hi[3],
                                                  const float o[3]) {
    bool finite = true;
    for (int i = 0; i < 3; ++i) finite = finite && fabsf(o[i]) <= 0x1p127f;
    const float S = cull_scale(lo, hi, o);
    const float M = fmaxf(S * 0x1p-17f, 0x1p-100f);
    CullOrigin c;
    for (int i = 0; i < 3; ++i) {
        c.a[i] = (lo[i] - M) - o[i];
        c.b[i] = (hi[i] + M) - o[i];
    }
    c.unsure = !finite || S > 0x1p40f;
    return c;
}

// written without branches: the pre-pass runs it for every pixel
template <class Rcp>
__host__ __device__ inline bool may_enter_dir(const CullOrigin &c,
                                              const float d[3], Rcp rcp) {
    float dmax = 0.0f;
    bool finite = true;
    for (int i = 0; i < 3; ++i) {
        finite = finite && fabsf(d[i]) <= 0x1p127f;
        dmax = fmaxf(dmax, fabsf(d[i]));
    }
    float t0 = 0.0f, t1 = INFINITY;
    for (int i = 0; i < 3; ++i) {
        const bool par = !(fabsf(d[i]) >= 0x1p-60f);
        const float r = rcp(par ? 1.0f : d[i]);
        float ta = c.a[i] * r, tb = c.b[i] * r;
        // parallel to the slab: all t, or none when outside it (an
        // interval before t = 0, so that t1 < 0 <= t0)
        const bool out = c.a[i] > 0.0f || c.b[i] < 0.0f;
        ta = par ? (out ? -1.0f : -INFINITY) : ta;
        tb = par ? (out ? -1.0f : INFINITY) : tb;
        t0 = fmaxf(t0, fminf(ta, tb));
        t1 = fminf(t1, fmaxf(ta, tb));
    }
    return c.unsure || !finite || dmax < 0x1p-30f || t0 <= t1;
}

template <class Rcp>
__host__ __device__ inline bool may_enter(const float lo[3], const float hi[3],
                                          const float o[3], const float d[3],
                                          Rcp rcp) {
    return may_enter_dir(cull_origin(lo, hi, o), d, rcp);
}

// ---- whole 4x4 blocks of camera rays ----
//
// In exact arithmetic screen2world is affine in the pixel coordinates (it
// rotates (mul x - xsub, 1, ysub - mul y) by a fixed quaternion), so the
// exact directions of a 4x4 block lie, component by component, between
// those of its four corners.  A computed direction differs from the exact
// one by less than E = 2^-16 (Q^2 + 1) V, Q the largest quaternion
// component and V the largest magnitude among the unrotated vectors'
// components: counted term by term its roundings add up to at most
// 2^-24 (68 Q^2 + 3) V.  So the box [min - 3E, max + 3E] of the corner
// directions holds every computed direction of the block (the extra E
// covers the rounding of the bounds).
struct DirBox {
    float lo[3], hi[3];
    bool unsure;  // a non-finite or huge input: every ray may enter
};

__host__ __device__ inline DirBox block_dirs(const float d[4][3], float Q,
                                             float V) {
    DirBox b;
    const float E = 0x1p-16f * (Q * Q + 1.0f) * V;
    bool finite = E <= 0x1p100f;
    for (int i = 0; i < 3; ++i) {
        float mn = d[0][i], mx = d[0][i];
        for (int k = 0; k < 4; ++k) {
            finite = finite && fabsf(d[k][i]) <= 0x1p100f;
            mn = fminf(mn, d[k][i]);
            mx = fmaxf(mx, d[k][i]);
        }
        b.lo[i] = mn - 3.0f * E;
        b.hi[i] = mx + 3.0f * E;
    }
    b.unsure = !finite;
    return b;
}

// A quotient moved outward past its rounding (and past gradual
// underflow); infinities stay as they are.
__host__ __device__ inline float cull_up(float q) {
    return (q > 0.0f ? q * (1.0f + 0x1p-20f) : q * (1.0f - 0x1p-20f)) +
           0x1p-126f;
}
__host__ __device__ inline float cull_down(float q) {
    return (q > 0.0f ? q * (1.0f - 0x1p-20f) : q * (1.0f + 0x1p-20f)) -
           0x1p-126f;
}

// False only when no ray o + t d, t >= 0, with d anywhere in the box b
// comes within the margin of may_enter of the scene box: then, by the
// argument above, hit_box rejects every ray of the block.  For each axis
// the set of t at which some d_i in [lo_i, hi_i] puts o_i + t d_i between
// the enlarged slab's ends is cut out by t lo_i <= b_i and t hi_i >= a_i
// (a, b as in CullOrigin); the block may enter when those half-lines and
// t >= 0 meet.  The quotients are exact divisions moved outward.  A block
// some of whose directions could have every component below 2^-30 (may
// _enter gives true for those) also gives true.
__host__ __device__ inline bool block_may_enter(const CullOrigin &c,
                                                const DirBox &db) {
    if (c.unsure || db.unsure) return true;
    float t0 = 0.0f, t1 = INFINITY, big = 0.0f;
    bool ok = true;
    for (int i = 0; i < 3; ++i) {
        const float l = db.lo[i], u = db.hi[i], a = c.a[i], b = c.b[i];
        big = fmaxf(big, l > 0.0f ? l : (u < 0.0f ? -u : 0.0f));
        if (l > 0.0f) t1 = fminf(t1, cull_up(b / l));
        else if (l < 0.0f) t0 = fmaxf(t0, cull_down(b / l));
        else ok = ok && b >= 0.0f;
        if (u > 0.0f) t0 = fmaxf(t0, cull_down(a / u));
        else if (u < 0.0f) t1 = fminf(t1, cull_up(a / u));
        else ok = ok && a <= 0.0f;
    }
    return big < 0x1p-29f || (ok && t0 <= t1);
}

// ---- boxes below the root ----
//
// The traversal reports a hit at a point it computed inside a voxel's box
// (a dense leaf: inside the leaf's box), so inside the box of every node
// above that leaf: a node's box is the tight box of its voxels.  Each point
// the traversal computes on the way -- a box test's entry point, RN(m d +
// e) with m > 0 and one coordinate set to the box face, or a plane point
// fma(k, d, e) with k > 0 (geom-sse.c:84-142) -- lies, coordinate by
// coordinate, within 4 u S (u = 2^-24) of the point e + m d of the line
// through the point e before it, S the largest coordinate magnitude of the
// camera origin o and the root box: every such point lies in a node's box
// or is o.  The first e is o, and a path from the root to a voxel computes
// at most 2 D + 1 points in a tree of D levels (an entry point and a plane
// point per level, then the voxel's entry point; a FAST leaf try computes
// two), so a reported hit lies within (8 D + 4) u S of a point o + T d,
// T >= 0, coordinate by coordinate.
//
// Hence, with the box of a node enlarged by M = 2^-14 S (1024 u S; the
// rounding of a = (lo - M) - o and b = (hi + M) - o takes at most 4 u S of
// it), a block of rays none of which comes within that margin of any box
// of a cut of the tree (nodes below which every voxel lies) gets no hit in
// any quality mode, provided 8 D + 12 < 1024: trees of at most
// kDeepCullMaxDepth levels.
constexpr unsigned kDeepCullMaxDepth = 120;

// The CullOrigin of a box below the root, S = cull_scale() of the root box
// and o (the caller has checked the root's CullOrigin is not unsure).
__host__ __device__ inline CullOrigin deep_origin(const float lo[3],
                                                  const float hi[3],
                                                  const float o[3], float S) {
    const float M = fmaxf(S * 0x1p-14f, 0x1p-100f);
    CullOrigin c;
    for (int i = 0; i < 3; ++i) {
        c.a[i] = (lo[i] - M) - o[i];
        c.b[i] = (hi[i] + M) - o[i];
    }
    c.unsure = false;
    return c;
}

// A block tested against many boxes: the direction box's bounds are
// inverted once (rcp, within 1 ulp as for may_enter_dir), and each slab
// bound b / l becomes a product b * (1 / l) * (1 +- 2^-20) that lies
// outside the exact quotient: the factor of larger magnitude where that
// moves the bound outward, the smaller one elsewhere (the reciprocal, the
// factor and the product are off by less than 4.01 2^-24 relative), and
// 2^-126 more for underflow, as cull_up and cull_down do for an exact
// quotient.  A bound below 2^-60 in magnitude counts as zero: when the
// test goes on at all, some component of every direction in the box is at
// least 2^-29 in magnitude, so a ray can stay within M of a box only for
// t <= 2.01 S / 2^-29, over which that component moves the ray less than
// 2^-30 S -- the margin above covers it.
struct DirRcp {
    // 1 / lo and 1 / hi times 1 + 2^-20 (big) and 1 - 2^-20 (small); all
    // 0 for a bound counted as zero
    float lbig[3], lsmall[3], ubig[3], usmall[3];
    int8_t ls[3], us[3];  // signs of the bounds (0: counted as zero)
    bool any;             // unsure, or tiny directions: every box may be entered
};

// the `any` of dir_rcp alone
__host__ __device__ inline bool dir_any(const DirBox &db) {
    float big = 0.0f;
    for (int i = 0; i < 3; ++i) {
        const float l = db.lo[i], u = db.hi[i];
        big = fmaxf(big, l > 0.0f ? l : (u < 0.0f ? -u : 0.0f));
    }
    return db.unsure || !(big >= 0x1p-29f);
}

template <class Rcp>
__host__ __device__ inline DirRcp dir_rcp(const DirBox &db, Rcp rcp) {
    DirRcp r;
    for (int i = 0; i < 3; ++i) {
        const float l = db.lo[i], u = db.hi[i];
        const bool lz = !(fabsf(l) >= 0x1p-60f), uz = !(fabsf(u) >= 0x1p-60f);
        const float rl = rcp(lz ? 1.0f : l), ru = rcp(uz ? 1.0f : u);
        r.lbig[i] = lz ? 0.0f : rl * (1.0f + 0x1p-20f);
        r.lsmall[i] = lz ? 0.0f : rl * (1.0f - 0x1p-20f);
        r.ubig[i] = uz ? 0.0f : ru * (1.0f + 0x1p-20f);
        r.usmall[i] = uz ? 0.0f : ru * (1.0f - 0x1p-20f);
        r.ls[i] = lz ? 0 : (l > 0.0f ? 1 : -1);
        r.us[i] = uz ? 0 : (u > 0.0f ? 1 : -1);
    }
    r.any = dir_any(db);
    return r;
}

// block_may_enter with the bounds of dir_rcp (false only when no direction
// of the box comes within the box's margin, as above).  Per axis: t lo <= b
// (lo > 0: t <= b / lo; lo < 0: t >= b / lo) and t hi >= a (hi > 0:
// t >= a / hi; hi < 0: t <= a / hi).
__host__ __device__ inline bool block_may_enter_rcp(const CullOrigin &c,
                                                    const DirRcp &r) {
    if (r.any || c.unsure) return true;
    float t0 = 0.0f, t1 = INFINITY;
    bool ok = true;
    for (int i = 0; i < 3; ++i) {
        const float a = c.a[i], b = c.b[i];
        const float ql = b * (b > 0.0f ? r.lbig[i] : r.lsmall[i]);
        const float qu = a * (a > 0.0f ? r.usmall[i] : r.ubig[i]);
        t1 = r.ls[i] > 0 ? fminf(t1, ql + 0x1p-126f) : t1;
        t0 = r.ls[i] < 0 ? fmaxf(t0, ql - 0x1p-126f) : t0;
        ok = ok && (r.ls[i] != 0 || b >= 0.0f);
        t0 = r.us[i] > 0 ? fmaxf(t0, qu - 0x1p-126f) : t0;
        t1 = r.us[i] < 0 ? fminf(t1, qu + 0x1p-126f) : t1;
        ok = ok && (r.us[i] != 0 || a <= 0.0f);
    }
    return ok && t0 <= t1;
}
