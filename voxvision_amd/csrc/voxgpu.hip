// libvoxgpu: HIP kernels (gfx950) and the C ABI declared in include/voxgpu.h.
//
// Ray traversal (vox_ray_tree_intersection, src/voxtrees/search.c:8-153).
// The reference recursion stops at the first leaf that reports a hit, so it
// is a depth-first search over (node, origin) pairs in a fixed order:
// the child holding the entry point first, then the children behind the
// dividing planes in L1-distance order.  Each lane keeps the pair it is
// working on in registers and the pending pairs on a stack.  Plane points
// depend only on the node and its entry point, so they are computed when
// the node is visited and pushed (last-visited first); the visit order,
// every origin and every rounding step are the reference's, which makes
// the results bit-identical.  The stack is a ring of K entries per lane in
// LDS; older entries spill to an HBM buffer and come back when the ring
// empties, so any depth stays exact.
//
// Work distribution: persistent waves.  A wave keeps 64 rays in flight and
// refills idle lanes from a global ray counter (one atomic per refill,
// lanes numbered with a ballot and mbcnt), so long rays do not hold a whole
// wave.  The counter resets itself when the last wave leaves, so a launch
// needs no memset and can be captured in a graph.
//
// Floating point: compiled with -ffp-contract=off; __builtin_fmaf marks the
// reference's fused sites (geom-sse.c:109,135; simple-camera.c:29;
// vect-ops-sse.c:24).  Division is IEEE correctly rounded (HIP default).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "cull.h"
#include "fastdiv.h"
#include "lights.h"
#include "tree_build.h"
#include "voxg_internal.h"
#include "../../include/voxgpu_testing.h"
#include "voxg_layout.h"

using namespace voxg;

// Entry bound of the traversal stack: one entry per inner node on the
// current path that still has a child to visit, so at most one per inner
// level.
static uint32_t stack_bound_of(const HostTree &h) {
    return h.depth > 1 ? h.depth - 1u : 1u;
}

// Bytes of one stack entry (ring and spill buffer): node index and entry
// point (16 B), then the node's pending list (4 B) in a second array.
constexpr size_t kEntryBytes = sizeof(uint4) + sizeof(uint32_t);

// The render pre-pass's cut of the tree: the nodes of the deepest level
// that has at most kCutMax of them, counting the leaves of the levels above
// (every voxel lies below exactly one), as boxes of 6 floats (min, max),
// most voxels first so that a block that may enter finds its box early.
// Empty (no test beyond the root box) when the cut would be the root alone
// or the tree is deeper than the bound of the test (cull.h).
constexpr size_t kCutMax = 64;  // the pre-pass tables at most 64 boxes

static std::vector<float> cut_of(const HostTree &h) {
    std::vector<float> out;
    if (h.nodes.empty() || h.depth > kDeepCullMaxDepth) return out;
    std::vector<uint32_t> cur{0u}, next;
    for (;;) {
        next.clear();
        bool inner = false;
        for (uint32_t i : cur) {
            const NodeRec &n = h.nodes[i];
            if (node_type(n.info) == NT_INNER) {
                inner = true;
                const uint32_t k = (uint32_t)__builtin_popcount(child_mask(n.info));
                for (uint32_t c = 0; c < k; ++c) next.push_back(n.ref + c);
            } else {
                next.push_back(i);
            }
        }
        if (!inner || next.size() > kCutMax) break;
        cur.swap(next);
    }
    if (cur.size() < 2) return out;
    if (h.counts.size() == h.nodes.size())
        std::stable_sort(cur.begin(), cur.end(), [&](uint32_t x, uint32_t y) {
            return h.counts[x] > h.counts[y];
        });
    for (uint32_t i : cur) {
        const NodeRec &n = h.nodes[i];
        out.insert(out.end(), {n.bmin[0], n.bmin[1], n.bmin[2], n.bmax[0],
                               n.bmax[1], n.bmax[2]});
    }
    return out;
}

// HBM of an uploaded tree: node records, dots, then the cut's boxes
struct TreeBytes {
    size_t dots_off, cut_off, total;
};

static TreeBytes tree_bytes(const HostTree &h, size_t cut_floats) {
    const size_t nb = h.nodes.size() * sizeof(NodeRec);
    const size_t db = h.dots.size() * sizeof(float);
    TreeBytes b;
    b.dots_off = (nb + 255) & ~(size_t)255;
    b.cut_off = (b.dots_off + db + 255) & ~(size_t)255;
    b.total = cut_floats ? b.cut_off + cut_floats * sizeof(float)
                         : b.dots_off + db;
    return b;
}

static size_t device_bytes_of(const HostTree &h) {
    return tree_bytes(h, cut_of(h).size()).total;
}

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                       \
    do {                                                                   \
        hipError_t e_ = (expr);                                            \
        if (e_ != hipSuccess)                                              \
            return fail(VOXG_EIO, std::string(#expr ": ") +                \
                                      hipGetErrorString(e_));              \
    } while (0)

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
namespace dev {

// MINPS / MAXPS lane semantics: the second operand wins unless the first
// compares strictly smaller / larger.
__device__ __forceinline__ float smin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float smax(float a, float b) { return a > b ? a : b; }

struct f3 {
    float x, y, z;
};

// geom-sse.c:84-116 (fit_into_box :34-40)
__device__ __forceinline__ bool hit_box(f3 bmin, f3 bmax, f3 o, f3 d, f3 &r) {
    float fx = smin(smax(o.x, bmin.x), bmax.x);
    float fy = smin(smax(o.y, bmin.y), bmax.y);
    float fz = smin(smax(o.z, bmin.z), bmax.z);
    float sx = fx - o.x, sy = fy - o.y, sz = fz - o.z;
    if (sx == 0.0f && sy == 0.0f && sz == 0.0f) {
        r = o;
        return true;
    }
    float tx = (d.x == 0.0f) ? 0.0f : sx / d.x;
    float ty = (d.y == 0.0f) ? 0.0f : sy / d.y;
    float tz = (d.z == 0.0f) ? 0.0f : sz / d.z;
    // lane-wise result of the shuffle/max network (:101-105)
    float m0 = smax(smax(tx, tz), smax(tx, ty));
    if (m0 <= 0.0f) return false;
    float m1 = smax(smax(ty, tx), smax(ty, tz));
    float m2 = smax(smax(tz, ty), smax(tz, tx));
    float hx = (m0 == tx) ? fx : __builtin_fmaf(m0, d.x, o.x);
    float hy = (m1 == ty) ? fy : __builtin_fmaf(m1, d.y, o.y);
    float hz = (m2 == tz) ? fz : __builtin_fmaf(m2, d.z, o.z);
    if (hx < bmin.x || hx > bmax.x || hy < bmin.y || hy > bmax.y ||
        hz < bmin.z || hz > bmax.z)
        return false;
    r = f3{hx, hy, hz};
    return true;
}

__device__ __forceinline__ float comp(f3 v, int i) {
    return i == 0 ? v.x : (i == 1 ? v.y : v.z);
}

// geom-sse.c:118-142
__device__ __forceinline__ bool hit_plane(f3 o, f3 d, f3 c, int p, f3 bmin,
                                          f3 bmax, f3 &r) {
    float cp = comp(c, p), op = comp(o, p), dp = comp(d, p);
    float k = cp - op;
    if (k == 0.0f || ((dp < 0.0f) != (k < 0.0f))) return false;
    k = k / dp;
    float rx = __builtin_fmaf(k, d.x, o.x);
    float ry = __builtin_fmaf(k, d.y, o.y);
    float rz = __builtin_fmaf(k, d.z, o.z);
    if (rx < bmin.x || rx > bmax.x || ry < bmin.y || ry > bmax.y ||
        rz < bmin.z || rz > bmax.z)
        return false;
    r = f3{rx, ry, rz};
    return true;
}

// geom-sse.c:4-10
__device__ __forceinline__ float abs_metric(f3 a, f3 b) {
    return (fabsf(a.x - b.x) + fabsf(a.y - b.y)) + fabsf(a.z - b.z);
}

// geom-sse.c:70-82
__device__ __forceinline__ int corrected_subspace(f3 c, f3 p, f3 d) {
    int bx = (p.x == c.x) ? (d.x < 0.0f) : (p.x < c.x);
    int by = (p.y == c.y) ? (d.y < 0.0f) : (p.y < c.y);
    int bz = (p.z == c.z) ? (d.z < 0.0f) : (p.z < c.z);
    return bx | (by << 1) | (bz << 2);
}

// geom-sse.c:144-149 with the radius squared precomputed (radius*radius,
// one rounding, as in the reference expression)
__device__ __forceinline__ bool box_ball(f3 bmin, f3 bmax, f3 c, float rr) {
    float fx = smin(smax(c.x, bmin.x), bmax.x);
    float fy = smin(smax(c.y, bmin.y), bmax.y);
    float fz = smin(smax(c.z, bmin.z), bmax.z);
    float dx = fx - c.x, dy = fy - c.y, dz = fz - c.z;
    dx = dx * dx;
    dy = dy * dy;
    dz = dz * dz;
    return ((dx + dy) + dz) < rr;
}

struct Tree {
    const NodeRec *__restrict__ nodes;
    const float *__restrict__ dots;
    float vx, vy, vz;
    uint32_t empty;
    float ivx, ivy, ivz;  // 1/V when V is a power of two, else 0
    uint32_t ncut;        // boxes of the render pre-pass's cut (0: none)
    const float *__restrict__ cut;  // 6 floats per box: min, max
};

// x / v, as a multiplication when v is a power of two (iv = 1/v exactly):
// both round the exact quotient once, so the bits are the same.
__device__ __forceinline__ float div_v(float x, float v, float iv) {
    return iv != 0.0f ? x * iv : x / v;
}

struct Node {
    f3 bmin, bmax, c;
    uint32_t ref, info, escape;
};

__device__ __forceinline__ Node load_node(const NodeRec *__restrict__ nodes,
                                          uint32_t i) {
    const float4 *p = reinterpret_cast<const float4 *>(nodes + i);
    float4 a = p[0], b = p[1], c = p[2];
    Node n;
    n.bmin = f3{a.x, a.y, a.z};
    n.bmax = f3{a.w, b.x, b.y};
    n.ref = __float_as_uint(b.z);
    n.info = __float_as_uint(b.w);
    n.c = f3{c.x, c.y, c.z};
    n.escape = __float_as_uint(c.w);
    return n;
}

// Voxel id of a dense-leaf hit: floor((res - min)/V) clamped to the leaf's
// dimensions (get_dimensions, geom.c:192-196).  The clamp is taken on the
// float, before the conversion, so that a hit point that is not finite
// (rays with infinite or NaN components) gives a clamped index too: NaN
// goes to 0 (maxNum), +-inf to the last or first voxel.
// The dimensions are precomputed at flattening and stored in the record's
// centre slot, which dense leaves do not use.
__device__ __forceinline__ uint32_t dense_voxel(const Node &n, f3 r,
                                                const Tree &t) {
    int dx = (int)__float_as_uint(n.c.x);
    int dy = (int)__float_as_uint(n.c.y);
    int dz = (int)__float_as_uint(n.c.z);
    auto index = [](float q, int dim) {
        return (int)fminf(fmaxf(floorf(q), 0.0f), (float)(dim - 1));
    };
    int ix = index(div_v(r.x - n.bmin.x, t.vx, t.ivx), dx);
    int iy = index(div_v(r.y - n.bmin.y, t.vy, t.ivy), dy);
    int iz = index(div_v(r.z - n.bmin.z, t.vz, t.ivz), dz);
    return n.ref + (uint32_t)((ix * dy + iy) * dz + iz);
}

__device__ __forceinline__ uint32_t child_of(uint32_t ref, uint32_t mask,
                                             int s) {
    return ((mask >> s) & 1u) ? ref + __popc(mask & ((1u << s) - 1u)) : kEnd;
}

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask,
                                                               0u));
}

// ---------------------------------------------------------------------------
// ray sources and sinks
// ---------------------------------------------------------------------------
struct SoaSrc {
    const float *ox, *oy, *oz, *dx, *dy, *dz;
    __device__ __forceinline__ void get(uint32_t r, f3 &o, f3 &d) const {
        o = f3{ox[r], oy[r], oz[r]};
        d = f3{dx[r], dy[r], dz[r]};
    }
};

struct SoaSink {
    uint32_t *voxel, *leaf;
    float *hx, *hy, *hz;
    __device__ __forceinline__ void put(uint32_t r, bool hit, uint32_t vid,
                                        uint32_t leafi, f3 p) const {
        if (voxel) voxel[r] = hit ? vid : kEnd;
        if (leaf) leaf[r] = hit ? leafi : kEnd;
        if (hit && hx) {
            hx[r] = p.x;
            hy[r] = p.y;
            hz[r] = p.z;
        }
    }
};

struct CamDev {
    float px, py, pz, qw, qi, qj, qk, mul, xsub, ysub, fov;
};

// simple-camera.c:21-31 + vect-ops-sse.c:16-26,58-64
__device__ __forceinline__ f3 screen2world(const CamDev &c, int sx, int sy) {
    float vx = __builtin_fmaf(c.mul, (float)sx, -c.xsub);
    float vy = 1.0f;
    float vz = __builtin_fmaf(-c.mul, (float)sy, c.ysub);
    float t1 = (c.qj * vz - vy * c.qk) * 2.0f;
    float t2 = (c.qk * vx - vz * c.qi) * 2.0f;
    float t3 = (c.qi * vy - vx * c.qj) * 2.0f;
    float c1 = c.qj * t3 - t2 * c.qk;
    float c2 = c.qk * t1 - t3 * c.qi;
    float c3 = c.qi * t2 - t1 * c.qj;
    return f3{__builtin_fmaf(c.qw, t1, vx) + c1,
              __builtin_fmaf(c.qw, t2, vy) + c2,
              __builtin_fmaf(c.qw, t3, vz) + c3};
}

__device__ __forceinline__ float dist2(const float *c, f3 p) {
    float x = c[0] - p.x, y = c[1] - p.y, z = c[2] - p.z;
    x = x * x;
    y = y * y;
    z = z * z;
    return (x + y) + z;
}

struct Shade {
    const uint8_t *tex;
    uint32_t rshift, gshift, bshift, amask;
    float vx, vy, vz;
    float ivx, ivy, ivz;
    // light manager (renderer.c:38-49), or lnodes null: none
    const LightNode *lnodes;
    const LightRec *lsph;
    uint32_t nlnodes;
    float amb[3];
    // vox_get_light (lights.c:103-116): the ambient light plus, for each
    // light sphere containing p in the order of the M-tree walk
    // (mtree.c:463-491: depth first, a one-sphere leaf reporting its sphere
    // on the leaf's test), (1 - min(r, |p - c|) / r) * colour, summed with
    // one fused multiply-add a channel (lights.c:91-101)
    __device__ __forceinline__ void light(f3 p, float L[3]) const {
        L[0] = amb[0];
        L[1] = amb[1];
        L[2] = amb[2];
        uint32_t n = 0;
        while (n < nlnodes) {
            const LightNode nd = lnodes[n];
            if (!(dist2(nd.c, p) < nd.sqr_radius)) {
                n = nd.escape;
                continue;
            }
            if (!(nd.info & 1u)) {  // inner: its first child is next
                n++;
                continue;
            }
            const uint32_t cnt = nd.info >> 8;
            for (uint32_t k = 0; k < cnt; k++) {
                const LightRec sp = lsph[nd.first + k];
                if (cnt == 1u || dist2(sp.c, p) < sp.sqr_radius) {
                    const float d = fminf(sp.radius, sqrtf(dist2(sp.c, p)));
                    const float a = 1.0f - d / sp.radius;
                    L[0] = __builtin_fmaf(a, sp.color[0], L[0]);
                    L[1] = __builtin_fmaf(a, sp.color[1], L[1]);
                    L[2] = __builtin_fmaf(a, sp.color[2], L[2]);
                }
            }
            n = nd.escape;
        }
    }
    // renderer.c:27-52: texture texel, scaled per channel by the light
    // (min(light, 1) * texel, truncated to 8 bits) when there is a manager
    __device__ __forceinline__ uint32_t color(f3 p) const {
        int x = abs((int)div_v(p.x, vx, ivx)) & 31;
        int y = abs((int)div_v(p.y, vy, ivy)) & 31;
        int z = abs((int)div_v(p.z, vz, ivz)) & 31;
        uint32_t c = tex[x * 1024 + y * 32 + z];
        if (!lnodes) return (c << rshift) | (c << gshift) | (c << bshift) | amask;
        float L[3];
        light(p, L);
        const float cf = (float)c;
        const uint32_t r = (uint32_t)(fminf(L[0], 1.0f) * cf);
        const uint32_t g = (uint32_t)(fminf(L[1], 1.0f) * cf);
        const uint32_t b = (uint32_t)(fminf(L[2], 1.0f) * cf);
        return (r << rshift) | (g << gshift) | (b << bshift) | amask;
    }
};

struct PixSink {
    uint32_t *squares, *voxel;
    float *hit;
    size_t plane;  // floats per hit plane
    Shade sh;
    __device__ __forceinline__ void put(uint32_t r, bool hit, uint32_t vid,
                                        uint32_t, f3 p) const {
        if (voxel) voxel[r] = hit ? vid : kEnd;
        if (hit) {
            if (this->hit) {
                this->hit[r] = p.x;
                this->hit[plane + r] = p.y;
                this->hit[2 * plane + r] = p.z;
            }
            if (squares) squares[r] = sh.color(p);
        }
    }
};

struct Ctrl {
    uint32_t rdone;    // workgroups of resolve_kernel that have finished
    uint32_t left;     // waves that have finished
    uint32_t pad0;
    uint32_t pad;
    unsigned long long visits, tests, rays, leaf_trav;
    unsigned long long fast_blocks, best_blocks, merge_blocks;
    unsigned long long spills; // stack entries moved from LDS to HBM
    // block queue built by block_cull_kernel, claimed with one 64-bit
    // atomic per workgroup: low word, blocks all of whose rays enter the
    // scene box (queued from the front); high word, the other entering
    // blocks (queued from the back)
    unsigned long long qpair;
    // FAST/ADAPTIVE renders with plain ray merging: blocks whose merging
    // stopped, so that the replay needs the odd pixels the first pass
    // skipped (low word; the high word stays 0), listed by resolve_kernel
    unsigned long long rpair;
};
// the counters cleared before a counting launch: visits .. spills
constexpr size_t kCounterBytes =
    offsetof(Ctrl, spills) + sizeof(unsigned long long) - offsetof(Ctrl, visits);

// ---------------------------------------------------------------------------
// work units driven by the persistent traversal kernel
//
//   begin(unit, st, cur, org, d)  -> starts the unit's first traversal
//   finish(st, hit, vid, node, e, cur, org, d) -> consumes a traversal's
//       result; returns true when it started another traversal
// ---------------------------------------------------------------------------

// A run of `chunk` consecutive rays per unit, traced one after the other
// by the same lane: a lane whose ray ends moves on at once instead of
// waiting for a wave-wide refill.
template <class Src, class Sink>
struct ChunkWork {
    // no occupancy floor asked of the register allocator (it fits 7 waves
    // per SIMD as it is; 8 would spill)
    static constexpr int kMinWavesPerSimd = 1;
    __host__ __device__ __forceinline__ bool keeps_queue() const { return false; }
    // XCD hand-out: one contiguous range of the units per XCD
    __device__ __forceinline__ uint32_t xcd_group() const { return 0; }
    __device__ __forceinline__ uint32_t unit_shift() const { return 0; }
    Src src;
    Sink sink;
    uint32_t n, chunk;
    __device__ __forceinline__ uint32_t prepare(uint32_t n) { return n; }
    struct State {
        uint32_t ray, end;
    };
    __device__ __forceinline__ void begin(uint32_t unit, State &st,
                                          uint32_t &cur, f3 &org, f3 &d) const {
        st.ray = unit * chunk;
        st.end = min(st.ray + chunk, n);
        src.get(st.ray, org, d);
        cur = 0;
    }
    __device__ __forceinline__ bool finish(State &st, bool hit, uint32_t vid,
                                           uint32_t node, f3 e, uint32_t &cur,
                                           f3 &org, f3 &d) const {
        sink.put(st.ray, hit, vid, node, e);
        if (++st.ray >= st.end) return false;
        src.get(st.ray, org, d);
        cur = 0;
        return true;
    }
    __device__ __forceinline__ void flush(const State &, Ctrl *) const {}
    __device__ __forceinline__ const Tree &tree(const State &,
                                                const Tree &t) const {
        return t;
    }
};

// One item of a multi-tree batch: a tree and a run of rays of the batch.
struct BatchItem {
    Tree t;
    uint32_t ray_begin, ray_count, unit_begin, pad;
};

// Many trees in one launch: units are ray chunks of all items; a lane
// finds its unit's item by binary search over the items' first units and
// traverses that item's tree.
template <class Src, class Sink>
struct MultiWork {
    // 6 waves per SIMD (at most 80 VGPRs) fit without spilling
    static constexpr int kMinWavesPerSimd = 6;
    __host__ __device__ __forceinline__ bool keeps_queue() const { return false; }
    __device__ __forceinline__ uint32_t xcd_group() const { return 0; }
    __device__ __forceinline__ uint32_t unit_shift() const { return 0; }
    Src src;
    Sink sink;
    const BatchItem *items;
    uint32_t n_items, chunk;
    __device__ __forceinline__ uint32_t prepare(uint32_t n) { return n; }
    struct State {
        uint32_t ray, end;
        Tree t;
    };
    __device__ __forceinline__ void begin(uint32_t unit, State &st,
                                          uint32_t &cur, f3 &org, f3 &d) const {
        uint32_t lo = 0, hi = n_items - 1;
        while (lo < hi) {
            uint32_t mid = (lo + hi + 1) >> 1;
            if (items[mid].unit_begin <= unit) lo = mid;
            else hi = mid - 1;
        }
        const BatchItem &it = items[lo];
        st.t = it.t;
        st.ray = it.ray_begin + (unit - it.unit_begin) * chunk;
        st.end = min(st.ray + chunk, it.ray_begin + it.ray_count);
        src.get(st.ray, org, d);
        cur = 0;
    }
    __device__ __forceinline__ bool finish(State &st, bool hit, uint32_t vid,
                                           uint32_t node, f3 e, uint32_t &cur,
                                           f3 &org, f3 &d) const {
        sink.put(st.ray, hit, vid, node, e);
        if (++st.ray >= st.end) return false;
        src.get(st.ray, org, d);
        cur = 0;
        return true;
    }
    __device__ __forceinline__ void flush(const State &, Ctrl *) const {}
    __device__ __forceinline__ const Tree &tree(const State &st,
                                                const Tree &) const {
        return st.t;
    }
};

// renderer.c:23-25
__device__ __forceinline__ int render_order(int i) {
    // 0, 1, 4, 5, 2, 3, 6, 7, 8, 9, 12, 13, 10, 11, 14, 15 packed by nibble
    const unsigned long long tab = 0xfebadc9876325410ull;
    return (int)((tab >> (4 * i)) & 0xfull);
}

__device__ __forceinline__ float sqr_metric(f3 a, f3 b) {
    float x = a.x - b.x, y = a.y - b.y, z = a.z - b.z;
    x = x * x;
    y = y * y;
    z = z * z;
    return (x + y) + z;
}

// search.c:31-81: the voxel of a leaf that the ray, entering the leaf's
// box at e, hits first: a dense leaf's voxel at e, or the voxel whose entry
// point is L1-closest to e (strictly, in stored order; an entry at e ends
// the search).  Writes the voxel id and the hit point; false: no voxel.
template <bool COUNT, bool PREFETCH = false>
__device__ __forceinline__ bool leaf_search(const Tree &T, const Node &nd,
                                            f3 e, f3 d, uint32_t &vid,
                                            f3 &bres,
                                            unsigned long long &nt) {
    bres = e;
    if ((nd.info & NT_TYPE_MASK) == NT_DENSE) {
        vid = dense_voxel(nd, e, T);
        return true;
    }
    const f3 V = f3{T.vx, T.vy, T.vz};
    const uint32_t cnt = nd.info >> 8;
    float best = INFINITY;
    bool hit = false;
    uint32_t k0 = 0;
    if constexpr (PREFETCH) {
        // the first 7 voxels' corners loaded together, not one round trip
        // per voxel
        float qx[7], qy[7], qz[7];
        const float *base = T.dots + 3u * nd.ref;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            qx[k] = qy[k] = qz[k] = 0.0f;
            if ((uint32_t)k < cnt) {
                qx[k] = base[3 * k];
                qy[k] = base[3 * k + 1];
                qz[k] = base[3 * k + 2];
            }
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            if ((uint32_t)k < cnt) {
                f3 vmin = f3{qx[k], qy[k], qz[k]};
                f3 vmax = f3{vmin.x + V.x, vmin.y + V.y, vmin.z + V.z};
                f3 far;
                if (COUNT) nt++;
                if (hit_box(vmin, vmax, e, d, far)) {
                    float m = abs_metric(e, far);
                    if (m == 0.0f) {
                        bres = far;
                        vid = nd.ref + (uint32_t)k;
                        return true;
                    }
                    if (m < best) {
                        best = m;
                        bres = far;
                        vid = nd.ref + (uint32_t)k;
                        hit = true;
                    }
                }
            }
        }
        k0 = 7;
    }
    for (uint32_t k = k0; k < cnt; k++) {
        const float *p = T.dots + 3u * (nd.ref + k);
        f3 vmin = f3{p[0], p[1], p[2]};
        f3 vmax = f3{vmin.x + V.x, vmin.y + V.y, vmin.z + V.z};
        f3 far;
        if (COUNT) nt++;
        if (hit_box(vmin, vmax, e, d, far)) {
            float m = abs_metric(e, far);
            if (m == 0.0f) {
                bres = far;
                vid = nd.ref + k;
                return true;
            }
            if (m < best) {
                best = m;
                bres = far;
                vid = nd.ref + k;
                hit = true;
            }
        }
    }
    return hit;
}

// The queue counts of the render pre-pass (Ctrl::qpair): entries from the
// front and in all, clamped to the queue's length, so that counts left by
// a render whose later launches failed cannot send a reader past the end.
__device__ __forceinline__ void queue_counts(unsigned long long pr,
                                             uint32_t qlen, uint32_t &qf,
                                             uint32_t &qt) {
    const uint32_t lo = (uint32_t)pr, hi = (uint32_t)(pr >> 32);
    qf = min(lo, qlen);
    qt = min(qf + min(hi, qlen), qlen);
}

// A render job: one tree, a run of camera poses of one size and quality,
// its output buffers.  Units are 4x4 blocks, `bpf` per frame.
struct RenderItem {
    Tree t;
    const CamDev *cams;
    uint32_t W, unit_begin;
    FastDiv ws, bpf;  // blocks per block row, blocks per frame
    int rnd_mode, merge_mode;
    PixSink sink;
};

// One 4x4 block per unit (renderer.c:214-318): in FAST / ADAPTIVE mode the
// leaf found for one pixel is tried first for the next, so the pixels of a
// block are traced in order by one lane; in BEST mode every pixel is a
// root search (a single BEST render uses BestWork instead).  MULTI: units
// of several render jobs (different trees, cameras, sizes, modes) in one
// launch.
template <bool MULTI>
struct BlockWork {
    // 6 waves per SIMD (at most 80 VGPRs) fit without spilling for one
    // render; the multi-job form would spill at that floor
    static constexpr int kMinWavesPerSimd = MULTI ? 1 : 6;
    __host__ __device__ __forceinline__ bool keeps_queue() const { return false; }
    // one render's block queue is ordered longest-first (block_cull_kernel)
    RenderItem one;             // the job when !MULTI
    const RenderItem *items;    // the jobs when MULTI, by unit_begin
    uint32_t n_items;
    const uint32_t *queue;      // !MULTI: units are queue entries if set:
    const unsigned long long *qpair;  // the first (low word) units from
    uint32_t qlen;              // the front, the rest from the back of qlen
    uint32_t qf;                // entries; qf: the low word, from prepare()
    uint32_t group;             // !MULTI: units dealt to the XCDs in turns
                                // of `group` (the queue is ordered
                                // longest-first, so every XCD should follow
                                // that order); MULTI: contiguous ranges
    __device__ __forceinline__ uint32_t xcd_group() const {
        if constexpr (MULTI) return 0;
        else return group;
    }
    __device__ __forceinline__ uint32_t unit_shift() const { return 0; }
    uint32_t units_per_block_host() const { return 1u; }
    // reads once per thread what the block pre-pass counted on the device;
    // returns the number of units
    __device__ __forceinline__ uint32_t prepare(uint32_t n) {
        if constexpr (!MULTI) {
            if (queue) {
                uint32_t qt;
                queue_counts(*qpair, qlen, qf, qt);
                return qt;
            }
        }
        return n;
    }
    enum { CORNER1 = 0, CORNER2 = 1, LOOP = 2 };
    struct State {
        uint32_t item, frame, base;  // job, frame in the job, output index
        // packed in one word: the pixel position i in the block's order and
        // its end, the stage, the block's mode, ray merging, a leaf try
        uint32_t i : 5, iend : 5, stage : 2, mode : 2, merge : 1,
            leaf_try : 1;
        uint32_t leaf, vid_last;
        f3 inter_last;  // at the second corner: the first corner's hit
        uint32_t xs : 16, ys : 16;  // block corner (W, H < 65536)
        uint32_t fast, best, merged;
    };

    __device__ __forceinline__ const RenderItem &item(const State &st) const {
        if constexpr (MULTI) return items[st.item];
        else return one;
    }
    __device__ __forceinline__ const CamDev &cam(const State &st) const {
        const RenderItem &it = item(st);
        return it.cams[st.frame];
    }
    __device__ __forceinline__ void emit(const State &st, int p, bool hit,
                                         uint32_t vid, f3 e) const {
        item(st).sink.put(st.base + (uint32_t)p, hit, vid, 0, e);
    }
    __device__ __forceinline__ void trace_pixel(const State &st, int p,
                                                uint32_t start, uint32_t &cur,
                                                f3 &org, f3 &d) const {
        const CamDev &c = cam(st);
        d = screen2world(c, st.xs + (p & 3), st.ys + (p >> 2));
        org = f3{c.px, c.py, c.pz};
        cur = start;
    }
    // renderer.c:286-316: advance to the next pixel that needs a traversal,
    // writing merged pixels on the way; false when the block is done.
    __device__ __forceinline__ bool next(State &st, uint32_t &cur, f3 &org,
                                         f3 &d) const {
        while (st.i < st.iend) {
            int p = render_order(st.i);
            bool m = st.merge && (st.i & 1);
            if (m) {
                emit(st, p, st.leaf != kEnd, st.vid_last, st.inter_last);
                st.i++;
                continue;
            }
            st.leaf_try = (st.mode == VOXG_QUALITY_FAST && st.leaf != kEnd);
            trace_pixel(st, p, st.leaf_try ? st.leaf : 0u, cur, org, d);
            return true;
        }
        st.fast += (st.mode == VOXG_QUALITY_FAST) ? 1u : 0u;
        st.best += (st.mode == VOXG_QUALITY_FAST) ? 0u : 1u;
        st.merged += st.merge ? 1u : 0u;
        return false;
    }
    __device__ __forceinline__ void begin(uint32_t unit, State &st,
                                          uint32_t &cur, f3 &org, f3 &d) const {
        st.item = 0;
        if constexpr (!MULTI) {
            if (queue) unit = queue[unit < qf ? unit : qlen - 1u - (unit - qf)];
        }
        if constexpr (MULTI) {
            uint32_t lo = 0, hi = n_items - 1;
            while (lo < hi) {
                uint32_t mid = (lo + hi + 1) >> 1;
                if (items[mid].unit_begin <= unit) lo = mid;
                else hi = mid - 1;
            }
            st.item = lo;
        }
        const RenderItem &it = item(st);
        uint32_t u = unit - it.unit_begin;
        // frames are stored one after the other in the square layout
        st.frame = it.bpf.div(u);
        const uint32_t cs = u - st.frame * it.bpf.d;
        const uint32_t row = it.ws.div(cs);
        st.base = u * 16u;
        st.xs = (int)(cs - row * it.ws.d) * 4;
        st.ys = (int)row * 4;
        st.mode = it.rnd_mode;
        st.merge = 0;
        st.leaf = kEnd;
        st.vid_last = kEnd;
        st.inter_last = f3{0, 0, 0};
        st.i = 0;
        st.iend = 16;
        if (st.mode == VOXG_QUALITY_ADAPTIVE) {
            // renderer.c:240-253
            st.merge = (it.merge_mode == VOXG_QUALITY_RAY_MERGE);
            st.i = 1;
            st.stage = CORNER1;
            st.leaf_try = 0;
            trace_pixel(st, 0, 0u, cur, org, d);
            return;
        }
        st.stage = LOOP;
        next(st, cur, org, d);  // a block always has a first pixel to trace
    }
    __device__ __forceinline__ bool finish(State &st, bool hit, uint32_t vid,
                                           uint32_t node, f3 e, uint32_t &cur,
                                           f3 &org, f3 &d) const {
        if (st.stage == CORNER1) {
            // renderer.c:252-259
            st.mode = VOXG_QUALITY_FAST;
            st.leaf = hit ? node : kEnd;
            emit(st, 0, hit, vid, e);
            if (hit) {
                st.vid_last = vid;
                st.inter_last = e;
                st.iend = 15;
                st.stage = CORNER2;
                trace_pixel(st, 15, 0u, cur, org, d);
                return true;
            }
            st.stage = LOOP;
            return next(st, cur, org, d);
        }
        if (st.stage == CORNER2) {
            // renderer.c:260-279
            emit(st, 15, hit, vid, e);
            if (hit) {
                // the first corner's hit point and direction: inter_last
                // has not changed since, and the direction is recomputed
                const CamDev &c = cam(st);
                f3 origin = f3{c.px, c.py, c.pz};
                float d1 = sqr_metric(st.inter_last, origin);
                f3 dir1 = screen2world(c, st.xs, st.ys);
                float d2 = (dir1.x * dir1.x + dir1.y * dir1.y) + dir1.z * dir1.z;
                float criteria = d1 / d2 * sqr_metric(dir1, d);
                float dist = sqr_metric(st.inter_last, e);
                if ((double)(dist / criteria) > 1.69) st.mode = VOXG_QUALITY_BEST;
                if (d1 > 22500.0f) {
                    st.merge = (st.mode == VOXG_QUALITY_FAST &&
                                item(st).merge_mode ==
                                    VOXG_QUALITY_RAY_MERGE_ACCURATE)
                                   ? 1
                                   : st.merge;
                } else {
                    st.merge = 0;
                }
            }
            st.stage = LOOP;
            return next(st, cur, org, d);
        }
        // renderer.c:293-315
        if (!hit && st.leaf_try) {
            st.leaf_try = 0;
            cur = 0;  // the leaf missed: search from the root
            const CamDev &c = cam(st);
            org = f3{c.px, c.py, c.pz};
            return true;
        }
        st.leaf = hit ? node : kEnd;
        if (hit) {
            st.vid_last = vid;
            st.inter_last = e;
        }
        emit(st, render_order(st.i), hit, vid, e);
        st.i++;
        return next(st, cur, org, d);
    }
    __device__ __forceinline__ const Tree &tree(const State &st,
                                                const Tree &t) const {
        if constexpr (MULTI) return items[st.item].t;
        else return t;
    }
    __device__ __forceinline__ void flush(const State &st, Ctrl *ctrl) const {
        atomicAdd(&ctrl->fast_blocks, (unsigned long long)st.fast);
        atomicAdd(&ctrl->best_blocks, (unsigned long long)st.best);
        atomicAdd(&ctrl->merge_blocks, (unsigned long long)st.merged);
    }
};

// Root-search results of the pixels of the queued blocks: the first pass
// of FAST and ADAPTIVE renders writes them, resolve_kernel reads them.
// Stored pixel-major, index p * stride + block (stride: the blocks of the
// render), so that the resolve threads, one per block, read consecutive
// words when they all take their blocks' pixel p.
struct SpecOut {
    uint32_t *vid, *node;  // voxel id or kEnd; leaf node of a hit
    float *hx, *hy, *hz;   // hit point
    uint32_t stride;
    // direct renders only: hit points of corners 0 and 15 (index
    // corner * stride + block); the replay's changed pixels, appended per
    // wave of the replay into the region of its 64 queue positions
    // (1024 entries from 16 x the region's first position; vid and hx..hz
    // hold the results, ekey the position within the region | pixel << 6)
    // and the number of entries of each region
    float *cx, *cy, *cz;
    uint32_t *ekey, *ecount;
    __device__ __forceinline__ void put(uint32_t blk, int px, bool hit,
                                        uint32_t v, uint32_t nd, f3 p) const {
        const uint32_t r = (uint32_t)px * stride + blk;
        vid[r] = hit ? v : kEnd;
        if (hit) {
            node[r] = nd;
            hx[r] = p.x;
            hy[r] = p.y;
            hz[r] = p.z;
        }
    }
    // direct FAST/ADAPTIVE renders: the leaf node of every pixel (kEnd on
    // a miss), a block's 16 together (index block * 16 + pixel: the
    // neighbouring blocks that a wave finishes together write whole lines),
    // and the hit points of corners 0 and 15 for ADAPTIVE's decision; the
    // results themselves go to the outputs
    __device__ __forceinline__ void put_node(uint32_t blk, int px, bool hit,
                                             uint32_t nd, f3 p) const {
        node[blk * 16u + (uint32_t)px] = hit ? nd : kEnd;
        if (hit && (px == 0 || px == 15)) {
            const uint32_t c = (px == 15 ? stride : 0u) + blk;
            cx[c] = p.x;
            cy[c] = p.y;
            cz[c] = p.z;
        }
    }
};

// BEST-mode render of one job: every pixel is a root search
// (renderer.c:300-308), so a unit is 16 >> split pixels of a 4x4 block,
// down to one pixel, taken in the block's order (renderer.c:23-25).  With
// one pixel per unit, the lanes that one refill starts trace neighbouring
// pixels together: coherent rays, the same nodes at the same time.  Keeps
// only the state BEST mode needs, which leaves registers for more waves
// than BlockWork.  MODE 0: results to the job's outputs (BEST renders);
// 1: root searches into SpecOut (FAST and ADAPTIVE renders with accurate
// ray merging); 2: the same for one pixel per unit from a set of a block's
// pixels (pset below), the passes of ADAPTIVE renders with plain ray
// merging; 3: results to the outputs and leaf nodes into SpecOut (FAST
// and ADAPTIVE renders without merging).  Separate instantiations, so
// that no unit pays for another mode's branches (a shared pixel-set branch
// cost the BEST render 1%).
template <int MODE>
struct BestWorkT {
    static constexpr bool SETS = MODE == 2;
    // 7 waves per SIMD (at most 72 VGPRs): the compiler fits this kernel
    // without spilling, and the extra wave hides node-load latency
    static constexpr int kMinWavesPerSimd = 7;
    RenderItem one;
    SpecOut spec;            // MODE 1, 2: the root searches, for
                             // resolve_kernel, instead of the job's outputs
    uint32_t keep_queue;     // the queue counts are still needed afterwards
    __host__ __device__ __forceinline__ bool keeps_queue() const {
        return keep_queue != 0;
    }
    const uint32_t *queue;  // block queue of the pre-pass, or null: the
    const unsigned long long *qpair;  // first (low word) entries from the
    uint32_t qlen, qf;      // front, the rest from the back of qlen entries
    uint32_t split;         // a unit is 16 >> split pixels
    uint32_t group;         // units dealt to the XCDs in turns of `group`
    // SETS only: one pixel per unit of 1: pixels 0, 2, ..., 14 and 15 of
    // the block's order (9 units per block) or 2: pixels 1, 3, ..., 13 (7
    // units); pdiv divides by the units per block
    uint32_t pset;
    FastDiv pdiv;
    __device__ __forceinline__ uint32_t xcd_group() const { return group; }
    __device__ __forceinline__ uint32_t unit_shift() const { return split; }
    uint32_t units_per_block_host() const {
        if constexpr (SETS) return pdiv.d;
        else return 1u << split;
    }
    // reads once per thread what the block pre-pass counted on the device;
    // returns the number of units
    __device__ __forceinline__ uint32_t prepare(uint32_t n) {
        if (!queue) return n;
        uint32_t qt;
        queue_counts(*qpair, qlen, qf, qt);
        if constexpr (SETS) return qt * pdiv.d;
        else return qt << split;
    }
    struct State {
        uint32_t frame, base;  // camera, output index of the block
        uint32_t xy;           // block corner: x | y << 16
        uint32_t pos;          // pixel i | end << 8 | own << 16, where
                               // pixels i..end-1 of the block's order are
                               // this unit's, and own: it counts the block
        uint32_t best;         // blocks counted (counting launch)
    };
    __device__ __forceinline__ void trace(const State &st, uint32_t &cur,
                                          f3 &org, f3 &d) const {
        const int p = render_order((int)(st.pos & 0xffu));
        const CamDev &c = one.cams[st.frame];
        d = screen2world(c, (int)(st.xy & 0xffffu) + (p & 3),
                         (int)(st.xy >> 16) + (p >> 2));
        org = f3{c.px, c.py, c.pz};
        cur = 0;
    }
    __device__ __forceinline__ void begin(uint32_t unit, State &st,
                                          uint32_t &cur, f3 &org, f3 &d) const {
        uint32_t b, i, len, own;
        if constexpr (SETS) {
            b = pdiv.div(unit);
            const uint32_t k = unit - b * pdiv.d;
            i = pset == 1u ? (k < 8u ? 2u * k : 15u) : 2u * k + 1u;
            len = 1u;
            own = k == 0u;
        } else {
            const uint32_t part = unit & ((1u << split) - 1u);
            b = unit >> split;
            len = 16u >> split;
            i = part * len;
            own = part == 0u;
        }
        if (queue) b = queue[b < qf ? b : qlen - 1u - (b - qf)];
        // frames are stored one after the other in the square layout
        st.frame = one.bpf.div(b);
        const uint32_t cs = b - st.frame * one.bpf.d;
        const uint32_t row = one.ws.div(cs);
        st.base = b * 16u;
        st.xy = ((cs - row * one.ws.d) * 4u) | ((row * 4u) << 16);
        st.pos = i | ((i + len) << 8) | (own << 16);
        trace(st, cur, org, d);
    }
    __device__ __forceinline__ bool finish(State &st, bool hit, uint32_t vid,
                                           uint32_t node, f3 e, uint32_t &cur,
                                           f3 &org, f3 &d) const {
        const int p = render_order((int)(st.pos & 0xffu));
        if constexpr (MODE == 1 || MODE == 2)
            spec.put(st.base >> 4, p, hit, vid, node, e);
        else
            one.sink.put(st.base + (uint32_t)p, hit, vid, 0, e);
        if constexpr (MODE == 3) spec.put_node(st.base >> 4, p, hit, node, e);
        st.pos++;
        if ((st.pos & 0xffu) < ((st.pos >> 8) & 0xffu)) {
            trace(st, cur, org, d);
            return true;
        }
        st.best += st.pos >> 16;
        return false;
    }
    __device__ __forceinline__ const Tree &tree(const State &,
                                                const Tree &t) const {
        return t;
    }
    __device__ __forceinline__ void flush(const State &st, Ctrl *ctrl) const {
        atomicAdd(&ctrl->best_blocks, (unsigned long long)st.best);
    }
};
using BestWork = BestWorkT<0>;
using BestSpecWork = BestWorkT<1>;
using BestSetWork = BestWorkT<2>;
using BestDirectWork = BestWorkT<3>;

// ---------------------------------------------------------------------------
// persistent traversal kernel
// ---------------------------------------------------------------------------
constexpr uint32_t kXcds = 8;        // XCDs of an MI355X
constexpr uint32_t kXcdStride = 64;  // counters 256 bytes apart
template <int K, class Work, bool COUNT>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(COUNT ? 1 : Work::kMinWavesPerSimd, 8)))
rti_kernel(Tree t, Work work_arg, uint32_t n, Ctrl *ctrl, uint32_t *ovf,
           uint32_t *xnext, uint4 *__restrict__ spill,
           uint32_t spill_slots, uint32_t refill, uint32_t fin_batch,
           uint32_t leaf_batch) {
    // stack ring: entries [K][blockDim.x], then their pending lists
    extern __shared__ uint4 ring[];
    uint32_t *const ring_st = reinterpret_cast<uint32_t *>(ring + K * blockDim.x);
    Work work = work_arg;
    // the unit count may come from the device (the render's block queue)
    n = work.prepare(n);
    // one contiguous range per XCD, or groups of xcd_group() units dealt
    // to the XCDs in turn when the order of the units matters
    const uint32_t xg = work.xcd_group();
    uint32_t xch = xg ? xg : (n + kXcds - 1) / kXcds;
    if (xch == 0) xch = 1;
    const uint32_t tid = threadIdx.x;
    const uint32_t bdim = blockDim.x;
    const uint32_t gtid = blockIdx.x * bdim + tid;
    const uint32_t gthreads = gridDim.x * bdim;
    const uint32_t lane = lane_id();
    // spill buffer: entries [spill_slots][gthreads], then their lists
    uint32_t *const spill_st =
        reinterpret_cast<uint32_t *>(spill + (size_t)spill_slots * gthreads);

    // XCD-aware hand-out: the units are cut into chunks of xch units and
    // chunk j goes to range j % kXcds, each range with a counter of its own
    // (own 256-byte line).  A wave draws from the range of the XCD it runs
    // on (workgroups go round-robin over the XCDs), so neighbouring units
    // share that XCD's L2 and no counter takes the atomics of the whole
    // chip, and moves on to the next range when its own is used up.  With
    // xch = n / kXcds each range is one contiguous eighth; a work type whose
    // units are ordered longest-first deals small chunks instead, so every
    // range follows that order.
    uint32_t xr = blockIdx.x % kXcds, xtries = 0;
    bool active = false;
    bool exhausted = false;
    uint32_t cur = 0;
    f3 d{0, 0, 0}, org{0, 0, 0};
    // lane phase: traversing, waiting at a leaf whose box it entered (entry
    // point in org), or ended with a result
    // in (p_hit, p_vid, cur, org) that waits to be handed to the work unit
    enum { PH_TRAV = 0, PH_LEAF = 1, PH_DONE = 2 };
    int ph = PH_TRAV;
    bool p_hit = false;
    uint32_t p_vid = kEnd;
    typename Work::State st;
    std::memset(&st, 0, sizeof(st));
    int sp = 0, lo = 0;
    unsigned long long nv = 0, nt = 0, nr = 0, nl = 0, ns = 0;

    for (;;) {
        uint64_t idle = __ballot(!active);
        if (!exhausted && idle != 0 &&
            (__popcll(idle) >= refill || idle == ~0ull)) {
            uint32_t cnt = (uint32_t)__popcll(idle);
            uint32_t leader = (uint32_t)(__ffsll((long long)idle) - 1);
            // range x owns chunks x, x + kXcds, ... of xch units
            const uint32_t x = xr;
            const uint32_t nch = (n + xch - 1) / xch;
            const uint32_t r_len =
                x < nch ? ((nch - x + kXcds - 1) / kXcds) * xch : 0u;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&xnext[x * kXcdStride], cnt);
            base = __shfl(base, (int)leader);
            base = __builtin_amdgcn_readfirstlane(base);
            if (base + cnt >= r_len) {  // range x is used up: go on to the next
                xr = (x + 1) % kXcds;
                if (++xtries == kXcds) exhausted = true;
            }
            if (!active) {
                const uint32_t my = base + lanes_below(idle);
                const uint32_t u = ((my / xch) * kXcds + x) * xch + my % xch;
                if (my < r_len && u < n) {
                    work.begin(u, st, cur, org, d);
                    sp = lo = 0;
                    ph = PH_TRAV;
                    active = true;
                    if (COUNT) {
                        if (cur == 0) nr++;
                        else nl++;
                    }
                }
            }
        }
        if (__ballot(active) == 0) {
            if (exhausted) break;
            continue;
        }
        if (!active) continue;

        const Tree &T = work.tree(st, t);
        bool popme = false;

        // phase 1: one node of every traversing lane (search.c:8-29, 84-148)
        if (ph == PH_TRAV) {
            f3 e = org;
            Node nd;
            bool in = false;
            if (!T.empty) {
                nd = load_node(T.nodes, cur);
                if (COUNT) nv++;
                in = hit_box(nd.bmin, nd.bmax, org, d, e);
            }
            if (!in) {
                popme = true;
            } else if ((nd.info & NT_TYPE_MASK) != NT_INNER) {
                // a leaf: wait at it with the entry point in org
                ph = PH_LEAF;
                org = e;
            } else {
                // search.c:84-148
                int s = corrected_subspace(nd.c, e, d);
                f3 P0, P1, P2;
                bool h0 = hit_plane(e, d, nd.c, 0, nd.bmin, nd.bmax, P0);
                bool h1 = hit_plane(e, d, nd.c, 1, nd.bmin, nd.bmax, P1);
                bool h2 = hit_plane(e, d, nd.c, 2, nd.bmin, nd.bmax, P2);
                // compact the hits in plane order (:108-113)
                int cntp = 0, i0 = 0, i1 = 0, i2 = 0;
                f3 Q0 = P0, Q1 = P1, Q2 = P2;
                if (h0) {
                    i0 = 0;
                    Q0 = P0;
                    cntp = 1;
                }
                if (h1) {
                    if (cntp == 0) {
                        i0 = 1;
                        Q0 = P1;
                    } else {
                        i1 = 1;
                        Q1 = P1;
                    }
                    cntp++;
                }
                if (h2) {
                    if (cntp == 0) {
                        i0 = 2;
                        Q0 = P2;
                    } else if (cntp == 1) {
                        i1 = 2;
                        Q1 = P2;
                    } else {
                        i2 = 2;
                        Q2 = P2;
                    }
                    cntp++;
                }
                // selection sort by L1 distance from the entry (:115-134)
                float m0 = abs_metric(e, Q0), m1 = abs_metric(e, Q1),
                      m2 = abs_metric(e, Q2);
                if (cntp >= 2 && m1 < m0) {
                    f3 tq = Q0; Q0 = Q1; Q1 = tq;
                    int ti = i0; i0 = i1; i1 = ti;
                    float tv = m0; m0 = m1; m1 = tv;
                }
                if (cntp >= 3 && m2 < m0) {
                    f3 tq = Q0; Q0 = Q2; Q2 = tq;
                    int ti = i0; i0 = i2; i2 = ti;
                    float tv = m0; m0 = m2; m2 = tv;
                }
                if (cntp >= 3 && m2 < m1) {
                    f3 tq = Q1; Q1 = Q2; Q2 = tq;
                    int ti = i1; i1 = i2; i2 = ti;
                }
                uint32_t mask = (nd.info >> 8) & 0xffu;
                int s1 = s ^ (1 << i0);
                int s2 = s1 ^ (1 << i1);
                int s3 = s2 ^ (1 << i2);
                uint32_t cs0 = child_of(nd.ref, mask, s);
                // the children behind the planes that exist, in visiting
                // order (:137-147), as (plane, child offset) pairs of a
                // pending list: count in bits 0-1, pair j in bits 2+5j..
                uint32_t lst = 0u, np = 0u;
                auto pend = [&](int sub, int plane) {
                    if ((mask >> sub) & 1u) {
                        const uint32_t off =
                            (uint32_t)__popc(mask & ((1u << sub) - 1u));
                        lst |= ((uint32_t)plane | (off << 2)) << (2u + 5u * np);
                        np++;
                    }
                };
                if (cntp >= 1) pend(s1, i0);
                if (cntp >= 2) pend(s2, i1);
                if (cntp >= 3) pend(s3, i2);
                // one stack entry for the node: its index and entry point,
                // the plane points are recomputed when the entry is popped
                if (np) {
                    if (sp - lo == K) {
                        // ring full: its oldest entry goes to the spill
                        // buffer (or is dropped past the bound: reported)
                        if ((uint32_t)lo >= spill_slots) {
                            *(volatile uint32_t *)ovf = 1u;
                        } else {
                            const size_t si = (size_t)lo * gthreads + gtid;
                            const uint32_t ri = (lo & (K - 1)) * bdim + tid;
                            spill[si] = ring[ri];
                            spill_st[si] = ring_st[ri];
                            if (COUNT) ns++;
                        }
                        lo++;
                    }
                    const uint32_t ri = (sp & (K - 1)) * bdim + tid;
                    ring[ri] = make_uint4(cur, __float_as_uint(e.x),
                                          __float_as_uint(e.y),
                                          __float_as_uint(e.z));
                    ring_st[ri] = lst | np;
                    sp++;
                }
                if (cs0 != kEnd) {
                    cur = cs0;
                    org = e;
                } else {
                    popme = true;
                }
            }
        }
        const uint64_t tm = __ballot(ph == PH_TRAV);
        const uint64_t lm = __ballot(ph == PH_LEAF);

        // phase 2: leaves (search.c:31-81), once enough lanes wait at one
        // or no lane is traversing
        if (ph == PH_LEAF && ((uint32_t)__popcll(lm) >= leaf_batch || tm == 0)) {
            const Node nd = load_node(T.nodes, cur);
            uint32_t vid = kEnd;
            f3 bres = org;
            const bool hit = leaf_search<COUNT>(T, nd, org, d, vid, bres, nt);
            if (hit) {
                ph = PH_DONE;
                p_hit = true;
                p_vid = vid;
                org = bres;
            } else {
                ph = PH_TRAV;
                popme = true;
            }
        }

        // phase 3: continue with the next pending child of the nearest
        // node that has one: the child behind the entry's next plane, with
        // the plane point as origin, computed from the node's entry point as
        // hit_plane computes it (geom-sse.c:118-142)
        if (popme) {
            cur = kEnd;
            bool have = sp > 0;
            if (have && sp == lo) {
                // the ring is empty: the top entry comes back from the
                // spill buffer (unless it was dropped past the bound)
                lo--;
                if ((uint32_t)lo < spill_slots) {
                    const size_t si = (size_t)lo * gthreads + gtid;
                    const uint32_t ri = (lo & (K - 1)) * bdim + tid;
                    ring[ri] = spill[si];
                    ring_st[ri] = spill_st[si];
                } else {
                    have = false;
                }
            }
            if (have) {
                const uint32_t ri = ((sp - 1) & (K - 1)) * bdim + tid;
                const uint4 v = ring[ri];
                const uint32_t lst = ring_st[ri];
                const uint32_t np = lst & 3u, pair = (lst >> 2) & 31u;
                if (np == 1u) sp--;
                else ring_st[ri] = (np - 1u) | ((lst >> 5) & ~3u);
                const int p = (int)(pair & 3u);
                const f3 e = f3{__uint_as_float(v.y), __uint_as_float(v.z),
                                __uint_as_float(v.w)};
                const uint32_t *rec =
                    reinterpret_cast<const uint32_t *>(T.nodes + v.x);
                const uint32_t ref = rec[offsetof(NodeRec, ref) / 4];
                const float cp =
                    __uint_as_float(rec[offsetof(NodeRec, c) / 4 + p]);
                const float k = (cp - comp(e, p)) / comp(d, p);
                org = f3{__builtin_fmaf(k, d.x, e.x),
                         __builtin_fmaf(k, d.y, e.y),
                         __builtin_fmaf(k, d.z, e.z)};
                cur = ref + (pair >> 2);
            }
            if (cur == kEnd) {
                ph = PH_DONE;
                p_hit = false;
                p_vid = kEnd;
            }
        }

        // phase 4: hand ended traversals to their work units, in batches
        const uint64_t dm = __ballot(ph == PH_DONE);
        const uint64_t am = __ballot(true);
        if (ph == PH_DONE && ((uint32_t)__popcll(dm) >= fin_batch || dm == am)) {
            ph = PH_TRAV;
            sp = lo = 0;
            if (work.finish(st, p_hit, p_vid, cur, org, cur, org, d)) {
                if (COUNT) {
                    if (cur == 0) nr++;
                    else nl++;
                }
            } else {
                active = false;
            }
        }
    }

    if (COUNT) {
        atomicAdd(&ctrl->visits, nv);
        atomicAdd(&ctrl->tests, nt);
        atomicAdd(&ctrl->rays, nr);
        atomicAdd(&ctrl->leaf_trav, nl);
        atomicAdd(&ctrl->spills, ns);
        work.flush(st, ctrl);
    }
    // self-reset of the unit counter by the last wave to leave
    __threadfence();
    if (lane == 0) {
        uint32_t total = gridDim.x * (bdim >> 6);
        uint32_t done = atomicAdd(&ctrl->left, 1u);
        if (done == total - 1) {
            for (uint32_t x = 0; x < kXcds; x++)
                atomicExch(&xnext[x * kXcdStride], 0u);
            if (!work.keeps_queue()) atomicExch(&ctrl->qpair, 0ull);
            atomicExch(&ctrl->left, 0u);
        }
    }
}

// Pre-pass of a render launch: one thread per 4x4 block tests the box that
// its four corner rays bound for all 16 directions against the root box,
// with the conservative block_may_enter (cull.h: false only when the
// traversal's exact hit_box rejects every ray of the block; testing each
// of the 16 rays instead made the pass 0.31 ms on the benchmark frames,
// bound by that per-pixel arithmetic).  A block that may enter the root box
// is then tested against the boxes of the tree's cut (cut_of) with the
// wider margin of deep_origin (cull.h), except in counting launches.  A
// block that no ray can enter, or that no ray can take near a box of the
// cut, is finished here -- every pixel is a miss in every quality mode --
// and the others go, in block order, to the queue the traversal kernel
// takes its blocks from.  A workgroup covers kCullTile pixels (1024
// blocks): it marks
// the entering blocks in LDS, compacts them with a workgroup scan and
// reserves its queue ranges with a single 64-bit atomic (the queue
// counters share one address, so their atomics are serialised: one per
// workgroup of 1024 blocks keeps them off the critical path).  In counting
// mode it adds what the loop would have counted for each finished block:
// one root search per traced pixel (ADAPTIVE with ray merging traces 8 of
// the 16 pixels of a block whose first corner misses), one node visit per
// search, and the block's final mode.
constexpr uint32_t kCullBlocks = 1024;
constexpr uint32_t kCullTile = kCullBlocks * 16;

template <bool COUNT>
__global__ void __launch_bounds__(256)
block_cull_kernel(RenderItem it, uint32_t npix, Ctrl *ctrl,
                  uint32_t *__restrict__ queue, uint32_t qlen) {
    __shared__ uint8_t enter[kCullBlocks];
    __shared__ uint16_t dlist[kCullBlocks];
    __shared__ float ddir[6][kCullBlocks];
    __shared__ float cutab[2][kCutMax][6];
    __shared__ uint32_t wave_sum[4];
    __shared__ uint32_t gbase[2];
    __shared__ uint32_t ndeep;
    __shared__ unsigned long long missed;
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    if (tid == 0) {
        missed = 0;
        ndeep = 0;
    }
    __syncthreads();
    const uint32_t tile0 = blockIdx.x * kCullTile;
    uint32_t my_missed = 0;
    // the host launches this kernel only for a tree with a root record
    const float4 *q = reinterpret_cast<const float4 *>(it.t.nodes);
    const float4 a = q[0], b = q[1];
    const float lo3[3] = {a.x, a.y, a.z}, hi3[3] = {a.w, b.x, b.y};
    // the camera of the tile's first pixel (tile0 < npix, so f0 is a frame
    // of the job); a block of another frame recomputes
    const uint32_t f0 = it.bpf.div(tile0 >> 4);
    CullOrigin co0;
    {
        const CamDev &c = it.cams[f0];
        const float o3[3] = {c.px, c.py, c.pz};
        co0 = cull_origin(lo3, hi3, o3);
    }
    // the miss ids of the finished blocks (64 contiguous bytes each) are
    // written after the tests, by the whole workgroup; a voxel plane that
    // is not 16-byte aligned takes them one by one from the testing thread
    uint32_t *const vox = it.sink.voxel;
    const bool vox16 = ((uintptr_t)vox & 15u) == 0;
    // a block's camera and the box of its 16 directions (block_dirs,
    // cull.h: the four corner rays bound the others)
    auto block_geom = [&](uint32_t blk, const CamDev *&cam, float (&dc)[4][3]) {
        const uint32_t f = it.bpf.div(blk);
        const uint32_t cs = blk - f * it.bpf.d;
        const uint32_t row = it.ws.div(cs);
        const CamDev &c = it.cams[f];
        cam = &c;
        const int x0 = (int)(cs - row * it.ws.d) * 4, y0 = (int)row * 4;
        for (int k = 0; k < 4; ++k) {
            const f3 d = screen2world(c, x0 + 3 * (k & 1), y0 + 3 * (k >> 1));
            dc[k][0] = d.x;
            dc[k][1] = d.y;
            dc[k][2] = d.z;
        }
        // the bound's scale: the largest quaternion component and the
        // largest unrotated component (screen2world's vector)
        const float Q = fmaxf(fmaxf(fabsf(c.qw), fabsf(c.qi)),
                              fmaxf(fabsf(c.qj), fabsf(c.qk)));
        float V = 1.0f;
        V = fmaxf(V, fabsf(__builtin_fmaf(c.mul, (float)x0, -c.xsub)));
        V = fmaxf(V, fabsf(__builtin_fmaf(c.mul, (float)(x0 + 3), -c.xsub)));
        V = fmaxf(V, fabsf(__builtin_fmaf(-c.mul, (float)y0, c.ysub)));
        V = fmaxf(V, fabsf(__builtin_fmaf(-c.mul, (float)(y0 + 3), c.ysub)));
        return block_dirs(dc, Q, V);
    };
    // (not when counting: the counts add one root box test per finished
    // ray, which a deeper test would not match)
    const bool deep = !COUNT && it.t.ncut != 0;
    // one thread per block, 4 blocks per thread
    for (uint32_t j = 0; j < kCullBlocks / 256; ++j) {
        const uint32_t lb = j * 256 + tid;
        const uint32_t blk = (tile0 >> 4) + lb;
        const bool valid = blk * 16u < npix;
        uint8_t cls = 0;
        bool listed = false;
        uint32_t fs = 0;
        DirBox db;
        if (valid && !it.t.empty) {
            const CamDev *c;
            float dc[4][3];
            db = block_geom(blk, c, dc);
            CullOrigin co = co0;
            fs = it.bpf.div(blk) - f0;
            if (fs != 0) {
                const float o3[3] = {c->px, c->py, c->pz};
                co = cull_origin(lo3, hi3, o3);
            }
            if (block_may_enter(co, db)) {
                // 2: the four corner rays enter the box, 1: some ray may
                // (the queue order only)
                bool all = true;
                for (int k = 0; k < 4; ++k)
                    all = all && may_enter_dir(co, dc[k], [](float v) {
                              return __builtin_amdgcn_rcpf(v);
                          });
                cls = all ? 2 : 1;
                // listed for the test against the cut's boxes when its
                // frame is the tile's first or second (the boxes are
                // tabled for those two; a tile of small frames holds
                // more, whose blocks stay queued)
                listed = deep && fs < 2u && !co.unsure && !db.unsure;
            }
        }
        if (deep) {
            // in lane order, one LDS atomic per wave: a wave's 64 list
            // entries then come from neighbouring blocks
            const unsigned long long m = __ballot(listed);
            if (m) {
                const uint32_t first = (uint32_t)(__ffsll((long long)m) - 1);
                uint32_t base = 0;
                if (lane == first) base = atomicAdd(&ndeep, (uint32_t)__popcll(m));
                base = __shfl(base, (int)first);
                if (listed) {
                    const uint32_t k = base + lanes_below(m);
                    dlist[k] = (uint16_t)(lb | (fs << 15));
                    for (int i = 0; i < 3; ++i) {
                        ddir[i][k] = db.lo[i];
                        ddir[3 + i][k] = db.hi[i];
                    }
                }
            }
        }
        if (valid && cls == 0 && vox && !vox16) {
            uint32_t *v = vox + (size_t)blk * 16u;
            for (int k = 0; k < 16; ++k) v[k] = kEnd;
        }
        enter[lb] = valid ? cls : 0;
        if (COUNT && valid && cls == 0) my_missed++;
    }
    if (deep) {
        // The listed blocks, one per lane: finished as well when none of
        // their rays comes near any box of the cut (deep_origin,
        // block_may_enter_rcp: cull.h).  The boxes' slab bounds relative to
        // the camera (a, b of deep_origin) are tabled for the tile's first
        // two frames.  A wave first tests every box once against the union
        // of its 64 blocks' direction boxes (lane j, box j): a box that no
        // direction of the union comes near is none of its blocks' concern.
        const uint32_t ncut = min(it.t.ncut, (uint32_t)kCutMax);
        const uint32_t nfr = npix / 16u / it.bpf.d;
        if (tid < 2u * ncut) {
            const uint32_t fs = tid / ncut, j = tid - fs * ncut;
            if (f0 + fs < nfr) {
                const CamDev &c = it.cams[f0 + fs];
                const float o3[3] = {c.px, c.py, c.pz};
                const float *bx = it.t.cut + 6u * j;
                const CullOrigin cj =
                    deep_origin(bx, bx + 3, o3, cull_scale(lo3, hi3, o3));
                for (int i = 0; i < 3; ++i) {
                    cutab[fs][j][i] = cj.a[i];
                    cutab[fs][j][3 + i] = cj.b[i];
                }
            }
        }
        __syncthreads();
        auto cut_box = [&](uint32_t fs, uint32_t j) {
            CullOrigin cj;
            for (int i = 0; i < 3; ++i) {
                cj.a[i] = cutab[fs][j][i];
                cj.b[i] = cutab[fs][j][3 + i];
            }
            cj.unsure = false;
            return cj;
        };
        const uint32_t nd = ndeep;
        for (uint32_t k0 = tid & ~63u; k0 < nd; k0 += 256) {
            const uint32_t k = k0 + lane;
            const bool have = k < nd;
            uint32_t e = 0;
            DirBox db;
            for (int i = 0; i < 3; ++i) {
                db.lo[i] = INFINITY;
                db.hi[i] = -INFINITY;
            }
            db.unsure = false;
            if (have) {
                e = dlist[k];
                for (int i = 0; i < 3; ++i) {
                    db.lo[i] = ddir[i][k];
                    db.hi[i] = ddir[3 + i][k];
                }
            }
            const uint32_t fs = e >> 15;
            DirBox ub;
            for (int i = 0; i < 3; ++i) {
                float lo = db.lo[i], hi = db.hi[i];
                for (int o = 32; o > 0; o >>= 1) {
                    lo = fminf(lo, __shfl_xor(lo, o));
                    hi = fmaxf(hi, __shfl_xor(hi, o));
                }
                ub.lo[i] = lo;
                ub.hi[i] = hi;
            }
            ub.unsure = false;
            auto rcp = [](float v) { return __builtin_amdgcn_rcpf(v); };
            const DirRcp ur = dir_rcp(ub, rcp);
            const bool two = __ballot(have && fs == 1u) != 0;
            const unsigned long long m0 =
                __ballot(lane < ncut && block_may_enter_rcp(cut_box(0, lane), ur));
            const unsigned long long m1 =
                two ? __ballot(lane < ncut &&
                               block_may_enter_rcp(cut_box(1, lane), ur))
                    : 0ull;
            if (have) {
                // (a block whose union leaves no box finishes without its
                // own reciprocals)
                unsigned long long m = fs ? m1 : m0;
                bool keep = dir_any(db);
                if (m != 0 && !keep) {
                    const DirRcp r = dir_rcp(db, rcp);
                    for (; m != 0 && !keep; m &= m - 1ull)
                        keep = block_may_enter_rcp(
                            cut_box(fs, (uint32_t)(__ffsll((long long)m) - 1)), r);
                }
                if (!keep) {
                    const uint32_t lb = e & 0x3ffu;
                    enter[lb] = 0;
                    if (vox && !vox16) {
                        uint32_t *v = vox + (size_t)((tile0 >> 4) + lb) * 16u;
                        for (int q = 0; q < 16; ++q) v[q] = kEnd;
                    }
                }
            }
        }
    }
    __syncthreads();
    if (vox && vox16) {
        // the finished blocks' miss ids, consecutive threads on
        // consecutive 16 bytes (a wave stores 1 KiB at a time)
        const uint4 m = make_uint4(kEnd, kEnd, kEnd, kEnd);
        uint4 *v4 = reinterpret_cast<uint4 *>(vox + (size_t)tile0);
        for (uint32_t e = tid; e < kCullBlocks * 4; e += 256) {
            const uint32_t lb = e >> 2;
            if (((tile0 >> 4) + lb) * 16u < npix && enter[lb] == 0) v4[e] = m;
        }
    }
    __syncthreads();
    // ordered compaction of the 1024 block flags (thread t owns blocks
    // 4t..4t+3): blocks whose rays all enter go to the front of the queue,
    // the other entering blocks to the back, so the longest work is handed
    // out first and the end of the launch is made of the cheaper blocks.
    // Both counts are scanned at once, packed in the two halves of a word.
    uint32_t cls[4], mine = 0;
    for (uint32_t j = 0; j < 4; ++j) {
        cls[j] = enter[4 * tid + j];
        mine += cls[j] == 2 ? 1u : (cls[j] == 1 ? 0x10000u : 0u);
    }
    uint32_t incl = mine;
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
    }
    if (lane == 63) wave_sum[w] = incl;
    if (COUNT && my_missed) atomicAdd(&missed, (unsigned long long)my_missed);
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (uint32_t i = 0; i < 4; ++i) {
        before += (i < w) ? wave_sum[i] : 0u;
        total += wave_sum[i];
    }
    if (tid == 0) {
        unsigned long long base = 0;
        if (total)
            base = atomicAdd(&ctrl->qpair,
                             (unsigned long long)(total & 0xffffu) |
                                 ((unsigned long long)(total >> 16) << 32));
        gbase[0] = (uint32_t)base;
        gbase[1] = (uint32_t)(base >> 32);
    }
    __syncthreads();
    uint32_t off = before + incl - mine;  // exclusive, both halves
    uint32_t of = gbase[0] + (off & 0xffffu), ob = gbase[1] + (off >> 16);
    // the counts start at 0 unless a render failed after its pre-pass:
    // entries beyond the queue are then dropped (the readers clamp the
    // counts), never written out of bounds
    for (uint32_t j = 0; j < 4; ++j) {
        const uint32_t id = tile0 / 16 + 4 * tid + j;
        if (cls[j] == 2 && of < qlen) queue[of] = id;
        if (cls[j] == 1 && ob < qlen) queue[qlen - 1u - ob] = id;
        of += cls[j] == 2;
        ob += cls[j] == 1;
    }
    if (COUNT && tid == 0 && missed) {
        const unsigned long long nb = missed;
        const bool halved = it.rnd_mode == VOXG_QUALITY_ADAPTIVE &&
                            it.merge_mode == VOXG_QUALITY_RAY_MERGE;
        const unsigned long long traced = nb * (halved ? 8ull : 16ull);
        atomicAdd(&ctrl->visits, traced);
        atomicAdd(&ctrl->rays, traced);
        if (it.rnd_mode == VOXG_QUALITY_BEST) atomicAdd(&ctrl->best_blocks, nb);
        else atomicAdd(&ctrl->fast_blocks, nb);
        if (halved) atomicAdd(&ctrl->merge_blocks, nb);
    }
}

// Second pass of a FAST or ADAPTIVE render (renderer.c:240-316).  The
// first pass traced every pixel of the queued blocks as a root search,
// coherently (BestWork into SpecOut); here one thread per queued block
// replays the block's pixels in the reference's order -- ADAPTIVE's two
// corners, its mode and merge decisions, then each pixel: a copy of its
// predecessor when merging, in FAST mode a try of the previous pixel's
// leaf first, else (or when the try misses) the root search -- taking
// every root search from the first pass.  A root search depends on its ray alone, so each pixel gets the
// result the reference computes; root searches that the reference skips
// are traced in vain.  A leaf try is the traversal from a leaf node: the
// leaf's box test, then its voxels.  Results stay in SpecOut (a leaf try
// that hits replaces the root search there); emit_kernel writes them out.
//
// With plain ray merging the first pass traced only the pixels a block
// that keeps merging needs (0, 2, ..., 14 and 15: the odd ones are copies,
// renderer.c:290).  A block whose merging stops (both corners hit within
// 150 units of the camera, renderer.c:274-278) is listed in `redo` instead
// of replayed (redo_pass 0); once its odd pixels are traced too, a second
// launch (redo_pass 1) replays the listed blocks.
//
// DIRECT (FAST and ADAPTIVE without merging): the first pass wrote its root
// searches to the outputs and only the leaf nodes (and the corners' hit
// points) to SpecOut.  Without copies no pixel can turn from a hit into a
// miss: the replay appends the leaf tries that hit, the only results that
// differ from the root searches, to its wave's region of SpecOut (a
// ballot and a prefix per step, so the stores of a step are contiguous)
// for emit_changed_kernel, which then reads only those.
template <bool DIRECT>
__global__ void __launch_bounds__(256)
resolve_kernel(RenderItem it, SpecOut spec, const uint32_t *__restrict__ queue,
               uint32_t qlen, Ctrl *ctrl, uint32_t *__restrict__ redo,
               int redo_pass) {
    uint32_t qf, qt;
    if (redo_pass) {
        qf = qt = min((uint32_t)ctrl->rpair, qlen);
    } else {
        queue_counts(ctrl->qpair, qlen, qf, qt);
    }
    const Tree &T = it.t;
    unsigned long long nt = 0;
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < qt;
         b += gridDim.x * blockDim.x) {
        const uint32_t id = redo_pass
                                ? redo[b]
                                : queue[b < qf ? b : qlen - 1u - (b - qf)];
        const uint32_t f = it.bpf.div(id);
        const uint32_t cs = id - f * it.bpf.d;
        const uint32_t row = it.ws.div(cs);
        const int xs = (int)(cs - row * it.ws.d) * 4, ys = (int)row * 4;
        const CamDev &c = it.cams[f];
        const f3 o = f3{c.px, c.py, c.pz};
        // the first pass's root search of pixel p (all five words loaded
        // together: the node and point are only meaningful on a hit); in a
        // direct render the leaf node alone, and the point for the corners
        auto root = [&](int p, uint32_t &vid, uint32_t &nd, f3 &pt,
                        bool corner) {
            const uint32_t r = (uint32_t)p * spec.stride + id;
            if constexpr (DIRECT) {
                nd = spec.node[id * 16u + (uint32_t)p];
                vid = kEnd;  // used by merge copies only
                if (corner) {
                    const uint32_t k = (p == 15 ? spec.stride : 0u) + id;
                    pt = f3{spec.cx[k], spec.cy[k], spec.cz[k]};
                }
                return nd != kEnd;
            }
            vid = spec.vid[r];
            nd = spec.node[r];
            pt = f3{spec.hx[r], spec.hy[r], spec.hz[r]};
            return vid != kEnd;
        };
        auto store = [&](int p, uint32_t vid, f3 pt) {
            const uint32_t r = (uint32_t)p * spec.stride + id;
            spec.vid[r] = vid;
            spec.hx[r] = pt.x;
            spec.hy[r] = pt.y;
            spec.hz[r] = pt.z;
        };
        int mode = it.rnd_mode, merge = 0;
        uint32_t leaf = kEnd, vid_last = kEnd;
        f3 inter_last = o;
        int i = 0, iend = 16;
        if (mode == VOXG_QUALITY_ADAPTIVE) {
            // renderer.c:240-281: corners 0 and 15, then FAST or BEST, and
            // whether every other pixel copies its predecessor
            mode = VOXG_QUALITY_FAST;
            merge = it.merge_mode == VOXG_QUALITY_RAY_MERGE;
            uint32_t v0, n0, v15, n15;
            f3 p0, p15;
            const bool h0 = root(0, v0, n0, p0, true);
            leaf = h0 ? n0 : kEnd;
            i = 1;
            if (h0) {
                vid_last = v0;
                inter_last = p0;
                iend = 15;
                if (root(15, v15, n15, p15, true)) {
                    const f3 dir1 = screen2world(c, xs, ys);
                    const f3 dir15 = screen2world(c, xs + 3, ys + 3);
                    const float d1 = sqr_metric(p0, o);
                    const float d2 =
                        (dir1.x * dir1.x + dir1.y * dir1.y) + dir1.z * dir1.z;
                    const float criteria = d1 / d2 * sqr_metric(dir1, dir15);
                    const float dist = sqr_metric(p0, p15);
                    if ((double)(dist / criteria) > 1.69)
                        mode = VOXG_QUALITY_BEST;
                    if (d1 > 22500.0f) {
                        if (mode == VOXG_QUALITY_FAST &&
                            it.merge_mode == VOXG_QUALITY_RAY_MERGE_ACCURATE)
                            merge = 1;
                    } else {
                        merge = 0;
                    }
                }
            }
        }
        if (redo && !redo_pass) {
            // the block's merging stopped: list it (one atomic per wave)
            const bool wait = it.merge_mode == VOXG_QUALITY_RAY_MERGE && !merge;
            const unsigned long long m = __ballot(wait);
            if (m) {
                const uint32_t lane = lane_id();
                const uint32_t leader = (uint32_t)(__ffsll((long long)m) - 1);
                uint32_t base = 0;
                if (lane == leader)
                    base = (uint32_t)atomicAdd(
                        &ctrl->rpair, (unsigned long long)__popcll(m));
                base = __shfl(base, (int)leader);
                const uint32_t slot =
                    base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                // (a count left by a failed render cannot push a write
                // past the list; the readers clamp the count)
                if (wait && slot < qlen) redo[slot] = id;
            }
            if (wait) continue;
        }
        // BEST without merging: the root searches are the results
        if (!DIRECT && mode != VOXG_QUALITY_FAST && !merge) continue;
        // the record of the leaf last tried (a FAST chain often keeps it)
        uint32_t loaded = kEnd;
        Node ln;
        if constexpr (DIRECT) {
            // every lane of the wave takes 16 steps (its own pixels are
            // steps i .. iend - 1; a BEST block of ADAPTIVE has none), so
            // that each step's appends see one count on all lanes
            const int i0 = i;
            int i1 = mode == VOXG_QUALITY_FAST ? iend : i;
            const uint32_t region = b & ~63u;
            uint32_t run = 0;
            // which pixels' root searches hit: all 16 loads in flight at
            // once, rather than one per step; a block with none has nothing
            // to replay (its leaf stays empty), and neither has a wave
            uint32_t hits = 0;
            if (i1 > i0) {
                const uint4 *n4 = reinterpret_cast<const uint4 *>(spec.node) + id * 4u;
                for (int q = 0; q < 4; ++q) {
                    const uint4 v = n4[q];
                    hits |= ((v.x != kEnd ? 1u : 0u) | (v.y != kEnd ? 2u : 0u) |
                             (v.z != kEnd ? 4u : 0u) | (v.w != kEnd ? 8u : 0u))
                            << (4 * q);
                }
            }
            if (hits == 0) i1 = i0;
            for (int k = 0; k < 16; ++k) {
                if (__ballot(i0 + k < i1) == 0) break;
                const int ii = i0 + k;
                const int p = render_order(ii & 15);
                bool st = false;
                uint32_t svid = kEnd;
                f3 spt = o;
                if (ii < i1) {
                    const bool rhit = (hits >> p) & 1u;
                    const uint32_t rnd =
                        rhit ? spec.node[id * 16u + (uint32_t)p] : kEnd;
                    bool kept = false;
                    if (leaf != kEnd) {
                        // renderer.c:294-300: the previous pixel's leaf first
                        const f3 d = screen2world(c, xs + (p & 3), ys + (p >> 2));
                        if (leaf != loaded) {
                            ln = load_node(T.nodes, leaf);
                            loaded = leaf;
                        }
                        f3 e;
                        if (hit_box(ln.bmin, ln.bmax, o, d, e) &&
                            leaf_search<false, true>(T, ln, e, d, svid, spt, nt)) {
                            st = true;
                            kept = true;  // the leaf stays the same
                        }
                    }
                    if (!kept) leaf = rhit ? rnd : kEnd;
                }
                // the leaf tries that hit, appended to the wave's region
                const unsigned long long m = __ballot(st);
                if (st) {
                    const uint32_t e = region * 16u + run + lanes_below(m);
                    spec.ekey[e] = (b - region) | ((uint32_t)p << 6);
                    spec.vid[e] = svid;
                    spec.hx[e] = spt.x;
                    spec.hy[e] = spt.y;
                    spec.hz[e] = spt.z;
                }
                run += (uint32_t)__popcll(m);
            }
            const unsigned long long act = __ballot(true);
            if (lane_id() == (uint32_t)(__ffsll((long long)act) - 1))
                spec.ecount[region >> 6] = run;
            continue;
        }
        for (; i < iend; ++i) {
            const int p = render_order(i);
            if (merge && (i & 1)) {
                // renderer.c:290, 311-314: a copy of the previous pixel
                store(p, leaf != kEnd ? vid_last : kEnd, inter_last);
                continue;
            }
            // the root search is loaded before the leaf try, so that the
            // two memory round trips overlap
            uint32_t rvid, rnd;
            f3 rpt;
            const bool rhit = root(p, rvid, rnd, rpt, false);
            if (mode == VOXG_QUALITY_FAST && leaf != kEnd) {
                // renderer.c:294-300: the previous pixel's leaf first
                const f3 d = screen2world(c, xs + (p & 3), ys + (p >> 2));
                if (leaf != loaded) {
                    ln = load_node(T.nodes, leaf);
                    loaded = leaf;
                }
                f3 e, pt;
                uint32_t vid = kEnd;
                if (hit_box(ln.bmin, ln.bmax, o, d, e) &&
                    leaf_search<false, true>(T, ln, e, d, vid, pt, nt)) {
                    store(p, vid, pt);
                    vid_last = vid;
                    inter_last = pt;
                    continue;  // the leaf stays the same
                }
            }
            if (rhit) {
                leaf = rnd;
                vid_last = rvid;
                inter_last = rpt;
            } else {
                leaf = kEnd;
            }
        }
    }
}

// Third pass: the final results of the queued blocks from SpecOut to the
// render's outputs (square layout).  A workgroup takes 64 blocks at a time
// and transposes through LDS: it reads the blocks' 16 pixels row by row of
// SpecOut (neighbouring lanes, neighbouring blocks: consecutive words) and
// writes each block's run of the outputs with 16 neighbouring lanes.
// (Without the transpose, 16 lanes per block reading 16 rows took 0.42 ms
// on the benchmark frames; one thread per block writing its 16 pixels left
// lines partly written in L2 and was 3.6x slower.)  The last workgroup
// clears the queue counts for the next render (every workgroup read them
// when it started).
constexpr uint32_t kEmitBlocks = 64;

__global__ void __launch_bounds__(256)
emit_kernel(RenderItem it, SpecOut spec, const uint32_t *__restrict__ queue,
            uint32_t qlen, Ctrl *ctrl) {
    __shared__ uint32_t sid[kEmitBlocks];
    __shared__ uint32_t svid[kEmitBlocks * 16];
    __shared__ float spx[kEmitBlocks * 16], spy[kEmitBlocks * 16],
        spz[kEmitBlocks * 16];
    uint32_t qf, qt;
    queue_counts(ctrl->qpair, qlen, qf, qt);
    const uint32_t tid = threadIdx.x;
    for (uint32_t b0 = blockIdx.x * kEmitBlocks; b0 < qt;
         b0 += gridDim.x * kEmitBlocks) {
        const uint32_t nb = min(kEmitBlocks, qt - b0);
        if (tid < nb) {
            const uint32_t b = b0 + tid;
            sid[tid] = queue[b < qf ? b : qlen - 1u - (b - qf)];
        }
        __syncthreads();
        // rows p of SpecOut, 4 at a time: lane k reads block k
        const uint32_t k = tid & (kEmitBlocks - 1u);
        for (uint32_t p = tid / kEmitBlocks; p < 16u; p += 256u / kEmitBlocks) {
            if (k < nb) {
                const uint32_t r = p * spec.stride + sid[k];
                const uint32_t v = spec.vid[r];
                svid[k * 16u + p] = v;
                if (v != kEnd) {
                    spx[k * 16u + p] = spec.hx[r];
                    spy[k * 16u + p] = spec.hy[r];
                    spz[k * 16u + p] = spec.hz[r];
                }
            }
        }
        __syncthreads();
        for (uint32_t j = tid; j < nb * 16u; j += 256u) {
            const uint32_t v = svid[j];
            const bool hit = v != kEnd;
            f3 pt = f3{0, 0, 0};
            if (hit) pt = f3{spx[j], spy[j], spz[j]};
            it.sink.put(sid[j >> 4] * 16u + (j & 15u), hit, v, 0, pt);
        }
        __syncthreads();
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&ctrl->rdone, 1u) == gridDim.x - 1u) {
            atomicExch(&ctrl->qpair, 0ull);
            atomicExch(&ctrl->rpair, 0ull);
            atomicExch(&ctrl->rdone, 0u);
        }
    }
}

// Third pass of a direct render: the pixels whose leaf try hit, from the
// regions the replay's waves appended them to, over the first pass's
// results in the outputs.  A workgroup takes one region (64 queued
// blocks) at a time: it places the region's entries in LDS by block and
// pixel, then writes each block's run of the outputs with 16 neighbouring
// lanes, as emit_kernel does.  The loads of a region (its block ids and up
// to 4 entries per thread, all independent) go out while the previous
// region is written, and its entry count one region earlier still, so a
// region costs no chain of memory round trips.  (Reading the changed
// blocks' scratch rows instead read about as much as a full output pass,
// 0.25 ms on the benchmark frames; storing from the entries directly, one
// pixel of 64 blocks per instruction, took 0.5 ms.)  Every changed pixel
// is a hit.  The last workgroup clears the queue counts for the next
// render.
__global__ void __launch_bounds__(256)
emit_changed_kernel(RenderItem it, SpecOut spec,
                    const uint32_t *__restrict__ queue, uint32_t qlen,
                    Ctrl *ctrl) {
    __shared__ uint32_t sid[64];
    __shared__ uint32_t svid[1024];
    __shared__ float spx[1024], spy[1024], spz[1024];
    uint32_t qf, qt;
    queue_counts(ctrl->qpair, qlen, qf, qt);
    const uint32_t nreg = (qt + 63u) / 64u;
    const uint32_t tid = threadIdx.x, G = gridDim.x;
    auto nblocks = [&](uint32_t r) { return min(64u, qt - 64u * r); };
    // a region's entry count as written (loaded unconditionally, from a
    // clamped index, so that no copy of it waits for the load), then
    // bounded: at most 15 entries per block of the region
    auto raw_count = [&](uint32_t r) { return spec.ecount[min(r, nreg - 1u)]; };
    auto count = [&](uint32_t r, uint32_t raw) -> uint32_t {
        return r < nreg ? min(raw, 15u * nblocks(r)) : 0u;
    };
    uint32_t bid = 0, key[4], vid[4];
    float x[4], y[4], z[4];
    auto fetch = [&](uint32_t r, uint32_t cnt) {
        if (r >= nreg) return;
        if (tid < nblocks(r)) {
            const uint32_t bq = 64u * r + tid;
            bid = queue[bq < qf ? bq : qlen - 1u - (bq - qf)];
        }
        const uint32_t base = r * 1024u;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t e = base + tid + 256u * k;
            if (tid + 256u * k < cnt) {
                key[k] = spec.ekey[e];
                vid[k] = spec.vid[e];
                x[k] = spec.hx[e];
                y[k] = spec.hy[e];
                z[k] = spec.hz[e];
            }
        }
    };
    uint32_t r = blockIdx.x;
    if (r >= nreg) r = nreg;  // (nreg may be 0: no loads at all)
    uint32_t cnt = 0, raw_next = 0;
    if (r < nreg) {
        cnt = count(r, raw_count(r));
        raw_next = raw_count(r + G);
        fetch(r, cnt);
    }
    for (; r < nreg; r += G) {
        const uint32_t nb = nblocks(r);
        for (uint32_t j = tid; j < 1024u; j += 256u) svid[j] = kEnd;
        if (tid < 64u) sid[tid] = bid;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t lb = key[k] & 63u, p = (key[k] >> 6) & 15u;
            if (tid + 256u * k < cnt && lb < nb) {
                const uint32_t j = lb * 16u + p;
                svid[j] = vid[k];
                spx[j] = x[k];
                spy[j] = y[k];
                spz[j] = z[k];
            }
        }
        __syncthreads();
        // the next region's loads, in flight during this one's stores
        cnt = count(r + G, raw_next);
        raw_next = raw_count(r + 2u * G);
        fetch(r + G, cnt);
        for (uint32_t j = tid; j < nb * 16u; j += 256u) {
            const uint32_t v = svid[j];
            if (v != kEnd)
                it.sink.put(sid[j >> 4] * 16u + (j & 15u), true, v, 0,
                            f3{spx[j], spy[j], spz[j]});
        }
        __syncthreads();
    }
    __syncthreads();
    if (tid == 0) {
        __threadfence();
        if (atomicAdd(&ctrl->rdone, 1u) == gridDim.x - 1u) {
            atomicExch(&ctrl->qpair, 0ull);
            atomicExch(&ctrl->rpair, 0ull);
            atomicExch(&ctrl->rdone, 0u);
        }
    }
}

// ---------------------------------------------------------------------------
// ball kernel
// ---------------------------------------------------------------------------
// search.c:155-183: is there a voxel whose box comes closer than r to c
// (rr = r * r)?  Any-hit depth-first walk over children 0..7 through the
// escape links; the answer does not depend on the order.
template <bool COUNT>
__device__ __forceinline__ uint8_t ball_collides(const Tree &t, f3 c, float rr,
                                                 unsigned long long &nv,
                                                 unsigned long long &nt) {
    const f3 V = f3{t.vx, t.vy, t.vz};
    uint32_t node = t.empty ? kEnd : 0u;
    while (node != kEnd) {
        Node nd = load_node(t.nodes, node);
        if (COUNT) nv++;
        if (box_ball(nd.bmin, nd.bmax, c, rr)) {
            uint32_t type = nd.info & NT_TYPE_MASK;
            if (type == NT_DENSE) return 1;
            if (type == NT_LEAF) {
                uint32_t cnt = nd.info >> 8;
                for (uint32_t k = 0; k < cnt; k++) {
                    const float *p = t.dots + 3u * (nd.ref + k);
                    f3 vmin = f3{p[0], p[1], p[2]};
                    f3 vmax = f3{vmin.x + V.x, vmin.y + V.y, vmin.z + V.z};
                    if (COUNT) nt++;
                    if (box_ball(vmin, vmax, c, rr)) return 1;
                }
                node = nd.escape;
            } else {
                node = (nd.ref != kEnd) ? nd.ref : nd.escape;
            }
        } else {
            node = nd.escape;
        }
    }
    return 0;
}

struct BallItem {
    Tree t;
    const float *cx, *cy, *cz, *r;
    uint8_t *out;
    uint32_t begin, count;  // queries [begin, begin+count) of the launch
};

template <bool COUNT, bool MULTI>
__global__ void __launch_bounds__(256)
ball_kernel(BallItem one, const BallItem *__restrict__ items,
            uint32_t n_items, uint32_t n, Ctrl *ctrl) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long nv = 0, nt = 0;
    if (i < n) {
        const BallItem *itp = &one;
        if constexpr (MULTI) {
            uint32_t lo = 0, hi = n_items - 1;
            while (lo < hi) {
                uint32_t mid = (lo + hi + 1) >> 1;
                if (items[mid].begin <= i) lo = mid;
                else hi = mid - 1;
            }
            itp = &items[lo];
        }
        const BallItem &it = *itp;
        uint32_t k0 = i - it.begin;
        f3 c = f3{it.cx[k0], it.cy[k0], it.cz[k0]};
        float r = it.r[k0];
        uint8_t res = ball_collides<COUNT>(it.t, c, r * r, nv, nt);
        if (it.out) it.out[k0] = res;
    }
    if (COUNT) {
        atomicAdd(&ctrl->visits, nv);
        atomicAdd(&ctrl->tests, nt);
        if (i < n) atomicAdd(&ctrl->rays, 1ull);
    }
}

// cd.c:38-55 for n bodies: a body that moved (vox_dot_equalp,
// voxvision.h.in:52: x, y and z all compare equal) and whose ball collides
// with the scene goes back to its last position; otherwise its last
// position becomes its position.
__global__ void __launch_bounds__(256)
collide_kernel(Tree t, uint32_t n, float *__restrict__ px,
               float *__restrict__ py, float *__restrict__ pz,
               float *__restrict__ lx, float *__restrict__ ly,
               float *__restrict__ lz, const float *__restrict__ radius,
               uint8_t *__restrict__ collided) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const f3 p = f3{px[i], py[i], pz[i]};
    const f3 l = f3{lx[i], ly[i], lz[i]};
    uint8_t moved_back = 0;
    if (!(p.x == l.x && p.y == l.y && p.z == l.z)) {
        const float r = radius[i];
        unsigned long long nv = 0, nt = 0;
        if (ball_collides<false>(t, p, r * r, nv, nt)) {
            px[i] = l.x;
            py[i] = l.y;
            pz[i] = l.z;
            moved_back = 1;
        } else {
            lx[i] = p.x;
            ly[i] = p.y;
            lz[i] = p.z;
        }
    }
    if (collided) collided[i] = moved_back;
}

// Scratch of voxg_ray_batch_noalloc: the control record, the per-XCD unit
// counters, the overflow word and the stack spill buffer of one launch, in
// a file-scope array that each GPU holds for the life of the process.  Host
// code reaches it only through the address hipGetSymbolAddress returns on
// each device (the array's name is not a device address in host code).
constexpr size_t kNoAllocBytes = 32u << 20;
constexpr size_t kNoAllocCtrl = 0;       // Ctrl
constexpr size_t kNoAllocXnext = 256;    // kXcds counters, 256 B apart
constexpr size_t kNoAllocOvf = 2304;     // overflow word
constexpr size_t kNoAllocSpill = 4096;   // spill entries to the end
static_assert(sizeof(Ctrl) <= kNoAllocXnext, "Ctrl fits its slot");
static_assert(kNoAllocXnext + kXcds * kXcdStride * 4 <= kNoAllocOvf,
              "counters fit their slot");
__device__ __align__(256) uint8_t scratch[kNoAllocBytes];

// Clears what a launch expects at zero: the control record, the counters
// and the overflow word (a kernel rather than a memset, so that the call can
// be captured in a graph).
__global__ void __launch_bounds__(256) noalloc_clear_kernel(uint8_t *base) {
    uint32_t *w = reinterpret_cast<uint32_t *>(base);
    for (uint32_t i = threadIdx.x; i < kNoAllocSpill / 4; i += blockDim.x)
        w[i] = 0u;
}

}  // namespace dev

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct voxg_tree {
    HostTree host;
    NodeRec *d_nodes = nullptr;
    float *d_dots = nullptr;
    float *d_cut = nullptr;
    uint32_t ncut = 0;
    void *d_mem = nullptr;
    size_t bytes = 0;
    int device = 0;
    uint32_t stack_bound = 0;
};

namespace {
struct Workspace;
}

// Light manager (src/voxrnd/lights.c): host M-tree and ambient light, and
// a device copy of the flattened tree refreshed when a render needs it
// after a change.
struct voxg_lights {
    voxg::LightTree tree;
    float ambient[3] = {1.0f, 1.0f, 1.0f};  // lights.c:20
    bool dirty = true;
    int device = -1;
    void *d_mem = nullptr;
    size_t d_bytes = 0;
    uint32_t nnodes = 0;
    const voxg::LightNode *d_nodes = nullptr;
    const voxg::LightRec *d_sph = nullptr;
};

struct voxg_batch {
    int device = 0;
    dev::BatchItem *d_items = nullptr;
    uint32_t n_items = 0, units = 0, chunk = 1, stack_bound = 1;
    uint64_t rays_end = 0;  // one past the highest ray index used
    // launch configuration and workspace fixed at creation, so that a graph
    // captured around voxg_batch_ray never refers to buffers that another
    // call could reallocate
    voxg_launch_config cfg;
    Workspace *ws = nullptr;
};

namespace {

std::mutex g_mu;
// stack_lds, block, waves_per_cu, refill, chunk, render_cull, leaf_batch,
// finish_batch, split_best per kind of launch (fastest measured on the
// 256^3 benchmark scene; every field is validated by voxg_set_*)
voxg_launch_config g_cfg[VOXG_LAUNCH_KINDS] = {
    {4, 256, 28, 8, 1, 1, 24, 16, 0},   // VOXG_LAUNCH_RAYS
    {4, 256, 24, 8, 4, 1, 16, 16, 0},   // VOXG_LAUNCH_RENDER
    {4, 256, 28, 48, 4, 1, 32, 64, 4},  // VOXG_LAUNCH_RENDER_BEST
    {4, 256, 24, 8, 1, 1, 16, 16, 0},   // VOXG_LAUNCH_BATCH
};

voxg_launch_config current_config(int kind) {
    std::lock_guard<std::mutex> g(g_mu);
    return g_cfg[kind];
}
thread_local int g_device = 0;

struct Workspace {
    dev::Ctrl *ctrl = nullptr;
    uint32_t *xnext = nullptr;  // per-XCD unit counters (kXcds x 256 bytes)
    uint4 *spill = nullptr;
    size_t spill_bytes = 0;
    uint32_t *queue = nullptr;  // block queue of the render pre-pass
    size_t queue_bytes = 0;
    void *aux = nullptr;        // scratch of the other translation units
    size_t aux_bytes = 0;
    // stack-overflow flag of the traversals launched on this stream: pinned,
    // mapped host memory that a kernel sets (it is never expected to); read
    // and cleared by the host at the next launch or synchronising call
    uint32_t *ovf_h = nullptr, *ovf_d = nullptr;
    // a render enqueued its pre-pass but not the launch that clears the
    // queue counts (a later step failed): the next render clears them first
    bool dirty = false;
};
std::map<std::pair<int, hipStream_t>, std::unique_ptr<Workspace>> g_ws;

void free_workspace(Workspace *w) {
    if (w->ctrl) (void)hipFree(w->ctrl);
    if (w->xnext) (void)hipFree(w->xnext);
    if (w->spill) (void)hipFree(w->spill);
    if (w->queue) (void)hipFree(w->queue);
    if (w->aux) (void)hipFree(w->aux);
    if (w->ovf_h) (void)hipHostFree(w->ovf_h);
}

// Test hook: the n-th host checkpoint of a render after this is armed
// returns VOXG_EIO as if a launch had failed (0 = disarmed).
thread_local int g_inject = 0;
bool injected_failure() {
    return g_inject > 0 && --g_inject == 0;
}
// Test hook: when nonzero, the stack bound given to traversal launches is
// at most this (so that a bound violation can be forced from the host; the
// kernel then drops the entries beyond it and sets the overflow flag).
std::atomic<uint32_t> g_test_bound{0};

int cu_count(int device) {
    static std::map<int, int> cache;
    std::lock_guard<std::mutex> g(g_mu);
    auto it = cache.find(device);
    if (it != cache.end()) return it->second;
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount,
                              device) != hipSuccess || n <= 0)
        n = 256;
    cache[device] = n;
    return n;
}

// Rays per work unit for a launch of n rays: the configured chunk, cut so
// that a small launch still has a unit for every persistent lane.
uint32_t effective_chunk(size_t n, int kind) {
    voxg_launch_config cfg = current_config(kind);
    int devid = 0;
    if (hipGetDevice(&devid) != hipSuccess) devid = 0;
    uint64_t lanes = (uint64_t)cu_count(devid) *
                     (uint64_t)std::max(1, cfg.waves_per_cu) * 64u;
    uint64_t c = (uint64_t)std::max(1, cfg.chunk);
    uint64_t fit = std::max<uint64_t>(1, (uint64_t)n / lanes);
    return (uint32_t)std::min(c, fit);
}

// The control record, the unit counters (cleared) and the overflow flag
// of a new workspace; on failure everything allocated is freed.
int init_workspace(Workspace *w) {
    const size_t xb = dev::kXcds * dev::kXcdStride * sizeof(uint32_t);
    hipError_t e = hipMalloc(&w->ctrl, sizeof(dev::Ctrl));
    if (e == hipSuccess) e = hipMemset(w->ctrl, 0, sizeof(dev::Ctrl));
    if (e == hipSuccess) e = hipMalloc(&w->xnext, xb);
    if (e == hipSuccess) e = hipMemset(w->xnext, 0, xb);
    if (e == hipSuccess)
        e = hipHostMalloc((void **)&w->ovf_h, 256,
                          hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) {
        *(volatile uint32_t *)w->ovf_h = 0u;
        e = hipHostGetDevicePointer((void **)&w->ovf_d, w->ovf_h, 0);
    }
    if (e != hipSuccess) {
        free_workspace(w);
        *w = Workspace();
        return fail(VOXG_EIO, std::string("workspace: ") +
                                  hipGetErrorString(e));
    }
    return 0;
}

int get_ws(hipStream_t s, size_t spill_bytes, Workspace **out) {
    int devid = 0;
    HIPCHK(hipGetDevice(&devid));
    std::lock_guard<std::mutex> g(g_mu);
    auto &slot = g_ws[{devid, s}];
    if (!slot) {
        // registered only once its buffers exist and are cleared
        std::unique_ptr<Workspace> w(new Workspace());
        int rc = init_workspace(w.get());
        if (rc) {
            g_ws.erase({devid, s});
            return rc;
        }
        slot = std::move(w);
    }
    if (spill_bytes > slot->spill_bytes) {
        // kernels already queued on the stream may still use the old buffer
        HIPCHK(hipStreamSynchronize(s));
        if (slot->spill) HIPCHK(hipFree(slot->spill));
        slot->spill = nullptr;
        slot->spill_bytes = 0;
        HIPCHK(hipMalloc(&slot->spill, spill_bytes));
        slot->spill_bytes = spill_bytes;
    }
    *out = slot.get();
    return 0;
}

int get_queue(Workspace *ws, hipStream_t s, size_t bytes) {
    std::lock_guard<std::mutex> g(g_mu);
    if (bytes > ws->queue_bytes) {
        HIPCHK(hipStreamSynchronize(s));
        if (ws->queue) HIPCHK(hipFree(ws->queue));
        ws->queue = nullptr;
        ws->queue_bytes = 0;
        HIPCHK(hipMalloc(&ws->queue, bytes));
        ws->queue_bytes = bytes;
    }
    return 0;
}

}  // namespace

namespace voxg_internal {
int fail(int code, const std::string &msg) { return ::fail(code, msg); }

int scratch(hipStream_t stream, size_t bytes, void **out) {
    Workspace *ws = nullptr;
    int rc = get_ws(stream, 16, &ws);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(g_mu);
    if (bytes > ws->aux_bytes) {
        HIPCHK(hipStreamSynchronize(stream));
        if (ws->aux) HIPCHK(hipFree(ws->aux));
        ws->aux = nullptr;
        ws->aux_bytes = 0;
        HIPCHK(hipMalloc(&ws->aux, bytes));
        ws->aux_bytes = bytes;
    }
    *out = ws->aux;
    return 0;
}

void release_stream(hipStream_t stream) {
    int devid = 0;
    if (hipGetDevice(&devid) != hipSuccess) return;
    (void)hipStreamSynchronize(stream);
    std::lock_guard<std::mutex> g(g_mu);
    auto it = g_ws.find({devid, stream});
    if (it == g_ws.end()) return;
    free_workspace(it->second.get());
    g_ws.erase(it);
}
}  // namespace voxg_internal

namespace {

// Reads and clears the stream's overflow flag: VOXG_EIO when a traversal
// launched on it exceeded its stack bound (the ray of that lane was left
// incomplete), else 0.
int take_overflow(Workspace *ws) {
    volatile uint32_t *f = ws->ovf_h;
    if (f && *f) {
        *f = 0u;
        return fail(VOXG_EIO, "a traversal launched on this stream exceeded "
                              "its stack bound; its results are incomplete");
    }
    return 0;
}

// Clears the block-queue counts of a render that failed after its
// pre-pass (a fill kernel, so that the call can be graph-captured).
__global__ void clear_queue_kernel(dev::Ctrl *c) {
    c->qpair = 0ull;
    c->rpair = 0ull;
    c->rdone = 0u;
}

int clear_dirty_queue(Workspace *ws, hipStream_t stream) {
    if (!ws->dirty) return 0;
    hipLaunchKernelGGL(clear_queue_kernel, dim3(1), dim3(1), 0, stream,
                       ws->ctrl);
    HIPCHK(hipGetLastError());
    ws->dirty = false;
    return 0;
}

}  // namespace

namespace {

// 1/v when v is a power of two (then x * (1/v) == x / v exactly), else 0.
float pow2_inverse(float v) {
    int e = 0;
    float m = std::frexp(v, &e);
    if (!(v > 0.0f) || m != 0.5f || e < -125 || e > 126) return 0.0f;
    return std::ldexp(1.0f, 1 - e);
}

dev::Tree dev_tree(const voxg_tree *t) {
    dev::Tree d;
    std::memset(&d, 0, sizeof(d));
    d.nodes = t->d_nodes;
    d.dots = t->d_dots;
    d.vx = t->host.voxel[0];
    d.vy = t->host.voxel[1];
    d.vz = t->host.voxel[2];
    d.ivx = pow2_inverse(d.vx);
    d.ivy = pow2_inverse(d.vy);
    d.ivz = pow2_inverse(d.vz);
    d.empty = t->host.nodes.empty() ? 1u : 0u;
    d.ncut = t->ncut;
    d.cut = t->d_cut;
    return d;
}

// Grid of a persistent traversal launch of n units: waves_per_cu waves per
// CU, never more lanes than units (rounded up to whole workgroups; n is only
// an upper bound when the count comes from the device).
int traversal_grid(const voxg_launch_config &cfg, size_t n, int devid) {
    const int cus = cu_count(devid);
    const int block = cfg.block;
    const int waves = std::max(1, cfg.waves_per_cu);
    const long long threads = (long long)cus * waves * 64;
    int grid = (int)std::max<long long>(1, threads / block);
    long long need_blocks = ((long long)n + block - 1) / block;
    if (need_blocks < 1) need_blocks = 1;
    if (grid > need_blocks) grid = (int)need_blocks;
    return grid;
}

struct NoPre {
    int operator()(Workspace *) const { return 0; }
};

// One traversal launch.  Everything that can fail on the host is reserved
// first (workspace, spill buffer, LDS attribute); `pre` then enqueues what
// must run before the traversal on the stream (the render pre-pass, which
// may also fill in the work's queue), and the launch follows.
template <int K, class Work, bool COUNT, class Pre>
int launch_k(const dev::Tree &dt, uint32_t stack_bound, Work &work, size_t n,
             hipStream_t stream, const voxg_launch_config &cfg,
             voxg_counts *counts, bool reset_counts, Workspace *own,
             Pre &&pre) {
    int devid = 0;
    HIPCHK(hipGetDevice(&devid));
    const int block = cfg.block;
    const int grid = traversal_grid(cfg, n, devid);
    uint32_t gthreads = (uint32_t)grid * (uint32_t)block;
    if (const uint32_t tb = g_test_bound.load()) stack_bound = std::min(stack_bound, tb);
    uint32_t slots = stack_bound > (uint32_t)K ? stack_bound - K : 0u;
    const size_t spill_need =
        (size_t)std::max<uint32_t>(slots, 1u) * gthreads * kEntryBytes;
    // the stream's workspace, or one the caller owns (sized in advance)
    Workspace *ws = own;
    int rc = 0;
    if (!ws) {
        rc = get_ws(stream, spill_need, &ws);
        if (rc) return rc;
    } else if (spill_need > ws->spill_bytes) {
        return fail(VOXG_EINVAL, "workspace too small for this launch");
    }
    rc = take_overflow(ws);
    if (rc) return rc;
    size_t lds = (size_t)K * block * kEntryBytes;
    auto kern = dev::rti_kernel<K, Work, COUNT>;
    if (lds > 65536) {
        HIPCHK(hipFuncSetAttribute((const void *)kern,
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
    }
    if (COUNT && reset_counts) {
        HIPCHK(hipMemsetAsync(&ws->ctrl->visits, 0,
                              dev::kCounterBytes, stream));
    }
    rc = pre(ws);
    if (rc) return rc;
    if (injected_failure()) return fail(VOXG_EIO, "injected failure");
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), lds, stream, dt, work,
                       (uint32_t)n, ws->ctrl, ws->ovf_d, ws->xnext,
                       ws->spill, slots,
                       (uint32_t)cfg.refill, (uint32_t)cfg.finish_batch,
                       (uint32_t)cfg.leaf_batch);
    HIPCHK(hipGetLastError());
    // the traversal's last wave clears the queue counts unless a later
    // pass still needs them
    if (!work.keeps_queue()) ws->dirty = false;
    if (COUNT) {
        dev::Ctrl h;
        HIPCHK(hipMemcpyAsync(&h, ws->ctrl, sizeof(h), hipMemcpyDeviceToHost,
                              stream));
        HIPCHK(hipStreamSynchronize(stream));
        rc = take_overflow(ws);
        if (rc) return rc;
        counts->node_visits = h.visits;
        counts->voxel_tests = h.tests;
        counts->rays = h.rays;
        counts->leaf_traversals = h.leaf_trav;
        counts->fast_blocks = h.fast_blocks;
        counts->best_blocks = h.best_blocks;
        counts->merge_blocks = h.merge_blocks;
        counts->stack_spills = h.spills;
    }
    return 0;
}

template <class Work, bool COUNT, class Pre = NoPre>
int launch_cfg(const voxg_launch_config &cfg, const dev::Tree &t,
               uint32_t stack_bound, Work &work, size_t n,
               hipStream_t stream, voxg_counts *counts,
               bool reset_counts = true, Workspace *own = nullptr,
               Pre &&pre = Pre()) {
    switch (cfg.stack_lds) {
    case 2: return launch_k<2, Work, COUNT>(t, stack_bound, work, n, stream, cfg, counts, reset_counts, own, pre);
    case 4: return launch_k<4, Work, COUNT>(t, stack_bound, work, n, stream, cfg, counts, reset_counts, own, pre);
    case 8: return launch_k<8, Work, COUNT>(t, stack_bound, work, n, stream, cfg, counts, reset_counts, own, pre);
    case 16: return launch_k<16, Work, COUNT>(t, stack_bound, work, n, stream, cfg, counts, reset_counts, own, pre);
    case 32: return launch_k<32, Work, COUNT>(t, stack_bound, work, n, stream, cfg, counts, reset_counts, own, pre);
    default: return fail(VOXG_EINVAL, "stack_lds must be 2, 4, 8, 16 or 32");
    }
}

template <class Work, bool COUNT>
int launch_tree(int kind, const dev::Tree &t, uint32_t stack_bound,
                Work &work, size_t n, hipStream_t stream,
                voxg_counts *counts, bool reset_counts = true) {
    return launch_cfg<Work, COUNT>(current_config(kind), t, stack_bound, work,
                                   n, stream, counts, reset_counts);
}

template <class Work, bool COUNT>
int launch(int kind, const voxg_tree *t, Work &work, size_t n,
           hipStream_t stream, voxg_counts *counts) {
    return launch_tree<Work, COUNT>(kind, dev_tree(t), t->stack_bound, work,
                                    n, stream, counts);
}

int check_tree(const voxg_tree *t) {
    if (!t) return fail(VOXG_EINVAL, "tree is NULL");
    int devid = 0;
    HIPCHK(hipGetDevice(&devid));
    if (devid != t->device)
        return fail(VOXG_EINVAL, "tree lives on another device");
    return 0;
}

int upload(voxg_tree *t) {
    const HostTree &h = t->host;
    size_t nb = h.nodes.size() * sizeof(NodeRec);
    size_t db = h.dots.size() * sizeof(float);
    std::vector<float> cut;
    try {
        cut = cut_of(h);
    } catch (const std::bad_alloc &) {
        return fail(VOXG_ENOMEM, "host allocation failed");
    }
    const TreeBytes tb = tree_bytes(h, cut.size());
    t->bytes = tb.total;
    HIPCHK(hipGetDevice(&t->device));
    if (t->bytes == 0) return 0;
    HIPCHK(hipMalloc(&t->d_mem, t->bytes));
    t->d_nodes = (NodeRec *)t->d_mem;
    t->d_dots = (float *)((char *)t->d_mem + tb.dots_off);
    if (nb) HIPCHK(hipMemcpy(t->d_nodes, h.nodes.data(), nb, hipMemcpyHostToDevice));
    if (db) HIPCHK(hipMemcpy(t->d_dots, h.dots.data(), db, hipMemcpyHostToDevice));
    if (!cut.empty()) {
        t->d_cut = (float *)((char *)t->d_mem + tb.cut_off);
        t->ncut = (uint32_t)(cut.size() / 6);
        HIPCHK(hipMemcpy(t->d_cut, cut.data(), cut.size() * sizeof(float),
                         hipMemcpyHostToDevice));
    }
    return 0;
}

int finish_tree(voxg_tree *t, voxg_tree **out) {
    t->stack_bound = stack_bound_of(t->host);
    int rc = upload(t);
    if (rc) {
        if (t->d_mem) (void)hipFree(t->d_mem);
        delete t;
        return rc;
    }
    *out = t;
    return 0;
}

}  // namespace

// Device address of dev::scratch on each device (hipGetSymbolAddress once
// per device, with that device current); 0 until set up.
std::mutex g_noalloc_mu;
std::map<int, uint8_t *> g_noalloc;

static int noalloc_setup(int device) {
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    if (g_noalloc.count(device)) return 0;
    void *ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(dev::scratch)));
    if (!ptr) return fail(VOXG_EIO, "no address for the scratch array");
    g_noalloc[device] = (uint8_t *)ptr;
    return 0;
}

extern "C" {

const char *voxg_last_error(void) { return g_err.c_str(); }

int voxg_synchronize(hipStream_t stream) {
    HIPCHK(hipStreamSynchronize(stream));
    int devid = 0;
    HIPCHK(hipGetDevice(&devid));
    Workspace *ws = nullptr;
    {
        std::lock_guard<std::mutex> g(g_mu);
        auto it = g_ws.find({devid, stream});
        if (it != g_ws.end()) ws = it->second.get();
    }
    return ws ? take_overflow(ws) : 0;
}

int voxg_release_stream(hipStream_t stream) {
    voxg_internal::release_stream(stream);
    return 0;
}

int voxg_test_fail_at(int checkpoint) {
    if (checkpoint < 0) return fail(VOXG_EINVAL, "checkpoint must be >= 0");
    g_inject = checkpoint;
    return 0;
}

int voxg_test_stack_bound(unsigned bound) {
    g_test_bound.store(bound);
    return 0;
}

int voxg_init(int device) {
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return fail(VOXG_ENODEV, "no such HIP device");
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, device));
    if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
        return fail(VOXG_ENODEV, std::string("device is ") + p.gcnArchName +
                                     ", this library is built for gfx950");
    g_device = device;
    return noalloc_setup(device);
}

int voxg_tree_build(const float *xyz, size_t n, size_t stride,
                    const float voxel[3], voxg_tree **out) {
    if (!out) return fail(VOXG_EINVAL, "out is NULL");
    *out = nullptr;
    if (n && !xyz) return fail(VOXG_EINVAL, "xyz is NULL");
    if (stride < 3) return fail(VOXG_EINVAL, "stride must be >= 3");
    float V[3] = {1.0f, 1.0f, 1.0f};
    if (voxel) std::memcpy(V, voxel, sizeof(V));
    if (!(V[0] > 0.0f && V[1] > 0.0f && V[2] > 0.0f))
        return fail(VOXG_EINVAL, "voxel size must be positive");
    std::unique_ptr<voxg_tree> t(new (std::nothrow) voxg_tree());
    if (!t) return fail(VOXG_ENOMEM, "host allocation failed");
    std::string err;
    int rc;
    try {
        rc = build_tree(xyz, n, stride, V, t->host, err);
    } catch (const std::bad_alloc &) {
        return fail(VOXG_ENOMEM, "host allocation failed");
    }
    if (rc) return fail(rc, err);
    return finish_tree(t.release(), out);
}

int voxg_tree_upload(const struct vox_node *root, const float voxel[3],
                     voxg_tree **out) {
    if (!out) return fail(VOXG_EINVAL, "out is NULL");
    *out = nullptr;
    float V[3] = {1.0f, 1.0f, 1.0f};
    if (voxel) std::memcpy(V, voxel, sizeof(V));
    if (!(V[0] > 0.0f && V[1] > 0.0f && V[2] > 0.0f))
        return fail(VOXG_EINVAL, "voxel size must be positive");
    std::unique_ptr<voxg_tree> t(new (std::nothrow) voxg_tree());
    if (!t) return fail(VOXG_ENOMEM, "host allocation failed");
    std::string err;
    int rc;
    try {
        rc = flatten_reference(root, V, t->host, err);
    } catch (const std::bad_alloc &) {
        return fail(VOXG_ENOMEM, "host allocation failed");
    }
    if (rc) return fail(rc, err);
    return finish_tree(t.release(), out);
}

int voxg_tree_release(voxg_tree *t) {
    if (!t) return 0;
    if (t->d_mem) HIPCHK(hipFree(t->d_mem));
    delete t;
    return 0;
}

static void fill_info(const HostTree &h, voxg_tree_info *o) {
    std::memset(o, 0, sizeof(*o));
    o->nodes = h.nodes.size();
    o->inner = h.inner;
    o->leaves = h.leaves;
    o->dense = h.dense;
    o->overflow = h.overflow;
    o->dots = h.ndots;
    o->voxel_ids = h.voxel_ids;
    o->depth = h.depth;
    o->stack_entries = stack_bound_of(h);
    o->device_bytes = h.nodes.empty() ? 0 : device_bytes_of(h);
    if (!h.nodes.empty()) {
        for (int l = 0; l < 3; l++) {
            o->bbox_min[l] = h.nodes[0].bmin[l];
            o->bbox_max[l] = h.nodes[0].bmax[l];
        }
    }
    for (int l = 0; l < 3; l++) o->voxel[l] = h.voxel[l];
}

int voxg_tree_get_info(const voxg_tree *t, voxg_tree_info *o) {
    if (!t || !o) return fail(VOXG_EINVAL, "NULL argument");
    fill_info(t->host, o);
    return 0;
}

int voxg_tree_plan(const float *xyz, size_t n, size_t stride,
                   const float voxel[3], voxg_tree_info *info,
                   uint64_t *hash) {
    if (n && !xyz) return fail(VOXG_EINVAL, "xyz is NULL");
    if (stride < 3) return fail(VOXG_EINVAL, "stride must be >= 3");
    float V[3] = {1.0f, 1.0f, 1.0f};
    if (voxel) std::memcpy(V, voxel, sizeof(V));
    if (!(V[0] > 0.0f && V[1] > 0.0f && V[2] > 0.0f))
        return fail(VOXG_EINVAL, "voxel size must be positive");
    HostTree h;
    std::string err;
    int rc;
    try {
        rc = build_tree(xyz, n, stride, V, h, err);
    } catch (const std::bad_alloc &) {
        return fail(VOXG_ENOMEM, "host allocation failed");
    }
    if (rc) return fail(rc, err);
    if (info) fill_info(h, info);
    if (hash) *hash = tree_hash(h);
    return 0;
}

int voxg_test_plan_cut(const float *xyz, size_t n, size_t stride,
                       const float voxel[3], float *out, size_t cap,
                       size_t *count) {
    if ((n && !xyz) || !count || (cap && !out))
        return fail(VOXG_EINVAL, "NULL argument");
    if (stride < 3) return fail(VOXG_EINVAL, "stride must be >= 3");
    float V[3] = {1.0f, 1.0f, 1.0f};
    if (voxel) std::memcpy(V, voxel, sizeof(V));
    if (!(V[0] > 0.0f && V[1] > 0.0f && V[2] > 0.0f))
        return fail(VOXG_EINVAL, "voxel size must be positive");
    HostTree h;
    std::string err;
    std::vector<float> cut;
    try {
        const int rc = build_tree(xyz, n, stride, V, h, err);
        if (rc) return fail(rc, err);
        cut = cut_of(h);
    } catch (const std::bad_alloc &) {
        return fail(VOXG_ENOMEM, "host allocation failed");
    }
    *count = cut.size() / 6;
    if (*count > cap) return fail(VOXG_EINVAL, "cap is too small");
    if (!cut.empty()) std::memcpy(out, cut.data(), cut.size() * sizeof(float));
    return 0;
}

int voxg_tree_plan_reference(const struct vox_node *root, const float voxel[3],
                             voxg_tree_info *info, uint64_t *hash) {
    float V[3] = {1.0f, 1.0f, 1.0f};
    if (voxel) std::memcpy(V, voxel, sizeof(V));
    if (!(V[0] > 0.0f && V[1] > 0.0f && V[2] > 0.0f))
        return fail(VOXG_EINVAL, "voxel size must be positive");
    HostTree h;
    std::string err;
    int rc;
    try {
        rc = flatten_reference(root, V, h, err);
    } catch (const std::bad_alloc &) {
        return fail(VOXG_ENOMEM, "host allocation failed");
    }
    if (rc) return fail(rc, err);
    if (info) fill_info(h, info);
    if (hash) *hash = tree_hash(h);
    return 0;
}

uint64_t voxg_tree_hash(const voxg_tree *t) {
    return t ? tree_hash(t->host) : 0;
}

int voxg_tree_voxel_coords(const voxg_tree *t, size_t n, const uint32_t *ids,
                           float *xyz) {
    if (!t || (n && (!ids || !xyz))) return fail(VOXG_EINVAL, "NULL argument");
    return voxel_coords(t->host, n, ids, xyz);
}

const struct vox_node *voxg_tree_node_ptr(const voxg_tree *t, uint32_t i) {
    if (!t || i >= t->host.src.size()) return nullptr;
    return (const struct vox_node *)t->host.src[i];
}

void voxg_camera_set_window_size(voxg_camera *cam, int w, int h) {
    // simple-camera.c:120-136
    if (!cam) return;
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

static int check_launch_config(const voxg_launch_config *cfg) {
    if (!cfg) return fail(VOXG_EINVAL, "cfg is NULL");
    int k = cfg->stack_lds;
    if (k != 2 && k != 4 && k != 8 && k != 16 && k != 32)
        return fail(VOXG_EINVAL, "stack_lds must be 2, 4, 8, 16 or 32");
    if (cfg->block < 64 || cfg->block > 256 || cfg->block % 64)
        return fail(VOXG_EINVAL, "block must be 64, 128, 192 or 256");
    if ((size_t)k * cfg->block * kEntryBytes > 160 * 1024)
        return fail(VOXG_EINVAL,
                    "stack_lds * block * 20 exceeds 160 KiB of LDS");
    if (cfg->waves_per_cu < 1 || cfg->waves_per_cu > 64)
        return fail(VOXG_EINVAL, "waves_per_cu must be in 1..64");
    if (cfg->refill < 1 || cfg->refill > 64)
        return fail(VOXG_EINVAL, "refill must be in 1..64");
    if (cfg->chunk < 1 || cfg->chunk > 4096)
        return fail(VOXG_EINVAL, "chunk must be in 1..4096");
    if (cfg->render_cull != 0 && cfg->render_cull != 1)
        return fail(VOXG_EINVAL, "render_cull must be 0 or 1");
    if (cfg->leaf_batch < 1 || cfg->leaf_batch > 64)
        return fail(VOXG_EINVAL, "leaf_batch must be in 1..64");
    if (cfg->finish_batch < 1 || cfg->finish_batch > 64)
        return fail(VOXG_EINVAL, "finish_batch must be in 1..64");
    if (cfg->split_best < 0 || cfg->split_best > 4)
        return fail(VOXG_EINVAL, "split_best must be in 0..4");
    return 0;
}

static int check_kind(int kind) {
    if (kind < 0 || kind >= VOXG_LAUNCH_KINDS)
        return fail(VOXG_EINVAL, "unknown kind of launch");
    return 0;
}

int voxg_set_launch_config(const voxg_launch_config *cfg) {
    int rc = check_launch_config(cfg);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(g_mu);
    for (int k = 0; k < VOXG_LAUNCH_KINDS; k++) g_cfg[k] = *cfg;
    return 0;
}

int voxg_get_launch_config(voxg_launch_config *cfg) {
    return voxg_get_kind_launch_config(VOXG_LAUNCH_RAYS, cfg);
}

int voxg_set_kind_launch_config(int kind, const voxg_launch_config *cfg) {
    int rc = check_kind(kind);
    if (!rc) rc = check_launch_config(cfg);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(g_mu);
    g_cfg[kind] = *cfg;
    return 0;
}

int voxg_get_kind_launch_config(int kind, voxg_launch_config *cfg) {
    int rc = check_kind(kind);
    if (rc) return rc;
    if (!cfg) return fail(VOXG_EINVAL, "cfg is NULL");
    std::lock_guard<std::mutex> g(g_mu);
    *cfg = g_cfg[kind];
    return 0;
}

static int ray_common(const voxg_tree *t, size_t n, const float *ox,
                      const float *oy, const float *oz, const float *dx,
                      const float *dy, const float *dz) {
    int rc = check_tree(t);
    if (rc) return rc;
    if (n >= 0x80000000ull) return fail(VOXG_ERANGE, "n must be < 2^31");
    if (n && (!ox || !oy || !oz || !dx || !dy || !dz))
        return fail(VOXG_EINVAL, "ray input pointer is NULL");
    return 0;
}

int voxg_ray_batch(const voxg_tree *t, size_t n, const float *ox,
                   const float *oy, const float *oz, const float *dx,
                   const float *dy, const float *dz, uint32_t *voxel_out,
                   uint32_t *leaf_out, float *hx, float *hy, float *hz,
                   hipStream_t stream) {
    int rc = ray_common(t, n, ox, oy, oz, dx, dy, dz);
    if (rc) return rc;
    if ((hx || hy || hz) && !(hx && hy && hz))
        return fail(VOXG_EINVAL, "hx, hy, hz must be all set or all NULL");
    if (n == 0) return 0;
    uint32_t c = effective_chunk(n, VOXG_LAUNCH_RAYS);
    dev::ChunkWork<dev::SoaSrc, dev::SoaSink> work{
        dev::SoaSrc{ox, oy, oz, dx, dy, dz},
        dev::SoaSink{voxel_out, leaf_out, hx, hy, hz}, (uint32_t)n, c};
    return launch<decltype(work), false>(VOXG_LAUNCH_RAYS, t, work,
                                         (n + c - 1) / c, stream, nullptr);
}

int voxg_ray_batch_count(const voxg_tree *t, size_t n, const float *ox,
                         const float *oy, const float *oz, const float *dx,
                         const float *dy, const float *dz, voxg_counts *out,
                         hipStream_t stream) {
    int rc = ray_common(t, n, ox, oy, oz, dx, dy, dz);
    if (rc) return rc;
    if (!out) return fail(VOXG_EINVAL, "out is NULL");
    std::memset(out, 0, sizeof(*out));
    if (n == 0) return 0;
    uint32_t c = effective_chunk(n, VOXG_LAUNCH_RAYS);
    dev::ChunkWork<dev::SoaSrc, dev::SoaSink> work{
        dev::SoaSrc{ox, oy, oz, dx, dy, dz},
        dev::SoaSink{nullptr, nullptr, nullptr, nullptr, nullptr},
        (uint32_t)n, c};
    return launch<decltype(work), true>(VOXG_LAUNCH_RAYS, t, work,
                                        (n + c - 1) / c, stream, out);
}

int voxg_ray_batch_noalloc(const voxg_tree *t, size_t n, const float *ox,
                           const float *oy, const float *oz, const float *dx,
                           const float *dy, const float *dz,
                           uint32_t *voxel_out, uint32_t *leaf_out, float *hx,
                           float *hy, float *hz, hipStream_t stream) {
    int rc = ray_common(t, n, ox, oy, oz, dx, dy, dz);
    if (rc) return rc;
    if ((hx || hy || hz) && !(hx && hy && hz))
        return fail(VOXG_EINVAL, "hx, hy, hz must be all set or all NULL");
    if (n == 0) return 0;
    int devid = 0;
    HIPCHK(hipGetDevice(&devid));
    rc = noalloc_setup(devid);
    if (rc) return rc;
    uint8_t *base = nullptr;
    {
        std::lock_guard<std::mutex> g(g_noalloc_mu);
        base = g_noalloc[devid];
    }
    voxg_launch_config cfg = current_config(VOXG_LAUNCH_RAYS);
    uint32_t bound = t->stack_bound;
    if (const uint32_t tb = g_test_bound.load()) bound = std::min(bound, tb);
    const uint32_t slots =
        std::max<uint32_t>(1u, bound > (uint32_t)cfg.stack_lds
                                   ? bound - (uint32_t)cfg.stack_lds
                                   : 0u);
    // as many persistent lanes as the scratch holds spill entries for: the
    // configured waves per CU when they fit, fewer otherwise
    const size_t entries = (dev::kNoAllocBytes - dev::kNoAllocSpill) /
                           kEntryBytes;
    const size_t lanes_fit = entries / slots;
    const size_t wave_all = (size_t)cu_count(devid) * 64;  // a wave per CU
    if (lanes_fit < wave_all)
        return fail(VOXG_ERANGE,
                    "tree too deep for the no-allocation scratch (" +
                        std::to_string(dev::kNoAllocBytes >> 20) +
                        " MiB holds " + std::to_string(lanes_fit) +
                        " lanes' spill entries, fewer than one wave per CU)");
    cfg.waves_per_cu = (int)std::min<size_t>(
        (size_t)std::max(1, cfg.waves_per_cu), lanes_fit / wave_all);
    uint32_t c = effective_chunk(n, VOXG_LAUNCH_RAYS);
    dev::ChunkWork<dev::SoaSrc, dev::SoaSink> work{
        dev::SoaSrc{ox, oy, oz, dx, dy, dz},
        dev::SoaSink{voxel_out, leaf_out, hx, hy, hz}, (uint32_t)n, c};
    // a workspace that points into the scratch array; nothing allocated
    Workspace ws;
    ws.ctrl = (dev::Ctrl *)(base + dev::kNoAllocCtrl);
    ws.xnext = (uint32_t *)(base + dev::kNoAllocXnext);
    ws.ovf_d = (uint32_t *)(base + dev::kNoAllocOvf);
    ws.spill = (uint4 *)(base + dev::kNoAllocSpill);
    ws.spill_bytes = entries * kEntryBytes;
    hipLaunchKernelGGL(dev::noalloc_clear_kernel, dim3(1), dim3(256), 0,
                       stream, base);
    HIPCHK(hipGetLastError());
    dev::Tree dt = dev_tree(t);
    rc = launch_cfg<decltype(work), false>(cfg, dt, bound, work,
                                           (n + c - 1) / c, stream, nullptr,
                                           true, &ws);
    // the Workspace destructor frees nothing: its pointers are the array's
    return rc;
}

static int ball_launch(const voxg_tree *t, size_t n, const float *cx,
                       const float *cy, const float *cz, const float *r,
                       uint8_t *out, voxg_counts *counts, hipStream_t stream) {
    int rc = check_tree(t);
    if (rc) return rc;
    if (n >= 0x80000000ull) return fail(VOXG_ERANGE, "n must be < 2^31");
    if (n && (!cx || !cy || !cz || !r))
        return fail(VOXG_EINVAL, "ball input pointer is NULL");
    if (n && !out && !counts) return fail(VOXG_EINVAL, "out is NULL");
    if (n == 0) return 0;
    Workspace *ws = nullptr;
    rc = get_ws(stream, 16, &ws);
    if (rc) return rc;
    int block = 256;
    int grid = (int)((n + block - 1) / block);
    dev::BallItem one;
    std::memset(&one, 0, sizeof(one));
    one.t = dev_tree(t);
    one.cx = cx;
    one.cy = cy;
    one.cz = cz;
    one.r = r;
    one.out = out;
    one.begin = 0;
    one.count = (uint32_t)n;
    if (counts) {
        HIPCHK(hipMemsetAsync(&ws->ctrl->visits, 0,
                              dev::kCounterBytes, stream));
        hipLaunchKernelGGL((dev::ball_kernel<true, false>), dim3(grid),
                           dim3(block), 0, stream, one, nullptr, 0u,
                           (uint32_t)n, ws->ctrl);
        HIPCHK(hipGetLastError());
        dev::Ctrl h;
        HIPCHK(hipMemcpyAsync(&h, ws->ctrl, sizeof(h), hipMemcpyDeviceToHost,
                              stream));
        HIPCHK(hipStreamSynchronize(stream));
        counts->node_visits = h.visits;
        counts->voxel_tests = h.tests;
        counts->rays = h.rays;
        return 0;
    }
    hipLaunchKernelGGL((dev::ball_kernel<false, false>), dim3(grid),
                       dim3(block), 0, stream, one, nullptr, 0u, (uint32_t)n,
                       ws->ctrl);
    HIPCHK(hipGetLastError());
    return 0;
}

int voxg_ball_batch(const voxg_tree *t, size_t n, const float *cx,
                    const float *cy, const float *cz, const float *r,
                    uint8_t *out, hipStream_t stream) {
    if (n && !out) return fail(VOXG_EINVAL, "out is NULL");
    return ball_launch(t, n, cx, cy, cz, r, out, nullptr, stream);
}

int voxg_cd_collide_batch(const voxg_tree *t, size_t n, float *px, float *py,
                          float *pz, float *lx, float *ly, float *lz,
                          const float *radius, uint8_t *collided,
                          hipStream_t stream) {
    int rc = check_tree(t);
    if (rc) return rc;
    if (n >= 0x80000000ull) return fail(VOXG_ERANGE, "n must be < 2^31");
    if (n == 0) return 0;
    if (!px || !py || !pz || !lx || !ly || !lz || !radius)
        return fail(VOXG_EINVAL, "body pointer is NULL");
    const uint32_t block = 256;
    const uint32_t grid = (uint32_t)((n + block - 1) / block);
    hipLaunchKernelGGL(dev::collide_kernel, dim3(grid), dim3(block), 0, stream,
                       dev_tree(t), (uint32_t)n, px, py, pz, lx, ly, lz,
                       radius, collided);
    HIPCHK(hipGetLastError());
    return 0;
}

int voxg_ball_batch_count(const voxg_tree *t, size_t n, const float *cx,
                          const float *cy, const float *cz, const float *r,
                          voxg_counts *out, hipStream_t stream) {
    if (!out) return fail(VOXG_EINVAL, "out is NULL");
    std::memset(out, 0, sizeof(*out));
    return ball_launch(t, n, cx, cy, cz, r, nullptr, out, stream);
}

static int render_common(const voxg_tree *t, size_t n_frames,
                         const voxg_camera *cams, int W, int H,
                         const voxg_render_params *params, uint32_t *squares,
                         size_t *npix) {
    int rc = check_tree(t);
    if (rc) return rc;
    if (W <= 0 || H <= 0 || (W & 0xf) || (H & 0x3))
        return fail(VOXG_EINVAL,
                    "W must be a positive multiple of 16 and H of 4");
    if (n_frames && !cams) return fail(VOXG_EINVAL, "cams is NULL");
    unsigned q = params ? params->quality : VOXG_QUALITY_BEST;
    // renderer.c:179-195
    if ((q & 0xc) && (q & 3) != VOXG_QUALITY_ADAPTIVE)
        return fail(VOXG_EINVAL, "ray merging needs the adaptive mode");
    if ((q & 3) > 2 || (q & 0xc) > 8 || (q & ~0xfu))
        return fail(VOXG_EINVAL, "reserved quality bits set");
    if (squares && (!params || !params->texture))
        return fail(VOXG_EINVAL, "colour output needs a texture");
    // block corners are kept in 16-bit fields
    if (W >= 65536 || H >= 65536)
        return fail(VOXG_ERANGE, "W and H must be < 65536");
    // n_frames * per < 2^31, tested without forming a product that could
    // wrap around
    const size_t per = (size_t)W * (size_t)H;
    if (n_frames > 0x7fffffffull / per)
        return fail(VOXG_ERANGE, "n_frames * W * H must be < 2^31");
    *npix = n_frames * per;
    return 0;
}

extern "C++" {
static dev::PixSink make_sink(const voxg_tree *t,
                              const voxg_render_params *params,
                              uint32_t *squares, uint32_t *voxel_px,
                              float *hit_px, size_t npix) {
    dev::PixSink sink;
    std::memset(&sink, 0, sizeof(sink));
    sink.squares = squares;
    sink.voxel = voxel_px;
    sink.hit = hit_px;
    sink.plane = npix;
    sink.sh.tex = params ? params->texture : nullptr;
    sink.sh.rshift = params ? params->rshift : 0;
    sink.sh.gshift = params ? params->gshift : 0;
    sink.sh.bshift = params ? params->bshift : 0;
    sink.sh.amask = params ? params->amask : 0;
    sink.sh.vx = t->host.voxel[0];
    sink.sh.vy = t->host.voxel[1];
    sink.sh.vz = t->host.voxel[2];
    sink.sh.ivx = pow2_inverse(sink.sh.vx);
    sink.sh.ivy = pow2_inverse(sink.sh.vy);
    sink.sh.ivz = pow2_inverse(sink.sh.vz);
    return sink;
}

// One render: the block pre-pass (when enabled) and the traversal over the
// blocks it queued, with units of `Work` (BlockWork<false> or BestWork).
// Everything that can fail on the host is reserved before the pre-pass is
// enqueued; a render that fails after its pre-pass leaves the stream's
// workspace dirty, and the next render clears the queue counts first.
template <class Work, bool COUNT>
static int render_run(const voxg_tree *t, Work &work,
                      const voxg_launch_config &cfg, size_t npix,
                      voxg_counts *counts, hipStream_t stream) {
    const size_t nblk = npix / 16;
    const size_t units = nblk * work.units_per_block_host();
    if (!cfg.render_cull || t->host.nodes.empty())
        return launch_cfg<Work, COUNT>(cfg, dev_tree(t), t->stack_bound, work,
                                       units, stream, counts);
    // pre-pass: finish the blocks whose rays all miss the root box and
    // queue the others for the traversal kernel
    Workspace *ws = nullptr;
    int rc = get_ws(stream, 16, &ws);
    if (rc) return rc;
    rc = get_queue(ws, stream, nblk * sizeof(uint32_t));
    if (rc) return rc;
    rc = clear_dirty_queue(ws, stream);
    if (rc) return rc;
    auto pre = [&](Workspace *w) -> int {
        if (COUNT)
            HIPCHK(hipMemsetAsync(&w->ctrl->visits, 0, dev::kCounterBytes,
                                  stream));
        w->dirty = true;
        hipLaunchKernelGGL((dev::block_cull_kernel<COUNT>),
                           dim3((npix + dev::kCullTile - 1) / dev::kCullTile),
                           dim3(256), 0, stream, work.one, (uint32_t)npix,
                           w->ctrl, w->queue, (uint32_t)nblk);
        HIPCHK(hipGetLastError());
        work.queue = w->queue;
        work.qpair = &w->ctrl->qpair;
        work.qlen = (uint32_t)nblk;
        return 0;
    };
    return launch_cfg<Work, COUNT>(cfg, dev_tree(t), t->stack_bound, work,
                                   units, stream, counts, false, nullptr,
                                   pre);
}

// FAST and ADAPTIVE renders, after the pre-pass: every pixel of the
// blocks that the pre-pass queues is root-searched as in a BEST render
// (coherent groups, BEST-render launch configuration) into a per-stream
// scratch buffer of 20 bytes per pixel, resolve_kernel replays each
// block's pixels in the reference's order, with its leaf tries, from those
// results, and emit_kernel writes the outputs.  With plain ray merging
// (skip_odd) the first pass traces 9 of a block's 16 pixels, the ones a
// block that keeps merging needs; the blocks whose merging stops get their
// other 7 traced and are replayed by two more launches before the output
// pass.  Without merging (direct) the first pass writes the outputs and
// only leaf nodes to the scratch, and the output pass writes only the
// pixels whose leaf try hit.  The output pass clears the queue counts;
// until it is enqueued the stream's workspace stays dirty.
static int render_speculative(const voxg_tree *t, const dev::RenderItem &one,
                              size_t npix, bool skip_odd, bool direct,
                              hipStream_t stream) {
    voxg_launch_config cfg = current_config(VOXG_LAUNCH_RENDER_BEST);
    cfg.render_cull = 1;  // the resolve pass reads the pre-pass's queue
    const size_t nblk = npix / 16;
    void *sp = nullptr;
    // root-search results (voxel id, leaf node, hit point: 20 bytes per
    // pixel), then the list of blocks whose merging stopped or (direct) the
    // appended entries' keys, the regions' counts and the corners' points
    const size_t nreg = (nblk + 63) / 64;
    const size_t bytes = npix * 20 + (direct ? npix * 4 + nreg * 4 + nblk * 24
                                             : nblk * 4);
    int rc = voxg_internal::scratch(stream, bytes, &sp);
    if (rc) return rc;
    dev::SpecOut spec;
    std::memset(&spec, 0, sizeof(spec));
    spec.vid = (uint32_t *)sp;
    spec.node = spec.vid + npix;
    spec.hx = (float *)(spec.node + npix);
    spec.hy = spec.hx + npix;
    spec.hz = spec.hy + npix;
    spec.stride = (uint32_t)nblk;
    uint32_t *redo = skip_odd ? (uint32_t *)(spec.hz + npix) : nullptr;
    if (direct) {
        spec.ekey = (uint32_t *)(spec.hz + npix);
        spec.ecount = spec.ekey + npix;
        spec.cx = (float *)(spec.ecount + nreg);
        spec.cy = spec.cx + 2 * nblk;
        spec.cz = spec.cy + 2 * nblk;
    }
    // one pixel per unit, into the scratch; the queue counts stay for the
    // passes below
    auto first_pass = [&](auto &work) {
        std::memset(&work, 0, sizeof(work));
        work.one = one;
        work.spec = spec;
        work.keep_queue = 1;
        work.split = 4;
    };
    if (direct) {
        dev::BestDirectWork work;
        first_pass(work);
        work.group = 256u << work.split;
        rc = render_run<dev::BestDirectWork, false>(t, work, cfg, npix,
                                                    nullptr, stream);
    } else if (skip_odd) {
        dev::BestSetWork work;
        first_pass(work);
        work.pset = 1;
        work.pdiv = make_fastdiv(9);
        work.group = 256u * 9u;
        rc = render_run<dev::BestSetWork, false>(t, work, cfg, npix, nullptr,
                                                 stream);
    } else {
        dev::BestSpecWork work;
        first_pass(work);
        work.group = 256u << work.split;
        rc = render_run<dev::BestSpecWork, false>(t, work, cfg, npix, nullptr,
                                                  stream);
    }
    if (rc) return rc;
    Workspace *ws = nullptr;
    rc = get_ws(stream, 16, &ws);
    if (rc) return rc;
    int devid = 0;
    HIPCHK(hipGetDevice(&devid));
    // workgroups of 256 threads per CU: 8 for the passes of the merging
    // modes; for the direct ones 16 for the replay (more blocks in flight
    // hide its leaf loads: 0.45 -> 0.38 ms on the benchmark frames) and 2
    // for the output pass (fewer regions in flight at a time write fewer
    // lines at once: 0.22 -> 0.12 ms)
    auto wg_grid = [&](size_t per_cu) {
        return std::max<size_t>(1, std::min<size_t>((nblk + 255) / 256,
                                                    (size_t)cu_count(devid) * per_cu));
    };
    const size_t grid = wg_grid(8);
    if (injected_failure()) return fail(VOXG_EIO, "injected failure");
    if (direct) {
        hipLaunchKernelGGL(dev::resolve_kernel<true>, dim3((unsigned)wg_grid(16)),
                           dim3(256), 0, stream, one, spec, ws->queue,
                           (uint32_t)nblk, ws->ctrl, nullptr, 0);
        HIPCHK(hipGetLastError());
        if (injected_failure()) return fail(VOXG_EIO, "injected failure");
        hipLaunchKernelGGL(dev::emit_changed_kernel, dim3((unsigned)wg_grid(2)),
                           dim3(256), 0, stream, one, spec, ws->queue,
                           (uint32_t)nblk, ws->ctrl);
        HIPCHK(hipGetLastError());
        ws->dirty = false;
        return 0;
    }
    hipLaunchKernelGGL(dev::resolve_kernel<false>, dim3((unsigned)grid),
                       dim3(256), 0, stream, one, spec, ws->queue,
                       (uint32_t)nblk, ws->ctrl, redo, 0);
    HIPCHK(hipGetLastError());
    if (skip_odd) {
        // the odd pixels of the listed blocks, then their replay (the
        // first pass reserved a spill buffer for at least as many lanes)
        dev::BestSetWork odd;
        first_pass(odd);
        odd.pset = 2;
        odd.pdiv = make_fastdiv(7);
        odd.group = 256u * 7u;
        odd.queue = redo;
        odd.qpair = &ws->ctrl->rpair;
        odd.qlen = (uint32_t)nblk;
        rc = launch_cfg<dev::BestSetWork, false>(
            cfg, dev_tree(t), t->stack_bound, odd, nblk * 7, stream, nullptr,
            false);
        if (rc) return rc;
        hipLaunchKernelGGL(dev::resolve_kernel<false>, dim3((unsigned)grid),
                           dim3(256), 0, stream, one, spec, ws->queue,
                           (uint32_t)nblk, ws->ctrl, redo, 1);
        HIPCHK(hipGetLastError());
    }
    if (injected_failure()) return fail(VOXG_EIO, "injected failure");
    hipLaunchKernelGGL(dev::emit_kernel, dim3((unsigned)grid), dim3(256), 0,
                       stream, one, spec, ws->queue, (uint32_t)nblk,
                       ws->ctrl);
    HIPCHK(hipGetLastError());
    ws->dirty = false;
    return 0;
}

template <bool COUNT>
static int render_launch(const voxg_tree *t, size_t n_frames,
                         const voxg_camera *cams, int W, int H,
                         const voxg_render_params *params,
                         const dev::PixSink &sink, size_t npix,
                         voxg_counts *counts, hipStream_t stream) {
    static_assert(sizeof(voxg_camera) == sizeof(dev::CamDev), "camera layout");
    unsigned q = params ? params->quality : VOXG_QUALITY_BEST;
    dev::RenderItem one;
    std::memset(&one, 0, sizeof(one));
    one.t = dev_tree(t);
    one.cams = (const dev::CamDev *)cams;
    one.W = (uint32_t)W;
    one.ws = make_fastdiv((uint32_t)W / 4u);
    one.bpf = make_fastdiv((uint32_t)(W / 4) * (uint32_t)(H / 4));
    one.unit_begin = 0;
    one.rnd_mode = (int)(q & 3);
    one.merge_mode = (int)(q & 0xc);
    one.sink = sink;
    if (one.rnd_mode == VOXG_QUALITY_BEST) {
        // every pixel is a root search: units of 16 >> split_best pixels,
        // down to one; the lanes that a refill starts together then trace
        // neighbouring pixels, coherent rays
        const voxg_launch_config cfg = current_config(VOXG_LAUNCH_RENDER_BEST);
        dev::BestWork work;
        std::memset(&work, 0, sizeof(work));
        work.one = one;
        work.split = (uint32_t)cfg.split_best;
        // units dealt to the XCDs in turns of 256 blocks' worth keep the
        // longest-first order of the block queue on every XCD, and spread
        // the refill atomics over the 8 per-XCD counters
        work.group = 256u << work.split;
        return render_run<dev::BestWork, COUNT>(t, work, cfg, npix, counts,
                                                stream);
    }
    const voxg_launch_config cfg = current_config(VOXG_LAUNCH_RENDER);
    // speculated passes (render_speculative) except for the counting
    // launch; with plain ray merging the first pass skips the odd pixels,
    // which most blocks copy
    if (!COUNT && cfg.render_cull && !t->host.nodes.empty())
        return render_speculative(t, one, npix,
                                  (q & VOXG_QUALITY_RAY_MERGE) != 0,
                                  one.merge_mode == 0, stream);
    // one 4x4 block per unit (renderer.c:214): the leaf of one pixel is
    // tried first for the next (the counting launch and renders without
    // the pre-pass take this path)
    dev::BlockWork<false> work;
    std::memset(&work, 0, sizeof(work));
    work.one = one;
    work.group = 256u;
    return render_run<dev::BlockWork<false>, COUNT>(t, work, cfg, npix,
                                                    counts, stream);
}

}  // extern "C++"

// Device copy of the manager's flattened M-tree on the current device,
// made again after a change (synchronous; a manager with no lights keeps
// no device memory, and its walk ends at once).
static int lights_upload(voxg_lights *l) {
    int devid = 0;
    HIPCHK(hipGetDevice(&devid));
    if (!l->dirty && l->device == devid) return 0;
    std::vector<voxg::LightNode> nodes;
    std::vector<voxg::LightRec> sph;
    try {
        l->tree.flatten(nodes, sph);
    } catch (const std::bad_alloc &) {
        return fail(VOXG_ENOMEM, "host allocation failed");
    }
    const size_t nb = nodes.size() * sizeof(voxg::LightNode);
    const size_t sb = sph.size() * sizeof(voxg::LightRec);
    if (l->d_mem && (l->device != devid || nb + sb > l->d_bytes)) {
        HIPCHK(hipDeviceSynchronize());  // renders may still read it
        HIPCHK(hipFree(l->d_mem));
        l->d_mem = nullptr;
        l->d_bytes = 0;
    }
    if (nb && !l->d_mem) {
        HIPCHK(hipMalloc(&l->d_mem, nb + sb));
        l->d_bytes = nb + sb;
    } else if (nb) {
        HIPCHK(hipDeviceSynchronize());  // earlier renders read the old copy
    }
    if (nb) {
        HIPCHK(hipMemcpy(l->d_mem, nodes.data(), nb, hipMemcpyHostToDevice));
        if (sb)
            HIPCHK(hipMemcpy((char *)l->d_mem + nb, sph.data(), sb,
                             hipMemcpyHostToDevice));
    }
    l->d_nodes = (const voxg::LightNode *)l->d_mem;
    l->d_sph = (const voxg::LightRec *)((char *)l->d_mem + nb);
    l->nnodes = (uint32_t)nodes.size();
    l->device = devid;
    l->dirty = false;
    return 0;
}

static bool light_color_ok(const float c[3]) {  // lights.c:81-89
    for (int i = 0; i < 3; i++)
        if (!(c[i] >= 0.0f && c[i] <= 1.0f)) return false;
    return true;
}

int voxg_lights_create(voxg_lights **out) {
    if (!out) return fail(VOXG_EINVAL, "out is NULL");
    *out = new (std::nothrow) voxg_lights();
    return *out ? 0 : fail(VOXG_ENOMEM, "host allocation failed");
}

int voxg_lights_release(voxg_lights *l) {
    if (!l) return 0;
    if (l->d_mem) {
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipFree(l->d_mem));
    }
    delete l;
    return 0;
}

int voxg_lights_set_ambient(voxg_lights *l, const float color[3], int *set) {
    if (!l || !color) return fail(VOXG_EINVAL, "NULL argument");
    const bool ok = light_color_ok(color);
    if (ok) std::memcpy(l->ambient, color, sizeof(l->ambient));
    if (set) *set = ok ? 1 : 0;
    return 0;
}

int voxg_lights_insert(voxg_lights *l, const float center[3], float radius,
                       const float color[3], int *inserted) {
    if (!l || !center || !color) return fail(VOXG_EINVAL, "NULL argument");
    int r = 0;
    if (light_color_ok(color)) {
        try {
            r = l->tree.insert(center, radius, color);
        } catch (const std::bad_alloc &) {
            return fail(VOXG_ENOMEM, "host allocation failed");
        }
    }
    if (r < 0)
        return fail(VOXG_EINVAL, "a light with this centre exists, or the "
                                 "centre or radius is not finite");
    if (r) l->dirty = true;
    if (inserted) *inserted = r;
    return 0;
}

int voxg_lights_delete(voxg_lights *l, const float center[3], float radius,
                       int *deleted) {
    if (!l || !center) return fail(VOXG_EINVAL, "NULL argument");
    const int r = l->tree.remove(center, radius);
    if (r) l->dirty = true;
    if (deleted) *deleted = r;
    return 0;
}

int voxg_lights_clear(voxg_lights *l) {
    if (!l) return fail(VOXG_EINVAL, "NULL argument");
    l->tree.clear();
    l->dirty = true;
    return 0;
}

size_t voxg_lights_count(const voxg_lights *l) {
    return l ? l->tree.count() : 0;
}

int voxg_lights_get(const voxg_lights *l, size_t n, const float *xyz,
                    float *rgb) {
    if (!l || (n && (!xyz || !rgb))) return fail(VOXG_EINVAL, "NULL argument");
    for (size_t i = 0; i < n; i++)
        l->tree.light(l->ambient, xyz + 3 * i, rgb + 3 * i);
    return 0;
}

int voxg_render(const voxg_tree *t, size_t n_frames, const voxg_camera *cams,
                int W, int H, const voxg_render_params *params,
                uint32_t *squares, uint32_t *voxel_px, float *hit_px,
                hipStream_t stream) {
    size_t n = 0;
    int rc = render_common(t, n_frames, cams, W, H, params, squares, &n);
    if (rc) return rc;
    if (n == 0) return 0;
    dev::PixSink sink = make_sink(t, params, squares, voxel_px, hit_px, n);
    if (squares && params->lights) {
        rc = lights_upload(const_cast<voxg_lights *>(params->lights));
        if (rc) return rc;
        const voxg_lights *l = params->lights;
        sink.sh.lnodes = l->d_nodes;
        sink.sh.lsph = l->d_sph;
        sink.sh.nlnodes = l->nnodes;
        for (int k = 0; k < 3; k++) sink.sh.amb[k] = l->ambient[k];
    }
    return render_launch<false>(t, n_frames, cams, W, H, params, sink, n,
                                nullptr, stream);
}

int voxg_render_count(const voxg_tree *t, size_t n_frames,
                      const voxg_camera *cams, int W, int H,
                      const voxg_render_params *params, voxg_counts *out,
                      hipStream_t stream) {
    if (!out) return fail(VOXG_EINVAL, "out is NULL");
    std::memset(out, 0, sizeof(*out));
    size_t n = 0;
    int rc = render_common(t, n_frames, cams, W, H, params, nullptr, &n);
    if (rc) return rc;
    if (n == 0) return 0;
    dev::PixSink sink = make_sink(t, nullptr, nullptr, nullptr, nullptr, n);
    return render_launch<true>(t, n_frames, cams, W, H, params, sink, n, out,
                               stream);
}


int voxg_batch_create(size_t n_items, const voxg_batch_item *items,
                      voxg_batch **out) {
    if (!out) return fail(VOXG_EINVAL, "out is NULL");
    *out = nullptr;
    if (n_items == 0 || !items) return fail(VOXG_EINVAL, "no items");
    if (n_items > (1u << 20)) return fail(VOXG_ERANGE, "too many items");
    int devid = 0;
    HIPCHK(hipGetDevice(&devid));
    uint64_t total = 0;  // only picks the chunk; each count is checked below
    for (size_t i = 0; i < n_items; i++)
        total += std::min<uint64_t>(items[i].count, 0x80000000ull);
    uint32_t chunk = effective_chunk((size_t)total, VOXG_LAUNCH_BATCH);
    std::vector<dev::BatchItem> h;
    uint64_t units = 0, rays_end = 0;
    uint32_t bound = 1;
    for (size_t i = 0; i < n_items; i++) {
        const voxg_batch_item &it = items[i];
        int rc = check_tree(it.tree);
        if (rc) return rc;
        // each term below 2^31 first, so that the sum cannot wrap
        if (it.first >= 0x80000000ull || it.count >= 0x80000000ull ||
            it.first + it.count >= 0x80000000ull)
            return fail(VOXG_ERANGE, "ray indices must stay below 2^31");
        if (it.count == 0) continue;
        dev::BatchItem b;
        std::memset(&b, 0, sizeof(b));
        b.t = dev_tree(it.tree);
        b.ray_begin = (uint32_t)it.first;
        b.ray_count = (uint32_t)it.count;
        b.unit_begin = (uint32_t)units;
        units += (it.count + chunk - 1) / chunk;
        rays_end = std::max<uint64_t>(rays_end, it.first + it.count);
        bound = std::max(bound, it.tree->stack_bound);
        h.push_back(b);
    }
    if (units >= 0x80000000ull) return fail(VOXG_ERANGE, "too many rays");
    std::unique_ptr<voxg_batch> b(new (std::nothrow) voxg_batch());
    if (!b) return fail(VOXG_ENOMEM, "host allocation failed");
    b->device = devid;
    b->n_items = (uint32_t)h.size();
    b->units = (uint32_t)units;
    b->chunk = chunk;
    b->stack_bound = bound;
    b->rays_end = rays_end;
    b->cfg = current_config(VOXG_LAUNCH_BATCH);
    std::unique_ptr<Workspace> ws(new (std::nothrow) Workspace());
    if (!ws) return fail(VOXG_ENOMEM, "host allocation failed");
    int rc = init_workspace(ws.get());
    if (rc) return rc;
    // spill buffer for the launch's lanes at this configuration (the
    // largest stack_lds ring is the configured one)
    const uint32_t gthreads =
        (uint32_t)traversal_grid(b->cfg, b->units, devid) *
        (uint32_t)b->cfg.block;
    const uint32_t k = (uint32_t)b->cfg.stack_lds;
    const uint32_t slots = bound > k ? bound - k : 1u;
    const size_t sb = (size_t)slots * gthreads * kEntryBytes;
    hipError_t e = hipMalloc(&ws->spill, sb);
    if (e == hipSuccess) {
        ws->spill_bytes = sb;
        if (!h.empty())
            e = hipMalloc(&b->d_items, h.size() * sizeof(dev::BatchItem));
    }
    if (e == hipSuccess && !h.empty())
        e = hipMemcpy(b->d_items, h.data(), h.size() * sizeof(dev::BatchItem),
                      hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_workspace(ws.get());
        if (b->d_items) (void)hipFree(b->d_items);
        return fail(VOXG_EIO, hipGetErrorString(e));
    }
    b->ws = ws.release();
    *out = b.release();
    return 0;
}

int voxg_batch_release(voxg_batch *b) {
    if (!b) return 0;
    if (b->ws) {
        // a launch of the batch may still be running on some stream
        HIPCHK(hipDeviceSynchronize());
        free_workspace(b->ws);
        delete b->ws;
    }
    if (b->d_items) HIPCHK(hipFree(b->d_items));
    delete b;
    return 0;
}

int voxg_batch_ray(const voxg_batch *b, const float *ox, const float *oy,
                   const float *oz, const float *dx, const float *dy,
                   const float *dz, uint32_t *voxel_out, uint32_t *leaf_out,
                   float *hx, float *hy, float *hz, hipStream_t stream) {
    if (!b) return fail(VOXG_EINVAL, "batch is NULL");
    int devid = 0;
    HIPCHK(hipGetDevice(&devid));
    if (devid != b->device)
        return fail(VOXG_EINVAL, "batch lives on another device");
    if (b->units == 0) return 0;
    if (!ox || !oy || !oz || !dx || !dy || !dz)
        return fail(VOXG_EINVAL, "ray input pointer is NULL");
    if ((hx || hy || hz) && !(hx && hy && hz))
        return fail(VOXG_EINVAL, "hx, hy, hz must be all set or all NULL");
    dev::MultiWork<dev::SoaSrc, dev::SoaSink> work{
        dev::SoaSrc{ox, oy, oz, dx, dy, dz},
        dev::SoaSink{voxel_out, leaf_out, hx, hy, hz}, b->d_items,
        b->n_items, b->chunk};
    dev::Tree none;
    std::memset(&none, 0, sizeof(none));
    return launch_cfg<decltype(work), false>(b->cfg, none, b->stack_bound,
                                             work, b->units, stream, nullptr,
                                             true, b->ws);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// serving loop
// ---------------------------------------------------------------------------
namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator over the server's device arena (offsets, 256-byte
// granules).  Admission is by free arena bytes.
class Arena {
  public:
    void init(size_t size) {
        size_ = size;
        free_.clear();
        if (size) free_[0] = size;
        used_ = 0;
    }
    bool alloc(size_t bytes, size_t *off) {
        bytes = (bytes + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                size_t rest = it->second - bytes;
                size_t start = it->first + bytes;
                free_.erase(it);
                if (rest) free_[start] = rest;
                used_ += bytes;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t used() const { return used_; }
    size_t size() const { return size_; }

  private:
    size_t size_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;
};

struct SrvRequest {
    uint64_t id = 0;
    int kind = 0;  // 0 render, 1 balls
    int tree = 0;
    dev::Tree t;   // the tree version current when the request was submitted
    const voxg_tree *th = nullptr;  // and its host handle
    voxg_render_request rr;
    size_t n = 0;
    std::vector<float> balls;  // cx | cy | cz | r, when not staged
    float *staged = nullptr;   // the same in the server's pinned staging
    size_t stage_off = 0;
    uint8_t *ball_out = nullptr;
    size_t bytes = 0, off = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    uint64_t group = 0;
    uint32_t group_size = 0;
    int status = 1;  // 1 pending, 0 ok, < 0 error
};

struct SrvGroup {
    uint64_t id = 0;
    int stream = 0;
    uint32_t stack_bound = 1;  // of the deepest tree, read at admission
    hipEvent_t done = nullptr;
    std::vector<SrvRequest *> reqs;
};

// Item tables of a group are written on the host into one of
// kTableSlots pinned slots of their stream, in turn, and copied to the
// stream's device table ahead of the group's kernels (which the next
// group's copy follows in stream order).  A slot is rewritten only once
// its copy has run; with one slot the scheduler thread waited for each
// stream's previous group to reach its copy, behind whatever that stream
// still had queued, and every stream stalled meanwhile.
constexpr int kTableSlots = 4;
constexpr size_t kTableBytes = 1u << 20;  // item tables of one group

struct SrvStream {
    hipStream_t s = nullptr;
    void *h_table = nullptr;  // kTableSlots pinned slots
    void *d_table = nullptr;
    hipEvent_t table_free[kTableSlots] = {};  // slot's copy has run
    bool table_pending[kTableSlots] = {};
    int slot = 0;
};

// Pinned host staging of ball batches: their coordinates are copied here
// at submission, so that the scheduler thread's copy to the device is
// asynchronous (a copy from pageable memory returns only once the stream
// has reached it).  A batch that does not fit stays in pageable memory.
constexpr size_t kStageBytes = (size_t)64 << 20;

}  // namespace

struct voxg_server {
    int device = 0;
    voxg_server_config cfg;
    Clock::time_point t0;
    std::vector<dev::Tree> trees;
    std::vector<const voxg_tree *> htrees;  // the same versions, host handles
    uint32_t stack_bound = 1;
    void *d_arena = nullptr;
    uint8_t *d_texture = nullptr;
    Arena arena;
    void *h_stage = nullptr;  // pinned staging of ball batches
    Arena stage;
    std::vector<SrvStream> streams;
    int next_stream = 0;
    std::mutex mu;
    std::condition_variable cv;     // wakes the scheduler
    std::condition_variable done_cv;  // wakes waiters
    std::deque<SrvRequest *> pending;
    std::map<uint64_t, SrvRequest *> reqs;
    std::vector<SrvGroup> groups;
    uint64_t next_id = 1, next_group = 1;
    uint32_t inflight = 0;
    bool stop = false;
    bool hold = false;  // test hook: admit nothing while set
    std::thread thr;
    voxg_server_stats st;
    std::string thread_error;

    double now() const {
        return std::chrono::duration<double>(Clock::now() - t0).count();
    }
};

namespace {

// Device bytes of a render request: voxel ids, then hit points and
// colours when they are asked for (fixed offsets 4*np and 16*np).
size_t render_bytes(const voxg_render_request &r, bool color) {
    size_t np = (size_t)r.W * (size_t)r.H;
    if (color && r.color_out) return np * 20 + 256;
    if (r.hit_out) return np * 16 + 256;
    return np * 4 + 256;
}

// Whether a traversal on server stream i flagged a stack overflow (read
// and cleared).
bool srv_stream_overflow(voxg_server *s, int i) {
    std::lock_guard<std::mutex> g(g_mu);
    auto it = g_ws.find({s->device, s->streams[i].s});
    if (it == g_ws.end() || !it->second->ovf_h) return false;
    volatile uint32_t *f = it->second->ovf_h;
    if (!*f) return false;
    *f = 0u;
    return true;
}

// Frames go through the render path of voxg_render, which spreads a
// frame's pixels over the whole device, one launch sequence per frame.
// Only when a group holds at least kSrvShareMin frames smaller than
// kSrvSmallPixels do those share one multi-job launch of one 4x4 block
// per lane (a lane then walks 16 pixels in turn, about 2 ms for a
// 320x240 frame, but one launch serves them all): under load that keeps
// the throughput 10% higher, while alone a small frame takes 0.6 ms
// through the render path (a Poisson stream of 1000 requests/s: p99
// 1.9 ms against 4.6 ms when every small frame took the shared launch).
constexpr size_t kSrvSmallPixels = (size_t)1 << 19;
constexpr size_t kSrvShareMin = 8;

bool srv_small(const SrvRequest *r) {
    return r->kind == 0 &&
           (size_t)r->rr.W * (size_t)r->rr.H < kSrvSmallPixels;
}

// Release a ball batch's pinned staging (server lock held; its copy to
// the device has run or will never run).
void srv_unstage(voxg_server *s, SrvRequest *r) {
    if (r->staged) {
        s->stage.release(r->stage_off, 16 * r->n);
        r->staged = nullptr;
    }
}

// Launch one group on its stream (scheduler thread, no lock held).
int srv_launch(voxg_server *s, SrvGroup &g) {
    SrvStream &ss = s->streams[g.stream];
    const int slot = ss.slot;
    ss.slot = (slot + 1) % kTableSlots;
    if (ss.table_pending[slot]) {
        HIPCHK(hipEventSynchronize(ss.table_free[slot]));
        ss.table_pending[slot] = false;
    }
    char *arena = (char *)s->d_arena;
    std::vector<dev::RenderItem> ritems;
    std::vector<dev::BallItem> bitems;
    std::vector<dev::CamDev> cams;
    uint32_t units = 0, balls = 0;
    bool color = s->d_texture != nullptr;
    // render requests in table order: the frames that share the
    // multi-job launch, then the others (the render path, one at a time)
    size_t smalls = 0;
    for (SrvRequest *r : g.reqs) smalls += srv_small(r) ? 1 : 0;
    const bool share = smalls >= kSrvShareMin;
    std::vector<SrvRequest *> order, big;
    for (SrvRequest *r : g.reqs)
        if (share && srv_small(r)) order.push_back(r);
    const size_t n_small = order.size();
    for (SrvRequest *r : g.reqs)
        if (r->kind == 0 && !(share && srv_small(r))) {
            order.push_back(r);
            big.push_back(r);
        }
    for (SrvRequest *r : g.reqs)
        if (r->kind != 0) order.push_back(r);
    uint32_t small_units = 0;
    for (SrvRequest *r : order) {
        if (r->kind == 0) {
            size_t np = (size_t)r->rr.W * r->rr.H;
            char *base = arena + r->off;
            dev::RenderItem it;
            std::memset(&it, 0, sizeof(it));
            it.t = r->t;
            it.cams = nullptr;  // set after the table layout is known
            it.W = (uint32_t)r->rr.W;
            it.ws = make_fastdiv((uint32_t)r->rr.W / 4u);
            it.bpf = make_fastdiv((uint32_t)(np / 16));
            it.unit_begin = units;
            it.rnd_mode = (int)(r->rr.quality & 3);
            it.merge_mode = (int)(r->rr.quality & 0xc);
            bool want_color = color && r->rr.color_out;
            it.sink.voxel = (uint32_t *)base;
            it.sink.hit = r->rr.hit_out ? (float *)(base + np * 4) : nullptr;
            it.sink.plane = np;
            it.sink.squares =
                want_color ? (uint32_t *)(base + np * 16) : nullptr;
            if (want_color)  // pixels that miss stay 0, as after vox_redraw
                HIPCHK(hipMemsetAsync(base + np * 16, 0, np * 4, ss.s));
            it.sink.sh.tex = s->d_texture;
            it.sink.sh.rshift = s->cfg.rshift;
            it.sink.sh.gshift = s->cfg.gshift;
            it.sink.sh.bshift = s->cfg.bshift;
            it.sink.sh.amask = s->cfg.amask;
            it.sink.sh.vx = it.t.vx;
            it.sink.sh.vy = it.t.vy;
            it.sink.sh.vz = it.t.vz;
            it.sink.sh.ivx = it.t.ivx;
            it.sink.sh.ivy = it.t.ivy;
            it.sink.sh.ivz = it.t.ivz;
            units += it.bpf.d;
            if (ritems.size() < n_small) small_units = units;
            ritems.push_back(it);
            dev::CamDev c;
            std::memcpy(&c, &r->rr.cam, sizeof(c));
            cams.push_back(c);
        } else {
            char *base = arena + r->off;
            size_t n = r->n;
            dev::BallItem it;
            std::memset(&it, 0, sizeof(it));
            it.t = r->t;
            it.cx = (const float *)base;
            it.cy = it.cx + n;
            it.cz = it.cy + n;
            it.r = it.cz + n;
            it.out = (uint8_t *)(base + 16 * n);
            it.begin = balls;
            it.count = (uint32_t)n;
            balls += (uint32_t)n;
            bitems.push_back(it);
            HIPCHK(hipMemcpyAsync(base, r->staged ? r->staged : r->balls.data(),
                                  16 * n, hipMemcpyHostToDevice, ss.s));
        }
    }
    // table layout: [render items][ball items][cameras]
    size_t rb = ritems.size() * sizeof(dev::RenderItem);
    size_t bb = bitems.size() * sizeof(dev::BallItem);
    size_t cb = cams.size() * sizeof(dev::CamDev);
    size_t o_b = (rb + 255) & ~(size_t)255;
    size_t o_c = (o_b + bb + 255) & ~(size_t)255;
    if (o_c + cb > kTableBytes) return fail(VOXG_ERANGE, "group too large");
    char *dt = (char *)ss.d_table;
    for (size_t i = 0; i < ritems.size(); i++)
        ritems[i].cams = (const dev::CamDev *)(dt + o_c) + i;
    char *ht = (char *)ss.h_table + (size_t)slot * kTableBytes;
    if (rb) std::memcpy(ht, ritems.data(), rb);
    if (bb) std::memcpy(ht + o_b, bitems.data(), bb);
    if (cb) std::memcpy(ht + o_c, cams.data(), cb);
    HIPCHK(hipMemcpyAsync(dt, ht, o_c + cb, hipMemcpyHostToDevice, ss.s));
    HIPCHK(hipEventRecord(ss.table_free[slot], ss.s));
    ss.table_pending[slot] = true;
    // Large frames go through the render path of voxg_render (block
    // pre-pass, coherent one-pixel units, the speculated FAST / ADAPTIVE
    // passes), one after the other on the group's stream; the small ones,
    // which would each be a few short launches, share one multi-job launch
    // of one 4x4 block per lane (their units were numbered first).  Both
    // give the reference's results.  (launch_k and the render path grow
    // the stream's buffers here, on the scheduler thread, the only thread
    // that launches on the server's streams.)
    if (small_units) {
        dev::BlockWork<true> work;
        std::memset(&work, 0, sizeof(work));
        work.items = (const dev::RenderItem *)dt;
        work.n_items = (uint32_t)n_small;
        dev::Tree none;
        std::memset(&none, 0, sizeof(none));
        int rc = launch_tree<dev::BlockWork<true>, false>(
            VOXG_LAUNCH_BATCH, none, g.stack_bound, work, small_units, ss.s,
            nullptr);
        if (rc) return rc;
    }
    for (size_t i = n_small; i < ritems.size(); i++) {
        SrvRequest *r = big[i - n_small];
        voxg_render_params p;
        std::memset(&p, 0, sizeof(p));
        p.quality = r->rr.quality;
        const size_t np = (size_t)r->rr.W * (size_t)r->rr.H;
        int rc = render_launch<false>(r->th, 1,
                                      (const voxg_camera *)ritems[i].cams,
                                      r->rr.W, r->rr.H, &p, ritems[i].sink,
                                      np, nullptr, ss.s);
        if (rc) return rc;
    }
    if (balls) {
        dev::BallItem one;
        std::memset(&one, 0, sizeof(one));
        Workspace *ws = nullptr;
        int rc = get_ws(ss.s, 16, &ws);
        if (rc) return rc;
        hipLaunchKernelGGL((dev::ball_kernel<false, true>),
                           dim3((balls + 255) / 256), dim3(256), 0, ss.s, one,
                           (const dev::BallItem *)(dt + o_b),
                           (uint32_t)bitems.size(), balls, ws->ctrl);
        HIPCHK(hipGetLastError());
    }
    // results back to the callers' buffers
    for (SrvRequest *r : g.reqs) {
        char *base = arena + r->off;
        if (r->kind == 0) {
            size_t np = (size_t)r->rr.W * r->rr.H;
            if (r->rr.voxel_out)
                HIPCHK(hipMemcpyAsync(r->rr.voxel_out, base, np * 4,
                                      hipMemcpyDeviceToHost, ss.s));
            if (r->rr.hit_out)
                HIPCHK(hipMemcpyAsync(r->rr.hit_out, base + np * 4, np * 12,
                                      hipMemcpyDeviceToHost, ss.s));
            if (r->rr.color_out && color)
                HIPCHK(hipMemcpyAsync(r->rr.color_out, base + np * 16, np * 4,
                                      hipMemcpyDeviceToHost, ss.s));
        } else {
            HIPCHK(hipMemcpyAsync(r->ball_out, base + 16 * r->n, r->n,
                                  hipMemcpyDeviceToHost, ss.s));
        }
    }
    HIPCHK(hipEventRecord(g.done, ss.s));
    return 0;
}

void srv_loop(voxg_server *s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        // retire finished groups
        for (size_t i = 0; i < s->groups.size();) {
            SrvGroup &g = s->groups[i];
            const hipError_t q = hipEventQuery(g.done);
            if (q != hipErrorNotReady) {
                double t = s->now();
                int status = 0;
                if (q != hipSuccess) {
                    // the group's work failed: make sure nothing of it still
                    // runs before its memory is handed out again
                    (void)hipStreamSynchronize(s->streams[g.stream].s);
                    status = VOXG_EIO;
                    s->thread_error = std::string("group failed: ") +
                                      hipGetErrorString(q);
                } else if (srv_stream_overflow(s, g.stream)) {
                    status = VOXG_EIO;
                    s->thread_error = "a traversal exceeded its stack bound";
                }
                for (SrvRequest *r : g.reqs) {
                    r->t_done = t;
                    r->status = status;
                    s->arena.release(r->off, r->bytes);
                    srv_unstage(s, r);
                    s->inflight--;
                    s->st.completed++;
                }
                (void)hipEventDestroy(g.done);
                s->groups.erase(s->groups.begin() + i);
                s->done_cv.notify_all();
            } else {
                i++;
            }
        }
        if (s->stop && s->pending.empty() && s->groups.empty()) break;
        // admit in arrival order while slots and memory allow
        SrvGroup g;
        while (!s->hold && !s->pending.empty() &&
               s->inflight < (uint32_t)s->cfg.max_inflight) {
            SrvRequest *r = s->pending.front();
            size_t off;
            if (!s->arena.alloc(r->bytes, &off)) break;
            s->pending.pop_front();
            r->off = off;
            r->t_admit = s->now();
            s->inflight++;
            g.reqs.push_back(r);
        }
        if (!g.reqs.empty()) {
            g.id = s->next_group++;
            g.stack_bound = s->stack_bound;
            g.stream = s->next_stream;
            s->next_stream = (s->next_stream + 1) % (int)s->streams.size();
            for (SrvRequest *r : g.reqs) {
                r->group = g.id;
                r->group_size = (uint32_t)g.reqs.size();
            }
            s->st.groups++;
            s->st.peak_inflight =
                std::max<uint64_t>(s->st.peak_inflight, s->inflight);
            s->st.peak_bytes =
                std::max<uint64_t>(s->st.peak_bytes, s->arena.used());
            lk.unlock();
            int rc = (hipEventCreateWithFlags(&g.done,
                                              hipEventDisableTiming) ==
                      hipSuccess)
                         ? srv_launch(s, g)
                         : VOXG_EIO;
            if (rc) {
                // part of the group may be queued already (copies,
                // kernels): let it drain before its memory is reused
                (void)hipStreamSynchronize(s->streams[g.stream].s);
            }
            lk.lock();
            if (rc) {
                s->thread_error = g_err;
                for (SrvRequest *r : g.reqs) {
                    r->status = rc;
                    r->t_done = s->now();
                    s->arena.release(r->off, r->bytes);
                    srv_unstage(s, r);
                    s->inflight--;
                }
                if (g.done) (void)hipEventDestroy(g.done);
                s->done_cv.notify_all();
            } else {
                s->groups.push_back(std::move(g));
                s->st.peak_groups_inflight = std::max<uint64_t>(
                    s->st.peak_groups_inflight, s->groups.size());
            }
            continue;
        }
        if (s->groups.empty()) {
            s->cv.wait(lk, [s] {
                return s->stop || (!s->hold && !s->pending.empty());
            });
        } else {
            s->cv.wait_for(lk, std::chrono::microseconds(20));
        }
    }
}

}  // namespace

extern "C" {

double voxg_server_now(const voxg_server *s) { return s ? s->now() : 0.0; }

int voxg_server_create(const voxg_server_config *cfg, voxg_server **out) {
    if (!out) return fail(VOXG_EINVAL, "out is NULL");
    *out = nullptr;
    std::unique_ptr<voxg_server> s(new (std::nothrow) voxg_server());
    if (!s) return fail(VOXG_ENOMEM, "host allocation failed");
    std::memset(&s->cfg, 0, sizeof(s->cfg));
    if (cfg) s->cfg = *cfg;
    if (s->cfg.max_inflight <= 0) s->cfg.max_inflight = 256;
    if (s->cfg.streams <= 0) s->cfg.streams = 4;
    if (s->cfg.streams > 16) return fail(VOXG_EINVAL, "at most 16 streams");
    std::memset(&s->st, 0, sizeof(s->st));
    HIPCHK(hipGetDevice(&s->device));
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    size_t budget = s->cfg.hbm_budget ? s->cfg.hbm_budget : fr / 2;
    if (budget > fr) return fail(VOXG_ENOMEM, "budget exceeds free HBM");
    budget &= ~(size_t)255;
    if (budget < 4096) return fail(VOXG_EINVAL, "budget too small");
    HIPCHK(hipMalloc(&s->d_arena, budget));
    s->arena.init(budget);
    s->st.budget = budget;
    if (s->cfg.texture) {
        HIPCHK(hipMalloc(&s->d_texture, 32 * 32 * 32));
        HIPCHK(hipMemcpy(s->d_texture, s->cfg.texture, 32 * 32 * 32,
                         hipMemcpyHostToDevice));
    }
    HIPCHK(hipHostMalloc(&s->h_stage, kStageBytes, hipHostMallocDefault));
    s->stage.init(kStageBytes);
    s->streams.resize(s->cfg.streams);
    for (auto &ss : s->streams) {
        HIPCHK(hipStreamCreateWithFlags(&ss.s, hipStreamNonBlocking));
        HIPCHK(hipHostMalloc(&ss.h_table, kTableSlots * kTableBytes,
                             hipHostMallocDefault));
        HIPCHK(hipMalloc(&ss.d_table, kTableBytes));
        for (int k = 0; k < kTableSlots; k++)
            HIPCHK(hipEventCreateWithFlags(&ss.table_free[k],
                                           hipEventDisableTiming));
        Workspace *ws = nullptr;
        int rc = get_ws(ss.s, 16, &ws);
        if (rc) return rc;
    }
    s->t0 = Clock::now();
    voxg_server *raw = s.release();
    raw->thr = std::thread(srv_loop, raw);
    *out = raw;
    return 0;
}

// Requests submitted from now on use tree t under this id; requests
// submitted before keep the version they were submitted with.  The
// streams' spill buffers are not touched here (the scheduler thread may be
// launching on them): a deeper tree makes the next launch grow them.
int voxg_server_replace_tree(voxg_server *s, int id, const voxg_tree *t) {
    if (!s) return fail(VOXG_EINVAL, "server is NULL");
    int rc = check_tree(t);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(s->mu);
    if (id < 0 || id >= (int)s->trees.size())
        return fail(VOXG_EINVAL, "unknown tree id");
    s->stack_bound = std::max(s->stack_bound, t->stack_bound);
    s->trees[id] = dev_tree(t);
    s->htrees[id] = t;
    return 0;
}

// The streams' spill buffers are grown by the scheduler thread when it
// launches a group of a deeper tree, never from the caller's thread.
int voxg_server_add_tree(voxg_server *s, const voxg_tree *t, int *id) {
    if (!s || !id) return fail(VOXG_EINVAL, "NULL argument");
    int rc = check_tree(t);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(s->mu);
    s->stack_bound = std::max(s->stack_bound, t->stack_bound);
    s->trees.push_back(dev_tree(t));
    s->htrees.push_back(t);
    *id = (int)s->trees.size() - 1;
    return 0;
}

static int srv_submit(voxg_server *s, SrvRequest *r, uint64_t *ticket) {
    std::lock_guard<std::mutex> g(s->mu);
    if (r->tree < 0 || r->tree >= (int)s->trees.size()) {
        srv_unstage(s, r);
        delete r;
        return fail(VOXG_EINVAL, "unknown tree id");
    }
    r->t = s->trees[r->tree];
    r->th = s->htrees[r->tree];
    if (r->bytes > s->arena.size()) {
        s->st.rejected++;
        srv_unstage(s, r);
        delete r;
        return fail(VOXG_ENOMEM, "request larger than the server budget");
    }
    r->id = s->next_id++;
    r->t_submit = s->now();
    s->reqs[r->id] = r;
    s->pending.push_back(r);
    s->st.submitted++;
    *ticket = r->id;
    s->cv.notify_one();
    return 0;
}

int voxg_server_submit_render(voxg_server *s, const voxg_render_request *rq,
                              uint64_t *ticket) {
    if (!s || !rq || !ticket) return fail(VOXG_EINVAL, "NULL argument");
    if (rq->W <= 0 || rq->H <= 0 || (rq->W & 0xf) || (rq->H & 0x3))
        return fail(VOXG_EINVAL,
                    "W must be a positive multiple of 16 and H of 4");
    unsigned q = rq->quality;
    if ((q & 0xc) && (q & 3) != VOXG_QUALITY_ADAPTIVE)
        return fail(VOXG_EINVAL, "ray merging needs the adaptive mode");
    if ((q & 3) > 2 || (q & 0xc) > 8 || (q & ~0xfu))
        return fail(VOXG_EINVAL, "reserved quality bits set");
    if (rq->color_out && !s->d_texture)
        return fail(VOXG_EINVAL, "colour output needs a server texture");
    if ((size_t)rq->W * rq->H >= 0x40000000ull || rq->W >= 65536 ||
        rq->H >= 65536)
        return fail(VOXG_ERANGE, "frame too large");
    SrvRequest *r = new (std::nothrow) SrvRequest();
    if (!r) return fail(VOXG_ENOMEM, "host allocation failed");
    r->kind = 0;
    r->tree = rq->tree;
    r->rr = *rq;
    r->bytes = render_bytes(*rq, s->d_texture != nullptr);
    return srv_submit(s, r, ticket);
}

int voxg_server_submit_balls(voxg_server *s, const voxg_ball_request *rq,
                             uint64_t *ticket) {
    if (!s || !rq || !ticket) return fail(VOXG_EINVAL, "NULL argument");
    if (rq->n == 0 || !rq->cx || !rq->cy || !rq->cz || !rq->r || !rq->out)
        return fail(VOXG_EINVAL, "empty or NULL ball request");
    if (rq->n >= 0x10000000ull) return fail(VOXG_ERANGE, "too many balls");
    SrvRequest *r = new (std::nothrow) SrvRequest();
    if (!r) return fail(VOXG_ENOMEM, "host allocation failed");
    r->kind = 1;
    r->tree = rq->tree;
    r->n = rq->n;
    {
        // pinned staging when it has room (the copy into it runs outside
        // the lock: the range is this request's)
        std::lock_guard<std::mutex> g(s->mu);
        size_t off = 0;
        if (s->stage.alloc(16 * rq->n, &off)) {
            r->stage_off = off;
            r->staged = (float *)((char *)s->h_stage + off);
        }
    }
    float *dst = r->staged;
    if (!dst) {
        try {
            r->balls.resize(4 * rq->n);
        } catch (const std::bad_alloc &) {
            delete r;
            return fail(VOXG_ENOMEM, "host allocation failed");
        }
        dst = r->balls.data();
    }
    std::memcpy(dst, rq->cx, rq->n * 4);
    std::memcpy(dst + rq->n, rq->cy, rq->n * 4);
    std::memcpy(dst + 2 * rq->n, rq->cz, rq->n * 4);
    std::memcpy(dst + 3 * rq->n, rq->r, rq->n * 4);
    r->ball_out = rq->out;
    r->bytes = 17 * rq->n + 256;
    return srv_submit(s, r, ticket);
}

int voxg_server_wait(voxg_server *s, uint64_t ticket, double timeout_s,
                     voxg_request_info *info) {
    if (!s) return fail(VOXG_EINVAL, "server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(VOXG_EINVAL, "unknown ticket");
    SrvRequest *r = it->second;
    auto ready = [r] { return r->status != 1; };
    if (timeout_s > 0) {
        if (!s->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s),
                                 ready))
            return fail(VOXG_EIO, "timed out");
    } else {
        s->done_cv.wait(lk, ready);
    }
    if (info) {
        info->status = r->status;
        info->t_submit = r->t_submit;
        info->t_admit = r->t_admit;
        info->t_done = r->t_done;
        info->group = r->group;
        info->group_size = r->group_size;
    }
    int st = r->status;
    s->reqs.erase(it);
    delete r;
    if (st) return fail(st, s->thread_error);
    return 0;
}

int voxg_server_get_stats(voxg_server *s, voxg_server_stats *out) {
    if (!s || !out) return fail(VOXG_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> g(s->mu);
    *out = s->st;
    return 0;
}

int voxg_test_server_hold(voxg_server *s, int hold) {
    if (!s) return fail(VOXG_EINVAL, "server is NULL");
    std::lock_guard<std::mutex> g(s->mu);
    s->hold = hold != 0;
    s->cv.notify_all();
    return 0;
}

int voxg_server_destroy(voxg_server *s) {
    if (!s) return 0;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->hold = false;
        s->stop = true;
        s->cv.notify_all();
    }
    if (s->thr.joinable()) s->thr.join();
    for (auto &kv : s->reqs) delete kv.second;
    for (auto &ss : s->streams) {
        if (ss.s) (void)hipStreamSynchronize(ss.s);
        if (ss.h_table) (void)hipHostFree(ss.h_table);
        if (ss.d_table) (void)hipFree(ss.d_table);
        for (int k = 0; k < kTableSlots; k++)
            if (ss.table_free[k]) (void)hipEventDestroy(ss.table_free[k]);
        if (ss.s) voxg_internal::release_stream(ss.s);
    }
    if (s->h_stage) (void)hipHostFree(s->h_stage);
    if (s->d_arena) (void)hipFree(s->d_arena);
    if (s->d_texture) (void)hipFree(s->d_texture);
    for (auto &ss : s->streams)
        if (ss.s) (void)hipStreamDestroy(ss.s);
    delete s;
    return 0;
}

}  // extern "C"
