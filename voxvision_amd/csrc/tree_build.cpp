// Host side of libvoxgpu: octree construction and flattening.
//
// build_tree() reproduces vox_make_tree (src/voxtrees/tree.c:249-338) node
// for node -- same bounding boxes, centres, leaf dot order, dense and
// overflow leaves -- but runs independent subtrees on several threads and
// keeps the tree in one working array instead of per-node allocations.
// flatten() turns either that tree or a tree in the reference's own layout
// (struct vox_node, src/voxtrees/tree.h:28-38) into the breadth-first
// device records of voxg_layout.h.
//
// Floating point: this file is compiled with -ffp-contract=off; the one
// place where the reference's clang build fuses a multiply-add on this
// path (dense_set_p, src/voxtrees/geom.c:166) uses fmaf explicitly.
#include "tree_build.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

namespace voxg {

namespace {

inline float smin(float a, float b) { return a < b ? a : b; }  // MINPS
inline float smax(float a, float b) { return a > b ? a : b; }  // MAXPS

struct BNode {
    float bmin[3], bmax[3], c[3];
    uint32_t flags;  // NT_INNER / NT_LEAF (| NT_OVERFLOW) / NT_DENSE
    uint32_t n;
    size_t off;  // leaf: first dot in the working array
    BNode *child[8];
};

// Lock-protected pool of node blocks; each thread bump-allocates from its
// current block.  A thread's cached block is tied to the pool through a
// process-unique pool id (a new pool may reuse a dead pool's address).
class NodePool {
  public:
    NodePool() : id_(next_id().fetch_add(1) + 1) {}

    BNode *alloc() {
        thread_local Block *cur = nullptr;
        thread_local size_t used = 0;
        thread_local uint64_t owner = 0;
        if (owner != id_ || cur == nullptr || used == kBlock) {
            std::lock_guard<std::mutex> g(mu_);
            blocks_.emplace_back();
            cur = &blocks_.back();
            used = 0;
            owner = id_;
        }
        BNode *n = &cur->nodes[used++];
        std::memset(n, 0, sizeof(*n));
        return n;
    }

  private:
    static std::atomic<uint64_t> &next_id() {
        static std::atomic<uint64_t> v{0};
        return v;
    }
    static constexpr size_t kBlock = 4096;
    struct Block {
        BNode nodes[kBlock];
    };
    const uint64_t id_;
    std::mutex mu_;
    std::deque<Block> blocks_;
};

struct Builder {
    float V[3];
    float *set;  // n x 4 working array (x, y, z, pad)
    NodePool pool;
    std::atomic<int> spare_threads{0};
    std::atomic<bool> too_big{false};

    // tree.c:57-72
    void bbox(size_t off, size_t n, float mn[3], float mx[3]) const {
        const float *p = set + 4 * off;
        for (int l = 0; l < 3; l++) mn[l] = mx[l] = p[l];
        for (size_t i = 1; i < n; i++) {
            const float *q = p + 4 * i;
            for (int l = 0; l < 3; l++) {
                mn[l] = smin(mn[l], q[l]);
                mx[l] = smax(mx[l], q[l]);
            }
        }
        for (int l = 0; l < 3; l++) mx[l] = mx[l] + V[l];
    }

    // geom.c:157-167
    bool dense(const float mn[3], const float mx[3], size_t n) const {
        float s0 = mx[0] - mn[0], s1 = mx[1] - mn[1], s2 = mx[2] - mn[2];
        float bbv = (s0 * s1) * s2;
        float vv = (V[0] * V[1]) * V[2];
        return std::fabs(std::fmaf((float)n, vv, -bbv)) < vv;
    }

    // tree.c:31-55: sequential float sum in working-array order
    void center(size_t off, size_t n, const float mn[3], float c[3]) const {
        const float *p = set + 4 * off;
        float len = (float)n;
        float s[3] = {0.0f, 0.0f, 0.0f};
        for (size_t i = 0; i < n; i++)
            for (int l = 0; l < 3; l++) s[l] = s[l] + (p[4 * i + l] - mn[l]);
        for (int l = 0; l < 3; l++) {
            float m = s[l] / len;
            m = m + mn[l];
            c[l] = std::ceil(m / V[l]) * V[l];
        }
    }

    BNode *build(size_t off, size_t n) {
        if (n == 0) return nullptr;
        BNode *node = pool.alloc();
        float mn[3], mx[3];
        bbox(off, n, mn, mx);
        std::memcpy(node->bmin, mn, sizeof(mn));
        std::memcpy(node->bmax, mx, sizeof(mx));
        node->n = (uint32_t)n;
        node->off = off;
        if (dense(mn, mx, n)) {
            node->flags = NT_DENSE;
            return node;
        }
        if (n <= 7) {
            node->flags = NT_LEAF;
            return node;
        }
        node->flags = NT_INNER;
        center(off, n, mn, node->c);

        // tree.c:77-102: eight in-place swap partitions over [offset, n),
        // one per subspace; the subspace code of a dot is the sign bits of
        // dot - centre.  Codes are computed once and swapped with the dots.
        std::vector<uint8_t> code(n);
        float *p = set + 4 * off;
        for (size_t i = 0; i < n; i++) {
            int b = 0;
            for (int l = 0; l < 3; l++)
                b |= (std::signbit(p[4 * i + l] - node->c[l]) ? 1 : 0) << l;
            code[i] = (uint8_t)b;
        }
        size_t start[8], len[8], offset = 0;
        for (int idx = 0; idx < 8; idx++) {
            size_t counter = offset;
            for (size_t i = offset; i < n; i++) {
                if (code[i] == idx) {
                    if (i != counter) {
                        float t[4];
                        std::memcpy(t, p + 4 * counter, sizeof(t));
                        std::memcpy(p + 4 * counter, p + 4 * i, sizeof(t));
                        std::memcpy(p + 4 * i, t, sizeof(t));
                        std::swap(code[counter], code[i]);
                    }
                    counter++;
                }
            }
            if (offset == 0 && counter == n) {
                // tree.c:281-305: no split possible -> overflow leaf
                node->flags = NT_LEAF | NT_OVERFLOW;
                if (n > kMaxLeafDots) too_big = true;
                return node;
            }
            start[idx] = offset;
            len[idx] = counter - offset;
            offset = counter;
        }
        std::vector<uint8_t>().swap(code);

        std::thread th[8];
        bool spawned[8] = {false};
        for (int idx = 0; idx < 8; idx++) {
            if (len[idx] >= (1u << 15) && spare_threads.fetch_sub(1) > 0) {
                spawned[idx] = true;
                th[idx] = std::thread([this, node, idx, &start, &len, off] {
                    node->child[idx] = build(off + start[idx], len[idx]);
                    spare_threads.fetch_add(1);
                });
            } else {
                if (len[idx] >= (1u << 15)) spare_threads.fetch_add(1);
                node->child[idx] = build(off + start[idx], len[idx]);
            }
        }
        for (int idx = 0; idx < 8; idx++)
            if (spawned[idx]) th[idx].join();
        return node;
    }
};

// Node views for the generic flattener.
struct BView {
    const float *set;
    using N = const BNode *;
    static bool null(N n) { return n == nullptr; }
    static uint32_t flags(N n) { return n->flags; }
    static uint32_t count(N n) { return n->n; }
    static const float *bmin(N n) { return n->bmin; }
    static const float *bmax(N n) { return n->bmax; }
    static const float *center(N n) { return n->c; }
    static N child(N n, int i) { return n->child[i]; }
    const float *dot(N n, uint32_t k) const { return set + 4 * (n->off + k); }
    static const void *src(N) { return nullptr; }
};

struct RView {
    using N = const struct vox_node *;
    static bool null(N n) { return n == nullptr; }
    static uint32_t flags(N n) {
        if (n->flags & VOX_DENSE_LEAF) return NT_DENSE;
        if (n->flags & VOX_LEAF)
            return NT_LEAF | ((n->flags & VOX_OVERFLOW) ? NT_OVERFLOW : 0u);
        return NT_INNER;
    }
    static uint32_t count(N n) { return n->dots_num; }
    static const float *bmin(N n) { return n->bounding_box.min; }
    static const float *bmax(N n) { return n->bounding_box.max; }
    static const float *center(N n) { return n->data.inner.center; }
    static N child(N n, int i) { return n->data.inner.children[i]; }
    const float *dot(N n, uint32_t k) const { return n->data.dots[k]; }
    static const void *src(N n) { return n; }
};

// geom.c:192-196 get_dimensions: float division, truncated
inline void dense_dims(const float *mn, const float *mx, const float V[3],
                       uint32_t dim[3]) {
    for (int l = 0; l < 3; l++) {
        float f = (mx[l] - mn[l]) / V[l];
        dim[l] = f > 0.0f ? (uint32_t)f : 0u;
    }
}

template <class View>
int flatten(const View &view, typename View::N root, const float V[3],
            HostTree &out, std::string &err) {
    using N = typename View::N;
    out = HostTree();
    for (int l = 0; l < 3; l++) out.voxel[l] = V[l];
    if (View::null(root)) return 0;

    // 1. depth-first preorder: voxel ids (dots first, then dense voxels)
    struct Acc {
        uint64_t dots = 0, dense_ids = 0;
    } acc;
    {
        std::vector<std::pair<N, uint32_t>> st{{root, 1}};
        while (!st.empty()) {
            N n = st.back().first;
            uint32_t d = st.back().second;
            st.pop_back();
            out.depth = std::max(out.depth, d);
            uint32_t f = View::flags(n);
            if ((f & NT_TYPE_MASK) == NT_LEAF) {
                if (View::count(n) > kMaxLeafDots) {
                    err = "leaf with more than 2^24-1 voxels";
                    return -34;
                }
                acc.dots += View::count(n);
            } else if ((f & NT_TYPE_MASK) == NT_DENSE) {
                uint32_t dim[3];
                dense_dims(View::bmin(n), View::bmax(n), V, dim);
                acc.dense_ids += (uint64_t)dim[0] * dim[1] * dim[2];
            } else {
                for (int i = 7; i >= 0; i--)
                    if (!View::null(View::child(n, i)))
                        st.push_back({View::child(n, i), d + 1});
            }
        }
    }
    if (acc.dots + acc.dense_ids >= 0xffffffffull) {
        err = "more than 2^32-1 voxels in one tree";
        return -34;
    }

    // 2. breadth-first records; children of a node are contiguous
    std::vector<N> q;
    q.push_back(root);
    out.nodes.reserve(1024);
    std::vector<uint32_t> esc(1, kEnd);
    for (size_t i = 0; i < q.size(); i++) {
        N n = q[i];
        NodeRec r;
        std::memset(&r, 0, sizeof(r));
        for (int l = 0; l < 3; l++) {
            r.bmin[l] = View::bmin(n)[l];
            r.bmax[l] = View::bmax(n)[l];
        }
        uint32_t f = View::flags(n);
        r.info = f & (NT_TYPE_MASK | NT_OVERFLOW);
        r.escape = esc[i];
        if ((f & NT_TYPE_MASK) == NT_INNER) {
            for (int l = 0; l < 3; l++) r.c[l] = View::center(n)[l];
            uint32_t mask = 0, first = (uint32_t)q.size(), k = 0;
            for (int c = 0; c < 8; c++)
                if (!View::null(View::child(n, c))) mask |= 1u << c;
            uint32_t nk = (uint32_t)__builtin_popcount(mask);
            for (int c = 0; c < 8; c++) {
                N ch = View::child(n, c);
                if (View::null(ch)) continue;
                q.push_back(ch);
                esc.push_back(k + 1 < nk ? first + k + 1 : esc[i]);
                k++;
            }
            r.ref = nk ? first : kEnd;
            r.info |= mask << 8;
            out.inner++;
        } else if ((f & NT_TYPE_MASK) == NT_LEAF) {
            r.info |= View::count(n) << 8;
            out.leaves++;
            if (f & NT_OVERFLOW) out.overflow++;
        } else {
            out.dense++;
        }
        out.nodes.push_back(r);
        out.counts.push_back(View::count(n));
        out.src.push_back(View::src(n));
        if (q.size() >= 0xfffffff0ull) {
            err = "too many nodes";
            return -34;
        }
    }

    // 3. leaf refs in canonical id order (second preorder walk)
    std::vector<uint32_t> index_of;  // preorder position -> record index
    {
        // map node -> record index through the BFS order
        out.dots.resize(3 * acc.dots);
        uint64_t dot_next = 0, dense_next = acc.dots;
        // preorder over records: children are [ref, ref+popcount)
        std::vector<uint32_t> st{0};
        while (!st.empty()) {
            uint32_t i = st.back();
            st.pop_back();
            NodeRec &r = out.nodes[i];
            uint32_t t = r.info & NT_TYPE_MASK;
            if (t == NT_LEAF) {
                N n = q[i];
                r.ref = (uint32_t)dot_next;
                uint32_t cnt = r.info >> 8;
                for (uint32_t k = 0; k < cnt; k++) {
                    const float *p = view.dot(n, k);
                    out.dots[3 * (dot_next + k) + 0] = p[0];
                    out.dots[3 * (dot_next + k) + 1] = p[1];
                    out.dots[3 * (dot_next + k) + 2] = p[2];
                }
                dot_next += cnt;
            } else if (t == NT_DENSE) {
                uint32_t dim[3];
                dense_dims(r.bmin, r.bmax, V, dim);
                r.ref = (uint32_t)dense_next;
                dense_next += (uint64_t)dim[0] * dim[1] * dim[2];
                // the device reads the dimensions from the centre slot
                std::memcpy(r.c, dim, sizeof(dim));
            } else {
                uint32_t nk = (uint32_t)__builtin_popcount((r.info >> 8) & 0xff);
                for (int k = (int)nk - 1; k >= 0; k--) st.push_back(r.ref + k);
            }
        }
    }
    out.ndots = acc.dots;
    out.voxel_ids = acc.dots + acc.dense_ids;
    return 0;
}

}  // namespace

int build_tree(const float *xyz, size_t n, size_t stride, const float V[3],
               HostTree &out, std::string &err) {
    if (n == 0) {
        out = HostTree();
        for (int l = 0; l < 3; l++) out.voxel[l] = V[l];
        return 0;
    }
    if (n >= 0xffffffffull) {
        err = "more than 2^32-1 voxels";
        return -34;
    }
    Builder b;
    for (int l = 0; l < 3; l++) b.V[l] = V[l];
    std::vector<float> work(4 * n);
    for (size_t i = 0; i < n; i++) {
        work[4 * i + 0] = xyz[stride * i + 0];
        work[4 * i + 1] = xyz[stride * i + 1];
        work[4 * i + 2] = xyz[stride * i + 2];
        work[4 * i + 3] = 0.0f;
    }
    b.set = work.data();
    int hw = (int)std::thread::hardware_concurrency();
    b.spare_threads = std::max(0, std::min(hw, 64) - 1);
    BNode *root = b.build(0, n);
    if (b.too_big) {
        err = "leaf with more than 2^24-1 voxels";
        return -34;
    }
    BView view{work.data()};
    return flatten(view, root, V, out, err);
}

// A BNode subtree as reference-layout nodes (struct vox_node): the same
// boxes, centres, flags and dot order; regular leaves get room for 7 dots
// (tree.c:267), overflow leaves exactly their own.
static struct vox_node *to_reference(const BNode *b, const float *set) {
    if (!b) return nullptr;
    auto *n = static_cast<struct vox_node *>(
        std::aligned_alloc(16, sizeof(struct vox_node)));
    if (!n) return nullptr;
    std::memset(n, 0, sizeof(*n));
    for (int l = 0; l < 3; l++) {
        n->bounding_box.min[l] = b->bmin[l];
        n->bounding_box.max[l] = b->bmax[l];
    }
    n->dots_num = b->n;
    if (b->flags == NT_DENSE) {
        n->flags = VOX_DENSE_LEAF;
    } else if (b->flags & NT_LEAF) {
        n->flags = VOX_LEAF | ((b->flags & NT_OVERFLOW) ? VOX_OVERFLOW : 0u);
        size_t cap = std::max<size_t>(b->n, 7);
        n->data.dots = static_cast<vox_dot *>(
            std::aligned_alloc(16, cap * sizeof(vox_dot)));
        if (!n->data.dots) {
            std::free(n);
            return nullptr;
        }
        std::memset(n->data.dots, 0, cap * sizeof(vox_dot));
        for (uint32_t k = 0; k < b->n; k++)
            for (int l = 0; l < 3; l++)
                n->data.dots[k][l] = set[4 * (b->off + k) + l];
    } else {
        n->flags = 0;
        for (int l = 0; l < 3; l++) n->data.inner.center[l] = b->c[l];
        for (int i = 0; i < 8; i++) {
            n->data.inner.children[i] = to_reference(b->child[i], set);
            if (b->child[i] && !n->data.inner.children[i]) {
                destroy_reference(n);
                return nullptr;
            }
        }
    }
    return n;
}

void destroy_reference(struct vox_node *t) {
    if (!t) return;
    if (t->flags & VOX_LEAF) {
        std::free(t->data.dots);
    } else if (!(t->flags & VOX_LEAF_MASK)) {
        for (int i = 0; i < 8; i++) destroy_reference(t->data.inner.children[i]);
    }
    std::free(t);
}

int build_reference(const float *xyz, size_t n, size_t stride, const float V[3],
                    struct vox_node **root, std::string &err) {
    *root = nullptr;
    if (n == 0) return 0;
    if (n >= 0xffffffffull) {
        err = "more than 2^32-1 voxels";
        return -34;
    }
    Builder b;
    for (int l = 0; l < 3; l++) b.V[l] = V[l];
    std::vector<float> work(4 * n);
    for (size_t i = 0; i < n; i++) {
        work[4 * i + 0] = xyz[stride * i + 0];
        work[4 * i + 1] = xyz[stride * i + 1];
        work[4 * i + 2] = xyz[stride * i + 2];
        work[4 * i + 3] = 0.0f;
    }
    b.set = work.data();
    int hw = (int)std::thread::hardware_concurrency();
    b.spare_threads = std::max(0, std::min(hw, 64) - 1);
    BNode *broot = b.build(0, n);
    *root = to_reference(broot, work.data());
    if (!*root) {
        err = "host allocation failed";
        return -12;
    }
    return 0;
}

int flatten_reference(const struct vox_node *root, const float V[3],
                      HostTree &out, std::string &err) {
    RView view;
    return flatten(view, root, V, out, err);
}

// Structure hash, walked depth-first over children 0..7 exactly like the
// oracle's hash of a pointer tree, so equal trees hash equally whatever
// produced them.
static uint64_t mix64(uint64_t h, uint64_t v) {
    h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    h *= 0xff51afd7ed558ccdull;
    return h ^ (h >> 33);
}
static uint64_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

static uint64_t hash_rec(const HostTree &t, uint32_t i, uint64_t h,
                         int slot) {
    const NodeRec &r = t.nodes[i];
    uint32_t type = r.info & NT_TYPE_MASK;
    h = mix64(h, 0x1000ull + (uint64_t)slot);
    h = mix64(h, r.info & 7u);
    h = mix64(h, t.counts[i]);
    for (int l = 0; l < 3; l++) {
        h = mix64(h, fbits(r.bmin[l]));
        h = mix64(h, fbits(r.bmax[l]));
    }
    if (type == NT_DENSE) return h;
    if (type == NT_LEAF) {
        uint32_t cnt = r.info >> 8;
        for (uint32_t k = 0; k < cnt; k++)
            for (int l = 0; l < 3; l++)
                h = mix64(h, fbits(t.dots[3 * ((size_t)r.ref + k) + l]));
        return h;
    }
    for (int l = 0; l < 3; l++) h = mix64(h, fbits(r.c[l]));
    uint32_t mask = (r.info >> 8) & 0xff, k = 0;
    for (int c = 0; c < 8; c++) {
        if (mask & (1u << c)) h = hash_rec(t, r.ref + k++, h, c);
        else h = mix64(h, 0xdead0000ull + (uint64_t)c);
    }
    return h;
}

uint64_t tree_hash(const HostTree &t) {
    if (t.nodes.empty()) return mix64(0x12345678ull, 0xdead0000ull + 8);
    return hash_rec(t, 0, 0x12345678ull, 8);
}

int voxel_coords(const HostTree &t, size_t n, const uint32_t *ids,
                 float *xyz) {
    // dense leaves in id order for the lookup
    std::vector<uint32_t> dense;
    for (uint32_t i = 0; i < t.nodes.size(); i++)
        if ((t.nodes[i].info & NT_TYPE_MASK) == NT_DENSE) dense.push_back(i);
    std::sort(dense.begin(), dense.end(), [&](uint32_t a, uint32_t b) {
        return t.nodes[a].ref < t.nodes[b].ref;
    });
    for (size_t k = 0; k < n; k++) {
        uint32_t id = ids[k];
        if (id == kEnd || id >= t.voxel_ids) {
            xyz[3 * k] = xyz[3 * k + 1] = xyz[3 * k + 2] = NAN;
            continue;
        }
        if (id < t.ndots) {
            for (int l = 0; l < 3; l++) xyz[3 * k + l] = t.dots[3 * (size_t)id + l];
            continue;
        }
        auto it = std::upper_bound(dense.begin(), dense.end(), id,
                                   [&](uint32_t v, uint32_t nd) {
                                       return v < t.nodes[nd].ref;
                                   });
        const NodeRec &r = t.nodes[*(it - 1)];
        uint32_t dim[3];
        dense_dims(r.bmin, r.bmax, t.voxel, dim);
        uint32_t local = id - r.ref;
        uint32_t iz = local % dim[2];
        uint32_t iy = (local / dim[2]) % dim[1];
        uint32_t ix = local / (dim[2] * dim[1]);
        uint32_t ixyz[3] = {ix, iy, iz};
        for (int l = 0; l < 3; l++)
            xyz[3 * k + l] = r.bmin[l] + (float)ixyz[l] * t.voxel[l];
    }
    return 0;
}

}  // namespace voxg
