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
#include <new>
#include <system_error>
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
    // an allocation failed on a helper thread (rethrown once all are done)
    std::atomic<bool> oom{false};
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
                try {
                    th[idx] = std::thread([this, node, idx, &start, &len, off] {
                        try {
                            node->child[idx] = build(off + start[idx], len[idx]);
                        } catch (const std::bad_alloc &) {
                            oom = true;
                        }
                        spare_threads.fetch_add(1);
                    });
                    spawned[idx] = true;
                } catch (const std::system_error &) {
                    // no thread to be had: this one builds the child
                    spare_threads.fetch_add(1);
                    node->child[idx] = build(off + start[idx], len[idx]);
                }
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
    const void *dots_ptr(N n) const { return set + 4 * n->off; }
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
    const void *dots_ptr(N n) const { return n->data.dots; }
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

// The flattener.  Node structures are scattered over the heap, so a walk
// over them is bound by cache misses, one or two per node.  The
// breadth-first order at each level is the concatenation, in order, of
// the nodes at that level of the subtrees below the first level that
// holds kSubtrees nodes; so the top levels are walked on one thread, and
// the subtrees below them on a thread pool, each into its own per-level
// lists, which are then placed at their offsets.
constexpr size_t kSubtrees = 256;
constexpr unsigned kMaxThreads = 16;

// Runs work(w) on the calling thread (w = 0) and on up to nt - 1 more
// threads; the work draws its items from a shared counter, so fewer
// threads (when one cannot be started) only take longer.  An allocation
// failure in any of them is rethrown here, as the callers expect.
template <class F>
void run_pool(unsigned nt, F &&work) {
    std::atomic<bool> oom{false};
    auto guarded = [&](unsigned w) {
        try {
            work(w);
        } catch (const std::bad_alloc &) {
            oom = true;
        }
    };
    std::vector<std::thread> pool;
    for (unsigned w = 1; w < nt; w++) {
        try {
            pool.emplace_back(guarded, w);
        } catch (const std::system_error &) {
            break;
        }
    }
    guarded(0);
    for (auto &th : pool) th.join();
    if (oom) throw std::bad_alloc();
}

// One node as a subtree walk records it, before its place is known.
template <class N>
struct Loc {
    NodeRec r;      // box, centre, info; ref and escape come later
    uint32_t nk;    // children (inner nodes)
    uint32_t count;
    const void *src;
    const float *dots;  // regular leaf: its dots, 4 floats apart
};

struct Tally {
    uint64_t inner = 0, leaves = 0, dense = 0, overflow = 0;
    uint64_t ndots = 0, dense_ids = 0;
    bool too_big = false;
    void add(const Tally &o) {
        inner += o.inner;
        leaves += o.leaves;
        dense += o.dense;
        overflow += o.overflow;
        ndots += o.ndots;
        dense_ids += o.dense_ids;
        too_big |= o.too_big;
    }
};

// The record of node n (ref and escape left for the caller); counts it.
template <class View>
Loc<typename View::N> record(const View &view, typename View::N n,
                             const float V[3], Tally &t) {
    Loc<typename View::N> l;
    std::memset(&l.r, 0, sizeof(l.r));
    for (int k = 0; k < 3; k++) {
        l.r.bmin[k] = View::bmin(n)[k];
        l.r.bmax[k] = View::bmax(n)[k];
    }
    const uint32_t f = View::flags(n);
    l.r.info = f & (NT_TYPE_MASK | NT_OVERFLOW);
    l.r.ref = kEnd;
    l.nk = 0;
    l.count = View::count(n);
    l.src = View::src(n);
    l.dots = nullptr;
    if ((f & NT_TYPE_MASK) == NT_INNER) {
        for (int k = 0; k < 3; k++) l.r.c[k] = View::center(n)[k];
        uint32_t mask = 0;
        for (int c = 0; c < 8; c++)
            if (!View::null(View::child(n, c))) mask |= 1u << c;
        l.nk = (uint32_t)__builtin_popcount(mask);
        l.r.info |= mask << 8;
        t.inner++;
    } else if ((f & NT_TYPE_MASK) == NT_LEAF) {
        if (l.count > kMaxLeafDots) t.too_big = true;
        l.r.info |= (l.count & kMaxLeafDots) << 8;
        if (l.count) l.dots = view.dot(n, 0);
        t.ndots += l.count;
        t.leaves++;
        if (f & NT_OVERFLOW) t.overflow++;
    } else {
        uint32_t dim[3];
        dense_dims(View::bmin(n), View::bmax(n), V, dim);
        t.dense_ids += (uint64_t)dim[0] * dim[1] * dim[2];
        t.dense++;
    }
    return l;
}

template <class View>
int flatten(const View &view, typename View::N root, const float V[3],
            HostTree &out, std::string &err) {
    using N = typename View::N;
    out = HostTree();
    for (int l = 0; l < 3; l++) out.voxel[l] = V[l];
    if (View::null(root)) return 0;
    Tally tally;

    // 1. the top levels on this thread, breadth first, until the next
    //    level holds kSubtrees nodes (or there is none): records and
    //    escape links of levels 1..lt; the next level's nodes are the
    //    subtree roots, with their escape links
    std::vector<Loc<N>> top;
    std::vector<uint32_t> top_esc{kEnd};
    std::vector<N> front{root}, next;
    std::vector<uint32_t> front_esc{kEnd}, next_esc;
    uint32_t lt = 0;
    for (;;) {
        if (front.size() >= kSubtrees && lt > 0) break;
        // move the front into the top and expand it
        const uint32_t base = (uint32_t)(top.size() + front.size());
        next.clear();
        next_esc.clear();
        for (size_t i = 0; i < front.size(); i++) {
            N n = front[i];
            top.push_back(record(view, n, V, tally));
            if (i > 0 || lt > 0) top_esc.push_back(front_esc[i]);
            const uint32_t nk = top.back().nk;
            uint32_t k = 0;
            for (int c = 0; c < 8 && nk; c++) {
                N ch = View::child(n, c);
                if (View::null(ch)) continue;
                next.push_back(ch);
                const uint32_t g = base + (uint32_t)next.size() - 1u;
                next_esc.push_back(k + 1 < nk ? g + 1u : front_esc[i]);
                k++;
            }
        }
        lt++;
        front.swap(next);
        front_esc.swap(next_esc);
        if (front.empty()) break;
    }
    const size_t T = top.size(), S = front.size();

    // 2. the subtrees, on a pool: per subtree, per level, its nodes'
    //    records in breadth-first order
    std::vector<std::vector<std::vector<Loc<N>>>> sub(S);
    std::vector<Tally> tallies;
    {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const unsigned nt =
            (unsigned)std::min<size_t>(std::min(hw, kMaxThreads), S);
        tallies.resize(std::max(1u, nt));
        std::atomic<size_t> next_sub{0};
        auto work = [&](unsigned w) {
            Tally &t = tallies[w];
            std::vector<N> lv, nx;
            for (;;) {
                const size_t s = next_sub.fetch_add(1);
                if (s >= S) break;
                auto &levels = sub[s];
                lv.assign(1, front[s]);
                while (!lv.empty()) {
                    levels.emplace_back();
                    auto &rec = levels.back();
                    rec.reserve(lv.size());
                    nx.clear();
                    for (N n : lv) {
                        rec.push_back(record(view, n, V, t));
                        if (rec.back().nk)
                            for (int c = 0; c < 8; c++) {
                                N ch = View::child(n, c);
                                if (!View::null(ch)) nx.push_back(ch);
                            }
                    }
                    lv.swap(nx);
                }
            }
        };
        run_pool(nt, work);
    }
    for (const Tally &t : tallies) tally.add(t);
    if (tally.too_big) {
        err = "leaf with more than 2^24-1 voxels";
        return -34;
    }
    if (tally.ndots + tally.dense_ids >= 0xffffffffull) {
        err = "more than 2^32-1 voxels in one tree";
        return -34;
    }

    // 3. places: level d of the subtrees is level lt + 1 + d of the tree;
    //    a subtree's nodes at a level follow those of the subtrees before
    size_t depth_sub = 0;
    for (const auto &levels : sub) depth_sub = std::max(depth_sub, levels.size());
    std::vector<std::vector<uint64_t>> off(S, std::vector<uint64_t>(depth_sub));
    uint64_t total = T;
    for (size_t d = 0; d < depth_sub; d++)
        for (size_t s = 0; s < S; s++) {
            off[s][d] = total;
            if (d < sub[s].size()) total += sub[s][d].size();
        }
    if (total >= 0xfffffff0ull) {
        err = "too many nodes";
        return -34;
    }
    out.depth = lt + (uint32_t)depth_sub;
    out.nodes.resize(total);
    out.counts.resize(total);
    out.src.resize(total);
    std::vector<const float *> dots(total, nullptr);
    auto place = [&](const Loc<N> &l, uint64_t g, uint32_t ref, uint32_t esc) {
        NodeRec r = l.r;
        r.ref = ref;
        r.escape = esc;
        out.nodes[g] = r;
        out.counts[g] = l.count;
        out.src[g] = l.src;
        dots[g] = l.dots;
    };
    // the top: children are numbered in order from 1 (breadth first)
    {
        uint32_t next_child = 1;
        for (size_t i = 0; i < T; i++) {
            const uint32_t ref = top[i].nk ? next_child : kEnd;
            next_child += top[i].nk;
            place(top[i], i, ref, top_esc[i]);
        }
    }
    // the subtrees, in parallel again
    {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const unsigned nt =
            (unsigned)std::min<size_t>(std::min(hw, kMaxThreads), S);
        std::atomic<size_t> next_sub{0};
        auto work = [&](unsigned) {
            std::vector<uint32_t> esc, nesc;
            for (;;) {
                const size_t s = next_sub.fetch_add(1);
                if (s >= S) break;
                const auto &levels = sub[s];
                esc.assign(1, front_esc[s]);
                for (size_t d = 0; d < levels.size(); d++) {
                    nesc.clear();
                    uint64_t child = d + 1 < levels.size() ? off[s][d + 1] : 0;
                    for (size_t p = 0; p < levels[d].size(); p++) {
                        const Loc<N> &l = levels[d][p];
                        const uint32_t ref = l.nk ? (uint32_t)child : kEnd;
                        for (uint32_t k = 0; k < l.nk; k++)
                            nesc.push_back(k + 1 < l.nk
                                               ? (uint32_t)(child + k + 1)
                                               : esc[p]);
                        child += l.nk;
                        place(l, off[s][d] + p, ref, esc[p]);
                    }
                    esc.swap(nesc);
                }
            }
        };
        run_pool(nt, work);
    }
    out.inner = tally.inner;
    out.leaves = tally.leaves;
    out.dense = tally.dense;
    out.overflow = tally.overflow;

    // 4. leaf refs and voxel ids in canonical order: a depth-first preorder
    //    over the records (children are [ref, ref + popcount)); dots first,
    //    then dense voxels.  The leaves' dots are copied afterwards, on the
    //    pool (each leaf's source is a cache miss of its own)
    const uint64_t ndots = tally.ndots;
    out.dots.resize(3 * ndots);
    uint64_t dot_next = 0, dense_next = ndots;
    std::vector<uint32_t> st{0}, leaves;
    leaves.reserve(tally.leaves);
    while (!st.empty()) {
        uint32_t i = st.back();
        st.pop_back();
        NodeRec &r = out.nodes[i];
        uint32_t t = r.info & NT_TYPE_MASK;
        if (t == NT_LEAF) {
            r.ref = (uint32_t)dot_next;
            dot_next += r.info >> 8;
            leaves.push_back(i);
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
    {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const size_t chunk = 1024;
        const unsigned nt = (unsigned)std::min<size_t>(
            std::min(hw, kMaxThreads), (leaves.size() + chunk - 1) / chunk);
        std::atomic<size_t> next_leaf{0};
        auto work = [&](unsigned) {
            for (;;) {
                const size_t b = next_leaf.fetch_add(chunk);
                if (b >= leaves.size()) break;
                const size_t e = std::min(leaves.size(), b + chunk);
                for (size_t j = b; j < e; j++) {
                    const NodeRec &r = out.nodes[leaves[j]];
                    const uint32_t cnt = r.info >> 8;
                    const float *p = dots[leaves[j]];
                    float *q = out.dots.data() + 3 * (size_t)r.ref;
                    for (uint32_t k = 0; k < cnt; k++) {
                        q[3 * k + 0] = p[4 * k + 0];
                        q[3 * k + 1] = p[4 * k + 1];
                        q[3 * k + 2] = p[4 * k + 2];
                    }
                }
            }
        };
        run_pool(nt, work);
    }
    out.ndots = ndots;
    out.voxel_ids = ndots + tally.dense_ids;
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
    if (b.oom) throw std::bad_alloc();
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

static int build_reference_(const float *xyz, size_t n, size_t stride,
                            const float V[3], struct vox_node **root,
                            std::string &err);

// The edit API calls this without a handler of its own: an allocation
// failure becomes a status here.
int build_reference(const float *xyz, size_t n, size_t stride, const float V[3],
                    struct vox_node **root, std::string &err) {
    try {
        return build_reference_(xyz, n, stride, V, root, err);
    } catch (const std::bad_alloc &) {
        err = "host allocation failed";
        return -12;
    }
}

static int build_reference_(const float *xyz, size_t n, size_t stride,
                            const float V[3], struct vox_node **root,
                            std::string &err) {
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
    if (b.oom) throw std::bad_alloc();
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
    // dense leaves in id order for the lookup (a leaf whose truncated
    // dimensions hold no voxel, possible only when coordinates are too
    // coarse for the voxel size, has no ids and is left out)
    std::vector<uint32_t> dense;
    for (uint32_t i = 0; i < t.nodes.size(); i++) {
        if ((t.nodes[i].info & NT_TYPE_MASK) != NT_DENSE) continue;
        uint32_t dim[3];
        dense_dims(t.nodes[i].bmin, t.nodes[i].bmax, t.voxel, dim);
        if (dim[0] && dim[1] && dim[2]) dense.push_back(i);
    }
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
        if (it == dense.begin()) {
            xyz[3 * k] = xyz[3 * k + 1] = xyz[3 * k + 2] = NAN;
            continue;
        }
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
