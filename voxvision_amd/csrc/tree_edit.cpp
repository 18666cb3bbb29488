// Editable host trees: vox_insert_voxel / vox_delete_voxel
// (src/voxtrees/tree.c:437-835) on a tree in the reference's own node
// layout, and new device versions of it (voxg_edit_commit).
//
// The edits follow the reference step for step -- the same node kinds,
// boxes, centres and dot order come out -- because the device traversal
// must see the tree the reference would have after the same edits.  The
// reference's SSE build decides subspaces in two ways: the edits by the
// sign bits of dot - centre (tree.c:77-81), the membership test by
// dot < centre (geom-sse.c:64-68); both are kept.  Compiled with
// -ffp-contract=off; dense_set_p fuses its multiply-add as clang does.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "tree_build.h"

using namespace voxg;

struct voxg_edit_tree {
    struct vox_node *root = nullptr;
    float V[3] = {1.0f, 1.0f, 1.0f};
    // set when an edit met a dense leaf whose box, divided by the voxel
    // size and truncated (geom.c:184-196), does not give its voxel count:
    // the reference stops there on an assertion (tree.c:389, 530, 671) and
    // its listing of that leaf's voxels is undefined; every later call on
    // the tree fails with this message
    std::string broken;
};

namespace {

constexpr unsigned kMaxDots = 7;  // VOX_MAX_DOTS, tree.h

inline float smin(float a, float b) { return a < b ? a : b; }  // MINPS
inline float smax(float a, float b) { return a > b ? a : b; }  // MAXPS

// n vox_dots in one 16-byte aligned allocation (operator new aligns to 16)
struct DotBuf {
    std::vector<float> f;
    explicit DotBuf(size_t n) : f(4 * (n ? n : 1), 0.0f) {}
    vox_dot *data() { return reinterpret_cast<vox_dot *>(f.data()); }
};

struct Editor {
    float V[3];
    std::string err;
    bool oom = false;
    mutable bool bad_dense = false;  // a dense leaf its box does not describe

    vox_node *alloc_node() {
        auto *n = static_cast<vox_node *>(std::aligned_alloc(16, sizeof(vox_node)));
        if (!n) {
            oom = true;
            return nullptr;
        }
        std::memset(n, 0, sizeof(*n));
        return n;
    }

    // geom.c:192-196
    void dims(const vox_box &b, size_t d[3]) const {
        for (int l = 0; l < 3; l++) d[l] = (size_t)((b.max[l] - b.min[l]) / V[l]);
    }
    // geom.c:157-167 (n*vol - bbvol contracted)
    bool dense_p(const vox_box &b, size_t n) const {
        float s0 = b.max[0] - b.min[0], s1 = b.max[1] - b.min[1],
              s2 = b.max[2] - b.min[2];
        float bbv = (s0 * s1) * s2;
        float vv = (V[0] * V[1]) * V[2];
        return std::fabs(std::fmaf((float)n, vv, -bbv)) < vv;
    }
    // geom.c:198-209
    bool stripe(const vox_box &b, int *which) const {
        size_t d[3];
        dims(b, d);
        int count = 0;
        for (int l = 0; l < 3; l++) {
            if (d[l] == 1) count++;
            else *which = l;
        }
        return count == 2;
    }
    // geom.c:169-181
    static bool divide(const vox_box &box, const float *c, vox_box &res, int idx) {
        res = box;
        for (int l = 0; l < 3; l++) {
            if (idx & (1 << l)) res.max[l] = c[l];
            else res.min[l] = c[l];
            if (res.min[l] == res.max[l]) return false;
        }
        return true;
    }
    // tree.c:77-81
    static int sub_sign(const float *c, const float *d) {
        int r = 0;
        for (int l = 0; l < 3; l++) r |= (std::signbit(d[l] - c[l]) ? 1 : 0) << l;
        return r;
    }
    // geom-sse.c:64-68
    static int sub_less(const float *c, const float *d) {
        int r = 0;
        for (int l = 0; l < 3; l++) r |= (d[l] < c[l] ? 1 : 0) << l;
        return r;
    }
    // voxvision.h.in:52
    static bool equal3(const float *a, const float *b) {
        return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
    }
    // tree.c:137-143
    void grow(vox_box &b, const float *d) const {
        for (int l = 0; l < 3; l++) {
            b.min[l] = smin(d[l], b.min[l]);
            b.max[l] = smax(d[l] + V[l], b.max[l]);
        }
    }

    // tree.c:229-240
    vox_node *dense_leaf(const vox_box &b) {
        vox_node *n = alloc_node();
        if (!n) return nullptr;
        n->bounding_box = b;
        n->flags = VOX_DENSE_LEAF;
        size_t d[3];
        dims(b, d);
        n->dots_num = (unsigned)(d[0] * d[1] * d[2]);
        return n;
    }

    // tree.c:364-420: dense voxels by repeated addition from the minimum.
    // Writes at most cap entries.  A dense leaf whose truncated dimensions
    // (dims) do not multiply to its voxel count -- the reference asserts
    // (tree.c:389) -- and a subtree with more voxels than cap set
    // bad_dense instead, and what is listed then is not used.
    size_t flatten(const vox_node *t, vox_dot *set, size_t cap) const {
        size_t count = 0;
        if (!t) return 0;
        if (t->dots_num > cap) {
            bad_dense = true;
            return 0;
        }
        if (t->flags & VOX_LEAF) {
            std::memcpy(set, t->data.dots, t->dots_num * sizeof(vox_dot));
            return t->dots_num;
        }
        if (t->flags & VOX_DENSE_LEAF) {
            size_t d[3];
            dims(t->bounding_box, d);
            if ((double)d[0] * (double)d[1] * (double)d[2] !=
                (double)t->dots_num) {
                bad_dense = true;
                return 0;
            }
            float cur[3];
            cur[0] = t->bounding_box.min[0];
            for (size_t i = 0; i < d[0]; i++) {
                cur[1] = t->bounding_box.min[1];
                for (size_t j = 0; j < d[1]; j++) {
                    cur[2] = t->bounding_box.min[2];
                    for (size_t k = 0; k < d[2]; k++) {
                        set[count][0] = cur[0];
                        set[count][1] = cur[1];
                        set[count][2] = cur[2];
                        set[count][3] = 0.0f;
                        cur[2] += V[2];
                        count++;
                    }
                    cur[1] += V[1];
                }
                cur[0] += V[0];
            }
            return count;
        }
        for (int i = 0; i < 8; i++) {
            size_t c = flatten(t->data.inner.children[i], set, cap - count);
            set += c;
            count += c;
        }
        return count;
    }

    // vox_make_tree over n dots (host builder, reference layout)
    vox_node *make(const vox_dot *set, size_t n) {
        vox_node *root = nullptr;
        if (build_reference(&set[0][0], n, 4, V, &root, err)) oom = true;
        return root;
    }

    // tree.c:437-458 (voxel_in_tree): the leaf or dense leaf that holds
    // v, or NULL when v is not in the tree
    const vox_node *holder(const vox_node *t, const float *v) const {
        while (t) {
            for (int l = 0; l < 3; l++)
                if (v[l] < t->bounding_box.min[l] || v[l] >= t->bounding_box.max[l])
                    return nullptr;
            if (t->flags & VOX_DENSE_LEAF) return t;
            if (t->flags & VOX_LEAF) {
                for (unsigned i = 0; i < t->dots_num; i++)
                    if (equal3(v, t->data.dots[i])) return t;
                return nullptr;
            }
            t = t->data.inner.children[sub_less(t->data.inner.center, v)];
        }
        return nullptr;
    }

    // v is on a dense leaf's voxel grid: along each axis its minimum plus
    // the voxel size added some number of times, as the leaf's voxels are
    // listed (tree.c:364-420).  A dense leaf holds every point of its box,
    // so a point off that grid passes voxel_in_tree; the reference's
    // deletion then finds no voxel to remove (an assertion, tree.c:808) or
    // divides the box off its grid (tree.c:683-745).
    bool on_grid(const vox_node *t, const float *v) const {
        for (int l = 0; l < 3; l++) {
            float cur = t->bounding_box.min[l];
            while (cur < v[l]) {
                const float next = cur + V[l];
                if (!(next > cur)) return false;
                cur = next;
            }
            if (cur != v[l]) return false;
        }
        return true;
    }

    // a leaf or small dense leaf remade with one more voxel
    vox_node *remake_with(const vox_node *t, const float *v) {
        DotBuf set(t->dots_num + 1);
        flatten(t, set.data(), t->dots_num);
        set.data()[t->dots_num][0] = v[0];
        set.data()[t->dots_num][1] = v[1];
        set.data()[t->dots_num][2] = v[2];
        set.data()[t->dots_num][3] = 0.0f;
        return make(set.data(), t->dots_num + 1);
    }

    // tree.c:461-501
    vox_node *insert_in_big_dense(vox_node *t, const float *v) {
        float c[3], mid[3];
        for (int l = 0; l < 3; l++) {
            // geom-sse.c:51-62 closest_vertex
            float d1 = std::fabs(t->bounding_box.min[l] - v[l]);
            float d2 = std::fabs(t->bounding_box.max[l] - v[l]);
            c[l] = (d1 < d2) ? t->bounding_box.min[l] : t->bounding_box.max[l];
            mid[l] = (t->bounding_box.min[l] + t->bounding_box.max[l]) / 2.0f;
        }
        int i1 = sub_sign(c, mid), i2 = sub_sign(c, v);
        if (i1 == i2) {
            // the new voxel falls on the leaf's own side of every plane
            // (coordinates too coarse for the voxel size): the reference
            // recurses into the same leaf without end
            bad_dense = true;
            return nullptr;
        }
        vox_node *n = alloc_node();
        if (!n) return nullptr;
        n->bounding_box = t->bounding_box;
        grow(n->bounding_box, v);
        n->flags = 0;
        n->dots_num = t->dots_num + 1;
        for (int l = 0; l < 3; l++) n->data.inner.center[l] = c[l];
        n->data.inner.children[i1] = t;
        insert_at(&n->data.inner.children[i2], v);
        return n;
    }

    // tree.c:504-605
    void insert_at(vox_node **slot, const float *v) {
        for (;;) {
            vox_node *t = *slot, *node = t;
            if (!t) {
                vox_dot one[1] = {{v[0], v[1], v[2], 0.0f}};
                *slot = make(one, 1);
                return;
            }
            vox_box saved = t->bounding_box;
            grow(t->bounding_box, v);
            if (dense_p(t->bounding_box, (size_t)t->dots_num + 1)) {
                if (t->flags & VOX_DENSE_LEAF) {
                    t->dots_num++;
                } else {
                    node = dense_leaf(t->bounding_box);
                    if (!node) return;
                    if (node->dots_num != t->dots_num + 1)  // tree.c:530
                        bad_dense = true;
                    destroy_reference(t);
                }
                *slot = node;
                return;
            }
            if (t->flags & VOX_DENSE_LEAF) t->bounding_box = saved;

            if (t->flags & VOX_DENSE_LEAF) {
                if (t->dots_num < kMaxDots) {
                    node = remake_with(t, v);
                    if (!node) return;
                    destroy_reference(t);
                } else {
                    node = insert_in_big_dense(t, v);
                    if (!node) return;
                }
                *slot = node;
                return;
            }
            if (t->flags & VOX_LEAF) {
                if (t->dots_num < kMaxDots) {
                    vox_dot &d = t->data.dots[t->dots_num];
                    d[0] = v[0];
                    d[1] = v[1];
                    d[2] = v[2];
                    d[3] = 0.0f;
                    t->dots_num++;
                } else {
                    // a full leaf (or an overflow leaf, for which the
                    // reference's behaviour is undefined) is remade
                    node = remake_with(t, v);
                    if (!node) return;
                    destroy_reference(t);
                    *slot = node;
                }
                return;
            }
            int idx = sub_sign(t->data.inner.center, v);
            t->dots_num++;
            slot = &t->data.inner.children[idx];
        }
    }

    // tree.c:628-681
    vox_node *delete_from_stripe(vox_node *t, const float *v) {
        int idx = 0;
        if (!stripe(t->bounding_box, &idx)) return nullptr;
        if (equal3(t->bounding_box.min, v)) {
            t->dots_num--;
            t->bounding_box.min[idx] += V[idx];
            return t;
        }
        float other[3] = {v[0] + V[0], v[1] + V[1], v[2] + V[2]};
        if (equal3(t->bounding_box.max, other)) {
            t->dots_num--;
            t->bounding_box.max[idx] -= V[idx];
            return t;
        }
        vox_node *n = alloc_node();
        if (!n) return nullptr;
        n->bounding_box = t->bounding_box;
        n->flags = 0;
        n->dots_num = t->dots_num - 1;
        for (int l = 0; l < 3; l++) n->data.inner.center[l] = v[l];
        vox_box bb;
        divide(t->bounding_box, v, bb, 1 << idx);
        n->data.inner.children[1 << idx] = dense_leaf(bb);
        divide(t->bounding_box, v, bb, 0);
        vox_node *c = dense_leaf(bb);
        n->data.inner.children[0] = c;
        if (!c || !n->data.inner.children[1 << idx]) return nullptr;
        if (c->dots_num + n->data.inner.children[1 << idx]->dots_num !=
            t->dots_num)  // tree.c:671
            bad_dense = true;
        c->dots_num--;
        c->bounding_box.min[idx] += V[idx];
        destroy_reference(t);
        return n;
    }

    // tree.c:683-745
    vox_node *delete_from_big_dense(vox_node *t, const float *v) {
        vox_node *n = delete_from_stripe(t, v);
        if (n || oom) return n;
        vox_box bb;
        if (equal3(t->bounding_box.min, v)) {
            float other[3] = {v[0] + V[0], v[1] + V[1], v[2] + V[2]};
            n = alloc_node();
            if (!n) return nullptr;
            n->bounding_box = t->bounding_box;
            n->flags = 0;
            n->dots_num = t->dots_num - 1;
            for (int l = 0; l < 3; l++) n->data.inner.center[l] = other[l];
            for (int i = 0; i < 7; i++)
                if (divide(t->bounding_box, other, bb, i))
                    n->data.inner.children[i] = dense_leaf(bb);
            destroy_reference(t);
            return n;
        }
        n = alloc_node();
        if (!n) return nullptr;
        n->bounding_box = t->bounding_box;
        n->flags = 0;
        n->dots_num = t->dots_num - 1;
        for (int l = 0; l < 3; l++) n->data.inner.center[l] = v[l];
        for (int i = 0; i < 8; i++)
            if (divide(t->bounding_box, v, bb, i))
                n->data.inner.children[i] = dense_leaf(bb);
        n->data.inner.children[0] = delete_from_big_dense(n->data.inner.children[0], v);
        destroy_reference(t);
        return n;
    }

    // tree.c:748-818
    void delete_at(vox_node **slot, const float *v) {
        for (;;) {
            vox_node *t = *slot;
            if (t->dots_num == 1) {
                destroy_reference(t);
                *slot = nullptr;
                return;
            }
            if (t->flags & VOX_LEAF) {
                unsigned i = 0;
                while (i < t->dots_num && !equal3(t->data.dots[i], v)) i++;
                t->dots_num--;
                std::memmove(t->data.dots + i, t->data.dots + i + 1,
                             sizeof(vox_dot) * (t->dots_num - i));
                return;
            }
            if (t->flags & VOX_DENSE_LEAF) {
                if (t->dots_num <= kMaxDots) {
                    vox_node *n = delete_from_stripe(t, v);
                    if (n) {
                        *slot = n;
                        return;
                    }
                    if (oom) return;
                    vox_dot set[kMaxDots] = {};
                    flatten(t, set, kMaxDots);
                    unsigned i = 0;
                    while (i < t->dots_num && !equal3(set[i], v)) i++;
                    if (i == t->dots_num) {  // not listed: see on_grid
                        bad_dense = true;
                        return;
                    }
                    std::memmove(set + i, set + i + 1,
                                 sizeof(vox_dot) * (t->dots_num - i - 1));
                    vox_node *m = make(set, t->dots_num - 1);
                    if (oom) return;
                    destroy_reference(t);
                    *slot = m;
                } else {
                    vox_node *m = delete_from_big_dense(t, v);
                    if (m) *slot = m;
                }
                return;
            }
            int idx = sub_sign(t->data.inner.center, v);
            vox_node *child = t->data.inner.children[idx];
            if (child->dots_num == t->dots_num) {
                // the only child takes its parent's place
                *slot = child;
                std::free(t);
                continue;
            }
            t->dots_num--;
            slot = &t->data.inner.children[idx];
        }
    }

    // tree.c:128-134
    void align(float *v) const {
        for (int l = 0; l < 3; l++) v[l] = std::floor(v[l] / V[l]) * V[l];
    }
};

// the thread's voxg_last_error() message is kept by the HIP translation unit
int fail_edit(const std::string &msg, int code) {
    return voxg_internal::fail(code, msg);
}

}  // namespace

extern "C" {

int voxg_edit_create(const float *xyz, size_t n, size_t stride,
                     const float voxel[3], voxg_edit_tree **out) {
    if (!out) return fail_edit("out is NULL", VOXG_EINVAL);
    *out = nullptr;
    if (n && !xyz) return fail_edit("xyz is NULL", VOXG_EINVAL);
    if (stride != 3 && stride != 4)
        return fail_edit("stride must be 3 or 4", VOXG_EINVAL);
    if (voxel && !(voxel[0] > 0.0f && voxel[1] > 0.0f && voxel[2] > 0.0f))
        return fail_edit("voxel size must be positive", VOXG_EINVAL);
    auto *t = new (std::nothrow) voxg_edit_tree();
    if (!t) return fail_edit("host allocation failed", VOXG_ENOMEM);
    if (voxel)
        for (int l = 0; l < 3; l++) t->V[l] = voxel[l];
    std::string err;
    int rc = build_reference(xyz, n, stride, t->V, &t->root, err);
    if (rc) {
        delete t;
        return fail_edit(err, rc);
    }
    *out = t;
    return 0;
}

// A dense leaf whose dimensions disagree with its voxel count (possible
// when the voxel size is not exact in binary, so that box extents divided
// by it truncate one short): the reference asserts; here the edit that met
// it fails, and so does every later call on the tree.
static int broken_dense(voxg_edit_tree *t) {
    t->broken =
        "an edit met a dense leaf that its box and the voxel size do not "
        "describe exactly (its box does not give its voxel count, it does "
        "not list the voxel, or a new voxel does not separate from it: a "
        "voxel size not exact in binary floating point, or coordinates too "
        "large for it); the tree is no longer usable: rebuild it from its "
        "dots";
    return fail_edit(t->broken, VOXG_ERANGE);
}

static int edit_op(voxg_edit_tree *t, const float xyz[3], int insert,
                   int *changed) {
    if (!t || !xyz) return fail_edit("NULL argument", VOXG_EINVAL);
    if (!t->broken.empty()) return fail_edit(t->broken, VOXG_ERANGE);
    Editor e;
    for (int l = 0; l < 3; l++) e.V[l] = t->V[l];
    float v[3] = {xyz[0], xyz[1], xyz[2]};
    e.align(v);
    const vox_node *h = e.holder(t->root, v);
    const bool present = h != nullptr;
    if (!insert && h && (h->flags & VOX_DENSE_LEAF) && !e.on_grid(h, v))
        return fail_edit("the voxel lies in a dense leaf but off its voxel "
                         "grid (the tree's dots are not multiples of the "
                         "voxel size); the tree is unchanged",
                         VOXG_EINVAL);
    int res = insert ? !present : present;
    if (res) {
        if (insert) e.insert_at(&t->root, v);
        else e.delete_at(&t->root, v);
        if (e.oom)
            return fail_edit(e.err.empty() ? "host allocation failed" : e.err,
                             VOXG_ENOMEM);
        if (e.bad_dense) return broken_dense(t);
    }
    if (changed) *changed = res;
    return 0;
}

int voxg_edit_insert(voxg_edit_tree *t, const float xyz[3], int *inserted) {
    return edit_op(t, xyz, 1, inserted);
}

int voxg_edit_delete(voxg_edit_tree *t, const float xyz[3], int *deleted) {
    return edit_op(t, xyz, 0, deleted);
}

int voxg_edit_rebuild(voxg_edit_tree *t) {
    if (!t) return fail_edit("NULL argument", VOXG_EINVAL);
    if (!t->broken.empty()) return fail_edit(t->broken, VOXG_ERANGE);
    if (!t->root) return 0;
    Editor e;
    for (int l = 0; l < 3; l++) e.V[l] = t->V[l];
    DotBuf set(t->root->dots_num);
    e.flatten(t->root, set.data(), t->root->dots_num);
    if (e.bad_dense) return broken_dense(t);
    vox_node *m = e.make(set.data(), t->root->dots_num);
    if (e.oom) return fail_edit(e.err, VOXG_ENOMEM);
    destroy_reference(t->root);
    t->root = m;
    return 0;
}

const struct vox_node *voxg_edit_root(const voxg_edit_tree *t) {
    return t ? t->root : nullptr;
}

size_t voxg_edit_voxels(const voxg_edit_tree *t) {
    return (t && t->root) ? t->root->dots_num : 0;
}

int voxg_edit_commit(const voxg_edit_tree *t, voxg_tree **out) {
    if (!t) return fail_edit("NULL argument", VOXG_EINVAL);
    if (!t->broken.empty()) return fail_edit(t->broken, VOXG_ERANGE);
    return voxg_tree_upload(t->root, t->V, out);
}

int voxg_edit_release(voxg_edit_tree *t) {
    if (t) {
        destroy_reference(t->root);
        delete t;
    }
    return 0;
}

}  // extern "C"
