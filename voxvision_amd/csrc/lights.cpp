// Host M-tree of the light manager (src/voxtrees/mtree.c:11-369 for the
// tree, src/voxrnd/lights.c:91-116 for the light at a point), and its
// depth-first flattening for the device.
//
// Floating point as the reference's clang build computes it (this file is
// compiled with -ffp-contract=off): squared distances (dx*dx + dy*dy) +
// dz*dz (geom-sse.c:12-17), pow(r, 2) as one rounded r*r, the merged
// bounding-sphere centre crit*(b - s) + s and the light sum
// colour + add*light_colour as fused multiply-adds (each is one expression
// of the reference, which clang contracts).
#include "lights.h"

#include <cmath>
#include <cstring>

namespace voxg {

namespace {

constexpr unsigned kFanout = 9;  // MTREE_MAX_CHILDREN (src/voxtrees/mtree.h)

float dist2(const float *a, const float *b) {
    float x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    x = x * x;
    y = y * y;
    z = z * z;
    return (x + y) + z;
}

bool same_center(const float *a, const float *b) {
    return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}

// mtree.c:22-56: grow `b` to cover `s`; false when it already does.
bool cover(const LightSphere &s, LightSphere &b) {
    const float dist = std::sqrt(dist2(s.center, b.center));
    const float crit = (1.0f + (b.radius - s.radius) / dist) / 2.0f;
    if (crit < 0.0f) {
        b = s;
        return true;
    }
    if (!(crit < 1.0f)) return false;
    b.radius = ((s.radius + b.radius) + dist) / 2.0f;
    b.sqr_radius = b.radius * b.radius;
    for (int l = 0; l < 3; l++)
        b.center[l] = std::fma(crit, b.center[l] - s.center[l], s.center[l]);
    return true;
}

}  // namespace

struct LightTree::Node {
    LightSphere bs;  // bounding sphere
    Node *parent = nullptr;
    bool leaf = false;
    unsigned num = 0;
    Node *kids[kFanout + 1] = {};
    LightSphere items[kFanout + 1];  // leaf entries

    const LightSphere &entry(unsigned i) const {
        return leaf ? items[i] : kids[i]->bs;
    }
    // mtree.c:59-76: the first entry, then each further entry merged in
    bool rebound() {
        bs = entry(0);
        bool changed = false;
        for (unsigned i = 1; i < num; i++) changed |= cover(entry(i), bs);
        return changed;
    }
};

LightTree::LightTree() : root_(nullptr) {}
LightTree::~LightTree() { destroy(root_); }

LightTree::Node *LightTree::alloc(bool leaf) {
    Node *n = new Node();
    n->leaf = leaf;
    return n;
}

void LightTree::destroy(Node *n) {
    if (!n) return;
    if (!n->leaf)
        for (unsigned i = 0; i < n->num; i++) destroy(n->kids[i]);
    delete n;
}

void LightTree::clear() {
    destroy(root_);
    root_ = nullptr;
}

// mtree.c:82-89: recompute bounding spheres upward while they change
static void propagate(LightTree::Node *n);

// mtree.c:117-189
LightTree::Node *LightTree::split(Node *node) {
    // mtree.c:91-115: the two entries whose centres are farthest apart,
    // the first such pair in entry order (centres are distinct here)
    const float *c1 = node->entry(0).center, *c2 = c1;
    float far = 0.0f;
    for (unsigned i = 0; i < node->num; i++)
        for (unsigned j = i + 1; j < node->num; j++) {
            const float d = dist2(node->entry(i).center, node->entry(j).center);
            if (d > far) {
                far = d;
                c1 = node->entry(i).center;
                c2 = node->entry(j).center;
            }
        }
    float a[3], b[3];
    std::memcpy(a, c1, sizeof(a));
    std::memcpy(b, c2, sizeof(b));
    Node *parent = node->parent, *res = parent;
    Node *half[2] = {alloc(node->leaf), alloc(node->leaf)};
    for (unsigned i = 0; i < node->num; i++) {
        const float *c = node->entry(i).center;
        Node *w = half[dist2(c, a) < dist2(c, b) ? 0 : 1];
        if (node->leaf) {
            w->items[w->num] = node->items[i];
        } else {
            w->kids[w->num] = node->kids[i];
            node->kids[i]->parent = w;
        }
        w->num++;
    }
    if (half[0]->num == 0) {
        // every entry's centre was the same (the reference leaves this
        // case undefined): the last entry alone makes the other half
        Node *h = half[1];
        h->num--;
        if (node->leaf) {
            half[0]->items[0] = h->items[h->num];
        } else {
            half[0]->kids[0] = h->kids[h->num];
            half[0]->kids[0]->parent = half[0];
        }
        half[0]->num = 1;
    }
    for (Node *h : half) {
        h->parent = parent;
        h->rebound();
    }
    if (parent) {
        for (unsigned i = 0; i < parent->num; i++)
            if (parent->kids[i] == node) parent->kids[i] = half[0];
        delete node;
        parent->kids[parent->num++] = half[1];
        propagate(parent);
        if (parent->num == kFanout + 1) res = split(parent);
    } else {
        delete node;
        res = alloc(false);
        res->num = 2;
        res->kids[0] = half[0];
        res->kids[1] = half[1];
        res->rebound();
        half[0]->parent = half[1]->parent = res;
    }
    return res;
}

static void propagate(LightTree::Node *n) {
    while (n && n->rebound()) n = n->parent;
}

// mtree.c:219-255
LightTree::Node *LightTree::add(Node *node, const LightSphere &s) {
    if (!node) {
        Node *n = alloc(true);
        n->bs = s;
        n->items[0] = s;
        n->num = 1;
        return n;
    }
    if (node->leaf) {
        node->items[node->num++] = s;
        propagate(node);
        return node->num == kFanout + 1 ? split(node) : node;
    }
    Node *best = nullptr;
    float bestd = INFINITY;
    for (unsigned i = 0; i < node->num; i++) {
        const float d = dist2(s.center, node->kids[i]->bs.center);
        if (d < bestd) {
            bestd = d;
            best = node->kids[i];
        }
    }
    return add(best, s);
}

// mtree.c:191-217: the leaf holding the sphere (centre and radius equal)
LightTree::Node *LightTree::find(Node *n, const float c[3], float r) const {
    if (!n || !(dist2(c, n->bs.center) < n->bs.sqr_radius)) return nullptr;
    if (n->leaf) {
        for (unsigned i = 0; i < n->num; i++)
            if (same_center(n->items[i].center, c) && n->items[i].radius == r)
                return n;
        return nullptr;
    }
    for (unsigned i = 0; i < n->num; i++)
        if (Node *f = find(n->kids[i], c, r)) return f;
    return nullptr;
}

bool LightTree::has_center(const Node *n, const float c[3]) const {
    if (!n) return false;
    for (unsigned i = 0; i < n->num; i++)
        if (n->leaf ? same_center(n->items[i].center, c)
                    : has_center(n->kids[i], c))
            return true;
    return false;
}

int LightTree::insert(const float center[3], float radius,
                      const float color[3]) {
    if (!(radius > 0.0f) || find(root_, center, radius)) return 0;
    for (int l = 0; l < 3; l++)
        if (!std::isfinite(center[l])) return -1;
    if (!std::isfinite(radius) || has_center(root_, center)) return -1;
    LightSphere s;
    std::memcpy(s.center, center, sizeof(s.center));
    std::memcpy(s.color, color, sizeof(s.color));
    s.radius = radius;
    s.sqr_radius = radius * radius;
    Node *n = add(root_, s);
    if (!n->parent) root_ = n;
    return 1;
}

// mtree.c:278-322: remove a leaf with one entry; a parent left with one
// child is replaced by that child
LightTree::Node *LightTree::drop_leaf(Node *leaf) {
    Node *parent = leaf->parent;
    if (!parent) {
        delete leaf;
        return nullptr;
    }
    unsigned i = 0;
    while (parent->kids[i] != leaf) i++;
    delete leaf;
    if (parent->num == 2) {
        Node *other = parent->kids[1 - i], *top = parent->parent;
        other->parent = top;
        if (!top) {
            delete parent;
            return other;
        }
        unsigned j = 0;
        while (top->kids[j] != parent) j++;
        top->kids[j] = other;
        delete parent;
        propagate(top);
        return top;
    }
    parent->num--;
    for (unsigned k = i; k < parent->num; k++)
        parent->kids[k] = parent->kids[k + 1];
    parent->kids[parent->num] = nullptr;
    propagate(parent);
    return parent;
}

// mtree.c:325-369
int LightTree::remove(const float center[3], float radius) {
    Node *leaf = find(root_, center, radius);
    if (!(radius > 0.0f) || !leaf) return 0;
    Node *last;
    if (leaf->num == 1) {
        last = drop_leaf(leaf);
    } else {
        unsigned i = 0;
        while (!(same_center(leaf->items[i].center, center) &&
                 leaf->items[i].radius == radius))
            i++;
        for (unsigned k = i; k + 1 < leaf->num; k++)
            leaf->items[k] = leaf->items[k + 1];
        leaf->num--;
        propagate(leaf);
        last = leaf;
    }
    if (!last || !last->parent) root_ = last;
    return 1;
}

static size_t count_of(const LightTree::Node *n);

size_t LightTree::count() const { return count_of(root_); }

static size_t count_of(const LightTree::Node *n) {
    if (!n) return 0;
    if (n->leaf) return n->num;
    size_t k = 0;
    for (unsigned i = 0; i < n->num; i++) k += count_of(n->kids[i]);
    return k;
}

// lights.c:91-101 with the walk of mtree.c:463-491
static void light_walk(const LightTree::Node *n, const float *p, float *col) {
    if (!n || !(dist2(n->bs.center, p) < n->bs.sqr_radius)) return;
    auto add_one = [&](const LightSphere &s) {
        const float d = std::fmin(s.radius, std::sqrt(dist2(s.center, p)));
        const float a = 1.0f - d / s.radius;
        for (int l = 0; l < 3; l++) col[l] = std::fma(a, s.color[l], col[l]);
    };
    if (n->leaf) {
        if (n->num == 1) {
            add_one(n->items[0]);
            return;
        }
        for (unsigned i = 0; i < n->num; i++)
            if (dist2(n->items[i].center, p) < n->items[i].sqr_radius)
                add_one(n->items[i]);
        return;
    }
    for (unsigned i = 0; i < n->num; i++) light_walk(n->kids[i], p, col);
}

void LightTree::light(const float ambient[3], const float p[3],
                      float out[3]) const {
    float col[3] = {ambient[0], ambient[1], ambient[2]};
    light_walk(root_, p, col);
    std::memcpy(out, col, sizeof(col));
}

static void flatten_node(const LightTree::Node *n,
                         std::vector<LightNode> &nodes,
                         std::vector<LightRec> &spheres) {
    const size_t me = nodes.size();
    nodes.emplace_back();
    LightNode r;
    std::memset(&r, 0, sizeof(r));
    std::memcpy(r.c, n->bs.center, sizeof(r.c));
    r.sqr_radius = n->bs.sqr_radius;
    if (n->leaf) {
        r.info = 1u | (n->num << 8);
        r.first = (uint32_t)spheres.size();
        for (unsigned i = 0; i < n->num; i++) {
            const LightSphere &s = n->items[i];
            LightRec q;
            std::memcpy(q.c, s.center, sizeof(q.c));
            q.radius = s.radius;
            std::memcpy(q.color, s.color, sizeof(q.color));
            q.sqr_radius = s.sqr_radius;
            spheres.push_back(q);
        }
    } else {
        r.info = 0u;
        for (unsigned i = 0; i < n->num; i++)
            flatten_node(n->kids[i], nodes, spheres);
    }
    r.escape = (uint32_t)nodes.size();
    nodes[me] = r;
}

void LightTree::flatten(std::vector<LightNode> &nodes,
                        std::vector<LightRec> &spheres) const {
    nodes.clear();
    spheres.clear();
    if (root_) flatten_node(root_, nodes, spheres);
    // the last record's escape (one past the end) marks the walk's end
}

}  // namespace voxg
