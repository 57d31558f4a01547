// Host-side scene preparation for the device renderer.
//   light distributions          reference scene.cpp:24-52, distribution.h:7-59
//   edge extraction + filtering  reference scene.cpp:54-97 (sorted by
//                                (shape, v0, v1), as the reference's std::map)
//   edge trees                   reference edge_tree.cpp:49-316 (median
//                                splits; the same build as the CPU oracle, so
//                                both sample the same edges)
//   primary-edge distribution    reference edge_tree.cpp:266-294
//   triangle BVH                 binned SAH, BVH2 with both child boxes in
//                                the parent node (replaces Embree)
#include "scene_build.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <tuple>

namespace drh {

namespace {

struct V3 {
    float x, y, z;
    float &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, float s) { float inv = 1.f / s; return a * inv; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline V3 normalize(V3 a) { return a / length(a); }
inline float luminance(V3 a) { return 0.212671f * a.x + 0.715160f * a.y + 0.072169f * a.z; }

struct Box {
    V3 lo, hi;
    Box() {
        float inf = std::numeric_limits<float>::infinity();
        lo = V3{inf, inf, inf};
        hi = V3{-inf, -inf, -inf};
    }
    Box(V3 a, V3 b) {
        lo = V3{std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)};
        hi = V3{std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)};
    }
    void grow(const Box &b) {
        lo = V3{std::min(lo.x, b.lo.x), std::min(lo.y, b.lo.y), std::min(lo.z, b.lo.z)};
        hi = V3{std::max(hi.x, b.hi.x), std::max(hi.y, b.hi.y), std::max(hi.z, b.hi.z)};
    }
    void grow(V3 p) { grow(Box(p, p)); }
    float area() const {
        if (lo.x > hi.x) return 0.f;
        V3 d = hi - lo;
        return 2.f * (d.x * d.y + d.x * d.z + d.y * d.z);
    }
};

// pbrt-style 1-D distribution with a float CDF accumulated in order
Dist make_dist(const std::vector<float> &f) {
    Dist d;
    int n = (int)f.size();
    d.func = f;
    d.cdf.assign(n + 1, 0.f);
    for (int i = 1; i < n + 1; i++) d.cdf[i] = d.cdf[i - 1] + d.func[i - 1] / n;
    d.func_int = d.cdf[n];
    if (d.func_int == 0.f) {
        for (int i = 1; i < n + 1; i++) d.cdf[i] = float(i) / float(n);
    } else {
        for (int i = 1; i < n + 1; i++) d.cdf[i] /= d.func_int;
    }
    return d;
}

V3 getv(const HostScene &hs, int shape, int v) {
    const float *p = &hs.verts[3 * (size_t)(hs.shapes[shape].vert_off + v)];
    return V3{p[0], p[1], p[2]};
}
V3 getn(const HostScene &hs, int shape, int v) {
    const float *p = &hs.normals[3 * (size_t)(hs.shapes[shape].vert_off + v)];
    return V3{p[0], p[1], p[2]};
}
int tri_index(const HostScene &hs, int shape, int t, int k) {
    return hs.tri_vidx[3 * (size_t)(hs.shapes[shape].tri_off + t) + k] - hs.shapes[shape].vert_off;
}
V3 face_normal(const HostScene &hs, int shape, int t) {
    V3 a = getv(hs, shape, tri_index(hs, shape, t, 0)), b = getv(hs, shape, tri_index(hs, shape, t, 1)),
       c = getv(hs, shape, tri_index(hs, shape, t, 2));
    return normalize(cross(b - a, c - a));
}

// --------------------------------------------------------- triangle BVH --
struct BPrim {
    Box box;
    V3 c;
    int id;
};
struct BNode {
    Box box;
    int left = -1, right = -1, first = 0, count = 0;
};

// Binned SAH split; below depth kSahDepth the split is the object median on
// the largest centroid axis, so the depth stays below kSahDepth + log2(n) and
// the device traversal stack (DR_STACK entries, one per level) cannot fill.
const int kSahDepth = 32;

int bvh_rec(std::vector<BNode> &nodes, std::vector<int> &order, std::vector<BPrim> &p, int b, int e, int depth,
            int &max_depth) {
    max_depth = std::max(max_depth, depth);
    int id = (int)nodes.size();
    nodes.emplace_back();
    Box box, cb;
    for (int i = b; i < e; i++) {
        box.grow(p[i].box);
        cb.grow(p[i].c);
    }
    nodes[id].box = box;
    int n = e - b;
    auto leaf = [&]() {
        nodes[id].first = (int)order.size();
        nodes[id].count = n;
        for (int i = b; i < e; i++) order.push_back(p[i].id);
    };
    // leaves of at most 2 triangles: on the heightfield's adjoint the edge
    // samples' side rays take 13 % less time than with leaves of 4 (single-
    // triangle leaves: 2 % more than with 2)
    if (n <= 2) {
        leaf();
        return id;
    }
    V3 ext = cb.hi - cb.lo;
    int axis = 0;
    if (ext.y > ext[axis]) axis = 1;
    if (ext.z > ext[axis]) axis = 2;
    int mid = (b + e) / 2;
    if (depth >= kSahDepth) {
        std::nth_element(p.begin() + b, p.begin() + mid, p.begin() + e,
                         [axis](const BPrim &x, const BPrim &y) { return x.c[axis] < y.c[axis]; });
    } else if (ext[axis] > 0.f) {
        const int NB = 32;
        Box bb[NB];
        int bc[NB] = {0};
        float scale = NB / ext[axis];
        auto bin = [&](const BPrim &q) { return std::min(std::max((int)((q.c[axis] - cb.lo[axis]) * scale), 0), NB - 1); };
        for (int i = b; i < e; i++) {
            int k = bin(p[i]);
            bb[k].grow(p[i].box);
            bc[k]++;
        }
        Box rb[NB];
        int rc[NB];
        Box acc;
        int cnt = 0;
        for (int k = NB - 1; k >= 0; k--) {
            acc.grow(bb[k]);
            cnt += bc[k];
            rb[k] = acc;
            rc[k] = cnt;
        }
        float best = std::numeric_limits<float>::infinity();
        int best_k = -1;
        Box lb;
        int lc = 0;
        for (int k = 1; k < NB; k++) {
            lb.grow(bb[k - 1]);
            lc += bc[k - 1];
            if (lc == 0 || rc[k] == 0) continue;
            float cost = lb.area() * lc + rb[k].area() * rc[k];
            if (cost < best) { best = cost; best_k = k; }
        }
        if (best_k > 0) {
            auto it = std::partition(p.begin() + b, p.begin() + e, [&](const BPrim &q) { return bin(q) < best_k; });
            mid = (int)(it - p.begin());
            if (mid == b || mid == e) mid = (b + e) / 2;
        }
    }
    int l = bvh_rec(nodes, order, p, b, mid, depth + 1, max_depth);
    int r = bvh_rec(nodes, order, p, mid, e, depth + 1, max_depth);
    nodes[id].left = l;
    nodes[id].right = r;
    return id;
}

void build_bvh(HostScene &hs) {
    int nt = (int)hs.tri_shape.size();
    hs.nodes.clear();
    hs.tris.clear();
    hs.bvh_depth = 0;
    if (nt == 0) return;
    std::vector<BPrim> p(nt);
    for (int g = 0; g < nt; g++) {
        const int *vi = &hs.tri_vidx[3 * (size_t)g];
        V3 a = V3{hs.verts[3 * (size_t)vi[0]], hs.verts[3 * (size_t)vi[0] + 1], hs.verts[3 * (size_t)vi[0] + 2]};
        V3 b = V3{hs.verts[3 * (size_t)vi[1]], hs.verts[3 * (size_t)vi[1] + 1], hs.verts[3 * (size_t)vi[1] + 2]};
        V3 c = V3{hs.verts[3 * (size_t)vi[2]], hs.verts[3 * (size_t)vi[2] + 1], hs.verts[3 * (size_t)vi[2] + 2]};
        Box bx(a, b);
        bx.grow(c);
        p[g].box = bx;
        p[g].c = (bx.lo + bx.hi) * 0.5f;
        p[g].id = g;
    }
    std::vector<BNode> bn;
    std::vector<int> order;
    bvh_rec(bn, order, p, 0, nt, 0, hs.bvh_depth);
    // flatten into 2-wide nodes holding both child boxes; leaves are folded
    // into their parent.  A root that is itself a leaf gets an empty sibling.
    hs.tris.resize(nt);
    for (int i = 0; i < nt; i++) {
        int g = order[i];
        const int *vi = &hs.tri_vidx[3 * (size_t)g];
        float v[3][3];
        for (int k = 0; k < 3; k++)
            for (int c = 0; c < 3; c++) v[k][c] = hs.verts[3 * (size_t)vi[k] + c];
        DrTri &T = hs.tris[i];
        for (int c = 0; c < 3; c++) {
            T.v0[c] = v[0][c];
            T.e1[c] = v[1][c] - v[0][c];
            T.e2[c] = v[2][c] - v[0][c];
        }
        std::memcpy(&T.v0[3], &g, 4);
        T.e1[3] = T.e2[3] = 0.f;
    }
    std::vector<int> map(bn.size(), -1);
    // assign output indices to interior nodes in preorder
    std::vector<int> stack = {0};
    std::vector<int> interior;
    while (!stack.empty()) {
        int n = stack.back();
        stack.pop_back();
        if (bn[n].left < 0) continue;
        map[n] = (int)interior.size();
        interior.push_back(n);
        stack.push_back(bn[n].right);
        stack.push_back(bn[n].left);
    }
    auto put_child = [&](DrNode &o, int slot, int c) {
        const BNode &ch = bn[c];
        float *lo, *hi;
        float box[6] = {ch.box.lo.x, ch.box.lo.y, ch.box.lo.z, ch.box.hi.x, ch.box.hi.y, ch.box.hi.z};
        (void)lo;
        (void)hi;
        if (slot == 0) {
            o.a[0] = box[0]; o.a[1] = box[1]; o.a[2] = box[2]; o.a[3] = box[3];
            o.b[0] = box[4]; o.b[1] = box[5];
        } else {
            o.b[2] = box[0]; o.b[3] = box[1];
            o.c[0] = box[2]; o.c[1] = box[3]; o.c[2] = box[4]; o.c[3] = box[5];
        }
        if (ch.left < 0) {
            o.d[slot] = ch.first;
            o.d[2 + slot] = ch.count;
        } else {
            o.d[slot] = map[c];
            o.d[2 + slot] = 0;
        }
    };
    if (interior.empty()) {
        DrNode o;
        std::memset(&o, 0, sizeof(o));
        put_child(o, 0, 0);
        float inf = std::numeric_limits<float>::infinity();
        o.b[2] = o.b[3] = o.c[0] = inf;
        o.c[1] = o.c[2] = o.c[3] = -inf;
        o.d[1] = 0;
        o.d[3] = 0;
        hs.nodes.push_back(o);
        return;
    }
    hs.nodes.resize(interior.size());
    for (size_t i = 0; i < interior.size(); i++) {
        DrNode o;
        std::memset(&o, 0, sizeof(o));
        put_child(o, 0, bn[interior[i]].left);
        put_child(o, 1, bn[interior[i]].right);
        hs.nodes[i] = o;
    }
}

// ------------------------------------------------------------ edge trees --
struct EPrim {
    int id;
    Box box;
    V3 c;
    V3 n0, n1;
    Box dbox;
    V3 dc;
};

int edge_rec(const std::vector<DrEdge> &edges, std::vector<EPrim> &p, int b, int e, bool directional,
             std::vector<DrEdgeNode> &nodes, std::vector<DrEdge> &ordered, std::vector<int> &child1,
             std::vector<V3> &on, int depth, int &max_depth) {
    max_depth = std::max(max_depth, depth);
    int id = (int)nodes.size();
    nodes.emplace_back();
    child1.push_back(-1);
    Box box, dbox;
    for (int i = b; i < e; i++) {
        box.grow(p[i].box);
        if (directional) dbox.grow(p[i].dbox);
    }
    // the normal cone is set bottom-up once the tree is built
    DrEdgeNode &N = nodes[id];
    N.lo[0] = box.lo.x; N.lo[1] = box.lo.y; N.lo[2] = box.lo.z;
    N.hi[0] = box.hi.x; N.hi[1] = box.hi.y; N.hi[2] = box.hi.z;
    N.nprims = 0;
    N.offset = 0;
    N.total_length = 0.f;
    N.pad = 0.f;
    int n = e - b;
    auto leaf = [&]() {
        nodes[id].offset = (int)ordered.size();
        nodes[id].nprims = n;
        for (int i = b; i < e; i++) {
            ordered.push_back(edges[p[i].id]);
            on.push_back(p[i].n0);
            on.push_back(p[i].n1);
        }
    };
    if (n == 1) {
        leaf();
        return id;
    }
    Box cb, db;
    for (int i = b; i < e; i++) {
        cb.grow(p[i].c);
        if (directional) db.grow(p[i].dc);
    }
    int dim = -1;
    float mext = -std::numeric_limits<float>::infinity();
    for (int i = 0; i < 3; i++) {
        float x = cb.hi[i] - cb.lo[i];
        if (x > mext) { dim = i; mext = x; }
    }
    if (directional) {
        for (int i = 0; i < 3; i++) {
            float x = db.hi[i] - db.lo[i];
            if (x > mext) { dim = i + 3; mext = x; }
        }
    }
    bool spatial = dim < 3;
    const Box &used = spatial ? cb : db;
    if (dim >= 3) dim -= 3;
    if (used.hi[dim] == used.lo[dim]) {
        leaf();
        return id;
    }
    int mid = (b + e) / 2;
    std::nth_element(p.begin() + b, p.begin() + mid, p.begin() + e, [dim, spatial](const EPrim &x, const EPrim &y) {
        return spatial ? x.c[dim] < y.c[dim] : x.dc[dim] < y.dc[dim];
    });
    edge_rec(edges, p, b, mid, directional, nodes, ordered, child1, on, depth + 1, max_depth);
    int r = edge_rec(edges, p, mid, e, directional, nodes, ordered, child1, on, depth + 1, max_depth);
    child1[id] = r;
    return id;
}

// Returns the depth of the deepest node (root: 0).  Median splits keep it at
// most ceil(log2(edges)).
int build_edge_tree(const HostScene &hs, const std::vector<DrEdge> &edges, bool directional,
                    std::vector<DrEdgeNode> &nodes, std::vector<DrEdge> &ordered) {
    nodes.clear();
    ordered.clear();
    if (edges.empty()) return 0;
    std::vector<EPrim> p(edges.size());
    Box world;
    for (size_t i = 0; i < edges.size(); i++) {
        p[i].id = (int)i;
        p[i].box = Box(getv(hs, edges[i].shape, edges[i].v0), getv(hs, edges[i].shape, edges[i].v1));
        p[i].c = .5f * p[i].box.lo + .5f * p[i].box.hi;
        p[i].n0 = V3{0.f, 0.f, 0.f};
        p[i].n1 = V3{0.f, 0.f, 0.f};
        world.grow(p[i].box);
    }
    if (directional) {
        float factor = length(world.hi - world.lo) / 8.f;
        for (size_t i = 0; i < edges.size(); i++) {
            const DrEdge &ed = edges[i];
            V3 sn = normalize(getn(hs, ed.shape, ed.v0) + getn(hs, ed.shape, ed.v1));
            V3 n0 = face_normal(hs, ed.shape, ed.f0);
            V3 n1 = face_normal(hs, ed.shape, ed.f1);
            if (dot(n0, sn) < 0.f) n0 = -n0;
            if (dot(n1, sn) < 0.f) n1 = -n1;
            p[i].n0 = n0;
            p[i].n1 = n1;
            p[i].dbox = Box(n0 * factor, n1 * factor);
            p[i].dc = 0.5f * (p[i].dbox.lo + p[i].dbox.hi);
        }
    }
    std::vector<int> child1;
    std::vector<V3> on;  // per ordered edge: its two face normals
    int depth = 0;
    edge_rec(edges, p, 0, (int)edges.size(), directional, nodes, ordered, child1, on, 0, depth);
    // Bottom-up (children follow parents): interior offset = second child,
    // total length, the sum of the face normals (leaf: in edge order;
    // interior: left + right) and the node's range of ordered edges; the
    // cone axis is the normalised sum, its cos_angle the least cosine to it
    // over the node's edges.  The device refit (refit.hip) repeats
    // these steps in the same order, so a refit to the same vertices
    // reproduces this tree.
    int nn = (int)nodes.size();
    std::vector<V3> nsum(nn);
    std::vector<int> first(nn), last(nn);
    for (int i = nn - 1; i >= 0; i--) {
        DrEdgeNode &n = nodes[i];
        if (n.nprims > 0) {
            n.total_length = 0.f;
            nsum[i] = V3{0.f, 0.f, 0.f};
            for (int k = 0; k < n.nprims; k++) {
                const DrEdge &ed = ordered[n.offset + k];
                n.total_length += length(getv(hs, ed.shape, ed.v1) - getv(hs, ed.shape, ed.v0));
                nsum[i] = nsum[i] + on[2 * (n.offset + k)];
                nsum[i] = nsum[i] + on[2 * (n.offset + k) + 1];
            }
            first[i] = n.offset;
            last[i] = n.offset + n.nprims;
        } else {
            n.offset = child1[i];
            n.total_length = nodes[i + 1].total_length + nodes[n.offset].total_length;
            nsum[i] = nsum[i + 1] + nsum[n.offset];
            first[i] = first[i + 1];
            last[i] = last[n.offset];
        }
        V3 axis = normalize(nsum[i]);
        float cmin = 1.f;
        for (int k = first[i]; k < last[i]; k++) {
            cmin = std::min(cmin, dot(on[2 * k], axis));
            cmin = std::min(cmin, dot(on[2 * k + 1], axis));
        }
        n.axis[0] = axis.x; n.axis[1] = axis.y; n.axis[2] = axis.z;
        n.cos_angle = cmin;
    }
    // sampling length of the learned proposal: the plain length until
    // weights are set
    for (DrEdgeNode &n : nodes) n.pad = n.total_length;
    return depth;
}


// ------------------------------------------------------- camera helpers --
V3 xfm_point(const float *a, V3 p) {
    float x = a[0] * p.x + a[1] * p.y + a[2] * p.z + a[3];
    float y = a[4] * p.x + a[5] * p.y + a[6] * p.z + a[7];
    float z = a[8] * p.x + a[9] * p.y + a[10] * p.z + a[11];
    float w = a[12] * p.x + a[13] * p.y + a[14] * p.z + a[15];
    float inv = 1.f / w;
    return V3{x * inv, y * inv, z * inv};
}
struct V2 {
    float x, y;
};
// camera.h:156-178
V2 project_local(const DrCamera &c, V3 l) {
    if (c.fisheye) {
        V3 d = normalize(l);
        float phi = std::atan2(d.y, d.x);
        float r = std::acos(d.z) * 2.f / float(M_PI);
        return V2{0.5f * (-r * std::cos(phi) + 1.f), 0.5f * (-r * std::sin(phi) + 1.f)};
    }
    return V2{(-l.x / (l.z * c.fov) + 1.f) * 0.5f, (-l.y / (l.z * c.fov * c.aspect) + 1.f) * 0.5f};
}
bool project_edge(const DrCamera &c, V3 p0, V3 p1, V2 &s0, V2 &s1) {
    V3 a = xfm_point(c.w2c, p0), b = xfm_point(c.w2c, p1);
    if (a.z < c.clip_near && b.z < c.clip_near) return false;
    if (a.z < c.clip_near) {
        V3 d = a - b;
        float t = -(b.z - c.clip_near) / d.z;
        a = b + t * d;
    } else if (b.z < c.clip_near) {
        V3 d = b - a;
        float t = -(a.z - c.clip_near) / d.z;
        b = a + t * d;
    }
    s0 = project_local(c, a);
    s1 = project_local(c, b);
    return true;
}
// Cohen-Sutherland clip to [0,1]^2 (reference line_clip.h:13-102)
bool clip_line(V2 &a, V2 &b) {
    auto code = [](const V2 &v) {
        int c = 0;
        if (v.x < 0.f) c |= 1; else if (v.x > 1.f) c |= 2;
        if (v.y < 0.f) c |= 4; else if (v.y > 1.f) c |= 8;
        return c;
    };
    int c0 = code(a), c1 = code(b);
    while (true) {
        if (!(c0 | c1)) return true;
        if (c0 & c1) return false;
        int out = c0 ? c0 : c1;
        V2 v = V2{0.f, 0.f};
        if (out & 8) { v.x = a.x + (b.x - a.x) * (1.f - a.y) / (b.y - a.y); v.y = 1.f; }
        else if (out & 4) { v.x = a.x + (b.x - a.x) * (0.f - a.y) / (b.y - a.y); v.y = 0.f; }
        else if (out & 2) { v.y = a.y + (b.y - a.y) * (1.f - a.x) / (b.x - a.x); v.x = 1.f; }
        else if (out & 1) { v.y = a.y + (b.y - a.y) * (0.f - a.x) / (b.x - a.x); v.x = 0.f; }
        if (out == c0) { a = v; c0 = code(a); } else { b = v; c1 = code(b); }
    }
}
// reference edge.h:59-79
bool is_silhouette(const HostScene &hs, V3 p, const DrEdge &e) {
    const DrShape &sh = hs.shapes[e.shape];
    if (!sh.has_n) return true;
    if (e.f0 == -1 || e.f1 == -1) return true;
    V3 v0 = getv(hs, e.shape, e.v0), v1 = getv(hs, e.shape, e.v1);
    auto ns = [&](int f) {
        for (int k = 0; k < 3; k++) {
            int id = tri_index(hs, e.shape, f, k);
            if (id != e.v0 && id != e.v1) return getv(hs, e.shape, id);
        }
        return V3{0.f, 0.f, 0.f};
    };
    V3 a = ns(e.f0), b = ns(e.f1);
    V3 n0 = normalize(cross(v0 - a, v1 - a));
    V3 n1 = normalize(cross(v1 - b, v0 - b));
    bool f0 = dot(p - v0, n0) > 0.f && dot(p - v1, n0) > 0.f && dot(p - a, n0) > 0.f;
    bool f1 = dot(p - v0, n1) > 0.f && dot(p - v1, n1) > 0.f && dot(p - b, n1) > 0.f;
    return f0 != f1;
}

// reference scene.cpp:80-96: drop degenerate edges and edges between
// coplanar faces
bool edge_kept(const HostScene &hs, const DrEdge &e) {
    V3 dv = getv(hs, e.shape, e.v0) - getv(hs, e.shape, e.v1);
    if (std::fabs(dv.x) + std::fabs(dv.y) + std::fabs(dv.z) < 1e-10f) return false;
    if (e.f1 != -1) {
        float ang = std::fabs(dot(face_normal(hs, e.shape, e.f0), face_normal(hs, e.shape, e.f1)));
        if (ang == 1.f || ang == -1.f) return false;
    }
    return true;
}

}  // namespace

void primary_edges_for(const HostScene &hs, const DrCamera &cam, std::vector<DrEdge> &out, Dist &dist) {
    out.clear();
    std::vector<float> lens;
    V3 org = xfm_point(cam.c2w, V3{0.f, 0.f, 0.f});
    for (const DrEdge &e : hs.edges) {
        float len = 0.f;
        V2 a, b;
        if (project_edge(cam, getv(hs, e.shape, e.v0), getv(hs, e.shape, e.v1), a, b)) {
            if (clip_line(a, b)) {
                float dx = b.x - a.x, dy = b.y - a.y;
                len = std::sqrt(dx * dx + dy * dy);
            }
        }
        if (!is_silhouette(hs, org, e)) len = 0.f;
        if (len > 0.f) {
            out.push_back(e);
            lens.push_back(len);
        }
    }
    dist = lens.empty() ? Dist() : make_dist(lens);
}

void build_primary_edges(HostScene &hs) { primary_edges_for(hs, hs.cam, hs.prim_edges, hs.prim); }

DrCamera make_camera(const dr_camera &c) {
    DrCamera o;
    std::memcpy(o.c2w, c.cam_to_world, sizeof(float) * 16);
    std::memcpy(o.w2c, c.world_to_cam, sizeof(float) * 16);
    o.fov = c.fov_factor;
    o.aspect = c.aspect_ratio;
    o.clip_near = c.clip_near;
    o.fisheye = c.fisheye ? 1 : 0;
    return o;
}

void set_camera(HostScene &hs, const dr_camera &c) { hs.cam = make_camera(c); }

std::string build_host_scene(const dr_scene_desc *d, HostScene &hs) {
    if (!d) return "null scene description";
    if (d->num_shapes < 0 || d->num_materials < 0 || d->num_lights < 0) return "negative counts";
    if (d->num_lights == 0) return "the scene has no light (the reference indexes an empty light list)";
    set_camera(hs, d->camera);
    // gradient layout
    std::vector<size_t> doff(d->num_materials), soff(d->num_materials), roff(d->num_materials);
    std::vector<size_t> voff(d->num_shapes), noff(d->num_shapes), ioff(d->num_lights);
    hs.grad_size = dr_grad_layout_compute(d, nullptr, doff.data(), soff.data(), roff.data(), voff.data(),
                                          noff.data(), ioff.data());
    // materials
    auto tex = [&](const dr_texture &t, int channels, size_t goff, DrTex &o) -> std::string {
        if (t.channels != channels) return "material channel count mismatch";
        if (t.height < 0 || t.width < 0 || (t.height > 0 && t.width <= 0)) return "bad texture size";
        if (!t.data) return "null material data";
        o.off = (int)hs.texels.size();
        o.h = t.height;
        o.w = t.width;
        o.c = channels;
        o.grad_off = (long long)goff;
        size_t n = dr_texture_size(&t);
        hs.texels.insert(hs.texels.end(), t.data, t.data + n);
        return "";
    };
    for (int i = 0; i < d->num_materials; i++) {
        const dr_material &m = d->materials[i];
        DrMaterial o;
        std::memset(&o, 0, sizeof(o));
        std::string err;
        if (!(err = tex(m.diffuse_reflectance, 3, doff[i], o.diffuse)).empty()) return err;
        if (!(err = tex(m.specular_reflectance, 3, soff[i], o.specular)).empty()) return err;
        if (!(err = tex(m.roughness, 1, roff[i], o.roughness)).empty()) return err;
        for (int k = 0; k < 2; k++) {
            o.duv[k] = m.diffuse_uv_scale[k];
            o.suv[k] = m.specular_uv_scale[k];
            o.ruv[k] = m.roughness_uv_scale[k];
        }
        o.two_sided = m.two_sided ? 1 : 0;
        hs.materials.push_back(o);
    }
    // shapes
    int vbase = 0, tbase = 0;
    for (int i = 0; i < d->num_shapes; i++) {
        const dr_shape &s = d->shapes[i];
        if (!s.vertices || !s.indices || s.num_vertices < 0 || s.num_triangles < 0) return "bad shape arrays";
        if (s.material_id < 0 || s.material_id >= d->num_materials) return "shape material id out of range";
        DrShape o;
        std::memset(&o, 0, sizeof(o));
        o.vert_off = vbase;
        o.tri_off = tbase;
        o.nv = s.num_vertices;
        o.nt = s.num_triangles;
        o.has_uv = s.uvs != nullptr;
        o.has_n = s.normals != nullptr;
        o.mat = s.material_id;
        o.light = -1;
        o.vgrad_off = (long long)voff[i];
        o.ngrad_off = s.normals ? (long long)noff[i] : -1;
        hs.verts.insert(hs.verts.end(), s.vertices, s.vertices + 3 * (size_t)s.num_vertices);
        if (s.uvs) hs.uvs.insert(hs.uvs.end(), s.uvs, s.uvs + 2 * (size_t)s.num_vertices);
        else hs.uvs.insert(hs.uvs.end(), 2 * (size_t)s.num_vertices, 0.f);
        if (s.normals) hs.normals.insert(hs.normals.end(), s.normals, s.normals + 3 * (size_t)s.num_vertices);
        else hs.normals.insert(hs.normals.end(), 3 * (size_t)s.num_vertices, 0.f);
        for (int t = 0; t < s.num_triangles; t++) {
            for (int k = 0; k < 3; k++) {
                int v = s.indices[3 * (size_t)t + k];
                if (v < 0 || v >= s.num_vertices) return "triangle vertex index out of range";
                hs.tri_vidx.push_back(vbase + v);
            }
            hs.tri_shape.push_back(i);
        }
        vbase += s.num_vertices;
        tbase += s.num_triangles;
        hs.shapes.push_back(o);
    }
    hs.light_shapes.clear();
    for (int i = 0; i < d->num_lights; i++) {
        const dr_light &l = d->lights[i];
        if (l.shape_id < 0 || l.shape_id >= d->num_shapes) return "light shape id out of range";
        if (hs.shapes[l.shape_id].nt == 0) return "light shape has no triangles";
        hs.light_shapes.push_back(l.shape_id);
        hs.light_I.insert(hs.light_I.end(), l.intensity, l.intensity + 3);
    }
    hs.light_goff = ioff;
    build_lights(hs);
    return build_structures(hs);
}

// lights (reference scene.cpp:24-52): selection pmf by luminance, per light a
// triangle pmf by area, total area and bounding sphere
void build_lights(HostScene &hs) {
    std::vector<float> imp;
    hs.lights.clear();
    hs.light_func.clear();
    hs.light_cdf.clear();
    hs.light_total_intensity = 0.f;
    for (size_t i = 0; i < hs.light_shapes.size(); i++) {
        int sid = hs.light_shapes[i];
        DrShape &sh = hs.shapes[sid];
        sh.light = (int)i;
        DrLight o;
        std::memset(&o, 0, sizeof(o));
        o.shape = sid;
        for (int k = 0; k < 3; k++) o.I[k] = hs.light_I[3 * i + k];
        float lum = luminance(V3{o.I[0], o.I[1], o.I[2]});
        o.lum = lum;
        imp.push_back(lum);
        hs.light_total_intensity += lum;
        std::vector<float> areas;
        Box bounds;
        float asum = 0.f;
        for (int t = 0; t < sh.nt; t++) {
            V3 a = getv(hs, sid, tri_index(hs, sid, t, 0));
            V3 b = getv(hs, sid, tri_index(hs, sid, t, 1));
            V3 c = getv(hs, sid, tri_index(hs, sid, t, 2));
            float ar = 0.5f * length(cross(b - a, c - a));
            asum += ar;
            areas.push_back(ar);
            Box tb(a, b);
            tb.grow(c);
            bounds.grow(tb);
        }
        Dist dd = make_dist(areas);
        o.ntri = sh.nt;
        o.func_off = (int)hs.light_func.size();
        o.cdf_off = (int)hs.light_cdf.size();
        o.func_int = dd.func_int;
        hs.light_func.insert(hs.light_func.end(), dd.func.begin(), dd.func.end());
        hs.light_cdf.insert(hs.light_cdf.end(), dd.cdf.begin(), dd.cdf.end());
        o.area = asum;
        V3 c = (bounds.lo + bounds.hi) / 2.f;
        o.sphere_c[0] = c.x; o.sphere_c[1] = c.y; o.sphere_c[2] = c.z;
        o.sphere_r = length(bounds.hi - c);
        o.grad_off = (long long)hs.light_goff[i];
        hs.lights.push_back(o);
    }
    hs.light_sel = make_dist(imp);
}

namespace {
// Nodes of a tree grouped by depth, deepest level first (a bottom-up refit
// handles one level at a time).  children(i, out) lists node i's children.
template <class Children>
void tree_levels(int n, Children children, std::vector<int> &order, std::vector<int> &off) {
    order.clear();
    off.clear();
    if (n == 0) return;
    std::vector<int> depth(n, 0);
    int maxd = 0;
    for (int i = 0; i < n; i++) {  // preorder: a parent precedes its children
        int c[2];
        int k = children(i, c);
        for (int j = 0; j < k; j++) {
            depth[c[j]] = depth[i] + 1;
            maxd = std::max(maxd, depth[i] + 1);
        }
    }
    std::vector<int> cnt(maxd + 1, 0);
    for (int i = 0; i < n; i++) cnt[maxd - depth[i]]++;
    off.assign(maxd + 2, 0);
    for (int l = 0; l <= maxd; l++) off[l + 1] = off[l] + cnt[l];
    order.assign(n, 0);
    std::vector<int> pos(off.begin(), off.end() - 1);
    for (int i = 0; i < n; i++) order[pos[maxd - depth[i]]++] = i;
}
}  // namespace

std::string adopt_device_bvh(HostScene &hs, const std::vector<DrNode> &built, std::vector<DrTri> &tris) {
    int m = (int)built.size();
    if (m == 0) return "empty device BVH";
    // relabel in preorder (a parent before its children, the left subtree
    // before the right one), as the host builder lays nodes out
    std::vector<int> map(m, -1), order;
    order.reserve(m);
    std::vector<int> stack = {0};
    while (!stack.empty()) {
        int i = stack.back();
        stack.pop_back();
        if (i < 0 || i >= m || map[i] >= 0) return "device BVH is not a tree";
        map[i] = (int)order.size();
        order.push_back(i);
        for (int sl = 1; sl >= 0; sl--)
            if (built[i].d[2 + sl] == 0) stack.push_back(built[i].d[sl]);
    }
    if ((int)order.size() != m) return "device BVH is not a tree";
    std::vector<DrNode> nodes(m);
    for (int k = 0; k < m; k++) {
        DrNode N = built[order[k]];
        for (int sl = 0; sl < 2; sl++)
            if (N.d[2 + sl] == 0) N.d[sl] = map[N.d[sl]];
        nodes[k] = N;
    }
    std::vector<int> lv, off;
    tree_levels(m, [&](int i, int *c) {
        int k = 0;
        for (int sl = 0; sl < 2; sl++)
            if (nodes[i].d[2 + sl] == 0 && nodes[i].d[sl] > 0) c[k++] = nodes[i].d[sl];
        return k;
    }, lv, off);
    // the ray traversal stack holds at most one far child per interior level
    int levels = (int)off.size() - 1;
    if (levels > DR_STACK) return "device BVH deeper than the traversal stack";
    hs.nodes.swap(nodes);
    hs.tris.swap(tris);
    hs.bvh_lv.swap(lv);
    hs.bvh_lv_off.swap(off);
    hs.bvh_depth = levels;
    return "";
}

// Triangle BVH, mesh edges (reference scene.cpp:54-97), edge trees, their
// per-edge geometry records, the primary-edge distribution and the refit
// tables, from the shapes' current vertices.
std::string build_structures(HostScene &hs) {
    build_bvh(hs);
    // candidate edges, sorted by (shape, v0, v1) as the reference's std::map
    std::map<std::tuple<int, int, int>, DrEdge> emap;
    for (int i = 0; i < (int)hs.shapes.size(); i++) {
        const DrShape &s = hs.shapes[i];
        for (int t = 0; t < s.nt; t++) {
            int a = tri_index(hs, i, t, 0), b = tri_index(hs, i, t, 1), c = tri_index(hs, i, t, 2);
            auto add = [&](int x, int y) {
                if (x > y) std::swap(x, y);
                auto key = std::make_tuple(i, x, y);
                auto it = emap.find(key);
                if (it == emap.end()) emap[key] = DrEdge{i, x, y, t, -1};
                else it->second.f1 = t;
            };
            add(a, b);
            add(b, c);
            add(c, a);
        }
    }
    hs.edges.clear();
    hs.cand_edges.clear();
    hs.cand_keep.clear();
    for (auto &kv : emap) {
        const DrEdge &e = kv.second;
        bool keep = edge_kept(hs, e);
        hs.cand_edges.push_back(e);
        hs.cand_keep.push_back(keep ? 1 : 0);
        if (keep || hs.all_candidates) hs.edges.push_back(e);
    }
    build_primary_edges(hs);
    std::vector<DrEdge> sil, nrm;
    for (const DrEdge &e : hs.edges) {
        if (e.f0 == -1 || e.f1 == -1 || !hs.shapes[e.shape].has_n) sil.push_back(e);
        else nrm.push_back(e);
    }
    // edge ordinals in the random-number streams of edge samples have 27 bits
    // (and a traversal stack entry holds a node index or an edge count in 29)
    if (hs.edges.size() >= (1u << DR_EDGE_ORDINAL_BITS))
        return "too many edges for the edge trees (at most 2^27 - 1)";
    hs.sil_depth = build_edge_tree(hs, sil, false, hs.sil_nodes, hs.sil_edges);
    hs.nrm_depth = build_edge_tree(hs, nrm, true, hs.nrm_nodes, hs.nrm_edges);
    // a depth-first walk holds at most one pending sibling per level plus the
    // two children just pushed: depth + 1 entries
    if (hs.sil_depth + 1 > DR_TREE_STACK || hs.nrm_depth + 1 > DR_TREE_STACK)
        return "edge tree deeper than the traversal stack";
    // the ray traversal stack holds at most one far child per interior level
    if (hs.bvh_depth > DR_STACK) return "triangle BVH deeper than the traversal stack";
    build_geo_records(hs);
    // refit tables: levels of the three trees, edge ranges of the edge-tree nodes
    tree_levels((int)hs.nodes.size(), [&](int i, int *c) {
        int k = 0;
        const DrNode &n = hs.nodes[i];
        for (int sl = 0; sl < 2; sl++)
            if (n.d[2 + sl] == 0 && n.d[sl] > 0) c[k++] = n.d[sl];
        return k;
    }, hs.bvh_lv, hs.bvh_lv_off);
    auto etree = [&](const std::vector<DrEdgeNode> &nodes, std::vector<int> &lv, std::vector<int> &off,
                     std::vector<int> &rng) {
        tree_levels((int)nodes.size(), [&](int i, int *c) {
            if (nodes[i].nprims > 0) return 0;
            c[0] = i + 1;
            c[1] = nodes[i].offset;
            return 2;
        }, lv, off);
        int nn = (int)nodes.size();
        rng.assign(2 * (size_t)nn, 0);
        for (int i = nn - 1; i >= 0; i--) {
            if (nodes[i].nprims > 0) {
                rng[2 * i] = nodes[i].offset;
                rng[2 * i + 1] = nodes[i].offset + nodes[i].nprims;
            } else {
                rng[2 * i] = rng[2 * (i + 1)];
                rng[2 * i + 1] = rng[2 * nodes[i].offset + 1];
            }
        }
    };
    etree(hs.sil_nodes, hs.sil_lv, hs.sil_lv_off, hs.sil_rng);
    etree(hs.nrm_nodes, hs.nrm_lv, hs.nrm_lv_off, hs.nrm_rng);
    return "";
}

// the vertices the silhouette test of a normal-tree edge reads (edge.h:59-79):
// endpoints, and each face's vertex off the edge; the endpoints of
// silhouette-tree edges
void build_geo_records(HostScene &hs) {
    hs.nrm_geo.assign(16 * hs.nrm_edges.size(), 0.f);
    for (size_t i = 0; i < hs.nrm_edges.size(); i++) {
        const DrEdge &e = hs.nrm_edges[i];
        const DrShape &sh = hs.shapes[e.shape];
        int g0 = sh.vert_off + e.v0, g1 = sh.vert_off + e.v1;
        const int *i0 = &hs.tri_vidx[3 * (size_t)(sh.tri_off + e.f0)];
        const int *i1 = &hs.tri_vidx[3 * (size_t)(sh.tri_off + e.f1)];
        int a = i0[0] != g0 && i0[0] != g1 ? i0[0] : (i0[1] != g0 && i0[1] != g1 ? i0[1] : i0[2]);
        int b = i1[0] != g0 && i1[0] != g1 ? i1[0] : (i1[1] != g0 && i1[1] != g1 ? i1[1] : i1[2]);
        const int gv[4] = {g0, g1, a, b};
        for (int k = 0; k < 4; k++)
            for (int c = 0; c < 3; c++) hs.nrm_geo[16 * i + 4 * k + c] = hs.verts[3 * (size_t)gv[k] + c];
    }
    hs.sil_geo.assign(16 * hs.sil_edges.size(), 0.f);
    for (size_t i = 0; i < hs.sil_edges.size(); i++) {
        const DrEdge &e = hs.sil_edges[i];
        const DrShape &sh = hs.shapes[e.shape];
        const int gv[2] = {sh.vert_off + e.v0, sh.vert_off + e.v1};
        for (int k = 0; k < 2; k++)
            for (int c = 0; c < 3; c++) hs.sil_geo[16 * i + 4 * k + c] = hs.verts[3 * (size_t)gv[k] + c];
    }
}

void edge_features(const HostScene &hs, const DrCamera &cam, std::vector<float> &out) {
    const int F = 24;
    size_t n = hs.sil_edges.size() + hs.nrm_edges.size();
    out.assign(n * F, 0.f);
    // scene scale and centre from the vertex bounds
    V3 lo{1e30f, 1e30f, 1e30f}, hi{-1e30f, -1e30f, -1e30f};
    for (size_t i = 0; i + 2 < hs.verts.size(); i += 3) {
        V3 p{hs.verts[i], hs.verts[i + 1], hs.verts[i + 2]};
        lo = V3{std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
        hi = V3{std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
    }
    V3 ctr = 0.5f * (lo + hi);
    float scale = std::max(1e-6f, length(hi - lo));
    auto face_normal = [&](int shape, int f) {
        V3 a = getv(hs, shape, tri_index(hs, shape, f, 0)), b = getv(hs, shape, tri_index(hs, shape, f, 1)),
           c = getv(hs, shape, tri_index(hs, shape, f, 2));
        V3 nn = cross(b - a, c - a);
        float l = length(nn);
        return l > 0.f ? nn / l : V3{0.f, 0.f, 0.f};
    };
    auto fill = [&](const std::vector<DrEdge> &edges, const std::vector<DrEdgeNode> &nodes, size_t base) {
        // leaf of every edge of the tree, for the node features
        std::vector<int> leaf(edges.size(), 0);
        for (size_t i = 0; i < nodes.size(); i++)
            if (nodes[i].nprims > 0)
                for (int k = 0; k < nodes[i].nprims; k++) leaf[nodes[i].offset + k] = (int)i;
        float root = nodes.empty() ? 1.f : std::max(nodes[0].total_length, 1e-12f);
        for (size_t i = 0; i < edges.size(); i++) {
            const DrEdge &e = edges[i];
            float *o = &out[(base + i) * F];
            V3 v0 = getv(hs, e.shape, e.v0), v1 = getv(hs, e.shape, e.v1);
            V2 s0, s1;
            if (project_edge(cam, v0, v1, s0, s1)) {
                o[0] = s0.x; o[1] = s0.y; o[2] = s1.x; o[3] = s1.y; o[4] = 1.f;
            }
            V3 n0 = face_normal(e.shape, e.f0);
            V3 n1 = e.f1 >= 0 ? face_normal(e.shape, e.f1) : n0;
            o[5] = n0.x; o[6] = n0.y; o[7] = n0.z;
            o[8] = n1.x; o[9] = n1.y; o[10] = n1.z;
            o[11] = dot(n0, n1);
            o[12] = e.f1 < 0 ? 1.f : 0.f;
            o[13] = length(v1 - v0) / scale;
            V3 mid = 0.5f * (v0 + v1);
            o[14] = (mid.x - ctr.x) / scale; o[15] = (mid.y - ctr.y) / scale; o[16] = (mid.z - ctr.z) / scale;
            float best = 1e30f;
            V3 ldir{0.f, 1.f, 0.f};
            for (const DrLight &L : hs.lights) {
                V3 c{L.sphere_c[0], L.sphere_c[1], L.sphere_c[2]};
                float d = length(c - mid);
                if (d < best) {
                    best = d;
                    ldir = d > 0.f ? (c - mid) / d : V3{0.f, 1.f, 0.f};
                }
            }
            o[17] = best / scale;
            o[18] = dot(n0, ldir);
            o[19] = dot(n1, ldir);
            const DrEdgeNode &lf = nodes[leaf[i]];
            o[20] = lf.total_length / root * (float)edges.size();
            o[21] = std::acos(lf.cos_angle);
            o[22] = hs.shapes[e.shape].has_n ? 1.f : 0.f;
            o[23] = 1.f;
        }
    };
    fill(hs.sil_edges, hs.sil_nodes, 0);
    fill(hs.nrm_edges, hs.nrm_nodes, hs.sil_edges.size());
}

}  // namespace drh
