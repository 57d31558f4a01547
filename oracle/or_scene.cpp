// ORACLE — test infrastructure only.  Nothing in the product links this.
#include "or_scene.h"

#include <cstdio>
#include <cstring>
#include <stdexcept>

namespace oracle {

static Tex make_tex(const dr_texture &t) {
    Tex r;
    r.h = t.height;
    r.w = t.width;
    r.c = t.channels;
    size_t n = dr_texture_size(&t);
    r.data.assign(t.data, t.data + n);
    return r;
}

static void load_m4(M4T<float> &m, const float *a) {
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) m.m[i][j] = a[i * 4 + j];
}

// ---------------------------------------------------------------- BVH ----
namespace {
struct BuildPrim {
    AABB box;
    V3f c;
    int id;
};

int build_rec(Scene &sc, std::vector<BuildPrim> &p, int b, int e) {
    int node_id = (int)sc.bvh.size();
    sc.bvh.emplace_back();
    AABB box, cbox;
    for (int i = b; i < e; i++) {
        box = merge(box, p[i].box);
        cbox = merge(cbox, p[i].c);
    }
    sc.bvh[node_id].box = box;
    int n = e - b;
    auto make_leaf = [&]() {
        sc.bvh[node_id].first = (int)sc.tri_order.size();
        sc.bvh[node_id].count = n;
        for (int i = b; i < e; i++) sc.tri_order.push_back(p[i].id);
    };
    if (n <= 4) {
        make_leaf();
        return node_id;
    }
    // binned SAH over the widest centroid axis
    int axis = 0;
    V3f ext = cbox.hi - cbox.lo;
    if (ext[1] > ext[axis]) axis = 1;
    if (ext[2] > ext[axis]) axis = 2;
    if (!(ext[axis] > 0.f)) {
        // identical centroids: split in the middle of the list
        int mid = (b + e) / 2;
        int l = build_rec(sc, p, b, mid);
        int r = build_rec(sc, p, mid, e);
        sc.bvh[node_id].left = l;
        sc.bvh[node_id].right = r;
        return node_id;
    }
    const int NB = 16;
    AABB bb[NB];
    int bc[NB] = {0};
    float scale = NB / ext[axis];
    auto bin_of = [&](const BuildPrim &q) {
        int k = (int)((q.c[axis] - cbox.lo[axis]) * scale);
        return std::min(std::max(k, 0), NB - 1);
    };
    for (int i = b; i < e; i++) {
        int k = bin_of(p[i]);
        bb[k] = merge(bb[k], p[i].box);
        bc[k]++;
    }
    auto area = [](const AABB &a) {
        if (a.lo[0] > a.hi[0]) return 0.f;
        V3f d = a.hi - a.lo;
        return 2.f * (d[0] * d[1] + d[0] * d[2] + d[1] * d[2]);
    };
    float best = std::numeric_limits<float>::infinity();
    int best_k = -1;
    for (int k = 1; k < NB; k++) {
        AABB l, r;
        int nl = 0, nr = 0;
        for (int j = 0; j < k; j++) { l = merge(l, bb[j]); nl += bc[j]; }
        for (int j = k; j < NB; j++) { r = merge(r, bb[j]); nr += bc[j]; }
        if (nl == 0 || nr == 0) continue;
        float cost = area(l) * nl + area(r) * nr;
        if (cost < best) { best = cost; best_k = k; }
    }
    int mid;
    if (best_k < 0) {
        mid = (b + e) / 2;
        std::nth_element(p.begin() + b, p.begin() + mid, p.begin() + e,
                         [axis](const BuildPrim &x, const BuildPrim &y) { return x.c[axis] < y.c[axis]; });
    } else {
        auto it = std::partition(p.begin() + b, p.begin() + e,
                                 [&](const BuildPrim &q) { return bin_of(q) < best_k; });
        mid = (int)(it - p.begin());
        if (mid == b || mid == e) mid = (b + e) / 2;
    }
    int l = build_rec(sc, p, b, mid);
    int r = build_rec(sc, p, mid, e);
    sc.bvh[node_id].left = l;
    sc.bvh[node_id].right = r;
    return node_id;
}

inline bool box_hit(const AABB &bx, const V3f &org, const V3f &inv, float tmin, float tmax) {
    float t0 = tmin, t1 = tmax;
    for (int a = 0; a < 3; a++) {
        float tn = (bx.lo[a] - org[a]) * inv[a];
        float tf = (bx.hi[a] - org[a]) * inv[a];
        if (tn > tf) std::swap(tn, tf);
        if (std::isnan(tn) || std::isnan(tf)) continue;
        // conservative widening so that the exact triangle test decides
        tn = tn - std::fabs(tn) * 1e-5f - 1e-7f;
        tf = tf + std::fabs(tf) * 1e-5f + 1e-7f;
        t0 = tn > t0 ? tn : t0;
        t1 = tf < t1 ? tf : t1;
        if (t0 > t1) return false;
    }
    return true;
}

// The hit rule shared with the device kernels: Moller-Trumbore values from
// the reference formula, inside test u>=0, v>=0, u+v<=1, tmin < t < tmax;
// closest = smallest t, ties broken by the smaller global triangle index.
inline bool tri_test(const Scene &sc, int g, const V3f &org, const V3f &dir,
                     float tmin, float tmax, float &u, float &v, float &t) {
    mt_uvt(sc.vtx(g, 0), sc.vtx(g, 1), sc.vtx(g, 2), org, dir, u, v, t);
    return u >= 0.f && v >= 0.f && u + v <= 1.f && t > tmin && t < tmax;
}
}  // namespace

static void build_bvh(Scene &sc) {
    int nt = (int)sc.tri_shape.size();
    std::vector<BuildPrim> p(nt);
    for (int g = 0; g < nt; g++) {
        p[g].box = AABB(sc.vtx(g, 0), sc.vtx(g, 1), sc.vtx(g, 2));
        p[g].c = p[g].box.center();
        p[g].id = g;
    }
    sc.bvh.clear();
    sc.tri_order.clear();
    if (nt > 0) build_rec(sc, p, 0, nt);
}

thread_local uint64_t or_thread_rays = 0;

Hit nearest_hit(const Scene &sc, const V3f &org, const V3f &dir, float tmin, float tmax) {
    or_thread_rays++;
    Hit best;
    float best_t = tmax;
    int best_g = -1;
    float bu = 0, bv = 0;
    auto consider = [&](int g) {
        float u, v, t;
        if (tri_test(sc, g, org, dir, tmin, tmax, u, v, t)) {
            if (best_g < 0 || t < best_t || (t == best_t && g < best_g)) {
                best_t = t;
                best_g = g;
                bu = u;
                bv = v;
            }
        }
    };
    int nt = (int)sc.tri_shape.size();
    if (!sc.use_bvh || sc.bvh.empty()) {
        for (int g = 0; g < nt; g++) consider(g);
    } else {
        V3f inv = v3(1.f / dir[0], 1.f / dir[1], 1.f / dir[2]);
        int stack[128];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const BVHNode &n = sc.bvh[stack[--sp]];
            if (!box_hit(n.box, org, inv, tmin, best_t)) continue;
            if (n.left < 0) {
                for (int i = 0; i < n.count; i++) consider(sc.tri_order[n.first + i]);
            } else {
                stack[sp++] = n.right;
                stack[sp++] = n.left;
            }
        }
    }
    if (best_g >= 0) {
        best.shape = sc.tri_shape[best_g];
        best.tri = sc.tri_local[best_g];
        best.u = bu;
        best.v = bv;
        best.t = best_t;
    }
    return best;
}

bool any_hit(const Scene &sc, const V3f &org, const V3f &dir, float tmin, float tmax) {
    or_thread_rays++;
    int nt = (int)sc.tri_shape.size();
    float u, v, t;
    if (!sc.use_bvh || sc.bvh.empty()) {
        for (int g = 0; g < nt; g++)
            if (tri_test(sc, g, org, dir, tmin, tmax, u, v, t)) return true;
        return false;
    }
    V3f inv = v3(1.f / dir[0], 1.f / dir[1], 1.f / dir[2]);
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const BVHNode &n = sc.bvh[stack[--sp]];
        if (!box_hit(n.box, org, inv, tmin, tmax)) continue;
        if (n.left < 0) {
            for (int i = 0; i < n.count; i++)
                if (tri_test(sc, sc.tri_order[n.first + i], org, dir, tmin, tmax, u, v, t)) return true;
        } else {
            stack[sp++] = n.right;
            stack[sp++] = n.left;
        }
    }
    return false;
}

// scene.cpp:147-175: tnear = 1e-3 * d, tfar = (1 - 1e-3) * d
bool occluded(const Scene &sc, const V3f &org, const V3f &dir, float max_t) {
    return any_hit(sc, org, dir, max_t * 1e-3f, max_t * (1.f - 1e-3f));
}

// ----------------------------------------------------------- edge trees ---
namespace {
struct EPrim {
    int id;
    AABB box;
    V3f c;
    V3f n0, n1;
    AABB dbox;
    V3f dc;
};

int build_edge_rec(const std::vector<Edge> &edges, std::vector<EPrim> &p, int b, int e,
                   bool directional, std::vector<EdgeNode> &nodes, std::vector<Edge> &ordered,
                   std::vector<int> &child1, std::vector<V3f> &on) {
    int id = (int)nodes.size();
    nodes.emplace_back();
    child1.push_back(-1);
    AABB box, dbox;
    for (int i = b; i < e; i++) {
        box = merge(box, p[i].box);
        if (directional) dbox = merge(dbox, p[i].dbox);
    }
    // the normal cone is set bottom-up once the tree is built
    nodes[id].bounds = box;
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
    AABB cb, db;
    for (int i = b; i < e; i++) {
        cb = merge(cb, p[i].c);
        if (directional) db = merge(db, p[i].dc);
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
    const AABB &used = spatial ? cb : db;
    if (dim >= 3) dim -= 3;
    if (used.hi[dim] == used.lo[dim]) {
        leaf();
        return id;
    }
    int mid = (b + e) / 2;
    std::nth_element(p.begin() + b, p.begin() + mid, p.begin() + e,
                     [dim, spatial](const EPrim &x, const EPrim &y) {
                         return spatial ? x.c[dim] < y.c[dim] : x.dc[dim] < y.dc[dim];
                     });
    nodes[id].nprims = 0;
    build_edge_rec(edges, p, b, mid, directional, nodes, ordered, child1, on);
    int r = build_edge_rec(edges, p, mid, e, directional, nodes, ordered, child1, on);
    child1[id] = r;
    return id;
}
}  // namespace

static EdgeTree build_edge_tree(const Scene &sc, const std::vector<Edge> &edges, bool directional) {
    EdgeTree tree;
    if (edges.empty()) return tree;
    std::vector<EPrim> p(edges.size());
    AABB world;
    for (size_t i = 0; i < edges.size(); i++) {
        const Shape &s = sc.shapes[edges[i].shape];
        p[i].id = (int)i;
        p[i].box = AABB(s.vert(edges[i].v0), s.vert(edges[i].v1));
        p[i].c = .5f * p[i].box.lo + .5f * p[i].box.hi;
        // silhouette edges carry no normals (their cone is never read)
        p[i].n0 = v3(0.f, 0.f, 0.f);
        p[i].n1 = v3(0.f, 0.f, 0.f);
        world = merge(world, p[i].box);
    }
    if (directional) {
        float factor = length(world.hi - world.lo) / 8.f;
        for (size_t i = 0; i < edges.size(); i++) {
            const Edge &ed = edges[i];
            const Shape &s = sc.shapes[ed.shape];
            V3f sn = normalize(s.normal(ed.v0) + s.normal(ed.v1));
            V3f n0 = s.face_normal(ed.f0);
            V3f n1 = s.face_normal(ed.f1);
            if (dot(n0, sn) < 0.f) n0 = -n0;
            if (dot(n1, sn) < 0.f) n1 = -n1;
            p[i].n0 = n0;
            p[i].n1 = n1;
            p[i].dbox = AABB(n0 * factor, n1 * factor);
            p[i].dc = 0.5f * (p[i].dbox.lo + p[i].dbox.hi);
        }
    }
    std::vector<int> child1;
    std::vector<V3f> on;  // per ordered edge: its two face normals
    build_edge_rec(edges, p, 0, (int)edges.size(), directional, tree.nodes, tree.edges, child1, on);
    // Bottom-up (children follow parents): interior offset = second child,
    // total length, the sum of the face normals (leaf: in edge order;
    // interior: left + right) and the range of the node's edges in the
    // ordered list.  The cone: axis = normalised sum, angle = acos of the
    // least cosine to it over the node's edges.
    int nn = (int)tree.nodes.size();
    std::vector<V3f> nsum(nn);
    std::vector<int> first(nn), last(nn);
    for (int i = nn - 1; i >= 0; i--) {
        EdgeNode &n = tree.nodes[i];
        if (n.nprims > 0) {
            n.total_length = 0.f;
            nsum[i] = v3(0.f, 0.f, 0.f);
            for (int k = 0; k < n.nprims; k++) {
                const Edge &ed = tree.edges[n.offset + k];
                const Shape &s = sc.shapes[ed.shape];
                n.total_length += length(s.vert(ed.v1) - s.vert(ed.v0));
                nsum[i] += on[2 * (n.offset + k)];
                nsum[i] += on[2 * (n.offset + k) + 1];
            }
            first[i] = n.offset;
            last[i] = n.offset + n.nprims;
        } else {
            n.offset = child1[i];
            n.total_length = tree.nodes[i + 1].total_length + tree.nodes[n.offset].total_length;
            nsum[i] = nsum[i + 1] + nsum[n.offset];
            first[i] = first[i + 1];
            last[i] = last[n.offset];
        }
        V3f axis = normalize(nsum[i]);
        float cmin = 1.f;
        for (int k = first[i]; k < last[i]; k++) {
            cmin = std::min(cmin, dot(on[2 * k], axis));
            cmin = std::min(cmin, dot(on[2 * k + 1], axis));
        }
        n.cone_axis = axis;
        n.cone_angle = std::acos(cmin);
    }
    return tree;
}

// ------------------------------------------------------- primary edges ---
// Cohen-Sutherland clip of a screen segment to [0,1]^2 (line_clip.h:13-102)
static bool clip_line(V2f &a, V2f &b) {
    auto code = [](const V2f &v) {
        int c = 0;
        if (v[0] < 0.f) c |= 1; else if (v[0] > 1.f) c |= 2;
        if (v[1] < 0.f) c |= 4; else if (v[1] > 1.f) c |= 8;
        return c;
    };
    int c0 = code(a), c1 = code(b);
    while (true) {
        if (!(c0 | c1)) return true;
        if (c0 & c1) return false;
        int out = c0 ? c0 : c1;
        V2f v = v2(0.f, 0.f);
        if (out & 8) { v[0] = a[0] + (b[0] - a[0]) * (1.f - a[1]) / (b[1] - a[1]); v[1] = 1.f; }
        else if (out & 4) { v[0] = a[0] + (b[0] - a[0]) * (0.f - a[1]) / (b[1] - a[1]); v[1] = 0.f; }
        else if (out & 2) { v[1] = a[1] + (b[1] - a[1]) * (1.f - a[0]) / (b[0] - a[0]); v[0] = 1.f; }
        else if (out & 1) { v[1] = a[1] + (b[1] - a[1]) * (0.f - a[0]) / (b[0] - a[0]); v[0] = 0.f; }
        if (out == c0) { a = v; c0 = code(a); } else { b = v; c1 = code(b); }
    }
}

bool probe_clip_line(V2f &a, V2f &b) { return clip_line(a, b); }

bool project_edge(const Camera &cam, V3f p0, V3f p1, V2f &s0, V2f &s1);
bool is_silhouette(const Scene &sc, const V3f &p, const Edge &e);

void build_primary_edges(Scene &sc) {
    sc.primary_edges.clear();
    std::vector<float> lens;
    V3f org = xfm_point(sc.cam.c2w, v3(0.f, 0.f, 0.f));
    for (const Edge &e : sc.edges) {
        const Shape &s = sc.shapes[e.shape];
        float dist = 0.f;
        V2f a, b;
        if (project_edge(sc.cam, s.vert(e.v0), s.vert(e.v1), a, b)) {
            if (clip_line(a, b)) dist = length2(b - a);
        }
        if (!is_silhouette(sc, org, e)) dist = 0.f;
        if (dist > 0.f) {
            sc.primary_edges.push_back(e);
            lens.push_back(dist);
        }
    }
    sc.primary_dist = lens.empty() ? Dist1D() : Dist1D(lens.data(), (int)lens.size());
}

// ------------------------------------------------------------ scene -----
void build_scene(Scene &sc, const dr_scene_desc *d) {
    const dr_camera &c = d->camera;
    load_m4(sc.cam.c2w, c.cam_to_world);
    load_m4(sc.cam.w2c, c.world_to_cam);
    load_m4(sc.cam.s2c, c.sample_to_cam);
    load_m4(sc.cam.c2s, c.cam_to_sample);
    sc.cam.fov = c.fov_factor;
    sc.cam.aspect = c.aspect_ratio;
    sc.cam.clip_near = c.clip_near;
    sc.cam.fisheye = c.fisheye != 0;
    for (int i = 0; i < d->num_materials; i++) {
        const dr_material &m = d->materials[i];
        Material o;
        o.diffuse = make_tex(m.diffuse_reflectance);
        o.specular = make_tex(m.specular_reflectance);
        o.roughness = make_tex(m.roughness);
        for (int k = 0; k < 2; k++) {
            o.duv[k] = m.diffuse_uv_scale[k];
            o.suv[k] = m.specular_uv_scale[k];
            o.ruv[k] = m.roughness_uv_scale[k];
        }
        o.two_sided = m.two_sided != 0;
        sc.mats.push_back(o);
    }
    for (int i = 0; i < d->num_shapes; i++) {
        const dr_shape &s = d->shapes[i];
        Shape o;
        o.nv = s.num_vertices;
        o.nt = s.num_triangles;
        o.V.assign(s.vertices, s.vertices + 3 * (size_t)s.num_vertices);
        o.I.assign(s.indices, s.indices + 3 * (size_t)s.num_triangles);
        if (s.uvs) { o.has_uv = true; o.UV.assign(s.uvs, s.uvs + 2 * (size_t)s.num_vertices); }
        if (s.normals) { o.has_n = true; o.N.assign(s.normals, s.normals + 3 * (size_t)s.num_vertices); }
        o.mat = s.material_id;
        o.light = s.light_id;
        sc.shapes.push_back(o);
    }
    for (int i = 0; i < d->num_lights; i++) {
        Light l;
        l.shape = d->lights[i].shape_id;
        for (int k = 0; k < 3; k++) l.I[k] = d->lights[i].intensity[k];
        sc.lights.push_back(l);
        sc.shapes[l.shape].light = i;
    }
    // lights (scene.cpp:24-52)
    std::vector<float> imp;
    sc.total_light_intensity = 0.f;
    for (const Light &l : sc.lights) {
        float lum = luminance(v3(l.I[0], l.I[1], l.I[2]));
        imp.push_back(lum);
        sc.total_light_intensity += lum;
        const Shape &s = sc.shapes[l.shape];
        std::vector<float> areas;
        AABB bounds;
        float asum = 0.f;
        for (int t = 0; t < s.nt; t++) {
            V3f a = s.vert(s.idx(t, 0)), b = s.vert(s.idx(t, 1)), cc = s.vert(s.idx(t, 2));
            float ar = 0.5f * length(cross(b - a, cc - a));
            asum += ar;
            areas.push_back(ar);
            bounds = merge(bounds, AABB(a, b, cc));
        }
        sc.light_tri.push_back(Dist1D(areas.data(), (int)areas.size()));
        sc.light_area.push_back(asum);
        sc.light_bounds.push_back(bounds);
    }
    if (!imp.empty()) sc.light_dist = Dist1D(imp.data(), (int)imp.size());
    // triangles
    int g = 0;
    for (int i = 0; i < (int)sc.shapes.size(); i++) {
        sc.tri_base.push_back(g);
        for (int t = 0; t < sc.shapes[i].nt; t++) {
            sc.tri_shape.push_back(i);
            sc.tri_local.push_back(t);
            g++;
        }
    }
    build_bvh(sc);
    // edges (scene.cpp:54-97): key (shape, min v, max v); f1 = second face
    std::map<std::tuple<int, int, int>, Edge> emap;
    for (int i = 0; i < (int)sc.shapes.size(); i++) {
        const Shape &s = sc.shapes[i];
        for (int t = 0; t < s.nt; t++) {
            int a = s.idx(t, 0), b = s.idx(t, 1), cc = s.idx(t, 2);
            auto add = [&](int x, int y) {
                if (x > y) std::swap(x, y);
                auto key = std::make_tuple(i, x, y);
                auto it = emap.find(key);
                if (it == emap.end()) emap[key] = Edge{i, x, y, t, -1};
                else it->second.f1 = t;
            };
            add(a, b);
            add(b, cc);
            add(cc, a);
        }
    }
    sc.edges.clear();
    for (auto &kv : emap) {
        const Edge &e = kv.second;
        const Shape &s = sc.shapes[e.shape];
        if (abs_sum(s.vert(e.v0) - s.vert(e.v1)) < 1e-10f) continue;
        if (e.f1 != -1) {
            float ang = std::fabs(dot(s.face_normal(e.f0), s.face_normal(e.f1)));
            if (ang == 1.f || ang == -1.f) continue;
        }
        sc.edges.push_back(e);
    }
    build_primary_edges(sc);
    std::vector<Edge> sil, nrm;
    for (const Edge &e : sc.edges) {
        if (e.f0 == -1 || e.f1 == -1 || !sc.shapes[e.shape].has_n) sil.push_back(e);
        else nrm.push_back(e);
    }
    sc.sil_tree = build_edge_tree(sc, sil, false);
    sc.nrm_tree = build_edge_tree(sc, nrm, true);
}

// ------------------------------------------------- cone / box overlap ----
// Box-cone test of Eberly's GeometricTools (IntrAlignedBox3Cone3), the
// variant the reference uses (aabb.cpp:98-311): quick rejection against the
// cone's supporting plane, then a spherical-polygon search over the box
// silhouette seen from the apex.  The polygon table (box corners in CCW order
// for each of the 27 apex regions) is geometric data and must be the same.
// Algorithm and case table after the Geometric Tools Engine, David Eberly,
// Geometric Tools LLC, distributed under the Boost Software License 1.0.
namespace {
struct ConePoly {
    int n;
    int idx[6];
};
const ConePoly kConePolys[27] = {
    {6, {1, 5, 4, 6, 2, 3}}, {6, {0, 2, 3, 1, 5, 4}}, {6, {0, 2, 3, 7, 5, 4}},
    {6, {0, 4, 6, 2, 3, 1}}, {4, {0, 2, 3, 1}},       {6, {0, 2, 3, 7, 5, 1}},
    {6, {0, 4, 6, 7, 3, 1}}, {6, {0, 2, 6, 7, 3, 1}}, {6, {0, 2, 6, 7, 5, 1}},
    {6, {0, 1, 5, 4, 6, 2}}, {4, {0, 1, 5, 4}},       {6, {0, 1, 3, 7, 5, 4}},
    {4, {0, 4, 6, 2}},       {0, {0}},                {4, {1, 3, 7, 5}},
    {6, {0, 4, 6, 7, 3, 2}}, {4, {2, 6, 7, 3}},       {6, {1, 3, 2, 6, 7, 5}},
    {6, {0, 1, 5, 7, 6, 2}}, {6, {0, 1, 5, 7, 6, 4}}, {6, {0, 1, 3, 7, 6, 4}},
    {6, {0, 4, 5, 7, 6, 2}}, {4, {4, 5, 7, 6}},       {6, {1, 3, 7, 6, 4, 5}},
    {6, {0, 4, 5, 7, 3, 2}}, {6, {2, 6, 4, 5, 7, 3}}, {6, {1, 3, 2, 6, 4, 5}}};

bool edge_max_test(const V3f &D, float c2, const V3f &maxX, const V3f &maxPmV, float maxSqrLen,
                   float maxDd, const V3f &CmV, int jA, int jB, float sign_flip) {
    int jDiff = jB - jA;
    float s = jDiff > 0 ? 1.f : -1.f;
    int k0 = std::abs(jDiff) >> 1;
    float fder = sign_flip * s * (D[k0] * maxSqrLen - maxDd * maxPmV[k0]);
    if (!(fder > 0)) return false;
    int k1 = (k0 + 1) % 3, k2 = (k1 + 1) % 3;
    float numer = maxPmV[k1] * maxPmV[k1] + maxPmV[k2] * maxPmV[k2];
    float denom = D[k1] * maxPmV[k1] + D[k2] * maxPmV[k2];
    V3f M;
    M[k0] = numer * D[k0];
    M[k1] = denom * (maxX[k1] + CmV[k1]);
    M[k2] = denom * (maxX[k2] + CmV[k2]);
    float DdM = dot(D, M);
    float q = DdM * DdM - c2 * dot(M, M);
    if (q > 0) return true;
    float det = s * (D[k1] * maxPmV[k2] - D[k2] * maxPmV[k1]);
    return det <= 0;
}
}  // namespace

bool aabb_cone_intersect(const AABB &box, const Cone &cone) {
    if (cone.cos_angle <= 0.f) return true;
    V3f center = 0.5f * (box.lo + box.hi);
    V3f ext = 0.5f * (box.hi - box.lo);
    V3f CmV = center - cone.org;
    const V3f &D = cone.dir;
    float DdCmV = dot(D, CmV);
    float radius = ext[0] * std::fabs(D[0]) + ext[1] * std::fabs(D[1]) + ext[2] * std::fabs(D[2]);
    if (DdCmV + radius <= 0) return false;
    int ix[3];
    for (int i = 0; i < 3; i++) ix[i] = CmV[i] < -ext[i] ? 2 : (CmV[i] > ext[i] ? 0 : 1);
    int lookup = ix[0] + 3 * ix[1] + 9 * ix[2];
    if (lookup == 13) return true;
    const ConePoly &poly = kConePolys[lookup];
    float c2 = cone.cos_angle * cone.cos_angle;
    V3f X[8], PmV[8];
    float DdPmV[8], sqrDd[8], sqrLen[8];
    int iMax = -1, jMax = -1;
    for (int i = 0; i < poly.n; i++) {
        int j = poly.idx[i];
        X[j] = v3((j & 1) ? ext[0] : -ext[0], (j & 2) ? ext[1] : -ext[1], (j & 4) ? ext[2] : -ext[2]);
        DdPmV[j] = dot(D, X[j]) + DdCmV;
        if (DdPmV[j] > 0) {
            PmV[j] = X[j] + CmV;
            sqrDd[j] = DdPmV[j] * DdPmV[j];
            sqrLen[j] = dot(PmV[j], PmV[j]);
            float q = sqrDd[j] - c2 * sqrLen[j];
            if (q > 0) return true;
            if (iMax == -1 || sqrDd[j] * sqrLen[jMax] > sqrDd[jMax] * sqrLen[j]) {
                iMax = i;
                jMax = j;
            }
        }
    }
    if (iMax == -1) return false;
    int iNext = iMax < poly.n - 1 ? iMax + 1 : 0;
    int jNext = poly.idx[iNext];
    {
        int jDiff = jNext - jMax;
        float s = jDiff > 0 ? 1.f : -1.f;
        int k0 = std::abs(jDiff) >> 1;
        float fder = s * (D[k0] * sqrLen[jMax] - DdPmV[jMax] * PmV[jMax][k0]);
        if (fder > 0) return edge_max_test(D, c2, X[jMax], PmV[jMax], sqrLen[jMax], DdPmV[jMax], CmV, jMax, jNext, 1.f);
    }
    int iPrev = iMax > 0 ? iMax - 1 : poly.n - 1;
    int jPrev = poly.idx[iPrev];
    {
        int jDiff = jMax - jPrev;
        float s = jDiff > 0 ? 1.f : -1.f;
        int k0 = std::abs(jDiff) >> 1;
        float fder = -s * (D[k0] * sqrLen[jMax] - DdPmV[jMax] * PmV[jMax][k0]);
        if (fder > 0) return edge_max_test(D, c2, X[jMax], PmV[jMax], sqrLen[jMax], DdPmV[jMax], CmV, jPrev, jMax, -1.f);
    }
    return false;
}

}  // namespace oracle
