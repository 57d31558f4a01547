// ORACLE — test infrastructure only.  Nothing in the product links this.
//
// Scene preprocessing restated from the reference:
//   light distributions          scene.cpp:24-52, distribution.h:7-59
//   edge extraction + filtering  scene.cpp:54-97
//   edge trees                   edge_tree.cpp:49-316
//   closest / any hit            scene.cpp:147-202 (Embree there; a binned-SAH
//                                BVH with an exact tie rule here, plus a
//                                brute-force mode used to check it)
#pragma once

#include "../include/delta_ray.h"
#include "or_math.h"

#include <algorithm>
#include <cfloat>
#include <cstdint>
#include <limits>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

namespace oracle {

struct Tex {
    std::vector<float> data;
    int h = 0, w = 0, c = 0;
};
struct Material {
    Tex diffuse, specular, roughness;
    float duv[2], suv[2], ruv[2];
    bool two_sided;
};
struct Shape {
    std::vector<float> V, UV, N;
    std::vector<int> I;
    int nv = 0, nt = 0;
    int mat = 0, light = -1;
    bool has_uv = false, has_n = false;
    V3f vert(int i) const { return v3(V[3 * i], V[3 * i + 1], V[3 * i + 2]); }
    V3f normal(int i) const { return v3(N[3 * i], N[3 * i + 1], N[3 * i + 2]); }
    V2f uv(int i) const { return v2(UV[2 * i], UV[2 * i + 1]); }
    int idx(int t, int k) const { return I[3 * t + k]; }
    V3f face_normal(int t) const {
        V3f a = vert(idx(t, 0)), b = vert(idx(t, 1)), c = vert(idx(t, 2));
        return normalize(cross(b - a, c - a));
    }
};
struct Light {
    int shape;
    float I[3];
};
struct Camera {
    M4T<float> c2w, w2c, s2c, c2s;
    float fov, aspect, clip_near;
    bool fisheye = false;  // equi-angular projection (camera.h:51-71)
};

// pbrt-style piecewise-constant 1-D distribution (distribution.h:7-59)
struct Dist1D {
    std::vector<float> func, cdf;
    float func_int = 0.f;
    int n = 0;
    Dist1D() {}
    Dist1D(const float *f, int count) {
        n = count;
        func.assign(f, f + count);
        cdf.assign(count + 1, 0.f);
        for (int i = 1; i < count + 1; i++) cdf[i] = cdf[i - 1] + func[i - 1] / count;
        func_int = cdf[count];
        if (func_int == 0.f) {
            for (int i = 1; i < count + 1; i++) cdf[i] = float(i) / float(count);
        } else {
            for (int i = 1; i < count + 1; i++) cdf[i] /= func_int;
        }
    }
    int sample_discrete(float u, float *pmf = nullptr) const {
        int off = (int)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) - 1;
        off = std::min(std::max(off, 0), n - 1);
        if (pmf) *pmf = func[off] / (func_int * n);
        return off;
    }
    float pmf(int off) const { return func[off] / (func_int * n); }
};

struct AABB {
    V3f lo, hi;
    AABB() {
        float inf = std::numeric_limits<float>::infinity();
        lo = v3(inf, inf, inf);
        hi = v3(-inf, -inf, -inf);
    }
    AABB(const V3f &a, const V3f &b) {
        for (int i = 0; i < 3; i++) {
            lo[i] = std::min(a[i], b[i]);
            hi[i] = std::max(a[i], b[i]);
        }
    }
    AABB(const V3f &a, const V3f &b, const V3f &c) {
        for (int i = 0; i < 3; i++) {
            lo[i] = std::min(std::min(a[i], b[i]), c[i]);
            hi[i] = std::max(std::max(a[i], b[i]), c[i]);
        }
    }
    V3f center() const { return 0.5f * (lo + hi); }
    V3f corner(int i) const {
        return v3((i & 1) ? hi[0] : lo[0], (i & 2) ? hi[1] : lo[1], (i & 4) ? hi[2] : lo[2]);
    }
    bool inside(const V3f &p) const {
        return p[0] >= lo[0] && p[0] <= hi[0] && p[1] >= lo[1] && p[1] <= hi[1] &&
               p[2] >= lo[2] && p[2] <= hi[2];
    }
    void sphere(V3f &c, float &r) const {
        c = (lo + hi) / 2.f;
        r = length(hi - c);
    }
};
inline AABB merge(const AABB &a, const AABB &b) {
    AABB r;
    for (int i = 0; i < 3; i++) {
        r.lo[i] = std::min(a.lo[i], b.lo[i]);
        r.hi[i] = std::max(a.hi[i], b.hi[i]);
    }
    return r;
}
inline AABB merge(const AABB &a, const V3f &p) { return merge(a, AABB(p, p)); }

struct Cone {
    V3f org, dir;
    float cos_angle;
};
bool aabb_cone_intersect(const AABB &box, const Cone &cone);

struct Edge {
    int shape, v0, v1, f0, f1;
};

struct EdgeNode {
    AABB bounds;
    V3f cone_axis;
    float cone_angle;
    int offset;  // leaf: first edge; interior: second child
    int nprims;  // 0 = interior
    float total_length;
};
struct EdgeTree {
    std::vector<EdgeNode> nodes;
    std::vector<Edge> edges;
};

// Triangle BVH (binned SAH, up to 4 triangles per leaf).
struct BVHNode {
    AABB box;
    int left = -1, right = -1;  // interior children
    int first = 0, count = 0;   // leaf range into tri_order
};

struct Scene;

struct Hit {
    int shape = -1, tri = -1;
    float u = 0.f, v = 0.f, t = 0.f;
};

struct Scene {
    Camera cam;
    std::vector<Shape> shapes;
    std::vector<Material> mats;
    std::vector<Light> lights;
    // lights
    std::vector<Dist1D> light_tri;
    Dist1D light_dist;
    std::vector<AABB> light_bounds;
    std::vector<float> light_area;
    float total_light_intensity = 0.f;
    // edges
    std::vector<Edge> edges;
    std::vector<Edge> primary_edges;
    Dist1D primary_dist;
    EdgeTree sil_tree, nrm_tree;
    // triangles (global index = tri_base[shape] + tri)
    std::vector<int> tri_base;
    std::vector<int> tri_shape, tri_local;
    std::vector<BVHNode> bvh;
    std::vector<int> tri_order;
    bool use_bvh = true;

    V3f vtx(int gtri, int k) const {
        const Shape &s = shapes[tri_shape[gtri]];
        return s.vert(s.idx(tri_local[gtri], k));
    }
};

void build_scene(Scene &sc, const dr_scene_desc *d);
void build_primary_edges(Scene &sc);
// ray queries made by this thread (nearest_hit + any_hit), for the CPU
// baseline's rays/s
extern thread_local uint64_t or_thread_rays;
Hit nearest_hit(const Scene &sc, const V3f &org, const V3f &dir, float tmin, float tmax);
bool any_hit(const Scene &sc, const V3f &org, const V3f &dir, float tmin, float tmax);
bool occluded(const Scene &sc, const V3f &org, const V3f &dir, float max_t);

// Moller-Trumbore exactly as the reference's intersect.h:71-90 (no inside
// test there; the caller applies one).
template <class T>
inline void mt_uvt(const V3<T> &v0, const V3<T> &v1, const V3<T> &v2,
                   const V3<T> &org, const V3<T> &dir, T &u, T &v, T &t) {
    V3<T> e1 = v1 - v0;
    V3<T> e2 = v2 - v0;
    V3<T> pvec = cross(dir, e2);
    T divisor = dot(pvec, e1);
    if (divisor == 0.f) divisor = T(1e-8f);
    V3<T> s = org - v0;
    u = dot(s, pvec) / divisor;
    V3<T> qvec = cross(s, e1);
    v = dot(dir, qvec) / divisor;
    t = dot(e2, qvec) / divisor;
}

}  // namespace oracle
