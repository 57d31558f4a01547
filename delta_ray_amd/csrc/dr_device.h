// Device functions of the forward estimator.  Each follows the reference
// function cited beside it; the adjoints live in dr_adjoint.h.
#pragma once

#include "dr_math.h"
#include "dr_types.h"

#define DR_INF __builtin_huge_valf()

struct Isect {
    int shape, tri;  // tri: shape-local id
    int gtri;        // global triangle id, -1 = miss
};
DR_INLINE Isect no_hit() { return Isect{-1, -1, -1}; }
DR_INLINE bool valid(const Isect &i) { return i.gtri >= 0; }

struct SPoint {
    F3 pos, ng;
    Frame fr;
    F2 uv;
};

// ---------------------------------------------------------------- BVH -----
// Moller-Trumbore values with the reference formula (intersect.h:71-90) on
// precomputed e1 = v1 - v0, e2 = v2 - v0 (same rounding as computing them
// here).  Inside test u>=0, v>=0, u+v<=1 and tmin < t < tmax; closest =
// smallest t, ties to the smaller global triangle id.  The CPU oracle applies
// the same rule, so hits agree bit for bit.
DR_INLINE bool tri_hit(const DrTri &T, F3 org, F3 dir, float tmin, float tmax, float &u, float &v, float &t) {
    F3 v0 = F3{T.v0[0], T.v0[1], T.v0[2]};
    F3 e1 = F3{T.e1[0], T.e1[1], T.e1[2]};
    F3 e2 = F3{T.e2[0], T.e2[1], T.e2[2]};
    F3 pvec = cross(dir, e2);
    float divisor = dot(pvec, e1);
    if (divisor == 0.f) divisor = 1e-8f;
    F3 s = org - v0;
    u = dot(s, pvec) / divisor;
    F3 qvec = cross(s, e1);
    v = dot(dir, qvec) / divisor;
    t = dot(e2, qvec) / divisor;
    return u >= 0.f && v >= 0.f && u + v <= 1.f && t > tmin && t < tmax;
}

// Slab test widened by a relative 1e-5 (+1e-7) so that box rounding can
// never hide a triangle the exact test would accept.  An axis whose slab
// distances are NaN (0 * inf: the ray lies in a slab plane), or whose entry
// distance is +inf (exit -inf), does not constrain the interval.  The
// widening is monotone, so the entry and exit are widened once, after the
// max / min over the axes, which is the same as widening every axis.
DR_INLINE bool box_hit(float lx, float ly, float lz, float hx, float hy, float hz, F3 org, F3 inv, float tmin,
                       float tmax, float &tenter) {
    tenter = tmax;
    // an empty slot (lo = +inf, hi = -inf) never hits
    if (!(lx <= hx)) return false;
    float tnx = (lx - org.x) * inv.x, tfx = (hx - org.x) * inv.x;
    float tny = (ly - org.y) * inv.y, tfy = (hy - org.y) * inv.y;
    float tnz = (lz - org.z) * inv.z, tfz = (hz - org.z) * inv.z;
    bool bx = tnx != tnx || tfx != tfx, by = tny != tny || tfy != tfy, bz = tnz != tnz || tfz != tfz;
    float mnx = fminf(tnx, tfx), mxx = fmaxf(tnx, tfx);
    float mny = fminf(tny, tfy), mxy = fmaxf(tny, tfy);
    float mnz = fminf(tnz, tfz), mxz = fmaxf(tnz, tfz);
    mnx = (bx || mnx == DR_INF) ? -DR_INF : mnx;
    mxx = (bx || mxx == -DR_INF) ? DR_INF : mxx;
    mny = (by || mny == DR_INF) ? -DR_INF : mny;
    mxy = (by || mxy == -DR_INF) ? DR_INF : mxy;
    mnz = (bz || mnz == DR_INF) ? -DR_INF : mnz;
    mxz = (bz || mxz == -DR_INF) ? DR_INF : mxz;
    float mn = fmaxf(fmaxf(mnx, mny), mnz), mx = fminf(fminf(mxx, mxy), mxz);
    mn = mn - fabsf(mn) * 1e-5f - 1e-7f;
    mx = mx + fabsf(mx) * 1e-5f + 1e-7f;
    float t0 = mn > tmin ? mn : tmin;
    float t1 = mx < tmax ? mx : tmax;
    tenter = t0;
    return t0 <= t1;
}

// Optional counters (rays, nodes visited, triangles tested) for the roofline
// figures; S.stats is NULL in timed runs.
DR_INLINE void count_trace(const DrSceneView &S, uint32_t nn, uint32_t nt) {
    if (S.stats) {
        atomicAdd(&S.stats[0], 1ull);
        atomicAdd(&S.stats[1], (unsigned long long)nn);
        atomicAdd(&S.stats[2], (unsigned long long)nt);
    }
}

// Traversal stack of a ray query.  On the GPU its first DR_TSTACK_LDS entries
// live in LDS (one column per lane of the workgroup, at most 256 lanes; a
// thread runs one query at a time, so every query of a kernel shares the
// array) and deeper ones in private memory.
#define DR_TSTACK_LDS 16
#if defined(__HIP_DEVICE_COMPILE__)
DR_INLINE DR_LDS int *trace_stack_lds() {
    __shared__ int s_tstack[DR_TSTACK_LDS * 256];
    return (DR_LDS int *)s_tstack + threadIdx.x;
}
struct TraceStack {
    DR_LDS int *lds;
    int stride;
    int priv[DR_STACK - DR_TSTACK_LDS];
    DR_INLINE void init() {
        lds = trace_stack_lds();
        stride = (int)blockDim.x;
    }
    DR_INLINE void put(int i, int v) {
        if (i < DR_TSTACK_LDS) lds[i * stride] = v;
        else priv[i - DR_TSTACK_LDS] = v;
    }
    DR_INLINE int get(int i) const { return i < DR_TSTACK_LDS ? lds[i * stride] : priv[i - DR_TSTACK_LDS]; }
};
#else
struct TraceStack {
    int a[DR_STACK];
    DR_INLINE void init() {}
    DR_INLINE void put(int i, int v) { a[i] = v; }
    DR_INLINE int get(int i) const { return a[i]; }
};
#endif

DR_INLINE Isect trace_closest(const DrSceneView &S, F3 org, F3 dir, float tmin, float tmax, float &ou, float &ov) {
    Isect best = no_hit();
    if (S.num_nodes == 0) return best;
    F3 inv = F3{1.f / dir.x, 1.f / dir.y, 1.f / dir.z};
    float best_t = tmax;
    int best_g = -1;
    float bu = 0.f, bv = 0.f;
    TraceStack stack;
    stack.init();
    int sp = 0;
    int node = 0;
    uint32_t nn = 0, nt = 0;
    // every node is entered at most once; the cap only guards against a
    // malformed tree
    const uint32_t cap = 2u * (uint32_t)S.num_nodes + 8u;
    const DR_GLOBAL DrNode *nodes = gptr(S.nodes);
    while (nn < cap) {
        const DR_GLOBAL DrNode &N = nodes[node];
        float4 A = *(const DR_GLOBAL float4 *)N.a, B = *(const DR_GLOBAL float4 *)N.b,
               C = *(const DR_GLOBAL float4 *)N.c;
        int4 D = *(const DR_GLOBAL int4 *)N.d;
        nn++;
        float t0, t1;
        bool h0 = box_hit(A.x, A.y, A.z, A.w, B.x, B.y, org, inv, tmin, best_t, t0);
        bool h1 = box_hit(B.z, B.w, C.x, C.y, C.z, C.w, org, inv, tmin, best_t, t1);
        // leaves first
        if (h0 && D.z > 0) {
            for (int i = 0; i < D.z; i++) {
                const DrTri T = gptr(S.tris)[D.x + i];
                float u, v, t;
                nt++;
                if (tri_hit(T, org, dir, tmin, tmax, u, v, t)) {
                    int g = __float_as_int(T.v0[3]);
                    if (best_g < 0 || t < best_t || (t == best_t && g < best_g)) {
                        best_t = t; best_g = g; bu = u; bv = v;
                    }
                }
            }
            h0 = false;
        }
        if (h1 && D.w > 0) {
            for (int i = 0; i < D.w; i++) {
                const DrTri T = gptr(S.tris)[D.y + i];
                float u, v, t;
                nt++;
                if (tri_hit(T, org, dir, tmin, tmax, u, v, t)) {
                    int g = __float_as_int(T.v0[3]);
                    if (best_g < 0 || t < best_t || (t == best_t && g < best_g)) {
                        best_t = t; best_g = g; bu = u; bv = v;
                    }
                }
            }
            h1 = false;
        }
        if (h0 && h1) {
            int near = D.x, far = D.y;
            if (t1 < t0) { near = D.y; far = D.x; }
            // the builder bounds the depth, so the stack cannot fill; a full
            // stack would be counted, not silently dropped
            if (sp < DR_STACK) stack.put(sp++, far);
            else if (S.stats) atomicAdd(&S.stats[7], 1ull);
            node = near;
        } else if (h0) {
            node = D.x;
        } else if (h1) {
            node = D.y;
        } else {
            if (sp == 0) break;
            node = stack.get(--sp);
        }
    }
    count_trace(S, nn, nt);
    if (best_g >= 0) {
        best.gtri = best_g;
        best.shape = S.tri_shape[best_g];
        best.tri = best_g - S.shapes[best.shape].tri_off;
        ou = bu;
        ov = bv;
    }
    return best;
}

// Closest hits of two rays from one origin in one walk of the BVH (the two
// side rays of an edge sample are a few 1e-3 rad apart and visit nearly the
// same nodes, so one node fetch serves both).  A node is entered when either
// ray's box test passes; each ray keeps its own closest hit and makes its
// own box and triangle tests with trace_closest's arithmetic and tie rule.
// Culling with a ray's current closest distance is conservative and the
// closest hit does not depend on the visiting order, so each result is the
// one trace_closest returns for that ray.
DR_INLINE void trace_closest2(const DrSceneView &S, F3 org, F3 da, F3 db, float tmin, float tmax, Isect &ha,
                              float &ua, float &va, Isect &hb, float &ub, float &vb) {
    ha = no_hit();
    hb = no_hit();
    if (S.num_nodes == 0) return;
    F3 ia = F3{1.f / da.x, 1.f / da.y, 1.f / da.z};
    F3 ib = F3{1.f / db.x, 1.f / db.y, 1.f / db.z};
    float ta_best = tmax, tb_best = tmax;
    int ga = -1, gb = -1;
    float au = 0.f, av = 0.f, bu = 0.f, bv = 0.f;
    TraceStack stack;
    stack.init();
    int sp = 0;
    int node = 0;
    uint32_t nn = 0, nt = 0;
    const uint32_t cap = 2u * (uint32_t)S.num_nodes + 8u;
    const DR_GLOBAL DrNode *nodes = gptr(S.nodes);
    auto test = [&](int first, int count, bool ta_on, bool tb_on) {
        for (int i = 0; i < count; i++) {
            const DrTri T = gptr(S.tris)[first + i];
            int g = __float_as_int(T.v0[3]);
            float u, v, t;
            nt++;
            if (ta_on && tri_hit(T, org, da, tmin, tmax, u, v, t) &&
                (ga < 0 || t < ta_best || (t == ta_best && g < ga))) {
                ta_best = t; ga = g; au = u; av = v;
            }
            if (tb_on && tri_hit(T, org, db, tmin, tmax, u, v, t) &&
                (gb < 0 || t < tb_best || (t == tb_best && g < gb))) {
                tb_best = t; gb = g; bu = u; bv = v;
            }
        }
    };
    while (nn < cap) {
        const DR_GLOBAL DrNode &N = nodes[node];
        float4 A = *(const DR_GLOBAL float4 *)N.a, B = *(const DR_GLOBAL float4 *)N.b,
               C = *(const DR_GLOBAL float4 *)N.c;
        int4 D = *(const DR_GLOBAL int4 *)N.d;
        nn++;
        float a0, a1, b0, b1;
        bool ha0 = box_hit(A.x, A.y, A.z, A.w, B.x, B.y, org, ia, tmin, ta_best, a0);
        bool ha1 = box_hit(B.z, B.w, C.x, C.y, C.z, C.w, org, ia, tmin, ta_best, a1);
        bool hb0 = box_hit(A.x, A.y, A.z, A.w, B.x, B.y, org, ib, tmin, tb_best, b0);
        bool hb1 = box_hit(B.z, B.w, C.x, C.y, C.z, C.w, org, ib, tmin, tb_best, b1);
        bool h0 = ha0 || hb0, h1 = ha1 || hb1;
        // leaves first
        if (h0 && D.z > 0) {
            test(D.x, D.z, ha0, hb0);
            h0 = false;
        }
        if (h1 && D.w > 0) {
            test(D.y, D.w, ha1, hb1);
            h1 = false;
        }
        if (h0 && h1) {
            // nearer child first, by the first ray that enters both
            float t0 = ha0 && ha1 ? a0 : b0, t1 = ha0 && ha1 ? a1 : b1;
            int near = D.x, far = D.y;
            if (t1 < t0) { near = D.y; far = D.x; }
            // the builder bounds the depth, so the stack cannot fill; a full
            // stack would be counted, not silently dropped
            if (sp < DR_STACK) stack.put(sp++, far);
            else if (S.stats) atomicAdd(&S.stats[7], 1ull);
            node = near;
        } else if (h0) {
            node = D.x;
        } else if (h1) {
            node = D.y;
        } else {
            if (sp == 0) break;
            node = stack.get(--sp);
        }
    }
    if (S.stats) {
        atomicAdd(&S.stats[0], 2ull);
        atomicAdd(&S.stats[1], (unsigned long long)nn);
        atomicAdd(&S.stats[2], (unsigned long long)nt);
    }
    if (ga >= 0) {
        ha.gtri = ga;
        ha.shape = S.tri_shape[ga];
        ha.tri = ga - S.shapes[ha.shape].tri_off;
        ua = au;
        va = av;
    }
    if (gb >= 0) {
        hb.gtri = gb;
        hb.shape = S.tri_shape[gb];
        hb.tri = gb - S.shapes[hb.shape].tri_off;
        ub = bu;
        vb = bv;
    }
}

DR_INLINE bool trace_any(const DrSceneView &S, F3 org, F3 dir, float tmin, float tmax) {
    if (S.num_nodes == 0) return false;
    F3 inv = F3{1.f / dir.x, 1.f / dir.y, 1.f / dir.z};
    TraceStack stack;
    stack.init();
    int sp = 0;
    int node = 0;
    uint32_t nn = 0, nt = 0;
    bool hit = false;
    const uint32_t cap = 2u * (uint32_t)S.num_nodes + 8u;
    const DR_GLOBAL DrNode *nodes = gptr(S.nodes);
    while (nn < cap) {
        const DR_GLOBAL DrNode &N = nodes[node];
        float4 A = *(const DR_GLOBAL float4 *)N.a, B = *(const DR_GLOBAL float4 *)N.b,
               C = *(const DR_GLOBAL float4 *)N.c;
        int4 D = *(const DR_GLOBAL int4 *)N.d;
        nn++;
        float t0, t1;
        bool h0 = box_hit(A.x, A.y, A.z, A.w, B.x, B.y, org, inv, tmin, tmax, t0);
        bool h1 = box_hit(B.z, B.w, C.x, C.y, C.z, C.w, org, inv, tmin, tmax, t1);
        if (h0 && D.z > 0) {
            for (int i = 0; i < D.z && !hit; i++) {
                float u, v, t;
                nt++;
                const DrTri T = gptr(S.tris)[D.x + i];
                hit = tri_hit(T, org, dir, tmin, tmax, u, v, t);
            }
            h0 = false;
        }
        if (h1 && D.w > 0 && !hit) {
            for (int i = 0; i < D.w && !hit; i++) {
                float u, v, t;
                nt++;
                const DrTri T = gptr(S.tris)[D.y + i];
                hit = tri_hit(T, org, dir, tmin, tmax, u, v, t);
            }
            h1 = false;
        }
        if (hit) break;
        if (h0 && h1) {
            if (sp < DR_STACK) stack.put(sp++, D.y);
            else if (S.stats) atomicAdd(&S.stats[7], 1ull);
            node = D.x;
        } else if (h0) {
            node = D.x;
        } else if (h1) {
            node = D.y;
        } else {
            if (sp == 0) break;
            node = stack.get(--sp);
        }
    }
    count_trace(S, nn, nt);
    return hit;
}

// scene.cpp:147-175
DR_INLINE bool occluded(const DrSceneView &S, F3 org, F3 dir, float max_t) {
    return trace_any(S, org, dir, max_t * 1e-3f, max_t * (1.f - 1e-3f));
}

// --------------------------------------------------------- geometry -------
DR_INLINE F3 vert(const DrSceneView &S, int gv) { return load3(S.verts + 3 * (size_t)gv); }

struct TriData {
    int gv[3];
    F3 v[3];
    F2 uv[3];
    F3 n[3];
    bool has_n;
};
DR_INLINE TriData load_tri(const DrSceneView &S, const Isect &is) {
    TriData d;
    const int *vi = S.tri_vidx + 3 * (size_t)is.gtri;
    const DrShape &sh = S.shapes[is.shape];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        d.gv[k] = vi[k];
        d.v[k] = vert(S, vi[k]);
    }
    if (sh.has_uv) {
#pragma unroll
        for (int k = 0; k < 3; k++) d.uv[k] = F2{S.uvs[2 * (size_t)vi[k]], S.uvs[2 * (size_t)vi[k] + 1]};
    } else {
        d.uv[0] = F2{0.f, 0.f}; d.uv[1] = F2{1.f, 0.f}; d.uv[2] = F2{1.f, 1.f};
    }
    d.has_n = sh.has_n != 0;
    if (d.has_n) {
#pragma unroll
        for (int k = 0; k < 3; k++) d.n[k] = load3(S.normals + 3 * (size_t)vi[k]);
    }
    return d;
}

// intersect.h:71-90 (plain arithmetic on v0, v1, v2).  rcp: divide by
// multiplying with the reciprocal, as the reference's tape type does
// (autodiff.h:166-175); its differentiated render computes every surface
// point that way, so the adjoint kernel must too to reproduce its decisions.
DR_INLINE void mt_uvt(F3 v0, F3 v1, F3 v2, F3 org, F3 dir, float &u, float &v, float &t, bool &div_zero,
                      bool rcp = false) {
    F3 e1 = v1 - v0;
    F3 e2 = v2 - v0;
    F3 pvec = cross(dir, e2);
    float divisor = dot(pvec, e1);
    div_zero = divisor == 0.f;
    if (div_zero) divisor = 1e-8f;
    F3 s = org - v0;
    F3 qvec = cross(s, e1);
    if (rcp) {
        float inv = 1.f / divisor;
        u = dot(s, pvec) * inv;
        v = dot(dir, qvec) * inv;
        t = dot(e2, qvec) * inv;
    } else {
        u = dot(s, pvec) / divisor;
        v = dot(dir, qvec) / divisor;
        t = dot(e2, qvec) / divisor;
    }
}

// intersect.h:100-148
DR_INLINE SPoint surface_point(const TriData &d, F3 org, F3 dir, bool rcp = false) {
    float u, v, t;
    bool dz;
    mt_uvt(d.v[0], d.v[1], d.v[2], org, dir, u, v, t, dz, rcp);
    float w = 1.f - (u + v);
    SPoint sp;
    sp.uv = F2{w * d.uv[0].x + u * d.uv[1].x + v * d.uv[2].x, w * d.uv[0].y + u * d.uv[1].y + v * d.uv[2].y};
    sp.pos = org + dir * t;
    F3 gn = normalize(cross(d.v[1] - d.v[0], d.v[2] - d.v[0]));
    F3 sn = gn;
    if (d.has_n) {
        sn = normalize(d.n[0] * w + d.n[1] * u + d.n[2] * v);
        if (dot(gn, sn) < 0.f) gn = -gn;
    }
    sp.ng = gn;
    sp.fr = make_frame(sn);
    return sp;
}

// scene.h:121-134: closest hit from org along dir (tnear 1e-3, tfar inf)
DR_INLINE Isect intersect(const DrSceneView &S, F3 org, F3 dir, SPoint &sp, bool rcp = false) {
    float u, v;
    Isect is = trace_closest(S, org, dir, 1e-3f, DR_INF, u, v);
    if (valid(is)) sp = surface_point(load_tri(S, is), org, dir, rcp);
    return is;
}
// intersect() of two rays from one origin, in one BVH walk
DR_INLINE void intersect2(const DrSceneView &S, F3 org, F3 da, F3 db, Isect &ia, SPoint &spa, Isect &ib,
                          SPoint &spb) {
    float ua, va, ub, vb;
    trace_closest2(S, org, da, db, 1e-3f, DR_INF, ia, ua, va, ib, ub, vb);
    if (valid(ia)) spa = surface_point(load_tri(S, ia), org, da, false);
    if (valid(ib)) spb = surface_point(load_tri(S, ib), org, db, false);
}

// --------------------------------------------------------- textures -------
DR_INLINE int modulo(int a, int b) {
    int r = a % b;
    return r < 0 ? r + b : r;
}

// bilinear lookup with wrap (material.h:108-239); returns corner indices
// (texel ids) and the fractional position for the adjoint
struct TexFetch {
    int xf, yf, xc, yc;  // wrapped
    float u, v;
    bool constant;
};
DR_INLINE TexFetch tex_locate(const DrTex &t, F2 uv) {
    TexFetch f;
    f.constant = t.h == 0;
    if (f.constant) return f;
    int xf = (int)floorf(uv.x * t.w);
    int yf = (int)floorf(uv.y * t.h);
    f.u = uv.x * (float)t.w - (float)xf;
    f.v = uv.y * (float)t.h - (float)yf;
    f.xf = modulo(xf, t.w);
    f.yf = modulo(yf, t.h);
    f.xc = modulo(xf + 1, t.w);
    f.yc = modulo(yf + 1, t.h);
    return f;
}
DR_INLINE float tex1(const DrSceneView &S, const DrTex &t, F2 uv) {
    if (t.h == 0) return S.texels[t.off];
    TexFetch f = tex_locate(t, uv);
    const float *p = S.texels + t.off;
    float ff = p[f.yf * t.w + f.xf], fc = p[f.yc * t.w + f.xf], cf = p[f.yf * t.w + f.xc], cc = p[f.yc * t.w + f.xc];
    return ff * (1.f - f.u) * (1.f - f.v) + fc * (1.f - f.u) * f.v + cf * f.u * (1.f - f.v) + cc * f.u * f.v;
}
DR_INLINE F3 tex3(const DrSceneView &S, const DrTex &t, F2 uv) {
    if (t.h == 0) return load3(S.texels + t.off);
    TexFetch f = tex_locate(t, uv);
    const float *p = S.texels + t.off;
    F3 ff = load3(p + 3 * (f.yf * t.w + f.xf)), fc = load3(p + 3 * (f.yc * t.w + f.xf));
    F3 cf = load3(p + 3 * (f.yf * t.w + f.xc)), cc = load3(p + 3 * (f.yc * t.w + f.xc));
    return ff * (1.f - f.u) * (1.f - f.v) + fc * (1.f - f.u) * f.v + cf * f.u * (1.f - f.v) + cc * f.u * f.v;
}
DR_INLINE F2 mul2(F2 a, const float *s) { return F2{a.x * s[0], a.y * s[1]}; }
DR_INLINE F3 get_diffuse(const DrSceneView &S, const DrMaterial &m, F2 uv) { return tex3(S, m.diffuse, mul2(uv, m.duv)); }
DR_INLINE F3 get_specular(const DrSceneView &S, const DrMaterial &m, F2 uv) { return tex3(S, m.specular, mul2(uv, m.suv)); }
DR_INLINE float get_roughness(const DrSceneView &S, const DrMaterial &m, F2 uv) {
    return tex1(S, m.roughness, mul2(uv, m.ruv));
}

// ---------------------------------------------------------- material ------
DR_INLINE float roughness_to_phong(float r) { return fmaxf(2.f / r, 0.f); }

// Smith G1 with the rational approximation of material.h:302-331
DR_INLINE float smith_g1(F3 v, F3 h, F3 n, const Frame &fr, float roughness) {
    float c = dot(v, n);
    if (dot(v, h) * c <= 0.f) return 0.f;
    float tan_t = sqrtf(fmaxf(1.f / (c * c) - 1.f, 0.f));
    if (tan_t == 0.f) return 1.f;
    F3 vl = to_local(fr, v);
    float cz = vl.z;
    float s2 = 1.f - cz * cz;
    float alpha;
    if (s2 <= 0.f) {
        alpha = sqrtf(roughness);
    } else {
        float inv = 1.f / s2;
        float cp2 = vl.x * vl.x * inv;
        float sp2 = vl.y * vl.y * inv;
        alpha = sqrtf(cp2 * roughness + sp2 * roughness);
    }
    float a = 1.f / (alpha * tan_t);
    if (a >= 1.6f) return 1.f;
    float a2 = a * a;
    return (3.535f * a + 2.181f * a2) / (1.0f + 2.276f * a + 2.577f * a2);
}

// material.h:257-351.  comp: -1 the whole BSDF, 0 its diffuse part, 1 its
// specular part.
DR_INLINE F3 bsdf(const DrSceneView &S, int mid, const SPoint &sp, F3 wi, F3 wo, int comp = -1) {
    const DrMaterial &m = S.materials[mid];
    F3 n = sp.fr.n;
    if (m.two_sided && dot(wi, n) < 0.f) n = -n;
    float bsdf_cos = dot(n, wo);
    if (bsdf_cos <= 5e-3f) return F3{0.f, 0.f, 0.f};
    F3 diffuse = get_diffuse(S, m, sp.uv);
    F3 specular = get_specular(S, m, sp.uv);
    float roughness = get_roughness(S, m, sp.uv);
    F3 result = comp == 1 ? F3{0.f, 0.f, 0.f} : diffuse * bsdf_cos / DR_PI;
    if (comp != 0 && sum3(specular) > 0.f && dot(wi, n) > 0.f) {
        F3 h = normalize(wi + wo);
        F3 hl = to_local(sp.fr, h);
        if (hl.z > 0.f) {
            float e = roughness_to_phong(roughness);
            float D = powf(fmaxf(hl.z, 0.f), e) * (e + 2.f) / (2.f * DR_PI);
            float G = smith_g1(wi, h, n, sp.fr, roughness) * smith_g1(wo, h, n, sp.fr, roughness);
            float cd = dot(h, wo);
            float sch = powf(fmaxf(1.f - cd, 0.f), 5.f);
            F3 F = F3{specular.x + (1.f - specular.x) * sch, specular.y + (1.f - specular.y) * sch,
                      specular.z + (1.f - specular.z) * sch};
            result += F * D * G / (4.f * dot(wi, n));
        }
    }
    return result;
}

// material.h:353-412; returns the direction, is_specular set
DR_INLINE F3 sample_bsdf_dir(const DrSceneView &S, int mid, const SPoint &sp, F3 wi, float u0, float u1, float uw,
                             bool &is_specular) {
    const DrMaterial &m = S.materials[mid];
    is_specular = false;
    F3 diffuse = get_diffuse(S, m, sp.uv);
    F3 specular = get_specular(S, m, sp.uv);
    float dw = luminance(diffuse), sw = luminance(specular);
    float ws = dw + sw;
    if (ws <= 0.f) return F3{0.f, 0.f, 0.f};
    float dpmf = dw / ws;
    if (uw <= dpmf) {
        if (dpmf <= 0.f) return F3{0.f, 0.f, 0.f};
        return to_world(sp.fr, cos_hemisphere(F2{u0, u1}));
    }
    if (sw <= 0.f) return F3{0.f, 0.f, 0.f};
    float roughness = get_roughness(S, m, sp.uv);
    float e = roughness_to_phong(roughness);
    float phi = 2.f * DR_PI * u1;
    float cos_t = powf(u0, 1.0f / (e + 2.0f));
    float sin_t = sqrtf(fmaxf(1.f - cos_t * cos_t, 0.f));
    F3 mw = to_world(sp.fr, F3{sin_t * cosf(phi), sin_t * sinf(phi), cos_t});
    is_specular = roughness <= 0.01f;
    return 2.f * dot(wi, mw) * mw - wi;
}

// material.h:414-462
DR_INLINE float bsdf_pdf(const DrSceneView &S, int mid, const SPoint &sp, F3 wi, F3 wo) {
    const DrMaterial &m = S.materials[mid];
    F3 diffuse = get_diffuse(S, m, sp.uv);
    F3 specular = get_specular(S, m, sp.uv);
    float dw = luminance(diffuse), sw = luminance(specular);
    float ws = dw + sw;
    if (ws <= 0.f) return 0.f;
    float dpmf = dw / ws, spmf = sw / ws;
    float dpdf = 0.f;
    if (dpmf > 0.f) {
        float c = dot(sp.fr.n, wo);
        c = m.two_sided ? fabsf(c) : fmaxf(c, 0.f);
        dpdf = dpmf * c / DR_PI;
    }
    float spdf = 0.f;
    if (spmf > 0.f) {
        F3 h = normalize(wi + wo);
        float hz = dot(h, sp.fr.n);
        if (hz > 0.f) {
            float e = roughness_to_phong(get_roughness(S, m, sp.uv));
            float D = powf(hz, e) * (e + 2.f) / (2.f * DR_PI);
            spdf = spmf * D / (4.f * fabsf(dot(wo, h)));
        }
    }
    return dpdf + spdf;
}

// material.h:464-476: nearest entry of the LTC table
DR_INLINE M3 get_ltc_matrix(const DrSceneView &S, int mid, const SPoint &sp, F3 wi) {
    float r = get_roughness(S, S.materials[mid], sp.uv);
    float theta = acosf(dot(wi, sp.fr.n));
    int rid = (int)(r * 127.f);
    int tid = (int)((double)theta / (3.14159265358979323846 / 2.0) * 127.0);
    rid = min(max(rid, 0), 127);
    tid = min(max(tid, 0), 127);
    const float *e = S.ltc + (size_t)(rid + tid * 128) * 9;
    M3 M;
#pragma unroll
    for (int i = 0; i < 9; i++) M.m[i / 3][i % 3] = e[i];
    return M;
}

// ---------------------------------------------------------- lights --------
// distribution.h:45-52 (upper_bound over the cdf, clamped)
DR_INLINE int sample_discrete(const float *cdf, int n, float u) {
    int lo = 0, hi = n + 1;  // first index with cdf[i] > u
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cdf[mid] <= u) lo = mid + 1;
        else hi = mid;
    }
    int off = lo - 1;
    return min(max(off, 0), n - 1);
}

struct LightSample {
    float u0, u1, light_sel, tri_sel;
};
DR_INLINE LightSample draw_light_sample(Philox &r) {
    LightSample s;
    s.u0 = philox_next(r);
    s.u1 = philox_next(r);
    s.light_sel = philox_next(r);
    s.tri_sel = philox_next(r);
    return s;
}
struct SecondarySample {
    LightSample ls;
    float b0, b1, bw, rr;
};
DR_INLINE SecondarySample draw_secondary(Philox &r) {
    SecondarySample s;
    s.ls = draw_light_sample(r);
    s.b0 = philox_next(r);
    s.b1 = philox_next(r);
    s.bw = philox_next(r);
    s.rr = philox_next(r);
    return s;
}

struct LightPoint {
    int light, shape, tri, gtri;
    int gv[3];
    F3 v[3];
    float b1, b2;
    F3 pos, n;
};
// scene.h:49-75, sample.h:6-21
DR_INLINE LightPoint sample_point_on_light(const DrSceneView &S, const LightSample &ls) {
    LightPoint L;
    L.light = sample_discrete(S.lights_cdf, S.num_lights, ls.light_sel);
    const DrLight &l = S.lights[L.light];
    L.tri = sample_discrete(S.light_cdf + l.cdf_off, l.ntri, ls.tri_sel);
    L.shape = l.shape;
    L.gtri = S.shapes[l.shape].tri_off + L.tri;
    const int *vi = S.tri_vidx + 3 * (size_t)L.gtri;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        L.gv[k] = vi[k];
        L.v[k] = vert(S, vi[k]);
    }
    float a = sqrtf(ls.u0);
    L.b1 = 1.f - a;
    L.b2 = a * ls.u1;
    F3 e1 = L.v[1] - L.v[0], e2 = L.v[2] - L.v[0];
    L.n = normalize(cross(e1, e2));
    L.pos = L.v[0] + e1 * L.b1 + e2 * L.b2;
    return L;
}

// scene.h:77-108 for a triangle of a light (area measure)
DR_INLINE float light_pdf_tri(const DrSceneView &S, int light, F3 v0, F3 v1, F3 v2, float *area_out = nullptr) {
    const DrLight &l = S.lights[light];
    float light_pmf = luminance(F3{l.I[0], l.I[1], l.I[2]}) / S.light_total_intensity;
    float area = 0.5f * length(cross(v1 - v0, v2 - v0));
    if (area_out) *area_out = area;
    float tri_pmf = area / l.area;
    return light_pmf * tri_pmf / area;
}
DR_INLINE float light_pdf(const DrSceneView &S, const Isect &is) {
    if (!valid(is)) return 0.f;
    int lid = S.shapes[is.shape].light;
    if (lid < 0) return 0.f;
    const int *vi = S.tri_vidx + 3 * (size_t)is.gtri;
    return light_pdf_tri(S, lid, vert(S, vi[0]), vert(S, vi[1]), vert(S, vi[2]));
}

// pathtrace.h:7-57 for an area-light point (visibility always tested)
DR_INLINE F3 eval_light(const DrSceneView &S, F3 wi, const Isect &sis, const SPoint &sp, int light, F3 lpos, F3 ln,
                        int comp = -1) {
    F3 zero = F3{0.f, 0.f, 0.f};
    if (!valid(sis) || light < 0) return zero;
    F3 d = lpos - sp.pos;
    float dist = length(d);
    if (dist == 0.f) return zero;
    F3 nd = d / dist;
    if (occluded(S, sp.pos, nd, dist)) return zero;
    F3 b = bsdf(S, S.shapes[sis.shape].mat, sp, wi, nd, comp);
    float lcos = dot(ln, -nd);
    if (lcos <= 0.f) return zero;
    const DrLight &l = S.lights[light];
    return b * (F3{l.I[0], l.I[1], l.I[2]} * (lcos / sq(dist)));
}

// pathtrace.h:59-102
DR_INLINE F3 nee(const DrSceneView &S, const LightSample &ls, F3 wi, const Isect &sis, const SPoint &sp) {
    if (!valid(sis)) return F3{0.f, 0.f, 0.f};
    LightPoint L = sample_point_on_light(S, ls);
    F3 ldir = normalize(L.pos - sp.pos);
    float lcos = fabsf(dot(ldir, L.n));
    float lpdf = light_pdf_tri(S, L.light, L.v[0], L.v[1], L.v[2]);
    float bpdf = bsdf_pdf(S, S.shapes[sis.shape].mat, sp, wi, ldir) * lcos / length_sq(L.pos - sp.pos);
    float mis = sq(lpdf) / (sq(bpdf) + sq(lpdf));
    return eval_light(S, wi, sis, sp, L.light, L.pos, L.n) * (mis / lpdf);
}

// pathtrace.h:104-162: sample a direction and find the next surface point
DR_INLINE Isect sample_bsdf_scene(const DrSceneView &S, const SecondarySample &ss, F3 wi, const Isect &is,
                                  const SPoint &sp_in, SPoint &out, bool &is_specular, F3 *dir_out = nullptr) {
    is_specular = false;
    if (!valid(is)) return no_hit();
    int mid = S.shapes[is.shape].mat;
    SPoint sp = sp_in;
    if (S.materials[mid].two_sided && dot(wi, sp.fr.n) < 0.f) {
        sp.fr = neg_frame(sp.fr);
        sp.ng = -sp.ng;
    }
    F3 dir = sample_bsdf_dir(S, mid, sp, wi, ss.b0, ss.b1, ss.bw, is_specular);
    if (dir_out) *dir_out = dir;
    if (length(dir) <= 0.f) {
        is_specular = false;
        return no_hit();
    }
    return intersect(S, sp.pos, dir, out);
}

// pathtrace.h:182-297, forward only
DR_INLINE F3 shade(const DrSceneView &S, F3 wi, const Isect &pis, const SPoint &psp, const LightSample &ls,
                   const SecondarySample *ss, int nb, bool include_primary_light) {
    F3 color = F3{0.f, 0.f, 0.f};
    if (!valid(pis)) return color;
    int plid = S.shapes[pis.shape].light;
    if (include_primary_light && plid >= 0 && dot(psp.fr.n, wi) > 0.f) {
        const DrLight &l = S.lights[plid];
        color += F3{l.I[0], l.I[1], l.I[2]};
    }
    F3 thr = F3{1.f, 1.f, 1.f};
    F3 cur_wi = wi;
    Isect cur = pis;
    SPoint csp = psp;
    for (int b = 0; b < nb; b++) {
        color += thr * nee(S, ls, wi, cur, csp);
        SPoint nsp;
        bool spec;
        Isect nis = sample_bsdf_scene(S, ss[b], cur_wi, cur, csp, nsp, spec);
        if (!valid(nis)) break;
        F3 wo = normalize(nsp.pos - csp.pos);
        int mid = S.shapes[cur.shape].mat;
        float pdf = bsdf_pdf(S, mid, csp, cur_wi, wo);
        if (pdf <= 0.f) break;
        F3 be = bsdf(S, mid, csp, cur_wi, wo);
        if (sum3(be) <= 0.f) break;
        int nlid = S.shapes[nis.shape].light;
        if (nlid >= 0 && dot(nsp.fr.n, -wo) > 0.f) {
            const DrLight &l = S.lights[nlid];
            F3 contrib = be * F3{l.I[0], l.I[1], l.I[2]};
            float lpdf = light_pdf(S, nis) * length_sq(nsp.pos - csp.pos) / fabsf(dot(wo, nsp.ng));
            float mis = sq(pdf) / (sq(lpdf) + sq(pdf));
            color += thr * contrib * (mis / pdf);
        }
        thr = thr * (be / pdf);
        if (b >= 5) {
            float q = fminf(max3(thr), 0.95f);
            if (ss[b].rr >= q) break;
            thr = thr / q;
        }
        cur_wi = -wo;
        cur = nis;
        csp = nsp;
    }
    return color;
}

// ------------------------------------------------------------ camera ------
// camera.h:48-84.  Pinhole: the film at z = 1, w = tan(fov / 2).  Fisheye
// (equi-angular): the screen disk's radius r maps linearly to the angle
// r pi / 2 from the view axis; a screen point outside the disk has no ray
// (false; the reference makes a ray without direction, which hits nothing).
// The fisheye direction does not depend on fov or aspect, so the camera's
// derivative through it reaches cam_to_world only (through ndir).
DR_INLINE F3 fisheye_dir(float x, float y) {
    float r = sqrtf(x * x + y * y);
    float phi = atan2f(y, x);
    float theta = r * DR_PI / 2.f;
    float sp = sinf(phi), cp = cosf(phi), st = sinf(theta), ct = cosf(theta);
    return F3{-cp * st, -sp * st, ct};
}
DR_INLINE bool sample_primary(const DrCamera &c, F2 s, F3 &org, F3 &dir, F3 *ndir = nullptr) {
    org = xfm_point(c.c2w, F3{0.f, 0.f, 0.f});
    F3 nd;
    if (c.fisheye) {
        float x = 2.f * (s.x - 0.5f), y = 2.f * (s.y - 0.5f);
        if (x * x + y * y > 1.f) {
            dir = F3{0.f, 0.f, 0.f};
            if (ndir) *ndir = dir;
            return false;
        }
        nd = normalize(fisheye_dir(x, y));
    } else {
        float nx = (s.x - 0.5f) * -2.f;
        float ny = (s.y - 0.5f) * -2.f / c.aspect;
        nd = normalize(F3{c.fov * nx, c.fov * ny, 1.f});
    }
    if (ndir) *ndir = nd;
    dir = xfm_vector(c.c2w, nd);
    return true;
}
// camera.h:86-111
DR_INLINE F3 unproject(const DrCamera &c, F2 s) {
    if (c.fisheye) return fisheye_dir(2.f * (s.x - 0.5f), 2.f * (s.y - 0.5f));
    float nx = (s.x - 0.5f) * -2.f;
    float ny = (s.y - 0.5f) * -2.f / c.aspect;
    return F3{c.fov * nx, c.fov * ny, 1.f};
}
// camera.h:113-153: d unproject / d screen x (dx) and y (dy)
DR_INLINE void d_unproject(const DrCamera &c, F2 s, F3 &dx, F3 &dy) {
    if (c.fisheye) {
        float x = 2.f * (s.x - 0.5f), y = 2.f * (s.y - 0.5f);
        float r = sqrtf(x * x + y * y);
        float phi = atan2f(y, x);
        float theta = r * DR_PI / 2.f;
        float sp = sinf(phi), cp = cosf(phi), st = sinf(theta), ct = cosf(theta);
        float dxdphi = sp * st, dxdth = -cp * ct;
        float dydphi = -cp * st, dydth = -sp * ct;
        float dzdth = -st;
        float dphidx = -y / (r * r), dphidy = x / (r * r);
        float dthdx = (DR_PI / 2.f) * x / r, dthdy = (DR_PI / 2.f) * y / r;
        dx = 2.f * F3{dxdphi * dphidx + dxdth * dthdx, dydphi * dphidx + dydth * dthdx, dzdth * dthdx};
        dy = 2.f * F3{dxdphi * dphidy + dxdth * dthdy, dydphi * dphidy + dydth * dthdy, dzdth * dthdy};
        return;
    }
    dx = F3{-2.f * c.fov, 0.f, 0.f};
    dy = F3{0.f, -2.f * c.fov / c.aspect, 0.f};
}
// camera.h:156-178
DR_INLINE F2 project_local(const DrCamera &c, F3 l) {
    if (c.fisheye) {
        F3 d = normalize(l);
        float phi = atan2f(d.y, d.x);
        float r = acosf(d.z) * 2.f / DR_PI;
        return F2{0.5f * (-r * cosf(phi) + 1.f), 0.5f * (-r * sinf(phi) + 1.f)};
    }
    return F2{(-l.x / (l.z * c.fov) + 1.f) * 0.5f, (-l.y / (l.z * c.fov * c.aspect) + 1.f) * 0.5f};
}
// camera.h:314-320
DR_INLINE bool in_screen(const DrCamera &c, F2 p) {
    if (!c.fisheye) return p.x >= 0.f && p.x < 1.f && p.y >= 0.f && p.y < 1.f;
    return (p.x - 0.5f) * (p.x - 0.5f) + (p.y - 0.5f) * (p.y - 0.5f) < 0.25f;
}
// camera.h:216-242
DR_INLINE bool project_edge(const DrCamera &c, F3 p0, F3 p1, F2 &s0, F2 &s1) {
    F3 a = xfm_point(c.w2c, p0), b = xfm_point(c.w2c, p1);
    if (a.z < c.clip_near && b.z < c.clip_near) return false;
    if (a.z < c.clip_near) {
        F3 d = a - b;
        float t = -(b.z - c.clip_near) / d.z;
        a = b + t * d;
    } else if (b.z < c.clip_near) {
        F3 d = b - a;
        float t = -(a.z - c.clip_near) / d.z;
        b = a + t * d;
    }
    s0 = project_local(c, a);
    s1 = project_local(c, b);
    return true;
}
// camera.h:180-214
DR_INLINE F3 d_project_local(const DrCamera &c, F3 l, float dx, float dy) {
    if (c.fisheye) {
        F3 d = normalize(l);
        float phi = atan2f(d.y, d.x);
        float r = acosf(d.z) * 2.f / DR_PI;
        float dr = -0.5f * (cosf(phi) * dx + sinf(phi) * dy);
        float dphi = 0.5f * r * sinf(phi) * dx - 0.5f * r * cosf(phi) * dy;
        float dtheta = dr * (2.f / DR_PI);
        float d_cos_theta = -dtheta / sqrtf(1.f - d.z * d.z);
        float a2 = d.x * d.x + d.y * d.y;
        F3 dd = F3{-dphi * d.y / a2, dphi * d.x / a2, d_cos_theta};
        // vector.h:454-464 in its arithmetic
        float len = length(l);
        float d_l = -dot(dd, l) / (len * len);
        return (dd + d_l * l) / len;
    }
    float zf = l.z * c.fov;
    float dxd0 = -1.f / zf;
    float dxd2 = c.fov * l.x / (zf * zf);
    float zfy = zf * c.aspect;
    float dyd1 = -1.f / zfy;
    float dyd2 = c.fov * c.aspect * l.y / (zfy * zfy);
    return F3{dx * dxd0, dy * dyd1, dx * dxd2 + dy * dyd2};
}
// camera.h:244-313.  clip_quirk: the reference's clip, t = -(z + clip_near)
// / dz with dt / d(dz) = z / dz^2 (DR_QUIRK_CLIP_DERIVATIVE); otherwise the
// derivative of project_edge's clip, t = -(z - clip_near) / dz.
DR_INLINE void d_project(const DrCamera &c, F3 p0, F3 p1, float d0x, float d0y, float d1x, float d1y, float *dw2c,
                         bool clip_quirk) {
    F3 a = xfm_point(c.w2c, p0), b = xfm_point(c.w2c, p1);
    const float cn = clip_quirk ? -c.clip_near : c.clip_near;  // t = -(z - cn) / dz
    F3 ca = a, cb = b;
    if (a.z < c.clip_near) {
        F3 d = a - b;
        float t = -(b.z - cn) / d.z;
        ca = b + t * d;
    } else if (b.z < c.clip_near) {
        F3 d = b - a;
        float t = -(a.z - cn) / d.z;
        cb = a + t * d;
    }
    F3 dca = d_project_local(c, ca, d0x, d0y);
    F3 dcb = d_project_local(c, cb, d1x, d1y);
    F3 da = F3{0.f, 0.f, 0.f}, db = F3{0.f, 0.f, 0.f};
    if (a.z < c.clip_near) {
        F3 d = a - b;
        float t = -(b.z - cn) / d.z;
        db += dca;
        float dt = dot(d, dca);
        F3 dd = t * dca;
        db.z += (-dt / d.z);
        dd.z += dt * ((clip_quirk ? b.z : b.z - cn) / (d.z * d.z));
        da += dd;
        db = db - dd;
        db += dcb;
    } else if (b.z < c.clip_near) {
        F3 d = b - a;
        float t = -(a.z - cn) / d.z;
        da += dcb;
        float dt = dot(d, dcb);
        F3 dd = t * dcb;
        da.z += (-dt / d.z);
        dd.z += dt * ((clip_quirk ? a.z : a.z - cn) / (d.z * d.z));
        db += dd;
        da = da - dd;
        da += dca;
    } else {
        da += dca;
        db += dcb;
    }
    d_xfm_point(c.w2c, p0, da, dw2c);
    d_xfm_point(c.w2c, p1, db, dw2c);
}
