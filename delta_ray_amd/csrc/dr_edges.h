// Boundary (visibility) terms: silhouette test, edge-tree importance
// sampling, the secondary (shading) edge sample with LTC line sampling, and the
// primary (screen) edge sample.
#pragma once

#include "dr_device.h"

// edge.h:59-79
DR_INLINE bool is_silhouette(const DrSceneView &S, F3 p, const DrEdge &e) {
    const DrShape &sh = S.shapes[e.shape];
    if (!sh.has_n) return true;
    if (e.f0 == -1 || e.f1 == -1) return true;
    F3 v0 = vert(S, sh.vert_off + e.v0), v1 = vert(S, sh.vert_off + e.v1);
    const int *i0 = S.tri_vidx + 3 * (size_t)(sh.tri_off + e.f0);
    const int *i1 = S.tri_vidx + 3 * (size_t)(sh.tri_off + e.f1);
    int g0 = sh.vert_off + e.v0, g1 = sh.vert_off + e.v1;
    int a = i0[0] != g0 && i0[0] != g1 ? i0[0] : (i0[1] != g0 && i0[1] != g1 ? i0[1] : i0[2]);
    int b = i1[0] != g0 && i1[0] != g1 ? i1[0] : (i1[1] != g0 && i1[1] != g1 ? i1[1] : i1[2]);
    F3 va = vert(S, a), vb = vert(S, b);
    F3 n0 = normalize(cross(v0 - va, v1 - va));
    F3 n1 = normalize(cross(v1 - vb, v0 - vb));
    bool f0 = dot(p - v0, n0) > 0.f && dot(p - v1, n0) > 0.f && dot(p - va, n0) > 0.f;
    bool f1 = dot(p - v0, n1) > 0.f && dot(p - v1, n1) > 0.f && dot(p - vb, n1) > 0.f;
    return f0 != f1;
}

// The same test for edge i of the normal tree, from its geometry record
// (one 64-byte read instead of edge -> shape -> face indices -> vertices);
// same operations on the same values.  Silhouette-tree edges (boundary or
// without shading normals) always pass.
DR_INLINE bool is_silhouette_nrm(const DrSceneView &S, F3 p, int i) {
    const DR_GLOBAL float4 *g = (const DR_GLOBAL float4 *)S.nrm_geo + 4 * (size_t)i;
    float4 A = g[0], B = g[1], C = g[2], D = g[3];
    F3 v0 = F3{A.x, A.y, A.z}, v1 = F3{B.x, B.y, B.z}, va = F3{C.x, C.y, C.z}, vb = F3{D.x, D.y, D.z};
    F3 n0 = normalize(cross(v0 - va, v1 - va));
    F3 n1 = normalize(cross(v1 - vb, v0 - vb));
    bool f0 = dot(p - v0, n0) > 0.f && dot(p - v1, n0) > 0.f && dot(p - va, n0) > 0.f;
    bool f1 = dot(p - v0, n1) > 0.f && dot(p - v1, n1) > 0.f && dot(p - vb, n1) > 0.f;
    return f0 != f1;
}

struct Cone {
    F3 org, dir;
    float cos_angle;
};

// Box-cone overlap of Eberly's GeometricTools (IntrAlignedBox3Cone3), as the
// reference uses it (aabb.cpp:98-311).  Algorithm and case table after the
// Geometric Tools Engine, David Eberly, Geometric Tools LLC, distributed
// under the Boost Software License 1.0.  Box corners in CCW order for each
// of the 27 apex regions; geometric data that must match the reference.
__constant__ static const signed char kConePolyN[27] = {6, 6, 6, 6, 4, 6, 6, 6, 6, 6, 4, 6, 4, 0,
                                                         4, 6, 4, 6, 6, 6, 6, 6, 4, 6, 6, 6, 6};
__constant__ static const signed char kConePolyIdx[27][6] = {
    {1, 5, 4, 6, 2, 3}, {0, 2, 3, 1, 5, 4}, {0, 2, 3, 7, 5, 4}, {0, 4, 6, 2, 3, 1}, {0, 2, 3, 1, 0, 0},
    {0, 2, 3, 7, 5, 1}, {0, 4, 6, 7, 3, 1}, {0, 2, 6, 7, 3, 1}, {0, 2, 6, 7, 5, 1}, {0, 1, 5, 4, 6, 2},
    {0, 1, 5, 4, 0, 0}, {0, 1, 3, 7, 5, 4}, {0, 4, 6, 2, 0, 0}, {0, 0, 0, 0, 0, 0}, {1, 3, 7, 5, 0, 0},
    {0, 4, 6, 7, 3, 2}, {2, 6, 7, 3, 0, 0}, {1, 3, 2, 6, 7, 5}, {0, 1, 5, 7, 6, 2}, {0, 1, 5, 7, 6, 4},
    {0, 1, 3, 7, 6, 4}, {0, 4, 5, 7, 6, 2}, {4, 5, 7, 6, 0, 0}, {1, 3, 7, 6, 4, 5}, {0, 4, 5, 7, 3, 2},
    {2, 6, 4, 5, 7, 3}, {1, 3, 2, 6, 4, 5}};

// component k (0, 1, 2) of v, by value (no runtime indexing of a vector)
DR_INLINE float comp(F3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

DR_INLINE bool cone_edge_test(F3 D, float c2, F3 maxX, F3 maxPmV, float maxSqrLen, float maxDd, F3 CmV, int jA,
                              int jB, float flip, bool &decided) {
    int jDiff = jB - jA;
    float s = jDiff > 0 ? 1.f : -1.f;
    int k0 = abs(jDiff) >> 1;
    float fder = flip * s * (comp(D, k0) * maxSqrLen - maxDd * comp(maxPmV, k0));
    decided = fder > 0;
    if (!decided) return false;
    int k1 = (k0 + 1) % 3, k2 = (k1 + 1) % 3;
    float p1 = comp(maxPmV, k1), p2 = comp(maxPmV, k2);
    float d1 = comp(D, k1), d2 = comp(D, k2);
    float numer = p1 * p1 + p2 * p2;
    float denom = d1 * p1 + d2 * p2;
    float m0 = numer * comp(D, k0);
    float m1 = denom * (comp(maxX, k1) + comp(CmV, k1));
    float m2 = denom * (comp(maxX, k2) + comp(CmV, k2));
    // M has m0 at k0, m1 at k1, m2 at k2
    F3 M = F3{k0 == 0 ? m0 : (k1 == 0 ? m1 : m2), k0 == 1 ? m0 : (k1 == 1 ? m1 : m2),
              k0 == 2 ? m0 : (k1 == 2 ? m1 : m2)};
    float DdM = dot(D, M);
    if (DdM * DdM - c2 * dot(M, M) > 0) return true;
    return s * (d1 * p2 - d2 * p1) <= 0;
}

DR_INLINE bool aabb_cone_intersect(const float *lo, const float *hi, const Cone &cone) {
    if (cone.cos_angle <= 0.f) return true;
    F3 center = 0.5f * (F3{lo[0], lo[1], lo[2]} + F3{hi[0], hi[1], hi[2]});
    F3 ext = 0.5f * (F3{hi[0], hi[1], hi[2]} - F3{lo[0], lo[1], lo[2]});
    F3 CmV = center - cone.org;
    F3 D = cone.dir;
    float DdCmV = dot(D, CmV);
    float radius = ext.x * fabsf(D.x) + ext.y * fabsf(D.y) + ext.z * fabsf(D.z);
    if (DdCmV + radius <= 0) return false;
    int ix = CmV.x < -ext.x ? 2 : (CmV.x > ext.x ? 0 : 1);
    int iy = CmV.y < -ext.y ? 2 : (CmV.y > ext.y ? 0 : 1);
    int iz = CmV.z < -ext.z ? 2 : (CmV.z > ext.z ? 0 : 1);
    int lookup = ix + 3 * iy + 9 * iz;
    if (lookup == 13) return true;
    int np = kConePolyN[lookup];
    float c2 = cone.cos_angle * cone.cos_angle;
    int iMax = -1, jMax = -1;
    float maxSqrDd = 0.f, maxSqrLen = 0.f, maxDd = 0.f;
    F3 maxX = F3{0.f, 0.f, 0.f}, maxPmV = F3{0.f, 0.f, 0.f};
    for (int i = 0; i < np; i++) {
        int j = kConePolyIdx[lookup][i];
        F3 X = F3{(j & 1) ? ext.x : -ext.x, (j & 2) ? ext.y : -ext.y, (j & 4) ? ext.z : -ext.z};
        float Dd = dot(D, X) + DdCmV;
        if (Dd > 0) {
            F3 PmV = X + CmV;
            float sqrDd = Dd * Dd;
            float sqrLen = dot(PmV, PmV);
            if (sqrDd - c2 * sqrLen > 0) return true;
            if (iMax == -1 || sqrDd * maxSqrLen > maxSqrDd * sqrLen) {
                iMax = i; jMax = j;
                maxSqrDd = sqrDd; maxSqrLen = sqrLen; maxDd = Dd; maxX = X; maxPmV = PmV;
            }
        }
    }
    if (iMax == -1) return false;
    int jNext = kConePolyIdx[lookup][iMax < np - 1 ? iMax + 1 : 0];
    bool decided;
    bool r = cone_edge_test(D, c2, maxX, maxPmV, maxSqrLen, maxDd, CmV, jMax, jNext, 1.f, decided);
    if (decided) return r;
    int jPrev = kConePolyIdx[lookup][iMax > 0 ? iMax - 1 : np - 1];
    r = cone_edge_test(D, c2, maxX, maxPmV, maxSqrLen, maxDd, CmV, jPrev, jMax, -1.f, decided);
    if (decided) return r;
    return false;
}

DR_INLINE float min_sq_bound(float lo, float hi) {
    if (lo <= 0.f && hi >= 0.f) return 0.f;
    if (lo < 0.f && hi < 0.f) return hi * hi;
    return lo * lo;
}
DR_INLINE float max_sq_bound(float lo, float hi) { return fmaxf(lo * lo, hi * hi); }

// The BSDF bound of edge_tree.cpp:345-480, split in two: the parts that
// depend only on the shading point (the frame around wi, the normal in it,
// the Phong exponent) are computed once per point in bound_ctx(); the per-node
// part is bsdf_bound().  The per-node part uses the hardware square root,
// reciprocal and exp2/log2 (about 1 ulp): the bound only sets node
// importances, and the same value serves for drawing and for weighting, so
// the estimator stays unbiased; against the CPU restatement's exact
// arithmetic the pmfs differ in the last bits (a split decision flips only
// when its uniform draw falls within them).
struct BoundCtx {
    Frame hf;
    float lx, lz, lx2, lz2;  // normal in hf, and its squares
    float e, e2;             // Phong exponent of the roughness, e + 2
    float dterm, slum;       // diffuse luminance / pi, specular luminance
};

DR_INLINE void bound_ctx(BoundCtx &B, F3 wi, const SPoint &sp, float roughness, float dlum, float slum) {
    B.dterm = dlum / DR_PI;
    B.slum = slum;
    if (slum > 0.f) {
        F3 hy = normalize(cross(sp.fr.n, wi));
        B.hf = Frame{normalize(cross(wi, hy)), hy, wi};
        F3 ln = to_local(B.hf, sp.fr.n);
        B.lx = ln.x;
        B.lz = ln.z;
        B.lx2 = ln.x * ln.x;
        B.lz2 = ln.z * ln.z;
        B.e = roughness_to_phong(roughness);
        B.e2 = B.e + 2.f;
    } else {
        B.hf = Frame{F3{0.f, 0.f, 0.f}, F3{0.f, 0.f, 0.f}, F3{0.f, 0.f, 0.f}};
        B.lx = B.lz = B.lx2 = B.lz2 = B.e = B.e2 = 0.f;
    }
}

// Over the box's 8 corners c, a linear function f(c - p) ranges over
// [f(lo - p) + sum_k min(0, f(ext_k)), f(lo - p) + sum_k max(0, f(ext_k))],
// ext_k the box's extent along world axis k: the bounds below take the
// corners' extreme values that way (the same values up to rounding).
DR_INLINE float bsdf_bound(const DrEdgeNode &node, const SPoint &sp, const BoundCtx &B) {
    const F3 lo = F3{node.lo[0], node.lo[1], node.lo[2]};
    const F3 ext = F3{node.hi[0], node.hi[1], node.hi[2]} - lo;
    const F3 b = lo - sp.pos;
    // some corner above the shading point's tangent plane (cos_max 1), else 0
    const F3 n = sp.fr.n;
    float up = dot(n, b) + fmaxf(n.x * ext.x, 0.f) + fmaxf(n.y * ext.y, 0.f) + fmaxf(n.z * ext.z, 0.f);
    if (!(up > 0.f)) return 0.f;
    float phong_resp = 0.f;
    if (B.slum > 0.f) {
        // the box in the frame around wi: per local axis j, base + extents
        const Frame &f = B.hf;
        F3 l0 = to_local(f, b);
        F3 ax = to_local(f, F3{ext.x, 0.f, 0.f}), ay = to_local(f, F3{0.f, ext.y, 0.f}),
           az = to_local(f, F3{0.f, 0.f, ext.z});
        float blo[3], bhi[3];
        blo[0] = l0.x + fminf(ax.x, 0.f) + fminf(ay.x, 0.f) + fminf(az.x, 0.f);
        bhi[0] = l0.x + fmaxf(ax.x, 0.f) + fmaxf(ay.x, 0.f) + fmaxf(az.x, 0.f);
        blo[1] = l0.y + fminf(ax.y, 0.f) + fminf(ay.y, 0.f) + fminf(az.y, 0.f);
        bhi[1] = l0.y + fmaxf(ax.y, 0.f) + fmaxf(ay.y, 0.f) + fmaxf(az.y, 0.f);
        blo[2] = l0.z + fminf(ax.z, 0.f) + fminf(ay.z, 0.f) + fminf(az.z, 0.f);
        bhi[2] = l0.z + fmaxf(ax.z, 0.f) + fmaxf(ay.z, 0.f) + fmaxf(az.z, 0.f);
        float cphm = 0.f;
        float hx = bhi[0];
        if (hx > 0.f) {
            cphm = hx * approx_rsqrt(min_sq_bound(blo[1], bhi[1]) + hx * hx);
        } else if (hx == 0.f) {
            cphm = fmaxf(cphm, 0.f);
        } else {
            cphm = hx * approx_rsqrt(max_sq_bound(blo[1], bhi[1]) + hx * hx);
        }
        float cto_min, cto_max;
        if (blo[2] > 0.f)
            cto_min = blo[2] * approx_rcp(max_sq_bound(blo[0], bhi[0]) + max_sq_bound(blo[1], bhi[1]) + blo[2] * blo[2]);
        else
            cto_min = blo[2] * approx_rcp(min_sq_bound(blo[0], bhi[0]) + min_sq_bound(blo[1], bhi[1]) + blo[2] * blo[2]);
        if (bhi[2] > 0.f)
            cto_max = bhi[2] * approx_rcp(min_sq_bound(blo[0], bhi[0]) + min_sq_bound(blo[1], bhi[1]) + bhi[2] * bhi[2]);
        else
            cto_max = bhi[2] * approx_rcp(max_sq_bound(blo[0], bhi[0]) + max_sq_bound(blo[1], bhi[1]) + bhi[2] * bhi[2]);
        float cth_min = approx_sqrt((cto_min + 1.f) * 0.5f);
        float cth_max = approx_sqrt((cto_max + 1.f) * 0.5f);
        float ch = cphm >= 0.f ? approx_sqrt(B.lz2 * approx_rcp(B.lz2 + B.lx2 * cphm * cphm)) : 1.f;
        if (ch > cth_max) ch = cth_max;
        else if (ch < cth_min) ch = cth_min;
        float sh = approx_sqrt(fmaxf(1.f - ch * ch, 0.f));
        float hn = fminf(fabsf(B.lx * sh * cphm) + B.lz * ch, 1.f);
        phong_resp = approx_pow01(hn, B.e) * B.e2 * (1.f / (2.f * DR_PI));
    }
    return B.dterm + B.slum * phong_resp;
}

// edge_tree.cpp:482-497: the node's edges can hold a silhouette seen from p
// when the normal cone widened by the box's angular radius seen from p, va =
// atan(r / |c - p|), reaches the plane perpendicular to the direction to the
// box: |acos(-d.axis) - pi/2| <= angle + va.  With both sides' cosines (the
// node stores cos(angle), the cosine of its normal cone's half angle): the
// difference lies in [0, pi/2], so the test is sin(acos(-d.axis)) >=
// cos(angle + va), which also holds whenever angle + va >= pi/2.  The same
// verdicts as the angles' up to rounding at the boundary.
DR_INLINE bool contains_silhouette(const DrEdgeNode &n, F3 p) {
    F3 lo = F3{n.lo[0], n.lo[1], n.lo[2]}, hi = F3{n.hi[0], n.hi[1], n.hi[2]};
    F3 c = (lo + hi) / 2.f;
    float r = length(hi - c);
    F3 d = c - p;
    float len2 = length_sq(d);
    float len = approx_sqrt(len2);
    float tv = r / len;
    if (tv >= 1.f) return true;
    float ca = -dot(d, F3{n.axis[0], n.axis[1], n.axis[2]}) / len;
    float s_ang = approx_sqrt(fmaxf(1.f - ca * ca, 0.f));
    float cA = n.cos_angle, sA = approx_sqrt(fmaxf(1.f - cA * cA, 0.f));
    // cos(angle + va) = (cos(angle) - sin(angle) tv) / sqrt(1 + tv^2)
    return s_ang * approx_sqrt(1.f + tv * tv) >= cA - sA * tv;
}

DR_INLINE bool node_inside(const DrEdgeNode &n, F3 p) {
    return p.x >= n.lo[0] && p.x <= n.hi[0] && p.y >= n.lo[1] && p.y <= n.hi[1] && p.z >= n.lo[2] && p.z <= n.hi[2];
}
DR_INLINE float node_importance(const DrEdgeNode &n, float bound, F3 p) {
    F3 c = 0.5f * (F3{n.lo[0], n.lo[1], n.lo[2]} + F3{n.hi[0], n.hi[1], n.hi[2]});
    return bound * n.total_length / length_sq(c - p);
}
// the same with another node length (the learned proposal's)
DR_INLINE float node_importance_len(const DrEdgeNode &n, float bound, F3 p, float len) {
    F3 c = 0.5f * (F3{n.lo[0], n.lo[1], n.lo[2]} + F3{n.hi[0], n.hi[1], n.hi[2]});
    return bound * len / length_sq(c - p);
}

// Edge-tree traversal (edge_tree.cpp:511-683), iterative.  The reference
// walks the tree twice per shading point, a "cone" pass and an "all" pass,
// depth first.  Here both passes walk together: a stack entry carries a pass
// mask, a node both passes reach (all of them where the point lies inside the
// boxes, which split in both passes) is expanded once, and each pass still
// meets its nodes in its own depth-first order.  A node's decisions are keyed
// draws of the point's tree stream (block = node index; output 0 for the
// cone pass, 1 for the "all" pass), so they do not depend on the order of the
// walk; a selected edge's ordinal is its index in the tree with bit 30 set
// for the "all" pass: the tree is stored in depth-first order (left child
// first, leaves' edges in leaf order), so ordinals sort like the reference's
// sequential order of selection.  The traversal is explicit state advanced
// one node per et_step() so that a persistent kernel can give a lane a new
// point as soon as its traversal ends; et_step() calls
// `emit(weight, in_normal_tree, index, ordinal)` per selected edge.
//
// The stack is an object of its own (owned by the caller), not a member of
// EdgeTraversal: its storage is indexed at run time, while every other field
// of the traversal state stays in registers.  An entry carries what the parent
// already knew of the node (child offset, edge count, what its bound and
// whether the point lies in its box decide), so popping an interior node needs
// only its two children's records, loaded side by side.  Two stacks:
// LdsStack (the selection kernel): the top entries live in LDS and older ones
// spill to private memory in halves; and PlainStack, an array, against which
// the stack self-test (k_tree_stack) checks it.
//
// An entry is 16 bytes.  What popping needs of the node's bound is two flags,
// kept in the sign bits of the two pmfs (a pmf is never negative):
//   code  (interior: node index; leaf: -1 - edge count) << 2 | pass mask
//         (bit 0: the cone pass reaches the node, bit 1: the "all" pass)
//   cpmf  the cone pass's pmf of the path to the node; sign bit: the "all"
//         pass takes both children (bound > 1, or the point lies in the box)
//   apmf  the "all" pass's pmf; sign bit: the bound is <= 0 (nothing to take)
//   off   the node's offset (interior: right child; leaf: first edge)
struct TreeEntry {
    int code;
    uint32_t cpmf, apmf;
    int off;
    DR_INLINE int mask() const { return code & 3; }
    DR_INLINE int node() const { return code >> 2; }
    DR_INLINE int nprims() const { return code < 0 ? -1 - (code >> 2) : 0; }
    DR_INLINE float cone_pmf() const { return __uint_as_float(cpmf & 0x7fffffffu); }
    DR_INLINE float all_pmf() const { return __uint_as_float(apmf & 0x7fffffffu); }
    DR_INLINE bool split_all() const { return (cpmf >> 31) != 0; }
    DR_INLINE bool empty() const { return (apmf >> 31) != 0; }
};

// The same path's pmfs under the other proposal (learned mixture only),
// kept beside the entry at the same depth.
struct TreeAlt {
    float cone, all;
};

struct PlainStack {
    TreeEntry e[DR_TREE_STACK];
    TreeAlt alt[DR_TREE_STACK];
    int sp;
    DR_INLINE void clear() { sp = 0; }
    DR_INLINE int size() const { return sp; }
    DR_INLINE void push(const TreeEntry &x, TreeAlt a, bool with_alt) {
        if (with_alt) alt[sp] = a;
        e[sp++] = x;
    }
    DR_INLINE TreeEntry pop(TreeAlt &a, bool with_alt) {
        --sp;
        if (with_alt) a = alt[sp];
        return e[sp];
    }
};

// entries per lane in LDS: 10 x 16 B x 256 lanes = 40 KB per workgroup, so
// that four workgroups (4 waves per SIMD) share a CU's 160 KB; 12 entries
// (48 KB, 3 waves per SIMD) made k_select 8 % slower
#define DR_LDS_STACK 10
#define DR_LDS_WORDS 4   // words per entry

// Top DR_LDS_STACK entries of each lane in LDS, one word of an entry per
// row of the block's lanes (conflict-free: lane l always uses bank l % 64);
// entries below them in private memory.  Entries [lo, sp) are in LDS, entry i
// in slot i % DR_LDS_STACK; [0, lo) in `spill`.  A push onto a full LDS part
// moves its bottom entry to `spill`; a pop below lo takes the entry from
// `spill` directly, so an entry goes to private memory at most once and
// comes back at most once (the walk is depth first: once entries below the
// LDS part are being popped, everything above them has been).
struct LdsStack {
    DR_LDS float *lds;  // the block's stack area + threadIdx.x (stride: blockDim.x)
    int stride;
    TreeEntry spill[DR_TREE_STACK];
    TreeAlt alt[DR_TREE_STACK];
    int sp, lo;
    DR_INLINE void clear() { sp = lo = 0; }
    DR_INLINE int size() const { return sp; }
    DR_INLINE void put(int slot, const TreeEntry &x) {
        DR_LDS float *w = lds + slot * DR_LDS_WORDS * stride;
        w[0] = __int_as_float(x.code);
        w[stride] = __uint_as_float(x.cpmf);
        w[2 * stride] = __uint_as_float(x.apmf);
        w[3 * stride] = __int_as_float(x.off);
    }
    DR_INLINE TreeEntry get(int slot) const {
        const DR_LDS float *w = lds + slot * DR_LDS_WORDS * stride;
        return TreeEntry{__float_as_int(w[0]), __float_as_uint(w[stride]), __float_as_uint(w[2 * stride]),
                         __float_as_int(w[3 * stride])};
    }
    DR_INLINE void push(const TreeEntry &x, TreeAlt a, bool with_alt) {
        if (sp - lo == DR_LDS_STACK) {
            spill[lo] = get(lo % DR_LDS_STACK);
            lo++;
        }
        if (with_alt) alt[sp] = a;
        put(sp % DR_LDS_STACK, x);
        sp++;
    }
    DR_INLINE TreeEntry pop(TreeAlt &a, bool with_alt) {
        --sp;
        if (with_alt) a = alt[sp];
        if (sp < lo) {
            lo = sp;
            return spill[sp];
        }
        return get(sp % DR_LDS_STACK);
    }
};

struct EdgeTraversal {
    const DrEdgeNode *nodes;
    const DrEdge *edges;
    int num_nodes;
    bool directional;
    BoundCtx B;
    Cone cone;
    Philox rc;  // the point's stream (the tree choice)
    Philox rt;  // the point's tree stream: keyed draws, block n = node n
    // proposal: 0 the reference's (node length = edge lengths); with the
    // learned mixture, 1 = this point uses the learned node lengths, 2 = the
    // plain ones (probabilities 0.8 / 0.2)
    int mode;
    int pops;
    uint32_t n_inner, n_edges;
    bool cone_pmf_quirk;  // DR_QUIRK_CONE_PMF
};

// The entry of node `node` (record n, bound computed by its parent).  Edge
// counts and node indices below 2^29 fit the code (checked by the builder).
DR_INLINE TreeEntry tree_entry(int node, int mask, const DrEdgeNode &n, float bound, float cone_pmf, float all_pmf,
                               F3 p) {
    int v = n.nprims > 0 ? -1 - n.nprims : node;
    bool split = bound > 1.f || node_inside(n, p);
    return TreeEntry{(int)((uint32_t)v << 2 | (uint32_t)mask),
                     (__float_as_uint(cone_pmf) & 0x7fffffffu) | (split ? 0x80000000u : 0u),
                     (__float_as_uint(all_pmf) & 0x7fffffffu) | (bound <= 0.f ? 0x80000000u : 0u), n.offset};
}

// edge_tree.cpp:685-725: choose the tree (one draw from rc when both exist)
// and push the root for both passes.  Returns false when there is nothing
// to traverse.
template <class Stack>
DR_INLINE bool et_init(EdgeTraversal &T, Stack &stack, const DrSceneView &S, F3 wi, const SPoint &sp,
                       const Cone &cone, float roughness, float dlum, float slum, int mode) {
    if (dot(wi, sp.fr.n) < 0.f) return false;
    bool has_n = S.num_nrm_nodes > 0, has_s = S.num_sil_nodes > 0;
    if (!has_n && !has_s) return false;
    // pmf: the tree's probability under the point's proposal; pmf_o: under
    // the other one (learned mixture: the proposals weigh the trees by their
    // plain or learned root lengths)
    float pmf = 1.f, pmf_o = 1.f;
    if (!has_n) {
        T.nodes = S.sil_nodes; T.edges = S.sil_edges; T.num_nodes = S.num_sil_nodes; T.directional = false;
    } else if (!has_s) {
        T.nodes = S.nrm_nodes; T.edges = S.nrm_edges; T.num_nodes = S.num_nrm_nodes; T.directional = true;
    } else {
        float ni = S.nrm_nodes[0].total_length, si = S.sil_nodes[0].total_length;
        float npmf = ni / (ni + si), spmf = si / (ni + si);
        float npmf_o = npmf, spmf_o = spmf;
        if (mode != 0) {
            float nl = S.nrm_nodes[0].pad, sl = S.sil_nodes[0].pad;
            float npl = npmf, spl = spmf;
            if (nl + sl > 0.f) {
                npl = nl / (nl + sl);
                spl = sl / (nl + sl);
            }
            if (mode == 1) {
                npmf_o = npmf; spmf_o = spmf; npmf = npl; spmf = spl;
            } else {
                npmf_o = npl; spmf_o = spl;
            }
        }
        float u = philox_next(T.rc);
        if (u < npmf) {
            T.nodes = S.nrm_nodes; T.edges = S.nrm_edges; T.num_nodes = S.num_nrm_nodes; T.directional = true;
            pmf = npmf;
            pmf_o = npmf_o;
        } else {
            T.nodes = S.sil_nodes; T.edges = S.sil_edges; T.num_nodes = S.num_sil_nodes; T.directional = false;
            pmf = spmf;
            pmf_o = spmf_o;
        }
    }
    bound_ctx(T.B, wi, sp, roughness, dlum, slum);
    T.cone = cone;
    const DrEdgeNode &root = T.nodes[0];
    float root_bound = bsdf_bound(root, sp, T.B);
    T.n_inner = 0;
    T.n_edges = 0;
    T.pops = 0;
    T.mode = mode;
    stack.clear();
    stack.push(tree_entry(0, 3, root, root_bound, pmf, pmf, sp.pos), TreeAlt{pmf_o, pmf_o}, mode != 0);
    return true;
}

#define DR_LEARNED_SHARE 0.8f

// Edge i of a leaf: emitted for each pass that reached the leaf when it is
// a silhouette at the point.
template <class Emit>
DR_INLINE void et_leaf_edge(const EdgeTraversal &T, const TreeEntry &e, TreeAlt ea, const DrSceneView &S,
                            const SPoint &sp, int i, Emit &emit) {
    const int mask = e.mask();
    const float cone_pmf = e.cone_pmf(), all_pmf = e.all_pmf();
    // power heuristic over the two passes, which both run
    // (edge_tree.cpp:528-536): pass t selected the edge with pmf p_t and
    // weighs it p_t / (p_cone^2 + p_all^2).  With the learned mixture the
    // point ran one proposal s of two (probabilities pi_s, pi_o); the
    // passes' weights are scaled by q_s / (pi_s q_s + pi_o q_o), q = p_cone
    // + p_all of each proposal (balance heuristic over the proposals: the
    // defensive mixture, unbiased for any q, and exactly the plain
    // estimator when both proposals agree).
    float den = cone_pmf * cone_pmf + all_pmf * all_pmf;
    float scale = 1.f;
    if (T.mode != 0) {
        float pi = T.mode == 1 ? DR_LEARNED_SHARE : 1.f - DR_LEARNED_SHARE;
        float qs = cone_pmf + all_pmf, qo = ea.cone + ea.all;
        scale = qs / (pi * qs + (1.f - pi) * qo);
    }
    if (S.edge_on && !S.edge_on[(T.directional ? S.num_sil_edges : 0) + e.off + i]) return;
    if (T.directional && !is_silhouette_nrm(S, sp.pos, e.off + i)) return;
    // ordinal: the edge's index in its tree (increasing in the depth-first
    // order of each pass), bit 30 for the "all" pass
    if (mask & 1) emit(scale * cone_pmf / den, T.directional, e.off + i, e.off + i);
    if (mask & 2) emit(scale * all_pmf / den, T.directional, e.off + i, DR_ORD_ALL_PASS | (e.off + i));
}

// A leaf's edges in order.
template <class Emit>
DR_INLINE void et_leaf(EdgeTraversal &T, const TreeEntry &e, TreeAlt ea, const DrSceneView &S, const SPoint &sp,
                       Emit &emit) {
    const int nprims = e.nprims();
    T.n_edges += nprims;
    for (int i = 0; i < nprims; i++) et_leaf_edge(T, e, ea, S, sp, i, emit);
}

// An interior node: the children the passes go on to, at most two, in the
// order a depth-first walk pushes them (right, then left).  Returns how
// many (c0 first, then c1).
DR_INLINE int et_expand(EdgeTraversal &T, const TreeEntry &e, TreeAlt ea, const DrSceneView &S, const SPoint &sp,
                        TreeEntry &c0, TreeAlt &a0, TreeEntry &c1, TreeAlt &a1) {
    const int ni = e.node(), mask = e.mask();
    const float cone_pmf = e.cone_pmf(), all_pmf = e.all_pmf();
    int li = ni + 1, ri = e.off;
    const DrEdgeNode L = gptr(T.nodes)[li], R = gptr(T.nodes)[ri];
    T.n_inner++;
    bool all_split = e.split_all();
    bool lcs = true, rcs = true;
    if (T.directional) {
        lcs = contains_silhouette(L, sp.pos);
        rcs = contains_silhouette(R, sp.pos);
    }
    float lb = lcs ? bsdf_bound(L, sp, T.B) : 0.f;
    float rb = rcs ? bsdf_bound(R, sp, T.B) : 0.f;
    float il = lcs ? node_importance(L, lb, sp.pos) : 0.f;
    float ir = rcs ? node_importance(R, rb, sp.pos) : 0.f;
    // with the learned mixture: importances of the point's proposal (il,
    // ir) and of the other one (il2, ir2), which only rescales node lengths
    float il2 = il, ir2 = ir;
    if (T.mode != 0) {
        float ll = lcs ? node_importance_len(L, lb, sp.pos, L.pad) : 0.f;
        float rl = rcs ? node_importance_len(R, rb, sp.pos, R.pad) : 0.f;
        if (T.mode == 1) {
            il = ll; ir = rl;
        } else {
            il2 = ll; ir2 = rl;
        }
    }
    if (il + ir <= 1e-10f) return 0;
    // the cone tests only matter while the cone pass can still reach the
    // node's subtree (its own mask bit or a nonzero pmf of either proposal);
    // otherwise every quantity they feed is multiplied by zero
    bool need_cone = (mask & 1) || cone_pmf != 0.f || (T.mode != 0 && ea.cone != 0.f);
    bool lhc = need_cone && aabb_cone_intersect(L.lo, L.hi, T.cone);
    bool rhc = need_cone && aabb_cone_intersect(R.lo, R.hi, T.cone);
    bool cone_split = all_split || (lhc && rhc && lcs && rcs && T.cone.cos_angle > 0.95f);
    float ilc = lhc ? il : 0.f, irc = rhc ? ir : 0.f;
    // The reference's cone pass stops where neither child meets the cone,
    // even where the split rule would take both (edge_tree.cpp:587-589
    // before 590-601), yet its all pass weighs the edges below with cone pmf
    // 1: without the quirk they get the cone pmf 0 the cone pass really has.
    if (!T.cone_pmf_quirk && ilc + irc <= 1e-10f) cone_split = false;
    float plc = 0.f, prc = 0.f, pla, pra;
    if (cone_split) {
        plc = prc = 1.f;
    } else if (ilc + irc > 1e-10f) {
        plc = ilc / (ilc + irc);
        prc = 1.f - plc;
    }
    if (all_split) {
        pla = pra = 1.f;
    } else {
        pla = il / (il + ir);
        pra = 1.f - pla;
    }
    // the other proposal's split probabilities on this path
    float plc2 = 0.f, prc2 = 0.f, pla2 = 0.f, pra2 = 0.f;
    if (T.mode != 0 && il2 + ir2 > 1e-10f) {
        float ilc2 = lhc ? il2 : 0.f, irc2 = rhc ? ir2 : 0.f;
        if (cone_split) {
            plc2 = prc2 = 1.f;
        } else if (ilc2 + irc2 > 1e-10f) {
            plc2 = ilc2 / (ilc2 + irc2);
            prc2 = 1.f - plc2;
        }
        if (all_split) {
            pla2 = pra2 = 1.f;
        } else {
            pla2 = il2 / (il2 + ir2);
            pra2 = 1.f - pla2;
        }
    }
    int lm = 0, rm = 0;
    // this node's decisions: keyed draws of the point's tree stream
    float uc = 0.f, ua = 0.f;
    if (((mask & 1) && !cone_split) || ((mask & 2) && !all_split)) philox_keyed2(T.rt, (uint32_t)ni, uc, ua);
    if ((mask & 1) && ilc + irc > 1e-10f) {
        if (cone_split) {
            if (rcs) rm |= 1;
            if (lcs) lm |= 1;
        } else if (uc < plc) {
            lm |= 1;
        } else {
            rm |= 1;
        }
    }
    if (mask & 2) {
        if (all_split) {
            if (rcs) rm |= 2;
            if (lcs) lm |= 2;
        } else if (ua < pla) {
            lm |= 2;
        } else {
            rm |= 2;
        }
    }
    int n = 0;
    if (rm) {
        c0 = tree_entry(ri, rm, R, rb, prc * cone_pmf, pra * all_pmf, sp.pos);
        a0 = TreeAlt{prc2 * ea.cone, pra2 * ea.all};
        n = 1;
    }
    if (lm) {
        TreeEntry x = tree_entry(li, lm, L, lb, plc * cone_pmf, pla * all_pmf, sp.pos);
        TreeAlt xa = TreeAlt{plc2 * ea.cone, pla2 * ea.all};
        if (n == 0) {
            c0 = x;
            a0 = xa;
        } else {
            c1 = x;
            a1 = xa;
        }
        n++;
    }
    return n;
}

// One node.  Returns false when the traversal is done.
template <class Stack, class Emit>
DR_INLINE bool et_step(EdgeTraversal &T, Stack &stack, const DrSceneView &S, const SPoint &sp, Emit &emit) {
    // every node is pushed at most once; the cap guards a malformed tree
    if (stack.size() == 0 || T.pops >= T.num_nodes + 8) {
        if (S.stats) {
            atomicAdd(&S.stats[5], (unsigned long long)T.n_inner);
            atomicAdd(&S.stats[6], (unsigned long long)T.n_edges);
        }
        return false;
    }
    T.pops++;
    TreeAlt ea{0.f, 0.f};
    TreeEntry e = stack.pop(ea, T.mode != 0);
    if (e.empty()) return true;
    if (e.nprims() > 0) {
        et_leaf(T, e, ea, S, sp, emit);
        return true;
    }
    TreeEntry c0, c1;
    TreeAlt a0{0.f, 0.f}, a1{0.f, 0.f};
    int n = et_expand(T, e, ea, S, sp, c0, a0, c1, a1);
    // the builder bounds the tree depth so that the stack cannot fill (a
    // full stack would be counted in the statistics, not silently dropped)
    if (n > 0) {
        if (stack.size() < DR_TREE_STACK) stack.push(c0, a0, T.mode != 0);
        else if (S.stats) atomicAdd(&S.stats[7], 1ull);
    }
    if (n > 1) {
        if (stack.size() < DR_TREE_STACK) stack.push(c1, a1, T.mode != 0);
        else if (S.stats) atomicAdd(&S.stats[7], 1ull);
    }
    return true;
}

struct EdgeResult {
    F3 dp, dv0, dv1;
};

// The power-heuristic weight of BSDF sampling for the emitter point hp (hit
// h) seen from sp, as the path estimate weighs an emitter a BSDF-sampled
// direction hits (pathtrace.h:139-150): pdfs in solid angle.
DR_INLINE float emitter_bsdf_mis(const DrSceneView &S, int mid, const SPoint &sp, F3 wi, const Isect &h,
                                 const SPoint &hp) {
    F3 d = hp.pos - sp.pos;
    float dsq = length_sq(d);
    if (dsq <= 0.f) return 0.f;
    F3 wo = normalize(d);
    float pdf = bsdf_pdf(S, mid, sp, wi, wo);
    if (pdf <= 0.f) return 0.f;
    float lpdf = light_pdf(S, h) * dsq / fabsf(dot(wo, hp.ng));
    return sq(pdf) / (sq(lpdf) + sq(pdf));
}

// delta_ray.cpp:45-365: one point on one silhouette edge, seen from a
// shading point; LTC line sampling, two rays across the edge, Dirac
// Jacobian.  ss / nss: bounces after the edge hit (empty when direct only).
// A point on the segment a-b (above z = 0) drawn with density proportional to
// the clamped cosine to +z over the solid angle (the line integral of a
// cosine lobe): closed-form CDF I(l), inverted by bisection + Newton, at most
// 20 iterations, |CDF - u| < 1e-5 (delta_ray.cpp:108-180).  The point is
// vo + l wt; pdf is the density in l.
struct LineSample {
    F3 vo, wt;
    float l, pdf;
};
DR_INLINE bool ltc_line_sample(F3 a, F3 b, float eu, LineSample &L) {
    F3 vd = b - a;
    if (abs_sum(vd) < 1e-10f) return false;
    F3 wt = normalize(vd);
    float l0 = dot(a, wt), l1 = dot(b, wt);
    F3 vo = a - l0 * wt;
    float d = length(vo);
    auto I = [&](float l) {
        return (l / (d * (d * d + l * l)) + atanf(l / d) / (d * d)) * vo.z + (l * l / (d * (d * d + l * l))) * wt.z;
    };
    float Il0 = I(l0), Il1 = I(l1);
    float norm = Il1 - Il0;
    if (norm <= 1e-10f) return false;
    auto pdf = [&](float l) {
        float dsq = d * d + l * l;
        return 2.f * d * (vo + l * wt).z / (norm * dsq * dsq);
    };
    float lb = l0, ub = l1;
    if (lb > ub) { float t = lb; lb = ub; ub = t; }
    float l = 0.5f * (lb + ub);
    for (int it = 0; it < 20; it++) {
        if (!(l >= lb && l <= ub)) l = 0.5f * (lb + ub);
        float value = (I(l) - Il0) / norm - eu;
        if (fabsf(value) < 1e-5f || it == 19) break;
        if (value > 0.f) ub = l;
        else lb = l;
        l -= value / pdf(l);
    }
    float pl = pdf(l);
    if (pl <= 0.f) return false;
    L.vo = vo;
    L.wt = wt;
    L.l = l;
    L.pdf = pl;
    return true;
}

DR_INLINE EdgeResult sample_shading_edge_point(const DrSceneView &S, float bsdf_component, float eu, const DrEdge &edge,
                                               const LightSample &ls, const SecondarySample *ss, int nss, F3 wi,
                                               const Isect &is, const SPoint &sp, F3 weight, F3 d_color,
                                               bool direct_only, bool *defer = nullptr,
                                               const float *geo = nullptr,
                                               unsigned quirks = DR_QUIRK_EDGE_BSDF_SUM | DR_QUIRK_EDGE_HIT_JACOBIAN) {
    EdgeResult R = EdgeResult{F3{0.f, 0.f, 0.f}, F3{0.f, 0.f, 0.f}, F3{0.f, 0.f, 0.f}};
    if (!valid(is)) return R;
    F3 v0, v1;
    if (geo) {
        // an edge the tree selected at this point (a silhouette there), with
        // its endpoints from the tree's geometry record
        const DR_GLOBAL float4 *g = (const DR_GLOBAL float4 *)geo;
        float4 A = g[0], B = g[1];
        v0 = F3{A.x, A.y, A.z};
        v1 = F3{B.x, B.y, B.z};
    } else {
        if (!is_silhouette(S, sp.pos, edge)) return R;
        const DrShape &es = S.shapes[edge.shape];
        v0 = vert(S, es.vert_off + edge.v0);
        v1 = vert(S, es.vert_off + edge.v1);
    }
    if (abs_sum(v1 - v0) == 1e-10f) return R;
    int mid = S.shapes[is.shape].mat;
    const DrMaterial &m = S.materials[mid];
    float dw = luminance(get_diffuse(S, m, sp.uv));
    float sw = luminance(get_specular(S, m, sp.uv));
    float ws = dw + sw;
    if (ws <= 0.f) return R;
    float dpmf = dw / ws, spmf = sw / ws;
    F3 n = sp.fr.n;
    F3 fx = normalize(wi - n * dot(wi, n));
    F3 fy = cross(n, fx);
    M3 iso = m3_rows(fx, fy, n);
    M3 M, Minv;
    float m_pmf;
    int comp;
    if (bsdf_component <= dpmf) {
        Minv = iso;
        M = inverse(Minv);
        m_pmf = dpmf;
        comp = 0;
    } else {
        Minv = mul(inverse(get_ltc_matrix(S, mid, sp, wi)), iso);
        M = inverse(Minv);
        m_pmf = spmf;
        comp = 1;
    }
    // The reference weighs the whole BSDF by the chosen component's pmf
    // (delta_ray.cpp:86-101, 249, 296-330), so each component's line
    // density integrates the whole BSDF: twice the derivative when both
    // components are nonzero.  Without that quirk a sample counts only the
    // part of the BSDF of the component it was drawn from (unbiased).
    if (quirks & DR_QUIRK_EDGE_BSDF_SUM) comp = -1;
    // The reference's Jacobian of a side ray's hit point w.r.t. its direction
    // treats the ray as leaving the world origin (intersect_jacobian of
    // vector.h:690-716 gets the absolute hit position, delta_ray.cpp:329-331),
    // which scales it by (x . n) / ((x - p) . n); without that quirk the
    // hit is taken relative to the shading point p.
    const bool from_origin = (quirks & DR_QUIRK_EDGE_HIT_JACOBIAN) != 0;
    F3 a = mul(Minv, v0 - sp.pos);
    F3 b = mul(Minv, v1 - sp.pos);
    if (a.z <= 0.f && b.z <= 0.f) return R;
    if (a.z < 0.f) a = (a * b.z - b * a.z) / (b.z - a.z);
    if (b.z < 0.f) b = (a * b.z - b * a.z) / (b.z - a.z);
    LineSample L;
    if (!ltc_line_sample(a, b, eu, L)) return R;
    float l = L.l, pl = L.pdf;
    F3 vo = L.vo, wt = L.wt;
    F3 sample_p = mul(M, vo + l * wt);
    F3 spn = normalize(cross(v0 - sp.pos, v1 - sp.pos));
    float delta = 1e-3f * length(sample_p);
    F3 sdir = normalize(sample_p);
    F3 rdir = normalize(sdir + delta * spn);
    F3 ldir = normalize(sdir - delta * spn);
    SPoint spr, spl;
    Isect ir, il;
    intersect2(S, sp.pos, rdir, ldir, ir, spr, il, spl);
    bool rc = ir.shape == edge.shape && (ir.tri == edge.f0 || ir.tri == edge.f1);
    bool lc = il.shape == edge.shape && (il.tri == edge.f0 || il.tri == edge.f1);
    if (rc == lc || (ir.shape == is.shape && ir.tri == is.tri) || (il.shape == is.shape && il.tri == is.tri))
        return R;
    F3 eb = bsdf(S, mid, sp, wi, sdir);
    if (sum3(eb) < 1e-6f) return R;
    if (comp >= 0) eb = bsdf(S, mid, sp, wi, sdir, comp);
    // with `defer`, a sample that passed the connectivity test stops here
    // and is reported: the caller runs it again in full in a launch of such
    // samples only (about one in five), so the evaluation and continuation
    // code runs on full waves
    if (defer) {
        *defer = true;
        return R;
    }
    F3 shr = F3{0.f, 0.f, 0.f}, shl = F3{0.f, 0.f, 0.f};
    if (!direct_only) {
        shr = shade(S, -sdir, ir, spr, ls, ss, nss, false);
        shl = shade(S, -sdir, il, spl, ls, ss, nss, false);
        if (valid(ir)) {
            F3 dd = spr.pos - sp.pos;
            float dsq = length_sq(dd);
            if (dsq > 0.f) {
                shr = shr / fmaxf(dsq, 1e-2f);
                shr = shr * fabsf(dot(normalize(dd), spr.ng));
            } else {
                shr = F3{0.f, 0.f, 0.f};
            }
        }
        if (valid(il)) {
            F3 dd = spl.pos - sp.pos;
            float dsq = length_sq(dd);
            if (dsq > 0.f) {
                shl = shl / fmaxf(dsq, 1e-2f);
                shl = shl * fabsf(dot(normalize(dd), spl.ng));
            } else {
                shl = F3{0.f, 0.f, 0.f};
            }
        }
    }
    int rl = valid(ir) ? S.shapes[ir.shape].light : -1;
    int ll = valid(il) ? S.shapes[il.shape].light : -1;
    F3 er_e = eval_light(S, wi, is, sp, rl, spr.pos, spr.fr.n, comp);
    F3 el_e = eval_light(S, wi, is, sp, ll, spl.pos, spl.fr.n, comp);
    if (!(quirks & DR_QUIRK_EDGE_EMITTER_FULL)) {
        // on an emitter's own outline (the side that hits one of the edge's
        // faces is the emitter) only the BSDF-sampled share of its light has
        // a boundary: the light-sampled share is an area integral over the
        // emitter.  Light seen past another object's edge keeps its full
        // weight (both shares are occluded there).
        if (rl >= 0 && rc) er_e = er_e * emitter_bsdf_mis(S, mid, sp, wi, ir, spr);
        if (ll >= 0 && lc) el_e = el_e * emitter_bsdf_mis(S, mid, sp, wi, il, spl);
    }
    F3 er = er_e + eb * shr;
    F3 el = el_e + eb * shl;
    if (valid(ir)) {
        F3 dd = spr.pos - sp.pos;
        if (length(dd) * fabsf(dot(normalize(dd), spr.ng)) < 1e-3f) return R;
    }
    if (valid(il)) {
        F3 dd = spl.pos - sp.pos;
        if (length(dd) * fabsf(dot(normalize(dd), spl.ng)) < 1e-3f) return R;
    }
    F3 d0 = v0 - sp.pos, d1 = v1 - sp.pos;
    float dirac_j = length(cross(d0, d1));
    auto accum = [&](const SPoint &hp, F3 ev) {
        M3 nj = normalize_jacobian(sample_p);
        M3 ij = intersect_jacobian(sdir, from_origin ? hp.pos : hp.pos - sp.pos, hp.ng);
        float line_j = length(mul(mul(mul(ij, nj), M), wt)) / (m_pmf * pl);
        float w = line_j / dirac_j;
        F3 x = hp.pos - sp.pos;
        F3 dp = w * (cross(d1, d0) + cross(x, d1) + cross(d0, x));
        F3 dv0 = w * cross(d1, x);
        F3 dv1 = w * cross(x, d0);
        float we = dot(ev * weight, d_color);
        R.dp += dp * we;
        R.dv0 += dv0 * we;
        R.dv1 += dv1 * we;
    };
    if (valid(ir) && abs_sum(er) > 0.f) accum(spr, er);
    if (valid(il) && abs_sum(el) > 0.f) accum(spl, -el);
    return R;
}

struct PrimaryResult {
    bool ok;
    F3 dv0, dv1;
    float dw2c[16];
};

// delta_ray.cpp:377-557 (world Jacobian by 12-point finite differences, as there)
DR_INLINE void sample_primary_edge_point(const DrSceneView &S, const DrEdge &edge, float eu, const LightSample &ls,
                                         const SecondarySample *ss, int nss, const float *d_image, int W, int H,
                                         float weight, unsigned flags, PrimaryResult &R) {
    R.ok = false;
    const DrCamera &cam = S.cam;
    const DrShape &es = S.shapes[edge.shape];
    F3 v0 = vert(S, es.vert_off + edge.v0), v1 = vert(S, es.vert_off + edge.v1);
    F2 v0s, v1s;
    if (!project_edge(cam, v0, v1, v0s, v1s)) return;
    F3 v0d = unproject(cam, v0s), v1d = unproject(cam, v1s);
    F3 vd3 = v1d - v0d;
    float t = eu;
    F3 v0l = xfm_point(cam.w2c, v0), v1l = xfm_point(cam.w2c, v1);
    F3 ep3 = v0d + t * vd3;
    F3 elocal = v0l + t * v1l;
    F2 ep = project_local(cam, ep3);
    if (!in_screen(cam, ep)) return;
    int xi = min((int)(ep.x * W), W - 1), yi = min((int)(ep.y * H), H - 1);
    F3 d_color = load3(d_image + ((size_t)yi * W + xi) * 3);
    float delta = 1e-4f * length(elocal);
    F3 spn = normalize(cross(v0d, v1d));
    F3 rdir = normalize(ep3 + delta * spn);
    F2 rpt = project_local(cam, rdir);
    F3 rorg, rd;
    SPoint spr, spl;
    // (a fisheye side point outside the screen disk has no ray: a miss)
    Isect ir = no_hit();
    if (sample_primary(cam, rpt, rorg, rd)) ir = intersect(S, rorg, rd, spr);
    F3 ldir = normalize(ep3 - delta * spn);
    F2 lpt = project_local(cam, ldir);
    F3 lorg, ld;
    Isect il = no_hit();
    if (sample_primary(cam, lpt, lorg, ld)) il = intersect(S, lorg, ld, spl);
    bool rc = ir.shape == edge.shape && (ir.tri == edge.f0 || ir.tri == edge.f1);
    bool lc = il.shape == edge.shape && (il.tri == edge.f0 || il.tri == edge.f1);
    if (!rc && !lc) return;
    F3 rcol = shade(S, -rd, ir, spr, ls, ss, nss, true);
    F3 lcol = shade(S, -ld, il, spl, ls, ss, nss, true);
    F3 diff = (rcol - lcol) * weight;
    F3 xdiff = lpt.x > rpt.x ? -diff : diff;
    F3 ydiff = lpt.y > rpt.y ? -diff : diff;
    F3 edir = unproject(cam, ep);
    // d unproject / d screen at both endpoints and at the sample
    F3 dx0, dy0, dx1, dy1, dxe, dye;
    d_unproject(cam, v0s, dx0, dy0);
    d_unproject(cam, v1s, dx1, dy1);
    d_unproject(cam, ep, dxe, dye);
    F3 c01 = cross(v0d, v1d);
    float dadx = dot(dxe, c01), dady = dot(dye, c01);
    float da0x = fabsf(dot(cross(v1d, edir), dx0));
    float da0y = fabsf(dot(cross(v1d, edir), dy0));
    float da1x = fabsf(dot(cross(edir, v0d), dx1));
    float da1y = fabsf(dot(cross(edir, v0d), dy1));
    float dirac_j = 1.f / sqrtf(dadx * dadx + dady * dady);
    F3 ep3d = v0d + (t + delta) * vd3;
    F2 epd = project_local(cam, ep3d);
    F2 dlt = epd - ep;
    float line_j = length2(F2{dlt.x / delta, dlt.y / delta});
    float jac = line_j * dirac_j;
    float d0x = dot(-xdiff * da0x * jac, d_color);
    float d0y = dot(-ydiff * da0y * jac, d_color);
    float d1x = dot(-xdiff * da1x * jac, d_color);
    float d1y = dot(-ydiff * da1y * jac, d_color);
    const float fd = 1e-3f;
    F3 ex = F3{fd, 0.f, 0.f}, ey = F3{0.f, fd, 0.f}, ez = F3{0.f, 0.f, fd};
    float g[6];
    for (int k = 0; k < 6; k++) {
        F3 e = (k % 3 == 0) ? ex : ((k % 3 == 1) ? ey : ez);
        F2 p0, p1;
        bool ok = k < 3 ? project_edge(cam, v0 + e, v1, p0, p1) : project_edge(cam, v0, v1 + e, p0, p1);
        if (!ok) return;
        F2 a = p0 - v0s, b = p1 - v1s;
        a = F2{a.x / fd, a.y / fd};
        b = F2{b.x / fd, b.y / fd};
        g[k] = d0x * a.x + d0y * a.y + d1x * b.x + d1y * b.y;
    }
    R.dv0 = F3{g[0], g[1], g[2]};
    R.dv1 = F3{g[3], g[4], g[5]};
    for (int i = 0; i < 16; i++) R.dw2c[i] = 0.f;
    // the exact camera derivative carries project_local's 0.5 (the pinhole
    // d_project_local drops it, the fisheye one keeps it)
    const float cs = ((flags & DR_QUIRK_PROJECT_SCALE) || cam.fisheye) ? 1.f : 0.5f;
    d_project(cam, v0, v1, cs * d0x, cs * d0y, cs * d1x, cs * d1y, R.dw2c, (flags & DR_QUIRK_CLIP_DERIVATIVE) != 0);
    R.ok = true;
}
