// Device refit of a scene whose vertex positions changed (no reference
// counterpart: the reference rebuilds its Scene, Embree BVH and edge trees
// inside every render() call, scene.cpp:19-136, edge_tree.cpp:263-316).
// The trees keep their topology; every quantity stored in them is recomputed
// from the new vertices with the host builder's arithmetic (scene_build.cpp)
// and in its order, so a refit to the vertices a scene was built from
// reproduces the built scene:
//   k_refit_tris       triangle records (v0, e1, e2) in leaf order
//   k_refit_bvh        BVH child boxes, bottom-up one level at a time
//   k_edge_filter      the degenerate / coplanar edge filter (scene.cpp:80-96);
//                      a change of its verdicts means the edge set changed and
//                      the host rebuilds
//   k_edge_flags       the same verdicts as per-edge flags, for trees built over
//                      every mesh edge (vertex updates then never rebuild)
//   k_refit_geo        per-tree-edge geometry records (endpoints, faces' third vertices)
//   k_edge_normals     face normals of normal-tree edges (edge_tree.cpp:218-249)
//   k_refit_etree      edge-tree boxes, lengths, normal sums, learned lengths
//   k_etree_cone       edge-tree normal cones (a wave per node)
//   k_prim_len         clipped screen length of camera silhouettes (edge_tree.cpp:266-294)
//   k_prim_build       the primary-edge list and its distribution (distribution.h:7-59)
// All of it is compiled with -ffp-contract=off, as the host builder.
#include "refit.h"

#include "dr_math.h"

#define RF_BLOCK 256
#define RF_TREE_BLOCK 1024

namespace {

DR_INLINE F3 vtx(const float *verts, int g) { return F3{verts[3 * (size_t)g], verts[3 * (size_t)g + 1], verts[3 * (size_t)g + 2]}; }
// global vertex k of local triangle t of shape sh
DR_INLINE int tri_vertex(const int *tri_vidx, const DrShape &sh, int t, int k) { return tri_vidx[3 * (size_t)(sh.tri_off + t) + k]; }
DR_INLINE F3 face_normal(const float *verts, const int *tri_vidx, const DrShape &sh, int t) {
    F3 a = vtx(verts, tri_vertex(tri_vidx, sh, t, 0)), b = vtx(verts, tri_vertex(tri_vidx, sh, t, 1)),
       c = vtx(verts, tri_vertex(tri_vidx, sh, t, 2));
    return normalize(cross(b - a, c - a));
}
DR_INLINE float minf_(float a, float b) { return b < a ? b : a; }  // std::min(a, b)
DR_INLINE float maxf_(float a, float b) { return a < b ? b : a; }  // std::max(a, b)

}  // namespace

__global__ void k_refit_tris(DrTri *tris, const float *verts, const int *tri_vidx, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    DrTri &T = tris[i];
    int g = __float_as_int(T.v0[3]);
    const int *vi = tri_vidx + 3 * (size_t)g;
    F3 a = vtx(verts, vi[0]), b = vtx(verts, vi[1]), c = vtx(verts, vi[2]);
    T.v0[0] = a.x; T.v0[1] = a.y; T.v0[2] = a.z;
    T.e1[0] = b.x - a.x; T.e1[1] = b.y - a.y; T.e1[2] = b.z - a.z;
    T.e2[0] = c.x - a.x; T.e2[1] = c.y - a.y; T.e2[2] = c.z - a.z;
}

// One workgroup; levels deepest first, so a node's children are final when
// it is reached.  A leaf slot's box is the union of its triangles' vertex
// boxes; an interior slot's the union of the child's two boxes; a slot with
// neither (the empty sibling of a root leaf) is left alone.
__global__ void __launch_bounds__(RF_TREE_BLOCK) k_refit_bvh(DrNode *nodes, const DrTri *tris, const float *verts,
                                                             const int *tri_vidx, const int *lv, const int *lv_off,
                                                             int nlev) {
    for (int l = 0; l < nlev; l++) {
        for (int k = lv_off[l] + (int)threadIdx.x; k < lv_off[l + 1]; k += blockDim.x) {
            DrNode &N = nodes[lv[k]];
            float box[2][6];
            bool set[2] = {false, false};
            for (int sl = 0; sl < 2; sl++) {
                int child = N.d[sl], count = N.d[2 + sl];
                float lo[3] = {__builtin_huge_valf(), __builtin_huge_valf(), __builtin_huge_valf()};
                float hi[3] = {-__builtin_huge_valf(), -__builtin_huge_valf(), -__builtin_huge_valf()};
                if (count > 0) {
                    for (int t = child; t < child + count; t++) {
                        const int *vi = tri_vidx + 3 * (size_t)__float_as_int(tris[t].v0[3]);
                        for (int v = 0; v < 3; v++) {
                            F3 p = vtx(verts, vi[v]);
                            lo[0] = minf_(lo[0], p.x); lo[1] = minf_(lo[1], p.y); lo[2] = minf_(lo[2], p.z);
                            hi[0] = maxf_(hi[0], p.x); hi[1] = maxf_(hi[1], p.y); hi[2] = maxf_(hi[2], p.z);
                        }
                    }
                    set[sl] = true;
                } else if (child > 0) {
                    const DrNode &C = nodes[child];
                    float c0lo[3] = {C.a[0], C.a[1], C.a[2]}, c0hi[3] = {C.a[3], C.b[0], C.b[1]};
                    float c1lo[3] = {C.b[2], C.b[3], C.c[0]}, c1hi[3] = {C.c[1], C.c[2], C.c[3]};
                    for (int q = 0; q < 3; q++) {
                        lo[q] = minf_(c0lo[q], c1lo[q]);
                        hi[q] = maxf_(c0hi[q], c1hi[q]);
                    }
                    set[sl] = true;
                }
                for (int q = 0; q < 3; q++) {
                    box[sl][q] = lo[q];
                    box[sl][3 + q] = hi[q];
                }
            }
            if (set[0]) {
                N.a[0] = box[0][0]; N.a[1] = box[0][1]; N.a[2] = box[0][2]; N.a[3] = box[0][3];
                N.b[0] = box[0][4]; N.b[1] = box[0][5];
            }
            if (set[1]) {
                N.b[2] = box[1][0]; N.b[3] = box[1][1];
                N.c[0] = box[1][2]; N.c[1] = box[1][3]; N.c[2] = box[1][4]; N.c[3] = box[1][5];
            }
        }
        __syncthreads();
    }
}

// reference scene.cpp:80-96: is mesh edge e kept on the current vertices
DR_INLINE bool filter_keeps(const DrEdge &e, const float *verts, const int *tri_vidx, const DrShape &sh) {
    F3 dv = vtx(verts, sh.vert_off + e.v0) - vtx(verts, sh.vert_off + e.v1);
    if (fabsf(dv.x) + fabsf(dv.y) + fabsf(dv.z) < 1e-10f) return false;
    if (e.f1 != -1) {
        float ang = fabsf(dot(face_normal(verts, tri_vidx, sh, e.f0), face_normal(verts, tri_vidx, sh, e.f1)));
        if (ang == 1.f || ang == -1.f) return false;
    }
    return true;
}

// the filter on the new vertices; counts edges whose verdict differs from
// the one the scene was built with
__global__ void k_edge_filter(const DrEdge *cand, const int *keep, int n, const float *verts, const int *tri_vidx,
                              const DrShape *shapes, unsigned int *changed) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DrEdge e = cand[i];
    if ((filter_keeps(e, verts, tri_vidx, shapes[e.shape]) ? 1 : 0) != keep[i]) atomicAdd(changed, 1u);
}

// the filter's verdicts as flags of the tree edges (trees over every mesh
// edge); counts the edges switched off
__global__ void k_edge_flags(const DrEdge *edges, int n, const float *verts, const int *tri_vidx,
                             const DrShape *shapes, unsigned char *on, unsigned int *off_count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DrEdge e = edges[i];
    bool k = filter_keeps(e, verts, tri_vidx, shapes[e.shape]);
    on[i] = k ? 1 : 0;
    if (!k) atomicAdd(off_count, 1u);
}

// per tree edge, 16 floats: endpoints, and for a normal-tree edge each face's
// vertex off the edge (xyz + pad each)
__global__ void k_refit_geo(float *geo, const DrEdge *edges, int n, bool faces, const float *verts,
                            const int *tri_vidx, const DrShape *shapes) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DrEdge e = edges[i];
    const DrShape sh = shapes[e.shape];
    int g0 = sh.vert_off + e.v0, g1 = sh.vert_off + e.v1;
    int gv[4] = {g0, g1, 0, 0};
    int nk = 2;
    if (faces) {
        const int *i0 = tri_vidx + 3 * (size_t)(sh.tri_off + e.f0);
        const int *i1 = tri_vidx + 3 * (size_t)(sh.tri_off + e.f1);
        gv[2] = i0[0] != g0 && i0[0] != g1 ? i0[0] : (i0[1] != g0 && i0[1] != g1 ? i0[1] : i0[2]);
        gv[3] = i1[0] != g0 && i1[0] != g1 ? i1[0] : (i1[1] != g0 && i1[1] != g1 ? i1[1] : i1[2]);
        nk = 4;
    }
    for (int k = 0; k < nk; k++) {
        F3 p = vtx(verts, gv[k]);
        geo[16 * (size_t)i + 4 * k] = p.x;
        geo[16 * (size_t)i + 4 * k + 1] = p.y;
        geo[16 * (size_t)i + 4 * k + 2] = p.z;
    }
}

// the two face normals of each normal-tree edge, turned to the side of the
// edge's summed shading normal (6 floats per edge)
__global__ void k_edge_normals(float *on, const DrEdge *edges, int n, const float *verts, const float *normals,
                               const int *tri_vidx, const DrShape *shapes) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DrEdge e = edges[i];
    const DrShape sh = shapes[e.shape];
    F3 sn = normalize(vtx(normals, sh.vert_off + e.v0) + vtx(normals, sh.vert_off + e.v1));
    F3 n0 = face_normal(verts, tri_vidx, sh, e.f0);
    F3 n1 = face_normal(verts, tri_vidx, sh, e.f1);
    if (dot(n0, sn) < 0.f) n0 = -n0;
    if (dot(n1, sn) < 0.f) n1 = -n1;
    float *o = on + 6 * (size_t)i;
    o[0] = n0.x; o[1] = n0.y; o[2] = n0.z;
    o[3] = n1.x; o[4] = n1.y; o[5] = n1.z;
}

// One workgroup, levels deepest first (children follow their parent in the
// node array).  Leaf: box of its edges' endpoints, summed edge lengths, summed
// face normals in edge order, summed learned lengths; interior: the union /
// sums of its two children.  nsum: 3 floats per node; wsum: 1 per node.
// flags (trees over every mesh edge): edges switched off count no length.
__global__ void __launch_bounds__(RF_TREE_BLOCK) k_refit_etree(DrEdgeNode *nodes, const float *geo, const float *on,
                                                               const float *w, const unsigned char *flags,
                                                               const int *lv, const int *lv_off, int nlev,
                                                               float *nsum, float *wsum) {
    for (int l = 0; l < nlev; l++) {
        for (int k = lv_off[l] + (int)threadIdx.x; k < lv_off[l + 1]; k += blockDim.x) {
            int i = lv[k];
            DrEdgeNode &N = nodes[i];
            if (N.nprims > 0) {
                float lo[3] = {__builtin_huge_valf(), __builtin_huge_valf(), __builtin_huge_valf()};
                float hi[3] = {-__builtin_huge_valf(), -__builtin_huge_valf(), -__builtin_huge_valf()};
                float len = 0.f, wl = 0.f;
                F3 ns = F3{0.f, 0.f, 0.f};
                for (int q = 0; q < N.nprims; q++) {
                    int ei = N.offset + q;
                    const float *g = geo + 16 * (size_t)ei;
                    F3 a = F3{g[0], g[1], g[2]}, b = F3{g[4], g[5], g[6]};
                    lo[0] = minf_(lo[0], minf_(a.x, b.x)); lo[1] = minf_(lo[1], minf_(a.y, b.y));
                    lo[2] = minf_(lo[2], minf_(a.z, b.z));
                    hi[0] = maxf_(hi[0], maxf_(a.x, b.x)); hi[1] = maxf_(hi[1], maxf_(a.y, b.y));
                    hi[2] = maxf_(hi[2], maxf_(a.z, b.z));
                    // an edge the filter switched off keeps its place in the
                    // box and the normal cone (conservative) but adds no length
                    float el = (!flags || flags[ei]) ? length(b - a) : 0.f;
                    len += el;
                    if (w) wl += maxf_(w[ei], 0.f) * el;
                    if (on) {
                        const float *o = on + 6 * (size_t)ei;
                        ns = ns + F3{o[0], o[1], o[2]};
                        ns = ns + F3{o[3], o[4], o[5]};
                    }
                }
                N.lo[0] = lo[0]; N.lo[1] = lo[1]; N.lo[2] = lo[2];
                N.hi[0] = hi[0]; N.hi[1] = hi[1]; N.hi[2] = hi[2];
                N.total_length = len;
                wsum[i] = wl;
                nsum[3 * (size_t)i] = ns.x; nsum[3 * (size_t)i + 1] = ns.y; nsum[3 * (size_t)i + 2] = ns.z;
            } else {
                const DrEdgeNode &L = nodes[i + 1], &R = nodes[N.offset];
                for (int q = 0; q < 3; q++) {
                    N.lo[q] = minf_(L.lo[q], R.lo[q]);
                    N.hi[q] = maxf_(L.hi[q], R.hi[q]);
                }
                N.total_length = L.total_length + R.total_length;
                wsum[i] = wsum[i + 1] + wsum[N.offset];
                for (int q = 0; q < 3; q++) nsum[3 * (size_t)i + q] = nsum[3 * (size_t)(i + 1) + q] + nsum[3 * (size_t)N.offset + q];
            }
        }
        __syncthreads();
    }
}

// A wave per node: axis = normalised normal sum, cos_angle = the least
// cosine to it over the node's edges [rng[2i], rng[2i+1]) (the minimum does
// not depend on the order; a NaN cosine never replaces it, as std::min);
// learned length = summed weighted lengths.
__global__ void __launch_bounds__(RF_BLOCK) k_etree_cone(DrEdgeNode *nodes, int n, const float *on, const int *rng,
                                                         const float *nsum, const float *wsum, bool learned) {
    const int lane = (int)__lane_id();
    int i = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) / 64);
    if (i >= n) return;
    DrEdgeNode &N = nodes[i];
    F3 axis = normalize(F3{nsum[3 * (size_t)i], nsum[3 * (size_t)i + 1], nsum[3 * (size_t)i + 2]});
    float cmin = 1.f;
    if (on) {
        for (int k = rng[2 * i] + lane; k < rng[2 * i + 1]; k += 64) {
            const float *o = on + 6 * (size_t)k;
            cmin = minf_(cmin, dot(F3{o[0], o[1], o[2]}, axis));
            cmin = minf_(cmin, dot(F3{o[3], o[4], o[5]}, axis));
        }
    }
    for (int s = 32; s > 0; s >>= 1) cmin = minf_(cmin, __shfl_xor(cmin, s));
    if (lane == 0) {
        N.axis[0] = axis.x; N.axis[1] = axis.y; N.axis[2] = axis.z;
        N.cos_angle = cmin;
        // unscaled, so that the two trees' learned totals also decide
        // between the trees (unit weights give the plain lengths exactly)
        N.pad = learned ? wsum[i] : N.total_length;
    }
}

// ---------------------------------------------------------- primary edges --
namespace {
// (xfm_point of dr_math.h has the host builder's arithmetic)
DR_INLINE F2 project_local(const DrCamera &c, F3 l) {
    if (c.fisheye) {
        F3 d = normalize(l);
        float phi = atan2f(d.y, d.x);
        float r = acosf(d.z) * 2.f / DR_PI;
        return F2{0.5f * (-r * cosf(phi) + 1.f), 0.5f * (-r * sinf(phi) + 1.f)};
    }
    return F2{(-l.x / (l.z * c.fov) + 1.f) * 0.5f, (-l.y / (l.z * c.fov * c.aspect) + 1.f) * 0.5f};
}
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
// Cohen-Sutherland clip to [0,1]^2 (reference line_clip.h:13-102)
DR_INLINE int clip_code(F2 v) {
    int c = 0;
    if (v.x < 0.f) c |= 1; else if (v.x > 1.f) c |= 2;
    if (v.y < 0.f) c |= 4; else if (v.y > 1.f) c |= 8;
    return c;
}
DR_INLINE bool clip_line(F2 &a, F2 &b) {
    int c0 = clip_code(a), c1 = clip_code(b);
    for (int it = 0; it < 8; it++) {  // each round fixes one outcode bit of one end
        if (!(c0 | c1)) return true;
        if (c0 & c1) return false;
        int out = c0 ? c0 : c1;
        F2 v = F2{0.f, 0.f};
        if (out & 8) { v.x = a.x + (b.x - a.x) * (1.f - a.y) / (b.y - a.y); v.y = 1.f; }
        else if (out & 4) { v.x = a.x + (b.x - a.x) * (0.f - a.y) / (b.y - a.y); v.y = 0.f; }
        else if (out & 2) { v.y = a.y + (b.y - a.y) * (1.f - a.x) / (b.x - a.x); v.x = 1.f; }
        else if (out & 1) { v.y = a.y + (b.y - a.y) * (0.f - a.x) / (b.x - a.x); v.x = 0.f; }
        if (out == c0) { a = v; c0 = clip_code(a); } else { b = v; c1 = clip_code(b); }
    }
    return !(c0 | c1);
}
// reference edge.h:59-79 at point p
DR_INLINE bool is_silhouette(const float *verts, const int *tri_vidx, const DrShape &sh, F3 p, const DrEdge &e) {
    if (!sh.has_n) return true;
    if (e.f0 == -1 || e.f1 == -1) return true;
    F3 v0 = vtx(verts, sh.vert_off + e.v0), v1 = vtx(verts, sh.vert_off + e.v1);
    auto third = [&](int f) {
        for (int k = 0; k < 3; k++) {
            int id = tri_vertex(tri_vidx, sh, f, k) - sh.vert_off;
            if (id != e.v0 && id != e.v1) return vtx(verts, sh.vert_off + id);
        }
        return F3{0.f, 0.f, 0.f};
    };
    F3 a = third(e.f0), b = third(e.f1);
    F3 n0 = normalize(cross(v0 - a, v1 - a));
    F3 n1 = normalize(cross(v1 - b, v0 - b));
    bool f0 = dot(p - v0, n0) > 0.f && dot(p - v1, n0) > 0.f && dot(p - a, n0) > 0.f;
    bool f1 = dot(p - v0, n1) > 0.f && dot(p - v1, n1) > 0.f && dot(p - b, n1) > 0.f;
    return f0 != f1;
}
}  // namespace

__global__ void k_prim_len(const DrEdge *edges, int n, DrCamera cam, const float *verts, const int *tri_vidx,
                           const DrShape *shapes, bool filter, float *len) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DrEdge e = edges[i];
    const DrShape sh = shapes[e.shape];
    if (filter && !filter_keeps(e, verts, tri_vidx, sh)) {
        len[i] = 0.f;
        return;
    }
    F3 org = xfm_point(cam.c2w, F3{0.f, 0.f, 0.f});
    float l = 0.f;
    F2 a, b;
    if (project_edge(cam, vtx(verts, sh.vert_off + e.v0), vtx(verts, sh.vert_off + e.v1), a, b)) {
        if (clip_line(a, b)) {
            float dx = b.x - a.x, dy = b.y - a.y;
            l = sqrtf(dx * dx + dy * dy);
        }
    }
    if (!is_silhouette(verts, tri_vidx, sh, org, e)) l = 0.f;
    len[i] = l;
}

// One workgroup.  The edges of positive length, in order, with a pbrt-style
// distribution of their lengths (make_dist of the host builder: cdf[i] =
// cdf[i-1] + func[i-1] / n accumulated in order, then divided by its last
// value, or i / n when that is 0).  out_n[0] = the count; out_int[0] = the
// integral.
__global__ void __launch_bounds__(RF_TREE_BLOCK) k_prim_build(const DrEdge *edges, const float *len, int n,
                                                              DrEdge *out_edges, float *func, float *cdf, int *out_n,
                                                              float *out_int) {
    __shared__ int s_scan[RF_TREE_BLOCK];
    __shared__ float s_q[RF_TREE_BLOCK];
    __shared__ int s_base;
    const int t = (int)threadIdx.x;
    if (t == 0) s_base = 0;
    __syncthreads();
    // compaction, chunk by chunk with a block-wide inclusive scan of the flags
    for (int c0 = 0; c0 < n; c0 += RF_TREE_BLOCK) {
        int i = c0 + t;
        int f = (i < n && len[i] > 0.f) ? 1 : 0;
        s_scan[t] = f;
        __syncthreads();
        for (int off = 1; off < RF_TREE_BLOCK; off <<= 1) {
            int v = t >= off ? s_scan[t - off] : 0;
            __syncthreads();
            s_scan[t] += v;
            __syncthreads();
        }
        int base = s_base;
        if (f) {
            int pos = base + s_scan[t] - 1;
            out_edges[pos] = edges[i];
            func[pos] = len[i];
        }
        __syncthreads();
        if (t == RF_TREE_BLOCK - 1) s_base = base + s_scan[t];
        __syncthreads();
    }
    const int m = s_base;
    if (m == 0) {
        if (t == 0) {
            out_n[0] = 0;
            out_int[0] = 0.f;
        }
        return;
    }
    // cdf[i] = cdf[i-1] + func[i-1] / m, in order (one lane adds, from LDS)
    float acc = 0.f;
    if (t == 0) cdf[0] = 0.f;
    for (int c0 = 0; c0 < m; c0 += RF_TREE_BLOCK) {
        int i = c0 + t;
        if (i < m) s_q[t] = func[i] / (float)m;
        __syncthreads();
        if (t == 0) {
            int cn = min(RF_TREE_BLOCK, m - c0);
            for (int k = 0; k < cn; k++) {
                acc = acc + s_q[k];
                s_q[k] = acc;
            }
        }
        __syncthreads();
        if (i < m) cdf[i + 1] = s_q[t];
        __syncthreads();
    }
    __shared__ float s_int;
    if (t == 0) s_int = acc;
    __syncthreads();
    const float fi = s_int;
    for (int i = 1 + t; i <= m; i += RF_TREE_BLOCK) {
        if (fi == 0.f) cdf[i] = float(i) / float(m);
        else cdf[i] = cdf[i] / fi;
    }
    if (t == 0) {
        out_n[0] = m;
        out_int[0] = fi;
    }
}

// ------------------------------------------------------------------ host --
namespace drrefit {

static unsigned grid(long long n, int b) { return (unsigned)((n + b - 1) / b); }

void refit_geometry(const RefitView &v, hipStream_t st) {
    if (v.num_tris > 0) k_refit_tris<<<grid(v.num_tris, RF_BLOCK), RF_BLOCK, 0, st>>>(v.tris, v.verts, v.tri_vidx, v.num_tris);
    if (v.bvh_nlev > 0)
        k_refit_bvh<<<1, RF_TREE_BLOCK, 0, st>>>(v.nodes, v.tris, v.verts, v.tri_vidx, v.bvh_lv, v.bvh_lv_off, v.bvh_nlev);
}

void edge_filter(const RefitView &v, unsigned int *changed, hipStream_t st) {
    if (v.num_cand > 0)
        k_edge_filter<<<grid(v.num_cand, RF_BLOCK), RF_BLOCK, 0, st>>>(v.cand, v.cand_keep, v.num_cand, v.verts,
                                                                       v.tri_vidx, v.shapes, changed);
}

void edge_flags(const RefitView &v, const DrEdge *edges, int n, unsigned char *on, unsigned int *off_count,
                hipStream_t st) {
    if (n > 0)
        k_edge_flags<<<grid(n, RF_BLOCK), RF_BLOCK, 0, st>>>(edges, n, v.verts, v.tri_vidx, v.shapes, on, off_count);
}

void refit_edge_tree(const RefitView &v, const EdgeTreeRefit &t, hipStream_t st) {
    if (t.num_nodes == 0) return;
    k_refit_geo<<<grid(t.num_edges, RF_BLOCK), RF_BLOCK, 0, st>>>(t.geo, t.edges, t.num_edges, t.directional, v.verts,
                                                                    v.tri_vidx, v.shapes);
    if (t.directional)
        k_edge_normals<<<grid(t.num_edges, RF_BLOCK), RF_BLOCK, 0, st>>>(t.on, t.edges, t.num_edges, v.verts,
                                                                          v.normals, v.tri_vidx, v.shapes);
    k_refit_etree<<<1, RF_TREE_BLOCK, 0, st>>>(t.nodes, t.geo, t.directional ? t.on : nullptr, t.weights, t.flags,
                                               t.lv, t.lv_off, t.nlev, t.nsum, t.wsum);
    k_etree_cone<<<grid(64ll * t.num_nodes, RF_BLOCK), RF_BLOCK, 0, st>>>(
        t.nodes, t.num_nodes, t.directional ? t.on : nullptr, t.rng, t.nsum, t.wsum, t.weights != nullptr);
}

void primary_edges(const RefitView &v, const DrCamera &cam, const PrimaryOut &o, hipStream_t st) {
    if (v.num_edges == 0) return;
    k_prim_len<<<grid(v.num_edges, RF_BLOCK), RF_BLOCK, 0, st>>>(v.edges, v.num_edges, cam, v.verts, v.tri_vidx,
                                                                 v.shapes, v.filter_edges, o.len);
    k_prim_build<<<1, RF_TREE_BLOCK, 0, st>>>(v.edges, o.len, v.num_edges, o.edges, o.func, o.cdf, o.count,
                                              o.func_int);
}

}  // namespace drrefit
