// Device-resident scene layout shared by the host builder (scene_build.cpp)
// and the kernels (render.hip).  Plain data only.
//
// HBM layout (all arrays 16-byte aligned, one allocation per array):
//   nodes      BVH2 nodes, 64 B: both child boxes in the parent so one node
//              fetch tests two boxes (one 64 B cache-line-friendly load)
//   tris       48 B per triangle in leaf order: v0, e1 = v1-v0, e2 = v2-v0
//              (float4 each; v0.w carries the global triangle id)
//   verts      float3 per global vertex (all shapes concatenated)
//   tri_vidx   int3 global vertex ids per global triangle
//   tri_shape  shape id per global triangle
//   uvs/normals per global vertex, present for shapes that have them
//   texels     every material texture, concatenated
//   edges      mesh edges (sorted by shape, v0, v1), edge-tree nodes and the
//              primary-edge distribution
//   ltc        16384 x 9 floats (the fitted LTC matrices)
#pragma once

#include <stdint.h>

#define DR_MAX_SHAPES 65536
// Traversal stack sizes of ray queries (entries, one per BVH level) and of
// the edge-tree walk (depth + 1 entries); the builder rejects deeper trees.
#define DR_STACK 64
#define DR_TREE_STACK 32
// The ordinal of a selected edge within its pass takes 27 bits of the
// random-number stream word (the pass flag is bit 30 of the ordinal).
#define DR_EDGE_ORDINAL_BITS 27
#define DR_ORD_ALL_PASS 0x40000000

struct DrNode {
    // child 0 box: a.xyz lo, a.w hi.x, b.xy hi.yz; child 1 box: b.zw lo.xy,
    // c.x lo.z, c.yzw hi.xyz; d = (child0, child1, count0, count1) where a
    // count > 0 marks a leaf whose child field is its first triangle slot
    float a[4], b[4], c[4];
    int d[4];
};

struct DrTri {
    float v0[4];  // w: global triangle id (int bits)
    float e1[4];
    float e2[4];
};

struct DrShape {
    int vert_off, tri_off, nv, nt;
    int has_uv, has_n, mat, light;
    long long vgrad_off, ngrad_off;  // gradient-buffer offsets (ngrad -1: none)
};

struct DrTex {
    int off, h, w, c;  // h == 0: constant of c floats at off
    long long grad_off;
};

struct DrMaterial {
    DrTex diffuse, specular, roughness;
    float duv[2], suv[2], ruv[2];
    int two_sided, pad;
};

struct DrLight {
    int shape, ntri, func_off, cdf_off;  // triangle-area distribution
    float I[3], lum;
    float area;          // total area of the light's triangles
    float func_int;      // integral of the triangle distribution
    float sphere_c[3];   // bounding sphere of the light's triangles
    float sphere_r;
    long long grad_off;
};

struct DrEdge {
    int shape, v0, v1, f0, f1;  // shape-local vertex / face ids (f1 -1: boundary)
};

struct DrEdgeNode {
    float lo[3], hi[3];
    float axis[3];
    float cos_angle;  // cosine of the normal cone's half angle (least cosine of a face normal to the axis)
    int offset;   // leaf: first edge in the tree's edge list; interior: second child
    int nprims;   // 0: interior
    float total_length;
    float pad;    // the learned proposal's length: sum of weight x edge length (total_length without weights)
};

struct DrCamera {
    float c2w[16], w2c[16];
    float fov, aspect, clip_near;
    int fisheye;  // equi-angular projection (the screen disk's radius maps linearly to the angle)
};

// Everything a kernel needs, passed by value.
// Device code reads the scene through pointers in the global address space,
// so its loads compile to global_* rather than generic flat_* instructions.
#ifndef DR_GLOBAL
#if defined(__HIP_DEVICE_COMPILE__)
#define DR_GLOBAL __attribute__((address_space(1)))
#else
#define DR_GLOBAL
#endif
#endif
struct DrSceneView {
    const DR_GLOBAL DrNode *nodes;
    const DR_GLOBAL DrTri *tris;
    const DR_GLOBAL float *verts;
    const DR_GLOBAL int *tri_vidx;
    const DR_GLOBAL int *tri_shape;
    const DR_GLOBAL float *uvs;
    const DR_GLOBAL float *normals;
    const DR_GLOBAL float *texels;
    const DR_GLOBAL DrShape *shapes;
    const DR_GLOBAL DrMaterial *materials;
    const DR_GLOBAL DrLight *lights;
    const DR_GLOBAL float *light_func;  // per light: triangle areas, then cdf (n+1)
    const DR_GLOBAL float *light_cdf;
    const DR_GLOBAL float *lights_func;  // light selection by luminance
    const DR_GLOBAL float *lights_cdf;
    float lights_func_int, light_total_intensity;
    int num_lights, num_shapes;
    int num_tris, num_nodes;
    const DR_GLOBAL DrEdge *sil_edges;
    const DR_GLOBAL DrEdgeNode *sil_nodes;
    const DR_GLOBAL DrEdge *nrm_edges;
    const DR_GLOBAL DrEdgeNode *nrm_nodes;
    const DR_GLOBAL float *nrm_geo;  // per normal-tree edge: v0, v1 and the faces' third vertices (float4 each)
    const DR_GLOBAL float *sil_geo;  // per silhouette-tree edge: v0, v1 (same layout)
    int num_sil_nodes, num_nrm_nodes;
    int num_sil_edges;
    // per tree edge (silhouette tree, then normal tree): 0 when the
    // degenerate / coplanar filter drops the edge on the current vertices;
    // NULL when the trees hold only the edges the filter kept at the build
    const DR_GLOBAL unsigned char *edge_on;
    float *edge_stat;  // per tree edge (silhouette tree, then normal tree): sum of |contribution|, count, sum of squares
    // per shape: 1 when the statistics count gradients of its vertices, 0
    // when not; NULL: every shape counts
    const DR_GLOBAL float *stat_shape;
    const DR_GLOBAL DrEdge *prim_edges;
    const DR_GLOBAL float *prim_func;
    const DR_GLOBAL float *prim_cdf;
    float prim_func_int;
    int num_prim_edges;
    const DR_GLOBAL float *ltc;
    DrCamera cam;
    unsigned long long *stats;  // NULL or 5 counters
};
