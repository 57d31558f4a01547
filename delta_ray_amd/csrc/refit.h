// Launchers of the device refit kernels (refit.hip), called by the C-ABI in
// render.hip.  Plain device pointers; every launch is ordered on the stream.
#pragma once

#include <hip/hip_runtime.h>

#include "dr_types.h"

namespace drrefit {

struct RefitView {
    // geometry (global vertex, triangle and shape arrays)
    const float *verts, *normals;
    const int *tri_vidx;
    const DrShape *shapes;
    DrTri *tris;
    int num_tris;
    DrNode *nodes;
    const int *bvh_lv, *bvh_lv_off;
    int bvh_nlev;
    // mesh edges kept by the filter (sorted) and all candidates with the
    // verdict they were built with
    const DrEdge *edges;
    int num_edges;
    const DrEdge *cand;
    const int *cand_keep;
    int num_cand;
    // edges holds every mesh edge (trees over all candidates): the primary
    // edges take the filter's verdict on the current vertices
    bool filter_edges;
};

struct EdgeTreeRefit {
    DrEdgeNode *nodes;
    int num_nodes;
    const DrEdge *edges;  // the tree's ordered edges
    int num_edges;
    bool directional;     // the normal tree (face normals and cones)
    float *geo;           // 16 floats per edge
    float *on;            // 6 floats per edge (directional only)
    const float *weights; // learned per-edge weights or NULL
    const unsigned char *flags;  // per edge: 0 = switched off by the filter (no length); NULL: all on
    const int *lv, *lv_off, *rng;
    int nlev;
    float *nsum, *wsum;   // scratch: 3 floats / 1 float per node
};

struct PrimaryOut {
    float *len;           // scratch, one float per mesh edge
    DrEdge *edges;        // capacity: the mesh edges
    float *func, *cdf;    // capacity: mesh edges / mesh edges + 1
    int *count;           // device: the number of primary edges
    float *func_int;      // device: the distribution's integral
};

// triangle records and BVH boxes
void refit_geometry(const RefitView &v, hipStream_t st);
// adds to *changed (device) the candidate edges whose filter verdict changed
void edge_filter(const RefitView &v, unsigned int *changed, hipStream_t st);
// the filter's verdicts of `edges` as flags on[i] (1: kept); adds the edges
// switched off to *off_count (device)
void edge_flags(const RefitView &v, const DrEdge *edges, int n, unsigned char *on, unsigned int *off_count,
                hipStream_t st);
// geometry records, boxes, lengths, cones and learned lengths of one tree
void refit_edge_tree(const RefitView &v, const EdgeTreeRefit &t, hipStream_t st);
// primary-edge list and distribution for a camera
void primary_edges(const RefitView &v, const DrCamera &cam, const PrimaryOut &o, hipStream_t st);

}  // namespace drrefit
