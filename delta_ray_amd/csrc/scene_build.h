// Host-side scene preparation for the device renderer: flattening of shapes
// into global arrays, the triangle BVH, edge extraction, the two edge trees
// and the light / primary-edge distributions.
#pragma once

#include "../../include/delta_ray.h"
#include "dr_types.h"

#include <string>
#include <vector>

namespace drh {

struct Dist {
    std::vector<float> func, cdf;
    float func_int = 0.f;
};

struct HostScene {
    DrCamera cam;
    std::vector<DrShape> shapes;
    std::vector<DrMaterial> materials;
    std::vector<DrLight> lights;
    std::vector<float> verts, uvs, normals, texels;
    std::vector<int> tri_vidx, tri_shape;
    std::vector<DrNode> nodes;
    std::vector<DrTri> tris;
    std::vector<float> light_func, light_cdf;  // per light, concatenated
    Dist light_sel;
    float light_total_intensity = 0.f;
    // trees over every mesh edge (the filter's verdicts become per-edge flags
    // on the device) instead of the edges the filter keeps
    bool all_candidates = false;
    std::vector<DrEdge> edges;  // all edges, sorted by (shape, v0, v1)
    std::vector<DrEdge> sil_edges, nrm_edges;
    std::vector<DrEdgeNode> sil_nodes, nrm_nodes;
    // per normal-tree edge (tree order), 16 floats: its endpoints and the
    // third vertices of its two faces, xyz + 1 pad each (the silhouette test);
    // per silhouette-tree edge the same layout with the endpoints only
    std::vector<float> nrm_geo, sil_geo;
    std::vector<DrEdge> prim_edges;
    Dist prim;
    size_t grad_size = 0;
    int bvh_depth = 0, sil_depth = 0, nrm_depth = 0;  // deepest node (root 0)
    // lights as given: shape, intensity (3 per light), gradient offset
    std::vector<int> light_shapes;
    std::vector<float> light_I;
    std::vector<size_t> light_goff;
    // every mesh edge before the degenerate / coplanar filter, in the order
    // of `edges`, and whether the filter kept it
    std::vector<DrEdge> cand_edges;
    std::vector<int> cand_keep;
    // refit tables: nodes by depth, deepest level first, with level offsets;
    // per edge-tree node the range [first, last) of its ordered edges
    std::vector<int> bvh_lv, bvh_lv_off, sil_lv, sil_lv_off, nrm_lv, nrm_lv_off;
    std::vector<int> sil_rng, nrm_rng;
};

// Returns an empty string on success, else the reason for rejecting desc.
std::string build_host_scene(const dr_scene_desc *desc, HostScene &hs);
// From the current vertices: the light distributions / the BVH, edges, edge
// trees, geometry records, primary edges and refit tables (a rebuild)
void build_lights(HostScene &hs);
std::string build_structures(HostScene &hs);
void build_geo_records(HostScene &hs);
// A triangle BVH built on the device (internal nodes, root 0, in any order;
// tris in leaf order) in place of the host-built one: relabelled in
// preorder, with refit levels.  Returns an empty string, or why it was
// refused (the host tree is then kept).
std::string adopt_device_bvh(HostScene &hs, const std::vector<DrNode> &built, std::vector<DrTri> &tris);
DrCamera make_camera(const dr_camera &c);
void set_camera(HostScene &hs, const dr_camera &c);
// Primary-edge list and distribution (pmf proportional to the clipped screen
// length of camera silhouettes) for the scene's camera / for a given camera
void build_primary_edges(HostScene &hs);
void primary_edges_for(const HostScene &hs, const DrCamera &cam, std::vector<DrEdge> &edges, Dist &dist);
// DR_MLP_FEATURES floats per tree edge (silhouette tree, then normal tree)
void edge_features(const HostScene &hs, const DrCamera &cam, std::vector<float> &out);

}  // namespace drh
