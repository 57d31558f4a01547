// Device build of the triangle BVH (lbvh.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "dr_types.h"
#include "dr_math.h"

namespace drlbvh {
// byte offsets of the build's working arrays in one scratch allocation
struct Layout {
    size_t bounds, keys_in, keys, child, node_parent, leaf_parent, arrived, sort, sort_bytes, total;
};
Layout layout(int n);
// n >= 2 triangles (global vertex indices tri_vidx[3 n] into verts[3 *]):
// nodes[n - 1] (internal node i in the DrNode layout, root 0; not in
// preorder) and tris[n] (leaf order), stream-ordered.
hipError_t build(const float *verts, const int *tri_vidx, int n, DrNode *nodes, DrTri *tris, void *scratch,
                 size_t scratch_bytes, hipStream_t hs);
}  // namespace drlbvh
