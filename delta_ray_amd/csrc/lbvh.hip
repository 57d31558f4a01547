// Device build of the triangle BVH (dr_scene_build_bvh_device; no reference
// counterpart: the reference builds an Embree BVH on the host in every
// render() call, scene.cpp:100-135).  A linear BVH after Karras, "Maximizing
// parallelism in the construction of BVHs, octrees, and k-d trees" (HPG
// 2012): triangles sorted by the 30-bit Morton code of their box centre in
// the scene's centre bounds (the triangle index below it makes keys unique),
// a binary radix tree over the sorted keys (internal node i covers a key
// range whose split is the first differing bit), then boxes bottom up, each
// node finished by the second of its two children's threads to arrive.  One
// triangle per leaf, folded into its parent's slot (count 1), in the DrNode
// layout of the host builder; the host relabels the nodes in preorder and
// derives the refit levels.  Ray hits do not depend on the tree (closest t,
// ties to the smaller triangle id), so they equal the host-built tree's.
#include <hipcub/hipcub.hpp>

#include "lbvh.h"

namespace {

#define LB_BLOCK 256

DR_INLINE F3 vtx3(const float *v, int g) { return F3{v[3 * (size_t)g], v[3 * (size_t)g + 1], v[3 * (size_t)g + 2]}; }

// orderable int of a float (a < b as floats <=> key(a) < key(b) as ints)
DR_INLINE int ord_key(float f) {
    int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
DR_INLINE float ord_val(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

// the 10 low bits of v spread to every third bit
DR_INLINE unsigned spread3(unsigned v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void k_lb_bounds(const float *verts, const int *tri_vidx, int n, int *bounds) {
    int lo[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, hi[3] = {-0x7fffffff - 1, -0x7fffffff - 1, -0x7fffffff - 1};
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        F3 a = vtx3(verts, tri_vidx[3 * t]), b = vtx3(verts, tri_vidx[3 * t + 1]), c = vtx3(verts, tri_vidx[3 * t + 2]);
        float cx = 0.5f * (fminf(fminf(a.x, b.x), c.x) + fmaxf(fmaxf(a.x, b.x), c.x));
        float cy = 0.5f * (fminf(fminf(a.y, b.y), c.y) + fmaxf(fmaxf(a.y, b.y), c.y));
        float cz = 0.5f * (fminf(fminf(a.z, b.z), c.z) + fmaxf(fmaxf(a.z, b.z), c.z));
        int k[3] = {ord_key(cx), ord_key(cy), ord_key(cz)};
        for (int q = 0; q < 3; q++) {
            lo[q] = min(lo[q], k[q]);
            hi[q] = max(hi[q], k[q]);
        }
    }
    for (int q = 0; q < 3; q++) {
        for (int o = 32; o > 0; o >>= 1) {
            lo[q] = min(lo[q], __shfl_xor(lo[q], o));
            hi[q] = max(hi[q], __shfl_xor(hi[q], o));
        }
    }
    if (__lane_id() == 0) {
        for (int q = 0; q < 3; q++) {
            atomicMin(&bounds[q], lo[q]);
            atomicMax(&bounds[3 + q], hi[q]);
        }
    }
}

__global__ void k_lb_keys(const float *verts, const int *tri_vidx, int n, const int *bounds, uint64_t *keys) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    F3 a = vtx3(verts, tri_vidx[3 * t]), b = vtx3(verts, tri_vidx[3 * t + 1]), c = vtx3(verts, tri_vidx[3 * t + 2]);
    float cen[3] = {0.5f * (fminf(fminf(a.x, b.x), c.x) + fmaxf(fmaxf(a.x, b.x), c.x)),
                    0.5f * (fminf(fminf(a.y, b.y), c.y) + fmaxf(fmaxf(a.y, b.y), c.y)),
                    0.5f * (fminf(fminf(a.z, b.z), c.z) + fmaxf(fmaxf(a.z, b.z), c.z))};
    unsigned code = 0;
    for (int q = 0; q < 3; q++) {
        float lo = ord_val(bounds[q]), hi = ord_val(bounds[3 + q]);
        float ext = hi - lo;
        float u = ext > 0.f ? (cen[q] - lo) / ext : 0.f;
        unsigned cell = (unsigned)fminf(fmaxf(u * 1024.f, 0.f), 1023.f);
        code |= spread3(cell) << (2 - q);
    }
    keys[t] = ((uint64_t)code << 32) | (uint64_t)(unsigned)t;
}

// length of the common prefix of keys i and j; -1 outside [0, n)
DR_INLINE int prefix(const uint64_t *k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    return __clzll((long long)(k[i] ^ k[j]));
}

// Internal node i: its key range [min(i, j), max(i, j)] and split g (Karras
// 2012, section 4).  Children: a code >= 0 is an internal node, -1 - p the
// leaf at sorted position p.
__global__ void k_lb_tree(const uint64_t *k, int n, int *child, int *node_parent, int *leaf_parent) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = prefix(k, n, i, i + 1) - prefix(k, n, i, i - 1) >= 0 ? 1 : -1;
    int dmin = prefix(k, n, i, i - d);
    int lmax = 2;
    while (prefix(k, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (prefix(k, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = prefix(k, n, i, j);
    int s = 0, step = l;
    do {
        step = (step + 1) >> 1;
        if (prefix(k, n, i, i + (s + step) * d) > dnode) s += step;
    } while (step > 1);
    int g = i + s * d + min(d, 0);
    int lo = min(i, j), hi = max(i, j);
    int left = lo == g ? -1 - g : g;
    int right = hi == g + 1 ? -1 - (g + 1) : g + 1;
    child[2 * i] = left;
    child[2 * i + 1] = right;
    if (left >= 0) node_parent[left] = i;
    else leaf_parent[-1 - left] = i;
    if (right >= 0) node_parent[right] = i;
    else leaf_parent[-1 - right] = i;
}

__global__ void k_lb_tris(const float *verts, const int *tri_vidx, int n, const uint64_t *keys, DrTri *tris) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int g = (int)(unsigned)(keys[p] & 0xffffffffull);
    F3 a = vtx3(verts, tri_vidx[3 * g]), b = vtx3(verts, tri_vidx[3 * g + 1]), c = vtx3(verts, tri_vidx[3 * g + 2]);
    DrTri T;
    T.v0[0] = a.x; T.v0[1] = a.y; T.v0[2] = a.z; T.v0[3] = __int_as_float(g);
    T.e1[0] = b.x - a.x; T.e1[1] = b.y - a.y; T.e1[2] = b.z - a.z; T.e1[3] = 0.f;
    T.e2[0] = c.x - a.x; T.e2[1] = c.y - a.y; T.e2[2] = c.z - a.z; T.e2[3] = 0.f;
    tris[p] = T;
}

DR_INLINE void tri_box(const float *verts, const int *tri_vidx, int g, float lo[3], float hi[3]) {
    for (int q = 0; q < 3; q++) {
        lo[q] = __builtin_huge_valf();
        hi[q] = -__builtin_huge_valf();
    }
    for (int v = 0; v < 3; v++) {
        F3 p = vtx3(verts, tri_vidx[3 * g + v]);
        lo[0] = fminf(lo[0], p.x); lo[1] = fminf(lo[1], p.y); lo[2] = fminf(lo[2], p.z);
        hi[0] = fmaxf(hi[0], p.x); hi[1] = fmaxf(hi[1], p.y); hi[2] = fmaxf(hi[2], p.z);
    }
}

// One thread per leaf walks up; the second thread to reach a node writes it
// (both children are final then).  Node boxes are read through volatile
// loads: another workgroup wrote them.
__global__ void k_lb_boxes(const float *verts, const int *tri_vidx, int n, const uint64_t *keys, const int *child,
                           const int *node_parent, const int *leaf_parent, unsigned *arrived, DrNode *nodes) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int node = leaf_parent[p];
    while (node >= 0) {
        __threadfence();
        if (atomicAdd(&arrived[node], 1u) == 0) return;
        DrNode N;
        for (int sl = 0; sl < 2; sl++) {
            int c = child[2 * node + sl];
            float lo[3], hi[3];
            if (c < 0) {
                int pos = -1 - c;
                tri_box(verts, tri_vidx, (int)(unsigned)(keys[pos] & 0xffffffffull), lo, hi);
                N.d[sl] = pos;
                N.d[2 + sl] = 1;
            } else {
                const volatile DrNode &C = nodes[c];
                float c0lo[3] = {C.a[0], C.a[1], C.a[2]}, c0hi[3] = {C.a[3], C.b[0], C.b[1]};
                float c1lo[3] = {C.b[2], C.b[3], C.c[0]}, c1hi[3] = {C.c[1], C.c[2], C.c[3]};
                for (int q = 0; q < 3; q++) {
                    lo[q] = fminf(c0lo[q], c1lo[q]);
                    hi[q] = fmaxf(c0hi[q], c1hi[q]);
                }
                N.d[sl] = c;
                N.d[2 + sl] = 0;
            }
            if (sl == 0) {
                N.a[0] = lo[0]; N.a[1] = lo[1]; N.a[2] = lo[2]; N.a[3] = hi[0];
                N.b[0] = hi[1]; N.b[1] = hi[2];
            } else {
                N.b[2] = lo[0]; N.b[3] = lo[1];
                N.c[0] = lo[2]; N.c[1] = hi[0]; N.c[2] = hi[1]; N.c[3] = hi[2];
            }
        }
        nodes[node] = N;
        __threadfence();
        node = node == 0 ? -1 : node_parent[node];
    }
}

__global__ void k_lb_fill_i(int *p, long long n, int v) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

}  // namespace

namespace drlbvh {

hipError_t build(const float *verts, const int *tri_vidx, int n, DrNode *nodes, DrTri *tris, void *scratch,
                 size_t scratch_bytes, hipStream_t hs) {
    if (n < 2) return hipErrorInvalidValue;
    Layout L = layout(n);
    if (scratch_bytes < L.total) return hipErrorInvalidValue;
    char *s = (char *)scratch;
    int *bounds = (int *)(s + L.bounds);
    uint64_t *keys_in = (uint64_t *)(s + L.keys_in), *keys = (uint64_t *)(s + L.keys);
    int *child = (int *)(s + L.child), *node_parent = (int *)(s + L.node_parent), *leaf_parent = (int *)(s + L.leaf_parent);
    unsigned *arrived = (unsigned *)(s + L.arrived);
    int nb = (n + LB_BLOCK - 1) / LB_BLOCK;
    int init[6] = {0x7fffffff, 0x7fffffff, 0x7fffffff, -0x7fffffff - 1, -0x7fffffff - 1, -0x7fffffff - 1};
    hipError_t e = hipMemcpyAsync(bounds, init, sizeof(init), hipMemcpyHostToDevice, hs);
    if (e != hipSuccess) return e;
    k_lb_bounds<<<std::min(nb, 1024), LB_BLOCK, 0, hs>>>(verts, tri_vidx, n, bounds);
    k_lb_keys<<<nb, LB_BLOCK, 0, hs>>>(verts, tri_vidx, n, bounds, keys_in);
    size_t tb = L.sort_bytes;
    e = hipcub::DeviceRadixSort::SortKeys(s + L.sort, tb, keys_in, keys, n, 0, 64, hs);
    if (e != hipSuccess) return e;
    k_lb_fill_i<<<256, LB_BLOCK, 0, hs>>>(node_parent, n, -1);
    k_lb_fill_i<<<256, LB_BLOCK, 0, hs>>>((int *)arrived, n, 0);
    k_lb_tree<<<nb, LB_BLOCK, 0, hs>>>(keys, n, child, node_parent, leaf_parent);
    k_lb_tris<<<nb, LB_BLOCK, 0, hs>>>(verts, tri_vidx, n, keys, tris);
    k_lb_boxes<<<nb, LB_BLOCK, 0, hs>>>(verts, tri_vidx, n, keys, child, node_parent, leaf_parent, arrived, nodes);
    return hipGetLastError();
}

Layout layout(int n) {
    Layout L;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        off = (off + 255) & ~(size_t)255;
        size_t at = off;
        off += bytes;
        return at;
    };
    L.bounds = take(6 * sizeof(int));
    L.keys_in = take((size_t)n * sizeof(uint64_t));
    L.keys = take((size_t)n * sizeof(uint64_t));
    L.child = take(2 * (size_t)n * sizeof(int));
    L.node_parent = take((size_t)n * sizeof(int));
    L.leaf_parent = take((size_t)n * sizeof(int));
    L.arrived = take((size_t)n * sizeof(unsigned));
    size_t tb = 0;
    (void)hipcub::DeviceRadixSort::SortKeys(nullptr, tb, (uint64_t *)nullptr, (uint64_t *)nullptr, n, 0, 64);
    L.sort_bytes = tb;
    L.sort = take(tb);
    L.total = off;
    return L;
}

}  // namespace drlbvh
