// Kernels and C-ABI of the MI355X path tracer (include/delta_ray.h).
//
// Kernels (256-thread workgroups = 4 wave64):
//   k_forward   radiance only, thread = (pixel, chunk of samples)
//               (reference render_sample_fast, delta_ray.cpp:640-692)
//   adjoint (reference render_sample, delta_ray.cpp:694-1002) as a wavefront
//   pipeline over batches of pixel samples:
//     k_path    camera path + radiance + interior adjoint; stores one
//               PointRec per (sample, bounce) shading point
//     k_select  edge-tree traversal per shading point -> queue of edges
//     k_edge    thread = one selected edge: secondary edge sample, vertex
//               gradients, Dirac position derivative dp of its point
//     k_dp_last keeps the last dp per point (set_adjoint overwrite)
//     k_dp      replays the path to propagate dp (2nd adjoint sweep)
//   k_primary   primary-edge samples (reference pass 2, delta_ray.cpp:1468-1550)
//   k_intersect ray queries for the parity tests
// Gradients: camera matrices are reduced in LDS per workgroup (64 floats);
// the other parameters are scattered with float atomics (zero adjoints are
// skipped).
#include "../../include/delta_ray.h"
#include "dr_sample.h"
#include "scene_build.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#define DR_BLOCK 256

// ---------------------------------------------------------------- kernels --
__global__ void k_fill(float *p, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// thread -> (pixel, chunk of `per` samples)
__global__ void __launch_bounds__(DR_BLOCK) k_forward(DrSceneView S, SampleCtx c, int per, int chunks, float *image) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long total = (long long)c.W * c.H * chunks;
    if (t >= total) return;
    int pix = (int)(t / chunks), ch = (int)(t % chunks);
    int xi = pix % c.W, yi = pix / c.W;
    F3 acc = F3{0.f, 0.f, 0.f};
    int s0 = ch * per, s1 = min(s0 + per, c.spp);
    for (int s = s0; s < s1; s++) acc += forward_sample(S, c, xi, yi, s);
    float *px = image + 3 * (size_t)pix;
    if (chunks == 1) {
        px[0] = acc.x; px[1] = acc.y; px[2] = acc.z;
    } else {
        atomicAdd(px, acc.x); atomicAdd(px + 1, acc.y); atomicAdd(px + 2, acc.z);
    }
}

// ---- wavefront adjoint: path (A) -> edge selection (B) -> edge samples (C)
// -> Dirac position sweep (D).  A thread index t of A and D enumerates
// (pixel, chunk of `per` samples); a point slot is ((t - t0) * per + k) * MAXB + b.
struct EdgeEntry {
    int slot;
    int edge;  // index into the tree's edge list; bit 31 set: normal tree
    float w;
    int j;     // ordinal of the edge in the point's selection
};

struct SlotDecode {
    int pix, xi, yi, s, b;
    uint64_t sid;
};
DR_INLINE SlotDecode decode_slot(const SampleCtx &c, int per, int chunks, long long t0, long long slot, int maxb) {
    SlotDecode d;
    d.b = (int)(slot % maxb);
    long long rest = slot / maxb;
    int k = (int)(rest % per);
    long long t = t0 + rest / per;
    d.pix = (int)(t / chunks);
    d.s = (int)(t % chunks) * per + k;
    d.xi = d.pix % c.W;
    d.yi = d.pix / c.W;
    d.sid = sample_id(c, d.xi, d.yi, d.s);
    return d;
}

template <int MAXB>
__global__ void __launch_bounds__(DR_BLOCK) k_path(DrSceneView S, SampleCtx c, int per, int chunks, long long t0,
                                                   long long t1, const float *d_image, float *image, float *grads,
                                                   PointRec *pts) {
    __shared__ float s_cam[64];
    __shared__ GradCache gc;
    if (threadIdx.x < 64) s_cam[threadIdx.x] = 0.f;
    gc_init(&gc);
    __syncthreads();
    long long t = t0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < t1) {
        int pix = (int)(t / chunks), ch = (int)(t % chunks);
        int xi = pix % c.W, yi = pix / c.W;
        F3 d_color = load3(d_image + 3 * (size_t)pix);
        GradSink G;
        G.grads = grads;
        G.cam = s_cam;
        G.flags = c.flags;
        G.color_mode = true;
        G.gc = &gc;
        F3 acc = F3{0.f, 0.f, 0.f};
        bool zero = is_zero(d_color);
        for (int k = 0; k < per; k++) {
            int s = ch * per + k;
            PointRec *P = pts + ((t - t0) * per + k) * MAXB;
            if (s >= c.spp || zero) {
                for (int b = 0; b < MAXB; b++) P[b].valid = 0;
                if (s < c.spp) acc += forward_sample(S, c, xi, yi, s);
                continue;
            }
            acc += stage_path<MAXB>(S, c, G, xi, yi, s, d_color, P);
        }
        float *px = image + 3 * (size_t)pix;
        if (chunks == 1) {
            px[0] = acc.x; px[1] = acc.y; px[2] = acc.z;
        } else {
            atomicAdd(px, acc.x); atomicAdd(px + 1, acc.y); atomicAdd(px + 2, acc.z);
        }
    }
    __syncthreads();
    if (threadIdx.x < 64 && s_cam[threadIdx.x] != 0.f) atomicAdd(grads + threadIdx.x, s_cam[threadIdx.x]);
    gc_flush(&gc, grads);
}

// B: edge selection, persistent.  Per-point traversal lengths vary by two
// orders of magnitude (a point inside many node boxes splits at every level),
// so a lane that has finished its point fetches the next one: idle lanes of a
// wave take a batch of point slots from `work` with one atomic, and every
// iteration advances each busy lane's traversal by one node.  Selected edges
// are appended to the queue in groups of up to 8 per lane (one queue atomic
// per group).  Every lane leaves the loop once `work` passes nslots.
__global__ void __launch_bounds__(DR_BLOCK) k_select(DrSceneView S, SampleCtx c, int per, int chunks, long long t0,
                                                     long long nslots, int maxb, const PointRec *pts, EdgeEntry *q,
                                                     unsigned long long *counter, long long cap, int *counts,
                                                     unsigned long long *work) {
    const int lane = (int)__lane_id();
    EdgeTraversal T;
    Philox prng;
    SPoint sp;
    F3 wi = F3{0.f, 0.f, 0.f};
    long long slot = -1;
    int j = 0;
    bool more = true;
    EdgeEntry buf[8];
    int nbuf = 0;
    auto flush = [&]() {
        unsigned long long base = atomicAdd(counter, (unsigned long long)nbuf);
        for (int e = 0; e < nbuf; e++)
            if ((long long)(base + e) < cap) q[base + e] = buf[e];
        nbuf = 0;
    };
    auto emit = [&](const DrEdge &ed, float w, bool nrm, int idx) {
        // zero-weight edges carry no work (reference delta_ray.cpp:906-908)
        if (w != 0.f) {
            buf[nbuf++] = EdgeEntry{(int)slot, idx | (nrm ? (int)0x80000000 : 0), w, j};
            if (nbuf == 8) flush();
        }
        j++;
    };
    while (true) {
        bool want = slot < 0 && more;
        unsigned long long idle = __ballot(want);
        if (idle) {
            int leader = __ffsll((long long)idle) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(work, (unsigned long long)__popcll(idle));
            base = __shfl(base, leader);
            if (want) {
                unsigned long long below = lane == 0 ? 0ull : (idle & (~0ull >> (64 - lane)));
                long long s = (long long)(base + __popcll(below));
                if (s >= nslots) {
                    more = false;
                } else {
                    const PointRec &p = pts[s];
                    bool busy = false;
                    if (p.valid) {
                        SlotDecode d = decode_slot(c, per, chunks, t0, s, maxb);
                        busy = select_begin(S, c, d.sid, d.b, p, T, prng);
                    }
                    if (busy) {
                        slot = s;
                        sp = p.sp;
                        wi = p.wi;
                        j = 0;
                    } else {
                        counts[s] = 0;
                    }
                }
            }
        }
        if (slot < 0) {
            if (!more) break;
            continue;
        }
        if (!et_step(T, S, wi, sp, prng, emit)) {
            if (nbuf > 0) flush();
            counts[slot] = j;
            if (S.stats) atomicAdd(&S.stats[3], (unsigned long long)j);
            slot = -1;
        }
    }
}

// C: one thread per selected edge
template <int MAXB>
DR_INLINE void edge_entry(DrSceneView S, SampleCtx c, int per, int chunks, long long t0, const EdgeEntry *q,
                          long long e, const PointRec *pts, const float *d_image, float *grads, float *pt_dp,
                          int *pt_lastj, float4 *entry_dp, GradCache *gcp) {
    EdgeEntry E = q[e];
    SlotDecode d = decode_slot(c, per, chunks, t0, E.slot, MAXB);
    const PointRec &p = pts[E.slot];
    int idx = E.edge & 0x7fffffff;
    const DrEdge &edge = (E.edge & 0x80000000) ? S.nrm_edges[idx] : S.sil_edges[idx];
    GradSink G;
    G.grads = grads;
    G.cam = nullptr;
    G.flags = c.flags;
    G.color_mode = true;
    G.gc = gcp;
    F3 d_color = load3(d_image + 3 * (size_t)d.pix);
    F3 dp = stage_edge<MAXB>(S, c, G, d.sid, d.b, p, edge, E.w, E.j, d_color);
    bool nz = !is_zero(dp);
    if (c.flags & DR_QUIRK_DP_LAST) {
        entry_dp[e] = make_float4(dp.x, dp.y, dp.z, nz ? __int_as_float(E.j) : __int_as_float(-1));
        if (nz) atomicMax(pt_lastj + E.slot, E.j);
    } else if (nz) {
        atomicAdd(pt_dp + 3ll * E.slot, dp.x);
        atomicAdd(pt_dp + 3ll * E.slot + 1, dp.y);
        atomicAdd(pt_dp + 3ll * E.slot + 2, dp.z);
    }
}

template <int MAXB>
__global__ void __launch_bounds__(DR_BLOCK) k_edge(DrSceneView S, SampleCtx c, int per, int chunks, long long t0,
                                                   const EdgeEntry *q, long long n, const PointRec *pts,
                                                   const float *d_image, float *grads, float *pt_dp, int *pt_lastj,
                                                   float4 *entry_dp) {
    __shared__ GradCache gc;
    gc_init(&gc);
    __syncthreads();
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) edge_entry<MAXB>(S, c, per, chunks, t0, q, e, pts, d_image, grads, pt_dp, pt_lastj, entry_dp, &gc);
    __syncthreads();
    gc_flush(&gc, grads);
}

// reference set_adjoint overwrite: of several dp at one point the last
// (largest ordinal) is the one propagated
__global__ void k_dp_last(const EdgeEntry *q, long long n, const float4 *entry_dp, const int *pt_lastj,
                          float *pt_dp) {
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    float4 v = entry_dp[e];
    int j = __float_as_int(v.w);
    int slot = q[e].slot;
    if (j >= 0 && j == pt_lastj[slot]) {
        pt_dp[3ll * slot] = v.x;
        pt_dp[3ll * slot + 1] = v.y;
        pt_dp[3ll * slot + 2] = v.z;
    }
}

// D: Dirac position sweep per pixel sample, and the dx_image edge counts
template <int MAXB>
__global__ void __launch_bounds__(DR_BLOCK) k_dp(DrSceneView S, SampleCtx c, int per, int chunks, long long t0,
                                                 long long t1, const PointRec *pts, const float *pt_dp,
                                                 const int *counts, float *grads, float *dx_image) {
    __shared__ float s_cam[64];
    __shared__ GradCache gc;
    if (threadIdx.x < 64) s_cam[threadIdx.x] = 0.f;
    gc_init(&gc);
    __syncthreads();
    long long t = t0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < t1) {
        int pix = (int)(t / chunks), ch = (int)(t % chunks);
        int xi = pix % c.W, yi = pix / c.W;
        GradSink G;
        G.grads = grads;
        G.cam = s_cam;
        G.flags = c.flags;
        G.color_mode = false;
        G.gc = &gc;
        float dx = 0.f;
        for (int k = 0; k < per; k++) {
            int s = ch * per + k;
            if (s >= c.spp) break;
            long long base = ((t - t0) * per + k) * MAXB;
            const PointRec *P = pts + base;
            F3 dps[MAXB];
            int last = -1;
            for (int b = 0; b < MAXB; b++) {
                dps[b] = load3(pt_dp + 3 * (base + b));
                if (P[b].valid) last = b;
            }
            if (last >= 0) dx += counts[base + last] / 1000.f;
            stage_dp<MAXB>(S, c, G, xi, yi, s, P, dps);
        }
        if (dx_image) {
            float *q = dx_image + 3 * (size_t)pix;
            if (chunks == 1) {
                q[0] = q[1] = q[2] = dx;
            } else {
                atomicAdd(q, dx); atomicAdd(q + 1, dx); atomicAdd(q + 2, dx);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < 64 && s_cam[threadIdx.x] != 0.f) atomicAdd(grads + threadIdx.x, s_cam[threadIdx.x]);
    gc_flush(&gc, grads);
}

__global__ void k_fill_i(int *p, size_t n, int v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

template <int MAXB>
__global__ void __launch_bounds__(DR_BLOCK) k_primary(DrSceneView S, SampleCtx c, long long N, const float *d_image,
                                                      float *grads) {
    __shared__ float s_w2c[16];
    if (threadIdx.x < 16) s_w2c[threadIdx.x] = 0.f;
    __syncthreads();
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < N && S.num_prim_edges > 0) {
        Philox rng;
        philox_init(rng, c.seed, 1, (uint64_t)k, DR_STREAM_MAIN);
        float u = philox_next(rng);
        int id = sample_discrete(S.prim_cdf, S.num_prim_edges, u);
        float pmf = S.prim_func[id] / (S.prim_func_int * S.num_prim_edges);
        float w = pmf > 0.f ? 1.f / pmf : 0.f;
        if (w > 0.f) {
            philox_next(rng);  // bsdf component of the edge sample (unused for primary edges)
            float eu = philox_next(rng);
            LightSample ls = draw_light_sample(rng);
            SecondarySample ss[MAXB];
            for (int b = 0; b < c.nb; b++) ss[b] = draw_secondary(rng);
            const DrEdge &e = S.prim_edges[id];
            PrimaryResult R;
            sample_primary_edge_point(S, e, eu, ls, ss, c.nb, d_image, c.W, c.H, w / c.spp, c.flags, R);
            if (R.ok) {
                const DrShape &es = S.shapes[e.shape];
                if (R.dv0.x != 0.f) atomicAdd(grads + es.vgrad_off + 3ll * e.v0, R.dv0.x);
                if (R.dv0.y != 0.f) atomicAdd(grads + es.vgrad_off + 3ll * e.v0 + 1, R.dv0.y);
                if (R.dv0.z != 0.f) atomicAdd(grads + es.vgrad_off + 3ll * e.v0 + 2, R.dv0.z);
                if (R.dv1.x != 0.f) atomicAdd(grads + es.vgrad_off + 3ll * e.v1, R.dv1.x);
                if (R.dv1.y != 0.f) atomicAdd(grads + es.vgrad_off + 3ll * e.v1 + 1, R.dv1.y);
                if (R.dv1.z != 0.f) atomicAdd(grads + es.vgrad_off + 3ll * e.v1 + 2, R.dv1.z);
                for (int i = 0; i < 16; i++)
                    if (R.dw2c[i] != 0.f) atomicAdd(&s_w2c[i], R.dw2c[i]);
            }
        }
        if (S.stats) atomicAdd(&S.stats[4], 1ull);
    }
    __syncthreads();
    if (threadIdx.x < 16 && s_w2c[threadIdx.x] != 0.f) atomicAdd(grads + 16 + threadIdx.x, s_w2c[threadIdx.x]);
}

__global__ void k_intersect(DrSceneView S, const float *rays, long long n, int *hits, int any) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *r = rays + 8 * i;
    F3 org = F3{r[0], r[1], r[2]}, dir = F3{r[4], r[5], r[6]};
    int *h = hits + 4 * i;
    if (any) {
        h[0] = trace_any(S, org, dir, r[3], r[7]) ? 1 : 0;
        return;
    }
    float u = 0.f, v = 0.f;
    Isect is = trace_closest(S, org, dir, r[3], r[7], u, v);
    h[0] = is.shape;
    h[1] = is.tri;
    h[2] = __float_as_int(u);
    h[3] = __float_as_int(v);
}

// The four line-integral estimators of the reference's test_line_integral
// (delta_ray.cpp:1572-1836): a constant emitter on the plane y = 3 seen along
// the segment (-10, 3, 0)-(10, 3, 0) from the origin, weighted by a cosine lobe
// about +y and by the glossy LTC lobe Minv = diag(3, 3, 1) * frame(+y);
// (0) uniform position on the segment, cosine lobe; (1) the LTC line sampler,
// cosine lobe; (2) uniform, glossy; (3) the LTC line sampler, glossy.
// out: per estimator the sum and the sum of squares (double).
__global__ void __launch_bounds__(256) k_line_integral(long long n, uint64_t seed, double *out) {
    __shared__ double red[8][256];
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < n) {
        const F3 p1 = F3{-10.f, 3.f, 0.f}, p2 = F3{10.f, 3.f, 0.f};
        const float y = 3.f;
        const F3 down = F3{0.f, -1.f, 0.f};
        Frame fr = make_frame(F3{0.f, 1.f, 0.f});
        M3 Fm = m3_rows(fr.x, fr.y, fr.n);
        M3 Ft = m3_rows(F3{fr.x.x, fr.y.x, fr.n.x}, F3{fr.x.y, fr.y.y, fr.n.y}, F3{fr.x.z, fr.y.z, fr.n.z});
        M3 S = m3_rows(F3{3.f, 0.f, 0.f}, F3{0.f, 3.f, 0.f}, F3{0.f, 0.f, 1.f});
        M3 Si = m3_rows(F3{1.f / 3.f, 0.f, 0.f}, F3{0.f, 1.f / 3.f, 0.f}, F3{0.f, 0.f, 1.f});
        M3 Minv_g = mul(S, Fm), M_g = mul(Ft, Si);
        const float det_g = 9.f;
        Philox rng;
        for (int k = 0; k < 4; k++) {
            philox_init(rng, seed, 2, (uint64_t)i, (uint32_t)k);
            float u = philox_next(rng);
            bool glossy = k >= 2;
            F3 p, wline;
            float inv_pdf = 1.f;
            if ((k & 1) == 0) {
                p = p1 + u * (p2 - p1);
                wline = p2 - p1;
            } else {
                const M3 &Minv = glossy ? Minv_g : Fm;
                LineSample L;
                if (!ltc_line_sample(mul(Minv, p1), mul(Minv, p2), u, L)) continue;
                F3 lp = L.vo + L.l * L.wt;
                const M3 &M = glossy ? M_g : Ft;
                p = mul(M, lp);
                wline = mul(M, L.wt);
                inv_pdf = 1.f / L.pdf;
            }
            F3 dir = normalize(p);
            F3 hit = dir * (y / dir.y);
            M3 J = mul(intersect_jacobian(dir, hit, down), normalize_jacobian(p));
            float jac = length(mul(J, wline)) * inv_pdf;
            float geom = dir.y / length_sq(hit);
            float brdf;
            if (!glossy) {
                brdf = dir.y / DR_PI;
            } else {
                F3 td = mul(Minv_g, dir);
                float tl = length(td);
                brdf = (td.z / tl) / DR_PI * (det_g / (tl * tl * tl));
            }
            v[k] = (double)(brdf * geom * jac);
        }
    }
    for (int k = 0; k < 4; k++) {
        red[2 * k][threadIdx.x] = v[k];
        red[2 * k + 1][threadIdx.x] = v[k] * v[k];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int j = 0; j < 8; j++) red[j][threadIdx.x] += red[j][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 8) atomicAdd(out + threadIdx.x, red[threadIdx.x][0]);
}

// ------------------------------------------------------------------- host --
extern "C" const float dr_ltc_blob[];
extern "C" const float dr_ltc_blob_end[];

static thread_local std::string g_err;

static dr_status fail(dr_status s, const std::string &msg) {
    g_err = msg;
    return s;
}

// kernel timing: event pairs recorded around launches when enabled
enum { TK_FORWARD = 0, TK_PATH = 1, TK_PRIMARY = 2, TK_SELECT = 3, TK_EDGE = 4, TK_DP = 5, TK_COUNT = 6 };

// scratch of the wavefront adjoint, grown on demand, owned by the scene
struct AdjBuffers {
    PointRec *pts = nullptr;
    float *pt_dp = nullptr;
    int *pt_lastj = nullptr;
    int *counts = nullptr;
    long long slot_cap = 0;
    EdgeEntry *q = nullptr;
    float4 *entry_dp = nullptr;
    long long q_cap = 0;
    unsigned long long *counter = nullptr;
    unsigned long long *counter_host = nullptr;
};
struct TimedLaunch {
    int kind;
    hipEvent_t a, b;
};

struct dr_scene {
    drh::HostScene hs;
    DrSceneView view;
    std::vector<void *> allocs;
    unsigned long long *stats_dev = nullptr;
    bool stats_on = false;
    bool timing_on = false;
    std::vector<TimedLaunch> timed;
    std::vector<hipEvent_t> event_pool;
    AdjBuffers adj;
    double build_ms = 0.0;
    void *prim_bufs[3] = {nullptr, nullptr, nullptr};
};

template <class T>
static bool upload(dr_scene *s, const std::vector<T> &v, const T **out) {
    if (v.empty()) {
        *out = nullptr;
        return true;
    }
    void *p = nullptr;
    if (hipMalloc(&p, v.size() * sizeof(T)) != hipSuccess) return false;
    s->allocs.push_back(p);
    if (hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return false;
    *out = (const T *)p;
    return true;
}

static bool upload_primary(dr_scene *s) {
    for (int i = 0; i < 3; i++) {
        if (s->prim_bufs[i]) (void)hipFree(s->prim_bufs[i]);
        s->prim_bufs[i] = nullptr;
    }
    const drh::HostScene &hs = s->hs;
    s->view.num_prim_edges = (int)hs.prim_edges.size();
    s->view.prim_func_int = hs.prim.func_int;
    s->view.prim_edges = nullptr;
    s->view.prim_func = nullptr;
    s->view.prim_cdf = nullptr;
    if (hs.prim_edges.empty()) return true;
    size_t sz[3] = {hs.prim_edges.size() * sizeof(DrEdge), hs.prim.func.size() * sizeof(float),
                    hs.prim.cdf.size() * sizeof(float)};
    const void *src[3] = {hs.prim_edges.data(), hs.prim.func.data(), hs.prim.cdf.data()};
    for (int i = 0; i < 3; i++) {
        if (hipMalloc(&s->prim_bufs[i], sz[i]) != hipSuccess) return false;
        if (hipMemcpy(s->prim_bufs[i], src[i], sz[i], hipMemcpyHostToDevice) != hipSuccess) return false;
    }
    s->view.prim_edges = (const DrEdge *)s->prim_bufs[0];
    s->view.prim_func = (const float *)s->prim_bufs[1];
    s->view.prim_cdf = (const float *)s->prim_bufs[2];
    return true;
}

static hipEvent_t get_event(dr_scene *s) {
    if (!s->event_pool.empty()) {
        hipEvent_t e = s->event_pool.back();
        s->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// record an event before (begin) / after (end) a launch when timing is on
struct LaunchTimer {
    dr_scene *s;
    int kind;
    hipStream_t st;
    hipEvent_t a = nullptr;
    LaunchTimer(dr_scene *s_, int k, hipStream_t st_) : s(s_), kind(k), st(st_) {
        if (s->timing_on) {
            a = get_event(s);
            if (a) (void)hipEventRecord(a, st);
        }
    }
    ~LaunchTimer() {
        if (!a) return;
        hipEvent_t b = get_event(s);
        if (b && hipEventRecord(b, st) == hipSuccess) s->timed.push_back(TimedLaunch{kind, a, b});
        else s->event_pool.push_back(a);
    }
};

static void free_adj(AdjBuffers &a) {
    void *ps[] = {a.pts, a.pt_dp, a.pt_lastj, a.counts, a.q, a.entry_dp, a.counter};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    if (a.counter_host) (void)hipHostFree(a.counter_host);
    a = AdjBuffers();
}

static void free_scene(dr_scene *s) {
    free_adj(s->adj);
    for (auto &t : s->timed) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
    for (auto e : s->event_pool) (void)hipEventDestroy(e);
    for (void *p : s->allocs) (void)hipFree(p);
    for (int i = 0; i < 3; i++)
        if (s->prim_bufs[i]) (void)hipFree(s->prim_bufs[i]);
    if (s->stats_dev) (void)hipFree(s->stats_dev);
    delete s;
}

// samples per thread: keep >= 2^20 threads in flight when the image allows
static void plan_chunks(const dr_render_options *o, int &per, int &chunks) {
    long long pixels = (long long)o->width * o->height;
    chunks = 1;
    while ((long long)chunks * 2 <= o->num_samples && pixels * chunks < (1ll << 20)) chunks *= 2;
    per = (o->num_samples + chunks - 1) / chunks;
    chunks = (o->num_samples + per - 1) / per;
}

static SampleCtx make_ctx(const dr_render_options *o) {
    SampleCtx c;
    c.seed = o->seed;
    c.W = o->width;
    c.H = o->height;
    c.spp = o->num_samples;
    c.nb = o->max_bounces;
    c.flags = o->flags;
    return c;
}

static bool grow(void **p, size_t bytes) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    return hipMalloc(p, bytes) == hipSuccess;
}

static dr_status ensure_slots(AdjBuffers &a, long long n) {
    if (n <= a.slot_cap) return DR_OK;
    if (!grow((void **)&a.pts, n * sizeof(PointRec)) || !grow((void **)&a.pt_dp, n * 3 * sizeof(float)) ||
        !grow((void **)&a.pt_lastj, n * sizeof(int)) || !grow((void **)&a.counts, n * sizeof(int))) {
        a.slot_cap = 0;
        return fail(DR_ERROR_OUT_OF_MEMORY, "adjoint point buffers");
    }
    a.slot_cap = n;
    return DR_OK;
}

static dr_status ensure_queue(AdjBuffers &a, long long n) {
    if (!a.counter) {
        if (hipMalloc((void **)&a.counter, 2 * sizeof(unsigned long long)) != hipSuccess ||
            hipHostMalloc((void **)&a.counter_host, sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess)
            return fail(DR_ERROR_OUT_OF_MEMORY, "adjoint queue counter");
    }
    if (n <= a.q_cap) return DR_OK;
    if (!grow((void **)&a.q, n * sizeof(EdgeEntry)) || !grow((void **)&a.entry_dp, n * sizeof(float4))) {
        a.q_cap = 0;
        return fail(DR_ERROR_OUT_OF_MEMORY, "adjoint edge queue");
    }
    a.q_cap = n;
    return DR_OK;
}

static unsigned grid_of(long long n) { return (unsigned)((n + DR_BLOCK - 1) / DR_BLOCK); }

// samples per batch of the wavefront adjoint (bounds the point and edge
// buffers; one host read of the edge count per batch)
#define DR_BATCH_SAMPLES (1ll << 22)

// Resident workgroups of the persistent selection kernel on this device.
static long long select_blocks() {
    static long long cached = 0;
    if (cached == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_select, DR_BLOCK, 0) != hipSuccess ||
            cus <= 0 || per_cu <= 0) {
            cus = 256;
            per_cu = 2;
        }
        cached = (long long)cus * per_cu;
    }
    return cached;
}

template <int MAXB>
static dr_status launch_adjoint(dr_scene *s, const dr_render_options *o, const float *d_image, float *image,
                                float *grads, float *dx, hipStream_t hs) {
    int per, chunks;
    plan_chunks(o, per, chunks);
    size_t n = 3ull * o->width * o->height;
    if (chunks > 1) {
        k_fill<<<1024, 256, 0, hs>>>(image, n, 0.f);
        if (dx) k_fill<<<1024, 256, 0, hs>>>(dx, n, 0.f);
    }
    long long total = (long long)o->width * o->height * chunks;
    long long bt = std::max(1ll, std::min(total, DR_BATCH_SAMPLES / per));
    SampleCtx c = make_ctx(o);
    AdjBuffers &A = s->adj;
    dr_status st = ensure_slots(A, bt * per * MAXB);
    if (st != DR_OK) return st;
    st = ensure_queue(A, std::max(A.q_cap, bt * per * 4));
    if (st != DR_OK) return st;
    const bool edges = !(o->flags & DR_SKIP_SECONDARY_EDGES);
    for (long long t0 = 0; t0 < total; t0 += bt) {
        long long t1 = std::min(total, t0 + bt);
        long long nslots = (t1 - t0) * per * MAXB;
        {
            LaunchTimer timer(s, TK_PATH, hs);
            k_path<MAXB><<<grid_of(t1 - t0), DR_BLOCK, 0, hs>>>(s->view, c, per, chunks, t0, t1, d_image, image,
                                                                 grads, A.pts);
        }
        if (!edges) continue;
        k_fill<<<1024, 256, 0, hs>>>(A.pt_dp, (size_t)nslots * 3, 0.f);
        k_fill_i<<<1024, 256, 0, hs>>>(A.pt_lastj, (size_t)nslots, -1);
        long long nedges = 0;
        for (int attempt = 0; attempt < 2; attempt++) {
            // queue counter and work counter: two 8-byte zeros
            k_fill<<<1, 64, 0, hs>>>((float *)A.counter, 4, 0.f);
            {
                LaunchTimer timer(s, TK_SELECT, hs);
                long long blocks = std::min<long long>(grid_of(nslots), select_blocks());
                k_select<<<(unsigned)blocks, DR_BLOCK, 0, hs>>>(s->view, c, per, chunks, t0, nslots, MAXB, A.pts,
                                                                A.q, A.counter, A.q_cap, A.counts, A.counter + 1);
            }
            if (hipMemcpyAsync(A.counter_host, A.counter, sizeof(unsigned long long), hipMemcpyDeviceToHost, hs) !=
                    hipSuccess ||
                hipStreamSynchronize(hs) != hipSuccess)
                return fail(DR_ERROR_DEVICE, "reading the edge queue size failed");
            nedges = (long long)*A.counter_host;
            if (nedges <= A.q_cap) break;
            // the selection is deterministic: grow the queue and select again
            st = ensure_queue(A, nedges + nedges / 4);
            if (st != DR_OK) return st;
        }
        if (nedges > 0) {
            LaunchTimer timer(s, TK_EDGE, hs);
            k_edge<MAXB><<<grid_of(nedges), DR_BLOCK, 0, hs>>>(s->view, c, per, chunks, t0, A.q, nedges, A.pts,
                                                                d_image, grads, A.pt_dp, A.pt_lastj, A.entry_dp);
        }
        if (nedges > 0 && (o->flags & DR_QUIRK_DP_LAST))
            k_dp_last<<<grid_of(nedges), DR_BLOCK, 0, hs>>>(A.q, nedges, A.entry_dp, A.pt_lastj, A.pt_dp);
        {
            LaunchTimer timer(s, TK_DP, hs);
            k_dp<MAXB><<<grid_of(t1 - t0), DR_BLOCK, 0, hs>>>(s->view, c, per, chunks, t0, t1, A.pts, A.pt_dp,
                                                               A.counts, grads, dx);
        }
    }
    if (dx && !edges) k_fill<<<1024, 256, 0, hs>>>(dx, n, 0.f);
    if (!(o->flags & DR_SKIP_PRIMARY_EDGES)) {
        LaunchTimer timer(s, TK_PRIMARY, hs);
        long long N = o->num_primary_edge_samples > 0 ? o->num_primary_edge_samples
                                                      : (long long)o->num_samples * o->width * o->height;
        k_primary<MAXB><<<grid_of(N), DR_BLOCK, 0, hs>>>(s->view, c, N, d_image, grads);
    }
    return DR_OK;
}

extern "C" {

const char *dr_last_error(void) { return g_err.c_str(); }

dr_status dr_scene_create(const dr_scene_desc *desc, dr_scene **out) {
    if (!out) return fail(DR_ERROR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    dr_scene *s = new (std::nothrow) dr_scene();
    if (!s) return fail(DR_ERROR_OUT_OF_MEMORY, "host allocation failed");
    std::string err;
    try {
        err = drh::build_host_scene(desc, s->hs);
    } catch (const std::exception &e) {
        err = std::string("scene build failed: ") + e.what();
    }
    if (!err.empty()) {
        delete s;
        return fail(err.find("fisheye") != std::string::npos ? DR_ERROR_UNSUPPORTED : DR_ERROR_INVALID_ARGUMENT, err);
    }
    drh::HostScene &hs = s->hs;
    DrSceneView &v = s->view;
    std::memset(&v, 0, sizeof(v));
    bool ok = upload(s, hs.nodes, &v.nodes) && upload(s, hs.tris, &v.tris) && upload(s, hs.verts, &v.verts) &&
              upload(s, hs.tri_vidx, &v.tri_vidx) && upload(s, hs.tri_shape, &v.tri_shape) &&
              upload(s, hs.uvs, &v.uvs) && upload(s, hs.normals, &v.normals) && upload(s, hs.texels, &v.texels) &&
              upload(s, hs.shapes, &v.shapes) && upload(s, hs.materials, &v.materials) &&
              upload(s, hs.lights, &v.lights) && upload(s, hs.light_func, &v.light_func) &&
              upload(s, hs.light_cdf, &v.light_cdf) && upload(s, hs.light_sel.func, &v.lights_func) &&
              upload(s, hs.light_sel.cdf, &v.lights_cdf) && upload(s, hs.sil_edges, &v.sil_edges) &&
              upload(s, hs.sil_nodes, &v.sil_nodes) && upload(s, hs.nrm_edges, &v.nrm_edges) &&
              upload(s, hs.nrm_nodes, &v.nrm_nodes);
    std::vector<float> ltc(dr_ltc_blob, dr_ltc_blob + 16384 * 9);
    if ((size_t)(dr_ltc_blob_end - dr_ltc_blob) != 16384 * 9) {
        free_scene(s);
        return fail(DR_ERROR_DEVICE, "embedded LTC table has the wrong size");
    }
    ok = ok && upload(s, ltc, &v.ltc) && upload_primary(s);
    if (!ok) {
        free_scene(s);
        return fail(DR_ERROR_OUT_OF_MEMORY, "device allocation or copy failed");
    }
    v.lights_func_int = hs.light_sel.func_int;
    v.light_total_intensity = hs.light_total_intensity;
    v.num_lights = (int)hs.lights.size();
    v.num_shapes = (int)hs.shapes.size();
    v.num_tris = (int)hs.tri_shape.size();
    v.num_nodes = (int)hs.nodes.size();
    v.num_sil_nodes = (int)hs.sil_nodes.size();
    v.num_nrm_nodes = (int)hs.nrm_nodes.size();
    v.cam = hs.cam;
    v.stats = nullptr;
    auto t1 = std::chrono::steady_clock::now();
    s->build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    *out = s;
    return DR_OK;
}

void dr_scene_destroy(dr_scene *s) {
    if (s) free_scene(s);
}

dr_status dr_scene_set_camera(dr_scene *s, const dr_camera *c) {
    if (!s || !c) return fail(DR_ERROR_INVALID_ARGUMENT, "null argument");
    if (c->fisheye) return fail(DR_ERROR_UNSUPPORTED, "fisheye cameras are not supported");
    drh::set_camera(s->hs, *c);
    drh::build_primary_edges(s->hs);
    if (!upload_primary(s)) return fail(DR_ERROR_OUT_OF_MEMORY, "device allocation failed");
    s->view.cam = s->hs.cam;
    return DR_OK;
}

size_t dr_scene_grad_size(const dr_scene *s) { return s ? s->hs.grad_size : 0; }

dr_status dr_scene_get_info(const dr_scene *s, dr_scene_info *o) {
    if (!s || !o) return fail(DR_ERROR_INVALID_ARGUMENT, "null argument");
    o->num_triangles = (int)s->hs.tri_shape.size();
    o->num_vertices = (int)(s->hs.verts.size() / 3);
    o->bvh_nodes = (int)s->hs.nodes.size();
    o->num_edges = (int)s->hs.edges.size();
    o->num_primary_edges = (int)s->hs.prim_edges.size();
    o->silhouette_tree_nodes = (int)s->hs.sil_nodes.size();
    o->normal_tree_nodes = (int)s->hs.nrm_nodes.size();
    o->build_ms = s->build_ms;
    return DR_OK;
}

dr_status dr_scene_stats(dr_scene *s, int enable, dr_stats *out) {
    if (!s) return fail(DR_ERROR_INVALID_ARGUMENT, "null scene");
    if (!s->stats_dev) {
        if (hipMalloc(&s->stats_dev, 8 * sizeof(unsigned long long)) != hipSuccess)
            return fail(DR_ERROR_OUT_OF_MEMORY, "stats allocation failed");
        (void)hipMemset(s->stats_dev, 0, 8 * sizeof(unsigned long long));
    }
    if (hipDeviceSynchronize() != hipSuccess) return fail(DR_ERROR_DEVICE, "device synchronize failed");
    if (out) {
        std::memset(out, 0, sizeof(*out));
        unsigned long long h[8];
        if (hipMemcpy(h, s->stats_dev, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess)
            return fail(DR_ERROR_DEVICE, "stats copy failed");
        out->rays = h[0];
        out->nodes_visited = h[1];
        out->tris_visited = h[2];
        out->secondary_edges = h[3];
        out->primary_edge_samples = h[4];
        out->edge_nodes_visited = h[5];
        out->edges_tested = h[6];
        for (auto &t : s->timed) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, t.a, t.b) != hipSuccess) ms = 0.f;
            if (t.kind < 0 || t.kind >= TK_COUNT) continue;
            out->kernel_ms[t.kind] += ms;
            out->kernel_launches[t.kind]++;
        }
    }
    for (auto &t : s->timed) {
        s->event_pool.push_back(t.a);
        s->event_pool.push_back(t.b);
    }
    s->timed.clear();
    (void)hipMemset(s->stats_dev, 0, 8 * sizeof(unsigned long long));
    s->timing_on = (enable & DR_STATS_TIMING) != 0;
    s->stats_on = (enable & DR_STATS_COUNTERS) != 0;
    s->view.stats = s->stats_on ? s->stats_dev : nullptr;
    return DR_OK;
}

static dr_status check_opts(const dr_scene *s, const dr_render_options *o) {
    if (!s || !o) return fail(DR_ERROR_INVALID_ARGUMENT, "null argument");
    if (o->width <= 0 || o->height <= 0 || o->num_samples <= 0 || o->max_bounces < 1)
        return fail(DR_ERROR_INVALID_ARGUMENT, "width, height, num_samples must be > 0 and max_bounces >= 1");
    if ((long long)o->width * o->height > (1ll << 28)) return fail(DR_ERROR_INVALID_ARGUMENT, "image too large");
    return DR_OK;
}



static dr_status launch_check(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(DR_ERROR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    return DR_OK;
}

dr_status dr_render(dr_scene *s, const dr_render_options *o, float *image, void *stream) {
    dr_status st = check_opts(s, o);
    if (st != DR_OK) return st;
    if (!image) return fail(DR_ERROR_INVALID_ARGUMENT, "image is NULL");
    hipStream_t hs = (hipStream_t)stream;
    int per, chunks;
    plan_chunks(o, per, chunks);
    size_t n = 3ull * o->width * o->height;
    if (chunks > 1) k_fill<<<1024, 256, 0, hs>>>(image, n, 0.f);
    long long total = (long long)o->width * o->height * chunks;
    LaunchTimer timer(s, TK_FORWARD, hs);
    k_forward<<<(unsigned)((total + DR_BLOCK - 1) / DR_BLOCK), DR_BLOCK, 0, hs>>>(s->view, make_ctx(o), per, chunks,
                                                                                image);
    return launch_check("k_forward");
}

dr_status dr_render_adjoint(dr_scene *s, const dr_render_options *o, const float *d_image, float *image, float *grads,
                            float *dx_image, void *stream) {
    dr_status st = check_opts(s, o);
    if (st != DR_OK) return st;
    if (!d_image || !image || !grads) return fail(DR_ERROR_INVALID_ARGUMENT, "d_image, image and grads are required");
    if (o->max_bounces > 8) return fail(DR_ERROR_UNSUPPORTED, "the adjoint supports max_bounces <= 8");
    hipStream_t hs = (hipStream_t)stream;
    k_fill<<<256, 256, 0, hs>>>(grads, s->hs.grad_size, 0.f);
    float *dx = (o->flags & DR_WRITE_DX_IMAGE) ? dx_image : nullptr;
    if (o->max_bounces <= 1) st = launch_adjoint<1>(s, o, d_image, image, grads, dx, hs);
    else if (o->max_bounces <= 2) st = launch_adjoint<2>(s, o, d_image, image, grads, dx, hs);
    else if (o->max_bounces <= 4) st = launch_adjoint<4>(s, o, d_image, image, grads, dx, hs);
    else st = launch_adjoint<8>(s, o, d_image, image, grads, dx, hs);
    if (st != DR_OK) return st;
    return launch_check("adjoint kernels");
}

dr_status dr_intersect(dr_scene *s, const float *rays, int64_t n, int32_t *hits, int any_hit, void *stream) {
    if (!s || (!rays && n > 0) || (!hits && n > 0) || n < 0) return fail(DR_ERROR_INVALID_ARGUMENT, "bad arguments");
    if (n == 0) return DR_OK;
    hipStream_t hs = (hipStream_t)stream;
    k_intersect<<<(unsigned)((n + 255) / 256), 256, 0, hs>>>(s->view, rays, n, hits, any_hit);
    return launch_check("k_intersect");
}

dr_status dr_test_line_integral(int64_t num_samples, uint64_t seed, double *out, void *stream) {
    if (num_samples <= 0 || !out) return fail(DR_ERROR_INVALID_ARGUMENT, "bad arguments");
    hipStream_t hs = (hipStream_t)stream;
    double *dev = nullptr;
    if (hipMalloc(&dev, 8 * sizeof(double)) != hipSuccess) return fail(DR_ERROR_OUT_OF_MEMORY, "hipMalloc");
    k_fill<<<1, 64, 0, hs>>>((float *)dev, 16, 0.f);
    k_line_integral<<<(unsigned)((num_samples + 255) / 256), 256, 0, hs>>>(num_samples, seed, dev);
    double sums[8];
    hipError_t e = hipMemcpyAsync(sums, dev, sizeof(sums), hipMemcpyDeviceToHost, hs);
    if (e == hipSuccess) e = hipStreamSynchronize(hs);
    hipFree(dev);
    if (e != hipSuccess) return fail(DR_ERROR_DEVICE, hipGetErrorString(e));
    for (int k = 0; k < 4; k++) {
        double mean = sums[2 * k] / (double)num_samples;
        out[k] = mean;
        out[4 + k] = sums[2 * k + 1] / (double)num_samples - mean * mean;
    }
    return launch_check("k_line_integral");
}

}  // extern "C"
