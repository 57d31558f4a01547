// Kernels and C-ABI of the MI355X path tracer (include/delta_ray.h).
//
// Kernels (one thread = one pixel sample-chunk or one edge sample; 256-thread
// workgroups = 4 wave64):
//   k_forward   radiance only (reference render_sample_fast,
//               delta_ray.cpp:640-692)
//   k_adjoint   radiance + interior adjoint + secondary edges + Dirac
//               position sweep (reference render_sample, delta_ray.cpp:694-1002)
//   k_primary   primary-edge samples (reference pass 2, delta_ray.cpp:1468-1550)
//   k_intersect ray queries for the parity tests
// Gradients: camera matrices are reduced in LDS per workgroup (64 floats),
// everything else goes through the workgroup's LDS hash (GradCache) so that
// a vertex touched by many lanes of a workgroup reaches HBM as one atomic.
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

template <int MAXB>
__global__ void __launch_bounds__(DR_BLOCK) k_adjoint(DrSceneView S, SampleCtx c, int per, int chunks,
                                                      const float *d_image, float *image, float *grads,
                                                      float *dx_image) {
    __shared__ float s_cam[64];
    if (threadIdx.x < 64) s_cam[threadIdx.x] = 0.f;
    __syncthreads();
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long total = (long long)c.W * c.H * chunks;
    if (t < total) {
        int pix = (int)(t / chunks), ch = (int)(t % chunks);
        int xi = pix % c.W, yi = pix / c.W;
        F3 d_color = load3(d_image + 3 * (size_t)pix);
        GradSink G;
        G.grads = grads;
        G.cam = s_cam;
        G.flags = c.flags;
        G.color_mode = true;
        F3 acc = F3{0.f, 0.f, 0.f};
        float dx = 0.f;
        int s0 = ch * per, s1 = min(s0 + per, c.spp);
        bool zero = is_zero(d_color);
        for (int s = s0; s < s1; s++) {
            if (zero) acc += forward_sample(S, c, xi, yi, s);
            else acc += adjoint_sample<MAXB>(S, c, G, xi, yi, s, d_color, dx);
        }
        float *px = image + 3 * (size_t)pix;
        if (chunks == 1) {
            px[0] = acc.x; px[1] = acc.y; px[2] = acc.z;
            if (dx_image) { float *q = dx_image + 3 * (size_t)pix; q[0] = q[1] = q[2] = dx; }
        } else {
            atomicAdd(px, acc.x); atomicAdd(px + 1, acc.y); atomicAdd(px + 2, acc.z);
            if (dx_image) {
                float *q = dx_image + 3 * (size_t)pix;
                atomicAdd(q, dx); atomicAdd(q + 1, dx); atomicAdd(q + 2, dx);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < 64 && s_cam[threadIdx.x] != 0.f) atomicAdd(grads + threadIdx.x, s_cam[threadIdx.x]);
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
            sample_primary_edge_point(S, e, eu, ls, ss, c.nb, d_image, c.W, c.H, w / c.spp, R);
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

// ------------------------------------------------------------------- host --
extern "C" const float dr_ltc_blob[];
extern "C" const float dr_ltc_blob_end[];

static thread_local std::string g_err;

static dr_status fail(dr_status s, const std::string &msg) {
    g_err = msg;
    return s;
}

// kernel timing: event pairs recorded around launches when enabled
enum { TK_FORWARD = 0, TK_ADJOINT = 1, TK_PRIMARY = 2 };
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

static void free_scene(dr_scene *s) {
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

template <int MAXB>
static void launch_adjoint(dr_scene *s, const dr_render_options *o, const float *d_image, float *image, float *grads,
                           float *dx, hipStream_t hs) {
    int per, chunks;
    plan_chunks(o, per, chunks);
    size_t n = 3ull * o->width * o->height;
    if (chunks > 1) {
        k_fill<<<1024, 256, 0, hs>>>(image, n, 0.f);
        if (dx) k_fill<<<1024, 256, 0, hs>>>(dx, n, 0.f);
    }
    long long total = (long long)o->width * o->height * chunks;
    SampleCtx c = make_ctx(o);
    {
        LaunchTimer timer(s, TK_ADJOINT, hs);
        k_adjoint<MAXB><<<(unsigned)((total + DR_BLOCK - 1) / DR_BLOCK), DR_BLOCK, 0, hs>>>(s->view, c, per, chunks,
                                                                                       d_image, image, grads, dx);
    }
    if (!(o->flags & DR_SKIP_PRIMARY_EDGES)) {
        LaunchTimer timer(s, TK_PRIMARY, hs);
        long long N = o->num_primary_edge_samples > 0 ? o->num_primary_edge_samples
                                                      : (long long)o->num_samples * o->width * o->height;
        k_primary<MAXB><<<(unsigned)((N + DR_BLOCK - 1) / DR_BLOCK), DR_BLOCK, 0, hs>>>(s->view, c, N, d_image,
                                                                                       grads);
    }
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
        for (auto &t : s->timed) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, t.a, t.b) != hipSuccess) ms = 0.f;
            if (t.kind == TK_FORWARD) { out->forward_ms += ms; out->forward_launches++; }
            if (t.kind == TK_ADJOINT) { out->adjoint_ms += ms; out->adjoint_launches++; }
            if (t.kind == TK_PRIMARY) { out->primary_ms += ms; out->primary_launches++; }
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
    if (o->max_bounces <= 1) launch_adjoint<1>(s, o, d_image, image, grads, dx, hs);
    else if (o->max_bounces <= 2) launch_adjoint<2>(s, o, d_image, image, grads, dx, hs);
    else if (o->max_bounces <= 4) launch_adjoint<4>(s, o, d_image, image, grads, dx, hs);
    else launch_adjoint<8>(s, o, d_image, image, grads, dx, hs);
    return launch_check("k_adjoint");
}

dr_status dr_intersect(dr_scene *s, const float *rays, int64_t n, int32_t *hits, int any_hit, void *stream) {
    if (!s || (!rays && n > 0) || (!hits && n > 0) || n < 0) return fail(DR_ERROR_INVALID_ARGUMENT, "bad arguments");
    if (n == 0) return DR_OK;
    hipStream_t hs = (hipStream_t)stream;
    k_intersect<<<(unsigned)((n + 255) / 256), 256, 0, hs>>>(s->view, rays, n, hits, any_hit);
    return launch_check("k_intersect");
}

}  // extern "C"
