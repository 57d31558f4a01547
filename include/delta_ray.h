/*
 * delta_ray (MI355X) — C-ABI of the differentiable path tracer.
 *
 * This is the boundary the Python front end (delta_ray_amd/delta_ray.py,
 * delta_ray_amd/render_pytorch.py) binds with ctypes.  It replaces the
 * reference's pybind11 module:
 *
 *   reference delta_ray.cpp:1838-1920   PYBIND11_MODULE(delta_ray): Material,
 *                                       Shape, Light, Camera, D* classes,
 *                                       RenderResult, render, test_line_integral
 *   reference delta_ray.cpp:1017-1026   render(camera, shapes, materials, lights,
 *                                       resolution, d_image, num_samples,
 *                                       max_bounces, seed, print_progress)
 *
 * The reference rebuilds its Scene (scene.cpp:19-136) inside every render()
 * call.  Here the scene is a handle: dr_scene_create uploads geometry,
 * materials, lights and camera once and builds the triangle BVH, the edge
 * list, the two edge trees and the light distributions; dr_render and
 * dr_render_adjoint then run on the device and may be called many times.
 *
 * Conventions
 *   - Scene description pointers are HOST pointers (the reference takes numpy
 *     buffers, shape.h:11-18, material.h:14-22).  The library copies them.
 *   - Image, d_image, gradient and ray buffers are DEVICE pointers, owned by
 *     the caller.  Work is ordered on the given HIP stream (a hipStream_t passed
 *     as void*; NULL = default stream).  Nothing blocks the host except
 *     dr_scene_create / dr_scene_set_* and the error query.
 *   - Every entry point returns DR_OK or an error code; dr_last_error()
 *     returns a thread-local message for the last failure.  No exception
 *     crosses the boundary.
 *   - Calls on one scene handle must not overlap (they share its scratch).
 */
#ifndef DELTA_RAY_H
#define DELTA_RAY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum dr_status {
    DR_OK = 0,
    DR_ERROR_INVALID_ARGUMENT = 1,
    DR_ERROR_OUT_OF_MEMORY = 2,
    DR_ERROR_DEVICE = 3,
    DR_ERROR_UNSUPPORTED = 4
} dr_status;

/* Camera (reference camera.h:12-21).  Matrices are row-major 4x4.
 * fov_factor = tan(fov/2).  fisheye != 0: the equi-angular projection of
 * camera.h:51-71 (the screen disk's radius maps linearly to the angle from
 * the view axis, pi/2 at the rim; fov and aspect are not used); screen
 * points outside the disk stay black. */
typedef struct dr_camera {
    float cam_to_world[16];
    float world_to_cam[16];
    float sample_to_cam[16];
    float cam_to_sample[16];
    float fov_factor;
    float aspect_ratio;
    float clip_near;
    int fisheye;
} dr_camera;

/* A material channel (reference material.h:14-22): either a constant
 * (height == 0: `channels` floats) or a texture of shape [height, width,
 * channels] sampled bilinearly with wrap-around (material.h:108-239). */
typedef struct dr_texture {
    const float *data;
    int height;
    int width;
    int channels;
} dr_texture;

typedef struct dr_material {
    dr_texture diffuse_reflectance;   /* 3 channels */
    dr_texture specular_reflectance;  /* 3 channels */
    dr_texture roughness;             /* 1 channel  */
    float diffuse_uv_scale[2];
    float specular_uv_scale[2];
    float roughness_uv_scale[2];
    int two_sided;
} dr_material;

/* Triangle mesh (reference shape.h:11-18).  uvs / normals may be NULL. */
typedef struct dr_shape {
    const float *vertices;   /* [num_vertices, 3] */
    int num_vertices;
    const int *indices;      /* [num_triangles, 3] */
    int num_triangles;
    const float *uvs;        /* [num_vertices, 2] or NULL */
    const float *normals;    /* [num_vertices, 3] or NULL */
    int material_id;
    int light_id;            /* index into lights, -1 if not emissive */
} dr_shape;

/* Area light attached to a shape (reference light.h:15-18). */
typedef struct dr_light {
    int shape_id;
    float intensity[3];
} dr_light;

typedef struct dr_scene_desc {
    dr_camera camera;
    const dr_shape *shapes;
    int num_shapes;
    const dr_material *materials;
    int num_materials;
    const dr_light *lights;
    int num_lights;
} dr_scene_desc;

/* Render options.  The reference's render() arguments (delta_ray.cpp:1017-
 * 1026) plus switches.  Quirk bits reproduce reference behaviours that are
 * not the exact derivative of the rendered image; they are ON in the default
 * (parity) mode, DR_FLAGS_PARITY. */
enum {
    /* sub-pixel jitter is U/spp instead of U (delta_ray.cpp:1236-1237) */
    DR_QUIRK_JITTER_OVER_SPP = 1u << 0,
    /* texel adjoints of corners (0,0) and (1,1) land on swapped texels
     * (delta_ray.cpp:1364-1375, 1391-1402, 1417-1424) */
    DR_QUIRK_TEXEL_SWAP = 1u << 1,
    /* shading-normal adjoint of vertex 0 is scattered to all three vertices
     * (delta_ray.cpp:1263-1269) */
    DR_QUIRK_NORMAL_DN0 = 1u << 2,
    /* of several Dirac position derivatives at one shading point, only the
     * last one is propagated (set_adjoint overwrites, autodiff.h:340-342,
     * delta_ray.cpp:977-982) */
    DR_QUIRK_DP_LAST = 1u << 3,
    /* the world_to_cam gradient of primary edges seen by a pinhole camera is
     * twice the derivative: d_project_local (camera.h:204-213) drops the 0.5
     * of project_local (camera.h:170-171); its fisheye branch keeps it */
    DR_QUIRK_PROJECT_SCALE = 1u << 4,
    /* a secondary-edge sample weighs the whole BSDF by the pmf of the
     * component (diffuse / specular) whose LTC density drew it
     * (delta_ray.cpp:86-101, 249, 296-330): twice the derivative on
     * materials with both components; without it a sample counts its own
     * component's part of the BSDF */
    DR_QUIRK_EDGE_BSDF_SUM = 1u << 5,
    /* the Jacobian of a secondary-edge side ray's hit point w.r.t. the ray
     * direction treats the ray as leaving the world origin (the absolute hit
     * position goes to intersect_jacobian, vector.h:690-716, delta_ray.cpp:
     * 329-331): derivatives scaled by (x . n) / ((x - p) . n) for a hit x of
     * normal n seen from p, zero on planes through the origin */
    DR_QUIRK_EDGE_HIT_JACOBIAN = 1u << 6,
    /* the edge tree's cone pass stops at a node where neither child meets
     * the light cone even when its split rule would take both children
     * (edge_tree.cpp:587-589 before 590-601), while the all pass weighs the
     * edges below as if the cone pass reached them: those edges are
     * under-weighted; without it the all pass uses the cone pmf 0 the cone
     * pass really has there */
    DR_QUIRK_CONE_PMF = 1u << 7,
    /* terms (for tests): skip the interior adjoint / secondary edges /
     * primary edges */
    DR_SKIP_INTERIOR = 1u << 8,
    DR_SKIP_SECONDARY_EDGES = 1u << 9,
    DR_SKIP_PRIMARY_EDGES = 1u << 10,
    /* write per-pixel secondary edge counts / 1000 into dx_image
     * (delta_ray.cpp:1251-1254) */
    DR_WRITE_DX_IMAGE = 1u << 11,
    /* secondary edges: mix the plain edge-tree proposal (weight 0.2) with the
     * learned one (node lengths weighted by dr_scene_set_edge_weights); no
     * reference counterpart */
    DR_LEARNED_EDGES = 1u << 12,
    /* accumulate per-edge statistics of the secondary-edge samples
     * (dr_scene_edge_stats) */
    DR_EDGE_STATS = 1u << 13,
    /* the world_to_cam gradient of a primary edge clipped at the near plane
     * differentiates another clip than the one drawn: d_project clips at
     * t = -(z + clip_near) / dz where project clips at -(z - clip_near) / dz,
     * and drops clip_near from dt / d(dz) (camera.h:231-236 vs 259-290);
     * without it the clipped endpoint's derivative is the exact one */
    DR_QUIRK_CLIP_DERIVATIVE = 1u << 14,
    /* a secondary-edge sample counts the whole emission an emitter sends
     * through one side of the edge (delta_ray.cpp:289-295), while the
     * render's estimate of that light is split by MIS between light sampling
     * (an integral over the emitter's area, whose derivative the interior
     * adjoint carries) and BSDF sampling (the only part with a boundary at
     * the emitter's outline): emitter outlines count about twice; without it
     * the emission seen across an emitter's own edge is weighted by the BSDF
     * strategy's MIS weight in that direction (pathtrace.h:139-150) */
    DR_QUIRK_EDGE_EMITTER_FULL = 1u << 15
};
#define DR_FLAGS_PARITY (DR_QUIRK_JITTER_OVER_SPP | DR_QUIRK_TEXEL_SWAP | \
                         DR_QUIRK_NORMAL_DN0 | DR_QUIRK_DP_LAST | DR_QUIRK_PROJECT_SCALE | \
                         DR_QUIRK_EDGE_BSDF_SUM | DR_QUIRK_EDGE_HIT_JACOBIAN | DR_QUIRK_CONE_PMF | \
                         DR_QUIRK_CLIP_DERIVATIVE | DR_QUIRK_EDGE_EMITTER_FULL)

typedef struct dr_render_options {
    int width;
    int height;
    int num_samples;      /* samples per pixel */
    int max_bounces;      /* >= 1 */
    uint64_t seed;
    uint32_t flags;
    /* >0: number of primary-edge samples; 0: num_samples * width * height,
     * as in delta_ray.cpp:1474-1475 */
    int64_t num_primary_edge_samples;
} dr_render_options;

/* Gradient layout.  All parameter gradients of one render live in ONE float
 * buffer; this computes the offsets (in floats).  The order matches the
 * argument order of the reference's RenderFunction.backward
 * (render_pytorch.py:283-333): camera, materials, shapes, lights. */
#define DR_MAX_LAYOUT_ENTRIES 4096
typedef struct dr_grad_layout {
    size_t cam_to_world, world_to_cam, sample_to_cam, cam_to_sample; /* 16 each */
    size_t total;
} dr_grad_layout;

static inline size_t dr_texture_size(const dr_texture *t) {
    return t->height > 0 ? (size_t)t->height * (size_t)t->width * (size_t)t->channels
                         : (size_t)t->channels;
}

/* Offsets of the per-object gradients.  Pass arrays of the scene's sizes (or
 * NULL to skip); normals_off[i] is (size_t)-1 when shape i has no normals. */
static inline size_t dr_grad_layout_compute(const dr_scene_desc *d,
                                            dr_grad_layout *lay,
                                            size_t *diffuse_off,
                                            size_t *specular_off,
                                            size_t *roughness_off,
                                            size_t *vertices_off,
                                            size_t *normals_off,
                                            size_t *intensity_off) {
    size_t o = 0;
    int i;
    if (lay) {
        lay->cam_to_world = 0; lay->world_to_cam = 16;
        lay->sample_to_cam = 32; lay->cam_to_sample = 48;
    }
    o = 64;
    for (i = 0; i < d->num_materials; i++) {
        if (diffuse_off) diffuse_off[i] = o;
        o += dr_texture_size(&d->materials[i].diffuse_reflectance);
        if (specular_off) specular_off[i] = o;
        o += dr_texture_size(&d->materials[i].specular_reflectance);
        if (roughness_off) roughness_off[i] = o;
        o += dr_texture_size(&d->materials[i].roughness);
    }
    for (i = 0; i < d->num_shapes; i++) {
        if (vertices_off) vertices_off[i] = o;
        o += (size_t)d->shapes[i].num_vertices * 3;
        if (d->shapes[i].normals) {
            if (normals_off) normals_off[i] = o;
            o += (size_t)d->shapes[i].num_vertices * 3;
        } else if (normals_off) {
            normals_off[i] = (size_t)-1;
        }
    }
    for (i = 0; i < d->num_lights; i++) {
        if (intensity_off) intensity_off[i] = o;
        o += 3;
    }
    if (lay) lay->total = o;
    return o;
}

typedef struct dr_scene dr_scene;

/* Build a device-resident scene (replaces Scene::Scene, scene.cpp:19-136,
 * and build_double_edge_tree, edge_tree.cpp:263-316). */
dr_status dr_scene_create(const dr_scene_desc *desc, dr_scene **out);
void dr_scene_destroy(dr_scene *scene);

/* Replace the camera (no rebuild of geometry; the primary-edge list and its
 * distribution, which depend on the camera, are recomputed on the device).
 * Synchronises the default stream. */
dr_status dr_scene_set_camera(dr_scene *scene, const dr_camera *camera);

/* New vertex positions (and shading normals) of shapes, from DEVICE buffers:
 * vertices [num_vertices, 3] of the shape's size, normals the same or NULL
 * (keep).  Replaces what the reference does by building a new Scene in every
 * render() call (scene.cpp:19-136, edge_tree.cpp:263-316, delta_ray.cpp:
 * 1027-1030) with a refit on the device: the triangle records and BVH boxes,
 * the edge trees' boxes, normal cones, lengths and geometry records, and the
 * primary-edge distributions of the scene's camera and views are recomputed
 * for the new vertices in the trees' existing topology, with the builder's
 * arithmetic (a refit to the vertices a scene was built from reproduces it).
 * When the new positions change which mesh edges the degenerate / coplanar
 * filter keeps (scene.cpp:80-96), the BVH, edges and edge trees are rebuilt
 * on the host instead (learned edge weights and edge statistics are then
 * cleared), unless the trees hold every mesh edge (dr_scene_set_edge_policy).  Light distributions are recomputed when an emissive shape
 * moves.  Ordered on `stream`; synchronises it (one read of the filter's
 * verdict and of the primary-edge count). */
typedef struct dr_shape_update {
    int shape_id;
    const float *vertices;
    const float *normals;
} dr_shape_update;
dr_status dr_scene_update_vertices(dr_scene *scene, const dr_shape_update *updates, int num_updates,
                                   void *stream);

/* Which mesh edges the edge trees hold (no reference counterpart: the
 * reference rebuilds per call).  DR_EDGES_FILTERED (the default): the edges
 * the degenerate / coplanar filter keeps (scene.cpp:80-96), as the
 * reference's trees; a vertex update that changes a verdict rebuilds on the
 * host.  The filter drops an edge whose faces are coplanar to the last bit
 * (|n0.n1| == 1 in float), and vertices moved by an optimiser step flip that
 * verdict for some nearly flat edge almost every step.  DR_EDGES_ALL_CANDIDATES:
 * the trees hold every mesh edge and the filter's verdicts are per-edge flags
 * recomputed on the device at every vertex update (an edge switched off adds
 * no length to its tree nodes and is never selected; it stays in the node
 * boxes and normal cones), so updates never rebuild.  The estimator is the
 * reference's in expectation; its samples differ from a fresh build's (other
 * trees).  Switching rebuilds the structures on the host (learned edge
 * weights and statistics are cleared). */
#define DR_EDGES_FILTERED 0
#define DR_EDGES_ALL_CANDIDATES 1
dr_status dr_scene_set_edge_policy(dr_scene *scene, int policy);

/* New material parameters (HOST arrays of the same sizes as the scene's) and
 * new light intensities (same lights, same shapes): what the reference reads
 * from its Material / Light objects at every render() call
 * (delta_ray.cpp:1027-1030, scene.cpp:24-52). */
dr_status dr_scene_update_materials(dr_scene *scene, const dr_material *materials, int num_materials);
dr_status dr_scene_update_lights(dr_scene *scene, const dr_light *lights, int num_lights);

/* Number of floats of the gradient buffer for this scene. */
size_t dr_scene_grad_size(const dr_scene *scene);

/* Forward render (reference render() with a 0-d d_image: pass 1 with
 * render_sample_fast, delta_ray.cpp:640-692).  image: device [H, W, 3]. */
dr_status dr_render(dr_scene *scene, const dr_render_options *opts,
                    float *image, void *stream);

/* Forward + adjoint render (reference render() with d_image [H, W, 3]:
 * The secondary-edge stage appends the selected edges of a batch of pixel
 * samples to a queue whose length the host reads (one stream synchronisation
 * per batch of 2^22 samples); a queue that overflows is grown and only the
 * deterministic selection of that batch is repeated.
 * pass 1 with render_sample, delta_ray.cpp:694-1005, then pass 2, the
 * primary-edge samples, delta_ray.cpp:1468-1550).  d_image: device [H,W,3].
 * image: device [H,W,3] (written).  grads: device, dr_scene_grad_size floats
 * (overwritten: the call clears it first).  dx_image: device [H,W,3] or NULL
 * (written when DR_WRITE_DX_IMAGE is set). */
dr_status dr_render_adjoint(dr_scene *scene, const dr_render_options *opts,
                            const float *d_image, float *image, float *grads,
                            float *dx_image, void *stream);

/* The same render as dr_render_adjoint for callers that cannot allocate
 * device memory (tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  It calls no hipMalloc / hipFree and no
 * allocating library routine: every device buffer besides the caller's inputs and outputs lives
 * in a fixed 256 MiB array of the library (one per GPU, held until the
 * process ends), so passes are smaller and a pass whose edge selection
 * exceeds the array's queue (one entry per shading point of a pass) is
 * selected and sampled in parts.  Fails with
 * DR_ERROR_UNSUPPORTED when a render does not fit (more than 4096 passes, or
 * one shading point selecting more edges than the queue holds).  Calls that
 * overlap on one GPU share the array and race: serialise them.  Synchronises
 * the stream (queue lengths are read per pass), so it cannot be captured
 * into a graph. */
dr_status dr_render_adjoint_noalloc(dr_scene *scene, const dr_render_options *opts,
                                    const float *d_image, float *image, float *grads,
                                    float *dx_image, void *stream);

/* Ray queries on the scene BVH (closest hit: tnear 1e-3, tfar inf,
 * scene.cpp:177-202; any hit: the caller's tmin/tmax, scene.cpp:147-175).
 * rays: device [n, 8] = org.xyz, tmin, dir.xyz, tmax.
 * hits: device [n, 4] int32 = shape id (-1 = miss), triangle id, u bits,
 * v bits (float32 bit patterns of intersect.h:71-90's u, v).  For any_hit
 * only hits[i*4] is written: 1 occluded, 0 free. */
dr_status dr_intersect(dr_scene *scene, const float *rays, int64_t n,
                       int32_t *hits, int any_hit, void *stream);

/* Statistics of the renders on this scene since the previous call (host;
 * the call synchronises the device).  enable: DR_STATS_COUNTERS counts rays
 * traced, BVH nodes and triangles visited, secondary edges and primary-edge
 * samples (device atomics: slows the kernels, for measurement runs only);
 * DR_STATS_TIMING records HIP events around every kernel on its launch
 * stream and reports the summed device time per kernel.  The counters and
 * times are cleared by every call. */
enum { DR_STATS_COUNTERS = 1, DR_STATS_TIMING = 2 };
/* kernel_ms / kernel_launches index: forward, path (radiance + interior
 * adjoint), primary edges, edge selection, edge samples, Dirac sweep, edge
 * samples with continuation paths */
enum { DR_K_FORWARD = 0, DR_K_PATH = 1, DR_K_PRIMARY = 2, DR_K_SELECT = 3, DR_K_EDGE = 4, DR_K_DP = 5,
       DR_K_SHADE = 6, DR_K_COUNT = 7 };
typedef struct dr_stats {
    uint64_t rays;
    uint64_t nodes_visited;
    uint64_t tris_visited;
    uint64_t secondary_edges;
    uint64_t primary_edge_samples;
    uint64_t edge_nodes_visited;           /* interior edge-tree nodes expanded */
    uint64_t edges_tested;                 /* leaf edges given the silhouette test */
    uint64_t stack_overflows;              /* traversal stack pushes that found the stack full (0: the
                                            * scene builder bounds tree depths by the stack sizes) */
    double kernel_ms[DR_K_COUNT];          /* summed kernel times */
    uint64_t kernel_launches[DR_K_COUNT];
} dr_stats;
dr_status dr_scene_stats(dr_scene *scene, int enable, dr_stats *out);

/* Replace the scene's triangle BVH (built on the host, binned SAH, up to 4
 * triangles per leaf) with one built on the device: a linear BVH (Karras,
 * HPG 2012: Morton codes of the triangles' box centres, radix sort, binary
 * radix tree, boxes bottom up), one triangle per leaf.  Ray hits are the
 * same (closest t, ties to the smaller triangle id); traversal costs differ.
 * Refits keep it; a host rebuild (the edge set changed) goes back to the
 * host builder.  Structures derived from the BVH are re-uploaded, so call it
 * before setting learned edge weights.  build_ms (may be NULL): the device
 * build's time.  No reference counterpart (the reference builds Embree's BVH
 * on the host, scene.cpp:100-135).  Synchronises the stream. */
dr_status dr_scene_build_bvh_device(dr_scene *scene, double *build_ms, void *stream);

/* Scene sizes (host). */
typedef struct dr_scene_info {
    int num_triangles;
    int num_vertices;
    int bvh_nodes;
    int num_edges;
    int num_primary_edges;
    int silhouette_tree_nodes;
    int normal_tree_nodes;
    double build_ms;
    int refits;              /* dr_scene_update_vertices calls refitted on the device */
    int rebuilds;            /* ... and rebuilt on the host (the edge set changed) */
    int edges_switched_off;  /* DR_EDGES_ALL_CANDIDATES: tree edges the filter drops on the current
                              * vertices (0 otherwise) */
    int bvh_depth;           /* deepest triangle-BVH node (root 0); the builder keeps it within the
                              * traversal stack (64 entries) */
    int edge_tree_depth;     /* deepest node of the silhouette and normal edge trees */
} dr_scene_info;
dr_status dr_scene_get_info(const dr_scene *scene, dr_scene_info *out);

/* ---------------------------------------------------------------- batches
 * Many renders of uneven sizes -- several scenes, several views of a scene,
 * different resolutions and sample counts -- in one stream-ordered launch
 * sequence: every stage (paths, edge selection, edge samples, Dirac sweep,
 * primary edges, forward-only renders) is one kernel launch over all jobs.
 * No reference counterpart: the reference renders one scene per call
 * (delta_ray.cpp:1017-1570). */

/* A scene seen by another camera (its own primary-edge list).  Must not
 * outlive its scene. */
typedef struct dr_view dr_view;
dr_status dr_view_create(dr_scene *scene, const dr_camera *camera, dr_view **out);
void dr_view_destroy(dr_view *view);

typedef struct dr_job {
    dr_scene *scene;
    dr_view *view;               /* NULL: the scene's own camera */
    dr_render_options opts;
    const float *d_image;        /* device [H][W][3]; NULL: forward only */
    float *image;                /* device [H][W][3] */
    float *grads;                /* device, dr_scene_grad_size floats (adjoint jobs); jobs that
                                  * share a buffer add into it (e.g. views of one scene) */
    float *dx_image;             /* device [H][W][3] or NULL (DR_WRITE_DX_IMAGE) */
} dr_job;

typedef struct dr_batch dr_batch;
dr_status dr_batch_create(dr_batch **out);
void dr_batch_destroy(dr_batch *batch);

/* DR_BATCH_GRAPH: capture the launches of this job list once in a hipGraph
 * and replay it on later calls with the same jobs (same scenes, views, sizes,
 * flags and buffers; seeds and cameras may change). */
enum { DR_BATCH_GRAPH = 1 };
/* Renders every job.  Every output buffer is cleared first.  When some job
 * samples secondary edges the call ends with one host synchronisation of the
 * stream (the edge queue's length is checked; an overflowing queue is grown
 * and the jobs rendered again). */
dr_status dr_render_batch(dr_batch *batch, const dr_job *jobs, int num_jobs, int flags, void *stream);

/* Pass sizes of the pipeline: pixel samples per pass and the least number
 * of threads a job is split into; select_refill: 1-64 = the depth-first
 * persistent edge selection, whose waves gather that many idle lanes before
 * starting new shading points (64 by default), 0 = the breadth-first
 * selection (one tree level per launch over every shading point of a pass;
 * the same selected edges).  dr_batch_autotune renders the job list with
 * each candidate (outputs are overwritten) and keeps the fastest. */
typedef struct dr_tuning {
    int64_t pass_samples;
    int64_t min_threads;
    int32_t select_refill;
    double best_ms;              /* out: time of the chosen candidate */
    int candidates;              /* out: candidates timed */
} dr_tuning;
dr_status dr_batch_autotune(dr_batch *batch, const dr_job *jobs, int num_jobs, void *stream, dr_tuning *out);
dr_status dr_batch_set_tuning(dr_batch *batch, const dr_tuning *tuning);

/* ------------------------------------------------ learned edge proposal
 * A small network, 24 per-edge features -> 64 -> 64 -> 1 (ReLU), whose
 * output is the log of a per-edge sampling weight.  No reference
 * counterpart.  Inference on the matrix cores (bf16 weights, fp32
 * accumulation); training on the GPU (fp32 master weights, Adam) against
 * target log-weights with per-edge loss weights (0 = no data).  Parameter
 * layout (fp32): W1 [24][64], b1 [64], W2 [64][64] (in x out), b2 [64],
 * w3 [64], b3. */
#define DR_MLP_FEATURES 24
#define DR_MLP_PARAMS 5825
typedef struct dr_edge_mlp dr_edge_mlp;
dr_status dr_mlp_create(uint64_t seed, dr_edge_mlp **out);
void dr_mlp_destroy(dr_edge_mlp *mlp);
/* get (host, DR_MLP_PARAMS floats) and / or set the fp32 parameters */
dr_status dr_mlp_params(dr_edge_mlp *mlp, float *get, const float *set);
/* logits[i] for features[i][24] (device buffers) */
dr_status dr_mlp_infer(dr_edge_mlp *mlp, const float *features, int64_t n, float *logits, void *stream);
/* steps Adam steps on random minibatches of `batch` rows (device buffers);
 * last_loss (host, may be NULL) receives the last step's weighted mean
 * squared error */
dr_status dr_mlp_train(dr_edge_mlp *mlp, const float *features, const float *targets, const float *weights,
                       int64_t n, int steps, int batch, float lr, uint64_t seed, float *last_loss, void *stream);

/* Edges of the scene's edge trees (silhouette tree, then normal tree): the
 * rows of the features, weights and statistics below. */
int64_t dr_scene_edge_count(const dr_scene *scene);
/* DR_MLP_FEATURES floats per edge (device buffer): screen-space endpoints
 * and visibility for the camera of `view` (NULL: the scene's), adjacent face
 * normals and their cosine, boundary flag, length, midpoint, distance and
 * direction cosines to the nearest light, the edge's leaf length share and
 * normal-cone angle, has-normals flag, 1.  Synchronises the stream. */
dr_status dr_scene_edge_features(dr_scene *scene, const dr_view *view, float *features, void *stream);
/* Per-edge sampling weights of the learned proposal (host, one per edge;
 * NULL: all 1): node lengths become sums of weight x edge length
 * (DR_LEARNED_EDGES renders use them). */
dr_status dr_scene_set_edge_weights(dr_scene *scene, const float *weights);
/* Statistics accumulated by DR_EDGE_STATS renders, three floats per edge
 * (device buffer `out`, may be NULL): over the edge's secondary-edge
 * samples, the sum of |gradient contribution| (weights included; what counts
 * is set by dr_scene_set_stat_shapes),
 * the number of samples and the sum of squared contributions.  reset: clear
 * them afterwards. */
dr_status dr_scene_edge_stats(dr_scene *scene, float *out, int reset, void *stream);
/* The shapes whose gradients the statistics above measure (the parameters
 * an optimisation changes): a sample's contribution is |dv0| + |dv1| of its
 * edge when the edge's shape is listed, plus |dp| (the Dirac derivative at
 * the shading point) when the shading point's shape is.  n = 0: every shape
 * (the default).  No reference counterpart. */
dr_status dr_scene_set_stat_shapes(dr_scene *scene, const int *shape_ids, int n);

/* The reference's test_line_integral (delta_ray.cpp:1572-1836) on the GPU
 * with the device's LTC line sampler: four estimates of one line integral
 * (uniform / LTC-sampled position, cosine / glossy lobe) from num_samples
 * samples each.  out (host, 8 doubles): the four means, then the four
 * (population) variances, in the reference's print order.  Synchronises the
 * stream. */
dr_status dr_test_line_integral(int64_t num_samples, uint64_t seed, double *out, void *stream);

/* Check of the edge-tree traversal stack the selection kernel keeps in LDS
 * (top entries in LDS, older ones spilled to private memory): every GPU lane
 * runs `ops` random pushes and pops on it and on a plain array stack.  out
 * (host, 2 int64): pops that differ (0 expected), deepest stack reached.  No
 * reference counterpart (a self-check of this implementation).  Synchronises
 * the stream. */
dr_status dr_test_tree_stack(int64_t ops, uint64_t seed, int64_t *out, void *stream);

/* ----------------------------------------------------------- scene files
 * Mitsuba ".serialized" meshes (reference load_serialized.cpp:1-277, a
 * pybind11 module returning (indices, vertices, uvs, normals) there).  Reads
 * shape `shape_index` of the file into host arrays that the library
 * allocates; dr_mesh_data_free releases them.  Unflagged attributes are NULL.
 * Colors are read (the file holds them) but no renderer input uses them. */
typedef struct dr_mesh_data {
    int num_vertices, num_triangles;
    float *vertices; /* [num_vertices][3] */
    float *normals;  /* [num_vertices][3] or NULL */
    float *uvs;      /* [num_vertices][2] or NULL */
    float *colors;   /* [num_vertices][3] or NULL */
    int *indices;    /* [num_triangles][3] */
} dr_mesh_data;
dr_status dr_load_serialized(const char *filename, int shape_index, dr_mesh_data *out);
void dr_mesh_data_free(dr_mesh_data *mesh);

/* ------------------------------------------------------------ LTC table
 * Regenerates the table of linearly transformed cosines the secondary-edge
 * sampler reads (the reference's fit/fitLTC.cpp:191-276 with the Blinn-Phong
 * BRDF of fit/brdf_blinn_phong.h:10-66; the table the renderer embeds is the
 * reference's ltc.inc, fitted that way).  tab: n x n entries of 9 floats
 * (a 3x3 matrix row by row, as ltc.inc), entry (a, t) at index a + t * n for roughness index a
 * (r = a / (n - 1)) and view angle index t; amplitude: n x n floats.  Fits
 * rows [row_begin, row_end) at view angles [t_begin, t_end); a fit that does
 * not start at (n - 1, 0) takes its first guesses from `guess` (a full
 * table: the entry before each one), else guess may be NULL.  flags:
 * DR_LTC_FIT_CONSISTENT fits the lobe the renderer evaluates at roughness r
 * (Phong exponent 2 / r) instead of the reference's alpha = r^2 (exponent
 * 2 / r^4).  nthreads <= 0: all host threads.  Host only; about a minute
 * for the full 128 x 128 table on 16 threads. */
#define DR_LTC_FIT_CONSISTENT 1
dr_status dr_ltc_fit(int n, int row_begin, int row_end, int t_begin, int t_end, const float *guess, int flags,
                     int nthreads, float *tab, float *amplitude);

const char *dr_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
