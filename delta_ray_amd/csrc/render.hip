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
#include "lbvh.h"
#include "refit.h"
#include "scene_build.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
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

// ------------------------------------------------------- jobs, ranges ----
// A job is one render: a scene seen by one camera with its options and its
// buffers.  A launch covers a list of ranges, each a contiguous piece of one
// job's threads; ranges are ordered and carry prefix offsets of their
// workgroups and point slots, so any mix of jobs (scenes, views, sample
// counts of uneven sizes) runs as one launch per stage.  Workgroups never
// straddle two ranges; a workgroup finds its range by binary search.
struct DrJob {
    DrSceneView S;          // the scene, with this job's camera and primary edges
    SampleCtx c;
    int per, chunks;        // samples per thread, threads per pixel
    long long prim_n;       // primary-edge samples
    const float *d_image;   // NULL: forward only
    float *image, *grads, *dx;
};
struct DrRange {
    int job, pad;
    long long t0, t1;       // (pixel, chunk) threads of the job, or primary-edge samples
    long long slot0;        // first point slot of the range in its pass
    long long block0;       // first workgroup of the range in its launch
};
DR_INLINE int range_of_block(const DrRange *r, int n, long long b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (r[mid].block0 <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

#define DR_GC_SLOTS_BIG 4096
#define DR_GC_SLOTS_SMALL 2048
// k_edge: 1024 slots (12 KB) + the ray stack (16 KB) leave room for five
// workgroups per CU, which its 96-VGPR budget also allows (5 % faster than
// 2048 slots at four)
#define DR_GC_SLOTS_EDGE 1024

// thread -> (pixel, chunk of `per` samples) of a forward-only job
__global__ void __launch_bounds__(DR_BLOCK) k_forward(const DrJob *__restrict__ jobs, const DrRange *__restrict__ R,
                                                      int nr) {
    const DrRange rg = R[range_of_block(R, nr, blockIdx.x)];
    const DrJob &J = jobs[rg.job];
    long long t = rg.t0 + (long long)(blockIdx.x - rg.block0) * blockDim.x + threadIdx.x;
    if (t >= rg.t1) return;
    const SampleCtx &c = J.c;
    int pix = (int)(t / J.chunks), ch = (int)(t % J.chunks);
    int xi = pix % c.W, yi = pix / c.W;
    F3 acc = F3{0.f, 0.f, 0.f};
    int s0 = ch * J.per, s1 = min(s0 + J.per, c.spp);
    for (int s = s0; s < s1; s++) acc += forward_sample(J.S, c, xi, yi, s);
    float *px = J.image + 3 * (size_t)pix;
    if (J.chunks == 1) {
        px[0] = acc.x; px[1] = acc.y; px[2] = acc.z;
    } else {
        atomicAdd(px, acc.x); atomicAdd(px + 1, acc.y); atomicAdd(px + 2, acc.z);
    }
}

// ---- wavefront adjoint: path (A) -> edge selection (B) -> edge samples (C)
// -> Dirac position sweep (D), over the ranges of one pass.  A thread of A
// and D is one (pixel, chunk) of a job; its point slots are
// slot0 + ((t - t0) * per + k) * MAXB + b.
struct EdgeEntry {
    int slot;  // point slot in the pass
    int edge;  // index into the tree's edge list; bit 31 set: normal tree
    float w;
    int j;     // ordinal of the edge in its pass (bit 30: the "all" pass)
};


template <int MAXB>
__global__ void __launch_bounds__(DR_BLOCK) k_path(const DrJob *__restrict__ jobs, const DrRange *__restrict__ R,
                                                   int nr, PointRec *pts) {
    __shared__ float s_cam[64];
    __shared__ unsigned long long gkey[DR_GC_SLOTS_BIG];
    __shared__ float gval[DR_GC_SLOTS_BIG];
    if (threadIdx.x < 64) s_cam[threadIdx.x] = 0.f;
    gc_init(gkey, gval, DR_GC_SLOTS_BIG);
    __syncthreads();
    const DrRange rg = R[range_of_block(R, nr, blockIdx.x)];
    const DrJob &J = jobs[rg.job];
    long long t = rg.t0 + (long long)(blockIdx.x - rg.block0) * blockDim.x + threadIdx.x;
    if (t < rg.t1) {
        const SampleCtx &c = J.c;
        const int per = J.per, chunks = J.chunks;
        int pix = (int)(t / chunks), ch = (int)(t % chunks);
        int xi = pix % c.W, yi = pix / c.W;
        F3 d_color = load3(J.d_image + 3 * (size_t)pix);
        GradSink G;
        G.grads = J.grads;
        G.cam = s_cam;
        G.flags = c.flags;
        G.color_mode = true;
        G.gkey = gkey;
        G.gval = gval;
        G.gmask = DR_GC_SLOTS_BIG - 1;
        G.job = (unsigned)rg.job;
        F3 acc = F3{0.f, 0.f, 0.f};
        bool zero = is_zero(d_color);
        for (int k = 0; k < per; k++) {
            int s = ch * per + k;
            PointRec *P = pts + rg.slot0 + ((t - rg.t0) * per + k) * MAXB;
            if (s >= c.spp || zero) {
                for (int b = 0; b < MAXB; b++) P[b].valid = 0;
                if (s < c.spp) acc += forward_sample(J.S, c, xi, yi, s);
                continue;
            }
            acc += stage_path<MAXB>(J.S, c, G, xi, yi, s, d_color, P, rg.job);
        }
        float *px = J.image + 3 * (size_t)pix;
        if (chunks == 1) {
            px[0] = acc.x; px[1] = acc.y; px[2] = acc.z;
        } else {
            atomicAdd(px, acc.x); atomicAdd(px + 1, acc.y); atomicAdd(px + 2, acc.z);
        }
    }
    __syncthreads();
    if (threadIdx.x < 64 && s_cam[threadIdx.x] != 0.f) atomicAdd(J.grads + threadIdx.x, s_cam[threadIdx.x]);
    gc_flush(gkey, gval, DR_GC_SLOTS_BIG, [&](int j) { return jobs[j].grads; });
}

// B: edge selection, persistent.  Per-point traversal lengths vary by two
// orders of magnitude (a point inside many node boxes splits at every level),
// so a lane that has finished its point takes the next one, and every
// iteration advances each busy lane's traversal by one node.  Starting a point
// (decoding its slot, the light cone, the material, the root bound) is a long
// stretch of code that the whole wave executes whenever any lane does, so
// idle lanes wait until at least `refill` of them are idle (or no lane is
// busy) and then start new points together.  In a refill, idle lanes take
// batches of point slots from `work` with one atomic per batch until each has
// a valid point or the slots run out, so empty slots cost no iteration.
// Selected edges are appended to the queue in groups of up to DR_EMIT_GROUP
// per lane (one queue atomic per group; a point's edges flush together at its
// end, so they sit side by side in the queue, which keeps k_edge's waves on
// rays from one point: 32 rather than 8 makes k_edge 6 % faster).  Every
// lane leaves the loop once `work` passes nslots.  cnt[0]: queue length (may
// exceed cap: the host grows the queue and renders again), cnt[1]: work
// counter.
#define DR_EMIT_GROUP 32
// Points are fetched in the order f in [f_begin, f_end) (a pass's points are
// f in [0, nslots); a part of them when a pass's selection runs in parts).
// Four waves per SIMD: the register budget of 128 VGPRs fits the kernel
// without spills and matches what its LDS stack leaves room for; five (96
// VGPRs) spill and are 40 % slower.
__global__ void __launch_bounds__(DR_BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
k_select(const DrJob *__restrict__ jobs, long long nslots, int maxb, const PointRec *pts, EdgeEntry *q,
         unsigned long long *cnt, long long cap, int *counts, int refill, long long f_begin, long long f_end) {
    const long long per_bounce = nslots / maxb;  // nslots is a multiple of maxb
    __shared__ float s_stack[DR_LDS_STACK * DR_LDS_WORDS * DR_BLOCK];
    const int lane = (int)__lane_id();
    EdgeTraversal T;
    LdsStack stack;
    stack.lds = (DR_LDS float *)s_stack + threadIdx.x;
    stack.stride = DR_BLOCK;
    stack.clear();
    SPoint sp;
    const DrSceneView *Sp = &jobs[0].S;
    long long slot = -1;
    int j = 0;
    bool more = true;
    EdgeEntry buf[DR_EMIT_GROUP];
    int nbuf = 0;
    auto flush = [&]() {
        unsigned long long base = atomicAdd(&cnt[0], (unsigned long long)nbuf);
        for (int e = 0; e < nbuf; e++)
            if ((long long)(base + e) < cap) q[base + e] = buf[e];
        nbuf = 0;
    };
    auto emit = [&](float w, bool nrm, int idx, int ord) {
        // zero-weight edges carry no work (reference delta_ray.cpp:906-908)
        if (w != 0.f) {
            buf[nbuf++] = EdgeEntry{(int)slot, idx | (nrm ? (int)0x80000000 : 0), w, ord};
            if (nbuf == DR_EMIT_GROUP) flush();
        }
        j++;
    };
    while (true) {
        unsigned long long idle = __ballot(slot < 0 && more);
        if (idle && (__popcll(idle) >= refill || __ballot(slot >= 0) == 0)) {
            // fetch: until every idle lane holds a valid point or the slots run out
            bool want = slot < 0 && more;
            long long cand = -1;
            while (true) {
                unsigned long long m = __ballot(want);
                if (!m) break;
                int leader = __ffsll((long long)m) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(&cnt[1], (unsigned long long)__popcll(m));
                base = __shfl(base, leader);
                if (want) {
                    unsigned long long below = lane == 0 ? 0ull : (m & (~0ull >> (64 - lane)));
                    long long f = f_begin + (long long)(base + __popcll(below));
                    // points in the order bounce-major: a wave's points are the
                    // same bounce of neighbouring samples and pixels, which
                    // walk similar tree paths (slots are bounce-minor)
                    long long s = f < f_end ? (f % per_bounce) * maxb + f / per_bounce : nslots;
                    if (s >= nslots) {
                        more = false;
                        want = false;
                    } else if (pts[s].valid) {  // (only jobs with secondary edges mark points valid)
                        cand = s;
                        want = false;
                    } else {
                        counts[s] = 0;
                    }
                }
            }
            // start the new points together
            if (cand >= 0) {
                const PointRec &p = pts[cand];
                const DrJob &J = jobs[p.job];
                Sp = &J.S;
                if (select_begin(*Sp, J.c, p, T, stack)) {
                    slot = cand;
                    sp = p.sp;
                    j = 0;
                } else {
                    counts[cand] = 0;
                }
            }
        }
        if (slot < 0) {
            if (!more) break;
            continue;
        }
        if (!et_step(T, stack, *Sp, sp, emit)) {
            if (nbuf > 0) flush();
            counts[slot] = j;
            if (Sp->stats) atomicAdd(&Sp->stats[3], (unsigned long long)j);
            slot = -1;
        }
    }
}

// B, breadth first: the same traversal (et_expand / et_leaf, the same keyed
// draws and ordinals, so the same selected edges with the same weights) run
// one tree level per launch over every point of the pass.  A lane takes one
// frontier entry (a node some point's passes reached) and expands it, so all
// lanes of a wave run the same code on full waves, where the per-point walk
// leaves most lanes of a wave idle while the longest walk of its 64 points
// finishes.  k_sel_start writes each point's traversal constants (SelPoint)
// and its root entry; k_sel_level expands level d into level d + 1 (interior
// children) and the leaf list; k_sel_leaves tests and emits the leaves'
// edges.  Entries are appended with one atomic per wave; an append beyond a
// buffer's capacity is dropped and the largest end is kept in cnt[4] (levels)
// or cnt[5] (leaves), so the host grows the buffer and selects again.
struct alignas(16) SelPoint {
    F3 pos, n;
    BoundCtx B;
    F3 cdir;
    float ccos;
    uint32_t sid_lo, sid_hi;
    int info;  // bit 0: normal tree; bits 1-2: proposal mode; bit 3: DR_QUIRK_CONE_PMF; bits 8-15: bounce
    int job;
};
static_assert(sizeof(SelPoint) == 128, "SelPoint is 8 x 16 bytes");
struct alignas(16) SelEntry {
    int slot, code;
    uint32_t cpmf, apmf;
    int off;
    float altc, alta;
    int pad;
};
static_assert(sizeof(SelEntry) == 32, "SelEntry is 2 x 16 bytes");
#define DR_SEL_LEAVES 64  // index of the leaf count among the level counters

DR_INLINE SelEntry sel_entry(int slot, const TreeEntry &e, TreeAlt a) {
    return SelEntry{slot, e.code, e.cpmf, e.apmf, e.off, a.cone, a.all, 0};
}
DR_INLINE TreeEntry sel_tree(const SelEntry &x) { return TreeEntry{x.code, x.cpmf, x.apmf, x.off}; }

// the point's traversal state rebuilt from its SelPoint (what et_expand and
// et_leaf read: tree, bound constants, light cone, proposal, keyed draws;
// of the surface point only its position and normal)
DR_INLINE void sel_ctx(const SelPoint &P, const DrJob &J, EdgeTraversal &T, SPoint &sp) {
    const DrSceneView &S = J.S;
    bool nrm = (P.info & 1) != 0;
    T.nodes = nrm ? S.nrm_nodes : S.sil_nodes;
    T.edges = nrm ? S.nrm_edges : S.sil_edges;
    T.num_nodes = nrm ? S.num_nrm_nodes : S.num_sil_nodes;
    T.directional = nrm;
    T.B = P.B;
    T.cone = Cone{P.pos, P.cdir, P.ccos};
    T.mode = (P.info >> 1) & 3;
    T.cone_pmf_quirk = (P.info & 8) != 0;
    T.pops = 0;
    T.n_inner = 0;
    T.n_edges = 0;
    uint64_t sid = (uint64_t)P.sid_lo | ((uint64_t)P.sid_hi << 32);
    philox_init(T.rt, J.c.seed, 0, sid, dr_stream_point_tree((P.info >> 8) & 0xff));
    sp.pos = P.pos;
    sp.ng = P.n;
    sp.fr = Frame{F3{0.f, 0.f, 0.f}, F3{0.f, 0.f, 0.f}, P.n};
    sp.uv = F2{0.f, 0.f};
}

// Appends k (0-2) items per lane with one atomic per wave; every lane of the
// wave must call it.  Returns the lane's first index; *end_max keeps the
// largest end of an append.
DR_INLINE long long wave_append(unsigned long long *counter, int k, unsigned long long *end_max) {
    const int lane = (int)__lane_id();
    unsigned long long m1 = __ballot(k >= 1), m2 = __ballot(k >= 2);
    unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    int before = __popcll(m1 & lt) + __popcll(m2 & lt);
    int total = __popcll(m1) + __popcll(m2);
    if (total == 0) return 0;
    int leader = __ffsll((long long)m1) - 1;
    unsigned long long base = 0;
    if (lane == leader) {
        base = atomicAdd(counter, (unsigned long long)total);
        atomicMax(end_max, base + (unsigned long long)total);
    }
    base = __shfl(base, leader);
    return (long long)base + before;
}

// a tree root to the frontier or to the leaf list
struct RootStack {
    TreeEntry e;
    TreeAlt a;
    int sp;
    DR_INLINE void clear() { sp = 0; }
    DR_INLINE int size() const { return sp; }
    DR_INLINE void push(const TreeEntry &x, TreeAlt al, bool) {
        e = x;
        a = al;
        sp = 1;
    }
};

__global__ void __launch_bounds__(DR_BLOCK) k_sel_start(const DrJob *__restrict__ jobs, long long nslots, int maxb,
                                                        const PointRec *pts, SelPoint *spt, SelEntry *front,
                                                        SelEntry *leaves, unsigned long long *lvl, long long fcap,
                                                        long long lcap, unsigned long long *cnt, int *counts) {
    const long long per_bounce = nslots / maxb;
    const int lane = (int)__lane_id();
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long f0 = (long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63); f0 < nslots; f0 += stride) {
        long long f = f0 + lane;
        int kf = 0, kl = 0;
        SelEntry x{};
        if (f < nslots) {
            // bounce-major, as the depth-first selection: neighbouring entries
            // are the same bounce of neighbouring samples
            long long s = (f % per_bounce) * maxb + f / per_bounce;
            counts[s] = 0;
            const PointRec &p = pts[s];
            if (p.valid) {
                const DrJob &J = jobs[p.job];
                EdgeTraversal T;
                RootStack st;
                st.sp = 0;
                if (select_begin(J.S, J.c, p, T, st) && st.sp == 1 && !st.e.empty()) {
                    SelPoint P;
                    P.pos = p.sp.pos;
                    P.n = p.sp.fr.n;
                    P.B = T.B;
                    P.cdir = p.cone_dir;
                    P.ccos = p.cone_cos;
                    P.sid_lo = (uint32_t)p.sid;
                    P.sid_hi = (uint32_t)(p.sid >> 32);
                    P.info = (T.directional ? 1 : 0) | (T.mode << 1) | (T.cone_pmf_quirk ? 8 : 0) | (p.b << 8);
                    P.job = p.job;
                    spt[s] = P;
                    x = sel_entry((int)s, st.e, st.a);
                    if (st.e.nprims() > 0) kl = 1;
                    else kf = 1;
                }
            }
        }
        long long bf = wave_append(lvl, kf, &cnt[4]);
        long long bl = wave_append(lvl + DR_SEL_LEAVES, kl, &cnt[5]);
        if (kf && bf < fcap) front[bf] = x;
        if (kl && bl < lcap) leaves[bl] = x;
    }
}

__global__ void __launch_bounds__(DR_BLOCK) k_sel_level(const DrJob *__restrict__ jobs, const SelPoint *spt,
                                                        const SelEntry *in, const unsigned long long *in_cnt,
                                                        SelEntry *out, unsigned long long *out_cnt, SelEntry *leaves,
                                                        unsigned long long *leaf_cnt, long long fcap, long long lcap,
                                                        unsigned long long *cnt) {
    const long long n = std::min<long long>((long long)*in_cnt, fcap);
    const int lane = (int)__lane_id();
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long b0 = (long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63); b0 < n; b0 += stride) {
        long long i = b0 + lane;
        SelEntry x0{}, x1{};
        bool v0 = false, v1 = false, l0 = false, l1 = false;
        if (i < n) {
            const SelEntry E = in[i];
            const TreeEntry te = sel_tree(E);
            const SelPoint P = spt[E.slot];
            const DrJob &J = jobs[P.job];
            EdgeTraversal T;
            SPoint sp;
            sel_ctx(P, J, T, sp);
            TreeEntry c0, c1;
            TreeAlt a0{0.f, 0.f}, a1{0.f, 0.f};
            int k = et_expand(T, te, TreeAlt{E.altc, E.alta}, J.S, sp, c0, a0, c1, a1);
            if (J.S.stats) atomicAdd(&J.S.stats[5], 1ull);
            // children whose bound is 0 take nothing (the depth-first walk
            // pops and drops them)
            if (k > 0 && !c0.empty()) {
                x0 = sel_entry(E.slot, c0, a0);
                v0 = true;
                l0 = c0.nprims() > 0;
            }
            if (k > 1 && !c1.empty()) {
                x1 = sel_entry(E.slot, c1, a1);
                v1 = true;
                l1 = c1.nprims() > 0;
            }
        }
        int kf = (v0 && !l0) + (v1 && !l1), kl = (v0 && l0) + (v1 && l1);
        long long bf = wave_append(out_cnt, kf, &cnt[4]);
        long long bl = wave_append(leaf_cnt, kl, &cnt[5]);
        if (v0) {
            if (l0) {
                if (bl < lcap) leaves[bl] = x0;
                bl++;
            } else {
                if (bf < fcap) out[bf] = x0;
                bf++;
            }
        }
        if (v1) {
            if (l1) {
                if (bl < lcap) leaves[bl] = x1;
            } else {
                if (bf < fcap) out[bf] = x1;
            }
        }
    }
}

// One leaf per lane; its edges in turn (as many rounds as the wave's largest
// leaf), each round's selections appended with one atomic per wave.
__global__ void __launch_bounds__(DR_BLOCK) k_sel_leaves(const DrJob *__restrict__ jobs, const SelPoint *spt,
                                                         const SelEntry *leaves, const unsigned long long *leaf_cnt,
                                                         long long lcap, EdgeEntry *q, unsigned long long *cnt,
                                                         long long cap, int *counts) {
    const long long n = std::min<long long>((long long)*leaf_cnt, lcap);
    const int lane = (int)__lane_id();
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long b0 = (long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63); b0 < n; b0 += stride) {
        long long i = b0 + lane;
        SelEntry E{};
        TreeEntry te{0, 0u, 0x80000000u, 0};
        int nprims = 0;
        if (i < n) {
            E = leaves[i];
            te = sel_tree(E);
            nprims = te.nprims();
        }
        int rounds = nprims;
        for (int o = 32; o > 0; o >>= 1) rounds = max(rounds, __shfl_xor(rounds, o));
        SelPoint P;
        EdgeTraversal T;
        SPoint sp;
        const DrJob *J = nullptr;
        if (nprims > 0) {
            P = spt[E.slot];
            J = &jobs[P.job];
            sel_ctx(P, *J, T, sp);
        }
        int calls = 0;
        for (int r = 0; r < rounds; r++) {
            EdgeEntry y0{}, y1{};
            int ke = 0;
            if (r < nprims) {
                int got = 0;
                auto emit = [&](float w, bool nrm, int idx, int ord) {
                    calls++;
                    if (w != 0.f) {
                        EdgeEntry y{E.slot, idx | (nrm ? (int)0x80000000 : 0), w, ord};
                        if (got == 0) y0 = y;
                        else y1 = y;
                        got++;
                    }
                };
                et_leaf_edge(T, te, TreeAlt{E.altc, E.alta}, J->S, sp, r, emit);
                ke = got;
            }
            long long bq = wave_append(&cnt[0], ke, &cnt[2]);
            if (ke > 0 && bq < cap) q[bq] = y0;
            if (ke > 1 && bq + 1 < cap) q[bq + 1] = y1;
        }
        if (nprims > 0 && calls > 0) {
            if (J->dx) atomicAdd(&counts[E.slot], calls);
            if (J->S.stats) {
                atomicAdd(&J->S.stats[3], (unsigned long long)calls);
            }
        }
        if (nprims > 0 && J->S.stats) atomicAdd(&J->S.stats[6], (unsigned long long)nprims);
    }
}

// C: one selected edge per iteration of a grid-stride loop over the queue
// (its length is read on the device, so no host round trip is needed).  Only
// about one edge sample in five passes the connectivity test of its two side
// rays (the rest stop at the horizon, at that test or at a zero BSDF), and
// what follows (evaluation at both hits, continuation paths, the Dirac
// Jacobians, the scatter) is the longest code of the kernel: run in the same
// waves it would leave most lanes idle.  So k_edge stops such samples there
// (returns true: nothing of theirs is written) and appends them to a list;
// k_edge_shade then runs them in full, on full waves.  Every sample is thus
// computed by exactly one launch, with the same operations as in one pass.
template <int MAXB>
DR_INLINE bool edge_entry(const DrJob *jobs, const EdgeEntry &E, long long e,
                          const PointRec *pts, float *pt_dp, int *pt_lastj, float4 *entry_dp,
                          unsigned long long *gkey, float *gval, int gslots, bool defer) {
    const PointRec &p = pts[E.slot];
    const DrJob &J = jobs[p.job];
    int idx = E.edge & 0x7fffffff;
    const DrEdge &edge = (E.edge & 0x80000000) ? J.S.nrm_edges[idx] : J.S.sil_edges[idx];
    GradSink G;
    G.grads = J.grads;
    G.cam = nullptr;
    G.flags = J.c.flags;
    G.color_mode = true;
    G.gkey = gkey;
    G.gval = gval;
    G.gmask = gslots - 1;
    G.job = (unsigned)p.job;
    F3 d_color = load3(J.d_image + 3 * (size_t)p.pix);
    float mag = 0.f;
    bool deferred = false;
    // the edge was selected at this point, so it is a silhouette there: its
    // endpoints come from the tree's geometry record
    const float *geo = ((E.edge & 0x80000000) ? J.S.nrm_geo : J.S.sil_geo) + 16 * (size_t)idx;
    F3 dp = stage_edge<MAXB>(J.S, J.c, G, p.sid, p.b, p, edge, E.w, E.j, d_color, &mag,
                             defer ? &deferred : nullptr, geo);
    if (deferred) return true;
    if ((J.c.flags & DR_EDGE_STATS) && J.S.edge_stat) {
        long long ge = (E.edge & 0x80000000) ? (long long)J.S.num_sil_edges + idx : (long long)idx;
        if (mag != 0.f) {
            atomicAdd(J.S.edge_stat + 3 * ge, mag);
            atomicAdd(J.S.edge_stat + 3 * ge + 2, mag * mag);
        }
        atomicAdd(J.S.edge_stat + 3 * ge + 1, 1.f);
    }
    bool nz = !is_zero(dp);
    if (J.c.flags & DR_QUIRK_DP_LAST) {
        entry_dp[e] = make_float4(dp.x, dp.y, dp.z, nz ? __int_as_float(E.j) : __int_as_float(-1));
        if (nz) atomicMax(pt_lastj + E.slot, E.j);
    } else {
        // a pass may mix jobs with and without the last-dp rule: mark the
        // entry as not taking part in it
        entry_dp[e] = make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
        if (nz) {
            atomicAdd(pt_dp + 3ll * E.slot, dp.x);
            atomicAdd(pt_dp + 3ll * E.slot + 1, dp.y);
            atomicAdd(pt_dp + 3ll * E.slot + 2, dp.z);
        }
    }
    return false;
}

// cnt[2] keeps the largest queue length seen, for the host's overflow check;
// cnt[3] counts the samples left for k_edge_shade (listed in `cont`, which
// has the queue's capacity).
template <int MAXB>
__global__ void __launch_bounds__(DR_BLOCK) __attribute__((amdgpu_waves_per_eu(5)))
k_edge(const DrJob *__restrict__ jobs, const EdgeEntry *q, unsigned long long *cnt, long long cap,
       const PointRec *pts, float *pt_dp, int *pt_lastj, float4 *entry_dp, int *cont) {
    __shared__ unsigned long long gkey[DR_GC_SLOTS_EDGE];
    __shared__ float gval[DR_GC_SLOTS_EDGE];
    gc_init(gkey, gval, DR_GC_SLOTS_EDGE);
    __syncthreads();
    unsigned long long len = cnt[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicMax(&cnt[2], len);
    long long n = (long long)len < cap ? (long long)len : cap;
    long long stride = (long long)gridDim.x * blockDim.x;
    const int lane = (int)__lane_id();
    for (long long e0 = (long long)blockIdx.x * blockDim.x + (threadIdx.x - lane); e0 < n; e0 += stride) {
        // whole waves iterate together (e0: the wave's first entry), so the
        // list append below is one atomic per wave
        long long e = e0 + lane;
        bool later = e < n && edge_entry<MAXB>(jobs, q[e], e, pts, pt_dp, pt_lastj, entry_dp, gkey, gval,
                                               DR_GC_SLOTS_EDGE, true);
        unsigned long long m = __ballot(later);
        if (m) {
            int leader = __ffsll((long long)m) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(&cnt[3], (unsigned long long)__popcll(m));
            base = __shfl(base, leader);
            if (later) cont[base + __popcll(m & ((1ull << lane) - 1ull))] = (int)e;
        }
    }
    __syncthreads();
    gc_flush(gkey, gval, DR_GC_SLOTS_EDGE, [&](int j) { return jobs[j].grads; });
}

// C': the samples k_edge left for the continuation paths, in full.
template <int MAXB>
__global__ void __launch_bounds__(DR_BLOCK) __attribute__((amdgpu_waves_per_eu(2)))
k_edge_shade(const DrJob *__restrict__ jobs, const EdgeEntry *q,
                                                         const unsigned long long *cnt, const PointRec *pts,
                                                         float *pt_dp, int *pt_lastj, float4 *entry_dp,
                                                         const int *cont) {
    __shared__ unsigned long long gkey[DR_GC_SLOTS_BIG];
    __shared__ float gval[DR_GC_SLOTS_BIG];
    gc_init(gkey, gval, DR_GC_SLOTS_BIG);
    __syncthreads();
    long long n = (long long)cnt[3];
    long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        long long e = cont[i];
        edge_entry<MAXB>(jobs, q[e], e, pts, pt_dp, pt_lastj, entry_dp, gkey, gval, DR_GC_SLOTS_BIG, false);
    }
    __syncthreads();
    gc_flush(gkey, gval, DR_GC_SLOTS_BIG, [&](int j) { return jobs[j].grads; });
}

// reference set_adjoint overwrite: of several dp at one point the last
// (largest ordinal) is the one propagated
__global__ void k_dp_last(const EdgeEntry *q, const unsigned long long *cnt, long long cap, const float4 *entry_dp,
                          const int *pt_lastj, float *pt_dp) {
    long long n = (long long)cnt[0] < cap ? (long long)cnt[0] : cap;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
        float4 v = entry_dp[e];
        int j = __float_as_int(v.w);
        int slot = q[e].slot;
        if (j >= 0 && j == pt_lastj[slot]) {
            pt_dp[3ll * slot] = v.x;
            pt_dp[3ll * slot + 1] = v.y;
            pt_dp[3ll * slot + 2] = v.z;
        }
    }
}

// D: Dirac position sweep per pixel sample, and the dx_image edge counts
template <int MAXB>
__global__ void __launch_bounds__(DR_BLOCK) k_dp(const DrJob *__restrict__ jobs, const DrRange *__restrict__ R, int nr,
                                                 const PointRec *pts, const float *pt_dp, const int *counts) {
    __shared__ float s_cam[64];
    __shared__ unsigned long long gkey[DR_GC_SLOTS_SMALL];
    __shared__ float gval[DR_GC_SLOTS_SMALL];
    if (threadIdx.x < 64) s_cam[threadIdx.x] = 0.f;
    gc_init(gkey, gval, DR_GC_SLOTS_SMALL);
    __syncthreads();
    const DrRange rg = R[range_of_block(R, nr, blockIdx.x)];
    const DrJob &J = jobs[rg.job];
    long long t = rg.t0 + (long long)(blockIdx.x - rg.block0) * blockDim.x + threadIdx.x;
    if (t < rg.t1) {
        const SampleCtx &c = J.c;
        const int per = J.per, chunks = J.chunks;
        int pix = (int)(t / chunks), ch = (int)(t % chunks);
        int xi = pix % c.W, yi = pix / c.W;
        GradSink G;
        G.grads = J.grads;
        G.cam = s_cam;
        G.flags = c.flags;
        G.color_mode = false;
        G.gkey = gkey;
        G.gval = gval;
        G.gmask = DR_GC_SLOTS_SMALL - 1;
        G.job = (unsigned)rg.job;
        float dx = 0.f;
        for (int k = 0; k < per; k++) {
            int s = ch * per + k;
            if (s >= c.spp) break;
            long long base = rg.slot0 + ((t - rg.t0) * per + k) * MAXB;
            const PointRec *P = pts + base;
            F3 dps[MAXB];
            int last = -1;
            for (int b = 0; b < MAXB; b++) {
                dps[b] = load3(pt_dp + 3 * (base + b));
                if (P[b].valid) last = b;
            }
            if (last >= 0) dx += counts[base + last] / 1000.f;
            stage_dp<MAXB>(J.S, c, G, xi, yi, s, P, dps);
        }
        if (J.dx) {
            float *qx = J.dx + 3 * (size_t)pix;
            if (chunks == 1) {
                qx[0] = qx[1] = qx[2] = dx;
            } else {
                atomicAdd(qx, dx); atomicAdd(qx + 1, dx); atomicAdd(qx + 2, dx);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < 64 && s_cam[threadIdx.x] != 0.f) atomicAdd(J.grads + threadIdx.x, s_cam[threadIdx.x]);
    gc_flush(gkey, gval, DR_GC_SLOTS_SMALL, [&](int j) { return jobs[j].grads; });
}

__global__ void k_fill_i(int *p, size_t n, int v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// primary-edge samples; a range is a piece [t0, t1) of one job's samples
template <int MAXB>
__global__ void __launch_bounds__(DR_BLOCK) k_primary(const DrJob *__restrict__ jobs, const DrRange *__restrict__ R,
                                                      int nr) {
    __shared__ float s_w2c[16];
    // vertex adjoints go through the workgroup's LDS cache: on a small scene
    // every sample hits the same few vertices, and direct global atomics on
    // them serialise (7.5 ms for 262 k samples of a single triangle)
    __shared__ unsigned long long gkey[DR_GC_SLOTS_SMALL];
    __shared__ float gval[DR_GC_SLOTS_SMALL];
    if (threadIdx.x < 16) s_w2c[threadIdx.x] = 0.f;
    gc_init(gkey, gval, DR_GC_SLOTS_SMALL);
    __syncthreads();
    const DrRange rg = R[range_of_block(R, nr, blockIdx.x)];
    const DrJob &J = jobs[rg.job];
    const DrSceneView &S = J.S;
    const SampleCtx &c = J.c;
    long long k = rg.t0 + (long long)(blockIdx.x - rg.block0) * blockDim.x + threadIdx.x;
    if (k < rg.t1 && S.num_prim_edges > 0) {
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
            PrimaryResult Rs;
            sample_primary_edge_point(S, e, eu, ls, ss, c.nb, J.d_image, c.W, c.H, w / c.spp, c.flags, Rs);
            if (Rs.ok) {
                GradSink G;
                G.grads = J.grads;
                G.cam = nullptr;
                G.flags = c.flags;
                G.color_mode = true;
                G.gkey = gkey;
                G.gval = gval;
                G.gmask = DR_GC_SLOTS_SMALL - 1;
                G.job = (unsigned)rg.job;
                const DrShape &es = S.shapes[e.shape];
                G.add3(es.vgrad_off + 3ll * e.v0, Rs.dv0);
                G.add3(es.vgrad_off + 3ll * e.v1, Rs.dv1);
                for (int i = 0; i < 16; i++)
                    if (Rs.dw2c[i] != 0.f) atomicAdd(&s_w2c[i], Rs.dw2c[i]);
            }
        }
        if (S.stats) atomicAdd(&S.stats[4], 1ull);
    }
    __syncthreads();
    if (threadIdx.x < 16 && s_w2c[threadIdx.x] != 0.f) atomicAdd(J.grads + 16 + threadIdx.x, s_w2c[threadIdx.x]);
    gc_flush(gkey, gval, DR_GC_SLOTS_SMALL, [&](int j) { return jobs[j].grads; });
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
// dr_test_tree_stack: every lane runs `ops` random pushes and pops on the
// selection kernel's LDS stack and on a plain array stack side by side (odd
// waves with the alternative pmfs, even waves without); out[0] counts pops
// that differ, out[1] the deepest stack reached.
__global__ void __launch_bounds__(DR_BLOCK) k_tree_stack(long long ops, uint64_t seed, unsigned long long *out) {
    __shared__ float s_stack[DR_LDS_STACK * DR_LDS_WORDS * DR_BLOCK];
    LdsStack A;
    A.lds = (DR_LDS float *)s_stack + threadIdx.x;
    A.stride = DR_BLOCK;
    A.clear();
    PlainStack B;
    B.clear();
    Philox r;
    philox_init(r, seed, 4, (uint64_t)blockIdx.x * blockDim.x + threadIdx.x, 0);
    unsigned long long bad = 0;
    int deepest = 0;
    // with or without the alternative pmfs for a lane's whole run, as a walk
    // of the selection kernel keeps its proposal mode: an entry pushed
    // without them must not be popped with them (the stacks' slots for them
    // would hold whatever private memory held before)
    const bool alt = (threadIdx.x >> 6) & 1;
    for (long long k = 0; k < ops; k++) {
        float u = philox_next(r);
        bool push = B.size() == 0 || (B.size() < DR_TREE_STACK && u < 0.55f);
        if (push) {
            int x = (int)(philox_next(r) * 16777216.f);
            TreeEntry e{x - 8388608, __float_as_uint(philox_next(r)) ^ (uint32_t)x << 8,
                        __float_as_uint(philox_next(r)) ^ (uint32_t)x << 11, x ^ 0x5555};
            TreeAlt t{philox_next(r), philox_next(r)};
            A.push(e, t, alt);
            B.push(e, t, alt);
            deepest = max(deepest, B.size());
        } else {
            TreeAlt ta{0.f, 0.f}, tb{0.f, 0.f};
            TreeEntry a = A.pop(ta, alt), b = B.pop(tb, alt);
            if (a.code != b.code || a.cpmf != b.cpmf || a.apmf != b.apmf || a.off != b.off ||
                (alt && (ta.cone != tb.cone || ta.all != tb.all)))
                bad++;
        }
        if (A.size() != B.size()) bad++;
    }
    if (bad) atomicAdd(&out[0], bad);
    atomicMax(&out[1], (unsigned long long)deepest);
}

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
// the message of a failing entry point compiled in another file (dr_last_error)
void dr_set_error(const std::string &msg) { g_err = msg; }

static dr_status fail(dr_status s, const std::string &msg) {
    g_err = msg;
    return s;
}
// error reporting for the library's other translation units
extern "C" dr_status dr_internal_fail(dr_status s, const char *msg) { return fail(s, msg ? msg : "?"); }

// kernel timing: event pairs recorded around launches when enabled
enum { TK_FORWARD = 0, TK_PATH = 1, TK_PRIMARY = 2, TK_SELECT = 3, TK_EDGE = 4, TK_DP = 5, TK_SHADE = 6, TK_COUNT = 7 };
static_assert(TK_FORWARD == DR_K_FORWARD && TK_EDGE == DR_K_EDGE && TK_SHADE == DR_K_SHADE && TK_COUNT == DR_K_COUNT,
              "kernel timing slots follow dr_stats");
struct TimedLaunch {
    int kind;
    hipEvent_t a, b;
};
struct Timing {
    bool on = false;
    std::vector<TimedLaunch> timed;
    std::vector<hipEvent_t> pool;
};
static hipEvent_t get_event(Timing &t) {
    if (!t.pool.empty()) {
        hipEvent_t e = t.pool.back();
        t.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
static void free_timing(Timing &t) {
    for (auto &x : t.timed) { (void)hipEventDestroy(x.a); (void)hipEventDestroy(x.b); }
    for (auto e : t.pool) (void)hipEventDestroy(e);
    t = Timing();
}
// record an event before (begin) / after (end) a launch when timing is on
struct LaunchTimer {
    Timing *t;
    int kind;
    hipStream_t st;
    hipEvent_t a = nullptr;
    LaunchTimer(Timing *t_, int k, hipStream_t st_) : t(t_), kind(k), st(st_) {
        if (t && t->on) {
            a = get_event(*t);
            if (a) (void)hipEventRecord(a, st);
        }
    }
    ~LaunchTimer() {
        if (!a) return;
        hipEvent_t b = get_event(*t);
        if (b && hipEventRecord(b, st) == hipSuccess) t->timed.push_back(TimedLaunch{kind, a, b});
        else t->pool.push_back(a);
    }
};

// Pass sizes: pixel samples per pass (bounds the point and queue buffers)
// and the least number of threads a job should offer (sets samples per
// thread).  Chosen by dr_batch_autotune, or these defaults.
struct Tuning {
    long long pass_samples = 1ll << 22;
    long long min_threads = 1ll << 20;
    int select_refill = 64;  // 1-64: depth first with that refill threshold; 0: breadth first
};

// Device working set of the pipeline, grown on demand, and the hipGraph of
// the last captured job list.
struct Workspace {
    PointRec *pts = nullptr;
    float *pt_dp = nullptr;
    int *pt_lastj = nullptr;
    int *counts = nullptr;
    long long slot_cap = 0;
    EdgeEntry *q = nullptr;
    float4 *entry_dp = nullptr;
    int *cont = nullptr;                     // queue entries left for the continuation launch
    long long q_cap = 0;
    // queue length, work counter, largest queue length, continuations,
    // largest frontier end, largest leaf-list end (breadth-first selection)
    unsigned long long *cnt = nullptr;
    unsigned long long *cnt_host = nullptr;  // pinned, 8 words
    SelPoint *spt = nullptr;                 // breadth-first selection: per point slot
    SelEntry *fr[2] = {nullptr, nullptr};    // frontier of the current / next level
    SelEntry *lv = nullptr;                  // leaves reached
    unsigned long long *lvl = nullptr;       // entries per level (0-63), leaves (DR_SEL_LEAVES)
    long long spt_cap = 0, f_cap = 0, l_cap = 0;
    DrJob *jobs = nullptr;
    int jobs_cap = 0;
    DrRange *ranges = nullptr;
    int ranges_cap = 0;
    hipStream_t cap_stream = nullptr;        // stream graphs are captured on
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned long long> graph_key;
    unsigned long long epoch = 0;            // counts reallocations (a captured graph is stale after one)
};

static void free_workspace(Workspace &w) {
    void *ps[] = {w.pts, w.pt_dp, w.pt_lastj, w.counts, w.q, w.entry_dp, w.cont, w.cnt, w.jobs, w.ranges,
                  w.spt, w.fr[0], w.fr[1], w.lv, w.lvl};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    if (w.cnt_host) (void)hipHostFree(w.cnt_host);
    if (w.exec) (void)hipGraphExecDestroy(w.exec);
    if (w.cap_stream) (void)hipStreamDestroy(w.cap_stream);
    w = Workspace();
}

// Device arrays of the scene's structures (BVH, edges, edge trees, their
// refit tables and scratch, primary-edge buffers): replaced together when the
// edge set changes and the structures are rebuilt.
struct SceneStruct {
    std::vector<void *> allocs;
    const int *bvh_lv = nullptr, *bvh_lv_off = nullptr;
    const int *sil_lv = nullptr, *sil_lv_off = nullptr, *sil_rng = nullptr;
    const int *nrm_lv = nullptr, *nrm_lv_off = nullptr, *nrm_rng = nullptr;
    int bvh_nlev = 0, sil_nlev = 0, nrm_nlev = 0;
    const DrEdge *edges = nullptr, *cand = nullptr;
    const int *cand_keep = nullptr;
    float *nrm_on = nullptr;              // 6 floats per normal-tree edge
    float *nsum = nullptr, *wsum = nullptr;  // 3 / 1 floats per node of the larger tree
    float *prim_len = nullptr;            // one float per mesh edge
    float *edge_w = nullptr;              // learned per-edge weights (tree order), or NULL
    unsigned char *edge_on = nullptr;     // per tree edge: the filter's verdict (trees over every mesh edge)
    size_t prim_cap = 0;                  // capacity of the primary-edge buffers (mesh edges)
    void *prim_bufs[3] = {nullptr, nullptr, nullptr};  // edges, func, cdf
};

struct dr_scene {
    drh::HostScene hs;
    unsigned long long serial = 0;  // unique per created scene (keys captured graphs)
    DrSceneView view;
    std::vector<void *> allocs;     // arrays whose sizes never change (vertices, textures, lights, LTC)
    SceneStruct st;
    unsigned long long *stats_dev = nullptr;
    bool stats_on = false;
    Timing tm;
    Workspace ws;
    double build_ms = 0.0;
    bool host_stale = false;        // device vertices newer than hs.verts / hs.normals
    void *small = nullptr;          // device, 4 words: edge-filter counts, or primary count and integral
    float *stat_shape = nullptr;    // device, one float per shape (dr_scene_set_stat_shapes)
    int refits = 0, rebuilds = 0;
    int edges_off = 0;              // tree edges the filter drops on the current vertices (all-candidate trees)
    std::vector<dr_view *> views;   // views of this scene (their primary edges follow its vertices)
};

// a scene seen by another camera: its own primary-edge list
struct dr_view {
    dr_scene *scene;                   // NULL once the scene is destroyed
    DrCamera cam;
    int num_prim_edges = 0;
    float prim_func_int = 0.f;
    void *bufs[3] = {nullptr, nullptr, nullptr};
    size_t cap = 0;                    // capacity of bufs (the scene's mesh edges when allocated)
};

struct dr_batch {
    Workspace ws;
    Timing tm;
    Tuning tune;
};

template <class T, class P>
static bool upload_to(std::vector<void *> &allocs, const std::vector<T> &v, P *out) {
    if (v.empty()) {
        *out = nullptr;
        return true;
    }
    void *p = nullptr;
    if (hipMalloc(&p, v.size() * sizeof(T)) != hipSuccess) return false;
    allocs.push_back(p);
    if (hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return false;
    *out = (P)p;
    return true;
}
template <class T, class P>
static bool upload(dr_scene *s, const std::vector<T> &v, P *out) {
    return upload_to(s->allocs, v, out);
}
static bool alloc_to(std::vector<void *> &allocs, size_t bytes, void **out) {
    *out = nullptr;
    if (bytes == 0) return true;
    if (hipMalloc(out, bytes) != hipSuccess) {
        *out = nullptr;
        return false;
    }
    allocs.push_back(*out);
    return true;
}
// copy into an existing device array of the same size
template <class T, class P>
static bool refresh(const std::vector<T> &v, P dst) {
    return v.empty() || hipMemcpy((void *)dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
}

// capacity buffers for a primary-edge list of at most `cap` edges
static bool alloc_prim(void *bufs[3], size_t cap) {
    for (int i = 0; i < 3; i++) {
        if (bufs[i]) (void)hipFree(bufs[i]);
        bufs[i] = nullptr;
    }
    size_t c = std::max<size_t>(cap, 1);
    size_t sz[3] = {c * sizeof(DrEdge), c * sizeof(float), (c + 1) * sizeof(float)};
    for (int i = 0; i < 3; i++)
        if (hipMalloc(&bufs[i], sz[i]) != hipSuccess) return false;
    return true;
}

static void free_struct(SceneStruct &st) {
    for (void *p : st.allocs) (void)hipFree(p);
    for (int i = 0; i < 3; i++)
        if (st.prim_bufs[i]) (void)hipFree(st.prim_bufs[i]);
    st = SceneStruct();
}

// Upload the structures of hs (BVH, triangles, edges, edge trees, geometry
// records, refit tables, primary edges from the host build, statistics) and
// point the view at them.  Frees the previous ones.
static bool upload_structures(dr_scene *s) {
    const drh::HostScene &hs = s->hs;
    DrSceneView &v = s->view;
    free_struct(s->st);
    SceneStruct &st = s->st;
    std::vector<void *> &A = st.allocs;
    bool ok = upload_to(A, hs.nodes, &v.nodes) && upload_to(A, hs.tris, &v.tris) &&
              upload_to(A, hs.sil_edges, &v.sil_edges) && upload_to(A, hs.sil_nodes, &v.sil_nodes) &&
              upload_to(A, hs.nrm_edges, &v.nrm_edges) && upload_to(A, hs.nrm_nodes, &v.nrm_nodes) &&
              upload_to(A, hs.nrm_geo, &v.nrm_geo) && upload_to(A, hs.sil_geo, &v.sil_geo) &&
              upload_to(A, hs.bvh_lv, &st.bvh_lv) && upload_to(A, hs.bvh_lv_off, &st.bvh_lv_off) &&
              upload_to(A, hs.sil_lv, &st.sil_lv) && upload_to(A, hs.sil_lv_off, &st.sil_lv_off) &&
              upload_to(A, hs.sil_rng, &st.sil_rng) && upload_to(A, hs.nrm_lv, &st.nrm_lv) &&
              upload_to(A, hs.nrm_lv_off, &st.nrm_lv_off) && upload_to(A, hs.nrm_rng, &st.nrm_rng) &&
              upload_to(A, hs.edges, &st.edges) && upload_to(A, hs.cand_edges, &st.cand) &&
              upload_to(A, hs.cand_keep, &st.cand_keep);
    if (!ok) return false;
    st.bvh_nlev = hs.bvh_lv_off.empty() ? 0 : (int)hs.bvh_lv_off.size() - 1;
    st.sil_nlev = hs.sil_lv_off.empty() ? 0 : (int)hs.sil_lv_off.size() - 1;
    st.nrm_nlev = hs.nrm_lv_off.empty() ? 0 : (int)hs.nrm_lv_off.size() - 1;
    size_t maxn = std::max(hs.sil_nodes.size(), hs.nrm_nodes.size());
    size_t ne = hs.sil_edges.size() + hs.nrm_edges.size();
    void *p = nullptr;
    ok = alloc_to(A, 6 * hs.nrm_edges.size() * sizeof(float), (void **)&st.nrm_on) &&
         alloc_to(A, 3 * maxn * sizeof(float), (void **)&st.nsum) &&
         alloc_to(A, maxn * sizeof(float), (void **)&st.wsum) &&
         alloc_to(A, hs.edges.size() * sizeof(float), (void **)&st.prim_len) &&
         alloc_to(A, 3 * std::max<size_t>(ne, 1) * sizeof(float), &p);
    if (!ok) return false;
    v.edge_stat = (float *)p;
    if (hipMemset(p, 0, 3 * std::max<size_t>(ne, 1) * sizeof(float)) != hipSuccess) return false;
    v.edge_on = nullptr;
    if (hs.all_candidates) {
        if (!alloc_to(A, std::max<size_t>(ne, 1), (void **)&st.edge_on)) return false;
        v.edge_on = (decltype(v.edge_on))st.edge_on;
    }
    // primary edges: capacity of every mesh edge, filled from the host build
    st.prim_cap = hs.edges.size();
    if (!alloc_prim(st.prim_bufs, st.prim_cap)) return false;
    if (!hs.prim_edges.empty() &&
        (hipMemcpy(st.prim_bufs[0], hs.prim_edges.data(), hs.prim_edges.size() * sizeof(DrEdge),
                   hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(st.prim_bufs[1], hs.prim.func.data(), hs.prim.func.size() * sizeof(float),
                   hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(st.prim_bufs[2], hs.prim.cdf.data(), hs.prim.cdf.size() * sizeof(float),
                   hipMemcpyHostToDevice) != hipSuccess))
        return false;
    v.num_prim_edges = (int)hs.prim_edges.size();
    v.prim_func_int = hs.prim.func_int;
    v.prim_edges = (decltype(v.prim_edges))st.prim_bufs[0];
    v.prim_func = (decltype(v.prim_func))st.prim_bufs[1];
    v.prim_cdf = (decltype(v.prim_cdf))st.prim_bufs[2];
    v.num_tris = (int)hs.tri_shape.size();
    v.num_nodes = (int)hs.nodes.size();
    v.num_sil_nodes = (int)hs.sil_nodes.size();
    v.num_nrm_nodes = (int)hs.nrm_nodes.size();
    v.num_sil_edges = (int)hs.sil_edges.size();
    return true;
}

// light arrays of hs into the existing device arrays (same sizes)
static bool refresh_lights(dr_scene *s) {
    const drh::HostScene &hs = s->hs;
    DrSceneView &v = s->view;
    bool ok = refresh(hs.lights, v.lights) && refresh(hs.light_func, v.light_func) &&
              refresh(hs.light_cdf, v.light_cdf) && refresh(hs.light_sel.func, v.lights_func) &&
              refresh(hs.light_sel.cdf, v.lights_cdf);
    v.lights_func_int = hs.light_sel.func_int;
    v.light_total_intensity = hs.light_total_intensity;
    return ok;
}

static void free_scene(dr_scene *s) {
    free_workspace(s->ws);
    free_timing(s->tm);
    for (void *p : s->allocs) (void)hipFree(p);
    free_struct(s->st);
    if (s->stats_dev) (void)hipFree(s->stats_dev);
    if (s->small) (void)hipFree(s->small);
    delete s;
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

// samples per thread: give a job >= min_threads threads when its image allows
static void plan_chunks(const dr_render_options *o, long long min_threads, int &per, int &chunks) {
    long long pixels = (long long)o->width * o->height;
    chunks = 1;
    while ((long long)chunks * 2 <= o->num_samples && pixels * chunks < min_threads) chunks *= 2;
    per = (o->num_samples + chunks - 1) / chunks;
    chunks = (o->num_samples + per - 1) / per;
}

static unsigned long long g_alloc_epoch = 0;
static bool grow(void **p, size_t bytes) {
    g_alloc_epoch++;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    return hipMalloc(p, bytes) == hipSuccess;
}

static unsigned grid_of(long long n) { return (unsigned)((n + DR_BLOCK - 1) / DR_BLOCK); }

static dr_status check_opts(const dr_render_options *o) {
    if (!o) return fail(DR_ERROR_INVALID_ARGUMENT, "null options");
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

// Resident workgroups of a kernel on this device (for persistent launches).
template <class K>
static long long resident_blocks(K kernel, long long &cache) {
    if (cache == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, DR_BLOCK, 0) != hipSuccess || cus <= 0 ||
            per_cu <= 0) {
            cus = 256;
            per_cu = 2;
        }
        cache = (long long)cus * per_cu;
    }
    return cache;
}

// ------------------------------------------------------------ planning ----
struct Pass {
    int r0, r1;
    long long nslots, nblocks;
    bool edges, dp_last;
};
struct Plan {
    std::vector<DrJob> jobs;
    std::vector<DrRange> ranges;
    std::vector<Pass> passes;
    int prim_r0 = 0, prim_r1 = 0, fwd_r0 = 0, fwd_r1 = 0;
    long long prim_blocks = 0, fwd_blocks = 0, max_slots = 0;
    int maxb = 1;
    int select_refill = 64;
    int sel_depth = 0;  // deepest edge-tree node of the adjoint jobs' scenes (levels of the breadth-first selection)
    std::vector<std::pair<float *, size_t>> zero;  // buffers cleared first
    std::vector<unsigned long long> key;           // what a captured graph depends on
};

static int bucket(int nb) { return nb <= 1 ? 1 : nb <= 2 ? 2 : nb <= 4 ? 4 : 8; }

static dr_status build_plan(const dr_job *jobs, int n, const Tuning &tn, Plan &P) {
    P = Plan();
    if (n <= 0 || !jobs) return fail(DR_ERROR_INVALID_ARGUMENT, "no jobs");
    std::vector<float *> grads_seen;
    for (int i = 0; i < n; i++) {
        const dr_job &jb = jobs[i];
        dr_scene *sc = jb.scene;
        if (!sc) return fail(DR_ERROR_INVALID_ARGUMENT, "job without a scene");
        dr_status st = check_opts(&jb.opts);
        if (st != DR_OK) return st;
        if (!jb.image) return fail(DR_ERROR_INVALID_ARGUMENT, "job without an image buffer");
        bool adj = jb.d_image != nullptr;
        if (adj && !jb.grads) return fail(DR_ERROR_INVALID_ARGUMENT, "adjoint job without a gradient buffer");
        if (adj && jb.opts.max_bounces > 8) return fail(DR_ERROR_UNSUPPORTED, "the adjoint supports max_bounces <= 8");
        if (jb.view && jb.view->scene != sc) return fail(DR_ERROR_INVALID_ARGUMENT, "view of another scene");
        DrJob dj;
        std::memset(&dj, 0, sizeof(dj));
        dj.S = sc->view;
        if (jb.view) {
            dj.S.cam = jb.view->cam;
            dj.S.num_prim_edges = jb.view->num_prim_edges;
            dj.S.prim_func_int = jb.view->prim_func_int;
            dj.S.prim_edges = (decltype(dj.S.prim_edges))jb.view->bufs[0];
            dj.S.prim_func = (decltype(dj.S.prim_func))jb.view->bufs[1];
            dj.S.prim_cdf = (decltype(dj.S.prim_cdf))jb.view->bufs[2];
        }
        dj.c = make_ctx(&jb.opts);
        plan_chunks(&jb.opts, tn.min_threads, dj.per, dj.chunks);
        dj.prim_n = jb.opts.num_primary_edge_samples > 0
                        ? jb.opts.num_primary_edge_samples
                        : (long long)jb.opts.num_samples * jb.opts.width * jb.opts.height;
        dj.d_image = jb.d_image;
        dj.image = jb.image;
        dj.grads = adj ? jb.grads : nullptr;
        dj.dx = (adj && (jb.opts.flags & DR_WRITE_DX_IMAGE)) ? jb.dx_image : nullptr;
        size_t npix = 3ull * jb.opts.width * jb.opts.height;
        if (dj.chunks > 1) P.zero.push_back({jb.image, npix});
        if (dj.dx) P.zero.push_back({dj.dx, npix});
        if (adj && std::find(grads_seen.begin(), grads_seen.end(), jb.grads) == grads_seen.end()) {
            grads_seen.push_back(jb.grads);
            P.zero.push_back({jb.grads, sc->hs.grad_size});
        }
        if (adj) P.maxb = std::max(P.maxb, bucket(jb.opts.max_bounces));
        if (adj) P.sel_depth = std::max(P.sel_depth, std::max(sc->hs.sil_depth, sc->hs.nrm_depth));
        P.jobs.push_back(dj);
        unsigned long long k[] = {(unsigned long long)(uintptr_t)sc, sc->serial, (unsigned long long)sc->hs.grad_size,
                                  (unsigned long long)(uintptr_t)jb.view,
                                  (unsigned long long)jb.opts.width, (unsigned long long)jb.opts.height,
                                  (unsigned long long)jb.opts.num_samples, (unsigned long long)jb.opts.max_bounces,
                                  (unsigned long long)jb.opts.flags, (unsigned long long)dj.prim_n,
                                  (unsigned long long)(uintptr_t)jb.d_image, (unsigned long long)(uintptr_t)jb.image,
                                  (unsigned long long)(uintptr_t)dj.grads, (unsigned long long)(uintptr_t)dj.dx,
                                  (unsigned long long)(uintptr_t)sc->view.stats};
        P.key.insert(P.key.end(), k, k + sizeof(k) / sizeof(k[0]));
    }
    // passes over the adjoint jobs' (pixel, chunk) threads
    long long cur_samples = 0, cur_slots = 0, cur_blocks = 0;
    Pass pass{0, 0, 0, 0, false, false};
    auto close = [&]() {
        if (pass.r1 > pass.r0) {
            pass.nslots = cur_slots;
            pass.nblocks = cur_blocks;
            P.passes.push_back(pass);
            P.max_slots = std::max(P.max_slots, cur_slots);
        }
        pass = Pass{(int)P.ranges.size(), (int)P.ranges.size(), 0, 0, false, false};
        cur_samples = cur_slots = cur_blocks = 0;
    };
    for (int i = 0; i < n; i++) {
        const DrJob &dj = P.jobs[i];
        if (!dj.d_image) continue;
        long long total = (long long)dj.c.W * dj.c.H * dj.chunks;
        for (long long t = 0; t < total;) {
            long long room = (tn.pass_samples - cur_samples) / dj.per;
            if (room <= 0 && cur_samples > 0) {
                close();
                continue;
            }
            long long take = std::min(total - t, std::max(room, 1ll));
            DrRange r{i, 0, t, t + take, cur_slots, cur_blocks};
            P.ranges.push_back(r);
            pass.r1 = (int)P.ranges.size();
            pass.edges |= !(dj.c.flags & DR_SKIP_SECONDARY_EDGES);
            pass.dp_last |= (dj.c.flags & DR_QUIRK_DP_LAST) != 0;
            cur_samples += take * dj.per;
            cur_slots += take * dj.per * P.maxb;
            cur_blocks += (take + DR_BLOCK - 1) / DR_BLOCK;
            t += take;
        }
    }
    close();
    // primary-edge samples of the adjoint jobs, one launch
    P.prim_r0 = (int)P.ranges.size();
    long long blocks = 0;
    for (int i = 0; i < n; i++) {
        const DrJob &dj = P.jobs[i];
        if (!dj.d_image || (dj.c.flags & DR_SKIP_PRIMARY_EDGES) || dj.prim_n <= 0) continue;
        P.ranges.push_back(DrRange{i, 0, 0, dj.prim_n, 0, blocks});
        blocks += (dj.prim_n + DR_BLOCK - 1) / DR_BLOCK;
    }
    P.prim_r1 = (int)P.ranges.size();
    P.prim_blocks = blocks;
    // forward-only jobs, one launch
    P.fwd_r0 = (int)P.ranges.size();
    blocks = 0;
    for (int i = 0; i < n; i++) {
        const DrJob &dj = P.jobs[i];
        if (dj.d_image) continue;
        long long total = (long long)dj.c.W * dj.c.H * dj.chunks;
        P.ranges.push_back(DrRange{i, 0, 0, total, 0, blocks});
        blocks += (total + DR_BLOCK - 1) / DR_BLOCK;
    }
    P.fwd_r1 = (int)P.ranges.size();
    P.fwd_blocks = blocks;
    // the breadth-first selection keeps a count per level for 63 levels;
    // deeper trees are walked depth first
    P.select_refill = tn.select_refill == 0 && P.sel_depth > 62 ? 64 : tn.select_refill;
    P.key.push_back((unsigned long long)P.sel_depth);
    P.key.push_back((unsigned long long)tn.pass_samples);
    P.key.push_back((unsigned long long)tn.min_threads);
    P.key.push_back((unsigned long long)tn.select_refill);
    return DR_OK;
}

static dr_status ensure_workspace_(Workspace &w, const Plan &P);
static bool has_edges(const Plan &P);
static dr_status grow_frontier(Workspace &w, long long n);
static dr_status grow_leaves(Workspace &w, long long n);
static dr_status ensure_workspace(Workspace &w, const Plan &P) {
    unsigned long long e0 = g_alloc_epoch;
    dr_status st = ensure_workspace_(w, P);
    if (g_alloc_epoch != e0) w.epoch++;
    return st;
}
static dr_status ensure_workspace_(Workspace &w, const Plan &P) {
    if (!w.cnt) {
        if (hipMalloc((void **)&w.cnt, 8 * sizeof(unsigned long long)) != hipSuccess ||
            hipHostMalloc((void **)&w.cnt_host, 8 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess)
            return fail(DR_ERROR_OUT_OF_MEMORY, "adjoint counters");
    }
    if (P.select_refill == 0 && has_edges(P)) {
        // breadth-first selection: per point slot its constants; frontier and
        // leaf lists start at 6 and 12 entries per slot (the heightfield at 2
        // bounces tests about 9 leaves per point) and grow on overflow
        if (!w.lvl && !grow((void **)&w.lvl, 72 * sizeof(unsigned long long)))
            return fail(DR_ERROR_OUT_OF_MEMORY, "selection counters");
        if (P.max_slots > w.spt_cap) {
            if (!grow((void **)&w.spt, P.max_slots * sizeof(SelPoint))) {
                w.spt_cap = 0;
                return fail(DR_ERROR_OUT_OF_MEMORY, "selection point buffer");
            }
            w.spt_cap = P.max_slots;
        }
        if (6 * P.max_slots > w.f_cap) {
            dr_status st = grow_frontier(w, 6 * P.max_slots);
            if (st != DR_OK) return st;
        }
        if (12 * P.max_slots > w.l_cap) {
            dr_status st = grow_leaves(w, 12 * P.max_slots);
            if (st != DR_OK) return st;
        }
    }
    if (P.max_slots > w.slot_cap) {
        long long n = P.max_slots;
        if (!grow((void **)&w.pts, n * sizeof(PointRec)) || !grow((void **)&w.pt_dp, n * 3 * sizeof(float)) ||
            !grow((void **)&w.pt_lastj, n * sizeof(int)) || !grow((void **)&w.counts, n * sizeof(int))) {
            w.slot_cap = 0;
            return fail(DR_ERROR_OUT_OF_MEMORY, "adjoint point buffers");
        }
        w.slot_cap = n;
    }
    // four queue entries per point slot (the heightfield at 2 bounces selects
    // about 3.5); a pass that selects more grows the queue
    long long qn = P.max_slots * 4;
    if (qn > w.q_cap) {
        if (!grow((void **)&w.q, qn * sizeof(EdgeEntry)) || !grow((void **)&w.entry_dp, qn * sizeof(float4)) ||
            !grow((void **)&w.cont, qn * sizeof(int))) {
            w.q_cap = 0;
            return fail(DR_ERROR_OUT_OF_MEMORY, "adjoint edge queue");
        }
        w.q_cap = qn;
    }
    if ((int)P.jobs.size() > w.jobs_cap) {
        if (!grow((void **)&w.jobs, P.jobs.size() * sizeof(DrJob))) return fail(DR_ERROR_OUT_OF_MEMORY, "job table");
        w.jobs_cap = (int)P.jobs.size();
    }
    if ((int)P.ranges.size() > w.ranges_cap) {
        if (!grow((void **)&w.ranges, std::max<size_t>(1, P.ranges.size()) * sizeof(DrRange)))
            return fail(DR_ERROR_OUT_OF_MEMORY, "range table");
        w.ranges_cap = (int)P.ranges.size();
    }
    return DR_OK;
}

static dr_status grow_frontier(Workspace &w, long long n) {
    w.epoch++;
    if (!grow((void **)&w.fr[0], n * sizeof(SelEntry)) || !grow((void **)&w.fr[1], n * sizeof(SelEntry))) {
        w.f_cap = 0;
        return fail(DR_ERROR_OUT_OF_MEMORY, "selection frontier");
    }
    w.f_cap = n;
    return DR_OK;
}
static dr_status grow_leaves(Workspace &w, long long n) {
    w.epoch++;
    if (!grow((void **)&w.lv, n * sizeof(SelEntry))) {
        w.l_cap = 0;
        return fail(DR_ERROR_OUT_OF_MEMORY, "selection leaf list");
    }
    w.l_cap = n;
    return DR_OK;
}

static dr_status grow_queue(Workspace &w, long long n) {
    w.epoch++;
    if (!grow((void **)&w.q, n * sizeof(EdgeEntry)) || !grow((void **)&w.entry_dp, n * sizeof(float4)) ||
        !grow((void **)&w.cont, n * sizeof(int))) {
        w.q_cap = 0;
        return fail(DR_ERROR_OUT_OF_MEMORY, "adjoint edge queue");
    }
    w.q_cap = n;
    return DR_OK;
}

static long long g_select_blocks = 0, g_edge_blocks[4] = {0, 0, 0, 0}, g_shade_blocks[4] = {0, 0, 0, 0};
static long long g_sel_blocks[3] = {0, 0, 0};

// All launches of a plan, stream-ordered, without host synchronisation (so
// they can be captured in a hipGraph).  The tables are already on the device.
static dr_status grow_queue(Workspace &w, long long n);

// After a pass's edge selection, when the host may wait (no graph capture):
// read the queue length and, if the queue overflowed, grow it and run the
// selection of that pass again (it only writes the queue and the per-point
// counts, so a second run replaces the first).
// Grows what overflowed (edge queue; breadth-first frontier and leaf list)
// after the counters cnt[0..5] were read into w.cnt_host; grew: whether
// something did.
static dr_status grow_overflowed(Workspace &w, long long qlen, bool bfs, bool &grew) {
    grew = false;
    if (qlen > w.q_cap) {
        dr_status st = grow_queue(w, qlen + qlen / 4);
        if (st != DR_OK) return st;
        grew = true;
    }
    if (bfs) {
        long long f = (long long)w.cnt_host[4], l = (long long)w.cnt_host[5];
        // at least twice as large: entries a full list dropped never reached
        // the next levels, which may then overflow in turn
        if (f > w.f_cap) {
            dr_status st = grow_frontier(w, std::max(2 * f, 2 * w.f_cap));
            if (st != DR_OK) return st;
            grew = true;
        }
        if (l > w.l_cap) {
            dr_status st = grow_leaves(w, std::max(2 * l, 2 * w.l_cap));
            if (st != DR_OK) return st;
            grew = true;
        }
    }
    return DR_OK;
}

template <class Launch>
static dr_status settle_selection(Workspace &w, hipStream_t hs, bool bfs, Launch &&select) {
    for (int attempt = 0; attempt < 8; attempt++) {
        if (hipMemcpyAsync(w.cnt_host, w.cnt, 6 * sizeof(unsigned long long), hipMemcpyDeviceToHost, hs) !=
                hipSuccess ||
            hipStreamSynchronize(hs) != hipSuccess)
            return fail(DR_ERROR_DEVICE, "reading the edge queue length failed");
        bool grew = false;
        dr_status st = grow_overflowed(w, (long long)w.cnt_host[0], bfs, grew);
        if (st != DR_OK) return st;
        if (!grew) return DR_OK;
        k_fill<<<1, 64, 0, hs>>>((float *)w.cnt, 4, 0.f);
        select();
    }
    return fail(DR_ERROR_DEVICE, "edge queue kept overflowing");
}

// sync_select: settle every pass's queue length on the host before its edge
// samples (uncaptured runs); otherwise the caller checks the largest length
// once at the end.
template <int MAXB>
static dr_status enqueue(const Plan &P, Workspace &w, Timing *tm, hipStream_t hs, bool sync_select) {
    for (auto &z : P.zero) k_fill<<<256, 256, 0, hs>>>(z.first, z.second, 0.f);
    k_fill<<<1, 64, 0, hs>>>((float *)(w.cnt + 2), 8, 0.f);  // largest queue, continuations, largest lists
    const bool bfs = P.select_refill == 0;
    const int bi = MAXB == 1 ? 0 : MAXB == 2 ? 1 : MAXB == 4 ? 2 : 3;
    for (const Pass &ps : P.passes) {
        const DrRange *R = w.ranges + ps.r0;
        int nr = ps.r1 - ps.r0;
        {
            LaunchTimer t(tm, TK_PATH, hs);
            k_path<MAXB><<<(unsigned)ps.nblocks, DR_BLOCK, 0, hs>>>(w.jobs, R, nr, w.pts);
        }
        if (!ps.edges) continue;
        k_fill<<<1024, 256, 0, hs>>>(w.pt_dp, (size_t)ps.nslots * 3, 0.f);
        k_fill_i<<<1024, 256, 0, hs>>>(w.pt_lastj, (size_t)ps.nslots, -1);
        k_fill<<<1, 64, 0, hs>>>((float *)w.cnt, 4, 0.f);        // queue length, work counter
        k_fill<<<1, 64, 0, hs>>>((float *)(w.cnt + 3), 2, 0.f);  // continuations
        auto select = [&]() {
            LaunchTimer t(tm, TK_SELECT, hs);
            if (bfs) {
                k_fill<<<1, 256, 0, hs>>>((float *)w.lvl, 2 * 72, 0.f);
                long long b0 = std::min<long long>(grid_of(ps.nslots), resident_blocks(k_sel_start, g_sel_blocks[0]));
                k_sel_start<<<(unsigned)b0, DR_BLOCK, 0, hs>>>(w.jobs, ps.nslots, MAXB, w.pts, w.spt, w.fr[0], w.lv,
                                                               w.lvl, w.f_cap, w.l_cap, w.cnt, w.counts);
                long long b1 = resident_blocks(k_sel_level, g_sel_blocks[1]);
                for (int d = 0; d < P.sel_depth; d++)
                    k_sel_level<<<(unsigned)b1, DR_BLOCK, 0, hs>>>(w.jobs, w.spt, w.fr[d & 1], w.lvl + d,
                                                                   w.fr[(d + 1) & 1], w.lvl + d + 1, w.lv,
                                                                   w.lvl + DR_SEL_LEAVES, w.f_cap, w.l_cap, w.cnt);
                long long b2 = resident_blocks(k_sel_leaves, g_sel_blocks[2]);
                k_sel_leaves<<<(unsigned)b2, DR_BLOCK, 0, hs>>>(w.jobs, w.spt, w.lv, w.lvl + DR_SEL_LEAVES, w.l_cap,
                                                                w.q, w.cnt, w.q_cap, w.counts);
            } else {
                long long blocks = std::min<long long>(grid_of(ps.nslots), resident_blocks(k_select, g_select_blocks));
                k_select<<<(unsigned)blocks, DR_BLOCK, 0, hs>>>(w.jobs, ps.nslots, MAXB, w.pts, w.q, w.cnt, w.q_cap,
                                                                w.counts, P.select_refill, 0ll, ps.nslots);
            }
        };
        select();
        if (sync_select) {
            dr_status st = settle_selection(w, hs, bfs, select);
            if (st != DR_OK) return st;
        }
        long long eb = resident_blocks(k_edge<MAXB>, g_edge_blocks[bi]);
        {
            LaunchTimer t(tm, TK_EDGE, hs);
            k_edge<MAXB><<<(unsigned)eb, DR_BLOCK, 0, hs>>>(w.jobs, w.q, w.cnt, w.q_cap, w.pts, w.pt_dp, w.pt_lastj,
                                                            w.entry_dp, w.cont);
        }
        {
            LaunchTimer t(tm, TK_SHADE, hs);
            long long sb = resident_blocks(k_edge_shade<MAXB>, g_shade_blocks[bi]);
            k_edge_shade<MAXB><<<(unsigned)sb, DR_BLOCK, 0, hs>>>(w.jobs, w.q, w.cnt, w.pts, w.pt_dp, w.pt_lastj,
                                                                  w.entry_dp, w.cont);
        }
        if (ps.dp_last) k_dp_last<<<(unsigned)eb, DR_BLOCK, 0, hs>>>(w.q, w.cnt, w.q_cap, w.entry_dp, w.pt_lastj, w.pt_dp);
        {
            LaunchTimer t(tm, TK_DP, hs);
            k_dp<MAXB><<<(unsigned)ps.nblocks, DR_BLOCK, 0, hs>>>(w.jobs, R, nr, w.pts, w.pt_dp, w.counts);
        }
    }
    if (P.prim_r1 > P.prim_r0) {
        LaunchTimer t(tm, TK_PRIMARY, hs);
        k_primary<MAXB><<<(unsigned)P.prim_blocks, DR_BLOCK, 0, hs>>>(w.jobs, w.ranges + P.prim_r0,
                                                                       P.prim_r1 - P.prim_r0);
    }
    if (P.fwd_r1 > P.fwd_r0) {
        LaunchTimer t(tm, TK_FORWARD, hs);
        k_forward<<<(unsigned)P.fwd_blocks, DR_BLOCK, 0, hs>>>(w.jobs, w.ranges + P.fwd_r0, P.fwd_r1 - P.fwd_r0);
    }
    return DR_OK;
}

static dr_status enqueue_any(const Plan &P, Workspace &w, Timing *tm, hipStream_t hs, bool sync_select) {
    if (P.maxb <= 1) return enqueue<1>(P, w, tm, hs, sync_select);
    if (P.maxb <= 2) return enqueue<2>(P, w, tm, hs, sync_select);
    if (P.maxb <= 4) return enqueue<4>(P, w, tm, hs, sync_select);
    return enqueue<8>(P, w, tm, hs, sync_select);
}

static bool has_edges(const Plan &P) {
    for (const Pass &ps : P.passes)
        if (ps.edges) return true;
    return false;
}

// Run a plan: tables to the device, then the launches.  Uncaptured, every
// pass's queue length is settled after its edge selection (an overflowing
// queue is grown and only that selection runs again).  With use_graph the
// launches are captured once and replayed as a hipGraph, and, when there are
// secondary edges, the largest queue length is read once at the end; a queue
// that overflowed is grown and the plan runs again (every output is
// rewritten from zero by a run).
static dr_status run_plan(const Plan &P, Workspace &w, Timing *tm, hipStream_t hs, bool use_graph) {
    dr_status st = ensure_workspace(w, P);
    if (st != DR_OK) return st;
    for (int attempt = 0; attempt < 8; attempt++) {
        if (hipMemcpyAsync(w.jobs, P.jobs.data(), P.jobs.size() * sizeof(DrJob), hipMemcpyHostToDevice, hs) !=
                hipSuccess ||
            (!P.ranges.empty() && hipMemcpyAsync(w.ranges, P.ranges.data(), P.ranges.size() * sizeof(DrRange),
                                                 hipMemcpyHostToDevice, hs) != hipSuccess))
            return fail(DR_ERROR_DEVICE, "job table upload failed");
        if (use_graph) {
            // every device buffer the captured launches use, and the
            // workspace's reallocation count
            std::vector<unsigned long long> key = P.key;
            const void *bufs[] = {w.pts, w.pt_dp, w.pt_lastj, w.counts, w.q, w.entry_dp, w.cont, w.cnt, w.jobs,
                                  w.ranges, w.spt, w.fr[0], w.fr[1], w.lv, w.lvl};
            for (const void *b : bufs) key.push_back((unsigned long long)(uintptr_t)b);
            key.push_back((unsigned long long)w.q_cap);
            key.push_back((unsigned long long)w.slot_cap);
            key.push_back((unsigned long long)w.f_cap);
            key.push_back((unsigned long long)w.l_cap);
            key.push_back(w.epoch);
            if (!w.exec || key != w.graph_key) {
                if (w.exec) (void)hipGraphExecDestroy(w.exec);
                w.exec = nullptr;
                w.graph_key.clear();
                if (!w.cap_stream && hipStreamCreate(&w.cap_stream) != hipSuccess)
                    return fail(DR_ERROR_DEVICE, "stream creation failed");
                hipGraph_t g = nullptr;
                if (hipStreamBeginCapture(w.cap_stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
                    return fail(DR_ERROR_DEVICE, "graph capture failed to start");
                (void)enqueue_any(P, w, nullptr, w.cap_stream, false);
                hipError_t e = hipStreamEndCapture(w.cap_stream, &g);
                if (e != hipSuccess || !g) return fail(DR_ERROR_DEVICE, "graph capture failed");
                e = hipGraphInstantiate(&w.exec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (e != hipSuccess) {
                    w.exec = nullptr;
                    return fail(DR_ERROR_DEVICE, "graph instantiation failed");
                }
                w.graph_key = key;
            }
            if (hipGraphLaunch(w.exec, hs) != hipSuccess) return fail(DR_ERROR_DEVICE, "graph launch failed");
        } else {
            st = enqueue_any(P, w, tm, hs, true);
            if (st != DR_OK) return st;
        }
        st = launch_check("render kernels");
        if (st != DR_OK) return st;
        if (!has_edges(P)) return DR_OK;
        if (hipMemcpyAsync(w.cnt_host, w.cnt, 6 * sizeof(unsigned long long), hipMemcpyDeviceToHost, hs) !=
                hipSuccess ||
            hipStreamSynchronize(hs) != hipSuccess)
            return fail(DR_ERROR_DEVICE, "reading the edge queue length failed");
        bool grew = false;
        st = grow_overflowed(w, (long long)w.cnt_host[2], P.select_refill == 0, grew);
        if (st != DR_OK) return st;
        if (!grew) return DR_OK;
    }
    return fail(DR_ERROR_DEVICE, "edge queue kept overflowing");
}

static dr_status single_job(dr_scene *s, const dr_render_options *o, const float *d_image, float *image, float *grads,
                            float *dx, hipStream_t hs) {
    dr_job jb;
    std::memset(&jb, 0, sizeof(jb));
    jb.scene = s;
    jb.opts = *o;
    jb.d_image = d_image;
    jb.image = image;
    jb.grads = grads;
    jb.dx_image = dx;
    Plan P;
    dr_status st = build_plan(&jb, 1, Tuning(), P);
    if (st != DR_OK) return st;
    return run_plan(P, s->ws, &s->tm, hs, false);
}

// ---------------------------------------------- allocation-free adjoint ---
// dr_render_adjoint_noalloc keeps every device buffer beyond the caller's
// inputs and outputs in `scratch`, a fixed array of this library (each GPU
// the library runs on holds its own copy until the process ends), for
// callers that cannot allocate.  Its passes are sized to fit the array; a
// pass whose edge selection overflows the queue part of it is selected and
// sampled in parts (a shading point's edges always fall in one part, so its
// Dirac derivatives and the last-ordinal rule see all of them).
#define DR_NOALLOC_BYTES (256ull << 20)
#define DR_NOALLOC_QUEUE_PER_SLOT 1
#define DR_NOALLOC_MAX_RANGES 4096
static __device__ __align__(256) unsigned char scratch[DR_NOALLOC_BYTES];
static std::mutex g_scratch_mu;
static void *g_scratch_dev[256];  // device address of `scratch`, per device

static dr_status scratch_address(void **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 256) return fail(DR_ERROR_DEVICE, "no current device");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(DR_ERROR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return DR_OK;
}

// slices of the scratch array (256-byte aligned)
struct Carver {
    char *base;
    size_t off, cap;
    template <class T>
    T *take(size_t n) {
        off = (off + 255) & ~(size_t)255;
        T *p = (T *)(base + off);
        off += n * sizeof(T);
        return off <= cap ? p : nullptr;
    }
};

template <int MAXB>
static dr_status run_noalloc(const Plan &P, Workspace &w, hipStream_t hs) {
    for (auto &z : P.zero) k_fill<<<256, 256, 0, hs>>>(z.first, z.second, 0.f);
    k_fill<<<1, 64, 0, hs>>>((float *)(w.cnt + 2), 2, 0.f);
    double per = -1.0;  // selected edges per point of the last part (sizes the next one)
    for (const Pass &ps : P.passes) {
        const DrRange *R = w.ranges + ps.r0;
        int nr = ps.r1 - ps.r0;
        k_path<MAXB><<<(unsigned)ps.nblocks, DR_BLOCK, 0, hs>>>(w.jobs, R, nr, w.pts);
        if (!ps.edges) continue;
        k_fill<<<1024, 256, 0, hs>>>(w.pt_dp, (size_t)ps.nslots * 3, 0.f);
        k_fill_i<<<1024, 256, 0, hs>>>(w.pt_lastj, (size_t)ps.nslots, -1);
        long long part = per > 0.0 ? std::max(1ll, (long long)(0.9 * (double)w.q_cap / per)) : ps.nslots;
        for (long long f0 = 0; f0 < ps.nslots;) {
            long long f1 = std::min(ps.nslots, f0 + part);
            k_fill<<<1, 64, 0, hs>>>((float *)w.cnt, 2 * 2, 0.f);       // queue length, work counter
            k_fill<<<1, 64, 0, hs>>>((float *)(w.cnt + 3), 2, 0.f);     // continuations
            long long blocks = std::min<long long>(grid_of(f1 - f0), resident_blocks(k_select, g_select_blocks));
            // (the depth-first selection: its working set fits the array)
            k_select<<<(unsigned)blocks, DR_BLOCK, 0, hs>>>(w.jobs, ps.nslots, MAXB, w.pts, w.q, w.cnt, w.q_cap,
                                                            w.counts, P.select_refill > 0 ? P.select_refill : 64,
                                                            f0, f1);
            unsigned long long len = 0;
            if (hipMemcpyAsync(&len, w.cnt, sizeof(len), hipMemcpyDeviceToHost, hs) != hipSuccess ||
                hipStreamSynchronize(hs) != hipSuccess)
                return fail(DR_ERROR_DEVICE, "reading the edge queue length failed");
            per = (double)len / (double)(f1 - f0);
            if ((long long)len > w.q_cap) {
                if (f1 - f0 <= 1)
                    return fail(DR_ERROR_UNSUPPORTED, "a shading point selects more edges than the allocation-free "
                                                      "queue holds");
                part = std::max(1ll, std::min((f1 - f0) / 2, (long long)(0.9 * (double)w.q_cap / per)));
                continue;
            }
            long long eb = resident_blocks(k_edge<MAXB>, g_edge_blocks[MAXB == 1 ? 0 : MAXB == 2 ? 1 : MAXB == 4 ? 2 : 3]);
            k_edge<MAXB><<<(unsigned)eb, DR_BLOCK, 0, hs>>>(w.jobs, w.q, w.cnt, w.q_cap, w.pts, w.pt_dp, w.pt_lastj,
                                                            w.entry_dp, w.cont);
            long long sb = resident_blocks(k_edge_shade<MAXB>,
                                           g_shade_blocks[MAXB == 1 ? 0 : MAXB == 2 ? 1 : MAXB == 4 ? 2 : 3]);
            k_edge_shade<MAXB><<<(unsigned)sb, DR_BLOCK, 0, hs>>>(w.jobs, w.q, w.cnt, w.pts, w.pt_dp, w.pt_lastj,
                                                                  w.entry_dp, w.cont);
            if (ps.dp_last)
                k_dp_last<<<(unsigned)eb, DR_BLOCK, 0, hs>>>(w.q, w.cnt, w.q_cap, w.entry_dp, w.pt_lastj, w.pt_dp);
            f0 = f1;
            if (per > 0.0) part = std::max(1ll, (long long)(0.9 * (double)w.q_cap / per));
        }
        k_dp<MAXB><<<(unsigned)ps.nblocks, DR_BLOCK, 0, hs>>>(w.jobs, R, nr, w.pts, w.pt_dp, w.counts);
    }
    if (P.prim_r1 > P.prim_r0)
        k_primary<MAXB><<<(unsigned)P.prim_blocks, DR_BLOCK, 0, hs>>>(w.jobs, w.ranges + P.prim_r0,
                                                                       P.prim_r1 - P.prim_r0);
    if (P.fwd_r1 > P.fwd_r0)
        k_forward<<<(unsigned)P.fwd_blocks, DR_BLOCK, 0, hs>>>(w.jobs, w.ranges + P.fwd_r0, P.fwd_r1 - P.fwd_r0);
    return launch_check("render kernels");
}

extern "C" {

const char *dr_last_error(void) { return g_err.c_str(); }

dr_status dr_scene_create(const dr_scene_desc *desc, dr_scene **out) {
    if (!out) return fail(DR_ERROR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    dr_scene *s = new (std::nothrow) dr_scene();
    if (!s) return fail(DR_ERROR_OUT_OF_MEMORY, "host allocation failed");
    static std::atomic<unsigned long long> serial{0};
    s->serial = ++serial;
    std::string err;
    try {
        err = drh::build_host_scene(desc, s->hs);
    } catch (const std::exception &e) {
        err = std::string("scene build failed: ") + e.what();
    }
    if (!err.empty()) {
        delete s;
        return fail(DR_ERROR_INVALID_ARGUMENT, err);
    }
    drh::HostScene &hs = s->hs;
    DrSceneView &v = s->view;
    std::memset(&v, 0, sizeof(v));
    bool ok = upload(s, hs.verts, &v.verts) && upload(s, hs.tri_vidx, &v.tri_vidx) &&
              upload(s, hs.tri_shape, &v.tri_shape) && upload(s, hs.uvs, &v.uvs) &&
              upload(s, hs.normals, &v.normals) && upload(s, hs.texels, &v.texels) &&
              upload(s, hs.shapes, &v.shapes) && upload(s, hs.materials, &v.materials) &&
              upload(s, hs.lights, &v.lights) && upload(s, hs.light_func, &v.light_func) &&
              upload(s, hs.light_cdf, &v.light_cdf) && upload(s, hs.light_sel.func, &v.lights_func) &&
              upload(s, hs.light_sel.cdf, &v.lights_cdf);
    std::vector<float> ltc(dr_ltc_blob, dr_ltc_blob + 16384 * 9);
    if ((size_t)(dr_ltc_blob_end - dr_ltc_blob) != 16384 * 9) {
        free_scene(s);
        return fail(DR_ERROR_DEVICE, "embedded LTC table has the wrong size");
    }
    ok = ok && upload(s, ltc, &v.ltc) && upload_structures(s) && hipMalloc(&s->small, 16) == hipSuccess;
    if (!ok) {
        free_scene(s);
        return fail(DR_ERROR_OUT_OF_MEMORY, "device allocation or copy failed");
    }
    v.lights_func_int = hs.light_sel.func_int;
    v.light_total_intensity = hs.light_total_intensity;
    v.num_lights = (int)hs.lights.size();
    v.num_shapes = (int)hs.shapes.size();
    v.cam = hs.cam;
    v.stats = nullptr;
    auto t1 = std::chrono::steady_clock::now();
    s->build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    *out = s;
    return DR_OK;
}

}  // extern "C"

// ------------------------------------------------------- device refit -----
static drrefit::RefitView refit_view(dr_scene *s) {
    const DrSceneView &v = s->view;
    const SceneStruct &st = s->st;
    drrefit::RefitView r;
    r.verts = v.verts;
    r.normals = v.normals;
    r.tri_vidx = v.tri_vidx;
    r.shapes = v.shapes;
    r.tris = (DrTri *)(v.tris);
    r.num_tris = (int)s->hs.tris.size();
    r.nodes = (DrNode *)(v.nodes);
    r.bvh_lv = st.bvh_lv;
    r.bvh_lv_off = st.bvh_lv_off;
    r.bvh_nlev = st.bvh_nlev;
    r.edges = st.edges;
    r.num_edges = (int)s->hs.edges.size();
    r.cand = st.cand;
    r.cand_keep = st.cand_keep;
    r.num_cand = (int)s->hs.cand_edges.size();
    r.filter_edges = s->hs.all_candidates;
    return r;
}

// The primary-edge list and distribution of camera `cam` on the device into
// the capacity buffers `bufs`; reads back the count and the integral
// (synchronises the stream).
static dr_status device_primary(dr_scene *s, const DrCamera &cam, void *bufs[3], int &count, float &fint,
                                hipStream_t hs) {
    drrefit::PrimaryOut o;
    o.len = s->st.prim_len;
    o.edges = (DrEdge *)bufs[0];
    o.func = (float *)bufs[1];
    o.cdf = (float *)bufs[2];
    o.count = (int *)s->small + 1;
    o.func_int = (float *)s->small + 2;
    k_fill<<<1, 64, 0, hs>>>((float *)s->small + 1, 2, 0.f);
    drrefit::primary_edges(refit_view(s), cam, o, hs);
    dr_status st = launch_check("primary-edge kernels");
    if (st != DR_OK) return st;
    int r[4];
    if (hipMemcpyAsync(r, s->small, 16, hipMemcpyDeviceToHost, hs) != hipSuccess || hipStreamSynchronize(hs) != hipSuccess)
        return fail(DR_ERROR_DEVICE, "reading the primary-edge count failed");
    count = r[1];
    std::memcpy(&fint, &r[2], sizeof(float));
    return DR_OK;
}

// device edge trees from the current device vertices (and learned weights)
static dr_status refit_edge_trees(dr_scene *s, hipStream_t hs) {
    const drh::HostScene &h = s->hs;
    const SceneStruct &st = s->st;
    DrSceneView &v = s->view;
    drrefit::RefitView rv = refit_view(s);
    drrefit::EdgeTreeRefit t;
    t.nsum = st.nsum;
    t.wsum = st.wsum;
    // silhouette tree: endpoints only, no cones
    t.nodes = (DrEdgeNode *)(v.sil_nodes);
    t.num_nodes = (int)h.sil_nodes.size();
    t.edges = v.sil_edges;
    t.num_edges = (int)h.sil_edges.size();
    t.directional = false;
    t.geo = (float *)(v.sil_geo);
    t.on = nullptr;
    t.weights = st.edge_w;
    t.flags = st.edge_on;
    t.lv = st.sil_lv;
    t.lv_off = st.sil_lv_off;
    t.rng = st.sil_rng;
    t.nlev = st.sil_nlev;
    drrefit::refit_edge_tree(rv, t, hs);
    t.nodes = (DrEdgeNode *)(v.nrm_nodes);
    t.num_nodes = (int)h.nrm_nodes.size();
    t.edges = v.nrm_edges;
    t.num_edges = (int)h.nrm_edges.size();
    t.directional = true;
    t.geo = (float *)(v.nrm_geo);
    t.on = st.nrm_on;
    t.weights = st.edge_w ? st.edge_w + h.sil_edges.size() : nullptr;
    t.flags = st.edge_on ? st.edge_on + h.sil_edges.size() : nullptr;
    t.lv = st.nrm_lv;
    t.lv_off = st.nrm_lv_off;
    t.rng = st.nrm_rng;
    t.nlev = st.nrm_nlev;
    drrefit::refit_edge_tree(rv, t, hs);
    return launch_check("edge-tree refit");
}

static dr_status sync_host_vertices(dr_scene *s) {
    if (!s->host_stale) return DR_OK;
    drh::HostScene &h = s->hs;
    if ((!h.verts.empty() &&
         hipMemcpy(h.verts.data(), s->view.verts, h.verts.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) ||
        (!h.normals.empty() && hipMemcpy(h.normals.data(), s->view.normals, h.normals.size() * sizeof(float),
                                         hipMemcpyDeviceToHost) != hipSuccess))
        return fail(DR_ERROR_DEVICE, "vertex read-back failed");
    s->host_stale = false;
    return DR_OK;
}

// Edge flags (trees over every mesh edge), edge trees and the scene camera's
// primary edges for the current device vertices (synchronises the stream).
static dr_status settle_edges(dr_scene *s, hipStream_t hs) {
    const drh::HostScene &h = s->hs;
    if (h.all_candidates) {
        drrefit::RefitView rv = refit_view(s);
        k_fill<<<1, 64, 0, hs>>>((float *)s->small, 1, 0.f);
        unsigned int *off = (unsigned int *)s->small;
        drrefit::edge_flags(rv, s->view.sil_edges, (int)h.sil_edges.size(), s->st.edge_on, off, hs);
        drrefit::edge_flags(rv, s->view.nrm_edges, (int)h.nrm_edges.size(), s->st.edge_on + h.sil_edges.size(), off,
                            hs);
        unsigned int n_off = 0;
        if (launch_check("edge flags") != DR_OK ||
            hipMemcpyAsync(&n_off, s->small, sizeof(n_off), hipMemcpyDeviceToHost, hs) != hipSuccess ||
            hipStreamSynchronize(hs) != hipSuccess)
            return fail(DR_ERROR_DEVICE, "edge flags failed");
        s->edges_off = (int)n_off;
    }
    dr_status st = refit_edge_trees(s, hs);
    if (st != DR_OK) return st;
    int n = 0;
    float fi = 0.f;
    st = device_primary(s, s->view.cam, s->st.prim_bufs, n, fi, hs);
    if (st != DR_OK) return st;
    s->view.num_prim_edges = n;
    s->view.prim_func_int = fi;
    return DR_OK;
}

// primary edges of every view of the scene, for its current vertices
static dr_status refresh_views(dr_scene *s, hipStream_t hs) {
    for (dr_view *v : s->views) {
        if (v->cap != s->st.prim_cap || !v->bufs[0]) {
            if (!alloc_prim(v->bufs, s->st.prim_cap)) return fail(DR_ERROR_OUT_OF_MEMORY, "device allocation failed");
            v->cap = s->st.prim_cap;
        }
        dr_status st = device_primary(s, v->cam, v->bufs, v->num_prim_edges, v->prim_func_int, hs);
        if (st != DR_OK) return st;
    }
    return DR_OK;
}

extern "C" {

void dr_scene_destroy(dr_scene *s) {
    if (!s) return;
    for (dr_view *v : s->views) v->scene = nullptr;
    free_scene(s);
}

dr_status dr_scene_set_camera(dr_scene *s, const dr_camera *c) {
    if (!s || !c) return fail(DR_ERROR_INVALID_ARGUMENT, "null argument");
    DrCamera cam = drh::make_camera(*c);
    int n = 0;
    float fi = 0.f;
    dr_status st = device_primary(s, cam, s->st.prim_bufs, n, fi, nullptr);
    if (st != DR_OK) return st;
    s->hs.cam = cam;
    s->view.cam = cam;
    s->view.num_prim_edges = n;
    s->view.prim_func_int = fi;
    return DR_OK;
}

dr_status dr_scene_update_vertices(dr_scene *s, const dr_shape_update *updates, int num_updates, void *stream) {
    if (!s || (num_updates > 0 && !updates) || num_updates < 0) return fail(DR_ERROR_INVALID_ARGUMENT, "bad arguments");
    hipStream_t hs = (hipStream_t)stream;
    drh::HostScene &h = s->hs;
    bool lights = false;
    for (int i = 0; i < num_updates; i++) {
        const dr_shape_update &u = updates[i];
        if (u.shape_id < 0 || u.shape_id >= (int)h.shapes.size()) return fail(DR_ERROR_INVALID_ARGUMENT, "shape id out of range");
        const DrShape &sh = h.shapes[u.shape_id];
        if (!u.vertices && sh.nv > 0) return fail(DR_ERROR_INVALID_ARGUMENT, "vertices are required");
        if (u.normals && !sh.has_n) return fail(DR_ERROR_INVALID_ARGUMENT, "the shape was built without normals");
        if (sh.nv > 0 && hipMemcpyAsync((float *)s->view.verts + 3 * (size_t)sh.vert_off, u.vertices,
                                        3 * (size_t)sh.nv * sizeof(float), hipMemcpyDeviceToDevice, hs) != hipSuccess)
            return fail(DR_ERROR_DEVICE, "vertex copy failed");
        if (u.normals && hipMemcpyAsync((float *)s->view.normals + 3 * (size_t)sh.vert_off, u.normals,
                                        3 * (size_t)sh.nv * sizeof(float), hipMemcpyDeviceToDevice, hs) != hipSuccess)
            return fail(DR_ERROR_DEVICE, "normal copy failed");
        lights |= sh.light >= 0;
    }
    s->host_stale = true;
    drrefit::RefitView rv = refit_view(s);
    bool rebuild = false;
    if (!h.all_candidates) {
        // the edge filter's verdicts on the new vertices
        k_fill<<<1, 64, 0, hs>>>((float *)s->small, 1, 0.f);
        drrefit::edge_filter(rv, (unsigned int *)s->small, hs);
        unsigned int changed = 0;
        if (launch_check("edge filter") != DR_OK ||
            hipMemcpyAsync(&changed, s->small, sizeof(changed), hipMemcpyDeviceToHost, hs) != hipSuccess ||
            hipStreamSynchronize(hs) != hipSuccess)
            return fail(DR_ERROR_DEVICE, "edge filter failed");
        rebuild = changed != 0;
    }
    if (rebuild) {
        // the edge set changed: rebuild the structures on the host
        dr_status st = sync_host_vertices(s);
        if (st != DR_OK) return st;
        std::string err = drh::build_structures(h);
        if (!err.empty()) return fail(DR_ERROR_INVALID_ARGUMENT, err);
        if (!upload_structures(s)) return fail(DR_ERROR_OUT_OF_MEMORY, "device allocation or copy failed");
        s->rebuilds++;
    } else {
        drrefit::refit_geometry(rv, hs);
        dr_status st = launch_check("BVH refit");
        if (st == DR_OK) st = settle_edges(s, hs);
        if (st != DR_OK) return st;
        s->refits++;
    }
    if (lights) {
        dr_status st = sync_host_vertices(s);
        if (st != DR_OK) return st;
        drh::build_lights(h);
        if (!refresh_lights(s)) return fail(DR_ERROR_DEVICE, "light upload failed");
    }
    return refresh_views(s, hs);
}

dr_status dr_scene_set_edge_policy(dr_scene *s, int policy) {
    if (!s || (policy != DR_EDGES_FILTERED && policy != DR_EDGES_ALL_CANDIDATES))
        return fail(DR_ERROR_INVALID_ARGUMENT, "unknown edge policy");
    drh::HostScene &h = s->hs;
    bool all = policy == DR_EDGES_ALL_CANDIDATES;
    if (all == h.all_candidates) return DR_OK;
    dr_status st = sync_host_vertices(s);
    if (st != DR_OK) return st;
    h.all_candidates = all;
    std::string err = drh::build_structures(h);
    if (!err.empty()) return fail(DR_ERROR_INVALID_ARGUMENT, err);
    if (!upload_structures(s)) return fail(DR_ERROR_OUT_OF_MEMORY, "device allocation or copy failed");
    s->edges_off = 0;
    if (all) {
        st = settle_edges(s, nullptr);
        if (st != DR_OK) return st;
    }
    return refresh_views(s, nullptr);
}

dr_status dr_scene_update_materials(dr_scene *s, const dr_material *materials, int num_materials) {
    if (!s || (num_materials > 0 && !materials)) return fail(DR_ERROR_INVALID_ARGUMENT, "bad arguments");
    drh::HostScene &h = s->hs;
    if (num_materials != (int)h.materials.size()) return fail(DR_ERROR_INVALID_ARGUMENT, "material count differs");
    std::vector<float> texels(h.texels.size());
    std::vector<DrMaterial> mats = h.materials;
    for (int i = 0; i < num_materials; i++) {
        const dr_material &m = materials[i];
        DrMaterial &o = mats[i];
        const dr_texture *t[3] = {&m.diffuse_reflectance, &m.specular_reflectance, &m.roughness};
        DrTex *d[3] = {&o.diffuse, &o.specular, &o.roughness};
        for (int k = 0; k < 3; k++) {
            if (t[k]->height != d[k]->h || (t[k]->height > 0 && t[k]->width != d[k]->w) || t[k]->channels != d[k]->c ||
                !t[k]->data)
                return fail(DR_ERROR_INVALID_ARGUMENT, "material texture sizes differ from the scene's");
            std::memcpy(&texels[d[k]->off], t[k]->data, dr_texture_size(t[k]) * sizeof(float));
        }
        for (int k = 0; k < 2; k++) {
            o.duv[k] = m.diffuse_uv_scale[k];
            o.suv[k] = m.specular_uv_scale[k];
            o.ruv[k] = m.roughness_uv_scale[k];
        }
        o.two_sided = m.two_sided ? 1 : 0;
    }
    h.texels.swap(texels);
    h.materials.swap(mats);
    if (!refresh(h.texels, s->view.texels) || !refresh(h.materials, s->view.materials))
        return fail(DR_ERROR_DEVICE, "material upload failed");
    return DR_OK;
}

dr_status dr_scene_update_lights(dr_scene *s, const dr_light *lights, int num_lights) {
    if (!s || (num_lights > 0 && !lights)) return fail(DR_ERROR_INVALID_ARGUMENT, "bad arguments");
    drh::HostScene &h = s->hs;
    if (num_lights != (int)h.light_shapes.size()) return fail(DR_ERROR_INVALID_ARGUMENT, "light count differs");
    for (int i = 0; i < num_lights; i++)
        if (lights[i].shape_id != h.light_shapes[i])
            return fail(DR_ERROR_INVALID_ARGUMENT, "a light's shape differs from the scene's");
    dr_status st = sync_host_vertices(s);
    if (st != DR_OK) return st;
    for (int i = 0; i < num_lights; i++)
        for (int k = 0; k < 3; k++) h.light_I[3 * i + k] = lights[i].intensity[k];
    drh::build_lights(h);
    if (!refresh_lights(s)) return fail(DR_ERROR_DEVICE, "light upload failed");
    return DR_OK;
}

size_t dr_scene_grad_size(const dr_scene *s) { return s ? s->hs.grad_size : 0; }

dr_status dr_scene_get_info(const dr_scene *s, dr_scene_info *o) {
    if (!s || !o) return fail(DR_ERROR_INVALID_ARGUMENT, "null argument");
    o->num_triangles = (int)s->hs.tri_shape.size();
    o->num_vertices = (int)(s->hs.verts.size() / 3);
    o->bvh_nodes = (int)s->hs.nodes.size();
    o->num_edges = (int)s->hs.edges.size();
    o->num_primary_edges = s->view.num_prim_edges;
    o->silhouette_tree_nodes = (int)s->hs.sil_nodes.size();
    o->normal_tree_nodes = (int)s->hs.nrm_nodes.size();
    o->build_ms = s->build_ms;
    o->refits = s->refits;
    o->rebuilds = s->rebuilds;
    o->edges_switched_off = s->edges_off;
    o->bvh_depth = s->hs.bvh_depth;
    o->edge_tree_depth = std::max(s->hs.sil_depth, s->hs.nrm_depth);
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
        out->stack_overflows = h[7];
        for (auto &t : s->tm.timed) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, t.a, t.b) != hipSuccess) ms = 0.f;
            if (t.kind < 0 || t.kind >= TK_COUNT) continue;
            out->kernel_ms[t.kind] += ms;
            out->kernel_launches[t.kind]++;
        }
    }
    for (auto &t : s->tm.timed) {
        s->tm.pool.push_back(t.a);
        s->tm.pool.push_back(t.b);
    }
    s->tm.timed.clear();
    (void)hipMemset(s->stats_dev, 0, 8 * sizeof(unsigned long long));
    s->tm.on = (enable & DR_STATS_TIMING) != 0;
    s->stats_on = (enable & DR_STATS_COUNTERS) != 0;
    s->view.stats = s->stats_on ? s->stats_dev : nullptr;
    return DR_OK;
}

dr_status dr_render(dr_scene *s, const dr_render_options *o, float *image, void *stream) {
    if (!s) return fail(DR_ERROR_INVALID_ARGUMENT, "null scene");
    dr_status st = check_opts(o);
    if (st != DR_OK) return st;
    if (!image) return fail(DR_ERROR_INVALID_ARGUMENT, "image is NULL");
    return single_job(s, o, nullptr, image, nullptr, nullptr, (hipStream_t)stream);
}

dr_status dr_render_adjoint(dr_scene *s, const dr_render_options *o, const float *d_image, float *image, float *grads,
                            float *dx_image, void *stream) {
    if (!s) return fail(DR_ERROR_INVALID_ARGUMENT, "null scene");
    dr_status st = check_opts(o);
    if (st != DR_OK) return st;
    if (!d_image || !image || !grads) return fail(DR_ERROR_INVALID_ARGUMENT, "d_image, image and grads are required");
    return single_job(s, o, d_image, image, grads, dx_image, (hipStream_t)stream);
}

dr_status dr_render_adjoint_noalloc(dr_scene *s, const dr_render_options *o, const float *d_image, float *image,
                                    float *grads, float *dx_image, void *stream) {
    if (!s) return fail(DR_ERROR_INVALID_ARGUMENT, "null scene");
    dr_status st = check_opts(o);
    if (st != DR_OK) return st;
    if (!d_image || !image || !grads) return fail(DR_ERROR_INVALID_ARGUMENT, "d_image, image and grads are required");
    if (o->max_bounces > 8) return fail(DR_ERROR_UNSUPPORTED, "the adjoint supports max_bounces <= 8");
    hipStream_t hs = (hipStream_t)stream;
    void *base = nullptr;
    st = scratch_address(&base);
    if (st != DR_OK) return st;
    dr_job jb;
    std::memset(&jb, 0, sizeof(jb));
    jb.scene = s;
    jb.opts = *o;
    jb.d_image = d_image;
    jb.image = image;
    jb.grads = grads;
    jb.dx_image = dx_image;
    // passes sized so that points and queue fit the array after the tables
    const size_t fixed = 64 * 1024 + DR_NOALLOC_MAX_RANGES * sizeof(DrRange) + 4 * 256;
    const size_t per_slot = sizeof(PointRec) + 3 * sizeof(float) + 2 * sizeof(int) +
                            DR_NOALLOC_QUEUE_PER_SLOT * (sizeof(EdgeEntry) + sizeof(float4) + sizeof(int));
    const int maxb = bucket(o->max_bounces);
    Tuning tn;
    tn.pass_samples = std::max<long long>(1024, (long long)((DR_NOALLOC_BYTES - fixed) / (per_slot * maxb)) & ~1023ll);
    tn.min_threads = std::min<long long>(tn.min_threads, tn.pass_samples);
    Plan P;
    st = build_plan(&jb, 1, tn, P);
    if (st != DR_OK) return st;
    if (P.ranges.size() > DR_NOALLOC_MAX_RANGES)
        return fail(DR_ERROR_UNSUPPORTED, "the render needs more passes than the allocation-free tables hold");
    Carver c{(char *)base, 0, DR_NOALLOC_BYTES};
    Workspace w;
    w.cnt = c.take<unsigned long long>(4);
    w.jobs = c.take<DrJob>(1);
    w.ranges = c.take<DrRange>(DR_NOALLOC_MAX_RANGES);
    long long slots = std::max<long long>(P.max_slots, 1);
    w.pts = c.take<PointRec>(slots);
    w.pt_dp = c.take<float>(3 * slots);
    w.pt_lastj = c.take<int>(slots);
    w.counts = c.take<int>(slots);
    long long qn = DR_NOALLOC_QUEUE_PER_SLOT * slots;
    w.q = c.take<EdgeEntry>(qn);
    w.entry_dp = c.take<float4>(qn);
    w.cont = c.take<int>(qn);
    if (!w.cont) return fail(DR_ERROR_UNSUPPORTED, "the render does not fit the allocation-free scratch array");
    w.slot_cap = slots;
    w.q_cap = qn;
    if (hipMemcpyAsync(w.jobs, P.jobs.data(), sizeof(DrJob), hipMemcpyHostToDevice, hs) != hipSuccess ||
        hipMemcpyAsync(w.ranges, P.ranges.data(), P.ranges.size() * sizeof(DrRange), hipMemcpyHostToDevice, hs) !=
            hipSuccess)
        return fail(DR_ERROR_DEVICE, "job table upload failed");
    if (P.maxb <= 1) st = run_noalloc<1>(P, w, hs);
    else if (P.maxb <= 2) st = run_noalloc<2>(P, w, hs);
    else if (P.maxb <= 4) st = run_noalloc<4>(P, w, hs);
    else st = run_noalloc<8>(P, w, hs);
    // the host tables must outlive their asynchronous upload
    if (hipStreamSynchronize(hs) != hipSuccess && st == DR_OK) st = fail(DR_ERROR_DEVICE, "stream synchronisation failed");
    return st;
}

int64_t dr_scene_edge_count(const dr_scene *s) {
    return s ? (int64_t)(s->hs.sil_edges.size() + s->hs.nrm_edges.size()) : 0;
}

dr_status dr_scene_edge_features(dr_scene *s, const dr_view *view, float *features, void *stream) {
    if (!s || !features) return fail(DR_ERROR_INVALID_ARGUMENT, "null argument");
    if (view && view->scene != s) return fail(DR_ERROR_INVALID_ARGUMENT, "view of another scene");
    dr_status st = sync_host_vertices(s);
    if (st != DR_OK) return st;
    std::vector<float> f;
    drh::edge_features(s->hs, view ? view->cam : s->hs.cam, f);
    if (f.empty()) return DR_OK;
    hipStream_t hs = (hipStream_t)stream;
    if (hipMemcpyAsync(features, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, hs) != hipSuccess ||
        hipStreamSynchronize(hs) != hipSuccess)
        return fail(DR_ERROR_DEVICE, "feature upload failed");
    return DR_OK;
}

dr_status dr_scene_build_bvh_device(dr_scene *s, double *build_ms, void *stream) {
    if (!s) return fail(DR_ERROR_INVALID_ARGUMENT, "null scene");
    hipStream_t hs = (hipStream_t)stream;
    dr_status st = sync_host_vertices(s);  // the structures are uploaded from the host copy
    if (st != DR_OK) return st;
    int n = (int)s->hs.tri_shape.size();
    if (n < 2) {
        if (build_ms) *build_ms = 0.0;
        return DR_OK;  // a single leaf: the host tree is that already
    }
    drlbvh::Layout L = drlbvh::layout(n);
    void *scratch = nullptr;
    DrNode *dn = nullptr;
    DrTri *dt = nullptr;
    if (hipMalloc(&scratch, L.total) != hipSuccess || hipMalloc((void **)&dn, (size_t)(n - 1) * sizeof(DrNode)) != hipSuccess ||
        hipMalloc((void **)&dt, (size_t)n * sizeof(DrTri)) != hipSuccess) {
        (void)hipFree(scratch);
        (void)hipFree(dn);
        return fail(DR_ERROR_OUT_OF_MEMORY, "device BVH build buffers");
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, hs);
    hipError_t e = drlbvh::build((const float *)s->view.verts, (const int *)s->view.tri_vidx, n, dn, dt, scratch,
                                 L.total, hs);
    (void)hipEventRecord(e1, hs);
    std::vector<DrNode> kn((size_t)n - 1);
    std::vector<DrTri> tris((size_t)n);
    if (e == hipSuccess)
        e = hipMemcpyAsync(kn.data(), dn, kn.size() * sizeof(DrNode), hipMemcpyDeviceToHost, hs);
    if (e == hipSuccess) e = hipMemcpyAsync(tris.data(), dt, tris.size() * sizeof(DrTri), hipMemcpyDeviceToHost, hs);
    if (e == hipSuccess) e = hipStreamSynchronize(hs);
    float ms = 0.f;
    if (e == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(scratch);
    (void)hipFree(dn);
    (void)hipFree(dt);
    if (e != hipSuccess) return fail(DR_ERROR_DEVICE, "device BVH build failed");
    std::string why = drh::adopt_device_bvh(s->hs, kn, tris);
    if (!why.empty()) return fail(DR_ERROR_UNSUPPORTED, why);
    if (!upload_structures(s)) return fail(DR_ERROR_OUT_OF_MEMORY, "device allocation or copy failed");
    if (build_ms) *build_ms = ms;
    return DR_OK;
}

dr_status dr_scene_set_stat_shapes(dr_scene *s, const int *shape_ids, int n) {
    if (!s) return fail(DR_ERROR_INVALID_ARGUMENT, "null scene");
    if (n < 0 || (n > 0 && !shape_ids)) return fail(DR_ERROR_INVALID_ARGUMENT, "bad shape list");
    int ns = s->view.num_shapes;
    for (int i = 0; i < n; i++)
        if (shape_ids[i] < 0 || shape_ids[i] >= ns) return fail(DR_ERROR_INVALID_ARGUMENT, "shape index out of range");
    if (n == 0) {
        s->view.stat_shape = nullptr;  // (the array stays with the scene)
        return DR_OK;
    }
    if (!s->stat_shape && !alloc_to(s->allocs, std::max(ns, 1) * sizeof(float), (void **)&s->stat_shape))
        return fail(DR_ERROR_OUT_OF_MEMORY, "device allocation failed");
    std::vector<float> m((size_t)std::max(ns, 1), 0.f);
    for (int i = 0; i < n; i++) m[shape_ids[i]] = 1.f;
    if (hipMemcpy(s->stat_shape, m.data(), m.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return fail(DR_ERROR_DEVICE, "mask upload failed");
    s->view.stat_shape = (decltype(s->view.stat_shape))s->stat_shape;
    return DR_OK;
}

dr_status dr_scene_set_edge_weights(dr_scene *s, const float *weights) {
    if (!s) return fail(DR_ERROR_INVALID_ARGUMENT, "null scene");
    SceneStruct &st = s->st;
    size_t ne = s->hs.sil_edges.size() + s->hs.nrm_edges.size();
    if (weights && ne > 0) {
        if (!st.edge_w && !alloc_to(st.allocs, ne * sizeof(float), (void **)&st.edge_w))
            return fail(DR_ERROR_OUT_OF_MEMORY, "device allocation failed");
        if (hipMemcpy(st.edge_w, weights, ne * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            return fail(DR_ERROR_DEVICE, "weight upload failed");
    } else {
        st.edge_w = nullptr;  // (the allocation stays with the structures)
    }
    // node lengths of the learned proposal, recomputed on the device with the
    // tree refit (which leaves the rest of the trees as they are)
    dr_status r = refit_edge_trees(s, nullptr);
    if (r != DR_OK) return r;
    if (hipStreamSynchronize(nullptr) != hipSuccess) return fail(DR_ERROR_DEVICE, "edge-tree refit failed");
    return DR_OK;
}

dr_status dr_scene_edge_stats(dr_scene *s, float *out, int reset, void *stream) {
    if (!s) return fail(DR_ERROR_INVALID_ARGUMENT, "null scene");
    hipStream_t hs = (hipStream_t)stream;
    size_t n = 3 * (size_t)dr_scene_edge_count(s);
    if (n == 0) return DR_OK;
    if (out && hipMemcpyAsync(out, s->view.edge_stat, n * sizeof(float), hipMemcpyDeviceToDevice, hs) != hipSuccess)
        return fail(DR_ERROR_DEVICE, "statistics copy failed");
    if (reset) k_fill<<<256, 256, 0, hs>>>(s->view.edge_stat, n, 0.f);
    return launch_check("dr_scene_edge_stats");
}

dr_status dr_view_create(dr_scene *s, const dr_camera *c, dr_view **out) {
    if (!s || !c || !out) return fail(DR_ERROR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    dr_view *v = new (std::nothrow) dr_view();
    if (!v) return fail(DR_ERROR_OUT_OF_MEMORY, "host allocation failed");
    v->scene = s;
    v->cam = drh::make_camera(*c);
    s->views.push_back(v);
    dr_status st = refresh_views(s, nullptr);
    if (st != DR_OK) {
        dr_view_destroy(v);
        return st;
    }
    *out = v;
    return DR_OK;
}

void dr_view_destroy(dr_view *v) {
    if (!v) return;
    if (v->scene) {
        auto &vs = v->scene->views;
        vs.erase(std::remove(vs.begin(), vs.end(), v), vs.end());
    }
    for (int i = 0; i < 3; i++)
        if (v->bufs[i]) (void)hipFree(v->bufs[i]);
    delete v;
}

dr_status dr_batch_create(dr_batch **out) {
    if (!out) return fail(DR_ERROR_INVALID_ARGUMENT, "out is NULL");
    *out = new (std::nothrow) dr_batch();
    if (!*out) return fail(DR_ERROR_OUT_OF_MEMORY, "host allocation failed");
    return DR_OK;
}

void dr_batch_destroy(dr_batch *b) {
    if (!b) return;
    free_workspace(b->ws);
    free_timing(b->tm);
    delete b;
}

dr_status dr_render_batch(dr_batch *b, const dr_job *jobs, int num_jobs, int flags, void *stream) {
    if (!b) return fail(DR_ERROR_INVALID_ARGUMENT, "null batch");
    Plan P;
    dr_status st = build_plan(jobs, num_jobs, b->tune, P);
    if (st != DR_OK) return st;
    return run_plan(P, b->ws, &b->tm, (hipStream_t)stream, (flags & DR_BATCH_GRAPH) != 0);
}

dr_status dr_batch_set_tuning(dr_batch *b, const dr_tuning *t) {
    if (!b || !t || t->pass_samples <= 0 || t->min_threads <= 0 || t->select_refill < 0 || t->select_refill > 64)
        return fail(DR_ERROR_INVALID_ARGUMENT, "bad tuning");
    b->tune.pass_samples = t->pass_samples;
    b->tune.min_threads = t->min_threads;
    b->tune.select_refill = t->select_refill;
    return DR_OK;
}

dr_status dr_batch_autotune(dr_batch *b, const dr_job *jobs, int num_jobs, void *stream, dr_tuning *out) {
    if (!b) return fail(DR_ERROR_INVALID_ARGUMENT, "null batch");
    hipStream_t hs = (hipStream_t)stream;
    static const long long kPass[] = {1ll << 20, 1ll << 21, 1ll << 22, 1ll << 23};
    static const long long kThreads[] = {1ll << 19, 1ll << 20, 1ll << 21};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return fail(DR_ERROR_DEVICE, "event creation failed");
    Plan P;
    dr_status st = build_plan(jobs, num_jobs, b->tune, P);
    if (st == DR_OK) st = run_plan(P, b->ws, nullptr, hs, false);  // warm-up: buffers, code objects
    double best = 1e300;
    Tuning bt = b->tune;
    int tried = 0;
    for (long long ps : kPass) {
        for (long long mt : kThreads) {
            if (st != DR_OK) break;
            Tuning t;
            t.pass_samples = ps;
            t.min_threads = mt;
            st = build_plan(jobs, num_jobs, t, P);
            if (st != DR_OK) break;
            (void)hipEventRecord(e0, hs);
            st = run_plan(P, b->ws, nullptr, hs, false);
            (void)hipEventRecord(e1, hs);
            if (st != DR_OK || hipEventSynchronize(e1) != hipSuccess) break;
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            tried++;
            if (ms < best) {
                best = ms;
                bt = t;
            }
        }
    }
    // then the refill threshold of the edge selection, at the chosen pass sizes
    static const int kRefill[] = {8, 16, 32, 0};
    if (has_edges(P)) {
        for (int rf : kRefill) {
            if (st != DR_OK) break;
            Tuning t = bt;
            t.select_refill = rf;
            st = build_plan(jobs, num_jobs, t, P);
            if (st != DR_OK) break;
            (void)hipEventRecord(e0, hs);
            st = run_plan(P, b->ws, nullptr, hs, false);
            (void)hipEventRecord(e1, hs);
            if (st != DR_OK || hipEventSynchronize(e1) != hipSuccess) break;
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            tried++;
            if (ms < best) {
                best = ms;
                bt = t;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != DR_OK) return st;
    b->tune = bt;
    if (out) {
        out->pass_samples = bt.pass_samples;
        out->min_threads = bt.min_threads;
        out->select_refill = bt.select_refill;
        out->best_ms = best;
        out->candidates = tried;
    }
    return DR_OK;
}

dr_status dr_intersect(dr_scene *s, const float *rays, int64_t n, int32_t *hits, int any_hit, void *stream) {
    if (!s || (!rays && n > 0) || (!hits && n > 0) || n < 0) return fail(DR_ERROR_INVALID_ARGUMENT, "bad arguments");
    if (n == 0) return DR_OK;
    hipStream_t hs = (hipStream_t)stream;
    k_intersect<<<(unsigned)((n + 255) / 256), 256, 0, hs>>>(s->view, rays, n, hits, any_hit);
    return launch_check("k_intersect");
}

dr_status dr_test_tree_stack(int64_t ops, uint64_t seed, int64_t *out, void *stream) {
    if (ops <= 0 || !out) return fail(DR_ERROR_INVALID_ARGUMENT, "bad arguments");
    hipStream_t hs = (hipStream_t)stream;
    unsigned long long *dev = nullptr;
    if (hipMalloc(&dev, 2 * sizeof(unsigned long long)) != hipSuccess)
        return fail(DR_ERROR_OUT_OF_MEMORY, "hipMalloc");
    k_fill<<<1, 64, 0, hs>>>((float *)dev, 4, 0.f);
    k_tree_stack<<<64, DR_BLOCK, 0, hs>>>(ops, seed, dev);
    unsigned long long r[2];
    hipError_t e = hipMemcpyAsync(r, dev, sizeof(r), hipMemcpyDeviceToHost, hs);
    if (e == hipSuccess) e = hipStreamSynchronize(hs);
    (void)hipFree(dev);
    if (e != hipSuccess) return fail(DR_ERROR_DEVICE, hipGetErrorString(e));
    out[0] = (int64_t)r[0];
    out[1] = (int64_t)r[1];
    return launch_check("k_tree_stack");
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
    (void)hipFree(dev);
    if (e != hipSuccess) return fail(DR_ERROR_DEVICE, hipGetErrorString(e));
    for (int k = 0; k < 4; k++) {
        double mean = sums[2 * k] / (double)num_samples;
        out[k] = mean;
        out[4 + k] = sums[2 * k + 1] / (double)num_samples - mean * mean;
    }
    return launch_check("k_line_integral");
}

}  // extern "C"
