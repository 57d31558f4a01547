// Per-sample device functions of the kernels in render.hip: draws of a
// pixel sample, forward radiance, the adjoint path record and its reverse
// sweep, and the secondary-edge loop of one pixel sample.
#pragma once

#include "../../include/delta_ray.h"
#include "dr_adjoint.h"

// ------------------------------------------------------------ sample API --
struct SampleCtx {
    uint64_t seed;
    int W, H, spp, nb;
    uint32_t flags;
};

DR_INLINE void pixel_sample_draws(const SampleCtx &c, int xi, int yi, int s, Philox &rng, F2 &screen,
                                  LightSample &ls) {
    uint64_t sid = ((uint64_t)yi * c.W + xi) * (uint64_t)c.spp + s;
    philox_init(rng, c.seed, 0, sid, DR_STREAM_MAIN);
    float ux = philox_next(rng), uy = philox_next(rng);
    float sx = (c.flags & DR_QUIRK_JITTER_OVER_SPP) ? ux / c.spp : ux;
    float sy = (c.flags & DR_QUIRK_JITTER_OVER_SPP) ? uy / c.spp : uy;
    screen = F2{float((xi + sx) / c.W), float((yi + sy) / c.H)};
    ls = draw_light_sample(rng);
}

// forward radiance of one pixel sample, secondary samples drawn in order
DR_INLINE F3 forward_sample(const DrSceneView &S, const SampleCtx &c, int xi, int yi, int s) {
    Philox rng;
    F2 scr;
    LightSample ls;
    pixel_sample_draws(c, xi, yi, s, rng, scr, ls);
    F3 org, dir;
    if (!sample_primary(S.cam, scr, org, dir)) return F3{0.f, 0.f, 0.f};
    SPoint sp;
    Isect is = intersect(S, org, dir, sp);
    if (!valid(is)) return F3{0.f, 0.f, 0.f};
    F3 wi = -dir;
    F3 color = F3{0.f, 0.f, 0.f};
    int plid = S.shapes[is.shape].light;
    if (plid >= 0 && dot(sp.fr.n, wi) > 0.f) color += F3{S.lights[plid].I[0], S.lights[plid].I[1], S.lights[plid].I[2]};
    F3 thr = F3{1.f, 1.f, 1.f};
    F3 cur_wi = wi;
    Isect cur = is;
    SPoint csp = sp;
    for (int b = 0; b < c.nb; b++) {
        SecondarySample ss = draw_secondary(rng);
        color += thr * nee(S, ls, wi, cur, csp);
        SPoint nsp;
        bool spec;
        Isect nis = sample_bsdf_scene(S, ss, cur_wi, cur, csp, nsp, spec);
        if (!valid(nis)) break;
        F3 wo = normalize(nsp.pos - csp.pos);
        int mid = S.shapes[cur.shape].mat;
        float pdf = bsdf_pdf(S, mid, csp, cur_wi, wo);
        if (pdf <= 0.f) break;
        F3 be = bsdf(S, mid, csp, cur_wi, wo);
        if (sum3(be) <= 0.f) break;
        int nlid = S.shapes[nis.shape].light;
        if (nlid >= 0 && dot(nsp.fr.n, -wo) > 0.f) {
            F3 contrib = be * F3{S.lights[nlid].I[0], S.lights[nlid].I[1], S.lights[nlid].I[2]};
            float lpdf = light_pdf(S, nis) * length_sq(nsp.pos - csp.pos) / fabsf(dot(wo, nsp.ng));
            float mis = sq(pdf) / (sq(lpdf) + sq(pdf));
            color += thr * contrib * (mis / pdf);
        }
        thr = thr * (be / pdf);
        if (b >= 5) {
            float q = fminf(max3(thr), 0.95f);
            if (ss.rr >= q) break;
            thr = thr / q;
        }
        cur_wi = -wo;
        cur = nis;
        csp = nsp;
    }
    return color * (1.f / c.spp);
}

// ------------------------------------------------------- adjoint record --
struct BRec {
    F3 cur_wi, thr, nee;
    F3 sdir;          // sampled direction (ray of the next surface point)
    F3 wo, be;
    float pdf, mis, q, b0, b1;
    int stage;        // 0: stopped after NEE, 1: full bounce
    int emis_light;   // light hit by the sampled ray with MIS emission, -1
    bool spec, flipped;
};

// adjoint of the primary surface point and of the camera ray
// (reference: adjoints of cam_to_world read after the sweep,
// delta_ray.cpp:770-781)
DR_INLINE void primary_adjoint(const DrSceneView &S, const GradSink &G, const Isect &is0, F3 cam_org, F3 cam_dir,
                               F3 cam_nd, const AdjSP &dsp0, F3 d_wi) {
    TriData td = load_tri(S, is0);
    F3 d_org = F3{0.f, 0.f, 0.f}, d_dir = F3{0.f, 0.f, 0.f};
    F3 dv[3] = {F3{0.f, 0.f, 0.f}, F3{0.f, 0.f, 0.f}, F3{0.f, 0.f, 0.f}};
    F3 dn[3] = {F3{0.f, 0.f, 0.f}, F3{0.f, 0.f, 0.f}, F3{0.f, 0.f, 0.f}};
    surface_point_adj(td, cam_org, cam_dir, dsp0, d_org, d_dir, dv, dn);
    scatter_tri(S, G, td, is0.shape, dv, dn);
    d_dir -= d_wi;
    const float *m = S.cam.c2w;
    float iw = 1.f / m[15];
    G.cam_add(3, d_org.x * iw);
    G.cam_add(7, d_org.y * iw);
    G.cam_add(11, d_org.z * iw);
    G.cam_add(15, -(d_org.x * m[3] + d_org.y * m[7] + d_org.z * m[11]) * iw * iw);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) G.cam_add(i * 4 + j, d_dir[i] * cam_nd[j]);
}

// Reverse sweep over the recorded path.  color_mode: seeds d_c on the colour
// (everything scattered); otherwise seeds dp at the shading-point positions
// (vertices and camera only, reference delta_ray.cpp:968-1002).
template <int MAXB>
DR_INLINE void reverse_path(const DrSceneView &S, const GradSink &G, const LightSample &ls, F3 cam_org, F3 cam_dir,
                            F3 cam_nd, const Isect *is, const SPoint *sp, const BRec *R, int nused, const bool *chain,
                            F3 d_c, const F3 *dp_seed, int primary_light) {
    AdjSP dsp[MAXB + 1];
    F3 d_cw[MAXB + 1];
    for (int b = 0; b <= MAXB; b++) {
        dsp[b] = adj_zero();
        d_cw[b] = F3{0.f, 0.f, 0.f};
    }
    F3 d_wi = F3{0.f, 0.f, 0.f};
    F3 d_thr_next = F3{0.f, 0.f, 0.f};
    const bool cm = G.color_mode;
    for (int b = nused - 1; b >= 0; b--) {
        const BRec &r = R[b];
        int mid = S.shapes[is[b].shape].mat;
        if (!cm) dsp[b].pos += dp_seed[b];
        F3 d_thr = F3{0.f, 0.f, 0.f};
        // a full bounce (stage 1) has a next surface point and a wo
        bool has_next = r.stage == 1;
        if (r.stage == 1) {
            F3 d_wo = F3{0.f, 0.f, 0.f};
            if (cm) {
                float ip = 1.f / r.pdf;
                F3 d_be = F3{0.f, 0.f, 0.f};
                float d_pdf = 0.f;
                F3 dtp = d_thr_next * (1.f / r.q);
                F3 rr = r.be * ip;
                d_thr += dtp * rr;
                F3 d_rr = dtp * r.thr;
                d_be += d_rr * ip;
                d_pdf += -dot(d_rr, r.be) * ip * ip;
                if (r.emis_light >= 0) {
                    const DrLight &l = S.lights[r.emis_light];
                    F3 I = F3{l.I[0], l.I[1], l.I[2]};
                    F3 contrib = r.be * I;
                    float k = r.mis / r.pdf;
                    d_thr += d_c * contrib * k;
                    F3 d_contrib = d_c * r.thr * k;
                    float d_k = dot(d_c, r.thr * contrib);
                    d_pdf += -d_k * r.mis / (r.pdf * r.pdf);
                    d_be += d_contrib * I;
                    G.add3(l.grad_off, d_contrib * r.be);
                }
                bsdf_adj(S, G, mid, sp[b], r.cur_wi, r.wo, d_be, dsp[b], d_cw[b], d_wo);
                bsdf_pdf_adj(S, G, mid, sp[b], r.cur_wi, r.wo, d_pdf, dsp[b], d_cw[b], d_wo);
            }
            if (b + 1 < nused) d_wo -= d_cw[b + 1];
            if (chain[b + 1] && has_next) {
                F3 dv = d_normalize(sp[b + 1].pos - sp[b].pos, d_wo);
                dsp[b + 1].pos += dv;
                dsp[b].pos -= dv;
            }
        }
        if (cm) {
            d_thr += d_c * r.nee;
            nee_adj(S, G, ls, cam_dir * -1.f, is[b], sp[b], d_c * r.thr, dsp[b], d_wi);
        }
        d_thr_next = d_thr;
        // the next surface point is complete: propagate through it
        if (chain[b + 1] && has_next) {
            TriData td = load_tri(S, is[b + 1]);
            F3 d_org = F3{0.f, 0.f, 0.f}, d_dir = F3{0.f, 0.f, 0.f};
            F3 dv[3] = {F3{0.f, 0.f, 0.f}, F3{0.f, 0.f, 0.f}, F3{0.f, 0.f, 0.f}};
            F3 dn[3] = {F3{0.f, 0.f, 0.f}, F3{0.f, 0.f, 0.f}, F3{0.f, 0.f, 0.f}};
            surface_point_adj(td, sp[b].pos, r.sdir, dsp[b + 1], d_org, d_dir, dv, dn);
            scatter_tri(S, G, td, is[b + 1].shape, dv, dn);
            dsp[b].pos += d_org;
            SPoint spf = sp[b];
            if (r.flipped) spf.fr = neg_frame(spf.fr);
            AdjSP df = adj_zero();
            sample_dir_adj(S, G, mid, spf, r.cur_wi, r.b0, r.b1, d_dir, df, d_cw[b]);
            float sg = r.flipped ? -1.f : 1.f;
            dsp[b].fr.x += df.fr.x * sg;
            dsp[b].fr.y += df.fr.y * sg;
            dsp[b].fr.n += df.fr.n * sg;
            dsp[b].uv = dsp[b].uv + df.uv;
        }
    }
    d_wi += d_cw[0];
    primary_adjoint(S, G, is[0], cam_org, cam_dir, cam_nd, dsp[0], d_wi);
    if (cm && primary_light >= 0) G.add3(S.lights[primary_light].grad_off, d_c);
}

// Forward path of one pixel sample with the bounded per-bounce record the
// reverse sweep needs (reference render_sample, delta_ray.cpp:694-766).
template <int MAXB>
struct PathRec {
    LightSample ls;
    F3 org, dir, nd;
    Isect is[MAXB + 1];
    SPoint sp[MAXB + 1];
    BRec R[MAXB];
    bool chain[MAXB + 1];
    int nused;
    int primary_light;
    F3 color;  // not yet divided by spp
};

template <int MAXB>
DR_INLINE bool trace_path(const DrSceneView &S, const SampleCtx &c, int xi, int yi, int s, PathRec<MAXB> &P) {
    Philox rng;
    F2 scr;
    pixel_sample_draws(c, xi, yi, s, rng, scr, P.ls);
    SecondarySample ss[MAXB];
    for (int b = 0; b < c.nb; b++) ss[b] = draw_secondary(rng);
    bool has_ray = sample_primary(S.cam, scr, P.org, P.dir, &P.nd);
    P.nused = 0;
    P.primary_light = -1;
    P.color = F3{0.f, 0.f, 0.f};
    if (!has_ray) return false;
    // the differentiated path computes surface points as the tape does
    P.is[0] = intersect(S, P.org, P.dir, P.sp[0], true);
    if (!valid(P.is[0])) return false;
    F3 wi = -P.dir;
    int plid = S.shapes[P.is[0].shape].light;
    if (plid >= 0 && dot(P.sp[0].fr.n, wi) > 0.f) {
        P.color += F3{S.lights[plid].I[0], S.lights[plid].I[1], S.lights[plid].I[2]};
        P.primary_light = plid;
    }
    F3 thr = F3{1.f, 1.f, 1.f};
    F3 cur_wi = wi;
    P.chain[0] = true;
    for (int b = 0; b < c.nb; b++) {
        BRec &r = P.R[b];
        P.nused = b + 1;
        r.cur_wi = cur_wi;
        r.thr = thr;
        r.nee = nee(S, P.ls, wi, P.is[b], P.sp[b]);
        P.color += thr * r.nee;
        r.stage = 0;
        r.emis_light = -1;
        r.q = 1.f;
        r.b0 = ss[b].b0;
        r.b1 = ss[b].b1;
        P.chain[b + 1] = false;
        int mid = S.shapes[P.is[b].shape].mat;
        SPoint spf = P.sp[b];
        r.flipped = S.materials[mid].two_sided && dot(cur_wi, spf.fr.n) < 0.f;
        if (r.flipped) {
            spf.fr = neg_frame(spf.fr);
            spf.ng = -spf.ng;
        }
        bool spec;
        r.sdir = sample_bsdf_dir(S, mid, spf, cur_wi, ss[b].b0, ss[b].b1, ss[b].bw, spec);
        r.spec = spec;
        if (length(r.sdir) <= 0.f) {
            r.spec = false;
            break;
        }
        P.is[b + 1] = intersect(S, P.sp[b].pos, r.sdir, P.sp[b + 1], true);
        P.chain[b + 1] = P.chain[b] && r.spec;
        if (!valid(P.is[b + 1])) break;
        F3 wo = normalize(P.sp[b + 1].pos - P.sp[b].pos);
        float pdf = bsdf_pdf(S, mid, P.sp[b], cur_wi, wo);
        if (pdf <= 0.f) break;
        F3 be = bsdf(S, mid, P.sp[b], cur_wi, wo);
        if (sum3(be) <= 0.f) break;
        r.stage = 1;
        r.wo = wo;
        r.pdf = pdf;
        r.be = be;
        int nlid = S.shapes[P.is[b + 1].shape].light;
        if (nlid >= 0 && dot(P.sp[b + 1].fr.n, -wo) > 0.f) {
            F3 contrib = be * F3{S.lights[nlid].I[0], S.lights[nlid].I[1], S.lights[nlid].I[2]};
            float lpdf =
                light_pdf(S, P.is[b + 1]) * length_sq(P.sp[b + 1].pos - P.sp[b].pos) / fabsf(dot(wo, P.sp[b + 1].ng));
            float mis = sq(pdf) / (sq(lpdf) + sq(pdf));
            P.color += thr * contrib * (mis / pdf);
            r.emis_light = nlid;
            r.mis = mis;
        }
        thr = thr * (be / pdf);
        if (b >= 5) {
            float q = fminf(max3(thr), 0.95f);
            r.q = q;
            if (ss[b].rr >= q) break;
            thr = thr / q;
        }
        cur_wi = -wo;
    }
    return true;
}

// A secondary shading point (reference SecondaryPoint, pathtrace.h:170-176):
// recorded for every bounce reached without a diffuse sample; the edge
// stages read it.  valid = 0 marks an empty slot.
// Edge-tree cone toward the light chosen by the primary light sample
// (reference delta_ray.cpp:876-890)
DR_INLINE Cone light_cone(const DrSceneView &S, float light_sel, F3 p) {
    int lid = sample_discrete(S.lights_cdf, S.num_lights, light_sel);
    const DrLight &L = S.lights[lid];
    F3 dl = F3{L.sphere_c[0], L.sphere_c[1], L.sphere_c[2]} - p;
    float tan_a = L.sphere_r / length(dl);
    float cos_a = tan_a <= 1.f ? 1.f / sqrtf(tan_a * tan_a + 1.f) : 0.f;
    return Cone{p, normalize(dl), cos_a};
}

DR_INLINE uint64_t sample_id(const SampleCtx &c, int xi, int yi, int s) {
    return ((uint64_t)yi * c.W + xi) * (uint64_t)c.spp + s;
}

// One secondary shading point, written by the path stage.  Besides the
// point itself it carries what starting its edge selection needs, computed
// once here (where all lanes of a wave do it together) instead of in the
// selection kernel: the sample's ids, the cone toward the light of the
// primary light sample and the material's luminances and roughness.
struct PointRec {
    uint64_t sid;        // sample id ((y * W + x) * spp + s)
    SPoint sp;
    F3 wi, thr;
    F3 cone_dir;
    float cone_cos;
    int shape, tri, gtri, valid;
    float roughness, dlum, slum;
    int job, b, pix;
};

// Stage A: forward path + interior adjoint of one pixel sample; fills the
// MAXB point slots of the sample.  Returns the colour / spp.
template <int MAXB>
DR_INLINE F3 stage_path(const DrSceneView &S, const SampleCtx &c, GradSink &G, int xi, int yi, int s, F3 d_color,
                        PointRec *pts, int job = 0) {
    PathRec<MAXB> P;
    for (int b = 0; b < MAXB; b++) pts[b].valid = 0;
    if (!trace_path<MAXB>(S, c, xi, yi, s, P)) return F3{0.f, 0.f, 0.f};
    float sw = 1.f / c.spp;
    if (!(c.flags & DR_SKIP_INTERIOR)) {
        G.color_mode = true;
        reverse_path<MAXB>(S, G, P.ls, P.org, P.dir, P.nd, P.is, P.sp, P.R, P.nused, P.chain, d_color * sw, nullptr,
                           P.primary_light);
    }
    if (!(c.flags & DR_SKIP_SECONDARY_EDGES)) {
        for (int b = 0; b < P.nused; b++) {
            if (!P.chain[b]) continue;
            PointRec &q = pts[b];
            q.sid = sample_id(c, xi, yi, s);
            q.sp = P.sp[b];
            q.wi = P.R[b].cur_wi;
            q.thr = P.R[b].thr;
            q.shape = P.is[b].shape;
            q.tri = P.is[b].tri;
            q.gtri = P.is[b].gtri;
            q.valid = 1;
            Cone cn = light_cone(S, P.ls.light_sel, q.sp.pos);
            q.cone_dir = cn.dir;
            q.cone_cos = cn.cos_angle;
            const DrMaterial &m = S.materials[S.shapes[q.shape].mat];
            q.roughness = get_roughness(S, m, q.sp.uv);
            q.dlum = luminance(get_diffuse(S, m, q.sp.uv));
            q.slum = luminance(get_specular(S, m, q.sp.uv));
            q.job = job;
            q.b = b;
            q.pix = yi * c.W + xi;
        }
    }
    return P.color * sw;
}

// The random streams, light cone and tree choice of one point's selection
// (the light cone and material values come with the point record); false
// when the point selects nothing.
template <class Stack>
DR_INLINE bool select_begin(const DrSceneView &S, const SampleCtx &c, const PointRec &p, EdgeTraversal &T,
                            Stack &stack) {
    philox_init(T.rc, c.seed, 0, p.sid, dr_stream_point(p.b));
    philox_next(T.rc);  // bsdf component of the point's edge sample
    philox_next(T.rc);  // position on the edge
    philox_init(T.rt, c.seed, 0, p.sid, dr_stream_point_tree(p.b));
    int mode = 0;
    if (c.flags & DR_LEARNED_EDGES) {
        Philox mix;
        philox_init(mix, c.seed, 0, p.sid, dr_stream_point_mix(p.b));
        mode = philox_next(mix) < DR_LEARNED_SHARE ? 1 : 2;
    }
    T.cone_pmf_quirk = (c.flags & DR_QUIRK_CONE_PMF) != 0;
    return et_init(T, stack, S, p.wi, p.sp, Cone{p.sp.pos, p.cone_dir, p.cone_cos}, p.roughness, p.dlum, p.slum,
                   mode);
}

// Stage C: one selected edge (reference delta_ray.cpp:904-966).  Scatters
// the vertex derivatives of the edge and returns dp (the Dirac derivative at
// the shading point).
template <int MAXB>
DR_INLINE F3 stage_edge(const DrSceneView &S, const SampleCtx &c, GradSink &G, uint64_t sid, int b,
                        const PointRec &p, const DrEdge &edge, float w, int j, F3 d_color, float *mag = nullptr,
                        bool *defer = nullptr, const float *geo = nullptr) {
    Philox prng;
    philox_init(prng, c.seed, 0, sid, dr_stream_point(b));
    float bcomp = philox_next(prng), eu = philox_next(prng);
    Philox erng;
    philox_init(erng, c.seed, 0, sid, dr_stream_edge(b, j));
    LightSample els = draw_light_sample(erng);
    SecondarySample ess[MAXB];
    int nss = c.nb - b - 1;
    for (int i = 0; i < nss; i++) ess[i] = draw_secondary(erng);
    F3 weight = p.thr * ((1.f / c.spp) * w);
    Isect is = Isect{p.shape, p.tri, p.gtri};
    EdgeResult er = sample_shading_edge_point(S, bcomp, eu, edge, els, ess, nss, p.wi, is, p.sp, weight, d_color,
                                              c.nb == 1, defer, geo, c.flags);
    if (defer && *defer) return F3{0.f, 0.f, 0.f};
    if (!is_zero(er.dv0) || !is_zero(er.dv1)) {
        const DrShape &es = S.shapes[edge.shape];
        G.add3(es.vgrad_off + 3ll * edge.v0, er.dv0);
        G.add3(es.vgrad_off + 3ll * edge.v1, er.dv1);
    }
    if (mag) {
        // the sample's contribution to the gradients the statistics measure:
        // the edge's vertex adjoints and the Dirac derivative at the shading
        // point (which k_dp hands to that point's vertices)
        float se = S.stat_shape ? S.stat_shape[edge.shape] : 1.f;
        float spt = S.stat_shape ? S.stat_shape[p.shape] : 1.f;
        *mag = se * (length(er.dv0) + length(er.dv1)) + spt * length(er.dp);
    }
    return er.dp;
}

// Stage D: propagate the Dirac position derivatives of a pixel sample's
// secondary points to vertices and camera (reference delta_ray.cpp:968-1002).
// A bounce-0 point depends only on the primary ray, rebuilt from the
// sample's draws; deeper (specular-chain) points replay the path.
template <int MAXB>
DR_INLINE void stage_dp(const DrSceneView &S, const SampleCtx &c, GradSink &G, int xi, int yi, int s,
                        const PointRec *pts, const F3 *dps) {
    bool deep = false, any = false;
    for (int b = 0; b < MAXB; b++) {
        if (!is_zero(dps[b])) {
            any = true;
            if (b > 0) deep = true;
        }
    }
    if (!any) return;
    G.color_mode = false;
    if (!deep) {
        Philox rng;
        F2 scr;
        LightSample ls;
        pixel_sample_draws(c, xi, yi, s, rng, scr, ls);
        F3 org, dir, nd;
        sample_primary(S.cam, scr, org, dir, &nd);
        AdjSP a = adj_zero();
        a.pos = dps[0];
        primary_adjoint(S, G, Isect{pts[0].shape, pts[0].tri, pts[0].gtri}, org, dir, nd, a, F3{0.f, 0.f, 0.f});
        return;
    }
    PathRec<MAXB> P;
    if (!trace_path<MAXB>(S, c, xi, yi, s, P)) return;
    reverse_path<MAXB>(S, G, P.ls, P.org, P.dir, P.nd, P.is, P.sp, P.R, P.nused, P.chain, F3{0.f, 0.f, 0.f}, dps,
                       -1);
}
