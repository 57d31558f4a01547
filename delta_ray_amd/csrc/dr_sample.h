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
    sample_primary(S.cam, scr, org, dir);
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
    // primary surface point
    TriData td = load_tri(S, is[0]);
    F3 d_org = F3{0.f, 0.f, 0.f}, d_dir = F3{0.f, 0.f, 0.f};
    F3 dv[3] = {F3{0.f, 0.f, 0.f}, F3{0.f, 0.f, 0.f}, F3{0.f, 0.f, 0.f}};
    F3 dn[3] = {F3{0.f, 0.f, 0.f}, F3{0.f, 0.f, 0.f}, F3{0.f, 0.f, 0.f}};
    surface_point_adj(td, cam_org, cam_dir, dsp[0], d_org, d_dir, dv, dn);
    scatter_tri(S, G, td, is[0].shape, dv, dn);
    d_dir -= d_wi;
    if (cm && primary_light >= 0) G.add3(S.lights[primary_light].grad_off, d_c);
    // camera (reference: adjoints of cam_to_world read after the sweep)
    const float *m = S.cam.c2w;
    float w = m[15];
    float iw = 1.f / w;
    G.cam_add(3, d_org.x * iw);
    G.cam_add(7, d_org.y * iw);
    G.cam_add(11, d_org.z * iw);
    G.cam_add(15, -(d_org.x * m[3] + d_org.y * m[7] + d_org.z * m[11]) * iw * iw);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) G.cam_add(i * 4 + j, d_dir[i] * cam_nd[j]);
}

template <int MAXB>
DR_INLINE F3 adjoint_sample(const DrSceneView &S, const SampleCtx &c, GradSink &G, int xi, int yi, int s, F3 d_color,
                            float &dx_count) {
    Philox rng;
    F2 scr;
    LightSample ls;
    pixel_sample_draws(c, xi, yi, s, rng, scr, ls);
    SecondarySample ss[MAXB];
    for (int b = 0; b < c.nb; b++) ss[b] = draw_secondary(rng);
    F3 org, dir, nd;
    sample_primary(S.cam, scr, org, dir, &nd);
    Isect is[MAXB + 1];
    SPoint sp[MAXB + 1];
    BRec R[MAXB];
    bool chain[MAXB + 1];
    // the differentiated path computes surface points as the tape does
    is[0] = intersect(S, org, dir, sp[0], true);
    if (!valid(is[0])) return F3{0.f, 0.f, 0.f};
    F3 wi = -dir;
    F3 color = F3{0.f, 0.f, 0.f};
    int plid = S.shapes[is[0].shape].light;
    int primary_light = -1;
    if (plid >= 0 && dot(sp[0].fr.n, wi) > 0.f) {
        color += F3{S.lights[plid].I[0], S.lights[plid].I[1], S.lights[plid].I[2]};
        primary_light = plid;
    }
    F3 thr = F3{1.f, 1.f, 1.f};
    F3 cur_wi = wi;
    chain[0] = true;
    int nused = 0;
    for (int b = 0; b < c.nb; b++) {
        BRec &r = R[b];
        nused = b + 1;
        r.cur_wi = cur_wi;
        r.thr = thr;
        r.nee = nee(S, ls, wi, is[b], sp[b]);
        color += thr * r.nee;
        r.stage = 0;
        r.emis_light = -1;
        r.q = 1.f;
        r.b0 = ss[b].b0;
        r.b1 = ss[b].b1;
        chain[b + 1] = false;
        int mid = S.shapes[is[b].shape].mat;
        SPoint spf = sp[b];
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
        is[b + 1] = intersect(S, sp[b].pos, r.sdir, sp[b + 1], true);
        chain[b + 1] = chain[b] && r.spec;
        if (!valid(is[b + 1])) break;
        F3 wo = normalize(sp[b + 1].pos - sp[b].pos);
        float pdf = bsdf_pdf(S, mid, sp[b], cur_wi, wo);
        if (pdf <= 0.f) break;
        F3 be = bsdf(S, mid, sp[b], cur_wi, wo);
        if (sum3(be) <= 0.f) break;
        r.stage = 1;
        r.wo = wo;
        r.pdf = pdf;
        r.be = be;
        int nlid = S.shapes[is[b + 1].shape].light;
        if (nlid >= 0 && dot(sp[b + 1].fr.n, -wo) > 0.f) {
            F3 contrib = be * F3{S.lights[nlid].I[0], S.lights[nlid].I[1], S.lights[nlid].I[2]};
            float lpdf = light_pdf(S, is[b + 1]) * length_sq(sp[b + 1].pos - sp[b].pos) / fabsf(dot(wo, sp[b + 1].ng));
            float mis = sq(pdf) / (sq(lpdf) + sq(pdf));
            color += thr * contrib * (mis / pdf);
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
    float sw = 1.f / c.spp;
    color = color * sw;
    if (!(c.flags & DR_SKIP_INTERIOR)) {
        G.color_mode = true;
        reverse_path<MAXB>(S, G, ls, org, dir, nd, is, sp, R, nused, chain, d_color * sw, nullptr,
                           primary_light);
    }
    // secondary edges (reference delta_ray.cpp:868-966)
    F3 dps[MAXB];
    bool any_dp = false;
    for (int b = 0; b < MAXB; b++) dps[b] = F3{0.f, 0.f, 0.f};
    int last_count = 0;
    if (!(c.flags & DR_SKIP_SECONDARY_EDGES)) {
        uint64_t sid = ((uint64_t)yi * c.W + xi) * (uint64_t)c.spp + s;
        for (int b = 0; b < nused; b++) {
            if (!chain[b]) continue;
            Philox prng;
            philox_init(prng, c.seed, 0, sid, dr_stream_point(b));
            float bcomp = philox_next(prng), eu = philox_next(prng);
            int lid = sample_discrete(S.lights_cdf, S.num_lights, ls.light_sel);
            const DrLight &L = S.lights[lid];
            F3 bc = F3{L.sphere_c[0], L.sphere_c[1], L.sphere_c[2]};
            F3 dl = bc - sp[b].pos;
            float tan_a = L.sphere_r / length(dl);
            float cos_a = tan_a <= 1.f ? 1.f / sqrtf(tan_a * tan_a + 1.f) : 0.f;
            Cone cone = Cone{sp[b].pos, normalize(dl), cos_a};
            int j = 0;
            const BRec &r = R[b];
            const SPoint &spb = sp[b];
            const Isect &isb = is[b];
            int nss = c.nb - b - 1;
            auto emit = [&](const DrEdge &edge, float w) {
                int jj = j++;
                if (w == 0.f) return;
                Philox erng;
                philox_init(erng, c.seed, 0, sid, dr_stream_edge(b, jj));
                LightSample els = draw_light_sample(erng);
                SecondarySample ess[MAXB];
                for (int i = 0; i < nss; i++) ess[i] = draw_secondary(erng);
                F3 weight = r.thr * (sw * w);
                EdgeResult er = sample_shading_edge_point(S, bcomp, eu, edge, els, ess, nss, r.cur_wi, isb, spb,
                                                          weight, d_color, c.nb == 1);
                if (!is_zero(er.dp)) {
                    if (c.flags & DR_QUIRK_DP_LAST) dps[b] = er.dp;
                    else dps[b] += er.dp;
                    any_dp = true;
                }
                if (!is_zero(er.dv0) || !is_zero(er.dv1)) {
                    const DrShape &es = S.shapes[edge.shape];
                    G.add3(es.vgrad_off + 3ll * edge.v0, er.dv0);
                    G.add3(es.vgrad_off + 3ll * edge.v1, er.dv1);
                }
            };
            sample_shading_edges(S, r.cur_wi, isb, spb, cone, prng, emit);
            last_count = j;
            if (S.stats) atomicAdd(&S.stats[3], (unsigned long long)j);
        }
    }
    dx_count += last_count / 1000.f;
    if (any_dp) {
        G.color_mode = false;
        reverse_path<MAXB>(S, G, ls, org, dir, nd, is, sp, R, nused, chain, F3{0.f, 0.f, 0.f}, dps,
                           -1);
    }
    return color;
}

