// Hand-written reverse-mode adjoint of the path tracer.  It replaces the
// reference's per-sample tape (autodiff.h, delta_ray.cpp:694-1002): the path
// is recomputed, a bounded per-bounce record is kept, and adjoints are swept
// back bounce by bounce.  What is and is not differentiated follows the
// reference exactly:
//   - MIS weights are constants (pathtrace.h:78-91, 269-279);
//   - the light's triangle pmf is a constant, 1/area is not (scene.h:102-107);
//   - after a diffuse (non-specular) BSDF sample the next surface point and
//     direction are constants (pathtrace.h:148-159, 245-247);
//   - next-event estimation at every bounce uses the primary wi and light
//     sample (pathtrace.h:213-216);
//   - uv scales, fov, aspect, screen position carry no gradient out.
#pragma once

#include "dr_edges.h"

// Workgroup gradient cache in LDS.  Neighbouring pixel samples of one
// workgroup hit the same triangles and texels, so their adjoints go to the
// same few hundred gradient entries: they are summed here with LDS atomics
// (open addressing, linear probing) and reach HBM as one global atomic per
// distinct entry when the workgroup ends.  Keys are (job, flat index), so one
// workgroup may serve several renders of a batch.  An adjoint that finds no
// free slot within DR_GC_PROBES goes straight to HBM.
#define DR_GC_PROBES 8
#define DR_GC_EMPTY 0xffffffffffffffffull
// every thread of the workgroup calls these, followed / preceded by a barrier
DR_INLINE void gc_init(unsigned long long *key, float *val, int slots) {
    for (int i = threadIdx.x; i < slots; i += blockDim.x) {
        key[i] = DR_GC_EMPTY;
        val[i] = 0.f;
    }
}
// grads_of(job) -> that job's gradient buffer
template <class GradsOf>
DR_INLINE void gc_flush(const unsigned long long *key, const float *val, int slots, GradsOf grads_of) {
    for (int i = threadIdx.x; i < slots; i += blockDim.x) {
        unsigned long long k = key[i];
        float v = val[i];
        if (k != DR_GC_EMPTY && v != 0.f) atomicAdd(grads_of((int)(k >> 32)) + (k & 0xffffffffull), v);
    }
}

// Gradient sink: where adjoints of scene parameters go.
struct GradSink {
    float *grads;      // flat buffer, include/delta_ray.h layout
    float *cam;        // 64 floats (LDS) for camera matrices
    uint32_t flags;
    bool color_mode;   // false: Dirac position sweep (vertices + camera only)
    unsigned long long *gkey = nullptr;  // LDS cache of the workgroup, or none
    float *gval = nullptr;
    int gmask = 0;
    unsigned job = 0;
    DR_INLINE void add(long long idx, float v) const {
        if (v == 0.f) return;
        if (gkey && idx < 0x100000000ll) {
            unsigned long long key = ((unsigned long long)job << 32) | (unsigned long long)idx;
            uint32_t h = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 40);
            for (int p = 0; p < DR_GC_PROBES; p++) {
                int slot = (int)((h + p) & (uint32_t)gmask);
                unsigned long long old = atomicCAS(&gkey[slot], DR_GC_EMPTY, key);
                if (old == DR_GC_EMPTY || old == key) {
                    atomicAdd(&gval[slot], v);
                    return;
                }
            }
        }
        atomicAdd(grads + idx, v);
    }
    DR_INLINE void add3(long long idx, F3 v) const {
        add(idx, v.x); add(idx + 1, v.y); add(idx + 2, v.z);
    }
    DR_INLINE void cam_add(int idx, float v) const {
        if (v != 0.f) atomicAdd(cam + idx, v);
    }
};

// adjoint of the surface point of a triangle
struct AdjSP {
    F3 pos;
    Frame fr;
    F2 uv;
};
DR_INLINE AdjSP adj_zero() {
    AdjSP a;
    a.pos = F3{0.f, 0.f, 0.f};
    a.fr = Frame{F3{0.f, 0.f, 0.f}, F3{0.f, 0.f, 0.f}, F3{0.f, 0.f, 0.f}};
    a.uv = F2{0.f, 0.f};
    return a;
}

// ------------------------------------------------------------ textures ----
// adjoint of a bilinear lookup: texel adjoints go to the sink, the uv
// adjoint is returned (in the unscaled uv)
DR_INLINE F2 tex_adj(const DrSceneView &S, const GradSink &G, const DrTex &t, const float *uvscale, F2 uv,
                     const float *d_val, int C) {
    if (t.h == 0) {
        if (G.color_mode) {
            for (int k = 0; k < C; k++) {
                float d = d_val[k];
                // constant roughness: the reference drops non-finite adjoints
                if (C == 1 && !isfinite(d)) continue;
                G.add(t.grad_off + k, d);
            }
        }
        return F2{0.f, 0.f};
    }
    F2 suv = F2{uv.x * uvscale[0], uv.y * uvscale[1]};
    TexFetch f = tex_locate(t, suv);
    const float *p = S.texels + t.off;
    float wff = (1.f - f.u) * (1.f - f.v), wfc = (1.f - f.u) * f.v, wcf = f.u * (1.f - f.v), wcc = f.u * f.v;
    int iff = f.yf * t.w + f.xf, ifc = f.yc * t.w + f.xf, icf = f.yf * t.w + f.xc, icc = f.yc * t.w + f.xc;
    float du = 0.f, dv = 0.f;
    for (int k = 0; k < C; k++) {
        float d = d_val[k];
        float vff = p[iff * C + k], vfc = p[ifc * C + k], vcf = p[icf * C + k], vcc = p[icc * C + k];
        du += d * (-(1.f - f.v) * vff - f.v * vfc + (1.f - f.v) * vcf + f.v * vcc);
        dv += d * (-(1.f - f.u) * vff + (1.f - f.u) * vfc - f.u * vcf + f.u * vcc);
    }
    if (G.color_mode) {
        // texel targets; the reference swaps the (f,f) and (c,c) corners
        int tff = iff, tcc = icc;
        if (G.flags & DR_QUIRK_TEXEL_SWAP) { tff = icc; tcc = iff; }
        for (int k = 0; k < C; k++) {
            float d = d_val[k];
            G.add(t.grad_off + (long long)tff * C + k, d * wff);
            G.add(t.grad_off + (long long)ifc * C + k, d * wfc);
            G.add(t.grad_off + (long long)icf * C + k, d * wcf);
            G.add(t.grad_off + (long long)tcc * C + k, d * wcc);
        }
    }
    return F2{du * (float)t.w * uvscale[0], dv * (float)t.h * uvscale[1]};
}
DR_INLINE F2 tex3_adj(const DrSceneView &S, const GradSink &G, const DrTex &t, const float *sc, F2 uv, F3 d) {
    float dv[3] = {d.x, d.y, d.z};
    return tex_adj(S, G, t, sc, uv, dv, 3);
}
DR_INLINE F2 tex1_adj(const DrSceneView &S, const GradSink &G, const DrTex &t, const float *sc, F2 uv, float d) {
    return tex_adj(S, G, t, sc, uv, &d, 1);
}

// ------------------------------------------------------------ material ----
// Blinn-Phong normal-distribution term and its partials
DR_INLINE float phong_D(float hz, float e) { return powf(hz, e) * (e + 2.f) / (2.f * DR_PI); }
DR_INLINE void phong_D_adj(float hz, float e, float dD, float &d_hz, float &d_e) {
    float P = powf(hz, e);
    float inv2pi = 1.f / (2.f * DR_PI);
    float dP = dD * (e + 2.f) * inv2pi;
    d_e += dD * P * inv2pi + dP * P * logf(fmaxf(hz, 1e-6f));
    d_hz += dP * e * powf(hz, e - 1.f);
}
// e = max(2/r, 0)
DR_INLINE float d_phong_dr(float r) { return 2.f / r > 0.f ? -2.f / (r * r) : 0.f; }

// adjoint of smith_g1 (dr_device.h)
DR_INLINE void smith_g1_adj(F3 v, F3 h, F3 n, const Frame &fr, float r, float dG, F3 &d_v, F3 &d_n, Frame &d_fr,
                            float &d_r) {
    float c = dot(v, n);
    if (dot(v, h) * c <= 0.f) return;
    float x = 1.f / (c * c) - 1.f;
    if (!(x >= 0.f)) return;
    float tan_t = sqrtf(x);
    if (tan_t == 0.f) return;
    F3 vl = to_local(fr, v);
    float cz = vl.z;
    float s2 = 1.f - cz * cz;
    float alpha, inv = 0.f, cp2 = 0.f, sp2 = 0.f;
    if (s2 <= 0.f) {
        alpha = sqrtf(r);
    } else {
        inv = 1.f / s2;
        cp2 = vl.x * vl.x * inv;
        sp2 = vl.y * vl.y * inv;
        alpha = sqrtf(cp2 * r + sp2 * r);
    }
    float q = alpha * tan_t;
    float a = 1.f / q;
    if (a >= 1.6f) return;
    float num = 3.535f * a + 2.181f * a * a;
    float den = 1.0f + 2.276f * a + 2.577f * a * a;
    float dGda = ((3.535f + 2.f * 2.181f * a) * den - num * (2.276f + 2.f * 2.577f * a)) / (den * den);
    float d_a = dG * dGda;
    float d_q = -d_a / (q * q);
    float d_alpha = d_q * tan_t;
    float d_tan = d_q * alpha;
    if (s2 <= 0.f) {
        d_r += d_alpha * 0.5f / alpha;
    } else {
        float d_qq = d_alpha * 0.5f / alpha;
        float d_cp2 = d_qq * r, d_sp2 = d_qq * r;
        d_r += d_qq * (cp2 + sp2);
        float d_vlx = 2.f * vl.x * inv * d_cp2;
        float d_vly = 2.f * vl.y * inv * d_sp2;
        float d_inv = vl.x * vl.x * d_cp2 + vl.y * vl.y * d_sp2;
        float d_s2 = -inv * inv * d_inv;
        float d_cz = -2.f * cz * d_s2;
        d_v += fr.x * d_vlx + fr.y * d_vly + fr.n * d_cz;
        d_fr.x += v * d_vlx;
        d_fr.y += v * d_vly;
        d_fr.n += v * d_cz;
    }
    float d_x = d_tan * 0.5f / tan_t;
    float d_c = d_x * (-2.f / (c * c * c));
    d_v += n * d_c;
    d_n += v * d_c;
}

// adjoint of bsdf (dr_device.h / material.h:257-351)
DR_INLINE void bsdf_adj(const DrSceneView &S, const GradSink &G, int mid, const SPoint &sp, F3 wi, F3 wo, F3 d_out,
                        AdjSP &dsp, F3 &d_wi, F3 &d_wo) {
    const DrMaterial &m = S.materials[mid];
    F3 n = sp.fr.n;
    float sgn = 1.f;
    if (m.two_sided && dot(wi, n) < 0.f) { n = -n; sgn = -1.f; }
    float c = dot(n, wo);
    if (c <= 5e-3f) return;
    F3 Dr = get_diffuse(S, m, sp.uv);
    F3 Sr = get_specular(S, m, sp.uv);
    float r = get_roughness(S, m, sp.uv);
    const float inv_pi = 1.f / DR_PI;
    F3 d_Dr = d_out * (c * inv_pi);
    float d_c = dot(d_out, Dr) * inv_pi;
    F3 d_n = wo * d_c;
    d_wo += n * d_c;
    F3 d_Sr = F3{0.f, 0.f, 0.f};
    float d_r = 0.f;
    if (sum3(Sr) > 0.f && dot(wi, n) > 0.f) {
        F3 hun = wi + wo;
        F3 h = normalize(hun);
        float hz = dot(h, sp.fr.n);
        if (hz > 0.f) {
            float e = roughness_to_phong(r);
            float D = phong_D(hz, e);
            float G1i = smith_g1(wi, h, n, sp.fr, r), G1o = smith_g1(wo, h, n, sp.fr, r);
            float Gv = G1i * G1o;
            float cd = dot(h, wo);
            float xs = 1.f - cd;
            float sch = powf(fmaxf(xs, 0.f), 5.f);
            F3 F = F3{Sr.x + (1.f - Sr.x) * sch, Sr.y + (1.f - Sr.y) * sch, Sr.z + (1.f - Sr.z) * sch};
            float den = 4.f * dot(wi, n);
            float inv = 1.f / den;
            F3 d_F = d_out * (D * Gv * inv);
            float dFs = dot(d_out, F);
            float d_D = dFs * Gv * inv;
            float d_G = dFs * D * inv;
            float d_den = -dFs * D * Gv * inv * inv;
            d_wi += n * (4.f * d_den);
            d_n += wi * (4.f * d_den);
            d_Sr += d_F * (1.f - sch);
            float d_sch = d_F.x * (1.f - Sr.x) + d_F.y * (1.f - Sr.y) + d_F.z * (1.f - Sr.z);
            F3 d_h = F3{0.f, 0.f, 0.f};
            if (xs >= 0.f) {
                float d_cd = -d_sch * 5.f * powf(xs, 4.f);
                d_h += wo * d_cd;
                d_wo += h * d_cd;
            }
            smith_g1_adj(wi, h, n, sp.fr, r, d_G * G1o, d_wi, d_n, dsp.fr, d_r);
            smith_g1_adj(wo, h, n, sp.fr, r, d_G * G1i, d_wo, d_n, dsp.fr, d_r);
            float d_hz = 0.f, d_e = 0.f;
            phong_D_adj(hz, e, d_D, d_hz, d_e);
            d_r += d_e * d_phong_dr(r);
            d_h += sp.fr.n * d_hz;
            dsp.fr.n += h * d_hz;
            F3 d_hun = d_normalize(hun, d_h);
            d_wi += d_hun;
            d_wo += d_hun;
        }
    }
    dsp.fr.n += d_n * sgn;
    F2 du = tex3_adj(S, G, m.diffuse, m.duv, sp.uv, d_Dr);
    F2 su = tex3_adj(S, G, m.specular, m.suv, sp.uv, d_Sr);
    F2 ru = tex1_adj(S, G, m.roughness, m.ruv, sp.uv, d_r);
    dsp.uv = dsp.uv + du + su + ru;
}

// adjoint of bsdf_pdf (material.h:414-462)
DR_INLINE void bsdf_pdf_adj(const DrSceneView &S, const GradSink &G, int mid, const SPoint &sp, F3 wi, F3 wo,
                            float d_pdf, AdjSP &dsp, F3 &d_wi, F3 &d_wo) {
    const DrMaterial &m = S.materials[mid];
    F3 Dr = get_diffuse(S, m, sp.uv);
    F3 Sr = get_specular(S, m, sp.uv);
    float dw = luminance(Dr), sw = luminance(Sr);
    float ws = dw + sw;
    if (ws <= 0.f) return;
    float dpmf = dw / ws, spmf = sw / ws;
    float d_dpmf = 0.f, d_spmf = 0.f, d_r = 0.f;
    if (dpmf > 0.f) {
        float c = dot(sp.fr.n, wo);
        float cc = m.two_sided ? fabsf(c) : fmaxf(c, 0.f);
        d_dpmf += d_pdf * cc / DR_PI;
        float d_cc = d_pdf * dpmf / DR_PI;
        float d_c = m.two_sided ? (c >= 0.f ? d_cc : -d_cc) : (c >= 0.f ? d_cc : 0.f);
        dsp.fr.n += wo * d_c;
        d_wo += sp.fr.n * d_c;
    }
    if (spmf > 0.f) {
        F3 hun = wi + wo;
        F3 h = normalize(hun);
        float hz = dot(h, sp.fr.n);
        if (hz > 0.f) {
            float r = get_roughness(S, m, sp.uv);
            float e = roughness_to_phong(r);
            float D = phong_D(hz, e);
            float dwh = dot(wo, h);
            float den = 4.f * fabsf(dwh);
            d_spmf += d_pdf * D / den;
            float d_D = d_pdf * spmf / den;
            float d_den = -d_pdf * spmf * D / (den * den);
            float d_dwh = 4.f * d_den * (dwh >= 0.f ? 1.f : -1.f);
            F3 d_h = wo * d_dwh;
            d_wo += h * d_dwh;
            float d_hz = 0.f, d_e = 0.f;
            phong_D_adj(hz, e, d_D, d_hz, d_e);
            d_r += d_e * d_phong_dr(r);
            d_h += sp.fr.n * d_hz;
            dsp.fr.n += h * d_hz;
            F3 d_hun = d_normalize(hun, d_h);
            d_wi += d_hun;
            d_wo += d_hun;
        }
    }
    float d_ws = -(d_dpmf * dw + d_spmf * sw) / (ws * ws);
    float d_dw = d_dpmf / ws + d_ws;
    float d_sw = d_spmf / ws + d_ws;
    F2 du = tex3_adj(S, G, m.diffuse, m.duv, sp.uv, lum_w() * d_dw);
    F2 su = tex3_adj(S, G, m.specular, m.suv, sp.uv, lum_w() * d_sw);
    F2 ru = tex1_adj(S, G, m.roughness, m.ruv, sp.uv, d_r);
    dsp.uv = dsp.uv + du + su + ru;
}

// adjoint of the specular branch of sample_bsdf_dir (material.h:390-411)
DR_INLINE void sample_dir_adj(const DrSceneView &S, const GradSink &G, int mid, const SPoint &sp, F3 wi, float u0,
                              float u1, F3 d_dir, AdjSP &dsp, F3 &d_wi) {
    const DrMaterial &m = S.materials[mid];
    float r = get_roughness(S, m, sp.uv);
    float e = roughness_to_phong(r);
    float phi = 2.f * DR_PI * u1;
    float cp = cosf(phi), sphi = sinf(phi);
    float ee = 1.0f / (e + 2.0f);
    float cos_t = powf(u0, ee);
    float xs = 1.f - cos_t * cos_t;
    float sin_t = sqrtf(fmaxf(xs, 0.f));
    F3 ml = F3{sin_t * cp, sin_t * sphi, cos_t};
    F3 mw = to_world(sp.fr, ml);
    float k = 2.f * dot(wi, mw);
    F3 d_mw = d_dir * k;
    float d_k = dot(d_dir, mw);
    d_wi += -d_dir;
    d_wi += mw * (2.f * d_k);
    d_mw += wi * (2.f * d_k);
    dsp.fr.x += d_mw * ml.x;
    dsp.fr.y += d_mw * ml.y;
    dsp.fr.n += d_mw * ml.z;
    F3 d_ml = F3{dot(d_mw, sp.fr.x), dot(d_mw, sp.fr.y), dot(d_mw, sp.fr.n)};
    float d_sin = d_ml.x * cp + d_ml.y * sphi;
    float d_cos = d_ml.z;
    if (xs >= 0.f && sin_t > 0.f) d_cos += d_sin * 0.5f / sin_t * (-2.f * cos_t);
    float d_ee = d_cos * cos_t * logf(fmaxf(u0, 1e-6f));
    float d_e = -d_ee * ee * ee;
    float d_r = d_e * d_phong_dr(r);
    dsp.uv = dsp.uv + tex1_adj(S, G, m.roughness, m.ruv, sp.uv, d_r);
}

// -------------------------------------------------------- surface point ---
// adjoint of surface_point (intersect.h:100-148)
DR_INLINE void surface_point_adj(const TriData &d, F3 org, F3 dir, const AdjSP &a, F3 &d_org, F3 &d_dir, F3 d_v[3],
                                 F3 d_n[3]) {
    F3 v0 = d.v[0], v1 = d.v[1], v2 = d.v[2];
    F3 e1 = v1 - v0, e2 = v2 - v0;
    F3 pvec = cross(dir, e2);
    float div = dot(pvec, e1);
    bool dz = div == 0.f;
    if (dz) div = 1e-8f;
    F3 s = org - v0;
    F3 qvec = cross(s, e1);
    float idiv = 1.f / div;
    float u = dot(s, pvec) * idiv, v = dot(dir, qvec) * idiv, t = dot(e2, qvec) * idiv;
    float w = 1.f - (u + v);
    // shading normal -> frame
    F3 d_sn = F3{0.f, 0.f, 0.f};
    F3 sn;
    F3 cr = cross(e1, e2);
    if (d.has_n) {
        F3 nun = d.n[0] * w + d.n[1] * u + d.n[2] * v;
        sn = normalize(nun);
        d_sn = d_make_frame(sn, a.fr);
        F3 d_nun = d_normalize(nun, d_sn);
        d_n[0] += d_nun * w;
        d_n[1] += d_nun * u;
        d_n[2] += d_nun * v;
        float d_w = dot(d_nun, d.n[0]);
        // folded into d_u, d_v below
        float d_u = dot(d_nun, d.n[1]) - d_w, d_v_ = dot(d_nun, d.n[2]) - d_w;
        // uv
        float dwu = a.uv.x * d.uv[0].x + a.uv.y * d.uv[0].y;
        d_u += a.uv.x * d.uv[1].x + a.uv.y * d.uv[1].y - dwu;
        d_v_ += a.uv.x * d.uv[2].x + a.uv.y * d.uv[2].y - dwu;
        // pos = org + dir * t
        d_org += a.pos;
        d_dir += a.pos * t;
        float d_t = dot(a.pos, dir);
        // Moller-Trumbore
        float d_div = dz ? 0.f : -(d_u * u + d_v_ * v + d_t * t) / div;
        F3 d_s = pvec * (d_u / div);
        F3 d_pvec = s * (d_u / div) + e1 * d_div;
        F3 d_e1 = pvec * d_div;
        d_dir += qvec * (d_v_ / div);
        F3 d_qvec = dir * (d_v_ / div) + e2 * (d_t / div);
        F3 d_e2 = qvec * (d_t / div);
        d_s += cross(e1, d_qvec);
        d_e1 += cross(d_qvec, s);
        d_dir += cross(e2, d_pvec);
        d_e2 += cross(d_pvec, dir);
        d_org += d_s;
        d_v[0] -= d_s;
        d_v[1] += d_e1;
        d_v[2] += d_e2;
        d_v[0] -= d_e1 + d_e2;
    } else {
        sn = normalize(cr);
        d_sn = d_make_frame(sn, a.fr);
        F3 d_cr = d_normalize(cr, d_sn);
        F3 d_e1 = cross(e2, d_cr);
        F3 d_e2 = cross(d_cr, e1);
        float dwu = a.uv.x * d.uv[0].x + a.uv.y * d.uv[0].y;
        float d_u = a.uv.x * d.uv[1].x + a.uv.y * d.uv[1].y - dwu;
        float d_v_ = a.uv.x * d.uv[2].x + a.uv.y * d.uv[2].y - dwu;
        d_org += a.pos;
        d_dir += a.pos * t;
        float d_t = dot(a.pos, dir);
        float d_div = dz ? 0.f : -(d_u * u + d_v_ * v + d_t * t) / div;
        F3 d_s = pvec * (d_u / div);
        F3 d_pvec = s * (d_u / div) + e1 * d_div;
        d_e1 += pvec * d_div;
        d_dir += qvec * (d_v_ / div);
        F3 d_qvec = dir * (d_v_ / div) + e2 * (d_t / div);
        d_e2 += qvec * (d_t / div);
        d_s += cross(e1, d_qvec);
        d_e1 += cross(d_qvec, s);
        d_dir += cross(e2, d_pvec);
        d_e2 += cross(d_pvec, dir);
        d_org += d_s;
        d_v[0] -= d_s;
        d_v[1] += d_e1;
        d_v[2] += d_e2;
        d_v[0] -= d_e1 + d_e2;
    }
}

// ------------------------------------------------------------- lights -----
// adjoint of nee (pathtrace.h:59-102) w.r.t. the shading point, the primary
// wi, the light triangle and the light intensity
DR_INLINE void nee_adj(const DrSceneView &S, const GradSink &G, const LightSample &ls, F3 wi, const Isect &sis,
                       const SPoint &sp, F3 d_out, AdjSP &dsp, F3 &d_wi) {
    if (!valid(sis)) return;
    LightPoint L = sample_point_on_light(S, ls);
    F3 ldir = normalize(L.pos - sp.pos);
    float lcos = fabsf(dot(ldir, L.n));
    float area;
    float lpdf = light_pdf_tri(S, L.light, L.v[0], L.v[1], L.v[2], &area);
    float bpdf = bsdf_pdf(S, S.shapes[sis.shape].mat, sp, wi, ldir) * lcos / length_sq(L.pos - sp.pos);
    float mis = sq(lpdf) / (sq(bpdf) + sq(lpdf));
    F3 E = eval_light(S, wi, sis, sp, L.light, L.pos, L.n);
    float k = mis / lpdf;
    F3 d_E = d_out * k;
    float d_k = dot(d_out, E);
    float d_lpdf = -d_k * mis / (lpdf * lpdf);
    float d_area = -d_lpdf * lpdf / area;
    F3 e1 = L.v[1] - L.v[0], e2 = L.v[2] - L.v[0];
    F3 cr = cross(e1, e2);
    F3 d_cr = cr * (d_area * 0.5f / length(cr));
    // eval
    F3 d_lp = F3{0.f, 0.f, 0.f};
    {
        F3 dd = L.pos - sp.pos;
        float dist = length(dd);
        if (dist != 0.f) {
            F3 nd = dd / dist;
            if (!occluded(S, sp.pos, nd, dist)) {
                int mid = S.shapes[sis.shape].mat;
                F3 b = bsdf(S, mid, sp, wi, nd);
                float lc = dot(L.n, -nd);
                if (lc > 0.f) {
                    const DrLight &l = S.lights[L.light];
                    F3 I = F3{l.I[0], l.I[1], l.I[2]};
                    float dist2 = sq(dist);
                    float f = lc / dist2;
                    F3 d_b = d_E * (I * f);
                    if (G.color_mode) G.add3(l.grad_off, d_E * b * f);
                    float d_f = dot(d_E * b, I);
                    float d_lc = d_f / dist2;
                    float d_dist2 = -d_f * lc / (dist2 * dist2);
                    float d_dist = d_dist2 * 2.f * dist;
                    F3 d_ln = -nd * d_lc;
                    F3 d_nd = -L.n * d_lc;
                    bsdf_adj(S, G, mid, sp, wi, nd, d_b, dsp, d_wi, d_nd);
                    F3 d_dd = d_nd / dist;
                    d_dist += -dot(d_nd, dd) / (dist * dist);
                    d_dd += dd * (d_dist / dist);
                    d_lp += d_dd;
                    dsp.pos -= d_dd;
                    d_cr += d_normalize(cr, d_ln);
                }
            }
        }
    }
    F3 d_e1 = cross(e2, d_cr) + d_lp * L.b1;
    F3 d_e2 = cross(d_cr, e1) + d_lp * L.b2;
    F3 d_v0 = d_lp - d_e1 - d_e2;
    const DrShape &lsh = S.shapes[L.shape];
    G.add3(lsh.vgrad_off + 3ll * (L.gv[0] - lsh.vert_off), d_v0);
    G.add3(lsh.vgrad_off + 3ll * (L.gv[1] - lsh.vert_off), d_e1);
    G.add3(lsh.vgrad_off + 3ll * (L.gv[2] - lsh.vert_off), d_e2);
}

// scatter the adjoint of a triangle's vertices (and shading normals)
DR_INLINE void scatter_tri(const DrSceneView &S, const GradSink &G, const TriData &d, int shape, const F3 d_v[3],
                           const F3 d_n[3]) {
    const DrShape &sh = S.shapes[shape];
    for (int k = 0; k < 3; k++) G.add3(sh.vgrad_off + 3ll * (d.gv[k] - sh.vert_off), d_v[k]);
    if (G.color_mode && d.has_n && sh.ngrad_off >= 0) {
        for (int k = 0; k < 3; k++) {
            F3 dn = (G.flags & DR_QUIRK_NORMAL_DN0) ? d_n[0] : d_n[k];
            G.add3(sh.ngrad_off + 3ll * (d.gv[k] - sh.vert_off), dn);
        }
    }
}
