// Regeneration of the LTC table the renderer reads (dr_ltc_fit; the
// reference's fit/fitLTC.cpp:191-276 with the Blinn-Phong BRDF of
// fit/brdf_blinn_phong.h:10-66, the LTC of fit/LTC.h and the downhill simplex
// of fit/nelder_mead.h).  For each roughness index a (row) and view angle
// index t of an n x n table, a linearly transformed cosine M = [X Y Z] *
// [[m11, 0, m13], [0, m22, m23], [0, 0, 1]] is fitted to the cosine-weighted
// BRDF by Nelder-Mead on the MIS estimate (50 x 50 stratified samples of
// each density) of the cubed absolute difference; entries (0,1), (1,0),
// (1,2), (2,1) are then zeroed and M is divided by M[2][2] (column-major
// indices; the table holds each matrix row by row).  The first guess of an
// entry is the fit of the previous view angle in the same row (t > 0, around
// the BRDF's mean
// direction) or of the next rougher row's first entry (t = 0): rows depend on
// each other only through their first entries, so those are fitted first, in
// order, and the rest of every row in parallel.  The matrix and vector
// algebra is written in the column-major conventions of the reference's GLM
// calls (M[c][r]), with their formulas, in single precision.
#include "../../include/delta_ray.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <string>
#include <thread>
#include <vector>

void dr_set_error(const std::string &msg);  // render.hip

namespace {

struct V3 {
    float x, y, z;
};
V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
float length(V3 a) { return std::sqrt(dot(a, a)); }
V3 normalize(V3 a) { return a * (1.f / std::sqrt(dot(a, a))); }

// column-major 3x3: c[col] is a column
struct M3 {
    V3 c[3];
    float at(int col, int row) const { return row == 0 ? c[col].x : (row == 1 ? c[col].y : c[col].z); }
    void set(int col, int row, float v) { (row == 0 ? c[col].x : (row == 1 ? c[col].y : c[col].z)) = v; }
};
V3 mul(const M3 &m, V3 v) {
    return V3{m.c[0].x * v.x + m.c[1].x * v.y + m.c[2].x * v.z, m.c[0].y * v.x + m.c[1].y * v.y + m.c[2].y * v.z,
              m.c[0].z * v.x + m.c[1].z * v.y + m.c[2].z * v.z};
}
M3 mul(const M3 &a, const M3 &b) {
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            r.set(i, j, a.at(0, j) * b.at(i, 0) + a.at(1, j) * b.at(i, 1) + a.at(2, j) * b.at(i, 2));
    return r;
}
float determinant(const M3 &m) {
    return +m.at(0, 0) * (m.at(1, 1) * m.at(2, 2) - m.at(2, 1) * m.at(1, 2)) -
           m.at(1, 0) * (m.at(0, 1) * m.at(2, 2) - m.at(2, 1) * m.at(0, 2)) +
           m.at(2, 0) * (m.at(0, 1) * m.at(1, 2) - m.at(1, 1) * m.at(0, 2));
}
M3 inverse(const M3 &m) {
    float od = 1.f / (+m.at(0, 0) * (m.at(1, 1) * m.at(2, 2) - m.at(2, 1) * m.at(1, 2)) -
                      m.at(1, 0) * (m.at(0, 1) * m.at(2, 2) - m.at(2, 1) * m.at(0, 2)) +
                      m.at(2, 0) * (m.at(0, 1) * m.at(1, 2) - m.at(1, 1) * m.at(0, 2)));
    M3 r;
    r.set(0, 0, +(m.at(1, 1) * m.at(2, 2) - m.at(2, 1) * m.at(1, 2)) * od);
    r.set(1, 0, -(m.at(1, 0) * m.at(2, 2) - m.at(2, 0) * m.at(1, 2)) * od);
    r.set(2, 0, +(m.at(1, 0) * m.at(2, 1) - m.at(2, 0) * m.at(1, 1)) * od);
    r.set(0, 1, -(m.at(0, 1) * m.at(2, 2) - m.at(2, 1) * m.at(0, 2)) * od);
    r.set(1, 1, +(m.at(0, 0) * m.at(2, 2) - m.at(2, 0) * m.at(0, 2)) * od);
    r.set(2, 1, -(m.at(0, 0) * m.at(2, 1) - m.at(2, 0) * m.at(0, 1)) * od);
    r.set(0, 2, +(m.at(0, 1) * m.at(1, 2) - m.at(1, 1) * m.at(0, 2)) * od);
    r.set(1, 2, -(m.at(0, 0) * m.at(1, 2) - m.at(1, 0) * m.at(0, 2)) * od);
    r.set(2, 2, +(m.at(0, 0) * m.at(1, 1) - m.at(1, 0) * m.at(0, 1)) * od);
    return r;
}

const int kSamples = 50;
const float kMinAlpha = 0.0001f;

// scaled hypotenuse of (a, b) (vector.h:736-746)
float hypot2(float a, float b) {
    if (std::fabs(a) > std::fabs(b)) {
        float r = b / a;
        return std::fabs(a) * std::sqrt(1.f + r * r);
    } else if (b != 0.f) {
        float r = a / b;
        return std::fabs(b) * std::sqrt(1.f + r * r);
    }
    return 0.f;
}

// cosine-weighted Blinn-Phong with the Smith term of the reference's material
// (phong exponent 2 / alpha^2), and its half-vector sampling
float brdf_eval(V3 V, V3 L, float alpha, float &pdf) {
    if (V.z <= 0) {
        pdf = 0;
        return 0;
    }
    V3 H = normalize(V + L);
    auto project_roughness = [&](V3 v) {
        float ct = v.z;
        float st2 = 1.f - ct * ct;
        if (st2 <= 0.f) return alpha;
        float inv = 1.f / st2;
        float cp2 = v.x * v.x * inv, sp2 = v.y * v.y * inv;
        return std::sqrt(cp2 + sp2) * alpha;
    };
    auto smith_g1 = [&](V3 v) {
        float ct = v.z;
        if (dot(v, H) * ct <= 0) return 0.f;
        float tt = std::sqrt(1.f / (ct * ct) - 1.f);
        if (tt == 0.0f) return 1.f;
        float root = project_roughness(v) * tt;
        return 2.0f / (1.0f + hypot2(1.0f, root));
    };
    float G = smith_g1(V) * smith_g1(L);
    float e = std::max(2.f / (alpha * alpha), 0.f);
    float D = std::pow(H.z, e) * (e + 2.f) / float(2 * M_PI);
    pdf = std::fabs(D * H.z / 4.0f / dot(V, H));
    return D * G / 4.0f / V.z;
}
V3 brdf_sample(V3 V, float alpha, float U1, float U2) {
    float phi = 2.0f * 3.14159f * U1;
    float e = std::max(2.f / (alpha * alpha), 0.f);
    float ct = std::pow(U2, 1.f / (e + 2.f));
    float r = std::sqrt(1.f / (ct * ct) - 1.f);
    V3 N = normalize(V3{r * std::cos(phi), r * std::sin(phi), 1.0f});
    return -V + 2.0f * N * dot(N, V);
}

struct LTC {
    float amplitude = 1.f;
    float m11 = 1.f, m22 = 1.f, m13 = 0.f, m23 = 0.f;
    V3 X{1.f, 0.f, 0.f}, Y{0.f, 1.f, 0.f}, Z{0.f, 0.f, 1.f};
    M3 M, invM;
    float detM = 1.f;
    LTC() { update(); }
    void update() {
        M3 B{{X, Y, Z}};
        M3 P{{V3{m11, 0.f, 0.f}, V3{0.f, m22, 0.f}, V3{m13, m23, 1.f}}};
        M = mul(B, P);
        invM = inverse(M);
        detM = std::fabs(determinant(M));
    }
    float eval(V3 L) const {
        V3 lo = normalize(mul(invM, L));
        V3 l2 = mul(M, lo);
        float l = length(l2);
        float jac = detM / (l * l * l);
        float D = 1.0f / 3.14159f * std::max(0.0f, lo.z);
        return amplitude * D / jac;
    }
    V3 sample(float U1, float U2) const {
        float theta = std::acos(std::sqrt(U1));
        float phi = 2.0f * 3.14159f * U2;
        return normalize(mul(M, V3{std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta)}));
    }
};

float compute_norm(V3 V, float alpha) {
    float norm = 0.0;
    for (int j = 0; j < kSamples; ++j)
        for (int i = 0; i < kSamples; ++i) {
            float U1 = (i + 0.5f) / (float)kSamples, U2 = (j + 0.5f) / (float)kSamples;
            V3 L = brdf_sample(V, alpha, U1, U2);
            float pdf;
            float ev = brdf_eval(V, L, alpha, pdf);
            norm += (pdf > 0) ? ev / pdf : 0.0f;
        }
    return norm / (float)(kSamples * kSamples);
}

V3 compute_average_dir(V3 V, float alpha) {
    V3 avg{0, 0, 0};
    for (int j = 0; j < kSamples; ++j)
        for (int i = 0; i < kSamples; ++i) {
            float U1 = (i + 0.5f) / (float)kSamples, U2 = (j + 0.5f) / (float)kSamples;
            V3 L = brdf_sample(V, alpha, U1, U2);
            float pdf;
            float ev = brdf_eval(V, L, alpha, pdf);
            avg = avg + ((pdf > 0) ? ev / pdf * L : V3{0, 0, 0});
        }
    avg.y = 0.0f;
    return normalize(avg);
}

float compute_error(const LTC &ltc, V3 V, float alpha) {
    double error = 0.0;
    for (int j = 0; j < kSamples; ++j)
        for (int i = 0; i < kSamples; ++i) {
            float U1 = (i + 0.5f) / (float)kSamples, U2 = (j + 0.5f) / (float)kSamples;
            for (int strategy = 0; strategy < 2; strategy++) {
                V3 L = strategy == 0 ? ltc.sample(U1, U2) : brdf_sample(V, alpha, U1, U2);
                float pdf_brdf;
                float eval_brdf = brdf_eval(V, L, alpha, pdf_brdf);
                float eval_ltc = ltc.eval(L);
                float pdf_ltc = eval_ltc / ltc.amplitude;
                double e = std::fabs(eval_brdf - eval_ltc);
                e = e * e * e;
                error += e / (pdf_ltc + pdf_brdf);
            }
        }
    return (float)error / (float)(kSamples * kSamples);
}

void apply_params(LTC &ltc, const float *p, bool isotropic) {
    float m11 = std::max<float>(p[0], kMinAlpha), m22 = std::max<float>(p[1], kMinAlpha);
    ltc.m11 = m11;
    ltc.m22 = isotropic ? m11 : m22;
    ltc.m13 = isotropic ? 0.0f : p[2];
    ltc.m23 = isotropic ? 0.0f : p[3];
    ltc.update();
}

// downhill simplex in 4 dimensions (reflection 1, expansion 2, contraction
// and shrink 0.5; stop when 2 |f_lo - f_hi| < (|f_lo| + |f_hi|) tol)
template <class F>
float nelder_mead(float *pmin, const float *start, float delta, float tol, int max_iters, F f_of) {
    const int D = 4, P = D + 1;
    float s[P][D], f[P];
    for (int i = 0; i < P; i++) {
        for (int k = 0; k < D; k++) s[i][k] = start[k];
        if (i > 0) s[i][i - 1] += delta;
    }
    for (int i = 0; i < P; i++) f[i] = f_of(s[i]);
    int lo = 0, hi, nh;
    for (int it = 0; it < max_iters; it++) {
        lo = hi = nh = 0;
        for (int i = 1; i < P; i++) {
            if (f[i] < f[lo]) lo = i;
            if (f[i] > f[hi]) {
                nh = hi;
                hi = i;
            } else if (f[i] > f[nh]) {
                nh = i;
            }
        }
        float a = std::fabs(f[lo]), b = std::fabs(f[hi]);
        if (2.0f * std::fabs(a - b) < (a + b) * tol) break;
        float o[D] = {0.f, 0.f, 0.f, 0.f};
        for (int i = 0; i < P; i++)
            if (i != hi)
                for (int k = 0; k < D; k++) o[k] += s[i][k];
        for (int k = 0; k < D; k++) o[k] /= D;
        float r[D];
        for (int k = 0; k < D; k++) r[k] = o[k] + 1.0f * (o[k] - s[hi][k]);
        float fr = f_of(r);
        if (fr < f[nh]) {
            if (fr < f[lo]) {
                float e[D];
                for (int k = 0; k < D; k++) e[k] = o[k] + 2.0f * (o[k] - s[hi][k]);
                float fe = f_of(e);
                if (fe < fr) {
                    for (int k = 0; k < D; k++) s[hi][k] = e[k];
                    f[hi] = fe;
                    continue;
                }
            }
            for (int k = 0; k < D; k++) s[hi][k] = r[k];
            f[hi] = fr;
            continue;
        }
        float c[D];
        for (int k = 0; k < D; k++) c[k] = o[k] - 0.5f * (o[k] - s[hi][k]);
        float fc = f_of(c);
        if (fc < f[hi]) {
            for (int k = 0; k < D; k++) s[hi][k] = c[k];
            f[hi] = fc;
            continue;
        }
        for (int q = 0; q < P; q++) {
            if (q == lo) continue;
            for (int k = 0; k < D; k++) s[q][k] = s[lo][k] + 0.5f * (s[q][k] - s[lo][k]);
            f[q] = f_of(s[q]);
        }
    }
    for (int k = 0; k < D; k++) pmin[k] = s[lo][k];
    return f[lo];
}

struct Fitter {
    int n;
    bool consistent;
    float alpha_of(int a) const {
        float r = a / float(n - 1);
        // the reference fits alpha = r^2 (Phong exponent 2 / r^4) while the
        // renderer's material uses the exponent 2 / r at roughness r;
        // consistent: alpha = sqrt(r), the lobe the renderer evaluates
        return std::max<float>(consistent ? std::sqrt(r) : r * r, kMinAlpha);
    }
    V3 view(int t) const {
        float theta = std::min<float>(1.57f, t / float(n - 1) * 1.57079f);
        return V3{std::sin(theta), 0, std::cos(theta)};
    }
    // entry (a, t) from the state `ltc` left by the entry fitted before it
    // (its parameters: the first guess); returns the stored matrix
    M3 fit_entry(LTC &ltc, int a, int t, float *amp) const {
        V3 V = view(t);
        float alpha = alpha_of(a);
        ltc.amplitude = compute_norm(V, alpha);
        V3 avg = compute_average_dir(V, alpha);
        bool iso;
        if (t == 0) {
            ltc.X = V3{1, 0, 0};
            ltc.Y = V3{0, 1, 0};
            ltc.Z = V3{0, 0, 1};
            ltc.m13 = 0;
            ltc.m23 = 0;
            ltc.update();
            iso = true;
        } else {
            V3 L = normalize(avg);
            ltc.X = V3{L.z, 0, -L.x};
            ltc.Y = V3{0, 1, 0};
            ltc.Z = L;
            ltc.update();
            iso = false;
        }
        float start[4] = {ltc.m11, ltc.m22, ltc.m13, ltc.m23}, best[4];
        nelder_mead(best, start, 0.05f, 1e-5f, 100, [&](const float *p) {
            apply_params(ltc, p, iso);
            return compute_error(ltc, V, alpha);
        });
        apply_params(ltc, best, iso);
        *amp = ltc.amplitude;
        M3 m = ltc.M;
        m.set(0, 1, 0);
        m.set(1, 0, 0);
        m.set(2, 1, 0);
        m.set(1, 2, 0);
        float s = 1.0f / m.at(2, 2);
        for (int c = 0; c < 3; c++) m.c[c] = s * m.c[c];
        return m;
    }
};

// the stored entry as the 9 floats of the table, row by row (the layout of
// the reference's ltc.inc, fit/export.h:5-60, which the renderer reads)
void store(const M3 &m, float *out) {
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) out[3 * r + c] = m.at(c, r);
}
M3 load(const float *e) {
    M3 m;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) m.set(c, r, e[3 * r + c]);
    return m;
}

}  // namespace

extern "C" {

dr_status dr_ltc_fit(int n, int row_begin, int row_end, int t_begin, int t_end, const float *guess, int flags,
                     int nthreads, float *tab, float *amplitude) {
    if (n < 2 || row_begin < 0 || row_end > n || row_begin >= row_end || t_begin < 0 || t_end > n ||
        t_begin >= t_end || !tab || !amplitude) {
        dr_set_error("dr_ltc_fit: bad arguments");
        return DR_ERROR_INVALID_ARGUMENT;
    }
    if ((t_begin > 0 || row_end < n) && !guess) {
        dr_set_error("dr_ltc_fit: a partial fit needs the table of first guesses");
        return DR_ERROR_INVALID_ARGUMENT;
    }
    Fitter F{n, (flags & DR_LTC_FIT_CONSISTENT) != 0};
    // the first guess of entry (a, t) from a stored table entry (a, t - 1):
    // the LTC's parameters are the stored matrix expressed in the frame the
    // previous entry was fitted in (its scale undone by the unit (2, 2)
    // parameter); the zeroed entries lose m23, which the isotropic BRDF
    // leaves at 0
    auto guess_state = [&](LTC &ltc, int a, int t) {
        if (t == 0) {
            // first entry of a row: the next rougher row's first entry, or 1
            if (a == n - 1) {
                ltc.m11 = ltc.m22 = 1.0f;
            } else {
                const float *g = (a + 1 < row_end && t_begin == 0) ? &tab[(size_t)(a + 1) * 9] : &guess[(size_t)(a + 1) * 9];
                ltc.m11 = std::max<float>(g[0], kMinAlpha);
                ltc.m22 = std::max<float>(g[4], kMinAlpha);
            }
            return;
        }
        M3 m = load(&guess[((size_t)a + (size_t)(t - 1) * n) * 9]);
        V3 L = normalize(compute_average_dir(F.view(t - 1), F.alpha_of(a)));
        M3 B{{V3{L.z, 0, -L.x}, V3{0, 1, 0}, L}};
        if (t - 1 == 0) B = M3{{V3{1, 0, 0}, V3{0, 1, 0}, V3{0, 0, 1}}};
        M3 P = mul(inverse(B), m);  // [[m11, 0, m13], [0, m22, m23], [0, 0, 1]] / s
        float s = 1.0f / P.at(2, 2);
        ltc.m11 = s * P.at(0, 0);
        ltc.m22 = s * P.at(1, 1);
        ltc.m13 = s * P.at(2, 0);
        ltc.m23 = s * P.at(2, 1);
    };
    // first entries of the rows, from the roughest down (each guesses from
    // the one before)
    if (t_begin == 0) {
        LTC ltc;
        for (int a = row_end - 1; a >= row_begin; --a) {
            guess_state(ltc, a, 0);
            store(F.fit_entry(ltc, a, 0, &amplitude[a]), &tab[(size_t)a * 9]);
        }
    }
    // the rest of every row, rows in parallel
    std::atomic<int> next(row_begin);
    auto worker = [&]() {
        for (int a = next.fetch_add(1); a < row_end; a = next.fetch_add(1)) {
            LTC ltc;
            int t0 = std::max(t_begin, 1);
            if (t0 == t_begin) {
                guess_state(ltc, a, t0);
            } else {
                // continue from the row's first entry just fitted
                guess_state(ltc, a, 0);
                float amp;
                F.fit_entry(ltc, a, 0, &amp);
            }
            for (int t = t0; t < t_end; ++t) {
                size_t k = (size_t)a + (size_t)t * n;
                store(F.fit_entry(ltc, a, t, &amplitude[k]), &tab[k * 9]);
            }
        }
    };
    int nt = std::max(1, std::min(nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency(),
                                  row_end - row_begin));
    std::vector<std::thread> th;
    for (int i = 0; i < nt; i++) th.emplace_back(worker);
    for (auto &x : th) x.join();
    return DR_OK;
}

}  // extern "C"
