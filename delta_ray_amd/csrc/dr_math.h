// Device math for the path tracer kernels: 3-vectors, frames, 3x3 matrices
// and the counter-based random streams.  The formulas follow the reference's
// vector.h (dot / cross / normalize / luminance / make_frame / cos_hemisphere
// / normalize_jacobian / intersect_jacobian / inverse) and transform.h.
//
// The library is compiled with -ffp-contract=off: the ray/triangle test must
// round exactly like the CPU oracle's so that hit ids agree bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define DR_INLINE __device__ __forceinline__
#define DR_PI 3.14159265358979323846f

// Address spaces in device code: loads through DR_GLOBAL pointers compile to
// global_* instructions and DR_LDS to ds_*, where the generic flat_* ones
// would also hold the LDS / scalar-load wait counter.
#if defined(__HIP_DEVICE_COMPILE__)
#ifndef DR_GLOBAL
#define DR_GLOBAL __attribute__((address_space(1)))
#endif
#define DR_LDS __attribute__((address_space(3)))
#else
#ifndef DR_GLOBAL
#define DR_GLOBAL
#endif
#define DR_LDS
#endif
// a pointer into device (global) memory, for loads
template <class T>
DR_INLINE const DR_GLOBAL T *gptr(const T *p) {
    return (const DR_GLOBAL T *)p;
}

struct F2 {
    float x, y;
};
struct F3 {
    float x, y, z;
    DR_INLINE float &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    DR_INLINE float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

DR_INLINE F3 f3(float a, float b, float c) { return F3{a, b, c}; }
DR_INLINE F2 f2(float a, float b) { return F2{a, b}; }
DR_INLINE F3 operator+(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
DR_INLINE F3 operator-(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
DR_INLINE F3 operator-(F3 a) { return F3{-a.x, -a.y, -a.z}; }
DR_INLINE F3 operator*(F3 a, float s) { return F3{a.x * s, a.y * s, a.z * s}; }
DR_INLINE F3 operator*(float s, F3 a) { return F3{s * a.x, s * a.y, s * a.z}; }
DR_INLINE F3 operator*(F3 a, F3 b) { return F3{a.x * b.x, a.y * b.y, a.z * b.z}; }
// vector / scalar as in vector.h:395-398: multiply by the reciprocal
DR_INLINE F3 operator/(F3 a, float s) { float inv = 1.f / s; return a * inv; }
DR_INLINE F3 &operator+=(F3 &a, F3 b) { a = a + b; return a; }
DR_INLINE F3 &operator-=(F3 &a, F3 b) { a = a - b; return a; }
DR_INLINE F3 &operator*=(F3 &a, float s) { a = a * s; return a; }
DR_INLINE F2 operator+(F2 a, F2 b) { return F2{a.x + b.x, a.y + b.y}; }
DR_INLINE F2 operator-(F2 a, F2 b) { return F2{a.x - b.x, a.y - b.y}; }
DR_INLINE F2 operator*(F2 a, float s) { return F2{a.x * s, a.y * s}; }
DR_INLINE F2 operator*(float s, F2 a) { return F2{s * a.x, s * a.y}; }

DR_INLINE float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DR_INLINE F3 cross(F3 a, F3 b) {
    return F3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
DR_INLINE float sq(float x) { return x * x; }
DR_INLINE float length_sq(F3 a) { return sq(a.x) + sq(a.y) + sq(a.z); }
DR_INLINE float length(F3 a) { return sqrtf(length_sq(a)); }
DR_INLINE float length2(F2 a) { return sqrtf(a.x * a.x + a.y * a.y); }
DR_INLINE F3 normalize(F3 a) { return a / length(a); }
DR_INLINE float sum3(F3 a) { return a.x + a.y + a.z; }
DR_INLINE float abs_sum(F3 a) { return fabsf(a.x) + fabsf(a.y) + fabsf(a.z); }
DR_INLINE float max3(F3 a) { return fmaxf(fmaxf(a.x, a.y), a.z); }
DR_INLINE bool is_zero(F3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
DR_INLINE float luminance(F3 a) { return 0.212671f * a.x + 0.715160f * a.y + 0.072169f * a.z; }
DR_INLINE F3 lum_w() { return F3{0.212671f, 0.715160f, 0.072169f}; }
// Hardware approximations (about 1 ulp) for sampling bounds, where only
// consistency between sampling and weighting matters, not the last bit:
// v_sqrt_f32, v_rsq_f32, v_rcp_f32, and pow as v_exp_f32(e * v_log_f32(x))
// for x in [0, 1].  The host build (CPU checks) uses the exact functions.
#if defined(__HIP_DEVICE_COMPILE__)
DR_INLINE float approx_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
DR_INLINE float approx_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
DR_INLINE float approx_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
DR_INLINE float approx_pow01(float x, float e) { return __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(x)); }
#else
DR_INLINE float approx_sqrt(float x) { return sqrtf(x); }
DR_INLINE float approx_rsqrt(float x) { return 1.f / sqrtf(x); }
DR_INLINE float approx_rcp(float x) { return 1.f / x; }
DR_INLINE float approx_pow01(float x, float e) { return powf(x, e); }
#endif

// three floats from device (global) memory
DR_INLINE F3 load3(const float *p) {
    const DR_GLOBAL float *q = gptr(p);
    return F3{q[0], q[1], q[2]};
}

// adjoint of normalize: d(v/|v|) given d_out
DR_INLINE F3 d_normalize(F3 v, F3 d_out) {
    float l = length(v);
    F3 n = v / l;
    return (d_out - n * dot(n, d_out)) * (1.f / l);
}

struct Frame {
    F3 x, y, n;
};
// vector.h:485-501
DR_INLINE Frame make_frame(F3 n) {
    if (n.z < -1.f + 1e-6f) return Frame{F3{0.f, -1.f, 0.f}, F3{-1.f, 0.f, 0.f}, n};
    float a = 1.f / (1.f + n.z);
    float b = -n.x * n.y * a;
    return Frame{F3{1.f - sq(n.x) * a, b, -n.x}, F3{b, 1.f - sq(n.y) * a, -n.y}, n};
}
// adjoint of make_frame: d_n from d(frame)
DR_INLINE F3 d_make_frame(F3 n, const Frame &d) {
    F3 dn = d.n;
    if (n.z < -1.f + 1e-6f) return dn;
    float a = 1.f / (1.f + n.z);
    float d_b = d.x.y + d.y.x;
    float d_a = -sq(n.x) * d.x.x - sq(n.y) * d.y.y - n.x * n.y * d_b;
    dn.x += -2.f * n.x * a * d.x.x - d.x.z - n.y * a * d_b;
    dn.y += -2.f * n.y * a * d.y.y - d.y.z - n.x * a * d_b;
    dn.z += -a * a * d_a;
    return dn;
}
DR_INLINE F3 to_local(const Frame &f, F3 v) { return F3{dot(v, f.x), dot(v, f.y), dot(v, f.n)}; }
DR_INLINE F3 to_world(const Frame &f, F3 v) { return f.x * v.x + f.y * v.y + f.n * v.z; }
DR_INLINE Frame neg_frame(const Frame &f) { return Frame{-f.x, -f.y, -f.n}; }

DR_INLINE F3 cos_hemisphere(F2 s) {
    float phi = 2.f * DR_PI * s.x;
    float tmp = sqrtf(1.f - s.y);
    return F3{cosf(phi) * tmp, sinf(phi) * tmp, sqrtf(s.y)};
}

struct M3 {
    float m[3][3];
};
DR_INLINE M3 m3_rows(F3 r0, F3 r1, F3 r2) {
    M3 r;
    r.m[0][0] = r0.x; r.m[0][1] = r0.y; r.m[0][2] = r0.z;
    r.m[1][0] = r1.x; r.m[1][1] = r1.y; r.m[1][2] = r1.z;
    r.m[2][0] = r2.x; r.m[2][1] = r2.y; r.m[2][2] = r2.z;
    return r;
}
DR_INLINE M3 mul(const M3 &a, const M3 &b) {
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float s = 0.f;
#pragma unroll
            for (int k = 0; k < 3; k++) s += a.m[i][k] * b.m[k][j];
            r.m[i][j] = s;
        }
    return r;
}
DR_INLINE F3 mul(const M3 &a, F3 v) {
    F3 r;
    r.x = 0.f; r.x += a.m[0][0] * v.x; r.x += a.m[0][1] * v.y; r.x += a.m[0][2] * v.z;
    r.y = 0.f; r.y += a.m[1][0] * v.x; r.y += a.m[1][1] * v.y; r.y += a.m[1][2] * v.z;
    r.z = 0.f; r.z += a.m[2][0] * v.x; r.z += a.m[2][1] * v.y; r.z += a.m[2][2] * v.z;
    return r;
}
// vector.h:646-666
DR_INLINE M3 inverse(const M3 &a) {
    const float(*m)[3] = a.m;
    float det = m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) -
                m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    float id = 1.f / det;
    M3 r;
    r.m[0][0] = (m[1][1] * m[2][2] - m[2][1] * m[1][2]) * id;
    r.m[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
    r.m[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
    r.m[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * id;
    r.m[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
    r.m[1][2] = (m[1][0] * m[0][2] - m[0][0] * m[1][2]) * id;
    r.m[2][0] = (m[1][0] * m[2][1] - m[2][0] * m[1][1]) * id;
    r.m[2][1] = (m[2][0] * m[0][1] - m[0][0] * m[2][1]) * id;
    r.m[2][2] = (m[0][0] * m[1][1] - m[1][0] * m[0][1]) * id;
    return r;
}
DR_INLINE M3 normalize_jacobian(F3 x) {
    float dsq = length_sq(x);
    float d = sqrtf(dsq);
    float inv = 1.f / (d * dsq);
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) r.m[i][j] = ((i == j ? dsq : 0.f) - x[i] * x[j]) * inv;
    return r;
}
DR_INLINE M3 intersect_jacobian(F3 dir, F3 p, F3 n) {
    M3 r;
    float d = -dot(p, n);
    float dn = dot(dir, n);
    if (fabsf(dn) < 1e-10f) {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) r.m[i][j] = 0.f;
        return r;
    }
    float inv2 = 1.f / (dn * dn);
    float t = -d / dn;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) r.m[i][j] = (i == j ? t : 0.f) + dir[i] * d * n[j] * inv2;
    return r;
}

// row-major 4x4 transforms (transform.h:9-21, 60-69)
DR_INLINE F3 xfm_point(const float *a, F3 p) {
    float x = a[0] * p.x + a[1] * p.y + a[2] * p.z + a[3];
    float y = a[4] * p.x + a[5] * p.y + a[6] * p.z + a[7];
    float z = a[8] * p.x + a[9] * p.y + a[10] * p.z + a[11];
    float w = a[12] * p.x + a[13] * p.y + a[14] * p.z + a[15];
    float inv = 1.f / w;
    return F3{x * inv, y * inv, z * inv};
}
DR_INLINE F3 xfm_vector(const float *a, F3 v) {
    return F3{a[0] * v.x + a[1] * v.y + a[2] * v.z, a[4] * v.x + a[5] * v.y + a[6] * v.z,
              a[8] * v.x + a[9] * v.y + a[10] * v.z};
}
// adjoint of xfm_point w.r.t. the matrix (transform.h:23-58), accumulated
DR_INLINE void d_xfm_point(const float *a, F3 p, F3 dout, float *dm) {
    float t[4];
#pragma unroll
    for (int i = 0; i < 4; i++) t[i] = a[i * 4] * p.x + a[i * 4 + 1] * p.y + a[i * 4 + 2] * p.z + a[i * 4 + 3];
    float inv = 1.f / t[3];
    float dt[4];
    dt[0] = dout.x * inv;
    dt[1] = dout.y * inv;
    dt[2] = dout.z * inv;
    float dinv = dout.x * t[0] + dout.y * t[1] + dout.z * t[2];
    dt[3] = -dinv * inv * inv;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        dm[i * 4 + 0] += dt[i] * p.x;
        dm[i * 4 + 1] += dt[i] * p.y;
        dm[i * 4 + 2] += dt[i] * p.z;
        dm[i * 4 + 3] += dt[i];
    }
}

// -------------------------------------------------------------- random ----
// Philox4x32-10 (Salmon et al., SC 2011).  Streams are keyed by the render
// seed and counted by (sample id, pass, sub-stream, block); the CPU oracle
// uses the same specification, so both draw identical numbers.
// The four outputs of a block are held in b0..b3 and shifted down as they are
// drawn (no indexed array: the state stays in registers).
struct Philox {
    uint32_t k0, k1, c0, c1, c2, c3;
    uint32_t b0, b1, b2, b3;
    int idx;
};
DR_INLINE void philox_init(Philox &r, uint64_t seed, uint32_t pass, uint64_t sample, uint32_t stream) {
    r.k0 = (uint32_t)seed;
    r.k1 = (uint32_t)(seed >> 32);
    r.c0 = (uint32_t)sample;
    r.c1 = (uint32_t)(sample >> 32) ^ (pass << 24);
    r.c2 = stream;
    r.c3 = 0;
    r.idx = 4;
}
DR_INLINE void philox_refill(Philox &r) {
    uint32_t x0 = r.c0, x1 = r.c1, x2 = r.c2, x3 = r.c3;
    uint32_t a = r.k0, b = r.k1;
#pragma unroll
    for (int i = 0; i < 10; i++) {
        uint32_t lo0 = 0xD2511F53u * x0, hi0 = __umulhi(0xD2511F53u, x0);
        uint32_t lo1 = 0xCD9E8D57u * x2, hi1 = __umulhi(0xCD9E8D57u, x2);
        uint32_t n0 = hi1 ^ x1 ^ a;
        uint32_t n2 = hi0 ^ x3 ^ b;
        x0 = n0; x1 = lo1; x2 = n2; x3 = lo0;
        a += 0x9E3779B9u;
        b += 0xBB67AE85u;
    }
    r.b0 = x0; r.b1 = x1; r.b2 = x2; r.b3 = x3;
    r.c3++;
    r.idx = 0;
}
// the first two outputs of counter block `block` of r's stream, as uniforms
// (independent of the draws made from r so far; r is not advanced)
DR_INLINE void philox_keyed2(const Philox &r, uint32_t block, float &u0, float &u1) {
    Philox t = r;
    t.c3 = block;
    philox_refill(t);
    u0 = (float)(t.b0 >> 8) * (1.0f / 16777216.0f);
    u1 = (float)(t.b1 >> 8) * (1.0f / 16777216.0f);
}
DR_INLINE float philox_next(Philox &r) {
    if (r.idx >= 4) philox_refill(r);
    uint32_t v = r.b0;
    r.b0 = r.b1;
    r.b1 = r.b2;
    r.b2 = r.b3;
    r.idx++;
    return (float)(v >> 8) * (1.0f / 16777216.0f);
}
#define DR_STREAM_MAIN 0u
DR_INLINE uint32_t dr_stream_point(int b) { return 1u + (uint32_t)b; }
// the edge-tree decisions at the point of bounce b, keyed by node: block n
// holds node n's draws (output 0: the cone pass, 1: the "all" pass), so a
// node's decisions do not depend on the order in which the tree is walked
DR_INLINE uint32_t dr_stream_point_tree(int b) { return 0x100u + (uint32_t)b; }
// the choice between the plain and the learned edge proposal at that point
DR_INLINE uint32_t dr_stream_point_mix(int b) { return 0x200u + (uint32_t)b; }
// the selected edge of ordinal j at bounce b (j: the edge's index in its
// tree, below 2^27; bit 30: the "all" pass): bit 31 marks an edge stream,
// bits 28-30 the bounce, bit 27 the pass, bits 0-26 the index
DR_INLINE uint32_t dr_stream_edge(int b, int j) {
    return 0x80000000u | ((uint32_t)b << 28) | ((((uint32_t)j >> 30) & 1u) << 27) | ((uint32_t)j & 0x7ffffffu);
}
