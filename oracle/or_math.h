// ORACLE — test infrastructure only.  Nothing in the product links this.
//
// Small vector / frame / matrix helpers templated on the scalar (float or the
// tape variable ADR).  Formulas follow the reference's vector.h (cross,
// dot, normalize, luminance: 451-456, make_frame: 485-501, cos_hemisphere:
// 540-545, normalize_jacobian: 718-733, intersect_jacobian: 690-716,
// inverse: 646-666) and transform.h (xfm_point / xfm_vector / d_xfm_point).
#pragma once

#include "or_ad.h"

#include <cmath>
#include <cstring>

namespace oracle {

template <class T>
struct V2 {
    T e[2];
    T &operator[](int i) { return e[i]; }
    const T &operator[](int i) const { return e[i]; }
};
template <class T>
struct V3 {
    T e[3];
    T &operator[](int i) { return e[i]; }
    const T &operator[](int i) const { return e[i]; }
};
using V2f = V2<float>;
using V3f = V3<float>;

template <class T>
inline V2<T> v2(const T &a, const T &b) { return V2<T>{{a, b}}; }
template <class T>
inline V3<T> v3(const T &a, const T &b, const T &c) { return V3<T>{{a, b, c}}; }

template <class TO, class TI>
inline V3<TO> conv3(const V3<TI> &v) { return v3(conv<TO>(v[0]), conv<TO>(v[1]), conv<TO>(v[2])); }
template <class TO, class TI>
inline V2<TO> conv2(const V2<TI> &v) { return v2(conv<TO>(v[0]), conv<TO>(v[1])); }

template <class T> inline V3<T> operator+(const V3<T> &a, const V3<T> &b) { return v3<T>(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
template <class T> inline V3<T> operator-(const V3<T> &a, const V3<T> &b) { return v3<T>(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
template <class T> inline V3<T> operator-(const V3<T> &a) { return v3<T>(-a[0], -a[1], -a[2]); }
template <class T, class S> inline V3<T> operator*(const V3<T> &a, const S &s) { return v3<T>(a[0] * s, a[1] * s, a[2] * s); }
template <class T, class S> inline V3<T> operator*(const S &s, const V3<T> &a) { return v3<T>(s * a[0], s * a[1], s * a[2]); }
template <class T> inline V3<T> operator*(const V3<T> &a, const V3<T> &b) { return v3<T>(a[0] * b[0], a[1] * b[1], a[2] * b[2]); }
template <class T, class S> inline V3<T> operator/(const V3<T> &a, const S &s) { auto inv = 1.f / s; return a * inv; }
template <class T> inline V3<T> &operator+=(V3<T> &a, const V3<T> &b) { return a = a + b; }
template <class T, class S> inline V3<T> &operator*=(V3<T> &a, const S &s) { return a = a * s; }
template <class T> inline V3<T> &operator*=(V3<T> &a, const V3<T> &b) { return a = a * b; }
template <class T, class S> inline V3<T> &operator/=(V3<T> &a, const S &s) { return a = a / s; }

template <class T> inline V2<T> operator+(const V2<T> &a, const V2<T> &b) { return v2<T>(a[0] + b[0], a[1] + b[1]); }
template <class T> inline V2<T> operator-(const V2<T> &a, const V2<T> &b) { return v2<T>(a[0] - b[0], a[1] - b[1]); }
template <class T, class S> inline V2<T> operator*(const V2<T> &a, const S &s) { return v2<T>(a[0] * s, a[1] * s); }
template <class T, class S> inline V2<T> operator*(const S &s, const V2<T> &a) { return v2<T>(s * a[0], s * a[1]); }
template <class T> inline V2<T> operator*(const V2<T> &a, const V2<T> &b) { return v2<T>(a[0] * b[0], a[1] * b[1]); }
template <class T, class S> inline V2<T> operator/(const V2<T> &a, const S &s) { return v2<T>(a[0] / s, a[1] / s); }

template <class T> inline T dot(const V3<T> &a, const V3<T> &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
template <class T> inline T dot2(const V2<T> &a, const V2<T> &b) { return a[0] * b[0] + a[1] * b[1]; }
template <class T> inline V3<T> cross(const V3<T> &a, const V3<T> &b) {
    return v3<T>(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
template <class T> inline T length_squared(const V3<T> &a) { return square(a[0]) + square(a[1]) + square(a[2]); }
template <class T> inline T length(const V3<T> &a) { using std::sqrt; return sqrt(length_squared(a)); }
inline float length2(const V2f &a) { return std::sqrt(a[0] * a[0] + a[1] * a[1]); }
template <class T> inline V3<T> normalize(const V3<T> &a) { return a / length(a); }
template <class T> inline T sum(const V3<T> &a) { return a[0] + a[1] + a[2]; }
inline float abs_sum(const V3f &a) { return std::fabs(a[0]) + std::fabs(a[1]) + std::fabs(a[2]); }
template <class T> inline T luminance(const V3<T> &a) { return 0.212671f * a[0] + 0.715160f * a[1] + 0.072169f * a[2]; }
inline float maxc(const V3f &a) { return std::max(std::max(a[0], a[1]), a[2]); }
inline bool is_zero(const V3f &a) { return a[0] == 0.f && a[1] == 0.f && a[2] == 0.f; }
inline bool isnan3(const V3f &a) { return std::isnan(a[0]) || std::isnan(a[1]) || std::isnan(a[2]); }

template <class T>
struct Frame {
    V3<T> e[3];
    V3<T> &operator[](int i) { return e[i]; }
    const V3<T> &operator[](int i) const { return e[i]; }
};

template <class T>
inline Frame<T> make_frame(const V3<T> &n) {
    if (n[2] < -1.f + 1e-6f) {
        return Frame<T>{{v3<T>(T(0.f), T(-1.f), T(0.f)), v3<T>(T(-1.f), T(0.f), T(0.f)), n}};
    }
    T a = 1.f / (1.f + n[2]);
    T b = -n[0] * n[1] * a;
    return Frame<T>{{v3<T>(1.f - square(n[0]) * a, b, -n[0]),
                     v3<T>(b, 1.f - square(n[1]) * a, -n[1]), n}};
}
template <class T>
inline V3<T> to_local(const Frame<T> &f, const V3<T> &v) { return v3<T>(dot(v, f[0]), dot(v, f[1]), dot(v, f[2])); }
template <class T>
inline V3<T> to_world(const Frame<T> &f, const V3<T> &v) { return f[0] * v[0] + f[1] * v[1] + f[2] * v[2]; }
template <class TO, class TI>
inline Frame<TO> conv_frame(const Frame<TI> &f) { return Frame<TO>{{conv3<TO>(f[0]), conv3<TO>(f[1]), conv3<TO>(f[2])}}; }

inline V3f cos_hemisphere(const V2f &s) {
    float phi = 2.f * float(M_PI) * s[0];
    float tmp = std::sqrt(1.f - s[1]);
    return v3(std::cos(phi) * tmp, std::sin(phi) * tmp, std::sqrt(s[1]));
}

struct M3 {
    float m[3][3];
    float &operator()(int i, int j) { return m[i][j]; }
    float operator()(int i, int j) const { return m[i][j]; }
};
inline M3 m3_from_frame(const Frame<float> &f) {
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i][j] = f[i][j];
    return r;
}
inline M3 operator*(const M3 &a, const M3 &b) {
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            r.m[i][j] = 0.f;
            for (int k = 0; k < 3; k++) r.m[i][j] += a.m[i][k] * b.m[k][j];
        }
    return r;
}
inline V3f operator*(const M3 &a, const V3f &v) {
    V3f r;
    for (int i = 0; i < 3; i++) {
        r[i] = 0.f;
        for (int j = 0; j < 3; j++) r[i] += a.m[i][j] * v[j];
    }
    return r;
}
inline M3 inverse(const M3 &m) {
    float det = m(0, 0) * (m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2)) -
                m(0, 1) * (m(1, 0) * m(2, 2) - m(1, 2) * m(2, 0)) +
                m(0, 2) * (m(1, 0) * m(2, 1) - m(1, 1) * m(2, 0));
    float id = 1.f / det;
    M3 r;
    r(0, 0) = (m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2)) * id;
    r(0, 1) = (m(0, 2) * m(2, 1) - m(0, 1) * m(2, 2)) * id;
    r(0, 2) = (m(0, 1) * m(1, 2) - m(0, 2) * m(1, 1)) * id;
    r(1, 0) = (m(1, 2) * m(2, 0) - m(1, 0) * m(2, 2)) * id;
    r(1, 1) = (m(0, 0) * m(2, 2) - m(0, 2) * m(2, 0)) * id;
    r(1, 2) = (m(1, 0) * m(0, 2) - m(0, 0) * m(1, 2)) * id;
    r(2, 0) = (m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1)) * id;
    r(2, 1) = (m(2, 0) * m(0, 1) - m(0, 0) * m(2, 1)) * id;
    r(2, 2) = (m(0, 0) * m(1, 1) - m(1, 0) * m(0, 1)) * id;
    return r;
}

// adjoint of normalize (vector.h:454-464, in its arithmetic)
inline V3f d_normalize(const V3f &v, const V3f &d_out) {
    float l = length(v);
    float d_inv_l = dot(d_out, v);
    float d_l = -d_inv_l / (l * l);
    return (d_out + d_l * v) / l;
}

// d normalize(x) / dx
inline M3 normalize_jacobian(const V3f &x) {
    float dsq = length_squared(x);
    float d = std::sqrt(dsq);
    float inv = 1.f / (d * dsq);
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i][j] = ((i == j ? dsq : 0.f) - x[i] * x[j]) * inv;
    return r;
}
// d (hit point on plane {p, n}) / d dir, for a ray from the origin
inline M3 intersect_jacobian(const V3f &dir, const V3f &p, const V3f &n) {
    M3 r;
    float d = -dot(p, n);
    float dn = dot(dir, n);
    if (std::fabs(dn) < 1e-10f) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) r.m[i][j] = 0.f;
        return r;
    }
    float inv2 = 1.f / (dn * dn);
    float t = -d / dn;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i][j] = (i == j ? t : 0.f) + dir[i] * d * n[j] * inv2;
    return r;
}

// 4x4 row-major transforms
struct M4 {
    float m[4][4];
};
template <class T>
struct M4T {
    T m[4][4];
};
template <class T>
inline V3<T> xfm_point(const M4T<T> &a, const V3<T> &p) {
    T x = a.m[0][0] * p[0] + a.m[0][1] * p[1] + a.m[0][2] * p[2] + a.m[0][3];
    T y = a.m[1][0] * p[0] + a.m[1][1] * p[1] + a.m[1][2] * p[2] + a.m[1][3];
    T z = a.m[2][0] * p[0] + a.m[2][1] * p[1] + a.m[2][2] * p[2] + a.m[2][3];
    T w = a.m[3][0] * p[0] + a.m[3][1] * p[1] + a.m[3][2] * p[2] + a.m[3][3];
    T inv = 1.f / w;
    return v3<T>(x * inv, y * inv, z * inv);
}
template <class T>
inline V3<T> xfm_vector(const M4T<T> &a, const V3<T> &v) {
    return v3<T>(a.m[0][0] * v[0] + a.m[0][1] * v[1] + a.m[0][2] * v[2],
                 a.m[1][0] * v[0] + a.m[1][1] * v[1] + a.m[1][2] * v[2],
                 a.m[2][0] * v[0] + a.m[2][1] * v[1] + a.m[2][2] * v[2]);
}
inline void d_xfm_point(const M4T<float> &a, const V3f &p, const V3f &dout, float dm[16]) {
    float t[4];
    for (int i = 0; i < 4; i++) t[i] = a.m[i][0] * p[0] + a.m[i][1] * p[1] + a.m[i][2] * p[2] + a.m[i][3];
    float inv = 1.f / t[3];
    float dt[4];
    for (int i = 0; i < 3; i++) dt[i] = dout[i] * inv;
    float dinv = dout[0] * t[0] + dout[1] * t[1] + dout[2] * t[2];
    dt[3] = -dinv * inv * inv;
    for (int i = 0; i < 4; i++) {
        dm[i * 4 + 0] += dt[i] * p[0];
        dm[i * 4 + 1] += dt[i] * p[1];
        dm[i * 4 + 2] += dt[i] * p[2];
        dm[i * 4 + 3] += dt[i];
    }
}

}  // namespace oracle
