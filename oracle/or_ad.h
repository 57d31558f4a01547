// ORACLE — test infrastructure only.  Nothing in the product links this.
//
// Reverse-mode tape used as the interior-gradient oracle.  It restates the
// semantics of the reference's tape (autodiff.h:28-358): every arithmetic
// result is a new variable; each variable remembers up to two parents with
// local partial derivatives; a sweep in reverse creation order accumulates
// adjoints.  Semantics that change results and are reproduced on purpose:
//   - a partial derivative that is exactly 0 is not recorded (AddEdge,
//     autodiff.h:77-82), so an inf/NaN adjoint never crosses it;
//   - a variable whose adjoint is exactly 0 does not propagate
//     (propagate_adjoint, autodiff.h:350-358);
//   - set_adjoint OVERWRITES (autodiff.h:340-342);
//   - fabs / max / min pick one operand by value, max(x, c) and min(x, c)
//     with a float c return a fresh constant when c wins (autodiff.h:295-324);
//   - pow(x, a) with variable a has d/da = x^a * log(max(x, 1e-6))
//     (autodiff.h:225-231).
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

namespace oracle {

struct TapeNode {
    int p0, p1;
    float w0, w1;
};

struct Tape {
    std::vector<TapeNode> nodes;
    std::vector<float> adj;
    void clear() { nodes.clear(); }
    void zero_adjoints() { adj.assign(nodes.size(), 0.f); }
    int push(int p0, float w0, int p1, float w1) {
        TapeNode n;
        n.p0 = (w0 != 0.f) ? p0 : -1;
        n.w0 = w0;
        n.p1 = (w1 != 0.f) ? p1 : -1;
        n.w1 = w1;
        nodes.push_back(n);
        return (int)nodes.size() - 1;
    }
    void propagate() {
        for (int i = (int)nodes.size() - 1; i >= 0; i--) {
            float a = adj[i];
            if (a == 0.f) continue;
            const TapeNode &n = nodes[i];
            if (n.p0 >= 0) adj[n.p0] += n.w0 * a;
            if (n.p1 >= 0) adj[n.p1] += n.w1 * a;
        }
    }
};

extern thread_local Tape *g_tape;

struct ADR {
    float v;
    int id;
    ADR() : v(0.f), id(-1) {}
    ADR(float x) : v(x), id(g_tape->push(-1, 0.f, -1, 0.f)) {}
    static ADR make(float x, int p0, float w0, int p1 = -1, float w1 = 0.f) {
        ADR r;
        r.v = x;
        r.id = g_tape->push(p0, w0, p1, w1);
        return r;
    }
};

inline float val(float x) { return x; }
inline float val(const ADR &x) { return x.v; }

inline void set_adjoint(const ADR &x, float a) {
    if (x.id >= 0) g_tape->adj[x.id] = a;
}
inline float get_adjoint(const ADR &x) {
    return x.id >= 0 ? g_tape->adj[x.id] : 0.f;
}

inline ADR operator+(const ADR &a, const ADR &b) { return ADR::make(a.v + b.v, a.id, 1.f, b.id, 1.f); }
inline ADR operator+(const ADR &a, float b) { return ADR::make(a.v + b, a.id, 1.f); }
inline ADR operator+(float a, const ADR &b) { return ADR::make(a + b.v, b.id, 1.f); }
inline ADR operator-(const ADR &a, const ADR &b) { return ADR::make(a.v - b.v, a.id, 1.f, b.id, -1.f); }
inline ADR operator-(const ADR &a, float b) { return ADR::make(a.v - b, a.id, 1.f); }
inline ADR operator-(float a, const ADR &b) { return ADR::make(a - b.v, b.id, -1.f); }
inline ADR operator-(const ADR &a) { return ADR::make(-a.v, a.id, -1.f); }
inline ADR operator*(const ADR &a, const ADR &b) { return ADR::make(a.v * b.v, a.id, b.v, b.id, a.v); }
inline ADR operator*(const ADR &a, float b) { return ADR::make(a.v * b, a.id, b); }
inline ADR operator*(float a, const ADR &b) { return ADR::make(a * b.v, b.id, a); }
inline ADR operator/(const ADR &a, const ADR &b) {
    float inv = 1.f / b.v;
    return ADR::make(a.v * inv, a.id, inv, b.id, -a.v * inv * inv);
}
inline ADR operator/(const ADR &a, float b) { return a * (1.f / b); }
inline ADR operator/(float a, const ADR &b) {
    float inv = 1.f / b.v;
    return ADR::make(a * inv, b.id, -a * inv * inv);
}
inline ADR &operator+=(ADR &a, const ADR &b) { return a = a + b; }
inline ADR &operator-=(ADR &a, const ADR &b) { return a = a - b; }
inline ADR &operator*=(ADR &a, const ADR &b) { return a = a * b; }
inline ADR &operator/=(ADR &a, const ADR &b) { return a = a / b; }
inline ADR &operator*=(ADR &a, float b) { return a = a * b; }
inline ADR &operator/=(ADR &a, float b) { return a = a / b; }

inline bool operator<(const ADR &a, float b) { return a.v < b; }
inline bool operator<=(const ADR &a, float b) { return a.v <= b; }
inline bool operator>(const ADR &a, float b) { return a.v > b; }
inline bool operator>=(const ADR &a, float b) { return a.v >= b; }
inline bool operator==(const ADR &a, float b) { return a.v == b; }
inline bool operator<(const ADR &a, const ADR &b) { return a.v < b.v; }
inline bool operator>(const ADR &a, const ADR &b) { return a.v > b.v; }

inline ADR sqrt(const ADR &x) {
    float s = std::sqrt(x.v);
    return ADR::make(s, x.id, 0.5f * (1.f / s));
}
inline ADR square(const ADR &x) { return ADR::make(x.v * x.v, x.id, 2.f * x.v); }
inline float square(float x) { return x * x; }
inline ADR pow(const ADR &x, float a) {
    return ADR::make(std::pow(x.v, a), x.id, a * std::pow(x.v, a - 1.f));
}
inline ADR pow(const ADR &x, const ADR &a) {
    float p = std::pow(x.v, a.v);
    return ADR::make(p, x.id, a.v * std::pow(x.v, a.v - 1.f), a.id,
                     p * std::log(std::max(x.v, 1e-6f)));
}
// The reference calls pow(float, AReal) through the implicit float->AReal
// constructor, i.e. pow(AReal, AReal) with a fresh leaf as base.
inline ADR pow(float x, const ADR &a) { return pow(ADR(x), a); }
inline ADR sin(const ADR &x) { return ADR::make(std::sin(x.v), x.id, std::cos(x.v)); }
inline ADR cos(const ADR &x) { return ADR::make(std::cos(x.v), x.id, -std::sin(x.v)); }
inline ADR acos(const ADR &x) {
    return ADR::make(std::acos(x.v), x.id, -std::sqrt(1.f / (1.f - x.v * x.v)));
}
inline ADR atan2(const ADR &y, const ADR &x) {
    float t = x.v * x.v + y.v * y.v;
    return ADR::make(std::atan2(y.v, x.v), y.id, x.v / t, x.id, -y.v / t);
}
inline ADR fabs(const ADR &x) { return x.v >= 0.f ? x : -x; }
inline ADR max(const ADR &x, const ADR &y) { return x.v > y.v ? x : y; }
inline ADR max(const ADR &x, float y) { return x.v >= y ? x : ADR(y); }
inline ADR min(const ADR &x, float y) { return x.v <= y ? x : ADR(y); }
inline float max(float x, float y) { return x > y ? x : y; }
inline float sqrt(float x) { return std::sqrt(x); }
inline float pow(float x, float a) { return std::pow(x, a); }
inline float fabs(float x) { return std::fabs(x); }
inline float sin(float x) { return std::sin(x); }
inline float cos(float x) { return std::cos(x); }
inline float acos(float x) { return std::acos(x); }
inline float min(float x, float y) { return x < y ? x : y; }

template <class TOut>
inline TOut conv(float x) { return TOut(x); }
template <class TOut>
inline TOut conv(const ADR &x);
template <>
inline float conv<float>(const ADR &x) { return x.v; }
template <>
inline ADR conv<ADR>(const ADR &x) { return x; }

}  // namespace oracle
