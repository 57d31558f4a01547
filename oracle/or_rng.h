// ORACLE — test infrastructure only.  Nothing in the product links this.
//
// Counter-based random numbers shared (as a specification, not as code) with
// the device renderer, so that the CPU oracle and the GPU kernels draw the
// same uniforms for the same (seed, pass, sample, stream) and their per-sample
// results can be compared directly.  Generator: Philox4x32-10 (Salmon, Moraes,
// Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC 2011).
// Uniform in [0, 1): top 24 bits * 2^-24.
//
// The reference draws from per-tile std::mt19937 streams
// (delta_ray.cpp:1183-1191, 1473) — not reproducible on a GPU and dependent
// on the core count; only the distribution of the draws matters for the
// estimator, and that is the same.
#pragma once

#include <cstdint>

namespace oracle {

struct Philox {
    uint32_t k0, k1;
    uint32_t c0, c1, c2, c3;
    uint32_t buf[4];
    int idx;

    static inline void mulhilo(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
        uint64_t p = (uint64_t)a * (uint64_t)b;
        hi = (uint32_t)(p >> 32);
        lo = (uint32_t)p;
    }
    void refill() {
        uint32_t x0 = c0, x1 = c1, x2 = c2, x3 = c3;
        uint32_t a = k0, b = k1;
        for (int r = 0; r < 10; r++) {
            uint32_t hi0, lo0, hi1, lo1;
            mulhilo(0xD2511F53u, x0, hi0, lo0);
            mulhilo(0xCD9E8D57u, x2, hi1, lo1);
            uint32_t n0 = hi1 ^ x1 ^ a;
            uint32_t n1 = lo1;
            uint32_t n2 = hi0 ^ x3 ^ b;
            uint32_t n3 = lo0;
            x0 = n0; x1 = n1; x2 = n2; x3 = n3;
            a += 0x9E3779B9u;
            b += 0xBB67AE85u;
        }
        buf[0] = x0; buf[1] = x1; buf[2] = x2; buf[3] = x3;
        c3++;
        idx = 0;
    }
    // pass: 0 = pixel samples, 1 = primary-edge samples.
    // stream: sub-stream of one sample (see the stream constants below).
    Philox(uint64_t seed, uint32_t pass, uint64_t sample, uint32_t stream) {
        k0 = (uint32_t)seed;
        k1 = (uint32_t)(seed >> 32);
        c0 = (uint32_t)sample;
        c1 = (uint32_t)(sample >> 32) ^ (pass << 24);
        c2 = stream;
        c3 = 0;
        idx = 4;
    }
    uint32_t next_u32() {
        if (idx >= 4) refill();
        return buf[idx++];
    }
    float next() { return (float)(next_u32() >> 8) * (1.0f / 16777216.0f); }
    // output k (0-3) of counter block `block`, independent of the draws made
    // so far (keyed draws: block = an edge-tree node)
    float keyed(uint32_t block, int k) const {
        Philox t = *this;
        t.c3 = block;
        t.refill();
        return (float)(t.buf[k] >> 8) * (1.0f / 16777216.0f);
    }
};

// Sub-streams of one pixel sample.
//   STREAM_MAIN: jitter (2), light sample (4), per bounce: light sample (4),
//                bsdf sample (3), russian roulette (1)
//   stream_point(b): secondary shading point at bounce b: edge sample (2),
//                edge-tree choice (1)
//   stream_point_tree(b): the edge-tree decisions at that point, keyed by
//                node: block n holds node n's draws, output 0 for the cone
//                pass, output 1 for the "all" pass (so a node's decisions do
//                not depend on the order in which the tree is walked)
//   stream_edge(b, j): the selected edge of ordinal j at bounce b (j: the
//                edge's index in its tree, bit 30 for the "all" pass): light
//                sample (4), per remaining bounce (8)
// A primary-edge sample uses STREAM_MAIN of pass 1: edge choice (1), edge
// sample (2), light sample (4), per bounce (8).
static const uint32_t STREAM_MAIN = 0;
inline uint32_t stream_point(int b) { return 1u + (uint32_t)b; }
inline uint32_t stream_point_tree(int b) { return 0x100u + (uint32_t)b; }
// bit 31: an edge stream; bits 28-30: bounce; bit 27: the "all" pass (bit 30
// of the ordinal); bits 0-26: the edge's index in its tree
inline uint32_t stream_edge(int b, int j) {
    return 0x80000000u | ((uint32_t)b << 28) | ((((uint32_t)j >> 30) & 1u) << 27) | ((uint32_t)j & 0x7ffffffu);
}

}  // namespace oracle
