// Host-side procedural scene generators (declared in include/clear_amd_scenes.h).
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/clear_amd_scenes.h"

namespace {

// xorshift128 with the reference host seeding (XORShift.hpp:23-35,
// XORShift.cpp:22-29): signed seed expansion, then 50 discarded draws.
class XorShift128 {
    uint32_t s[4];

public:
    explicit XorShift128(int seed) {
        for (uint32_t i = 0; i < 4; ++i) {
            // 1812433253U * (seed ^ (seed >> 30)) + i, stored back into an int.
            uint32_t v = 1812433253u * (uint32_t)(seed ^ (seed >> 30)) + i;
            s[i] = v;
            seed = (int)v;
        }
        for (int i = 0; i < 50; ++i) next();
    }
    uint32_t next() {
        uint32_t t = s[0] ^ (s[0] << 11);
        s[0] = s[1];
        s[1] = s[2];
        s[2] = s[3];
        return s[3] = (s[3] ^ (s[3] >> 19)) ^ (t ^ (t >> 8));
    }
    // getFloat0cTo1o: 23 mantissa bits, [0, 1).
    float uniform() {
        uint32_t b = (next() >> 9) | 0x3f800000u;
        float f;
        std::memcpy(&f, &b, 4);
        return f - 1.0f;
    }
};

constexpr float kPi = 3.14159265358979323846f;

void unitDir(XorShift128& r, float* d) {
    float z = 1.0f - 2.0f * r.uniform();
    float phi = 2.0f * kPi * r.uniform();
    float rr = std::sqrt(std::fmax(0.0f, 1.0f - z * z));
    d[0] = rr * std::cos(phi);
    d[1] = rr * std::sin(phi);
    d[2] = z;
}

}  // namespace

extern "C" {

int clr_gen_cluster_soup(uint32_t numTris, int seed, uint32_t numClusters, float sigma,
                         float eMin, float eMax, float* out) {
    if (!out || numClusters == 0) return 1;
    XorShift128 rng(seed);
    float* centres = new float[3 * (size_t)numClusters];
    for (uint32_t c = 0; c < 3 * numClusters; ++c) centres[c] = rng.uniform();
    for (uint32_t i = 0; i < numTris; ++i) {
        uint32_t c = (uint32_t)(rng.uniform() * (float)numClusters);
        if (c >= numClusters) c = numClusters - 1;
        float g[4];
        for (int k = 0; k < 2; ++k) {
            float u1 = 1.0f - rng.uniform();  // (0, 1]
            float u2 = rng.uniform();
            float rad = std::sqrt(-2.0f * std::log(u1));
            g[2 * k] = rad * std::cos(2.0f * kPi * u2);
            g[2 * k + 1] = rad * std::sin(2.0f * kPi * u2);
        }
        float ctr[3];
        for (int k = 0; k < 3; ++k) ctr[k] = centres[3 * c + k] + sigma * g[k];
        float e1[3], e2[3];
        unitDir(rng, e1);
        unitDir(rng, e2);
        float l1 = eMin + (eMax - eMin) * rng.uniform();
        float l2 = eMin + (eMax - eMin) * rng.uniform();
        float* t = out + 9 * (size_t)i;
        for (int k = 0; k < 3; ++k) {
            e1[k] *= l1;
            e2[k] *= l2;
            float v0 = ctr[k] - (e1[k] + e2[k]) / 3.0f;
            t[k] = v0;
            t[3 + k] = v0 + e1[k];
            t[6 + k] = v0 + e2[k];
        }
    }
    delete[] centres;
    return 0;
}

int clr_gen_random_rays(uint32_t numRays, int seed, float* out) {
    if (!out) return 1;
    XorShift128 rng(seed);
    const size_t R = numRays;
    for (uint32_t i = 0; i < numRays; ++i) {
        out[i] = rng.uniform();
        out[R + i] = rng.uniform();
        out[2 * R + i] = rng.uniform();
        float d[3];
        unitDir(rng, d);
        out[3 * R + i] = d[0];
        out[4 * R + i] = d[1];
        out[5 * R + i] = d[2];
        out[6 * R + i] = INFINITY;
    }
    return 0;
}

}  // extern "C"
