// CPU oracle for post-processing — TEST INFRASTRUCTURE ONLY.
//
// Serial restatement of OpenCL_PathTracer/OpenCL/post_processing.cl in the
// reference's own scatter form: every source pixel (row-major order) adds its
// taps at clamped coordinates, as the kernel's atomic_add_f32 calls do
// (extra_atomics.cl:12-21), here in a fixed order. Loaded by tests/ only.
// Pixels are float[4] (clr_float3). -ffp-contract=off.
//
// Pinned in tests/test_oracle_post.py by closed forms (impulse responses,
// clamped-border mass conservation, the tone curve at chosen values) and
// layout invariants (BGR order, bottom-up rows, zero padding).
#include <algorithm>
#include <cmath>
#include <cstdint>

namespace {
const float kGauss9x9[81] = {
    0.000763447f, 0.001831415f, 0.003421534f, 0.004978302f, 0.005641155f, 0.004978302f, 0.003421534f, 0.001831415f, 0.000763447f,
    0.001831415f, 0.004393336f, 0.008207832f, 0.011942326f, 0.013532428f, 0.011942326f, 0.008207832f, 0.004393336f, 0.001831415f,
    0.003421534f, 0.008207832f, 0.01533425f,  0.022311201f, 0.025281903f, 0.022311201f, 0.01533425f,  0.008207832f, 0.003421534f,
    0.004978302f, 0.011942326f, 0.022311201f, 0.032462606f, 0.036784952f, 0.032462606f, 0.022311201f, 0.011942326f, 0.004978302f,
    0.005641155f, 0.013532428f, 0.025281903f, 0.036784952f, 0.041682812f, 0.036784952f, 0.025281903f, 0.013532428f, 0.005641155f,
    0.004978302f, 0.011942326f, 0.022311201f, 0.032462606f, 0.036784952f, 0.032462606f, 0.022311201f, 0.011942326f, 0.004978302f,
    0.003421534f, 0.008207832f, 0.01533425f,  0.022311201f, 0.025281903f, 0.022311201f, 0.01533425f,  0.008207832f, 0.003421534f,
    0.001831415f, 0.004393336f, 0.008207832f, 0.011942326f, 0.013532428f, 0.011942326f, 0.008207832f, 0.004393336f, 0.001831415f,
    0.000763447f, 0.001831415f, 0.003421534f, 0.004978302f, 0.005641155f, 0.004978302f, 0.003421534f, 0.001831415f, 0.000763447f,
};
const float kThreshold = 30.0f;
inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
}  // namespace

extern "C" {

// scaling (post_processing.cl:20-26)
int orc_post_scaling(uint32_t w, uint32_t h, uint32_t spp, const float* src, float* out) {
    for (size_t i = 0; i < (size_t)w * h; ++i)
        for (int c = 0; c < 3; ++c) out[4 * i + c] = src[4 * i + c] * (1.0f / spp);
    return 0;
}

// gaussianScattering (post_processing.cl:47-66)
int orc_post_gaussian(uint32_t w, uint32_t h, uint32_t spp, const float* src, float* out) {
    const int R = 4;
    for (int y = 0; y < (int)h; ++y)
        for (int x = 0; x < (int)w; ++x) {
            const float* s = src + 4 * ((size_t)y * w + x);
            for (int i = -R; i <= R; ++i)
                for (int j = -R; j <= R; ++j) {
                    const int px = clampi(x + j, 0, (int)w - 1), py = clampi(y + i, 0, (int)h - 1);
                    float* d = out + 4 * ((size_t)py * w + px);
                    const float f = kGauss9x9[(2 * R + 1) * (i + R) + (j + R)] * (1.0f / spp);
                    for (int c = 0; c < 3; ++c) d[c] += f * s[c];
                }
        }
    return 0;
}

// bloom (post_processing.cl:70-101)
int orc_post_bloom(uint32_t w, uint32_t h, uint32_t spp, const float* src, float* out) {
    for (int y = 0; y < (int)h; ++y)
        for (int x = 0; x < (int)w; ++x) {
            const float* s = src + 4 * ((size_t)y * w + x);
            for (int c = 0; c < 3; ++c) {
                const float v = s[c] * (1.0f / spp);
                if (v > kThreshold) {
                    const float excess = v - kThreshold;
                    const int r = excess > 7.0f ? 7 : (int)excess;
                    for (int dy = -r; dy <= r; ++dy)
                        for (int dx = -r; dx <= r; ++dx) {
                            const float dist = std::sqrt((float)(dx * dx + dy * dy));
                            const int px = clampi(x + dx, 0, (int)w - 1), py = clampi(y + dy, 0, (int)h - 1);
                            float* d = out + 4 * ((size_t)py * w + px);
                            if (dx != 0 || dy != 0) d[c] += std::exp(-dist) * excess;
                            else d[c] += kThreshold;
                        }
                } else {
                    out[4 * ((size_t)y * w + x) + c] += v;
                }
            }
        }
    return 0;
}

// toneMapping (post_processing.cl:103-117), luminance and scale in float.
int orc_post_tone(uint32_t w, uint32_t h, uint32_t rowBytes, const float* src, uint8_t* dst) {
    for (uint32_t y = 0; y < h; ++y) {
        uint8_t* row = dst + (size_t)rowBytes * (h - 1 - y);
        for (uint32_t x = 0; x < w; ++x) {
            const float* p = src + 4 * ((size_t)y * w + x);
            float s[3];
            for (int c = 0; c < 3; ++c) s[c] = p[c] > kThreshold ? kThreshold : p[c];
            const float Y = 0.2989f * s[0] + 0.5866f * s[1] + 0.1144f * s[2];
            float scale = (1.0f - std::exp(-Y)) / Y;
            if (Y == 0.0f) scale = 1.0f;
            uint8_t b[3];
            for (int c = 0; c < 3; ++c) {
                const float v = 255.0f * std::pow(scale * s[c], 1.0f / 2.2f);
                const float cl = std::fmin(std::fmax(v, 0.0f), 255.0f);
                b[c] = (uint8_t)(uint32_t)cl;
            }
            row[3 * x + 0] = b[2];
            row[3 * x + 1] = b[1];
            row[3 * x + 2] = b[0];
        }
        for (uint32_t k = 3 * w; k < rowBytes; ++k) row[k] = 0;
    }
    return 0;
}

}  // extern "C"
