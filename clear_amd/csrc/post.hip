// Post-processing of the accumulated image for gfx950: the kernels of
// OpenCL_PathTracer/OpenCL/post_processing.cl (clear, scaling,
// gaussianScattering, bloom, toneMapping) and the 24-bit BMP writer of
// BMPExporter.cpp:14-53.
//
// The reference's gaussianScattering and bloom scatter every source pixel into
// its clamped neighbourhood with float atomics (extra_atomics.cl:12-21), so
// their sums depend on arrival order. Here each output pixel gathers the same
// terms instead: a 16x16 tile of outputs stages its sources (halo R = 4 or 7)
// in LDS and adds, for every source q and offset d within the source's
// radius, the term the reference would add at clamp(q + d). Along one axis
// the offsets of a source qx that land on output px form one interval:
//   0 < px < W-1:  d = px - qx
//   px == 0:       d <= -qx          (everything pushed off the left edge)
//   px == W-1:     d >= W-1 - qx     (everything pushed off the right edge)
// (for W == 1 every offset lands on px = 0). Products are formed as the
// reference forms them (filter weight * (1 / spp) then * value); only the
// order of the additions differs. Results are deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "internal.h"

namespace clr {
namespace {

constexpr int kPostTile = 16;
constexpr float kBloomThreshold = 30.0f;  // `threshold`, post_processing.cl:69

__constant__ float c_gauss9x9[81] = {
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

// Offsets d in [-r, r] with clamp(q + d, 0, n - 1) == p, as [lo, hi]
// (empty when lo > hi).
__device__ __forceinline__ void landing(int q, int p, int n, int r, int& lo, int& hi) {
    if (n == 1) {
        lo = -r;
        hi = r;
    } else if (p == 0) {
        lo = -r;
        hi = min(r, -q);
    } else if (p == n - 1) {
        lo = max(-r, n - 1 - q);
        hi = r;
    } else {
        lo = hi = p - q;
        if (lo < -r || lo > r) hi = lo - 1;
    }
}

__global__ void k_post_clear(uint32_t n, float4* __restrict__ buf) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        buf[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__global__ void k_post_scaling(uint32_t n, float inv, const float4* __restrict__ src, float4* __restrict__ dst) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 v = src[i];
        dst[i] = make_float4(v.x * inv, v.y * inv, v.z * inv, 0.0f);
    }
}

// Stages the (16 + 2R)^2 sources of this tile (zero outside the image; those
// slots are never summed because landing() only admits in-image sources).
template <int R>
__device__ __forceinline__ void stageTile(const float4* __restrict__ src, int W, int H, int x0, int y0,
                                          float4* tile) {
    constexpr int S = kPostTile + 2 * R;
    for (int k = threadIdx.y * kPostTile + threadIdx.x; k < S * S; k += kPostTile * kPostTile) {
        const int qx = x0 - R + k % S, qy = y0 - R + k / S;
        tile[k] = (qx >= 0 && qx < W && qy >= 0 && qy < H) ? src[(size_t)qy * W + qx]
                                                          : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// gaussianScattering (post_processing.cl:47-66): output += the 9x9 Gaussian
// (weights * 1/spp) of src, each source's taps clamped to the image.
__global__ __launch_bounds__(kPostTile* kPostTile) void k_post_gaussian(int W, int H, float inv,
                                                                        const float4* __restrict__ src,
                                                                        float4* __restrict__ out) {
    constexpr int R = 4, S = kPostTile + 2 * R;
    __shared__ float4 tile[S * S];
    const int x0 = blockIdx.x * kPostTile, y0 = blockIdx.y * kPostTile;
    stageTile<R>(src, W, H, x0, y0, tile);
    __syncthreads();
    const int px = x0 + threadIdx.x, py = y0 + threadIdx.y;
    if (px >= W || py >= H) return;
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    for (int qy = max(py - R, 0); qy <= min(py + R, H - 1); ++qy) {
        int ylo, yhi;
        landing(qy, py, H, R, ylo, yhi);
        for (int qx = max(px - R, 0); qx <= min(px + R, W - 1); ++qx) {
            int xlo, xhi;
            landing(qx, px, W, R, xlo, xhi);
            const float4 v = tile[(qy - y0 + R) * S + (qx - x0 + R)];
            for (int i = ylo; i <= yhi; ++i)
                for (int j = xlo; j <= xhi; ++j) {
                    const float f = c_gauss9x9[9 * (i + R) + (j + R)] * inv;
                    ax += f * v.x;
                    ay += f * v.y;
                    az += f * v.z;
                }
        }
    }
    float4 o = out[(size_t)py * W + px];
    o.x += ax;
    o.y += ay;
    o.z += az;
    out[(size_t)py * W + px] = o;
}

// bloom (post_processing.cl:70-101), per channel: a value v = src * 1/spp
// above the threshold spreads exp(-|d|) * (v - 30) over the square of radius
// min(7, (int)(v - 30)) around it (30 at its own offset 0); otherwise v is
// added to its own pixel.
__global__ __launch_bounds__(kPostTile* kPostTile) void k_post_bloom(int W, int H, float inv,
                                                                     const float4* __restrict__ src,
                                                                     float4* __restrict__ out) {
    constexpr int R = 7, S = kPostTile + 2 * R;
    __shared__ float4 tile[S * S];
    const int x0 = blockIdx.x * kPostTile, y0 = blockIdx.y * kPostTile;
    stageTile<R>(src, W, H, x0, y0, tile);
    __syncthreads();
    const int px = x0 + threadIdx.x, py = y0 + threadIdx.y;
    if (px >= W || py >= H) return;
    float acc[3] = {0.0f, 0.0f, 0.0f};
    for (int qy = max(py - R, 0); qy <= min(py + R, H - 1); ++qy) {
        for (int qx = max(px - R, 0); qx <= min(px + R, W - 1); ++qx) {
            const float4 v4 = tile[(qy - y0 + R) * S + (qx - x0 + R)];
            const float vc[3] = {v4.x, v4.y, v4.z};
            for (int c = 0; c < 3; ++c) {
                const float v = vc[c] * inv;
                if (!(v > kBloomThreshold)) {
                    if (qx == px && qy == py) acc[c] += v;
                    continue;
                }
                const float excess = v - kBloomThreshold;
                const int r = excess > 7.0f ? 7 : (int)excess;
                int ylo, yhi, xlo, xhi;
                landing(qy, py, H, r, ylo, yhi);
                landing(qx, px, W, r, xlo, xhi);
                for (int dy = ylo; dy <= yhi; ++dy)
                    for (int dx = xlo; dx <= xhi; ++dx) {
                        if (dx != 0 || dy != 0) {
                            const float dist = sqrtf((float)(dx * dx + dy * dy));
                            acc[c] += expf(-dist) * excess;
                        } else {
                            acc[c] += kBloomThreshold;
                        }
                    }
            }
        }
    }
    float4 o = out[(size_t)py * W + px];
    o.x += acc[0];
    o.y += acc[1];
    o.z += acc[2];
    out[(size_t)py * W + px] = o;
}

// toneMapping (post_processing.cl:103-117): channels saturated at 30,
// Y = 0.2989 R + 0.5866 G + 0.1144 B, scale (1 - e^-Y) / Y (1 at Y = 0),
// bytes = trunc(clamp(255 * (scale * v)^(1/2.2), 0, 255)) stored B, G, R into
// bottom-up rows of rowBytes; each row's padding bytes are written as 0.
__global__ void k_post_tone(uint32_t W, uint32_t H, uint32_t rowBytes, const float4* __restrict__ src,
                            uint8_t* __restrict__ dst) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (y >= H) return;
    uint8_t* row = dst + (size_t)rowBytes * (H - 1 - y);
    if (x < W) {
        const float4 p = src[(size_t)y * W + x];
        const float s[3] = {p.x > kBloomThreshold ? kBloomThreshold : p.x, p.y > kBloomThreshold ? kBloomThreshold : p.y,
                            p.z > kBloomThreshold ? kBloomThreshold : p.z};
        const float Y = 0.2989f * s[0] + 0.5866f * s[1] + 0.1144f * s[2];
        float scale = (1.0f - expf(-Y)) / Y;
        if (Y == 0.0f) scale = 1.0f;
        uint8_t b[3];
        for (int c = 0; c < 3; ++c) {
            const float v = 255.0f * powf(scale * s[c], 1.0f / 2.2f);
            const float cl = fminf(fmaxf(v, 0.0f), 255.0f);
            b[c] = (uint8_t)(uint32_t)cl;
        }
        row[3 * x + 0] = b[2];
        row[3 * x + 1] = b[1];
        row[3 * x + 2] = b[0];
    }
    for (uint32_t k = 3 * W + x; k < rowBytes; k += blockDim.x * gridDim.x) row[k] = 0;
}

unsigned gridFor1d(uint32_t n) {
    return (unsigned)std::min<uint32_t>((n + 255) / 256, 65535u * 4u);
}

}  // namespace

int postLaunch(clr_ctx* ctx, int op, uint32_t width, uint32_t height, uint32_t spp, const clr_float3* src,
               void* dst, uint32_t rowBytes, hipStream_t stream) {
    const uint32_t n = width * height;
    const float inv = spp ? 1.0f / (float)spp : 0.0f;
    const float4* s4 = reinterpret_cast<const float4*>(src);
    float4* d4 = static_cast<float4*>(dst);
    const dim3 tiles((width + kPostTile - 1) / kPostTile, (height + kPostTile - 1) / kPostTile);
    const dim3 tb(kPostTile, kPostTile);
    switch (op) {
        case 0: k_post_clear<<<gridFor1d(n), 256, 0, stream>>>(n, d4); break;
        case 1: k_post_scaling<<<gridFor1d(n), 256, 0, stream>>>(n, inv, s4, d4); break;
        case 2: k_post_gaussian<<<tiles, tb, 0, stream>>>((int)width, (int)height, inv, s4, d4); break;
        case 3: k_post_bloom<<<tiles, tb, 0, stream>>>((int)width, (int)height, inv, s4, d4); break;
        case 4:
            k_post_tone<<<dim3((width + 255) / 256, height), 256, 0, stream>>>(width, height, rowBytes, s4,
                                                                               static_cast<uint8_t*>(dst));
            break;
        default: return fail(ctx, CLR_ERR_INVALID_ARG, "unknown post-processing operation");
    }
    return hipCheck(ctx, hipGetLastError(), "post-processing launch");
}

}  // namespace clr

namespace {
// Shared argument checks of the post-processing entry points. src and the
// output must not overlap (the reference reads and writes separate buffers).
int postEntry(clr_ctx* ctx, int op, uint32_t width, uint32_t height, uint32_t spp, const clr_float3* src, void* dst,
              uint32_t rowBytes, void* stream) {
    using namespace clr;
    if (!ctx) return CLR_ERR_INVALID_ARG;
    if (!dst || (op != 0 && !src)) return fail(ctx, CLR_ERR_INVALID_ARG, "null image buffer");
    if ((uint64_t)width * height >= (1ull << 31)) return fail(ctx, CLR_ERR_UNSUPPORTED, "image too large");
    if ((op == 1 || op == 2 || op == 3) && spp == 0) return fail(ctx, CLR_ERR_INVALID_ARG, "spp must be >= 1");
    if (op == 4 && rowBytes < 3u * width) return fail(ctx, CLR_ERR_INVALID_ARG, "rowBytes below 3 * width");
    if ((src && (reinterpret_cast<uintptr_t>(src) & 15u)) || (op != 4 && (reinterpret_cast<uintptr_t>(dst) & 15u)))
        return fail(ctx, CLR_ERR_INVALID_ARG, "float3 buffers must be 16-byte aligned");
    if (width == 0 || height == 0) return CLR_OK;
    DeviceScope ds(ctx->device);
    if (ds.err != hipSuccess) return hipCheck(ctx, ds.err, "set device");
    return postLaunch(ctx, op, width, height, spp, src, dst, rowBytes, static_cast<hipStream_t>(stream));
}
}  // namespace

extern "C" {

int clr_post_clear(clr_ctx* ctx, uint32_t width, uint32_t height, clr_float3* buffer, void* stream) {
    return postEntry(ctx, 0, width, height, 0, nullptr, buffer, 0, stream);
}

int clr_post_scaling(clr_ctx* ctx, uint32_t width, uint32_t height, uint32_t spp, const clr_float3* src,
                     clr_float3* output, void* stream) {
    return postEntry(ctx, 1, width, height, spp, src, output, 0, stream);
}

int clr_post_gaussian_scattering(clr_ctx* ctx, uint32_t width, uint32_t height, uint32_t spp,
                                 const clr_float3* src, clr_float3* output, void* stream) {
    return postEntry(ctx, 2, width, height, spp, src, output, 0, stream);
}

int clr_post_bloom(clr_ctx* ctx, uint32_t width, uint32_t height, uint32_t spp, const clr_float3* src,
                   clr_float3* output, void* stream) {
    return postEntry(ctx, 3, width, height, spp, src, output, 0, stream);
}

int clr_post_tone_mapping(clr_ctx* ctx, uint32_t width, uint32_t height, uint32_t rowBytes, const clr_float3* src,
                          uint8_t* dst, void* stream) {
    return postEntry(ctx, 4, width, height, 0, src, dst, rowBytes, stream);
}

uint32_t clr_bmp_row_bytes(uint32_t width) { return 3 * width + width % 4; }

int clr_save_bmp(const char* filename, const void* pixels, uint32_t width, uint32_t height) {
    if (!filename || !pixels) return CLR_ERR_INVALID_ARG;
    const uint32_t rowBytes = clr_bmp_row_bytes(width);
    const uint64_t dataSize64 = (uint64_t)rowBytes * height;
    if (dataSize64 + 54 > 0xFFFFFFFFull) return CLR_ERR_UNSUPPORTED;
    const uint32_t dataSize = (uint32_t)dataSize64, fileSize = dataSize + 54, dataOffset = 54, infoSize = 40;
    const uint16_t planes = 1, bits = 24;
    const uint32_t compress = 0, ppm = 1;
    uint8_t header[54] = {};
    header[0] = 'B';
    header[1] = 'M';
    std::memcpy(header + 2, &fileSize, 4);
    std::memcpy(header + 10, &dataOffset, 4);
    std::memcpy(header + 14, &infoSize, 4);
    std::memcpy(header + 18, &width, 4);
    std::memcpy(header + 22, &height, 4);
    std::memcpy(header + 26, &planes, 2);
    std::memcpy(header + 28, &bits, 2);
    std::memcpy(header + 30, &compress, 4);
    std::memcpy(header + 34, &dataSize, 4);
    std::memcpy(header + 38, &ppm, 4);
    std::memcpy(header + 42, &ppm, 4);
    FILE* fp = std::fopen(filename, "wb");
    if (!fp) return CLR_ERR_INVALID_ARG;
    const bool ok = std::fwrite(header, 1, 54, fp) == 54 && std::fwrite(pixels, 1, dataSize, fp) == dataSize;
    return (std::fclose(fp) == 0 && ok) ? CLR_OK : CLR_ERR_INVALID_ARG;
}

}  // extern "C"
