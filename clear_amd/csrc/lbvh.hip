// LBVH construction for gfx950 (wave64).
//
// Replaces LBVHBuilder::perform (OpenCL_PathTracer/LBVHBuilder.cpp:160-375)
// and the kernels of OpenCL/lbvh_construction.cl. Same outputs bit for bit
// (Morton quantisation, stable order with index tie-break, Karras topology and
// numbering, fmin/fmax boxes), different schedule:
//
//   init        clear histogram / lookback / counter scratch (fill kernel, no
//               hipMemset, so the sequence can be graph-captured)
//   bounds      per-triangle box, block min/max, one ordered-int atomic per
//               block and component: the scene box never leaves the device
//               (the reference reads it back to the host, LBVHBuilder.cpp:229)
//   morton      centroid quantisation (lbvh_construction.cl:169-181) into one
//               interleaved key (bit 3k+0 = x_k, 3k+1 = y_k, 3k+2 = z_k, the
//               order of the reference's digit x + 2y + 4z), plus the 8-bit
//               digit histograms of every radix pass
//   onesweep    one launch per 8-bit digit: stable LSD pass with wave64
//               ballot ranking and decoupled look-back (the reference spends
//               numBitsPerDim passes of 8 launches on 3-bit digits)
//   hierarchy   one bottom-up pass that derives the Karras 2012 binary radix
//               tree (node numbering of constructBinaryRadixTree,
//               lbvh_construction.cl:342-410) and the refit boxes
//               (calcNodeAABBs, :415-492) together: each leaf climbs, the
//               second arrival at a parent merges fmin/fmax(left, right) and
//               continues; release/acquire on the arrival word orders the box
//               writes (the reference's atomic_inc has no fence).
//
// A node covering sorted range [l, r] is the left child of its parent when the
// common prefix across its right boundary, delta(r, r+1), exceeds the one
// across its left boundary, delta(l-1, l); the parent splits at r then, else at
// l-1. Karras names a left child by its right end and a right child by its left
// end, which fixes every node's index on the way up.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "internal.h"

namespace clr {
namespace {

constexpr int kBlock = 256;
constexpr int kSortBlock = 512;
constexpr int kWaves = kSortBlock / 64;
static_assert(kSortBlock >= 256 && kSortBlock % 64 == 0, "a sort block owns all 256 digits");
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagInc = 2u << 30;
constexpr uint32_t kFlagMask = 3u << 30;
constexpr uint32_t kCountMask = (1u << 30) - 1;
constexpr int kMaxPasses = 8;

template <typename KeyT>
struct SortCfg;
template <>
struct SortCfg<uint32_t> {
    static constexpr int kItems = 16;
};
template <>
struct SortCfg<uint64_t> {
    static constexpr int kItems = 6;
};

// OpenCL fmin/fmax with a fixed answer for +-0 ties (left operand) and NaN
// (the other operand); the oracle uses the same rule.
__device__ __forceinline__ float omin(float a, float b) { return (b < a || a != a) ? b : a; }
__device__ __forceinline__ float omax(float a, float b) { return (b > a || a != a) ? b : a; }

// Total order on floats as unsigned ints (for atomicMin/Max).
__device__ __forceinline__ uint32_t ordOf(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float floatOf(uint32_t o) {
    uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    return __uint_as_float(u);
}

struct Box {
    float lo[3], hi[3];
};

__device__ __forceinline__ Box boxOf3(const float* p0, const float* p1, const float* p2) {
    Box b;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = omin(p0[k], omin(p1[k], p2[k]));
        b.hi[k] = omax(p0[k], omax(p1[k], p2[k]));
    }
    return b;
}

// Primitive source: a 36-B triangle soup, or (the drop-in path) a box array
// of 32-B records (lo.xyz, hi.x | hi.yz, triangle index, 0) in input order,
// written by the bounds pass. The hierarchy pass gathers one triangle (soup)
// or record per leaf by the sorted index.
struct PrimSrc {
    const float* tris;  // soup (null on the drop-in path)
    float4* boxes;      // box array
};

__device__ __forceinline__ Box soupBox(const float* tris, uint32_t i) {
    const float* t = tris + 9 * (size_t)i;
    float p0[3] = {t[0], t[1], t[2]};
    float p1[3] = {t[3], t[4], t[5]};
    float p2[3] = {t[6], t[7], t[8]};
    return boxOf3(p0, p1, p2);
}

// Streaming stores for records the next kernel does not read (box records,
// gathered only by the hierarchy pass; leaf and node records, the build's
// output): nontemporal, so they do not push the next kernel's inputs out of
// L2 and the Infinity Cache (the Morton pass's centres, the sort's keys).
typedef float StreamF4 __attribute__((ext_vector_type(4)));
typedef unsigned int StreamU4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void streamStore(float4* p, float4 v) {
    const StreamF4 x = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(x, reinterpret_cast<StreamF4*>(p));
}
__device__ __forceinline__ void streamStore(uint4* p, uint4 v) {
    const StreamU4 x = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(x, reinterpret_cast<StreamU4*>(p));
}

__device__ __forceinline__ void storeBoxRec(float4* boxes, uint32_t pos, const Box& b, uint32_t obj) {
    streamStore(boxes + 2 * (size_t)pos, make_float4(b.lo[0], b.lo[1], b.lo[2], b.hi[0]));
    streamStore(boxes + 2 * (size_t)pos + 1, make_float4(b.hi[1], b.hi[2], __uint_as_float(obj), 0.0f));
}

__device__ __forceinline__ float waveMin(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = omin(v, __shfl_xor(v, off));
    return v;
}
__device__ __forceinline__ float waveMax(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = omax(v, __shfl_xor(v, off));
    return v;
}

// ---------------------------------------------------------------- init ----
// Clears `count` words (histograms, look-back words) and sets the scene box
// to empty. With `epochWord`: when the build epoch is about to wrap to 0
// (every 2^32 - 1 builds) it also clears the hierarchy's exchange words, so
// that no word left by an earlier build carries the next epoch. No kernel
// writes the epoch word while this one runs, so every block reads one value.
__global__ __launch_bounds__(kBlock) void k_init(uint32_t* words, size_t count, uint32_t* sceneOrd,
                                                 const uint32_t* epochWord = nullptr,
                                                 uint32_t* xchgWords = nullptr, size_t xchgCount = 0) {
    size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t j = i; j < count; j += stride) words[j] = 0u;
    if (epochWord && *epochWord == 0xFFFFFFFFu)
        for (size_t j = i; j < xchgCount; j += stride) xchgWords[j] = 0u;
    if (sceneOrd && blockIdx.x == 0 && threadIdx.x < 6)
        sceneOrd[threadIdx.x] = threadIdx.x < 3 ? ordOf(INFINITY) : ordOf(-INFINITY);
}

// The build epoch tags the hierarchy's box hand-off words. It lives in device
// memory and advances once per build on the build's own stream (in the
// kernel after k_init), never 0, so a build replayed from a captured graph
// still gets a fresh epoch.
__device__ __forceinline__ void advanceEpoch(uint32_t* epochWord) {
    const uint32_t e = *epochWord;
    *epochWord = e == 0xFFFFFFFFu ? 1u : e + 1u;
}

// -------------------------------------------------------------- bounds ----
// One pass over the triangles (soup, or faces gathering vertices): per
// triangle its 32-B box record and its box centre (SoA, three floats) for the
// Morton pass, and the block's min / max towards the scene box. The centre is
// (lo + hi) * 0.5 per axis, the value calcMortonCodes quantises
// (lbvh_construction.cl:169-181), so the Morton pass reads 12 B per triangle
// instead of the triangle again.
template <bool SOUP>
__global__ __launch_bounds__(kBlock) void k_bounds(const float* __restrict__ tris,
                                                   const clr_float3* __restrict__ verts,
                                                   const clr_face* __restrict__ faces, uint32_t n,
                                                   float4* __restrict__ boxesOut,
                                                   float* __restrict__ cx, float* __restrict__ cy,
                                                   float* __restrict__ cz,
                                                   uint32_t* __restrict__ sceneOrd,
                                                   uint32_t* __restrict__ epochWord) {
    if (blockIdx.x == 0 && threadIdx.x == 0) advanceEpoch(epochWord);
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        Box b;
        if constexpr (SOUP) {
            b = soupBox(tris, i);
        } else {
            const uint32_t* f = reinterpret_cast<const uint32_t*>(faces + i);
            float4 v0 = *reinterpret_cast<const float4*>(verts + f[0]);
            float4 v1 = *reinterpret_cast<const float4*>(verts + f[1]);
            float4 v2 = *reinterpret_cast<const float4*>(verts + f[2]);
            float p0[3] = {v0.x, v0.y, v0.z}, p1[3] = {v1.x, v1.y, v1.z},
                  p2[3] = {v2.x, v2.y, v2.z};
            b = boxOf3(p0, p1, p2);
        }
        if constexpr (!SOUP) storeBoxRec(boxesOut, i, b, i);  // a soup's hierarchy pass gathers the triangle
        cx[i] = (b.lo[0] + b.hi[0]) * 0.5f;
        cy[i] = (b.lo[1] + b.hi[1]) * 0.5f;
        cz[i] = (b.lo[2] + b.hi[2]) * 0.5f;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = omin(lo[k], b.lo[k]);
            hi[k] = omax(hi[k], b.hi[k]);
        }
    }
    __shared__ float s_red[kBlock / 64][6];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = waveMin(lo[k]);
        hi[k] = waveMax(hi[k]);
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            s_red[wave][k] = lo[k];
            s_red[wave][3 + k] = hi[k];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        float v = s_red[0][k];
        for (int w = 1; w < kBlock / 64; ++w) v = k < 3 ? omin(v, s_red[w][k]) : omax(v, s_red[w][k]);
        if (k < 3)
            atomicMin(sceneOrd + k, ordOf(v));
        else
            atomicMax(sceneOrd + k, ordOf(v));
    }
}

// -------------------------------------------------------------- morton ----
template <typename KeyT>
__device__ __forceinline__ KeyT spread3(uint32_t v);
template <>
__device__ __forceinline__ uint32_t spread3<uint32_t>(uint32_t x) {
    x &= 0x3FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}
template <>
__device__ __forceinline__ uint64_t spread3<uint64_t>(uint32_t v) {
    uint64_t x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// min(convert_uint_rtz(v), numDivs - 1), NaN and non-positive values to 0.
__device__ __forceinline__ uint32_t quantize(float v, uint32_t numDivs) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return numDivs - 1u;
    uint32_t q = (uint32_t)v;
    return q < numDivs - 1u ? q : numDivs - 1u;
}

// Exclusive scan of one value per digit over threads 0..255 (digit t in
// thread t); every thread of the block takes part in the barriers, and the
// result of threads from 256 on is meaningless.
__device__ __forceinline__ uint32_t digitExclusiveScan(uint32_t v, uint32_t* s_tmp) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63 && wave < 4) s_tmp[wave] = x;
    __syncthreads();
    uint32_t pre = 0;
    for (int w = 0; w < wave && w < 4; ++w) pre += s_tmp[w];
    __syncthreads();
    return pre + x - v;
}

// Keys from the box centres and the digit histograms of every radix pass, in
// input order.
template <typename KeyT>
__global__ __launch_bounds__(kBlock) void k_morton(const float* __restrict__ cx, const float* __restrict__ cy,
                                                   const float* __restrict__ cz, uint32_t n, uint32_t bits,
                                                   const uint32_t* __restrict__ sceneOrd,
                                                   KeyT* __restrict__ keys,
                                                   uint32_t* __restrict__ hist, int numPasses) {
    __shared__ uint32_t s_hist[kMaxPasses * 256];
    for (int j = threadIdx.x; j < numPasses * 256; j += kBlock) s_hist[j] = 0u;
    __syncthreads();

    float smin[3], ssize[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        smin[k] = floatOf(sceneOrd[k]);
        ssize[k] = floatOf(sceneOrd[3 + k]) - smin[k];
    }
    const uint32_t numDivs = 1u << bits;
    const float fDivs = (float)numDivs;
    auto emit = [&](uint32_t i, const float* c) {
        uint32_t q[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float np = (c[k] - smin[k]) / ssize[k];
            q[k] = quantize(np * fDivs, numDivs);
        }
        KeyT key = spread3<KeyT>(q[0]) | (spread3<KeyT>(q[1]) << 1) | (spread3<KeyT>(q[2]) << 2);
        keys[i] = key;
        for (int p = 0; p < numPasses; ++p)
            atomicAdd(&s_hist[p * 256 + (uint32_t)((key >> (8 * p)) & 0xFFu)], 1u);
    };
    // Four triangles per thread per round, their centres loaded up front so
    // twelve loads are in flight before the first key is formed.
    const uint32_t stride = gridDim.x * kBlock;
    uint32_t i0 = blockIdx.x * kBlock + threadIdx.x;
    for (; i0 + 3 * stride < n; i0 += 4 * stride) {
        float cc[4][3];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            cc[u][0] = cx[i0 + u * stride];
            cc[u][1] = cy[i0 + u * stride];
            cc[u][2] = cz[i0 + u * stride];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) emit(i0 + u * stride, cc[u]);
    }
    for (uint32_t i = i0; i < n; i += stride) {
        const float c[3] = {cx[i], cy[i], cz[i]};
        emit(i, c);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < numPasses * 256; j += kBlock)
        if (s_hist[j]) atomicAdd(hist + j, s_hist[j]);
}

// ------------------------------------------------------------ onesweep ----
__device__ __forceinline__ void writeLeaf(clr_lbvh_leaf* leaves, uint32_t i, const Box& b,
                                          uint32_t obj) {
    uint4* o = reinterpret_cast<uint4*>(leaves + i);
    o[0] = make_uint4(__float_as_uint(b.lo[0]), __float_as_uint(b.lo[1]),
                      __float_as_uint(b.lo[2]), 0u);
    o[1] = make_uint4(__float_as_uint(b.hi[0]), __float_as_uint(b.hi[1]),
                      __float_as_uint(b.hi[2]), 0u);
    o[2] = make_uint4(obj, 0u, 0u, 0u);
}

// One stable LSD pass on the 8-bit digit at `shift`. FIRST generates the
// payload (original index) instead of reading it.
template <typename KeyT, bool FIRST>
__global__ __launch_bounds__(kSortBlock) void k_onesweep(const KeyT* __restrict__ keysIn,
                                                         const uint32_t* __restrict__ valsIn,
                                                         KeyT* __restrict__ keysOut,
                                                         uint32_t* __restrict__ valsOut,
                                                         uint32_t n, int shift,
                                                         const uint32_t* __restrict__ passHist,
                                                         uint32_t* __restrict__ status,
                                                         uint32_t* __restrict__ partCounter,
                                                         uint32_t* __restrict__ devStatus) {
    constexpr int kItems = SortCfg<KeyT>::kItems;
    constexpr int kTile = kSortBlock * kItems;
    __shared__ uint32_t s_part;
    __shared__ uint32_t s_waveHist[kWaves][256];
    __shared__ uint32_t s_localStart[256];
    __shared__ uint32_t s_digitBase[256];
    __shared__ uint32_t s_tmp[kWaves];
    __shared__ KeyT s_keys[kTile];
    __shared__ uint32_t s_vals[kTile];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int j = tid; j < kWaves * 256; j += kSortBlock) (&s_waveHist[0][0])[j] = 0u;
    if (tid == 0) s_part = atomicAdd(partCounter, 1u);
    __syncthreads();
    const uint32_t part = s_part;
    const uint32_t tileBase = part * (uint32_t)kTile;
    const uint32_t waveBase = tileBase + (uint32_t)wave * (64u * kItems);

    KeyT keys[kItems];
    uint32_t vals[kItems];
    uint32_t rank[kItems];
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        uint32_t idx = waveBase + (uint32_t)i * 64u + (uint32_t)lane;
        if (idx < n) {
            keys[i] = keysIn[idx];
            vals[i] = FIRST ? idx : valsIn[idx];
        } else {
            keys[i] = ~KeyT(0);  // digit 0xFF at every shift, ranked last
            vals[i] = 0u;
        }
    }

    // Rank each item among equal digits of its wave (wave-level multisplit):
    // 8 ballots give the set of lanes holding the same digit; the lowest such
    // lane owns the wave's running counter for that digit.
    const uint64_t ltMask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const uint32_t d = (uint32_t)(keys[i] >> shift) & 0xFFu;
        uint64_t peers = ~0ull;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const int leader = __ffsll((unsigned long long)peers) - 1;
        const uint32_t below = (uint32_t)__popcll(peers & ltMask);
        uint32_t base = 0;
        if (lane == leader) {
            base = s_waveHist[wave][d];
            s_waveHist[wave][d] = base + (uint32_t)__popcll(peers);
        }
        base = __shfl(base, leader);
        rank[i] = base + below;
    }
    __syncthreads();

    // Per digit: wave offsets, tile count, look-back for the tile's global
    // offset. Thread t < 256 owns digit t.
    const bool owner = tid < 256;
    uint32_t tileCount = 0;
    if (owner) {
        uint32_t sum = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            uint32_t c = s_waveHist[w][tid];
            s_waveHist[w][tid] = sum;
            sum += c;
        }
        tileCount = sum;
    }
    // Decoupled look-back, kWin predecessors per round trip. The first
    // window is requested before the tile-local scatter so that its latency
    // overlaps that work. Earlier partitions took their ids before this one,
    // so they are resident and will publish; the bound only guards against a
    // bug.
    constexpr int kWin = 4;
    uint32_t exclusive = 0;
    uint32_t win[kWin];
    int64_t look = (int64_t)part - 1;
    auto loadWindow = [&]() {
#pragma unroll
        for (int k = 0; k < kWin; ++k) {
            const int64_t q = look - k;
            const uint32_t v = __hip_atomic_load(status + (size_t)(q < 0 ? 0 : q) * 256 + tid, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            win[k] = q < 0 ? kFlagInc : v;  // never reached: partition 0 is inclusive
        }
    };
    if (owner) {
        if (part == 0) {
            __hip_atomic_store(status + tid, kFlagInc | tileCount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(status + (size_t)part * 256 + tid, kFlagAgg | tileCount, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            loadWindow();
        }
    }
    const uint32_t localStart = digitExclusiveScan(tileCount, s_tmp);
    const uint32_t globalStart = digitExclusiveScan(owner ? passHist[tid] : 0u, s_tmp);
    if (owner) s_localStart[tid] = localStart;
    __syncthreads();

    // Tile-local order in LDS (does not depend on the global offsets).
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const uint32_t d = (uint32_t)(keys[i] >> shift) & 0xFFu;
        const uint32_t pos = s_localStart[d] + s_waveHist[wave][d] + rank[i];
        s_keys[pos] = keys[i];
        s_vals[pos] = vals[i];
    }

    if (owner && part > 0) {
        uint32_t spins = 0;
        while (true) {
            int stop = kWin;  // first entry not yet published
            bool inc = false;
            uint32_t add = 0;
#pragma unroll
            for (int k = 0; k < kWin; ++k) {
                if (stop == kWin && !inc) {
                    const uint32_t v = win[k], flag = v & kFlagMask;
                    if (flag == 0u) {
                        stop = k;
                    } else {
                        add += v & kCountMask;
                        inc = flag == kFlagInc;
                    }
                }
            }
            exclusive += add;
            if (inc) break;
            look -= stop;
            if (stop < kWin && ++spins >= kSpinLimit) {
                atomicOr(devStatus, kStatusLookbackSpin);
                break;
            }
            loadWindow();
        }
        __hip_atomic_store(status + (size_t)part * 256 + tid, kFlagInc | (exclusive + tileCount),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (owner) s_digitBase[tid] = globalStart + exclusive;
    __syncthreads();
    const uint32_t valid = min((uint32_t)kTile, n - tileBase);
    for (uint32_t j = tid; j < valid; j += kSortBlock) {
        const KeyT k = s_keys[j];
        const uint32_t d = (uint32_t)(k >> shift) & 0xFFu;
        const uint32_t g = s_digitBase[d] + (j - s_localStart[d]);
        keysOut[g] = k;
        valsOut[g] = s_vals[j];
    }
}

// ----------------------------------------------------------- hierarchy ----
template <typename KeyT>
__device__ __forceinline__ int clzKey(KeyT v);
template <>
__device__ __forceinline__ int clzKey<uint32_t>(uint32_t v) {
    return __clz((int)v);
}
template <>
__device__ __forceinline__ int clzKey<uint64_t>(uint64_t v) {
    return __clzll((long long)v);
}

// delta(x, x+1) over extended keys (key, sorted position); the reference's
// _numCommonBits (lbvh_construction.cl:322-332) for adjacent positions.
// `base` is the first sorted position of x's scene (0 for a single build):
// equal keys fall back to the scene-local positions, so a scene built inside a
// segmented build gets the tree it gets alone.
template <typename KeyT>
__device__ __forceinline__ int deltaAdj(const KeyT* __restrict__ k, uint32_t x, int keyBits, uint32_t base = 0) {
    const KeyT a = k[x], b = k[x + 1];
    if (a != b) return clzKey<KeyT>(a ^ b) - (int)(8 * sizeof(KeyT)) + keyBits;
    return keyBits + __clz((int)((x - base) ^ (x + 1u - base)));
}

// Segment view of a segmented build: per sorted position its scene, per scene
// its first position (segOff[s]) and end (segOff[s + 1]). Scenes keep their
// input ranges through the sort (scene id is the key's top digit), so a
// position's scene is that of the input triangle at the same index.
struct SegView {
    const uint32_t* segOf;
    const uint32_t* segOff;
    const uint32_t* valid;  // 0: the scene offsets were rejected, nothing is built
};

// delta(x, x+1), -1 outside [0, n) and, in a segmented build, across scenes.
template <typename KeyT, bool SEG>
__device__ __forceinline__ int deltaAt(const KeyT* __restrict__ k, int64_t x, uint32_t n, int keyBits,
                                       const SegView& sv) {
    if (x < 0 || x + 1 >= (int64_t)n) return -1;
    if constexpr (SEG) {
        const uint32_t s = sv.segOf[x];
        if (sv.segOf[x + 1] != s) return -1;
        return deltaAdj<KeyT>(k, (uint32_t)x, keyBits, sv.segOff[s]);
    } else {
        return deltaAdj<KeyT>(k, (uint32_t)x, keyBits);
    }
}

// deltaAt with unconditional loads (an index outside [0, n - 1) reads
// position 0 and the result is -1), so that a thread's loads for several x
// can be in flight together.
template <typename KeyT, bool SEG>
__device__ __forceinline__ int deltaAtLoaded(const KeyT* __restrict__ k, int64_t x, uint32_t n, int keyBits,
                                             const SegView& sv) {
    const bool valid = x >= 0 && x + 1 < (int64_t)n;
    const uint32_t x0 = valid ? (uint32_t)x : 0u;
    const uint32_t x1 = valid ? x0 + 1u : 0u;
    const KeyT a = k[x0], b = k[x1];
    uint32_t base = 0;
    bool cross = false;
    if constexpr (SEG) {
        const uint32_t s0 = sv.segOf[x0], s1 = sv.segOf[x1];
        cross = s0 != s1;
        base = sv.segOff[s0];
    }
    const int dk = clzKey<KeyT>(a ^ b) - (int)(8 * sizeof(KeyT)) + keyBits;
    const int di = keyBits + __clz((int)((x0 - base) ^ (x1 - base)));
    return (!valid || cross) ? -1 : (a != b ? dk : di);
}

// Node record; children are split - base and split + 1 - base (scene-local in
// a segmented build).
__device__ __forceinline__ void writeNode(clr_lbvh_inode* nodes, uint32_t idx, const Box& b,
                                          uint32_t split, uint32_t l, uint32_t r, uint32_t base = 0) {
    uint4* o = reinterpret_cast<uint4*>(nodes + idx);
    o[0] = make_uint4(__float_as_uint(b.lo[0]), __float_as_uint(b.lo[1]),
                      __float_as_uint(b.lo[2]), 0u);
    o[1] = make_uint4(__float_as_uint(b.hi[0]), __float_as_uint(b.hi[1]),
                      __float_as_uint(b.hi[2]), 0u);
    const uint32_t flags = (split == l ? 1u : 0u) | (split + 1u == r ? 0x100u : 0u);
    o[2] = make_uint4(flags, split - base, split + 1u - base, 0u);
}

__device__ __forceinline__ Box leafBox(const clr_lbvh_leaf* leaves, uint32_t i) {
    const float4* p = reinterpret_cast<const float4*>(leaves + i);
    const float4 a = p[0], b = p[1];
    Box r;
    r.lo[0] = a.x; r.lo[1] = a.y; r.lo[2] = a.z;
    r.hi[0] = b.x; r.hi[1] = b.y; r.hi[2] = b.z;
    return r;
}

// Box hand-off between the two children of a node. Every internal node
// publishes its box before it arrives at its parent, as six 64-bit words
// (float bits | epoch) written with relaxed agent-scope atomics (plain sc1
// stores on gfx950; an acq_rel exchange would instead write back and
// invalidate the whole per-XCD L2 at every step). The second arrival polls the
// sibling's words until they carry this build's epoch: the sibling issued its
// stores before its own arrival, which this thread has already observed, so
// the wait ends; kSpinLimit only guards against a bug.
__device__ __forceinline__ void publishBox(uint64_t* slot, const Box& b, uint32_t epoch) {
    const uint64_t tag = (uint64_t)epoch << 32;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        __hip_atomic_store(slot + k, tag | __float_as_uint(b.lo[k]), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(slot + 3 + k, tag | __float_as_uint(b.hi[k]), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
}

__device__ __forceinline__ Box awaitBox(const uint64_t* slot, uint32_t epoch,
                                        uint32_t* devStatus) {
    uint64_t w[6];
    // All six words in flight at once; re-poll only while some are stale.
#pragma unroll
    for (int k = 0; k < 6; ++k)
        w[k] = __hip_atomic_load(slot + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t spins = 0;
    while (true) {
        bool fresh = true;
#pragma unroll
        for (int k = 0; k < 6; ++k) fresh &= (uint32_t)(w[k] >> 32) == epoch;
        if (fresh) break;
        if (++spins >= kSpinLimit) {
            atomicOr(devStatus, kStatusExchangeSpin);
            break;
        }
#pragma unroll
        for (int k = 0; k < 6; ++k)
            if ((uint32_t)(w[k] >> 32) != epoch)
                w[k] = __hip_atomic_load(slot + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    Box b;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = __uint_as_float((uint32_t)w[k]);
        b.hi[k] = __uint_as_float((uint32_t)w[3 + k]);
    }
    return b;
}

// Bottom-up topology + refit in two phases.
//
// Phase 1 (k_hier_local): block b owns leaves [bs, be) (kLocalLeaves per
// block, kHierThreads threads) and completes every node whose range lies
// inside them. Arrival slots for splits bs..be-2, boxes and boundary deltas
// live in LDS. The climb runs in synchronous steps over a queue of climbers
// (a climber holds a completed node [l, r]; at first every leaf): in a step
// each climber finds its parent (split r if delta(r) > delta(l-1), else l-1),
// exchanges its far end into the parent's slot, and the second arrival
// completes the parent (box = left child's box joined with the right's),
// writes its record at its index (r if the parent is itself a left child,
// else l; the root of the range at 0) and enqueues it. The queue is compacted
// every step, so lanes are spent on live climbs only (a lane per leaf would
// keep each wave stepping as long as its longest chain). Boxes: slot k holds
// leaf k's box until internal node k is completed (leaf k lies inside node k
// and was joined into its parent before). A climb whose parent splits at a
// block boundary (bs-1 or be-1) stops and leaves a pending record (at most two
// per block); after the last step the block writes the global arrival words
// of every slot it did not complete: the far end of the one child that
// arrived (publishing that child's box) or empty, and empty for the split
// between it and the next block. Those are all the words phase 2 reads, so no
// earlier pass clears them.
// Phase 2 (k_hier_global): one thread per pending record continues with the
// global epoch-tagged protocol. Only nodes whose range crosses a block
// boundary (a small fraction) are finished there.
constexpr int kLocalLeaves = 1024;
constexpr int kHierThreads = 256;
constexpr int kHierPer = kLocalLeaves / kHierThreads;
static_assert(kLocalLeaves % kHierThreads == 0 && kLocalLeaves <= 65536, "hierarchy block shape");
constexpr uint32_t kDone = 0xFFFFFFFEu;
constexpr uint32_t kAuxFlagShift = 10;  // log2(kLocalLeaves)
constexpr uint32_t kAuxEmpty = 0xFFFFu;
static_assert((1u << kAuxFlagShift) == (uint32_t)kLocalLeaves, "completed-node record layout");

struct Pending {
    uint32_t valid, l, r, split, internal;  // split: unused (a pending node is never a root)
    int dl, dr;
    float box[6];
    uint32_t pad[3];
};
static_assert(sizeof(Pending) == 64, "pending record layout");

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup g runs on
// XCD g % 8). This numbering gives each XCD one contiguous range of blocks,
// so the records and nodes of neighbouring leaves meet in one L2.
__device__ __forceinline__ uint32_t xcdContiguous(uint32_t g, uint32_t G) {
    constexpr uint32_t X = 8;
    const uint32_t x = g % X, k = g / X, q = G / X, r = G % X;
    return x * q + min(x, r) + k;
}

// SEG: segmented build (sv describes the scenes): deltas are -1 across
// scenes, a scene's root is the node covering its whole range (stored at its
// first inode slot), inodes of scene s are numbered locally and stored at
// segOff[s] - s + local, children and leaf object indices are scene-local.
// A range [l, r] is a scene's root exactly when delta(l-1) and delta(r) are
// both -1.
template <bool SOUP, typename KeyT, bool SEG = false>
__global__ __launch_bounds__(kHierThreads) void k_hier_local(
    PrimSrc src, const KeyT* __restrict__ skeys, const uint32_t* __restrict__ sidx, uint32_t n,
    int keyBits, uint32_t* __restrict__ otherEnd, uint64_t* __restrict__ xchg, const uint32_t* __restrict__ epochWord,
    clr_lbvh_inode* __restrict__ inodes, clr_lbvh_leaf* __restrict__ leaves,
    Pending* __restrict__ pend, uint32_t* __restrict__ devStatus, SegView sv = SegView{nullptr, nullptr, nullptr}) {
    if constexpr (SEG)
        if (*sv.valid == 0u) return;
    const uint32_t epoch = *epochWord;
    constexpr int B = kLocalLeaves;
    __shared__ uint32_t s_other[B];
    __shared__ float s_box[6][B];
    // Deltas fit in 8 bits (at most keyBits + 31 <= 95) and completed-node
    // records in 16 (split - bs and two flags): 39 KB, four blocks per CU.
    __shared__ int8_t s_delta[B + 1];  // delta(x) for x = bs-1 .. be-1 at [x - bs + 1]
    __shared__ uint32_t s_q[2][B];     // climbers: l | r << 16 (block-local)
    __shared__ uint16_t s_aux[B];      // node completed here at slot j: split - bs | flags << kAuxFlagShift
    __shared__ uint32_t s_qc[3];
    __shared__ uint32_t s_seg[SEG ? B : 1];

    const uint32_t t = threadIdx.x;
    const uint32_t blk = xcdContiguous(blockIdx.x, gridDim.x);
    const uint32_t bs = blk * (uint32_t)B;
    const uint32_t be = min(bs + (uint32_t)B, n);

    // Leaves: gather each sorted primitive's record, write its LeafNode. The
    // loads are unconditional (indices clamped into the block) so that every
    // gather of a thread is in flight at once.
    uint32_t rp[kHierPer];
#pragma unroll
    for (int k = 0; k < kHierPer; ++k) {
        const uint32_t j = (uint32_t)k * kHierThreads + t;
        s_other[j] = kEmpty;
        s_aux[j] = kAuxEmpty;
        rp[k] = sidx[min(bs + j, be - 1u)];
    }
    float4 ra[kHierPer], rb[kHierPer];
#pragma unroll
    for (int k = 0; k < kHierPer; ++k) {
        if constexpr (SOUP) {
            // The triangle itself (36 B: one 128-B line, two for the 27 % of
            // records that straddle one), its box formed as the bounds pass
            // forms it: the soup path writes no box records.
            const Box b = soupBox(src.tris, rp[k]);
            ra[k] = make_float4(b.lo[0], b.lo[1], b.lo[2], b.hi[0]);
            rb[k] = make_float4(b.hi[1], b.hi[2], __uint_as_float(rp[k]), 0.0f);
        } else {
            ra[k] = src.boxes[2 * (size_t)rp[k]];
            rb[k] = src.boxes[2 * (size_t)rp[k] + 1];
        }
    }
    uint32_t segS[kHierPer];
#pragma unroll
    for (int k = 0; k < kHierPer; ++k) {
        segS[k] = 0;
        if constexpr (SEG) {
            const uint32_t j = (uint32_t)k * kHierThreads + t;
            const uint32_t seg = sv.segOf[min(bs + j, be - 1u)];
            s_seg[j] = seg;
            segS[k] = sv.segOff[seg];
        }
    }
#pragma unroll
    for (int k = 0; k < kHierPer; ++k) {
        const uint32_t j = (uint32_t)k * kHierThreads + t, i = bs + j;
        if (i < be) {
            s_box[0][j] = ra[k].x; s_box[1][j] = ra[k].y; s_box[2][j] = ra[k].z;
            s_box[3][j] = ra[k].w; s_box[4][j] = rb[k].x; s_box[5][j] = rb[k].y;
            s_q[0][j] = __float_as_uint(rb[k].z) - segS[k];  // the leaf's object index
        }
    }
    // delta(x) for x = bs-1 .. be-1 (B + 1 values; thread 0 takes the last).
    {
        int dv[kHierPer + 1];
#pragma unroll
        for (int k = 0; k <= kHierPer; ++k) {
            const uint32_t j = k < kHierPer ? (uint32_t)k * kHierThreads + t : (uint32_t)B;
            const int64_t x = (int64_t)bs - 1 + j;
            dv[k] = deltaAtLoaded<KeyT, SEG>(skeys, x, n, keyBits, sv);
        }
#pragma unroll
        for (int k = 0; k < kHierPer; ++k) s_delta[(uint32_t)k * kHierThreads + t] = dv[k];
        if (t == 0) s_delta[B] = dv[kHierPer];
    }
    if (t < 2) pend[2 * blk + t].valid = 0u;
    if (t < 3) s_qc[t] = 0u;
    __syncthreads();
    // Leaf records [bs, be) as consecutive 16-B parts (whole lines per store).
    const uint32_t nl = be - bs;
    for (uint32_t q = t; q < 3u * nl; q += kHierThreads) {
        const uint32_t j = q / 3u, part = q - 3u * j;
        uint4 v;
        if (part == 2u)
            v = make_uint4(s_q[0][j], 0u, 0u, 0u);
        else
            v = make_uint4(__float_as_uint(s_box[3 * part][j]), __float_as_uint(s_box[3 * part + 1][j]),
                           __float_as_uint(s_box[3 * part + 2][j]), 0u);
        streamStore(reinterpret_cast<uint4*>(leaves + bs) + q, v);
    }
    __syncthreads();

    const int lane = (int)(t & 63u);
    const uint64_t ltMask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t cnt = be > bs ? be - bs : 0u;  // step 0: climber c is leaf c
    for (uint32_t it = 0; cnt > 0; ++it) {
        if (it >= (uint32_t)B) {  // a valid climb completes at least one node per step
            if (t == 0) atomicOr(devStatus, kStatusHierarchy);
            break;
        }
        if (t == 0) s_qc[(it + 2) % 3] = 0u;  // the counter of step it + 1's pushes
        const uint32_t* qin = s_q[it & 1u];
        uint32_t* qout = s_q[(it + 1u) & 1u];
        uint32_t* qcount = &s_qc[(it + 1) % 3];
        for (uint32_t c = t; c < cnt; c += kHierThreads) {
            uint32_t l = c, r = c;
            if (it > 0) {
                const uint32_t q = qin[c];
                l = q & 0xFFFFu;
                r = q >> 16;
            }
            const uint32_t L = bs + l, R = bs + r;
            const int dl = s_delta[l], dr = s_delta[r + 1];  // delta(L-1), delta(R)
            bool push = false;
            uint32_t packed = 0;
            if (dl != -1 || dr != -1) {  // else a one-leaf scene: nothing above it
                const bool isLeft = dr > dl;
                const uint32_t ps = isLeft ? R : L - 1u;  // the parent's split
                const uint32_t self = isLeft ? r : l;     // this node's box slot
                if (ps < bs || ps + 1u >= be) {
                    // Split on a block boundary: hand the climb to phase 2.
                    Pending& P = pend[2 * blk + (isLeft ? 1 : 0)];
                    P.l = L; P.r = R; P.split = 0u; P.internal = l != r ? 1u : 0u;
                    P.dl = dl; P.dr = dr;
#pragma unroll
                    for (int k = 0; k < 6; ++k) P.box[k] = s_box[k][self];
                    P.valid = 1u;
                } else {
                    const uint32_t other =
                        __hip_atomic_exchange(&s_other[ps - bs], isLeft ? L : R, __ATOMIC_ACQ_REL,
                                              __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (other != kEmpty && (other < bs || other >= be)) {
                        // A third arrival: the keys were not sorted. Stop this climb.
                        atomicOr(devStatus, kStatusHierarchy);
                    } else if (other != kEmpty) {
                        s_other[ps - bs] = kDone;  // both children arrived
                        const uint32_t sib = (isLeft ? ps + 1u : ps) - bs;
                        const uint32_t lo = isLeft ? self : sib, hi = isLeft ? sib : self;
                        Box pb;
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            pb.lo[k] = omin(s_box[k][lo], s_box[k][hi]);
                            pb.hi[k] = omax(s_box[3 + k][lo], s_box[3 + k][hi]);
                        }
                        const uint32_t pL = isLeft ? L : other, pR = isLeft ? other : R;
                        const int pdl = isLeft ? dl : s_delta[other - bs];
                        const int pdr = isLeft ? s_delta[other - bs + 1] : dr;
                        // The parent's slot: its index r or l, or for a scene's
                        // root the scene's first position (stored at the scene's
                        // first inode). Its record is written after the climb.
                        const bool root = pdl == -1 && pdr == -1;
                        const uint32_t pSelf = root ? pL : (pdr > pdl ? pR : pL);
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            s_box[k][pSelf - bs] = pb.lo[k];
                            s_box[3 + k][pSelf - bs] = pb.hi[k];
                        }
                        s_aux[pSelf - bs] = (ps - bs) | (ps == pL ? 1u << kAuxFlagShift : 0u) |
                                          (ps + 1u == pR ? 1u << (kAuxFlagShift + 1) : 0u);
                        if (!root) {
                            push = true;
                            packed = (pL - bs) | ((pR - bs) << 16);
                        }
                    }
                }
            }
            // Wave-aggregated enqueue.
            const uint64_t pm = __ballot(push);
            if (pm) {
                const int leader = __ffsll((unsigned long long)pm) - 1;
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(qcount, (uint32_t)__popcll(pm));
                base = __shfl(base, leader);
                if (push) qout[base + (uint32_t)__popcll(pm & ltMask)] = packed;
            }
        }
        __syncthreads();
        cnt = *qcount;
    }
    // The loop ends with a barrier. Export the arrival words of phase 2: the
    // slots where exactly one child arrived (phase 2 brings the other) hold
    // that child's far end, the slots where neither did are empty, and so is
    // the split between this block and the next. Completed slots are never
    // reached again.
    if (t == 0 && be < n) otherEnd[be - 1] = kEmpty;
#pragma unroll
    for (int k = 0; k < kHierPer; ++k) {
        const uint32_t j = (uint32_t)k * kHierThreads + t;
        if (bs + j + 1 < be) {
            const uint32_t v = s_other[j];
            if (v != kDone) otherEnd[bs + j] = v;
            if (v != kEmpty && v != kDone) {
                const uint32_t ps = bs + j;
                const bool leftArrived = v <= ps;
                const uint32_t idx = leftArrived ? ps : ps + 1u;
                const bool isLeaf = leftArrived ? (v == ps) : (v == ps + 1u);
                if (!isLeaf) {
                    const uint64_t tag = (uint64_t)epoch << 32;
                    uint64_t* slot = xchg + 6 * (size_t)idx;
#pragma unroll
                    for (int c = 0; c < 6; ++c) slot[c] = tag | __float_as_uint(s_box[c][idx - bs]);
                }
            }
        }
    }
    // Records of the nodes completed here, as consecutive 16-B parts. A node
    // at slot j is stored at j minus its scene index; children are split and
    // split + 1 relative to the scene's first position.
    for (uint32_t q = t; q < 3u * nl; q += kHierThreads) {
        const uint32_t j = q / 3u, part = q - 3u * j;
        const uint32_t aux = s_aux[j];
        if (aux == kAuxEmpty) continue;
        uint32_t seg = 0, segS = 0;
        if constexpr (SEG) {
            seg = s_seg[j];
            segS = sv.segOff[seg];
        }
        uint4 v;
        if (part == 2u) {
            const uint32_t split = bs + (aux & ((1u << kAuxFlagShift) - 1u));
            const uint32_t flags = ((aux >> kAuxFlagShift) & 1u) | (((aux >> (kAuxFlagShift + 1)) & 1u) << 8);
            v = make_uint4(flags, split - segS, split + 1u - segS, 0u);
        } else {
            v = make_uint4(__float_as_uint(s_box[3 * part][j]), __float_as_uint(s_box[3 * part + 1][j]),
                           __float_as_uint(s_box[3 * part + 2][j]), 0u);
        }
        streamStore(reinterpret_cast<uint4*>(inodes + (bs + j - seg)) + part, v);
    }
}

template <typename KeyT, bool SEG = false>
__global__ __launch_bounds__(kBlock) void k_hier_global(
    const KeyT* __restrict__ skeys, uint32_t n, int keyBits, uint32_t* __restrict__ otherEnd,
    uint64_t* __restrict__ xchg, const uint32_t* __restrict__ epochWord, uint32_t* __restrict__ devStatus,
    clr_lbvh_inode* __restrict__ inodes, const clr_lbvh_leaf* __restrict__ leaves,
    const Pending* __restrict__ pend, uint32_t numPending, SegView sv = SegView{nullptr, nullptr, nullptr}) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= numPending) return;
    if constexpr (SEG)
        if (*sv.valid == 0u) return;
    const uint32_t epoch = *epochWord;
    const Pending P = pend[p];
    if (!P.valid) return;
    uint32_t l = P.l, r = P.r, split = P.split;
    bool internal = P.internal != 0u;
    int dl = P.dl, dr = P.dr;
    Box box;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        box.lo[k] = P.box[k];
        box.hi[k] = P.box[3 + k];
    }
    uint32_t seg = 0, segS = 0, segE = n;
    if constexpr (SEG) {
        seg = sv.segOf[l];
        segS = sv.segOff[seg];
        segE = sv.segOff[seg + 1];
    }
    bool written = true;  // phase 1 already wrote this node's record
    while (true) {
        if (l == segS && r == segE - 1) {
            if (internal) writeNode(inodes, segS - seg, box, split, l, r, segS);
            return;
        }
        const bool isLeft = dr > dl;
        const uint32_t parentSplit = isLeft ? r : l - 1;
        if (internal) {
            const uint32_t self = isLeft ? r : l;
            if (!written) writeNode(inodes, self - seg, box, split, l, r, segS);
            publishBox(xchg + 6 * (size_t)self, box, epoch);
        }
        written = false;
        // Keep the publish ahead of the arrival in program order: a sibling in
        // the same wave may be the one that waits for it.
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        const uint32_t other = __hip_atomic_exchange(otherEnd + parentSplit, isLeft ? l : r,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (other == kEmpty) return;  // first arrival: the sibling finishes the parent
        const uint32_t s = isLeft ? parentSplit + 1u : parentSplit;
        int dNew;
        if (isLeft)
            dNew = deltaAt<KeyT, SEG>(skeys, (int64_t)other, n, keyBits, sv);
        else
            dNew = deltaAt<KeyT, SEG>(skeys, (int64_t)other - 1, n, keyBits, sv);
        const Box sb = (s == other) ? leafBox(leaves, s)
                                    : awaitBox(xchg + 6 * (size_t)s, epoch, devStatus);
        if (isLeft) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                box.lo[k] = omin(box.lo[k], sb.lo[k]);
                box.hi[k] = omax(box.hi[k], sb.hi[k]);
            }
            r = other;
            dr = dNew;
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                box.lo[k] = omin(sb.lo[k], box.lo[k]);
                box.hi[k] = omax(sb.hi[k], box.hi[k]);
            }
            l = other;
            dl = dNew;
        }
        split = parentSplit;
        internal = true;
    }
}

struct Layout {
    uint32_t* sceneOrd;
    uint32_t* hist;
    uint32_t* partCounters;
    uint32_t* status;
    size_t clearWords;  // hist + partCounters + status are contiguous
    float4* boxes;
    void* keysA;
    void* keysB;
    uint32_t* valsA;
    uint32_t* valsB;
    uint32_t* otherEnd;
    Pending* pend;
    size_t bytes;
};

inline int numPassesFor(uint32_t bits) { return (int)((3 * bits + 7) / 8); }
inline uint32_t tileFor(uint32_t bits) {
    return bits <= 10 ? (uint32_t)(kSortBlock * SortCfg<uint32_t>::kItems)
                      : (uint32_t)(kSortBlock * SortCfg<uint64_t>::kItems);
}

Layout layoutFor(Carve& c, uint32_t n, size_t keyBytes, int P, uint32_t tile) {
    Layout L{};
    const uint32_t parts = (n + tile - 1) / tile;
    L.sceneOrd = c.take<uint32_t>(8);
    // hist, counters and status carved back to back so one fill clears them
    L.hist = c.take<uint32_t>((size_t)kMaxPasses * 256 + 64 + (size_t)P * parts * 256);
    L.partCounters = L.hist + (size_t)kMaxPasses * 256;
    L.status = L.partCounters + 64;
    L.clearWords = (size_t)kMaxPasses * 256 + 64 + (size_t)P * parts * 256;
    L.boxes = c.take<float4>(2 * (size_t)n);
    L.keysA = c.take<char>(keyBytes * n);
    L.keysB = c.take<char>(keyBytes * n);
    L.valsA = c.take<uint32_t>(n);
    L.valsB = c.take<uint32_t>(n);
    L.otherEnd = c.take<uint32_t>(n);
    L.pend = c.take<Pending>(2 * (size_t)((n + kLocalLeaves - 1) / kLocalLeaves));
    L.bytes = c.off + 256;
    return L;
}
Layout layoutFor(void* base, uint32_t n, uint32_t bits) {
    Carve c(base);
    return layoutFor(c, n, bits <= 10 ? 4 : 8, numPassesFor(bits), tileFor(bits));
}

inline unsigned gridFor(uint32_t n, unsigned cap) {
    unsigned g = (unsigned)((n + kBlock - 1) / kBlock);
    return g < 1 ? 1 : (g > cap ? cap : g);
}

// Working memory of one build: the layout region, the hierarchy's exchange
// words (six per internal node) and the epoch word that tags them. The
// context's pool and exchange buffer, or the no-allocation entry point's
// static array.
struct BuildScratch {
    void* base;
    uint64_t* xchg;
    size_t xchgNodes;
    uint32_t* epochWord;
};

template <typename KeyT, bool SOUP>
int runBuild(clr_ctx* ctx, const LbvhInput& in, uint32_t n, uint32_t bits,
             clr_lbvh_inode* inodes, clr_lbvh_leaf* leaves, hipStream_t stream, const BuildScratch& W) {
    uint32_t* epochWord = W.epochWord;
    Layout L = layoutFor(W.base, n, bits);
    const int P = numPassesFor(bits);
    const uint32_t tile = tileFor(bits);
    const uint32_t parts = (n + tile - 1) / tile;
    const int keyBits = (int)(3 * bits);
    const bool prof = ctx->profiling && ctx->profiledBuilds < clr_ctx::kMaxProfiledBuilds;
    auto ev = [&](int k) {
        if (prof) (void)hipEventRecord(ctx->events[ctx->profiledBuilds][k], stream);
    };
    ev(0);
    k_init<<<gridFor((uint32_t)std::min<size_t>(L.clearWords, 0xFFFFFFFFu), 2048), kBlock, 0, stream>>>(
        L.hist, L.clearWords, L.sceneOrd, epochWord, reinterpret_cast<uint32_t*>(W.xchg), W.xchgNodes * 12);
    // One set of same-address atomics per block: a few blocks per CU.
    const unsigned reduceGrid = (unsigned)std::max(1, ctx->numCUs) * 4u;
    // Box centres live in buffers the first radix pass has not written yet:
    // the second key buffer and both payload buffers (pass 0 generates its
    // payload instead of reading it).
    float* cx = reinterpret_cast<float*>(L.keysB);
    float* cy = reinterpret_cast<float*>(L.valsA);
    float* cz = reinterpret_cast<float*>(L.valsB);
    k_bounds<SOUP><<<gridFor(n, reduceGrid), kBlock, 0, stream>>>(in.tris, in.vertices, in.faces, n, L.boxes, cx,
                                                                  cy, cz, L.sceneOrd, epochWord);
    ev(1);
    PrimSrc src{in.tris, L.boxes};
    k_morton<KeyT><<<gridFor(n, reduceGrid), kBlock, 0, stream>>>(
        cx, cy, cz, n, bits, L.sceneOrd, static_cast<KeyT*>(L.keysA), L.hist, P);
    ev(2);
    KeyT* kin = static_cast<KeyT*>(L.keysA);
    KeyT* kout = static_cast<KeyT*>(L.keysB);
    uint32_t* vin = L.valsA;
    uint32_t* vout = L.valsB;
    for (int p = 0; p < P; ++p) {
        uint32_t* st = L.status + (size_t)p * parts * 256;
        if (p == 0)
            k_onesweep<KeyT, true><<<parts, kSortBlock, 0, stream>>>(
                kin, vin, kout, vout, n, 8 * p, L.hist + 256 * p, st, L.partCounters + p,
                ctx->devStatus);
        else
            k_onesweep<KeyT, false><<<parts, kSortBlock, 0, stream>>>(
                kin, vin, kout, vout, n, 8 * p, L.hist + 256 * p, st, L.partCounters + p,
                ctx->devStatus);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    ev(3);
    const uint32_t localBlocks = (n + kLocalLeaves - 1) / kLocalLeaves;
    k_hier_local<SOUP, KeyT><<<localBlocks, kHierThreads, 0, stream>>>(
        src, kin, vin, n, keyBits, L.otherEnd, W.xchg, epochWord, inodes, leaves, L.pend, ctx->devStatus);
    ev(4);
    k_hier_global<KeyT><<<(2 * localBlocks + kBlock - 1) / kBlock, kBlock, 0, stream>>>(
        kin, n, keyBits, L.otherEnd, W.xchg, epochWord, ctx->devStatus, inodes, leaves, L.pend,
        2 * localBlocks);
    ev(5);
    if (prof) ++ctx->profiledBuilds;
    return hipCheck(ctx, hipGetLastError(), "LBVH build launch");
}

// ------------------------------------------------------ segmented build ----
// Many scenes (contiguous ranges of one triangle soup) in one set of
// launches. Every scene is split into chunks of at most kSegChunk triangles
// by k_seg_chunks, on the device, from the caller's device offsets; a block
// handles one chunk, so it knows its scene without a search. Keys are
// (scene << 3 * bits) | Morton code within the scene's own box, so one LSD
// sort orders scenes by index and each scene by its codes with the index
// tie-break, exactly as a build of that scene alone would.
constexpr uint32_t kSegChunk = 4096;
constexpr int kSegScanThreads = 1024;
// Control words of a segmented build: the number of chunks, and whether the
// scene offsets were accepted.
constexpr uint32_t kSegWordChunks = 0, kSegWordValid = 1;

// At most this many chunks: sum over scenes of ceil(len / kSegChunk).
inline uint32_t segChunkBound(uint32_t n, uint32_t numSegments) { return n / kSegChunk + numSegments; }

// Chunk table (one block): checks that the offsets start at 0, grow strictly
// and end at n; if not, it raises kStatusSegments and publishes no chunk and
// valid = 0, so the later kernels of the build write nothing. Otherwise chunk
// k is (scene, first, end) in scene order: a block-wide scan of the scenes'
// chunk counts, 1024 scenes per round.
__global__ __launch_bounds__(kSegScanThreads) void k_seg_chunks(const uint32_t* __restrict__ segOff,
                                                                uint32_t numSegments, uint32_t n,
                                                                uint4* __restrict__ chunks,
                                                                uint32_t* __restrict__ words,
                                                                uint32_t* __restrict__ devStatus) {
    __shared__ uint32_t s_wave[kSegScanThreads / 64];
    __shared__ uint32_t s_bad;
    if (threadIdx.x == 0) s_bad = (segOff[0] != 0u || segOff[numSegments] != n) ? 1u : 0u;
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < numSegments; s += kSegScanThreads)
        if (segOff[s + 1] <= segOff[s]) s_bad = 1u;  // every writer stores the same value
    __syncthreads();
    if (s_bad) {
        if (threadIdx.x == 0) {
            words[kSegWordChunks] = 0u;
            words[kSegWordValid] = 0u;
            atomicOr(devStatus, kStatusSegments);
        }
        return;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t base = 0;
    for (uint32_t s0 = 0; s0 < numSegments; s0 += kSegScanThreads) {
        const uint32_t sc = s0 + threadIdx.x;
        uint32_t lo = 0, hi = 0, c = 0;
        if (sc < numSegments) {
            lo = segOff[sc];
            hi = segOff[sc + 1];
            c = (hi - lo + kSegChunk - 1) / kSegChunk;
        }
        uint32_t x = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) s_wave[wave] = x;
        __syncthreads();
        uint32_t pre = 0, tot = 0;
        for (int w = 0; w < kSegScanThreads / 64; ++w) {
            if (w < wave) pre += s_wave[w];
            tot += s_wave[w];
        }
        __syncthreads();  // s_wave is rewritten in the next round
        uint32_t k = base + pre + x - c;
        for (uint32_t b = lo; b < hi; b += kSegChunk, ++k) chunks[k] = make_uint4(sc, b, min(b + kSegChunk, hi), 0u);
        base += tot;
    }
    if (threadIdx.x == 0) {
        words[kSegWordChunks] = base;
        words[kSegWordValid] = 1u;
    }
}

__global__ __launch_bounds__(kBlock) void k_seg_init(uint32_t* segBox, uint32_t numSegments,
                                                     uint32_t* epochWord) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i == 0) advanceEpoch(epochWord);
    if (i < 6 * numSegments) segBox[i] = (i % 6) < 3 ? ordOf(INFINITY) : ordOf(-INFINITY);
}

// Per-scene box of the triangle boxes (block reduction per chunk, then
// ordered-int atomics on the scene's six words).
__global__ __launch_bounds__(kBlock) void k_seg_bounds(const float* __restrict__ tris,
                                                       const uint4* __restrict__ chunks,
                                                       const uint32_t* __restrict__ words,
                                                       uint32_t* __restrict__ segBox) {
    if (blockIdx.x >= words[kSegWordChunks]) return;  // the grid covers the bound on chunks
    const uint4 ch = chunks[blockIdx.x];
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = ch.y + threadIdx.x; i < ch.z; i += kBlock) {
        const Box b = soupBox(tris, i);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = omin(lo[k], b.lo[k]);
            hi[k] = omax(hi[k], b.hi[k]);
        }
    }
    __shared__ float s_red[kBlock / 64][6];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = waveMin(lo[k]);
        hi[k] = waveMax(hi[k]);
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            s_red[wave][k] = lo[k];
            s_red[wave][3 + k] = hi[k];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        float v = s_red[0][k];
        for (int w = 1; w < kBlock / 64; ++w) v = k < 3 ? omin(v, s_red[w][k]) : omax(v, s_red[w][k]);
        uint32_t* box = segBox + 6 * (size_t)ch.x;
        if (k < 3) atomicMin(box + k, ordOf(v));
        else atomicMax(box + k, ordOf(v));
    }
}

// Scene-tagged 64-bit keys, per-position scene ids and the digit histograms
// of all passes.
__global__ __launch_bounds__(kBlock) void k_seg_morton(const float* __restrict__ tris, float4* __restrict__ boxes,
                                                       uint32_t n, uint32_t bits, const uint4* __restrict__ chunks,
                                                       const uint32_t* __restrict__ words,
                                                       const uint32_t* __restrict__ segBox,
                                                       uint64_t* __restrict__ keys, uint32_t* __restrict__ segOf,
                                                       uint32_t* __restrict__ hist, int numPasses) {
    if (blockIdx.x >= words[kSegWordChunks]) return;  // the grid covers the bound on chunks
    __shared__ uint32_t s_hist[kMaxPasses * 256];
    for (int j = threadIdx.x; j < numPasses * 256; j += kBlock) s_hist[j] = 0u;
    __syncthreads();
    const uint4 ch = chunks[blockIdx.x];
    const uint32_t* box = segBox + 6 * (size_t)ch.x;
    float smin[3], ssize[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        smin[k] = floatOf(box[k]);
        ssize[k] = floatOf(box[3 + k]) - smin[k];
    }
    const uint32_t numDivs = 1u << bits;
    const float fDivs = (float)numDivs;
    const uint64_t tag = (uint64_t)ch.x << (3 * bits);
    for (uint32_t i = ch.y + threadIdx.x; i < ch.z; i += kBlock) {
        const Box b = soupBox(tris, i);
        uint32_t q[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float c = (b.lo[k] + b.hi[k]) * 0.5f;
            const float np = (c - smin[k]) / ssize[k];
            q[k] = quantize(np * fDivs, numDivs);
        }
        const uint64_t key =
            tag | spread3<uint64_t>(q[0]) | (spread3<uint64_t>(q[1]) << 1) | (spread3<uint64_t>(q[2]) << 2);
        keys[i] = key;
        segOf[i] = ch.x;
        for (int p = 0; p < numPasses; ++p)
            atomicAdd(&s_hist[p * 256 + (uint32_t)((key >> (8 * p)) & 0xFFu)], 1u);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < numPasses * 256; j += kBlock)
        if (s_hist[j]) atomicAdd(hist + j, s_hist[j]);
}

struct SegLayout {
    Layout L;
    uint32_t* segOf;
    uint32_t* words;
    uint32_t* segBox;
    uint4* chunks;
    size_t bytes;
};

int segKeyBits(uint32_t bits, uint32_t numSegments) {
    int sb = 1;
    while (sb < 32 && (1u << sb) < numSegments) ++sb;
    return (int)(3 * bits) + sb;
}

SegLayout segLayoutFor(void* base, uint32_t n, uint32_t numSegments, int P) {
    Carve c(base);
    SegLayout S;
    S.L = layoutFor(c, n, 8, P, (uint32_t)(kSortBlock * SortCfg<uint64_t>::kItems));
    S.segOf = c.take<uint32_t>(n);
    S.words = c.take<uint32_t>(64);
    S.segBox = c.take<uint32_t>(6 * (size_t)numSegments);
    S.chunks = c.take<uint4>(segChunkBound(n, numSegments));
    S.bytes = c.off + 256;
    return S;
}

}  // namespace

size_t lbvhSegScratchBytes(uint32_t totalTris, uint32_t numSegments, uint32_t bitsPerDim) {
    const int P = (segKeyBits(bitsPerDim, numSegments) + 7) / 8;
    return segLayoutFor(nullptr, totalTris, numSegments, P).bytes;
}

int lbvhBuildSegmented(clr_ctx* ctx, const float* tris, const uint32_t* segOffsets, uint32_t numSegments,
                       uint32_t n, uint32_t bits, clr_lbvh_inode* inodes, clr_lbvh_leaf* leaves,
                       hipStream_t stream) {
    const int keyBits = segKeyBits(bits, numSegments);
    const int P = (keyBits + 7) / 8;
    if (P > kMaxPasses) return fail(ctx, CLR_ERR_UNSUPPORTED, "too many segments for 64-bit keys");
    int rc = ensurePool(ctx, lbvhSegScratchBytes(n, numSegments, bits));
    if (rc) return rc;
    rc = ensureStatus(ctx);
    if (rc) return rc;
    rc = ensureExchange(ctx, n);
    if (rc) return rc;
    uint32_t* epochWord = ctx->devStatus + kWordEpoch;
    SegLayout S = segLayoutFor(ctx->pool, n, numSegments, P);
    Layout& L = S.L;
    const uint32_t maxChunks = segChunkBound(n, numSegments);
    const uint32_t tile = (uint32_t)(kSortBlock * SortCfg<uint64_t>::kItems);
    const uint32_t parts = (n + tile - 1) / tile;
    k_init<<<gridFor((uint32_t)std::min<size_t>(L.clearWords, 0xFFFFFFFFu), 2048), kBlock, 0, stream>>>(
        L.hist, L.clearWords, nullptr, epochWord, reinterpret_cast<uint32_t*>(ctx->xchg), ctx->xchgNodes * 12);
    k_seg_init<<<(6 * numSegments + kBlock - 1) / kBlock, kBlock, 0, stream>>>(S.segBox, numSegments, epochWord);
    k_seg_chunks<<<1, kSegScanThreads, 0, stream>>>(segOffsets, numSegments, n, S.chunks, S.words, ctx->devStatus);
    k_seg_bounds<<<maxChunks, kBlock, 0, stream>>>(tris, S.chunks, S.words, S.segBox);
    k_seg_morton<<<maxChunks, kBlock, 0, stream>>>(tris, L.boxes, n, bits, S.chunks, S.words, S.segBox,
                                                   static_cast<uint64_t*>(L.keysA), S.segOf, L.hist, P);
    uint64_t* kin = static_cast<uint64_t*>(L.keysA);
    uint64_t* kout = static_cast<uint64_t*>(L.keysB);
    uint32_t* vin = L.valsA;
    uint32_t* vout = L.valsB;
    for (int p = 0; p < P; ++p) {
        uint32_t* st = L.status + (size_t)p * parts * 256;
        if (p == 0)
            k_onesweep<uint64_t, true><<<parts, kSortBlock, 0, stream>>>(kin, vin, kout, vout, n, 8 * p,
                                                                         L.hist + 256 * p, st, L.partCounters + p,
                                                                         ctx->devStatus);
        else
            k_onesweep<uint64_t, false><<<parts, kSortBlock, 0, stream>>>(kin, vin, kout, vout, n, 8 * p,
                                                                          L.hist + 256 * p, st, L.partCounters + p,
                                                                          ctx->devStatus);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    const SegView sv{S.segOf, segOffsets, S.words + kSegWordValid};
    PrimSrc src{tris, L.boxes};
    const uint32_t localBlocks = (n + kLocalLeaves - 1) / kLocalLeaves;
    k_hier_local<true, uint64_t, true><<<localBlocks, kHierThreads, 0, stream>>>(
        src, kin, vin, n, keyBits, L.otherEnd, ctx->xchg, epochWord, inodes, leaves, L.pend, ctx->devStatus, sv);
    k_hier_global<uint64_t, true><<<(2 * localBlocks + kBlock - 1) / kBlock, kBlock, 0, stream>>>(
        kin, n, keyBits, L.otherEnd, ctx->xchg, epochWord, ctx->devStatus, inodes, leaves, L.pend, 2 * localBlocks,
        sv);
    return hipCheck(ctx, hipGetLastError(), "segmented LBVH build launch");
}

void fillZero(uint32_t* words, size_t count, hipStream_t stream) {
    k_init<<<gridFor((uint32_t)std::min<size_t>(count, 0xFFFFFFFFu), 2048), kBlock, 0, stream>>>(
        words, count, nullptr);
}

size_t lbvhScratchBytes(uint32_t n, uint32_t bitsPerDim) {
    return layoutFor(nullptr, n, bitsPerDim).bytes;
}

int lbvhBuild(clr_ctx* ctx, const LbvhInput& in, uint32_t n, uint32_t bits,
              clr_lbvh_inode* inodes, clr_lbvh_leaf* leaves, hipStream_t stream) {
    int rc = ensurePool(ctx, lbvhScratchBytes(n, bits));
    if (!rc) rc = ensureStatus(ctx);
    if (!rc) rc = ensureExchange(ctx, n);
    if (rc) return rc;
    const BuildScratch W{ctx->pool, ctx->xchg, ctx->xchgNodes, ctx->devStatus + kWordEpoch};
    const bool soup = in.tris != nullptr;
    if (bits <= 10)
        return soup ? runBuild<uint32_t, true>(ctx, in, n, bits, inodes, leaves, stream, W)
                    : runBuild<uint32_t, false>(ctx, in, n, bits, inodes, leaves, stream, W);
    return soup ? runBuild<uint64_t, true>(ctx, in, n, bits, inodes, leaves, stream, W)
                : runBuild<uint64_t, false>(ctx, in, n, bits, inodes, leaves, stream, W);
}

// ---- the build that allocates nothing ----------------------------------------
// clr_lbvh_build_soup_noalloc's working memory: one static array per device
// (each device that loads this code object holds its own copy until the
// process ends), laid out as 256 B of control words (word 0: the build epoch
// of this array's exchange words), the exchange words for kNoAllocMaxTris
// triangles, then the build layout for kNoAllocMaxTris triangles with 30-bit
// keys. Host code never passes `scratch` itself to a kernel (its host-side
// name is not a device address): lbvhNoAllocInit looks up its device address
// once per device and every later use goes through that pointer.
constexpr uint32_t kNoAllocMaxTris = 1u << 20;
constexpr size_t kNoAllocCtrlBytes = 256;
constexpr size_t kNoAllocXchgBytes = (size_t)kNoAllocMaxTris * 6 * sizeof(uint64_t);
constexpr size_t kNoAllocLayoutBytes = (size_t)56 << 20;  // >= layoutFor(kNoAllocMaxTris, 10 bits), checked
constexpr size_t kNoAllocScratchBytes = kNoAllocCtrlBytes + kNoAllocXchgBytes + kNoAllocLayoutBytes;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_noAllocMutex;
void* g_noAllocPtr[kMaxDevices] = {};

int lbvhNoAllocInit(clr_ctx* ctx) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
        return fail(ctx, CLR_ERR_UNSUPPORTED, "device ordinal beyond the no-allocation table");
    if (layoutFor(nullptr, kNoAllocMaxTris, 10).bytes > kNoAllocLayoutBytes)
        return fail(ctx, CLR_ERR_UNSUPPORTED, "no-allocation scratch smaller than its layout");
    std::lock_guard<std::mutex> lock(g_noAllocMutex);
    void*& slot = g_noAllocPtr[ctx->device];
    if (!slot) {
        void* ptr = nullptr;
        int rc = hipCheck(ctx, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "no-allocation scratch address");
        if (rc) return rc;
        // Once per device: the epoch starts at 0 and no exchange word carries
        // a later epoch.
        fillZero(static_cast<uint32_t*>(ptr), (kNoAllocCtrlBytes + kNoAllocXchgBytes) / 4, nullptr);
        rc = hipCheck(ctx, hipDeviceSynchronize(), "no-allocation scratch clear");
        if (rc) return rc;
        slot = ptr;
    }
    ctx->noAllocScratch = slot;
    return CLR_OK;
}

uint32_t lbvhNoAllocMaxTris() { return kNoAllocMaxTris; }

int lbvhBuildSoupNoAlloc(clr_ctx* ctx, const float* tris, uint32_t n, uint32_t bits, clr_lbvh_inode* inodes,
                         clr_lbvh_leaf* leaves, hipStream_t stream) {
    if (!ctx->noAllocScratch) return fail(ctx, CLR_ERR_HIP, "no-allocation scratch unavailable on this device");
    if (n > kNoAllocMaxTris)
        return fail(ctx, CLR_ERR_UNSUPPORTED, "numTris " + std::to_string(n) + " exceeds the no-allocation limit of " +
                                                  std::to_string(kNoAllocMaxTris));
    if (bits > 10) return fail(ctx, CLR_ERR_UNSUPPORTED, "the no-allocation build supports numBitsPerDim 1..10");
    char* base = static_cast<char*>(ctx->noAllocScratch);
    const BuildScratch W{base + kNoAllocCtrlBytes + kNoAllocXchgBytes, reinterpret_cast<uint64_t*>(base + kNoAllocCtrlBytes),
                         kNoAllocMaxTris, reinterpret_cast<uint32_t*>(base)};
    LbvhInput in;
    in.tris = tris;
    return runBuild<uint32_t, true>(ctx, in, n, bits, inodes, leaves, stream, W);
}

}  // namespace clr
