// Wavefront path tracer for gfx950: the same integrator as the megakernel
// (pathtracing, OpenCL_PathTracer/OpenCL/pathtracer.cl:25-197) split into
// queue-driven kernels, each running only the lanes that have that kind of
// work:
//
//   extend  closest hit for every ray of the extension queue; the entry goes
//           to the hit queue (surface material work) or the miss queue
//           (environment material work);
//   shade   hit queue: emission with MIS, bounce cap, Russian roulette, BSDF
//           set-up, light sample (-> shadow queue, with its contribution
//           already weighted), BSDF sample (-> next extension queue);
//   miss    miss queue: environment emission with MIS, end of the sample;
//   shadow  shadow queue: occlusion (any-hit for environment samples, closest
//           hit compared with the light's face for area samples), adds the
//           contribution, ends a sample whose BSDF sample failed.
//
// One path slot per pixel of a chunk; a slot runs its pixel's samples one
// after the other and starts the next camera path as soon as one ends, so the
// queues stay full until the last samples drain. Per pixel, the random draws
// and the order of radiance additions are those of the megakernel (emission
// of a vertex, then its light sample's contribution, then the next vertex),
// so both renderers produce the same image for the same RNG states.
//
// Kernels grid-stride over device-side queue counts, so an iteration can be
// launched without knowing its size: a render is a fixed sequence of
// iterations (an upper bound of spp * (maxDepth + 1)), captured in a hipGraph
// of G iterations and replayed, with an optional early stop when the
// extension queue runs dry. Block shapes, resident waves per CU, G and the
// early-stop period come from WfConfig, picked by clr_wavefront_autotune.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.h"
#include "render_device.h"

namespace clr {

int prepareRenderScene(clr_ctx* ctx, const clr_scene_dev& sd, dev::RScene& S, hipStream_t stream);

namespace {

using namespace dev;

constexpr int kMinLdsDepth = 16;  // fewest traversal stack slots kept in LDS (spill sized for it)
constexpr uint32_t kFlagEndAfterShadow = 1u << 16;

// Device-side counters (one 256-B line).
enum : uint32_t {
    kCntExt0 = 0, kCntExt1 = 1, kCntShadow = 2, kCntHit = 3, kCntMiss = 4,
    kCntChunk = 8,
};

struct WfBufs {
    uint4* rng;      // [P] xorshift128 state
    float4* alpha;   // [P] xyz path throughput, w initial luminance
    float4* pix;     // [P] xyz radiance of the current sample, w dirPDF
    float4* acc;     // [P] xyz sum over finished samples, w samples left (bits)
    uint32_t* meta;  // [P] depth | sampledType << 8 | flags
    uint32_t* rqSlot[2];
    float4* rqOrg[2];
    float4* rqDir[2];
    uint4* hit;      // [P] face, t, b1, b2 (extension-queue order)
    uint32_t* hitQ;  // [P] extension-queue indices that hit
    uint32_t* missQ; // [P] extension-queue indices that missed
    uint32_t* sqSlot;
    float4* sqOrg;   // xyz origin, w target face (bits; kNone = environment)
    float4* sqDir;
    float4* sqC;     // weighted contribution
    uint32_t* counts;
    uint32_t* spill; // traversal stack overflow regions, 2 x (kStackPairs - kMinLdsDepth) x lanes
    uint32_t lanes;
    uint32_t slots;  // P: capacity of every queue
};

struct WfArgs {
    RScene S;                 // single-scene renders
    const RScene* table;      // batched renders: resolved scenes, one per image
    WfBufs B;
    uint32_t width, height, maxDepth, flags, refillMin;
    uint32_t posPerScene;     // tile-major positions per image
    uint32_t* states;         // [scene][pixel]
    float4* pixels;           // [scene][pixel]
};

// The scene and image position of a slot. BATCH: positions are scene-major
// (posPerScene per image) and the scene comes from the table; otherwise S is
// the kernel's own resolved copy of A.S.
template <bool BATCH>
__device__ __forceinline__ void slotScene(const WfArgs& A, uint32_t v, RScene& S, uint32_t& vLocal,
                                          uint32_t& pixBase) {
    if constexpr (BATCH) {
        const uint32_t sc = v / A.posPerScene;
        vLocal = v - sc * A.posPerScene;
        pixBase = sc * A.width * A.height;
        S = A.table[sc];
    } else {
        vLocal = v;
        pixBase = 0;
    }
}

// Scene a traversal walker starts bound to: the kernel's private copy in
// single-scene mode, table entry 0 in batch mode (where every ray rebinds).
// A pointer that could be either would be a generic address and force the
// private copy into scratch memory.
template <bool BATCH>
__device__ __forceinline__ const RScene* firstScene(const WfArgs& A, const RScene& S) {
    if constexpr (BATCH) return A.table;
    else return &S;
}

size_t wfBytes(uint32_t P, uint32_t lanes) {
    size_t b = 0;
    auto add = [&](size_t x) { b += (x + 255) & ~size_t(255); };
    add(16ull * P); add(16ull * P); add(16ull * P); add(16ull * P); add(4ull * P);
    for (int q = 0; q < 2; ++q) { add(4ull * P); add(16ull * P); add(16ull * P); }
    add(16ull * P); add(4ull * P); add(4ull * P);
    add(4ull * P); add(16ull * P); add(16ull * P); add(16ull * P);
    add(256);
    add(4ull * 2 * (kStackPairs - kMinLdsDepth) * lanes);
    return b;
}

WfBufs carveWf(void* pool, uint32_t P, uint32_t lanes) {
    Carve c(pool);
    WfBufs B;
    B.rng = c.take<uint4>(P);
    B.alpha = c.take<float4>(P);
    B.pix = c.take<float4>(P);
    B.acc = c.take<float4>(P);
    B.meta = c.take<uint32_t>(P);
    for (int q = 0; q < 2; ++q) {
        B.rqSlot[q] = c.take<uint32_t>(P);
        B.rqOrg[q] = c.take<float4>(P);
        B.rqDir[q] = c.take<float4>(P);
    }
    B.hit = c.take<uint4>(P);
    B.hitQ = c.take<uint32_t>(P);
    B.missQ = c.take<uint32_t>(P);
    B.sqSlot = c.take<uint32_t>(P);
    B.sqOrg = c.take<float4>(P);
    B.sqDir = c.take<float4>(P);
    B.sqC = c.take<float4>(P);
    B.counts = c.take<uint32_t>(64);
    B.spill = c.take<uint32_t>((size_t)2 * (kStackPairs - kMinLdsDepth) * lanes);
    B.lanes = lanes;
    B.slots = P;
    return B;
}

// Block-aggregated queue append: one atomic per block and call (appends to
// one address serialise in the memory system, so their number matters more
// than their size); returns this lane's position (meaningful where `want` is
// set). Every thread of the block must call it (blocks of at most 256).
__device__ __forceinline__ uint32_t appendIndex(uint32_t* counter, bool want) {
    __shared__ uint32_t s_part[5];
    const uint64_t mask = __ballot(want);
    const uint32_t wave = threadIdx.x / 64u, lane = __lane_id();
    if (lane == 0) s_part[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (uint32_t w = 0; w < blockDim.x / 64u; ++w) {
            const uint32_t c = s_part[w];
            s_part[w] = total;
            total += c;
        }
        s_part[4] = total ? atomicAdd(counter, total) : 0u;
    }
    __syncthreads();
    const uint32_t r = s_part[4] + s_part[wave] + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    __syncthreads();  // s_part is reused by the next call
    return r;
}

// Work distribution of the persistent traversal kernels, without atomics:
// each wave owns one contiguous range of the queue and hands its entries to
// idle lanes (ballot + popcount ranks) once at least refillMin lanes are idle
// (or none is busy), so that ray set-up runs with most lanes together.
// Ranges are grouped by XCD (workgroups are dispatched to the 8 XCDs
// round-robin, so block b runs on XCD b % 8): the waves of one XCD own one
// contiguous eighth of the queue, keeping neighbouring pixels' rays in one
// L2. The grid is a multiple of 8 blocks. (Chunks of 64 entries taken from a
// per-XCD atomic counter balance the waves better but rendered 12 % slower:
// a wave's contiguous range keeps neighbouring rays together.)
struct WavePool {
    uint32_t next, end;
    __device__ __forceinline__ explicit WavePool(uint32_t n) {
        const uint32_t wavesPerBlock = blockDim.x / 64u;
        const uint32_t blocksPerXcd = gridDim.x / 8u;
        const uint32_t logicalBlock = (blockIdx.x % 8u) * blocksPerXcd + blockIdx.x / 8u;
        const uint32_t wave = logicalBlock * wavesPerBlock + threadIdx.x / 64u;
        const uint64_t waves = (uint64_t)gridDim.x * wavesPerBlock;
        next = (uint32_t)((uint64_t)n * wave / waves);
        end = (uint32_t)((uint64_t)n * (wave + 1u) / waves);
    }
    // Returns true in lanes that got an entry (index in `j`). Every active
    // lane of the wave must call it.
    __device__ __forceinline__ bool refill(uint32_t refillMin, bool busy, uint32_t& j) {
        if (next >= end) return false;
        const uint64_t idle = __ballot(!busy);
        const uint32_t nIdle = (uint32_t)__popcll(idle);
        if (nIdle == 0 || (nIdle < refillMin && __ballot(busy) != 0)) return false;
        const uint32_t rank = (uint32_t)__popcll(idle & ((1ull << __lane_id()) - 1ull));
        const uint32_t take = min(nIdle, end - next);
        const bool got = !busy && rank < take;
        if (got) j = next + rank;
        next += take;
        return got;
    }
    __device__ __forceinline__ bool drained() const { return next >= end; }
};

__device__ __forceinline__ float4 f4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
__device__ __forceinline__ V3 xyz(float4 v) { return {v.x, v.y, v.z}; }

// Path slots walk the image in 8x8 tiles (tile-major, row-major inside a
// tile), as the megakernel's waves do, so that a wave's rays start coherent.
// Returns false for positions of partial edge tiles outside the image.
__device__ __forceinline__ bool tilePixel(uint32_t v, uint32_t width, uint32_t height, uint32_t& pixel) {
    const uint32_t tilesX = (width + 7u) / 8u;
    const uint32_t tile = v / 64u, in = v % 64u;
    const uint32_t x = (tile % tilesX) * 8u + in % 8u, y = (tile / tilesX) * 8u + in / 8u;
    pixel = y * width + x;
    return x < width && y < height;
}

// Starts the next camera path of `pixel` (cameraRay consumes 4 draws).
__device__ __forceinline__ void startPath(const RScene& S, uint32_t pixel, uint32_t width, Rng& rng, V3& org,
                                          V3& dir, V3& alpha, float& initY, float& dirPDF) {
    float w;
    cameraRay(S, pixel % width, pixel / width, rng, org, dir, w, dirPDF);
    alpha = v3s(1.0f) * w;
    initY = lum(alpha);
}

// End of one sample of `slot`: acc += pix; start the next sample (returns
// true with the new ray) or, after the last one, write the pixel and state.
__device__ __forceinline__ bool endSample(const WfArgs& A, const RScene& S, uint32_t slot, uint32_t vLocal,
                                          uint32_t pixBase, Rng& rng, V3 pix, V3& org, V3& dir) {
    const WfBufs& B = A.B;
    float4 acc = B.acc[slot];
    acc.x += pix.x;
    acc.y += pix.y;
    acc.z += pix.z;
    const uint32_t left = __float_as_uint(acc.w) - 1u;
    acc.w = __uint_as_float(left);
    B.acc[slot] = acc;
    uint32_t pixel;
    tilePixel(vLocal, A.width, A.height, pixel);
    if (left > 0) {
        V3 alpha;
        float initY, dirPDF;
        startPath(S, pixel, A.width, rng, org, dir, alpha, initY, dirPDF);
        B.alpha[slot] = f4(alpha, initY);
        B.pix[slot] = make_float4(0.0f, 0.0f, 0.0f, dirPDF);
        B.meta[slot] = 0u;
        B.rng[slot] = make_uint4(rng.s0, rng.s1, rng.s2, rng.s3);
        return true;
    }
    float4 p = A.pixels[pixBase + pixel];
    p.x += acc.x;
    p.y += acc.y;
    p.z += acc.z;
    A.pixels[pixBase + pixel] = p;
    reinterpret_cast<uint4*>(A.states)[pixBase + pixel] = make_uint4(rng.s0, rng.s1, rng.s2, rng.s3);
    return false;
}

__device__ __forceinline__ void pushExt(const WfBufs& B, uint32_t q, bool want, uint32_t slot, V3 org, V3 dir) {
    const uint32_t k = appendIndex(B.counts + q, want);
    if (want) {
        B.rqSlot[q][k] = slot;
        B.rqOrg[q][k] = f4(org, 0.0f);
        B.rqDir[q][k] = f4(dir, 0.0f);
    }
}

// First camera path of every slot of the chunk [chunk, chunk + count).
// Clears the device counters before a chunk (separate launch: k_wf_init
// appends to the first extension queue from every block).
__global__ void k_wf_reset(WfArgs A, uint32_t chunk) {
    const uint32_t i = threadIdx.x;
    if (i < 64) A.B.counts[i] = i == kCntChunk ? chunk : 0u;
}

// First camera path of every in-image position of the chunk
// [chunk, chunk + count) of tile-major positions.
template <bool BATCH>
__global__ void k_wf_init(WfArgs A, uint32_t chunk, uint32_t count, uint32_t spp) {
    const WfBufs& B = A.B;
    RScene S = A.S;
    if (!BATCH) finishScene(S);
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t pixel = 0, vLocal = 0, pixBase = 0;
    if (slot < count) slotScene<BATCH>(A, chunk + slot, S, vLocal, pixBase);
    const bool valid = slot < count && tilePixel(vLocal, A.width, A.height, pixel);
    V3 org = v3z(), dir = v3z();
    if (valid) {
        const uint4 s = reinterpret_cast<const uint4*>(A.states)[pixBase + pixel];
        Rng rng{s.x, s.y, s.z, s.w};
        V3 alpha;
        float initY, dirPDF;
        startPath(S, pixel, A.width, rng, org, dir, alpha, initY, dirPDF);
        B.rng[slot] = make_uint4(rng.s0, rng.s1, rng.s2, rng.s3);
        B.alpha[slot] = f4(alpha, initY);
        B.pix[slot] = make_float4(0.0f, 0.0f, 0.0f, dirPDF);
        B.acc[slot] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(spp));
        B.meta[slot] = 0u;
    }
    pushExt(B, 0, valid, slot, org, dir);
}

// Closest hit of extension queue `in`; resets the counters the shade / miss /
// shadow kernels of this iteration append to.
template <bool ORDERED, int LD, bool BATCH, int K = 1>
__global__ __launch_bounds__(256) void k_wf_extend(WfArgs A, uint32_t in) {
    extern __shared__ uint32_t s_stack[];
    const WfBufs& B = A.B;
    const RScene S = A.S;
    const uint32_t chunk = B.counts[kCntChunk];
    const uint32_t out = in ^ 1u;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lanes = gridDim.x * blockDim.x;
    if (gid == 0) {
        B.counts[out] = 0u;
        B.counts[kCntShadow] = 0u;
    }
    const uint32_t n = min(B.counts[in], B.slots);
    const SpillStack<LD> stk{s_stack + threadIdx.x, (int)blockDim.x, B.spill + gid, lanes};
    bool over = false;
    if constexpr (ORDERED) {
        // Persistent waves refilling idle lanes from a wave pool; hit / miss
        // queues are built by k_wf_compact. With K = 2 each lane walks two
        // rays at once, stepping them in turn, so that one ray's node fetch
        // overlaps the other's tests; each walker has its own stack.
        const RScene* S0 = firstScene<BATCH>(A, S);
        const SpillStack<LD> stk1{s_stack + LD * blockDim.x + threadIdx.x, (int)blockDim.x,
                                  B.spill + (size_t)(kStackPairs - kMinLdsDepth) * lanes + gid, lanes};
        using W = Wide<false, TexAlpha, SpillStack<LD>, kStackPairs>;
        W w0(S0->g, stk, TexAlpha{S0}), w1(S0->g, stk1, TexAlpha{S0});
        WavePool pool(n);
        bool busy0 = false, busy1 = false;
        uint32_t j0 = 0, j1 = 0;
        auto start = [&](W& w, uint32_t j) {
            if constexpr (BATCH) {
                const RScene* Sp = A.table + (chunk + B.rqSlot[in][j]) / A.posPerScene;
                w.rebind(Sp->g, TexAlpha{Sp});
            }
            w.begin(xyz(B.rqOrg[in][j]), xyz(B.rqDir[in][j]), INFINITY);
        };
        auto advance = [&](W& w, uint32_t j, bool& busy) {
            if (!w.done) w.step();
            if (w.done) {
                over |= w.over;
                const bool isHit = w.best != kNone && !isinf(w.t);
                B.hit[j] = make_uint4(isHit ? w.best : kNone, __float_as_uint(w.t), __float_as_uint(w.b1),
                                      __float_as_uint(w.b2));
                busy = false;
            }
        };
        while (true) {
            if (pool.refill(A.refillMin, busy0, j0)) {
                start(w0, j0);
                busy0 = true;
            }
            if constexpr (K == 2) {
                if (pool.refill(A.refillMin, busy1, j1)) {
                    start(w1, j1);
                    busy1 = true;
                }
            }
            if (__ballot(busy0 || busy1) == 0 && pool.drained()) break;
            if (busy0) advance(w0, j0, busy0);
            if constexpr (K == 2)
                if (busy1) advance(w1, j1, busy1);
        }
        if (over) atomicOr(S.status, kStatusStackOverflow);
        return;
    }
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += lanes) {
        const uint32_t j = base + threadIdx.x;
        const bool active = j < n;
        bool isHit = false;
        if (active) {
            const V3 o = xyz(B.rqOrg[in][j]), d = xyz(B.rqDir[in][j]);
            float t = INFINITY, b1 = 0.0f, b2 = 0.0f;
            TraceStats st{0u, 0u};
            const uint32_t face =
                closestHit<false, TexAlpha>(S.g, o, d, t, stk, st, &b1, &b2, TexAlpha{&S}, &over);
            isHit = face != kNone && !isinf(t);
            B.hit[j] = make_uint4(isHit ? face : kNone, __float_as_uint(t), __float_as_uint(b1),
                                  __float_as_uint(b2));
        }
        const uint32_t kh = appendIndex(B.counts + kCntHit, active && isHit);
        const uint32_t km = appendIndex(B.counts + kCntMiss, active && !isHit);
        if (active) {
            if (isHit) B.hitQ[kh] = j;
            else B.missQ[km] = j;
        }
    }
    if (over) atomicOr(S.status, kStatusStackOverflow);
}

// Surface vertices (hit queue of extension queue `in`).
template <bool BATCH>
__global__ __launch_bounds__(256, 3) void k_wf_shade(WfArgs A, uint32_t in) {
    RScene S = A.S;
    if (!BATCH) finishScene(S);
    const WfBufs& B = A.B;
    const uint32_t chunk = B.counts[kCntChunk];
    const uint32_t out = in ^ 1u;
    const uint32_t lanes = gridDim.x * blockDim.x;
    const uint32_t n = min(B.counts[kCntHit], B.slots);
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += lanes) {
        const uint32_t i = base + threadIdx.x;
        const bool active = i < n;
        bool wantRay = false, wantShadow = false;
        uint32_t slot = 0, target = kNone;
        V3 org = v3z(), dir = v3z(), sorg = v3z(), sdir = v3z(), contrib = v3z();
        if (active) {
            const uint32_t j = B.hitQ[i];
            slot = B.rqSlot[in][j];
            uint32_t vLocal, pixBase;
            slotScene<BATCH>(A, chunk + slot, S, vLocal, pixBase);
            org = xyz(B.rqOrg[in][j]);
            dir = xyz(B.rqDir[in][j]);
            const uint4 h = B.hit[j];
            const uint4 rs = B.rng[slot];
            Rng rng{rs.x, rs.y, rs.z, rs.w};
            const float4 a4 = B.alpha[slot];
            const float4 p4 = B.pix[slot];
            V3 alpha = xyz(a4), pix = xyz(p4);
            const float initY = a4.w;
            float dirPDF = p4.w;
            uint32_t meta = B.meta[slot];
            uint32_t depth = meta & 0xFFu, sampledType = (meta >> 8) & 0xFFu;

            SurfPt sp;
            hitParams(S, org, dir, __uint_as_float(h.y), h.x, __uint_as_float(h.z), __uint_as_float(h.w), sp);
            const clr_face F = S.g.faces[h.x];
            const V3 vout = -dir;
            EDF E;
            if (F.lightPtr != 0xFFFFu) {
                edfArea(S, F.lightPtr, sp, sp.u, sp.v, E);
                float mis = 1.0f;
                if (depth > 0 && (sampledType & kNonSingular)) {
                    const V3 dd = sp.p - org;
                    const float d2 = dd.x * dd.x + dd.y * dd.y + dd.z * dd.z;
                    const float lightPDF = areaPdfOfFace(S, h.x) * d2 / absCosEDF(E, vout);
                    mis = emissionHitMis(A.flags, dirPDF, lightPDF);
                }
                if (depth == 0) pix += alpha * evalLe(E, vout);
                else pix += (mis * alpha) * evalLe(E, vout);
            }
            bool live = true;
            if (depth > 0) {
                if (depth >= A.maxDepth) {
                    live = false;
                } else if (!(A.flags & CLR_RENDER_NO_RR)) {
                    const float cp = fminf(lum(alpha) * (1.0f / initY), 1.0f);
                    if (rng.uniform() < cp) {
                        alpha = alpha * (1.0f / cp);
                        dirPDF *= cp;
                    } else {
                        live = false;
                    }
                }
            }
            if (live) {
                BSDF Bs;
                bsdfAlloc<true>(S, F.matPtr, sp, Bs);
                if (hasNonSpecular(Bs)) {
                    const float lu = rng.uniform(), l0 = rng.uniform(), l1 = rng.uniform();
                    SurfPt lp;
                    bool atInf = false;
                    float lightPDF;
                    sampleLight(S, lu, l0, l1, lp, atInf, E, lightPDF);
                    V3 vinL = lp.p - sp.p;
                    float dist2 = dot(vinL, vinL);
                    vinL = vinL * (1.0f / sqrtf(dist2));
                    if (atInf) {
                        vinL = lp.p;
                        dist2 = 1.0f;
                    }
                    // With BSDF samples alone no shadow ray is cast (the numbers are drawn).
                    if (lightPDF > 0.0f && lightPDF < INFINITY && !(A.flags & CLR_RENDER_BSDF_SAMPLING_ONLY)) {
                        const float fsPDF = evalFsPdf<true>(Bs, vout, vinL) * absCosEDF(E, vinL) / dist2;
                        const float mis = lightSampleMis(A.flags, lightPDF, fsPDF);
                        const float wgt = mis * absCosEDF(E, vinL) * fabsf(dot(Bs.N, vinL)) / dist2 / lightPDF;
                        contrib = wgt * ((alpha * evalLe(E, -vinL)) * evalFs<true>(Bs, vout, vinL));
                        sorg = sp.p + vinL * kRayEps;
                        sdir = vinL;
                        target = atInf ? kNone : lp.face;
                        wantShadow = true;
                    }
                }
                const float bc = rng.uniform(), b0 = rng.uniform(), b1 = rng.uniform();
                V3 nd = v3z();
                const V3 f = sampleFs<true>(Bs, vout, bc, b0, b1, nd, dirPDF, sampledType);
                if ((f.x == 0.0f && f.y == 0.0f && f.z == 0.0f) || dirPDF == 0.0f) {
                    live = false;
                } else {
                    alpha = alpha * (f * (fabsf(dot(Bs.N, nd)) / dirPDF));
                    dir = nd;
                    org = sp.p + dir * kRayEps;
                    ++depth;
                    wantRay = true;
                }
            }
            meta = depth | (sampledType << 8);
            if (!live && wantShadow) {
                // The light sample's contribution comes first; the shadow
                // kernel ends the sample.
                meta |= kFlagEndAfterShadow;
                B.pix[slot] = f4(pix, dirPDF);
                B.rng[slot] = make_uint4(rng.s0, rng.s1, rng.s2, rng.s3);
            } else if (!live) {
                wantRay = endSample(A, S, slot, vLocal, pixBase, rng, pix, org, dir);
                meta = 0u;
            } else {
                B.alpha[slot] = f4(alpha, initY);
                B.pix[slot] = f4(pix, dirPDF);
                B.rng[slot] = make_uint4(rng.s0, rng.s1, rng.s2, rng.s3);
            }
            if (live || wantShadow) B.meta[slot] = meta;
        }
        pushExt(B, out, active && wantRay, slot, org, dir);
        const uint32_t ks = appendIndex(B.counts + kCntShadow, active && wantShadow);
        if (active && wantShadow) {
            B.sqSlot[ks] = slot;
            B.sqOrg[ks] = f4(sorg, __uint_as_float(target));
            B.sqDir[ks] = f4(sdir, 0.0f);
            B.sqC[ks] = f4(contrib, 0.0f);
        }
    }
}

// Environment vertices (miss queue of extension queue `in`).
template <bool BATCH>
__global__ __launch_bounds__(256) void k_wf_miss(WfArgs A, uint32_t in) {
    RScene S = A.S;
    if (!BATCH) finishScene(S);
    const WfBufs& B = A.B;
    const uint32_t chunk = B.counts[kCntChunk];
    const uint32_t out = in ^ 1u;
    const uint32_t lanes = gridDim.x * blockDim.x;
    const uint32_t n = min(B.counts[kCntMiss], B.slots);
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += lanes) {
        const uint32_t i = base + threadIdx.x;
        const bool active = i < n;
        bool wantRay = false;
        uint32_t slot = 0;
        V3 org = v3z(), dir = v3z();
        if (active) {
            const uint32_t j = B.missQ[i];
            slot = B.rqSlot[in][j];
            uint32_t vLocal, pixBase;
            slotScene<BATCH>(A, chunk + slot, S, vLocal, pixBase);
            dir = xyz(B.rqDir[in][j]);
            const uint4 rs = B.rng[slot];
            Rng rng{rs.x, rs.y, rs.z, rs.w};
            const V3 alpha = xyz(B.alpha[slot]);
            const float4 p4 = B.pix[slot];
            V3 pix = xyz(p4);
            const float dirPDF = p4.w;
            const uint32_t meta = B.meta[slot];
            const uint32_t depth = meta & 0xFFu, sampledType = (meta >> 8) & 0xFFu;
            const V3 vout = -dir;
            const float theta = acosf(clampf(dir.y, -1.0f, 1.0f));
            const float phi = fmodf(atan2f(-dir.x, dir.z) + 2.0f * kPiF, 2 * kPiF);
            const float u = phi / (2.0f * kPiF), v = theta / kPiF;
            EDF E;
            edfEnv(S, u, v, E);
            float mis = 1.0f;
            if ((sampledType & kNonSingular) != 0 && depth > 0) {
                const float lightPDF = areaPdfEnv(S, u, v) * 1.0f / absCosEDF(E, vout);
                mis = emissionHitMis(A.flags, dirPDF, lightPDF);
            }
            pix += (mis * alpha) * evalLe(E, vout);
            wantRay = endSample(A, S, slot, vLocal, pixBase, rng, pix, org, dir);
        }
        pushExt(B, out, active && wantRay, slot, org, dir);
    }
}

// Shadow rays of this iteration; resets the hit / miss counters for the next
// iteration's extend kernel.
template <bool ORDERED, int LD, bool BATCH, int K = 1>
__global__ __launch_bounds__(256) void k_wf_shadow(WfArgs A, uint32_t in) {
    RScene S = A.S;
    if (!BATCH) finishScene(S);
    const uint32_t chunk = A.B.counts[kCntChunk];
    extern __shared__ uint32_t s_stack[];
    const WfBufs& B = A.B;
    const uint32_t out = in ^ 1u;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lanes = gridDim.x * blockDim.x;
    if (gid == 0) {
        B.counts[kCntHit] = 0u;
        B.counts[kCntMiss] = 0u;
    }
    const uint32_t n = min(B.counts[kCntShadow], B.slots);
    const SpillStack<LD> stk{s_stack + threadIdx.x, (int)blockDim.x, B.spill + gid, lanes};
    bool over = false;
    if constexpr (ORDERED) {
        // Persistent waves as in k_wf_extend; environment samples use the
        // occlusion walk, area samples the closest hit compared with the
        // light's face. Samples ending here are finished by k_wf_shadow_end.
        // K walkers per lane, as in k_wf_extend (a path slot has at most one
        // shadow ray per iteration, so the pixel updates never collide).
        const RScene* S0 = firstScene<BATCH>(A, S);
        const SpillStack<LD> stk1{s_stack + LD * blockDim.x + threadIdx.x, (int)blockDim.x,
                                  B.spill + (size_t)(kStackPairs - kMinLdsDepth) * lanes + gid, lanes};
        using W = Wide<false, TexAlpha, SpillStack<LD>, kStackPairs>;
        W w0(S0->g, stk, TexAlpha{S0}), w1(S0->g, stk1, TexAlpha{S0});
        WavePool pool(n);
        bool busy0 = false, busy1 = false;
        uint32_t i0 = 0, i1 = 0, target0 = kNone, target1 = kNone;
        auto start = [&](W& w, uint32_t i, uint32_t& target) {
            if constexpr (BATCH) {
                const RScene* Sp = A.table + (chunk + B.sqSlot[i]) / A.posPerScene;
                w.rebind(Sp->g, TexAlpha{Sp});
            }
            const float4 o4 = B.sqOrg[i];
            target = __float_as_uint(o4.w);
            w.begin(xyz(o4), xyz(B.sqDir[i]), INFINITY, target == kNone);
        };
        auto advance = [&](W& w, uint32_t i, uint32_t target, bool& busy) {
            if (!w.done) w.step();
            if (w.done) {
                over |= w.over;
                const bool hit = w.best != kNone && !isinf(w.t);
                const bool visible = target == kNone ? !hit : (hit && w.best == target);
                if (visible) {
                    const uint32_t slot = B.sqSlot[i];
                    const float4 c = B.sqC[i];
                    float4 p4 = B.pix[slot];
                    p4.x += c.x;
                    p4.y += c.y;
                    p4.z += c.z;
                    B.pix[slot] = p4;
                }
                busy = false;
            }
        };
        while (true) {
            if (pool.refill(A.refillMin, busy0, i0)) {
                start(w0, i0, target0);
                busy0 = true;
            }
            if constexpr (K == 2) {
                if (pool.refill(A.refillMin, busy1, i1)) {
                    start(w1, i1, target1);
                    busy1 = true;
                }
            }
            if (__ballot(busy0 || busy1) == 0 && pool.drained()) break;
            if (busy0) advance(w0, i0, target0, busy0);
            if constexpr (K == 2)
                if (busy1) advance(w1, i1, target1, busy1);
        }
        if (over) atomicOr(S.status, kStatusStackOverflow);
        return;
    }
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += lanes) {
        const uint32_t i = base + threadIdx.x;
        const bool active = i < n;
        bool wantRay = false;
        uint32_t slot = 0;
        V3 org = v3z(), dir = v3z();
        if (active) {
            slot = B.sqSlot[i];
            const float4 o4 = B.sqOrg[i];
            const uint32_t target = __float_as_uint(o4.w);
            const V3 so = xyz(o4), sd = xyz(B.sqDir[i]);
            bool visible;
            float t = INFINITY, b1, b2;
            TraceStats st{0u, 0u};
            if (target == kNone) {
                visible = !anyHit<TexAlpha>(S.g, so, sd, stk, TexAlpha{&S}, &over);
            } else {
                const uint32_t face =
                    closestHit<false, TexAlpha>(S.g, so, sd, t, stk, st, &b1, &b2, TexAlpha{&S}, &over);
                visible = face != kNone && !isinf(t) && face == target;
            }
            float4 p4 = B.pix[slot];
            if (visible) {
                const float4 c = B.sqC[i];
                p4.x += c.x;
                p4.y += c.y;
                p4.z += c.z;
            }
            if (B.meta[slot] & kFlagEndAfterShadow) {
                const uint4 rs = B.rng[slot];
                Rng rng{rs.x, rs.y, rs.z, rs.w};
                wantRay = endSample(A, S, slot, chunk + slot, 0u, rng, xyz(p4), org, dir);
            } else {
                B.pix[slot] = p4;
            }
        }
        pushExt(B, out, active && wantRay, slot, org, dir);
    }
    if (over) atomicOr(S.status, kStatusStackOverflow);
}

// Hit / miss queues of extension queue `in` from the ordered extend kernel's
// records.
__global__ __launch_bounds__(256) void k_wf_compact(WfArgs A, uint32_t in) {
    // Each block takes one contiguous range of the queue, counts its hits
    // and misses, claims room in both queues with one atomic each, then
    // writes its entries in order (block-wide scans). Appends to one counter
    // serialise in the memory system, so one pair per block instead of one
    // per block and 256 entries is what this kernel's time is made of.
    __shared__ uint32_t s_sum[2][4];
    __shared__ uint32_t s_base[2];
    const WfBufs& B = A.B;
    const uint32_t n = min(B.counts[in], B.slots);
    const uint32_t per = (n + gridDim.x - 1) / gridDim.x;
    const uint32_t lo = min(n, blockIdx.x * per), hi = min(n, lo + per);
    if (lo >= hi) return;  // the whole block: lo and hi are block-uniform
    const uint32_t t = threadIdx.x, lane = __lane_id(), wave = t / 64u, waves = blockDim.x / 64u;
    auto isHitAt = [&](uint32_t j) { return reinterpret_cast<const uint32_t*>(B.hit + j)[0] != kNone; };
    uint32_t hits = 0;
    for (uint32_t j = lo + t; j < hi; j += blockDim.x) hits += isHitAt(j) ? 1u : 0u;
    // Block totals, then one atomic per queue.
    for (int off = 32; off > 0; off >>= 1) hits += __shfl_xor(hits, off);
    if (lane == 0) s_sum[0][wave] = hits;
    __syncthreads();
    if (t == 0) {
        uint32_t h = 0;
        for (uint32_t w = 0; w < waves; ++w) h += s_sum[0][w];
        s_base[0] = h ? atomicAdd(B.counts + kCntHit, h) : 0u;
        s_base[1] = (hi - lo) - h ? atomicAdd(B.counts + kCntMiss, (hi - lo) - h) : 0u;
    }
    __syncthreads();
    uint32_t baseH = s_base[0], baseM = s_base[1];
    for (uint32_t j0 = lo; j0 < hi; j0 += blockDim.x) {
        const uint32_t j = j0 + t;
        const bool active = j < hi;
        const bool isHit = active && isHitAt(j);
        const uint64_t mh = __ballot(isHit), mm = __ballot(active && !isHit);
        if (lane == 0) {
            s_sum[0][wave] = (uint32_t)__popcll(mh);
            s_sum[1][wave] = (uint32_t)__popcll(mm);
        }
        __syncthreads();
        uint32_t preH = 0, preM = 0, totH = 0, totM = 0;
        for (uint32_t w = 0; w < waves; ++w) {
            if (w < wave) {
                preH += s_sum[0][w];
                preM += s_sum[1][w];
            }
            totH += s_sum[0][w];
            totM += s_sum[1][w];
        }
        const uint64_t below = (1ull << lane) - 1ull;
        if (isHit) B.hitQ[baseH + preH + (uint32_t)__popcll(mh & below)] = j;
        else if (active) B.missQ[baseM + preM + (uint32_t)__popcll(mm & below)] = j;
        baseH += totH;
        baseM += totM;
        __syncthreads();  // s_sum is rewritten by the next chunk
    }
}

// Ends the samples whose last vertex had a light sample but no continuation
// (their radiance is complete once the shadow kernel ran).
template <bool BATCH>
__global__ __launch_bounds__(256) void k_wf_shadow_end(WfArgs A, uint32_t in) {
    RScene S = A.S;
    if (!BATCH) finishScene(S);
    const WfBufs& B = A.B;
    const uint32_t chunk = B.counts[kCntChunk];
    const uint32_t out = in ^ 1u;
    const uint32_t n = min(B.counts[kCntShadow], B.slots);
    const uint32_t lanes = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += lanes) {
        const uint32_t i = base + threadIdx.x;
        bool wantRay = false;
        uint32_t slot = 0;
        V3 org = v3z(), dir = v3z();
        if (i < n) {
            slot = B.sqSlot[i];
            if (B.meta[slot] & kFlagEndAfterShadow) {
                uint32_t vLocal, pixBase;
                slotScene<BATCH>(A, chunk + slot, S, vLocal, pixBase);
                const uint4 rs = B.rng[slot];
                Rng rng{rs.x, rs.y, rs.z, rs.w};
                wantRay = endSample(A, S, slot, vLocal, pixBase, rng, xyz(B.pix[slot]), org, dir);
            }
        }
        pushExt(B, out, wantRay, slot, org, dir);
    }
}

struct Launch {
    uint32_t grid, block;
};

// Scene table entries written from kernel arguments, blob headers resolved.
constexpr uint32_t kSceneChunk = 16;
struct SceneChunk {
    RScene s[kSceneChunk];
};
static_assert(sizeof(SceneChunk) + 16 <= 4096, "scene table chunk exceeds the kernel argument space");
__global__ void k_put_scenes(RScene* __restrict__ table, SceneChunk chunk, uint32_t count) {
    const uint32_t i = threadIdx.x;
    if (i >= count) return;
    RScene S = chunk.s[i];
    finishScene(S);
    table[i] = S;
}

}  // namespace

// ---- host side --------------------------------------------------------------

struct WfState {
    clr_wavefront_config cfg{};
    bool tuned = false;
    hipStream_t capture = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    std::vector<uint8_t> key;  // launch arguments the captured graph was made for
    uint32_t* hostCount = nullptr;  // pinned
    int numCUs = 0;
};

namespace {

clr_wavefront_config defaultConfig() {
    clr_wavefront_config c{};
    c.traceBlock = 128;
    c.traceWavesPerCU = 16;
    c.shadeBlock = 128;
    c.shadeWavesPerCU = 16;
    c.itersPerGraph = 8;
    c.checkEvery = 4;
    c.maxSlots = 1u << 22;
    c.useGraph = 1;
    c.refillMin = 16;
    c.ldsDepth = 32;
    c.raysPerLane = 1;
    return c;
}

int validConfig(const clr_wavefront_config& c) {
    auto okBlock = [](uint32_t b) { return b == 64 || b == 128 || b == 256; };
    // Up to 128 waves per CU: beyond what a CU holds (at most 32), the
    // dispatcher starts the remaining waves as resident ones finish.
    return okBlock(c.traceBlock) && okBlock(c.shadeBlock) && c.traceWavesPerCU >= 1 &&
           c.traceWavesPerCU <= 128 && c.shadeWavesPerCU >= 1 && c.shadeWavesPerCU <= 128 &&
           c.itersPerGraph >= 2 && c.itersPerGraph <= 64 && (c.itersPerGraph % 2) == 0 && c.maxSlots >= 64 &&
           c.refillMin >= 1 && c.refillMin <= 64 && (c.ldsDepth == 16 || c.ldsDepth == 32) &&
           (c.raysPerLane == 1 || c.raysPerLane == 2);
}

int ensureWfState(clr_ctx* ctx) {
    if (ctx->wf) return CLR_OK;
    WfState* w = new (std::nothrow) WfState();
    if (!w) return fail(ctx, CLR_ERR_OUT_OF_MEMORY, "wavefront state");
    w->cfg = defaultConfig();
    int rc = hipCheck(ctx, hipDeviceGetAttribute(&w->numCUs, hipDeviceAttributeMultiprocessorCount, ctx->device),
                      "CU count");
    if (!rc) rc = hipCheck(ctx, hipStreamCreateWithFlags(&w->capture, hipStreamNonBlocking), "capture stream");
    if (!rc) rc = hipCheck(ctx, hipHostMalloc(reinterpret_cast<void**>(&w->hostCount), 64, hipHostMallocDefault),
                           "pinned counter");
    ctx->wf = w;
    return rc;
}

Launch traceLaunch(const WfState& w) {
    const uint32_t b = w.cfg.traceBlock;
    const uint32_t g = std::max(1u, (uint32_t)w.numCUs * w.cfg.traceWavesPerCU * 64u / b);
    return {(g + 7u) / 8u * 8u, b};  // whole XCD groups (WavePool)
}
Launch shadeLaunch(const WfState& w) {
    const uint32_t b = w.cfg.shadeBlock;
    return {std::max(1u, (uint32_t)w.numCUs * w.cfg.shadeWavesPerCU * 64u / b), b};
}

template <bool BATCH, int K>
void launchIteration(const WfArgs& A, const Launch& tl, const Launch& sl, size_t lds, bool ld16, bool ordered,
                     uint32_t in, hipStream_t st) {
    if (ordered) {  // K walkers per lane: K stacks
        if (ld16) k_wf_extend<true, 16, BATCH, K><<<tl.grid, tl.block, K * lds, st>>>(A, in);
        else k_wf_extend<true, 32, BATCH, K><<<tl.grid, tl.block, K * lds, st>>>(A, in);
        k_wf_compact<<<sl.grid, sl.block, 0, st>>>(A, in);
    } else if constexpr (!BATCH) {
        if (ld16) k_wf_extend<false, 16, false><<<tl.grid, tl.block, lds, st>>>(A, in);
        else k_wf_extend<false, 32, false><<<tl.grid, tl.block, lds, st>>>(A, in);
    }
    k_wf_shade<BATCH><<<sl.grid, sl.block, 0, st>>>(A, in);
    k_wf_miss<BATCH><<<sl.grid, sl.block, 0, st>>>(A, in);
    if (ordered) {
        if (ld16) k_wf_shadow<true, 16, BATCH, K><<<tl.grid, tl.block, K * lds, st>>>(A, in);
        else k_wf_shadow<true, 32, BATCH, K><<<tl.grid, tl.block, K * lds, st>>>(A, in);
        k_wf_shadow_end<BATCH><<<sl.grid, sl.block, 0, st>>>(A, in);
    } else if constexpr (!BATCH) {
        if (ld16) k_wf_shadow<false, 16, false><<<tl.grid, tl.block, lds, st>>>(A, in);
        else k_wf_shadow<false, 32, false><<<tl.grid, tl.block, lds, st>>>(A, in);
    }
}

void enqueueIterations(const WfArgs& A, const WfState& w, uint32_t iters, hipStream_t st) {
    const Launch tl = traceLaunch(w), sl = shadeLaunch(w);
    const bool ordered = (A.flags & CLR_RENDER_REFERENCE_TRAVERSAL) == 0;
    const bool ld16 = w.cfg.ldsDepth == 16;
    const size_t lds = sizeof(uint32_t) * (ld16 ? 16 : 32) * tl.block;
    const bool two = w.cfg.raysPerLane == 2;
    for (uint32_t it = 0; it < iters; ++it) {
        if (A.table) {
            if (two) launchIteration<true, 2>(A, tl, sl, lds, ld16, ordered, it & 1u, st);
            else launchIteration<true, 1>(A, tl, sl, lds, ld16, ordered, it & 1u, st);
        } else {
            if (two) launchIteration<false, 2>(A, tl, sl, lds, ld16, ordered, it & 1u, st);
            else launchIteration<false, 1>(A, tl, sl, lds, ld16, ordered, it & 1u, st);
        }
    }
}

void dropGraph(WfState& w) {
    if (w.exec) (void)hipGraphExecDestroy(w.exec);
    if (w.graph) (void)hipGraphDestroy(w.graph);
    w.exec = nullptr;
    w.graph = nullptr;
    w.key.clear();
}

// Captures G iterations for these arguments unless the cached graph matches.
int ensureGraph(clr_ctx* ctx, WfState& w, const WfArgs& A) {
    std::vector<uint8_t> key(sizeof(WfArgs) + sizeof(clr_wavefront_config));
    std::memcpy(key.data(), &A, sizeof(WfArgs));
    std::memcpy(key.data() + sizeof(WfArgs), &w.cfg, sizeof(clr_wavefront_config));
    if (w.exec && key == w.key) return CLR_OK;
    dropGraph(w);
    int rc = hipCheck(ctx, hipStreamBeginCapture(w.capture, hipStreamCaptureModeThreadLocal), "begin capture");
    if (rc) return rc;
    enqueueIterations(A, w, w.cfg.itersPerGraph, w.capture);
    hipGraph_t g = nullptr;
    rc = hipCheck(ctx, hipStreamEndCapture(w.capture, &g), "end capture");
    if (rc) return rc;
    w.graph = g;
    rc = hipCheck(ctx, hipGraphInstantiate(&w.exec, g, nullptr, nullptr, 0), "graph instantiate");
    if (rc) {
        dropGraph(w);
        return rc;
    }
    w.key = key;
    return CLR_OK;
}

}  // namespace

// Whether `stream` is being captured into a graph (by the caller).
int streamCapturing(clr_ctx* ctx, hipStream_t stream, bool& capturing) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    const int rc = hipCheck(ctx, hipStreamIsCapturing(stream, &cs), "capture status");
    capturing = cs != hipStreamCaptureStatusNone;
    return rc;
}

void destroyWfState(clr_ctx* ctx) {
    if (!ctx->wf) return;
    WfState* w = ctx->wf;
    dropGraph(*w);
    if (w->capture) (void)hipStreamDestroy(w->capture);
    if (w->hostCount) (void)hipHostFree(w->hostCount);
    delete w;
    ctx->wf = nullptr;
}

int wavefrontGetConfig(clr_ctx* ctx, clr_wavefront_config* cfg) {
    int rc = ensureWfState(ctx);
    if (rc) return rc;
    *cfg = ctx->wf->cfg;
    return CLR_OK;
}

int wavefrontSetConfig(clr_ctx* ctx, const clr_wavefront_config* cfg) {
    if (!validConfig(*cfg)) return fail(ctx, CLR_ERR_INVALID_ARG, "invalid wavefront configuration");
    int rc = ensureWfState(ctx);
    if (rc) return rc;
    ctx->wf->cfg = *cfg;
    ctx->wf->tuned = true;
    return CLR_OK;
}

// Runs the iterations over `numPos` tile-major positions (all scenes) in
// chunks of at most maxSlots path slots. A carries the scene(s) and outputs.
int runChunks(clr_ctx* ctx, WfState& w, WfArgs& A, uint64_t numPos, const clr_render_params& rp,
              hipStream_t stream) {
    bool capturing = false;
    int rc = streamCapturing(ctx, stream, capturing);
    if (rc) return rc;
    const uint32_t P = (uint32_t)std::min<uint64_t>(numPos, w.cfg.maxSlots);
    const Launch tl = traceLaunch(w);
    const uint32_t lanes = tl.grid * tl.block;
    const size_t need = wfBytes(P, lanes);
    if (ctx->wfPoolBytes < need) {
        if (capturing)
            return fail(ctx, CLR_ERR_INVALID_ARG,
                        "the wavefront pool must be sized by an eager render of this size before capture");
        rc = hipCheck(ctx, hipDeviceSynchronize(), "wavefront pool regrow sync");
        if (rc) return rc;
        if (ctx->wfPool) (void)hipFree(ctx->wfPool);
        ctx->wfPool = nullptr;
        ctx->wfPoolBytes = 0;
        dropGraph(w);
        hipError_t e = hipMalloc(&ctx->wfPool, need);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(ctx, CLR_ERR_OUT_OF_MEMORY, "wavefront pool of " + std::to_string(need) + " bytes");
        }
        ctx->wfPoolBytes = need;
    }
    A.B = carveWf(ctx->wfPool, P, lanes);
    A.width = rp.width;
    A.height = rp.height;
    A.maxDepth = rp.maxDepth;
    A.flags = rp.flags & (CLR_RENDER_REFERENCE_TRAVERSAL | CLR_RENDER_NO_RR | CLR_RENDER_LIGHT_SAMPLING_ONLY |
                          CLR_RENDER_BSDF_SAMPLING_ONLY);
    A.refillMin = w.cfg.refillMin;
    // Each sample takes at most maxDepth + 1 iterations (vertices 0..maxDepth).
    const uint64_t bound = (uint64_t)rp.samplesPerPixel * (rp.maxDepth + 1ull) + 1ull;
    const uint32_t G = w.cfg.itersPerGraph;
    // Under the caller's stream capture the iterations are captured as plain
    // launches (no nested graph) and all `bound` of them run: the early stop
    // reads a count back to the host. Iterations with empty queues return at
    // once.
    const bool useGraph = w.cfg.useGraph && !capturing;
    if (useGraph) {
        rc = ensureGraph(ctx, w, A);
        if (rc) return rc;
    }
    for (uint64_t chunk = 0; chunk < numPos; chunk += P) {
        const uint32_t count = (uint32_t)std::min<uint64_t>(P, numPos - chunk);
        k_wf_reset<<<1, 64, 0, stream>>>(A, (uint32_t)chunk);
        if (A.table)
            k_wf_init<true><<<(count + 255) / 256, 256, 0, stream>>>(A, (uint32_t)chunk, count, rp.samplesPerPixel);
        else
            k_wf_init<false><<<(count + 255) / 256, 256, 0, stream>>>(A, (uint32_t)chunk, count, rp.samplesPerPixel);
        rc = hipCheck(ctx, hipGetLastError(), "wavefront init launch");
        if (rc) return rc;
        uint64_t launched = 0, batches = 0;
        while (launched < bound) {
            if (useGraph) {
                rc = hipCheck(ctx, hipGraphLaunch(w.exec, stream), "graph launch");
                if (rc) return rc;
            } else {
                enqueueIterations(A, w, G, stream);
                rc = hipCheck(ctx, hipGetLastError(), "wavefront launch");
                if (rc) return rc;
            }
            launched += G;
            ++batches;
            if (!capturing && w.cfg.checkEvery && batches % w.cfg.checkEvery == 0 && launched < bound) {
                // Early stop: G is even, so the next iteration reads queue 0.
                rc = hipCheck(ctx, hipMemcpyAsync(w.hostCount, A.B.counts + kCntExt0, 4, hipMemcpyDeviceToHost, stream),
                              "count readback");
                if (!rc) rc = hipCheck(ctx, hipStreamSynchronize(stream), "count sync");
                if (rc) return rc;
                if (*w.hostCount == 0) break;
            }
        }
    }
    return CLR_OK;
}

int renderWavefront(clr_ctx* ctx, const clr_scene_dev& sd, const clr_render_params& rp, uint32_t* states,
                    clr_float3* pixels, hipStream_t stream) {
    int rc = ensureWfState(ctx);
    if (rc) return rc;
    WfState& w = *ctx->wf;
    WfArgs A;
    std::memset(&A, 0, sizeof(A));  // padding too: the graph cache compares the bytes
    rc = prepareRenderScene(ctx, sd, A.S, stream);
    if (rc) return rc;
    // Tile-major positions: whole 8x8 tiles, edge-tile positions outside the
    // image are skipped by k_wf_init.
    const uint64_t numPos = (uint64_t)((rp.width + 7) / 8) * ((rp.height + 7) / 8) * 64;
    A.posPerScene = (uint32_t)numPos;
    A.states = states;
    A.pixels = reinterpret_cast<float4*>(pixels);
    return runChunks(ctx, w, A, numPos, rp, stream);
}

// Many scenes with one image size and one set of render parameters in one
// pass: every scene's traversal layout and face-light map go into one
// context buffer, a device table holds the resolved scenes, and path slots
// cover all scenes' tile-major positions (scene-major), so the queues mix
// scenes and stay full.
// A batch's per-scene traversal arrays (built on `stream`) and its device
// scene table, blob headers resolved, in the context's batch buffer (regrown
// only outside a capture). *firstOut: scene 0's host-side entry.
int prepareBatchScenes(clr_ctx* ctx, const clr_scene_dev* scenes, uint32_t numScenes, hipStream_t stream,
                       RScene** tableOut, RScene* firstOut) {
    int rc = ensureWfState(ctx);
    if (rc) return rc;
    rc = ensureStatus(ctx);
    if (rc) return rc;
    WfState& w = *ctx->wf;
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    size_t bytes = align(sizeof(RScene) * numScenes) + align(sizeof(int) * numScenes);
    for (uint32_t s = 0; s < numScenes; ++s) {
        const uint32_t n = scenes[s].numFaces;
        bytes += accelNodeBytes(n, scenes[s].bvhNodes != nullptr) + accelWideBytes(n, scenes[s].bvhNodes != nullptr) +
                 align(48ull * n) + align(4ull * n);
    }
    if (ctx->batchBytes < bytes) {
        bool capturing = false;
        rc = streamCapturing(ctx, stream, capturing);
        if (rc) return rc;
        if (capturing)
            return fail(ctx, CLR_ERR_INVALID_ARG,
                        "the batch buffers must be sized by an eager render of this batch before capture");
        rc = hipCheck(ctx, hipDeviceSynchronize(), "batch buffer regrow sync");
        if (rc) return rc;
        if (ctx->batch) (void)hipFree(ctx->batch);
        ctx->batch = nullptr;
        ctx->batchBytes = 0;
        dropGraph(w);
        hipError_t e = hipMalloc(&ctx->batch, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(ctx, CLR_ERR_OUT_OF_MEMORY, "batch buffers of " + std::to_string(bytes) + " bytes");
        }
        ctx->batchBytes = bytes;
    }
    Carve c(ctx->batch);
    RScene* dTable = c.take<RScene>(numScenes);
    int* envWords = c.take<int>(numScenes);
    std::vector<RScene> table(numScenes);
    // Per-scene arrays grouped by kind: every scene's 4-wide records, then
    // every scene's triangle records, binary records and face-light words.
    // The walks read the first two kinds of all scenes at once. Packed
    // together, away from the binary records that the renderers' walk does
    // not read, they span fewer pages: on the 256-scene batch the traversal
    // kernels' L1 address-translation misses fell from 25 % to 3 % of
    // requests, and the render took a third less time than with each scene's
    // arrays side by side.
    std::vector<float4*> wns(numScenes), ats(numScenes), ans(numScenes);
    std::vector<uint32_t*> fls(numScenes);
    for (uint32_t s = 0; s < numScenes; ++s)
        wns[s] = c.take<float4>(accelWideBytes(scenes[s].numFaces, scenes[s].bvhNodes != nullptr) / sizeof(float4));
    for (uint32_t s = 0; s < numScenes; ++s) ats[s] = c.take<float4>(3ull * scenes[s].numFaces);
    for (uint32_t s = 0; s < numScenes; ++s)
        ans[s] = c.take<float4>(accelNodeBytes(scenes[s].numFaces, scenes[s].bvhNodes != nullptr) / sizeof(float4));
    for (uint32_t s = 0; s < numScenes; ++s) fls[s] = c.take<uint32_t>(scenes[s].numFaces);
    for (uint32_t s = 0; s < numScenes; ++s) {
        const clr_scene_dev& sd = scenes[s];
        const uint32_t n = sd.numFaces;
        float4 *an = ans[s], *wn = wns[s], *at = ats[s];
        uint32_t* fl = fls[s];
        launchAccelBuild(sd.internalNodes, sd.leafNodes, sd.bvhNodes, n, sd.vertices, sd.faces, an, at, stream);
        launchAccelWide(sd.internalNodes, sd.leafNodes, sd.bvhNodes, n, wn, stream);
        launchFaceLights(sd.lights, sd.numLights, n, fl, envWords + s, stream);
        RScene& S = table[s];
        std::memset(&S, 0, sizeof(S));
        S.g = GeomScene{sd.internalNodes, sd.leafNodes, n, sd.vertices, sd.faces, an, at,
                        reinterpret_cast<const float4*>(sd.bvhNodes), wn};
        S.normals = sd.normals;
        S.tangents = sd.tangents;
        S.uvs = sd.uvs;
        S.lights = sd.lights;
        S.mat = sd.materialsData;
        S.tex = sd.texturesData;
        S.oth = sd.otherResources;
        S.faceLight = fl;
        S.envLightPtr = envWords + s;
        S.envLight = -1;
        S.status = ctx->devStatus;
    }
    // The table travels in kernel arguments (no host copy to wait for, and a
    // captured graph keeps the values).
    for (uint32_t s0 = 0; s0 < numScenes; s0 += kSceneChunk) {
        SceneChunk chunk;
        std::memset(&chunk, 0, sizeof(chunk));
        const uint32_t count = std::min<uint32_t>(kSceneChunk, numScenes - s0);
        for (uint32_t k = 0; k < count; ++k) chunk.s[k] = table[s0 + k];
        k_put_scenes<<<1, kSceneChunk, 0, stream>>>(dTable + s0, chunk, count);
    }
    *tableOut = dTable;
    *firstOut = table[0];
    return hipCheck(ctx, hipGetLastError(), "batch set-up launch");
}

int renderWavefrontBatch(clr_ctx* ctx, const clr_scene_dev* scenes, uint32_t numScenes,
                         const clr_render_params& rp, uint32_t* states, clr_float3* pixels, hipStream_t stream) {
    RScene* dTable = nullptr;
    RScene first;
    int rc = prepareBatchScenes(ctx, scenes, numScenes, stream, &dTable, &first);
    if (rc) return rc;
    WfState& w = *ctx->wf;
    WfArgs A;
    std::memset(&A, 0, sizeof(A));
    A.S = first;  // status word for the kernels' flags
    A.table = dTable;
    A.posPerScene = ((rp.width + 7) / 8) * ((rp.height + 7) / 8) * 64;
    A.states = states;
    A.pixels = reinterpret_cast<float4*>(pixels);
    return runChunks(ctx, w, A, (uint64_t)A.posPerScene * numScenes, rp, stream);
}

// Times launch-shape candidates by coordinate descent on renders of the
// caller's scene(s) (numScenes > 1 or batch: clr_render_batch renders) into
// scratch images, and keeps the fastest in the context.
int wavefrontAutotune(clr_ctx* ctx, const clr_scene_dev* scenes, uint32_t numScenes, bool batch,
                      const clr_render_params& rp, clr_wavefront_config* out, hipStream_t stream) {
    int rc = ensureWfState(ctx);
    if (rc) return rc;
    WfState& w = *ctx->wf;
    const uint64_t numPix = (uint64_t)rp.width * rp.height;
    const uint32_t images = batch ? numScenes : 1u;
    uint32_t* states = nullptr;
    clr_float3* pixels = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    rc = hipCheck(ctx, hipMalloc(&states, 16 * numPix * images), "autotune states");
    if (!rc) rc = hipCheck(ctx, hipMalloc(&pixels, 16 * numPix * images), "autotune pixels");
    if (!rc) rc = hipCheck(ctx, hipEventCreate(&e0), "event");
    if (!rc) rc = hipCheck(ctx, hipEventCreate(&e1), "event");
    auto renderAll = [&]() -> int {
        return batch ? renderWavefrontBatch(ctx, scenes, numScenes, rp, states, pixels, stream)
                     : renderWavefront(ctx, scenes[0], rp, states, pixels, stream);
    };
    auto timeOne = [&](const clr_wavefront_config& c, float* ms) -> int {
        w.cfg = c;
        int r = CLR_OK;
        for (uint32_t i = 0; i < images && !r; ++i)
            r = seedPixelStates(ctx, rp.width, rp.height, states + 4 * numPix * i, stream);
        // One untimed render (pool growth and graph capture happen here),
        // then the better of two timed renders.
        if (!r) r = renderAll();
        *ms = 0.0f;
        for (int rep = 0; rep < 2 && !r; ++rep) {
            float t = 0.0f;
            r = hipCheck(ctx, hipEventRecord(e0, stream), "event record");
            if (!r) r = renderAll();
            if (!r) r = hipCheck(ctx, hipEventRecord(e1, stream), "event record");
            if (!r) r = hipCheck(ctx, hipEventSynchronize(e1), "event sync");
            if (!r) r = hipCheck(ctx, hipEventElapsedTime(&t, e0, e1), "event time");
            if (!r && (rep == 0 || t < *ms)) *ms = t;
        }
        return r;
    };
    clr_wavefront_config best = w.tuned ? w.cfg : defaultConfig();
    float bestMs = 0.0f;
    if (!rc) rc = timeOne(best, &bestMs);
    // Coordinate descent over the trace kernels' shape, the shade kernels'
    // shape, then the graph length and the early-stop period.
    auto tryVariants = [&](auto mutate, int count) {
        for (int v = 0; v < count && !rc; ++v) {
            clr_wavefront_config c = best;
            mutate(c, v);
            if (!validConfig(c) || std::memcmp(&c, &best, sizeof c) == 0) continue;
            float ms = 0.0f;
            rc = timeOne(c, &ms);
            if (!rc && ms < bestMs) {
                bestMs = ms;
                best = c;
            }
        }
    };
    const uint32_t blocks[3] = {64, 128, 256}, waves[4] = {8, 12, 16, 24};
    tryVariants([&](clr_wavefront_config& c, int v) { c.traceBlock = blocks[v / 4]; c.traceWavesPerCU = waves[v % 4]; }, 12);
    tryVariants([&](clr_wavefront_config& c, int v) { c.shadeBlock = blocks[v / 4]; c.shadeWavesPerCU = waves[v % 4]; }, 12);
    const uint32_t refills[4] = {8, 16, 32, 64};
    tryVariants([&](clr_wavefront_config& c, int v) { c.refillMin = refills[v]; }, 4);
    tryVariants([&](clr_wavefront_config& c, int v) {
        c.raysPerLane = c.raysPerLane == 1 ? 2u : 1u;
        c.traceWavesPerCU = waves[v];
    }, 4);
    tryVariants([&](clr_wavefront_config& c, int v) {
        c.ldsDepth = v < 4 ? 16u : 32u;
        c.traceWavesPerCU = (v % 4 + 1) * 8u;
    }, 8);
    const uint32_t iters[3] = {4, 8, 16}, checks[3] = {0, 2, 8};
    tryVariants([&](clr_wavefront_config& c, int v) { c.itersPerGraph = iters[v / 3]; c.checkEvery = checks[v % 3]; }, 9);
    tryVariants([&](clr_wavefront_config& c, int v) { c.maxSlots = 1u << (20 + v); }, 3);
    w.cfg = best;
    w.tuned = true;
    if (out) *out = best;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(stream);
    if (states) (void)hipFree(states);
    if (pixels) (void)hipFree(pixels);
    return rc;
}

}  // namespace clr
