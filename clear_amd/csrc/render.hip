// Path tracing on gfx950: the megakernel drop-in for `pathtracing`
// (OpenCL_PathTracer/OpenCL/pathtracer.cl:25-197) and the per-call scene
// preparation it needs.
//
// Schedule: one lane per pixel, each wave an 8x8 pixel tile (the reference's
// 8x8 local size, main.cpp:402) for coherent primary rays; the lane keeps the
// pixel's xorshift128 state in registers for all samples of the call and adds
// the summed radiance to the pixel once (the reference launches one sample
// per pixel per call, 16 x 256 tile launches per render, main.cpp:477-494,
// and round-trips the RNG state through global memory at every draw).
#include <hip/hip_runtime.h>

#include "internal.h"
#include "render_device.h"

namespace clr {
namespace dev {

// Observer of a path's radiance terms; the forward renderer's is empty (and
// compiled away), the adjoint's differentiates them.
struct NoHook {
    __device__ __forceinline__ void emission(V3, uint32_t) {}
    __device__ __forceinline__ void lightSample(V3, uint32_t, uint32_t, const BSDF&, V3, V3, const SurfPt&) {}
    __device__ __forceinline__ void bounce(uint32_t, const BSDF&, V3, V3, const SurfPt&, V3) {}
};

// One camera path for pixel (gx, gy) (pathtracer.cl:48-196). flags: the
// CLR_RENDER_* word (NO_RR disables Russian roulette, the bounce cap still
// applies; the SAMPLING_ONLY bits select one strategy).
template <int WALK, class Stack, class Hook = NoHook>
__device__ V3 tracePath(const RScene& S, uint32_t gx, uint32_t gy, uint32_t maxDepth, uint32_t flags, Rng& rng,
                        const Stack& stack, Hook& hook) {
    const bool noRR = (flags & CLR_RENDER_NO_RR) != 0;
    V3 pix = v3z();
    V3 org, dir;
    float w, dirPDF;
    cameraRay(S, gx, gy, rng, org, dir, w, dirPDF);
    V3 alpha = v3s(1.0f) * w;
    const float initY = lum(alpha);

    SurfPt sp;
    BSDF B;
    EDF E;
    bool traceContinue = true;
    uint32_t sampledType = 0, depth = 0;
    V3 vout = v3z();
    if (intersect<WALK>(S, org, dir, stack, sp)) {
        uint32_t face = sp.face;
        clr_face F = S.g.faces[face];
        vout = -dir;
        if (F.lightPtr != 0xFFFFu) {
            edfArea(S, F.lightPtr, sp, sp.u, sp.v, E);
            const V3 term = alpha * evalLe(E, vout);
            pix += term;
            hook.emission(term, F.lightPtr);
        }
        while (true) {
            bsdfAlloc(S, F.matPtr, sp, B);
            if (hasNonSpecular(B)) {
                const float lu = rng.uniform(), l0 = rng.uniform(), l1 = rng.uniform();
                // With BSDF samples alone the light sample is skipped (its numbers are drawn).
                if (!(flags & CLR_RENDER_BSDF_SAMPLING_ONLY)) {
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
                    const V3 sorg = sp.p + vinL * kRayEps;
                    SurfPt lsp;
                    const bool hit = intersect<WALK>(S, sorg, vinL, stack, lsp);
                    const bool visible = atInf ? !hit : (hit && lsp.face == lp.face);
                    // A zero or infinite light density (an environment sample at a
                    // pole, where sin(theta) = 0) contributes nothing.
                    if (visible && lightPDF > 0.0f && lightPDF < INFINITY) {
                        const V3 vinLrev = -vinL;
                        const float fsPDF = evalFsPdf(B, vout, vinL) * absCosEDF(E, vinL) / dist2;
                        const float mis = lightSampleMis(flags, lightPDF, fsPDF);
                        const float wgt = mis * absCosEDF(E, vinL) * fabsf(dot(B.N, vinL)) / dist2 / lightPDF;
                        const V3 term = wgt * ((alpha * evalLe(E, vinLrev)) * evalFs(B, vout, vinL));
                        pix += term;
                        hook.lightSample(term, atInf ? 0xFFFFFFFFu : (uint32_t)S.g.faces[lp.face].lightPtr, F.matPtr, B,
                                         vout, vinL, sp);
                    }
                }
            }
            traceContinue = true;
            const float bc = rng.uniform(), b0 = rng.uniform(), b1 = rng.uniform();
            V3 nd = v3z();
            const V3 f = sampleFs(B, vout, bc, b0, b1, nd, dirPDF, sampledType);
            if ((f.x == 0.0f && f.y == 0.0f && f.z == 0.0f) || dirPDF == 0.0f) {
                traceContinue = false;
                break;
            }
            hook.bounce(F.matPtr, B, vout, nd, sp, pix);
            alpha = alpha * (f * (fabsf(dot(B.N, nd)) / dirPDF));
            dir = nd;
            org = sp.p + dir * kRayEps;
            ++depth;
            if (!intersect<WALK>(S, org, dir, stack, sp)) break;
            face = sp.face;
            F = S.g.faces[face];
            vout = -dir;
            if (F.lightPtr != 0xFFFFu) {
                edfArea(S, F.lightPtr, sp, sp.u, sp.v, E);
                float mis = 1.0f;
                if (sampledType & kNonSingular) {
                    const V3 dd = sp.p - org;
                    const float d2 = dd.x * dd.x + dd.y * dd.y + dd.z * dd.z;
                    const float lightPDF = areaPdfOfFace(S, face) * d2 / absCosEDF(E, vout);
                    mis = emissionHitMis(flags, dirPDF, lightPDF);
                }
                const V3 term = (mis * alpha) * evalLe(E, vout);
                pix += term;
                hook.emission(term, F.lightPtr);
            }
            if (depth >= maxDepth) {
                traceContinue = false;
                break;
            }
            if (noRR) continue;
            const float cp = fminf(lum(alpha) * (1.0f / initY), 1.0f);
            if (rng.uniform() < cp) {
                alpha = alpha * (1.0f / cp);
                dirPDF *= cp;
            } else {
                traceContinue = false;
                break;
            }
        }
    }
    if (traceContinue) {
        vout = -dir;
        const float theta = acosf(clampf(dir.y, -1.0f, 1.0f));
        const float phi = fmodf(atan2f(-dir.x, dir.z) + 2.0f * kPiF, 2 * kPiF);
        const float u = phi / (2.0f * kPiF), v = theta / kPiF;
        edfEnv(S, u, v, E);
        float mis = 1.0f;
        if ((sampledType & kNonSingular) != 0 && depth > 0) {
            const float lightPDF = areaPdfEnv(S, u, v) * 1.0f / absCosEDF(E, vout);
            mis = emissionHitMis(flags, dirPDF, lightPDF);
        }
        pix += (mis * alpha) * evalLe(E, vout);
    }
    return pix;
}

// All spp camera paths of pixel (gx, gy), summed: the same terms, in the same
// order and from the same random numbers as spp calls of tracePath, so the
// result is bit-identical. The two loops of those calls (samples, and bounces
// within a path) are merged into one: a lane whose path ends starts its next
// sample at once, and every pass of the loop traces exactly one ray per active
// lane (camera, shadow or extension) at one call site. A wave then runs as
// long as its lanes' longest sum of paths rather than the sum of its longest
// paths, and its traversals keep their lanes busy.
template <int WALK, class Stack>
__device__ V3 tracePixel(const RScene& S, uint32_t gx, uint32_t gy, uint32_t spp, uint32_t maxDepth,
                         uint32_t flags, Rng& rng, const Stack& stack) {
    const bool noRR = (flags & CLR_RENDER_NO_RR) != 0;
    V3 acc = v3z();
    if (spp == 0) return acc;
    uint32_t k = 0;
    // path state
    V3 pix, alpha, org, dir, vout = v3z();
    float initY, dirPDF;
    uint32_t sampledType, depth;
    bool first, shadow;
    SurfPt sp;
    BSDF B;
    EDF E;
    clr_face F;
    // pending light sample: while `shadow`, org and dir hold the shadow ray
    // (dir = vinL); the extension ray's are not needed again after it
    SurfPt lp;
    bool atInf = false;
    float lightPDF = 0.0f, dist2 = 0.0f;

    auto startSample = [&]() {
        float w;
        cameraRay(S, gx, gy, rng, org, dir, w, dirPDF);
        alpha = v3s(1.0f) * w;
        initY = lum(alpha);
        pix = v3z();
        sampledType = 0;
        depth = 0;
        first = true;
        shadow = false;
    };
    // The path ends; envTerm: it left the scene (tracePath's traceContinue).
    auto finishSample = [&](bool envTerm) {
        if (envTerm) {
            vout = -dir;
            const float theta = acosf(clampf(dir.y, -1.0f, 1.0f));
            const float phi = fmodf(atan2f(-dir.x, dir.z) + 2.0f * kPiF, 2 * kPiF);
            const float u = phi / (2.0f * kPiF), v = theta / kPiF;
            edfEnv(S, u, v, E);
            float mis = 1.0f;
            if ((sampledType & kNonSingular) != 0 && depth > 0) {
                const float lightPDF = areaPdfEnv(S, u, v) * 1.0f / absCosEDF(E, vout);
                mis = emissionHitMis(flags, dirPDF, lightPDF);
            }
            pix += (mis * alpha) * evalLe(E, vout);
        }
        acc += pix;
        if (++k < spp) startSample();
    };
    // BSDF sample at sp: the next extension ray, or the end of the path.
    auto sampleBounce = [&]() {
        const float bc = rng.uniform(), b0 = rng.uniform(), b1 = rng.uniform();
        V3 nd = v3z();
        const V3 f = sampleFs(B, vout, bc, b0, b1, nd, dirPDF, sampledType);
        if ((f.x == 0.0f && f.y == 0.0f && f.z == 0.0f) || dirPDF == 0.0f) {
            finishSample(false);
            return;
        }
        alpha = alpha * (f * (fabsf(dot(B.N, nd)) / dirPDF));
        dir = nd;
        org = sp.p + dir * kRayEps;
        ++depth;
        shadow = false;
    };

    startSample();
    while (k < spp) {
        SurfPt hsp;
        const bool hit = intersect<WALK>(S, org, dir, stack, hsp);
        if (shadow) {
            const V3 vinL = dir;
            const bool visible = atInf ? !hit : (hit && hsp.face == lp.face);
            // A zero or infinite light density (an environment sample at a
            // pole, where sin(theta) = 0) contributes nothing.
            if (visible && lightPDF > 0.0f && lightPDF < INFINITY) {
                const V3 vinLrev = -vinL;
                const float fsPDF = evalFsPdf(B, vout, vinL) * absCosEDF(E, vinL) / dist2;
                const float mis = lightSampleMis(flags, lightPDF, fsPDF);
                const float wgt = mis * absCosEDF(E, vinL) * fabsf(dot(B.N, vinL)) / dist2 / lightPDF;
                const V3 term = wgt * ((alpha * evalLe(E, vinLrev)) * evalFs(B, vout, vinL));
                pix += term;
            }
            sampleBounce();
            continue;
        }
        if (!hit) {
            finishSample(true);
            continue;
        }
        sp = hsp;
        F = S.g.faces[sp.face];
        vout = -dir;
        if (F.lightPtr != 0xFFFFu) {
            edfArea(S, F.lightPtr, sp, sp.u, sp.v, E);
            if (first) {
                pix += alpha * evalLe(E, vout);
            } else {
                float mis = 1.0f;
                if (sampledType & kNonSingular) {
                    const V3 dd = sp.p - org;
                    const float d2 = dd.x * dd.x + dd.y * dd.y + dd.z * dd.z;
                    const float lightPDF = areaPdfOfFace(S, sp.face) * d2 / absCosEDF(E, vout);
                    mis = emissionHitMis(flags, dirPDF, lightPDF);
                }
                pix += (mis * alpha) * evalLe(E, vout);
            }
        }
        if (!first) {
            if (depth >= maxDepth) {
                finishSample(false);
                continue;
            }
            if (!noRR) {
                const float cp = fminf(lum(alpha) * (1.0f / initY), 1.0f);
                if (rng.uniform() < cp) {
                    alpha = alpha * (1.0f / cp);
                    dirPDF *= cp;
                } else {
                    finishSample(false);
                    continue;
                }
            }
        }
        first = false;
        bsdfAlloc(S, F.matPtr, sp, B);
        if (hasNonSpecular(B)) {
            const float lu = rng.uniform(), l0 = rng.uniform(), l1 = rng.uniform();
            // With BSDF samples alone the light sample is skipped (its numbers are drawn).
            if (!(flags & CLR_RENDER_BSDF_SAMPLING_ONLY)) {
                sampleLight(S, lu, l0, l1, lp, atInf, E, lightPDF);
                V3 vinL = lp.p - sp.p;
                dist2 = dot(vinL, vinL);
                vinL = vinL * (1.0f / sqrtf(dist2));
                if (atInf) {
                    vinL = lp.p;
                    dist2 = 1.0f;
                }
                org = sp.p + vinL * kRayEps;
                dir = vinL;
                shadow = true;
                continue;
            }
        }
        sampleBounce();
    }
    return acc;
}

// ---- path-replay adjoint ----------------------------------------------------
// Parameters: per differentiable material (a single Oren-Nayar / Lambert
// element, reflection.cl:239-255) its albedo rho (RGB) and roughness sigma,
// and the emission M (RGB) of one emitter (Le = M / pi, light.cl:382-393).
// Sampling, light selection and MIS weights do not depend on them (cosine
// sampling, power-proportional light choice), and Russian roulette
// probabilities are held fixed (replayed), so with the random states fixed
// the estimate is a sum of path terms, each a product of BSDF values, one
// emission and parameter-free factors. Per term:
//   d/d rho_c   = term_c / rho_c    for every vertex of that material on it,
//   d/d sigma   = term . adj * dlog f / d sigma at that vertex,
//   d/d M_c     = term_c / M_c      when it ends on the emitter.
// A replay records, at each bounce from a differentiable vertex, the radiance
// accumulated so far; at the end every later term carries that vertex's
// sampled BSDF value, so their sum is L - prefix (path replay
// backpropagation, Vicini et al. 2021). Light-sample terms are differentiated
// where they occur, with the BSDF value of the light direction.
constexpr int kMaxAdjVertices = 16;
constexpr uint32_t kMaxDiffMaterials = 256;

struct AdjAcc {
    float alb[kMaxDiffMaterials * 3];
    float sig[kMaxDiffMaterials];
    float em[3];
};

// Oren-Nayar angular factor X of f = rho / pi * (A + B X) (same terms as fx).
__device__ __forceinline__ float orenNayarX(V3 wo, V3 wi) {
    const float sti = sinTh(wi), sto = sinTh(wo);
    float maxCos = 0.0f;
    if (sti > 1e-4f && sto > 1e-4f) maxCos = fmaxf(0.0f, cosPhi(wi) * cosPhi(wo) + sinPhi(wi) * sinPhi(wo));
    float sinA, tanB;
    if (fabsf(wi.z) > fabsf(wo.z)) {
        sinA = sto;
        tanB = sti / fabsf(wi.z);
    } else {
        sinA = sti;
        tanB = sto / fabsf(wo.z);
    }
    return maxCos * sinA * tanB;
}

struct AdjHook {
    const RScene& S;
    const uint16_t* table;  // matPtr -> parameter index + 1 (0: not differentiable)
    uint32_t lightProp;     // lightPtr of the differentiable emitter
    V3 M, adj;              // its emission; dLoss/dpixel / spp
    AdjAcc* acc;            // block accumulators (LDS)
    int n;
    uint32_t recM[kMaxAdjVertices];
    V3 recPrefix[kMaxAdjVertices], recRho[kMaxAdjVertices];
    float recDlog[kMaxAdjVertices];

    __device__ __forceinline__ uint32_t param(uint32_t matPtr, const BSDF& B) const {
        if (B.n != 1 || B.c[0].id != 0u) return 0u;
        return table[matPtr & 0xFFFFu];
    }
    __device__ __forceinline__ float dlogSigma(uint32_t matPtr, const BSDF& B, const SurfPt& sp, V3 vout,
                                               V3 vin) const {
        const BxDF& f = B.c[0];
        if (f.a == 1.0f) return 0.0f;  // sigma = 0: f = rho / pi
        const uint8_t* e = S.mat + ((matPtr + 8u + 3u) & ~3u);
        const float sigma = evalFloat(S.tex + ld<uint32_t>(e + 8), sp.u, sp.v);
        const V3 wo = toLocal(B.S, B.T, B.N, vout), wi = toLocal(B.S, B.T, B.N, vin);
        const float X = orenNayarX(wo, wi);
        const float s2 = sigma * sigma, qa = s2 + 0.33f, qb = s2 + 0.09f;
        const float dA = -0.33f * sigma / (qa * qa), dB = 0.081f * sigma / (qb * qb);
        return (dA + dB * X) / (f.a + f.b * X);
    }
    __device__ __forceinline__ void addAlbedo(uint32_t m, V3 term, V3 rho) const {
        if (rho.x != 0.0f) atomicAdd(acc->alb + 3 * m + 0, adj.x * term.x / rho.x);
        if (rho.y != 0.0f) atomicAdd(acc->alb + 3 * m + 1, adj.y * term.y / rho.y);
        if (rho.z != 0.0f) atomicAdd(acc->alb + 3 * m + 2, adj.z * term.z / rho.z);
    }
    __device__ __forceinline__ void addEmission(V3 term) const {
        if (M.x != 0.0f) atomicAdd(acc->em + 0, adj.x * term.x / M.x);
        if (M.y != 0.0f) atomicAdd(acc->em + 1, adj.y * term.y / M.y);
        if (M.z != 0.0f) atomicAdd(acc->em + 2, adj.z * term.z / M.z);
    }
    __device__ __forceinline__ void emission(V3 term, uint32_t lightPtr) {
        if (lightPtr == lightProp) addEmission(term);
    }
    __device__ __forceinline__ void lightSample(V3 term, uint32_t lightPtr, uint32_t matPtr, const BSDF& B, V3 vout,
                                                V3 vinL, const SurfPt& sp) {
        if (lightPtr == lightProp) addEmission(term);
        const uint32_t m = param(matPtr, B);
        if (m) {
            addAlbedo(m - 1, term, B.c[0].R);
            atomicAdd(acc->sig + (m - 1), dot(adj, term) * dlogSigma(matPtr, B, sp, vout, vinL));
        }
    }
    __device__ __forceinline__ void bounce(uint32_t matPtr, const BSDF& B, V3 vout, V3 vin, const SurfPt& sp,
                                           V3 pixSoFar) {
        const uint32_t m = param(matPtr, B);
        if (!m || n >= kMaxAdjVertices) return;
        recM[n] = m - 1;
        recPrefix[n] = pixSoFar;
        recRho[n] = B.c[0].R;
        recDlog[n] = dlogSigma(matPtr, B, sp, vout, vin);
        ++n;
    }
    __device__ __forceinline__ void finish(V3 L) {
        for (int k = 0; k < n; ++k) {
            const V3 after = L - recPrefix[k];
            addAlbedo(recM[k], after, recRho[k]);
            atomicAdd(acc->sig + recM[k], dot(adj, after) * recDlog[k]);
        }
        n = 0;
    }
};

}  // namespace dev

namespace {

constexpr int kTile = 8;  // 8x8 pixels per wave

// Occupancy of the ordered megakernel, which waits on memory. The ordered
// walk's stack keeps its first kMegaLds words per lane in LDS and the rest in
// a per-lane global region, so LDS holds 20 waves per CU (8 KiB each) where
// the whole stack in LDS (16 KiB) held 10, and the register budget is set for
// five waves per SIMD (96 VGPRs, a few spilled). 16 spp of the 1080p outdoor
// scene: 134.6 ms at 10 waves per CU, 109.9 at 12, 92.9 at 16, 89.5 at 20.
// Only deep walks touch the spill. The reference walks keep their 64-entry
// LDS stack (10 waves per CU).
constexpr int kMegaLds = 32;
constexpr uint32_t kMegaWavesPerCU = 20;  // persistent grid
// The merged sample/bounce loop keeps more state alive across its one trace
// call: at four waves per SIMD (128 VGPRs) it beat five (96 VGPRs, more
// spills): 1080p 341-345 -> 330-334 ms, the 256-scene batch 289-294 -> 284-286.
// Three (168 VGPRs) took 378-381 and 325-326 ms.
constexpr uint32_t kMegaMergedWavesPerCU = 16;
// The adjoint's waves also hold 4 KiB of LDS gradient sums, so its stack
// keeps 16 words in LDS (8 KiB per wave with the sums) and its register
// budget is set for four waves per SIMD: 512x512, 32 spp, depth 5 of the
// differentiation scene, 99 ms at 8 waves per CU (20 KiB of LDS each),
// 100 ms at 13 (a 32-word stack), 68.6 ms at 16.
constexpr uint32_t kAdjWavesPerCU = 16;
constexpr int kAdjLds = 16;

__global__ void k_mega_reset(uint32_t* counter) {
    if (threadIdx.x == 0) *counter = 0u;
}

// One 8x8 tile (tile index `tile` of an image tilesX tiles wide), all spp
// samples of its pixels; states and pixels are the image's own arrays.
template <int WALK, int LDS, bool MERGED>
__device__ __forceinline__ void renderTile(const dev::RScene& S, uint32_t tile, uint32_t tilesX, uint32_t width,
                                           uint32_t height, uint32_t spp, uint32_t maxDepth, uint32_t flags,
                                           uint32_t* __restrict__ states, clr_float3* __restrict__ pixels,
                                           uint32_t* __restrict__ spill, uint32_t* s_stack) {
    const uint32_t lanes = gridDim.x * 64u;
    const uint32_t lane = blockIdx.x * 64u + threadIdx.x;
    const uint32_t gx = (tile % tilesX) * kTile + (threadIdx.x % kTile);
    const uint32_t gy = (tile / tilesX) * kTile + (threadIdx.x / kTile);
    if (gx >= width || gy >= height) return;
    const size_t pi = (size_t)gy * width + gx;
    uint4 st = reinterpret_cast<uint4*>(states)[pi];
    dev::Rng rng{st.x, st.y, st.z, st.w};
    dev::V3 acc = dev::v3z();
    dev::NoHook hook;
    // MERGED: the sample and bounce loops merged (tracePixel).
    if constexpr (LDS < dev::kStackDepth) {
        const dev::SpillStack<LDS> stk{s_stack + threadIdx.x, 64, spill + lane, lanes};
        if constexpr (MERGED)
            acc = dev::tracePixel<WALK>(S, gx, gy, spp, maxDepth, flags, rng, stk);
        else
            for (uint32_t k = 0; k < spp; ++k) acc += dev::tracePath<WALK>(S, gx, gy, maxDepth, flags, rng, stk, hook);
    } else {
        const dev::LdsStack stk{s_stack + threadIdx.x, 64};
        if constexpr (MERGED)
            acc = dev::tracePixel<WALK>(S, gx, gy, spp, maxDepth, flags, rng, stk);
        else
            for (uint32_t k = 0; k < spp; ++k) acc += dev::tracePath<WALK>(S, gx, gy, maxDepth, flags, rng, stk, hook);
    }
    reinterpret_cast<uint4*>(states)[pi] = make_uint4(rng.s0, rng.s1, rng.s2, rng.s3);
    float4 p = reinterpret_cast<float4*>(pixels)[pi];
    p.x += acc.x;
    p.y += acc.y;
    p.z += acc.z;
    reinterpret_cast<float4*>(pixels)[pi] = p;
}

// The next tile of a persistent wave: lane 0 takes it from the counter.
__device__ __forceinline__ uint32_t nextTile(uint32_t* tileCounter) {
    uint32_t t = 0;
    if (threadIdx.x == 0) t = atomicAdd(tileCounter, 1u);
    return __builtin_amdgcn_readfirstlane(__shfl((int)t, 0));
}

// Persistent waves: each takes the next 8x8 tile from a device counter
// (zeroed by k_mega_reset on the same stream) until the image is done, so
// the spill region is sized by the grid, not by the image.
// LDS: stack words per lane kept in LDS (the ordered walk spills the rest;
// the reference walks need all kStackDepth).
template <int WALK, int LDS, bool MERGED>
__device__ __forceinline__ void renderTiles(const dev::RScene& scene, uint32_t width, uint32_t height, uint32_t spp,
                                            uint32_t maxDepth, uint32_t flags, uint32_t* __restrict__ states,
                                            clr_float3* __restrict__ pixels, uint32_t* __restrict__ tileCounter,
                                            uint32_t* __restrict__ spill, uint32_t* s_stack) {
    dev::RScene S = scene;  // resolved copy; the kernel argument itself stays read-only
    dev::finishScene(S);
    const uint32_t tilesX = (width + kTile - 1) / kTile;
    const uint32_t tiles = tilesX * ((height + kTile - 1) / kTile);
    for (uint32_t tile = nextTile(tileCounter); tile < tiles; tile = nextTile(tileCounter))
        renderTile<WALK, LDS, MERGED>(S, tile, tilesX, width, height, spp, maxDepth, flags, states, pixels, spill,
                                      s_stack);
}

// A batch of equally sized images: the tiles of all scenes share one
// counter, tile t being tile t % tilesPerScene of scene t / tilesPerScene,
// whose states and pixels start at scene * width * height. Consecutive tiles
// stay in one scene, so the waves resident at a time walk a few scenes'
// arrays. The table's entries have their blob headers resolved.
template <int WALK, int LDS, bool MERGED>
__device__ __forceinline__ void renderBatchTiles(const dev::RScene* __restrict__ table, uint32_t numScenes,
                                                 uint32_t width, uint32_t height, uint32_t spp, uint32_t maxDepth,
                                                 uint32_t flags, uint32_t* __restrict__ states,
                                                 clr_float3* __restrict__ pixels, uint32_t* __restrict__ tileCounter,
                                                 uint32_t* __restrict__ spill, uint32_t* s_stack) {
    const uint32_t tilesX = (width + kTile - 1) / kTile;
    const uint32_t tilesPerScene = tilesX * ((height + kTile - 1) / kTile);
    const uint32_t tiles = tilesPerScene * numScenes;
    const size_t npix = (size_t)width * height;
    for (uint32_t tile = nextTile(tileCounter); tile < tiles; tile = nextTile(tileCounter)) {
        const uint32_t s = tile / tilesPerScene;
        const dev::RScene S = table[s];
        renderTile<WALK, LDS, MERGED>(S, tile - s * tilesPerScene, tilesX, width, height, spp, maxDepth, flags,
                                      states + 4 * npix * s, pixels + npix * s, spill, s_stack);
    }
}

// Large images: five waves per SIMD (four with MERGED). MERGED: the ordered walk's merged
// sample and bounce loop (tracePixel), used when Russian roulette is on.
// Without it paths run to the bounce cap unless they leave the scene, little
// imbalance is left to recover, and the per-sample loop needs fewer registers
// (the 512x512 differentiation scene without roulette: 56 ms per-sample
// against 61 ms merged).
template <int WALK, bool MERGED>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MERGED ? 4 : 5))) void k_render_mega(
    dev::RScene scene, uint32_t width, uint32_t height, uint32_t spp, uint32_t maxDepth, uint32_t flags,
    uint32_t* __restrict__ states, clr_float3* __restrict__ pixels, uint32_t* __restrict__ tileCounter,
    uint32_t* __restrict__ spill) {
    constexpr int kLds = WALK == dev::kWalkOrdered ? kMegaLds : dev::kStackDepth;
    __shared__ uint32_t s_stack[kLds * 64];
    renderTiles<WALK, kLds, MERGED>(scene, width, height, spp, maxDepth, flags, states, pixels, tileCounter, spill,
                                    s_stack);
}

// Images of at most kMegaFewTiles tiles per CU cannot fill many waves per
// CU, so their waves keep the whole stack in LDS and all the registers they
// want (no spills).
constexpr uint32_t kMegaFewTiles = 4;
template <int WALK, bool MERGED>
__global__ __launch_bounds__(64) void k_render_mega_few(dev::RScene scene, uint32_t width, uint32_t height,
                                                        uint32_t spp, uint32_t maxDepth, uint32_t flags,
                                                        uint32_t* __restrict__ states, clr_float3* __restrict__ pixels,
                                                        uint32_t* __restrict__ tileCounter,
                                                        uint32_t* __restrict__ spill) {
    __shared__ uint32_t s_stack[dev::kStackDepth * 64];
    renderTiles<WALK, dev::kStackDepth, MERGED>(scene, width, height, spp, maxDepth, flags, states, pixels,
                                                tileCounter, spill, s_stack);
}

// Batches: the same two shapes.
template <bool MERGED>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MERGED ? 4 : 5))) void k_render_mega_batch(
    const dev::RScene* __restrict__ table, uint32_t numScenes, uint32_t width, uint32_t height, uint32_t spp,
    uint32_t maxDepth, uint32_t flags, uint32_t* __restrict__ states, clr_float3* __restrict__ pixels,
    uint32_t* __restrict__ tileCounter, uint32_t* __restrict__ spill) {
    __shared__ uint32_t s_stack[kMegaLds * 64];
    renderBatchTiles<dev::kWalkOrdered, kMegaLds, MERGED>(table, numScenes, width, height, spp, maxDepth, flags,
                                                          states, pixels, tileCounter, spill, s_stack);
}
template <bool MERGED>
__global__ __launch_bounds__(64) void k_render_mega_batch_few(
    const dev::RScene* __restrict__ table, uint32_t numScenes, uint32_t width, uint32_t height, uint32_t spp,
    uint32_t maxDepth, uint32_t flags, uint32_t* __restrict__ states, clr_float3* __restrict__ pixels,
    uint32_t* __restrict__ tileCounter, uint32_t* __restrict__ spill) {
    __shared__ uint32_t s_stack[dev::kStackDepth * 64];
    renderBatchTiles<dev::kWalkOrdered, dev::kStackDepth, MERGED>(table, numScenes, width, height, spp, maxDepth,
                                                                  flags, states, pixels, tileCounter, spill, s_stack);
}

// bvh_visualizer (bvh_visualizer.cl:25-84): one ray through each pixel
// centre from the lens centre; pixel += the base colour of the hit surface
// (sample_We's weight is 1), then += 0.025 * the ray's box hits. Box hits are
// the reference walk's count (rayIntersectionVis, bvh_traversal.cl:216-319:
// internal and leaf boxes passed for the LBVH, node boxes passed for the
// midpoint BVH), or for the ordered traversal its entered nodes plus
// triangle tests.
template <int WALK>
__global__ __launch_bounds__(64) void k_bvh_visualizer(dev::RScene scene, uint32_t width, uint32_t height,
                                                       clr_float3* __restrict__ pixels) {
    __shared__ uint32_t s_stack[dev::kStackDepth * 64];
    dev::RScene S = scene;
    dev::finishScene(S);
    const uint32_t tilesX = (width + kTile - 1) / kTile;
    const uint32_t gx = (blockIdx.x % tilesX) * kTile + (threadIdx.x % kTile);
    const uint32_t gy = (blockIdx.x / tilesX) * kTile + (threadIdx.x / kTile);
    if (gx >= width || gy >= height) return;
    dev::V3 org, dir;
    dev::cameraRayAt(S, 0.5f, 0.5f, gx + 0.5f, gy + 0.5f, org, dir);
    float t = INFINITY, b1 = 0.0f, b2 = 0.0f;
    dev::TraceStats st{0u, 0u};
    bool over = false;
    const dev::LdsStack stk{s_stack + threadIdx.x, 64};
    uint32_t face;
    if constexpr (WALK == dev::kWalkOrdered)
        face = dev::orderedHit<true, false, dev::TexAlpha>(S.g, org, dir, t, stk, st, &b1, &b2, dev::TexAlpha{&S},
                                                          &over);
    else if constexpr (WALK == dev::kWalkMidpoint)
        face = dev::bvhWalk<true, false>(S.g, org, dir, t, stk, st, &b1, &b2, dev::TexAlpha{&S}, &over);
    else
        face = dev::lbvhClosest<true, dev::TexAlpha>(S.g, org, dir, t, stk, st, &b1, &b2, dev::TexAlpha{&S}, &over);
    if (over) atomicOr(S.status, kStatusStackOverflow);
    const size_t pi = (size_t)gy * width + gx;
    float4 p = reinterpret_cast<float4*>(pixels)[pi];
    if (face != dev::kNone && !isinf(t)) {
        dev::SurfPt sp;
        dev::hitParams(S, org, dir, t, face, b1, b2, sp);
        dev::BSDF B;
        dev::bsdfAlloc(S, S.g.faces[face].matPtr, sp, B);
        const dev::V3 c = dev::baseColor(B);
        p.x += c.x;
        p.y += c.y;
        p.z += c.z;
    }
    const float k = (float)(st.visits + st.tests) * 0.025f;
    p.x += k;
    p.y += k;
    p.z += k;
    reinterpret_cast<float4*>(pixels)[pi] = p;
}

// Adjoint of k_render_mega: the same pixels, samples and random states, each
// sample replayed with the adjoint hook. Persistent waves take tiles from the
// same kind of counter, with a split stack, and sum their gradients in LDS
// over all their tiles; the sums go to global memory once per wave.
template <int WALK>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_render_adjoint(
    dev::RScene scene, uint32_t width, uint32_t height, uint32_t spp, uint32_t maxDepth, uint32_t flags,
    const uint32_t* __restrict__ states, const float4* __restrict__ dLdI, const uint16_t* __restrict__ table,
    uint32_t numMat, uint32_t lightProp, float* __restrict__ gAlb, float* __restrict__ gSig, float* __restrict__ gEm,
    uint32_t* __restrict__ tileCounter, uint32_t* __restrict__ spill) {
    constexpr int kLds = WALK == dev::kWalkOrdered ? kAdjLds : dev::kStackDepth;
    __shared__ uint32_t s_stack[kLds * 64];
    __shared__ dev::AdjAcc s_acc;
    for (uint32_t i = threadIdx.x; i < dev::kMaxDiffMaterials * 3; i += 64) s_acc.alb[i] = 0.0f;
    for (uint32_t i = threadIdx.x; i < dev::kMaxDiffMaterials; i += 64) s_acc.sig[i] = 0.0f;
    if (threadIdx.x < 3) s_acc.em[threadIdx.x] = 0.0f;
    __syncthreads();
    dev::RScene S = scene;
    dev::finishScene(S);
    dev::V3 M = dev::v3z();
    if (lightProp != 0xFFFFFFFFu) {
        const uint8_t* e = S.mat + ((lightProp + 1u + 3u) & ~3u);
        M = dev::evalColor(S.tex + dev::ld<uint32_t>(e + 4), 0.0f, 0.0f);
    }
    const uint32_t tilesX = (width + kTile - 1) / kTile;
    const uint32_t tiles = tilesX * ((height + kTile - 1) / kTile);
    const uint32_t lanes = gridDim.x * 64u;
    const uint32_t lane = blockIdx.x * 64u + threadIdx.x;
    auto nextTile = [&]() {
        uint32_t t = 0;
        if (threadIdx.x == 0) t = atomicAdd(tileCounter, 1u);
        return (uint32_t)__shfl((int)t, 0);
    };
    const float inv = 1.0f / (float)spp;
    for (uint32_t tile = nextTile(); tile < tiles; tile = nextTile()) {
        const uint32_t gx = (tile % tilesX) * kTile + (threadIdx.x % kTile);
        const uint32_t gy = (tile / tilesX) * kTile + (threadIdx.x / kTile);
        if (gx >= width || gy >= height) continue;
        const size_t pi = (size_t)gy * width + gx;
        const uint4 st = reinterpret_cast<const uint4*>(states)[pi];
        dev::Rng rng{st.x, st.y, st.z, st.w};
        const float4 a = dLdI[pi];
        dev::AdjHook hook{S, table, lightProp, M, {a.x * inv, a.y * inv, a.z * inv}, &s_acc, 0};
        if constexpr (WALK == dev::kWalkOrdered) {
            const dev::SpillStack<kLds> stk{s_stack + threadIdx.x, 64, spill + lane, lanes};
            for (uint32_t k = 0; k < spp; ++k) hook.finish(dev::tracePath<WALK>(S, gx, gy, maxDepth, flags, rng, stk, hook));
        } else {
            const dev::LdsStack stk{s_stack + threadIdx.x, 64};
            for (uint32_t k = 0; k < spp; ++k) hook.finish(dev::tracePath<WALK>(S, gx, gy, maxDepth, flags, rng, stk, hook));
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < numMat * 3; i += 64)
        if (s_acc.alb[i] != 0.0f) atomicAdd(gAlb + i, s_acc.alb[i]);
    for (uint32_t i = threadIdx.x; i < numMat; i += 64)
        if (s_acc.sig[i] != 0.0f) atomicAdd(gSig + i, s_acc.sig[i]);
    if (threadIdx.x < 3 && gEm && s_acc.em[threadIdx.x] != 0.0f) atomicAdd(gEm + threadIdx.x, s_acc.em[threadIdx.x]);
}

// Clears the matPtr -> parameter table and the gradient outputs, then (second
// kernel) enters the differentiable materials.
__global__ void k_adj_clear(uint16_t* table, float* gAlb, float* gSig, float* gEm, uint32_t numMat) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 65536u) table[i] = 0;
    if (i < numMat * 3) gAlb[i] = 0.0f;
    if (i < numMat) gSig[i] = 0.0f;
    if (i < 3 && gEm) gEm[i] = 0.0f;
}
__global__ void k_adj_fill(uint16_t* table, const uint32_t* offsets, uint32_t numMat) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < numMat) table[offsets[i]] = (uint16_t)(i + 1);
}

// faceLight[f] = first light index whose reference is f; envLight = first
// light at infinity (the linear searches of getAreaPDF, light.cl:274-289).
__global__ void k_face_lights_clear(uint32_t* faceLight, uint32_t numFaces, int* envLight) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < numFaces) faceLight[i] = 0xFFFFFFFFu;
    if (i == 0) *envLight = 0x7FFFFFFF;
}
__global__ void k_face_lights_fill(const clr_light_info* lights, uint32_t numLights, uint32_t numFaces,
                                   uint32_t* faceLight, int* envLight) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= numLights) return;
    const clr_light_info L = lights[i];
    if (L.atInfinity)
        atomicMin(envLight, (int)i);
    else if (L.reference < numFaces)
        atomicMin(faceLight + L.reference, i);
}
__global__ void k_seed_states(uint32_t width, uint32_t height, uint32_t* states) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= width * height) return;
    uint32_t seed = i ^ 0x9E3779B9u;
    uint32_t s[4];
    for (uint32_t k = 0; k < 4; ++k) s[k] = seed = 1812433253u * (seed ^ (seed >> 30)) + k;
    reinterpret_cast<uint4*>(states)[i] = make_uint4(s[0], s[1], s[2], s[3]);
}

}  // namespace

void launchFaceLights(const clr_light_info* lights, uint32_t numLights, uint32_t numFaces, uint32_t* faceLight,
                      int* envLight, hipStream_t stream) {
    k_face_lights_clear<<<(numFaces + 255) / 256 + 1, 256, 0, stream>>>(faceLight, numFaces, envLight);
    if (numLights)
        k_face_lights_fill<<<(numLights + 255) / 256, 256, 0, stream>>>(lights, numLights, numFaces, faceLight,
                                                                        envLight);
}

// Resolves the blob-header fields of a device table of scenes in place.
__global__ void k_resolve_scenes(dev::RScene* table, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    dev::RScene S = table[i];
    dev::finishScene(S);
    table[i] = S;
}
void launchResolveScenes(dev::RScene* table, uint32_t count, hipStream_t stream) {
    k_resolve_scenes<<<(count + 63) / 64, 64, 0, stream>>>(table, count);
}

// Device-side scene view. The blob header (otherResources[0..2] = offsets of
// the camera, the environment head and the light power distribution,
// Scene.cpp:39-41) is read by the kernels themselves; the face->light map and
// the first environment light are built on the stream. No host round trip, so
// a render call can be captured into a graph.
int prepareRenderScene(clr_ctx* ctx, const clr_scene_dev& sd, dev::RScene& S, hipStream_t stream) {
    int rc = ensureRenderScratch(ctx, sd.numFaces);
    if (rc) return rc;
    rc = ensureStatus(ctx);
    if (rc) return rc;
    uint32_t* faceLight = ctx->faceLight;
    int* envLight = reinterpret_cast<int*>(ctx->renderWords);
    k_face_lights_clear<<<(sd.numFaces + 255) / 256 + 1, 256, 0, stream>>>(faceLight, sd.numFaces, envLight);
    if (sd.numLights)
        k_face_lights_fill<<<(sd.numLights + 255) / 256, 256, 0, stream>>>(sd.lights, sd.numLights,
                                                                           sd.numFaces, faceLight, envLight);
    const float4 *anodes = nullptr, *atris = nullptr, *wnodes = nullptr;
    rc = buildAccel(ctx, sd.internalNodes, sd.leafNodes, sd.bvhNodes, sd.numFaces, sd.vertices, sd.faces, &anodes,
                    &atris, stream, &wnodes);
    if (rc) return rc;
    S.g = dev::GeomScene{sd.internalNodes, sd.leafNodes, sd.numFaces, sd.vertices, sd.faces, anodes, atris,
                         reinterpret_cast<const float4*>(sd.bvhNodes), wnodes};
    S.normals = sd.normals;
    S.tangents = sd.tangents;
    S.uvs = sd.uvs;
    S.lights = sd.lights;
    S.mat = sd.materialsData;
    S.tex = sd.texturesData;
    S.oth = sd.otherResources;
    S.camera = nullptr;
    S.envLP = nullptr;
    S.lightDist = nullptr;
    S.faceLight = faceLight;
    S.envLightPtr = envLight;
    S.envLight = -1;
    S.status = ctx->devStatus;
    return hipCheck(ctx, hipGetLastError(), "scene prepare launch");
}

int renderMegakernel(clr_ctx* ctx, const clr_scene_dev& sd, const clr_render_params& rp,
                     uint32_t* states, clr_float3* pixels, hipStream_t stream) {
    dev::RScene S;
    int rc = prepareRenderScene(ctx, sd, S, stream);
    if (rc) return rc;
    const uint32_t tiles = ((rp.width + kTile - 1) / kTile) * ((rp.height + kTile - 1) / kTile);
    const uint32_t flags = rp.flags;
    int cus = 0;
    rc = hipCheck(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device), "CU count");
    if (rc) return rc;
    const bool merged = !(flags & CLR_RENDER_NO_RR) && !(flags & CLR_RENDER_REFERENCE_TRAVERSAL);
    const uint32_t grid =
        std::max(1u, std::min(tiles, (uint32_t)cus * (merged ? kMegaMergedWavesPerCU : kMegaWavesPerCU)));
    rc = ensureTraceSpill(ctx, (size_t)grid * 64 * (dev::kStackDepth - kMegaLds) * sizeof(uint32_t));
    if (rc) return rc;
    uint32_t* counter = ctx->devStatus + kWordMegaTiles;
    k_mega_reset<<<1, 64, 0, stream>>>(counter);
    auto launch = [&](auto kernel) {
        kernel<<<grid, 64, 0, stream>>>(S, rp.width, rp.height, rp.samplesPerPixel, rp.maxDepth, flags, states, pixels,
                                        counter, ctx->traceSpill);
    };
    const bool few = tiles <= (uint32_t)cus * kMegaFewTiles;
    if (!(rp.flags & CLR_RENDER_REFERENCE_TRAVERSAL)) {
        if (few)
            launch(merged ? k_render_mega_few<dev::kWalkOrdered, true> : k_render_mega_few<dev::kWalkOrdered, false>);
        else
            launch(merged ? k_render_mega<dev::kWalkOrdered, true> : k_render_mega<dev::kWalkOrdered, false>);
    } else if (sd.bvhNodes)
        launch(k_render_mega<dev::kWalkMidpoint, false>);
    else
        launch(k_render_mega<dev::kWalkLbvh, false>);
    return hipCheck(ctx, hipGetLastError(), "render launch");
}

int renderMegakernelBatch(clr_ctx* ctx, const clr_scene_dev* scenes, uint32_t numScenes,
                          const clr_render_params& rp, uint32_t* states, clr_float3* pixels, hipStream_t stream) {
    int cus = 0;
    int rc = hipCheck(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device), "CU count");
    if (rc) return rc;
    const uint64_t tiles = (uint64_t)((rp.width + kTile - 1) / kTile) * ((rp.height + kTile - 1) / kTile) * numScenes;
    const bool merged = !(rp.flags & CLR_RENDER_NO_RR);
    const uint32_t grid = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>(tiles, (uint64_t)cus * (merged ? kMegaMergedWavesPerCU : kMegaWavesPerCU)));
    const size_t spillBytes = (size_t)grid * 64 * (dev::kStackDepth - kMegaLds) * sizeof(uint32_t);
    if (ctx->traceSpillBytes < spillBytes) {
        bool capturing = false;
        rc = streamCapturing(ctx, stream, capturing);
        if (rc) return rc;
        if (capturing)
            return fail(ctx, CLR_ERR_INVALID_ARG,
                        "the trace spill must be sized by an eager render of this batch before capture");
    }
    rc = ensureTraceSpill(ctx, spillBytes);
    if (rc) return rc;
    dev::RScene* table = nullptr;
    dev::RScene first;
    rc = prepareBatchScenes(ctx, scenes, numScenes, stream, &table, &first);
    if (rc) return rc;
    uint32_t* counter = ctx->devStatus + kWordMegaTiles;
    k_mega_reset<<<1, 64, 0, stream>>>(counter);
    auto launch = [&](auto kernel) {
        kernel<<<grid, 64, 0, stream>>>(table, numScenes, rp.width, rp.height, rp.samplesPerPixel, rp.maxDepth,
                                        rp.flags, states, pixels, counter, ctx->traceSpill);
    };
    if (tiles <= (uint64_t)cus * kMegaFewTiles)
        launch(merged ? k_render_mega_batch_few<true> : k_render_mega_batch_few<false>);
    else
        launch(merged ? k_render_mega_batch<true> : k_render_mega_batch<false>);
    return hipCheck(ctx, hipGetLastError(), "batch render launch");
}

int renderBvhVisualizer(clr_ctx* ctx, const clr_scene_dev& sd, const clr_render_params& rp, clr_float3* pixels,
                        hipStream_t stream) {
    dev::RScene S;
    int rc = prepareRenderScene(ctx, sd, S, stream);
    if (rc) return rc;
    const uint32_t tiles = ((rp.width + kTile - 1) / kTile) * ((rp.height + kTile - 1) / kTile);
    if (!(rp.flags & CLR_RENDER_REFERENCE_TRAVERSAL))
        k_bvh_visualizer<dev::kWalkOrdered><<<tiles, 64, 0, stream>>>(S, rp.width, rp.height, pixels);
    else if (sd.bvhNodes)
        k_bvh_visualizer<dev::kWalkMidpoint><<<tiles, 64, 0, stream>>>(S, rp.width, rp.height, pixels);
    else
        k_bvh_visualizer<dev::kWalkLbvh><<<tiles, 64, 0, stream>>>(S, rp.width, rp.height, pixels);
    return hipCheck(ctx, hipGetLastError(), "visualizer launch");
}

int renderAdjoint(clr_ctx* ctx, const clr_scene_dev& sd, const clr_render_params& rp, const clr_diff_params& dp,
                  const uint32_t* states, const clr_float3* dLdI, float* gAlb, float* gSig, float* gEm,
                  hipStream_t stream) {
    if (!ctx->adjTable) {
        hipError_t e = hipMalloc(&ctx->adjTable, 65536 * sizeof(uint16_t) + 256 * sizeof(uint32_t));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            ctx->adjTable = nullptr;
            return fail(ctx, CLR_ERR_OUT_OF_MEMORY, "adjoint parameter table");
        }
    }
    uint16_t* table = static_cast<uint16_t*>(ctx->adjTable);
    uint32_t* offsets = reinterpret_cast<uint32_t*>(table + 65536);
    int rc = CLR_OK;
    if (dp.numMaterials)
        rc = hipCheck(ctx, hipMemcpyAsync(offsets, dp.matOffsets, dp.numMaterials * sizeof(uint32_t),
                                          hipMemcpyHostToDevice, stream), "parameter offsets");
    if (rc) return rc;
    k_adj_clear<<<256, 256, 0, stream>>>(table, gAlb, gSig, gEm, dp.numMaterials);
    if (dp.numMaterials) k_adj_fill<<<1, 256, 0, stream>>>(table, offsets, dp.numMaterials);
    dev::RScene S;
    rc = prepareRenderScene(ctx, sd, S, stream);
    if (rc) return rc;
    const uint32_t tiles = ((rp.width + kTile - 1) / kTile) * ((rp.height + kTile - 1) / kTile);
    const uint32_t flags = rp.flags;
    const float4* adj = reinterpret_cast<const float4*>(dLdI);
    int cus = 0;
    rc = hipCheck(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device), "CU count");
    if (rc) return rc;
    const uint32_t grid = std::max(1u, std::min(tiles, (uint32_t)cus * kAdjWavesPerCU));
    rc = ensureTraceSpill(ctx, (size_t)grid * 64 * (dev::kStackDepth - kAdjLds) * sizeof(uint32_t));
    if (rc) return rc;
    uint32_t* counter = ctx->devStatus + kWordMegaTiles;
    k_mega_reset<<<1, 64, 0, stream>>>(counter);
    if (!(rp.flags & CLR_RENDER_REFERENCE_TRAVERSAL))
        k_render_adjoint<dev::kWalkOrdered><<<grid, 64, 0, stream>>>(
            S, rp.width, rp.height, rp.samplesPerPixel, rp.maxDepth, flags, states, adj, table, dp.numMaterials,
            dp.lightProp, gAlb, gSig, gEm, counter, ctx->traceSpill);
    else if (sd.bvhNodes)
        k_render_adjoint<dev::kWalkMidpoint><<<grid, 64, 0, stream>>>(
            S, rp.width, rp.height, rp.samplesPerPixel, rp.maxDepth, flags, states, adj, table, dp.numMaterials,
            dp.lightProp, gAlb, gSig, gEm, counter, ctx->traceSpill);
    else
        k_render_adjoint<dev::kWalkLbvh><<<grid, 64, 0, stream>>>(
            S, rp.width, rp.height, rp.samplesPerPixel, rp.maxDepth, flags, states, adj, table, dp.numMaterials,
            dp.lightProp, gAlb, gSig, gEm, counter, ctx->traceSpill);
    return hipCheck(ctx, hipGetLastError(), "adjoint launch");
}

int seedPixelStates(clr_ctx* ctx, uint32_t width, uint32_t height, uint32_t* states, hipStream_t stream) {
    const uint32_t n = width * height;
    if (n == 0) return CLR_OK;
    k_seed_states<<<(n + 255) / 256, 256, 0, stream>>>(width, height, states);
    return hipCheck(ctx, hipGetLastError(), "seed launch");
}

}  // namespace clr
