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

struct PathOut {
    V3 L;
};

// One camera path for pixel (gx, gy) (pathtracer.cl:48-196).
template <bool ORDERED>
__device__ V3 tracePath(const RScene& S, uint32_t gx, uint32_t gy, uint32_t maxDepth, Rng& rng,
                        uint32_t* stack, int stride) {
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
    if (intersect<ORDERED>(S, org, dir, stack, stride, sp)) {
        uint32_t face = sp.face;
        clr_face F = S.g.faces[face];
        vout = -dir;
        if (F.lightPtr != 0xFFFFu) {
            edfArea(S, F.lightPtr, sp, sp.u, sp.v, E);
            pix += alpha * evalLe(E, vout);
        }
        while (true) {
            bsdfAlloc(S, F.matPtr, sp, B);
            if (hasNonSpecular(B)) {
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
                const V3 sorg = sp.p + vinL * kRayEps;
                SurfPt lsp;
                const bool hit = intersect<ORDERED>(S, sorg, vinL, stack, stride, lsp);
                const bool visible = atInf ? !hit : (hit && lsp.face == lp.face);
                // A zero or infinite light density (an environment sample at a
                // pole, where sin(theta) = 0) contributes nothing.
                if (visible && lightPDF > 0.0f && lightPDF < INFINITY) {
                    const V3 vinLrev = -vinL;
                    const float fsPDF = evalFsPdf(B, vout, vinL) * absCosEDF(E, vinL) / dist2;
                    const float mis = (lightPDF * lightPDF) / (lightPDF * lightPDF + fsPDF * fsPDF);
                    const float wgt = mis * absCosEDF(E, vinL) * fabsf(dot(B.N, vinL)) / dist2 / lightPDF;
                    pix += wgt * ((alpha * evalLe(E, vinLrev)) * evalFs(B, vout, vinL));
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
            alpha = alpha * (f * (fabsf(dot(B.N, nd)) / dirPDF));
            dir = nd;
            org = sp.p + dir * kRayEps;
            ++depth;
            if (!intersect<ORDERED>(S, org, dir, stack, stride, sp)) break;
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
                    mis = (dirPDF * dirPDF) / (lightPDF * lightPDF + dirPDF * dirPDF);
                }
                pix += (mis * alpha) * evalLe(E, vout);
            }
            if (depth >= maxDepth) {
                traceContinue = false;
                break;
            }
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
            mis = (dirPDF * dirPDF) / (lightPDF * lightPDF + dirPDF * dirPDF);
        }
        pix += (mis * alpha) * evalLe(E, vout);
    }
    return pix;
}

}  // namespace dev

namespace {

constexpr int kTile = 8;  // 8x8 pixels per wave

template <bool ORDERED>
__global__ __launch_bounds__(64) void k_render_mega(dev::RScene scene, uint32_t width, uint32_t height,
                                                    uint32_t spp, uint32_t maxDepth,
                                                    uint32_t* __restrict__ states,
                                                    clr_float3* __restrict__ pixels) {
    __shared__ uint32_t s_stack[dev::kStackDepth * 64];
    dev::RScene S = scene;  // resolved copy; the kernel argument itself stays read-only
    dev::finishScene(S);
    const uint32_t tilesX = (width + kTile - 1) / kTile;
    const uint32_t gx = (blockIdx.x % tilesX) * kTile + (threadIdx.x % kTile);
    const uint32_t gy = (blockIdx.x / tilesX) * kTile + (threadIdx.x / kTile);
    if (gx >= width || gy >= height) return;
    const size_t pi = (size_t)gy * width + gx;
    uint4 st = reinterpret_cast<uint4*>(states)[pi];
    dev::Rng rng{st.x, st.y, st.z, st.w};
    dev::V3 acc = dev::v3z();
    for (uint32_t k = 0; k < spp; ++k)
        acc += dev::tracePath<ORDERED>(S, gx, gy, maxDepth, rng, s_stack + threadIdx.x, 64);
    reinterpret_cast<uint4*>(states)[pi] = make_uint4(rng.s0, rng.s1, rng.s2, rng.s3);
    float4 p = reinterpret_cast<float4*>(pixels)[pi];
    p.x += acc.x;
    p.y += acc.y;
    p.z += acc.z;
    reinterpret_cast<float4*>(pixels)[pi] = p;
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
    const float4 *anodes = nullptr, *atris = nullptr;
    rc = buildAccel(ctx, sd.internalNodes, sd.leafNodes, sd.numFaces, sd.vertices, sd.faces, &anodes, &atris,
                    stream);
    if (rc) return rc;
    S.g = dev::GeomScene{sd.internalNodes, sd.leafNodes, sd.numFaces, sd.vertices, sd.faces, anodes, atris};
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
    if (rp.flags & CLR_RENDER_REFERENCE_TRAVERSAL)
        k_render_mega<false><<<tiles, 64, 0, stream>>>(S, rp.width, rp.height, rp.samplesPerPixel, rp.maxDepth,
                                                       states, pixels);
    else
        k_render_mega<true><<<tiles, 64, 0, stream>>>(S, rp.width, rp.height, rp.samplesPerPixel, rp.maxDepth,
                                                      states, pixels);
    return hipCheck(ctx, hipGetLastError(), "render launch");
}

int seedPixelStates(clr_ctx* ctx, uint32_t width, uint32_t height, uint32_t* states, hipStream_t stream) {
    const uint32_t n = width * height;
    if (n == 0) return CLR_OK;
    k_seed_states<<<(n + 255) / 256, 256, 0, stream>>>(width, height, states);
    return hipCheck(ctx, hipGetLastError(), "seed launch");
}

}  // namespace clr
