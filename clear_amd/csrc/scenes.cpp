// Procedural benchmark scenes built with the drop-in host Scene API
// (include/clear_amd_scene.hpp) and handed to callers as the reference's byte
// blobs (declared in include/clear_amd_scenes.h).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/clear_amd_loaders.hpp"
#include "../../include/clear_amd_scene.hpp"
#include "../../include/clear_amd_scenes.h"

using namespace clr::host;

struct clr_host_scene {
    Scene scene;
    uint32_t width = 0, height = 0;
    // Differentiable parameters (CLR_SCENE_ADJOINT): per material its byte
    // offset in materialsData and the texturesData offsets of its albedo
    // (3 floats) and roughness (1 float); the emitter's light-property offset
    // and the texturesData offset of its emission (3 floats).
    std::vector<uint32_t> materialOffsets, albedoOffsets, sigmaOffsets;
    uint32_t lightProp = UINT32_MAX, emissionOffset = UINT32_MAX;
};

namespace {

constexpr float kPiF = 3.14159265358979323846f;

class Rng {
    uint32_t s[4];

public:
    explicit Rng(int seed) {
        for (uint32_t i = 0; i < 4; ++i) {
            uint32_t v = 1812433253u * (uint32_t)(seed ^ (seed >> 30)) + i;
            s[i] = v;
            seed = (int)v;
        }
        for (int i = 0; i < 50; ++i) next();
    }
    uint32_t next() {
        uint32_t t = s[0] ^ (s[0] << 11);
        s[0] = s[1]; s[1] = s[2]; s[2] = s[3];
        return s[3] = (s[3] ^ (s[3] >> 19)) ^ (t ^ (t >> 8));
    }
    float uniform() {
        uint32_t b = (next() >> 9) | 0x3f800000u;
        float f;
        std::memcpy(&f, &b, 4);
        return f - 1.0f;
    }
};

struct Mesh {
    std::vector<float> p, n, uv;  // xyz, xyz, uv
    std::vector<uint32_t> tri;
};

// Icosahedron subdivided `level` times, projected to the unit sphere.
Mesh icosphere(int level) {
    const float t = (1.0f + std::sqrt(5.0f)) / 2.0f;
    std::vector<float> v = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                            0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
    std::vector<uint32_t> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4,
                               11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8,
                               3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    auto normalizeAt = [&](size_t i) {
        float l = std::sqrt(v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
        v[3 * i] /= l; v[3 * i + 1] /= l; v[3 * i + 2] /= l;
    };
    for (size_t i = 0; i < v.size() / 3; ++i) normalizeAt(i);
    for (int l = 0; l < level; ++l) {
        std::map<uint64_t, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            uint32_t idx = (uint32_t)(v.size() / 3);
            for (int k = 0; k < 3; ++k) v.push_back(0.5f * (v[3 * a + k] + v[3 * b + k]));
            normalizeAt(idx);
            mid.emplace(key, idx);
            return idx;
        };
        std::vector<uint32_t> nf;
        for (size_t i = 0; i < f.size(); i += 3) {
            uint32_t a = f[i], b = f[i + 1], c = f[i + 2];
            uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            uint32_t add[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nf.insert(nf.end(), add, add + 12);
        }
        f.swap(nf);
    }
    Mesh m;
    m.p = v;
    m.n = v;
    m.tri = f;
    m.uv.resize(2 * (v.size() / 3));
    for (size_t i = 0; i < v.size() / 3; ++i) {
        m.uv[2 * i] = 0.5f + std::atan2(v[3 * i + 2], v[3 * i]) / (2 * kPiF);
        m.uv[2 * i + 1] = std::acos(std::fmax(-1.0f, std::fmin(1.0f, v[3 * i + 1]))) / kPiF;
    }
    return m;
}

// Latitude/longitude sphere with nu segments and nv rings: 2 nu (nv - 1) triangles.
Mesh uvsphere(uint32_t nu, uint32_t nv) {
    Mesh m;
    for (uint32_t j = 0; j <= nv; ++j) {
        float th = kPiF * j / nv;
        for (uint32_t i = 0; i <= nu; ++i) {
            float ph = 2 * kPiF * i / nu;
            float x = std::sin(th) * std::cos(ph), y = std::cos(th), z = std::sin(th) * std::sin(ph);
            m.p.insert(m.p.end(), {x, y, z});
            m.n.insert(m.n.end(), {x, y, z});
            m.uv.insert(m.uv.end(), {(float)i / nu, (float)j / nv});
        }
    }
    auto id = [&](uint32_t i, uint32_t j) { return j * (nu + 1) + i; };
    for (uint32_t j = 0; j < nv; ++j)
        for (uint32_t i = 0; i < nu; ++i) {
            uint32_t a = id(i, j), b = id(i + 1, j), c = id(i, j + 1), d = id(i + 1, j + 1);
            if (j != 0) m.tri.insert(m.tri.end(), {a, b, d});
            if (j != nv - 1) m.tri.insert(m.tri.end(), {a, d, c});
        }
    return m;
}

// Adds a mesh as one object; per-face material from `matOf(faceIndex)`.
template <class MatFn>
void addMesh(Scene& s, const Mesh& m, bool normals, bool uvs, bool tangents, MatFn matOf,
             uint16_t light = UINT16_MAX) {
    s.beginObject();
    for (size_t i = 0; i < m.p.size() / 3; ++i) {
        s.addVertex(m.p[3 * i], m.p[3 * i + 1], m.p[3 * i + 2]);
        if (normals) s.addNormal(m.n[3 * i], m.n[3 * i + 1], m.n[3 * i + 2]);
        if (uvs) s.addUV(m.uv[2 * i], m.uv[2 * i + 1]);
        if (tangents) {
            float nx = m.n[3 * i], nz = m.n[3 * i + 2];
            float l = std::sqrt(nx * nx + nz * nz);
            if (l < 1e-6f) s.addTangent(1.0f, 0.0f, 0.0f);
            else s.addTangent(-nz / l, 0.0f, nx / l);
        }
    }
    for (size_t f = 0; f < m.tri.size() / 3; ++f) {
        uint32_t a = m.tri[3 * f], b = m.tri[3 * f + 1], c = m.tri[3 * f + 2];
        clr_face face = Face_P(a, b, c, matOf(f), light);
        if (normals) { face.vn0 = a; face.vn1 = b; face.vn2 = c; }
        if (uvs) { face.uv0 = a; face.uv1 = b; face.uv2 = c; }
        if (tangents) { face.vt0 = a; face.vt1 = b; face.vt2 = c; }
        s.addFace(face);
    }
    s.endObject();
}

// Axis-aligned room [-1,1]^3 (the reference's room coordinates, main.cpp:126-133).
// `front` adds the z = +1 wall (a closed box). mats: back, left, right, floor,
// ceiling, front. Floor and back wall carry UVs.
void addRoom(Scene& s, const uint16_t mats[6], bool front) {
    s.beginObject();
    const float P[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                           {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    for (auto& p : P) s.addVertex(p[0], p[1], p[2]);
    const float UV[4][2] = {{0, 0}, {2, 0}, {2, 2}, {0, 2}};
    for (auto& t : UV) s.addUV(t[0], t[1]);
    s.addFace(Face_P_UV(1, 0, 3, 0, 1, 2, mats[0]));
    s.addFace(Face_P_UV(1, 3, 2, 0, 2, 3, mats[0]));
    s.addFace(Face_P(0, 4, 7, mats[1]));
    s.addFace(Face_P(0, 7, 3, mats[1]));
    s.addFace(Face_P(5, 1, 2, mats[2]));
    s.addFace(Face_P(5, 2, 6, mats[2]));
    s.addFace(Face_P_UV(4, 5, 1, 0, 1, 2, mats[3]));
    s.addFace(Face_P_UV(4, 1, 0, 0, 2, 3, mats[3]));
    s.addFace(Face_P(2, 3, 7, mats[4]));
    s.addFace(Face_P(2, 7, 6, mats[4]));
    if (front) {
        s.addFace(Face_P(4, 6, 7, mats[5]));
        s.addFace(Face_P(4, 5, 6, mats[5]));
    }
    s.endObject();
}

// Ceiling light: quad [-h, h]^2 at height y, emitting downwards (-y).
void addCeilingLight(Scene& s, float h, float y, uint16_t mat, uint16_t light) {
    s.beginObject();
    s.addVertex(-h, y, -h);
    s.addVertex(h, y, -h);
    s.addVertex(h, y, h);
    s.addVertex(-h, y, h);
    s.addFace(Face_P(0, 1, 2, mat, light));
    s.addFace(Face_P(0, 2, 3, mat, light));
    s.endObject();
}

void setPerspective(Scene& s, clr_host_scene* hs, uint32_t w, uint32_t h, const float eye[3],
                    const float at[3], float fovYDeg, float lensRadius, float focus) {
    const float up[3] = {0, 1, 0};
    Mat4 c2w = Mat4::cameraToWorld(eye, at, up);
    PerspectiveInfo pc = makePerspectiveCamera(w, h, c2w, fovYDeg * kPiF / 180.0f, lensRadius, focus);
    s.setCamera(addDataAligned(&s.otherResouces, pc, 64));
    hs->width = w;
    hs->height = h;
}

// Procedural HDR sky: gradient from horizon to zenith plus a Gaussian sun.
std::vector<float> skyImage(uint32_t w, uint32_t h, float sunPeak, float sunTheta, float sunPhi, float sunWidth) {
    std::vector<float> px(3 * (size_t)w * h);
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            float th = kPiF * (y + 0.5f) / h, ph = 2 * kPiF * (x + 0.5f) / w;
            float up = std::cos(th);
            float g = up > 0 ? up : 0.0f;
            float r = 0.6f + 0.4f * (1 - g), gg = 0.7f + 0.3f * (1 - g), b = 1.0f;
            float ground = up < 0 ? 0.3f : 1.0f;
            float d0[3] = {std::sin(th) * std::cos(ph), std::cos(th), std::sin(th) * std::sin(ph)};
            float d1[3] = {std::sin(sunTheta) * std::cos(sunPhi), std::cos(sunTheta), std::sin(sunTheta) * std::sin(sunPhi)};
            float c = d0[0] * d1[0] + d0[1] * d1[1] + d0[2] * d1[2];
            float ang = std::acos(std::fmax(-1.0f, std::fmin(1.0f, c)));
            float sun = sunPeak * std::exp(-(ang * ang) / (2 * sunWidth * sunWidth));
            float* o = &px[3 * ((size_t)y * w + x)];
            o[0] = ground * r + sun;
            o[1] = ground * gg + sun;
            o[2] = ground * b + 0.95f * sun;
        }
    return px;
}

// RGB8 normal map (4-byte texels) of gentle bumps from seeded noise.
std::vector<uint8_t> bumpImage(uint32_t w, uint32_t h, int seed) {
    Rng rng(seed);
    std::vector<float> hgt((size_t)w * h);
    for (auto& v : hgt) v = rng.uniform();
    for (int pass = 0; pass < 3; ++pass) {  // box blur
        std::vector<float> o(hgt.size());
        for (uint32_t y = 0; y < h; ++y)
            for (uint32_t x = 0; x < w; ++x) {
                float s = 0;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) s += hgt[((y + h + dy) % h) * w + (x + w + dx) % w];
                o[(size_t)y * w + x] = s / 9.0f;
            }
        hgt.swap(o);
    }
    std::vector<uint8_t> px(4 * (size_t)w * h);
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            float dx = (hgt[y * w + (x + 1) % w] - hgt[y * w + (x + w - 1) % w]) * 4.0f;
            float dy = (hgt[((y + 1) % h) * w + x] - hgt[((y + h - 1) % h) * w + x]) * 4.0f;
            float n[3] = {-dx, -dy, 1.0f};
            float l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            for (int k = 0; k < 3; ++k) px[4 * ((size_t)y * w + x) + k] = (uint8_t)std::lround((n[k] / l * 0.5f + 0.5f) * 255.0f);
            px[4 * ((size_t)y * w + x) + 3] = 255;
        }
    return px;
}

void buildCornell(clr_host_scene* hs, const clr_scene_params& p) {
    Scene& s = hs->scene;
    MaterialCreator& mc = MaterialCreator::sharedInstance();
    mc.setScene(&s);
    mc.createFloat3ConstantTexture("R_white", 0.75f, 0.75f, 0.75f);
    mc.createFloat3ConstantTexture("R_red", 0.6f, 0.0f, 0.0f);
    mc.createFloat3ConstantTexture("R_green", 0.0f, 0.6f, 0.0f);
    mc.createFloatConstantTexture("sigma_lambert", 0.0f);
    mc.createMatteMaterial("mat_white", nullptr, "R_white", "sigma_lambert");
    mc.createMatteMaterial("mat_red", nullptr, "R_red", "sigma_lambert");
    mc.createMatteMaterial("mat_green", nullptr, "R_green", "sigma_lambert");
    mc.createFloat3ConstantTexture("M_light", 15.0f, 15.0f, 15.0f);
    mc.createDiffuseLightProperty("light_top", "M_light");
    uint16_t W = (uint16_t)s.idxOfMat("mat_white"), R = (uint16_t)s.idxOfMat("mat_red"),
             G = (uint16_t)s.idxOfMat("mat_green");
    const uint16_t mats[6] = {W, R, G, W, W, W};
    addRoom(s, mats, true);
    s.localToWorld.push();
    s.localToWorld *= Mat4::scale(0.3f, 0.3f, 0.3f);
    Mesh sph = icosphere(p.detail ? (int)p.detail : 5);
    addMesh(s, sph, true, false, false, [&](size_t) { return W; });
    s.localToWorld.pop();
    addCeilingLight(s, 0.25f, 0.9999f, W, (uint16_t)s.idxOfLight("light_top"));
    const float eye[3] = {0.0f, 0.0f, 0.95f}, at[3] = {0.0f, 0.0f, 0.0f};
    setPerspective(s, hs, p.width ? p.width : 256, p.height ? p.height : 256, eye, at, 60.0f, 0.0f, 1.0f);
}

// Every BxDF, texture type, normal mapping, alpha cut-outs, an image-based
// environment through the open front and a thin lens.
void buildShowcase(clr_host_scene* hs, const clr_scene_params& p) {
    Scene& s = hs->scene;
    MaterialCreator& mc = MaterialCreator::sharedInstance();
    mc.setScene(&s);
    const uint32_t ew = 64, eh = 32;
    std::vector<float> sky = skyImage(ew, eh, 40.0f, 0.9f, 1.2f, 0.12f);
    mc.createImageTexture("env", TextureType::ColorImageRGBF32x3, ew, eh, sky.data());
    mc.createImageBasedEnvLightPropety("IBL", "env", 1.0f);
    EnvironmentHead eh0{(uint32_t)s.idxOfLight("IBL")};
    s.setEnvironment(addDataAligned(&s.otherResouces, eh0, 4));

    mc.createFloat3ConstantTexture("R_white", 0.75f, 0.75f, 0.75f);
    mc.createFloat3ConstantTexture("R_red", 0.7f, 0.2f, 0.2f);
    mc.createFloatConstantTexture("sigma0", 0.0f);
    mc.createFloatConstantTexture("sigma_rough", 0.35f);
    mc.createFloat3CheckerBoardTexture("R_floor", 0.8f, 0.8f, 0.8f, 0.2f, 0.2f, 0.2f);
    mc.createFloat3CheckerBoardBumpTexture("bump_floor", 0.08f, false);
    std::vector<uint8_t> bump = bumpImage(64, 64, 7);
    mc.createNormalMapTexture("bump_back", 64, 64, bump.data());
    std::vector<uint8_t> tex(4 * 32 * 32);
    for (uint32_t y = 0; y < 32; ++y)
        for (uint32_t x = 0; x < 32; ++x) {
            uint8_t* t = &tex[4 * (y * 32 + x)];
            t[0] = (uint8_t)(x * 8); t[1] = (uint8_t)(y * 8); t[2] = 128; t[3] = ((x / 4 + y / 4) % 3) ? 255 : 0;
        }
    mc.createImageTexture("R_back", TextureType::ColorImageRGBA8x4, 32, 32, tex.data());
    mc.createFloat3ConstantTexture("Rs_back", 0.2f, 0.2f, 0.2f);
    mc.createFloatConstantTexture("nu", 60.0f);
    mc.createFloatConstantTexture("nv", 20.0f);
    mc.createFloat3ConstantTexture("R_ward", 0.3f, 0.5f, 0.8f);
    mc.createFloatConstantTexture("ax", 0.15f);
    mc.createFloatConstantTexture("ay", 0.4f);
    mc.createFloat3ConstantTexture("R_one", 1.0f, 1.0f, 1.0f);
    mc.createFloat3ConstantTexture("R_coat", 0.9f, 0.9f, 0.9f);
    mc.createFloatConstantTexture("alpha_coat", 0.25f);
    std::vector<uint8_t> alpha(16 * 16);
    for (uint32_t y = 0; y < 16; ++y)
        for (uint32_t x = 0; x < 16; ++x) alpha[y * 16 + x] = ((x / 2 + y / 2) % 2) ? 255 : 0;
    mc.createImageTexture("alpha_grid", TextureType::GrayImage8, 16, 16, alpha.data());

    mc.createMatteMaterial("mat_white", nullptr, "R_white", "sigma0");
    mc.createMatteMaterial("mat_rough", nullptr, "R_red", "sigma_rough");
    mc.createMatteMaterial("mat_floor", "bump_floor", "R_floor", "sigma0");
    mc.createAshikhminMaterial("mat_back", "bump_back", "Rs_back", "R_back", "nu", "nv");
    mc.createNewWardMaterial("mat_ward", nullptr, "R_ward", "ax", "ay");
    mc.createGlassMaterial("mat_glass", nullptr, "R_one", "R_one", 1.0f, 1.5f);
    mc.createMetalMaterial("mat_gold", nullptr, "R_one", 0.143f, 0.374f, 1.442f, 3.983f, 2.385f, 1.603f);
    mc.createCoatedMaterial("mat_coat", nullptr, "R_red", "sigma0", "R_coat", "alpha_coat", 1.5f);
    mc.createFloat3ConstantTexture("M_light", 8.0f, 8.0f, 8.0f);
    mc.createDiffuseLightProperty("light_top", "M_light");

    auto M = [&](const char* n) { return (uint16_t)s.idxOfMat(n); };
    const uint16_t mats[6] = {M("mat_back"), M("mat_rough"), M("mat_ward"), M("mat_floor"), M("mat_white"), 0};
    addRoom(s, mats, false);
    addCeilingLight(s, 0.3f, 0.9999f, M("mat_white"), (uint16_t)s.idxOfLight("light_top"));
    Mesh ico = icosphere(p.detail ? (int)p.detail : 3);
    struct Obj { float x, y, z, r; const char* mat; bool n, uv, tan; };
    const Obj objs[4] = {{-0.5f, -0.6f, -0.2f, 0.35f, "mat_glass", true, false, false},
                         {0.45f, -0.65f, 0.1f, 0.3f, "mat_gold", true, true, true},
                         {0.0f, -0.1f, -0.6f, 0.28f, "mat_coat", true, true, false},
                         {-0.45f, 0.35f, -0.5f, 0.2f, "mat_ward", false, false, false}};
    for (const Obj& o : objs) {
        s.localToWorld.push();
        s.localToWorld *= Mat4::translate(o.x, o.y, o.z) * Mat4::scale(o.r, o.r, o.r);
        uint16_t mm = M(o.mat);
        addMesh(s, ico, o.n, o.uv, o.tan, [&](size_t) { return mm; });
        s.localToWorld.pop();
    }
    // Alpha-cut quad in front of the back wall.
    s.beginObject();
    s.addVertex(-0.2f, -0.9f, -0.9f); s.addVertex(0.6f, -0.9f, -0.9f);
    s.addVertex(0.6f, -0.1f, -0.9f);  s.addVertex(-0.2f, -0.1f, -0.9f);
    s.addUV(0, 0); s.addUV(1, 0); s.addUV(1, 1); s.addUV(0, 1);
    uint32_t at = (uint32_t)s.idxOfTex("alpha_grid");
    s.addFace(Face_P_UV(0, 1, 2, 0, 1, 2, M("mat_white"), UINT16_MAX, at));
    s.addFace(Face_P_UV(0, 2, 3, 0, 2, 3, M("mat_white"), UINT16_MAX, at));
    s.endObject();
    const float eye[3] = {0.0f, 0.0f, 2.8f}, tgt[3] = {0.0f, -0.2f, 0.0f};
    setPerspective(s, hs, p.width ? p.width : 128, p.height ? p.height : 128, eye, tgt, 50.0f,
                   p.lensRadius, p.focusDistance > 0 ? p.focusDistance : 2.8f);
}

void buildFurnace(clr_host_scene* hs, const clr_scene_params& p) {
    Scene& s = hs->scene;
    MaterialCreator& mc = MaterialCreator::sharedInstance();
    mc.setScene(&s);
    const uint32_t ew = 8, eh = 4;
    std::vector<float> env(3 * ew * eh, 1.0f);
    mc.createImageTexture("env", TextureType::ColorImageRGBF32x3, ew, eh, env.data());
    mc.createImageBasedEnvLightPropety("IBL", "env", 1.0f);
    EnvironmentHead eh0{(uint32_t)s.idxOfLight("IBL")};
    s.setEnvironment(addDataAligned(&s.otherResouces, eh0, 4));
    const float a = p.albedo;
    mc.createFloat3ConstantTexture("R", a, a, a);
    mc.createFloatConstantTexture("sigma0", 0.0f);
    mc.createMatteMaterial("mat", nullptr, "R", "sigma0");
    uint16_t m = (uint16_t)s.idxOfMat("mat");
    Mesh sph = icosphere(p.detail ? (int)p.detail : 4);
    s.localToWorld.push();
    s.localToWorld *= Mat4::scale(0.5f, 0.5f, 0.5f);
    addMesh(s, sph, true, false, false, [&](size_t) { return m; });
    s.localToWorld.pop();
    const float eye[3] = {0.0f, 0.0f, 2.0f}, at[3] = {0.0f, 0.0f, 0.0f};
    setPerspective(s, hs, p.width ? p.width : 64, p.height ? p.height : 64, eye, at,
                   p.fovY > 0 ? p.fovY : 40.0f, 0.0f, 1.0f);
}

// Value-noise fBm on a wrapped 256^2 lattice of seeded values, smoothstep
// interpolation, `octaves` octaves of doubling frequency and halving
// amplitude; result roughly in [-1, 1].
struct Fbm {
    std::vector<float> lat;
    explicit Fbm(int seed) : lat(256 * 256) {
        Rng r(seed);
        for (auto& v : lat) v = 2.0f * r.uniform() - 1.0f;
    }
    float noise(float x, float y) const {
        const float fx = std::floor(x), fy = std::floor(y);
        const int ix = (int)fx, iy = (int)fy;
        float tx = x - fx, ty = y - fy;
        tx = tx * tx * (3 - 2 * tx);
        ty = ty * ty * (3 - 2 * ty);
        auto at = [&](int a, int b) { return lat[(size_t)(b & 255) * 256 + (a & 255)]; };
        const float a = at(ix, iy) + tx * (at(ix + 1, iy) - at(ix, iy));
        const float b = at(ix, iy + 1) + tx * (at(ix + 1, iy + 1) - at(ix, iy + 1));
        return a + ty * (b - a);
    }
    float operator()(float x, float y, int octaves) const {
        float s = 0.0f, amp = 0.5f, f = 1.0f;
        for (int o = 0; o < octaves; ++o) {
            s += amp * noise(x * f, y * f);
            amp *= 0.5f;
            f *= 2.0f;
        }
        return s;
    }
};

// Outdoor scene: a (detail x detail)-vertex fBm heightfield over [-10,10]^2
// (default 1024 -> 2,093,058 triangles) with a tiled 512^2 RGB8 normal map,
// 200 icospheres (subdivision 4, 5,120 triangles each) resting on it, every
// object a Lambert base under a Beckmann coat (alpha U[0.05,0.5]) with a
// dielectric Fresnel of eta 1.5, lit by a 1024x512 FP32 sky with a Gaussian
// sun of peak 5e4, seen through a thin lens (default radius 0.02).
void buildTerrain(clr_host_scene* hs, const clr_scene_params& p) {
    Scene& s = hs->scene;
    MaterialCreator& mc = MaterialCreator::sharedInstance();
    mc.setScene(&s);
    const int seed = p.seed ? p.seed : 2;
    Rng rng(seed);
    const uint32_t ew = 1024, eh = 512;
    {
        std::vector<float> sky = skyImage(ew, eh, 5.0e4f, 0.55f, 2.2f, 0.03f);
        mc.createImageTexture("env", TextureType::ColorImageRGBF32x3, ew, eh, sky.data());
    }
    mc.createImageBasedEnvLightPropety("IBL", "env", 1.0f);
    EnvironmentHead eh0{(uint32_t)s.idxOfLight("IBL")};
    s.setEnvironment(addDataAligned(&s.otherResouces, eh0, 4));

    std::vector<uint8_t> nm = bumpImage(512, 512, seed + 1);
    mc.createNormalMapTexture("terrain_nm", 512, 512, nm.data());
    mc.createFloatConstantTexture("sigma0", 0.0f);
    mc.createFloat3ConstantTexture("R_coat", 1.0f, 1.0f, 1.0f);
    auto coat = [&](const std::string& name, const char* bump, float r, float g, float b) {
        const std::string rd = name + "_Rd", al = name + "_alpha";
        mc.createFloat3ConstantTexture(rd.c_str(), r, g, b);
        mc.createFloatConstantTexture(al.c_str(), 0.05f + 0.45f * rng.uniform());
        mc.createCoatedMaterial(name.c_str(), bump, rd.c_str(), "sigma0", "R_coat", al.c_str(), 1.5f);
        return (uint16_t)s.idxOfMat(name);
    };
    const uint16_t terrainMat = coat("terrain", "terrain_nm", 0.42f, 0.38f, 0.30f);

    const uint32_t G = p.detail >= 2 ? p.detail : 1024;
    const float ext = 10.0f, amp = 1.6f;
    Fbm fbm(seed);
    auto height = [&](float x, float z) { return amp * fbm(x * 0.25f + 37.0f, z * 0.25f + 11.0f, 6); };
    {
        Mesh t;
        t.p.resize(3 * (size_t)G * G);
        t.n.resize(3 * (size_t)G * G);
        t.uv.resize(2 * (size_t)G * G);
        const float step = 2 * ext / (G - 1);
        for (uint32_t j = 0; j < G; ++j)
            for (uint32_t i = 0; i < G; ++i) {
                const float x = -ext + step * i, z = -ext + step * j;
                const size_t k = (size_t)j * G + i;
                t.p[3 * k] = x;
                t.p[3 * k + 1] = height(x, z);
                t.p[3 * k + 2] = z;
                const float dx = height(x + step, z) - height(x - step, z);
                const float dz = height(x, z + step) - height(x, z - step);
                float n[3] = {-dx, 2 * step, -dz};
                const float l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                t.n[3 * k] = n[0] / l;
                t.n[3 * k + 1] = n[1] / l;
                t.n[3 * k + 2] = n[2] / l;
                t.uv[2 * k] = 8.0f * i / (G - 1);
                t.uv[2 * k + 1] = 8.0f * j / (G - 1);
            }
        t.tri.reserve(6 * (size_t)(G - 1) * (G - 1));
        for (uint32_t j = 0; j + 1 < G; ++j)
            for (uint32_t i = 0; i + 1 < G; ++i) {
                const uint32_t a = j * G + i, b = a + 1, c = a + G, d = c + 1;
                t.tri.insert(t.tri.end(), {a, c, b, b, c, d});
            }
        addMesh(s, t, true, true, true, [&](size_t) { return terrainMat; });
    }
    Mesh ico = icosphere(4);
    for (int k = 0; k < 200; ++k) {
        const float x = -8.0f + 16.0f * rng.uniform(), z = -8.0f + 16.0f * rng.uniform();
        const float r = 0.2f + 0.4f * rng.uniform();
        const float cr = 0.15f + 0.8f * rng.uniform(), cg = 0.15f + 0.8f * rng.uniform(),
                    cb = 0.15f + 0.8f * rng.uniform();
        const uint16_t m = coat("sphere" + std::to_string(k), nullptr, cr, cg, cb);
        s.localToWorld.push();
        s.localToWorld *= Mat4::translate(x, height(x, z) + 0.8f * r, z) * Mat4::scale(r, r, r);
        addMesh(s, ico, true, false, false, [&](size_t) { return m; });
        s.localToWorld.pop();
    }
    const float eye[3] = {0.0f, 3.2f, 11.0f}, at[3] = {0.0f, 0.0f, 0.0f};
    const float lens = p.lensRadius > 0 ? p.lensRadius : (p.lensRadius < 0 ? 0.0f : 0.02f);
    const float focus = p.focusDistance > 0 ? p.focusDistance : std::sqrt(3.2f * 3.2f + 11.0f * 11.0f);
    setPerspective(s, hs, p.width ? p.width : 1920, p.height ? p.height : 1080, eye, at,
                   p.fovY > 0 ? p.fovY : 45.0f, lens, focus);
}

// Differentiation scene: the closed Cornell room (walls fixed) with a
// 200,000-triangle UV sphere (500 segments x 201 rings) of radius 0.45 whose
// face f uses material f mod 64; each of the 64 materials is Oren-Nayar with
// albedo U[0.2,0.8]^3 and roughness sigma U[0.1,0.6] drawn from the host RNG
// (seed p.seed, 4 when 0). The ceiling light (emission 15) is differentiable
// too. 512x512 by default.
void buildAdjoint(clr_host_scene* hs, const clr_scene_params& p) {
    Scene& s = hs->scene;
    MaterialCreator& mc = MaterialCreator::sharedInstance();
    mc.setScene(&s);
    Rng rng(p.seed ? p.seed : 4);
    mc.createFloat3ConstantTexture("R_white", 0.75f, 0.75f, 0.75f);
    mc.createFloat3ConstantTexture("R_red", 0.6f, 0.0f, 0.0f);
    mc.createFloat3ConstantTexture("R_green", 0.0f, 0.6f, 0.0f);
    mc.createFloatConstantTexture("sigma_lambert", 0.0f);
    mc.createMatteMaterial("mat_white", nullptr, "R_white", "sigma_lambert");
    mc.createMatteMaterial("mat_red", nullptr, "R_red", "sigma_lambert");
    mc.createMatteMaterial("mat_green", nullptr, "R_green", "sigma_lambert");
    const uint32_t numMats = 64;
    std::vector<uint16_t> sphereMat(numMats);
    for (uint32_t k = 0; k < numMats; ++k) {
        const std::string r = "R_" + std::to_string(k), sg = "S_" + std::to_string(k), m = "mat_" + std::to_string(k);
        const float a0 = 0.2f + 0.6f * rng.uniform(), a1 = 0.2f + 0.6f * rng.uniform(), a2 = 0.2f + 0.6f * rng.uniform();
        mc.createFloat3ConstantTexture(r.c_str(), a0, a1, a2);
        mc.createFloatConstantTexture(sg.c_str(), 0.1f + 0.5f * rng.uniform());
        mc.createMatteMaterial(m.c_str(), nullptr, r.c_str(), sg.c_str());
        sphereMat[k] = (uint16_t)s.idxOfMat(m);
        hs->materialOffsets.push_back((uint32_t)s.idxOfMat(m));
        hs->albedoOffsets.push_back((uint32_t)s.idxOfTex(r) + 16u);
        hs->sigmaOffsets.push_back((uint32_t)s.idxOfTex(sg) + 4u);
    }
    mc.createFloat3ConstantTexture("M_light", 15.0f, 15.0f, 15.0f);
    mc.createDiffuseLightProperty("light_top", "M_light");
    hs->lightProp = (uint32_t)s.idxOfLight("light_top");
    hs->emissionOffset = (uint32_t)s.idxOfTex("M_light") + 16u;
    const uint16_t W = (uint16_t)s.idxOfMat("mat_white"), R = (uint16_t)s.idxOfMat("mat_red"),
                   G = (uint16_t)s.idxOfMat("mat_green");
    const uint16_t mats[6] = {W, R, G, W, W, W};
    addRoom(s, mats, true);
    s.localToWorld.push();
    s.localToWorld *= Mat4::scale(0.45f, 0.45f, 0.45f);
    Mesh sph = uvsphere(500, 201);
    addMesh(s, sph, true, true, false, [&](size_t f) { return sphereMat[f % numMats]; });
    s.localToWorld.pop();
    addCeilingLight(s, 0.25f, 0.9999f, W, (uint16_t)s.idxOfLight("light_top"));
    const float eye[3] = {0.0f, 0.0f, 0.95f}, at[3] = {0.0f, 0.0f, 0.0f};
    setPerspective(s, hs, p.width ? p.width : 512, p.height ? p.height : 512, eye, at,
                   p.fovY > 0 ? p.fovY : 60.0f, 0.0f, 1.0f);
}

// Batch-render scene: the cluster soup (faces 0..n-1, material i mod 8),
// the area light quad (faces n, n+1), a white floor (faces n+2, n+3).
void buildClusters(clr_host_scene* hs, const clr_scene_params& p) {
    Scene& s = hs->scene;
    MaterialCreator& mc = MaterialCreator::sharedInstance();
    mc.setScene(&s);
    const int seed = p.seed ? p.seed : 100;
    Rng rng(seed);
    mc.createFloatConstantTexture("sigma0", 0.0f);
    mc.createFloat3ConstantTexture("R_coat", 1.0f, 1.0f, 1.0f);
    mc.createFloat3ConstantTexture("R_white", 0.75f, 0.75f, 0.75f);
    mc.createMatteMaterial("mat_white", nullptr, "R_white", "sigma0");
    uint16_t mats[8];
    for (int k = 0; k < 8; ++k) {
        const std::string rd = "Rd_" + std::to_string(k), m = "mat_" + std::to_string(k);
        const float r = 0.3f + 0.6f * rng.uniform(), g = 0.3f + 0.6f * rng.uniform(), b = 0.3f + 0.6f * rng.uniform();
        mc.createFloat3ConstantTexture(rd.c_str(), r, g, b);
        if (k < 4) {
            mc.createMatteMaterial(m.c_str(), nullptr, rd.c_str(), "sigma0");
        } else {
            const std::string al = "alpha_" + std::to_string(k);
            mc.createFloatConstantTexture(al.c_str(), 0.05f + 0.45f * rng.uniform());
            mc.createCoatedMaterial(m.c_str(), nullptr, rd.c_str(), "sigma0", "R_coat", al.c_str(), 1.5f);
        }
        mats[k] = (uint16_t)s.idxOfMat(m);
    }
    mc.createFloat3ConstantTexture("M_light", 10.0f, 10.0f, 10.0f);
    mc.createDiffuseLightProperty("light_top", "M_light");
    const uint32_t n = p.numTriangles ? p.numTriangles : 10000;
    Mesh soup;
    soup.p.resize(9 * (size_t)n);
    if (clr_gen_cluster_soup(n, seed, 64, 0.02f, 1e-3f, 1e-2f, soup.p.data()) != 0)
        throw std::runtime_error("cluster soup");
    soup.tri.resize(3 * (size_t)n);
    for (size_t i = 0; i < soup.tri.size(); ++i) soup.tri[i] = (uint32_t)i;
    addMesh(s, soup, false, false, false, [&](size_t f) { return mats[f % 8]; });
    const uint16_t white = (uint16_t)s.idxOfMat("mat_white");
    s.localToWorld.push();
    s.localToWorld *= Mat4::translate(0.5f, 1.3f, 0.5f);
    addCeilingLight(s, 0.3f, 0.0f, white, (uint16_t)s.idxOfLight("light_top"));
    s.localToWorld.pop();
    // Floor [-1.5, 2.5]^2 at y = -0.1, facing up.
    s.beginObject();
    s.addVertex(-1.5f, -0.1f, -1.5f);
    s.addVertex(2.5f, -0.1f, -1.5f);
    s.addVertex(2.5f, -0.1f, 2.5f);
    s.addVertex(-1.5f, -0.1f, 2.5f);
    s.addFace(Face_P(0, 2, 1, white));
    s.addFace(Face_P(0, 3, 2, white));
    s.endObject();
    const float eye[3] = {0.5f, 0.6f, 2.0f}, at[3] = {0.5f, 0.5f, 0.5f};
    setPerspective(s, hs, p.width ? p.width : 128, p.height ? p.height : 128, eye, at,
                   p.fovY > 0 ? p.fovY : 45.0f, 0.0f, 1.0f);
}

// Direct-lighting scene: a Lambert floor under diffuse emitters (see
// CLR_SCENE_AREA_LIGHTS): the quad, the quad plus a coplanar triangle, or a
// tilted triangle.
void buildAreaLights(clr_host_scene* hs, const clr_scene_params& p) {
    Scene& s = hs->scene;
    MaterialCreator& mc = MaterialCreator::sharedInstance();
    mc.setScene(&s);
    const float rho = p.albedo > 0.0f ? p.albedo : 0.5f;
    mc.createFloatConstantTexture("sigma0", 0.0f);
    mc.createFloat3ConstantTexture("R_floor", rho, rho, rho);
    mc.createFloat3ConstantTexture("R_white", 0.75f, 0.75f, 0.75f);
    mc.createMatteMaterial("mat_floor", nullptr, "R_floor", "sigma0");
    mc.createMatteMaterial("mat_white", nullptr, "R_white", "sigma0");
    mc.createFloat3ConstantTexture("M_light", 4.0f, 2.0f, 1.0f);
    mc.createDiffuseLightProperty("light", "M_light");
    const uint16_t floorMat = (uint16_t)s.idxOfMat("mat_floor"), white = (uint16_t)s.idxOfMat("mat_white");
    const uint16_t light = (uint16_t)s.idxOfLight("light");
    s.beginObject();
    s.addVertex(-2.0f, 0.0f, -2.0f);
    s.addVertex(2.0f, 0.0f, -2.0f);
    s.addVertex(2.0f, 0.0f, 2.0f);
    s.addVertex(-2.0f, 0.0f, 2.0f);
    s.addFace(Face_P(0, 2, 1, floorMat));
    s.addFace(Face_P(0, 3, 2, floorMat));
    s.endObject();
    if (p.detail == 0 || p.detail == 1) {
        // [-0.1, 0.7] x [-0.45, 0.05] at y = 0.6: the unit quad scaled and moved.
        s.localToWorld.push();
        s.localToWorld *= Mat4::translate(0.3f, 0.6f, -0.2f);
        s.localToWorld *= Mat4::scale(0.4f, 1.0f, 0.25f);
        addCeilingLight(s, 1.0f, 0.0f, white, light);
        s.localToWorld.pop();
    }
    if (p.detail == 1) {
        s.beginObject();
        s.addVertex(-0.9f, 0.6f, 0.1f);
        s.addVertex(-0.3f, 0.6f, 0.5f);
        s.addVertex(-0.8f, 0.6f, 0.7f);
        s.addFace(Face_P(0, 1, 2, white, light));
        s.endObject();
    } else if (p.detail == 2) {
        // Circumradius 0.5 about (0, 0.6, 0) in the plane of normal n = (1, -1, 0) / sqrt 2,
        // spanned by e1 = (0, 0, 1) and e2 = n x e1 = (-1, -1, 0) / sqrt 2.
        const float r2 = 0.70710678f;
        s.beginObject();
        for (int k = 0; k < 3; ++k) {
            const float a = (90.0f + 120.0f * k) * kPiF / 180.0f, c = 0.5f * std::cos(a), d = 0.5f * std::sin(a);
            s.addVertex(-d * r2, 0.6f - d * r2, c);
        }
        // (p1 - p0) x (p2 - p0) points along +n for this order.
        s.addFace(Face_P(0, 1, 2, white, light));
        s.endObject();
    }
    const float eye[3] = {0.0f, 2.5f, 2.5f}, at[3] = {0.0f, 0.0f, 0.0f};
    setPerspective(s, hs, p.width ? p.width : 32, p.height ? p.height : 32, eye, at,
                   p.fovY > 0 ? p.fovY : 45.0f, 0.0f, 1.0f);
}

}  // namespace

extern "C" {

int clr_scene_build(int kind, const clr_scene_params* params, clr_host_scene** out) {
    if (!out) return CLR_ERR_INVALID_ARG;
    *out = nullptr;
    clr_scene_params p{};
    if (params) p = *params;
    clr_host_scene* hs = new (std::nothrow) clr_host_scene();
    if (!hs) return CLR_ERR_OUT_OF_MEMORY;
    try {
        hs->scene.buildCpuBvh = p.buildCpuBvh != 0;
        switch (kind) {
            case CLR_SCENE_CORNELL: buildCornell(hs, p); break;
            case CLR_SCENE_SHOWCASE: buildShowcase(hs, p); break;
            case CLR_SCENE_FURNACE: buildFurnace(hs, p); break;
            case CLR_SCENE_TERRAIN: buildTerrain(hs, p); break;
            case CLR_SCENE_ADJOINT: buildAdjoint(hs, p); break;
            case CLR_SCENE_CLUSTERS: buildClusters(hs, p); break;
            case CLR_SCENE_AREA_LIGHTS: buildAreaLights(hs, p); break;
            default: delete hs; return CLR_ERR_UNSUPPORTED;
        }
        hs->scene.build();
    } catch (const std::exception&) {
        delete hs;
        return CLR_ERR_INVALID_ARG;
    }
    *out = hs;
    return CLR_OK;
}

int clr_bvh_build_midpoint(const clr_float3* vertices, const clr_face* faces, uint32_t numFaces,
                           clr_bvh_node* out) {
    if (!vertices || !faces || !out || numFaces == 0) return CLR_ERR_INVALID_ARG;
    try {
        BVH bvh;
        bvh.leaves.reserve(numFaces);
        bvh.lIdx.reserve(numFaces);
        for (uint32_t i = 0; i < numFaces; ++i) {
            BBox b;
            b.unionP(&vertices[faces[i].p0].x);
            b.unionP(&vertices[faces[i].p1].x);
            b.unionP(&vertices[faces[i].p2].x);
            bvh.addLeaf(b);
        }
        bvh.build();
        static_assert(sizeof(BVHNode) == sizeof(clr_bvh_node), "BVH node layout");
        std::memcpy(out, bvh.nodes.data(), bvh.nodes.size() * sizeof(BVHNode));
    } catch (const std::exception&) {
        return CLR_ERR_OUT_OF_MEMORY;
    }
    return CLR_OK;
}

namespace {
thread_local std::string g_sceneError;
}

const char* clr_scene_last_error(void) { return g_sceneError.c_str(); }

int clr_image_load(const char* path, int gammaCorrection, uint32_t* width, uint32_t* height, uint32_t* colorType,
                   void* pixels, uint64_t capacity, uint64_t* bytes) {
    g_sceneError.clear();
    if (!path || !width || !height || !colorType || !bytes) return CLR_ERR_INVALID_ARG;
    std::vector<uint8_t> data;
    ColorChannel::Value ct;
    try {
        if (!loadImage(path, &data, width, height, &ct, gammaCorrection != 0, &g_sceneError))
            return CLR_ERR_INVALID_ARG;
    } catch (const std::exception& e) {
        g_sceneError = e.what();
        return CLR_ERR_OUT_OF_MEMORY;
    }
    *colorType = (uint32_t)ct;
    *bytes = data.size();
    if (pixels) {
        if (capacity < data.size()) {
            g_sceneError = "pixel buffer too small";
            return CLR_ERR_INVALID_ARG;
        }
        std::memcpy(pixels, data.data(), data.size());
    }
    return CLR_OK;
}

int clr_scene_build_files(const char* const* objPaths, const float* emission, uint32_t numModels,
                          const char* envPath, float envMultiplier, const clr_scene_params* params,
                          clr_host_scene** out) {
    g_sceneError.clear();
    if (!out || (numModels && !objPaths)) return CLR_ERR_INVALID_ARG;
    *out = nullptr;
    clr_scene_params p{};
    if (params) p = *params;
    clr_host_scene* hs = new (std::nothrow) clr_host_scene();
    if (!hs) return CLR_ERR_OUT_OF_MEMORY;
    try {
        Scene& s = hs->scene;
        s.buildCpuBvh = p.buildCpuBvh != 0;
        MaterialCreator& mc = MaterialCreator::sharedInstance();
        mc.setScene(&s);
        for (uint32_t k = 0; k < numModels; ++k) {
            uint64_t light = UINT64_MAX;
            if (emission && (emission[3 * k] != 0.0f || emission[3 * k + 1] != 0.0f || emission[3 * k + 2] != 0.0f)) {
                const std::string m = "Model" + std::to_string(k) + "_M", lp = "Model" + std::to_string(k) + "_Light";
                mc.createFloat3ConstantTexture(m.c_str(), emission[3 * k], emission[3 * k + 1], emission[3 * k + 2]);
                mc.createDiffuseLightProperty(lp.c_str(), m.c_str());
                light = s.idxOfLight(lp);
            }
            std::string err;
            if (!loadModel(objPaths[k], &s, UINT64_MAX, light, &err)) throw std::runtime_error(err);
        }
        if (s.vertices.empty()) throw std::runtime_error("the models have no vertices");
        if (envPath) {
            createImageTexture(mc, s, "EnvImage", envPath);
            mc.createImageBasedEnvLightPropety("EnvLight", "EnvImage", envMultiplier);
            EnvironmentHead eh{(uint32_t)s.idxOfLight("EnvLight")};
            s.setEnvironment(addDataAligned(&s.otherResouces, eh, 4));
        }
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (const clr_float3& v : s.vertices) {
            const float c[3] = {v.x, v.y, v.z};
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::fmin(lo[a], c[a]);
                hi[a] = std::fmax(hi[a], c[a]);
            }
        }
        const float fov = p.fovY > 0 ? p.fovY : 45.0f;
        const float at[3] = {0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2])};
        const float r = 0.5f * std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                         (hi[2] - lo[2]) * (hi[2] - lo[2]));
        const float dist = std::fmax(r, 1e-3f) / std::sin(0.5f * fov * kPiF / 180.0f);
        const float eye[3] = {at[0], at[1], at[2] + dist};
        setPerspective(s, hs, p.width ? p.width : 256, p.height ? p.height : 256, eye, at, fov,
                       p.lensRadius > 0 ? p.lensRadius : 0.0f, p.focusDistance > 0 ? p.focusDistance : dist);
        s.build();
    } catch (const std::exception& e) {
        g_sceneError = e.what();
        delete hs;
        return CLR_ERR_INVALID_ARG;
    }
    *out = hs;
    return CLR_OK;
}

int clr_scene_free(clr_host_scene* s) {
    delete s;
    return CLR_OK;
}

int clr_scene_info(const clr_host_scene* s, uint32_t* numFaces, uint32_t* numLights, uint32_t* width,
                   uint32_t* height) {
    if (!s) return CLR_ERR_INVALID_ARG;
    if (numFaces) *numFaces = (uint32_t)s->scene.faces.size();
    if (numLights) *numLights = (uint32_t)s->scene.lightInfos.size();
    if (width) *width = s->width;
    if (height) *height = s->height;
    return CLR_OK;
}

int clr_scene_diff_params(const clr_host_scene* s, uint32_t* numMaterials, uint32_t* matOffsets,
                          uint32_t* albedoOffsets, uint32_t* sigmaOffsets, uint32_t* lightProp,
                          uint32_t* emissionOffset) {
    if (!s || !numMaterials) return CLR_ERR_INVALID_ARG;
    const uint32_t n = (uint32_t)s->materialOffsets.size();
    if ((matOffsets || albedoOffsets || sigmaOffsets) && *numMaterials < n) {
        *numMaterials = n;
        return CLR_ERR_INVALID_ARG;
    }
    *numMaterials = n;
    for (uint32_t k = 0; k < n; ++k) {
        if (matOffsets) matOffsets[k] = s->materialOffsets[k];
        if (albedoOffsets) albedoOffsets[k] = s->albedoOffsets[k];
        if (sigmaOffsets) sigmaOffsets[k] = s->sigmaOffsets[k];
    }
    if (lightProp) *lightProp = s->lightProp;
    if (emissionOffset) *emissionOffset = s->emissionOffset;
    return CLR_OK;
}

int clr_scene_array(const clr_host_scene* s, int which, const void** data, uint64_t* bytes) {
    if (!s || !data || !bytes) return CLR_ERR_INVALID_ARG;
    const Scene& sc = s->scene;
    switch (which) {
        case CLR_ARRAY_VERTICES: *data = sc.vertices.data(); *bytes = sc.vertices.size() * sizeof(clr_float3); break;
        case CLR_ARRAY_NORMALS: *data = sc.normals.data(); *bytes = sc.normals.size() * sizeof(clr_float3); break;
        case CLR_ARRAY_TANGENTS: *data = sc.tangents.data(); *bytes = sc.tangents.size() * sizeof(clr_float3); break;
        case CLR_ARRAY_UVS: *data = sc.uvs.data(); *bytes = sc.uvs.size() * sizeof(float); break;
        case CLR_ARRAY_FACES: *data = sc.faces.data(); *bytes = sc.faces.size() * sizeof(clr_face); break;
        case CLR_ARRAY_LIGHTS: *data = sc.lightInfos.data(); *bytes = sc.lightInfos.size() * sizeof(clr_light_info); break;
        case CLR_ARRAY_MATERIALS: *data = sc.materialsData.data(); *bytes = sc.materialsData.size(); break;
        case CLR_ARRAY_TEXTURES: *data = sc.texturesData.data(); *bytes = sc.texturesData.size(); break;
        case CLR_ARRAY_OTHERS: *data = sc.otherResouces.data(); *bytes = sc.otherResouces.size(); break;
        case CLR_ARRAY_BVH_NODES: *data = sc.bvh.nodes.data(); *bytes = sc.bvh.nodes.size() * sizeof(BVHNode); break;
        default: return CLR_ERR_INVALID_ARG;
    }
    return CLR_OK;
}

}  // extern "C"
