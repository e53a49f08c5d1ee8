// clear_amd_scene.hpp — host-side scene assembly (header-only C++17).
//
// Drop-in for the reference's host scene ABI: Scene (OpenCL_PathTracer/Scene.h:29-295,
// Scene.cpp:12-42), MaterialCreator (CreateFunctions.hpp:16-101,
// CreateFunctions.cpp), the blob packers (CommonSources/clUtility.cpp:554-571)
// and the CPU midpoint BVH (BVH.hpp:94-252). Method names, argument meaning and
// the byte layout of the three blobs (materialsData, texturesData,
// otherResources) follow the reference, so scenes built with the reference's
// calling code produce the same bytes and render identically on the GPU.
//
// Additions (no reference counterpart): an FP32 RGB image texture type
// (TextureType::ColorImageRGBF32x3) used for HDR environments built in memory,
// and a Beckmann microfacet element (MatElem::Beckmann) for coated materials.
// Image files (PNG, OpenEXR) and OBJ/MTL models are read by clear_amd_loaders.hpp; images
// can also be supplied as in-memory pixels. Differences: calcLightPowerDistribution does not write PMF[numItems] (the reference
// writes one element past the end, Scene.h:281-284).
#pragma once

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "clear_amd.h"

namespace clr {
namespace host {

// ---- enums (MaterialStructures.hpp:14-35, TextureStructures.hpp:14-50) ----
namespace MatElem {
enum Value : uint8_t {
    Diffuse = 0, SpecularReflection, SpecularTransmission, NewWard, AshikhminS, AshikhminD,
    Beckmann  // addition
};
}
namespace LPElem { enum Value : uint8_t { DiffuseEmission = 0 }; }
namespace EnvLPElem { enum Value : uint8_t { ImageBased = 0 }; }
namespace TextureType {
enum Value : uint8_t {
    ColorConstant = 0, ColorImageRGB8x3, ColorImageRGBA8x4, ColorImageRGBA16Fx4, ColorProcedural,
    GrayImage8, FloatConstant, FloatImage, FloatProcedural,
    ColorImageRGBF32x3  // addition
};
}
namespace ColorProcedureType { enum Value : uint8_t { CheckerBoard = 0, CheckerBoardBump }; }
namespace FloatProcedureType { enum Value : uint8_t { CheckerBoard = 0 }; }
namespace FresnelID { enum Value : uint8_t { NoOp = 0, Conductor, Dielectric }; }
namespace DistributionType { enum Value : uint8_t { Discrete1D = 0, ContinuousConsts1D, ContinuousConsts2D_H }; }

// ---- blob records (byte layouts of the reference host structs) ----
struct MaterialInfo { uint8_t numBxDFs; uint8_t hasBump; uint8_t dum0[2]; uint32_t idx_bump; };
struct DiffuseRElem { uint8_t id; uint8_t dum0[3]; uint32_t idx_R; uint32_t idx_sigma; };
struct SpecularRElem { uint8_t id; uint8_t dum0[3]; uint32_t idx_R; uint32_t idx_Fresnel; };
struct SpecularTElem { uint8_t id; uint8_t dum0[3]; uint32_t idx_T; float etaExt, etaInt; uint32_t idx_Fresnel; };
struct NewWardElem { uint8_t id; uint8_t dum0[3]; uint32_t idx_R, idx_anisoX, idx_anisoY; };
struct AshikhminSElem { uint8_t id; uint8_t dum0[3]; uint32_t idx_Rs, idx_nu, idx_nv; };
struct AshikhminDElem { uint8_t id; uint8_t dum0[3]; uint32_t idx_Rs, idx_Rd; };
struct BeckmannElem { uint8_t id; uint8_t dum0[3]; uint32_t idx_R, idx_alpha, idx_Fresnel; };
struct LightPropertyInfo { uint8_t numEEDFs; };
struct DiffuseLElem { uint8_t id; uint8_t dum0[3]; uint32_t idx_M; };
struct ImageBasedEnvLElem { uint8_t id; uint8_t dum0[3]; uint32_t idx_Le; float multiplier; uint32_t idx_Dist2D; };

struct Float3ConstantTexture { uint8_t texType; uint8_t dum0[15]; float value[4]; };
struct FloatConstantTexture { uint8_t texType; uint8_t dum0[3]; float value; };
struct ImageTexture { uint8_t texType; uint8_t dum0[3]; uint32_t width, height; int32_t offsetData; };
struct Float3CheckerBoardTexture { uint8_t texType, procedureType; uint8_t dum0[14]; float c[2][4]; };
struct Float3CheckerBoardBumpTexture { uint8_t texType, procedureType; uint8_t dum0[2]; float width; uint8_t reverse; uint8_t dum1[3]; };
struct FloatCheckerBoardTexture { uint8_t texType, procedureType; uint8_t dum0[2]; float v[2]; };
struct FresnelNoOp { uint8_t fresnelType; };
struct FresnelConductor { uint8_t fresnelType; uint8_t dum0[15]; float eta[4]; float k[4]; };
struct FresnelDielectric { uint8_t fresnelType; uint8_t dum0[3]; float etaExt, etaInt; };

struct Discrete1D { uint8_t type; uint8_t dum0[3]; uint32_t numItems; int32_t offsetPMF, offsetCDF; };
struct ContinuousConsts1D { uint8_t type; uint8_t dum0[3]; uint32_t numValues; float startDomain, endDomain, widthStratum; int32_t offsetPDF, offsetCDF; };
struct ContinuousConsts2D_H { uint8_t type; uint8_t dum0[3]; int32_t offsetChildren; ContinuousConsts1D distParent; };
static_assert(sizeof(Discrete1D) == 16 && sizeof(ContinuousConsts1D) == 28 && sizeof(ContinuousConsts2D_H) == 36, "distribution layout");

// PerspectiveInfo (camera.cl:27-35; packed by main.cpp:52-90), 256 bytes.
struct PerspectiveInfo {
    uint32_t width, height; uint8_t dum0[56];
    float localToWorld[16];  // column-major
    uint8_t id; uint8_t dum1[3];
    float virtualPlaneArea, lensRadius, objPDistance; uint8_t dum2[48];
    float rasterToCamera[16];
};
static_assert(sizeof(PerspectiveInfo) == 256, "camera layout");
struct EnvironmentHead { uint32_t idx_envLightProperty; };

// ---- blob packers (clUtility.cpp:554-571) ----
inline uint64_t alignBlob(std::vector<uint8_t>* v, uint32_t a) {
    size_t fill = ((v->size() + (a - 1)) & ~size_t(a - 1)) - v->size();
    v->insert(v->end(), fill, 0);
    return v->size();
}
inline uint64_t addDataAligned(std::vector<uint8_t>* v, const void* data, size_t bytes, uint32_t a) {
    alignBlob(v, a);
    size_t at = v->size();
    v->insert(v->end(), bytes, 0);
    std::memcpy(v->data() + at, data, bytes);
    return at;
}
template <typename T>
inline uint64_t addDataAligned(std::vector<uint8_t>* v, const T& d, uint32_t a = sizeof(T)) {
    return addDataAligned(v, &d, sizeof(T), a);
}
inline uint64_t fillZerosAligned(std::vector<uint8_t>* v, uint32_t bytes, uint32_t a) {
    alignBlob(v, a);
    size_t at = v->size();
    v->insert(v->end(), bytes, 0);
    return at;
}

// ---- 4x4 column-major matrices (host transforms for scene setup) ----
struct Mat4 {
    float m[16];  // m[col * 4 + row]
    static Mat4 identity() { Mat4 r{}; r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f; return r; }
    float& at(int row, int col) { return m[col * 4 + row]; }
    float at(int row, int col) const { return m[col * 4 + row]; }
    Mat4 operator*(const Mat4& b) const {
        Mat4 r{};
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                float s = 0.0f;
                for (int k = 0; k < 4; ++k) s += at(i, k) * b.at(k, j);
                r.at(i, j) = s;
            }
        return r;
    }
    void point(const float* p, float* out) const {
        float w = at(3, 0) * p[0] + at(3, 1) * p[1] + at(3, 2) * p[2] + at(3, 3);
        for (int i = 0; i < 3; ++i) out[i] = (at(i, 0) * p[0] + at(i, 1) * p[1] + at(i, 2) * p[2] + at(i, 3)) / w;
    }
    void vector(const float* v, float* out) const {
        for (int i = 0; i < 3; ++i) out[i] = at(i, 0) * v[0] + at(i, 1) * v[1] + at(i, 2) * v[2];
    }
    Mat4 transpose() const { Mat4 r{}; for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) r.at(i, j) = at(j, i); return r; }
    Mat4 inverse() const {
        // Gauss-Jordan with partial pivoting in double precision.
        double a[4][8];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 8; ++j) a[i][j] = j < 4 ? at(i, j) : (j - 4 == i ? 1.0 : 0.0);
        for (int c = 0; c < 4; ++c) {
            int p = c;
            for (int r = c + 1; r < 4; ++r) if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
            if (a[p][c] == 0.0) throw std::runtime_error("singular matrix");
            for (int j = 0; j < 8; ++j) std::swap(a[c][j], a[p][j]);
            double inv = 1.0 / a[c][c];
            for (int j = 0; j < 8; ++j) a[c][j] *= inv;
            for (int r = 0; r < 4; ++r)
                if (r != c) { double f = a[r][c]; for (int j = 0; j < 8; ++j) a[r][j] -= f * a[c][j]; }
        }
        Mat4 r{};
        for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) r.at(i, j) = (float)a[i][j + 4];
        return r;
    }
    static Mat4 translate(float x, float y, float z) { Mat4 r = identity(); r.at(0, 3) = x; r.at(1, 3) = y; r.at(2, 3) = z; return r; }
    static Mat4 scale(float x, float y, float z) { Mat4 r = identity(); r.at(0, 0) = x; r.at(1, 1) = y; r.at(2, 2) = z; return r; }
    static Mat4 rotate(float angle, float x, float y, float z) {
        float l = std::sqrt(x * x + y * y + z * z); x /= l; y /= l; z /= l;
        float c = std::cos(angle), s = std::sin(angle), t = 1.0f - c;
        Mat4 r = identity();
        r.at(0, 0) = t * x * x + c;     r.at(0, 1) = t * x * y - s * z; r.at(0, 2) = t * x * z + s * y;
        r.at(1, 0) = t * x * y + s * z; r.at(1, 1) = t * y * y + c;     r.at(1, 2) = t * y * z - s * x;
        r.at(2, 0) = t * x * z - s * y; r.at(2, 1) = t * y * z + s * x; r.at(2, 2) = t * z * z + c;
        return r;
    }
    // Camera-to-world of a camera at `eye` looking at `at` (camera looks down -z).
    static Mat4 cameraToWorld(const float* eye, const float* at, const float* up) {
        float f[3] = {at[0] - eye[0], at[1] - eye[1], at[2] - eye[2]};
        float fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
        for (float& v : f) v /= fl;
        float s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
        float sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        for (float& v : s) v /= sl;
        float u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
        Mat4 r = identity();
        for (int i = 0; i < 3; ++i) { r.at(i, 0) = s[i]; r.at(i, 1) = u[i]; r.at(i, 2) = -f[i]; r.at(i, 3) = eye[i]; }
        return r;
    }
};

class Mat4Stack {
    std::vector<Mat4> st_{Mat4::identity()};
public:
    const Mat4& top() const { return st_.back(); }
    void push() { st_.push_back(st_.back()); }
    void pop() { if (st_.size() > 1) st_.pop_back(); }
    Mat4Stack& operator=(const Mat4& m) { st_.back() = m; return *this; }
    Mat4Stack& operator*=(const Mat4& m) { st_.back() = st_.back() * m; return *this; }
};

// ---- CPU midpoint BVH (BVH.hpp:94-252) ----
struct BBox {
    float min[4] = {INFINITY, INFINITY, INFINITY, 0.0f};
    float max[4] = {-INFINITY, -INFINITY, -INFINITY, 0.0f};
    float centerOfAxis(int a) const { return (min[a] + max[a]) * 0.5f; }
    int widestAxis() const {
        float dx = max[0] - min[0], dy = max[1] - min[1], dz = max[2] - min[2];
        if (dx > dy && dx > dz) return 0;
        return dy > dz ? 1 : 2;
    }
    void unionP(const float* p) { for (int k = 0; k < 3; ++k) { min[k] = std::min(min[k], p[k]); max[k] = std::max(max[k], p[k]); } }
    void unionBB(const BBox& b) { for (int k = 0; k < 3; ++k) { min[k] = std::min(min[k], b.min[k]); max[k] = std::max(max[k], b.max[k]); } }
};
// 48 bytes: bbox (32) + children[2] + 2 pads; leaf iff children[1] == UINT32_MAX.
struct BVHNode {
    BBox bbox;
    uint32_t children[2] = {UINT32_MAX, UINT32_MAX};
    uint32_t dum0[2] = {0, 0};
};
static_assert(sizeof(BVHNode) == 48, "BVH node layout");

struct BVH {
    std::vector<BBox> leaves;
    std::vector<uint32_t> lIdx;
    std::vector<BVHNode> nodes;
    void addLeaf(const BBox& b) { leaves.push_back(b); lIdx.push_back((uint32_t)leaves.size() - 1); }
    // Hoare partition of [start, end) around `pivot` on `axis`; the split
    // index is never start (BVH.hpp:195-214). The scan is bounded by the
    // range (the reference can run past it when every centre is below the
    // pivot, BVH.hpp:203).
    uint32_t divide(uint32_t start, uint32_t end, int axis, float pivot) {
        if (end - start <= 2) return (start + end) / 2;
        int64_t left = (int64_t)start - 1, right = end;
        while (true) {
            do { ++left; } while (left < (int64_t)end - 1 && leaves[lIdx[left]].centerOfAxis(axis) < pivot);
            do { --right; } while (right > (int64_t)start && leaves[lIdx[right]].centerOfAxis(axis) > pivot);
            if (left < right) std::swap(lIdx[left], lIdx[right]);
            else return (uint32_t)std::max<int64_t>(left, (int64_t)start + 1);
        }
    }
    void buildRecursive(uint32_t parent, uint32_t start, uint32_t end) {
        BBox bbox;
        for (uint32_t i = start; i < end; ++i) bbox.unionBB(leaves[lIdx[i]]);
        nodes[parent].bbox = bbox;
        if (end - start < 2) { nodes[parent].children[0] = lIdx[start]; nodes[parent].children[1] = UINT32_MAX; return; }
        int axis = bbox.widestAxis();
        uint32_t split = divide(start, end, axis, bbox.centerOfAxis(axis));
        nodes.emplace_back();
        nodes[parent].children[0] = (uint32_t)nodes.size() - 1;
        buildRecursive((uint32_t)nodes.size() - 1, start, split);
        nodes.emplace_back();
        nodes[parent].children[1] = (uint32_t)nodes.size() - 1;
        buildRecursive((uint32_t)nodes.size() - 1, split, end);
    }
    void build() {
        nodes.clear();
        if (leaves.empty()) return;
        nodes.reserve(2 * leaves.size());
        nodes.emplace_back();
        buildRecursive(0, 0, (uint32_t)lIdx.size());
    }
};

// ---- Scene (Scene.h:29-295) ----
class Scene {
public:
    std::vector<clr_float3> vertices, normals, tangents;
    std::vector<float> uvs;  // pairs
    std::vector<clr_face> faces;
    std::vector<clr_light_info> lightInfos;
    std::vector<float> lightPowers;
    std::vector<uint8_t> materialsData, texturesData, otherResouces;
    std::map<std::string, uint64_t> materialsRef, lightPropsRef, texturesRef, othersRef;
    std::map<std::string, uint64_t> texFilesDatabase;  // image file -> texture head (Scene.h:129-132)
    std::vector<std::string> modelDatabase;           // loaded model files (Scene.h:137-139)
    uint64_t numModelsLoaded = 0;                     // prefixes the names of each model's materials
    BVH bvh;
    bool buildCpuBvh = true;

    bool immediateMode = false;
    uint64_t idxBaseVertices = 0, idxBaseNormals = 0, idxBaseTangents = 0, idxBaseUVs = 0;
    Mat4Stack localToWorld;

    Scene() { otherResouces.insert(otherResouces.end(), sizeof(uint32_t) * 3, 0); }

    void addVertex(float x, float y, float z) {
        float p[3] = {x, y, z}, o[3];
        localToWorld.top().point(p, o);
        vertices.push_back({o[0], o[1], o[2], 0.0f});
    }
    void addNormal(float x, float y, float z) {
        Mat4 nm = localToWorld.top().inverse().transpose();
        float v[3] = {x, y, z}, o[3];
        nm.vector(v, o);
        normals.push_back({o[0], o[1], o[2], 0.0f});
    }
    void addTangent(float x, float y, float z) {
        float v[3] = {x, y, z}, o[3];
        localToWorld.top().vector(v, o);
        tangents.push_back({o[0], o[1], o[2], 0.0f});
    }
    void addUV(float u, float v) { uvs.push_back(u); uvs.push_back(v); }
    void addFace(const clr_face& face) {
        clr_face f = face;
        if (immediateMode) {
            f.p0 += (uint32_t)idxBaseVertices; f.p1 += (uint32_t)idxBaseVertices; f.p2 += (uint32_t)idxBaseVertices;
            if (face.vn0 != UINT32_MAX) { f.vn0 += (uint32_t)idxBaseNormals; f.vn1 += (uint32_t)idxBaseNormals; f.vn2 += (uint32_t)idxBaseNormals; }
            if (face.vt0 != UINT32_MAX) { f.vt0 += (uint32_t)idxBaseTangents; f.vt1 += (uint32_t)idxBaseTangents; f.vt2 += (uint32_t)idxBaseTangents; }
            if (face.uv0 != UINT32_MAX) { f.uv0 += (uint32_t)idxBaseUVs; f.uv1 += (uint32_t)idxBaseUVs; f.uv2 += (uint32_t)idxBaseUVs; }
        }
        faces.push_back(f);
        if (face.lightPtr != UINT16_MAX) {
            clr_light_info li{};
            li.atInfinity = 0;
            li.reference = (uint32_t)faces.size() - 1;
            lightInfos.push_back(li);
            lightPowers.push_back(1.0f);
        }
    }
    bool addMaterial(uint64_t idx, const char* name) { return materialsRef.emplace(name, idx).second; }
    bool addLightProp(uint64_t idx, const char* name) { return lightPropsRef.emplace(name, idx).second; }
    bool addTexture(uint64_t idx, const char* name) { return texturesRef.emplace(name, idx).second; }
    bool addOtherResouce(uint64_t idx, const char* name) { return othersRef.emplace(name, idx).second; }
    bool addTexFileToDB(uint64_t idx, const char* filename) { return texFilesDatabase.emplace(filename, idx).second; }
    void addModelToDB(const char* filename) { modelDatabase.emplace_back(filename); }
    bool texFilehasLoaded(const std::string& filename, uint64_t* idx) const {
        auto it = texFilesDatabase.find(filename);
        *idx = it == texFilesDatabase.end() ? 0 : it->second;
        return it != texFilesDatabase.end();
    }
    bool setCamera(uint64_t idx) { return othersRef.emplace("Camera", idx).second; }
    bool setEnvironment(uint64_t idx) {
        bool ret = othersRef.emplace("Environment", idx).second;
        clr_light_info li{};
        li.atInfinity = 1;
        li.reference = UINT32_MAX;
        lightInfos.push_back(li);
        lightPowers.push_back(1.0f);
        return ret;
    }
    void beginObject() {
        immediateMode = true;
        idxBaseVertices = vertices.size(); idxBaseNormals = normals.size();
        idxBaseTangents = tangents.size(); idxBaseUVs = uvs.size() / 2;
    }
    void endObject() { immediateMode = false; }

    uint64_t numVertices() const { return vertices.size(); }
    uint64_t numFaces() const { return faces.size(); }
    uint64_t numLights() const { return lightInfos.size(); }
    static uint64_t lookup(const std::map<std::string, uint64_t>& m, const std::string& n, const char* what) {
        auto it = m.find(n);
        if (it == m.end()) throw std::runtime_error(std::string("unknown ") + what + ": " + n);
        return it->second;
    }
    uint64_t idxOfMat(const std::string& n) const { return lookup(materialsRef, n, "material"); }
    uint64_t idxOfLight(const std::string& n) const { return lookup(lightPropsRef, n, "light property"); }
    uint64_t idxOfTex(const std::string& n) const { return lookup(texturesRef, n, "texture"); }
    uint64_t idxOfOther(const std::string& n) const { return lookup(othersRef, n, "resource"); }
    uint8_t* ptrOfTex(const std::string& n) { return texturesData.data() + idxOfTex(n); }

    void calcLightPowerDistribution() {
        uint64_t head = fillZerosAligned(&otherResouces, sizeof(Discrete1D), 4);
        Discrete1D d{};
        d.type = DistributionType::Discrete1D;
        d.numItems = (uint32_t)lightPowers.size();
        std::vector<float> PMF(lightPowers.size()), CDF(lightPowers.size() + 1);
        CDF[0] = 0.0f;
        for (size_t i = 1; i < CDF.size(); ++i) CDF[i] = CDF[i - 1] + lightPowers[i - 1];
        float sum = CDF.back();
        for (size_t i = 0; i < CDF.size(); ++i) CDF[i] /= sum;
        for (size_t i = 0; i < PMF.size(); ++i) PMF[i] = lightPowers[i] / sum;
        uint64_t cdfHead = addDataAligned(&otherResouces, CDF.data(), sizeof(float) * CDF.size(), 4);
        uint64_t pmfHead = addDataAligned(&otherResouces, PMF.data(), sizeof(float) * PMF.size(), 4);
        d.offsetCDF = (int32_t)cdfHead - (int32_t)head;
        d.offsetPMF = (int32_t)pmfHead - (int32_t)head;
        std::memcpy(otherResouces.data() + head, &d, sizeof(d));
        addOtherResouce(head, "LightPowerDistribution");
    }

    void build();  // defined after MaterialCreator
};

// ---- MaterialCreator (CreateFunctions.hpp:16-101) ----
class MaterialCreator {
    Scene* scene = nullptr;
    std::string matName;
    uint8_t numBxDFs = 0, numEEDFs = 0;
    uint64_t matHead = 0, litHead = 0;

public:
    static MaterialCreator& sharedInstance() { static MaterialCreator mc; return mc; }
    void setScene(Scene* s) { scene = s; }

    void beginMaterial(const char* name, const char* bump = nullptr) {
        matName = name;
        numBxDFs = 0;
        auto* md = &scene->materialsData;
        matHead = addDataAligned<uint8_t>(md, 0);
        scene->addMaterial(matHead, name);
        bool hasBump = bump != nullptr;
        addDataAligned<uint8_t>(md, hasBump ? 1 : 0);
        addDataAligned<uint32_t>(md, hasBump ? (uint32_t)scene->idxOfTex(bump) : 0u);
    }
    void endMaterial() {
        if (numBxDFs == 0 || numBxDFs > 4) throw std::runtime_error("a material needs 1..4 BxDFs");
        scene->materialsData[matHead] = numBxDFs;
    }
    void beginLightProperty(const char* name) {
        numEEDFs = 0;
        litHead = addDataAligned<uint8_t>(&scene->materialsData, 0);
        scene->addLightProp(litHead, name);
    }
    void endLightProperty() { scene->materialsData[litHead] = numEEDFs; }

    // textures (CreateFunctions.cpp:18-123)
    void createFloat3ConstantTexture(const char* name, float r, float g, float b) {
        Float3ConstantTexture t{};
        t.texType = TextureType::ColorConstant;
        t.value[0] = r; t.value[1] = g; t.value[2] = b;
        scene->addTexture(addDataAligned(&scene->texturesData, t, 16), name);
    }
    void createFloatConstantTexture(const char* name, float v) {
        FloatConstantTexture t{};
        t.texType = TextureType::FloatConstant;
        t.value = v;
        scene->addTexture(addDataAligned(&scene->texturesData, t, 4), name);
    }
    // In-memory image (the reference loads files, CreateFunctions.cpp:38-69).
    // RGB8 pixels are stored with a 4-byte stride (ImageLoader.cpp:146 fills
    // the alpha byte), FP32 RGB with 12 bytes, gray8 with 1.
    void createImageTexture(const char* name, TextureType::Value type, uint32_t w, uint32_t h,
                            const void* pixels) {
        auto* td = &scene->texturesData;
        uint64_t head = fillZerosAligned(td, sizeof(ImageTexture), 4);
        uint64_t data = alignBlob(td, 128);
        size_t bpp = type == TextureType::ColorImageRGB8x3 || type == TextureType::ColorImageRGBA8x4 ? 4
                   : type == TextureType::ColorImageRGBF32x3 ? 12
                   : type == TextureType::FloatImage ? 4
                   : type == TextureType::GrayImage8 ? 1 : 8;
        td->insert(td->end(), pixels ? (const uint8_t*)pixels : nullptr,
                   pixels ? (const uint8_t*)pixels + bpp * w * h : nullptr);
        if (!pixels) td->insert(td->end(), bpp * w * h, 0);
        ImageTexture* im = reinterpret_cast<ImageTexture*>(td->data() + head);
        im->texType = type;
        im->width = w;
        im->height = h;
        im->offsetData = (int32_t)data - (int32_t)head;
        scene->addTexture(head, name);
    }
    // Normal map from in-memory RGB8 texels (CreateFunctions.cpp:71-86).
    void createNormalMapTexture(const char* name, uint32_t w, uint32_t h, const uint8_t* rgbx) {
        createImageTexture(name, TextureType::ColorImageRGB8x3, w, h, rgbx);
    }
    void createFloat3CheckerBoardTexture(const char* name, float r0, float g0, float b0, float r1, float g1, float b1) {
        Float3CheckerBoardTexture t{};
        t.texType = TextureType::ColorProcedural;
        t.procedureType = ColorProcedureType::CheckerBoard;
        t.c[0][0] = r0; t.c[0][1] = g0; t.c[0][2] = b0;
        t.c[1][0] = r1; t.c[1][1] = g1; t.c[1][2] = b1;
        scene->addTexture(addDataAligned(&scene->texturesData, t, 16), name);
    }
    void createFloat3CheckerBoardBumpTexture(const char* name, float width = 0.01f, bool reverse = false) {
        Float3CheckerBoardBumpTexture t{};
        t.texType = TextureType::ColorProcedural;
        t.procedureType = ColorProcedureType::CheckerBoardBump;
        t.width = width;
        t.reverse = reverse ? 1 : 0;
        scene->addTexture(addDataAligned(&scene->texturesData, t, 4), name);
    }
    void createFloatCheckerBoardTexture(const char* name, float v0, float v1) {
        FloatCheckerBoardTexture t{};
        t.texType = TextureType::FloatProcedural;  // the reference tags it ColorProcedural (CreateFunctions.cpp:117)
        t.procedureType = FloatProcedureType::CheckerBoard;
        t.v[0] = v0; t.v[1] = v1;
        scene->addTexture(addDataAligned(&scene->texturesData, t, 4), name);
    }

    // Fresnel (CreateFunctions.cpp:126-156)
    void createFresnelNoOp(const char* name) {
        FresnelNoOp f{FresnelID::NoOp};
        scene->addOtherResouce(addDataAligned(&scene->otherResouces, f, 1), name);
    }
    void createFresnelConductor(const char* name, float er, float eg, float eb, float kr, float kg, float kb) {
        FresnelConductor f{};
        f.fresnelType = FresnelID::Conductor;
        f.eta[0] = er; f.eta[1] = eg; f.eta[2] = eb;
        f.k[0] = kr; f.k[1] = kg; f.k[2] = kb;
        scene->addOtherResouce(addDataAligned(&scene->otherResouces, f, 16), name);
    }
    void createFresnelDielectric(const char* name, float etaExt, float etaInt) {
        FresnelDielectric f{};
        f.fresnelType = FresnelID::Dielectric;
        f.etaExt = etaExt; f.etaInt = etaInt;
        scene->addOtherResouce(addDataAligned(&scene->otherResouces, f, 4), name);
    }

    static float halfBitsToFloat(uint16_t h) {
        const uint32_t s = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 0x1F, m = h & 0x3FF;
        float mag = e == 0 ? std::ldexp((float)m, -24) : e == 31 ? (m ? NAN : INFINITY) : std::ldexp((float)(m | 0x400), (int)e - 25);
        return s ? -mag : mag;
    }
    // averageLuminance (CreateFunctions.cpp:159-244) of an RGBA16F image over the region
    // [xl, xr) x [yt, yb) in pixels: partial corner and edge pixels weighted by their
    // covered fraction, in the reference's order of additions.
    float averageLuminanceHalf(const ImageTexture* im, const uint16_t* px, float xl, float xr, float yt, float yb) const {
        const uint32_t x0 = (uint32_t)xl, x1 = (uint32_t)std::ceil(xr) - 1, y0 = (uint32_t)yt, y1 = (uint32_t)std::ceil(yb) - 1;
        auto lum = [&](uint64_t off) {
            const float R = halfBitsToFloat(px[off]), G = halfBitsToFloat(px[off + 1]), B = halfBitsToFloat(px[off + 2]);
            return 0.2126f * R + 0.7152f * G + 0.0722f * B;
        };
        const uint64_t w = im->width;
        float sum = 0.0f;
        const uint64_t corners[4] = {4 * (y0 * w + x0), 4 * (y0 * w + x1), 4 * (y1 * w + x0), 4 * (y1 * w + x1)};
        const float cc[4] = {(x0 + 1 - xl) * (y0 + 1 - yt), (xr - x1) * (y0 + 1 - yt), (x0 + 1 - xl) * (yb - y1), (xr - x1) * (yb - y1)};
        for (int k = 0; k < 4; ++k) sum += cc[k] * lum(corners[k]);
        const float ce[4] = {y0 + 1 - yt, x0 + 1 - xl, xr - x1, yb - y1};
        for (uint32_t x = x0 + 1; x < x1; ++x) {
            sum += ce[0] * lum(4 * (y0 * w + x));
            sum += ce[3] * lum(4 * (y1 * w + x));
        }
        for (uint32_t y = y0 + 1; y < y1; ++y) {
            sum += ce[1] * lum(4 * (y * w + x0));
            sum += ce[2] * lum(4 * (y * w + x1));
        }
        for (uint32_t y = y0 + 1; y < y1; ++y)
            for (uint32_t x = x0 + 1; x < x1; ++x) sum += lum(4 * (y * w + x));
        return sum / ((xr - xl) * (yb - yt));
    }

    // Environment importance distribution from an image: one stratum per pixel. RGBA16F
    // images (OpenEXR files) follow the reference exactly: averageLuminance of each pixel's
    // region times the zenith correction cos(pi (0.5 - (i + 0.5) / H)) (CreateFunctions.cpp:
    // 246-349). FP32 RGB images (built in memory) take the pixel's luminance times the same
    // correction.
    void createContinuousConsts2D_H_FromImageTexture(const char* name, const char* image, bool zenithCorrection) {
        auto* od = &scene->otherResouces;
        const ImageTexture* im = reinterpret_cast<const ImageTexture*>(scene->texturesData.data() + scene->idxOfTex(image));
        const bool half = im->texType == TextureType::ColorImageRGBA16Fx4;
        if (im->texType != TextureType::ColorImageRGBF32x3 && !half)
            throw std::runtime_error("environment distribution needs an FP32 RGB or RGBA16F image");
        const uint32_t sx = im->width, sy = im->height;
        const uint64_t texHead = scene->idxOfTex(image);
        const int32_t offData = im->offsetData;
        uint64_t head = fillZerosAligned(od, sizeof(ContinuousConsts2D_H), 4);
        uint64_t pPDFHead = fillZerosAligned(od, sizeof(float) * sy, 4);
        uint64_t pCDFHead = fillZerosAligned(od, sizeof(float) * (sy + 1), 4);
        uint64_t childrenHead = fillZerosAligned(od, sizeof(ContinuousConsts1D) * sy, 4);
        uint64_t cPDFsHead = fillZerosAligned(od, sizeof(float) * sx * sy, 4);
        uint64_t cCDFsHead = fillZerosAligned(od, sizeof(float) * (sx + 1) * sy, 4);
        const float* px = reinterpret_cast<const float*>(scene->texturesData.data() + texHead + offData);
        for (uint32_t i = 0; i < sy; ++i) {
            ContinuousConsts1D c{};
            c.type = DistributionType::ContinuousConsts1D;
            c.startDomain = 0.0f; c.endDomain = 1.0f;
            c.widthStratum = (c.endDomain - c.startDomain) / sx;
            c.numValues = sx;
            uint64_t childHead = childrenHead + i * sizeof(ContinuousConsts1D);
            uint64_t cPDFHead = cPDFsHead + sizeof(float) * sx * i;
            uint64_t cCDFHead = cCDFsHead + sizeof(float) * (sx + 1) * i;
            c.offsetPDF = (int32_t)cPDFHead - (int32_t)childHead;
            c.offsetCDF = (int32_t)cCDFHead - (int32_t)childHead;
            float* cPDF = reinterpret_cast<float*>(od->data() + cPDFHead);
            float* cCDF = reinterpret_cast<float*>(od->data() + cCDFHead);
            if (half) {
                const uint16_t* hp = reinterpret_cast<const uint16_t*>(scene->texturesData.data() + texHead + offData);
                const float yBase = ((float)im->height / sy) * i, yNext = ((float)im->height / sy) * (i + 1);
                const float corr = zenithCorrection ? cosf(M_PI * (0.5f - (i + 0.5f) / sy)) : 1.0f;
                for (uint32_t j = 0; j < sx; ++j) {
                    const float xBase = ((float)im->width / sx) * j, xNext = ((float)im->width / sx) * (j + 1);
                    cPDF[j] = averageLuminanceHalf(im, hp, xBase, xNext, yBase, yNext) * corr;
                }
            } else {
                float corr = zenithCorrection ? std::cos((float)M_PI * (0.5f - (i + 0.5f) / sy)) : 1.0f;
                for (uint32_t j = 0; j < sx; ++j) {
                    const float* p = px + 3 * ((size_t)i * sx + j);
                    cPDF[j] = (0.2126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2]) * corr;
                }
            }
            cCDF[0] = 0.0f;
            for (uint32_t j = 1; j <= sx; ++j) cCDF[j] = cCDF[j - 1] + cPDF[j - 1] * c.widthStratum;
            float* pPDF = reinterpret_cast<float*>(od->data() + pPDFHead) + i;
            *pPDF = cCDF[sx];
            float norm = 1.0f / cCDF[sx];
            for (uint32_t j = 0; j < sx; ++j) { cPDF[j] *= norm; cCDF[j + 1] *= norm; }
            // x * (1 / x) can round below 1; a draw u in [CDF[n], 1) would then
            // select stratum n and read CDF[n + 1] past the table.
            cCDF[sx] = 1.0f;
            std::memcpy(od->data() + childHead, &c, sizeof(c));
        }
        ContinuousConsts2D_H d{};
        d.type = DistributionType::ContinuousConsts2D_H;
        d.offsetChildren = (int32_t)childrenHead - (int32_t)head;
        d.distParent.type = DistributionType::ContinuousConsts1D;
        d.distParent.startDomain = 0.0f; d.distParent.endDomain = 1.0f;
        d.distParent.widthStratum = 1.0f / sy;
        d.distParent.numValues = sy;
        // Offsets of the marginal are relative to the 2D header, as the
        // reference host writes them (CreateFunctions.cpp:325-326).
        d.distParent.offsetPDF = (int32_t)pPDFHead - (int32_t)head;
        d.distParent.offsetCDF = (int32_t)pCDFHead - (int32_t)head;
        float* pPDF = reinterpret_cast<float*>(od->data() + pPDFHead);
        float* pCDF = reinterpret_cast<float*>(od->data() + pCDFHead);
        pCDF[0] = 0.0f;
        for (uint32_t i = 1; i <= sy; ++i) pCDF[i] = pCDF[i - 1] + pPDF[i - 1] * d.distParent.widthStratum;
        float norm = 1.0f / pCDF[sy];
        for (uint32_t j = 0; j < sy; ++j) { pPDF[j] *= norm; pCDF[j + 1] *= norm; }
        pCDF[sy] = 1.0f;
        std::memcpy(od->data() + head, &d, sizeof(d));
        scene->addOtherResouce(head, name);
    }

    // BxDF elements (CreateFunctions.cpp:352-423)
    template <typename T>
    T* newElem() {
        ++numBxDFs;
        auto* md = &scene->materialsData;
        uint64_t head = alignBlob(md, 4);
        md->insert(md->end(), sizeof(T), 0);
        return reinterpret_cast<T*>(md->data() + head);
    }
    void createDiffuseRElem(const char* R, const char* sigma) {
        uint32_t r = (uint32_t)scene->idxOfTex(R), s = (uint32_t)scene->idxOfTex(sigma);
        auto* e = newElem<DiffuseRElem>(); e->id = MatElem::Diffuse; e->idx_R = r; e->idx_sigma = s;
    }
    void createSpecularRElem(const char* R, const char* fresnel) {
        uint32_t r = (uint32_t)scene->idxOfTex(R), f = (uint32_t)scene->idxOfOther(fresnel);
        auto* e = newElem<SpecularRElem>(); e->id = MatElem::SpecularReflection; e->idx_R = r; e->idx_Fresnel = f;
    }
    void createSpecularTElem(const char* T, float etaExt, float etaInt) {
        std::string fn = "internalFresnelDielectric_" + matName;
        createFresnelDielectric(fn.c_str(), etaExt, etaInt);
        uint32_t t = (uint32_t)scene->idxOfTex(T), f = (uint32_t)scene->idxOfOther(fn);
        auto* e = newElem<SpecularTElem>(); e->id = MatElem::SpecularTransmission;
        e->idx_T = t; e->etaExt = etaExt; e->etaInt = etaInt; e->idx_Fresnel = f;
    }
    void createNewWardElem(const char* R, const char* ax, const char* ay) {
        uint32_t r = (uint32_t)scene->idxOfTex(R), x = (uint32_t)scene->idxOfTex(ax), y = (uint32_t)scene->idxOfTex(ay);
        auto* e = newElem<NewWardElem>(); e->id = MatElem::NewWard; e->idx_R = r; e->idx_anisoX = x; e->idx_anisoY = y;
    }
    void createAshikhminSElem(const char* Rs, const char* nu, const char* nv) {
        uint32_t r = (uint32_t)scene->idxOfTex(Rs), u = (uint32_t)scene->idxOfTex(nu), v = (uint32_t)scene->idxOfTex(nv);
        auto* e = newElem<AshikhminSElem>(); e->id = MatElem::AshikhminS; e->idx_Rs = r; e->idx_nu = u; e->idx_nv = v;
    }
    void createAshikhminDElem(const char* Rd, const char* Rs) {
        uint32_t d = (uint32_t)scene->idxOfTex(Rd), s = (uint32_t)scene->idxOfTex(Rs);
        auto* e = newElem<AshikhminDElem>(); e->id = MatElem::AshikhminD; e->idx_Rd = d; e->idx_Rs = s;
    }
    void createBeckmannElem(const char* R, const char* alpha, const char* fresnel) {
        uint32_t r = (uint32_t)scene->idxOfTex(R), a = (uint32_t)scene->idxOfTex(alpha), f = (uint32_t)scene->idxOfOther(fresnel);
        auto* e = newElem<BeckmannElem>(); e->id = MatElem::Beckmann; e->idx_R = r; e->idx_alpha = a; e->idx_Fresnel = f;
    }

    // materials (CreateFunctions.cpp:426-466)
    void createMatteMaterial(const char* name, const char* bump, const char* R, const char* sigma) {
        beginMaterial(name, bump); createDiffuseRElem(R, sigma); endMaterial();
    }
    void createGlassMaterial(const char* name, const char* bump, const char* R, const char* T, float etaExt, float etaInt) {
        beginMaterial(name, bump);
        std::string fn = "GlassInternalFresnelDielectric_" + matName;
        createFresnelDielectric(fn.c_str(), etaExt, etaInt);
        createSpecularRElem(R, fn.c_str());
        createSpecularTElem(T, etaExt, etaInt);
        endMaterial();
    }
    void createMetalMaterial(const char* name, const char* bump, const char* R, float er, float eg, float eb, float kr, float kg, float kb) {
        beginMaterial(name, bump);
        std::string fn = "MetalInternalFresnelConductor_" + matName;
        createFresnelConductor(fn.c_str(), er, eg, eb, kr, kg, kb);
        createSpecularRElem(R, fn.c_str());
        endMaterial();
    }
    void createNewWardMaterial(const char* name, const char* bump, const char* R, const char* ax, const char* ay) {
        beginMaterial(name, bump); createNewWardElem(R, ax, ay); endMaterial();
    }
    void createAshikhminMaterial(const char* name, const char* bump, const char* Rs, const char* Rd, const char* nu, const char* nv) {
        beginMaterial(name, bump); createAshikhminSElem(Rs, nu, nv); createAshikhminDElem(Rd, Rs); endMaterial();
    }
    // Lambert base + Beckmann coat with a dielectric Fresnel (addition).
    void createCoatedMaterial(const char* name, const char* bump, const char* Rd, const char* sigma,
                              const char* Rs, const char* alpha, float eta) {
        beginMaterial(name, bump);
        std::string fn = "CoatFresnelDielectric_" + matName;
        createFresnelDielectric(fn.c_str(), 1.0f, eta);
        createDiffuseRElem(Rd, sigma);
        createBeckmannElem(Rs, alpha, fn.c_str());
        endMaterial();
    }

    // lights (CreateFunctions.cpp:469-505)
    void createDiffuseLElem(const char* M) {
        ++numEEDFs;
        auto* md = &scene->materialsData;
        uint32_t m = (uint32_t)scene->idxOfTex(M);
        uint64_t head = alignBlob(md, 4);
        md->insert(md->end(), sizeof(DiffuseLElem), 0);
        auto* e = reinterpret_cast<DiffuseLElem*>(md->data() + head);
        e->id = LPElem::DiffuseEmission; e->idx_M = m;
    }
    void createDiffuseLightProperty(const char* name, const char* M) {
        beginLightProperty(name); createDiffuseLElem(M); endLightProperty();
    }
    void createImageBasedEnvLElem(const char* radiance, float multiplier = 1.0f) {
        std::string dn = std::string("InternalDistribution2D_") + radiance;
        createContinuousConsts2D_H_FromImageTexture(dn.c_str(), radiance, true);
        ++numEEDFs;
        auto* md = &scene->materialsData;
        uint32_t le = (uint32_t)scene->idxOfTex(radiance), dist = (uint32_t)scene->idxOfOther(dn);
        uint64_t head = alignBlob(md, 4);
        md->insert(md->end(), sizeof(ImageBasedEnvLElem), 0);
        auto* e = reinterpret_cast<ImageBasedEnvLElem*>(md->data() + head);
        e->id = EnvLPElem::ImageBased; e->idx_Le = le; e->multiplier = multiplier; e->idx_Dist2D = dist;
    }
    void createImageBasedEnvLightPropety(const char* name, const char* radiance, float multiplier = 1.0f) {
        beginLightProperty(name); createImageBasedEnvLElem(radiance, multiplier); endLightProperty();
    }
};

// Scene::build (Scene.cpp:12-42): default "Dark" environment, light power
// distribution, CPU midpoint BVH, and the three header offsets.
inline void Scene::build() {
    if (othersRef.count("Environment") == 0) {
        MaterialCreator& mc = MaterialCreator::sharedInstance();
        mc.setScene(this);
        mc.beginLightProperty("Dark");
        mc.endLightProperty();
        EnvironmentHead eh{(uint32_t)idxOfLight("Dark")};
        othersRef.emplace("Environment", addDataAligned(&otherResouces, eh, 4));
    }
    calcLightPowerDistribution();
    if (buildCpuBvh) {
        bvh = BVH();
        for (const clr_face& f : faces) {
            BBox b;
            b.unionP(&vertices[f.p0].x);
            b.unionP(&vertices[f.p1].x);
            b.unionP(&vertices[f.p2].x);
            bvh.addLeaf(b);
        }
        bvh.build();
    }
    uint32_t h[3] = {(uint32_t)idxOfOther("Camera"), (uint32_t)idxOfOther("Environment"),
                     (uint32_t)idxOfOther("LightPowerDistribution")};
    std::memcpy(otherResouces.data(), h, sizeof(h));
}

// Face factories (Face.hpp:24-72).
inline clr_face makeFace() {
    clr_face f;
    std::memset(&f, 0xFF, sizeof(f));
    return f;
}
inline clr_face Face_P(uint32_t p0, uint32_t p1, uint32_t p2, uint16_t mat, uint16_t light = UINT16_MAX,
                       uint32_t alpha = UINT32_MAX) {
    clr_face f = makeFace();
    f.p0 = p0; f.p1 = p1; f.p2 = p2; f.alphaTexPtr = alpha; f.matPtr = mat; f.lightPtr = light;
    return f;
}
inline clr_face Face_P_N(uint32_t p0, uint32_t p1, uint32_t p2, uint32_t n0, uint32_t n1, uint32_t n2, uint16_t mat,
                         uint16_t light = UINT16_MAX, uint32_t alpha = UINT32_MAX) {
    clr_face f = Face_P(p0, p1, p2, mat, light, alpha);
    f.vn0 = n0; f.vn1 = n1; f.vn2 = n2;
    return f;
}
inline clr_face Face_P_UV(uint32_t p0, uint32_t p1, uint32_t p2, uint32_t t0, uint32_t t1, uint32_t t2, uint16_t mat,
                          uint16_t light = UINT16_MAX, uint32_t alpha = UINT32_MAX) {
    clr_face f = Face_P(p0, p1, p2, mat, light, alpha);
    f.uv0 = t0; f.uv1 = t1; f.uv2 = t2;
    return f;
}
inline clr_face Face_P_N_UV(uint32_t p0, uint32_t p1, uint32_t p2, uint32_t n0, uint32_t n1, uint32_t n2, uint32_t t0,
                            uint32_t t1, uint32_t t2, uint16_t mat, uint16_t light = UINT16_MAX,
                            uint32_t alpha = UINT32_MAX) {
    clr_face f = Face_P_N(p0, p1, p2, n0, n1, n2, mat, light, alpha);
    f.uv0 = t0; f.uv1 = t1; f.uv2 = t2;
    return f;
}

// Perspective camera record (main.cpp:52-90): raster -> camera through the
// inverse projection, the lens (radius, focus distance) and the image-plane
// area at unit distance.
inline PerspectiveInfo makePerspectiveCamera(uint32_t width, uint32_t height, const Mat4& camToWorld,
                                             float fovY, float lensRadius, float objPDistance) {
    PerspectiveInfo pc{};
    pc.width = width;
    pc.height = height;
    std::memcpy(pc.localToWorld, camToWorld.m, sizeof(pc.localToWorld));
    pc.id = 0;
    float aspect = (float)width / (float)height;
    float near = 1.0f, far = 100.0f;
    pc.virtualPlaneArea = aspect * std::pow(std::tan(fovY / 2), 2.0f);
    pc.lensRadius = lensRadius;
    pc.objPDistance = objPDistance;
    Mat4 proj{};
    proj.at(0, 0) = 1.0f / (aspect * std::tan(fovY / 2));
    proj.at(1, 1) = 1.0f / std::tan(fovY / 2);
    proj.at(2, 2) = -(far + near) / (far - near);
    proj.at(3, 2) = -1.0f;
    proj.at(2, 3) = -2.0f * far * near / (far - near);
    Mat4 rasterToScreen{};
    rasterToScreen.at(0, 0) = 2.0f / width;
    rasterToScreen.at(1, 1) = -2.0f / height;
    rasterToScreen.at(2, 2) = 2.0f;
    rasterToScreen.at(0, 3) = -1.0f;
    rasterToScreen.at(1, 3) = 1.0f;
    rasterToScreen.at(2, 3) = -1.0f;
    rasterToScreen.at(3, 3) = 1.0f;
    Mat4 r2c = proj.inverse() * rasterToScreen;
    std::memcpy(pc.rasterToCamera, r2c.m, sizeof(pc.rasterToCamera));
    return pc;
}

}  // namespace host
}  // namespace clr
