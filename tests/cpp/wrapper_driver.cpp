// GPU driver for the C++ wrappers of include/clear_amd.hpp (run by
// tests/test_gpu_cpp_wrappers.py, which checks its outputs against the
// oracle). Usage: wrapper_driver <output directory>
//
//  1. clr::LBVHBuilder builds a 50,000-triangle cluster soup (seed 7) given as
//     16-B vertices + Face records (vertex 3i + k = corner k of triangle i),
//     with profiling on; the nodes go to lbvh_inodes.bin / lbvh_leaves.bin.
//  2. The Cornell box (48 x 48, sphere subdivision 3) is built on the host,
//     uploaded, given an LBVH by the same builder and rendered with
//     clr::render (4 spp, depth 5) by the megakernel and by the wavefront
//     renderer from freshly seeded states: render_mega.bin, render_wave.bin.
//  3. A build with numBitsPerDim 0 must throw clr::Error(CLR_ERR_UNSUPPORTED).
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "clear_amd.hpp"
#include "clear_amd_scenes.h"

namespace {

void hipOk(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Device buffer of `bytes` (zeroed), filled from `host` when it is not null.
template <class T>
T* upload(const void* host, size_t bytes) {
    void* d = nullptr;
    const size_t alloc = bytes < 16 ? 16 : bytes;  // empty arrays become small dummies, as the reference's host does
    hipOk(hipMalloc(&d, alloc), "hipMalloc");
    hipOk(hipMemset(d, 0, alloc), "hipMemset");
    if (host && bytes) hipOk(hipMemcpy(d, host, bytes, hipMemcpyHostToDevice), "upload");
    return static_cast<T*>(d);
}

void save(const std::string& path, const void* dev, size_t bytes) {
    std::vector<char> h(bytes);
    hipOk(hipMemcpy(h.data(), dev, bytes, hipMemcpyDeviceToHost), "download");
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(h.data(), 1, bytes, f) != bytes) throw std::runtime_error("cannot write " + path);
    std::fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s <output directory>\n", argv[0]);
        return 2;
    }
    const std::string out = argv[1];
    try {
        clr::Context ctx(0);

        // 1. LBVH build from vertices + faces through LBVHBuilder.
        const uint32_t n = 50000;
        std::vector<float> soup(9 * (size_t)n);
        if (clr_gen_cluster_soup(n, 7, 64, 0.02f, 1e-3f, 1e-2f, soup.data()) != CLR_OK) throw std::runtime_error("soup");
        std::vector<clr_float3> verts(3 * (size_t)n);
        std::vector<clr_face> faces(n);
        for (uint32_t i = 0; i < n; ++i) {
            for (int k = 0; k < 3; ++k)
                verts[3 * i + k] = {soup[9 * i + 3 * k], soup[9 * i + 3 * k + 1], soup[9 * i + 3 * k + 2], 0.0f};
            clr_face& f = faces[i];
            std::memset(&f, 0xFF, sizeof(f));
            f.p0 = 3 * i;
            f.p1 = 3 * i + 1;
            f.p2 = 3 * i + 2;
        }
        auto* dv = upload<clr_float3>(verts.data(), verts.size() * sizeof(clr_float3));
        auto* df = upload<clr_face>(faces.data(), faces.size() * sizeof(clr_face));
        auto* di = upload<clr_lbvh_inode>(nullptr, (n - 1) * sizeof(clr_lbvh_inode));
        auto* dl = upload<clr_lbvh_leaf>(nullptr, n * sizeof(clr_lbvh_leaf));
        clr::LBVHBuilder builder(ctx, n);
        builder.perform(nullptr, dv, df, n, 10, di, dl, true);
        const std::vector<float> stages = builder.lastProfile();
        if (stages.size() != 5) throw std::runtime_error("expected 5 profiled stages");
        save(out + "/lbvh_inodes.bin", di, (n - 1) * sizeof(clr_lbvh_inode));
        save(out + "/lbvh_leaves.bin", dl, n * sizeof(clr_lbvh_leaf));
        std::printf("lbvh ok, stages (ms):");
        for (float s : stages) std::printf(" %.4f", s);
        std::printf("\n");

        // 2. Cornell box through LBVHBuilder and clr::render.
        clr_scene_params sp{};
        sp.width = sp.height = 48;
        sp.detail = 3;
        clr_host_scene* hs = nullptr;
        if (clr_scene_build(CLR_SCENE_CORNELL, &sp, &hs) != CLR_OK) throw std::runtime_error("scene build");
        uint32_t nf = 0, nl = 0, w = 0, h = 0;
        clr_scene_info(hs, &nf, &nl, &w, &h);
        void* arr[9];
        for (int a = 0; a < 9; ++a) {
            const void* data = nullptr;
            uint64_t bytes = 0;
            clr_scene_array(hs, a, &data, &bytes);
            arr[a] = upload<uint8_t>(data, bytes);
        }
        clr_scene_free(hs);
        auto* sin = upload<clr_lbvh_inode>(nullptr, (nf - 1) * sizeof(clr_lbvh_inode));
        auto* slf = upload<clr_lbvh_leaf>(nullptr, nf * sizeof(clr_lbvh_leaf));
        clr::LBVHBuilder sceneBuilder(ctx, nf);
        sceneBuilder.perform(nullptr, static_cast<clr_float3*>(arr[CLR_ARRAY_VERTICES]),
                             static_cast<clr_face*>(arr[CLR_ARRAY_FACES]), nf, 10, sin, slf);
        clr_scene_dev scene{};
        scene.vertices = static_cast<const clr_float3*>(arr[CLR_ARRAY_VERTICES]);
        scene.normals = static_cast<const clr_float3*>(arr[CLR_ARRAY_NORMALS]);
        scene.tangents = static_cast<const clr_float3*>(arr[CLR_ARRAY_TANGENTS]);
        scene.uvs = static_cast<const float*>(arr[CLR_ARRAY_UVS]);
        scene.faces = static_cast<const clr_face*>(arr[CLR_ARRAY_FACES]);
        scene.lights = static_cast<const clr_light_info*>(arr[CLR_ARRAY_LIGHTS]);
        scene.materialsData = static_cast<const uint8_t*>(arr[CLR_ARRAY_MATERIALS]);
        scene.texturesData = static_cast<const uint8_t*>(arr[CLR_ARRAY_TEXTURES]);
        scene.otherResources = static_cast<const uint8_t*>(arr[CLR_ARRAY_OTHERS]);
        scene.internalNodes = sin;
        scene.leafNodes = slf;
        scene.numFaces = nf;
        scene.numLights = nl;
        const clr_render_params rp{w, h, 4, 5, 0};
        auto* states = upload<uint32_t>(nullptr, (size_t)w * h * 16);
        for (int mode : {CLR_RENDER_MEGAKERNEL, CLR_RENDER_WAVEFRONT}) {
            auto* pixels = upload<clr_float3>(nullptr, (size_t)w * h * sizeof(clr_float3));
            clr::seedPixelStates(ctx, w, h, states);
            clr::render(ctx, scene, rp, states, pixels, mode);
            save(out + (mode == CLR_RENDER_MEGAKERNEL ? "/render_mega.bin" : "/render_wave.bin"), pixels,
                 (size_t)w * h * sizeof(clr_float3));
            hipOk(hipFree(pixels), "hipFree");
        }
        if (ctx.deviceStatus() != 0) throw std::runtime_error("device status raised");
        std::printf("render ok (%u faces, %ux%u)\n", nf, w, h);

        // 3. Errors surface as clr::Error.
        try {
            builder.perform(nullptr, dv, df, n, 0, di, dl);
            throw std::runtime_error("numBitsPerDim 0 was accepted");
        } catch (const clr::Error& e) {
            if (e.status() != CLR_ERR_UNSUPPORTED) throw;
            std::printf("error path ok: %s\n", e.what());
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "wrapper_driver failed: %s\n", e.what());
        return 1;
    }
    std::printf("wrapper driver ok\n");
    return 0;
}
