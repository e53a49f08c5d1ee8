// clear_amd_loaders.hpp — host-side asset ingest (header-only C++17; link with -lz).
//
// Drop-in for the reference's model and image loaders:
//   loadModel  (OpenCL_PathTracer/ModelLoader.hpp:15, ModelLoader.cpp:15-508): OBJ with MTL
//              materials into a Scene;
//   loadImage  (ImageLoader.h:15-24, ImageLoader.cpp:27-232): PNG (the reference uses libpng)
//              and OpenEXR (the reference uses Imf::RgbaInputFile);
//   MaterialCreator::createImageTexture / createNormalMapTexture from files
//              (CreateFunctions.cpp:38-86).
// PNG is decoded with zlib alone, applying the transforms the reference asks libpng for
// (3-channel rows padded with a 0xFF filler byte, 16-bit samples stripped to their high
// byte, sBIT shifts, and libpng's gamma table for the requested screen gamma). EXR scanline
// files are read directly: NONE, RLE, ZIPS and ZIP compression, HALF and FLOAT channels,
// returned as RGBA half floats (missing colour channels 0, missing alpha 1) like RgbaInputFile.
// Not supported (an error is returned): interlaced, palette and grey+alpha PNGs and sample
// depths below 8 bits (the reference asserts on the middle two); tiled, deep or multi-part
// EXR files and the PIZ, PXR24, B44 and DWA codecs.
#pragma once

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "clear_amd_scene.hpp"

namespace clr {
namespace host {

namespace ColorChannel {  // ImageLoader.h:15-21
enum Value { RGB8x3 = 0, RGBA8x4, RGBA16Fx4, Gray8 };
}

namespace loaders_detail {

inline bool readWholeFile(const char* fileName, std::vector<uint8_t>& out) {
    FILE* fp = std::fopen(fileName, "rb");
    if (!fp) return false;
    std::fseek(fp, 0, SEEK_END);
    const long size = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    out.resize(size > 0 ? (size_t)size : 0);
    const size_t got = size > 0 ? std::fread(out.data(), 1, (size_t)size, fp) : 0;
    std::fclose(fp);
    return got == out.size();
}

inline uint32_t be32(const uint8_t* p) {
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}
inline uint32_t le32(const uint8_t* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
inline uint64_t le64(const uint8_t* p) { return (uint64_t)le32(p) | ((uint64_t)le32(p + 4) << 32); }

// zlib stream -> exactly `expect` bytes.
inline bool inflateAll(const uint8_t* src, size_t n, std::vector<uint8_t>& out, size_t expect) {
    out.assign(expect, 0);
    z_stream zs{};
    if (inflateInit(&zs) != Z_OK) return false;
    zs.next_in = const_cast<Bytef*>(src);
    zs.avail_in = (uInt)n;
    zs.next_out = out.data();
    zs.avail_out = (uInt)expect;
    const int rc = inflate(&zs, Z_FINISH);
    const bool ok = (rc == Z_STREAM_END || rc == Z_OK || rc == Z_BUF_ERROR) && zs.total_out == expect;
    inflateEnd(&zs);
    return ok;
}

// libpng's fixed-point gamma arithmetic (png.c: png_product2, png_reciprocal2,
// png_gamma_significant, png_gamma_8bit_correct, png_gamma_16bit_correct), values in
// units of 1e-5, restated from its published source.
inline int32_t pngFixed(double v) { return (int32_t)std::floor(v * 100000.0 + 0.5); }
inline bool pngGammaSignificant(int32_t g) { return g < 100000 - 5000 || g > 100000 + 5000; }
inline int32_t pngProduct2(int32_t a, int32_t b) {
    double r = a * 1e-5;
    r *= b;
    return (int32_t)std::floor(r + 0.5);
}
inline int32_t pngReciprocal2(int32_t a, int32_t b) {
    double r = 1e15 / a;
    r /= b;
    return (int32_t)std::floor(r + 0.5);
}
inline uint32_t pngGamma8(uint32_t v, int32_t g) {
    if (v > 0 && v < 255) return (uint32_t)std::floor(255.0 * std::pow((int)v / 255.0, g * 0.00001) + 0.5);
    return v;
}
inline uint32_t pngGamma16(uint32_t v, int32_t g) {
    if (v > 0 && v < 65535) return (uint32_t)std::floor(65535.0 * std::pow(v / 65535.0, g * 0.00001) + 0.5);
    return v;
}

inline uint32_t paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    return (uint32_t)(pa <= pb && pa <= pc ? a : (pb <= pc ? b : c));
}

// IEEE half <-> float, exact (half -> float) and round-to-nearest-even (float -> half).
inline float halfToFloat(uint16_t h) {
    const uint32_t s = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 0x1F, m = h & 0x3FF;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) {
            bits = s;
        } else {  // subnormal: normalise
            int ee = -1;
            uint32_t mm = m;
            do {
                ++ee;
                mm <<= 1;
            } while ((mm & 0x400) == 0);
            bits = s | ((uint32_t)(127 - 15 - ee) << 23) | ((mm & 0x3FF) << 13);
        }
    } else if (e == 31) {
        bits = s | 0x7F800000u | (m << 13);
    } else {
        bits = s | ((e + 112) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}
inline uint16_t floatToHalf(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t s = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7FFFFFFFu;
    if (ax >= 0x7F800000u) return (uint16_t)(s | 0x7C00u | (ax > 0x7F800000u ? 0x200u | ((ax >> 13) & 0x3FFu) : 0u));
    if (ax >= 0x477FF000u) return (uint16_t)(s | 0x7C00u);  // rounds to >= 65520: infinity
    if (ax < 0x38800000u) {                                 // result subnormal or zero
        if (ax < 0x33000000u) return (uint16_t)s;          // below half the smallest subnormal
        const uint32_t e = ax >> 23, m = (ax & 0x7FFFFFu) | 0x800000u;
        const uint32_t shift = 126 - e;  // 14 + (113 - e)
        uint32_t r = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (r & 1u))) ++r;
        return (uint16_t)(s | r);
    }
    uint32_t r = ((ax >> 13) - (112u << 10));
    const uint32_t rem = ax & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) ++r;
    return (uint16_t)(s | r);
}

}  // namespace loaders_detail

// ---- PNG (ImageLoader.cpp:72-201) ----------------------------------------------------------
// Rows top to bottom; RGB -> 4 bytes per pixel (R, G, B, 0xFF), RGBA -> 4, grey -> 1.
// gammaCorrection = false: screen gamma 1 (so a file gamma of 0.45455, the default when the
// file has no gAMA chunk, is undone: samples come out linear); true: screen gamma 2.2 or the
// SCREEN_GAMMA environment variable (samples stay as stored).
inline bool loadPNG(const char* fileName, std::vector<uint8_t>* storage, uint32_t* width, uint32_t* height,
                    ColorChannel::Value* color, bool gammaCorrection, std::string* error = nullptr) {
    using namespace loaders_detail;
    auto fail = [&](const char* what) {
        if (error) *error = std::string(fileName) + ": " + what;
        return false;
    };
    std::vector<uint8_t> f;
    if (!readWholeFile(fileName, f)) return fail("cannot read the file");
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) return fail("not a PNG file");
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    bool haveGamma = false, haveSBit = false;
    uint32_t fileGammaFixed = 0;
    uint8_t sbit[4] = {0, 0, 0, 0};
    std::vector<uint8_t> idat;
    size_t pos = 8;
    bool end = false;
    while (!end) {
        if (pos + 12 > f.size()) return fail("truncated chunk");
        const uint32_t len = be32(&f[pos]);
        if (len > f.size() - pos - 12) return fail("truncated chunk");
        const uint8_t* type = &f[pos + 4];
        const uint8_t* d = &f[pos + 8];
        if ((uint32_t)crc32(crc32(0L, Z_NULL, 0), type, len + 4) != be32(d + len)) return fail("chunk CRC mismatch");
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len != 13) return fail("bad IHDR");
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
            if (d[10] != 0 || d[11] != 0) return fail("unknown compression or filter method");
        } else if (!std::memcmp(type, "gAMA", 4) && len == 4) {
            haveGamma = true;
            fileGammaFixed = be32(d);
        } else if (!std::memcmp(type, "sBIT", 4) && len >= 1 && len <= 4) {
            haveSBit = true;
            for (uint32_t k = 0; k < len; ++k) sbit[k] = d[k];
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), d, d + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            end = true;
        } else if (!std::memcmp(type, "PLTE", 4) || (type[0] & 0x20) == 0) {
            if (std::memcmp(type, "PLTE", 4)) return fail("unknown critical chunk");
        }
        pos += 12 + len;
    }
    if (ctype < 0) return fail("no IHDR");
    if (w == 0 || h == 0 || w > 5120 || h > 5120) return fail("image size outside 1..5120 (png_set_user_limits)");
    if (ctype == 3 || ctype == 4) return fail("palette and grey+alpha images are not supported");
    if (ctype != 0 && ctype != 2 && ctype != 6) return fail("unknown colour type");
    if (interlace != 0) return fail("interlaced images are not supported");
    if (depth != 8 && depth != 16) return fail("sample depths other than 8 and 16 bits are not supported");
    const int ch = ctype == 0 ? 1 : ctype == 2 ? 3 : 4;
    const size_t bps = (size_t)depth / 8, bpp = ch * bps, rowBytes = (size_t)w * bpp;
    std::vector<uint8_t> raw;
    if (!inflateAll(idat.data(), idat.size(), raw, (rowBytes + 1) * h)) return fail("corrupt image data");
    // Undo the per-row filters.
    std::vector<uint8_t> px(rowBytes * h);
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t ft = raw[y * (rowBytes + 1)];
        const uint8_t* src = &raw[y * (rowBytes + 1) + 1];
        uint8_t* dst = &px[y * rowBytes];
        const uint8_t* up = y ? &px[(y - 1) * rowBytes] : nullptr;
        for (size_t i = 0; i < rowBytes; ++i) {
            const uint32_t a = i >= bpp ? dst[i - bpp] : 0, b = up ? up[i] : 0, c = (up && i >= bpp) ? up[i - bpp] : 0;
            uint32_t v = src[i];
            switch (ft) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) >> 1; break;
                case 4: v += paeth((int)a, (int)b, (int)c); break;
                default: return fail("unknown row filter");
            }
            dst[i] = (uint8_t)v;
        }
    }
    // Transforms in libpng's order (pngrtran.c png_do_read_transformations): gamma on the
    // colour samples, 16 -> 8 by the high byte, sBIT unshift, then the filler byte.
    double screenGamma = 1.0;
    if (gammaCorrection) {
        const char* gs = std::getenv("SCREEN_GAMMA");
        screenGamma = gs ? std::atof(gs) : 2.2;
    }
    const int32_t screenFixed = pngFixed(screenGamma);
    const int32_t fileFixed = haveGamma ? (int32_t)fileGammaFixed : pngFixed(0.45455);
    const bool doGamma = fileFixed > 0 && pngGammaSignificant(pngProduct2(fileFixed, screenFixed));
    const int32_t g = doGamma ? pngReciprocal2(fileFixed, screenFixed) : 100000;
    const bool gammaTable = doGamma && pngGammaSignificant(g);
    int shift[4] = {0, 0, 0, 0};
    if (haveSBit)
        for (int c = 0; c < ch; ++c) shift[c] = sbit[c] > 0 && sbit[c] < 8 ? 8 - sbit[c] : 0;
    const int outStride = ch == 1 ? 1 : 4;
    const size_t base = storage->size();
    storage->resize(base + (size_t)w * h * outStride);
    uint8_t* out = storage->data() + base;
    for (size_t p = 0; p < (size_t)w * h; ++p) {
        for (int c = 0; c < ch; ++c) {
            const uint8_t* s = &px[p * bpp + c * bps];
            const bool alpha = ch == 4 && c == 3;
            uint32_t v;
            if (depth == 8) {
                v = s[0];
                if (gammaTable && !alpha) v = pngGamma8(v, g);
            } else {
                v = ((uint32_t)s[0] << 8) | s[1];
                if (gammaTable && !alpha) v = pngGamma16(v, g);
                v >>= 8;
            }
            out[p * outStride + c] = (uint8_t)(v >> shift[c]);
        }
        if (ch == 3) out[p * outStride + 3] = 0xFF;
    }
    *width = w;
    *height = h;
    *color = ch == 1 ? ColorChannel::Gray8 : ch == 3 ? ColorChannel::RGB8x3 : ColorChannel::RGBA8x4;
    return true;
}

// ---- OpenEXR (ImageLoader.cpp:27-56) -----------------------------------------------------
// Scanline images: RGBA half floats, 8 bytes per pixel, rows from the data window's top.
inline bool loadEXR(const char* fileName, std::vector<uint8_t>* storage, uint32_t* width, uint32_t* height,
                    std::string* error = nullptr) {
    using namespace loaders_detail;
    auto fail = [&](const char* what) {
        if (error) *error = std::string(fileName) + ": " + what;
        return false;
    };
    std::vector<uint8_t> f;
    if (!readWholeFile(fileName, f)) return fail("cannot read the file");
    if (f.size() < 8 || le32(f.data()) != 20000630u) return fail("not an OpenEXR file");
    const uint32_t ver = le32(&f[4]);
    if ((ver & 0xFF) != 2) return fail("unsupported OpenEXR version");
    if (ver & (0x200u | 0x800u | 0x1000u)) return fail("tiled, deep and multi-part files are not supported");
    struct Chan { std::string name; int type; };
    std::vector<Chan> chans;
    int compression = -1;
    int32_t box[4] = {0, 0, -1, -1};
    size_t pos = 8;
    auto cstr = [&](std::string& s) {
        const size_t e = std::find(f.begin() + (long)pos, f.end(), (uint8_t)0) - f.begin();
        if (e >= f.size()) return false;
        s.assign((const char*)&f[pos], e - pos);
        pos = e + 1;
        return true;
    };
    while (true) {
        std::string name, type;
        if (!cstr(name)) return fail("truncated header");
        if (name.empty()) break;
        if (!cstr(type) || pos + 4 > f.size()) return fail("truncated header");
        const uint32_t size = le32(&f[pos]);
        pos += 4;
        if (size > f.size() - pos) return fail("truncated header");
        const uint8_t* v = &f[pos];
        if (name == "channels" && type == "chlist") {
            size_t q = 0;
            while (q < size && v[q] != 0) {
                const size_t e = std::find(v + q, v + size, (uint8_t)0) - v;
                if (e + 17 > size) return fail("bad channel list");
                Chan c{std::string((const char*)v + q, e - q), (int)le32(v + e + 1)};
                if (le32(v + e + 9) != 1 || le32(v + e + 13) != 1) return fail("subsampled channels are not supported");
                chans.push_back(c);
                q = e + 17;
            }
        } else if (name == "compression" && size >= 1) {
            compression = v[0];
        } else if (name == "dataWindow" && size >= 16) {
            for (int k = 0; k < 4; ++k) box[k] = (int32_t)le32(v + 4 * k);
        }
        pos += size;
    }
    if (box[2] < box[0] || box[3] < box[1]) return fail("empty data window");
    // Window extents in 64-bit arithmetic (int32 corners can be 2^32 apart),
    // capped like the PNG reader's limit.
    const int64_t w64 = (int64_t)box[2] - box[0] + 1, h64 = (int64_t)box[3] - box[1] + 1;
    if (w64 > 16384 || h64 > 16384) return fail("image size outside 1..16384");
    const uint32_t w = (uint32_t)w64, h = (uint32_t)h64;
    if (compression < 0 || compression > 3) return fail("only NONE, RLE, ZIPS and ZIP compression are supported");
    const uint32_t linesPerChunk = compression == 3 ? 16u : 1u;
    size_t lineBytes = 0;
    for (const Chan& c : chans) {
        if (c.type != 1 && c.type != 2) return fail("only HALF and FLOAT channels are supported");
        lineBytes += (size_t)w * (c.type == 1 ? 2 : 4);
    }
    const uint32_t chunks = (h + linesPerChunk - 1) / linesPerChunk;
    if (pos + 8ull * chunks > f.size()) return fail("truncated offset table");
    const size_t tableAt = pos;
    // Destination: RGBA halves; defaults R = G = B = 0, A = 1 (RgbaInputFile).
    const size_t base = storage->size();
    storage->resize(base + (size_t)w * h * 8);
    uint16_t* out = reinterpret_cast<uint16_t*>(storage->data() + base);
    for (size_t p = 0; p < (size_t)w * h; ++p) {
        out[4 * p] = out[4 * p + 1] = out[4 * p + 2] = 0;
        out[4 * p + 3] = 0x3C00;
    }
    std::vector<uint8_t> tmp, unpacked;
    for (uint32_t ci = 0; ci < chunks; ++ci) {
        const uint64_t off = le64(&f[tableAt + 8ull * ci]);
        if (f.size() < 8 || off > f.size() - 8) return fail("bad chunk offset");  // no wrap near 2^64
        const int32_t y0 = (int32_t)le32(&f[off]);
        const uint32_t dsize = le32(&f[off + 4]);
        if (dsize > f.size() - off - 8) return fail("truncated chunk");
        if (y0 < box[1] || y0 > box[3]) return fail("chunk outside the data window");
        const uint32_t lines = (uint32_t)std::min<int64_t>(linesPerChunk, (int64_t)box[3] - y0 + 1);
        const size_t expect = lineBytes * lines;
        const uint8_t* d = &f[off + 8];
        const uint8_t* data = d;
        if (dsize != expect) {  // compressed (data that would not shrink is stored as is)
            if (compression == 0) return fail("chunk size mismatch");
            if (compression == 1) {  // RLE: a negative count n copies -n bytes, else n + 1 repeats
                tmp.clear();
                size_t q = 0;
                while (q < dsize) {
                    const int8_t n = (int8_t)d[q++];
                    if (n < 0) {
                        if (q + (size_t)(-n) > dsize) return fail("corrupt RLE data");
                        tmp.insert(tmp.end(), d + q, d + q + (-n));
                        q += (size_t)(-n);
                    } else {
                        if (q >= dsize) return fail("corrupt RLE data");
                        tmp.insert(tmp.end(), (size_t)n + 1, d[q++]);
                    }
                }
                if (tmp.size() != expect) return fail("corrupt RLE data");
            } else if (!inflateAll(d, dsize, tmp, expect)) {
                return fail("corrupt ZIP data");
            }
            // Predictor, then the two byte halves interleaved (ImfZip / ImfRle).
            for (size_t q = 1; q < tmp.size(); ++q) tmp[q] = (uint8_t)(tmp[q - 1] + tmp[q] - 128);
            unpacked.resize(expect);
            const size_t halfN = (expect + 1) / 2;
            for (size_t q = 0; q < expect; ++q) unpacked[q] = (q & 1) ? tmp[halfN + q / 2] : tmp[q / 2];
            data = unpacked.data();
        }
        // Per line, per channel (in the list's order), w samples.
        size_t q = 0;
        for (uint32_t l = 0; l < lines; ++l) {
            const size_t row = (size_t)(y0 - box[1] + (int32_t)l) * w;
            for (const Chan& c : chans) {
                const int slot = c.name == "R" ? 0 : c.name == "G" ? 1 : c.name == "B" ? 2 : c.name == "A" ? 3 : -1;
                for (uint32_t x = 0; x < w; ++x) {
                    uint16_t hv;
                    if (c.type == 1) {
                        hv = (uint16_t)(data[q] | (data[q + 1] << 8));
                        q += 2;
                    } else {
                        float fv;
                        uint32_t bits = le32(data + q);
                        std::memcpy(&fv, &bits, 4);
                        hv = floatToHalf(fv);
                        q += 4;
                    }
                    if (slot >= 0) out[4 * (row + x) + slot] = hv;
                }
            }
        }
    }
    *width = w;
    *height = h;
    return true;
}

// Dispatch on the extension (ImageLoader.cpp:208-232). The reference cuts the extension at
// the first '.' of the path; here at the last, so that paths such as "../a.png" work.
inline bool loadImage(const char* fileName, std::vector<uint8_t>* storage, uint32_t* width, uint32_t* height,
                      ColorChannel::Value* color, bool gammaCorrection, std::string* error = nullptr) {
    const std::string fn = fileName;
    const size_t dot = fn.find_last_of('.');
    const std::string ext = dot == std::string::npos ? "" : fn.substr(dot + 1);
    if (ext == "png") return loadPNG(fileName, storage, width, height, color, gammaCorrection, error);
    if (ext == "exr") {
        *color = ColorChannel::RGBA16Fx4;
        return loadEXR(fileName, storage, width, height, error);
    }
    if (error) *error = fn + ": unsupported image format (png and exr are read)";
    return false;
}

// MaterialCreator::createImageTexture(name, filename) (CreateFunctions.cpp:38-69): a file
// already loaded into this scene's textures is shared.
inline void createImageTexture(MaterialCreator& mc, Scene& scene, const char* name, const char* filename) {
    (void)mc;
    uint64_t texHead = 0;
    if (scene.texFilehasLoaded(filename, &texHead)) {
        scene.addTexture(texHead, name);
        return;
    }
    std::vector<uint8_t>* td = &scene.texturesData;
    texHead = fillZerosAligned(td, sizeof(ImageTexture), 4);
    const uint64_t imageHead = alignBlob(td, 128);
    uint32_t w = 0, h = 0;
    ColorChannel::Value ct;
    std::string err;
    if (!loadImage(filename, td, &w, &h, &ct, false, &err)) throw std::runtime_error(err);
    ImageTexture* im = reinterpret_cast<ImageTexture*>(td->data() + texHead);
    im->width = w;
    im->height = h;
    im->texType = ct == ColorChannel::RGB8x3     ? TextureType::ColorImageRGB8x3
                  : ct == ColorChannel::RGBA8x4  ? TextureType::ColorImageRGBA8x4
                  : ct == ColorChannel::RGBA16Fx4 ? TextureType::ColorImageRGBA16Fx4
                                                 : TextureType::GrayImage8;
    im->offsetData = (int32_t)imageHead - (int32_t)texHead;
    scene.addTexture(texHead, name);
    scene.addTexFileToDB(texHead, filename);
}

// MaterialCreator::createNormalMapTexture(name, filename) (CreateFunctions.cpp:71-86): the
// image is read with gamma correction on (samples as stored) and tagged RGB8.
inline void createNormalMapTexture(MaterialCreator& mc, Scene& scene, const char* name, const char* filename) {
    (void)mc;
    std::vector<uint8_t>* td = &scene.texturesData;
    const uint64_t texHead = fillZerosAligned(td, sizeof(ImageTexture), 4);
    const uint64_t imageHead = alignBlob(td, 128);
    uint32_t w = 0, h = 0;
    ColorChannel::Value ct;
    std::string err;
    if (!loadImage(filename, td, &w, &h, &ct, true, &err)) throw std::runtime_error(err);
    ImageTexture* im = reinterpret_cast<ImageTexture*>(td->data() + texHead);
    im->texType = TextureType::ColorImageRGB8x3;
    im->width = w;
    im->height = h;
    im->offsetData = (int32_t)imageHead - (int32_t)texHead;
    scene.addTexture(texHead, name);
}

// ---- OBJ / MTL (ModelLoader.cpp:36-508) -------------------------------------------------
namespace obj_detail {

struct Material {
    std::string name;
    float Kd[3] = {0, 0, 0}, Ks[3] = {0, 0, 0}, Ka[3] = {0, 0, 0};
    float Ns = 100.0f, Ni = 1.0f;
    std::string mapKd, mapKs, mapKa, mapBump;
};
struct Group {
    enum Format { Unknown, VTN, VN };
    std::string name, material;
    Format format = Unknown;
    uint32_t smooth = 0;
    std::vector<uint32_t> pos, nrm, uv;
};
struct Object {
    std::string name;
    std::vector<float> positions, normals, uvs;
    std::vector<Group> groups;
};

// The reference strips leading bytes of the string "　 ¥t" (UTF-8: E3 80 80, space, C2 A5,
// 't'; ModelLoader.cpp:106,214) before looking at a line's first character.
inline void stripLead(std::string& s) {
    static const char kLead[] = "\xE3\x80\x80 \xC2\xA5t";
    s.erase(0, s.find_first_not_of(kLead));
}
inline std::string dirOf(const std::string& path) { return path.substr(0, path.find_last_of('/') + 1); }

inline bool loadMtl(const std::string& fileName, std::vector<Material>* materials,
                    std::map<std::string, uint32_t>* indices) {
    FILE* fp = std::fopen(fileName.c_str(), "rb");
    if (!fp) return false;
    const std::string directory = dirOf(fileName);
    Material* cur = nullptr;
    char buf[4096];
    while (std::fgets(buf, sizeof buf, fp)) {
        std::string line(buf);
        while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
        stripLead(line);
        if (line.empty()) continue;
        switch (line[0]) {
            case 'K': {
                if (line.size() < 2 || !cur) break;
                char fmt[32];
                std::snprintf(fmt, sizeof fmt, "K%c %%f %%f %%f", line[1]);
                float r, g, b;
                if (std::sscanf(line.c_str(), fmt, &r, &g, &b) != 3) r = g = b = 0.0f;
                float* dst = line[1] == 'd' ? cur->Kd : line[1] == 's' ? cur->Ks : line[1] == 'a' ? cur->Ka : nullptr;
                if (dst) {
                    dst[0] = r;
                    dst[1] = g;
                    dst[2] = b;
                }
                break;
            }
            case 'N': {
                float v;
                if (cur && std::sscanf(line.c_str(), "Ni %f", &v) == 1) cur->Ni = v;
                break;
            }
            case 'm': {
                char tex[256];
                if (!cur) break;
                if (std::sscanf(line.c_str(), "map_Kd %255s", tex) == 1) cur->mapKd = directory + tex;
                else if (std::sscanf(line.c_str(), "map_Ks %255s", tex) == 1) cur->mapKs = directory + tex;
                else if (std::sscanf(line.c_str(), "map_Ka %255s", tex) == 1) cur->mapKa = directory + tex;
                else if (std::sscanf(line.c_str(), "map_bump %255s", tex) == 1) cur->mapKd = directory + tex;  // as the reference
                break;
            }
            case 'n': {
                char nm[256];
                if (std::sscanf(line.c_str(), "newmtl %255s", nm) != 1) break;
                indices->emplace(nm, (uint32_t)materials->size());
                materials->emplace_back();
                cur = &materials->back();
                cur->name = nm;
                break;
            }
            default:
                break;
        }
    }
    std::fclose(fp);
    return true;
}

}  // namespace obj_detail

// loadModel(fileName, scene, matOverride, lightOverride) (ModelLoader.hpp:15): OBJ files
// only. Materials: Ni == 1 gives a matte material whose reflectance is map_Kd (image) or Kd,
// otherwise glass (R = T = 0.95, etaExt = Ni, etaInt = 1), named "<model number>_<name>"
// with the model number counted per scene (the reference counts per process). Faces take
// "p/t/n" or "p//n" triples (the group's first face fixes the form; other forms are skipped)
// and only the first three vertices of a polygon; "vt" v is flipped (1 - v); a group whose
// material is unknown takes the first material, and a file without materials gets a grey
// matte one. matOverride / lightOverride (byte offsets) replace every face's material /
// light property. Returns false (with `error`) when the file cannot be read.
inline bool loadModel(const char* fileName, Scene* scene, uint64_t matOverride = UINT64_MAX,
                      uint64_t lightOverride = UINT64_MAX, std::string* error = nullptr) {
    using namespace obj_detail;
    const std::string sFileName = fileName;
    const size_t dot = sFileName.find_last_of('.');
    if (dot == std::string::npos || sFileName.substr(dot + 1) != "obj") {
        if (error) *error = sFileName + ": not an .obj file";
        return false;
    }
    FILE* fp = std::fopen(fileName, "rb");
    if (!fp) {
        if (error) *error = sFileName + ": cannot read the file";
        return false;
    }
    const std::string directory = dirOf(sFileName);
    MaterialCreator& mc = MaterialCreator::sharedInstance();
    mc.setScene(scene);

    std::vector<Object> objects;
    std::vector<Material> materials;
    std::map<std::string, uint32_t> matIdx;
    Group* curGroup = nullptr;
    std::string curGroupName, curMatName;
    uint32_t smooth = 0;
    bool phaseF = false;
    auto obj = [&]() -> Object& {
        if (objects.empty()) objects.emplace_back();  // vertices before any "o" line
        return objects.back();
    };
    char buf[4096];
    while (std::fgets(buf, sizeof buf, fp)) {
        std::string line(buf);
        while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
        stripLead(line);
        if (line.empty()) {
            phaseF = false;
            continue;
        }
        if (phaseF && line[0] != 'f') phaseF = false;
        switch (line[0]) {
            case 'o': {
                char nm[256];
                if (std::sscanf(line.c_str(), "o %255s", nm) != 1) break;
                objects.emplace_back();
                objects.back().name = nm;
                break;
            }
            case 'g': {
                char nm[256];
                if (std::sscanf(line.c_str(), "g %255s", nm) == 1) curGroupName = nm;
                break;
            }
            case 's': {
                if (line.compare(0, 5, "s off") == 0) {
                    smooth = 0;
                    break;
                }
                unsigned s;
                if (std::sscanf(line.c_str(), "s %u", &s) == 1) smooth = s;
                break;
            }
            case 'v': {
                if (line.size() < 2) break;
                float a = 0, b = 0, c = 0;
                if (line[1] == ' ') {
                    std::sscanf(line.c_str(), "v %f %f %f", &a, &b, &c);
                    Object& o = obj();
                    o.positions.insert(o.positions.end(), {a, b, c});
                } else if (line[1] == 't') {
                    std::sscanf(line.c_str(), "vt %f %f", &a, &b);
                    Object& o = obj();
                    o.uvs.insert(o.uvs.end(), {a, 1.0f - b});
                } else if (line[1] == 'n') {
                    std::sscanf(line.c_str(), "vn %f %f %f", &a, &b, &c);
                    Object& o = obj();
                    o.normals.insert(o.normals.end(), {a, b, c});
                }
                break;
            }
            case 'f': {
                unsigned p[3], t[3], n[3];
                const char* vtn = "f %u/%u/%u %u/%u/%u %u/%u/%u";
                const char* vn = "f %u//%u %u//%u %u//%u";
                if (!phaseF) {
                    phaseF = true;
                    Object& o = obj();
                    o.groups.emplace_back();
                    curGroup = &o.groups.back();
                    curGroup->name = curGroupName;
                    curGroup->smooth = smooth;
                    curGroup->material = curMatName;
                    if (std::sscanf(line.c_str(), vtn, &p[0], &t[0], &n[0], &p[1], &t[1], &n[1], &p[2], &t[2], &n[2]) == 9)
                        curGroup->format = Group::VTN;
                    else if (std::sscanf(line.c_str(), vn, &p[0], &n[0], &p[1], &n[1], &p[2], &n[2]) == 6)
                        curGroup->format = Group::VN;
                }
                if (curGroup->format == Group::VTN &&
                    std::sscanf(line.c_str(), vtn, &p[0], &t[0], &n[0], &p[1], &t[1], &n[1], &p[2], &t[2], &n[2]) == 9) {
                    for (int k = 0; k < 3; ++k) {
                        curGroup->pos.push_back(p[k] - 1);
                        curGroup->uv.push_back(t[k] - 1);
                        curGroup->nrm.push_back(n[k] - 1);
                    }
                } else if (curGroup->format == Group::VN &&
                           std::sscanf(line.c_str(), vn, &p[0], &n[0], &p[1], &n[1], &p[2], &n[2]) == 6) {
                    for (int k = 0; k < 3; ++k) {
                        curGroup->pos.push_back(p[k] - 1);
                        curGroup->nrm.push_back(n[k] - 1);
                    }
                }
                break;
            }
            case 'm': {
                char nm[256];
                if (line.compare(0, 6, "mtllib") == 0 && std::sscanf(line.c_str(), "mtllib %255s", nm) == 1)
                    loadMtl(directory + nm, &materials, &matIdx);
                break;
            }
            case 'u': {
                char nm[256];
                if (line.compare(0, 6, "usemtl") == 0 && std::sscanf(line.c_str(), "usemtl %255s", nm) == 1)
                    curMatName = nm;
                break;
            }
            default:
                break;
        }
    }
    std::fclose(fp);

    char prefix[32];
    std::snprintf(prefix, sizeof prefix, "%05llu", (unsigned long long)scene->numModelsLoaded);
    if (materials.empty()) {  // no usable material: a grey matte one
        materials.emplace_back();
        materials.back().name = "default";
        materials.back().Kd[0] = materials.back().Kd[1] = materials.back().Kd[2] = 0.5f;
    }
    for (const Material& m : materials) {
        const std::string matName = std::string(prefix) + "_" + m.name;
        if (m.Ni == 1.0f) {
            const std::string kd = matName + "_Kd", sigma = matName + "_Kd_sigma";
            if (!m.mapKd.empty()) createImageTexture(mc, *scene, kd.c_str(), m.mapKd.c_str());
            else mc.createFloat3ConstantTexture(kd.c_str(), m.Kd[0], m.Kd[1], m.Kd[2]);
            mc.createFloatConstantTexture(sigma.c_str(), 0.0f);
            mc.createMatteMaterial(matName.c_str(), nullptr, kd.c_str(), sigma.c_str());
        } else {
            const std::string r = matName + "_Reflectance", t = matName + "_Transmittance";
            mc.createFloat3ConstantTexture(r.c_str(), 0.95f, 0.95f, 0.95f);
            mc.createFloat3ConstantTexture(t.c_str(), 0.95f, 0.95f, 0.95f);
            mc.createGlassMaterial(matName.c_str(), nullptr, r.c_str(), t.c_str(), m.Ni, 1.0f);
        }
    }

    scene->beginObject();
    for (const Object& o : objects) {
        for (size_t j = 0; j + 2 < o.positions.size(); j += 3)
            scene->addVertex(o.positions[j], o.positions[j + 1], o.positions[j + 2]);
        for (size_t j = 0; j + 1 < o.uvs.size(); j += 2) scene->addUV(o.uvs[j], o.uvs[j + 1]);
        for (size_t j = 0; j + 2 < o.normals.size(); j += 3)
            scene->addNormal(o.normals[j], o.normals[j + 1], o.normals[j + 2]);
        for (const Group& g : o.groups) {
            const auto it = matIdx.find(g.material);
            const Material& m = materials[it == matIdx.end() ? 0 : it->second];
            const std::string matName = std::string(prefix) + "_" + m.name;
            uint32_t alpha = UINT32_MAX;
            if (m.Ni == 1.0f && !m.mapKd.empty()) {
                const std::string kd = matName + "_Kd";
                const ImageTexture* im = reinterpret_cast<const ImageTexture*>(scene->ptrOfTex(kd));
                if (im->texType == TextureType::ColorImageRGBA8x4 || im->texType == TextureType::ColorImageRGBA16Fx4)
                    alpha = (uint32_t)scene->idxOfTex(kd);
            }
            const uint16_t mat = (uint16_t)(matOverride == UINT64_MAX ? scene->idxOfMat(matName) : matOverride);
            const uint16_t light = (uint16_t)(lightOverride == UINT64_MAX ? UINT16_MAX : lightOverride);
            for (size_t k = 0; k + 2 < g.pos.size(); k += 3) {
                if (g.format == Group::VTN)
                    scene->addFace(Face_P_N_UV(g.pos[k], g.pos[k + 1], g.pos[k + 2], g.nrm[k], g.nrm[k + 1], g.nrm[k + 2],
                                               g.uv[k], g.uv[k + 1], g.uv[k + 2], mat, light, alpha));
                else
                    scene->addFace(Face_P_N(g.pos[k], g.pos[k + 1], g.pos[k + 2], g.nrm[k], g.nrm[k + 1], g.nrm[k + 2],
                                            mat, light, alpha));
            }
        }
    }
    scene->endObject();
    scene->addModelToDB(fileName);
    ++scene->numModelsLoaded;
    return true;
}

}  // namespace host
}  // namespace clr
