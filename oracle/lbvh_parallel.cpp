// Multi-threaded CPU LBVH build — TEST INFRASTRUCTURE ONLY (the all-cores CPU
// baseline of bench.py). Same output as orc_lbvh_build in lbvh_oracle.cpp
// (the serial restatement of lbvh_construction.cl, driven by
// LBVHBuilder.cpp:160-375), byte for byte; tests/test_oracle_lbvh.py checks
// that. Threads split each stage:
//  * triangle boxes, and the scene box as per-thread folds combined in thread
//    order (omin / omax ignore NaN and keep the left operand on ties, so the
//    chunked fold equals the serial left fold);
//  * Morton codes, per triangle;
//  * the sort as an LSD radix sort of (interleaved key, index) with 8-bit
//    digits, per-thread digit counts and a stable scatter: the same
//    permutation as the reference's stable 3-bit-per-level passes;
//  * Karras nodes, per node (each writes only its own record and its
//    children's parent slots);
//  * the bottom-up refit as parallel climbs with an atomic arrival counter
//    per node (acq_rel: the second arrival sees the first child's box); the
//    merge order is fixed (left, right), so boxes do not depend on arrivals.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../include/clear_amd.h"

namespace {

struct F3 {
    float x, y, z;
};
inline float omin(float a, float b) { return (b < a || a != a) ? b : a; }
inline float omax(float a, float b) { return (b > a || a != a) ? b : a; }
inline F3 vmin(F3 a, F3 b) { return {omin(a.x, b.x), omin(a.y, b.y), omin(a.z, b.z)}; }
inline F3 vmax(F3 a, F3 b) { return {omax(a.x, b.x), omax(a.y, b.y), omax(a.z, b.z)}; }
inline uint32_t clz32(uint32_t v) { return v ? (uint32_t)__builtin_clz(v) : 32u; }

inline uint32_t quantize(float v, uint32_t numDivs) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return numDivs - 1u;
    uint32_t q = (uint32_t)v;
    return q < numDivs - 1u ? q : numDivs - 1u;
}

// Runs fn(t, begin, end) on `threads` threads over [0, n) in contiguous chunks.
template <class Fn>
void parallelFor(int threads, uint64_t n, Fn fn) {
    std::vector<std::thread> ts;
    for (int t = 0; t < threads; ++t) {
        const uint64_t b = n * t / threads, e = n * (t + 1) / threads;
        ts.emplace_back([=, &fn] { fn(t, b, e); });
    }
    for (auto& th : ts) th.join();
}

// The interleaved key x0 y0 z0 x1 y1 z1 ... (bit 3k: x bit k, 3k+1: y, 3k+2: z),
// whose order is the reference's digit order (x + 2y + 4z per bit level).
inline uint64_t interleave(const uint32_t* m, uint32_t bits) {
    uint64_t k = 0;
    for (uint32_t b = 0; b < bits; ++b)
        k |= (uint64_t)((m[0] >> b) & 1u) << (3 * b) | (uint64_t)((m[1] >> b) & 1u) << (3 * b + 1) |
             (uint64_t)((m[2] >> b) & 1u) << (3 * b + 2);
    return k;
}

// _numCommonBits (lbvh_construction.cl:322-332) over sorted positions; -1 outside [0, n).
inline int commonBits(const uint32_t* mc, const uint32_t* idx, uint32_t n, uint32_t bits, const uint32_t* mci,
                      int64_t i0, int64_t i1) {
    if (i1 < 0 || i1 >= (int64_t)n) return -1;
    const uint32_t j = idx[i1];
    const int lx = (int)(clz32(mci[0] ^ mc[3 * j + 0]) - (32u - bits));
    const int ly = (int)(clz32(mci[1] ^ mc[3 * j + 1]) - (32u - bits));
    const int lz = (int)(clz32(mci[2] ^ mc[3 * j + 2]) - (32u - bits));
    const int r = std::min(lz * 3 + 0, std::min(ly * 3 + 1, lx * 3 + 2));
    if ((uint32_t)r != bits * 3u) return r;
    return r + (int)clz32((uint32_t)((int32_t)i0 ^ (int32_t)i1));
}

}  // namespace

extern "C" {

int orc_lbvh_build_parallel(const clr_float3* verts, const clr_face* faces, uint32_t n, uint32_t bits,
                            clr_lbvh_inode* inodes, clr_lbvh_leaf* leaves, int threads) {
    if (n == 0 || bits < 1 || bits > 21) return 1;
    if (threads < 1) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    threads = (int)std::min<uint64_t>((uint64_t)threads, n);

    // calcAABBs and unifyAABBs.
    std::vector<F3> bmin(n), bmax(n);
    std::vector<F3> tmin(threads, F3{INFINITY, INFINITY, INFINITY}), tmax(threads, F3{-INFINITY, -INFINITY, -INFINITY});
    parallelFor(threads, n, [&](int t, uint64_t b, uint64_t e) {
        F3 lo = tmin[t], hi = tmax[t];
        for (uint64_t i = b; i < e; ++i) {
            const clr_face& f = faces[i];
            const F3 p0{verts[f.p0].x, verts[f.p0].y, verts[f.p0].z};
            const F3 p1{verts[f.p1].x, verts[f.p1].y, verts[f.p1].z};
            const F3 p2{verts[f.p2].x, verts[f.p2].y, verts[f.p2].z};
            bmin[i] = vmin(p0, vmin(p1, p2));
            bmax[i] = vmax(p0, vmax(p1, p2));
            lo = vmin(lo, bmin[i]);
            hi = vmax(hi, bmax[i]);
        }
        tmin[t] = lo;
        tmax[t] = hi;
    });
    F3 smin{INFINITY, INFINITY, INFINITY}, smax{-INFINITY, -INFINITY, -INFINITY};
    for (int t = 0; t < threads; ++t) {
        smin = vmin(smin, tmin[t]);
        smax = vmax(smax, tmax[t]);
    }
    const F3 ssize{smax.x - smin.x, smax.y - smin.y, smax.z - smin.z};

    // calcMortonCodes and the interleaved sort keys.
    const uint32_t numDivs = 1u << bits;
    const float fDivs = (float)numDivs;
    std::vector<uint32_t> mc(3 * (size_t)n);
    std::vector<uint64_t> key(n);
    parallelFor(threads, n, [&](int, uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            const float cx = (bmin[i].x + bmax[i].x) * 0.5f, cy = (bmin[i].y + bmax[i].y) * 0.5f,
                        cz = (bmin[i].z + bmax[i].z) * 0.5f;
            mc[3 * i + 0] = quantize((cx - smin.x) / ssize.x * fDivs, numDivs);
            mc[3 * i + 1] = quantize((cy - smin.y) / ssize.y * fDivs, numDivs);
            mc[3 * i + 2] = quantize((cz - smin.z) / ssize.z * fDivs, numDivs);
            key[i] = interleave(&mc[3 * i], bits);
        }
    });

    // Stable LSD radix sort of indices by key, 8-bit digits.
    std::vector<uint32_t> idx(n), tmp(n);
    std::vector<uint64_t> kin(key), kout(n);
    for (uint32_t i = 0; i < n; ++i) idx[i] = i;
    const int passes = (int)((3 * bits + 7) / 8);
    std::vector<uint64_t> cnt((size_t)threads * 256);
    for (int p = 0; p < passes; ++p) {
        const int sh = 8 * p;
        std::fill(cnt.begin(), cnt.end(), 0);
        parallelFor(threads, n, [&](int t, uint64_t b, uint64_t e) {
            uint64_t* c = &cnt[(size_t)t * 256];
            for (uint64_t i = b; i < e; ++i) ++c[(kin[i] >> sh) & 0xFF];
        });
        uint64_t run = 0;
        for (int d = 0; d < 256; ++d)
            for (int t = 0; t < threads; ++t) {
                const uint64_t c = cnt[(size_t)t * 256 + d];
                cnt[(size_t)t * 256 + d] = run;
                run += c;
            }
        parallelFor(threads, n, [&](int t, uint64_t b, uint64_t e) {
            uint64_t* o = &cnt[(size_t)t * 256];
            for (uint64_t i = b; i < e; ++i) {
                const uint64_t dst = o[(kin[i] >> sh) & 0xFF]++;
                kout[dst] = kin[i];
                tmp[dst] = idx[i];
            }
        });
        kin.swap(kout);
        idx.swap(tmp);
    }

    // Leaves and Karras nodes (constructBinaryRadixTree, lbvh_construction.cl:342-410).
    std::vector<uint32_t> parents(2 * (size_t)n, 0xFFFFFFFFu);
    parallelFor(threads, n, [&](int, uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            clr_lbvh_leaf& L = leaves[i];
            std::memset(&L, 0, sizeof(L));
            const uint32_t obj = idx[i];
            L.min[0] = bmin[obj].x; L.min[1] = bmin[obj].y; L.min[2] = bmin[obj].z;
            L.max[0] = bmax[obj].x; L.max[1] = bmax[obj].y; L.max[2] = bmax[obj].z;
            L.objIdx = obj;
        }
    });
    parallelFor(threads, n > 1 ? n - 1 : 0, [&](int, uint64_t b, uint64_t e) {
        for (uint64_t ii = b; ii < e; ++ii) {
            const int64_t i = (int64_t)ii;
            const uint32_t* mci = &mc[3 * (size_t)idx[i]];
            auto ncb = [&](int64_t j) { return commonBits(mc.data(), idx.data(), n, bits, mci, i, j); };
            const int d = ncb(i + 1) - ncb(i - 1) > 0 ? 1 : -1;
            const int minBits = ncb(i - d);
            uint64_t lMax = 2;
            while (ncb(i + (int64_t)lMax * d) > minBits) lMax <<= 1;
            uint64_t l = 0;
            for (uint64_t t = lMax >> 1; t >= 1; t >>= 1)
                if (ncb(i + (int64_t)(l + t) * d) > minBits) l += t;
            const int64_t other = i + (int64_t)l * d;
            const int nodeBits = ncb(other);
            uint64_t s = 0;
            float scale = 0.5f;
            while (true) {
                const uint64_t t = (uint64_t)std::ceil((float)l * scale);
                if (ncb(i + (int64_t)(s + t) * d) > nodeBits) s += t;
                if (t <= 1) break;
                scale *= 0.5f;
            }
            const int64_t split = i + (int64_t)s * d + std::min(d, 0);
            clr_lbvh_inode& N = inodes[i];
            std::memset(&N, 0, sizeof(N));
            const bool leftLeaf = std::min(i, other) == split, rightLeaf = std::max(i, other) == split + 1;
            N.isChild[0] = leftLeaf;
            N.isChild[1] = rightLeaf;
            N.c[0] = (uint32_t)split;
            N.c[1] = (uint32_t)(split + 1);
            parents[(leftLeaf ? 0 : n) + (size_t)split] = (uint32_t)i;
            parents[(rightLeaf ? 0 : n) + (size_t)split + 1] = (uint32_t)i;
        }
    });

    // calcNodeAABBs (lbvh_construction.cl:415-492) as parallel climbs.
    if (n >= 2) {
        std::vector<std::atomic<uint32_t>> arrivals(n - 1);
        for (auto& a : arrivals) a.store(0, std::memory_order_relaxed);
        parallelFor(threads, n, [&](int, uint64_t b, uint64_t e) {
            for (uint64_t g = b; g < e; ++g) {
                uint32_t tgt = parents[g];
                while (true) {
                    if (arrivals[tgt].fetch_add(1, std::memory_order_acq_rel) == 0) break;
                    clr_lbvh_inode& T = inodes[tgt];
                    auto cmin = [&](int k) -> F3 {
                        const uint32_t c = T.c[k];
                        return T.isChild[k] ? F3{leaves[c].min[0], leaves[c].min[1], leaves[c].min[2]}
                                            : F3{inodes[c].min[0], inodes[c].min[1], inodes[c].min[2]};
                    };
                    auto cmax = [&](int k) -> F3 {
                        const uint32_t c = T.c[k];
                        return T.isChild[k] ? F3{leaves[c].max[0], leaves[c].max[1], leaves[c].max[2]}
                                            : F3{inodes[c].max[0], inodes[c].max[1], inodes[c].max[2]};
                    };
                    // Both children are final: the climber's own box was written
                    // before it arrived, the sibling's before its arrival.
                    const F3 mn = vmin(cmin(0), cmin(1)), mx = vmax(cmax(0), cmax(1));
                    T.min[0] = mn.x; T.min[1] = mn.y; T.min[2] = mn.z;
                    T.max[0] = mx.x; T.max[1] = mx.y; T.max[2] = mx.z;
                    if (tgt == 0) break;
                    tgt = parents[n + tgt];
                }
            }
        });
    }
    return 0;
}

}  // extern "C"
