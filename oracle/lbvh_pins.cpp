// Independent checks for the LBVH oracle — TEST INFRASTRUCTURE ONLY.
//
// Each function computes a quantity the reference's build also defines, by a
// different route, so that a mistake in lbvh_oracle.cpp (a swapped axis in
// the digit, a wrong shift in the common-prefix count, a reversed split)
// shows up as a mismatch:
//   pin_interleave     Morton triples -> one interleaved key (bit 3k+0 = x_k,
//                      3k+1 = y_k, 3k+2 = z_k), the order blockwiseSort's
//                      digit x + 2y + 4z (lbvh_construction.cl:245-246)
//                      induces over LSB-first passes.
//   pin_stable_order   std::stable_sort of iota by that key.
//   pin_radix_tree     top-down binary radix tree over the extended keys
//                      (key, sorted position), split at the highest differing
//                      bit, numbered as Karras 2012 numbers it (a left child
//                      is named by its right end, a right child by its left
//                      end, the root is 0).
//   pin_range_boxes    every internal node's box as a plain linear min/max
//                      fold over the leaves of its range.
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <vector>

#include "../include/clear_amd.h"

namespace {
inline int clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }
inline int clz32(uint32_t v) { return v ? __builtin_clz(v) : 32; }

// Length of the common prefix of extended keys at sorted positions i, j.
inline int prefixExt(const uint64_t* k, uint32_t keyBits, uint32_t i, uint32_t j) {
    if (k[i] != k[j]) return clz64(k[i] ^ k[j]) - (64 - (int)keyBits);
    return (int)keyBits + clz32(i ^ j);
}
}  // namespace

extern "C" {

int pin_interleave(const uint32_t* mc, uint32_t n, uint32_t bits, uint64_t* keys) {
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t k = 0;
        for (uint32_t b = 0; b < bits; ++b) {
            k |= (uint64_t)((mc[3 * (size_t)i + 0] >> b) & 1u) << (3 * b + 0);
            k |= (uint64_t)((mc[3 * (size_t)i + 1] >> b) & 1u) << (3 * b + 1);
            k |= (uint64_t)((mc[3 * (size_t)i + 2] >> b) & 1u) << (3 * b + 2);
        }
        keys[i] = k;
    }
    return 0;
}

int pin_stable_order(const uint64_t* keys, uint32_t n, uint32_t* order) {
    std::vector<uint32_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0u);
    std::stable_sort(idx.begin(), idx.end(),
                     [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    std::copy(idx.begin(), idx.end(), order);
    return 0;
}

// sortedKeys: keys in sorted order. Outputs per internal node index (n-1):
// c0, c1, leaf flags and the covered range [lo, hi].
int pin_radix_tree(const uint64_t* sortedKeys, uint32_t n, uint32_t keyBits, uint32_t* c0,
                   uint32_t* c1, uint8_t* leaf0, uint8_t* leaf1, uint32_t* lo, uint32_t* hi) {
    if (n < 2) return 0;
    struct Item {
        uint32_t lo, hi, index;
    };
    std::vector<Item> stack;
    stack.push_back({0, n - 1, 0});
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        int p = prefixExt(sortedKeys, keyBits, it.lo, it.hi);
        // Largest g in [lo, hi-1] whose key still shares more than p bits
        // with key[lo]: binary search (the shared-prefix predicate is
        // monotone over a sorted range).
        uint32_t a = it.lo, b = it.hi - 1;
        while (a < b) {
            uint32_t m = a + (b - a + 1) / 2;
            if (prefixExt(sortedKeys, keyBits, it.lo, m) > p)
                a = m;
            else
                b = m - 1;
        }
        uint32_t g = a;
        c0[it.index] = g;
        c1[it.index] = g + 1;
        leaf0[it.index] = (g == it.lo);
        leaf1[it.index] = (g + 1 == it.hi);
        lo[it.index] = it.lo;
        hi[it.index] = it.hi;
        if (g > it.lo) stack.push_back({it.lo, g, g});
        if (g + 1 < it.hi) stack.push_back({g + 1, it.hi, g + 1});
    }
    return 0;
}

// leafBoxes: float[n][6] in leaf order. out: float[n-1][6].
int pin_range_boxes(const float* leafBoxes, const uint32_t* lo, const uint32_t* hi, uint32_t n,
                    float* out) {
    for (uint32_t i = 0; i + 1 < n; ++i) {
        float m[6];
        for (int k = 0; k < 6; ++k) m[k] = leafBoxes[6 * (size_t)lo[i] + k];
        for (uint32_t j = lo[i] + 1; j <= hi[i]; ++j)
            for (int k = 0; k < 3; ++k) {
                m[k] = std::min(m[k], leafBoxes[6 * (size_t)j + k]);
                m[3 + k] = std::max(m[3 + k], leafBoxes[6 * (size_t)j + 3 + k]);
            }
        for (int k = 0; k < 6; ++k) out[6 * (size_t)i + k] = m[k];
    }
    return 0;
}

}  // extern "C"
