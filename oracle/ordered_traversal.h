// CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// Restatement of the library's ordered LBVH traversal (the alternative to the
// reference's fixed-order walk of bvh_traversal.cl:182-215), shared by the
// trace and render oracles:
//   * per ray the reciprocal direction, with components below 1e-20 in
//     magnitude replaced by +-1e-20;
//   * padded slab test: per axis t0 = (min - o) * inv, t1 = (max - o) * inv;
//     tmin = max(min(t0, t1) over axes, 0), tmax = min(max(t0, t1) over axes);
//     the box is accepted when e = tmin * (1 - 2^-16) <= min(tmax * (1 + 2^-16), t),
//     and e is kept as its entry distance (re-checked as e <= t once t shrinks);
//   * a node whose box passed tests both children's boxes; leaf children are
//     tested first (nearer entry first, child 0 on equal entries), then the
//     nearer internal child is entered and the farther one pushed; a popped
//     node's box is tested again against the current t;
//   * a triangle replaces the hit when its distance is smaller, or equal with
//     a smaller face index, so the result is the lexicographic minimum of
//     (t, face) over accepted triangles, independent of the visiting order.
// With the padding, a box is never rejected while it holds a triangle whose
// intersection is accepted, so the result equals a brute-force loop with the
// same (t, face) rule (tests/test_oracle_traverse.py checks this).
#pragma once

#include <cmath>
#include <cstdint>
#include <utility>

#include "../include/clear_amd.h"

namespace orc_ordered {

constexpr float kPadDown = 1.0f - 1.0f / 65536.0f;
constexpr float kPadUp = 1.0f + 1.0f / 65536.0f;

inline float safeInv(float d) { return 1.0f / (std::fabs(d) > 1e-20f ? d : std::copysign(1e-20f, d)); }

// Padded entry distance e, or NaN when the box is missed within tBound (so
// that every later "e <= t" comparison fails, also for t = +inf).
inline float slabEnter(const float* mn, const float* mx, const float o[3], const float inv[3], float tBound) {
    const float x0 = (mn[0] - o[0]) * inv[0], x1 = (mx[0] - o[0]) * inv[0];
    const float y0 = (mn[1] - o[1]) * inv[1], y1 = (mx[1] - o[1]) * inv[1];
    const float z0 = (mn[2] - o[2]) * inv[2], z1 = (mx[2] - o[2]) * inv[2];
    const float tmin = std::fmax(std::fmax(std::fmin(x0, x1), std::fmin(y0, y1)), std::fmax(std::fmin(z0, z1), 0.0f));
    const float tmax = std::fmin(std::fmin(std::fmax(x0, x1), std::fmax(y0, y1)), std::fmax(z0, z1)) * kPadUp;
    const float e = tmin * kPadDown;
    return e <= std::fmin(tmax, tBound) ? e : NAN;
}

struct Counts {
    uint32_t visits = 0, tests = 0;
};

// tri(face, t, &tt) runs the triangle test (and any alpha test) against bound
// t and returns true with the distance tt when the triangle is accepted.
// anyHit stops at the first accepted triangle with a finite distance.
// Returns the face (0xFFFFFFFF when none) and updates t.
template <class TriFn>
uint32_t traverse(const clr_lbvh_inode* inodes, const clr_lbvh_leaf* leaves, uint32_t n, const float o[3],
                  const float d[3], float& t, TriFn tri, Counts& cnt, bool anyHit, bool* overflow) {
    const float inv[3] = {safeInv(d[0]), safeInv(d[1]), safeInv(d[2])};
    uint32_t best = 0xFFFFFFFFu;
    bool done = false;
    auto leaf = [&](uint32_t li) {
        const uint32_t obj = leaves[li].objIdx;
        ++cnt.tests;
        float tt;
        if (!tri(obj, t, &tt)) return;
        if (tt == t && obj > best) return;
        t = tt;
        best = obj;
        if (anyHit && !std::isinf(tt)) done = true;
    };
    if (n == 0) return best;
    if (n == 1) {
        if (slabEnter(leaves[0].min, leaves[0].max, o, inv, t) <= t) leaf(0);
        return best;
    }
    if (!(slabEnter(inodes[0].min, inodes[0].max, o, inv, t) <= t)) return best;
    uint32_t stack[64];
    int depth = 0;
    uint32_t node = 0;
    while (true) {
        ++cnt.visits;
        const clr_lbvh_inode& N = inodes[node];
        const bool l0 = N.isChild[0] != 0, l1 = N.isChild[1] != 0;
        const float* mn0 = l0 ? leaves[N.c[0]].min : inodes[N.c[0]].min;
        const float* mx0 = l0 ? leaves[N.c[0]].max : inodes[N.c[0]].max;
        const float* mn1 = l1 ? leaves[N.c[1]].min : inodes[N.c[1]].min;
        const float* mx1 = l1 ? leaves[N.c[1]].max : inodes[N.c[1]].max;
        float e0 = slabEnter(mn0, mx0, o, inv, t), e1 = slabEnter(mn1, mx1, o, inv, t);
        const bool first1 = e1 < e0;
        for (int k = 0; k < 2 && !done; ++k) {
            const int c = (k == 0) == first1 ? 1 : 0;
            const bool isLeaf = c ? l1 : l0;
            const float e = c ? e1 : e0;
            if (isLeaf && e <= t) leaf(N.c[c]);
        }
        if (done) break;
        const bool go0 = !l0 && e0 <= t, go1 = !l1 && e1 <= t;
        if (go0 && go1) {
            const uint32_t nearC = first1 ? N.c[1] : N.c[0], farC = first1 ? N.c[0] : N.c[1];
            if (depth < 64) stack[depth++] = farC;
            else if (overflow) *overflow = true;
            node = nearC;
            continue;
        }
        if (go0 || go1) {
            node = go0 ? N.c[0] : N.c[1];
            continue;
        }
        bool found = false;
        while (depth > 0) {
            const uint32_t cand = stack[--depth];
            if (slabEnter(inodes[cand].min, inodes[cand].max, o, inv, t) <= t) {
                node = cand;
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    return best;
}

// 4-wide form of the same traversal (the renderers' walk; the library keeps
// records of up to four grandchildren per internal node). A visit of node N
// gathers the children of each internal child of N, or a leaf child itself,
// in the order left.left, left.right, right.left, right.right; it tests all
// of their boxes against t, sorts them by entry distance (a five-exchange
// network that swaps only on a strictly smaller entry), tests the leaf
// entries in that order, then continues with the nearest internal entry whose
// entry is still <= t and pushes the other such entries (with their entry
// distance) farthest first; a pop skips entries beyond the current t. Each
// gathered box lies inside its parent's, so the result is the same
// lexicographic minimum of (t, face); only the visit counts (one per record)
// differ. stackPairs bounds the stack (the library's: 32 pairs in the
// megakernels' LDS stack, 64 in the wavefront and trace kernels); a push
// beyond it is dropped and flagged.
template <class TriFn>
uint32_t traverseWide(const clr_lbvh_inode* inodes, const clr_lbvh_leaf* leaves, uint32_t n, const float o[3],
                      const float d[3], float& t, TriFn tri, Counts& cnt, bool anyHit, bool* overflow,
                      int stackPairs = 64) {
    const float inv[3] = {safeInv(d[0]), safeInv(d[1]), safeInv(d[2])};
    uint32_t best = 0xFFFFFFFFu;
    bool done = false;
    auto leaf = [&](uint32_t li) {
        const uint32_t obj = leaves[li].objIdx;
        ++cnt.tests;
        float tt;
        if (!tri(obj, t, &tt)) return;
        if (tt == t && obj > best) return;
        t = tt;
        best = obj;
        if (anyHit && !std::isinf(tt)) done = true;
    };
    if (n == 0) return best;
    if (n == 1) {
        if (slabEnter(leaves[0].min, leaves[0].max, o, inv, t) <= t) leaf(0);
        return best;
    }
    if (!(slabEnter(inodes[0].min, inodes[0].max, o, inv, t) <= t)) return best;
    constexpr uint32_t kNone = 0xFFFFFFFFu;
    uint32_t stackRef[64];
    float stackE[64];
    int depth = 0;
    uint32_t node = 0;
    while (true) {
        ++cnt.visits;
        const clr_lbvh_inode& N = inodes[node];
        // The up to four entries: index, leaf flag, box.
        uint32_t ref[4] = {kNone, kNone, kNone, kNone};
        bool isLeaf[4] = {false, false, false, false};
        float k[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
        for (int c = 0; c < 2; ++c) {
            if (N.isChild[c]) {
                ref[2 * c] = N.c[c];
                isLeaf[2 * c] = true;
                continue;
            }
            const clr_lbvh_inode& C = inodes[N.c[c]];
            for (int m = 0; m < 2; ++m) {
                ref[2 * c + m] = C.c[m];
                isLeaf[2 * c + m] = C.isChild[m] != 0;
            }
        }
        for (int c = 0; c < 4; ++c) {
            if (ref[c] == kNone) continue;
            const float* mn = isLeaf[c] ? leaves[ref[c]].min : inodes[ref[c]].min;
            const float* mx = isLeaf[c] ? leaves[ref[c]].max : inodes[ref[c]].max;
            const float e = slabEnter(mn, mx, o, inv, t);
            if (e <= t) k[c] = e;
            else ref[c] = kNone;
        }
        auto cx = [&](int a, int b) {
            if (k[b] < k[a]) {
                std::swap(k[a], k[b]);
                std::swap(ref[a], ref[b]);
                std::swap(isLeaf[a], isLeaf[b]);
            }
        };
        cx(0, 1);
        cx(2, 3);
        cx(0, 2);
        cx(1, 3);
        cx(1, 2);
        for (int c = 0; c < 4 && !done; ++c)
            if (ref[c] != kNone && isLeaf[c] && k[c] <= t) leaf(ref[c]);
        if (done) break;
        uint32_t next = kNone;
        float enext = 0.0f;
        for (int c = 3; c >= 0; --c) {
            if (ref[c] == kNone || isLeaf[c] || !(k[c] <= t)) continue;
            if (next != kNone) {
                if (depth < stackPairs && depth < 64) {
                    stackRef[depth] = next;
                    stackE[depth] = enext;
                    ++depth;
                } else if (overflow) {
                    *overflow = true;
                }
            }
            next = ref[c];
            enext = k[c];
        }
        if (next != kNone) {
            node = next;
            continue;
        }
        bool found = false;
        while (depth > 0) {
            --depth;
            if (stackE[depth] <= t) {
                node = stackRef[depth];
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    return best;
}

}  // namespace orc_ordered
