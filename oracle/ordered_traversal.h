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

}  // namespace orc_ordered
