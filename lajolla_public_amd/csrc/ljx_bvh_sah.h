// Binned-SAH decisions shared by the host and the device BVH builders.
//
// Both builders call these functions, compiled without floating-point
// contraction, so that on the same primitive boxes they choose the same
// splits and the same 4-wide collapse: the device tree is the host tree, byte
// for byte.  The BVH replaces the one Embree builds in rtcCommitScene
// (src/scene.cpp:20-27).
#pragma once

#include <hip/hip_runtime.h>

#include "ljx_device.h"

namespace ljx {

constexpr int kBvhBins = 32;         // bins per axis over the centroid box
constexpr int kBvhMedianDepth = 80;  // from this depth on, ranges are halved in order (depth bound)

// Half the surface area of a box; an empty box (lo > hi) has none.
__host__ __device__ inline float box_half_area(const float lo[3], const float hi[3]) {
#pragma clang fp contract(off)
    float d0 = hi[0] - lo[0], d1 = hi[1] - lo[1], d2 = hi[2] - lo[2];
    if (!(d0 >= 0)) return 0;
    return d0 * d1 + d1 * d2 + d2 * d0;
}

__host__ __device__ inline void box_grow(float lo[3], float hi[3], const float blo[3], const float bhi[3]) {
    for (int k = 0; k < 3; k++) {
        lo[k] = blo[k] < lo[k] ? blo[k] : lo[k];
        hi[k] = bhi[k] > hi[k] ? bhi[k] : hi[k];
    }
}

// Bin of a centroid coordinate c in the centroid box [lo, lo + ext] (ext > 0):
// floor(32 (c - lo) / ext), at most 31 (and 0 for a NaN).
__host__ __device__ inline int sah_bin(float c, float lo, float ext) {
#pragma clang fp contract(off)
    float t = float(kBvhBins) * (c - lo) / ext;
    if (!(t > 0)) return 0;
    if (t >= float(kBvhBins - 1)) return kBvhBins - 1;
    return int(t);
}

// Cost of a split: count x half area of each side.
__host__ __device__ inline float sah_pair_cost(int nl, const float llo[3], const float lhi[3], int nr,
                                               const float rlo[3], const float rhi[3]) {
#pragma clang fp contract(off)
    return float(nl) * box_half_area(llo, lhi) + float(nr) * box_half_area(rlo, rhi);
}

struct SahSplit {
    int axis = -1;     // -1: no binned split (all centroids coincide, or past kBvhMedianDepth)
    int bin = -1;      // primitives in bins <= bin go left
    int nleft = 0;
    float cost = __builtin_inff();  // count x half area, summed over both sides
};

// Best split along one axis: the first bin (lowest index) of least cost,
// over bins with primitives on both sides.  Bins provides count(k, b) and
// box(k, b, lo, hi) (an empty bin: count 0, lo = +inf, hi = -inf).
template <class Bins>
__host__ __device__ inline SahSplit sah_axis_best(const Bins &B, int k) {
#pragma clang fp contract(off)
    SahSplit s;
    float rarea[kBvhBins];
    int rcnt[kBvhBins];
    const float inf = __builtin_inff();
    float alo[3] = {inf, inf, inf}, ahi[3] = {-inf, -inf, -inf};
    int c = 0;
#pragma unroll
    for (int b = kBvhBins - 1; b > 0; b--) {
        float lo[3], hi[3];
        B.box(k, b, lo, hi);
        box_grow(alo, ahi, lo, hi);
        c += B.count(k, b);
        rarea[b] = box_half_area(alo, ahi);
        rcnt[b] = c;
    }
    float llo[3] = {inf, inf, inf}, lhi[3] = {-inf, -inf, -inf};
    c = 0;
#pragma unroll
    for (int b = 0; b < kBvhBins - 1; b++) {
        float lo[3], hi[3];
        B.box(k, b, lo, hi);
        box_grow(llo, lhi, lo, hi);
        c += B.count(k, b);
        if (c == 0 || rcnt[b + 1] == 0) continue;
        float cost = float(c) * box_half_area(llo, lhi) + float(rcnt[b + 1]) * rarea[b + 1];
        if (cost < s.cost) {
            s.cost = cost;
            s.axis = k;
            s.bin = b;
            s.nleft = c;
        }
    }
    return s;
}

// Combines the per-axis results in axis order; ties keep the lower axis.
__host__ __device__ inline void sah_take(SahSplit &best, const SahSplit &s) {
    if (s.axis >= 0 && s.cost < best.cost) best = s;
}

// Leaf or split: SAH in units of one primitive test, a node visit costing
// node_cost tests; ranges above max_leaf always split.
__host__ __device__ inline bool sah_make_leaf(int n, int max_leaf, const SahSplit &s, float half_area, float node_cost) {
#pragma clang fp contract(off)
    float split_cost = half_area > 0 ? node_cost + s.cost / half_area : __builtin_inff();
    float leaf_cost = float(n);
    return n <= max_leaf && (s.axis < 0 || leaf_cost <= split_cost);
}

// Leaf reference: ~((first slot << 3) | (count - 1)).
__host__ __device__ inline int bvh_leaf_ref(int first, int count) { return ~((first << 3) | (count - 1)); }

// ---- 4-wide collapse: one child of a 4-wide node
struct Kid {
    int ref;
    float lo[3], hi[3];
};

__host__ __device__ inline float kid_half_area(const Kid &c) {
#pragma clang fp contract(off)
    float d0 = c.hi[0] - c.lo[0], d1 = c.hi[1] - c.lo[1], d2 = c.hi[2] - c.lo[2];
    return d0 * d1 + d1 * d2 + d2 * d0;
}

__host__ __device__ inline void binary_kids(const BvhNode &nd, Kid &a, Kid &b) {
    a.ref = nd.ch.x;
    a.lo[0] = nd.lxy.x, a.hi[0] = nd.lxy.y, a.lo[1] = nd.lxy.z, a.hi[1] = nd.lxy.w, a.lo[2] = nd.zz.x, a.hi[2] = nd.zz.y;
    b.ref = nd.ch.y;
    b.lo[0] = nd.rxy.x, b.hi[0] = nd.rxy.y, b.lo[1] = nd.rxy.z, b.hi[1] = nd.rxy.w, b.lo[2] = nd.zz.z, b.hi[2] = nd.zz.w;
}

// The children of the 4-wide node that binary node b becomes: starting from
// b's two children, the inner child of largest area (the first of equal
// ones) is replaced by its two children until there are four or only
// leaves.  Returns the number of children.
__host__ __device__ inline int open4(const BvhNode *bin, int b, Kid kids[4]) {
    binary_kids(bin[b], kids[0], kids[1]);
    int nk = 2;
    while (nk < 4) {
        int best = -1;
        float area = -1;
        for (int i = 0; i < nk; i++) {
            if (kids[i].ref < 0) continue;
            float a = kid_half_area(kids[i]);
            if (a > area) area = a, best = i;
        }
        if (best < 0) break;
        Kid a, c;
        binary_kids(bin[kids[best].ref], a, c);
        for (int i = nk; i > best + 1; i--) kids[i] = kids[i - 1];
        kids[best] = a;
        kids[best + 1] = c;
        nk++;
    }
    return nk;
}

// The 4-wide node record: child boxes in SoA form, empty slots with inverted
// boxes and kNoChild; references are copied from kids (callers remap inner
// references to 4-wide indices).
__host__ __device__ inline Bvh4Node make_bvh4(const Kid kids[4], int nk) {
    const float inf = __builtin_inff();
    float lo[3][4], hi[3][4];
    int ch[4];
    for (int c = 0; c < 4; c++) {
        for (int k = 0; k < 3; k++) {
            lo[k][c] = c < nk ? kids[c].lo[k] : inf;
            hi[k][c] = c < nk ? kids[c].hi[k] : -inf;
        }
        ch[c] = c < nk ? kids[c].ref : kNoChild;
    }
    Bvh4Node nd;
    nd.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
    nd.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
    nd.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
    nd.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
    nd.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
    nd.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
    nd.ch = make_int4(ch[0], ch[1], ch[2], ch[3]);
    nd.meta = make_int4(nk, 0, 0, 0);
    return nd;
}

}  // namespace ljx
