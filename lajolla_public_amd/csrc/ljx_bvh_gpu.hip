// Binned-SAH BVH builder on the device (gfx950), and the 4-wide collapse.
//
// Replaces the BVH build inside Embree's rtcCommitScene (src/scene.cpp:20-27).
// The split decisions are the host builder's (ljx_bvh_sah.h, no FP
// contraction) and every partition is stable, so the tree equals the host
// builder's byte for byte; only the work schedule differs:
//
//  * wide ranges (more than kSmall primitives) are split level by level, all
//    ranges of a level at once.  A level runs over chunks of kChunk
//    primitives (one 256-lane block each): the range box and centroid box
//    (block reduction, then atomic min / max on order-preserving integer
//    images of the floats), the 3 x 32 bins (LDS atomics, merged into the
//    range's bins with one global atomic per non-empty bin and value), the
//    split (one lane per range), then a stable partition by per-chunk counts,
//    their scan per range and a block-level scan of the flags;
//  * narrow ranges are built to their leaves by one wave each: bounds by wave
//    reduction, bins in LDS, the split by three lanes (one per axis), a
//    ballot-ranked stable partition staged through LDS, and a stack of
//    pending ranges (the larger pushed, the smaller taken next, so it stays
//    logarithmic);
//  * the collapse to 4-wide nodes runs level by level over the nodes that
//    become 4-wide nodes (open4), counts each one's 4-wide subtree bottom-up
//    and numbers the nodes top-down in depth-first preorder, the host
//    collapse's order.
#include "ljx_bvh_gpu.h"

#include "ljx_bvh_sah.h"
#include "ljx_host.h"
#include "ljx_trace.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <vector>

namespace ljx {

namespace {

constexpr int kChunk = 2048;  // primitives per block in the wide-range kernels
constexpr int kSmall = 512;   // ranges of at most this many primitives are built by one wave
constexpr int kWaveStack = 32;
constexpr int kBinInts = 3 * kBvhBins * 7;  // per bin: count, lo xyz, hi xyz

// Order-preserving map of non-NaN floats to ints (for atomic min / max).
__device__ __forceinline__ int f2o(float f) {
    int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float o2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

constexpr int kPosInfO = 0x7f800000;                  // f2o(+inf)
constexpr int kNegInfO = int(0xff800000u ^ 0x7fffffffu);  // f2o(-inf)

struct Item {
    int begin, end, depth, parent;  // parent: node * 2 + side, -1 for the root
};

struct LevelSplit {
    int axis, bin, mid;
    float clo, cext;  // centroid box along the axis
};

// Bins stored as ordered ints: [k][b][0] count, [1..3] lo, [4..6] hi.
struct OrderedBins {
    const int *p;
    __device__ int count(int k, int b) const { return p[(k * kBvhBins + b) * 7]; }
    __device__ void box(int k, int b, float lo[3], float hi[3]) const {
        const int *q = p + (k * kBvhBins + b) * 7;
        for (int d = 0; d < 3; d++) lo[d] = o2f(q[1 + d]), hi[d] = o2f(q[4 + d]);
    }
};

__device__ __forceinline__ float pick3(float a, float b, float c, int k) { return k == 0 ? a : (k == 1 ? b : c); }

// The best split of a range from its ordered-int bins, by one whole wave:
// lanes 0-31 and 32-63 take one axis each (axes 0 and 1, then axis 2), a lane
// per bin.  Inclusive prefix and suffix unions over the bins (exact min / max
// and integer counts) give lane b the two sides of the split after bin b,
// whose cost is computed as sah_axis_best computes it; the first least cost
// of each axis (lowest bin) is then taken in axis order with a strict
// comparison.  The result is sah_axis_best + sah_take's, on every lane.
__device__ SahSplit wave_sah(const int *bins, float ex, float ey, float ez) {
    const int lane = threadIdx.x & 63, b = lane & 31;
    const float inf = __builtin_inff();
    SahSplit best;
    for (int pass = 0; pass < 2; pass++) {
        const int k = 2 * pass + (lane >> 5);
        const bool active = k < 3 && pick3(ex, ey, ez, k) > 0;
        int cnt = 0;
        float lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
        if (active) {
            const int *q = bins + (k * kBvhBins + b) * 7;
            cnt = q[0];
            for (int d = 0; d < 3; d++) lo[d] = o2f(q[1 + d]), hi[d] = o2f(q[4 + d]);
        }
        int lc = cnt, rc = cnt;
        float llo[3], lhi[3], rlo[3], rhi[3];
        for (int d = 0; d < 3; d++) llo[d] = rlo[d] = lo[d], lhi[d] = rhi[d] = hi[d];
        for (int o = 1; o < 32; o <<= 1) {
            int uc = __shfl_up(lc, o, 32), dc = __shfl_down(rc, o, 32);
            float ul[3], uh[3], dl[3], dh[3];
            for (int d = 0; d < 3; d++) {
                ul[d] = __shfl_up(llo[d], o, 32), uh[d] = __shfl_up(lhi[d], o, 32);
                dl[d] = __shfl_down(rlo[d], o, 32), dh[d] = __shfl_down(rhi[d], o, 32);
            }
            if (b >= o) {
                lc += uc;
                for (int d = 0; d < 3; d++) llo[d] = fminf(llo[d], ul[d]), lhi[d] = fmaxf(lhi[d], uh[d]);
            }
            if (b + o < 32) {
                rc += dc;
                for (int d = 0; d < 3; d++) rlo[d] = fminf(rlo[d], dl[d]), rhi[d] = fmaxf(rhi[d], dh[d]);
            }
        }
        // the right side of the split after bin b is the suffix from bin b + 1
        int nrc = __shfl_down(rc, 1, 32);
        float nrl[3], nrh[3];
        for (int d = 0; d < 3; d++) nrl[d] = __shfl_down(rlo[d], 1, 32), nrh[d] = __shfl_down(rhi[d], 1, 32);
        float cost = inf;
        if (active && b < kBvhBins - 1 && lc > 0 && nrc > 0) {
            cost = sah_pair_cost(lc, llo, lhi, nrc, nrl, nrh);
            if (!(cost < inf)) cost = inf;  // NaN or overflow: never chosen (as cost < best fails)
        }
        int bb = b, nl = lc;
        for (int o = 16; o > 0; o >>= 1) {
            float c2 = __shfl_xor(cost, o, 32);
            int b2 = __shfl_xor(bb, o, 32), n2 = __shfl_xor(nl, o, 32);
            if (c2 < cost || (c2 == cost && b2 < bb)) cost = c2, bb = b2, nl = n2;
        }
        for (int h = 0; h < 2; h++) {
            const int kk = 2 * pass + h;
            float hc = __shfl(cost, 32 * h);
            int hb = __shfl(bb, 32 * h), hn = __shfl(nl, 32 * h);
            if (kk < 3 && hc < best.cost) {
                best.cost = hc;
                best.axis = kk;
                best.bin = hb;
                best.nleft = hn;
            }
        }
    }
    return best;
}

__device__ __forceinline__ void bins_clear(int *bins, int t, int nt) {
    for (int i = t; i < kBinInts; i += nt) {
        int r = i % 7;
        bins[i] = r == 0 ? 0 : (r <= 3 ? kPosInfO : kNegInfO);
    }
}

// Writes a child (box and reference) into its parent's slot; the root, when
// it is a leaf, becomes node 0 with that leaf as both children.
__device__ void write_child(BvhNode *nodes, int parent, const float lo[3], const float hi[3], int ref) {
    if (parent < 0) {
        if (ref >= 0) return;
        BvhNode nd;
        nd.lxy = make_float4(lo[0], hi[0], lo[1], hi[1]);
        nd.rxy = nd.lxy;
        nd.zz = make_float4(lo[2], hi[2], lo[2], hi[2]);
        nd.ch = make_int4(ref, ref, 0, 0);
        nodes[0] = nd;
        return;
    }
    BvhNode &nd = nodes[parent >> 1];
    if ((parent & 1) == 0) {
        nd.lxy = make_float4(lo[0], hi[0], lo[1], hi[1]);
        reinterpret_cast<float2 *>(&nd.zz)[0] = make_float2(lo[2], hi[2]);
        nd.ch.x = ref;
    } else {
        nd.rxy = make_float4(lo[0], hi[0], lo[1], hi[1]);
        reinterpret_cast<float2 *>(&nd.zz)[1] = make_float2(lo[2], hi[2]);
        nd.ch.y = ref;
    }
}

struct BuildCtx {
    const float4 *blo, *bhi, *cen;  // per primitive: box and centroid
    int *ids, *tmp;
    BvhNode *nodes;
    int *counters;
    int max_leaf;
    float node_cost;
};
// counters: binary nodes, next wide ranges, narrow ranges, max depth, stack
// (narrow-range overflow, then the traversal stack need), level size, error
enum { CN_NODES = 0, CN_WIDE = 1, CN_NARROW = 2, CN_DEPTH = 3, CN_STACK = 4, CN_LEVEL = 5, CN_ERR = 6, CN_NUM = 8 };

// ---------------------------------------------------------------- set-up
__global__ void k_init_prims(const float4 *blo, const float4 *bhi, float4 *cen, int *ids, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 l = blo[i], h = bhi[i];
    cen[i] = make_float4(0.5f * (l.x + h.x), 0.5f * (l.y + h.y), 0.5f * (l.z + h.z), 0.f);
    ids[i] = i;
}

// --------------------------------------------------------- wide ranges
// Chunk table of a level: cstart[i] = first chunk of item i (exclusive scan
// of ceil(size / kChunk)); one block.
__global__ void k_chunk_scan(const Item *items, int ni, int *cstart) {
    __shared__ int part[1024];
    int carry = 0;
    for (int base = 0; base < ni; base += 1024) {
        int i = base + threadIdx.x;
        int v = i < ni ? (items[i].end - items[i].begin + kChunk - 1) / kChunk : 0;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int a = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += a;
            __syncthreads();
        }
        if (i < ni) cstart[i] = carry + part[threadIdx.x] - v;
        int tot = part[1023];
        __syncthreads();
        carry += tot;
    }
    if (threadIdx.x == 0) cstart[ni] = carry;
}

__global__ void k_item_clear(int *ibox, int *ibins, int ni) {
    int it = blockIdx.x;
    if (it >= ni) return;
    for (int i = threadIdx.x; i < 12; i += blockDim.x) ibox[it * 12 + i] = (i % 6) < 3 ? kPosInfO : kNegInfO;
    bins_clear(ibins + size_t(it) * kBinInts, threadIdx.x, blockDim.x);
}

// The item and chunk of this block, or false past the last chunk.
__device__ bool chunk_of(const Item *items, int ni, const int *cstart, int &item, int &b0, int &b1) {
    int c = blockIdx.x;
    if (c >= cstart[ni]) return false;
    int lo = 0, hi = ni - 1;
    while (lo < hi) {  // last item with cstart <= c
        int mid = (lo + hi + 1) >> 1;
        if (cstart[mid] <= c) lo = mid;
        else hi = mid - 1;
    }
    item = lo;
    const Item it = items[lo];
    b0 = it.begin + (c - cstart[lo]) * kChunk;
    b1 = min(it.end, b0 + kChunk);
    return true;
}

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

__global__ __launch_bounds__(256) void k_wide_bounds(BuildCtx B, const Item *items, int ni, const int *cstart,
                                                     int *ibox) {
    int item, b0, b1;
    if (!chunk_of(items, ni, cstart, item, b0, b1)) return;
    const float inf = __builtin_inff();
    float v[12];
    for (int k = 0; k < 6; k++) v[k] = inf, v[6 + k] = -inf;  // lo: bb xyz, cb xyz; hi: bb xyz, cb xyz
    for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
        int id = B.ids[i];
        float4 l = B.blo[id], h = B.bhi[id], c = B.cen[id];
        v[0] = fminf(v[0], l.x), v[1] = fminf(v[1], l.y), v[2] = fminf(v[2], l.z);
        v[3] = fminf(v[3], c.x), v[4] = fminf(v[4], c.y), v[5] = fminf(v[5], c.z);
        v[6] = fmaxf(v[6], h.x), v[7] = fmaxf(v[7], h.y), v[8] = fmaxf(v[8], h.z);
        v[9] = fmaxf(v[9], c.x), v[10] = fmaxf(v[10], c.y), v[11] = fmaxf(v[11], c.z);
    }
    for (int k = 0; k < 6; k++) v[k] = wave_min(v[k]), v[6 + k] = wave_max(v[6 + k]);
    __shared__ int part[4][12];
    int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
        for (int k = 0; k < 12; k++) part[w][k] = f2o(v[k]);
    __syncthreads();
    if (threadIdx.x < 12) {
        int k = threadIdx.x;
        int r = part[0][k];
        for (int q = 1; q < 4; q++) r = k < 6 ? min(r, part[q][k]) : max(r, part[q][k]);
        // ibox layout: bb lo xyz, bb hi xyz, cb lo xyz, cb hi xyz
        int dst = k < 3 ? k : (k < 6 ? 6 + (k - 3) : (k < 9 ? 3 + (k - 6) : 9 + (k - 9)));
        if (k < 6) atomicMin(&ibox[item * 12 + dst], r);
        else atomicMax(&ibox[item * 12 + dst], r);
    }
}

__device__ __forceinline__ void item_boxes(const int *ib, float bl[3], float bh[3], float cl[3], float ch[3]) {
    for (int k = 0; k < 3; k++) {
        bl[k] = o2f(ib[k]);
        bh[k] = o2f(ib[3 + k]);
        cl[k] = o2f(ib[6 + k]);
        ch[k] = o2f(ib[9 + k]);
    }
}

__global__ __launch_bounds__(256) void k_wide_bin(BuildCtx B, const Item *items, int ni, const int *cstart,
                                                  const int *ibox, int *ibins) {
    int item, b0, b1;
    if (!chunk_of(items, ni, cstart, item, b0, b1)) return;
    if (items[item].depth >= kBvhMedianDepth) return;
    __shared__ int bins[kBinInts];
    bins_clear(bins, threadIdx.x, blockDim.x);
    float bl[3], bh[3], cl[3], chh[3], ext[3];
    item_boxes(ibox + item * 12, bl, bh, cl, chh);
    for (int k = 0; k < 3; k++) ext[k] = chh[k] - cl[k];
    __syncthreads();
    for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
        int id = B.ids[i];
        float4 l = B.blo[id], h = B.bhi[id], c = B.cen[id];
        float cc[3] = {c.x, c.y, c.z};
        int ol[3] = {f2o(l.x), f2o(l.y), f2o(l.z)}, oh[3] = {f2o(h.x), f2o(h.y), f2o(h.z)};
        for (int k = 0; k < 3; k++) {
            if (!(ext[k] > 0)) continue;
            int *q = bins + (k * kBvhBins + sah_bin(cc[k], cl[k], ext[k])) * 7;
            atomicAdd(q, 1);
            for (int d = 0; d < 3; d++) {
                atomicMin(q + 1 + d, ol[d]);
                atomicMax(q + 4 + d, oh[d]);
            }
        }
    }
    __syncthreads();
    int *g = ibins + size_t(item) * kBinInts;
    for (int bk = threadIdx.x; bk < 3 * kBvhBins; bk += blockDim.x) {
        const int *q = bins + bk * 7;
        if (q[0] == 0) continue;
        atomicAdd(g + bk * 7, q[0]);
        for (int d = 0; d < 3; d++) {
            atomicMin(g + bk * 7 + 1 + d, q[1 + d]);
            atomicMax(g + bk * 7 + 4 + d, q[4 + d]);
        }
    }
}

// One wave per wide range: the split, the node, the two child ranges.
__global__ __launch_bounds__(64) void k_wide_split(BuildCtx B, const Item *items, int ni, const int *ibox,
                                                   const int *ibins, LevelSplit *splits, Item *next_wide,
                                                   Item *narrow) {
    const int i = blockIdx.x;
    if (i >= ni) return;
    const int lane = threadIdx.x;
    const Item it = items[i];
    float bl[3], bh[3], cl[3], ch[3], ext[3];
    item_boxes(ibox + i * 12, bl, bh, cl, ch);
    for (int k = 0; k < 3; k++) ext[k] = ch[k] - cl[k];
    const int n = it.end - it.begin;
    SahSplit s;
    if (it.depth < kBvhMedianDepth) s = wave_sah(ibins + size_t(i) * kBinInts, ext[0], ext[1], ext[2]);
    if (lane != 0) return;
    atomicMax(&B.counters[CN_DEPTH], it.depth);
    LevelSplit ls{-1, -1, it.begin, 0.f, 0.f};
    if (n <= 1 || sah_make_leaf(n, B.max_leaf, s, box_half_area(bl, bh), B.node_cost)) {
        // (wide ranges exceed max_leaf, so this does not happen; kept for completeness)
        write_child(B.nodes, it.parent, bl, bh, bvh_leaf_ref(it.begin, n));
        ls.mid = -1;
        splits[i] = ls;
        return;
    }
    int node = it.parent < 0 ? 0 : atomicAdd(&B.counters[CN_NODES], 1);
    write_child(B.nodes, it.parent, bl, bh, node);
    ls.axis = s.axis;
    ls.bin = s.bin;
    ls.mid = s.axis >= 0 ? it.begin + s.nleft : it.begin + n / 2;
    if (s.axis >= 0)
        ls.clo = pick3(cl[0], cl[1], cl[2], s.axis), ls.cext = pick3(ext[0], ext[1], ext[2], s.axis);
    splits[i] = ls;
    Item kids[2] = {{it.begin, ls.mid, it.depth + 1, node * 2}, {ls.mid, it.end, it.depth + 1, node * 2 + 1}};
    for (int c = 0; c < 2; c++) {
        if (kids[c].end - kids[c].begin > kSmall) next_wide[atomicAdd(&B.counters[CN_WIDE], 1)] = kids[c];
        else narrow[atomicAdd(&B.counters[CN_NARROW], 1)] = kids[c];
    }
}

__device__ __forceinline__ bool goes_left(const BuildCtx &B, const LevelSplit &ls, int pos, int id) {
    if (ls.axis < 0) return pos < ls.mid;
    float4 c = B.cen[id];
    float v = ls.axis == 0 ? c.x : (ls.axis == 1 ? c.y : c.z);
    return sah_bin(v, ls.clo, ls.cext) <= ls.bin;
}

// Exclusive block scan of one int per thread (256 threads); returns the total.
__device__ int block_scan(int v, int &excl) {
    __shared__ int wsum[4];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int before = 0, tot = 0;
    for (int q = 0; q < 4; q++) {
        if (q < w) before += wsum[q];
        tot += wsum[q];
    }
    excl = before + x - v;
    __syncthreads();
    return tot;
}

constexpr int kPerThread = kChunk / 256;

__global__ __launch_bounds__(256) void k_wide_count(BuildCtx B, const Item *items, int ni, const int *cstart,
                                                    const LevelSplit *splits, int *ccount) {
    int item, b0, b1;
    if (!chunk_of(items, ni, cstart, item, b0, b1)) return;
    const LevelSplit ls = splits[item];
    int cnt = 0;
    if (ls.mid >= 0)
        for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x) cnt += goes_left(B, ls, i, B.ids[i]) ? 1 : 0;
    int excl;
    int tot = block_scan(cnt, excl);
    if (threadIdx.x == 0) ccount[blockIdx.x] = tot;
}

// Per range: exclusive scan of its chunks' left counts (one lane per range).
__global__ void k_wide_offsets(const Item *items, int ni, const int *cstart, int *ccount) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ni) return;
    int run = 0;
    for (int c = cstart[i]; c < cstart[i + 1]; c++) {
        int v = ccount[c];
        ccount[c] = run;
        run += v;
    }
}

__global__ __launch_bounds__(256) void k_wide_scatter(BuildCtx B, const Item *items, int ni, const int *cstart,
                                                      const LevelSplit *splits, const int *coff) {
    int item, b0, b1;
    if (!chunk_of(items, ni, cstart, item, b0, b1)) return;
    const LevelSplit ls = splits[item];
    if (ls.mid < 0) return;
    const Item it = items[item];
    // thread t takes kPerThread consecutive primitives, so ranks keep the order
    int base = b0 + threadIdx.x * kPerThread;
    int ids[kPerThread];
    bool f[kPerThread];
    int cnt = 0;
    for (int j = 0; j < kPerThread; j++) {
        int p = base + j;
        ids[j] = p < b1 ? B.ids[p] : 0;
        f[j] = p < b1 && goes_left(B, ls, p, ids[j]);
        cnt += f[j] ? 1 : 0;
    }
    int excl;
    block_scan(cnt, excl);
    int left = coff[blockIdx.x] + excl;  // lefts of the range before this thread's first primitive
    int done = base - it.begin;                            // primitives of the range before this thread
    for (int j = 0; j < kPerThread; j++) {
        int p = base + j;
        if (p >= b1) break;
        int pos;
        if (f[j]) pos = it.begin + left++;
        else pos = ls.mid + (done + j - left);
        if (pos < it.begin || pos >= it.end) {
            atomicOr(&B.counters[CN_ERR], 1);
            continue;
        }
        B.tmp[pos] = ids[j];
    }
}

__global__ __launch_bounds__(256) void k_wide_copy(BuildCtx B, const Item *items, int ni, const int *cstart,
                                                   const LevelSplit *splits) {
    int item, b0, b1;
    if (!chunk_of(items, ni, cstart, item, b0, b1)) return;
    if (splits[item].mid < 0) return;
    for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x) B.ids[i] = B.tmp[i];
}

// -------------------------------------------------------- narrow ranges
// One wave builds one range to its leaves.  The range's primitive boxes are
// staged in LDS once (SoA), with a local index permutation that the
// partitions reorder; centroids are recomputed from the boxes as k_init_prims
// computes them.  The global slot order is written back at the end.
__global__ __launch_bounds__(64) void k_narrow(BuildCtx B, const Item *items, int ni) {
    __shared__ int bins[kBinInts];
    __shared__ float pb[6][kSmall];  // lo xyz, hi xyz
    __shared__ int gid[kSmall];
    __shared__ int idx[kSmall];
    __shared__ Item stack[kWaveStack];
    if (blockIdx.x >= ni) return;
    const int lane = threadIdx.x;
    const float inf = __builtin_inff();
    const Item root = items[blockIdx.x];
    const int base = root.begin, nr = root.end - root.begin;
    for (int j = lane; j < nr; j += 64) {
        const int g = B.ids[base + j];
        const float4 l = B.blo[g], h = B.bhi[g];
        gid[j] = g;
        idx[j] = j;
        pb[0][j] = l.x, pb[1][j] = l.y, pb[2][j] = l.z;
        pb[3][j] = h.x, pb[4][j] = h.y, pb[5][j] = h.z;
    }
    if (lane == 0) stack[0] = Item{0, nr, root.depth, root.parent};
    int sp = 1;
    // the range needs at most nr - 1 inner nodes: one reservation per wave
    // (the root range of the whole scene is node 0 and needs one fewer)
    int next_node = 0;
    if (lane == 0 && nr > 1) next_node = atomicAdd(&B.counters[CN_NODES], root.parent < 0 ? nr - 2 : nr - 1);
    next_node = __shfl(next_node, 0);
    __syncthreads();
    int maxd = 0;
    while (sp > 0) {
        const Item it = stack[sp - 1];  // local range [begin, end)
        sp--;
        __syncthreads();
        maxd = max(maxd, it.depth);
        const int n = it.end - it.begin;
        // bounds
        float v[12];
        for (int k = 0; k < 6; k++) v[k] = inf, v[6 + k] = -inf;
        for (int j = it.begin + lane; j < it.end; j += 64) {
            const int p = idx[j];
            float l[3] = {pb[0][p], pb[1][p], pb[2][p]}, h[3] = {pb[3][p], pb[4][p], pb[5][p]};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float c = 0.5f * (l[k] + h[k]);
                v[k] = fminf(v[k], l[k]);
                v[3 + k] = fminf(v[3 + k], c);
                v[6 + k] = fmaxf(v[6 + k], h[k]);
                v[9 + k] = fmaxf(v[9 + k], c);
            }
        }
        for (int k = 0; k < 6; k++) v[k] = wave_min(v[k]), v[6 + k] = wave_max(v[6 + k]);
        float bl[3] = {v[0], v[1], v[2]}, bh[3] = {v[6], v[7], v[8]};
        float cl[3] = {v[3], v[4], v[5]}, ext[3] = {v[9] - v[3], v[10] - v[4], v[11] - v[5]};
        if (n <= 1) {
            if (lane == 0) write_child(B.nodes, it.parent, bl, bh, bvh_leaf_ref(base + it.begin, n));
            continue;
        }
        SahSplit s;
        if (it.depth < kBvhMedianDepth) {
            bins_clear(bins, lane, 64);
            __syncthreads();
            for (int j = it.begin + lane; j < it.end; j += 64) {
                const int p = idx[j];
                float l[3] = {pb[0][p], pb[1][p], pb[2][p]}, h[3] = {pb[3][p], pb[4][p], pb[5][p]};
                int ol[3] = {f2o(l[0]), f2o(l[1]), f2o(l[2])}, oh[3] = {f2o(h[0]), f2o(h[1]), f2o(h[2])};
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    if (!(ext[k] > 0)) continue;
                    const float c = 0.5f * (l[k] + h[k]);
                    int *q = bins + (k * kBvhBins + sah_bin(c, cl[k], ext[k])) * 7;
                    atomicAdd(q, 1);
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        atomicMin(q + 1 + d, ol[d]);
                        atomicMax(q + 4 + d, oh[d]);
                    }
                }
            }
            __syncthreads();
            s = wave_sah(bins, ext[0], ext[1], ext[2]);
        }
        if (sah_make_leaf(n, B.max_leaf, s, box_half_area(bl, bh), B.node_cost)) {
            if (lane == 0) write_child(B.nodes, it.parent, bl, bh, bvh_leaf_ref(base + it.begin, n));
            continue;
        }
        const int node = it.parent < 0 ? 0 : next_node++;
        if (lane == 0) write_child(B.nodes, it.parent, bl, bh, node);
        const int mid = s.axis >= 0 ? it.begin + s.nleft : it.begin + n / 2;
        // stable partition of the local indices: read the whole range into
        // registers, then write every index to its new place
        constexpr int kPer = kSmall / 64;
        const int ax = s.axis < 0 ? 0 : s.axis;
        const float a_lo = pick3(cl[0], cl[1], cl[2], ax), a_ext = pick3(ext[0], ext[1], ext[2], ax);
        int val[kPer];
        unsigned long long mask[kPer];
        for (int q = 0; q < kPer; q++) {
            const int j = it.begin + 64 * q + lane;
            const bool valid = j < it.end;
            const int p = valid ? idx[j] : 0;
            bool f = false;
            if (valid) {
                if (s.axis < 0) {
                    f = j < it.begin + n / 2;
                } else {
                    const float c = 0.5f * (pb[ax][p] + pb[3 + ax][p]);
                    f = sah_bin(c, a_lo, a_ext) <= s.bin;
                }
            }
            val[q] = p;
            mask[q] = __ballot(f);
        }
        __syncthreads();
        int nl = 0;
        for (int q = 0; q < kPer; q++) {
            const int j = it.begin + 64 * q + lane;
            const bool f = (mask[q] >> lane) & 1ull;
            const int rank = __popcll(mask[q] & ((1ull << lane) - 1ull));
            if (j < it.end) {
                const int pos = f ? it.begin + nl + rank : mid + ((j - it.begin) - (nl + rank));
                if (pos >= it.begin && pos < it.end) idx[pos] = val[q];
                else atomicOr(&B.counters[CN_ERR], 2);
            }
            nl += __popcll(mask[q]);
        }
        __syncthreads();
        // pending ranges: the larger below, the smaller on top
        const bool left_small = (mid - it.begin) <= (it.end - mid);
        if (sp + 2 > kWaveStack) {
            if (lane == 0) atomicMax(&B.counters[CN_STACK], 1);
            break;
        }
        if (lane == 0) {
            const int lb = it.begin, le = mid, rb = mid, re = it.end;
            stack[sp] = Item{left_small ? rb : lb, left_small ? re : le, it.depth + 1,
                             node * 2 + (left_small ? 1 : 0)};
            stack[sp + 1] = Item{left_small ? lb : rb, left_small ? le : re, it.depth + 1,
                                 node * 2 + (left_small ? 0 : 1)};
        }
        sp += 2;
        __syncthreads();
    }
    __syncthreads();
    for (int j = lane; j < nr; j += 64) B.ids[base + j] = gid[idx[j]];
    if (lane == 0) atomicMax(&B.counters[CN_DEPTH], maxd);
}

// ------------------------------------------------------------- collapse
struct CollapseCtx {
    const BvhNode *bin;
    int nbin;         // binary inner nodes
    Bvh4Node *w4;     // per binary node that becomes a 4-wide node: its record (binary references)
    int *stacked;     // stack entries held on arrival
    int *size4;       // 4-wide nodes in its subtree
    int *idx4;        // its 4-wide index
    int *lvl;         // nodes of each level, concatenated
    int *counters;
};

__global__ void k_open(CollapseCtx C, int off, int cnt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    int b = C.lvl[off + i];
    // every inner reference must name a binary node; open4 follows two levels
    const BvhNode nd = C.bin[b];
    for (int c = 0; c < 2; c++) {
        int r = c == 0 ? nd.ch.x : nd.ch.y;
        if (r >= C.nbin || r == 0) {
            atomicOr(&C.counters[CN_ERR], 4);
            return;
        }
        if (r > 0) {
            const BvhNode g = C.bin[r];
            if (g.ch.x >= C.nbin || g.ch.y >= C.nbin || g.ch.x == 0 || g.ch.y == 0) {
                atomicOr(&C.counters[CN_ERR], 4);
                return;
            }
        }
    }
    Kid kids[4];
    int nk = open4(C.bin, b, kids);
    C.w4[b] = make_bvh4(kids, nk);
    int held = C.stacked[b] + nk - 1;
    atomicMax(&C.counters[CN_STACK], held);
    for (int c = 0; c < nk; c++) {
        if (kids[c].ref < 0) continue;
        int slot = off + cnt + atomicAdd(&C.counters[CN_LEVEL], 1);
        if (slot >= C.nbin) {
            atomicOr(&C.counters[CN_ERR], 8);
            return;
        }
        C.stacked[kids[c].ref] = held;
        C.lvl[slot] = kids[c].ref;
    }
}

__device__ __forceinline__ int child_ref(const Bvh4Node &nd, int c) {
    return c == 0 ? nd.ch.x : (c == 1 ? nd.ch.y : (c == 2 ? nd.ch.z : nd.ch.w));
}

__global__ void k_size(CollapseCtx C, int off, int cnt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    int b = C.lvl[off + i];
    const Bvh4Node nd = C.w4[b];
    int s = 1;
    for (int c = 0; c < nd.meta.x; c++) {
        int r = child_ref(nd, c);
        if (r >= 0) s += C.size4[r];
    }
    C.size4[b] = s;
}

__global__ void k_index(CollapseCtx C, int off, int cnt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    int b = C.lvl[off + i];
    const Bvh4Node nd = C.w4[b];
    int run = C.idx4[b] + 1;
    for (int c = 0; c < nd.meta.x; c++) {
        int r = child_ref(nd, c);
        if (r < 0) continue;
        C.idx4[r] = run;
        run += C.size4[r];
    }
}

__global__ void k_emit(CollapseCtx C, int total, Bvh4Node *out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    int b = C.lvl[i];
    Bvh4Node nd = C.w4[b];
    int r[4] = {nd.ch.x, nd.ch.y, nd.ch.z, nd.ch.w};
    for (int c = 0; c < 4; c++)
        if (c < nd.meta.x && r[c] >= 0) r[c] = C.idx4[r[c]];
    nd.ch = make_int4(r[0], r[1], r[2], r[3]);
    out[C.idx4[b]] = nd;
}

// Temporaries carved from one allocation: a first pass of get() calls only
// counts (and returns null), commit() allocates, a second pass hands out.
struct Arena {
    char *base = nullptr;
    size_t used = 0, cap = 0;
    bool counting = true;
    template <class T>
    T *get(size_t count) {
        const size_t bytes = (std::max<size_t>(count * sizeof(T), 16) + 255) & ~size_t(255);
        T *p = counting ? nullptr : reinterpret_cast<T *>(base + used);
        used += bytes;
        if (!counting && used > cap) fail(LJX_ERR_HIP, "BVH build: temporary arena overflow");
        return p;
    }
    void commit() {
        cap = used;
        used = 0;
        counting = false;
        HIPX(hipMalloc(reinterpret_cast<void **>(&base), std::max<size_t>(cap, 256)));
    }
    ~Arena() {
        if (base) (void)hipFree(base);
    }
};

int blocks(long long n, int per) { return int((n + per - 1) / per); }

// ------------------------------------------------ primitive records
__global__ void k_prim_boxes(const float *vpos, const int *tri_vidx, int ntri, const double4 *spheres, int nsph,
                             float4 *lo, float4 *hi) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntri + nsph) return;
    float l[3], h[3];
    if (i < ntri) {
        const int a = tri_vidx[3 * i], b = tri_vidx[3 * i + 1], c = tri_vidx[3 * i + 2];
        for (int k = 0; k < 3; k++) {
            float pa = vpos[3 * a + k], pb = vpos[3 * b + k], pc = vpos[3 * c + k];
            l[k] = fminf(pa, fminf(pb, pc)) + 0.0f;
            h[k] = fmaxf(pa, fmaxf(pb, pc)) + 0.0f;
        }
    } else {
        const double4 s = spheres[i - ntri];
        const double c[3] = {s.x, s.y, s.z};
        for (int k = 0; k < 3; k++) {
            l[k] = float(c[k] - s.w) + 0.0f;
            h[k] = float(c[k] + s.w) + 0.0f;
        }
    }
    lo[i] = make_float4(l[0], l[1], l[2], 0.f);
    hi[i] = make_float4(h[0], h[1], h[2], 0.f);
}

__global__ void k_slot_prims(const int *slot_prim, int n, const float *vpos, const int *tri_vidx, int ntri,
                             const double4 *spheres, float4 *prims) {
    int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= n) return;
    const int gid = slot_prim[slot];
    if (gid < ntri) {
        for (int k = 0; k < 3; k++) {
            const float *p = vpos + 3 * tri_vidx[3 * gid + k];
            float w = k == 0 ? __int_as_float(gid) : (k == 1 ? __int_as_float(kPrimTri) : 0.f);
            prims[3 * slot + k] = make_float4(p[0], p[1], p[2], w);
        }
    } else {
        const double4 s = spheres[gid - ntri];
        prims[3 * slot] = make_float4(float(s.x), float(s.y), float(s.z), __int_as_float(gid));
        prims[3 * slot + 1] = make_float4(__int_as_float(gid - ntri), 0.f, 0.f, __int_as_float(kPrimSphere));
        prims[3 * slot + 2] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

__global__ void k_slot_buckets(const int *slot_prim, int n, const int *prim_shape, const int8_t *shape_bucket,
                               int8_t *out) {
    int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot < n) out[slot] = shape_bucket[prim_shape[slot_prim[slot]]];
}

}  // namespace

void prim_boxes_device(const float *vpos, const int *tri_vidx, int ntri, const double4 *spheres, int nsph,
                       float4 *lo, float4 *hi, hipStream_t st) {
    if (ntri + nsph <= 0) return;
    hipLaunchKernelGGL(k_prim_boxes, dim3(blocks(ntri + nsph, 256)), dim3(256), 0, st, vpos, tri_vidx, ntri, spheres,
                       nsph, lo, hi);
    HIPX(hipGetLastError());
}

void slot_prims_device(const int *slot_prim, int n, const float *vpos, const int *tri_vidx, int ntri,
                       const double4 *spheres, float4 *prims, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_slot_prims, dim3(blocks(n, 256)), dim3(256), 0, st, slot_prim, n, vpos, tri_vidx, ntri,
                       spheres, prims);
    HIPX(hipGetLastError());
}

void slot_buckets_device(const int *slot_prim, int n, const int *prim_shape, const int8_t *shape_bucket, int8_t *out,
                         hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_slot_buckets, dim3(blocks(n, 256)), dim3(256), 0, st, slot_prim, n, prim_shape,
                       shape_bucket, out);
    HIPX(hipGetLastError());
}

void build_bvh_device(const float4 *box_lo, const float4 *box_hi, int n, int max_leaf, float node_cost, hipStream_t st,
                      const std::function<void *(size_t)> &keep, DeviceBvh &out) {
    auto t0 = std::chrono::steady_clock::now();
    out = DeviceBvh{};
    if (n <= 0) return;
    if (n >= (1 << 27)) fail(LJX_ERR_UNSUPPORTED, "more primitives than a leaf reference can address");
    if (max_leaf < 1 || max_leaf > 8) max_leaf = 8;
    BuildCtx B;
    // build-phase temporaries
    const int nb_cap = std::max(n, 2);
    const int max_items = n / kSmall + 2;
    const int max_chunks = n / kChunk + max_items + 1;
    Arena A;
    float4 *cen;
    Item *wide[2], *narrow;
    int *cstart, *ccount, *ibox, *ibins;
    LevelSplit *splits;
    for (int pass = 0; pass < 2; pass++) {
        cen = A.get<float4>(n);
        B.tmp = A.get<int>(n);
        B.nodes = A.get<BvhNode>(nb_cap);
        B.counters = A.get<int>(CN_NUM);
        wide[0] = A.get<Item>(max_items);
        wide[1] = A.get<Item>(max_items);
        narrow = A.get<Item>(size_t(n) + 1);
        cstart = A.get<int>(max_items + 1);
        ccount = A.get<int>(max_chunks);
        ibox = A.get<int>(size_t(max_items) * 12);
        ibins = A.get<int>(size_t(max_items) * kBinInts);
        splits = A.get<LevelSplit>(max_items);
        if (pass == 0) A.commit();
    }
    B.blo = box_lo;
    B.bhi = box_hi;
    B.cen = cen;
    B.ids = static_cast<int *>(keep(size_t(n) * sizeof(int)));
    B.max_leaf = max_leaf;
    B.node_cost = node_cost;
    HIPX(hipMemsetAsync(B.nodes, 0, sizeof(BvhNode) * nb_cap, st));
    int init_ctr[CN_NUM] = {1, 0, 0, 0, 0, 0, 0, 0};  // node 0 is the root
    HIPX(hipMemcpyAsync(B.counters, init_ctr, sizeof(init_ctr), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_init_prims, dim3(blocks(n, 256)), dim3(256), 0, st, box_lo, box_hi, cen, B.ids, n);
    HIPX(hipGetLastError());

    // ---- wide ranges, level by level
    Item root{0, n, 1, -1};
    int nwide = 0, nnarrow = 0;
    if (n > kSmall) {
        HIPX(hipMemcpyAsync(wide[0], &root, sizeof(Item), hipMemcpyHostToDevice, st));
        nwide = 1;
    } else {
        HIPX(hipMemcpyAsync(narrow, &root, sizeof(Item), hipMemcpyHostToDevice, st));
        nnarrow = 1;
        HIPX(hipMemcpyAsync(B.counters + CN_NARROW, &nnarrow, sizeof(int), hipMemcpyHostToDevice, st));
    }
    int cur = 0;
    while (nwide > 0) {
        const Item *items = wide[cur];
        Item *next = wide[cur ^ 1];
        const int grid = std::min(max_chunks, n / kChunk + nwide + 1);
        hipLaunchKernelGGL(k_chunk_scan, dim3(1), dim3(1024), 0, st, items, nwide, cstart);
        hipLaunchKernelGGL(k_item_clear, dim3(nwide), dim3(256), 0, st, ibox, ibins, nwide);
        hipLaunchKernelGGL(k_wide_bounds, dim3(grid), dim3(256), 0, st, B, items, nwide, cstart, ibox);
        hipLaunchKernelGGL(k_wide_bin, dim3(grid), dim3(256), 0, st, B, items, nwide, cstart, ibox, ibins);
        HIPX(hipMemsetAsync(B.counters + CN_WIDE, 0, sizeof(int), st));
        hipLaunchKernelGGL(k_wide_split, dim3(nwide), dim3(64), 0, st, B, items, nwide, ibox, ibins, splits, next,
                           narrow);
        hipLaunchKernelGGL(k_wide_count, dim3(grid), dim3(256), 0, st, B, items, nwide, cstart, splits, ccount);
        hipLaunchKernelGGL(k_wide_offsets, dim3(blocks(nwide, 64)), dim3(64), 0, st, items, nwide, cstart, ccount);
        hipLaunchKernelGGL(k_wide_scatter, dim3(grid), dim3(256), 0, st, B, items, nwide, cstart, splits, ccount);
        hipLaunchKernelGGL(k_wide_copy, dim3(grid), dim3(256), 0, st, B, items, nwide, cstart, splits);
        HIPX(hipGetLastError());
        HIPX(hipMemcpyAsync(&nwide, B.counters + CN_WIDE, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPX(hipStreamSynchronize(st));
        if (nwide > max_items) fail(LJX_ERR_HIP, "BVH build: wide-range list overflow");
        cur ^= 1;
    }
    HIPX(hipMemcpyAsync(&nnarrow, B.counters + CN_NARROW, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
    if (nnarrow > 0) hipLaunchKernelGGL(k_narrow, dim3(nnarrow), dim3(64), 0, st, B, narrow, nnarrow);
    HIPX(hipGetLastError());
    int ctr[CN_NUM];
    HIPX(hipMemcpyAsync(ctr, B.counters, sizeof(ctr), hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
    if (ctr[CN_STACK]) fail(LJX_ERR_HIP, "BVH build: narrow-range stack overflow");
    if (ctr[CN_ERR]) fail(LJX_ERR_HIP, "BVH build: inconsistent partition (" + std::to_string(ctr[CN_ERR]) + ")");
    out.max_binary_depth = ctr[CN_DEPTH];
    if (out.max_binary_depth > kMaxTreeDepth) fail(LJX_ERR_UNSUPPORTED, "BVH deeper than the build bound");
    const int nbin = ctr[CN_NODES];

    // ---- collapse to 4-wide nodes, depth-first preorder
    CollapseCtx C;
    C.bin = B.nodes;
    C.nbin = nbin;
    Arena A2;
    for (int pass = 0; pass < 2; pass++) {
        C.w4 = A2.get<Bvh4Node>(nbin);
        C.stacked = A2.get<int>(nbin);
        C.size4 = A2.get<int>(nbin);
        C.idx4 = A2.get<int>(nbin);
        C.lvl = A2.get<int>(nbin);
        if (pass == 0) A2.commit();
    }
    C.counters = B.counters;
    HIPX(hipMemsetAsync(C.stacked, 0, sizeof(int), st));
    HIPX(hipMemsetAsync(C.idx4, 0, sizeof(int), st));
    HIPX(hipMemsetAsync(C.counters + CN_STACK, 0, sizeof(int), st));
    int zero = 0;
    HIPX(hipMemcpyAsync(C.lvl, &zero, sizeof(int), hipMemcpyHostToDevice, st));
    std::vector<int> off{0}, cnt{1};
    while (true) {
        HIPX(hipMemsetAsync(C.counters + CN_LEVEL, 0, sizeof(int), st));
        hipLaunchKernelGGL(k_open, dim3(blocks(cnt.back(), 256)), dim3(256), 0, st, C, off.back(), cnt.back());
        HIPX(hipGetLastError());
        int nx = 0;
        HIPX(hipMemcpyAsync(&nx, C.counters + CN_LEVEL, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPX(hipStreamSynchronize(st));
        int err = 0;
        HIPX(hipMemcpy(&err, C.counters + CN_ERR, sizeof(int), hipMemcpyDeviceToHost));
        if (err) fail(LJX_ERR_HIP, "BVH collapse: malformed binary tree (" + std::to_string(err) + ")");
        if (nx == 0) break;
        if (off.back() + cnt.back() + nx > nbin) fail(LJX_ERR_HIP, "BVH collapse: more 4-wide nodes than binary ones");
        off.push_back(off.back() + cnt.back());
        cnt.push_back(nx);
    }
    const int levels = int(cnt.size());
    const int total = off.back() + cnt.back();
    for (int l = levels - 1; l >= 0; l--)
        hipLaunchKernelGGL(k_size, dim3(blocks(cnt[l], 256)), dim3(256), 0, st, C, off[l], cnt[l]);
    for (int l = 0; l < levels; l++)
        hipLaunchKernelGGL(k_index, dim3(blocks(cnt[l], 256)), dim3(256), 0, st, C, off[l], cnt[l]);
    out.nodes = static_cast<Bvh4Node *>(keep(size_t(total) * sizeof(Bvh4Node)));
    hipLaunchKernelGGL(k_emit, dim3(blocks(total, 256)), dim3(256), 0, st, C, total, out.nodes);
    HIPX(hipGetLastError());
    int sn = 0;
    HIPX(hipMemcpyAsync(&sn, C.counters + CN_STACK, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPX(hipStreamSynchronize(st));
    out.num_nodes = total;
    out.slot_prim = B.ids;
    out.num_slots = n;
    out.depth = levels;
    out.stack_need = sn;
    out.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace ljx
