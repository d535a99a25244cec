// Binned-SAH BVH builder (host).  Top-down, 32 bins per axis over primitive
// centroids; subtrees above a size threshold are built on separate threads
// into private node arrays and spliced, so the output is independent of the
// thread count.  Split decisions come from ljx_bvh_sah.h, which the device
// builder (ljx_bvh_gpu.hip) shares, and partitions are stable, so both build
// the same tree.  Ranges without a binned split (all centroids equal, or
// deeper than kBvhMedianDepth) are halved in order, which bounds the depth
// (kMaxTreeDepth).
#include "ljx_bvh.h"
#include "ljx_bvh_sah.h"
#include "ljx_trace.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <thread>

namespace ljx {

namespace {

constexpr float kInf = std::numeric_limits<float>::infinity();

struct Box {
    float lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
    void grow(const float *l, const float *h) {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], l[k]);
            hi[k] = std::max(hi[k], h[k]);
        }
    }
    float half_area() const { return box_half_area(lo, hi); }
};

struct HostBins {
    int cnt[3][kBvhBins];
    float lo[3][kBvhBins][3], hi[3][kBvhBins][3];
    HostBins() {
        for (int k = 0; k < 3; k++)
            for (int b = 0; b < kBvhBins; b++) {
                cnt[k][b] = 0;
                for (int d = 0; d < 3; d++) lo[k][b][d] = kInf, hi[k][b][d] = -kInf;
            }
    }
    __host__ __device__ int count(int k, int b) const { return cnt[k][b]; }
    __host__ __device__ void box(int k, int b, float l[3], float h[3]) const {
        for (int d = 0; d < 3; d++) l[d] = lo[k][b][d], h[d] = hi[k][b][d];
    }
};

struct Builder {
    const std::vector<PrimBox> &boxes;
    std::vector<float> cen;  // 3 per primitive
    std::vector<int> ids;
    int max_leaf;
    float node_cost;
    std::atomic<int> threads_left;
    std::atomic<int> max_depth{0};

    Builder(const std::vector<PrimBox> &b, int ml, float nc, int threads)
        : boxes(b), max_leaf(ml), node_cost(nc), threads_left(threads) {
        cen.resize(b.size() * 3);
        ids.resize(b.size());
        for (size_t i = 0; i < b.size(); i++) {
            for (int k = 0; k < 3; k++) cen[3 * i + k] = 0.5f * (b[i].lo[k] + b[i].hi[k]);
            ids[i] = int(i);
        }
    }

    // Builds [begin, end) into out (local node indices); returns the child
    // reference and writes the range's box.
    int build(int begin, int end, int depth, std::vector<BvhNode> &out, Box &box_out) {
        int n = end - begin;
        Box bb, cb;
        for (int i = begin; i < end; i++) {
            const PrimBox &p = boxes[ids[i]];
            bb.grow(p.lo, p.hi);
            const float *c = &cen[3 * ids[i]];
            cb.grow(c, c);
        }
        box_out = bb;
        int cur = max_depth.load();
        while (depth > cur && !max_depth.compare_exchange_weak(cur, depth)) {
        }
        if (n <= 1) return bvh_leaf_ref(begin, n);

        float ext[3];
        for (int k = 0; k < 3; k++) ext[k] = cb.hi[k] - cb.lo[k];
        SahSplit split;
        if (depth < kBvhMedianDepth) {
            HostBins bins;
            for (int i = begin; i < end; i++) {
                const float *c = &cen[3 * ids[i]];
                const PrimBox &p = boxes[ids[i]];
                for (int k = 0; k < 3; k++) {
                    if (!(ext[k] > 0)) continue;
                    int b = sah_bin(c[k], cb.lo[k], ext[k]);
                    bins.cnt[k][b]++;
                    box_grow(bins.lo[k][b], bins.hi[k][b], p.lo, p.hi);
                }
            }
            for (int k = 0; k < 3; k++)
                if (ext[k] > 0) sah_take(split, sah_axis_best(bins, k));
        }
        if (sah_make_leaf(n, max_leaf, split, bb.half_area(), node_cost)) return bvh_leaf_ref(begin, n);

        int mid;
        if (split.axis >= 0) {
            const int axis = split.axis, sb = split.bin;
            const float clo = cb.lo[axis], cext = ext[axis];
            std::stable_partition(ids.begin() + begin, ids.begin() + end,
                                  [&](int id) { return sah_bin(cen[3 * id + axis], clo, cext) <= sb; });
            mid = begin + split.nleft;
        } else {
            mid = begin + n / 2;
        }

        int me = int(out.size());
        out.push_back(BvhNode{});
        Box lb, rb;
        int lref, rref;
        bool par = n > 100000 && threads_left.fetch_sub(1) > 1;
        if (par) {
            std::vector<BvhNode> L;
            std::thread t([&]() { lref = build(begin, mid, depth + 1, L, lb); });
            std::vector<BvhNode> R;
            rref = build(mid, end, depth + 1, R, rb);
            t.join();
            threads_left.fetch_add(1);
            lref = splice(out, L, lref);
            rref = splice(out, R, rref);
        } else {
            if (n > 100000) threads_left.fetch_add(1);
            lref = build(begin, mid, depth + 1, out, lb);
            rref = build(mid, end, depth + 1, out, rb);
        }
        BvhNode &nd = out[me];
        nd.lxy = make_float4(lb.lo[0], lb.hi[0], lb.lo[1], lb.hi[1]);
        nd.rxy = make_float4(rb.lo[0], rb.hi[0], rb.lo[1], rb.hi[1]);
        nd.zz = make_float4(lb.lo[2], lb.hi[2], rb.lo[2], rb.hi[2]);
        nd.ch = make_int4(lref, rref, 0, 0);
        return me;
    }

    static int splice(std::vector<BvhNode> &out, const std::vector<BvhNode> &sub, int ref) {
        int off = int(out.size());
        for (BvhNode nd : sub) {
            if (nd.ch.x >= 0) nd.ch.x += off;
            if (nd.ch.y >= 0) nd.ch.y += off;
            out.push_back(nd);
        }
        return ref >= 0 ? ref + off : ref;
    }
};

}  // namespace

namespace {

struct Collapser {
    const std::vector<BvhNode> &bin;
    std::vector<Bvh4Node> out;
    int depth = 0, stack_need = 0;

    // Emits the 4-wide node for binary node b in depth-first preorder;
    // returns its index.  level and stacked: depth of this node and stack
    // entries held on arrival.
    int emit(int b, int level, int stacked) {
        Kid kids[4];
        int nk = open4(bin.data(), b, kids);
        int me = int(out.size());
        out.emplace_back();
        depth = std::max(depth, level);
        // a traversal entering one child may have pushed all the others
        int held = stacked + nk - 1;
        stack_need = std::max(stack_need, held);
        for (int c = 0; c < nk; c++)
            if (kids[c].ref >= 0) kids[c].ref = emit(kids[c].ref, level + 1, held);
        out[me] = make_bvh4(kids, nk);
        return me;
    }
};

}  // namespace

Bvh4Result collapse_bvh4(const BvhBuildResult &bin) {
    Bvh4Result r;
    if (bin.nodes.empty()) return r;
    Collapser c{bin.nodes, {}, 0, 0};
    c.out.reserve(bin.nodes.size() / 2 + 1);
    c.emit(0, 1, 0);
    r.nodes.swap(c.out);
    r.depth = c.depth;
    r.stack_need = c.stack_need;
    return r;
}

BvhBuildResult build_bvh_sah(const std::vector<PrimBox> &boxes, int max_leaf, float node_cost, int threads) {
    BvhBuildResult res;
    if (boxes.empty()) return res;
    if (max_leaf < 1 || max_leaf > 8) max_leaf = 8;
    if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
    Builder b(boxes, max_leaf, node_cost, threads);
    std::vector<BvhNode> nodes;
    nodes.reserve(boxes.size());
    Box root_box;
    int ref = b.build(0, int(boxes.size()), 1, nodes, root_box);
    if (ref < 0) {
        // a single leaf: wrap it in a root whose two children are that leaf
        BvhNode nd;
        nd.lxy = make_float4(root_box.lo[0], root_box.hi[0], root_box.lo[1], root_box.hi[1]);
        nd.rxy = nd.lxy;
        nd.zz = make_float4(root_box.lo[2], root_box.hi[2], root_box.lo[2], root_box.hi[2]);
        nd.ch = make_int4(ref, ref, 0, 0);
        nodes.push_back(nd);
    }
    res.nodes.swap(nodes);
    res.slot_prim = b.ids;
    res.max_depth = b.max_depth.load();
    if (res.max_depth > kMaxTreeDepth) throw std::runtime_error("BVH deeper than the build bound");
    return res;
}

}  // namespace ljx
