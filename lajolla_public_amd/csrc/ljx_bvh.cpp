// Binned-SAH BVH builder (host).  Top-down, 32 bins per axis over primitive
// centroids; subtrees above a size threshold are built on separate threads
// into private node arrays and spliced, so the output is independent of the
// thread count.  Deep ranges fall back to object-median splits so the tree
// depth stays bounded (kMaxTreeDepth).
#include "ljx_bvh.h"
#include "ljx_trace.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <thread>

namespace ljx {

namespace {

constexpr int kBins = 32;
constexpr float kInf = std::numeric_limits<float>::infinity();
constexpr int kMedianDepth = 80;

struct Box {
    float lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
    void grow(const float *l, const float *h) {
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], l[k]);
            hi[k] = std::max(hi[k], h[k]);
        }
    }
    void grow(const Box &b) { grow(b.lo, b.hi); }
    float half_area() const {
        float d0 = hi[0] - lo[0], d1 = hi[1] - lo[1], d2 = hi[2] - lo[2];
        if (!(d0 >= 0)) return 0;
        return d0 * d1 + d1 * d2 + d2 * d0;
    }
};

struct Builder {
    const std::vector<PrimBox> &boxes;
    std::vector<float> cen;  // 3 per primitive
    std::vector<int> ids;
    int max_leaf;
    std::atomic<int> threads_left;
    std::atomic<int> max_depth{0};

    Builder(const std::vector<PrimBox> &b, int ml, int threads) : boxes(b), max_leaf(ml), threads_left(threads) {
        cen.resize(b.size() * 3);
        ids.resize(b.size());
        for (size_t i = 0; i < b.size(); i++) {
            for (int k = 0; k < 3; k++) cen[3 * i + k] = 0.5f * (b[i].lo[k] + b[i].hi[k]);
            ids[i] = int(i);
        }
    }

    static int leaf_ref(int first, int cnt) { return ~((first << 3) | (cnt - 1)); }

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
        if (n <= 1) return leaf_ref(begin, n);

        int axis = -1, split_bin = -1;
        float best = kInf;
        float ext[3];
        for (int k = 0; k < 3; k++) ext[k] = cb.hi[k] - cb.lo[k];
        if (depth < kMedianDepth) {
            int cnt[3][kBins] = {};
            Box bins[3][kBins];
            for (int i = begin; i < end; i++) {
                const float *c = &cen[3 * ids[i]];
                const PrimBox &p = boxes[ids[i]];
                for (int k = 0; k < 3; k++) {
                    if (!(ext[k] > 0)) continue;
                    int b = std::min(kBins - 1, int(kBins * (c[k] - cb.lo[k]) / ext[k]));
                    cnt[k][b]++;
                    bins[k][b].grow(p.lo, p.hi);
                }
            }
            for (int k = 0; k < 3; k++) {
                if (!(ext[k] > 0)) continue;
                float rarea[kBins];
                int rcnt[kBins];
                Box acc;
                int c = 0;
                for (int b = kBins - 1; b > 0; b--) {
                    acc.grow(bins[k][b]);
                    c += cnt[k][b];
                    rarea[b] = acc.half_area();
                    rcnt[b] = c;
                }
                Box lacc;
                c = 0;
                for (int b = 0; b < kBins - 1; b++) {
                    lacc.grow(bins[k][b]);
                    c += cnt[k][b];
                    if (c == 0 || rcnt[b + 1] == 0) continue;
                    float cost = c * lacc.half_area() + rcnt[b + 1] * rarea[b + 1];
                    if (cost < best) {
                        best = cost;
                        axis = k;
                        split_bin = b;
                    }
                }
            }
        }
        float pa = bb.half_area();
        // SAH in units of one primitive test; a node visit costs one test.
        float split_cost = pa > 0 ? 1.0f + best / pa : kInf;
        float leaf_cost = float(n);
        if (n <= max_leaf && (axis < 0 || leaf_cost <= split_cost)) return leaf_ref(begin, n);

        int mid;
        if (axis >= 0) {
            int *b0 = ids.data() + begin;
            int *b1 = ids.data() + end;
            int *m = std::partition(b0, b1, [&](int id) {
                float c = cen[3 * id + axis];
                int b = std::min(kBins - 1, int(kBins * (c - cb.lo[axis]) / ext[axis]));
                return b <= split_bin;
            });
            mid = int(m - ids.data());
        } else {
            mid = begin;
        }
        if (mid == begin || mid == end) {
            // object median on the widest centroid axis
            int k = 0;
            if (ext[1] > ext[k]) k = 1;
            if (ext[2] > ext[k]) k = 2;
            mid = begin + n / 2;
            std::nth_element(ids.begin() + begin, ids.begin() + mid, ids.begin() + end,
                             [&](int a, int b) { return cen[3 * a + k] < cen[3 * b + k]; });
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

struct Child {
    int ref;
    float lo[3], hi[3];
    float half_area() const {
        float d0 = hi[0] - lo[0], d1 = hi[1] - lo[1], d2 = hi[2] - lo[2];
        return d0 * d1 + d1 * d2 + d2 * d0;
    }
};

void binary_children(const BvhNode &nd, Child &a, Child &b) {
    a.ref = nd.ch.x;
    a.lo[0] = nd.lxy.x, a.hi[0] = nd.lxy.y, a.lo[1] = nd.lxy.z, a.hi[1] = nd.lxy.w, a.lo[2] = nd.zz.x, a.hi[2] = nd.zz.y;
    b.ref = nd.ch.y;
    b.lo[0] = nd.rxy.x, b.hi[0] = nd.rxy.y, b.lo[1] = nd.rxy.z, b.hi[1] = nd.rxy.w, b.lo[2] = nd.zz.z, b.hi[2] = nd.zz.w;
}

struct Collapser {
    const std::vector<BvhNode> &bin;
    std::vector<Bvh4Node> out;
    int depth = 0, stack_need = 0;

    // Emits the 4-wide node for binary node b; returns its index.  level and
    // stacked: depth of this node and stack entries held on arrival.
    int emit(int b, int level, int stacked) {
        std::vector<Child> kids(2);
        binary_children(bin[b], kids[0], kids[1]);
        while (kids.size() < 4) {
            int best = -1;
            float area = -1;
            for (size_t i = 0; i < kids.size(); i++)
                if (kids[i].ref >= 0 && kids[i].half_area() > area) area = kids[i].half_area(), best = int(i);
            if (best < 0) break;
            Child a, c;
            binary_children(bin[kids[best].ref], a, c);
            kids[best] = a;
            kids.insert(kids.begin() + best + 1, c);
        }
        int me = int(out.size());
        out.emplace_back();
        depth = std::max(depth, level);
        // a traversal entering one child may have pushed all the others
        int held = stacked + int(kids.size()) - 1;
        stack_need = std::max(stack_need, held);
        float lo[3][4], hi[3][4];
        int ch[4];
        for (int c = 0; c < 4; c++) {
            for (int k = 0; k < 3; k++) lo[k][c] = kInf, hi[k][c] = -kInf;
            ch[c] = kNoChild;
        }
        for (size_t c = 0; c < kids.size(); c++) {
            for (int k = 0; k < 3; k++) lo[k][c] = kids[c].lo[k], hi[k][c] = kids[c].hi[k];
            ch[c] = kids[c].ref >= 0 ? emit(kids[c].ref, level + 1, held) : kids[c].ref;
        }
        Bvh4Node &nd = out[me];
        nd.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
        nd.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
        nd.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
        nd.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
        nd.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
        nd.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
        nd.ch = make_int4(ch[0], ch[1], ch[2], ch[3]);
        nd.meta = make_int4(int(kids.size()), 0, 0, 0);
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

BvhBuildResult build_bvh_sah(const std::vector<PrimBox> &boxes, int max_leaf, int threads) {
    BvhBuildResult res;
    if (boxes.empty()) return res;
    if (max_leaf < 1 || max_leaf > 8) max_leaf = 8;
    if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
    Builder b(boxes, max_leaf, threads);
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
