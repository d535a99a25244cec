// Binned-SAH BVH builder (host, multi-threaded) producing the two-box-per-node
// layout the HIP traversal reads.  Replaces the BVH build inside Embree's
// rtcCommitScene (src/scene.cpp:20-27).
#pragma once

#include <cstdint>
#include <vector>

#include "ljx_device.h"

namespace ljx {

struct PrimBox {
    float lo[3], hi[3];
};

struct BvhBuildResult {
    std::vector<BvhNode> nodes;   // node 0 is the root (always an inner node)
    std::vector<int> slot_prim;   // leaf slot -> primitive index
    int max_depth = 0;
};

// Builds over n primitive boxes.  max_leaf <= 8; node_cost: the SAH cost of
// a node visit in primitive tests.  threads <= 0: all cores.
BvhBuildResult build_bvh_sah(const std::vector<PrimBox> &boxes, int max_leaf, float node_cost, int threads);

struct Bvh4Result {
    std::vector<Bvh4Node> nodes;  // node 0 is the root
    int depth = 0;                // nodes on the longest root-to-leaf path
    int stack_need = 0;           // most stack entries a traversal can hold
};

// Collapses a binary BVH into a 4-wide one: each node takes its children's
// largest-area inner nodes apart until it has four children or only leaves.
// Leaves (primitive slot ranges) are unchanged.
Bvh4Result collapse_bvh4(const BvhBuildResult &bin);

}  // namespace ljx
