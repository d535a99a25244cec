// The BVH built on the device (ljx_bvh_gpu.hip): the same tree as the host
// builder (ljx_bvh.h), from primitive boxes already in device memory.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>

#include "ljx_device.h"

namespace ljx {

struct DeviceBvh {
    Bvh4Node *nodes = nullptr;  // 4-wide nodes, depth-first preorder, node 0 the root
    int num_nodes = 0;
    int *slot_prim = nullptr;   // leaf slot -> primitive index
    int num_slots = 0;
    int depth = 0;              // 4-wide nodes on the longest root-to-leaf path
    int stack_need = 0;         // most stack entries a traversal can hold
    int max_binary_depth = 0;
    double ms = 0;              // wall time of the build
};

// Builds over n primitive boxes (lo, hi; w unused); max_leaf and node_cost
// as for build_bvh_sah (ljx_bvh.h).  keep(bytes) allocates
// device memory that outlives the build (the nodes and the slot table);
// temporaries are freed before returning.  Work is queued on stream, with a
// host wait per level.
void build_bvh_device(const float4 *box_lo, const float4 *box_hi, int n, int max_leaf, float node_cost, hipStream_t stream,
                      const std::function<void *(size_t)> &keep, DeviceBvh &out);

// Boxes of the scene's primitives: triangles 0..ntri-1 from the global vertex
// positions (3 floats each) and vertex ids (3 per triangle), then the spheres
// (centre, radius in FP64).  Signed zeros are made positive.
void prim_boxes_device(const float *vpos, const int *tri_vidx, int ntri, const double4 *spheres, int nsph,
                       float4 *lo, float4 *hi, hipStream_t stream);

// Leaf-slot records, 3 float4 per slot: a triangle's vertices with the
// primitive id and kind in w (v0|gid, v1|kind, v2|0); a sphere's centre|gid,
// sphere index|kind, 0.
void slot_prims_device(const int *slot_prim, int n, const float *vpos, const int *tri_vidx, int ntri,
                       const double4 *spheres, float4 *prims, hipStream_t stream);

// Per slot: the shading bucket of its shape (shape_bucket), through prim_shape.
void slot_buckets_device(const int *slot_prim, int n, const int *prim_shape, const int8_t *shape_bucket,
                         int8_t *out, hipStream_t stream);

}  // namespace ljx
