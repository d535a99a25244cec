// BVH traversal for gfx950: closest hit and any hit, FP32, one ray per lane.
//
// Replaces Embree's rtcIntersect1 / rtcOccluded1 as lajolla calls them
// (src/intersection.cpp:7-85).  Node layout: a 4-wide BVH, the four child
// boxes of a node as outward-rounded half offsets from a common corner, in
// 80 bytes (ljx_device.h Bvh4QNode, quantised from the builders' FP32
// Bvh4Node), loaded as five 16-byte vector loads.  For closest hits the hit
// children are ordered near to far with a five-comparator sorting network;
// any-hit rays take them in slot order.  The traversal stack lives in LDS
// (kStackLds entries per lane, lane-interleaved so a wave's pushes hit 64
// consecutive banks) and spills to a per-lane global area beyond that, sized
// per scene from the tree's worst-case stack need.  Triangles use the
// watertight test of Woop, Benthin and Wald (JCGT 2013) with an FP64 re-test
// of edge hits.
#pragma once

#include "ljx_device.h"

namespace ljx {

// 20 LDS entries x 256 lanes = 20 KB per block: eight traversal blocks fit
// the 160 KB of LDS of a CU, i.e. eight waves per SIMD.
constexpr int kStackLds = 20;       // per-lane LDS stack entries
constexpr int kMaxStackSpill = 256; // largest per-lane spill a scene may need
constexpr int kMaxTreeDepth = 118;  // binary build depth bound (object-median splits below)
constexpr int kPrimTri = 0, kPrimSphere = 1;

struct TraceHit {
    float t, u, v;
    int slot;
};

struct RayPre {
    float ox, oy, oz;
    float idx, idy, idz;
    float oidx, oidy, oidz;  // o * (1/d)
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

__device__ __forceinline__ float sel3(float a, float b, float c, int k) { return k == 0 ? a : (k == 1 ? b : c); }

__device__ __forceinline__ RayPre ray_setup(float ox, float oy, float oz, float dx, float dy, float dz) {
    RayPre r;
    r.ox = ox;
    r.oy = oy;
    r.oz = oz;
    // keep 1/d finite so that 0 * inf never arises in the slab test
    const float tiny = 1e-20f;
    float sdx = fabsf(dx) < tiny ? copysignf(tiny, dx) : dx;
    float sdy = fabsf(dy) < tiny ? copysignf(tiny, dy) : dy;
    float sdz = fabsf(dz) < tiny ? copysignf(tiny, dz) : dz;
    r.idx = 1.0f / sdx;
    r.idy = 1.0f / sdy;
    r.idz = 1.0f / sdz;
    r.oidx = ox * r.idx;
    r.oidy = oy * r.idy;
    r.oidz = oz * r.idz;
    float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    float dkz = sel3(dx, dy, dz, kz);
    if (dkz < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    r.Sx = sel3(dx, dy, dz, kx) / dkz;
    r.Sy = sel3(dx, dy, dz, ky) / dkz;
    r.Sz = 1.0f / dkz;
    return r;
}

// Half value of bits [16*HI, 16*HI+16) of w, widened to FP32 (exactly).  As
// an FMA operand the widening folds into v_fma_mix_f32.
template <int HI>
__device__ __forceinline__ float half_at(uint32_t w) {
    return float(__builtin_bit_cast(_Float16, uint16_t(HI ? (w >> 16) : (w & 0xffffu))));
}

// The half planes of one axis of a quantised node, ordered for the ray:
// n = the near planes' words (children 0|1, 2|3), f = the far planes'.  Along
// a positive direction the lower planes are the near ones, along a negative
// one the upper planes; this is the order min/max of the two slab distances
// would give, since an FMA with one positive (negative) factor is monotone
// in the other.
struct AxisPlanes {
    uint32_t n0, n1, f0, f1;
};
__device__ __forceinline__ AxisPlanes axis_planes(uint4 q, bool neg) {
    return AxisPlanes{neg ? q.z : q.x, neg ? q.w : q.y, neg ? q.x : q.z, neg ? q.y : q.w};
}

// Slab test of child C of a quantised node (Bvh4QNode).  s*: 1/d scaled by
// the node's scale; f*: the distance to the node's lower corner plane
// (org * 1/d - o * 1/d).  Returns the entry distance, or +inf on a miss.
// The far distance is widened by 1 + 2 * gamma(4) in FP32 (Ize 2013 widens by
// 1 + 2 * gamma(3)); the extra rounding of f* is covered by that and by the
// quantiser's margin.
template <int C>
__device__ __forceinline__ float qbox_entry(float sx, float fx, float sy, float fy, float sz, float fz,
                                            const AxisPlanes &px, const AxisPlanes &py, const AxisPlanes &pz,
                                            float tmin, float tmax) {
    constexpr int W = C >> 1, H = C & 1;
    auto nr = [](const AxisPlanes &p) { return half_at<H>(W ? p.n1 : p.n0); };
    auto fr = [](const AxisPlanes &p) { return half_at<H>(W ? p.f1 : p.f0); };
    const float tnx = __builtin_fmaf(nr(px), sx, fx), tfx = __builtin_fmaf(fr(px), sx, fx);
    const float tny = __builtin_fmaf(nr(py), sy, fy), tfy = __builtin_fmaf(fr(py), sy, fy);
    const float tnz = __builtin_fmaf(nr(pz), sz, fz), tfz = __builtin_fmaf(fr(pz), sz, fz);
    const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tmin));
    float tf = fminf(fminf(tfx, tfy), fminf(tfz, tmax));
    tf *= 1.00000048f;
    return tn <= tf ? tn : __builtin_huge_valf();
}

// a * b - c * d with both products rounded.  The fused form the compiler
// would make (one product exact, the other rounded) is not antisymmetric:
// for an edge two triangles share it gives both the same small nonzero value
// instead of exact negatives, and a ray through the edge misses both.
__device__ __forceinline__ float edge_fn(float a, float b, float c, float d) {
#pragma clang fp contract(off)
    return a * b - c * d;
}

// Watertight ray/triangle test (Woop, Benthin and Wald 2013).  On a hit
// inside (tmin, tmax] writes t and the barycentrics of v1 (u) and v2 (v).
__device__ __forceinline__ bool tri_test(const RayPre &r, float4 a4, float4 b4, float4 c4, float tmin, float tmax,
                                         float &t_out, float &u_out, float &v_out) {
    float Ax0 = a4.x - r.ox, Ay0 = a4.y - r.oy, Az0 = a4.z - r.oz;
    float Bx0 = b4.x - r.ox, By0 = b4.y - r.oy, Bz0 = b4.z - r.oz;
    float Cx0 = c4.x - r.ox, Cy0 = c4.y - r.oy, Cz0 = c4.z - r.oz;
    float Akz = sel3(Ax0, Ay0, Az0, r.kz), Bkz = sel3(Bx0, By0, Bz0, r.kz), Ckz = sel3(Cx0, Cy0, Cz0, r.kz);
    float Ax = sel3(Ax0, Ay0, Az0, r.kx) - r.Sx * Akz;
    float Ay = sel3(Ax0, Ay0, Az0, r.ky) - r.Sy * Akz;
    float Bx = sel3(Bx0, By0, Bz0, r.kx) - r.Sx * Bkz;
    float By = sel3(Bx0, By0, Bz0, r.ky) - r.Sy * Bkz;
    float Cx = sel3(Cx0, Cy0, Cz0, r.kx) - r.Sx * Ckz;
    float Cy = sel3(Cx0, Cy0, Cz0, r.ky) - r.Sy * Ckz;
    float U = edge_fn(Cx, By, Cy, Bx);
    float V = edge_fn(Ax, Cy, Ay, Cx);
    float W = edge_fn(Bx, Ay, By, Ax);
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        // edge case: redo the 2D edge functions in FP64 for watertightness
        U = float(double(Cx) * double(By) - double(Cy) * double(Bx));
        V = float(double(Ax) * double(Cy) - double(Ay) * double(Cx));
        W = float(double(Bx) * double(Ay) - double(By) * double(Ax));
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= tmin && t <= tmax)) return false;
    t_out = t;
    u_out = V * rcp;
    v_out = W * rcp;
    return true;
}

// Sphere user primitive (src/shapes/sphere.inl:12-106): the stable quadratic
// in FP64 on the FP32 ray and the FP64 centre and radius, nearest root in
// [tmin, tmax).
__device__ __forceinline__ bool sphere_roots(double ox, double oy, double oz, double dx, double dy, double dz,
                                             double4 c, double tmin, double tmax, double &t_out) {
    double vx = ox - c.x, vy = oy - c.y, vz = oz - c.z;
    double A = dx * dx + dy * dy + dz * dz;
    double B = 2 * (dx * vx + dy * vy + dz * vz);
    double Cc = vx * vx + vy * vy + vz * vz - c.w * c.w;
    double t0, t1;
    if (A == 0) {
        if (B == 0) return false;
        t0 = t1 = -Cc / B;
    } else {
        double disc = B * B - 4 * A * Cc;
        if (disc < 0) return false;
        double rd = sqrt(disc);
        if (B >= 0) {
            t0 = (-B - rd) / (2 * A);
            t1 = 2 * Cc / (-B - rd);
        } else {
            t0 = 2 * Cc / (-B + rd);
            t1 = (-B + rd) / (2 * A);
        }
    }
    if (t0 > t1) {
        double tt = t0;
        t0 = t1;
        t1 = tt;
    }
    double t = -1;
    if (t0 >= tmin && t0 < tmax) t = t0;
    if (t1 >= tmin && t1 < tmax && t < 0) t = t1;
    if (!(t >= tmin && t < tmax)) return false;
    t_out = t;
    return true;
}

struct TraceScene {
    const Bvh4QNode *__restrict__ nodes;
    const float4 *__restrict__ prims;    // 3 per leaf slot
    const double4 *__restrict__ spheres; // centre xyz, radius
    const int8_t *__restrict__ skip = nullptr;  // OPAQUE traversals: per slot, nonzero to pass through
};

// Traversal state of one ray, advanced one inner node or one leaf per step.
// stack: this lane's LDS column base (entries at stack[k*stride]); spill:
// this lane's global spill column.  SPH: the scene has sphere primitives
// (otherwise every leaf slot is a triangle).  OPAQUE (any hit only): slots
// that ts.skip marks are passed through (index-matched medium boundaries).
template <bool ANY, bool STATS, bool SPH, bool OPAQUE = false>
struct Tracer {
    RayPre r;
    float dx, dy, dz, tmin, tmax;
    int node, sp;
    bool found;
    bool passed;  // OPAQUE: a marked slot was hit and passed through
    TraceHit hit;

    __device__ __forceinline__ void start(float ox, float oy, float oz, float dx_, float dy_, float dz_, float tmin_,
                                          float tmax_) {
        r = ray_setup(ox, oy, oz, dx_, dy_, dz_);
        dx = dx_;
        dy = dy_;
        dz = dz_;
        tmin = tmin_;
        tmax = tmax_;
        node = 0;  // the root is an inner node
        sp = 0;
        found = false;
        passed = false;
        hit.slot = -1;
    }

    // One inner node or one leaf.  Returns true when the ray is finished: the
    // stack ran empty, or (any hit) an occluder was found.
    __device__ __forceinline__ bool step(const TraceScene &ts, int *stack, int stride, int *spill, int spill_stride,
                                         uint32_t &n_nodes, uint32_t &n_prims) {
        const float kMiss = __builtin_huge_valf();
        // compare-exchange by selects (the branchy form compiles to
        // exec-masked moves, two more operations each)
        auto cswap = [](float &ta, int &ca, float &tb, int &cb) {
            const bool sw = tb < ta;
            const float t0 = sw ? tb : ta, t1 = sw ? ta : tb;
            const int k0 = sw ? cb : ca, k1 = sw ? ca : cb;
            ta = t0;
            tb = t1;
            ca = k0;
            cb = k1;
        };
        if (node >= 0) {
            const Bvh4QNode *nd = ts.nodes + node;
            const float4 org = nd->org;
            const uint4 qx = nd->qx, qy = nd->qy, qz = nd->qz;
            int4 ch = nd->ch;
            if (STATS) n_nodes++;
            const float sx = r.idx * org.w, sy = r.idy * org.w, sz = r.idz * org.w;
            const float fx = __builtin_fmaf(org.x, r.idx, -r.oidx), fy = __builtin_fmaf(org.y, r.idy, -r.oidy),
                        fz = __builtin_fmaf(org.z, r.idz, -r.oidz);
            const AxisPlanes px = axis_planes(qx, r.idx < 0.f), py = axis_planes(qy, r.idy < 0.f),
                             pz = axis_planes(qz, r.idz < 0.f);
            float t0 = qbox_entry<0>(sx, fx, sy, fy, sz, fz, px, py, pz, tmin, tmax);
            float t1 = qbox_entry<1>(sx, fx, sy, fy, sz, fz, px, py, pz, tmin, tmax);
            float t2 = qbox_entry<2>(sx, fx, sy, fy, sz, fz, px, py, pz, tmin, tmax);
            float t3 = qbox_entry<3>(sx, fx, sy, fy, sz, fz, px, py, pz, tmin, tmax);
            int c0 = ch.x, c1 = ch.y, c2 = ch.z, c3 = ch.w;
            if (c2 == kNoChild) t2 = kMiss;  // empty slots carry inverted boxes, which
            if (c3 == kNoChild) t3 = kMiss;  // a slab test alone does not reject
            auto push = [&](int v) {
                if (sp < kStackLds) stack[sp * stride] = v;
                else spill[(sp - kStackLds) * spill_stride] = v;
                sp++;
            };
            if (ANY) {
                // any hit: children in slot order, no sorting (an occluder
                // ends the ray wherever it lies; sorting the children near
                // to far measured slower)
                // the volumetric occlusion test pushes without the per-push
                // spill test while its LDS stack has room (faster there; the
                // shadow-ray kernel measured slower with it)
                if (OPAQUE && sp <= kStackLds - 3) {
                    int k = sp;
                    if (t3 != kMiss) stack[(k++) * stride] = c3;
                    if (t2 != kMiss) stack[(k++) * stride] = c2;
                    if (t1 != kMiss) stack[(k++) * stride] = c1;
                    sp = k;
                } else {
                    if (t3 != kMiss) push(c3);
                    if (t2 != kMiss) push(c2);
                    if (t1 != kMiss) push(c1);
                }
                if (t0 != kMiss) {
                    node = c0;
                    return false;
                }
            } else {
                // closest hit: five comparators order the four children,
                // nearest taken next, the others pushed far to near
                cswap(t0, c0, t1, c1);
                cswap(t2, c2, t3, c3);
                cswap(t0, c0, t2, c2);
                cswap(t1, c1, t3, c3);
                cswap(t1, c1, t2, c2);
                if (t0 != kMiss) {
                    if (sp <= kStackLds - 3) {
                        // room for all three in LDS: no spill test per push
                        int k = sp;
                        if (t3 != kMiss) stack[(k++) * stride] = c3;
                        if (t2 != kMiss) stack[(k++) * stride] = c2;
                        if (t1 != kMiss) stack[(k++) * stride] = c1;
                        sp = k;
                    } else {
                        if (t3 != kMiss) push(c3);
                        if (t2 != kMiss) push(c2);
                        if (t1 != kMiss) push(c1);
                    }
                    node = c0;
                    return false;
                }
            }
        } else {
            const int code = ~node;
            const int first = code >> 3, cnt = (code & 7) + 1;
            for (int k = 0; k < cnt; k++) {
                const int slot = first + k;
                float4 p0 = ts.prims[3 * slot], p1 = ts.prims[3 * slot + 1];
                float t, u = 0, v = 0;
                bool ok;
                if (STATS) n_prims++;
                if (!SPH || __float_as_int(p1.w) == kPrimTri) {
                    float4 p2 = ts.prims[3 * slot + 2];
                    ok = tri_test(r, p0, p1, p2, tmin, tmax, t, u, v);
                } else {
                    double td;
                    ok = sphere_roots(r.ox, r.oy, r.oz, dx, dy, dz, ts.spheres[__float_as_int(p1.x)], tmin, tmax, td);
                    t = float(td);
                }
                if (OPAQUE && ok && ts.skip[slot]) {
                    ok = false;
                    passed = true;
                }
                if (ok) {
                    found = true;
                    hit.t = t;
                    hit.u = u;
                    hit.v = v;
                    hit.slot = slot;
                    if (ANY) return true;
                    tmax = t;
                }
            }
        }
        if (sp == 0) return true;
        sp--;
        node = sp < kStackLds ? stack[sp * stride] : spill[(sp - kStackLds) * spill_stride];
        return false;
    }

};

// Closest hit on a triangle: the FP32 distance loses relative accuracy when
// the hit is close to the origin next to a large triangle (T cancels);
// recompute it once in FP64 from the triangle's plane.  Sphere hits keep
// their distance (sphere_hit_uv() re-solves them).
template <bool SPH>
__device__ __forceinline__ float refine_t(const float4 *__restrict__ prims, int slot, float t, float ox, float oy,
                                          float oz, float dx, float dy, float dz) {
    float4 p0 = prims[3 * slot], p1 = prims[3 * slot + 1];
    if (!SPH || __float_as_int(p1.w) == kPrimTri) {
        float4 p2 = prims[3 * slot + 2];
        double e1x = double(p1.x) - p0.x, e1y = double(p1.y) - p0.y, e1z = double(p1.z) - p0.z;
        double e2x = double(p2.x) - p0.x, e2y = double(p2.y) - p0.y, e2z = double(p2.z) - p0.z;
        double nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
        double den = nx * dx + ny * dy + nz * dz;
        if (den != 0.0) {
            double num = nx * (double(p0.x) - ox) + ny * (double(p0.y) - oy) + nz * (double(p0.z) - oz);
            t = float(num / den);
        }
    }
    return t;
}

// Trace one ray to the end.  Returns true on a hit.
template <bool ANY, bool STATS, bool SPH, bool OPAQUE = false>
__device__ __forceinline__ bool trace_ray(const TraceScene &ts, float ox, float oy, float oz, float dx, float dy,
                                          float dz, float tmin, float tmax, int *stack, int stride, int *spill,
                                          int spill_stride, TraceHit &hit, uint32_t &n_nodes, uint32_t &n_prims) {
    Tracer<ANY, STATS, SPH, OPAQUE> tr;
    tr.start(ox, oy, oz, dx, dy, dz, tmin, tmax);
    while (!tr.step(ts, stack, stride, spill, spill_stride, n_nodes, n_prims)) {
    }
    hit = tr.hit;
    if (!ANY && tr.found) hit.t = refine_t<SPH>(ts.prims, hit.slot, hit.t, ox, oy, oz, dx, dy, dz);
    return tr.found;
}

// For a sphere hit: the FP64 distance and the spherical coordinates used as
// u, v (src/shapes/sphere.inl:85-104).
__device__ __forceinline__ void sphere_hit_uv(double4 c, double ox, double oy, double oz, double dx, double dy,
                                           double dz, double tmin, float t_hit, double &t, double &u, double &v) {
    t = double(t_hit);
    sphere_roots(ox, oy, oz, dx, dy, dz, c, tmin, double(t_hit) * (1 + 1e-6) + 1e-30, t);
    double px = ox + t * dx - c.x, py = oy + t * dy - c.y, pz = oz + t * dz - c.z;
    double el = acos(clampd(py / c.w, -1.0, 1.0));
    double az = atan2(pz / c.w, px / c.w);
    u = az * kInvTwoPi;
    v = el * kInvPi;
}

}  // namespace ljx
