// Hit reconstruction, light sampling and wave-level queue helpers shared by
// the wavefront kernels.
#pragma once

#include "ljx_device.h"
#include "ljx_trace.h"

namespace ljx {

// ------------------------------------------------------------ wave helpers
// Appends the lanes with pred set to a queue guarded by *counter: one atomic
// per wave (ballot + popcount), lane offsets from mbcnt.  Every lane of the
// wave must call it.  Returns the lane's queue index, or -1.
__device__ __forceinline__ int wave_append(bool pred, int *counter) {
    uint64_t mask = __ballot(pred);
    if (mask == 0) return -1;
    int lane = __lane_id();
    int leader = __ffsll((unsigned long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader);
    int below = __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0));
    return pred ? base + below : -1;
}

__device__ __forceinline__ D3 f3d(float4 v) { return D3{double(v.x), double(v.y), double(v.z)}; }

// ------------------------------------------------------------- path vertex
struct Vertex {
    D3 pos;
    ShadeGeom g;
    int shape_id;
    int material_id;
    int area_light_id;
    int interior_medium_id;
    int exterior_medium_id;
};

// Rebuilds lajolla's PathVertex for a hit (src/intersection.cpp:35-62 with
// compute_shading_info, src/shapes/triangle_mesh.inl:77-169 and
// src/shapes/sphere.inl:243-268).  Ray differentials are not carried: every
// texture here is constant, so they cannot change a value.
template <bool SPH>
__device__ __forceinline__ Vertex make_vertex(const DeviceScene &S, int slot, float t, float u, float v, D3 org,
                                              D3 dir) {
    Vertex vx;
    float4 p0 = S.prims[3 * slot], p1 = S.prims[3 * slot + 1];
    int gid = __float_as_int(p0.w);
    int sid = S.prim_shape[gid];
    const ShapeRec &sh = S.shapes[sid];
    vx.shape_id = sid;
    vx.material_id = sh.material_id;
    vx.area_light_id = sh.area_light_id;
    vx.interior_medium_id = sh.interior_medium_id;
    vx.exterior_medium_id = sh.exterior_medium_id;
    vx.pos = org + dir * double(t);
    D3 ng;
    Frame sf;
    if (!SPH || __float_as_int(p1.w) == kPrimTri) {
        float4 p2 = S.prims[3 * slot + 2];
        D3 a = f3d(p0), b = f3d(p1), c = f3d(p2);
        ng = normalize(cross(b - a, c - a));
        double s0 = double(u), s1 = double(v);
        double u0x = 0, u0y = 0, u1x = 1, u1y = 0, u2x = 1, u2y = 1;
        int i0 = 0, i1 = 0, i2 = 0;
        if (sh.has_normals || sh.has_uvs) {
            i0 = S.tri_vidx[3 * gid];
            i1 = S.tri_vidx[3 * gid + 1];
            i2 = S.tri_vidx[3 * gid + 2];
        }
        if (sh.has_uvs) {
            u0x = S.vuv[2 * i0];
            u0y = S.vuv[2 * i0 + 1];
            u1x = S.vuv[2 * i1];
            u1y = S.vuv[2 * i1 + 1];
            u2x = S.vuv[2 * i2];
            u2y = S.vuv[2 * i2 + 1];
        }
        double dsx = u2x - u0x, dsy = u2y - u0y;  // duv/ds
        double dtx = u2x - u1x, dty = u2y - u1y;  // duv/dt
        double det = dsx * dty - dtx * dsy;
        D3 dpdu;
        if (fabs(det) > double(1e-8f)) {
            double dsdu = dty / det, dtdu = -dsy / det;
            dpdu = (c - a) * dsdu + (c - b) * dtdu;
        } else {
            dpdu = frame_from_n(ng).x;
        }
        D3 n = ng;
        if (sh.has_normals) {
            D3 n0{S.vnrm[3 * i0], S.vnrm[3 * i0 + 1], S.vnrm[3 * i0 + 2]};
            D3 n1{S.vnrm[3 * i1], S.vnrm[3 * i1 + 1], S.vnrm[3 * i1 + 2]};
            D3 n2{S.vnrm[3 * i2], S.vnrm[3 * i2 + 1], S.vnrm[3 * i2 + 2]};
            n = normalize((1 - s0 - s1) * n0 + s0 * n1 + s1 * n2);
        }
        D3 tg = normalize(dpdu - n * dot(n, dpdu));
        sf = Frame{tg, normalize(cross(n, tg)), n};
    } else {
        D3 c{sh.center[0], sh.center[1], sh.center[2]};
        double r = sh.radius;
        double td, su, sv;
        sphere_hit_uv(make_double4(c.x, c.y, c.z, r), org.x, org.y, org.z, dir.x, dir.y, dir.z, 0.0, t, td, su, sv);
        // the reference passes the FP32 distance and u, v through Embree's hit record
        su = double(float(su));
        sv = double(float(sv));
        vx.pos = org + dir * double(float(td));
        ng = normalize(vx.pos - c);
        D3 dpdu{-r * sin(su) * sin(sv), r * cos(su) * sin(sv), 0.0};
        D3 tg = normalize(dpdu - ng * dot(ng, dpdu));
        sf = Frame{tg, normalize(cross(ng, tg)), ng};
    }
    if (dot(ng, sf.n) < 0) ng = -ng;
    vx.g.ng = ng;
    vx.g.sf = sf;
    return vx;
}

// Mean curvature of a hit (compute_shading_info, src/shapes/triangle_mesh.inl:
// 133-160 and src/shapes/sphere.inl:264): 1/r on spheres, 0 on meshes without
// vertex normals, otherwise (dn/du . tangent + dn/dv . bitangent) / 2 with
// the parameter derivatives written as the reference writes them.
__device__ inline double vertex_mean_curvature(const DeviceScene &S, int slot, double s0, double s1,
                                               const Vertex &vx) {
    const ShapeRec &sh = S.shapes[vx.shape_id];
    if (sh.type == 0) return 1.0 / sh.radius;
    if (!sh.has_normals) return 0.0;
    int gid = __float_as_int(S.prims[3 * slot].w);
    int i0 = S.tri_vidx[3 * gid], i1 = S.tri_vidx[3 * gid + 1], i2 = S.tri_vidx[3 * gid + 2];
    double u0x = 0, u0y = 0, u1x = 1, u1y = 0, u2x = 1, u2y = 1;
    if (sh.has_uvs) {
        u0x = S.vuv[2 * i0], u0y = S.vuv[2 * i0 + 1];
        u1x = S.vuv[2 * i1], u1y = S.vuv[2 * i1 + 1];
        u2x = S.vuv[2 * i2], u2y = S.vuv[2 * i2 + 1];
    }
    double dsx = u2x - u0x, dsy = u2y - u0y, dtx = u2x - u1x, dty = u2y - u1y;
    double det = dsx * dty - dtx * dsy;
    double dsdu = dty / det, dtdu = -dsy / det, dsdv = dtx / det, dtdv = -dsx / det;
    D3 n0{S.vnrm[3 * i0], S.vnrm[3 * i0 + 1], S.vnrm[3 * i0 + 2]};
    D3 n1{S.vnrm[3 * i1], S.vnrm[3 * i1 + 1], S.vnrm[3 * i1 + 2]};
    D3 n2{S.vnrm[3 * i2], S.vnrm[3 * i2 + 1], S.vnrm[3 * i2 + 2]};
    D3 dnds = n2 - n0, dndt = n2 - n1;
    D3 dndu = dnds * dsdu + dndt * dtdu, dndv = dnds * dsdv + dndt * dtdv;
    return (dot(dndu, vx.g.sf.x) + dot(dndv, vx.g.sf.y)) / 2;
}

// ------------------------------------------------------------------ lights
struct LightPoint {
    D3 pos, n;
};

// sample_point_on_shape for triangle meshes and spheres
// (src/shapes/triangle_mesh.inl:24-58, src/shapes/sphere.inl:164-212).
__device__ __forceinline__ LightPoint sample_shape(const DeviceScene &S, const ShapeRec &sh, D3 ref, double u0,
                                                  double u1, double w) {
    if (sh.type == 0) {
        D3 c{sh.center[0], sh.center[1], sh.center[2]};
        double r = sh.radius;
        double s, co;
        if (dist2(ref, c) < r * r) {
            double z = 1 - 2 * u0;
            double r_ = sqrt(fmax(0.0, 1 - z * z));
            sincos(2 * kPi * u1, &s, &co);
            D3 off{r_ * co, r_ * s, z};
            return {c + r * off, off};
        }
        D3 dc_dir = normalize(c - ref);
        Frame f = frame_from_n(dc_dir);
        double smax2 = r * r / dist2(ref, c);
        double cmax = sqrt(fmax(0.0, 1 - smax2));
        double ce = (1 - u0) + u0 * cmax;
        double se = sqrt(fmax(0.0, 1 - ce * ce));
        double dc = sqrt(dist2(ref, c));
        double ds = dc * ce - sqrt(fmax(0.0, r * r - dc * dc * se * se));
        double ca = (dc * dc + r * r - ds * ds) / (2 * dc * r);
        double sa = sqrt(fmax(0.0, 1 - ca * ca));
        sincos(u1 * 2 * kPi, &s, &co);
        D3 n = -to_world(f, D3{sa * co, sa * s, ca});
        return {r * n + c, n};
    }
    int tri = sample_cdf(S.tri_cdf + sh.cdf_offset, sh.num_tris, w);
    int gid = sh.tri_offset + tri;
    int i0 = S.tri_vidx[3 * gid], i1 = S.tri_vidx[3 * gid + 1], i2 = S.tri_vidx[3 * gid + 2];
    D3 v0{S.vpos[3 * i0], S.vpos[3 * i0 + 1], S.vpos[3 * i0 + 2]};
    D3 v1{S.vpos[3 * i1], S.vpos[3 * i1 + 1], S.vpos[3 * i1 + 2]};
    D3 v2{S.vpos[3 * i2], S.vpos[3 * i2 + 1], S.vpos[3 * i2 + 2]};
    D3 e1 = v1 - v0, e2 = v2 - v0;
    double a = sqrt(clampd(u0, 0, 1));
    double b1 = 1 - a, b2 = a * u1;
    D3 ng = normalize(cross(e1, e2));
    if (sh.has_normals) {
        D3 n0{S.vnrm[3 * i0], S.vnrm[3 * i0 + 1], S.vnrm[3 * i0 + 2]};
        D3 n1{S.vnrm[3 * i1], S.vnrm[3 * i1 + 1], S.vnrm[3 * i1 + 2]};
        D3 n2{S.vnrm[3 * i2], S.vnrm[3 * i2 + 1], S.vnrm[3 * i2 + 2]};
        D3 sn = normalize((1 - b1 - b2) * n0 + b1 * n1 + b2 * n2);
        if (dot(ng, sn) < 0) ng = -ng;
    }
    return {v0 + e1 * b1 + e2 * b2, ng};
}
// pdf_point_on_shape (src/shapes/triangle_mesh.inl:54-58, sphere.inl:218-238)
__device__ __forceinline__ double pdf_shape(const ShapeRec &sh, const LightPoint &lp, D3 ref) {
    if (sh.type == 0) {
        D3 c{sh.center[0], sh.center[1], sh.center[2]};
        double r = sh.radius;
        if (dist2(ref, c) < r * r) return 1 / (4 * kPi * r * r);
        double smax2 = r * r / dist2(ref, c);
        double cmax = sqrt(fmax(0.0, 1 - smax2));
        double pdf_sa = 1 / (2 * kPi * (1 - cmax));
        D3 dd = normalize(lp.pos - ref);
        return pdf_sa * fabs(dot(lp.n, dd)) / dist2(ref, lp.pos);
    }
    return 1 / sh.total_area;
}
// Envmap on a 1x1 image (src/lights/envmap.inl:11-42): (u, v) uniform,
// direction from lat-long with y up, pdf = 1 / (2 pi^2 sin(elevation)).
__device__ __forceinline__ LightPoint sample_env(const LightRec &l, double u0, double u1) {
    double az = u0 * (2 * kPi), el = u1 * kPi;
    double sa, ca, se, ce;
    sincos(az, &sa, &ca);
    sincos(el, &se, &ce);
    D3 ld{sa * se, ce, -ca * se};
    D3 wd = xform3(l.to_world, ld);
    return {{0, 0, 0}, -wd};
}
__device__ __forceinline__ double pdf_env(const LightRec &l, D3 world_dir) {
    D3 ld = xform3(l.to_local, world_dir);
    double ce = ld.y;
    double se = sqrt(clampd(1 - ce * ce, 0, 1));
    if (se <= 0) return 0;
    return 1.0 / (2 * kPi * kPi * se);
}
__device__ __forceinline__ LightPoint sample_light_point(const DeviceScene &S, const LightRec &l, D3 ref, double u0,
                                                        double u1, double w) {
    if (l.type == 1) return sample_env(l, u0, u1);
    return sample_shape(S, S.shapes[l.shape_id], ref, u0, u1, w);
}
__device__ __forceinline__ double pdf_light_point(const DeviceScene &S, const LightRec &l, const LightPoint &lp,
                                                 D3 ref) {
    if (l.type == 1) return pdf_env(l, -lp.n);
    return pdf_shape(S.shapes[l.shape_id], lp, ref);
}
// emission (src/lights/diffuse_area_light.inl:15-20; envmap value * scale)
__device__ __forceinline__ D3 light_emission(const LightRec &l, D3 view_dir, const LightPoint &lp) {
    if (l.type == 1) return l.env_value * l.scale;
    if (dot(lp.n, view_dir) <= 0) return {0, 0, 0};
    return l.intensity;
}

}  // namespace ljx
