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

// Lanes with key >= 0 append to the queue counters[key]: one atomic per
// distinct key, all issued by one instruction (the lowest lane of each key
// adds its key's count).  Returns the lane's index in its queue, or -1.
__device__ __forceinline__ int wave_append_keyed(int key, int *counters) {
    uint64_t pending = __ballot(key >= 0);
    uint64_t mine = 0;
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int k = __builtin_amdgcn_readlane(key, leader);
        const uint64_t m = __ballot(key == k);
        if (key == k) mine = m;
        pending &= ~m;
    }
    const int lane = __lane_id();
    const int first = mine ? __ffsll((unsigned long long)mine) - 1 : lane;
    int base = 0;
    if (key >= 0 && lane == first) base = atomicAdd(&counters[key], __popcll(mine));
    base = __shfl(base, first);
    const int below = __builtin_amdgcn_mbcnt_hi(uint32_t(mine >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mine), 0));
    return key >= 0 ? base + below : -1;
}

__device__ __forceinline__ D3 f3d(float4 v) { return D3{double(v.x), double(v.y), double(v.z)}; }

// ------------------------------------------------------------- path vertex
struct Vertex {
    D3 pos;
    ShadeGeom g;
    double tu, tv;   // texture coordinates (src/intersection.h:22)
    double inv_uv;   // |dp/duv|: the footprint's divisor (intersection.cpp:56-57)
    int shape_id;
    int material_id;
    int area_light_id;
    int interior_medium_id;
    int exterior_medium_id;
};

// Rebuilds lajolla's PathVertex for a hit (src/intersection.cpp:35-62 with
// compute_shading_info, src/shapes/triangle_mesh.inl:77-169 and
// src/shapes/sphere.inl:243-268), with the texture coordinates and the
// size of dp/duv that the footprint of a camera ray divides by.
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
        vx.tu = (1 - s0 - s1) * u0x + s0 * u1x + s1 * u2x;
        vx.tv = (1 - s0 - s1) * u0y + s0 * u1y + s1 * u2y;
        double dsx = u2x - u0x, dsy = u2y - u0y;  // duv/ds
        double dtx = u2x - u1x, dty = u2y - u1y;  // duv/dt
        double det = dsx * dty - dtx * dsy;
        D3 dpdu;
        if (fabs(det) > double(1e-8f)) {
            double dsdu = dty / det, dtdu = -dsy / det;
            double dsdv = dtx / det, dtdv = -dsx / det;  // the reference's signs (triangle_mesh.inl:117-120)
            dpdu = (c - a) * dsdu + (c - b) * dtdu;
            D3 dpdv = (c - a) * dsdv + (c - b) * dtdv;
            vx.inv_uv = fmax(length(dpdu), length(dpdv));
        } else {
            dpdu = frame_from_n(ng).x;
            vx.inv_uv = 1;  // coordinate_system: two unit vectors
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
        // the FP64 root the traversal found, searched for just below its FP32 distance: from 0, a ray
        // leaving the sphere's own surface (a refraction into it) would find its near root, ~0, instead
        sphere_hit_uv(make_double4(c.x, c.y, c.z, r), org.x, org.y, org.z, dir.x, dir.y, dir.z,
                      double(t) * (1 - 1e-6), t, td, su, sv);
        // the reference passes the FP32 distance and u, v through Embree's hit record
        su = double(float(su));
        sv = double(float(sv));
        vx.pos = org + dir * double(float(td));
        ng = normalize(vx.pos - c);
        D3 dpdu{-r * sin(su) * sin(sv), r * cos(su) * sin(sv), 0.0};
        D3 dpdv{r * cos(su) * cos(sv), r * sin(su) * cos(sv), -r * sin(sv)};
        vx.tu = su;
        vx.tv = sv;
        vx.inv_uv = (length(dpdu) + length(dpdv)) / 2;
        D3 tg = normalize(dpdu - ng * dot(ng, dpdu));
        sf = Frame{tg, normalize(cross(ng, tg)), ng};
    }
    if (dot(ng, sf.n) < 0) ng = -ng;
    vx.g.ng = ng;
    vx.g.sf = sf;
    return vx;
}

// The material with its texture fields evaluated at (u, v) with the given
// footprint, as each BSDF's eval(texture, vertex.uv, vertex.uv_screen_size)
// calls do (e.g. src/materials/lambertian.inl:11-13).
__device__ __forceinline__ void resolve_textures(const DeviceScene &S, MatRec &m, double u, double v, double fp) {
    auto spec = [&](int k, D3 &dst) {
        if (m.tex[k]) dst = tex_eval(S.tex_data, S.textures[m.tex[k] - 1], u, v, fp);
    };
    auto real = [&](int k, double &dst) {
        if (m.tex[k]) dst = tex_eval1(S.tex_data, S.textures[m.tex[k] - 1], u, v, fp);
    };
    spec(0, m.a);
    spec(1, m.b);
    real(2, m.roughness);
    real(3, m.specular_transmission);
    real(4, m.metallic);
    real(5, m.subsurface);
    real(6, m.specular);
    real(7, m.specular_tint);
    real(8, m.anisotropic);
    real(9, m.sheen);
    real(10, m.sheen_tint);
    real(11, m.clearcoat);
    real(12, m.clearcoat_gloss);
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
// TableDist2D sampling and pdf (src/table_dist.cpp:97-151) over the envmap's texels.
__device__ __forceinline__ void sample_env_uv(const LightRec &l, double rx, double ry, double &u, double &v) {
    const int w = l.env_w, h = l.env_h;
    const int yo = sample_cdf(l.cdf_marginals, h, ry);
    double dy = ry - l.cdf_marginals[yo];
    if (l.cdf_marginals[yo + 1] - l.cdf_marginals[yo] > 0) dy /= (l.cdf_marginals[yo + 1] - l.cdf_marginals[yo]);
    const double *c = l.cdf_rows + size_t(yo) * (w + 1);
    const int xo = sample_cdf(c, w, rx);
    double dx = rx - c[xo];
    if (c[xo + 1] - c[xo] > 0) dx /= (c[xo + 1] - c[xo]);
    u = (xo + dx) / w;
    v = (yo + dy) / h;
}
__device__ __forceinline__ double pdf_env_uv(const LightRec &l, double u, double v) {
    const int w = l.env_w, h = l.env_h;
    const int x = int(clampd(u * w, 0, w - 1)), y = int(clampd(v * h, 0, h - 1));
    return l.pdf_marginals[y] * l.pdf_rows[size_t(y) * w + x] * w * h;
}
// lat-long coordinates of a local direction, y up (src/lights/envmap.inl:25-31)
__device__ __forceinline__ void env_uv(D3 ld, double &u, double &v) {
    u = atan2(ld.x, -ld.z) * (1 / (2 * kPi));
    v = acos(clampd(ld.y, -1, 1)) * (1 / kPi);
    if (u < 0) u += 1;
}
// src/lights/envmap.inl:7-42
__device__ __forceinline__ LightPoint sample_env(const LightRec &l, double u0, double u1) {
    double eu, ev;
    sample_env_uv(l, u0, u1, eu, ev);
    double az = eu * (2 * kPi), el = ev * kPi;
    double sa, ca, se, ce;
    sincos(az, &sa, &ca);
    sincos(el, &se, &ce);
    D3 ld{sa * se, ce, -ca * se};
    D3 wd = xform3(l.to_world, ld);
    return {{0, 0, 0}, -wd};
}
__device__ __forceinline__ double pdf_env(const LightRec &l, D3 world_dir) {
    D3 ld = xform3(l.to_local, world_dir);
    double u, v;
    env_uv(ld, u, v);
    double ce = ld.y;
    double se = sqrt(clampd(1 - ce * ce, 0, 1));
    if (se <= 0) return 0;
    return pdf_env_uv(l, u, v) / (2 * kPi * kPi * se);
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
// Level-0 bilinear radiance of the envmap at lat-long (u, v) (src/mipmap.h:50-72; its
// texels are the envmap's FP32 values).
__device__ __forceinline__ D3 env_texel_bilinear(const LightRec &l, double u, double v) {
    const int W = l.env_w, H = l.env_h;
    u = u * W - 0.5;
    v = v * H - 0.5;
    const int ufi = imod(int(u), W), vfi = imod(int(v), H);
    const int uci = imod(ufi + 1, W), vci = imod(vfi + 1, H);
    const double uo = u - ufi, vo = v - vfi;
    auto tx = [&](int x, int y) {
        const float *p = l.env_rgb + 3 * (size_t(y) * W + x);
        return D3{double(p[0]), double(p[1]), double(p[2])};
    };
    return tx(ufi, vfi) * (1 - uo) * (1 - vo) + tx(ufi, vci) * (1 - uo) * vo + tx(uci, vfi) * uo * (1 - vo) +
           tx(uci, vci) * uo * vo;
}
// Next-event estimation towards the envmap (src/lights/envmap.inl:7-72): the
// TableDist2D sample's lat-long coordinates give the direction, its pdf and
// the radiance without mapping the direction back to (u, v).
__device__ __forceinline__ void env_nee(const LightRec &l, double u0, double u1, D3 &dir, double &pdf, D3 &L) {
    double eu, ev;
    sample_env_uv(l, u0, u1, eu, ev);
    double sa, ca, se, ce;
    sincos(eu * (2 * kPi), &sa, &ca);
    sincos(ev * kPi, &se, &ce);
    dir = xform3(l.to_world, D3{sa * se, ce, -ca * se});
    pdf = se > 0 ? pdf_env_uv(l, eu, ev) / (2 * kPi * kPi * se) : 0.0;
    L = env_texel_bilinear(l, eu, ev) * l.scale;
}

// emission (src/lights/diffuse_area_light.inl:15-20; src/lights/envmap.inl:44-72).
// The envmap's footprint, min(|du/dw|, dv/dw), is never positive (dv/dw < 0),
// so the reference reads the image bilinearly on level 0; only along the exact
// poles, where that footprint is NaN, would it read the last level instead.
__device__ __forceinline__ D3 light_emission(const DeviceScene &S, const LightRec &l, D3 view_dir,
                                            const LightPoint &lp) {
    if (l.type == 1) {
        D3 w = xform3(l.to_local, -view_dir);
        double u, v;
        env_uv(w, u, v);
        return env_texel_bilinear(l, u, v) * l.scale;
    }
    if (dot(lp.n, view_dir) <= 0) return {0, 0, 0};
    return l.intensity;
}

}  // namespace ljx
