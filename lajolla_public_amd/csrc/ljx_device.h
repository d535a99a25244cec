// Device-side shading substrate of the MI355X path-tracing core.
//
// Shading runs in FP64 like lajolla's Real (src/lajolla.h:23): CDNA4 issues
// FP64 vector FMAs at half the FP32 rate, and the BSDFs have ill-conditioned
// terms (GTR2 at alpha^2 = 1e-8, Fresnel near total internal reflection)
// that FP32 cannot evaluate to 1e-5.  Traversal stays FP32 (ljx_trace.h),
// as Embree's does.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define LJX_HD __host__ __device__ __forceinline__

namespace ljx {

constexpr double kPi = 3.14159265358979323846;
constexpr double kInvPi = 1.0 / kPi;
constexpr double kTwoPi = 2.0 * kPi;
constexpr double kInvTwoPi = 1.0 / kTwoPi;
constexpr double kInvFourPi = 1.0 / (4.0 * kPi);

// ------------------------------------------------------------------ vectors
struct D3 {
    double x, y, z;
};
LJX_HD D3 d3(double x, double y, double z) { return D3{x, y, z}; }
LJX_HD D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
LJX_HD D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
LJX_HD D3 operator-(D3 a) { return {-a.x, -a.y, -a.z}; }
LJX_HD D3 operator*(D3 a, D3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
LJX_HD D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
LJX_HD D3 operator*(double s, D3 a) { return {a.x * s, a.y * s, a.z * s}; }
LJX_HD D3 operator/(D3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
LJX_HD D3 operator/(D3 a, D3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
LJX_HD D3 operator+(D3 a, double s) { return {a.x + s, a.y + s, a.z + s}; }
LJX_HD D3 operator+(double s, D3 a) { return {a.x + s, a.y + s, a.z + s}; }
LJX_HD D3 operator-(double s, D3 a) { return {s - a.x, s - a.y, s - a.z}; }
LJX_HD double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LJX_HD D3 cross(D3 a, D3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
LJX_HD double length(D3 a) { return sqrt(dot(a, a)); }
LJX_HD double dist2(D3 a, D3 b) { return dot(a - b, a - b); }
// Zero-length vectors normalise to zero (src/vector.h:249-257).
LJX_HD D3 normalize(D3 a) {
    double l = length(a);
    return l <= 0 ? D3{0, 0, 0} : a / l;
}
LJX_HD double maxc(D3 a) { return fmax(fmax(a.x, a.y), a.z); }
LJX_HD double avgc(D3 a) { return (a.x + a.y + a.z) / 3; }
LJX_HD double comp(D3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
LJX_HD D3 vexp(D3 a) { return {exp(a.x), exp(a.y), exp(a.z)}; }
LJX_HD D3 vsqrt0(D3 a) { return {sqrt(fmax(a.x, 0.0)), sqrt(fmax(a.y, 0.0)), sqrt(fmax(a.z, 0.0))}; }
LJX_HD double luminance(D3 s) { return s.x * 0.212671 + s.y * 0.715160 + s.z * 0.072169; }
LJX_HD double clampd(double v, double lo, double hi) { return fmin(fmax(v, lo), hi); }
LJX_HD double pow5(double x) {
    double x2 = x * x;
    return x2 * x2 * x;
}

// ------------------------------------------------------------------- frames
struct Frame {
    D3 x, y, n;
};
// Frisvad basis with the n.z < -1 + 1e-6 special case (src/frame.h:11-21).
LJX_HD Frame frame_from_n(D3 n) {
    Frame f;
    f.n = n;
    if (n.z < -1.0 + 1e-6) {
        f.x = {0, -1, 0};
        f.y = {-1, 0, 0};
    } else {
        double a = 1 / (1 + n.z);
        double b = -n.x * n.y * a;
        f.x = {1 - n.x * n.x * a, b, -n.x};
        f.y = {b, 1 - n.y * n.y * a, -n.y};
    }
    return f;
}
LJX_HD Frame flip(const Frame &f) { return Frame{-f.x, -f.y, -f.n}; }
LJX_HD D3 to_local(const Frame &f, D3 v) { return {dot(v, f.x), dot(v, f.y), dot(v, f.n)}; }
LJX_HD D3 to_world(const Frame &f, D3 v) { return f.x * v.x + f.y * v.y + f.n * v.z; }

// ---------------------------------------------------------------------- PCG
// XSH-RR with lajolla's constants (src/pcg.h:16-68).  The increment is
// derived from the stream id; only the 64-bit state travels with a path.
LJX_HD uint64_t pcg_inc(uint64_t stream) { return (stream << 1u) | 1u; }
LJX_HD uint32_t pcg_next(uint64_t &state, uint64_t inc) {
    uint64_t old = state;
    state = old * 6364136223846793005ULL + (inc | 1);
    uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = uint32_t(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
}
// The PCG stream of a camera path from its id (sample * W * H + pixel): a
// bijection on 63-bit values, two multiply-xorshift rounds with the SplitMix64
// finaliser's constants (Steele, Lea and Flood 2014), masked to 63 bits so
// that distinct paths keep distinct increments.  Consecutive ids used as
// increments give correlated PCG streams: neighbouring paths' draws are
// related, which leaves an image's mean alone but not its error.
LJX_HD uint64_t path_stream(uint64_t id) {
    const uint64_t m = (uint64_t(1) << 63) - 1;
    uint64_t z = id & m;
    z = ((z ^ (z >> 31)) * 0xbf58476d1ce4e5b9ULL) & m;
    z = ((z ^ (z >> 27)) * 0x94d049bb133111ebULL) & m;
    return z ^ (z >> 31);
}
LJX_HD uint64_t pcg_init_state(uint64_t stream, uint64_t seed) {
    uint64_t inc = pcg_inc(stream);
    uint64_t st = 0;
    pcg_next(st, inc);
    st += seed;
    pcg_next(st, inc);
    return st;
}
LJX_HD double pcg_double(uint64_t &state, uint64_t inc) {
    uint64_t bits = (uint64_t(pcg_next(state, inc)) << 20) | 0x3ff0000000000000ULL;
    return __builtin_bit_cast(double, bits) - 1.0;
}
struct Rng {
    uint64_t state, inc;
    LJX_HD double next() { return pcg_double(state, inc); }
    LJX_HD uint32_t next_u32() { return pcg_next(state, inc); }
};

// -------------------------------------------------------------- device scene
// Material record: constant-texture parameters in FP64 (src/material.h:12-98).
struct MatRec {
    int type;
    int textured;  // some field reads a texture
    D3 a, b;
    double roughness, eta;
    double specular_transmission, metallic, subsurface, specular, specular_tint, anisotropic, sheen,
        sheen_tint, clearcoat, clearcoat_gloss;
    int tex[13];  // per field (ljx_material_desc order): 0 the constant, k > 0 texture k - 1
};
// ImageTexture / CheckerboardTexture (src/texture.h:82-97) with the image's
// mipmap (src/mipmap.h:7-48): level l is w[l] x h[l] pixels of channels
// doubles at tex_data + off[l], row-major from the top row.
constexpr int kMaxMipLevels = 8;
struct TexRec {
    int type;  // 1 image, 2 checkerboard
    int channels;
    int levels;
    int pad;
    int w[kMaxMipLevels], h[kMaxMipLevels];
    long long off[kMaxMipLevels];
    D3 color0, color1;
    double uscale, vscale, uoffset, voffset;
};
// Shape record (src/shape.h:20-51).
struct ShapeRec {
    int type;  // 0 sphere, 1 triangle mesh
    int material_id;
    int area_light_id;
    int interior_medium_id;
    int exterior_medium_id;
    int has_normals;
    int has_uvs;
    int tri_offset;   // first global triangle id
    int num_tris;
    int cdf_offset;   // triangle-area CDF (num_tris + 1 doubles), -1 if not a light
    double center[3];
    double radius;
    double total_area;
};
// Light record (src/light.h:12-30).  An envmap reads texture env_tex and
// samples directions from a TableDist2D over its texels (src/table_dist.cpp:40-151).
struct LightRec {
    int type;  // 0 area, 1 envmap
    int shape_id;
    D3 intensity;
    double to_world[9], to_local[9];  // 3x3 (vectors only)
    double scale;
    int env_tex;
    int env_w, env_h;
    int pad;
    const float *env_rgb;         // level 0 of the envmap's texture, FP32 as given
    const double *cdf_rows;       // env_h x (env_w + 1)
    const double *pdf_rows;       // env_h x env_w
    const double *cdf_marginals;  // env_h + 1
    const double *pdf_marginals;  // env_h
};
// A spectrum-valued volume (src/volume.h:13-31): a constant, or a grid of
// FP32 voxels (1 channel: grey, 3: RGB; x fastest) over a box, read with
// FP64 trilinear interpolation.  max_value is scale times the per-channel
// maximum over the voxels (at least 0), or scale * value.
struct VolumeRec {
    int type;  // 0 constant, 1 grid
    int channels;
    int rx, ry, rz;
    D3 value;
    D3 p_min, p_max;
    double scale;
    D3 max_value;
    const float *data;
};
struct MediumRec {  // src/medium.h:10-22
    int type;  // 0 homogeneous, 1 heterogeneous
    D3 sigma_a, sigma_s;
    int phase_type;
    double g;
    VolumeRec albedo, density;
};
struct CameraRec {  // src/camera.h:7-22
    double sample_to_cam[16];
    double cam_to_world[16];
    int width, height;
    int filter_type;
    double filter_param;
    int medium_id;
};

// BVH node: the two children's boxes (Aila & Laine layout), 64 bytes.
//   lxy = (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y)
//   rxy = (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y)
//   zz  = (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)
//   ch  = (child0, child1, -, -); child >= 0: inner node; child < 0: leaf
//         ~child = (first_slot << 3) | (count - 1)
// 4-wide node (128 bytes): the boxes of up to four children in SoA form
// (child c in lane c of each float4), their references (inner node index >= 0,
// leaf ~((first_slot << 3) | (count - 1)), kNoChild for an empty slot).
constexpr int kNoChild = 0x7fffffff;
struct alignas(16) Bvh4Node {
    float4 lox, hix, loy, hiy, loz, hiz;
    int4 ch;
    int4 meta;  // meta.x: number of children
};

struct alignas(16) BvhNode {
    float4 lxy, rxy, zz;
    int4 ch;
};

// The traversal's copy of a 4-wide node (80 bytes, five 16-byte loads instead
// of seven): the child boxes as IEEE half offsets from the children's common
// lower corner, in units of a power-of-two scale,
//   plane = org.xyz + half * org.w,
// rounded outwards (lower planes down, upper planes up) with a margin of
// 2^-18 of the node's extent, so every quantised box holds its FP32 box and
// the arithmetic of the quantised slab test stays on the conservative side.
// Per axis, qa = (lo0 | lo1 << 16, lo2 | lo3 << 16, hi0 | hi1 << 16,
// hi2 | hi3 << 16); ch as in Bvh4Node.  A half operand of an FMA is one
// v_fma_mix_f32 on gfx950, so the slab test costs what the FP32 one does.
struct alignas(16) Bvh4QNode {
    float4 org;
    uint4 qx, qy, qz;
    int4 ch;
};
static_assert(sizeof(Bvh4QNode) == 80, "Bvh4QNode layout");

// IEEE binary16 bits of the largest half <= v (DOWN) or the smallest half >= v.
template <bool DOWN>
__host__ __device__ inline uint32_t half_bits_rounded(double v) {
    if (v < 0) return half_bits_rounded<!DOWN>(-v) ^ 0x8000u;
    if (v == 0) return 0u;
    if (v > 65504.0) return DOWN ? 0x7bffu : 0x7c00u;
    int k;
    (void)frexp(v, &k);  // v = f * 2^k, f in [0.5, 1)
    const int e = k - 1;  // 2^e <= v < 2^(e+1)
    if (e < -14) {        // subnormal: units of 2^-24, bits = units
        const double u = ldexp(v, 24);
        return uint32_t(DOWN ? floor(u) : ceil(u));
    }
    const double q = ldexp(v, 10 - e);  // in [1024, 2048)
    const uint32_t m = uint32_t(DOWN ? floor(q) : ceil(q));
    return (uint32_t(e + 15) << 10) + m - 1024u;  // m == 2048 carries into the exponent
}

// Value of IEEE binary16 bits b.
__host__ __device__ inline double half_value(uint32_t b) {
    const int e = (b >> 10) & 31, m = b & 1023;
    const double v = e == 0 ? ldexp(double(m), -24) : (e == 31 ? (m ? NAN : INFINITY) : ldexp(double(1024 + m), e - 25));
    return (b & 0x8000u) ? -v : v;
}

__host__ __device__ inline Bvh4QNode quantize_bvh4(const Bvh4Node &n) {
    const float lo[3][4] = {{n.lox.x, n.lox.y, n.lox.z, n.lox.w}, {n.loy.x, n.loy.y, n.loy.z, n.loy.w},
                            {n.loz.x, n.loz.y, n.loz.z, n.loz.w}};
    const float hi[3][4] = {{n.hix.x, n.hix.y, n.hix.z, n.hix.w}, {n.hiy.x, n.hiy.y, n.hiy.z, n.hiy.w},
                            {n.hiz.x, n.hiz.y, n.hiz.z, n.hiz.w}};
    const int ch[4] = {n.ch.x, n.ch.y, n.ch.z, n.ch.w};
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
    bool any = false;
    for (int c = 0; c < 4; c++) {
        if (ch[c] == kNoChild) continue;
        for (int a = 0; a < 3; a++) {
            bmin[a] = any ? fminf(bmin[a], lo[a][c]) : lo[a][c];
            bmax[a] = any ? fmaxf(bmax[a], hi[a][c]) : hi[a][c];
        }
        any = true;
    }
    double ext = 0;
    for (int a = 0; a < 3; a++) ext = fmax(ext, double(bmax[a]) - double(bmin[a]));
    int e = 0;
    if (ext > 0) {
        int k;
        (void)frexp(ext, &k);
        e = k - 14;  // ext / 2^e in [2^13, 2^14)
        e = e < -60 ? -60 : (e > 60 ? 60 : e);
    }
    const double s = ldexp(1.0, e), margin = ldexp(ext, -18);
    uint32_t q[3][8] = {};
    for (int c = 0; c < 4; c++) {
        if (ch[c] == kNoChild) continue;
        for (int a = 0; a < 3; a++) {
            q[a][c] = half_bits_rounded<true>((double(lo[a][c]) - double(bmin[a]) - margin) / s);
            q[a][4 + c] = half_bits_rounded<false>((double(hi[a][c]) - double(bmin[a]) + margin) / s);
        }
    }
    Bvh4QNode o;
    o.org = make_float4(bmin[0], bmin[1], bmin[2], float(s));
    uint4 *qa[3] = {&o.qx, &o.qy, &o.qz};
    for (int a = 0; a < 3; a++)
        *qa[a] = make_uint4(q[a][0] | q[a][1] << 16, q[a][2] | q[a][3] << 16, q[a][4] | q[a][5] << 16,
                            q[a][6] | q[a][7] << 16);
    o.ch = n.ch;
    return o;
}

struct DeviceScene {
    // traversal
    const Bvh4Node *nodes;
    const Bvh4QNode *qnodes;  // the traversal's copy of nodes (quantize_bvh4)
    const float4 *prims;      // 3 float4 per leaf slot: v0|gid, v1|kind, v2|-
    const double4 *spheres;   // centre, radius per sphere primitive
    int num_nodes;
    int num_slots;
    // geometry for shading / light sampling
    const float *vpos;        // xyz per global vertex
    const float *vnrm;        // xyz per global vertex (0 where the shape has none)
    const float *vuv;         // uv per global vertex
    const int *tri_vidx;      // 3 global vertex ids per global triangle
    const int *prim_shape;    // shape id per global primitive id
    const int8_t *slot_bucket; // per leaf slot: shading bucket of its material when hits are sorted
    const int8_t *slot_passthru; // per leaf slot: 1 for an index-matched medium boundary (no material)
    const int *prim_local;    // primitive id inside its shape
    const ShapeRec *shapes;
    const MatRec *materials;
    const LightRec *lights;
    const MediumRec *media;
    const TexRec *textures;
    const double *tex_data;   // every texture level's pixels
    const double *light_pmf;  // num_lights
    const double *light_cdf;  // num_lights + 1
    const int *light_guide;   // light_guide_m + 1: guide table of light_cdf (sample_cdf_guided)
    const double *tri_cdf;    // concatenated per-light-shape triangle CDFs
    int num_lights;
    int envmap_light_id;
    int num_triangles;
    int light_guide_m;        // a power of two
    CameraRec camera;
    double bound_radius;
    double eps;  // min(1e-5 * r, 0.01) (src/scene.h:99-105)
};

// ---------------------------------------------------------------- textures
// src/lajolla.h:48-61
LJX_HD int imod(int a, int b) {
    int r = a % b;
    return r < 0 ? r + b : r;
}
LJX_HD double rmod(double a, double b) {
    double r = fmod(a, b);
    return r < 0 ? r + b : r;
}
LJX_HD D3 texel(const double *data, const TexRec &t, int level, int x, int y) {
    const double *p = data + t.off[level] + size_t(t.channels) * (size_t(y) * t.w[level] + x);
    return t.channels == 1 ? D3{p[0], p[0], p[0]} : D3{p[0], p[1], p[2]};
}
// src/mipmap.h:50-72: bilinear on one level, pixel centres at half-integers
LJX_HD D3 mip_lookup(const double *data, const TexRec &t, double u, double v, int level) {
    const int w = t.w[level], h = t.h[level];
    u = u * w - 0.5;
    v = v * h - 0.5;
    const int ufi = imod(int(u), w), vfi = imod(int(v), h);
    const int uci = imod(ufi + 1, w), vci = imod(vfi + 1, h);
    const double uo = u - ufi, vo = v - vfi;
    return texel(data, t, level, ufi, vfi) * (1 - uo) * (1 - vo) + texel(data, t, level, ufi, vci) * (1 - uo) * vo +
           texel(data, t, level, uci, vfi) * uo * (1 - vo) + texel(data, t, level, uci, vci) * uo * vo;
}
// src/mipmap.h:74-88: trilinear between levels
LJX_HD D3 mip_lookup_level(const double *data, const TexRec &t, double u, double v, double level) {
    const int n = t.levels;
    if (level <= 0) return mip_lookup(data, t, u, v, 0);
    if (level < double(n - 1)) {
        const int fl = min(max(int(floor(level)), 0), n - 1), cl = min(max(fl + 1, 0), n - 1);
        const double lo = level - fl;
        return mip_lookup(data, t, u, v, fl) * (1 - lo) + mip_lookup(data, t, u, v, cl) * lo;
    }
    return mip_lookup(data, t, u, v, n - 1);
}
// src/texture.h:131-154
LJX_HD D3 tex_eval(const double *data, const TexRec &t, double u, double v, double footprint) {
    const double lu = rmod(u * t.uscale + t.uoffset, 1.0), lv = rmod(v * t.vscale + t.voffset, 1.0);
    if (t.type == 2) {
        const int x = 2 * imod(int(lu * 2), 2) - 1, y = 2 * imod(int(lv * 2), 2) - 1;
        return x * y == 1 ? t.color0 : t.color1;
    }
    const double scaled = max(t.w[0], t.h[0]) * fmax(t.uscale, t.vscale) * footprint;
    const double level = log2(scaled < double(1e-8f) ? double(1e-8f) : scaled);  // std::max(scaled, 1e-8f)
    return mip_lookup_level(data, t, lu, lv, level);
}
// a Texture<Real> field: a 1-channel image as is, otherwise the average
LJX_HD double tex_eval1(const double *data, const TexRec &t, double u, double v, double footprint) {
    D3 c = tex_eval(data, t, u, v, footprint);
    return (t.type == 1 && t.channels == 1) ? c.x : (c.x + c.y + c.z) / 3;
}

// ---------------------------------------------------------------- sampling
// TableDist1D sampling: upper_bound over the CDF, clamped (src/table_dist.cpp:27-33).
LJX_HD int sample_cdf(const double *cdf, int n, double u) {
    int lo = 0, hi = n + 1;  // search [0, n+1) for the first cdf > u
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cdf[mid] <= u) lo = mid + 1;
        else hi = mid;
    }
    int k = lo - 1;
    return k < 0 ? 0 : (k > n - 1 ? n - 1 : k);
}

// The same index through a guide table (the cutpoint method of Chen and
// Asau, 1974): guide[k] is the upper bound of k / m in the CDF, so for u in
// [k / m, (k + 1) / m) the first cdf > u lies in [guide[k], guide[k + 1]] and
// the search runs there only (one or two steps when m is near n, against
// log2(n) for the whole table).  m is a power of two, so k = floor(u m) is
// exact.
LJX_HD int sample_cdf_guided(const double *cdf, int n, double u, const int *guide, int m) {
    int k = int(u * m);
    k = k < 0 ? 0 : (k > m - 1 ? m - 1 : k);
    int lo = guide[k], hi = guide[k + 1];
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cdf[mid] <= u) lo = mid + 1;
        else hi = mid;
    }
    int r = lo - 1;
    return r < 0 ? 0 : (r > n - 1 ? n - 1 : r);
}

// Builds the guide table of a CDF of n1 = n + 1 entries for m (a power of
// two) cells: guide[k] = the first index whose cdf exceeds k / m (n1 if
// none), k = 0..m.
inline void cdf_guide(const double *cdf, int n1, int m, int *guide) {
    int i = 0;
    for (int k = 0; k <= m; k++) {
        const double x = double(k) / m;
        while (i < n1 && cdf[i] <= x) i++;
        guide[k] = i;
    }
}

// Cosine hemisphere (src/material.cpp:4-11).
LJX_HD D3 sample_cos_hemisphere(double u0, double u1) {
    double phi = kTwoPi * u0;
    double tmp = sqrt(clampd(1 - u1, 0, 1));
    double s, c;
    sincos(phi, &s, &c);
    return {c * tmp, s * tmp, sqrt(clampd(u1, 0, 1))};
}

// ------------------------------------------------------------- microfacets
// src/microfacet.h:23-114
LJX_HD double fresnel_dielectric(double n_dot_i, double n_dot_t, double eta) {
    double rs = (n_dot_i - eta * n_dot_t) / (n_dot_i + eta * n_dot_t);
    double rp = (eta * n_dot_i - n_dot_t) / (eta * n_dot_i + n_dot_t);
    return (rs * rs + rp * rp) / 2;
}
LJX_HD double fresnel_dielectric(double n_dot_i, double eta) {
    double n_dot_t_sq = 1 - (1 - n_dot_i * n_dot_i) / (eta * eta);
    if (n_dot_t_sq < 0) return 1;
    return fresnel_dielectric(fabs(n_dot_i), sqrt(n_dot_t_sq), eta);
}
LJX_HD double gtr2(double n_dot_h, double roughness) {
    double alpha = roughness * roughness;
    double a2 = alpha * alpha;
    double t = 1 + (a2 - 1) * n_dot_h * n_dot_h;
    return a2 / (kPi * t * t);
}
LJX_HD double smith_gtr2(D3 v, double roughness) {
    double alpha = roughness * roughness;
    double a2 = alpha * alpha;
    double Lambda = (-1 + sqrt(1 + (v.x * v.x * a2 + v.y * v.y * a2) / (v.z * v.z))) / 2;
    return 1 / (1 + Lambda);
}
// Anisotropic GGX D and Smith G1 (handouts/homework1.tex:175-259).
LJX_HD double ggx_aniso_D(D3 hl, double ax, double ay) {
    double t = hl.x * hl.x / (ax * ax) + hl.y * hl.y / (ay * ay) + hl.z * hl.z;
    return 1 / (kPi * ax * ay * t * t);
}
LJX_HD double ggx_aniso_G1(D3 wl, double ax, double ay) {
    double a = (wl.x * ax) * (wl.x * ax) + (wl.y * ay) * (wl.y * ay);
    double Lambda = (sqrt(1 + a / (wl.z * wl.z)) - 1) / 2;
    return 1 / (1 + Lambda);
}
// Visible-normal sampling (Heitz 2018), src/microfacet.h:85-114, with
// separate x/y roughness for the anisotropic Disney lobes.
LJX_HD D3 sample_vndf(D3 in, double ax, double ay, double u0, double u1) {
    double sgn = 1;
    if (in.z < 0) {
        in = -in;
        sgn = -1;
    }
    D3 hemi = normalize(D3{ax * in.x, ay * in.y, in.z});
    double r = sqrt(u0);
    double s, c;
    sincos(2 * kPi * u1, &s, &c);
    double t1 = r * c, t2 = r * s;
    double sh = (1 + hemi.z) / 2;
    t2 = (1 - sh) * sqrt(1 - t1 * t1) + sh * t2;
    D3 disk{t1, t2, sqrt(fmax(0.0, 1 - t1 * t1 - t2 * t2))};
    Frame hf = frame_from_n(hemi);
    D3 hn = to_world(hf, disk);
    return sgn * normalize(D3{ax * hn.x, ay * hn.y, fmax(0.0, hn.z)});
}
LJX_HD D3 reflect_about(D3 in, D3 h) { return normalize(-in + 2 * dot(in, h) * h); }

// -------------------------------------------------------------- BSDF kit
struct ShadeGeom {
    D3 ng;      // geometric normal, flipped to the shading side
    Frame sf;   // shading frame
};
struct BsdfSample {
    D3 dir;
    double eta;
    double roughness;
    int valid;
};
LJX_HD Frame front_frame(const ShadeGeom &g, D3 in) {
    return dot(g.sf.n, in) < 0 ? flip(g.sf) : g.sf;
}
LJX_HD Frame dielectric_frame(const ShadeGeom &g, D3 in) {
    return dot(g.sf.n, in) * dot(g.ng, in) < 0 ? flip(g.sf) : g.sf;
}
LJX_HD double clamp_rough(double r) { return clampd(r, 0.01, 1.0); }

// Lambertian (src/materials/lambertian.inl:1-50); eval includes the cosine.
LJX_HD D3 lambert_eval(D3 refl, D3 in, D3 out, const ShadeGeom &g) {
    if (dot(g.ng, in) < 0 || dot(g.ng, out) < 0) return {0, 0, 0};
    Frame f = front_frame(g, in);
    return fmax(dot(f.n, out), 0.0) * refl / kPi;
}
LJX_HD double cos_pdf(D3 in, D3 out, const ShadeGeom &g) {
    if (dot(g.ng, in) < 0 || dot(g.ng, out) < 0) return 0;
    Frame f = front_frame(g, in);
    return fmax(dot(f.n, out), 0.0) / kPi;
}
LJX_HD BsdfSample cos_sample(D3 in, const ShadeGeom &g, double u0, double u1, double rough) {
    if (dot(g.ng, in) < 0) return BsdfSample{{0, 0, 0}, 0, 0, 0};
    Frame f = front_frame(g, in);
    return BsdfSample{to_world(f, sample_cos_hemisphere(u0, u1)), 0, rough, 1};
}

// RoughPlastic (src/materials/roughplastic.inl:3-161).
LJX_HD D3 plastic_eval(const MatRec &m, D3 in, D3 out, const ShadeGeom &g) {
    if (dot(g.ng, in) < 0 || dot(g.ng, out) < 0) return {0, 0, 0};
    Frame f = front_frame(g, in);
    D3 h = normalize(in + out);
    double ndh = dot(f.n, h), ndi = dot(f.n, in), ndo = dot(f.n, out);
    if (ndo <= 0 || ndh <= 0) return {0, 0, 0};
    double r = clamp_rough(m.roughness);
    double Fo = fresnel_dielectric(dot(h, out), m.eta);
    double D = gtr2(ndh, r);
    double G = smith_gtr2(to_local(f, in), r) * smith_gtr2(to_local(f, out), r);
    D3 spec = m.b * (G * Fo * D) / (4 * ndi * ndo);
    double Fi = fresnel_dielectric(dot(h, in), m.eta);
    D3 diff = m.a * ((1 - Fo) * (1 - Fi) / kPi);
    return (spec + diff) * ndo;
}
LJX_HD double plastic_pdf(const MatRec &m, D3 in, D3 out, const ShadeGeom &g) {
    if (dot(g.ng, in) < 0 || dot(g.ng, out) < 0) return 0;
    Frame f = front_frame(g, in);
    D3 h = normalize(in + out);
    double ndi = dot(f.n, in), ndo = dot(f.n, out), ndh = dot(f.n, h);
    if (ndo <= 0 || ndh <= 0) return 0;
    double lS = luminance(m.b), lR = luminance(m.a);
    if (lS + lR <= 0) return 0;
    double r = clamp_rough(m.roughness);
    double sp = lS / (lS + lR);
    double dp = 1 - sp;
    sp *= smith_gtr2(to_local(f, in), r) * gtr2(ndh, r) / (4 * ndi);
    dp *= ndo / kPi;
    return sp + dp;
}
LJX_HD BsdfSample plastic_sample(const MatRec &m, D3 in, const ShadeGeom &g, double u0, double u1, double w) {
    if (dot(g.ng, in) < 0) return BsdfSample{{0, 0, 0}, 0, 0, 0};
    Frame f = front_frame(g, in);
    double lS = luminance(m.b), lR = luminance(m.a);
    if (lS + lR <= 0) return BsdfSample{{0, 0, 0}, 0, 0, 0};
    if (w < lS / (lS + lR)) {
        double r = clamp_rough(m.roughness);
        double alpha = r * r;
        D3 h = to_world(f, sample_vndf(to_local(f, in), alpha, alpha, u0, u1));
        return BsdfSample{reflect_about(in, h), 0, r, 1};
    }
    return BsdfSample{to_world(f, sample_cos_hemisphere(u0, u1)), 0, 1, 1};
}

// Microfacet dielectric lobe shared by RoughDielectric
// (src/materials/roughdielectric.inl:3-177) and the Disney glass lobe
// (handouts/homework1.tex:336-404).
struct Glass {
    D3 Kr, Kt;
    double eta;
    int aniso;
    double rough;   // clamped (isotropic) or raw Disney roughness
    double ax, ay;
};
LJX_HD double glass_D(const Glass &gl, const Frame &f, D3 h) {
    return gl.aniso ? ggx_aniso_D(to_local(f, h), gl.ax, gl.ay) : gtr2(dot(f.n, h), gl.rough);
}
LJX_HD double glass_G1(const Glass &gl, const Frame &f, D3 w) {
    return gl.aniso ? ggx_aniso_G1(to_local(f, w), gl.ax, gl.ay) : smith_gtr2(to_local(f, w), gl.rough);
}
LJX_HD D3 glass_eval(const Glass &gl, D3 in, D3 out, const ShadeGeom &g, bool to_view) {
    bool refl = dot(g.ng, in) * dot(g.ng, out) > 0;
    Frame f = dielectric_frame(g, in);
    double eta = dot(g.ng, in) > 0 ? gl.eta : 1 / gl.eta;
    D3 h = refl ? normalize(in + out) : normalize(in + out * eta);
    if (dot(h, f.n) < 0) h = -h;
    double hdi = dot(h, in);
    double F = fresnel_dielectric(hdi, eta);
    double D = glass_D(gl, f, h);
    double G = glass_G1(gl, f, in) * glass_G1(gl, f, out);
    if (refl) return gl.Kr * (F * D * G / (4 * fabs(dot(f.n, in))));
    double eta_factor = to_view ? 1.0 : 1 / (eta * eta);
    double hdo = dot(h, out);
    double sd = hdi + eta * hdo;
    return gl.Kt * (eta_factor * (1 - F) * D * G * eta * eta * fabs(hdo * hdi) / (fabs(dot(f.n, in)) * sd * sd));
}
LJX_HD double glass_pdf(const Glass &gl, D3 in, D3 out, const ShadeGeom &g) {
    bool refl = dot(g.ng, in) * dot(g.ng, out) > 0;
    Frame f = dielectric_frame(g, in);
    double eta = dot(g.ng, in) > 0 ? gl.eta : 1 / gl.eta;
    D3 h = refl ? normalize(in + out) : normalize(in + out * eta);
    if (dot(h, f.n) < 0) h = -h;
    double hdi = dot(h, in);
    double F = fresnel_dielectric(hdi, eta);
    double D = glass_D(gl, f, h);
    double Gi = glass_G1(gl, f, in);
    if (refl) return F * D * Gi / (4 * fabs(dot(f.n, in)));
    double hdo = dot(h, out);
    double sd = hdi + eta * hdo;
    double dh_dout = eta * eta * hdo / (sd * sd);
    return (1 - F) * D * Gi * fabs(dh_dout * hdi / dot(f.n, in));
}
LJX_HD BsdfSample glass_sample(const Glass &gl, D3 in, const ShadeGeom &g, double u0, double u1, double w) {
    double eta = dot(g.ng, in) > 0 ? gl.eta : 1 / gl.eta;
    Frame f = dielectric_frame(g, in);
    double ax = gl.aniso ? gl.ax : gl.rough * gl.rough;
    double ay = gl.aniso ? gl.ay : gl.rough * gl.rough;
    D3 h = to_world(f, sample_vndf(to_local(f, in), ax, ay, u0, u1));
    if (dot(h, f.n) < 0) h = -h;
    double hdi = dot(h, in);
    double F = fresnel_dielectric(hdi, eta);
    if (w <= F) return BsdfSample{reflect_about(in, h), 0, gl.rough, 1};
    double hdo2 = 1 - (1 - hdi * hdi) / (eta * eta);
    if (hdo2 <= 0) return BsdfSample{{0, 0, 0}, 0, 0, 0};
    if (hdi < 0) h = -h;
    double hdo = sqrt(hdo2);
    D3 refr = -in / eta + (fabs(hdi) / eta - hdo) * h;
    return BsdfSample{refr, eta, gl.rough, 1};
}
LJX_HD Glass dielectric_glass(const MatRec &m) {
    Glass gl;
    gl.Kr = m.a;
    gl.Kt = m.b;
    gl.eta = m.eta;
    gl.aniso = 0;
    gl.rough = clamp_rough(m.roughness);
    gl.ax = gl.ay = gl.rough * gl.rough;
    return gl;
}

// ---------------------------------------------------------------- Disney
// handouts/homework1.tex:93-548 (the reference's Disney materials are stubs).
LJX_HD void disney_alphas(double rough, double aniso, double &ax, double &ay) {
    double aspect = sqrt(1 - 0.9 * aniso);
    ax = fmax(1e-4, rough * rough / aspect);
    ay = fmax(1e-4, rough * rough * aspect);
}
LJX_HD Glass disney_glass(D3 base, double rough, double aniso, double eta) {
    Glass gl;
    gl.Kr = base;
    gl.Kt = vsqrt0(base);
    gl.eta = eta;
    gl.aniso = 1;
    gl.rough = rough;
    disney_alphas(rough, aniso, gl.ax, gl.ay);
    return gl;
}
LJX_HD D3 tint_color(D3 base) {
    double l = luminance(base);
    return l > 0 ? base / l : D3{1, 1, 1};
}
LJX_HD D3 disney_diffuse_eval(D3 base, double rough, double subsurface, D3 in, D3 out, const ShadeGeom &g) {
    if (dot(g.ng, in) < 0 || dot(g.ng, out) < 0) return {0, 0, 0};
    Frame f = front_frame(g, in);
    D3 h = normalize(in + out);
    double ni = fabs(dot(f.n, in)), no = fabs(dot(f.n, out));
    double ho = fabs(dot(h, out));
    double fd90 = 0.5 + 2 * rough * ho * ho;
    double fdi = 1 + (fd90 - 1) * pow5(1 - ni), fdo = 1 + (fd90 - 1) * pow5(1 - no);
    D3 base_d = base * (fdi * fdo * no / kPi);
    double fss90 = rough * ho * ho;
    double fsi = 1 + (fss90 - 1) * pow5(1 - ni), fso = 1 + (fss90 - 1) * pow5(1 - no);
    D3 ss = base * (1.25 / kPi * (fsi * fso * (1 / (ni + no) - 0.5) + 0.5) * no);
    return (1 - subsurface) * base_d + subsurface * ss;
}
LJX_HD D3 disney_metal_eval(D3 F0, double ax, double ay, D3 in, D3 out, const ShadeGeom &g) {
    if (dot(g.ng, in) < 0 || dot(g.ng, out) < 0) return {0, 0, 0};
    Frame f = front_frame(g, in);
    D3 h = normalize(in + out);
    double ndi = dot(f.n, in), ndo = dot(f.n, out), ndh = dot(f.n, h);
    if (ndo <= 0 || ndh <= 0) return {0, 0, 0};
    D3 Fm = F0 + (1 - F0) * pow5(1 - fabs(dot(h, out)));
    double D = ggx_aniso_D(to_local(f, h), ax, ay);
    double G = ggx_aniso_G1(to_local(f, in), ax, ay) * ggx_aniso_G1(to_local(f, out), ax, ay);
    return Fm * (D * G / (4 * fabs(ndi)));
}
LJX_HD double disney_metal_pdf(double ax, double ay, D3 in, D3 out, const ShadeGeom &g) {
    if (dot(g.ng, in) < 0 || dot(g.ng, out) < 0) return 0;
    Frame f = front_frame(g, in);
    D3 h = normalize(in + out);
    double ndi = dot(f.n, in), ndo = dot(f.n, out), ndh = dot(f.n, h);
    if (ndo <= 0 || ndh <= 0) return 0;
    return ggx_aniso_D(to_local(f, h), ax, ay) * ggx_aniso_G1(to_local(f, in), ax, ay) / (4 * fabs(ndi));
}
LJX_HD BsdfSample disney_metal_sample(double ax, double ay, double rough, D3 in, const ShadeGeom &g, double u0,
                                      double u1) {
    if (dot(g.ng, in) < 0) return BsdfSample{{0, 0, 0}, 0, 0, 0};
    Frame f = front_frame(g, in);
    D3 h = to_world(f, sample_vndf(to_local(f, in), ax, ay, u0, u1));
    return BsdfSample{reflect_about(in, h), 0, rough, 1};
}
LJX_HD double cc_alpha(double gloss) { return (1 - gloss) * 0.1 + gloss * 0.001; }
LJX_HD double cc_D(double hz, double ag) {
    double a2 = ag * ag;
    return (a2 - 1) / (kPi * log(a2) * (1 + (a2 - 1) * hz * hz));
}
LJX_HD D3 disney_cc_eval(double gloss, D3 in, D3 out, const ShadeGeom &g) {
    if (dot(g.ng, in) < 0 || dot(g.ng, out) < 0) return {0, 0, 0};
    Frame f = front_frame(g, in);
    D3 h = normalize(in + out);
    double ndi = dot(f.n, in), ndo = dot(f.n, out), ndh = dot(f.n, h);
    if (ndo <= 0 || ndh <= 0) return {0, 0, 0};
    const double R0 = 0.04;  // ((1.5 - 1) / (1.5 + 1))^2
    double Fc = R0 + (1 - R0) * pow5(1 - fabs(dot(h, out)));
    double Dc = cc_D(to_local(f, h).z, cc_alpha(gloss));
    double Gc = ggx_aniso_G1(to_local(f, in), 0.25, 0.25) * ggx_aniso_G1(to_local(f, out), 0.25, 0.25);
    double v = Fc * Dc * Gc / (4 * fabs(ndi));
    return {v, v, v};
}
LJX_HD double disney_cc_pdf(double gloss, D3 in, D3 out, const ShadeGeom &g) {
    if (dot(g.ng, in) < 0 || dot(g.ng, out) < 0) return 0;
    Frame f = front_frame(g, in);
    D3 h = normalize(in + out);
    double ndo = dot(f.n, out), ndh = dot(f.n, h);
    if (ndo <= 0 || ndh <= 0) return 0;
    return cc_D(to_local(f, h).z, cc_alpha(gloss)) * fabs(ndh) / (4 * fabs(dot(h, out)));
}
LJX_HD BsdfSample disney_cc_sample(double gloss, D3 in, const ShadeGeom &g, double u0, double u1) {
    if (dot(g.ng, in) < 0) return BsdfSample{{0, 0, 0}, 0, 0, 0};
    Frame f = front_frame(g, in);
    double ag = cc_alpha(gloss);
    double a2 = ag * ag;
    double ch = sqrt(fmax(0.0, (1 - pow(a2, 1 - u0)) / (1 - a2)));
    double sh = sqrt(fmax(0.0, 1 - ch * ch));
    double s, c;
    sincos(2 * kPi * u1, &s, &c);
    D3 h = to_world(f, D3{sh * c, sh * s, ch});
    return BsdfSample{reflect_about(in, h), 0, sqrt(ag), 1};
}
LJX_HD D3 disney_sheen_eval(D3 base, double tint, D3 in, D3 out, const ShadeGeom &g) {
    if (dot(g.ng, in) < 0 || dot(g.ng, out) < 0) return {0, 0, 0};
    Frame f = front_frame(g, in);
    D3 h = normalize(in + out);
    D3 C = (1 - tint) + tint * tint_color(base);
    return C * (pow5(1 - fabs(dot(h, out))) * fabs(dot(f.n, out)));
}
struct DisneyLobes {
    double w_diffuse, w_sheen, w_metal, w_cc, w_glass;
};
LJX_HD DisneyLobes disney_lobes(const MatRec &m) {
    DisneyLobes l;
    l.w_diffuse = (1 - m.specular_transmission) * (1 - m.metallic);
    l.w_sheen = (1 - m.metallic) * m.sheen;
    l.w_metal = 1 - m.specular_transmission * (1 - m.metallic);
    l.w_cc = 0.25 * m.clearcoat;
    l.w_glass = (1 - m.metallic) * m.specular_transmission;
    return l;
}
LJX_HD D3 disney_F0(const MatRec &m) {
    double R0 = (m.eta - 1) * (m.eta - 1) / ((m.eta + 1) * (m.eta + 1));
    D3 Ks = (1 - m.specular_tint) + m.specular_tint * tint_color(m.a);
    return m.specular * R0 * (1 - m.metallic) * Ks + m.metallic * m.a;
}
LJX_HD D3 disney_bsdf_eval(const MatRec &m, D3 in, D3 out, const ShadeGeom &g, bool to_view) {
    double ax, ay;
    disney_alphas(m.roughness, m.anisotropic, ax, ay);
    Glass gl = disney_glass(m.a, m.roughness, m.anisotropic, m.eta);
    DisneyLobes w = disney_lobes(m);
    if (dot(g.ng, in) <= 0) return w.w_glass * glass_eval(gl, in, out, g, to_view);
    D3 r = w.w_diffuse * disney_diffuse_eval(m.a, m.roughness, m.subsurface, in, out, g);
    r = r + w.w_sheen * disney_sheen_eval(m.a, m.sheen_tint, in, out, g);
    r = r + w.w_metal * disney_metal_eval(disney_F0(m), ax, ay, in, out, g);
    r = r + w.w_cc * disney_cc_eval(m.clearcoat_gloss, in, out, g);
    r = r + w.w_glass * glass_eval(gl, in, out, g, to_view);
    return r;
}
LJX_HD void disney_probs(const MatRec &m, bool inside, double p[4]) {
    if (inside) {
        p[0] = p[1] = p[3] = 0;
        p[2] = 1;
        return;
    }
    DisneyLobes w = disney_lobes(m);
    double tot = w.w_diffuse + w.w_metal + w.w_glass + w.w_cc;
    p[0] = w.w_diffuse / tot;
    p[1] = w.w_metal / tot;
    p[2] = w.w_glass / tot;
    p[3] = w.w_cc / tot;
}
LJX_HD double disney_bsdf_pdf(const MatRec &m, D3 in, D3 out, const ShadeGeom &g) {
    double ax, ay;
    disney_alphas(m.roughness, m.anisotropic, ax, ay);
    Glass gl = disney_glass(m.a, m.roughness, m.anisotropic, m.eta);
    bool inside = dot(g.ng, in) <= 0;
    double p[4];
    disney_probs(m, inside, p);
    if (inside) return glass_pdf(gl, in, out, g);
    double r = 0;
    if (p[0] > 0) r += p[0] * cos_pdf(in, out, g);
    if (p[1] > 0) r += p[1] * disney_metal_pdf(ax, ay, in, out, g);
    if (p[2] > 0) r += p[2] * glass_pdf(gl, in, out, g);
    if (p[3] > 0) r += p[3] * disney_cc_pdf(m.clearcoat_gloss, in, out, g);
    return r;
}
LJX_HD BsdfSample disney_bsdf_sample(const MatRec &m, D3 in, const ShadeGeom &g, double u0, double u1,
                                     double w) {
    double ax, ay;
    disney_alphas(m.roughness, m.anisotropic, ax, ay);
    Glass gl = disney_glass(m.a, m.roughness, m.anisotropic, m.eta);
    bool inside = dot(g.ng, in) <= 0;
    double p[4];
    disney_probs(m, inside, p);
    if (inside) return glass_sample(gl, in, g, u0, u1, w);
    double c0 = p[0], c1 = c0 + p[1], c2 = c1 + p[2];
    if (w < c0) return cos_sample(in, g, u0, u1, m.roughness);
    if (w < c1) return disney_metal_sample(ax, ay, m.roughness, in, g, u0, u1);
    if (w < c2) return glass_sample(gl, in, g, u0, u1, (w - c1) / p[2]);
    return disney_cc_sample(m.clearcoat_gloss, in, g, u0, u1);
}

// -------------------------------------------------- material dispatch (MT)
// MT is the material type tag of include/lajolla_amd.h; MT = -1 dispatches
// at run time.  Shading kernels are instantiated per type so that each
// sorted bucket runs straight-line code.
// The material type a BSDF call switches on: MT itself in a kernel of one
// type; the record's own type under runtime dispatch (MT = -1); and under
// runtime dispatch over lajolla's three basic materials only (MT =
// kMtBasic: Lambertian, RoughPlastic, RoughDielectric; the host picks it
// only for scenes that use no other type) the type clamped to them, so the
// Disney lobes drop out of the kernel and its register budget.
constexpr int kMtBasic = -2;
template <int MT>
LJX_HD int dispatch_type(const MatRec &m) {
    if constexpr (MT >= 0) return MT;
    else if constexpr (MT == kMtBasic) return m.type < 2 ? m.type : 2;
    else return m.type;
}

template <int MT>
LJX_HD D3 bsdf_eval(const MatRec &m, D3 in, D3 out, const ShadeGeom &g, bool to_view = false) {
    const int t = dispatch_type<MT>(m);
    switch (t) {
    case 0: return lambert_eval(m.a, in, out, g);
    case 1: return plastic_eval(m, in, out, g);
    case 2: return glass_eval(dielectric_glass(m), in, out, g, to_view);
    case 3: return disney_diffuse_eval(m.a, m.roughness, m.subsurface, in, out, g);
    case 4: {
        double ax, ay;
        disney_alphas(m.roughness, m.anisotropic, ax, ay);
        return disney_metal_eval(m.a, ax, ay, in, out, g);
    }
    case 5: return glass_eval(disney_glass(m.a, m.roughness, m.anisotropic, m.eta), in, out, g, to_view);
    case 6: return disney_cc_eval(m.clearcoat_gloss, in, out, g);
    case 7: return disney_sheen_eval(m.a, m.sheen_tint, in, out, g);
    default: return disney_bsdf_eval(m, in, out, g, to_view);
    }
}
template <int MT>
LJX_HD double bsdf_pdf(const MatRec &m, D3 in, D3 out, const ShadeGeom &g) {
    const int t = dispatch_type<MT>(m);
    switch (t) {
    case 0: return cos_pdf(in, out, g);
    case 1: return plastic_pdf(m, in, out, g);
    case 2: return glass_pdf(dielectric_glass(m), in, out, g);
    case 3: return cos_pdf(in, out, g);
    case 4: {
        double ax, ay;
        disney_alphas(m.roughness, m.anisotropic, ax, ay);
        return disney_metal_pdf(ax, ay, in, out, g);
    }
    case 5: return glass_pdf(disney_glass(m.a, m.roughness, m.anisotropic, m.eta), in, out, g);
    case 6: return disney_cc_pdf(m.clearcoat_gloss, in, out, g);
    case 7: return cos_pdf(in, out, g);
    default: return disney_bsdf_pdf(m, in, out, g);
    }
}
template <int MT>
LJX_HD BsdfSample bsdf_sample(const MatRec &m, D3 in, const ShadeGeom &g, double u0, double u1, double w) {
    const int t = dispatch_type<MT>(m);
    switch (t) {
    case 0: return cos_sample(in, g, u0, u1, 1.0);
    case 1: return plastic_sample(m, in, g, u0, u1, w);
    case 2: return glass_sample(dielectric_glass(m), in, g, u0, u1, w);
    case 3: return cos_sample(in, g, u0, u1, m.roughness);
    case 4: {
        double ax, ay;
        disney_alphas(m.roughness, m.anisotropic, ax, ay);
        return disney_metal_sample(ax, ay, m.roughness, in, g, u0, u1);
    }
    case 5: return glass_sample(disney_glass(m.a, m.roughness, m.anisotropic, m.eta), in, g, u0, u1, w);
    case 6: return disney_cc_sample(m.clearcoat_gloss, in, g, u0, u1);
    case 7: return cos_sample(in, g, u0, u1, 1.0);
    default: return disney_bsdf_sample(m, in, g, u0, u1, w);
    }
}

// ----------------------------------------------------------------- media
LJX_HD D3 voxel(const VolumeRec &v, int x, int y, int z) {
    const float *q = v.data + size_t(v.channels) * ((size_t(z) * v.ry + y) * v.rx + x);
    if (v.channels == 1) {
        double a = q[0];
        return D3{a, a, a};
    }
    return D3{double(q[0]), double(q[1]), double(q[2])};
}

// src/volume.h:40-81: trilinear interpolation, zero outside the box
LJX_HD D3 vol_lookup(const VolumeRec &v, D3 p) {
    if (v.type == 0) return v.scale * v.value;
    D3 pn = (p - v.p_min) / (v.p_max - v.p_min);
    if (pn.x < 0 || pn.x > 1 || pn.y < 0 || pn.y > 1 || pn.z < 0 || pn.z > 1) return D3{0, 0, 0};
    pn.x *= double(v.rx - 1);
    pn.y *= double(v.ry - 1);
    pn.z *= double(v.rz - 1);
    const int x0 = min(max(int(pn.x), 0), v.rx - 1), x1 = min(x0 + 1, v.rx - 1);
    const int y0 = min(max(int(pn.y), 0), v.ry - 1), y1 = min(y0 + 1, v.ry - 1);
    const int z0 = min(max(int(pn.z), 0), v.rz - 1), z1 = min(z0 + 1, v.rz - 1);
    const double dx = pn.x - x0, dy = pn.y - y0, dz = pn.z - z0;
    D3 acc = voxel(v, x0, y0, z0) * ((1 - dx) * (1 - dy) * (1 - dz)) + voxel(v, x1, y0, z0) * (dx * (1 - dy) * (1 - dz)) +
             voxel(v, x0, y1, z0) * ((1 - dx) * dy * (1 - dz)) + voxel(v, x1, y1, z0) * (dx * dy * (1 - dz)) +
             voxel(v, x0, y0, z1) * ((1 - dx) * (1 - dy) * dz) + voxel(v, x1, y0, z1) * (dx * (1 - dy) * dz) +
             voxel(v, x0, y1, z1) * ((1 - dx) * dy * dz) + voxel(v, x1, y1, z1) * (dx * dy * dz);
    return v.scale * acc;
}

// src/volume.h:120-144: does the ray's [0, tfar] reach the grid box (pbrt's slab test)
LJX_HD bool vol_hit(const VolumeRec &v, D3 o, D3 d, double tfar) {
    if (v.type == 0) return true;
    double t0 = 0, t1 = tfar;
    for (int i = 0; i < 3; i++) {
        double tn = (comp(v.p_min, i) - comp(o, i)) / comp(d, i);
        double tf = (comp(v.p_max, i) - comp(o, i)) / comp(d, i);
        if (tn > tf) {
            double t = tn;
            tn = tf;
            tf = t;
        }
        t0 = tn > t0 ? tn : t0;
        t1 = tf < t1 ? tf : t1;
        if (t0 > t1) return false;
    }
    return true;
}

// src/medium.cpp:27-37, media/homogeneous.inl:1-11, media/heterogeneous.inl:3-21
LJX_HD D3 medium_majorant(const MediumRec &m, D3 o, D3 d, double tfar) {
    if (m.type == 0) return m.sigma_a + m.sigma_s;
    return vol_hit(m.density, o, d, tfar) ? m.density.max_value : D3{0, 0, 0};
}
// sigma_s and sigma_a at p
LJX_HD void medium_sigmas(const MediumRec &m, D3 p, D3 &sig_s, D3 &sig_a) {
    if (m.type == 0) {
        sig_s = m.sigma_s;
        sig_a = m.sigma_a;
        return;
    }
    D3 dens = vol_lookup(m.density, p), alb = vol_lookup(m.albedo, p);
    sig_s = dens * alb;
    sig_a = dens * (1.0 - alb);
}

// ----------------------------------------------------------------- phase
// src/phase_functions/isotropic.inl:1-15, henyeygreenstein.inl:1-48
LJX_HD double phase_value(const MediumRec &m, D3 in, D3 out) {
    if (m.phase_type == 0) return kInvFourPi;
    double g = m.g;
    return kInvFourPi * (1 - g * g) / pow(1 + g * g + 2 * g * dot(in, out), 1.5);
}
LJX_HD D3 phase_sample(const MediumRec &m, D3 in, double u0, double u1) {
    double s, c;
    if (m.phase_type == 0 || fabs(m.g) < 1e-3) {
        double z = 1 - 2 * u0;
        double r = sqrt(fmax(0.0, 1 - z * z));
        sincos(2 * kPi * u1, &s, &c);
        return {r * c, r * s, z};
    }
    double g = m.g;
    double tmp = (g * g - 1) / (2 * u0 * g - (g + 1));
    double ce = (tmp * tmp - (1 + g * g)) / (2 * g);
    double se = sqrt(fmax(1 - ce * ce, 0.0));
    sincos(2 * kPi * u1, &s, &c);
    return to_world(frame_from_n(in), D3{se * c, se * s, ce});
}

// ----------------------------------------------------------------- camera
LJX_HD D3 xform_point(const double *m, D3 p) {
    double t0 = m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3];
    double t1 = m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7];
    double t2 = m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11];
    double t3 = m[12] * p.x + m[13] * p.y + m[14] * p.z + m[15];
    double iw = 1 / t3;
    return {t0 * iw, t1 * iw, t2 * iw};
}
LJX_HD D3 xform_vector(const double *m, D3 v) {
    return {m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
            m[8] * v.x + m[9] * v.y + m[10] * v.z};
}
LJX_HD D3 xform3(const double *m, D3 v) {
    return {m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
            m[6] * v.x + m[7] * v.y + m[8] * v.z};
}
// Pixel filters (src/filters/{box,tent,gaussian}.inl).
LJX_HD void filter_sample(int type, double param, double u0, double u1, double &ox, double &oy) {
    if (type == 0) {
        ox = (2 * u0 - 1) * (param / 2);
        oy = (2 * u1 - 1) * (param / 2);
    } else if (type == 1) {
        double h = param / 2;
        ox = u0 < 0.5 ? h * (sqrt(2 * u0) - 1) : h * (1 - sqrt(1 - 2 * (u0 - 0.5)));
        oy = u1 < 0.5 ? h * (sqrt(2 * u1) - 1) : h * (1 - sqrt(1 - 2 * (u1 - 0.5)));
    } else {
        double r = param * sqrt(-2 * log(fmax(u0, 1e-8)));
        double s, c;
        sincos(2 * kPi * u1, &s, &c);
        ox = r * c;
        oy = r * s;
    }
}
// sample_primary (src/camera.cpp:23-47).
LJX_HD void sample_primary(const CameraRec &c, double sx, double sy, D3 &org, D3 &dir) {
    double px = sx * c.width, py = sy * c.height;
    double fx = floor(px), fy = floor(py);
    double ox, oy;
    filter_sample(c.filter_type, c.filter_param, px - fx, py - fy, ox, oy);
    D3 rp{(fx + 0.5 + ox) / c.width, (fy + 0.5 + oy) / c.height, 0};
    D3 pt = xform_point(c.sample_to_cam, rp);
    D3 d = normalize(pt);
    org = xform_point(c.cam_to_world, D3{0, 0, 0});
    dir = normalize(xform_vector(c.cam_to_world, d));
}

}  // namespace ljx
