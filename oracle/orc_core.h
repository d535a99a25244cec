// ORACLE -- test infrastructure only.  Nothing in the product links, loads or
// calls this code; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it, as the checker.
//
// FP64 restatement of lajolla's math substrate, RNG, sampling tables,
// camera, filters, BSDFs and phase functions.  Each function cites the
// reference file:line it follows (paths relative to the reference's src/).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <optional>
#include <utility>
#include <vector>

namespace orc {

using Real = double;  // lajolla.h:23
constexpr Real kPi = 3.14159265358979323846;  // lajolla.h:26-33
constexpr Real kInvPi = 1.0 / kPi;
constexpr Real kTwoPi = 2.0 * kPi;
constexpr Real kInvTwoPi = 1.0 / kTwoPi;
constexpr Real kInvFourPi = 1.0 / (4.0 * kPi);
inline Real inf() { return std::numeric_limits<Real>::infinity(); }

// Planted mistakes, for tests that show their checks can fail: 0 (the
// default) is the restatement as written; k > 0 swaps in one wrong formula
// (the list is at orc_set_plant in orc_render.cpp).  Never set outside tests.
inline int g_plant = 0;
inline bool planted(int k) { return g_plant == k; }

// ---------------------------------------------------------------- vectors
struct V3 {
    Real x = 0, y = 0, z = 0;
    V3() = default;
    V3(Real a, Real b, Real c) : x(a), y(b), z(c) {}
    Real &operator[](int i) { return (&x)[i]; }
    Real operator[](int i) const { return (&x)[i]; }
};
struct V2 {
    Real x = 0, y = 0;
    V2() = default;
    V2(Real a, Real b) : x(a), y(b) {}
    Real operator[](int i) const { return i == 0 ? x : y; }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator*(V3 a, Real s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(Real s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, Real s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 operator/(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline V3 operator+(V3 a, Real s) { return {a.x + s, a.y + s, a.z + s}; }
inline V3 operator+(Real s, V3 a) { return {a.x + s, a.y + s, a.z + s}; }
inline V3 operator-(Real s, V3 a) { return {s - a.x, s - a.y, s - a.z}; }
inline V3 &operator+=(V3 &a, V3 b) { a = a + b; return a; }
inline V3 &operator*=(V3 &a, V3 b) { a = a * b; return a; }
inline V3 &operator*=(V3 &a, Real s) { a = a * s; return a; }
inline V3 &operator/=(V3 &a, Real s) { a = a / s; return a; }
inline Real dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline Real length(V3 a) { return std::sqrt(dot(a, a)); }
inline Real distance_squared(V3 a, V3 b) { return dot(a - b, a - b); }
inline Real distance(V3 a, V3 b) { return std::sqrt(distance_squared(a, b)); }
// vector.h:249-257: zero-length vectors normalise to zero.
inline V3 normalize(V3 a) {
    Real l = length(a);
    if (l <= 0) return {0, 0, 0};
    return a / l;
}
inline Real max3(V3 a) { return std::max(std::max(a.x, a.y), a.z); }  // vector.h:265
inline V3 vmax(V3 a, V3 b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }
inline Real avg(V3 a) { return (a.x + a.y + a.z) / 3; }               // spectrum.h:36
inline V3 vexp(V3 a) { return {std::exp(a.x), std::exp(a.y), std::exp(a.z)}; }
inline V3 vsqrt0(V3 a) {  // spectrum.h:21-25
    return {std::sqrt(std::max(a.x, 0.0)), std::sqrt(std::max(a.y, 0.0)),
            std::sqrt(std::max(a.z, 0.0))};
}
inline bool isfinite3(V3 a) {
    return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z);
}
// spectrum.h:32-34
inline Real luminance(V3 s) { return s.x * 0.212671 + s.y * 0.715160 + s.z * 0.072169; }

// ------------------------------------------------------------------ frame
// frame.h:11-21 (Frisvad, with the n.z < -1+1e-6 special case)
inline std::pair<V3, V3> coordinate_system(V3 n) {
    if (n.z < Real(-1 + 1e-6)) return {V3{0, -1, 0}, V3{-1, 0, 0}};
    Real a = 1 / (1 + n.z);
    Real b = -n.x * n.y * a;
    return {V3{1 - n.x * n.x * a, b, -n.x}, V3{b, 1 - n.y * n.y * a, -n.y}};
}
struct Frame {  // frame.h:25-42
    V3 x, y, n;
    Frame() = default;
    Frame(V3 a, V3 b, V3 c) : x(a), y(b), n(c) {}
    explicit Frame(V3 nn) : n(nn) { std::tie(x, y) = coordinate_system(nn); }
};
inline Frame flip(const Frame &f) { return Frame(-f.x, -f.y, -f.n); }  // frame.h:45-47
inline V3 to_local(const Frame &f, V3 v) { return {dot(v, f.x), dot(v, f.y), dot(v, f.n)}; }
inline V3 to_world(const Frame &f, V3 v) { return f.x * v.x + f.y * v.y + f.n * v.z; }

// -------------------------------------------------------------------- PCG
// pcg.h:16-68 (XSH-RR, multiplier 6364136223846793005)
struct Pcg {
    uint64_t state = 0, inc = 0;
};
inline uint32_t pcg_next(Pcg &r) {
    uint64_t old = r.state;
    r.state = old * 6364136223846793005ULL + (r.inc | 1);
    uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = uint32_t(old >> 59u);
    return uint32_t((xs >> rot) | (xs << ((-rot) & 31)));
}
constexpr uint64_t kDefaultSeed = 0x31e241f862a1fb5eULL;  // pcg.h:33
inline Pcg pcg_init(uint64_t stream, uint64_t seed = kDefaultSeed) {
    Pcg s;
    s.state = 0;
    s.inc = (stream << 1u) | 1u;
    pcg_next(s);
    s.state += seed;
    pcg_next(s);
    return s;
}
// Plan 1's stream of a camera path: a bijection on 63-bit values of the path
// id (sample * W * H + pixel), two multiply-xorshift rounds with the
// SplitMix64 finaliser's constants (Steele, Lea and Flood 2014), masked to 63
// bits so that distinct paths keep distinct PCG increments.  Consecutive ids
// as increments give correlated PCG streams (neighbouring paths' draws are
// related), which changes the error of an image though not its mean.
inline uint64_t path_stream(uint64_t id) {
    const uint64_t m = (uint64_t(1) << 63) - 1;
    uint64_t z = id & m;
    z = ((z ^ (z >> 31)) * 0xbf58476d1ce4e5b9ULL) & m;
    z = ((z ^ (z >> 27)) * 0x94d049bb133111ebULL) & m;
    return z ^ (z >> 31);
}
inline double pcg_double(Pcg &r) {  // pcg.h:60-68
    union {
        uint64_t u;
        double d;
    } x;
    x.u = (uint64_t(pcg_next(r)) << 20) | 0x3ff0000000000000ULL;
    return x.d - 1.0;
}
inline float pcg_float(Pcg &r) {  // pcg.h:49-57
    union {
        uint32_t u;
        float f;
    } x;
    x.u = (pcg_next(r) >> 9) | 0x3f800000u;
    return x.f - 1.0f;
}

// ------------------------------------------------------------- TableDist
// table_dist.cpp:3-38
struct TableDist1D {
    std::vector<Real> pmf, cdf;
};
inline TableDist1D make_table_dist_1d(const std::vector<Real> &f) {
    TableDist1D t;
    t.pmf = f;
    t.cdf.assign(f.size() + 1, 0);
    for (size_t i = 0; i < f.size(); i++) t.cdf[i + 1] = t.cdf[i] + t.pmf[i];
    Real total = t.cdf.back();
    if (total > 0) {
        for (size_t i = 0; i < t.pmf.size(); i++) {
            t.pmf[i] /= total;
            t.cdf[i] /= total;
        }
    } else {
        for (size_t i = 0; i < t.pmf.size(); i++) {
            t.pmf[i] = Real(1) / Real(t.pmf.size());
            t.cdf[i] = Real(i) / Real(t.pmf.size());
        }
        t.cdf.back() = 1;
    }
    return t;
}
inline int sample(const TableDist1D &t, Real u) {
    int size = int(t.pmf.size());
    const Real *p = std::upper_bound(t.cdf.data(), t.cdf.data() + size + 1, u);
    return std::clamp(int(p - t.cdf.data() - 1), 0, size - 1);
}
inline Real pmf(const TableDist1D &t, int id) { return t.pmf[id]; }

// table_dist.cpp:40-151
struct TableDist2D {
    std::vector<Real> cdf_rows, pdf_rows, cdf_marginals, pdf_marginals;
    Real total_values = 0;
    int width = 0, height = 0;
};
inline TableDist2D make_table_dist_2d(const std::vector<Real> &f, int w, int h) {
    TableDist2D t;
    t.width = w;
    t.height = h;
    t.cdf_rows.assign(size_t(h) * (w + 1), 0);
    t.pdf_rows.assign(size_t(h) * w, 0);
    for (int y = 0; y < h; y++) {
        Real *c = &t.cdf_rows[size_t(y) * (w + 1)];
        c[0] = 0;
        for (int x = 0; x < w; x++) c[x + 1] = c[x] + f[size_t(y) * w + x];
        Real integral = c[w];
        if (integral > 0) {
            for (int x = 0; x < w; x++) c[x] /= integral;
            for (int x = 0; x < w; x++) t.pdf_rows[size_t(y) * w + x] = f[size_t(y) * w + x] / integral;
        } else {
            for (int x = 0; x < w; x++) {
                t.pdf_rows[size_t(y) * w + x] = Real(1) / Real(w);
                c[x] = Real(x) / Real(w);
            }
            c[w] = 1;
        }
    }
    t.cdf_marginals.assign(h + 1, 0);
    t.pdf_marginals.assign(h, 0);
    for (int y = 0; y < h; y++)
        t.cdf_marginals[y + 1] = t.cdf_marginals[y] + t.cdf_rows[size_t(y) * (w + 1) + w];
    t.total_values = t.cdf_marginals.back();
    if (t.total_values > 0) {
        for (int y = 0; y < h; y++) t.cdf_marginals[y] /= t.total_values;
        t.cdf_marginals[h] = 1;
        for (int y = 0; y < h; y++)
            t.pdf_marginals[y] = t.cdf_rows[size_t(y) * (w + 1) + w] / t.total_values;
    } else {
        for (int y = 0; y < h; y++) {
            t.pdf_marginals[y] = Real(1) / Real(h);
            t.cdf_marginals[y] = Real(y) / Real(h);
        }
        t.cdf_marginals[h] = 1;
    }
    for (int y = 0; y < h; y++) t.cdf_rows[size_t(y) * (w + 1) + w] = 1;
    return t;
}
inline V2 sample(const TableDist2D &t, V2 r) {
    int w = t.width, h = t.height;
    const Real *yp = std::upper_bound(t.cdf_marginals.data(), t.cdf_marginals.data() + h + 1, r.y);
    int yo = std::clamp(int(yp - t.cdf_marginals.data() - 1), 0, h - 1);
    Real dy = r.y - t.cdf_marginals[yo];
    if (t.cdf_marginals[yo + 1] - t.cdf_marginals[yo] > 0)
        dy /= (t.cdf_marginals[yo + 1] - t.cdf_marginals[yo]);
    const Real *c = &t.cdf_rows[size_t(yo) * (w + 1)];
    const Real *xp = std::upper_bound(c, c + w + 1, r.x);
    int xo = std::clamp(int(xp - c - 1), 0, w - 1);
    Real dx = r.x - c[xo];
    if (c[xo + 1] - c[xo] > 0) dx /= (c[xo + 1] - c[xo]);
    return {(xo + dx) / w, (yo + dy) / h};
}
inline Real pdf(const TableDist2D &t, V2 xy) {
    int w = t.width, h = t.height;
    int x = int(std::clamp(xy.x * w, Real(0), Real(w - 1)));
    int y = int(std::clamp(xy.y * h, Real(0), Real(h - 1)));
    return t.pdf_marginals[y] * t.pdf_rows[size_t(y) * w + x] * w * h;
}

// ------------------------------------------------------------ 4x4 matrix
struct M4 {
    Real m[4][4] = {{0}};
    Real &operator()(int i, int j) { return m[i][j]; }
    Real operator()(int i, int j) const { return m[i][j]; }
};
inline M4 mat_mul(const M4 &a, const M4 &b) {
    M4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            Real s = 0;
            for (int k = 0; k < 4; k++) s += a(i, k) * b(k, j);
            r(i, j) = s;
        }
    return r;
}
// matrix.h:79+ : inverse through the adjugate (cofactor) matrix.
M4 inverse(const M4 &m);
// transform.cpp:4-18, 72-100
inline M4 m_translate(V3 d) {
    M4 r;
    r(0, 0) = r(1, 1) = r(2, 2) = r(3, 3) = 1;
    r(0, 3) = d.x; r(1, 3) = d.y; r(2, 3) = d.z;
    return r;
}
inline M4 m_scale(V3 s) {
    M4 r;
    r(0, 0) = s.x; r(1, 1) = s.y; r(2, 2) = s.z; r(3, 3) = 1;
    return r;
}
inline M4 m_perspective(Real fov) {
    Real cot = Real(1) / std::tan((kPi / Real(180)) * (fov / 2.0));
    M4 r;
    r(0, 0) = cot; r(1, 1) = cot; r(2, 2) = 1; r(2, 3) = -1; r(3, 2) = 1;
    return r;
}
inline V3 xform_point(const M4 &x, V3 p) {
    Real t0 = x(0, 0) * p.x + x(0, 1) * p.y + x(0, 2) * p.z + x(0, 3);
    Real t1 = x(1, 0) * p.x + x(1, 1) * p.y + x(1, 2) * p.z + x(1, 3);
    Real t2 = x(2, 0) * p.x + x(2, 1) * p.y + x(2, 2) * p.z + x(2, 3);
    Real t3 = x(3, 0) * p.x + x(3, 1) * p.y + x(3, 2) * p.z + x(3, 3);
    Real iw = Real(1) / t3;
    return {t0 * iw, t1 * iw, t2 * iw};
}
inline V3 xform_vector(const M4 &x, V3 v) {
    return {x(0, 0) * v.x + x(0, 1) * v.y + x(0, 2) * v.z,
            x(1, 0) * v.x + x(1, 1) * v.y + x(1, 2) * v.z,
            x(2, 0) * v.x + x(2, 1) * v.y + x(2, 2) * v.z};
}

// ------------------------------------------------------------ camera/filter
struct Filter {
    int type = 0;       // 0 box, 1 tent, 2 gaussian (filter.h:38)
    Real param = 1;     // width or stddev
};
// filters/box.inl:1-4, tent.inl:1-46, gaussian.inl:1-8
inline V2 filter_sample(const Filter &f, V2 r) {
    if (f.type == 0) {
        return {(2 * r.x - 1) * (f.param / 2), (2 * r.y - 1) * (f.param / 2)};
    } else if (f.type == 1) {
        Real h = f.param / 2;
        Real x = r.x < 0.5 ? h * (std::sqrt(2 * r.x) - 1) : h * (1 - std::sqrt(1 - 2 * (r.x - Real(0.5))));
        Real y = r.y < 0.5 ? h * (std::sqrt(2 * r.y) - 1) : h * (1 - std::sqrt(1 - 2 * (r.y - Real(0.5))));
        return {x, y};
    } else {
        Real rr = f.param * std::sqrt(-2 * std::log(std::max(r.x, Real(1e-8))));
        return {rr * std::cos(2 * kPi * r.y), rr * std::sin(2 * kPi * r.y)};
    }
}
struct Ray {
    V3 org, dir;
    Real tnear = 0, tfar = inf();
};
struct Camera {  // camera.h:7-22, camera.cpp:7-21
    M4 sample_to_cam, cam_to_sample, cam_to_world, world_to_cam;
    int width = 0, height = 0;
    Filter filter;
    int medium_id = -1;
};
inline Camera make_camera(const M4 &cam_to_world, Real fov, int w, int h, Filter f, int medium) {
    Camera c;
    c.cam_to_world = cam_to_world;
    c.world_to_cam = inverse(cam_to_world);
    c.width = w;
    c.height = h;
    c.filter = f;
    c.medium_id = medium;
    Real aspect = Real(w) / Real(h);
    c.cam_to_sample = mat_mul(mat_mul(m_scale(V3(-0.5, -0.5 * aspect, 1.0)),
                                      m_translate(V3(-1.0, -1.0 / aspect, 0.0))),
                              m_perspective(fov));
    c.sample_to_cam = inverse(c.cam_to_sample);
    return c;
}
// camera.cpp:23-47
inline Ray sample_primary(const Camera &c, V2 screen) {
    V2 pp{screen.x * c.width, screen.y * c.height};
    Real dx = pp.x - std::floor(pp.x), dy = pp.y - std::floor(pp.y);
    V2 off = filter_sample(c.filter, V2{dx, dy});
    V2 rp{(std::floor(pp.x) + Real(0.5) + off.x) / c.width,
          (std::floor(pp.y) + Real(0.5) + off.y) / c.height};
    V3 pt = xform_point(c.sample_to_cam, V3(rp.x, rp.y, 0));
    V3 dir = normalize(pt);
    return Ray{xform_point(c.cam_to_world, V3{0, 0, 0}), normalize(xform_vector(c.cam_to_world, dir)),
               0, inf()};
}

// ------------------------------------------------------------ microfacet
// microfacet.h:23-114
inline Real fresnel_dielectric(Real n_dot_i, Real n_dot_t, Real eta) {
    Real rs = (n_dot_i - eta * n_dot_t) / (n_dot_i + eta * n_dot_t);
    Real rp = (eta * n_dot_i - n_dot_t) / (eta * n_dot_i + n_dot_t);
    return (rs * rs + rp * rp) / 2;
}
inline Real fresnel_dielectric(Real n_dot_i, Real eta) {
    Real n_dot_t_sq = 1 - (1 - n_dot_i * n_dot_i) / (eta * eta);
    if (n_dot_t_sq < 0) return 1;
    return fresnel_dielectric(std::fabs(n_dot_i), std::sqrt(n_dot_t_sq), eta);
}
inline Real GTR2(Real n_dot_h, Real roughness) {
    Real alpha = roughness * roughness;
    Real a2 = alpha * alpha;
    Real t = 1 + (a2 - 1) * n_dot_h * n_dot_h;
    return a2 / (kPi * t * t);
}
inline Real smith_masking_gtr2(V3 v, Real roughness) {
    Real alpha = roughness * roughness;
    Real a2 = alpha * alpha;
    V3 v2 = v * v;
    Real Lambda = (-1 + std::sqrt(1 + (v2.x * a2 + v2.y * a2) / v2.z)) / 2;
    return 1 / (1 + Lambda);
}
// Heitz 2018 VNDF with independent x/y roughness; the isotropic call is the
// reference's sample_visible_normals (microfacet.h:85-114).
inline V3 sample_visible_normals_aniso(V3 in, Real ax, Real ay, V2 r) {
    if (in.z < 0) return -sample_visible_normals_aniso(-in, ax, ay, r);
    V3 hemi = normalize(V3{ax * in.x, ay * in.y, in.z});
    Real rr = std::sqrt(r.x);
    Real phi = 2 * kPi * r.y;
    Real t1 = rr * std::cos(phi);
    Real t2 = rr * std::sin(phi);
    Real s = (1 + hemi.z) / 2;
    t2 = (1 - s) * std::sqrt(1 - t1 * t1) + s * t2;
    V3 disk{t1, t2, std::sqrt(std::max(Real(0), 1 - t1 * t1 - t2 * t2))};
    Frame hf(hemi);
    V3 hn = to_world(hf, disk);
    return normalize(V3{ax * hn.x, ay * hn.y, std::max(Real(0), hn.z)});
}
inline V3 sample_visible_normals(V3 in, Real alpha, V2 r) {
    return sample_visible_normals_aniso(in, alpha, alpha, r);
}
// material.cpp:4-11
inline V3 sample_cos_hemisphere(V2 r) {
    Real phi = kTwoPi * r.x;
    Real tmp = std::sqrt(std::clamp(1 - r.y, Real(0), Real(1)));
    return {std::cos(phi) * tmp, std::sin(phi) * tmp, std::sqrt(std::clamp(r.y, Real(0), Real(1)))};
}

// ------------------------------------------------------------- materials
struct Material {  // material.h:12-98
    int type = 0;
    V3 a, b;  // see ljx_material_desc
    Real roughness = 0, eta = 1;
    Real specular_transmission = 0, metallic = 0, subsurface = 0, specular = 0,
         specular_tint = 0, anisotropic = 0, sheen = 0, sheen_tint = 0, clearcoat = 0,
         clearcoat_gloss = 0;
    int tex[13] = {0};  // per field: 0 the constant above, k > 0 the scene's texture k - 1
    bool textured = false;
};
struct ShadePoint {  // the PathVertex fields a BSDF reads (intersection.h:15-35)
    V3 geometric_normal;
    Frame shading_frame;
};
struct BSDFSample {  // material.h:133-138
    V3 dir_out;
    Real eta = 0, roughness = 0;
};
V3 bsdf_eval(const Material &m, V3 dir_in, V3 dir_out, const ShadePoint &v, bool to_view = false);
Real bsdf_pdf(const Material &m, V3 dir_in, V3 dir_out, const ShadePoint &v, bool to_view = false);
std::optional<BSDFSample> bsdf_sample(const Material &m, V3 dir_in, const ShadePoint &v, V2 rnd_uv,
                                      Real rnd_w, bool to_view = false);

// ------------------------------------------------------- phase functions
struct Phase {
    int type = 0;  // 0 isotropic, 1 HG (phase_function.h:16)
    Real g = 0;
};
// phase_functions/isotropic.inl:1-15, henyeygreenstein.inl:1-48
inline V3 phase_eval(const Phase &p, V3 in, V3 out) {
    if (p.type == 0) return V3{kInvFourPi, kInvFourPi, kInvFourPi};
    Real v = kInvFourPi * (1 - p.g * p.g) /
             std::pow((1 + p.g * p.g + 2 * p.g * dot(in, out)), Real(3) / Real(2));
    return V3{v, v, v};
}
inline Real phase_pdf(const Phase &p, V3 in, V3 out) {
    if (p.type == 0) return kInvFourPi;
    return kInvFourPi * (1 - p.g * p.g) /
           std::pow((1 + p.g * p.g + 2 * p.g * dot(in, out)), Real(3) / Real(2));
}
inline std::optional<V3> phase_sample(const Phase &p, V3 in, V2 r) {
    if (p.type == 0 || std::fabs(p.g) < Real(1e-3)) {
        Real z = 1 - 2 * r.x;
        Real rr = std::sqrt(std::fmax(Real(0), 1 - z * z));
        Real phi = 2 * kPi * r.y;
        return V3{rr * std::cos(phi), rr * std::sin(phi), z};
    }
    Real tmp = (p.g * p.g - 1) / (2 * r.x * p.g - (p.g + 1));
    Real ce = (tmp * tmp - (1 + p.g * p.g)) / (2 * p.g);
    Real se = std::sqrt(std::max(1 - ce * ce, Real(0)));
    Real az = 2 * kPi * r.y;
    Frame f(in);
    return to_world(f, V3{se * std::cos(az), se * std::sin(az), ce});
}

}  // namespace orc
