// ORACLE -- test infrastructure only (see orc_core.h).
//
// FP64 restatement of lajolla's BSDFs.  Lambertian, RoughPlastic and
// RoughDielectric follow src/materials/{lambertian,roughplastic,
// roughdielectric}.inl.  The six Disney materials are stubs in the reference
// (src/materials/disney_*.inl return zero / nullopt); they follow the handout
// spec handouts/homework1.tex:93-548, and are pinned only by closed forms,
// sample/pdf Jacobian consistency and energy bounds (no reference values
// exist).
#include "orc_core.h"

namespace orc {

namespace {

constexpr int kLambertian = 0, kRoughPlastic = 1, kRoughDielectric = 2, kDisneyDiffuse = 3,
              kDisneyMetal = 4, kDisneyGlass = 5, kDisneyClearcoat = 6, kDisneySheen = 7,
              kDisneyBSDF = 8;

inline Real pow5(Real x) { Real x2 = x * x; return x2 * x2 * x; }
inline Real clamp_rough(Real r) { return std::clamp(r, Real(0.01), Real(1)); }

// Frame flip used by the one-sided BRDFs (lambertian.inl:10-13).
inline Frame front_frame(const ShadePoint &v, V3 dir_in) {
    Frame f = v.shading_frame;
    if (dot(f.n, dir_in) < 0) f = flip(f);
    return f;
}
// Frame flip used by the transmissive BSDFs (roughdielectric.inl:7-10).
inline Frame dielectric_frame(const ShadePoint &v, V3 dir_in) {
    Frame f = v.shading_frame;
    if (dot(f.n, dir_in) * dot(v.geometric_normal, dir_in) < 0) f = flip(f);
    return f;
}

// ---------------------------------------------------------------- Lambertian
V3 lambertian_eval(const Material &m, V3 in, V3 out, const ShadePoint &v) {
    if (dot(v.geometric_normal, in) < 0 || dot(v.geometric_normal, out) < 0) return {};
    Frame f = front_frame(v, in);
    return std::fmax(dot(f.n, out), Real(0)) * m.a / kPi;
}
Real lambertian_pdf(V3 in, V3 out, const ShadePoint &v) {
    if (dot(v.geometric_normal, in) < 0 || dot(v.geometric_normal, out) < 0) return 0;
    Frame f = front_frame(v, in);
    return std::fmax(dot(f.n, out), Real(0)) / kPi;
}
std::optional<BSDFSample> cos_sample(V3 in, const ShadePoint &v, V2 r, Real roughness) {
    if (dot(v.geometric_normal, in) < 0) return std::nullopt;
    Frame f = front_frame(v, in);
    return BSDFSample{to_world(f, sample_cos_hemisphere(r)), 0, roughness};
}

// ------------------------------------------------------------- RoughPlastic
V3 plastic_eval(const Material &m, V3 in, V3 out, const ShadePoint &v) {
    if (dot(v.geometric_normal, in) < 0 || dot(v.geometric_normal, out) < 0) return {};
    Frame f = front_frame(v, in);
    V3 h = normalize(in + out);
    Real n_dot_h = dot(f.n, h), n_dot_in = dot(f.n, in), n_dot_out = dot(f.n, out);
    if (n_dot_out <= 0 || n_dot_h <= 0) return {};
    V3 Kd = m.a, Ks = m.b;
    Real rough = clamp_rough(m.roughness);
    Real F_o = fresnel_dielectric(dot(h, out), m.eta);
    Real D = GTR2(n_dot_h, rough);
    Real G = smith_masking_gtr2(to_local(f, in), rough) * smith_masking_gtr2(to_local(f, out), rough);
    V3 spec = Ks * (G * F_o * D) / (4 * n_dot_in * n_dot_out);
    Real F_i = fresnel_dielectric(dot(h, in), m.eta);
    V3 diff = Kd * (Real(1) - F_o) * (Real(1) - F_i) / kPi;
    return (spec + diff) * n_dot_out;
}
Real plastic_pdf(const Material &m, V3 in, V3 out, const ShadePoint &v) {
    if (dot(v.geometric_normal, in) < 0 || dot(v.geometric_normal, out) < 0) return 0;
    Frame f = front_frame(v, in);
    V3 h = normalize(in + out);
    Real n_dot_in = dot(f.n, in), n_dot_out = dot(f.n, out), n_dot_h = dot(f.n, h);
    if (n_dot_out <= 0 || n_dot_h <= 0) return 0;
    Real lS = luminance(m.b), lR = luminance(m.a);
    if (lS + lR <= 0) return 0;
    Real rough = clamp_rough(m.roughness);
    Real spec_prob = lS / (lS + lR);
    Real diff_prob = 1 - spec_prob;
    Real G = smith_masking_gtr2(to_local(f, in), rough);
    Real D = GTR2(n_dot_h, rough);
    spec_prob *= (G * D) / (4 * n_dot_in);
    diff_prob *= n_dot_out / kPi;
    return spec_prob + diff_prob;
}
std::optional<BSDFSample> plastic_sample(const Material &m, V3 in, const ShadePoint &v, V2 r, Real w) {
    if (dot(v.geometric_normal, in) < 0) return std::nullopt;
    Frame f = front_frame(v, in);
    Real lS = luminance(m.b), lR = luminance(m.a);
    if (lS + lR <= 0) return std::nullopt;
    Real spec_prob = lS / (lS + lR);
    if (w < spec_prob) {
        V3 lin = to_local(f, in);
        Real rough = clamp_rough(m.roughness);
        Real alpha = rough * rough;
        V3 h = to_world(f, sample_visible_normals(lin, alpha, r));
        V3 refl = normalize(-in + 2 * dot(in, h) * h);
        return BSDFSample{refl, 0, rough};
    }
    return BSDFSample{to_world(f, sample_cos_hemisphere(r)), 0, 1};
}

// ---------------------------------------------------------- microfacet glass
// Shared by RoughDielectric (isotropic GTR2, roughdielectric.inl) and
// DisneyGlass (anisotropic GGX, homework1.tex:336-404).
struct GlassLobe {
    V3 Kr, Kt;
    Real eta_param;
    bool aniso;
    Real rough;      // clamped roughness (isotropic) or Disney roughness
    Real ax, ay;     // anisotropic alphas
};
inline Real D_aniso(V3 hl, Real ax, Real ay) {
    Real t = hl.x * hl.x / (ax * ax) + hl.y * hl.y / (ay * ay) + hl.z * hl.z;
    return 1 / (kPi * ax * ay * t * t);
}
inline Real G1_aniso(V3 wl, Real ax, Real ay) {
    Real a = (wl.x * ax) * (wl.x * ax) + (wl.y * ay) * (wl.y * ay);
    Real Lambda = (std::sqrt(1 + a / (wl.z * wl.z)) - 1) / 2;
    return 1 / (1 + Lambda);
}
inline Real lobe_D(const GlassLobe &g, const Frame &f, V3 h) {
    if (!g.aniso) return GTR2(dot(f.n, h), g.rough);
    return D_aniso(to_local(f, h), g.ax, g.ay);
}
inline Real lobe_G1(const GlassLobe &g, const Frame &f, V3 w) {
    if (!g.aniso) return smith_masking_gtr2(to_local(f, w), g.rough);
    return G1_aniso(to_local(f, w), g.ax, g.ay);
}
V3 glass_eval(const GlassLobe &g, V3 in, V3 out, const ShadePoint &v, bool to_view) {
    bool reflect = dot(v.geometric_normal, in) * dot(v.geometric_normal, out) > 0;
    Frame f = dielectric_frame(v, in);
    Real eta = dot(v.geometric_normal, in) > 0 ? g.eta_param : 1 / g.eta_param;
    V3 h = reflect ? normalize(in + out) : normalize(in + out * eta);
    if (dot(h, f.n) < 0) h = -h;
    Real h_dot_in = dot(h, in);
    Real F = fresnel_dielectric(h_dot_in, eta);
    Real D = lobe_D(g, f, h);
    Real G = lobe_G1(g, f, in) * lobe_G1(g, f, out);
    if (reflect) return g.Kr * (F * D * G) / (4 * std::fabs(dot(f.n, in)));
    Real eta_factor = to_view ? 1 : (1 / (eta * eta));
    Real h_dot_out = dot(h, out);
    Real sd = h_dot_in + eta * h_dot_out;
    return g.Kt * (eta_factor * (1 - F) * D * G * eta * eta * std::fabs(h_dot_out * h_dot_in)) /
           (std::fabs(dot(f.n, in)) * sd * sd);
}
Real glass_pdf(const GlassLobe &g, V3 in, V3 out, const ShadePoint &v) {
    bool reflect = dot(v.geometric_normal, in) * dot(v.geometric_normal, out) > 0;
    Frame f = dielectric_frame(v, in);
    Real eta = dot(v.geometric_normal, in) > 0 ? g.eta_param : 1 / g.eta_param;
    V3 h = reflect ? normalize(in + out) : normalize(in + out * eta);
    if (dot(h, f.n) < 0) h = -h;
    Real h_dot_in = dot(h, in);
    Real F = fresnel_dielectric(h_dot_in, eta);
    Real D = lobe_D(g, f, h);
    Real G_in = lobe_G1(g, f, in);
    if (reflect) return (F * D * G_in) / (4 * std::fabs(dot(f.n, in)));
    Real h_dot_out = dot(h, out);
    Real sd = h_dot_in + eta * h_dot_out;
    Real dh_dout = eta * eta * h_dot_out / (sd * sd);
    return (1 - F) * D * G_in * std::fabs(dh_dout * h_dot_in / dot(f.n, in));
}
std::optional<BSDFSample> glass_sample(const GlassLobe &g, V3 in, const ShadePoint &v, V2 r, Real w) {
    Real eta = dot(v.geometric_normal, in) > 0 ? g.eta_param : 1 / g.eta_param;
    Frame f = dielectric_frame(v, in);
    V3 lin = to_local(f, in);
    V3 hl = g.aniso ? sample_visible_normals_aniso(lin, g.ax, g.ay, r)
                    : sample_visible_normals(lin, g.rough * g.rough, r);
    V3 h = to_world(f, hl);
    if (dot(h, f.n) < 0) h = -h;
    Real h_dot_in = dot(h, in);
    Real F = fresnel_dielectric(h_dot_in, eta);
    Real rough_out = g.rough;
    if (w <= F) {
        V3 refl = normalize(-in + 2 * dot(in, h) * h);
        return BSDFSample{refl, 0, rough_out};
    }
    Real h_dot_out_sq = 1 - (1 - h_dot_in * h_dot_in) / (eta * eta);
    if (h_dot_out_sq <= 0) return std::nullopt;
    if (h_dot_in < 0) h = -h;
    Real h_dot_out = std::sqrt(h_dot_out_sq);
    V3 refr = -in / eta + (std::fabs(h_dot_in) / eta - h_dot_out) * h;
    return BSDFSample{refr, eta, rough_out};
}
GlassLobe dielectric_lobe(const Material &m) {
    GlassLobe g;
    g.Kr = m.a;
    g.Kt = m.b;
    g.eta_param = m.eta;
    g.aniso = false;
    g.rough = clamp_rough(m.roughness);
    g.ax = g.ay = g.rough * g.rough;
    return g;
}

// -------------------------------------------------------------- Disney lobes
// homework1.tex:175-259: aspect, alpha_x, alpha_y with alpha_min = 1e-4.
inline void disney_alphas(Real roughness, Real anisotropic, Real &ax, Real &ay) {
    Real aspect = std::sqrt(1 - Real(0.9) * anisotropic);
    ax = std::max(Real(1e-4), roughness * roughness / aspect);
    ay = std::max(Real(1e-4), roughness * roughness * aspect);
}
GlassLobe disney_glass_lobe(V3 base, Real roughness, Real anisotropic, Real eta) {
    GlassLobe g;
    g.Kr = base;
    g.Kt = planted(7) ? base : vsqrt0(base);
    g.eta_param = eta;
    g.aniso = true;
    g.rough = roughness;
    disney_alphas(roughness, anisotropic, g.ax, g.ay);
    return g;
}
inline V3 tint_color(V3 base) {  // homework1.tex:411-415 C_tint
    Real l = luminance(base);
    return l > 0 ? base / l : V3{1, 1, 1};
}

// homework1.tex:93-174 -- f_diffuse with the subsurface blend.
V3 disney_diffuse_eval(V3 base, Real roughness, Real subsurface, V3 in, V3 out, const ShadePoint &v) {
    if (dot(v.geometric_normal, in) < 0 || dot(v.geometric_normal, out) < 0) return {};
    Frame f = front_frame(v, in);
    V3 h = normalize(in + out);
    Real n_in = std::fabs(dot(f.n, in)), n_out = std::fabs(dot(f.n, out));
    Real h_out = std::fabs(dot(h, out));
    Real FD90 = Real(0.5) + (planted(3) ? 1 : 2) * roughness * h_out * h_out;
    Real FD_in = 1 + (FD90 - 1) * pow5(1 - n_in);
    Real FD_out = 1 + (FD90 - 1) * pow5(1 - n_out);
    V3 base_diffuse = base / kPi * (FD_in * FD_out * n_out);
    Real FSS90 = roughness * h_out * h_out;
    Real FSS_in = 1 + (FSS90 - 1) * pow5(1 - n_in);
    Real FSS_out = 1 + (FSS90 - 1) * pow5(1 - n_out);
    V3 ss = (planted(6) ? Real(1) : Real(1.25)) * base / kPi *
            (FSS_in * FSS_out * (1 / (n_in + n_out) - Real(0.5)) + Real(0.5)) * n_out;
    return (1 - subsurface) * base_diffuse + subsurface * ss;
}
// homework1.tex:175-259 -- metal; Fm is given as a colour F0 for the Schlick blend.
V3 disney_metal_eval(V3 F0, Real ax, Real ay, V3 in, V3 out, const ShadePoint &v) {
    if (dot(v.geometric_normal, in) < 0 || dot(v.geometric_normal, out) < 0) return {};
    Frame f = front_frame(v, in);
    V3 h = normalize(in + out);
    Real n_dot_in = dot(f.n, in), n_dot_out = dot(f.n, out), n_dot_h = dot(f.n, h);
    if (n_dot_out <= 0 || n_dot_h <= 0) return {};
    V3 Fm = F0 + (1 - F0) * pow5(1 - std::fabs(dot(h, out)));
    Real D = D_aniso(to_local(f, h), ax, ay);
    Real G = G1_aniso(to_local(f, in), ax, ay) * G1_aniso(to_local(f, out), ax, ay);
    return Fm * (D * G) / (4 * std::fabs(n_dot_in));
}
Real disney_metal_pdf(Real ax, Real ay, V3 in, V3 out, const ShadePoint &v) {
    if (dot(v.geometric_normal, in) < 0 || dot(v.geometric_normal, out) < 0) return 0;
    Frame f = front_frame(v, in);
    V3 h = normalize(in + out);
    Real n_dot_in = dot(f.n, in), n_dot_out = dot(f.n, out), n_dot_h = dot(f.n, h);
    if (n_dot_out <= 0 || n_dot_h <= 0) return 0;
    Real D = D_aniso(to_local(f, h), ax, ay);
    Real G_in = G1_aniso(to_local(f, in), ax, ay);
    return D * G_in / (4 * std::fabs(n_dot_in));
}
std::optional<BSDFSample> disney_metal_sample(Real ax, Real ay, Real roughness, V3 in,
                                              const ShadePoint &v, V2 r) {
    if (dot(v.geometric_normal, in) < 0) return std::nullopt;
    Frame f = front_frame(v, in);
    V3 h = to_world(f, sample_visible_normals_aniso(to_local(f, in), ax, ay, r));
    V3 refl = normalize(-in + 2 * dot(in, h) * h);
    return BSDFSample{refl, 0, roughness};
}
// homework1.tex:260-335 -- clearcoat.
inline Real clearcoat_alpha(Real gloss) { return (1 - gloss) * Real(0.1) + gloss * Real(0.001); }
inline Real clearcoat_D(Real hz, Real ag) {
    Real a2 = ag * ag;
    return (a2 - 1) / (kPi * std::log(a2) * (1 + (a2 - 1) * hz * hz));
}
V3 disney_clearcoat_eval(Real gloss, V3 in, V3 out, const ShadePoint &v) {
    if (dot(v.geometric_normal, in) < 0 || dot(v.geometric_normal, out) < 0) return {};
    Frame f = front_frame(v, in);
    V3 h = normalize(in + out);
    Real n_dot_in = dot(f.n, in), n_dot_out = dot(f.n, out), n_dot_h = dot(f.n, h);
    if (n_dot_out <= 0 || n_dot_h <= 0) return {};
    Real R0 = (Real(1.5) - 1) * (Real(1.5) - 1) / ((Real(1.5) + 1) * (Real(1.5) + 1));
    Real Fc = R0 + (1 - R0) * pow5(1 - std::fabs(dot(h, out)));
    Real Dc = clearcoat_D(to_local(f, h).z, clearcoat_alpha(gloss));
    Real Gc = G1_aniso(to_local(f, in), Real(0.25), Real(0.25)) *
              G1_aniso(to_local(f, out), Real(0.25), Real(0.25));
    Real val = Fc * Dc * Gc / (4 * std::fabs(n_dot_in));
    return V3{val, val, val};
}
Real disney_clearcoat_pdf(Real gloss, V3 in, V3 out, const ShadePoint &v) {
    if (dot(v.geometric_normal, in) < 0 || dot(v.geometric_normal, out) < 0) return 0;
    Frame f = front_frame(v, in);
    V3 h = normalize(in + out);
    Real n_dot_out = dot(f.n, out), n_dot_h = dot(f.n, h);
    if (n_dot_out <= 0 || n_dot_h <= 0) return 0;
    Real Dc = clearcoat_D(to_local(f, h).z, clearcoat_alpha(gloss));
    return Dc * std::fabs(n_dot_h) / (4 * std::fabs(dot(h, out)));
}
std::optional<BSDFSample> disney_clearcoat_sample(Real gloss, V3 in, const ShadePoint &v, V2 r) {
    if (dot(v.geometric_normal, in) < 0) return std::nullopt;
    Frame f = front_frame(v, in);
    Real ag = clearcoat_alpha(gloss);
    Real a2 = ag * ag;
    Real cos_h = std::sqrt(std::max(Real(0), (1 - std::pow(a2, 1 - r.x)) / (1 - a2)));
    Real sin_h = std::sqrt(std::max(Real(0), 1 - cos_h * cos_h));
    Real phi = 2 * kPi * r.y;
    V3 hl{sin_h * std::cos(phi), sin_h * std::sin(phi), cos_h};
    V3 h = to_world(f, hl);
    V3 refl = normalize(-in + 2 * dot(in, h) * h);
    return BSDFSample{refl, 0, std::sqrt(ag)};
}
// homework1.tex:405-459 -- sheen.
V3 disney_sheen_eval(V3 base, Real sheen_tint, V3 in, V3 out, const ShadePoint &v) {
    if (dot(v.geometric_normal, in) < 0 || dot(v.geometric_normal, out) < 0) return {};
    Frame f = front_frame(v, in);
    V3 h = normalize(in + out);
    V3 C = (1 - sheen_tint) + sheen_tint * tint_color(base);
    Real s5 = pow5(1 - std::fabs(dot(h, out)));
    if (planted(1)) s5 *= 1 - std::fabs(dot(h, out));
    return C * (s5 * std::fabs(dot(f.n, out)));
}

// homework1.tex:460-548 -- the combined BSDF and its lobe weights.
struct DisneyWeights {
    Real diffuse, sheen, metal, clearcoat, glass;
};
DisneyWeights disney_weights(const Material &m) {
    DisneyWeights w;
    w.diffuse = (1 - m.specular_transmission) * (1 - m.metallic);
    w.sheen = (planted(2) ? 1 : (1 - m.metallic)) * m.sheen;
    w.metal = (1 - m.specular_transmission * (1 - m.metallic));
    w.clearcoat = (planted(4) ? Real(0.5) : Real(0.25)) * m.clearcoat;
    w.glass = (1 - m.metallic) * m.specular_transmission;
    return w;
}
V3 disney_metal_F0(const Material &m) {
    Real R0 = (m.eta - 1) * (m.eta - 1) / ((m.eta + 1) * (m.eta + 1));
    V3 Ks = (1 - m.specular_tint) + m.specular_tint * tint_color(m.a);
    return m.specular * R0 * (planted(9) ? 1 : (1 - m.metallic)) * Ks + m.metallic * m.a;
}
V3 disney_bsdf_eval(const Material &m, V3 in, V3 out, const ShadePoint &v, bool to_view) {
    Real ax, ay;
    disney_alphas(m.roughness, m.anisotropic, ax, ay);
    GlassLobe g = disney_glass_lobe(m.a, m.roughness, m.anisotropic, m.eta);
    DisneyWeights w = disney_weights(m);
    if (dot(v.geometric_normal, in) <= 0) return w.glass * glass_eval(g, in, out, v, to_view);
    V3 r = w.diffuse * disney_diffuse_eval(m.a, m.roughness, m.subsurface, in, out, v);
    r += w.sheen * disney_sheen_eval(m.a, m.sheen_tint, in, out, v);
    r += w.metal * disney_metal_eval(disney_metal_F0(m), ax, ay, in, out, v);
    r += w.clearcoat * disney_clearcoat_eval(m.clearcoat_gloss, in, out, v);
    r += w.glass * glass_eval(g, in, out, v, to_view);
    return r;
}
// Sampling probabilities: diffuse, metal, glass, clearcoat (sheen ignored).
void disney_probs(const Material &m, bool inside, Real p[4]) {
    DisneyWeights w = disney_weights(m);
    if (inside) {
        p[0] = p[1] = p[3] = 0;
        p[2] = 1;
        return;
    }
    if (planted(8)) w.clearcoat *= 2;
    Real tot = w.diffuse + w.metal + w.glass + w.clearcoat;
    p[0] = w.diffuse / tot;
    p[1] = w.metal / tot;
    p[2] = w.glass / tot;
    p[3] = w.clearcoat / tot;
}
Real disney_bsdf_pdf(const Material &m, V3 in, V3 out, const ShadePoint &v) {
    Real ax, ay;
    disney_alphas(m.roughness, m.anisotropic, ax, ay);
    GlassLobe g = disney_glass_lobe(m.a, m.roughness, m.anisotropic, m.eta);
    bool inside = dot(v.geometric_normal, in) <= 0;
    Real p[4];
    disney_probs(m, inside, p);
    if (inside) return glass_pdf(g, in, out, v);
    Real r = 0;
    if (p[0] > 0) r += p[0] * lambertian_pdf(in, out, v);
    if (p[1] > 0) r += p[1] * disney_metal_pdf(ax, ay, in, out, v);
    if (p[2] > 0) r += p[2] * glass_pdf(g, in, out, v);
    if (p[3] > 0) r += p[3] * disney_clearcoat_pdf(m.clearcoat_gloss, in, out, v);
    return r;
}
std::optional<BSDFSample> disney_bsdf_sample(const Material &m, V3 in, const ShadePoint &v, V2 r,
                                             Real w) {
    Real ax, ay;
    disney_alphas(m.roughness, m.anisotropic, ax, ay);
    GlassLobe g = disney_glass_lobe(m.a, m.roughness, m.anisotropic, m.eta);
    bool inside = dot(v.geometric_normal, in) <= 0;
    Real p[4];
    disney_probs(m, inside, p);
    if (inside) return glass_sample(g, in, v, r, w);
    Real c0 = p[0], c1 = c0 + p[1], c2 = c1 + p[2];
    if (w < c0) return cos_sample(in, v, r, m.roughness);
    if (w < c1) return disney_metal_sample(ax, ay, m.roughness, in, v, r);
    if (w < c2) return glass_sample(g, in, v, r, (w - c1) / p[2]);
    return disney_clearcoat_sample(m.clearcoat_gloss, in, v, r);
}

}  // namespace

V3 bsdf_eval(const Material &m, V3 in, V3 out, const ShadePoint &v, bool to_view) {
    switch (m.type) {
    case kLambertian: return lambertian_eval(m, in, out, v);
    case kRoughPlastic: return plastic_eval(m, in, out, v);
    case kRoughDielectric: return glass_eval(dielectric_lobe(m), in, out, v, to_view);
    case kDisneyDiffuse: return disney_diffuse_eval(m.a, m.roughness, m.subsurface, in, out, v);
    case kDisneyMetal: {
        Real ax, ay;
        disney_alphas(m.roughness, m.anisotropic, ax, ay);
        return disney_metal_eval(m.a, ax, ay, in, out, v);
    }
    case kDisneyGlass:
        return glass_eval(disney_glass_lobe(m.a, m.roughness, m.anisotropic, m.eta), in, out, v, to_view);
    case kDisneyClearcoat: return disney_clearcoat_eval(m.clearcoat_gloss, in, out, v);
    case kDisneySheen: return disney_sheen_eval(m.a, m.sheen_tint, in, out, v);
    case kDisneyBSDF: return disney_bsdf_eval(m, in, out, v, to_view);
    }
    return {};
}

Real bsdf_pdf(const Material &m, V3 in, V3 out, const ShadePoint &v, bool) {
    switch (m.type) {
    case kLambertian: return lambertian_pdf(in, out, v);
    case kRoughPlastic: return plastic_pdf(m, in, out, v);
    case kRoughDielectric: return glass_pdf(dielectric_lobe(m), in, out, v);
    case kDisneyDiffuse: return lambertian_pdf(in, out, v);
    case kDisneyMetal: {
        Real ax, ay;
        disney_alphas(m.roughness, m.anisotropic, ax, ay);
        return disney_metal_pdf(ax, ay, in, out, v);
    }
    case kDisneyGlass:
        return glass_pdf(disney_glass_lobe(m.a, m.roughness, m.anisotropic, m.eta), in, out, v);
    case kDisneyClearcoat: return disney_clearcoat_pdf(m.clearcoat_gloss, in, out, v);
    case kDisneySheen: return lambertian_pdf(in, out, v);
    case kDisneyBSDF: return disney_bsdf_pdf(m, in, out, v);
    }
    return 0;
}

std::optional<BSDFSample> bsdf_sample(const Material &m, V3 in, const ShadePoint &v, V2 r, Real w,
                                      bool) {
    switch (m.type) {
    case kLambertian: return cos_sample(in, v, r, 1);
    case kRoughPlastic: return plastic_sample(m, in, v, r, w);
    case kRoughDielectric: return glass_sample(dielectric_lobe(m), in, v, r, w);
    case kDisneyDiffuse: return cos_sample(in, v, r, m.roughness);
    case kDisneyMetal: {
        Real ax, ay;
        disney_alphas(m.roughness, m.anisotropic, ax, ay);
        return disney_metal_sample(ax, ay, m.roughness, in, v, r);
    }
    case kDisneyGlass:
        return glass_sample(disney_glass_lobe(m.a, m.roughness, m.anisotropic, m.eta), in, v, r, w);
    case kDisneyClearcoat: return disney_clearcoat_sample(m.clearcoat_gloss, in, v, r);
    case kDisneySheen: return cos_sample(in, v, r, 1);
    case kDisneyBSDF: return disney_bsdf_sample(m, in, v, r, w);
    }
    return std::nullopt;
}

// matrix.h:79-220 restated: inverse = adjugate / determinant.
M4 inverse(const M4 &m) {
    auto minor3 = [&](int r, int c) {
        Real a[3][3];
        int ii = 0;
        for (int i = 0; i < 4; i++) {
            if (i == r) continue;
            int jj = 0;
            for (int j = 0; j < 4; j++) {
                if (j == c) continue;
                a[ii][jj++] = m(i, j);
            }
            ii++;
        }
        return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
               a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
               a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    };
    M4 inv;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            Real sgn = ((i + j) % 2 == 0) ? 1 : -1;
            inv(j, i) = sgn * minor3(i, j);  // transpose of the cofactor matrix
        }
    Real det = m(0, 0) * inv(0, 0) + m(0, 1) * inv(1, 0) + m(0, 2) * inv(2, 0) + m(0, 3) * inv(3, 0);
    Real idet = Real(1) / det;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) inv(i, j) *= idet;
    return inv;
}

}  // namespace orc
