// ORACLE -- test infrastructure only.
//
// Known-answer-test driver for the reference renderer.  It is compiled in the
// authoring container together with the reference's own sources
// (camera.cpp, transform.cpp, filter.cpp, table_dist.cpp, material.cpp,
// phase_function.cpp; see oracle/Makefile, target `ref`) into oracle/_ref/,
// and prints seeded inputs with the reference's outputs as JSON.
// tests/golden/make_golden.py stores that JSON as fixtures.  Only parts of
// the path whose objects link without Embree are covered here.
#include "camera.h"
#include "filter.h"
#include "material.h"
#include "intersection.h"
#include "pcg.h"
#include "phase_function.h"
#include "table_dist.h"
#include "texture.h"
#include "transform.h"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

static void jd(double v) {
    if (std::isnan(v)) std::printf("\"nan\"");
    else if (std::isinf(v)) std::printf(v > 0 ? "\"inf\"" : "\"-inf\"");
    else std::printf("%.17g", v);
}
static void jv(const Vector3 &v) {
    std::printf("[");
    jd(v.x); std::printf(","); jd(v.y); std::printf(","); jd(v.z);
    std::printf("]");
}

// Seeded generator for KAT inputs (PCG from the reference, stream 12345).
static pcg32_state g_rng = init_pcg32(12345, 0x0123456789abcdefULL);
static double U() { return next_pcg32_real<double>(g_rng); }
static Vector3 unit_sphere() {
    double z = 1 - 2 * U(), r = sqrt(std::max(0.0, 1 - z * z)), phi = 2 * c_PI * U();
    return Vector3{r * cos(phi), r * sin(phi), z};
}

static void kat_pcg() {
    std::printf("{\"pcg\":[");
    uint64_t seeds[2] = {0x31e241f862a1fb5eULL, 42ULL};
    bool first = true;
    for (uint64_t stream = 0; stream < 6; stream++)
        for (int si = 0; si < 2; si++) {
            if (!first) std::printf(",");
            first = false;
            std::printf("{\"stream\":%llu,\"seed\":%llu,\"u32\":[", (unsigned long long)stream,
                        (unsigned long long)seeds[si]);
            pcg32_state r = init_pcg32(stream, seeds[si]);
            for (int i = 0; i < 48; i++) std::printf("%s%u", i ? "," : "", next_pcg32(r));
            std::printf("],\"f64\":[");
            r = init_pcg32(stream, seeds[si]);
            for (int i = 0; i < 48; i++) { if (i) std::printf(","); jd(next_pcg32_real<double>(r)); }
            std::printf("],\"f32\":[");
            r = init_pcg32(stream, seeds[si]);
            for (int i = 0; i < 48; i++) { if (i) std::printf(","); jd(next_pcg32_real<float>(r)); }
            std::printf("]}");
        }
    std::printf("]}\n");
}

static void kat_table() {
    std::printf("{\"table1d\":[");
    std::vector<std::vector<Real>> fs = {
        {1, 2, 3, 4}, {0, 0, 0}, {0.5}, {3, 0, 0, 1, 0, 7, 2}, {}};
    {
        std::vector<Real> big;
        for (int i = 0; i < 1025; i++) big.push_back(pow(10.0, -1 + 3 * U()));
        fs.back() = big;
    }
    for (size_t k = 0; k < fs.size(); k++) {
        TableDist1D t = make_table_dist_1d(fs[k]);
        std::printf("%s{\"f\":[", k ? "," : "");
        for (size_t i = 0; i < fs[k].size(); i++) { if (i) std::printf(","); jd(fs[k][i]); }
        std::printf("],\"u\":[");
        std::vector<double> us;
        for (int i = 0; i < 64; i++) us.push_back(U());
        us.push_back(0.0);
        us.push_back(1.0);
        for (size_t i = 0; i < us.size(); i++) { if (i) std::printf(","); jd(us[i]); }
        std::printf("],\"idx\":[");
        for (size_t i = 0; i < us.size(); i++) std::printf("%s%d", i ? "," : "", sample(t, us[i]));
        std::printf("],\"pmf\":[");
        for (size_t i = 0; i < t.pmf.size(); i++) { if (i) std::printf(","); jd(t.pmf[i]); }
        std::printf("],\"cdf\":[");
        for (size_t i = 0; i < t.cdf.size(); i++) { if (i) std::printf(","); jd(t.cdf[i]); }
        std::printf("]}");
    }
    std::printf("],\"table2d\":[");
    struct G { int w, h; std::vector<Real> f; };
    std::vector<G> gs;
    {
        G g{5, 4, {}};
        for (int i = 0; i < 20; i++) g.f.push_back(U() * 3);
        for (int i = 5; i < 10; i++) g.f[i] = 0;  // an all-zero row
        gs.push_back(g);
        gs.push_back(G{1, 1, {0.7}});
        gs.push_back(G{3, 2, {0, 0, 0, 0, 0, 0}});
    }
    for (size_t k = 0; k < gs.size(); k++) {
        TableDist2D t = make_table_dist_2d(gs[k].f, gs[k].w, gs[k].h);
        std::printf("%s{\"w\":%d,\"h\":%d,\"f\":[", k ? "," : "", gs[k].w, gs[k].h);
        for (size_t i = 0; i < gs[k].f.size(); i++) { if (i) std::printf(","); jd(gs[k].f[i]); }
        std::printf("],\"uv\":[");
        std::vector<double> uv;
        for (int i = 0; i < 128; i++) uv.push_back(U());
        for (size_t i = 0; i < uv.size(); i++) { if (i) std::printf(","); jd(uv[i]); }
        std::printf("],\"xy\":[");
        std::vector<double> pd;
        for (size_t i = 0; i < uv.size() / 2; i++) {
            Vector2 r = sample(t, Vector2{uv[2 * i], uv[2 * i + 1]});
            if (i) std::printf(",");
            jd(r.x); std::printf(","); jd(r.y);
            pd.push_back(pdf(t, r));
        }
        std::printf("],\"pdf\":[");
        for (size_t i = 0; i < pd.size(); i++) { if (i) std::printf(","); jd(pd[i]); }
        std::printf("],\"total\":");
        jd(t.total_values);
        std::printf("}");
    }
    std::printf("]}\n");
}

static void kat_camera() {
    std::printf("{\"camera\":[");
    struct C { Vector3 pos, look, up; double fov; int w, h; int ft; double fp; };
    std::vector<C> cs;
    cs.push_back(C{Vector3{0.0, 1.0, 6.8}, Vector3{0.0, 1.0, 0.0}, Vector3{0.0, 1.0, 0.0}, 40, 256, 256, 0, 1.0});
    cs.push_back(C{Vector3{3.0, 4.0, -5.0}, Vector3{0.5, 0.0, 0.2}, Vector3{0.0, 0.0, 1.0}, 65, 320, 180, 1, 2.0});
    cs.push_back(C{Vector3{-2.0, 0.5, 1.0}, Vector3{1.0, 0.2, -3.0}, Vector3{0.2, 1.0, 0.0}, 25, 100, 300, 2, 0.5});
    for (size_t k = 0; k < cs.size(); k++) {
        const C &c = cs[k];
        Matrix4x4 to_world = look_at(c.pos, c.look, c.up);
        Filter f = c.ft == 0 ? Filter{Box{c.fp}} : (c.ft == 1 ? Filter{Tent{c.fp}} : Filter{Gaussian{c.fp}});
        Camera cam(to_world, c.fov, c.w, c.h, f, -1);
        std::printf("%s{\"cam_to_world\":[", k ? "," : "");
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) { if (i || j) std::printf(","); jd(to_world(i, j)); }
        std::printf("],\"fov\":"); jd(c.fov);
        std::printf(",\"width\":%d,\"height\":%d,\"filter_type\":%d,\"filter_param\":", c.w, c.h, c.ft);
        jd(c.fp);
        std::printf(",\"screen\":[");
        std::vector<double> sc;
        for (int i = 0; i < 64; i++) sc.push_back(U());
        for (size_t i = 0; i < sc.size(); i++) { if (i) std::printf(","); jd(sc[i]); }
        std::printf("],\"rays\":[");
        for (size_t i = 0; i < sc.size() / 2; i++) {
            Ray r = sample_primary(cam, Vector2{sc[2 * i], sc[2 * i + 1]});
            if (i) std::printf(",");
            jv(r.org); std::printf(","); jv(r.dir);
        }
        std::printf("]}");
    }
    std::printf("],\"filters\":[");
    for (int ft = 0; ft < 3; ft++) {
        double fp = ft == 2 ? 0.7 : 1.5;
        Filter f = ft == 0 ? Filter{Box{fp}} : (ft == 1 ? Filter{Tent{fp}} : Filter{Gaussian{fp}});
        std::printf("%s{\"type\":%d,\"param\":", ft ? "," : "", ft);
        jd(fp);
        std::printf(",\"rnd\":[");
        std::vector<double> rn;
        for (int i = 0; i < 64; i++) rn.push_back(U());
        for (size_t i = 0; i < rn.size(); i++) { if (i) std::printf(","); jd(rn[i]); }
        std::printf("],\"out\":[");
        for (size_t i = 0; i < rn.size() / 2; i++) {
            Vector2 o = sample(f, Vector2{rn[2 * i], rn[2 * i + 1]});
            if (i) std::printf(",");
            jd(o.x); std::printf(","); jd(o.y);
        }
        std::printf("]}");
    }
    std::printf("]}\n");
}

// Material KATs: constant textures, seeded shading frames and directions.
static void kat_materials() {
    struct M { const char *name; Material m; int type; double a[3], b[3], rough, eta; };
    std::vector<M> ms;
    auto cs = [](double r, double g, double b) { return ConstantTexture<Spectrum>{Vector3{r, g, b}}; };
    auto cr = [](double v) { return ConstantTexture<Real>{v}; };
    ms.push_back({"lambertian", Lambertian{cs(0.5, 0.3, 0.8)}, 0, {0.5, 0.3, 0.8}, {0, 0, 0}, 0, 1});
    ms.push_back({"plastic_r0.3", RoughPlastic{cs(0.6, 0.2, 0.1), cs(1, 1, 1), cr(0.3), 1.5}, 1,
                  {0.6, 0.2, 0.1}, {1, 1, 1}, 0.3, 1.5});
    ms.push_back({"plastic_r0.05", RoughPlastic{cs(0.2, 0.7, 0.4), cs(0.8, 0.9, 1), cr(0.05), 1.33}, 1,
                  {0.2, 0.7, 0.4}, {0.8, 0.9, 1}, 0.05, 1.33});
    ms.push_back({"plastic_r0.004", RoughPlastic{cs(0.5, 0.5, 0.5), cs(0.04, 0.04, 0.04), cr(0.004), 1.5}, 1,
                  {0.5, 0.5, 0.5}, {0.04, 0.04, 0.04}, 0.004, 1.5});
    ms.push_back({"plastic_r0.9", RoughPlastic{cs(0.9, 0.1, 0.3), cs(1, 1, 1), cr(0.9), 2.0}, 1,
                  {0.9, 0.1, 0.3}, {1, 1, 1}, 0.9, 2.0});
    ms.push_back({"dielectric_r0.3", RoughDielectric{cs(1, 1, 1), cs(1, 1, 1), cr(0.3), 1.5}, 2,
                  {1, 1, 1}, {1, 1, 1}, 0.3, 1.5});
    ms.push_back({"dielectric_r0.1", RoughDielectric{cs(0.9, 0.8, 1), cs(0.7, 1, 0.9), cr(0.1), 1.33}, 2,
                  {0.9, 0.8, 1}, {0.7, 1, 0.9}, 0.1, 1.33});
    ms.push_back({"dielectric_r0.6", RoughDielectric{cs(1, 1, 1), cs(1, 1, 1), cr(0.6), 1 / 1.5}, 2,
                  {1, 1, 1}, {1, 1, 1}, 0.6, 1 / 1.5});
    std::printf("{\"materials\":[");
    for (size_t k = 0; k < ms.size(); k++) {
        const M &mm = ms[k];
        std::printf("%s{\"name\":\"%s\",\"type\":%d,\"a\":", k ? "," : "", mm.name, mm.type);
        jv(Vector3{mm.a[0], mm.a[1], mm.a[2]});
        std::printf(",\"b\":");
        jv(Vector3{mm.b[0], mm.b[1], mm.b[2]});
        std::printf(",\"roughness\":"); jd(mm.rough);
        std::printf(",\"eta\":"); jd(mm.eta);
        std::printf(",\"queries\":[");
        for (int q = 0; q < 160; q++) {
            PathVertex v;
            Vector3 ng = unit_sphere();
            // shading normal: equal to ng, or perturbed (sometimes across the horizon)
            Vector3 sn = ng;
            if (q % 4 == 1) sn = normalize(ng + 0.3 * unit_sphere());
            if (q % 4 == 2) sn = normalize(ng + 1.2 * unit_sphere());
            v.geometric_normal = ng;
            v.shading_frame = Frame(sn);
            v.uv = Vector2{0, 0};
            v.uv_screen_size = 0;
            Vector3 din = unit_sphere(), dout = unit_sphere();
            if (q % 3 == 0) {  // bias toward the upper hemisphere of ng
                if (dot(din, ng) < 0) din = -din;
                if (dot(dout, ng) < 0) dout = -dout;
            }
            if (q % 16 == 5) dout = normalize(-din + 2 * dot(din, sn) * sn);  // mirror direction
            Vector2 uv{U(), U()};
            Real w = U();
            TransportDirection td = (q % 8 == 7) ? TransportDirection::TO_VIEW : TransportDirection::TO_LIGHT;
            Spectrum f = eval(mm.m, din, dout, v, TexturePool{}, td);
            Real p = pdf_sample_bsdf(mm.m, din, dout, v, TexturePool{}, td);
            std::optional<BSDFSampleRecord> s = sample_bsdf(mm.m, din, v, TexturePool{}, uv, w, td);
            std::printf("%s{\"ng\":", q ? "," : "");
            jv(ng);
            std::printf(",\"fx\":"); jv(v.shading_frame.x);
            std::printf(",\"fy\":"); jv(v.shading_frame.y);
            std::printf(",\"fn\":"); jv(v.shading_frame.n);
            std::printf(",\"din\":"); jv(din);
            std::printf(",\"dout\":"); jv(dout);
            std::printf(",\"uv\":["); jd(uv.x); std::printf(","); jd(uv.y);
            std::printf("],\"w\":"); jd(w);
            std::printf(",\"to_view\":%d,\"f\":", td == TransportDirection::TO_VIEW ? 1 : 0);
            jv(f);
            std::printf(",\"pdf\":"); jd(p);
            std::printf(",\"valid\":%d", s ? 1 : 0);
            if (s) {
                std::printf(",\"sdir\":"); jv(s->dir_out);
                std::printf(",\"seta\":"); jd(s->eta);
                std::printf(",\"srough\":"); jd(s->roughness);
            }
            std::printf("}");
        }
        std::printf("]}");
    }
    std::printf("]}\n");
}

static void kat_phase() {
    std::printf("{\"phase\":[");
    struct P { int type; double g; };
    std::vector<P> ps = {{0, 0}, {1, -0.3}, {1, 0.0}, {1, 0.7}, {1, 0.0005}, {1, 0.95}};
    for (size_t k = 0; k < ps.size(); k++) {
        PhaseFunction pf = ps[k].type == 0 ? PhaseFunction{IsotropicPhase{}} : PhaseFunction{HenyeyGreenstein{ps[k].g}};
        std::printf("%s{\"type\":%d,\"g\":", k ? "," : "", ps[k].type);
        jd(ps[k].g);
        std::printf(",\"q\":[");
        for (int q = 0; q < 64; q++) {
            Vector3 a = unit_sphere(), b = unit_sphere();
            Vector2 r{U(), U()};
            Spectrum e = eval(pf, a, b);
            Real p = pdf_sample_phase(pf, a, b);
            std::optional<Vector3> s = sample_phase_function(pf, a, r);
            std::printf("%s{\"din\":", q ? "," : "");
            jv(a);
            std::printf(",\"dout\":"); jv(b);
            std::printf(",\"rnd\":["); jd(r.x); std::printf(","); jd(r.y);
            std::printf("],\"val\":"); jd(e.x);
            std::printf(",\"pdf\":"); jd(p);
            std::printf(",\"smp\":"); jv(*s);
            std::printf("}");
        }
        std::printf("]}");
    }
    std::printf("]}\n");
}

// Mipmaps and texture lookups (mipmap.h, texture.h), on power-of-two square
// images: on other sizes make_mipmap reads outside the image at its last
// levels, so there is no known answer to record.
static void kat_texture() {
    std::printf("{\"images\":[");
    TexturePool pool;
    const int sizes[3] = {8, 16, 4};
    const int chans[3] = {3, 1, 3};
    for (int k = 0; k < 3; k++) {
        int n = sizes[k];
        std::printf("%s{\"size\":%d,\"channels\":%d,\"data\":[", k ? "," : "", n, chans[k]);
        if (chans[k] == 3) {
            Image3 img(n, n);
            for (int i = 0; i < n * n; i++) {
                img(i) = Vector3{float(2 * U()), float(2 * U()), float(2 * U())};
                std::printf("%s", i ? "," : "");
                jd(img(i).x); std::printf(","); jd(img(i).y); std::printf(","); jd(img(i).z);
            }
            insert_image3(pool, "t" + std::to_string(k), img);
        } else {
            Image1 img(n, n);
            for (int i = 0; i < n * n; i++) {
                img(i) = float(2 * U());
                std::printf("%s", i ? "," : "");
                jd(img(i));
            }
            insert_image1(pool, "t" + std::to_string(k), img);
        }
        std::printf("],\"lookups\":[");
        for (int q = 0; q < 64; q++) {
            double us = q % 4 == 0 ? 1.0 : 0.25 + 3 * U(), vs = q % 4 == 0 ? 1.0 : 0.25 + 3 * U();
            double uo = q % 3 == 0 ? 0.0 : 4 * U() - 2, vo = q % 3 == 0 ? 0.0 : 4 * U() - 2;
            Vector2 uv{6 * U() - 3, 6 * U() - 3};
            double fp = q % 8 == 0 ? 0.0 : pow(10.0, -4 + 4 * U());
            Vector3 out;
            if (chans[k] == 3) {
                ImageTexture<Spectrum> t{int(pool.image3s.size()) - 1, us, vs, uo, vo};
                out = eval(Texture<Spectrum>{t}, uv, fp, pool);
            } else {
                ImageTexture<Real> t{int(pool.image1s.size()) - 1, us, vs, uo, vo};
                Real r = eval(Texture<Real>{t}, uv, fp, pool);
                out = Vector3{r, r, r};
            }
            std::printf("%s{\"scale\":[", q ? "," : "");
            jd(us); std::printf(","); jd(vs); std::printf("],\"offset\":["); jd(uo); std::printf(","); jd(vo);
            std::printf("],\"uv\":["); jd(uv.x); std::printf(","); jd(uv.y);
            std::printf("],\"footprint\":"); jd(fp); std::printf(",\"out\":"); jv(out); std::printf("}");
        }
        // every level of the mipmap
        std::printf("],\"levels\":[");
        if (chans[k] == 3) {
            const Mipmap3 &m = pool.image3s.back();
            for (size_t l = 0; l < m.images.size(); l++) {
                std::printf("%s[", l ? "," : "");
                for (int i = 0; i < m.images[l].width * m.images[l].height; i++) {
                    if (i) std::printf(",");
                    jv(m.images[l](i));
                }
                std::printf("]");
            }
        } else {
            const Mipmap1 &m = pool.image1s.back();
            for (size_t l = 0; l < m.images.size(); l++) {
                std::printf("%s[", l ? "," : "");
                for (int i = 0; i < m.images[l].width * m.images[l].height; i++) {
                    if (i) std::printf(",");
                    jd(m.images[l](i));
                }
                std::printf("]");
            }
        }
        std::printf("]}");
    }
    std::printf("],\"checker\":[");
    for (int q = 0; q < 64; q++) {
        Vector3 c0{U(), U(), U()}, c1{U(), U(), U()};
        double us = 0.5 + 4 * U(), vs = 0.5 + 4 * U(), uo = 2 * U() - 1, vo = 2 * U() - 1;
        Vector2 uv{4 * U() - 2, 4 * U() - 2};
        CheckerboardTexture<Spectrum> t{c0, c1, us, vs, uo, vo};
        Vector3 out = eval(Texture<Spectrum>{t}, uv, Real(0), pool);
        CheckerboardTexture<Real> t1{avg(c0), avg(c1), us, vs, uo, vo};
        Real out1 = eval(Texture<Real>{t1}, uv, Real(0), pool);
        std::printf("%s{\"color0\":", q ? "," : "");
        jv(c0); std::printf(",\"color1\":"); jv(c1);
        std::printf(",\"scale\":["); jd(us); std::printf(","); jd(vs);
        std::printf("],\"offset\":["); jd(uo); std::printf(","); jd(vo);
        std::printf("],\"uv\":["); jd(uv.x); std::printf(","); jd(uv.y);
        std::printf("],\"out\":"); jv(out); std::printf(",\"out1\":"); jd(out1); std::printf("}");
    }
    std::printf("]}\n");
}

int main(int argc, char **argv) {
    std::string mode = argc > 1 ? argv[1] : "";
    if (mode == "pcg") kat_pcg();
    else if (mode == "table") kat_table();
    else if (mode == "camera") kat_camera();
    else if (mode == "materials") kat_materials();
    else if (mode == "phase") kat_phase();
    else if (mode == "texture") kat_texture();
    else {
        std::fprintf(stderr, "usage: ref_kat pcg|table|camera|materials|phase|texture\n");
        return 2;
    }
    return 0;
}
