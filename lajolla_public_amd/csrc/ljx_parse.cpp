// Scene files: the Mitsuba-style XML subset lajolla's parser reads
// (src/parsers/parse_scene.cpp) with Wavefront OBJ (src/parsers/parse_obj.cpp),
// Mitsuba .serialized (load_serialized.cpp) and PLY (parse_ply.cpp) meshes,
// turned into an ljx_scene_desc.
//
// Covered: <default>, integrators path / volpath / direct and the auxiliary
// views (depth, shading normal, mean curvature, ray differential, mipmap
// level), the perspective sensor with fovAxis, film and reconstruction
// filter, sampler sampleCount, every BSDF of the reference with constant
// textures, obj / serialized / ply / sphere / rectangle shapes, area / point /
// directional emitters, homogeneous and heterogeneous media (constvolume and
// gridvolume from Mitsuba .vol files), and nested transforms.  Image textures
// and envmap images are reported as LJX_ERR_UNSUPPORTED.
//
// Numbers are read as FP32 and widened, as the reference's std::stof does
// (parse_scene.cpp:90-96); relative file names resolve against the scene
// file's folder (parse_scene.cpp:1606-1612).
#include "../../include/lajolla_amd.h"
#include "ljx_host.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include <zlib.h>

namespace ljx {
namespace {

constexpr double kPiD = 3.14159265358979323846;

// ------------------------------------------------------------------ XML
struct XNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<XNode> kids;
    const char *attr(const char *k) const {
        for (auto &a : attrs)
            if (a.first == k) return a.second.c_str();
        return nullptr;
    }
    std::string get(const char *k) const {
        const char *v = attr(k);
        return v ? std::string(v) : std::string();
    }
};

class XmlReader {
  public:
    explicit XmlReader(const std::string &text) : s_(text) {}
    XNode document() {
        XNode root;
        root.name = "#document";
        for (;;) {
            skip_misc();
            if (pos_ >= s_.size()) break;
            if (s_[pos_] != '<') err("expected '<'");
            root.kids.push_back(element());
        }
        return root;
    }

  private:
    const std::string &s_;
    size_t pos_ = 0;
    [[noreturn]] void err(const std::string &m) {
        size_t line = 1;
        for (size_t i = 0; i < pos_ && i < s_.size(); i++) line += s_[i] == '\n';
        fail(LJX_ERR_INVALID_ARGUMENT, "XML parse error at line " + std::to_string(line) + ": " + m);
    }
    bool starts(const char *p) const { return s_.compare(pos_, std::strlen(p), p) == 0; }
    void skip_ws() {
        while (pos_ < s_.size() && std::isspace((unsigned char)s_[pos_])) pos_++;
    }
    void skip_until(const char *end) {
        size_t e = s_.find(end, pos_);
        if (e == std::string::npos) err(std::string("unterminated construct, expected ") + end);
        pos_ = e + std::strlen(end);
    }
    // whitespace, comments, declarations, processing instructions, text
    void skip_misc() {
        for (;;) {
            skip_ws();
            if (pos_ >= s_.size()) return;
            if (starts("<!--")) skip_until("-->");
            else if (starts("<?")) skip_until("?>");
            else if (starts("<!")) skip_until(">");
            else if (s_[pos_] != '<') {
                while (pos_ < s_.size() && s_[pos_] != '<') pos_++;  // character data: unused
            } else return;
        }
    }
    std::string name() {
        size_t b = pos_;
        while (pos_ < s_.size() && (std::isalnum((unsigned char)s_[pos_]) || std::strchr("_-:.", s_[pos_]))) pos_++;
        if (b == pos_) err("expected a name");
        return s_.substr(b, pos_ - b);
    }
    static std::string unescape(const std::string &v) {
        std::string o;
        for (size_t i = 0; i < v.size(); i++) {
            if (v[i] != '&') {
                o += v[i];
                continue;
            }
            size_t e = v.find(';', i);
            std::string ent = e == std::string::npos ? "" : v.substr(i + 1, e - i - 1);
            if (ent == "amp") o += '&';
            else if (ent == "lt") o += '<';
            else if (ent == "gt") o += '>';
            else if (ent == "quot") o += '"';
            else if (ent == "apos") o += '\'';
            else {
                o += '&';
                continue;
            }
            i = e;
        }
        return o;
    }
    XNode element() {
        XNode n;
        pos_++;  // '<'
        n.name = name();
        for (;;) {
            skip_ws();
            if (pos_ >= s_.size()) err("unterminated tag <" + n.name + ">");
            if (starts("/>")) {
                pos_ += 2;
                return n;
            }
            if (s_[pos_] == '>') {
                pos_++;
                break;
            }
            std::string k = name();
            skip_ws();
            if (pos_ >= s_.size() || s_[pos_] != '=') err("expected '=' after attribute " + k);
            pos_++;
            skip_ws();
            char q = pos_ < s_.size() ? s_[pos_] : 0;
            if (q != '"' && q != '\'') err("expected a quoted value for attribute " + k);
            size_t e = s_.find(q, pos_ + 1);
            if (e == std::string::npos) err("unterminated attribute value");
            n.attrs.push_back({k, unescape(s_.substr(pos_ + 1, e - pos_ - 1))});
            pos_ = e + 1;
        }
        for (;;) {
            skip_misc();
            if (pos_ >= s_.size()) err("missing </" + n.name + ">");
            if (starts("</")) {
                pos_ += 2;
                std::string c = name();
                if (c != n.name) err("</" + c + "> closes <" + n.name + ">");
                skip_ws();
                if (pos_ >= s_.size() || s_[pos_] != '>') err("expected '>'");
                pos_++;
                return n;
            }
            n.kids.push_back(element());
        }
    }
};

// ---------------------------------------------------------------- values
struct M4 {
    double m[16];
    static M4 identity() {
        M4 r{};
        for (int i = 0; i < 4; i++) r.m[i * 5] = 1;
        return r;
    }
    double &operator()(int i, int j) { return m[i * 4 + j]; }
    double operator()(int i, int j) const { return m[i * 4 + j]; }
};
M4 operator*(const M4 &a, const M4 &b) {
    M4 r{};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double s = 0;
            for (int k = 0; k < 4; k++) s += a(i, k) * b(k, j);
            r(i, j) = s;
        }
    return r;
}
bool invert(const M4 &m, M4 &out) {
    double a[4][8];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 8; j++) a[i][j] = j < 4 ? m(i, j) : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; c++) {
        int piv = c;
        for (int r = c + 1; r < 4; r++)
            if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
        if (a[piv][c] == 0) return false;
        if (piv != c)
            for (int j = 0; j < 8; j++) std::swap(a[c][j], a[piv][j]);
        double inv = 1.0 / a[c][c];
        for (int j = 0; j < 8; j++) a[c][j] *= inv;
        for (int r = 0; r < 4; r++)
            if (r != c && a[r][c] != 0) {
                double f = a[r][c];
                for (int j = 0; j < 8; j++) a[r][j] -= f * a[c][j];
            }
    }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) out(i, j) = a[i][j + 4];
    return true;
}
struct V3 {
    double x = 0, y = 0, z = 0;
};
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double length(V3 a) { return std::sqrt(dot(a, a)); }
// vector.h:249-257
V3 normalize(V3 a) {
    double l = length(a);
    return l > 0 ? V3{a.x / l, a.y / l, a.z / l} : V3{};
}
V3 xform_point(const M4 &m, V3 p) {
    double x = m(0, 0) * p.x + m(0, 1) * p.y + m(0, 2) * p.z + m(0, 3);
    double y = m(1, 0) * p.x + m(1, 1) * p.y + m(1, 2) * p.z + m(1, 3);
    double z = m(2, 0) * p.x + m(2, 1) * p.y + m(2, 2) * p.z + m(2, 3);
    double w = m(3, 0) * p.x + m(3, 1) * p.y + m(3, 2) * p.z + m(3, 3);
    return {x / w, y / w, z / w};
}
V3 xform_vector(const M4 &m, V3 v) {
    return {m(0, 0) * v.x + m(0, 1) * v.y + m(0, 2) * v.z, m(1, 0) * v.x + m(1, 1) * v.y + m(1, 2) * v.z,
            m(2, 0) * v.x + m(2, 1) * v.y + m(2, 2) * v.z};
}
// normal through the inverse transpose, given the inverse
V3 xform_normal(const M4 &inv, V3 n) {
    return normalize({inv(0, 0) * n.x + inv(1, 0) * n.y + inv(2, 0) * n.z,
                      inv(0, 1) * n.x + inv(1, 1) * n.y + inv(2, 1) * n.z,
                      inv(0, 2) * n.x + inv(1, 2) * n.y + inv(2, 2) * n.z});
}

using Defaults = std::map<std::string, std::string>;

std::string resolve(const std::string &v, const Defaults &d) {
    if (!v.empty() && v[0] == '$') {
        auto it = d.find(v.substr(1));
        if (it != d.end()) return it->second;
    }
    return v;
}
double to_real(const std::string &t) {
    // FP32 first, like std::stof, then widened
    char *end = nullptr;
    float f = std::strtof(t.c_str(), &end);
    if (end == t.c_str()) fail(LJX_ERR_INVALID_ARGUMENT, "not a number: '" + t + "'");
    return double(f);
}
double real(const std::string &v, const Defaults &d) { return to_real(resolve(v, d)); }
int integer(const std::string &v, const Defaults &d) {
    std::string t = resolve(v, d);
    char *end = nullptr;
    long x = std::strtol(t.c_str(), &end, 10);
    if (end == t.c_str()) fail(LJX_ERR_INVALID_ARGUMENT, "not an integer: '" + t + "'");
    return int(x);
}
bool boolean(const std::string &v, const Defaults &d) { return resolve(v, d) == "true"; }
std::vector<std::string> tokens(const std::string &s, const char *seps) {
    std::vector<std::string> out;
    std::string cur;
    for (char c : s) {
        if (std::strchr(seps, c)) {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
        } else cur += c;
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}
V3 vec3(const std::string &v, const Defaults &d) {
    auto t = tokens(resolve(v, d), ", \t\n");
    if (t.size() == 1) {
        double a = to_real(t[0]);
        return {a, a, a};
    }
    if (t.size() != 3) fail(LJX_ERR_INVALID_ARGUMENT, "expected 1 or 3 numbers: '" + v + "'");
    return {to_real(t[0]), to_real(t[1]), to_real(t[2])};
}

// Spectrum to linear RGB: the CIE 1931 matching functions as the
// multi-lobe Gaussian fits of Wyman, Sloan and Shirley (JCGT 2(2), 2013),
// integrated from 400 to 700 nm in 1 nm steps over the piecewise-linear
// measurement (clamped at the ends), normalised by the Y integral, then XYZ
// to linear sRGB primaries (src/spectrum.h:45-120).
double lobe(double w, double mu, double s_lo, double s_hi) {
    double t = (w - mu) * (w < mu ? s_lo : s_hi);
    return std::exp(-0.5 * t * t);
}
V3 cie_xyz(double w) {
    return {0.362 * lobe(w, 442.0, 0.0624, 0.0374) + 1.056 * lobe(w, 599.8, 0.0264, 0.0323) -
                0.065 * lobe(w, 501.1, 0.0490, 0.0382),
            0.821 * lobe(w, 568.8, 0.0213, 0.0247) + 0.286 * lobe(w, 530.9, 0.0613, 0.0322),
            1.217 * lobe(w, 437.0, 0.0845, 0.0278) + 0.681 * lobe(w, 459.0, 0.0385, 0.0725)};
}
V3 xyz_to_rgb(V3 c) {
    return {3.240479 * c.x - 1.537150 * c.y - 0.498535 * c.z, -0.969256 * c.x + 1.875991 * c.y + 0.041556 * c.z,
            0.055648 * c.x - 0.204043 * c.y + 1.057311 * c.z};
}
V3 spectrum_xyz(const std::vector<std::pair<double, double>> &d) {
    V3 acc;
    if (d.empty()) return acc;
    for (int w = 400; w <= 700; w++) {
        double m;
        if (w <= d.front().first) m = d.front().second;
        else if (w >= d.back().first) m = d.back().second;
        else {
            size_t k = 1;
            while (k < d.size() && d[k].first <= w) k++;
            const auto &a = d[k - 1], &b = d[k];
            m = a.second * (b.first - w) / (b.first - a.first) + b.second * (w - a.first) / (b.first - a.first);
        }
        acc = acc + cie_xyz(double(w)) * m;
    }
    return acc * (1.0 / 106.856895);
}
std::vector<std::pair<double, double>> spectrum_list(const std::string &v, const Defaults &d) {
    auto t = tokens(resolve(v, d), ", \t\n");
    std::vector<std::pair<double, double>> s;
    if (t.size() == 1 && t[0].find(':') == std::string::npos) {
        s.push_back({-1.0, to_real(t[0])});
        return s;
    }
    for (auto &p : t) {
        auto kv = tokens(p, ":");
        if (kv.size() < 2) fail(LJX_ERR_INVALID_ARGUMENT, "bad spectrum entry '" + p + "'");
        s.push_back({to_real(kv[0]), to_real(kv[1])});
    }
    return s;
}
V3 srgb_hex(const std::string &v) {
    if (v.size() != 7 || v[0] != '#') fail(LJX_ERR_INVALID_ARGUMENT, "unknown sRGB value '" + v + "'");
    char *end = nullptr;
    long c = std::strtol(v.c_str() + 1, &end, 16);
    if (*end) fail(LJX_ERR_INVALID_ARGUMENT, "unknown sRGB value '" + v + "'");
    V3 s{double(float(((c >> 16) & 255) / 255.0f)), double(float(((c >> 8) & 255) / 255.0f)),
         double(float((c & 255) / 255.0f))};
    auto lin = [](double x) { return x <= 0.04045 ? x / 12.92 : std::pow((x + 0.055) / 1.055, 2.4); };
    return {lin(s.x), lin(s.y), lin(s.z)};
}
// reflectance-like colours (parse_scene.cpp:286-311)
V3 color(const XNode &n, const Defaults &d) {
    const std::string v = n.get("value");
    if (n.name == "spectrum") {
        auto s = spectrum_list(v, d);
        if (s.size() > 1) return xyz_to_rgb(spectrum_xyz(s));
        if (s.size() == 1) return {1, 1, 1};
        return {};
    }
    if (n.name == "rgb") return vec3(v, d);
    if (n.name == "srgb") return srgb_hex(resolve(v, d));
    if (n.name == "float") {
        double a = real(v, d);
        return {a, a, a};
    }
    if (n.name == "texture" || n.name == "ref")
        fail(LJX_ERR_UNSUPPORTED, "image and procedural textures are not supported by this build");
    fail(LJX_ERR_INVALID_ARGUMENT, "unknown colour element <" + n.name + ">");
}
// emitted radiance: a single-value spectrum uses the light white point
// XYZ(0.9505, 1, 1.0888) (parse_scene.cpp:500-526)
V3 intensity(const XNode &n, const Defaults &d) {
    const std::string v = n.get("value");
    if (n.name == "spectrum") {
        auto s = spectrum_list(v, d);
        if (s.size() == 1) return xyz_to_rgb(V3{0.9505, 1.0, 1.0888} * s[0].second);
        return xyz_to_rgb(spectrum_xyz(s));
    }
    if (n.name == "rgb") return vec3(v, d);
    if (n.name == "srgb") return srgb_hex(resolve(v, d));
    return {1, 1, 1};
}
double scalar(const XNode &n, const Defaults &d) {
    if (n.name == "float") return real(n.get("value"), d);
    V3 c = color(n, d);
    return (c.x + c.y + c.z) / 3;
}

M4 translate_m(V3 t) {
    M4 m = M4::identity();
    m(0, 3) = t.x, m(1, 3) = t.y, m(2, 3) = t.z;
    return m;
}
M4 scale_m(V3 s) {
    M4 m = M4::identity();
    m(0, 0) = s.x, m(1, 1) = s.y, m(2, 2) = s.z;
    return m;
}
// rotation by angle (degrees) about an axis (Rodrigues)
M4 rotate_m(double deg, V3 axis) {
    V3 a = normalize(axis);
    double s = std::sin(deg * kPiD / 180), c = std::cos(deg * kPiD / 180);
    M4 m = M4::identity();
    m(0, 0) = a.x * a.x + (1 - a.x * a.x) * c;
    m(0, 1) = a.x * a.y * (1 - c) - a.z * s;
    m(0, 2) = a.x * a.z * (1 - c) + a.y * s;
    m(1, 0) = a.x * a.y * (1 - c) + a.z * s;
    m(1, 1) = a.y * a.y + (1 - a.y * a.y) * c;
    m(1, 2) = a.y * a.z * (1 - c) - a.x * s;
    m(2, 0) = a.x * a.z * (1 - c) - a.y * s;
    m(2, 1) = a.y * a.z * (1 - c) + a.x * s;
    m(2, 2) = a.z * a.z + (1 - a.z * a.z) * c;
    return m;
}
// camera frame: columns left, up, forward, position (src/transform.cpp:46-70)
M4 look_at_m(V3 pos, V3 target, V3 up) {
    V3 dir = normalize(target - pos);
    V3 left = normalize(cross(normalize(up), dir));
    V3 nup = cross(dir, left);
    M4 m = M4::identity();
    m(0, 0) = left.x, m(1, 0) = left.y, m(2, 0) = left.z;
    m(0, 1) = nup.x, m(1, 1) = nup.y, m(2, 1) = nup.z;
    m(0, 2) = dir.x, m(1, 2) = dir.y, m(2, 2) = dir.z;
    m(0, 3) = pos.x, m(1, 3) = pos.y, m(2, 3) = pos.z;
    return m;
}
V3 xyz_attrs(const XNode &n, const Defaults &d, V3 def) {
    V3 v = def;
    if (n.attr("x")) v.x = real(n.get("x"), d);
    if (n.attr("y")) v.y = real(n.get("y"), d);
    if (n.attr("z")) v.z = real(n.get("z"), d);
    return v;
}
std::string lower(std::string s) {
    for (auto &c : s) c = char(std::tolower((unsigned char)c));
    return s;
}
// each child pre-multiplies the transform so far (parse_scene.cpp:209-282)
M4 transform(const XNode &n, const Defaults &d) {
    M4 t = M4::identity();
    for (const XNode &c : n.kids) {
        std::string nm = lower(c.name);
        M4 step = M4::identity();
        if (nm == "scale") {
            V3 s = xyz_attrs(c, d, {1, 1, 1});
            if (c.attr("value")) s = vec3(c.get("value"), d);
            step = scale_m(s);
        } else if (nm == "translate") {
            V3 v = xyz_attrs(c, d, {0, 0, 0});
            if (c.attr("value")) v = vec3(c.get("value"), d);
            step = translate_m(v);
        } else if (nm == "rotate") {
            V3 a = xyz_attrs(c, d, {0, 0, 0});
            step = rotate_m(c.attr("angle") ? real(c.get("angle"), d) : 0.0, a);
        } else if (nm == "lookat") {
            step = look_at_m(vec3(c.get("origin"), d), vec3(c.get("target"), d), vec3(c.get("up"), d));
        } else if (nm == "matrix") {
            auto t16 = tokens(resolve(c.get("value"), d), ", \t\n");
            if (t16.size() != 16) fail(LJX_ERR_INVALID_ARGUMENT, "matrix needs 16 values");
            for (int k = 0; k < 16; k++) step.m[k] = to_real(t16[k]);
        } else {
            continue;
        }
        t = step * t;
    }
    return t;
}

// ------------------------------------------------------------------- mesh
struct Mesh {
    std::vector<V3> pos, nrm;
    std::vector<std::pair<double, double>> uv;
    std::vector<int> idx;
};

// Angle-weighted vertex normals (Max 1999), zero for isolated vertices
// (src/parsers/shape_utils.h:15-50).
std::vector<V3> vertex_normals(const Mesh &m) {
    std::vector<V3> n(m.pos.size());
    // The corner weight is the reference's unit_angle (shape_utils.h:8-13),
    // including its (pi - 2) * asin(...) for obtuse corners, where Max's
    // angle is pi - 2 asin(...): normals computed here are the reference's.
    auto unit_angle = [](V3 u, V3 v) {
        return dot(u, v) < 0 ? (kPiD - 2) * std::asin(0.5 * length(v + u)) : 2 * std::asin(0.5 * length(v - u));
    };
    for (size_t f = 0; f + 2 < m.idx.size(); f += 3) {
        V3 fn;
        for (int i = 0; i < 3; i++) {
            const V3 &v0 = m.pos[m.idx[f + i]], &v1 = m.pos[m.idx[f + (i + 1) % 3]], &v2 = m.pos[m.idx[f + (i + 2) % 3]];
            V3 s1 = v1 - v0, s2 = v2 - v0;
            if (i == 0) {
                fn = cross(s1, s2);
                double l = length(fn);
                if (l == 0) break;
                fn = {fn.x / l, fn.y / l, fn.z / l};
            }
            n[m.idx[f + i]] = n[m.idx[f + i]] + fn * unit_angle(normalize(s1), normalize(s2));
        }
    }
    for (auto &v : n) {
        double l = length(v);
        v = l != 0 ? V3{v.x / l, v.y / l, v.z / l} : V3{0, 0, 0};
    }
    return n;
}

// OBJ: v (with optional w), vt (t flipped to 1 - t), vn (normalised),
// triangles and quads (split 0-1-2, 0-2-3); one output vertex per distinct
// (v, vt, vn) triple (src/parsers/parse_obj.cpp:91-185).
Mesh load_obj(const std::string &file, const M4 &to_world) {
    std::ifstream in(file);
    if (!in) fail(LJX_ERR_INVALID_ARGUMENT, "cannot open OBJ file " + file);
    M4 inv;
    if (!invert(to_world, inv)) fail(LJX_ERR_INVALID_ARGUMENT, "singular toWorld transform");
    std::vector<V3> P, N;
    std::vector<std::pair<double, double>> T;
    std::map<std::tuple<int, int, int>, int> ids;
    Mesh m;
    auto vid = [&](const std::string &tok) {
        int v = 0, t = 0, n = 0;
        size_t a = tok.find('/');
        v = std::atoi(tok.substr(0, a).c_str());
        if (a != std::string::npos) {
            size_t b = tok.find('/', a + 1);
            std::string ts = tok.substr(a + 1, b == std::string::npos ? std::string::npos : b - a - 1);
            if (!ts.empty()) t = std::atoi(ts.c_str());
            if (b != std::string::npos && b + 1 < tok.size()) n = std::atoi(tok.c_str() + b + 1);
        }
        if (v <= 0 || v > int(P.size()) || t < 0 || t > int(T.size()) || n < 0 || n > int(N.size()))
            fail(LJX_ERR_INVALID_ARGUMENT, "OBJ face index out of range in " + file);
        auto key = std::make_tuple(v - 1, t - 1, n - 1);
        auto it = ids.find(key);
        if (it != ids.end()) return it->second;
        int id = int(m.pos.size());
        m.pos.push_back(xform_point(to_world, P[v - 1]));
        if (t > 0) m.uv.push_back(T[t - 1]);
        if (n > 0) m.nrm.push_back(xform_normal(inv, N[n - 1]));
        ids[key] = id;
        return id;
    };
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string tok;
        if (!(ss >> tok) || tok[0] == '#') continue;
        if (tok == "v") {
            double x = 0, y = 0, z = 0, w = 1;
            ss >> x >> y >> z;
            if (!(ss >> w)) w = 1;
            P.push_back({x / w, y / w, z / w});
        } else if (tok == "vt") {
            double s = 0, t = 0;
            ss >> s >> t;
            T.push_back({s, 1 - t});
        } else if (tok == "vn") {
            double x = 0, y = 0, z = 0;
            ss >> x >> y >> z;
            N.push_back(normalize({x, y, z}));
        } else if (tok == "f") {
            std::vector<std::string> f;
            while (ss >> tok) f.push_back(tok);
            if (f.size() < 3) fail(LJX_ERR_INVALID_ARGUMENT, "OBJ face with fewer than 3 vertices in " + file);
            if (f.size() > 4) fail(LJX_ERR_UNSUPPORTED, "OBJ n-gons (n > 4) are not supported: " + file);
            int a = vid(f[0]), b = vid(f[1]), c = vid(f[2]);
            m.idx.insert(m.idx.end(), {a, b, c});
            if (f.size() == 4) {
                int e = vid(f[3]);
                m.idx.insert(m.idx.end(), {a, c, e});
            }
        }
    }
    if (!m.uv.empty() && m.uv.size() != m.pos.size()) m.uv.clear();
    if (!m.nrm.empty() && m.nrm.size() != m.pos.size()) m.nrm.clear();
    return m;
}

// ------------------------------------------------------- .serialized
// Mitsuba's .serialized meshes (load_serialized.cpp:102-256): a u16 magic and
// a u16 version (3 or 4).  Shape k > 0 starts at an offset from the table at
// the end of the file (version 4: u64 offsets, version 3: u32 offsets, then
// a u32 count), past its own 4-byte header.  Each shape is one zlib stream:
// u32 flags, (version 4) a NUL-terminated name, u64 vertex and triangle
// counts, positions, then normals, texture coordinates and colours as the
// flags say (FP64 when flag 0x2000 is set, else FP32; colours are skipped),
// then u32 x 3 indices per triangle.  Texture coordinates are not flipped.
Mesh load_serialized(const std::string &file, int shape_index, const M4 &to_world) {
    std::ifstream in(file, std::ios::binary);
    if (!in) fail(LJX_ERR_INVALID_ARGUMENT, "cannot open .serialized file " + file);
    std::vector<unsigned char> buf((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    auto rd = [&](size_t at, void *dst, size_t n) {
        if (at + n > buf.size()) fail(LJX_ERR_INVALID_ARGUMENT, "truncated .serialized file " + file);
        std::memcpy(dst, buf.data() + at, n);
    };
    uint16_t version = 0;
    rd(2, &version, 2);
    if (version != 3 && version != 4) fail(LJX_ERR_INVALID_ARGUMENT, "unknown .serialized version in " + file);
    size_t start = 4;
    if (shape_index > 0) {
        uint32_t count = 0;
        rd(buf.size() - 4, &count, 4);
        if (uint32_t(shape_index) >= count)
            fail(LJX_ERR_INVALID_ARGUMENT, "shapeIndex " + std::to_string(shape_index) + " out of range in " + file);
        if (version == 4) {
            uint64_t off = 0;
            rd(buf.size() - 4 - 8 * size_t(count - shape_index), &off, 8);
            start = size_t(off) + 4;
        } else {
            uint32_t off = 0;
            rd(buf.size() - 4 * size_t(count - shape_index + 1), &off, 4);
            start = size_t(off) + 4;
        }
        if (start > buf.size()) fail(LJX_ERR_INVALID_ARGUMENT, "bad shape offset in " + file);
    }
    // inflate the shape's stream
    std::vector<unsigned char> z;
    z_stream zs{};
    if (inflateInit2(&zs, 15) != Z_OK) fail(LJX_ERR_INVALID_ARGUMENT, "zlib initialisation failed");
    zs.next_in = buf.data() + start;
    zs.avail_in = uInt(buf.size() - start);
    unsigned char chunk[1 << 16];
    int ret = Z_OK;
    while (ret != Z_STREAM_END) {
        zs.next_out = chunk;
        zs.avail_out = sizeof(chunk);
        ret = inflate(&zs, Z_NO_FLUSH);
        if (ret != Z_OK && ret != Z_STREAM_END) {
            inflateEnd(&zs);
            fail(LJX_ERR_INVALID_ARGUMENT, "corrupt zlib stream in " + file);
        }
        z.insert(z.end(), chunk, chunk + (sizeof(chunk) - zs.avail_out));
        if (ret == Z_OK && zs.avail_in == 0 && zs.avail_out != 0) break;  // input ended early
    }
    inflateEnd(&zs);
    size_t at = 0;
    auto take = [&](void *dst, size_t n) {
        if (at + n > z.size()) fail(LJX_ERR_INVALID_ARGUMENT, "mesh data ends early in " + file);
        std::memcpy(dst, z.data() + at, n);
        at += n;
    };
    uint32_t flags = 0;
    take(&flags, 4);
    if (version == 4) {
        char c = 1;
        while (c != 0) take(&c, 1);
    }
    uint64_t nv = 0, nt = 0;
    take(&nv, 8);
    take(&nt, 8);
    if (nv > (1ull << 31) || nt > (1ull << 31)) fail(LJX_ERR_INVALID_ARGUMENT, "mesh too large in " + file);
    const bool dbl = flags & 0x2000;
    auto real = [&]() {
        if (dbl) {
            double v;
            take(&v, 8);
            return v;
        }
        float v;
        take(&v, 4);
        return double(v);
    };
    M4 inv;
    if (!invert(to_world, inv)) fail(LJX_ERR_INVALID_ARGUMENT, "singular toWorld transform");
    Mesh m;
    m.pos.resize(nv);
    for (auto &p : m.pos) {
        double x = real(), y = real(), zz = real();
        p = xform_point(to_world, {x, y, zz});
    }
    if (flags & 0x1) {
        m.nrm.resize(nv);
        for (auto &v : m.nrm) {
            double x = real(), y = real(), zz = real();
            v = xform_normal(inv, {x, y, zz});
        }
    }
    if (flags & 0x2) {
        m.uv.resize(nv);
        for (auto &t : m.uv) {
            double u = real(), v = real();
            t = {u, v};
        }
    }
    if (flags & 0x8)
        for (size_t k = 0; k < 3 * nv; k++) real();
    m.idx.resize(3 * nt);
    for (auto &i : m.idx) {
        uint32_t v;
        take(&v, 4);
        if (v >= nv) fail(LJX_ERR_INVALID_ARGUMENT, "vertex index out of range in " + file);
        i = int(v);
    }
    return m;
}

// ------------------------------------------------------------------- PLY
// Stanford PLY meshes (parse_ply.cpp:9-123): ascii, binary_little_endian or
// binary_big_endian; vertex x y z (required), u v and nx ny nz when present,
// face vertex_indices lists; other elements and properties are skipped.
// Positions and normals take the toWorld transform, texture coordinates are
// not flipped.  Faces with more than three corners are split into fans.
Mesh load_ply(const std::string &file, const M4 &to_world) {
    std::ifstream in(file, std::ios::binary);
    if (!in) fail(LJX_ERR_INVALID_ARGUMENT, "cannot open PLY file " + file);
    struct Prop {
        std::string name;
        int type = 0;       // size in bytes; negative for signed ints; 4f/8f below
        bool fp = false;
        bool list = false;
        int count_type = 0;
        bool count_fp = false;
    };
    struct Elem {
        std::string name;
        size_t count = 0;
        std::vector<Prop> props;
    };
    auto type_of = [&](const std::string &t, int &size, bool &fp) {
        fp = false;
        if (t == "char" || t == "int8") size = -1;
        else if (t == "uchar" || t == "uint8") size = 1;
        else if (t == "short" || t == "int16") size = -2;
        else if (t == "ushort" || t == "uint16") size = 2;
        else if (t == "int" || t == "int32") size = -4;
        else if (t == "uint" || t == "uint32") size = 4;
        else if (t == "float" || t == "float32") size = 4, fp = true;
        else if (t == "double" || t == "float64") size = 8, fp = true;
        else fail(LJX_ERR_INVALID_ARGUMENT, "unknown PLY property type " + t + " in " + file);
    };
    std::string line;
    std::getline(in, line);
    if (line.rfind("ply", 0) != 0) fail(LJX_ERR_INVALID_ARGUMENT, "not a PLY file: " + file);
    int format = -1;  // 0 ascii, 1 little endian, 2 big endian
    std::vector<Elem> elems;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::istringstream ls(line);
        std::string kw;
        ls >> kw;
        if (kw == "format") {
            std::string f;
            ls >> f;
            format = f == "ascii" ? 0 : f == "binary_little_endian" ? 1 : f == "binary_big_endian" ? 2 : -1;
            if (format < 0) fail(LJX_ERR_INVALID_ARGUMENT, "unknown PLY format " + f + " in " + file);
        } else if (kw == "element") {
            Elem e;
            ls >> e.name >> e.count;
            elems.push_back(e);
        } else if (kw == "property") {
            if (elems.empty()) fail(LJX_ERR_INVALID_ARGUMENT, "PLY property before any element in " + file);
            Prop p;
            std::string t;
            ls >> t;
            if (t == "list") {
                std::string ct, it;
                ls >> ct >> it >> p.name;
                p.list = true;
                type_of(ct, p.count_type, p.count_fp);
                type_of(it, p.type, p.fp);
            } else {
                ls >> p.name;
                type_of(t, p.type, p.fp);
            }
            elems.back().props.push_back(p);
        } else if (kw == "end_header") {
            break;
        }
    }
    if (format < 0) fail(LJX_ERR_INVALID_ARGUMENT, "PLY file without a format line: " + file);
    const bool swap = format == 2;
    auto read_bin = [&](int type, bool fp) -> double {
        unsigned char b[8];
        int n = std::abs(type);
        in.read(reinterpret_cast<char *>(b), n);
        if (!in) fail(LJX_ERR_INVALID_ARGUMENT, "PLY data ends early in " + file);
        if (swap) std::reverse(b, b + n);
        if (fp) {
            if (n == 4) {
                float v;
                std::memcpy(&v, b, 4);
                return v;
            }
            double v;
            std::memcpy(&v, b, 8);
            return v;
        }
        switch (type) {
        case -1: return double(int8_t(b[0]));
        case 1: return double(b[0]);
        case -2: { int16_t v; std::memcpy(&v, b, 2); return v; }
        case 2: { uint16_t v; std::memcpy(&v, b, 2); return v; }
        case -4: { int32_t v; std::memcpy(&v, b, 4); return v; }
        default: { uint32_t v; std::memcpy(&v, b, 4); return v; }
        }
    };
    auto read_ascii = [&]() -> double {
        std::string tok;
        if (!(in >> tok)) fail(LJX_ERR_INVALID_ARGUMENT, "PLY data ends early in " + file);
        return std::strtod(tok.c_str(), nullptr);
    };
    auto value = [&](int type, bool fp) { return format == 0 ? read_ascii() : read_bin(type, fp); };
    M4 inv;
    if (!invert(to_world, inv)) fail(LJX_ERR_INVALID_ARGUMENT, "singular toWorld transform");
    Mesh m;
    bool have_pos = false, have_faces = false;
    for (const Elem &e : elems) {
        const bool vert = e.name == "vertex", face = e.name == "face";
        int ix = -1, iy = -1, iz = -1, inx = -1, iny = -1, inz = -1, iu = -1, iv = -1, ifc = -1;
        for (int k = 0; k < int(e.props.size()); k++) {
            const std::string &n = e.props[k].name;
            if (vert && !e.props[k].list) {
                if (n == "x") ix = k;
                else if (n == "y") iy = k;
                else if (n == "z") iz = k;
                else if (n == "nx") inx = k;
                else if (n == "ny") iny = k;
                else if (n == "nz") inz = k;
                else if (n == "u") iu = k;
                else if (n == "v") iv = k;
            }
            if (face && e.props[k].list && n == "vertex_indices") ifc = k;
        }
        const bool pos = ix >= 0 && iy >= 0 && iz >= 0, nrm = inx >= 0 && iny >= 0 && inz >= 0, uv = iu >= 0 && iv >= 0;
        if (vert) have_pos = pos;
        if (face) have_faces = ifc >= 0;
        std::vector<double> vals(e.props.size());
        std::vector<int> corners;
        for (size_t r = 0; r < e.count; r++) {
            corners.clear();
            for (int k = 0; k < int(e.props.size()); k++) {
                const Prop &p = e.props[k];
                if (p.list) {
                    const double cnt = value(p.count_type, p.count_fp);
                    if (cnt < 0 || cnt > 1e6) fail(LJX_ERR_INVALID_ARGUMENT, "bad PLY list length in " + file);
                    for (int j = 0; j < int(cnt); j++) {
                        double v = value(p.type, p.fp);
                        if (k == ifc) corners.push_back(int(v));
                    }
                } else {
                    vals[k] = value(p.type, p.fp);
                }
            }
            if (vert && pos) {
                m.pos.push_back(xform_point(to_world, {vals[ix], vals[iy], vals[iz]}));
                if (nrm) m.nrm.push_back(xform_normal(inv, {vals[inx], vals[iny], vals[inz]}));
                if (uv) m.uv.push_back({vals[iu], vals[iv]});
            }
            if (face && ifc >= 0)
                for (size_t j = 2; j < corners.size(); j++)
                    m.idx.insert(m.idx.end(), {corners[0], corners[j - 1], corners[j]});
        }
    }
    if (!have_pos) fail(LJX_ERR_INVALID_ARGUMENT, "vertex positions not found in " + file);
    if (!have_faces) fail(LJX_ERR_INVALID_ARGUMENT, "vertex indices not found in " + file);
    for (int i : m.idx)
        if (i < 0 || size_t(i) >= m.pos.size()) fail(LJX_ERR_INVALID_ARGUMENT, "vertex index out of range in " + file);
    return m;
}

// ------------------------------------------------------------------ scene
struct Parsed {
    ljx_scene_desc desc{};
    std::vector<ljx_shape_desc> shapes;
    std::vector<ljx_material_desc> materials;
    std::vector<ljx_light_desc> lights;
    std::vector<ljx_medium_desc> media;
    std::vector<std::unique_ptr<std::vector<float>>> fbufs;
    std::vector<std::unique_ptr<std::vector<int32_t>>> ibufs;
    std::string film_filename = "image.exr";
};

class SceneReader {
  public:
    SceneReader(Parsed &out, std::string dir) : P(out), dir_(std::move(dir)) {}

    void read(const XNode &scene) {
        // defaults of the reference (src/scene.h:24-31, parse_scene.cpp:13-16)
        ljx_render_options &o = P.desc.options;
        o.integrator = LJX_INTEGRATOR_PATH;
        o.samples_per_pixel = 4;
        o.max_depth = -1;
        o.rr_depth = 5;
        o.vol_path_version = 0;
        o.max_null_collisions = 1000;
        o.seed = 0x31e241f862a1fb5eULL;
        camera_ = {};
        M4 id = M4::identity();
        std::memcpy(camera_.cam_to_world, id.m, sizeof(id.m));
        camera_.fov = 45.0;
        camera_.width = camera_.height = 256;
        camera_.filter_type = LJX_FILTER_BOX;
        camera_.filter_param = 1.0;
        camera_.medium_id = -1;
        P.desc.envmap_light_id = -1;
        for (const XNode &c : scene.kids) {
            if (c.name == "default") {
                if (c.attr("name") && c.attr("value")) defaults_[c.get("name")] = c.get("value");
            } else if (c.name == "integrator") integrator(c);
            else if (c.name == "sensor") sensor(c);
            else if (c.name == "bsdf") {
                auto [id2, m] = bsdf(c, "");
                if (!id2.empty()) {
                    material_ids_[id2] = int(P.materials.size());
                    P.materials.push_back(m);
                }
            } else if (c.name == "shape") shape(c);
            else if (c.name == "texture") fail(LJX_ERR_UNSUPPORTED, "textures are not supported by this build");
            else if (c.name == "emitter") emitter(c);
            else if (c.name == "medium") {
                auto [id2, m] = medium(c);
                if (!id2.empty()) {
                    medium_ids_[id2] = int(P.media.size());
                    P.media.push_back(m);
                }
            }
        }
        P.desc.camera = camera_;
        P.desc.shapes = P.shapes.data();
        P.desc.num_shapes = int(P.shapes.size());
        P.desc.materials = P.materials.data();
        P.desc.num_materials = int(P.materials.size());
        P.desc.lights = P.lights.data();
        P.desc.num_lights = int(P.lights.size());
        P.desc.media = P.media.data();
        P.desc.num_media = int(P.media.size());
    }

  private:
    Parsed &P;
    std::string dir_;
    Defaults defaults_;
    std::map<std::string, int> material_ids_, medium_ids_;
    ljx_camera_desc camera_{};

    std::string path(const std::string &f) const {
        if (f.empty() || f[0] == '/') return f;
        return dir_.empty() ? f : dir_ + "/" + f;
    }

    void integrator(const XNode &n) {
        ljx_render_options &o = P.desc.options;
        std::string t = n.get("type");
        auto each = [&](auto fn) {
            for (const XNode &c : n.kids) fn(c.get("name"), c);
        };
        if (t == "path" || t == "volpath") {
            o.integrator = t == "path" ? LJX_INTEGRATOR_PATH : LJX_INTEGRATOR_VOLPATH;
            each([&](const std::string &nm, const XNode &c) {
                if (nm == "maxDepth" || nm == "max_depth") o.max_depth = integer(c.get("value"), defaults_);
                else if (nm == "rrDepth" || nm == "rr_depth") o.rr_depth = integer(c.get("value"), defaults_);
                else if (t == "volpath" && nm == "version") o.vol_path_version = integer(c.get("value"), defaults_);
                else if (t == "volpath" && (nm == "maxNullCollisions" || nm == "max_null_collisions"))
                    o.max_null_collisions = integer(c.get("value"), defaults_);
            });
        } else if (t == "direct") {
            o.integrator = LJX_INTEGRATOR_PATH;
            o.max_depth = 2;
        } else if (t == "depth") o.integrator = LJX_INTEGRATOR_DEPTH;
        else if (t == "shadingNormal" || t == "shading_normal") o.integrator = LJX_INTEGRATOR_SHADING_NORMAL;
        else if (t == "meanCurvature" || t == "mean_curvature") o.integrator = LJX_INTEGRATOR_MEAN_CURVATURE;
        else if (t == "rayDifferential" || t == "ray_differential") o.integrator = LJX_INTEGRATOR_RAY_DIFFERENTIAL;
        else if (t == "mipmapLevel" || t == "mipmap_level") o.integrator = LJX_INTEGRATOR_MIPMAP_LEVEL;
        else fail(LJX_ERR_UNSUPPORTED, "unsupported integrator: " + t);
    }

    int medium_ref(const std::string &id) {
        auto it = medium_ids_.find(id);
        if (it == medium_ids_.end()) fail(LJX_ERR_INVALID_ARGUMENT, "medium reference " + id + " not found");
        return it->second;
    }

    void sensor(const XNode &n) {
        if (n.get("type") != "perspective") fail(LJX_ERR_UNSUPPORTED, "unsupported sensor: " + n.get("type"));
        double fov = 45.0;
        std::string axis = "x";
        M4 to_world = M4::identity();
        for (const XNode &c : n.kids) {
            std::string nm = c.get("name");
            if (nm == "fov") fov = real(c.get("value"), defaults_);
            else if (nm == "toWorld" || nm == "to_world") to_world = transform(c, defaults_);
            else if (nm == "fovAxis" || nm == "fov_axis") {
                axis = c.get("value");
                if (axis != "x" && axis != "y" && axis != "diagonal" && axis != "smaller" && axis != "larger")
                    fail(LJX_ERR_INVALID_ARGUMENT, "unknown fovAxis value: " + axis);
            }
        }
        int W = 256, H = 256;
        for (const XNode &c : n.kids) {
            if (c.name == "film") {
                for (const XNode &f : c.kids) {
                    std::string nm = f.get("name");
                    if (nm == "width") W = integer(f.get("value"), defaults_);
                    else if (nm == "height") H = integer(f.get("value"), defaults_);
                    else if (nm == "filename") P.film_filename = resolve(f.get("value"), defaults_);
                    if (f.name == "rfilter") filter(f);
                }
            } else if (c.name == "sampler") {
                for (const XNode &s : c.kids) {
                    std::string nm = s.get("name");
                    if (nm == "sampleCount" || nm == "sample_count")
                        P.desc.options.samples_per_pixel = integer(s.get("value"), defaults_);
                }
            } else if (c.name == "ref") {
                camera_.medium_id = medium_ref(c.get("id"));
            } else if (c.name == "medium") {
                auto [id, m] = medium(c);
                if (!id.empty()) medium_ids_[id] = int(P.media.size());
                camera_.medium_id = int(P.media.size());
                P.media.push_back(m);
            }
        }
        // fov along x (parse_scene.cpp:830-843)
        double aspect = double(W) / double(H);
        auto rad = [](double deg) { return deg * kPiD / 180; };
        auto deg = [](double r) { return r * 180 / kPiD; };
        if (axis == "y" || (axis == "smaller" && H < W) || (axis == "larger" && W < H)) {
            fov = deg(2 * std::atan(std::tan(rad(fov) / 2) * aspect));
        } else if (axis == "diagonal") {
            double diagonal = 2 * std::tan(rad(fov) / 2);
            double width = diagonal / std::sqrt(1 + 1 / (aspect * aspect));
            fov = deg(2 * std::atan(width / 2));
        }
        std::memcpy(camera_.cam_to_world, to_world.m, sizeof(to_world.m));
        camera_.fov = fov;
        camera_.width = W;
        camera_.height = H;
    }

    void filter(const XNode &f) {
        std::string t = f.get("type");
        auto param = [&](const char *name, double def) {
            for (const XNode &g : f.kids)
                if (g.get("name") == name) return real(g.get("value"), defaults_);
            return def;
        };
        if (t == "box") {
            camera_.filter_type = LJX_FILTER_BOX;
            camera_.filter_param = param("width", 1.0);
        } else if (t == "tent") {
            camera_.filter_type = LJX_FILTER_TENT;
            camera_.filter_param = param("width", 2.0);
        } else if (t == "gaussian") {
            camera_.filter_type = LJX_FILTER_GAUSSIAN;
            camera_.filter_param = param("stddev", 0.5);
        }
    }

    void phase(const XNode &c, ljx_medium_desc &m) {
        std::string pt = c.get("type");
        if (pt == "isotropic") m.phase_type = LJX_PHASE_ISOTROPIC;
        else if (pt == "hg") {
            m.phase_type = LJX_PHASE_HENYEY_GREENSTEIN;
            for (const XNode &g : c.kids)
                if (g.get("name") == "g") m.g = real(g.get("value"), defaults_);
        } else fail(LJX_ERR_INVALID_ARGUMENT, "unrecognised phase function: " + pt);
    }

    // load_volume (volume.cpp:6-114): Mitsuba's .vol, version 3, float32
    // voxels with 1 or 3 channels, x fastest, then the box.
    void load_vol(const std::string &file, ljx_volume_desc &v) {
        std::ifstream f(file, std::ios::binary);
        if (!f) fail(LJX_ERR_INVALID_ARGUMENT, "cannot open volume file " + file);
        char hdr[4] = {0, 0, 0, 0};
        f.read(hdr, 4);
        if (!f || hdr[0] != 'V' || hdr[1] != 'O' || hdr[2] != 'L' || uint8_t(hdr[3]) != 3)
            fail(LJX_ERR_INVALID_ARGUMENT, "not a version-3 .vol file: " + file);
        int32_t head[5];
        f.read(reinterpret_cast<char *>(head), sizeof(head));
        if (!f) fail(LJX_ERR_INVALID_ARGUMENT, "truncated .vol header: " + file);
        if (head[0] != 1) fail(LJX_ERR_UNSUPPORTED, "only float32 .vol grids are supported: " + file);
        const int rx = head[1], ry = head[2], rz = head[3], ch = head[4];
        if (ch != 1 && ch != 3) fail(LJX_ERR_UNSUPPORTED, ".vol grids have 1 or 3 channels: " + file);
        if (rx < 1 || ry < 1 || rz < 1 || double(rx) * ry * rz * ch > double(1u << 31))
            fail(LJX_ERR_INVALID_ARGUMENT, "bad .vol resolution: " + file);
        float box[6];
        f.read(reinterpret_cast<char *>(box), sizeof(box));
        auto buf = std::make_unique<std::vector<float>>(size_t(rx) * ry * rz * ch);
        f.read(reinterpret_cast<char *>(buf->data()), std::streamsize(buf->size() * sizeof(float)));
        if (!f) fail(LJX_ERR_INVALID_ARGUMENT, "truncated .vol data: " + file);
        v.type = LJX_VOLUME_GRID;
        v.channels = ch;
        v.resolution[0] = rx, v.resolution[1] = ry, v.resolution[2] = rz;
        for (int k = 0; k < 3; k++) v.p_min[k] = box[k], v.p_max[k] = box[3 + k];
        v.scale = 1;
        v.data = buf->data();
        P.fbufs.push_back(std::move(buf));
    }

    // parse_volume_spectrum (parse_scene.cpp:643-672)
    ljx_volume_desc volume(const XNode &n) {
        ljx_volume_desc v{};
        v.scale = 1;
        std::string t = n.get("type");
        if (t == "constvolume") {
            v.type = LJX_VOLUME_CONSTANT;
            for (const XNode &c : n.kids)
                if (c.get("name") == "value") {
                    V3 x = color(c, defaults_);
                    v.value[0] = x.x, v.value[1] = x.y, v.value[2] = x.z;
                }
        } else if (t == "gridvolume") {
            std::string file;
            for (const XNode &c : n.kids)
                if (c.get("name") == "filename") file = resolve(c.get("value"), defaults_);
            if (file.empty()) fail(LJX_ERR_INVALID_ARGUMENT, "gridvolume without a filename");
            load_vol(path(file), v);
        } else {
            fail(LJX_ERR_INVALID_ARGUMENT, "unknown volume type: " + t);
        }
        return v;
    }

    // parse_medium (parse_scene.cpp:692-746)
    std::pair<std::string, ljx_medium_desc> medium(const XNode &n) {
        ljx_medium_desc m{};
        std::string t = n.get("type");
        m.phase_type = LJX_PHASE_ISOTROPIC;
        double scale = 1;
        if (t == "homogeneous") {
            m.type = LJX_MEDIUM_HOMOGENEOUS;
            V3 sa{0.5, 0.5, 0.5}, ss{0.5, 0.5, 0.5};
            for (const XNode &c : n.kids) {
                std::string nm = c.get("name");
                if (nm == "sigmaA" || nm == "sigma_a") sa = color(c, defaults_);
                else if (nm == "sigmaS" || nm == "sigma_s") ss = color(c, defaults_);
                else if (nm == "scale") scale = real(c.get("value"), defaults_);
                else if (c.name == "phase") phase(c, m);
            }
            // the parser's scale multiplies both coefficients (parse_scene.cpp:704-722)
            m.sigma_a[0] = sa.x * scale, m.sigma_a[1] = sa.y * scale, m.sigma_a[2] = sa.z * scale;
            m.sigma_s[0] = ss.x * scale, m.sigma_s[1] = ss.y * scale, m.sigma_s[2] = ss.z * scale;
        } else if (t == "heterogeneous") {
            m.type = LJX_MEDIUM_HETEROGENEOUS;
            // both volumes default to the constant 1 (parse_scene.cpp:724-725)
            for (ljx_volume_desc *v : {&m.albedo, &m.density}) {
                v->type = LJX_VOLUME_CONSTANT;
                v->value[0] = v->value[1] = v->value[2] = 1;
                v->scale = 1;
            }
            for (const XNode &c : n.kids) {
                std::string nm = c.get("name");
                if (nm == "albedo") m.albedo = volume(c);
                else if (nm == "density") m.density = volume(c);
                else if (nm == "scale") scale = real(c.get("value"), defaults_);
                else if (c.name == "phase") phase(c, m);
            }
            // the scale applies to the density only (parse_scene.cpp:739-740)
            m.density.scale = scale;
        } else {
            fail(LJX_ERR_INVALID_ARGUMENT, "unknown medium type: " + t);
        }
        return {n.get("id"), m};
    }

    // roughness from a Beckmann-style alpha: sqrt (parse_scene.cpp:849-913)
    double alpha_roughness(const XNode &c) { return std::sqrt(scalar(c, defaults_)); }

    std::pair<std::string, ljx_material_desc> bsdf(const XNode &n, const std::string &parent_id) {
        std::string t = n.get("type");
        std::string id = n.attr("id") ? n.get("id") : parent_id;
        ljx_material_desc m{};
        auto set3 = [](double *dst, V3 v) {
            dst[0] = v.x, dst[1] = v.y, dst[2] = v.z;
        };
        if (t == "twosided") {
            // every BSDF is two-sided here (parse_scene.cpp:926-932)
            for (const XNode &c : n.kids)
                if (c.name == "bsdf") return bsdf(c, id);
            fail(LJX_ERR_INVALID_ARGUMENT, "twosided without a nested bsdf");
        }
        if (t == "diffuse") {
            m.type = LJX_MAT_LAMBERTIAN;
            set3(m.color_a, {0.5, 0.5, 0.5});
            for (const XNode &c : n.kids)
                if (c.get("name") == "reflectance") set3(m.color_a, color(c, defaults_));
        } else if (t == "roughplastic" || t == "plastic") {
            m.type = LJX_MAT_ROUGH_PLASTIC;
            set3(m.color_a, {0.5, 0.5, 0.5});
            set3(m.color_b, {1, 1, 1});
            m.roughness = t == "plastic" ? 0.01 : 0.1;
            double ior_in = to_real("1.49"), ior_out = to_real("1.000277");
            for (const XNode &c : n.kids) {
                std::string nm = c.get("name");
                if (nm == "diffuseReflectance" || nm == "diffuse_reflectance") set3(m.color_a, color(c, defaults_));
                else if (nm == "specularReflectance" || nm == "specular_reflectance") set3(m.color_b, color(c, defaults_));
                else if (nm == "alpha") m.roughness = alpha_roughness(c);
                else if (nm == "roughness") m.roughness = scalar(c, defaults_);
                else if (nm == "intIOR" || nm == "int_ior") ior_in = real(c.get("value"), defaults_);
                else if (nm == "extIOR" || nm == "ext_ior") ior_out = real(c.get("value"), defaults_);
            }
            m.eta = ior_in / ior_out;
        } else if (t == "roughdielectric" || t == "dielectric") {
            m.type = LJX_MAT_ROUGH_DIELECTRIC;
            set3(m.color_a, {1, 1, 1});
            set3(m.color_b, {1, 1, 1});
            m.roughness = t == "dielectric" ? 0.01 : 0.1;
            double ior_in = to_real("1.5046"), ior_out = to_real("1.000277");
            for (const XNode &c : n.kids) {
                std::string nm = c.get("name");
                if (nm == "specularReflectance" || nm == "specular_reflectance") set3(m.color_a, color(c, defaults_));
                else if (nm == "specularTransmittance" || nm == "specular_transmittance")
                    set3(m.color_b, color(c, defaults_));
                else if (nm == "alpha") m.roughness = alpha_roughness(c);
                else if (nm == "roughness") m.roughness = scalar(c, defaults_);
                else if (nm == "intIOR" || nm == "int_ior") ior_in = real(c.get("value"), defaults_);
                else if (nm == "extIOR" || nm == "ext_ior") ior_out = real(c.get("value"), defaults_);
            }
            m.eta = ior_in / ior_out;
        } else if (t.rfind("disney", 0) == 0 || t == "principled") {
            // defaults of parse_scene.cpp:1009-1164
            set3(m.color_a, {0.5, 0.5, 0.5});
            m.roughness = 0.5;
            m.eta = 1.5;
            m.specular = 0.5;
            m.sheen_tint = 0.5;
            m.clearcoat_gloss = 1.0;
            if (t == "disneydiffuse") m.type = LJX_MAT_DISNEY_DIFFUSE;
            else if (t == "disneymetal") m.type = LJX_MAT_DISNEY_METAL;
            else if (t == "disneyglass") m.type = LJX_MAT_DISNEY_GLASS;
            else if (t == "disneyclearcoat") m.type = LJX_MAT_DISNEY_CLEARCOAT;
            else if (t == "disneysheen") m.type = LJX_MAT_DISNEY_SHEEN;
            else if (t == "disneybsdf" || t == "principled") m.type = LJX_MAT_DISNEY_BSDF;
            else fail(LJX_ERR_INVALID_ARGUMENT, "unknown BSDF: " + t);
            for (const XNode &c : n.kids) {
                std::string nm = c.get("name");
                if (nm == "baseColor" || nm == "base_color") set3(m.color_a, color(c, defaults_));
                else if (nm == "roughness") m.roughness = scalar(c, defaults_);
                else if (nm == "anisotropic") m.anisotropic = scalar(c, defaults_);
                else if (nm == "subsurface") m.subsurface = scalar(c, defaults_);
                else if (nm == "metallic") m.metallic = scalar(c, defaults_);
                else if (nm == "specular") m.specular = scalar(c, defaults_);
                else if (nm == "specularTint" || nm == "specular_tint" || nm == "specTint" || nm == "spec_tint")
                    m.specular_tint = scalar(c, defaults_);
                else if (nm == "specularTransmission" || nm == "specular_transmission" || nm == "specTrans" ||
                         nm == "spec_trans")
                    m.specular_transmission = scalar(c, defaults_);
                else if (nm == "sheen") m.sheen = scalar(c, defaults_);
                else if (nm == "sheenTint" || nm == "sheen_tint") m.sheen_tint = scalar(c, defaults_);
                else if (nm == "clearcoat") m.clearcoat = scalar(c, defaults_);
                else if (nm == "clearcoatGloss" || nm == "clearcoat_gloss") m.clearcoat_gloss = scalar(c, defaults_);
                else if (nm == "eta") m.eta = real(c.get("value"), defaults_);
            }
        } else if (t == "null") {
            // a black Lambertian, as the reference has (parse_scene.cpp:1166-1169)
            m.type = LJX_MAT_LAMBERTIAN;
        } else {
            fail(LJX_ERR_INVALID_ARGUMENT, "unknown BSDF: " + t);
        }
        return {id, m};
    }

    void add_mesh(ljx_shape_desc &s, const Mesh &m) {
        auto fp = std::make_unique<std::vector<float>>();
        for (auto &p : m.pos) fp->insert(fp->end(), {float(p.x), float(p.y), float(p.z)});
        auto ip = std::make_unique<std::vector<int32_t>>(m.idx.begin(), m.idx.end());
        s.type = LJX_SHAPE_TRIANGLE_MESH;
        s.positions = fp->data();
        s.indices = ip->data();
        s.num_vertices = m.pos.size();
        s.num_triangles = m.idx.size() / 3;
        P.fbufs.push_back(std::move(fp));
        P.ibufs.push_back(std::move(ip));
        if (!m.nrm.empty()) {
            auto np = std::make_unique<std::vector<float>>();
            for (auto &v : m.nrm) np->insert(np->end(), {float(v.x), float(v.y), float(v.z)});
            s.normals = np->data();
            P.fbufs.push_back(std::move(np));
        }
        if (!m.uv.empty()) {
            auto up = std::make_unique<std::vector<float>>();
            for (auto &v : m.uv) up->insert(up->end(), {float(v.first), float(v.second)});
            s.uvs = up->data();
            P.fbufs.push_back(std::move(up));
        }
    }

    void shape(const XNode &n) {
        ljx_shape_desc s{};
        s.material_id = -1;
        s.area_light_id = -1;
        s.interior_medium_id = -1;
        s.exterior_medium_id = -1;
        for (const XNode &c : n.kids) {
            if (c.name == "ref") {
                std::string role = c.get("name"), id = c.get("id");
                if (id.empty()) fail(LJX_ERR_INVALID_ARGUMENT, "reference without an id");
                if (role == "interior") s.interior_medium_id = medium_ref(id);
                else if (role == "exterior") s.exterior_medium_id = medium_ref(id);
                else {
                    auto it = material_ids_.find(id);
                    if (it == material_ids_.end()) fail(LJX_ERR_INVALID_ARGUMENT, "material reference " + id + " not found");
                    s.material_id = it->second;
                }
            } else if (c.name == "bsdf") {
                auto [id, m] = bsdf(c, "");
                if (!id.empty()) material_ids_[id] = int(P.materials.size());
                s.material_id = int(P.materials.size());
                P.materials.push_back(m);
            } else if (c.name == "medium") {
                auto [id, m] = medium(c);
                if (!id.empty()) medium_ids_[id] = int(P.media.size());
                std::string role = c.get("name");
                if (role == "interior") s.interior_medium_id = int(P.media.size());
                else if (role == "exterior") s.exterior_medium_id = int(P.media.size());
                else fail(LJX_ERR_INVALID_ARGUMENT, "unrecognised medium name: " + role);
                P.media.push_back(m);
            }
        }
        std::string t = n.get("type");
        auto to_world_of = [&]() {
            for (const XNode &c : n.kids)
                if ((c.get("name") == "toWorld" || c.get("name") == "to_world") && c.name == "transform")
                    return transform(c, defaults_);
            return M4::identity();
        };
        auto flag = [&](const char *a, const char *b) {
            for (const XNode &c : n.kids)
                if (c.get("name") == a || c.get("name") == b) return boolean(c.get("value"), defaults_);
            return false;
        };
        if (t == "obj") {
            std::string file;
            for (const XNode &c : n.kids)
                if (c.get("name") == "filename") file = resolve(c.get("value"), defaults_);
            Mesh m = load_obj(path(file), to_world_of());
            if (flag("faceNormals", "face_normals")) m.nrm.clear();
            else if (m.nrm.empty()) m.nrm = vertex_normals(m);
            add_mesh(s, m);
        } else if (t == "sphere") {
            V3 c{0, 0, 0};
            double r = 1;
            for (const XNode &k : n.kids) {
                std::string nm = k.get("name");
                if (nm == "center") c = xyz_attrs(k, defaults_, {0, 0, 0});
                else if (nm == "radius") r = real(k.get("value"), defaults_);
            }
            s.type = LJX_SHAPE_SPHERE;
            s.center[0] = c.x, s.center[1] = c.y, s.center[2] = c.z;
            s.radius = r;
        } else if (t == "rectangle") {
            Mesh m;
            m.pos = {{-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0}};
            m.idx = {0, 1, 2, 0, 2, 3};
            m.uv = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
            double nz = flag("flipNormals", "flip_normals") ? -1 : 1;
            m.nrm = {{0, 0, nz}, {0, 0, nz}, {0, 0, nz}, {0, 0, nz}};
            M4 tw = to_world_of(), inv;
            if (!invert(tw, inv)) fail(LJX_ERR_INVALID_ARGUMENT, "singular toWorld transform");
            for (auto &p : m.pos) p = xform_point(tw, p);
            for (auto &v : m.nrm) v = xform_normal(inv, v);
            add_mesh(s, m);
        } else if (t == "serialized" || t == "ply") {
            std::string file;
            int index = 0;
            for (const XNode &c : n.kids) {
                std::string nm = c.get("name");
                if (nm == "filename") file = resolve(c.get("value"), defaults_);
                else if (nm == "shapeIndex" || nm == "shape_index") index = integer(c.get("value"), defaults_);
            }
            Mesh m = t == "ply" ? load_ply(path(file), to_world_of())
                                : load_serialized(path(file), index, to_world_of());
            if (flag("faceNormals", "face_normals")) m.nrm.clear();
            else if (m.nrm.empty()) m.nrm = vertex_normals(m);
            add_mesh(s, m);
        } else {
            fail(LJX_ERR_INVALID_ARGUMENT, "unknown shape: " + t);
        }
        for (const XNode &c : n.kids) {
            if (c.name != "emitter") continue;
            V3 rad{1, 1, 1};
            for (const XNode &g : c.kids)
                if (g.get("name") == "radiance") rad = intensity(g, defaults_);
            ljx_light_desc l{};
            l.type = LJX_LIGHT_DIFFUSE_AREA;
            l.shape_id = int(P.shapes.size());
            l.intensity[0] = rad.x, l.intensity[1] = rad.y, l.intensity[2] = rad.z;
            s.area_light_id = int(P.lights.size());
            P.lights.push_back(l);
        }
        P.shapes.push_back(s);
    }

    int black_material() {
        ljx_material_desc m{};
        m.type = LJX_MAT_LAMBERTIAN;
        P.materials.push_back(m);
        return int(P.materials.size()) - 1;
    }

    void emitter(const XNode &n) {
        std::string t = n.get("type");
        if (t == "envmap") {
            fail(LJX_ERR_UNSUPPORTED, "image environment maps are not supported by this build");
        } else if (t == "point") {
            // a tiny spherical emitter of the same power (parse_scene.cpp:1487-1519)
            V3 pos{0, 0, 0}, I{1, 1, 1};
            for (const XNode &g : n.kids) {
                std::string nm = g.get("name");
                if (nm == "position") pos = xyz_attrs(g, defaults_, {0, 0, 0});
                else if (nm == "intensity") I = intensity(g, defaults_);
            }
            double r = 1e-4;
            double k = 4 * kPiD / (4 * kPiD * r * r);
            ljx_shape_desc s{};
            s.type = LJX_SHAPE_SPHERE;
            s.center[0] = pos.x, s.center[1] = pos.y, s.center[2] = pos.z;
            s.radius = r;
            s.material_id = black_material();
            s.interior_medium_id = s.exterior_medium_id = -1;
            ljx_light_desc l{};
            l.type = LJX_LIGHT_DIFFUSE_AREA;
            l.shape_id = int(P.shapes.size());
            l.intensity[0] = I.x * k, l.intensity[1] = I.y * k, l.intensity[2] = I.z * k;
            s.area_light_id = int(P.lights.size());
            P.lights.push_back(l);
            P.shapes.push_back(s);
        } else if (t == "directional") {
            // a tiny far-away quad facing the direction (parse_scene.cpp:1520-1566)
            V3 dir{0, 0, 1}, I{1, 1, 1};
            for (const XNode &g : n.kids) {
                std::string nm = g.get("name");
                if (nm == "direction") dir = xyz_attrs(g, defaults_, dir);
                else if (nm == "toWorld" || nm == "to_world") dir = xform_vector(transform(g, defaults_), dir);
                else if (nm == "irradiance") I = intensity(g, defaults_);
            }
            dir = normalize(dir);
            V3 nd = dir * -1;
            // Duff et al. basis around -dir (src/frame.h:11-21)
            double sg = std::copysign(1.0, nd.z);
            double a = -1 / (sg + nd.z), b = nd.x * nd.y * a;
            V3 tg{1 + sg * nd.x * nd.x * a, sg * b, -sg * nd.x}, bt{b, sg + nd.y * nd.y * a, -nd.y};
            double len = 1e-3, dist = 1e3;
            Mesh m;
            m.pos = {(tg * -1 + bt * -1) * (0.5 * len) - dir * dist, (tg + bt * -1) * (0.5 * len) - dir * dist,
                     (tg + bt) * (0.5 * len) - dir * dist, (tg * -1 + bt) * (0.5 * len) - dir * dist};
            m.idx = {0, 1, 2, 0, 2, 3};
            m.nrm = {dir, dir, dir, dir};
            ljx_shape_desc s{};
            add_mesh(s, m);
            s.material_id = black_material();
            s.interior_medium_id = s.exterior_medium_id = -1;
            double k = dist * dist / (len * len);
            ljx_light_desc l{};
            l.type = LJX_LIGHT_DIFFUSE_AREA;
            l.shape_id = int(P.shapes.size());
            l.intensity[0] = I.x * k, l.intensity[1] = I.y * k, l.intensity[2] = I.z * k;
            s.area_light_id = int(P.lights.size());
            P.lights.push_back(l);
            P.shapes.push_back(s);
        } else {
            fail(LJX_ERR_INVALID_ARGUMENT, "unknown emitter type: " + t);
        }
    }
};

}  // namespace
}  // namespace ljx

using namespace ljx;

extern "C" {

int ljx_scene_parse(const char *xml_path, ljx_scene_desc **desc_out) {
    return guarded([&]() {
        if (!xml_path || !desc_out) fail(LJX_ERR_INVALID_ARGUMENT, "NULL argument");
        *desc_out = nullptr;
        std::ifstream in(xml_path, std::ios::binary);
        if (!in) fail(LJX_ERR_INVALID_ARGUMENT, std::string("cannot open scene file ") + xml_path);
        std::stringstream ss;
        ss << in.rdbuf();
        std::string text = ss.str();
        XNode doc = XmlReader(text).document();
        const XNode *scene = nullptr;
        for (auto &k : doc.kids)
            if (k.name == "scene") scene = &k;
        if (!scene) fail(LJX_ERR_INVALID_ARGUMENT, "no <scene> element");
        std::string p(xml_path);
        size_t slash = p.find_last_of('/');
        auto parsed = std::make_unique<Parsed>();
        SceneReader(*parsed, slash == std::string::npos ? std::string() : p.substr(0, slash)).read(*scene);
        // the descriptor is the first member: its address identifies the owner
        *desc_out = &parsed.release()->desc;
    });
}

int ljx_scene_desc_free(ljx_scene_desc *desc) {
    return guarded([&]() {
        if (desc) delete reinterpret_cast<Parsed *>(desc);
    });
}

int ljx_scene_load(const char *xml_path, int device, void **scene_out) {
    ljx_scene_desc *d = nullptr;
    int rc = ljx_scene_parse(xml_path, &d);
    if (rc != LJX_OK) return rc;
    rc = ljx_scene_create(d, device, scene_out);
    ljx_scene_desc_free(d);
    return rc;
}

}  // extern "C"
