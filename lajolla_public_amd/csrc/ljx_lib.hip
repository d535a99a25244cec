// C ABI of the MI355X path-tracing core (include/lajolla_amd.h): scene
// upload, BVH build, the wavefront render loop and the parity entry points.
#include "../../include/lajolla_amd.h"
#include "ljx_bvh.h"
#include "ljx_bvh_gpu.h"
#include "ljx_guide_net.h"
#include "ljx_guide_render.h"
#include "ljx_host.h"
#include "ljx_kernels.h"
#include "ljx_vol.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

using namespace ljx;

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// ---------------------------------------------------------------- matrices
void mat_mul4(const double *a, const double *b, double *r) {
    double t[16];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double s = 0;
            for (int k = 0; k < 4; k++) s += a[i * 4 + k] * b[k * 4 + j];
            t[i * 4 + j] = s;
        }
    std::memcpy(r, t, sizeof(t));
}
// Gauss-Jordan elimination with partial pivoting.
bool mat_inv4(const double *m, double *out) {
    double a[4][8];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 8; j++) a[i][j] = j < 4 ? m[i * 4 + j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; c++) {
        int piv = c;
        for (int r = c + 1; r < 4; r++)
            if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
        if (a[piv][c] == 0) return false;
        if (piv != c)
            for (int j = 0; j < 8; j++) std::swap(a[c][j], a[piv][j]);
        double inv = 1.0 / a[c][c];
        for (int j = 0; j < 8; j++) a[c][j] *= inv;
        for (int r = 0; r < 4; r++) {
            if (r == c) continue;
            double f = a[r][c];
            if (f == 0) continue;
            for (int j = 0; j < 8; j++) a[r][j] -= f * a[c][j];
        }
    }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) out[i * 4 + j] = a[i][j + 4];
    return true;
}

double luminance_h(const double *c) { return c[0] * 0.212671 + c[1] * 0.715160 + c[2] * 0.072169; }

// TableDist1D construction (src/table_dist.cpp:3-24).
// make_mipmap (src/mipmap.h:25-48): 2x2 box filter in FP64, at most 8
// levels.  Where a level's width or height is already 1 the reference reads
// the pixel past the row or column (outside the image on its last row); that
// read is clamped to the image here.
void build_mipmap(TexRec &t, std::vector<double> &data, const float *px, int w, int h, int ch) {
    t.w[0] = w;
    t.h[0] = h;
    t.off[0] = (long long)data.size();
    for (size_t i = 0; i < size_t(w) * h * ch; i++) data.push_back(px[i]);
    int num_levels = std::min(int(std::ceil(std::log2(double(std::max(w, h))) + 1)), kMaxMipLevels);
    t.levels = num_levels;
    for (int l = 1; l < num_levels; l++) {
        const int pw = t.w[l - 1], ph = t.h[l - 1];
        const int nw = std::max(pw / 2, 1), nh = std::max(ph / 2, 1);
        const long long po = t.off[l - 1];
        t.w[l] = nw;
        t.h[l] = nh;
        t.off[l] = (long long)data.size();
        data.resize(data.size() + size_t(nw) * nh * ch);
        for (int y = 0; y < nh; y++)
            for (int x = 0; x < nw; x++)
                for (int c = 0; c < ch; c++) {
                    auto at = [&](int xx, int yy) {
                        return data[po + size_t(ch) * (size_t(std::min(yy, ph - 1)) * pw + std::min(xx, pw - 1)) + c];
                    };
                    data[t.off[l] + size_t(ch) * (size_t(y) * nw + x) + c] =
                        (at(2 * x, 2 * y) + at(2 * x + 1, 2 * y) + at(2 * x, 2 * y + 1) + at(2 * x + 1, 2 * y + 1)) / 4.0;
                }
    }
}

// make_table_dist_2d (src/table_dist.cpp:40-95)
struct Dist2D {
    std::vector<double> cdf_rows, pdf_rows, cdf_marginals, pdf_marginals;
    double total = 0;
};
Dist2D make_dist_2d(const std::vector<double> &f, int w, int h) {
    Dist2D t;
    t.cdf_rows.assign(size_t(h) * (w + 1), 0);
    t.pdf_rows.assign(size_t(h) * w, 0);
    for (int y = 0; y < h; y++) {
        double *c = &t.cdf_rows[size_t(y) * (w + 1)];
        c[0] = 0;
        for (int x = 0; x < w; x++) c[x + 1] = c[x] + f[size_t(y) * w + x];
        double integral = c[w];
        if (integral > 0) {
            for (int x = 0; x < w; x++) c[x] /= integral;
            for (int x = 0; x < w; x++) t.pdf_rows[size_t(y) * w + x] = f[size_t(y) * w + x] / integral;
        } else {
            for (int x = 0; x < w; x++) {
                t.pdf_rows[size_t(y) * w + x] = 1.0 / double(w);
                c[x] = double(x) / double(w);
            }
            c[w] = 1;
        }
    }
    t.cdf_marginals.assign(h + 1, 0);
    t.pdf_marginals.assign(h, 0);
    for (int y = 0; y < h; y++) t.cdf_marginals[y + 1] = t.cdf_marginals[y] + t.cdf_rows[size_t(y) * (w + 1) + w];
    t.total = t.cdf_marginals[h];
    if (t.total > 0) {
        for (int y = 0; y < h; y++) t.cdf_marginals[y] /= t.total;
        t.cdf_marginals[h] = 1;
        for (int y = 0; y < h; y++) t.pdf_marginals[y] = t.cdf_rows[size_t(y) * (w + 1) + w] / t.total;
    } else {
        for (int y = 0; y < h; y++) {
            t.pdf_marginals[y] = 1.0 / double(h);
            t.cdf_marginals[y] = double(y) / double(h);
        }
        t.cdf_marginals[h] = 1;
    }
    for (int y = 0; y < h; y++) t.cdf_rows[size_t(y) * (w + 1) + w] = 1;
    return t;
}

void make_dist(const std::vector<double> &f, std::vector<double> &pmf, std::vector<double> &cdf) {
    size_t n = f.size();
    pmf = f;
    cdf.assign(n + 1, 0.0);
    for (size_t i = 0; i < n; i++) cdf[i + 1] = cdf[i] + pmf[i];
    double total = cdf[n];
    if (total > 0) {
        for (size_t i = 0; i < n; i++) {
            pmf[i] /= total;
            cdf[i] /= total;
        }
    } else {
        for (size_t i = 0; i < n; i++) {
            pmf[i] = 1.0 / double(n);
            cdf[i] = double(i) / double(n);
        }
        cdf[n] = 1;
    }
}

// --------------------------------------------------------------- the scene
// Buffers of guided rendering, allocated on the first guided render.
struct GuideCtx {
    int capacity = 0, rec_depth = 0, sample_cap = 0;
    float4 *glob = nullptr, *gdir = nullptr, *rec_feat = nullptr, *rec_dir = nullptr, *rec_beta = nullptr,
           *rec_c = nullptr;
    float4 *s_feat = nullptr, *s_dir = nullptr;  // training samples of the last recording pass
    int *s_count = nullptr;
    int *s_blk = nullptr;  // per block of 256 paths: sample count, then offset
    std::vector<DevBuf> bufs;
};

struct RenderCtx {
    int capacity = 0;
    int trace_grid = 0;
    PathBufs P{};
    std::vector<DevBuf> bufs;
    std::unique_ptr<GuideCtx> gc;
    ~RenderCtx() {
        if (gc)
            for (auto &b : gc->bufs)
                if (b.p) hipFree(b.p);
    }
};

struct Scene {
    int device = 0;
    hipStream_t stream = nullptr;
    DeviceScene ds{};
    std::vector<DevBuf> bufs;
    ljx_render_options options{};
    int width = 0, height = 0;
    int num_cu = 0;
    int num_materials = 0;
    bool has_spheres = false;
    unsigned mat_types_present = 0;
    bool has_textured = false;  // some material reads a texture: shaded by the runtime-dispatch kernel
    bool has_media = false;
    bool has_grid_media = false;  // a heterogeneous medium: long delta-tracking loops
    int num_media = 0;
    double bound_center[3] = {0, 0, 0};
    double bound_radius = 0;
    float box_lo[3] = {0, 0, 0}, box_hi[3] = {0, 0, 0};  // axis-aligned bounds of every primitive
    int bvh_depth = 0;
    int spill_entries = 1;  // per-lane global stack entries beyond the LDS ones
    int stack_need = 0;     // most traversal stack entries a ray can hold
    int binary_depth = 0;
    int bvh_builder = LJX_BVH_BUILDER_GPU;
    int max_leaf = 8;
    float node_cost = 1.0f;  // SAH cost of a node visit in primitive tests
    int num_slots = 0;
    const int *slot_prim = nullptr;  // device: leaf slot -> primitive
    double bvh_ms = 0;
    double build_ms = 0;
    std::unique_ptr<RenderCtx> rc;
    int rc_alloc_W = 0, rc_alloc_H = 0;
    bool tuned = false;
    ljx_exec_options tuned_exec{};  // autotuned launch settings

    // side streams the shading kernels of one bounce fan out to (default
    // priority, like the scene's stream), with the events that fork and join them
    hipStream_t side[2] = {nullptr, nullptr};
    hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};

    ~Scene() {
        if (device >= 0) hipSetDevice(device);
        if (stream) hipStreamSynchronize(stream);
        release_ctx();
        for (auto &b : bufs)
            if (b.p) hipFree(b.p);
        for (int k = 0; k < 2; k++) {
            if (side[k]) hipStreamDestroy(side[k]);
            if (ev_join[k]) hipEventDestroy(ev_join[k]);
        }
        if (ev_fork) hipEventDestroy(ev_fork);
        if (stream) hipStreamDestroy(stream);
    }
    void release_ctx() {
        if (rc) {
            for (auto &b : rc->bufs)
                if (b.p) hipFree(b.p);
            rc.reset();
        }
    }
    template <class T>
    T *upload(const std::vector<T> &v, std::vector<DevBuf> *into = nullptr) {
        size_t bytes = std::max<size_t>(v.size() * sizeof(T), 16);
        void *p = nullptr;
        HIPX(hipMalloc(&p, bytes));
        (into ? *into : bufs).push_back(DevBuf{p, bytes});
        if (!v.empty()) HIPX(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return static_cast<T *>(p);
    }
    template <class T>
    T *alloc(size_t count, std::vector<DevBuf> &into) {
        size_t bytes = std::max<size_t>(count * sizeof(T), 16);
        void *p = nullptr;
        HIPX(hipMalloc(&p, bytes));
        into.push_back(DevBuf{p, bytes});
        return static_cast<T *>(p);
    }
};

// Camera record with lajolla's sample-to-camera matrix (src/camera.cpp:7-21).
CameraRec make_camera(const ljx_camera_desc &c) {
    if (c.width <= 0 || c.height <= 0) fail(LJX_ERR_INVALID_ARGUMENT, "camera width/height must be positive");
    if (c.filter_type > 2) fail(LJX_ERR_INVALID_ARGUMENT, "unknown filter type");
    CameraRec cam;
    std::memcpy(cam.cam_to_world, c.cam_to_world, sizeof(cam.cam_to_world));
    cam.width = c.width;
    cam.height = c.height;
    cam.filter_type = int(c.filter_type);
    cam.filter_param = c.filter_param;
    cam.medium_id = c.medium_id;
    double aspect = double(cam.width) / double(cam.height);
    double sm[16] = {-0.5, 0, 0, 0, 0, -0.5 * aspect, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    double tm[16] = {1, 0, 0, -1.0, 0, 1, 0, -1.0 / aspect, 0, 0, 1, 0, 0, 0, 0, 1};
    double cot = 1.0 / std::tan((kPi / 180.0) * (c.fov / 2.0));
    double pm[16] = {cot, 0, 0, 0, 0, cot, 0, 0, 0, 0, 1, -1, 0, 0, 1, 0};
    double st[16], c2s[16];
    mat_mul4(sm, tm, st);
    mat_mul4(st, pm, c2s);
    if (!mat_inv4(c2s, cam.sample_to_cam)) fail(LJX_ERR_INVALID_ARGUMENT, "camera projection is singular");
    return cam;
}

void validate(const ljx_scene_desc *d) {
    if (!d) fail(LJX_ERR_INVALID_ARGUMENT, "scene descriptor is NULL");
    const ljx_camera_desc &c = d->camera;
    if (c.width <= 0 || c.height <= 0) fail(LJX_ERR_INVALID_ARGUMENT, "camera width/height must be positive");
    if (c.filter_type > 2) fail(LJX_ERR_INVALID_ARGUMENT, "unknown filter type");
    if (d->num_shapes < 0 || d->num_materials < 0 || d->num_lights < 0 || d->num_media < 0)
        fail(LJX_ERR_INVALID_ARGUMENT, "negative element count");
    if ((d->num_shapes && !d->shapes) || (d->num_materials && !d->materials) || (d->num_lights && !d->lights) ||
        (d->num_media && !d->media))
        fail(LJX_ERR_INVALID_ARGUMENT, "NULL array with a nonzero count");
    if (c.medium_id < -1 || c.medium_id >= d->num_media) fail(LJX_ERR_INVALID_ARGUMENT, "camera medium id out of range");
    if (d->num_textures < 0 || (d->num_textures > 0 && !d->textures))
        fail(LJX_ERR_INVALID_ARGUMENT, "NULL texture array with a nonzero count");
    for (int i = 0; i < d->num_textures; i++) {
        const ljx_texture_desc &t = d->textures[i];
        if (t.type == LJX_TEXTURE_IMAGE) {
            if (!t.data || t.width < 1 || t.height < 1 || (t.channels != 1 && t.channels != 3) ||
                double(t.width) * t.height > double(1u << 28))
                fail(LJX_ERR_INVALID_ARGUMENT, "texture " + std::to_string(i) + ": an image needs pixels, a size and 1 or 3 channels");
        } else if (t.type != LJX_TEXTURE_CHECKERBOARD) {
            fail(LJX_ERR_INVALID_ARGUMENT, "texture " + std::to_string(i) + ": unknown type");
        }
    }
    for (int i = 0; i < d->num_materials; i++) {
        if (d->materials[i].type >= LJX_MAT_COUNT) fail(LJX_ERR_INVALID_ARGUMENT, "unknown material type");
        for (int k = 0; k < LJX_MAT_TEXTURE_SLOTS; k++)
            if (d->materials[i].texture[k] < 0 || d->materials[i].texture[k] > d->num_textures)
                fail(LJX_ERR_INVALID_ARGUMENT, "material " + std::to_string(i) + ": texture index out of range");
    }
    for (int i = 0; i < d->num_media; i++) {
        const ljx_medium_desc &m = d->media[i];
        if (m.type > LJX_MEDIUM_HETEROGENEOUS) fail(LJX_ERR_INVALID_ARGUMENT, "unknown medium type");
        if (m.phase_type > 1) fail(LJX_ERR_INVALID_ARGUMENT, "unknown phase function");
        if (m.type == LJX_MEDIUM_HETEROGENEOUS) {
            for (const ljx_volume_desc *v : {&m.albedo, &m.density}) {
                if (v->type > LJX_VOLUME_GRID) fail(LJX_ERR_INVALID_ARGUMENT, "unknown volume type");
                if (v->type == LJX_VOLUME_GRID &&
                    (!v->data || (v->channels != 1 && v->channels != 3) || v->resolution[0] < 1 ||
                     v->resolution[1] < 1 || v->resolution[2] < 1 ||
                     double(v->resolution[0]) * v->resolution[1] * v->resolution[2] * v->channels > double(1u << 31)))
                    fail(LJX_ERR_INVALID_ARGUMENT,
                         "medium " + std::to_string(i) + ": grid volumes need data, 1 or 3 channels and a positive resolution");
            }
        }
    }
    for (int i = 0; i < d->num_shapes; i++) {
        const ljx_shape_desc &s = d->shapes[i];
        if (s.type > 1) fail(LJX_ERR_INVALID_ARGUMENT, "unknown shape type");
        if (s.material_id < -1 || s.material_id >= d->num_materials)
            fail(LJX_ERR_INVALID_ARGUMENT, "shape " + std::to_string(i) + ": material id out of range");
        if (s.area_light_id < -1 || s.area_light_id >= d->num_lights)
            fail(LJX_ERR_INVALID_ARGUMENT, "shape " + std::to_string(i) + ": area light id out of range");
        if (s.interior_medium_id < -1 || s.interior_medium_id >= d->num_media || s.exterior_medium_id < -1 ||
            s.exterior_medium_id >= d->num_media)
            fail(LJX_ERR_INVALID_ARGUMENT, "shape " + std::to_string(i) + ": medium id out of range");
        if (s.type == LJX_SHAPE_TRIANGLE_MESH) {
            if (s.num_triangles && (!s.positions || !s.indices))
                fail(LJX_ERR_INVALID_ARGUMENT, "shape " + std::to_string(i) + ": NULL positions or indices");
            for (uint64_t t = 0; t < 3 * s.num_triangles; t++)
                if (s.indices[t] < 0 || uint64_t(s.indices[t]) >= s.num_vertices)
                    fail(LJX_ERR_INVALID_ARGUMENT, "shape " + std::to_string(i) + ": vertex index out of range");
        } else if (!(s.radius > 0)) {
            fail(LJX_ERR_INVALID_ARGUMENT, "shape " + std::to_string(i) + ": sphere radius must be positive");
        }
    }
    for (int i = 0; i < d->num_lights; i++) {
        const ljx_light_desc &l = d->lights[i];
        if (l.type == LJX_LIGHT_DIFFUSE_AREA) {
            if (l.shape_id < 0 || l.shape_id >= d->num_shapes)
                fail(LJX_ERR_INVALID_ARGUMENT, "light " + std::to_string(i) + ": shape id out of range");
            const ljx_shape_desc &s = d->shapes[l.shape_id];
            if (s.type == LJX_SHAPE_TRIANGLE_MESH && s.num_triangles == 0)
                fail(LJX_ERR_INVALID_ARGUMENT, "light " + std::to_string(i) + ": emitting mesh has no triangles");
        } else if (l.type == LJX_LIGHT_ENVMAP) {
            if (l.env_width < 1 || l.env_height < 1 || !l.env_rgb ||
                double(l.env_width) * l.env_height > double(1u << 28))
                fail(LJX_ERR_INVALID_ARGUMENT, "light " + std::to_string(i) + ": environment map needs texels and a size");
        } else {
            fail(LJX_ERR_INVALID_ARGUMENT, "unknown light type");
        }
    }
    if (d->envmap_light_id < -1 || d->envmap_light_id >= d->num_lights)
        fail(LJX_ERR_INVALID_ARGUMENT, "envmap light id out of range");
}

// Primitive boxes on the host, as the device computes them (signed zeros made
// positive): triangles from the global vertex positions, then spheres.
std::vector<PrimBox> host_prim_boxes(const std::vector<float> &vpos, const std::vector<int> &tri_vidx, int ntri,
                                     const std::vector<double4> &spheres) {
    std::vector<PrimBox> boxes(size_t(ntri) + spheres.size());
    for (int t = 0; t < ntri; t++) {
        const float *p0 = &vpos[3 * tri_vidx[3 * t]], *p1 = &vpos[3 * tri_vidx[3 * t + 1]],
                    *p2 = &vpos[3 * tri_vidx[3 * t + 2]];
        for (int k = 0; k < 3; k++) {
            boxes[t].lo[k] = std::min(p0[k], std::min(p1[k], p2[k])) + 0.0f;
            boxes[t].hi[k] = std::max(p0[k], std::max(p1[k], p2[k])) + 0.0f;
        }
    }
    for (size_t i = 0; i < spheres.size(); i++) {
        const double4 &c = spheres[i];
        const double cc[3] = {c.x, c.y, c.z};
        for (int k = 0; k < 3; k++) {
            boxes[ntri + i].lo[k] = float(cc[k] - c.w) + 0.0f;
            boxes[ntri + i].hi[k] = float(cc[k] + c.w) + 0.0f;
        }
    }
    return boxes;
}

__global__ __launch_bounds__(256) void k_quantize_nodes(const Bvh4Node *__restrict__ in, int n,
                                                        Bvh4QNode *__restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = quantize_bvh4(in[i]);
}

// Builds the 4-wide BVH over the scene's primitives (on the device, or on the
// host when the scene asks for it; both give the same tree), then the leaf-slot
// records and the shading bucket of every slot.  The geometry arrays are
// already on the device.
void build_bvh(Scene &sc, const ljx_scene_desc *d, const std::vector<MatRec> &mats, int ntri, int nsph,
               const std::vector<float> &vpos, const std::vector<int> &tri_vidx,
               const std::vector<double4> &spheres) {
    DeviceScene &S = sc.ds;
    const auto t0 = std::chrono::steady_clock::now();
    const int nprim = ntri + nsph;
    sc.num_slots = nprim;
    if (nprim == 0) {
        S.nodes = sc.upload(std::vector<Bvh4Node>());
        S.qnodes = sc.upload(std::vector<Bvh4QNode>());
        S.num_nodes = 0;
        S.prims = sc.upload(std::vector<float4>());
        S.num_slots = 0;
        S.slot_bucket = sc.upload(std::vector<int8_t>());
        S.slot_passthru = sc.upload(std::vector<int8_t>());
        sc.slot_prim = sc.upload(std::vector<int>());
        sc.bvh_depth = 0;
        sc.spill_entries = 1;
        return;
    }
    auto keep = [&](size_t bytes) -> void * {
        void *p = nullptr;
        HIPX(hipMalloc(&p, std::max<size_t>(bytes, 16)));
        sc.bufs.push_back(DevBuf{p, bytes});
        return p;
    };
    int stack_need = 0;
    if (sc.bvh_builder == LJX_BVH_BUILDER_GPU) {
        float4 *box = nullptr;
        HIPX(hipMalloc(&box, sizeof(float4) * 2 * size_t(nprim)));
        struct Free {
            void *p;
            ~Free() { hipFree(p); }
        } fbox{box};
        prim_boxes_device(S.vpos, S.tri_vidx, ntri, S.spheres, nsph, box, box + nprim, sc.stream);
        DeviceBvh dv;
        build_bvh_device(box, box + nprim, nprim, sc.max_leaf, sc.node_cost, sc.stream, keep, dv);
        S.nodes = dv.nodes;
        S.num_nodes = dv.num_nodes;
        sc.slot_prim = dv.slot_prim;
        sc.bvh_depth = dv.depth;
        sc.binary_depth = dv.max_binary_depth;
        stack_need = dv.stack_need;
    } else {
        BvhBuildResult bvh = build_bvh_sah(host_prim_boxes(vpos, tri_vidx, ntri, spheres), sc.max_leaf, sc.node_cost, 0);
        Bvh4Result bvh4 = collapse_bvh4(bvh);
        S.nodes = sc.upload(bvh4.nodes);
        S.num_nodes = int(bvh4.nodes.size());
        sc.slot_prim = sc.upload(bvh.slot_prim);
        sc.bvh_depth = bvh4.depth;
        sc.binary_depth = bvh.max_depth;
        stack_need = bvh4.stack_need;
    }
    {
        Bvh4QNode *q = static_cast<Bvh4QNode *>(keep(size_t(S.num_nodes) * sizeof(Bvh4QNode)));
        const int n = S.num_nodes;
        hipLaunchKernelGGL(k_quantize_nodes, dim3((n + 255) / 256), dim3(256), 0, sc.stream, S.nodes, n, q);
        HIPX(hipGetLastError());
        S.qnodes = q;
    }
    sc.stack_need = stack_need;
    sc.spill_entries = std::max(1, stack_need - kStackLds);
    if (sc.spill_entries > kMaxStackSpill) fail(LJX_ERR_UNSUPPORTED, "BVH needs a deeper traversal stack than this build has");
    // the bucket a closest hit on each slot joins when hits are sorted by
    // material: its material type, or the runtime-dispatch bucket for
    // textured materials, or the no-material bucket
    std::vector<int8_t> shape_bucket(d->num_shapes), shape_passthru(d->num_shapes);
    for (int i = 0; i < d->num_shapes; i++) {
        const int mid = d->shapes[i].material_id;
        shape_bucket[i] = int8_t(mid < 0 ? kBucketNoMat : (mats[mid].textured ? kBucketAny : mats[mid].type));
        shape_passthru[i] = int8_t(mid < 0 ? 1 : 0);
    }
    std::vector<DevBuf> tmp;
    const int8_t *sb_dev = sc.upload(shape_bucket, &tmp);
    const int8_t *sp_dev = sc.upload(shape_passthru, &tmp);
    struct FreeAll {
        std::vector<DevBuf> &v;
        ~FreeAll() {
            for (auto &b : v) hipFree(b.p);
        }
    } ftmp{tmp};
    float4 *prims = static_cast<float4 *>(keep(sizeof(float4) * 3 * size_t(nprim)));
    int8_t *slot_bucket = static_cast<int8_t *>(keep(size_t(nprim)));
    int8_t *slot_passthru = static_cast<int8_t *>(keep(size_t(nprim)));
    slot_prims_device(sc.slot_prim, nprim, S.vpos, S.tri_vidx, ntri, S.spheres, prims, sc.stream);
    slot_buckets_device(sc.slot_prim, nprim, S.prim_shape, sb_dev, slot_bucket, sc.stream);
    slot_buckets_device(sc.slot_prim, nprim, S.prim_shape, sp_dev, slot_passthru, sc.stream);
    HIPX(hipStreamSynchronize(sc.stream));
    S.prims = prims;
    S.num_slots = nprim;
    S.slot_bucket = slot_bucket;
    S.slot_passthru = slot_passthru;
    sc.bvh_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void build_scene(Scene &sc, const ljx_scene_desc *d) {
    validate(d);
    auto t0 = std::chrono::steady_clock::now();
    DeviceScene &S = sc.ds;
    sc.options = d->options;
    sc.num_materials = d->num_materials;
    sc.width = d->camera.width;
    sc.height = d->camera.height;
    sc.has_media = d->num_media > 0;
    sc.num_media = d->num_media;
    sc.has_grid_media = false;
    for (int i = 0; i < d->num_media; i++) sc.has_grid_media = sc.has_grid_media || d->media[i].type == LJX_MEDIUM_HETEROGENEOUS;

    // ---- flatten geometry
    std::vector<float> vpos, vnrm, vuv;
    std::vector<int> tri_vidx, prim_shape, prim_local;
    std::vector<ShapeRec> shapes(d->num_shapes);
    int ntri = 0;
    size_t nvert = 0;
    for (int i = 0; i < d->num_shapes; i++)
        if (d->shapes[i].type == LJX_SHAPE_TRIANGLE_MESH) {
            ntri += int(d->shapes[i].num_triangles);
            nvert += size_t(d->shapes[i].num_vertices);
        }
    vpos.reserve(3 * nvert);
    vnrm.reserve(3 * nvert);
    vuv.reserve(2 * nvert);
    tri_vidx.reserve(3 * size_t(ntri));
    prim_shape.reserve(size_t(ntri) + size_t(d->num_shapes));
    prim_local.reserve(size_t(ntri) + size_t(d->num_shapes));
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    std::vector<double> tri_cdf;
    std::vector<char> light_shape(d->num_shapes, 0);
    for (int i = 0; i < d->num_lights; i++)
        if (d->lights[i].type == LJX_LIGHT_DIFFUSE_AREA) light_shape[d->lights[i].shape_id] = 1;
    for (int i = 0; i < d->num_shapes; i++) {
        const ljx_shape_desc &s = d->shapes[i];
        ShapeRec &r = shapes[i];
        std::memset(&r, 0, sizeof(r));
        r.type = int(s.type);
        r.material_id = s.material_id;
        r.area_light_id = s.area_light_id;
        r.interior_medium_id = s.interior_medium_id;
        r.exterior_medium_id = s.exterior_medium_id;
        r.cdf_offset = -1;
        if (s.type == LJX_SHAPE_TRIANGLE_MESH) {
            int voff = int(vpos.size() / 3);
            r.has_normals = s.normals != nullptr;
            r.has_uvs = s.uvs != nullptr;
            r.tri_offset = int(tri_vidx.size() / 3);
            r.num_tris = int(s.num_triangles);
            for (uint64_t v = 0; v < s.num_vertices; v++) {
                for (int k = 0; k < 3; k++) {
                    float p = s.positions[3 * v + k];
                    vpos.push_back(p);
                    lo[k] = std::min(lo[k], p);
                    hi[k] = std::max(hi[k], p);
                    vnrm.push_back(s.normals ? s.normals[3 * v + k] : 0.f);
                }
                vuv.push_back(s.uvs ? s.uvs[2 * v] : 0.f);
                vuv.push_back(s.uvs ? s.uvs[2 * v + 1] : 0.f);
            }
            for (uint64_t t = 0; t < s.num_triangles; t++) {
                tri_vidx.push_back(s.indices[3 * t] + voff);
                tri_vidx.push_back(s.indices[3 * t + 1] + voff);
                tri_vidx.push_back(s.indices[3 * t + 2] + voff);
                prim_shape.push_back(i);
                prim_local.push_back(int(t));
            }
            // triangle areas in FP64 (src/shapes/triangle_mesh.inl:60-75): the
            // light's pdf and its triangle table, so emitting shapes only
            const bool emits = s.area_light_id >= 0 || light_shape[i];
            std::vector<double> areas(emits ? s.num_triangles : 0);
            double total = 0;
            for (uint64_t t = 0; emits && t < s.num_triangles; t++) {
                int a = s.indices[3 * t], b = s.indices[3 * t + 1], c = s.indices[3 * t + 2];
                double e1[3], e2[3];
                for (int k = 0; k < 3; k++) {
                    float pa = s.positions[3 * a + k], pbk = s.positions[3 * b + k], pc = s.positions[3 * c + k];
                    e1[k] = double(pbk) - double(pa);
                    e2[k] = double(pc) - double(pa);
                }
                double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                       cz = e1[0] * e2[1] - e1[1] * e2[0];
                areas[t] = std::sqrt(cx * cx + cy * cy + cz * cz) / 2;
                total += areas[t];
            }
            r.total_area = total;
            if (emits && s.num_triangles > 0) {
                std::vector<double> pmf, cdf;
                make_dist(areas, pmf, cdf);
                r.cdf_offset = int(tri_cdf.size());
                tri_cdf.insert(tri_cdf.end(), cdf.begin(), cdf.end());
            }
        } else {
            r.center[0] = s.center[0];
            r.center[1] = s.center[1];
            r.center[2] = s.center[2];
            r.radius = s.radius;
            r.total_area = 4 * kPi * s.radius * s.radius;
        }
    }
    int nsph = 0;
    std::vector<double4> spheres;
    for (int i = 0; i < d->num_shapes; i++) {
        const ljx_shape_desc &s = d->shapes[i];
        if (s.type != LJX_SHAPE_SPHERE) continue;
        spheres.push_back(make_double4(s.center[0], s.center[1], s.center[2], s.radius));
        prim_shape.push_back(i);
        prim_local.push_back(0);
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], float(s.center[k] - s.radius));
            hi[k] = std::max(hi[k], float(s.center[k] + s.radius));
        }
        nsph++;
    }
    // ---- scene bounds (src/scene.cpp:30-34)
    const int nprim = ntri + nsph;
    if (nprim == 0) {
        for (int k = 0; k < 3; k++) lo[k] = hi[k] = 0;
    }
    double dx = double(hi[0]) - lo[0], dy = double(hi[1]) - lo[1], dz = double(hi[2]) - lo[2];
    sc.bound_radius = std::sqrt(dx * dx + dy * dy + dz * dz) / 2;
    for (int k = 0; k < 3; k++) sc.bound_center[k] = (double(lo[k]) + double(hi[k])) / 2;
    for (int k = 0; k < 3; k++) {
        sc.box_lo[k] = lo[k];
        sc.box_hi[k] = hi[k];
    }
    S.bound_radius = sc.bound_radius;
    S.eps = std::min(sc.bound_radius * 1e-5, 0.01);

    // ---- materials
    std::vector<MatRec> mats(d->num_materials);
    for (int i = 0; i < d->num_materials; i++) {
        const ljx_material_desc &m = d->materials[i];
        MatRec &r = mats[i];
        std::memset(&r, 0, sizeof(r));
        r.type = int(m.type);
        r.a = D3{m.color_a[0], m.color_a[1], m.color_a[2]};
        r.b = D3{m.color_b[0], m.color_b[1], m.color_b[2]};
        r.roughness = m.roughness;
        r.eta = m.eta;
        r.specular_transmission = m.specular_transmission;
        r.metallic = m.metallic;
        r.subsurface = m.subsurface;
        r.specular = m.specular;
        r.specular_tint = m.specular_tint;
        r.anisotropic = m.anisotropic;
        r.sheen = m.sheen;
        r.sheen_tint = m.sheen_tint;
        r.clearcoat = m.clearcoat;
        r.clearcoat_gloss = m.clearcoat_gloss;
        for (int k = 0; k < LJX_MAT_TEXTURE_SLOTS; k++) {
            r.tex[k] = m.texture[k];
            if (m.texture[k]) r.textured = 1;
        }
    }
    // ---- textures (src/texture.h:13-19): the scene's, then each envmap's image
    std::vector<TexRec> texs;
    std::vector<double> tex_data;
    auto add_image = [&](const float *px, int w, int h, int ch, double us, double vs, double uo, double vo) {
        TexRec t{};
        t.type = 1;
        t.channels = ch;
        t.uscale = us, t.vscale = vs, t.uoffset = uo, t.voffset = vo;
        build_mipmap(t, tex_data, px, w, h, ch);
        texs.push_back(t);
        return int(texs.size()) - 1;
    };
    for (int i = 0; i < d->num_textures; i++) {
        const ljx_texture_desc &t = d->textures[i];
        if (t.type == LJX_TEXTURE_IMAGE) {
            add_image(t.data, t.width, t.height, t.channels, t.uscale, t.vscale, t.uoffset, t.voffset);
        } else {
            TexRec r{};
            r.type = 2;
            r.color0 = D3{t.color0[0], t.color0[1], t.color0[2]};
            r.color1 = D3{t.color1[0], t.color1[1], t.color1[2]};
            r.uscale = t.uscale, r.vscale = t.vscale, r.uoffset = t.uoffset, r.voffset = t.voffset;
            texs.push_back(r);
        }
    }
    sc.mat_types_present = 0;
    sc.has_textured = false;
    for (int i = 0; i < d->num_materials; i++) sc.has_textured = sc.has_textured || mats[i].textured;
    for (int i = 0; i < d->num_shapes; i++) {
        int mid = d->shapes[i].material_id;
        if (mid >= 0) sc.mat_types_present |= 1u << d->materials[mid].type;
    }

    // ---- lights and their selection table (src/scene.cpp:38-52)
    std::vector<LightRec> lights(d->num_lights);
    std::vector<double> power(d->num_lights);
    for (int i = 0; i < d->num_lights; i++) {
        const ljx_light_desc &l = d->lights[i];
        LightRec &r = lights[i];
        std::memset(&r, 0, sizeof(r));
        r.type = int(l.type);
        r.shape_id = l.shape_id;
        r.intensity = D3{l.intensity[0], l.intensity[1], l.intensity[2]};
        if (l.type == LJX_LIGHT_ENVMAP) {
            const int w = l.env_width, h = l.env_height;
            r.env_tex = add_image(l.env_rgb, w, h, 3, 1, 1, 0, 0);
            r.env_w = w;
            r.env_h = h;
            r.env_rgb = sc.upload(std::vector<float>(l.env_rgb, l.env_rgb + size_t(w) * h * 3));
            double inv[16];
            if (!mat_inv4(l.to_world, inv)) fail(LJX_ERR_INVALID_ARGUMENT, "envmap to_world is singular");
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) {
                    r.to_world[a * 3 + b] = l.to_world[a * 4 + b];
                    r.to_local[a * 3 + b] = inv[a * 4 + b];
                }
            r.scale = l.scale;
            // envmap.inl:75-98: luminance x sin(elevation) at the texel centres of level 0
            std::vector<double> f(size_t(w) * h);
            const TexRec &et = texs[r.env_tex];
            for (int y = 0, k = 0; y < h; y++) {
                double v = (y + 0.5) / double(h);
                double se = std::sin(kPi * v);
                for (int x = 0; x < w; x++) {
                    double u = (x + 0.5) / double(w);
                    f[k++] = luminance(mip_lookup(tex_data.data(), et, u, v, 0)) * se;
                }
            }
            Dist2D dd = make_dist_2d(f, w, h);
            r.cdf_rows = sc.upload(dd.cdf_rows);
            r.pdf_rows = sc.upload(dd.pdf_rows);
            r.cdf_marginals = sc.upload(dd.cdf_marginals);
            r.pdf_marginals = sc.upload(dd.pdf_marginals);
            // light_power (envmap.inl:1-5): pi r^2 * total / (w h)
            power[i] = kPi * sc.bound_radius * sc.bound_radius * dd.total / (double(w) * h);
        } else {
            double area = shapes[l.shape_id].total_area;
            power[i] = luminance_h(l.intensity) * area * kPi;
        }
    }
    std::vector<double> lpmf, lcdf;
    if (d->num_lights > 0) make_dist(power, lpmf, lcdf);

    // ---- media
    std::vector<MediumRec> media(d->num_media);
    auto volume_rec = [&](const ljx_volume_desc &v) {
        VolumeRec r{};
        r.type = int(v.type);
        r.scale = v.scale;
        r.value = D3{v.value[0], v.value[1], v.value[2]};
        r.rx = r.ry = r.rz = 1;
        r.channels = 3;
        if (v.type == LJX_VOLUME_CONSTANT) {
            r.max_value = r.scale * r.value;
            return r;
        }
        r.channels = v.channels;
        r.rx = v.resolution[0], r.ry = v.resolution[1], r.rz = v.resolution[2];
        r.p_min = D3{v.p_min[0], v.p_min[1], v.p_min[2]};
        r.p_max = D3{v.p_max[0], v.p_max[1], v.p_max[2]};
        const size_t n = size_t(r.rx) * r.ry * r.rz;
        std::vector<float> data(v.data, v.data + n * r.channels);
        // per-channel maximum from 0, as load_volume keeps it (volume.cpp:70-93)
        D3 mx{0, 0, 0};
        for (size_t k = 0; k < n; k++) {
            const float *q = &data[k * r.channels];
            D3 x = r.channels == 1 ? D3{q[0], q[0], q[0]} : D3{q[0], q[1], q[2]};
            mx = D3{std::max(mx.x, x.x), std::max(mx.y, x.y), std::max(mx.z, x.z)};
        }
        r.max_value = r.scale * mx;
        r.data = sc.upload(data);
        return r;
    };
    for (int i = 0; i < d->num_media; i++) {
        const ljx_medium_desc &m = d->media[i];
        media[i].type = int(m.type);
        media[i].sigma_a = D3{m.sigma_a[0], m.sigma_a[1], m.sigma_a[2]};
        media[i].sigma_s = D3{m.sigma_s[0], m.sigma_s[1], m.sigma_s[2]};
        media[i].phase_type = int(m.phase_type);
        media[i].g = m.g;
        if (m.type == LJX_MEDIUM_HETEROGENEOUS) {
            media[i].albedo = volume_rec(m.albedo);
            media[i].density = volume_rec(m.density);
        }
    }

    // ---- camera (src/camera.cpp:7-21)
    S.camera = make_camera(d->camera);

    // ---- geometry to the device
    S.vpos = sc.upload(vpos);
    S.vnrm = sc.upload(vnrm);
    S.vuv = sc.upload(vuv);
    S.tri_vidx = sc.upload(tri_vidx);
    S.prim_shape = sc.upload(prim_shape);
    S.spheres = sc.upload(spheres);

    // ---- BVH (the HIGH-quality build of rtcCommitScene, src/scene.cpp:20-27)
    build_bvh(sc, d, mats, ntri, nsph, vpos, tri_vidx, spheres);
    S.prim_local = sc.upload(prim_local);
    S.shapes = sc.upload(shapes);
    S.materials = sc.upload(mats);
    S.lights = sc.upload(lights);
    S.media = sc.upload(media);
    S.textures = sc.upload(texs);
    S.tex_data = sc.upload(tex_data);
    S.light_pmf = sc.upload(lpmf);
    S.light_cdf = sc.upload(lcdf);
    {
        // guide table of the light CDF: guide[k] = first index with cdf > k / m
        int m = 1;
        while (m < d->num_lights && m < 4096) m <<= 1;
        std::vector<int> guide(size_t(m) + 1, 0);
        if (!lcdf.empty()) cdf_guide(lcdf.data(), int(lcdf.size()), m, guide.data());
        S.light_guide = sc.upload(guide);
        S.light_guide_m = m;
    }
    S.tri_cdf = sc.upload(tri_cdf);
    S.num_lights = d->num_lights;
    S.envmap_light_id = d->envmap_light_id;
    S.num_triangles = ntri;
    sc.has_spheres = nsph > 0;
    sc.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// ------------------------------------------------------------ trace kernels
// SPH: the scene has spheres; without, the triangle-only traversal the
// render's closest-hit and any-hit kernels run in such scenes.
template <bool ANY, bool SPH>
__global__ __launch_bounds__(256) void k_trace_batch(DeviceScene S, const ljx_ray *rays, size_t n, ljx_hit *hits,
                                                   uint8_t *occ, int *spill, unsigned long long *stat, int stats) {
    __shared__ int stack[kStackLds * 256];
    int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    int stride = gridDim.x * blockDim.x;
    uint32_t nn = 0, np = 0;
    for (size_t i = gtid; i < n; i += stride) {
        ljx_ray r = rays[i];
        TraceHit h;
        bool ok = false;
        if (S.num_nodes > 0) {
            if (stats)
                ok = trace_ray<ANY, true, SPH>(TraceScene{S.qnodes, S.prims, S.spheres}, r.org[0], r.org[1], r.org[2], r.dir[0], r.dir[1],
                                          r.dir[2], r.tnear, r.tfar, stack + threadIdx.x, blockDim.x,
                                          spill + gtid, stride, h, nn, np);
            else
                ok = trace_ray<ANY, false, SPH>(TraceScene{S.qnodes, S.prims, S.spheres}, r.org[0], r.org[1], r.org[2], r.dir[0], r.dir[1],
                                           r.dir[2], r.tnear, r.tfar, stack + threadIdx.x, blockDim.x,
                                           spill + gtid, stride, h, nn, np);
        }
        if (ANY) {
            occ[i] = ok ? 1 : 0;
        } else {
            ljx_hit o;
            if (!ok) {
                o.t = 0;
                o.u = o.v = 0;
                o.shape_id = -1;
                o.prim_id = -1;
                o.ng[0] = o.ng[1] = o.ng[2] = 0;
            } else {
                float4 p0 = S.prims[3 * h.slot], p1 = S.prims[3 * h.slot + 1];
                int gid = __float_as_int(p0.w);
                o.t = h.t;
                o.u = h.u;
                o.v = h.v;
                o.shape_id = S.prim_shape[gid];
                o.prim_id = S.prim_local[gid];
                if (!SPH || __float_as_int(p1.w) == kPrimTri) {
                    float4 p2 = S.prims[3 * h.slot + 2];
                    float e1x = p1.x - p0.x, e1y = p1.y - p0.y, e1z = p1.z - p0.z;
                    float e2x = p2.x - p0.x, e2y = p2.y - p0.y, e2z = p2.z - p0.z;
                    o.ng[0] = e1y * e2z - e1z * e2y;
                    o.ng[1] = e1z * e2x - e1x * e2z;
                    o.ng[2] = e1x * e2y - e1y * e2x;
                } else {
                    double4 c = S.spheres[__float_as_int(p1.x)];
                    double td, su, sv;
                    sphere_hit_uv(c, r.org[0], r.org[1], r.org[2], r.dir[0], r.dir[1], r.dir[2], r.tnear, h.t, td,
                                  su, sv);
                    o.t = float(td);
                    o.u = float(su);
                    o.v = float(sv);
                    o.ng[0] = float(r.org[0] + td * r.dir[0] - c.x);
                    o.ng[1] = float(r.org[1] + td * r.dir[1] - c.y);
                    o.ng[2] = float(r.org[2] + td * r.dir[2] - c.z);
                }
            }
            hits[i] = o;
        }
    }
    if (stats) {
        atomicAdd(&stat[S_NODES], (unsigned long long)nn);
        atomicAdd(&stat[S_PRIMS], (unsigned long long)np);
    }
}

__global__ void k_bsdf_batch(DeviceScene S, const ljx_bsdf_query *q, size_t n, ljx_bsdf_result *out) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ljx_bsdf_query a = q[i];
    const MatRec &m = S.materials[a.material_id];
    ShadeGeom g;
    g.ng = D3{a.geometric_normal[0], a.geometric_normal[1], a.geometric_normal[2]};
    g.sf = Frame{D3{a.frame_x[0], a.frame_x[1], a.frame_x[2]}, D3{a.frame_y[0], a.frame_y[1], a.frame_y[2]},
                 D3{a.frame_n[0], a.frame_n[1], a.frame_n[2]}};
    D3 din{a.dir_in[0], a.dir_in[1], a.dir_in[2]}, dout{a.dir_out[0], a.dir_out[1], a.dir_out[2]};
    bool tv = a.transport_to_view != 0;
    D3 f = bsdf_eval<-1>(m, din, dout, g, tv);
    double pdf = bsdf_pdf<-1>(m, din, dout, g);
    BsdfSample s = bsdf_sample<-1>(m, din, g, a.rnd_uv[0], a.rnd_uv[1], a.rnd_w);
    ljx_bsdf_result r;
    r.f[0] = f.x;
    r.f[1] = f.y;
    r.f[2] = f.z;
    r.pdf = pdf;
    r.sample_valid = s.valid;
    r.sample_dir[0] = s.valid ? s.dir.x : 0;
    r.sample_dir[1] = s.valid ? s.dir.y : 0;
    r.sample_dir[2] = s.valid ? s.dir.z : 0;
    r.sample_eta = s.valid ? s.eta : 0;
    r.sample_roughness = s.valid ? s.roughness : 0;
    out[i] = r;
}

Scene *as_scene(void *p) {
    if (!p) fail(LJX_ERR_INVALID_ARGUMENT, "scene handle is NULL");
    return static_cast<Scene *>(p);
}

// Persistent traversal grid: blocks_per_cu blocks of 256 lanes per CU (the
// LDS stack allows eight).
int trace_grid_for(const Scene &sc, int blocks_per_cu = 0) {
    int b = blocks_per_cu > 0 ? std::min(blocks_per_cu, 16) : 8;
    return std::max(1, sc.num_cu * b);
}

struct SpillBuf {
    int *p = nullptr;
    ~SpillBuf() {
        if (p) hipFree(p);
    }
};

// ---------------------------------------------- allocation-free closest hit
// Device memory of ljx_trace_closest_device_noalloc: the per-lane traversal
// stack spill of its persistent grid (256 lanes x the scene's spill entries
// per block), 16 MiB.  Each device holds its own copy from module load until
// the process ends.  Nothing in it needs clearing: a lane writes a spill
// entry before it reads it.  Calls that overlap on one device share it and
// race.
constexpr size_t kScratchInts = size_t(4) << 20;
static __device__ __align__(256) int scratch[kScratchInts];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (the symbol's host-side name is not a device address).
int *scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) fail(LJX_ERR_INVALID_ARGUMENT, "device id out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        HIPX(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) fail(LJX_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[device] = static_cast<int *>(p);
    }
    return g_scratch_ptr[device];
}

// ------------------------------------------------------------ render loop
struct Timer {
    bool on;
    hipStream_t s;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[6];  // extend, shadow, shade, other, guide, train
    void begin(int k, hipEvent_t &a) {
        if (!on) return;
        HIPX(hipEventCreate(&a));
        HIPX(hipEventRecord(a, s));
    }
    void end(int k, hipEvent_t a) {
        if (!on) return;
        hipEvent_t b;
        HIPX(hipEventCreate(&b));
        HIPX(hipEventRecord(b, s));
        ev[k].push_back({a, b});
    }
    double total(int k) {
        double t = 0;
        for (auto &e : ev[k]) {
            float ms = 0;
            HIPX(hipEventElapsedTime(&ms, e.first, e.second));
            t += ms;
            hipEventDestroy(e.first);
            hipEventDestroy(e.second);
        }
        ev[k].clear();
        return t;
    }
};

void ensure_ctx(Scene &sc, int capacity, int trace_grid) {
    if (sc.rc && sc.rc->capacity >= capacity && sc.rc->trace_grid == trace_grid && sc.rc_alloc_W == sc.width &&
        sc.rc_alloc_H == sc.height)
        return;
    sc.release_ctx();
    auto rc = std::make_unique<RenderCtx>();
    rc->capacity = capacity;
    rc->trace_grid = trace_grid;
    PathBufs &P = rc->P;
    size_t C = size_t(capacity);
    P.capacity = capacity;
    P.rng = sc.alloc<uint64_t>(C, rc->bufs);
    P.tp = sc.alloc<float4>(C, rc->bufs);
    P.rad = sc.alloc<float4>(C, rc->bufs);
    P.prev = sc.alloc<float4>(C, rc->bufs);
    P.depth = sc.alloc<int>(C, rc->bufs);
    P.ray_o = sc.alloc<float4>(C, rc->bufs);
    P.ray_d = sc.alloc<float4>(C, rc->bufs);
    P.hit = sc.alloc<float4>(C, rc->bufs);
    P.extq0 = sc.alloc<uint32_t>(C, rc->bufs);
    P.extq1 = sc.alloc<uint32_t>(C, rc->bufs);
    P.bucket = sc.alloc<uint32_t>(C * kNumBuckets, rc->bufs);
    P.shq = sc.alloc<uint32_t>(C, rc->bufs);
    P.sh_o = sc.alloc<float4>(C, rc->bufs);
    P.sh_d = sc.alloc<float4>(C, rc->bufs);
    P.sh_c = sc.alloc<float4>(C, rc->bufs);
    P.sh_occ = sc.alloc<uint8_t>(C, rc->bufs);
    const size_t nblk = (C + kScanBlock - 1) / kScanBlock;
    P.alive = sc.alloc<uint8_t>(nblk * kScanBlock, rc->bufs);
    P.alive_blk = sc.alloc<int>(nblk, rc->bufs);
    P.ctr = sc.alloc<int>(C_TOTAL, rc->bufs);
    P.stat = sc.alloc<unsigned long long>(S_NUM, rc->bufs);
    P.spill = sc.alloc<int>(size_t(rc->trace_grid) * 256 * sc.spill_entries, rc->bufs);
    P.accum = sc.alloc<double>(size_t(sc.width) * sc.height * 3, rc->bufs);
    vol_alloc(sc.has_media, C, P, rc->bufs, [&](size_t bytes) {
        void *p = nullptr;
        HIPX(hipMalloc(&p, bytes));
        return p;
    });
    sc.rc_alloc_W = sc.width;
    sc.rc_alloc_H = sc.height;
    sc.rc = std::move(rc);
}

template <class K, class... A>
void launch(K kernel, int grid, int block, hipStream_t s, A... args) {
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, s, args...);
    HIPX(hipGetLastError());
}

// The shading kernels of one bounce work on disjoint buckets (their only
// shared writes are atomic appends to the shadow queue), so with fan set
// they run on three streams at once: a long kernel at one wave per SIMD
// (DisneyBSDF) leaves most of each SIMD to the others, and the launches'
// tails overlap.  Captured work (batches) keeps them on one stream.
template <bool SPH, bool GUIDE>
void shade_all_t(Scene &sc, PathBufs &P, int cur, int grid, int block, hipStream_t s, bool fan) {
    const DeviceScene &S = sc.ds;
    unsigned m = sc.mat_types_present;
    hipStream_t st[3] = {s, s, s};
    if (fan) {
        for (int k = 0; k < 2; k++) {
            if (!sc.side[k]) HIPX(hipStreamCreateWithFlags(&sc.side[k], hipStreamNonBlocking));
            if (!sc.ev_join[k]) HIPX(hipEventCreateWithFlags(&sc.ev_join[k], hipEventDisableTiming));
        }
        if (!sc.ev_fork) HIPX(hipEventCreateWithFlags(&sc.ev_fork, hipEventDisableTiming));
        HIPX(hipEventRecord(sc.ev_fork, s));
        for (int k = 0; k < 2; k++) {
            HIPX(hipStreamWaitEvent(sc.side[k], sc.ev_fork, 0));
            st[k + 1] = sc.side[k];
        }
    }
    int next = 0;
    auto on = [&]() { return st[next++ % 3]; };
    // the heaviest kernels first, one per stream
    if (m & (1u << 1)) launch(k_shade_w2<1, SPH, GUIDE>, grid, block, on(), S, P, cur);
    if (m & (1u << 8)) launch(k_shade_w2<8, SPH, GUIDE>, grid, block, on(), S, P, cur);
    if (m & (1u << 0)) launch(k_shade_w2<0, SPH, GUIDE>, grid, block, on(), S, P, cur);
    if (m & (1u << 2)) launch(k_shade<2, SPH, GUIDE>, grid, block, on(), S, P, cur);
    if (m & (1u << 3)) launch(k_shade_w2<3, SPH, GUIDE>, grid, block, on(), S, P, cur);
    if (m & (1u << 4)) launch(k_shade_w2<4, SPH, GUIDE>, grid, block, on(), S, P, cur);
    if (m & (1u << 5)) launch(k_shade<5, SPH, GUIDE>, grid, block, on(), S, P, cur);
    if (m & (1u << 6)) launch(k_shade<6, SPH, GUIDE>, grid, block, on(), S, P, cur);
    if (m & (1u << 7)) launch(k_shade_w2<7, SPH, GUIDE>, grid, block, on(), S, P, cur);
    // unsorted bounces and textured materials go to the runtime-dispatch kernel
    if (P.sort_threshold > 0 || sc.has_textured) launch(k_shade<-1, SPH, GUIDE>, grid, block, on(), S, P, cur);
    launch(k_miss, grid, block, on(), S, P);
    if (fan) {
        for (int k = 0; k < 2; k++) {
            HIPX(hipEventRecord(sc.ev_join[k], sc.side[k]));
            HIPX(hipStreamWaitEvent(s, sc.ev_join[k], 0));
        }
    }
}
void shade_all(Scene &sc, PathBufs &P, int cur, int grid, hipStream_t s, bool guide, int block = 256,
               bool fan = false) {
    if (guide) {
        if (sc.has_spheres) shade_all_t<true, true>(sc, P, cur, grid, block, s, fan);
        else shade_all_t<false, true>(sc, P, cur, grid, block, s, fan);
    } else {
        if (sc.has_spheres) shade_all_t<true, false>(sc, P, cur, grid, block, s, fan);
        else shade_all_t<false, false>(sc, P, cur, grid, block, s, fan);
    }
}
template <bool STATS, bool SPH>
void trace_stage_t(Scene &sc, PathBufs &P, int cur, int grid, int sgrid, hipStream_t s, bool shadow) {
    if (shadow) {
        launch(k_shadow<STATS, SPH>, grid, 256, s, sc.ds, P);
        launch(k_shadow_resolve, sgrid, 256, s, P);
    } else {
        launch(k_extend<STATS, SPH>, grid, 256, s, sc.ds, P, cur);
    }
}
// Paths per pass when the caller gives none: the most a shadow ray's packed
// path id can name (67,108,864).  Traversal and shading launches need many
// rays per lane to amortise their tails: a 1080p render runs at 359, 639, 693
// and 723 Msamples/s with 4M, 16M, 32M and 64M paths per pass.  The path
// state is ~0.5 KB per path (31 GB for a 1080p pass of 32 samples per pixel).
constexpr int kDefaultPathsInFlight = 1 << kSlotBits;

// The next extension queue from the flags shading left (k_alive_*).
void build_ext_queue(PathBufs &P, int n, int cur, hipStream_t s) {
    const int nblk = (n + kScanBlock - 1) / kScanBlock;
    if (nblk > 16384) fail(LJX_ERR_UNSUPPORTED, "pass larger than the queue builder handles");
    launch(k_alive_count, nblk, 256, s, P, n);
    launch(k_alive_scan, 1, 1024, s, P, nblk, cur);
    launch(k_alive_scatter, nblk, 256, s, P, n, cur);
}

void trace_stage(Scene &sc, PathBufs &P, int cur, int grid, int sgrid, hipStream_t s, bool stats, bool shadow) {
    if (stats) {
        if (sc.has_spheres) trace_stage_t<true, true>(sc, P, cur, grid, sgrid, s, shadow);
        else trace_stage_t<true, false>(sc, P, cur, grid, sgrid, s, shadow);
    } else {
        if (sc.has_spheres) trace_stage_t<false, true>(sc, P, cur, grid, sgrid, s, shadow);
        else trace_stage_t<false, false>(sc, P, cur, grid, sgrid, s, shadow);
    }
}

void ensure_guide(Scene &sc, RenderCtx &rc, int rec_depth) {
    int cap = rc.capacity;
    if (rc.gc && rc.gc->capacity == cap && rc.gc->rec_depth >= rec_depth) return;
    if (rc.gc)
        for (auto &b : rc.gc->bufs)
            if (b.p) hipFree(b.p);
    rc.gc = std::make_unique<GuideCtx>();
    GuideCtx &g = *rc.gc;
    g.capacity = cap;
    g.rec_depth = rec_depth;
    size_t C = size_t(cap), R = size_t(rec_depth);
    g.glob = sc.alloc<float4>(C * guide::kLobes, g.bufs);
    g.gdir = sc.alloc<float4>(C, g.bufs);
    g.rec_feat = sc.alloc<float4>(2 * R * C, g.bufs);
    g.rec_dir = sc.alloc<float4>(R * C, g.bufs);
    g.rec_beta = sc.alloc<float4>(R * C, g.bufs);
    g.rec_c = sc.alloc<float4>(R * C, g.bufs);
    g.sample_cap = int(std::min<size_t>(R * C, size_t(1) << 24));
    g.s_feat = sc.alloc<float4>(2 * size_t(g.sample_cap), g.bufs);
    g.s_dir = sc.alloc<float4>(size_t(g.sample_cap), g.bufs);
    g.s_count = sc.alloc<int>(4, g.bufs);
    g.s_blk = sc.alloc<int>(C / 256 + 1, g.bufs);
}

void guide_infer(Scene &sc, PathBufs &P, int cur, int grid, hipStream_t s, const guide::Net &net) {
    const auto *ff = (const guide::bf16x8 *)net.ffrag;
    const auto *fb = (const float *)net.fbias;
    if (sc.has_spheres) launch(k_guide_infer<true>, grid, 512, s, sc.ds, P, cur, ff, fb);
    else launch(k_guide_infer<false>, grid, 512, s, sc.ds, P, cur, ff, fb);
}

// The auxiliary integrators: one launch over the pixels, no accumulation.
void aux_render(Scene &sc, const ljx_render_options &opt, const ljx_exec_options *ex, float *out_dev,
                ljx_render_stats *stats, std::chrono::steady_clock::time_point t_start) {
    ljx_exec_options e{};
    if (ex) e = *ex;
    HIPX(hipSetDevice(sc.device));
    ensure_ctx(sc, sc.rc ? sc.rc->capacity : 64, trace_grid_for(sc, e.trace_blocks_per_cu));
    PathBufs P = sc.rc->P;
    P.W = sc.width;
    P.H = sc.height;
    int grid = sc.rc->trace_grid;
    int mode = int(opt.integrator);
    if (sc.has_spheres) launch(k_aux<true>, grid, 256, sc.stream, sc.ds, P, mode, out_dev);
    else launch(k_aux<false>, grid, 256, sc.stream, sc.ds, P, mode, out_dev);
    HIPX(hipStreamSynchronize(sc.stream));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->camera_paths = uint64_t(sc.width) * sc.height;
        stats->extension_rays = stats->camera_paths;
        stats->build_ms = sc.build_ms;
        stats->bvh_depth = sc.bvh_depth;
        stats->node_bytes = sizeof(Bvh4QNode);
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    }
}

// The wavefront render loop.  With net set (path integrator only), surface
// directions come from the BSDF / guide mixture, and the first train_spp
// samples per pixel are recording passes, each followed by Adam steps.
void render_impl(Scene &sc, const ljx_render_options *o, const ljx_exec_options *ex, float *out_dev,
                 ljx_render_stats *stats, guide::Net *net = nullptr, const ljx_guide_options *go = nullptr) {
    auto t_start = std::chrono::steady_clock::now();
    ljx_render_options opt = o ? *o : sc.options;
    if (opt.integrator <= LJX_INTEGRATOR_MIPMAP_LEVEL) {
        if (net) fail(LJX_ERR_UNSUPPORTED, "guided rendering supports the path integrator only");
        aux_render(sc, opt, ex, out_dev, stats, t_start);
        return;
    }
    if (opt.samples_per_pixel <= 0) fail(LJX_ERR_INVALID_ARGUMENT, "samples_per_pixel must be positive");
    bool vol = opt.integrator == LJX_INTEGRATOR_VOLPATH;
    if (opt.integrator != LJX_INTEGRATOR_PATH && !vol)
        fail(LJX_ERR_UNSUPPORTED, "unknown integrator");
    if (vol && !kVolpathEnabled) fail(LJX_ERR_UNSUPPORTED, "volpath is not enabled in this build");
    if (vol && sc.ds.envmap_light_id >= 0)
        fail(LJX_ERR_UNSUPPORTED, "volpath assumes no environment map (handouts/homework2.tex:180)");
    if (opt.max_depth < -1) fail(LJX_ERR_INVALID_ARGUMENT, "max_depth must be >= -1");
    const bool guided = net != nullptr;
    ljx_guide_options gopt{};
    if (guided) {
        if (vol) fail(LJX_ERR_UNSUPPORTED, "guided rendering supports the path integrator only");
        if (net->device != sc.device) fail(LJX_ERR_INVALID_ARGUMENT, "guide and scene live on different devices");
        if (go) gopt = *go;
        else gopt.mix_prob = 0.5f;
        if (!(gopt.mix_prob >= 0.f && gopt.mix_prob < 1.f))
            fail(LJX_ERR_INVALID_ARGUMENT, "mix_prob must be in [0, 1)");
        if (gopt.train_spp < 0 || gopt.train_spp_per_pass < 0 || gopt.steps_per_pass < 0)
            fail(LJX_ERR_INVALID_ARGUMENT, "negative guide option");
        if (gopt.train_spp_per_pass == 0) gopt.train_spp_per_pass = 1;
        int md = opt.max_depth > 0 ? opt.max_depth : 16;
        gopt.max_record_depth = std::max(1, std::min(gopt.max_record_depth > 0 ? gopt.max_record_depth : md, 16));
        // shadow rays carry the index of their vertex clamped to the bits above the path slot
        static_assert((1 << (32 - kSlotBits)) - 1 >= 16, "a recorded vertex index must fit beside the path slot");
    }
    ljx_exec_options e{};
    if (ex) e = *ex;
    const int shade_block = e.block_size > 0 ? e.block_size : 256;
    if (shade_block % 64 || shade_block > 256) fail(LJX_ERR_INVALID_ARGUMENT, "block_size must be 64, 128, 192 or 256");
    HIPX(hipSetDevice(sc.device));
    hipStream_t s = sc.stream;
    int W = sc.width, H = sc.height;
    int ntx8 = (W + 7) / 8, nty8 = (H + 7) / 8;
    int npix = ntx8 * nty8 * 64;
    int spp = opt.samples_per_pixel;
    long long want = e.max_paths_in_flight > 0 ? e.max_paths_in_flight : kDefaultPathsInFlight;
    want = std::min<long long>(std::max<long long>(want, 64), 1 << kSlotBits);
    int capacity;
    int ns_per_pass, nq_per_pass;
    if (npix <= want) {
        ns_per_pass = int(std::min<long long>(spp, want / npix));
        nq_per_pass = npix;
    } else {
        ns_per_pass = 1;
        nq_per_pass = int((want / 64) * 64);
    }
    capacity = ns_per_pass * nq_per_pass;
    ensure_ctx(sc, std::max(capacity, sc.rc ? sc.rc->capacity : 0), trace_grid_for(sc, e.trace_blocks_per_cu));
    RenderCtx &rc = *sc.rc;
    PathBufs P = rc.P;
    P.W = W;
    P.H = H;
    P.ntx8 = ntx8;
    P.spp = spp;
    P.max_depth = opt.max_depth;
    P.rr_depth = opt.rr_depth;
    P.seed = opt.seed;
    P.volpath = vol ? 1 : 0;
    P.vol_flight = vol && sc.has_grid_media ? 1 : 0;
    P.max_null = opt.max_null_collisions;
    P.rec_depth = 0;
    P.jobs = nullptr;
    P.njobs = 0;
    P.work0 = 0;
    P.sort_threshold = e.sort_threshold < 0 ? INT_MAX : e.sort_threshold;
    GuideCtx *gc = nullptr;
    if (guided) {
        ensure_guide(sc, rc, gopt.max_record_depth);
        gc = rc.gc.get();
        P.glob = gc->glob;
        P.gdir = gc->gdir;
        P.rec_feat = gc->rec_feat;
        P.rec_dir = gc->rec_dir;
        P.rec_beta = gc->rec_beta;
        P.rec_c = gc->rec_c;
        P.guide_mix = gopt.mix_prob;
        P.guide_mask = 0;
        for (int t = 0; t < LJX_MAT_COUNT; t++)
            if (guidable(t)) P.guide_mask |= 1u << t;
        for (int k = 0; k < 3; k++) {
            float ext = sc.box_hi[k] - sc.box_lo[k];
            P.box_lo[k] = sc.box_lo[k];
            P.box_scale[k] = ext > 0 ? 1.f / ext : 0.f;
        }
    }
    const DeviceScene &S = sc.ds;
    int shade_grid = std::max(1, std::min(capacity / shade_block + 1, sc.num_cu * 16 * 256 / shade_block));
    int guide_grid = std::max(1, std::min(capacity / 512 + 1, sc.num_cu * 2));
    int tgrid = rc.trace_grid;
    Timer tm{e.collect_timing != 0, s};
    bool stats_on = e.collect_traversal != 0;

    hipEvent_t ev;
    tm.begin(3, ev);
    launch(k_fill_f64, 1024, 256, s, P.accum, size_t(W) * H * 3);
    launch(k_reset_counters, 1, 64, s, P.ctr, 0, C_TOTAL);
    hipLaunchKernelGGL(k_reset_stats, dim3(1), dim3(64), 0, s, P.stat);
    HIPX(hipGetLastError());
    tm.end(3, ev);

    uint64_t passes = 0, train_samples = 0, train_steps = 0, train_step_samples = 0;
    int max_bounces = opt.max_depth >= 0 ? std::max(opt.max_depth, 1) : 1 << 20;
    std::vector<int> host_ctr(C_TOTAL);
    const int train_spp = guided ? std::min(gopt.train_spp, spp) : 0;
    for (int s0 = 0; s0 < spp;) {
        const bool recording = s0 < train_spp;
        int ns = recording ? std::min({gopt.train_spp_per_pass, train_spp - s0, ns_per_pass})
                           : std::min(ns_per_pass, spp - s0);
        P.rec_depth = recording ? gopt.max_record_depth : 0;
        for (int q0 = 0; q0 < npix; q0 += nq_per_pass) {
            int nq = std::min(nq_per_pass, npix - q0);
            P.pass_s0 = s0;
            P.pass_ns = ns;
            P.pass_q0 = q0;
            P.pass_nq = nq;
            P.rec_stride = ns * nq;
            passes++;
            tm.begin(3, ev);
            launch(k_reset_counters, 1, 64, s, P.ctr, 0, C_TOTAL);
            launch(k_raygen, shade_grid, 256, s, S, P, (W % 8 || H % 8) ? 1 : 0);
            launch(k_alive_clear, shade_grid, 256, s, P, ns * nq);
            if (vol) launch(k_vol_init, shade_grid, 256, s, S, P);
            if (recording)
                launch(k_guide_clear, shade_grid, 256, s, P.rec_dir, P.rec_beta, P.rec_c,
                       size_t(P.rec_depth) * P.rec_stride);
            tm.end(3, ev);
            int cur = 0;
            for (int b = 0; b < max_bounces; b++) {
                if (opt.max_depth < 0 && b >= 8) {
                    // RR-only paths: stop once the extension queue drains
                    HIPX(hipMemcpyAsync(host_ctr.data(), P.ctr, sizeof(int) * C_TOTAL, hipMemcpyDeviceToHost, s));
                    HIPX(hipStreamSynchronize(s));
                    if (host_ctr[C_EXT + cur] == 0) break;
                }
                tm.begin(0, ev);
                trace_stage(sc, P, cur, tgrid, shade_grid, s, stats_on, false);
                tm.end(0, ev);
                if (guided) {
                    tm.begin(4, ev);
                    guide_infer(sc, P, cur, guide_grid, s, *net);
                    tm.end(4, ev);
                }
                tm.begin(2, ev);
                if (vol) vol_shade_all(S, P, cur, shade_grid, s, sc.has_spheres, (sc.mat_types_present & ~7u) == 0);
                else shade_all(sc, P, cur, shade_grid, s, guided, shade_block, /*fan=*/true);
                build_ext_queue(P, ns * nq, cur, s);
                tm.end(2, ev);
                tm.begin(1, ev);
                if (vol) vol_shadow(S, P, tgrid, s, stats_on, sc.has_spheres);
                else trace_stage(sc, P, cur, tgrid, shade_grid, s, stats_on, true);
                tm.end(1, ev);
                cur ^= 1;
            }
            tm.begin(3, ev);
            launch(k_accumulate, std::max(1, nq / 256 + 1), 256, s, P);
            tm.end(3, ev);
            if (recording) {
                // training samples of this pass, then Adam steps on them
                tm.begin(5, ev);
                const int cblk = (ns * nq + 255) / 256;
                launch(k_guide_count, cblk, 256, s, P, gc->s_blk);
                launch(k_guide_scan, 1, 1024, s, gc->s_blk, cblk, gc->s_count);
                launch(k_guide_collect, cblk, 256, s, P, gc->s_blk, gc->s_feat, gc->s_dir, gc->sample_cap);
                int cnt = 0;
                HIPX(hipMemcpyAsync(&cnt, gc->s_count, sizeof(int), hipMemcpyDeviceToHost, s));
                HIPX(hipStreamSynchronize(s));
                int navail = std::min(cnt, gc->sample_cap);
                train_samples += uint64_t(navail);
                for (int k = 0; k < gopt.steps_per_pass && navail > 0; k++) {
                    long long start = (static_cast<long long>(k) * net->batch_cap) % navail;
                    int cntk = int(std::min<long long>(net->batch_cap, navail - start));
                    guide::net_step(*net, gc->s_feat + 2 * start, gc->s_dir + start, cntk, true, s, nullptr);
                    train_steps++;
                    train_step_samples += uint64_t(cntk);
                }
                tm.end(5, ev);
            }
        }
        s0 += ns;
    }
    tm.begin(3, ev);
    size_t nout = size_t(W) * H * 3;
    launch(k_resolve, std::max<int>(1, int(std::min<size_t>(nout / 256 + 1, 4096))), 256, s, (const double *)P.accum,
           out_dev, nout, 1.0 / double(spp));
    tm.end(3, ev);
    unsigned long long st[S_NUM];
    HIPX(hipMemcpyAsync(st, P.stat, sizeof(st), hipMemcpyDeviceToHost, s));
    HIPX(hipStreamSynchronize(s));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->camera_paths = uint64_t(W) * H * spp;
        stats->extension_rays = st[S_EXT_RAYS];
        stats->shadow_rays = st[S_SH_RAYS];
        stats->passes = passes;
        stats->extend_launches = tm.ev[0].size();
        stats->shadow_launches = tm.ev[1].size();
        stats->extend_ms = tm.total(0);
        stats->shadow_ms = tm.total(1);
        stats->shade_ms = tm.total(2);
        stats->other_ms = tm.total(3);
        stats->guide_ms = tm.total(4);
        stats->train_ms = tm.total(5);
        stats->train_samples = train_samples;
        stats->train_steps = train_steps;
        stats->train_step_samples = train_step_samples;
        stats->guide_queries = st[S_GUIDE_Q];
        stats->node_visits = st[S_NODES];
        stats->tri_tests = st[S_PRIMS];
        stats->shadow_node_visits = st[S_SH_NODES];
        stats->shadow_tri_tests = st[S_SH_PRIMS];
        stats->instrumented_rays = stats_on ? st[S_EXT_RAYS] + st[S_SH_RAYS] : 0;
        stats->build_ms = sc.build_ms;
        stats->bvh_depth = sc.bvh_depth;
        stats->node_bytes = sizeof(Bvh4QNode);  // corner and scale, three half-box vectors, child references
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    } else {
        for (int k = 0; k < 6; k++) tm.total(k);
    }
}

// ------------------------------------------------------------- autotuner
// Times short renders of the scene's own camera (about 8M camera paths each)
// for candidate settings and keeps the fastest: paths in flight x shading
// block size, then the sort threshold at the best pair.
ljx_exec_options autotune_impl(Scene &sc, const ljx_render_options *o) {
    ljx_render_options opt = o ? *o : sc.options;
    if (opt.max_depth < 0) opt.max_depth = 8;
    const long long npix = (long long)sc.width * sc.height;
    // ~32M samples: enough that pass sizes up to 32M paths are told apart
    opt.samples_per_pixel = int(std::max<long long>(1, std::min<long long>(1 << 12, (32ll << 20) / npix)));
    float *out = nullptr;
    HIPX(hipMalloc(&out, size_t(npix) * 3 * sizeof(float)));
    struct Free {
        void *p;
        ~Free() { hipFree(p); }
    } f{out};
    auto trial = [&](const ljx_exec_options &e) {
        ljx_render_stats st;
        render_impl(sc, &opt, &e, out, &st);  // warm (allocation, first launch)
        render_impl(sc, &opt, &e, out, &st);
        return st.total_ms;
    };
    ljx_exec_options best{};
    double best_ms = 1e300;
    // larger passes first; a smaller one must be 2% faster to be taken (the
    // batches these settings serve are far larger than the trial)
    for (int cap : {1 << 26, 1 << 25, 1 << 24, 1 << 23, 1 << 22})
        for (int blk : {256, 128}) {
            ljx_exec_options e{};
            e.max_paths_in_flight = cap;
            e.block_size = blk;
            double ms = trial(e);
            if (ms < 0.98 * best_ms) best_ms = ms, best = e;
        }
    for (int thr : {4096, 65536}) {
        ljx_exec_options e = best;
        e.sort_threshold = thr;
        double ms = trial(e);
        if (ms < best_ms) best_ms = ms, best = e;
    }
    for (int bpc : {4, 6}) {
        ljx_exec_options e = best;
        e.trace_blocks_per_cu = bpc;
        double ms = trial(e);
        if (ms < best_ms) best_ms = ms, best = e;
    }
    sc.tuned = true;
    sc.tuned_exec = best;
    return best;
}

// ------------------------------------------------------------ batch render
// Device wall-clock stamp (the constant-rate real-time counter), written by
// one lane; stamps inside a captured graph time the jobs of a batch.
__global__ void k_stamp(unsigned long long *dst) {
    if (threadIdx.x == 0) dst[0] = wall_clock64();
}

void batch_impl(const ljx_batch_job *jobs, size_t njobs, const ljx_exec_options *ex, double *per_job_ms,
                ljx_batch_stats *bst) {
    auto t0 = std::chrono::steady_clock::now();
    if (njobs == 0) {
        if (bst) std::memset(bst, 0, sizeof(*bst));
        return;
    }
    if (!jobs) fail(LJX_ERR_INVALID_ARGUMENT, "jobs is NULL");
    // group by scene, keeping first-appearance order
    std::vector<std::pair<Scene *, std::vector<int>>> groups;
    for (size_t i = 0; i < njobs; i++) {
        const ljx_batch_job &J = jobs[i];
        Scene *sc = as_scene(J.scene);
        if (J.samples_per_pixel <= 0) fail(LJX_ERR_INVALID_ARGUMENT, "job " + std::to_string(i) + ": spp must be positive");
        if (J.max_depth < 0) fail(LJX_ERR_INVALID_ARGUMENT, "job " + std::to_string(i) + ": max_depth must be >= 0");
        if (!J.rgb_out_device) fail(LJX_ERR_INVALID_ARGUMENT, "job " + std::to_string(i) + ": output is NULL");
        if (J.camera.medium_id < -1 || J.camera.medium_id >= sc->num_media)
            fail(LJX_ERR_INVALID_ARGUMENT, "job " + std::to_string(i) + ": camera medium id out of range");
        make_camera(J.camera);  // validates
        auto it = std::find_if(groups.begin(), groups.end(), [&](auto &g) { return g.first == sc; });
        if (it == groups.end()) groups.push_back({sc, {int(i)}});
        else it->second.push_back(int(i));
    }
    const int dev = groups[0].first->device;
    for (auto &g : groups) {
        if (g.first->device != dev) fail(LJX_ERR_INVALID_ARGUMENT, "all scenes of a batch must be on one device");
        for (int j : g.second)
            if (jobs[j].max_depth != jobs[g.second[0]].max_depth)
                fail(LJX_ERR_INVALID_ARGUMENT, "jobs of one scene must share max_depth");
        const ljx_render_options &so = g.first->options;
        if (so.integrator != LJX_INTEGRATOR_PATH && so.integrator != LJX_INTEGRATOR_VOLPATH)
            fail(LJX_ERR_UNSUPPORTED, "batched jobs need the path or volpath integrator");
        if (so.integrator == LJX_INTEGRATOR_VOLPATH && g.first->ds.envmap_light_id >= 0)
            fail(LJX_ERR_UNSUPPORTED, "volpath assumes no environment map (handouts/homework2.tex:180)");
    }
    HIPX(hipSetDevice(dev));
    // launch settings per scene
    std::vector<ljx_exec_options> gex(groups.size());
    for (size_t k = 0; k < groups.size(); k++) {
        Scene &sc = *groups[k].first;
        if (ex && ex->max_paths_in_flight > 0) gex[k] = *ex;
        else {
            if (!sc.tuned) autotune_impl(sc, nullptr);
            gex[k] = sc.tuned_exec;
        }
        gex[k].collect_timing = 0;
        gex[k].collect_traversal = 0;
    }
    // device job tables and accumulators
    struct Buf {
        void *p = nullptr;
        ~Buf() {
            if (p) hipFree(p);
        }
    };
    std::vector<std::unique_ptr<Buf>> keep;
    auto dmalloc = [&](size_t bytes) {
        keep.push_back(std::make_unique<Buf>());
        HIPX(hipMalloc(&keep.back()->p, std::max<size_t>(bytes, 16)));
        return keep.back()->p;
    };
    struct GroupPlan {
        std::vector<int> order;         // job indices, smallest work first
        std::vector<BatchJob> table;
        BatchJob *dtable = nullptr;
        long long total = 0;
        int capacity = 0, block = 256;
    };
    std::vector<GroupPlan> plans(groups.size());
    uint64_t total_samples = 0;
    for (size_t k = 0; k < groups.size(); k++) {
        GroupPlan &gp = plans[k];
        gp.order = groups[k].second;
        auto work = [&](int j) {
            const ljx_camera_desc &c = jobs[j].camera;
            return (long long)((c.width + 7) / 8) * ((c.height + 7) / 8) * 64 * jobs[j].samples_per_pixel;
        };
        std::stable_sort(gp.order.begin(), gp.order.end(), [&](int a, int b) { return work(a) < work(b); });
        for (int j : gp.order) {
            const ljx_batch_job &J = jobs[j];
            BatchJob b{};
            b.camera = make_camera(J.camera);
            b.W = J.camera.width;
            b.H = J.camera.height;
            b.ntx8 = (b.W + 7) / 8;
            b.npix8 = b.ntx8 * ((b.H + 7) / 8) * 64;
            b.spp = J.samples_per_pixel;
            b.seed = J.seed;
            b.work_base = gp.total;
            b.accum = static_cast<double *>(dmalloc(size_t(b.W) * b.H * 3 * sizeof(double)));
            b.out = J.rgb_out_device;
            gp.total += work(j);
            total_samples += uint64_t(b.W) * b.H * b.spp;
            gp.table.push_back(b);
        }
        gp.dtable = static_cast<BatchJob *>(dmalloc(gp.table.size() * sizeof(BatchJob)));
        HIPX(hipMemcpy(gp.dtable, gp.table.data(), gp.table.size() * sizeof(BatchJob), hipMemcpyHostToDevice));
        long long want = gex[k].max_paths_in_flight > 0 ? gex[k].max_paths_in_flight : kDefaultPathsInFlight;
        want = std::min<long long>(std::max<long long>(want, 64), 1 << kSlotBits);
        gp.capacity = int(std::min<long long>(want, gp.total));
        gp.block = gex[k].block_size > 0 ? gex[k].block_size : 256;
        if (gp.block % 64 || gp.block > 256) fail(LJX_ERR_INVALID_ARGUMENT, "block_size must be 64, 128, 192 or 256");
    }
    hipStream_t cs = nullptr;
    HIPX(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { hipStreamDestroy(s); }
    } sg{cs};
    // stamps: [0] graph start, [1 + i] completion of job i
    auto *stamps = static_cast<unsigned long long *>(dmalloc((njobs + 1) * sizeof(unsigned long long)));
    int clock_khz = 0;
    HIPX(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, dev));
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    HIPX(hipEventCreate(&ev_start));
    HIPX(hipEventCreate(&ev_stop));
    struct EvGuard {
        hipEvent_t a, b;
        ~EvGuard() {
            hipEventDestroy(a);
            hipEventDestroy(b);
        }
    } eg{ev_start, ev_stop};
    // contexts first: allocation is not allowed during capture
    for (size_t k = 0; k < groups.size(); k++)
        ensure_ctx(*groups[k].first, plans[k].capacity, trace_grid_for(*groups[k].first, gex[k].trace_blocks_per_cu));
    uint64_t passes = 0;
    hipGraph_t graph = nullptr;
    HIPX(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    try {
        launch(k_stamp, 1, 64, cs, stamps);
        for (size_t k = 0; k < groups.size(); k++) {
            Scene &sc = *groups[k].first;
            GroupPlan &gp = plans[k];
            const int J0 = gp.order[0];
            const bool vol = sc.options.integrator == LJX_INTEGRATOR_VOLPATH;
            PathBufs P = sc.rc->P;
            P.spp = 0;
            P.max_depth = jobs[J0].max_depth;
            P.rr_depth = sc.options.rr_depth;
            P.seed = 0;
            P.volpath = vol ? 1 : 0;
            P.vol_flight = vol && sc.has_grid_media ? 1 : 0;
            P.max_null = sc.options.max_null_collisions;
            P.rec_depth = 0;
            P.jobs = gp.dtable;
            P.njobs = int(gp.table.size());
            P.sort_threshold = gex[k].sort_threshold < 0 ? INT_MAX : gex[k].sort_threshold;
            P.W = P.H = P.ntx8 = 0;
            const DeviceScene &S = sc.ds;
            const int cap = gp.capacity;
            const int sgrid = std::max(1, std::min(cap / gp.block + 1, sc.num_cu * 16 * 256 / gp.block));
            const int tgrid = sc.rc->trace_grid;
            for (int q = 0; q < P.njobs; q++) launch(k_job_clear, 256, 256, cs, (const BatchJob *)gp.dtable, q);
            size_t next_done = 0;  // jobs complete in work order
            for (long long w0 = 0; w0 < gp.total; w0 += cap) {
                int n = int(std::min<long long>(cap, gp.total - w0));
                P.work0 = w0;
                P.pass_s0 = 0;
                P.pass_ns = 1;
                P.pass_q0 = 0;
                P.pass_nq = n;
                passes++;
                launch(k_reset_counters, 1, 64, cs, P.ctr, 0, C_TOTAL);
                launch(k_raygen, sgrid, 256, cs, S, P, 1);
                launch(k_alive_clear, sgrid, 256, cs, P, n);
                if (vol) launch(k_vol_init, sgrid, 256, cs, S, P);
                int cur = 0;
                for (int b = 0; b < std::max(P.max_depth, 1); b++) {
                    trace_stage(sc, P, cur, tgrid, sgrid, cs, false, false);
                    if (vol) vol_shade_all(S, P, cur, sgrid, cs, sc.has_spheres, (sc.mat_types_present & ~7u) == 0);
                    else shade_all(sc, P, cur, sgrid, cs, false, gp.block);
                    build_ext_queue(P, n, cur, cs);
                    if (vol) vol_shadow(S, P, tgrid, cs, false, sc.has_spheres);
                    else trace_stage(sc, P, cur, tgrid, sgrid, cs, false, true);
                    cur ^= 1;
                }
                launch(k_accumulate_batch, std::max(1, std::min(n / 256 + 1, sc.num_cu * 16)), 256, cs, P, n);
                while (next_done < gp.table.size()) {
                    const BatchJob &bj = gp.table[next_done];
                    long long end = bj.work_base + (long long)bj.npix8 * bj.spp;
                    if (end > w0 + n) break;
                    launch(k_job_resolve, 256, 256, cs, (const BatchJob *)gp.dtable, int(next_done));
                    launch(k_stamp, 1, 64, cs, stamps + 1 + gp.order[next_done]);
                    next_done++;
                }
            }
        }
    } catch (...) {
        hipGraph_t g = nullptr;
        hipStreamEndCapture(cs, &g);
        if (g) hipGraphDestroy(g);
        throw;
    }
    HIPX(hipStreamEndCapture(cs, &graph));
    struct GraphGuard {
        hipGraph_t g;
        hipGraphExec_t x = nullptr;
        ~GraphGuard() {
            if (x) hipGraphExecDestroy(x);
            if (g) hipGraphDestroy(g);
        }
    } gg{graph};
    size_t nodes = 0;
    HIPX(hipGraphGetNodes(graph, nullptr, &nodes));
    HIPX(hipGraphInstantiate(&gg.x, graph, nullptr, nullptr, 0));
    auto t_built = std::chrono::steady_clock::now();
    HIPX(hipEventRecord(ev_start, cs));
    HIPX(hipGraphLaunch(gg.x, cs));
    HIPX(hipEventRecord(ev_stop, cs));
    HIPX(hipStreamSynchronize(cs));
    float launch_ms = 0;
    HIPX(hipEventElapsedTime(&launch_ms, ev_start, ev_stop));
    if (per_job_ms) {
        std::vector<unsigned long long> st(njobs + 1);
        HIPX(hipMemcpy(st.data(), stamps, st.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < njobs; i++)
            per_job_ms[i] = clock_khz > 0 ? double(st[1 + i] - st[0]) / double(clock_khz) : -1.0;
    }
    if (bst) {
        std::memset(bst, 0, sizeof(*bst));
        bst->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        bst->build_ms = std::chrono::duration<double, std::milli>(t_built - t0).count();
        bst->launch_ms = launch_ms;
        bst->samples = total_samples;
        bst->passes = passes;
        bst->graph_nodes = nodes;
        bst->num_scenes = int(groups.size());
    }
}

template <bool ANY>
void trace_device(Scene &sc, const ljx_ray *rays, size_t n, void *out, const ljx_exec_options *ex, double *ms,
                  ljx_render_stats *stats) {
    HIPX(hipSetDevice(sc.device));
    if (n == 0) {
        if (ms) *ms = 0;
        return;
    }
    int grid = trace_grid_for(sc);
    SpillBuf spill;
    HIPX(hipMalloc(&spill.p, size_t(grid) * 256 * sc.spill_entries * sizeof(int)));
    unsigned long long *stat = nullptr;
    HIPX(hipMalloc(&stat, sizeof(unsigned long long) * S_NUM));
    struct Free {
        void *p;
        ~Free() { hipFree(p); }
    } fstat{stat};
    HIPX(hipMemsetAsync(stat, 0, sizeof(unsigned long long) * S_NUM, sc.stream));
    int st_on = ex && ex->collect_traversal ? 1 : 0;
    hipEvent_t a, b;
    HIPX(hipEventCreate(&a));
    HIPX(hipEventCreate(&b));
    HIPX(hipEventRecord(a, sc.stream));
    auto kern = sc.has_spheres ? k_trace_batch<ANY, true> : k_trace_batch<ANY, false>;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, sc.stream, sc.ds, rays, n,
                       ANY ? nullptr : static_cast<ljx_hit *>(out), ANY ? static_cast<uint8_t *>(out) : nullptr,
                       spill.p, stat, st_on);
    HIPX(hipGetLastError());
    HIPX(hipEventRecord(b, sc.stream));
    HIPX(hipEventSynchronize(b));
    float el = 0;
    HIPX(hipEventElapsedTime(&el, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (ms) *ms = el;
    if (stats) {
        unsigned long long h[S_NUM];
        HIPX(hipMemcpy(h, stat, sizeof(h), hipMemcpyDeviceToHost));
        std::memset(stats, 0, sizeof(*stats));
        if (ANY) {
            stats->shadow_rays = n;
            stats->shadow_node_visits = h[S_NODES];
            stats->shadow_tri_tests = h[S_PRIMS];
        } else {
            stats->extension_rays = n;
            stats->node_visits = h[S_NODES];
            stats->tri_tests = h[S_PRIMS];
        }
        stats->build_ms = sc.build_ms;
        stats->bvh_depth = sc.bvh_depth;
        stats->node_bytes = sizeof(Bvh4QNode);  // corner and scale, three half-box vectors, child references
        stats->instrumented_rays = st_on ? n : 0;
    }
}

}  // namespace

extern "C" {

int ljx_abi_version(void) { return LJX_ABI_VERSION; }

int ljx_last_error(char *buf, size_t cap) {
    if (!buf || cap == 0) return LJX_ERR_INVALID_ARGUMENT;
    std::snprintf(buf, cap, "%s", g_last_error.c_str());
    return LJX_OK;
}

int ljx_device_count(int *count) {
    return guarded([&]() {
        if (!count) fail(LJX_ERR_INVALID_ARGUMENT, "count is NULL");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        int ok = 0;
        for (int i = 0; i < n; i++) {
            hipDeviceProp_t p;
            if (hipGetDeviceProperties(&p, i) == hipSuccess && std::strncmp(p.gcnArchName, "gfx950", 6) == 0) ok++;
        }
        *count = ok;
    });
}

int ljx_scene_create(const ljx_scene_desc *desc, int device, void **scene_out) {
    return ljx_scene_create_ex(desc, device, nullptr, scene_out);
}

int ljx_scene_create_ex(const ljx_scene_desc *desc, int device, const ljx_build_options *build, void **scene_out) {
    return guarded([&]() {
        if (!scene_out) fail(LJX_ERR_INVALID_ARGUMENT, "scene_out is NULL");
        *scene_out = nullptr;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) fail(LJX_ERR_NO_DEVICE, "no HIP device");
        if (device < 0 || device >= n) fail(LJX_ERR_INVALID_ARGUMENT, "device index out of range");
        HIPX(hipSetDevice(device));
        auto sc = std::make_unique<Scene>();
        sc->device = device;
        hipDeviceProp_t prop;
        HIPX(hipGetDeviceProperties(&prop, device));
        sc->num_cu = prop.multiProcessorCount;
        HIPX(hipStreamCreateWithFlags(&sc->stream, hipStreamNonBlocking));
        if (build) {
            if (build->bvh_builder != LJX_BVH_BUILDER_GPU && build->bvh_builder != LJX_BVH_BUILDER_HOST)
                fail(LJX_ERR_INVALID_ARGUMENT, "unknown BVH builder");
            if (build->max_leaf_size < 0 || build->max_leaf_size > 8)
                fail(LJX_ERR_INVALID_ARGUMENT, "max_leaf_size must be in [0, 8]");
            sc->bvh_builder = build->bvh_builder;
            if (build->max_leaf_size > 0) sc->max_leaf = build->max_leaf_size;
            if (build->node_cost < 0 || !(build->node_cost < 64)) fail(LJX_ERR_INVALID_ARGUMENT, "node_cost out of range");
            if (build->node_cost > 0) sc->node_cost = build->node_cost;
        }
        build_scene(*sc, desc);
        *scene_out = sc.release();
    });
}

int ljx_scene_bvh_info(void *scene, ljx_bvh_info *info) {
    return guarded([&]() {
        Scene *sc = as_scene(scene);
        if (!info) fail(LJX_ERR_INVALID_ARGUMENT, "info is NULL");
        HIPX(hipSetDevice(sc->device));
        std::memset(info, 0, sizeof(*info));
        info->builder = sc->bvh_builder;
        info->depth = sc->bvh_depth;
        info->stack_need = sc->stack_need;
        info->binary_depth = sc->binary_depth;
        info->num_nodes = uint64_t(sc->ds.num_nodes);
        info->num_slots = uint64_t(sc->num_slots);
        info->bvh_ms = sc->bvh_ms;
        info->build_ms = sc->build_ms;
        std::vector<Bvh4Node> nodes(sc->ds.num_nodes);
        std::vector<int> slots(sc->num_slots);
        if (!nodes.empty())
            HIPX(hipMemcpy(nodes.data(), sc->ds.nodes, nodes.size() * sizeof(Bvh4Node), hipMemcpyDeviceToHost));
        if (!slots.empty())
            HIPX(hipMemcpy(slots.data(), sc->slot_prim, slots.size() * sizeof(int), hipMemcpyDeviceToHost));
        auto fnv = [](const void *p, size_t n) {
            uint64_t h = 1469598103934665603ull;
            const unsigned char *c = static_cast<const unsigned char *>(p);
            for (size_t i = 0; i < n; i++) h = (h ^ c[i]) * 1099511628211ull;
            return h;
        };
        info->node_hash = fnv(nodes.data(), nodes.size() * sizeof(Bvh4Node));
        info->slot_hash = fnv(slots.data(), slots.size() * sizeof(int));
        // SAH cost with unit node and primitive costs: every node (by the
        // area of its box as its parent stores it, relative to the root's)
        // plus every leaf's primitive count times its relative area
        if (!nodes.empty()) {
            auto area = [](double lx, double hx, double ly, double hy, double lz, double hz) {
                double a = hx - lx, b = hy - ly, c = hz - lz;
                return a * b + b * c + c * a;
            };
            auto child = [&](const Bvh4Node &nd, int c, double &a) {
                const float *lox = &nd.lox.x, *hix = &nd.hix.x, *loy = &nd.loy.x, *hiy = &nd.hiy.x,
                            *loz = &nd.loz.x, *hiz = &nd.hiz.x;
                a = area(lox[c], hix[c], loy[c], hiy[c], loz[c], hiz[c]);
                return (&nd.ch.x)[c];
            };
            double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            const Bvh4Node &r = nodes[0];
            for (int c = 0; c < r.meta.x; c++) {
                lo[0] = std::min<double>(lo[0], (&r.lox.x)[c]), hi[0] = std::max<double>(hi[0], (&r.hix.x)[c]);
                lo[1] = std::min<double>(lo[1], (&r.loy.x)[c]), hi[1] = std::max<double>(hi[1], (&r.hiy.x)[c]);
                lo[2] = std::min<double>(lo[2], (&r.loz.x)[c]), hi[2] = std::max<double>(hi[2], (&r.hiz.x)[c]);
            }
            const double root = area(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]);
            double cost = 1;
            for (const Bvh4Node &nd : nodes)
                for (int c = 0; c < nd.meta.x; c++) {
                    double a;
                    int ref = child(nd, c, a);
                    cost += (ref >= 0 ? 1.0 : double((~ref & 7) + 1)) * (root > 0 ? a / root : 1.0);
                }
            info->sah_cost = cost;
        }
        // the traversal's quantised copy: its fingerprint, and every box plane
        // checked against the FP32 plane it stands for
        std::vector<Bvh4QNode> qn(sc->ds.num_nodes);
        if (!qn.empty())
            HIPX(hipMemcpy(qn.data(), sc->ds.qnodes, qn.size() * sizeof(Bvh4QNode), hipMemcpyDeviceToHost));
        info->qnode_hash = fnv(qn.data(), qn.size() * sizeof(Bvh4QNode));
        uint64_t bad = 0;
        double loose = 0;
        for (size_t i = 0; i < qn.size(); i++) {
            const Bvh4Node &nd = nodes[i];
            const Bvh4QNode &q = qn[i];
            const float *lo[3] = {&nd.lox.x, &nd.loy.x, &nd.loz.x}, *hi[3] = {&nd.hix.x, &nd.hiy.x, &nd.hiz.x};
            const uint32_t *qa[3] = {&q.qx.x, &q.qy.x, &q.qz.x};
            const double org[3] = {q.org.x, q.org.y, q.org.z};
            double bmin[3] = {INFINITY, INFINITY, INFINITY}, bmax[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int c = 0; c < 4; c++)
                if ((&nd.ch.x)[c] != kNoChild)
                    for (int a = 0; a < 3; a++) bmin[a] = std::min<double>(bmin[a], lo[a][c]), bmax[a] = std::max<double>(bmax[a], hi[a][c]);
            const double ext = std::max({bmax[0] - bmin[0], bmax[1] - bmin[1], bmax[2] - bmin[2]});
            if ((&q.ch.x)[0] != nd.ch.x || q.ch.y != nd.ch.y || q.ch.z != nd.ch.z || q.ch.w != nd.ch.w) bad++;
            for (int c = 0; c < 4; c++) {
                if ((&nd.ch.x)[c] == kNoChild) continue;
                for (int a = 0; a < 3; a++) {
                    const double L = org[a] + half_value((qa[a][c >> 1] >> (16 * (c & 1))) & 0xffffu) * q.org.w;
                    const double H = org[a] + half_value((qa[a][2 + (c >> 1)] >> (16 * (c & 1))) & 0xffffu) * q.org.w;
                    if (!(L <= lo[a][c])) bad++;
                    if (!(H >= hi[a][c])) bad++;
                    if (ext > 0) loose = std::max(loose, std::max(lo[a][c] - L, H - hi[a][c]) / ext);
                }
            }
        }
        info->qnode_bad_planes = bad;
        info->qnode_max_loose = loose;
    });
}

// Random node fetches over the scene's 4-wide BVH: each of `lanes` lanes
// of every wave reads whole nodes (the traversal's five 16-byte loads) at
// hashed indices, the next index depending on the data just loaded, on the
// traversal's grid.  This is the traversal's memory access pattern with no
// coherence between the lanes of a wave and no arithmetic in the way (a
// traversal of rays that share no nodes, at the vector-memory pipeline's
// rate).
__device__ __forceinline__ uint32_t fetch_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__global__ __launch_bounds__(256) void k_node_fetch(const Bvh4QNode *__restrict__ nodes, uint32_t n, int iters,
                                                    int lanes, float *sink) {
    uint32_t h = fetch_hash(blockIdx.x * 256u + threadIdx.x);
    float acc = 0.f;
    if (int(threadIdx.x & 63) < lanes) {
        for (int it = 0; it < iters; it++) {
            h = fetch_hash(h + uint32_t(it));
            const Bvh4QNode *nd = nodes + h % n;
            const float4 org = nd->org;
            const uint4 qx = nd->qx, qy = nd->qy, qz = nd->qz;
            const int4 ch = nd->ch;
            acc += org.x + float(qx.y & 1) + float(qy.z & 1) + float(qz.w & 1) + float(ch.z & 1);
            h ^= __float_as_uint(acc) & 1u;
        }
    }
    if (acc == 1234.5f) sink[0] = acc;  // keeps the loads; the sum never matters
}

int ljx_scene_node_fetch_rate(void *scene, int lanes, double *fetches_per_s) {
    return guarded([&]() {
        Scene *sc = as_scene(scene);
        if (!fetches_per_s) fail(LJX_ERR_INVALID_ARGUMENT, "fetches_per_s is NULL");
        if (lanes < 1 || lanes > 64) fail(LJX_ERR_INVALID_ARGUMENT, "lanes must be in 1..64");
        if (sc->ds.num_nodes <= 0) fail(LJX_ERR_INVALID_ARGUMENT, "the scene has no BVH nodes");
        HIPX(hipSetDevice(sc->device));
        float *sink = nullptr;
        HIPX(hipMalloc(&sink, sizeof(float)));
        hipEvent_t e0, e1;
        HIPX(hipEventCreate(&e0));
        HIPX(hipEventCreate(&e1));
        const int grid = trace_grid_for(*sc), iters = 64;
        float best = 0;
        for (int rep = 0; rep < 5; rep++) {
            HIPX(hipEventRecord(e0, 0));
            hipLaunchKernelGGL(k_node_fetch, dim3(grid), dim3(256), 0, 0, sc->ds.qnodes, uint32_t(sc->ds.num_nodes),
                               iters, lanes, sink);
            HIPX(hipGetLastError());
            HIPX(hipEventRecord(e1, 0));
            HIPX(hipEventSynchronize(e1));
            float ms = 0;
            HIPX(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0 && (best == 0 || ms < best)) best = ms;  // the first launch warms up
        }
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        HIPX(hipFree(sink));
        const double fetches = double(grid) * 4 * lanes * iters;
        *fetches_per_s = best > 0 ? fetches / (best * 1e-3) : 0.0;
    });
}

int ljx_scene_destroy(void *scene) {
    return guarded([&]() { delete static_cast<Scene *>(scene); });
}

int ljx_scene_bounds(void *scene, double *cr) {
    return guarded([&]() {
        Scene *sc = as_scene(scene);
        if (!cr) fail(LJX_ERR_INVALID_ARGUMENT, "output is NULL");
        cr[0] = sc->bound_center[0];
        cr[1] = sc->bound_center[1];
        cr[2] = sc->bound_center[2];
        cr[3] = sc->bound_radius;
    });
}

int ljx_render_device(void *scene, const ljx_render_options *opts, const ljx_exec_options *exec, float *rgb_out_device,
                      ljx_render_stats *stats) {
    return guarded([&]() {
        Scene *sc = as_scene(scene);
        if (!rgb_out_device) fail(LJX_ERR_INVALID_ARGUMENT, "output is NULL");
        render_impl(*sc, opts, exec, rgb_out_device, stats);
    });
}

int ljx_render(void *scene, const ljx_render_options *opts, const ljx_exec_options *exec, float *rgb_out,
               ljx_render_stats *stats) {
    return guarded([&]() {
        Scene *sc = as_scene(scene);
        if (!rgb_out) fail(LJX_ERR_INVALID_ARGUMENT, "output is NULL");
        HIPX(hipSetDevice(sc->device));
        size_t bytes = size_t(sc->width) * sc->height * 3 * sizeof(float);
        float *d = nullptr;
        HIPX(hipMalloc(&d, bytes));
        struct Free {
            void *p;
            ~Free() { hipFree(p); }
        } f{d};
        render_impl(*sc, opts, exec, d, stats);
        HIPX(hipMemcpy(rgb_out, d, bytes, hipMemcpyDeviceToHost));
    });
}

int ljx_render_guided_device(void *scene, void *guide_h, const ljx_render_options *opts, const ljx_exec_options *exec,
                             const ljx_guide_options *gopts, float *rgb_out_device, ljx_render_stats *stats) {
    return guarded([&]() {
        Scene *sc = as_scene(scene);
        if (!guide_h) fail(LJX_ERR_INVALID_ARGUMENT, "guide handle is NULL");
        if (!rgb_out_device) fail(LJX_ERR_INVALID_ARGUMENT, "output is NULL");
        render_impl(*sc, opts, exec, rgb_out_device, stats, static_cast<guide::Net *>(guide_h), gopts);
    });
}

int ljx_render_batch_device(const ljx_batch_job *jobs, size_t njobs, const ljx_exec_options *exec,
                            double *per_job_ms, ljx_batch_stats *stats) {
    return guarded([&]() { batch_impl(jobs, njobs, exec, per_job_ms, stats); });
}

int ljx_autotune(void *scene, const ljx_render_options *opts, ljx_exec_options *best) {
    return guarded([&]() {
        Scene *sc = as_scene(scene);
        HIPX(hipSetDevice(sc->device));
        ljx_exec_options b = autotune_impl(*sc, opts);
        if (best) *best = b;
    });
}

int ljx_trace_closest_device(void *scene, const ljx_ray *rays, size_t n, ljx_hit *hits, const ljx_exec_options *exec,
                             double *elapsed_ms, ljx_render_stats *stats) {
    return guarded([&]() {
        Scene *sc = as_scene(scene);
        if (n && (!rays || !hits)) fail(LJX_ERR_INVALID_ARGUMENT, "NULL buffer");
        trace_device<false>(*sc, rays, n, hits, exec, elapsed_ms, stats);
    });
}

int ljx_trace_closest_device_noalloc(void *scene, const ljx_ray *rays, size_t n, ljx_hit *hits) {
    return guarded([&]() {
        Scene *sc = as_scene(scene);
        if (n && (!rays || !hits)) fail(LJX_ERR_INVALID_ARGUMENT, "NULL buffer");
        HIPX(hipSetDevice(sc->device));
        int *spill = scratch_on(sc->device);
        if (n == 0) return;
        const size_t per_block = size_t(256) * size_t(sc->spill_entries);
        if (per_block > kScratchInts)
            fail(LJX_ERR_UNSUPPORTED, "the scene's traversal needs " + std::to_string(per_block) +
                                          " stack-spill entries per block, more than the " +
                                          std::to_string(kScratchInts) + " of the allocation-free scratch");
        // the same persistent grid as ljx_trace_closest_device, capped by what
        // the scratch holds; rays do not depend on the grid
        const int grid = int(std::min<size_t>(size_t(trace_grid_for(*sc)), kScratchInts / per_block));
        auto kern = sc->has_spheres ? k_trace_batch<false, true> : k_trace_batch<false, false>;
        hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, sc->stream, sc->ds, rays, n, hits, nullptr, spill, nullptr,
                           0);
        HIPX(hipGetLastError());
        HIPX(hipStreamSynchronize(sc->stream));
    });
}

int ljx_trace_any_device(void *scene, const ljx_ray *rays, size_t n, uint8_t *occ, const ljx_exec_options *exec,
                         double *elapsed_ms, ljx_render_stats *stats) {
    return guarded([&]() {
        Scene *sc = as_scene(scene);
        if (n && (!rays || !occ)) fail(LJX_ERR_INVALID_ARGUMENT, "NULL buffer");
        trace_device<true>(*sc, rays, n, occ, exec, elapsed_ms, stats);
    });
}

int ljx_trace_closest(void *scene, const ljx_ray *rays, size_t n, ljx_hit *hits) {
    return guarded([&]() {
        Scene *sc = as_scene(scene);
        if (n == 0) return;
        if (!rays || !hits) fail(LJX_ERR_INVALID_ARGUMENT, "NULL buffer");
        HIPX(hipSetDevice(sc->device));
        void *dr = nullptr, *dh = nullptr;
        HIPX(hipMalloc(&dr, n * sizeof(ljx_ray)));
        struct Free {
            void *p;
            ~Free() {
                if (p) hipFree(p);
            }
        } f1{dr};
        HIPX(hipMalloc(&dh, n * sizeof(ljx_hit)));
        Free f2{dh};
        HIPX(hipMemcpy(dr, rays, n * sizeof(ljx_ray), hipMemcpyHostToDevice));
        trace_device<false>(*sc, static_cast<ljx_ray *>(dr), n, dh, nullptr, nullptr, nullptr);
        HIPX(hipMemcpy(hits, dh, n * sizeof(ljx_hit), hipMemcpyDeviceToHost));
    });
}

int ljx_trace_any(void *scene, const ljx_ray *rays, size_t n, uint8_t *occluded) {
    return guarded([&]() {
        Scene *sc = as_scene(scene);
        if (n == 0) return;
        if (!rays || !occluded) fail(LJX_ERR_INVALID_ARGUMENT, "NULL buffer");
        HIPX(hipSetDevice(sc->device));
        void *dr = nullptr, *dh = nullptr;
        HIPX(hipMalloc(&dr, n * sizeof(ljx_ray)));
        struct Free {
            void *p;
            ~Free() {
                if (p) hipFree(p);
            }
        } f1{dr};
        HIPX(hipMalloc(&dh, n));
        Free f2{dh};
        HIPX(hipMemcpy(dr, rays, n * sizeof(ljx_ray), hipMemcpyHostToDevice));
        trace_device<true>(*sc, static_cast<ljx_ray *>(dr), n, dh, nullptr, nullptr, nullptr);
        HIPX(hipMemcpy(occluded, dh, n, hipMemcpyDeviceToHost));
    });
}

int ljx_bsdf_batch(void *scene, const ljx_bsdf_query *queries, size_t n, ljx_bsdf_result *results) {
    return guarded([&]() {
        Scene *sc = as_scene(scene);
        if (n == 0) return;
        if (!queries || !results) fail(LJX_ERR_INVALID_ARGUMENT, "NULL buffer");
        HIPX(hipSetDevice(sc->device));
        for (size_t i = 0; i < n; i++)
            if (queries[i].material_id < 0 || queries[i].material_id >= sc->num_materials)
                fail(LJX_ERR_INVALID_ARGUMENT, "material id out of range");
        void *dq = nullptr, *dr = nullptr;
        HIPX(hipMalloc(&dq, n * sizeof(ljx_bsdf_query)));
        struct Free {
            void *p;
            ~Free() {
                if (p) hipFree(p);
            }
        } f1{dq};
        HIPX(hipMalloc(&dr, n * sizeof(ljx_bsdf_result)));
        Free f2{dr};
        HIPX(hipMemcpy(dq, queries, n * sizeof(ljx_bsdf_query), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_bsdf_batch, dim3(unsigned((n + 127) / 128)), dim3(128), 0, sc->stream, sc->ds,
                           static_cast<const ljx_bsdf_query *>(dq), n, static_cast<ljx_bsdf_result *>(dr));
        HIPX(hipGetLastError());
        HIPX(hipStreamSynchronize(sc->stream));
        HIPX(hipMemcpy(results, dr, n * sizeof(ljx_bsdf_result), hipMemcpyDeviceToHost));
    });
}

int ljx_device_alloc(int device, size_t bytes, void **ptr) {
    return guarded([&]() {
        if (!ptr) fail(LJX_ERR_INVALID_ARGUMENT, "ptr is NULL");
        HIPX(hipSetDevice(device));
        HIPX(hipMalloc(ptr, bytes));
    });
}

int ljx_device_free(void *ptr) {
    return guarded([&]() { HIPX(hipFree(ptr)); });
}

int ljx_copy_to_device(void *dst, const void *src, size_t bytes) {
    return guarded([&]() { HIPX(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}

int ljx_copy_to_host(void *dst, const void *src, size_t bytes) {
    return guarded([&]() { HIPX(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}

}  // extern "C"
