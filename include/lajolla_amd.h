/*
 * lajolla_amd.h -- C ABI of the MI355X (gfx950) path-tracing core.
 *
 * The reference renderer (lajolla) exposes three C++ entry points to its
 * callers and one native dependency under its hot path:
 *
 *   std::unique_ptr<Scene> parse_scene(const fs::path&, const RTCDevice&)
 *                                                  src/parsers/parse_scene.h:9
 *   Image3 render(const Scene&)                    src/render.h:9
 *   std::optional<PathVertex> intersect(const Scene&, const Ray&, ...)
 *   bool occluded(const Scene&, const Ray&)        src/intersection.h:39-44
 *     (both sit on Embree's rtcIntersect1 / rtcOccluded1,
 *      embree/include/embree4/rtcore_scene.h:160,198)
 *
 * This header replaces them with plain-pointer entry points.  Closed
 * std::variant types (Material src/material.h:102-110, Shape src/shape.h:53,
 * Light src/light.h:34, Medium src/medium.h:22, PhaseFunction
 * src/phase_function.h:16, Filter src/filter.h:38) become {tag, payload}
 * POD records.  Every call returns an int status (LJX_OK == 0); no C++
 * exception crosses the boundary, and ljx_last_error() gives the message of
 * the last failure on the calling thread.
 *
 * Ownership: the caller owns every input array; ljx_scene_create deep-copies
 * the descriptor into device memory and keeps no pointer into caller memory.
 * A scene handle owns its device buffers, BVH and path queues.  Calls on one
 * handle must be externally synchronised; different handles may be used from
 * different threads.
 */
#ifndef LAJOLLA_AMD_H
#define LAJOLLA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LJX_ABI_VERSION 1

/* ---- status codes -------------------------------------------------------- */
enum {
    LJX_OK = 0,
    LJX_ERR_INVALID_ARGUMENT = 1, /* malformed descriptor or argument */
    LJX_ERR_OUT_OF_MEMORY = 2,    /* device or host allocation failed */
    LJX_ERR_HIP = 3,              /* a HIP runtime call failed */
    LJX_ERR_QUEUE_OVERFLOW = 4,   /* a path queue exceeded its capacity */
    LJX_ERR_UNSUPPORTED = 5,      /* a feature outside this build's scope */
    LJX_ERR_NO_DEVICE = 6         /* no usable gfx950 device */
};

/* ---- variant tags (same order as the reference's std::variant lists) ----- */
/* Shape = variant<Sphere, TriangleMesh>                  src/shape.h:53 */
enum { LJX_SHAPE_SPHERE = 0, LJX_SHAPE_TRIANGLE_MESH = 1 };

/* Material = variant<Lambertian, RoughPlastic, RoughDielectric, DisneyDiffuse,
 * DisneyMetal, DisneyGlass, DisneyClearcoat, DisneySheen, DisneyBSDF>
 *                                                        src/material.h:102-110 */
enum {
    LJX_MAT_LAMBERTIAN = 0,
    LJX_MAT_ROUGH_PLASTIC = 1,
    LJX_MAT_ROUGH_DIELECTRIC = 2,
    LJX_MAT_DISNEY_DIFFUSE = 3,
    LJX_MAT_DISNEY_METAL = 4,
    LJX_MAT_DISNEY_GLASS = 5,
    LJX_MAT_DISNEY_CLEARCOAT = 6,
    LJX_MAT_DISNEY_SHEEN = 7,
    LJX_MAT_DISNEY_BSDF = 8,
    LJX_MAT_COUNT = 9
};

/* Light = variant<DiffuseAreaLight, Envmap>              src/light.h:34 */
enum { LJX_LIGHT_DIFFUSE_AREA = 0, LJX_LIGHT_ENVMAP = 1 };

/* Medium = variant<HomogeneousMedium, HeterogeneousMedium> src/medium.h:22 */
enum { LJX_MEDIUM_HOMOGENEOUS = 0, LJX_MEDIUM_HETEROGENEOUS = 1 };

/* Texture<T> = variant<ConstantTexture, ImageTexture, CheckerboardTexture>
 *                                                       src/texture.h:113 */
enum { LJX_TEXTURE_IMAGE = 1, LJX_TEXTURE_CHECKERBOARD = 2 };
#define LJX_MAT_TEXTURE_SLOTS 13

/* Volume<T> = variant<ConstantVolume<T>, GridVolume<T>>   src/volume.h:13-31 */
enum { LJX_VOLUME_CONSTANT = 0, LJX_VOLUME_GRID = 1 };

/* PhaseFunction = variant<IsotropicPhase, HenyeyGreenstein> src/phase_function.h:16 */
enum { LJX_PHASE_ISOTROPIC = 0, LJX_PHASE_HENYEY_GREENSTEIN = 1 };

/* Filter = variant<Box, Tent, Gaussian>                  src/filter.h:38 */
enum { LJX_FILTER_BOX = 0, LJX_FILTER_TENT = 1, LJX_FILTER_GAUSSIAN = 2 };

/* Integrator                                             src/scene.h:14-22 */
enum {
    LJX_INTEGRATOR_DEPTH = 0,
    LJX_INTEGRATOR_SHADING_NORMAL = 1,
    LJX_INTEGRATOR_MEAN_CURVATURE = 2,
    LJX_INTEGRATOR_RAY_DIFFERENTIAL = 3,
    LJX_INTEGRATOR_MIPMAP_LEVEL = 4,
    LJX_INTEGRATOR_PATH = 5,
    LJX_INTEGRATOR_VOLPATH = 6
};

/* ---- scene descriptor ---------------------------------------------------- */

/* One shape (src/shape.h:20-51).  Triangle meshes point at caller arrays;
 * positions/normals are xyz per vertex, uvs are uv per vertex, indices are
 * three int32 per triangle.  normals and uvs may be NULL. */
typedef struct ljx_shape_desc {
    uint32_t type;                 /* LJX_SHAPE_* */
    int32_t material_id;           /* -1: index-matched interface (volpath) */
    int32_t area_light_id;         /* -1: not emissive */
    int32_t interior_medium_id;    /* -1: vacuum */
    int32_t exterior_medium_id;
    /* triangle mesh */
    const float *positions;
    const int32_t *indices;
    const float *normals;
    const float *uvs;
    uint64_t num_vertices;
    uint64_t num_triangles;
    /* sphere */
    double center[3];
    double radius;
} ljx_shape_desc;

/* One material with constant textures (src/material.h:12-98).  Field use
 * per tag:
 *   LAMBERTIAN        color_a = reflectance
 *   ROUGH_PLASTIC     color_a = diffuse_reflectance, color_b = specular_reflectance,
 *                     roughness, eta
 *   ROUGH_DIELECTRIC  color_a = specular_reflectance, color_b = specular_transmittance,
 *                     roughness, eta
 *   DISNEY_*          color_a = base_color, the scalar Disney parameters, eta
 */
typedef struct ljx_material_desc {
    uint32_t type;                 /* LJX_MAT_* */
    double color_a[3];
    double color_b[3];
    double roughness;
    double eta;
    double specular_transmission;
    double metallic;
    double subsurface;
    double specular;
    double specular_tint;
    double anisotropic;
    double sheen;
    double sheen_tint;
    double clearcoat;
    double clearcoat_gloss;
    /* Texture<T> fields (src/material.h:12-98): 0 takes the constant above;
     * k > 0 takes textures[k - 1] of the scene, in the order color_a,
     * color_b, roughness, specular_transmission, metallic, subsurface,
     * specular, specular_tint, anisotropic, sheen, sheen_tint, clearcoat,
     * clearcoat_gloss.  A scalar field reads a 1-channel image as is and the
     * average of a 3-channel image or of the checkerboard colours. */
    int32_t texture[LJX_MAT_TEXTURE_SLOTS];
} ljx_material_desc;

/* One light (src/light.h:12-30).  An environment map is an image of
 * env_width x env_height RGB texels (row-major from the top row, float) in
 * lat-long layout (src/lights/envmap.inl), sampled by a TableDist2D over
 * luminance x sin(elevation) at texel centres; a 1x1 image is a constant
 * environment. */
typedef struct ljx_light_desc {
    uint32_t type;                 /* LJX_LIGHT_* */
    int32_t shape_id;              /* diffuse area light: emitting shape */
    double intensity[3];           /* diffuse area light radiance */
    const float *env_rgb;          /* envmap texels */
    int32_t env_width;
    int32_t env_height;
    double to_world[16];           /* envmap frame, row-major 4x4 */
    double scale;                  /* envmap radiance scale */
} ljx_light_desc;

/* A spectrum-valued volume (src/volume.h:13-31).  A lookup at p returns
 * scale * value for a constant volume; for a grid, zero outside the box
 * [p_min, p_max] and otherwise scale times the trilinear interpolation of the
 * voxels (src/volume.h:40-81), a voxel being channels floats (1: grey, the
 * same value in R, G and B; 3: RGB) at data[channels * ((z * ry + y) * rx + x)].
 * The majorant of a grid is scale times the per-channel maximum over its
 * voxels (at least 0).  The caller owns data; ljx_scene_create copies it. */
typedef struct ljx_volume_desc {
    uint32_t type;                 /* LJX_VOLUME_* */
    int32_t channels;              /* grid: 1 or 3 */
    int32_t resolution[3];         /* grid: rx, ry, rz >= 1 */
    int32_t pad;
    double value[3];               /* constant */
    double p_min[3];               /* grid box */
    double p_max[3];
    double scale;                  /* set_scale (src/volume.h:104-118) */
    const float *data;             /* grid voxels */
} ljx_volume_desc;

/* One participating medium (src/medium.h:10-22).  Homogeneous media use
 * sigma_a and sigma_s; heterogeneous ones sigma_s = density * albedo and
 * sigma_a = density * (1 - albedo) at each point, with the majorant of the
 * density volume (src/media/heterogeneous.inl:3-21). */
typedef struct ljx_medium_desc {
    uint32_t type;                 /* LJX_MEDIUM_* */
    double sigma_a[3];
    double sigma_s[3];
    uint32_t phase_type;           /* LJX_PHASE_* */
    double g;                      /* Henyey-Greenstein asymmetry */
    ljx_volume_desc albedo;        /* heterogeneous */
    ljx_volume_desc density;       /* heterogeneous */
} ljx_medium_desc;

/* Pinhole camera (src/camera.h:7-22).  fov is the x-axis field of view in
 * degrees, after the parser's fovAxis conversion (parse_scene.cpp:830-843). */
typedef struct ljx_camera_desc {
    double cam_to_world[16];       /* row-major 4x4 */
    double fov;
    int32_t width;
    int32_t height;
    uint32_t filter_type;          /* LJX_FILTER_* */
    double filter_param;           /* box/tent width, gaussian stddev */
    int32_t medium_id;             /* -1: camera in vacuum */
} ljx_camera_desc;

/* RenderOptions (src/scene.h:24-31) plus the sample seed. */
typedef struct ljx_render_options {
    uint32_t integrator;           /* LJX_INTEGRATOR_* */
    int32_t samples_per_pixel;
    int32_t max_depth;             /* -1: Russian roulette only */
    int32_t rr_depth;
    int32_t vol_path_version;
    int32_t max_null_collisions;
    uint64_t seed;                 /* mixed into every (pixel, sample) stream */
} ljx_render_options;

/* An image or checkerboard texture (src/texture.h:82-154).  An image is
 * width x height pixels of channels floats (1 or 3), row-major from the top
 * row; ljx_scene_create builds its mipmap (2x2 box filter, at most 8 levels,
 * src/mipmap.h:25-48).  A lookup at uv reads the image at
 * (uv * scale + offset) mod 1, bilinearly on the level the hit's footprint
 * selects (trilinear between levels), or the checkerboard colour of the
 * half-cell (src/texture.h:141-154). */
typedef struct ljx_texture_desc {
    uint32_t type;                 /* LJX_TEXTURE_* */
    int32_t channels;              /* image: 1 or 3 */
    int32_t width;                 /* image */
    int32_t height;
    const float *data;             /* image pixels; the caller owns them */
    double color0[3];              /* checkerboard */
    double color1[3];
    double uscale, vscale, uoffset, voffset;
} ljx_texture_desc;

typedef struct ljx_scene_desc {
    ljx_camera_desc camera;
    const ljx_shape_desc *shapes;
    int32_t num_shapes;
    const ljx_material_desc *materials;
    int32_t num_materials;
    const ljx_light_desc *lights;
    int32_t num_lights;
    const ljx_medium_desc *media;
    int32_t num_media;
    int32_t envmap_light_id;       /* -1 if none (src/scene.h:66) */
    ljx_render_options options;
    const ljx_texture_desc *textures;  /* TexturePool (src/texture.h:13-19) */
    int32_t num_textures;
} ljx_scene_desc;

/* ---- rays and hits for the intersection boundary ------------------------ */

/* A ray in FP32, as Embree receives it (src/intersection.cpp:15-31). */
typedef struct ljx_ray {
    float org[3];
    float tnear;
    float dir[3];
    float tfar;
} ljx_ray;

/* Closest hit: shape_id == -1 for a miss.  u weights vertex 1 and v weights
 * vertex 2, as Embree's RTCHit does (src/intersection.cpp:48). */
typedef struct ljx_hit {
    float t;
    float u;
    float v;
    int32_t shape_id;
    int32_t prim_id;
    float ng[3];                   /* unnormalised geometric normal */
} ljx_hit;

/* ---- BSDF evaluation boundary (material-level parity) ------------------- */

/* One query of eval / pdf_sample_bsdf / sample_bsdf (src/material.h:126-166)
 * at a shading point with the given frame.  Directions point away from the
 * surface.  All FP64, like the reference's Real. */
typedef struct ljx_bsdf_query {
    int32_t material_id;
    int32_t transport_to_view;     /* 0: TO_LIGHT (default), 1: TO_VIEW */
    double dir_in[3];
    double dir_out[3];
    double geometric_normal[3];
    double frame_x[3];
    double frame_y[3];
    double frame_n[3];
    double rnd_uv[2];
    double rnd_w;
} ljx_bsdf_query;

typedef struct ljx_bsdf_result {
    double f[3];                   /* eval(dir_in, dir_out) incl. cosine */
    double pdf;                    /* pdf_sample_bsdf(dir_in, dir_out) */
    double sample_dir[3];          /* sample_bsdf(dir_in, rnd) */
    double sample_eta;
    double sample_roughness;
    int32_t sample_valid;          /* 0 where the reference returns nullopt */
} ljx_bsdf_result;

/* ---- render statistics -------------------------------------------------- */
typedef struct ljx_render_stats {
    uint64_t camera_paths;         /* W*H*spp */
    uint64_t extension_rays;       /* closest-hit queries traced */
    uint64_t shadow_rays;          /* any-hit queries traced */
    uint64_t passes;               /* wavefront passes launched */
    double total_ms;               /* wall time of the render call */
    double extend_ms;              /* HIP-event time in the closest-hit kernel */
    uint64_t extend_launches;
    double shadow_ms;              /* HIP-event time in the any-hit kernel */
    uint64_t shadow_launches;
    double shade_ms;               /* HIP-event time in the shading kernels */
    double other_ms;               /* raygen, accumulate, bookkeeping kernels */
    uint64_t node_visits;          /* closest-hit kernel; instrumented runs only */
    uint64_t tri_tests;            /* primitive tests in the closest-hit kernel */
    uint64_t shadow_node_visits;   /* any-hit kernel; instrumented runs only */
    uint64_t shadow_tri_tests;
    uint64_t instrumented_rays;
    double build_ms;               /* scene upload + BVH build of this scene */
    int32_t bvh_depth;             /* nodes on the longest root-to-leaf path */
    int32_t node_bytes;            /* bytes a traversal reads per node visit */
    double guide_ms;               /* HIP-event time in guide inference (guided renders) */
    double train_ms;               /* sample collection + Adam steps (guided renders) */
    uint64_t train_samples;        /* training samples recorded */
    uint64_t train_steps;          /* Adam steps taken */
    uint64_t train_step_samples;   /* samples summed over the Adam steps' batches */
    uint64_t guide_queries;        /* network evaluations of guided vertices (guided renders) */
} ljx_render_stats;

/* Render-call switches that do not change results. */
typedef struct ljx_exec_options {
    int32_t collect_timing;        /* 1: HIP events around every kernel */
    int32_t collect_traversal;     /* 1: count node visits / tri tests */
    int32_t max_paths_in_flight;   /* 0: library default (67,108,864, also the maximum) */
    int32_t block_size;            /* 0: library default (multiple of 64) */
    int32_t sort_threshold;        /* 0: library default; <0: never sort */
    int32_t trace_blocks_per_cu;   /* traversal blocks per CU (0: library default, 8) */
} ljx_exec_options;

/* ---- entry points -------------------------------------------------------- */

/* Library ABI version (LJX_ABI_VERSION).  No reference counterpart. */
int ljx_abi_version(void);

/* Copies the message of the last failed call on this thread (NUL-terminated,
 * truncated to cap bytes).  Replaces fl_exception::what() (src/flexception.h:8-24). */
int ljx_last_error(char *buf, size_t cap);

/* Number of usable gfx950 devices.  Replaces rtcNewDevice (src/main.cpp:30). */
int ljx_device_count(int *count);

/* Reads a scene file (the Mitsuba-style XML subset of lajolla's parser, with
 * OBJ meshes, constant textures and homogeneous media) into a descriptor
 * owned by the library; free it with ljx_scene_desc_free.  Replaces the
 * parsing half of parse_scene (src/parsers/parse_scene.cpp:1600-1617). */
int ljx_scene_parse(const char *xml_path, ljx_scene_desc **desc_out);
int ljx_scene_desc_free(ljx_scene_desc *desc);

/* ljx_scene_parse followed by ljx_scene_create: the whole of parse_scene
 * (src/parsers/parse_scene.h:9). */
int ljx_scene_load(const char *xml_path, int device, void **scene_out);

/* Builds the device scene: uploads geometry, builds the SAH BVH on the
 * device, light and triangle sampling tables and the scene bounding sphere.
 * Replaces Scene::Scene (src/scene.cpp:4-53) including rtcCommitScene. */
int ljx_scene_create(const ljx_scene_desc *desc, int device, void **scene_out);

/* BVH build choices (rtcSetSceneBuildQuality, src/scene.cpp:20-27, is the
 * reference's one choice).  Both builders give the same tree. */
enum { LJX_BVH_BUILDER_GPU = 0, LJX_BVH_BUILDER_HOST = 1 };
typedef struct ljx_build_options {
    int32_t bvh_builder;           /* LJX_BVH_BUILDER_* (default GPU) */
    int32_t max_leaf_size;         /* primitives per leaf at most, 1..8 (0: 8) */
    float node_cost;               /* SAH cost of a node visit in primitive tests (0: the library default) */
    int32_t pad;
} ljx_build_options;

/* ljx_scene_create with build options (NULL: the defaults). */
int ljx_scene_create_ex(const ljx_scene_desc *desc, int device, const ljx_build_options *build,
                        void **scene_out);

/* The scene's BVH: sizes, quality and a fingerprint.  No reference
 * counterpart (Embree's tree is opaque). */
typedef struct ljx_bvh_info {
    int32_t builder;               /* LJX_BVH_BUILDER_* that built it */
    int32_t depth;                 /* 4-wide nodes on the longest root-to-leaf path */
    int32_t stack_need;            /* most traversal stack entries a ray can hold */
    int32_t binary_depth;          /* depth of the binary SAH tree before the collapse */
    uint64_t num_nodes;            /* 4-wide nodes */
    uint64_t num_slots;            /* leaf slots (= primitives) */
    uint64_t node_hash;            /* FNV-1a (64-bit) of the node bytes */
    uint64_t slot_hash;            /* FNV-1a (64-bit) of the slot-to-primitive table (int32) */
    double sah_cost;               /* expected node visits + primitive tests of a ray through the root box */
    double bvh_ms;                 /* wall time of the BVH build (boxes, tree, slot records) */
    double build_ms;               /* wall time of the whole scene build */
    uint64_t qnode_hash;           /* FNV-1a (64-bit) of the traversal's quantised nodes (80 B each) */
    uint64_t qnode_bad_planes;     /* quantised box planes that do not hold their FP32 plane (0 expected) */
    double qnode_max_loose;        /* largest outward move of a quantised plane, over its node's extent */
} ljx_bvh_info;

/* Fills info; downloads the nodes (both the built FP32 nodes and the
 * traversal's quantised copy, which it checks against them), so it is for
 * tests and tools, not loops. */
int ljx_scene_bvh_info(void *scene, ljx_bvh_info *info);

/* Measurement helper with no counterpart in the reference: the rate
 * (fetches per second, best of four timed launches) at which the device
 * reads whole 4-wide nodes of the scene's BVH at random, `lanes` lanes of
 * every wave (1..64) each fetching its own node with the next index
 * depending on the data loaded, on the traversal's grid: the rate the
 * vector-memory pipeline sustains for a traversal's fetches when the lanes
 * of a wave share no nodes.  Allocates 4 bytes; blocks until done. */
int ljx_scene_node_fetch_rate(void *scene, int lanes, double *fetches_per_s);

/* Frees every device buffer of the scene.  Replaces Scene::~Scene
 * (src/scene.cpp:55-59). */
int ljx_scene_destroy(void *scene);

/* Scene bounding sphere (src/scene.cpp:30-34): center xyz, radius. */
int ljx_scene_bounds(void *scene, double *center_radius4);

/* Renders into rgb_out (host memory, height*width*3 floats, HWC, the pixel
 * mean over samples_per_pixel).  Replaces render() (src/render.cpp:155-170);
 * the auxiliary integrators (depth ... mipmap level) take one pixel-centre
 * ray per pixel and ignore samples_per_pixel, as aux_render
 * (src/render.cpp:12-69).  opts may be NULL (scene options are used); exec
 * and stats may be NULL. */
int ljx_render(void *scene, const ljx_render_options *opts,
               const ljx_exec_options *exec, float *rgb_out,
               ljx_render_stats *stats);

/* Same as ljx_render, but rgb_out is device memory on the scene's device;
 * no host copy is made.  For callers that keep the image in HBM. */
int ljx_render_device(void *scene, const ljx_render_options *opts,
                      const ljx_exec_options *exec, float *rgb_out_device,
                      ljx_render_stats *stats);

/* Closest hit for n rays in host memory.  Replaces rtcIntersect1 as called
 * by intersect() (src/intersection.cpp:7-33). */
int ljx_trace_closest(void *scene, const ljx_ray *rays, size_t n,
                      ljx_hit *hits);

/* Any hit for n rays in host memory; occluded[i] = 1 if blocked within
 * [tnear, tfar].  Replaces rtcOccluded1 (src/intersection.cpp:67-85). */
int ljx_trace_any(void *scene, const ljx_ray *rays, size_t n,
                  uint8_t *occluded);

/* Device-memory variants of the two calls above (rays, hits and flags are
 * device pointers on the scene's device).  elapsed_ms (may be NULL) gets the
 * HIP-event time of the traversal kernel alone; stats (may be NULL) counts
 * node visits and triangle tests when exec->collect_traversal is set. */
int ljx_trace_closest_device(void *scene, const ljx_ray *rays, size_t n,
                             ljx_hit *hits, const ljx_exec_options *exec,
                             double *elapsed_ms, ljx_render_stats *stats);
int ljx_trace_any_device(void *scene, const ljx_ray *rays, size_t n,
                         uint8_t *occluded, const ljx_exec_options *exec,
                         double *elapsed_ms, ljx_render_stats *stats);

/* Closest hit like ljx_trace_closest_device (device pointers, identical
 * hits), for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): it
 * calls no allocator.  It has no counterpart in the reference.  Its only
 * device memory beyond rays and hits is a fixed 16 MiB static array per
 * device (the traversal stack spill of its grid), held until the process
 * ends; a scene whose traversal would need more than that per block is
 * rejected with LJX_ERR_UNSUPPORTED.  No timing or traversal counters.
 * Blocks until the hits are written.  Calls that overlap on one device share
 * that array and race: serialise them. */
int ljx_trace_closest_device_noalloc(void *scene, const ljx_ray *rays, size_t n,
                                     ljx_hit *hits);

/* Evaluates eval / pdf_sample_bsdf / sample_bsdf of the scene's materials on
 * the device for n queries in host memory, with the same device code the
 * shading kernels run.  Replaces the three dispatchers in
 * src/material.cpp:90-119. */
int ljx_bsdf_batch(void *scene, const ljx_bsdf_query *queries, size_t n,
                   ljx_bsdf_result *results);

/* ---- learned path guiding ------------------------------------------------
 * The reference has no path guiding: these calls extend the render path with
 * a learned directional proposal that is mixed with BSDF sampling by
 * one-sample MIS (the estimator stays unbiased).  The network maps a surface
 * vertex (position, normal, outgoing direction) to a mixture of 8 von
 * Mises-Fisher lobes; LJX_GUIDE_NUM_PARAMS FP32 master parameters in the
 * canonical layout: per layer l = 0..4, W_l[in][out] row-major, then b_l[out],
 * with (in, out) = (160, 64), (64, 64) x 3, (64, 40).  Output o = 5 * lobe +
 * k: k = 0..2 mean direction (normalised), 3 log kappa (clamped to [-3, 8]),
 * 4 mixture logit (softmax over lobes).
 *
 * Features (8 floats per sample, the last unused), each in [0, 1]: position
 * normalised to the scene box, normal (theta / pi, phi / 2 pi), outgoing
 * direction (theta / pi, phi / 2 pi); directions use theta = acos(z),
 * phi = atan2(y, x) wrapped to [0, 2 pi).
 *
 * Training minimises -sum_i w_i log q(dir_i | x_i) / sum_i w_i with Adam;
 * w_i is the sample's radiance-weighted contribution divided by the pdf it
 * was sampled with, so q approaches the product of BSDF, cosine and incident
 * radiance. */
#define LJX_GUIDE_NUM_PARAMS 25384
#define LJX_GUIDE_NUM_OUTPUTS 40

typedef struct ljx_guide_config {
    float learning_rate;           /* Adam step size (1e-3) */
    float beta1;                   /* 0.9 */
    float beta2;                   /* 0.999 */
    float epsilon;                 /* 1e-8 */
    int32_t batch_size;            /* samples per Adam step (1 << 20) */
    int32_t pad;
    uint64_t seed;                 /* parameter initialisation */
} ljx_guide_config;

/* Guided rendering: the first train_spp samples per pixel are rendered in
 * passes of train_spp_per_pass, each followed by steps_per_pass Adam steps on
 * the samples it recorded; the remaining samples use the trained network.
 * All samples enter the image with equal weight. */
typedef struct ljx_guide_options {
    float mix_prob;                /* probability of sampling the guide (0.5) */
    int32_t train_spp;             /* 0: no training (the network is used as is) */
    int32_t train_spp_per_pass;    /* 1 */
    int32_t steps_per_pass;        /* Adam steps after each training pass */
    int32_t max_record_depth;      /* vertices per path recorded for training (<= 16) */
    int32_t pad;
} ljx_guide_options;

/* Creates a network with seeded initial parameters on the device. */
int ljx_guide_create(int device, const ljx_guide_config *cfg, void **guide_out);
int ljx_guide_destroy(void *guide);
/* Copies the FP32 master parameters (n must be LJX_GUIDE_NUM_PARAMS). */
int ljx_guide_get_params(void *guide, float *params, size_t n);
/* Replaces the master parameters and clears the Adam state. */
int ljx_guide_set_params(void *guide, const float *params, size_t n);
/* Network outputs (n x 40, canonical order, before normalisation and
 * softmax) for n feature rows (n x 8 floats); device pointers. */
int ljx_guide_forward_device(void *guide, const float *features, size_t n, float *raw_out);
/* Gradient of the normalised loss with respect to the master parameters for
 * a batch (features n x 8, dirs n x 4 = direction xyz and weight w); device
 * pointers.  grad_out: LJX_GUIDE_NUM_PARAMS floats; loss_out may be NULL.
 * n must not exceed the configured batch_size.  The parameters are not
 * changed. */
int ljx_guide_gradient_device(void *guide, const float *features, const float *dirs, size_t n,
                              float *grad_out, double *loss_out);
/* One Adam step on a batch (same layout as ljx_guide_gradient_device). */
int ljx_guide_train_step_device(void *guide, const float *features, const float *dirs, size_t n,
                                double *loss_out);
/* Renders like ljx_render_device, sampling directions at surface vertices
 * from the mixture of the BSDF and the network (path integrator only), and
 * trains the network on the fly as gopts says.  stats may be NULL. */
int ljx_render_guided_device(void *scene, void *guide, const ljx_render_options *opts,
                             const ljx_exec_options *exec, const ljx_guide_options *gopts,
                             float *rgb_out_device, ljx_render_stats *stats);

/* ---- batched rendering ---------------------------------------------------
 * Many render jobs of uneven size in one captured HIP graph: the jobs of a
 * scene share one wavefront loop whose passes are packed with path slots
 * from several jobs, so small images still fill the device.  A job's image
 * equals ljx_render_device of the same camera, spp and seed up to the order
 * of its per-pixel sums.  No reference counterpart: lajolla renders one
 * scene per call (src/main.cpp:33-45). */
typedef struct ljx_batch_job {
    void *scene;                   /* scene handle: geometry, materials, lights, integrator */
    ljx_camera_desc camera;        /* this job's camera and resolution */
    int32_t samples_per_pixel;
    int32_t max_depth;             /* >= 0; uniform across the jobs of one scene */
    uint64_t seed;
    float *rgb_out_device;         /* height * width * 3 floats on the scene's device */
} ljx_batch_job;

typedef struct ljx_batch_stats {
    double total_ms;               /* wall time of the call */
    double build_ms;               /* graph capture + instantiation (+ autotuning) */
    double launch_ms;              /* HIP-event time of the graph launch */
    uint64_t samples;              /* camera paths of all jobs */
    uint64_t passes;
    uint64_t graph_nodes;
    int32_t num_scenes;
    int32_t pad;
} ljx_batch_stats;

/* Renders every job; per_job_ms (may be NULL, njobs entries) gets each job's
 * completion time from the start of the launch.  exec may be NULL: each
 * scene then uses its autotuned settings (ljx_autotune, run on first use). */
int ljx_render_batch_device(const ljx_batch_job *jobs, size_t njobs, const ljx_exec_options *exec,
                            double *per_job_ms, ljx_batch_stats *stats);

/* Times candidate launch settings (paths in flight, shading block size, sort
 * threshold) on short renders of the scene with opts (NULL: scene options)
 * and keeps the fastest for the scene's batched renders; best (may be NULL)
 * receives it.  Results do not depend on these settings. */
int ljx_autotune(void *scene, const ljx_render_options *opts, ljx_exec_options *best);

/* Device memory helpers for callers without their own allocator. */
int ljx_device_alloc(int device, size_t bytes, void **ptr);
int ljx_device_free(void *ptr);
int ljx_copy_to_device(void *dst, const void *src, size_t bytes);
int ljx_copy_to_host(void *dst, const void *src, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* LAJOLLA_AMD_H */
