"""ctypes mirror of include/lajolla_amd.h and the loader for the HIP library.

The structures here must match the C header field for field; the
``test_capi_layout`` test checks their sizes against the compiled library.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblajolla_amd.so")

# status codes
LJX_OK = 0
(ERR_INVALID_ARGUMENT, ERR_OUT_OF_MEMORY, ERR_HIP, ERR_QUEUE_OVERFLOW, ERR_UNSUPPORTED,
 ERR_NO_DEVICE) = range(1, 7)
ERR_NAMES = {
    1: "invalid argument",
    2: "out of memory",
    3: "HIP error",
    4: "queue overflow",
    5: "unsupported",
    6: "no device",
}

# variant tags (src/shape.h:53, src/material.h:102-110, src/light.h:34,
# src/medium.h:22, src/phase_function.h:16, src/filter.h:38, src/scene.h:14-22)
SHAPE_SPHERE, SHAPE_TRIANGLE_MESH = 0, 1
(MAT_LAMBERTIAN, MAT_ROUGH_PLASTIC, MAT_ROUGH_DIELECTRIC, MAT_DISNEY_DIFFUSE,
 MAT_DISNEY_METAL, MAT_DISNEY_GLASS, MAT_DISNEY_CLEARCOAT, MAT_DISNEY_SHEEN,
 MAT_DISNEY_BSDF) = range(9)
LIGHT_DIFFUSE_AREA, LIGHT_ENVMAP = 0, 1
MEDIUM_HOMOGENEOUS, MEDIUM_HETEROGENEOUS = 0, 1
VOLUME_CONSTANT, VOLUME_GRID = 0, 1
PHASE_ISOTROPIC, PHASE_HENYEY_GREENSTEIN = 0, 1
FILTER_BOX, FILTER_TENT, FILTER_GAUSSIAN = 0, 1, 2
(INTEGRATOR_DEPTH, INTEGRATOR_SHADING_NORMAL, INTEGRATOR_MEAN_CURVATURE,
 INTEGRATOR_RAY_DIFFERENTIAL, INTEGRATOR_MIPMAP_LEVEL, INTEGRATOR_PATH,
 INTEGRATOR_VOLPATH) = range(7)

DEFAULT_SEED = 0x31E241F862A1FB5E  # src/pcg.h:33

D3 = C.c_double * 3
D16 = C.c_double * 16


class ShapeDesc(C.Structure):
    _fields_ = [
        ("type", C.c_uint32),
        ("material_id", C.c_int32),
        ("area_light_id", C.c_int32),
        ("interior_medium_id", C.c_int32),
        ("exterior_medium_id", C.c_int32),
        ("positions", C.POINTER(C.c_float)),
        ("indices", C.POINTER(C.c_int32)),
        ("normals", C.POINTER(C.c_float)),
        ("uvs", C.POINTER(C.c_float)),
        ("num_vertices", C.c_uint64),
        ("num_triangles", C.c_uint64),
        ("center", D3),
        ("radius", C.c_double),
    ]


class MaterialDesc(C.Structure):
    _fields_ = [
        ("type", C.c_uint32),
        ("color_a", D3),
        ("color_b", D3),
        ("roughness", C.c_double),
        ("eta", C.c_double),
        ("specular_transmission", C.c_double),
        ("metallic", C.c_double),
        ("subsurface", C.c_double),
        ("specular", C.c_double),
        ("specular_tint", C.c_double),
        ("anisotropic", C.c_double),
        ("sheen", C.c_double),
        ("sheen_tint", C.c_double),
        ("clearcoat", C.c_double),
        ("clearcoat_gloss", C.c_double),
        ("texture", C.c_int32 * 13),
    ]


# texture slots of ljx_material_desc.texture, in order
TEXTURE_SLOTS = ("color_a", "color_b", "roughness", "specular_transmission", "metallic", "subsurface", "specular",
                 "specular_tint", "anisotropic", "sheen", "sheen_tint", "clearcoat", "clearcoat_gloss")
TEXTURE_IMAGE, TEXTURE_CHECKERBOARD = 1, 2


class TextureDesc(C.Structure):
    _fields_ = [
        ("type", C.c_uint32),
        ("channels", C.c_int32),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("data", C.POINTER(C.c_float)),
        ("color0", D3),
        ("color1", D3),
        ("uscale", C.c_double),
        ("vscale", C.c_double),
        ("uoffset", C.c_double),
        ("voffset", C.c_double),
    ]


class LightDesc(C.Structure):
    _fields_ = [
        ("type", C.c_uint32),
        ("shape_id", C.c_int32),
        ("intensity", D3),
        ("env_rgb", C.POINTER(C.c_float)),
        ("env_width", C.c_int32),
        ("env_height", C.c_int32),
        ("to_world", D16),
        ("scale", C.c_double),
    ]


class VolumeDesc(C.Structure):
    _fields_ = [
        ("type", C.c_uint32),
        ("channels", C.c_int32),
        ("resolution", C.c_int32 * 3),
        ("pad", C.c_int32),
        ("value", D3),
        ("p_min", D3),
        ("p_max", D3),
        ("scale", C.c_double),
        ("data", C.POINTER(C.c_float)),
    ]


class MediumDesc(C.Structure):
    _fields_ = [
        ("type", C.c_uint32),
        ("sigma_a", D3),
        ("sigma_s", D3),
        ("phase_type", C.c_uint32),
        ("g", C.c_double),
        ("albedo", VolumeDesc),
        ("density", VolumeDesc),
    ]


class CameraDesc(C.Structure):
    _fields_ = [
        ("cam_to_world", D16),
        ("fov", C.c_double),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("filter_type", C.c_uint32),
        ("filter_param", C.c_double),
        ("medium_id", C.c_int32),
    ]


class RenderOptions(C.Structure):
    _fields_ = [
        ("integrator", C.c_uint32),
        ("samples_per_pixel", C.c_int32),
        ("max_depth", C.c_int32),
        ("rr_depth", C.c_int32),
        ("vol_path_version", C.c_int32),
        ("max_null_collisions", C.c_int32),
        ("seed", C.c_uint64),
    ]


class SceneDescC(C.Structure):
    _fields_ = [
        ("camera", CameraDesc),
        ("shapes", C.POINTER(ShapeDesc)),
        ("num_shapes", C.c_int32),
        ("materials", C.POINTER(MaterialDesc)),
        ("num_materials", C.c_int32),
        ("lights", C.POINTER(LightDesc)),
        ("num_lights", C.c_int32),
        ("media", C.POINTER(MediumDesc)),
        ("num_media", C.c_int32),
        ("envmap_light_id", C.c_int32),
        ("options", RenderOptions),
        ("textures", C.POINTER(TextureDesc)),
        ("num_textures", C.c_int32),
    ]


class Ray(C.Structure):
    _fields_ = [("org", C.c_float * 3), ("tnear", C.c_float),
                ("dir", C.c_float * 3), ("tfar", C.c_float)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("u", C.c_float), ("v", C.c_float),
                ("shape_id", C.c_int32), ("prim_id", C.c_int32),
                ("ng", C.c_float * 3)]


class BSDFQuery(C.Structure):
    _fields_ = [
        ("material_id", C.c_int32),
        ("transport_to_view", C.c_int32),
        ("dir_in", D3),
        ("dir_out", D3),
        ("geometric_normal", D3),
        ("frame_x", D3),
        ("frame_y", D3),
        ("frame_n", D3),
        ("rnd_uv", C.c_double * 2),
        ("rnd_w", C.c_double),
    ]


class BSDFResult(C.Structure):
    _fields_ = [
        ("f", D3),
        ("pdf", C.c_double),
        ("sample_dir", D3),
        ("sample_eta", C.c_double),
        ("sample_roughness", C.c_double),
        ("sample_valid", C.c_int32),
    ]


class RenderStats(C.Structure):
    _fields_ = [
        ("camera_paths", C.c_uint64),
        ("extension_rays", C.c_uint64),
        ("shadow_rays", C.c_uint64),
        ("passes", C.c_uint64),
        ("total_ms", C.c_double),
        ("extend_ms", C.c_double),
        ("extend_launches", C.c_uint64),
        ("shadow_ms", C.c_double),
        ("shadow_launches", C.c_uint64),
        ("shade_ms", C.c_double),
        ("other_ms", C.c_double),
        ("node_visits", C.c_uint64),
        ("tri_tests", C.c_uint64),
        ("shadow_node_visits", C.c_uint64),
        ("shadow_tri_tests", C.c_uint64),
        ("instrumented_rays", C.c_uint64),
        ("build_ms", C.c_double),
        ("bvh_depth", C.c_int32),
        ("node_bytes", C.c_int32),
        ("guide_ms", C.c_double),
        ("train_ms", C.c_double),
        ("train_samples", C.c_uint64),
        ("train_steps", C.c_uint64),
        ("train_step_samples", C.c_uint64),
        ("guide_queries", C.c_uint64),
    ]


class ExecOptions(C.Structure):
    _fields_ = [
        ("collect_timing", C.c_int32),
        ("collect_traversal", C.c_int32),
        ("max_paths_in_flight", C.c_int32),
        ("block_size", C.c_int32),
        ("sort_threshold", C.c_int32),
        ("trace_blocks_per_cu", C.c_int32),
    ]


GUIDE_NUM_PARAMS = 25384
GUIDE_NUM_OUTPUTS = 40


class GuideConfig(C.Structure):
    _fields_ = [
        ("learning_rate", C.c_float),
        ("beta1", C.c_float),
        ("beta2", C.c_float),
        ("epsilon", C.c_float),
        ("batch_size", C.c_int32),
        ("pad", C.c_int32),
        ("seed", C.c_uint64),
    ]


class GuideOptions(C.Structure):
    _fields_ = [
        ("mix_prob", C.c_float),
        ("train_spp", C.c_int32),
        ("train_spp_per_pass", C.c_int32),
        ("steps_per_pass", C.c_int32),
        ("max_record_depth", C.c_int32),
        ("pad", C.c_int32),
    ]


class BatchJob(C.Structure):
    _fields_ = [
        ("scene", C.c_void_p),
        ("camera", CameraDesc),
        ("samples_per_pixel", C.c_int32),
        ("max_depth", C.c_int32),
        ("seed", C.c_uint64),
        ("rgb_out_device", C.c_void_p),
    ]


class BuildOptions(C.Structure):
    _fields_ = [("bvh_builder", C.c_int32), ("max_leaf_size", C.c_int32), ("node_cost", C.c_float), ("pad", C.c_int32)]


class BvhInfo(C.Structure):
    _fields_ = [
        ("builder", C.c_int32),
        ("depth", C.c_int32),
        ("stack_need", C.c_int32),
        ("binary_depth", C.c_int32),
        ("num_nodes", C.c_uint64),
        ("num_slots", C.c_uint64),
        ("node_hash", C.c_uint64),
        ("slot_hash", C.c_uint64),
        ("sah_cost", C.c_double),
        ("bvh_ms", C.c_double),
        ("build_ms", C.c_double),
        ("qnode_hash", C.c_uint64),
        ("qnode_bad_planes", C.c_uint64),
        ("qnode_max_loose", C.c_double),
    ]


BVH_BUILDER_GPU, BVH_BUILDER_HOST = 0, 1


class BatchStats(C.Structure):
    _fields_ = [
        ("total_ms", C.c_double),
        ("build_ms", C.c_double),
        ("launch_ms", C.c_double),
        ("samples", C.c_uint64),
        ("passes", C.c_uint64),
        ("graph_nodes", C.c_uint64),
        ("num_scenes", C.c_int32),
        ("pad", C.c_int32),
    ]


# Every symbol include/lajolla_amd.h declares, with its ctypes signature.
VP = C.c_void_p
SIGNATURES = {
    "ljx_abi_version": ([], C.c_int),
    "ljx_last_error": ([C.c_char_p, C.c_size_t], C.c_int),
    "ljx_device_count": ([C.POINTER(C.c_int)], C.c_int),
    "ljx_scene_parse": ([C.c_char_p, C.POINTER(C.POINTER(SceneDescC))], C.c_int),
    "ljx_scene_desc_free": ([C.POINTER(SceneDescC)], C.c_int),
    "ljx_scene_load": ([C.c_char_p, C.c_int, C.POINTER(VP)], C.c_int),
    "ljx_scene_create": ([C.POINTER(SceneDescC), C.c_int, C.POINTER(VP)], C.c_int),
    "ljx_scene_create_ex": ([C.POINTER(SceneDescC), C.c_int, C.POINTER(BuildOptions), C.POINTER(VP)], C.c_int),
    "ljx_scene_bvh_info": ([VP, C.POINTER(BvhInfo)], C.c_int),
    "ljx_scene_node_fetch_rate": ([VP, C.c_int, C.POINTER(C.c_double)], C.c_int),
    "ljx_scene_destroy": ([VP], C.c_int),
    "ljx_scene_bounds": ([VP, C.POINTER(C.c_double)], C.c_int),
    "ljx_render": ([VP, C.POINTER(RenderOptions), C.POINTER(ExecOptions),
                    C.POINTER(C.c_float), C.POINTER(RenderStats)], C.c_int),
    "ljx_render_device": ([VP, C.POINTER(RenderOptions), C.POINTER(ExecOptions),
                           VP, C.POINTER(RenderStats)], C.c_int),
    "ljx_trace_closest": ([VP, C.POINTER(Ray), C.c_size_t, C.POINTER(Hit)], C.c_int),
    "ljx_trace_any": ([VP, C.POINTER(Ray), C.c_size_t, C.POINTER(C.c_uint8)], C.c_int),
    "ljx_trace_closest_device": ([VP, VP, C.c_size_t, VP, C.POINTER(ExecOptions),
                                  C.POINTER(C.c_double), C.POINTER(RenderStats)], C.c_int),
    "ljx_trace_closest_device_noalloc": ([VP, VP, C.c_size_t, VP], C.c_int),
    "ljx_trace_any_device": ([VP, VP, C.c_size_t, VP, C.POINTER(ExecOptions),
                              C.POINTER(C.c_double), C.POINTER(RenderStats)], C.c_int),
    "ljx_bsdf_batch": ([VP, C.POINTER(BSDFQuery), C.c_size_t, C.POINTER(BSDFResult)], C.c_int),
    "ljx_guide_create": ([C.c_int, C.POINTER(GuideConfig), C.POINTER(VP)], C.c_int),
    "ljx_guide_destroy": ([VP], C.c_int),
    "ljx_guide_get_params": ([VP, C.POINTER(C.c_float), C.c_size_t], C.c_int),
    "ljx_guide_set_params": ([VP, C.POINTER(C.c_float), C.c_size_t], C.c_int),
    "ljx_guide_forward_device": ([VP, VP, C.c_size_t, VP], C.c_int),
    "ljx_guide_gradient_device": ([VP, VP, VP, C.c_size_t, VP, C.POINTER(C.c_double)], C.c_int),
    "ljx_guide_train_step_device": ([VP, VP, VP, C.c_size_t, C.POINTER(C.c_double)], C.c_int),
    "ljx_render_guided_device": ([VP, VP, C.POINTER(RenderOptions), C.POINTER(ExecOptions),
                                  C.POINTER(GuideOptions), VP, C.POINTER(RenderStats)], C.c_int),
    "ljx_render_batch_device": ([C.POINTER(BatchJob), C.c_size_t, C.POINTER(ExecOptions), C.POINTER(C.c_double),
                                 C.POINTER(BatchStats)], C.c_int),
    "ljx_autotune": ([VP, C.POINTER(RenderOptions), C.POINTER(ExecOptions)], C.c_int),
    "ljx_device_alloc": ([C.c_int, C.c_size_t, C.POINTER(VP)], C.c_int),
    "ljx_device_free": ([VP], C.c_int),
    "ljx_copy_to_device": ([VP, VP, C.c_size_t], C.c_int),
    "ljx_copy_to_host": ([VP, VP, C.c_size_t], C.c_int),
}

_lib = None


class LjxError(RuntimeError):
    pass


def load(path=None):
    """Loads the HIP library; raises if it is missing (there is no fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise LjxError(
            f"{p} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(p)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if path is None:
        _lib = lib
    return lib


def check(rc, lib=None):
    if rc != LJX_OK:
        lib = lib or load()
        buf = C.create_string_buffer(1024)
        lib.ljx_last_error(buf, len(buf))
        raise LjxError(f"{ERR_NAMES.get(rc, rc)}: {buf.value.decode(errors='replace')}")
