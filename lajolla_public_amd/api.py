"""Python host API mirroring lajolla's render path over the C ABI.

    scene = Scene(desc)                 # Scene::Scene (src/scene.cpp:4-53)
    img = render(scene)                 # render() (src/render.cpp:155-170) -> Image3 (H, W, 3)
    hits = scene.trace_closest(rays)    # intersect() on Embree's rtcIntersect1
    occ = scene.trace_any(rays)         # occluded() on rtcOccluded1

Every call goes through liblajolla_amd.so; there is no CPU fallback.
"""
import ctypes as C

import numpy as np

from . import _capi as capi
from .scenes import SceneDesc, render_options


class Scene:
    """Device scene: geometry, SAH BVH, light tables, owned by one handle."""

    def __init__(self, desc: SceneDesc, device: int = 0, bvh_builder=None, max_leaf_size=0, node_cost=0.0):
        """bvh_builder: None (the library default, the device builder),
        capi.BVH_BUILDER_GPU or capi.BVH_BUILDER_HOST; max_leaf_size and
        node_cost: 0 for the library defaults."""
        self.lib = capi.load()
        self.desc = desc
        d, keep = desc.to_c()
        h = C.c_void_p()
        if bvh_builder is None and not max_leaf_size and not node_cost:
            capi.check(self.lib.ljx_scene_create(C.byref(d), device, C.byref(h)), self.lib)
        else:
            bo = capi.BuildOptions(capi.BVH_BUILDER_GPU if bvh_builder is None else bvh_builder, max_leaf_size,
                                   node_cost, 0)
            capi.check(self.lib.ljx_scene_create_ex(C.byref(d), device, C.byref(bo), C.byref(h)), self.lib)
        self.h = h
        self.device = device
        self.width = desc.camera["width"]
        self.height = desc.camera["height"]

    def close(self):
        if getattr(self, "h", None):
            self.lib.ljx_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bvh_info(self):
        info = capi.BvhInfo()
        capi.check(self.lib.ljx_scene_bvh_info(self.h, C.byref(info)), self.lib)
        return {k: getattr(info, k) for k, _ in capi.BvhInfo._fields_}

    def node_fetch_rate(self, lanes=64):
        """Random whole-node fetches per second over this scene's BVH (the
        vector-memory pipeline's rate for lanes that share no nodes)."""
        r = C.c_double(0)
        capi.check(self.lib.ljx_scene_node_fetch_rate(self.h, int(lanes), C.byref(r)), self.lib)
        return r.value

    def bounds(self):
        out = (C.c_double * 4)()
        capi.check(self.lib.ljx_scene_bounds(self.h, out), self.lib)
        return np.array(out[:])

    def render(self, options=None, exec_options=None, stats=False, **kw):
        """Returns the (H, W, 3) float32 image (and a stats dict if asked)."""
        opts = options
        if opts is None and kw:
            o = dict(self.desc.options)
            o.update(kw)
            opts = render_options(**o)
        img = np.zeros((self.height, self.width, 3), np.float32)
        st = capi.RenderStats()
        ex = exec_options
        capi.check(self.lib.ljx_render(self.h, C.byref(opts) if opts is not None else None,
                                       C.byref(ex) if ex is not None else None,
                                       img.ctypes.data_as(C.POINTER(C.c_float)), C.byref(st)), self.lib)
        if stats:
            return img, stats_dict(st)
        return img

    def trace_closest(self, rays):
        n = len(rays)
        hits = (capi.Hit * n)()
        capi.check(self.lib.ljx_trace_closest(self.h, rays, n, hits), self.lib)
        return hits

    def trace_any(self, rays):
        n = len(rays)
        occ = np.zeros(n, np.uint8)
        capi.check(self.lib.ljx_trace_any(self.h, rays, n, occ.ctypes.data_as(C.POINTER(C.c_uint8))), self.lib)
        return occ

    def bsdf(self, queries):
        n = len(queries)
        res = (capi.BSDFResult * n)()
        capi.check(self.lib.ljx_bsdf_batch(self.h, queries, n, res), self.lib)
        return res


def stats_dict(st):
    return {name: getattr(st, name) for name, _ in capi.RenderStats._fields_}


def exec_options(collect_timing=0, collect_traversal=0, max_paths_in_flight=0, block_size=0, sort_threshold=0,
                 trace_blocks_per_cu=0):
    e = capi.ExecOptions()
    e.trace_blocks_per_cu = trace_blocks_per_cu
    e.collect_timing = collect_timing
    e.collect_traversal = collect_traversal
    e.max_paths_in_flight = max_paths_in_flight
    e.block_size = block_size
    e.sort_threshold = sort_threshold
    return e


def render(scene: Scene, **kw):
    """render(scene) -> Image3 as an (H, W, 3) float32 array (src/render.h:9)."""
    return scene.render(**kw)


class DeviceBuffer:
    """Raw device allocation through the C ABI (no torch dependency)."""

    def __init__(self, nbytes, device=0):
        self.lib = capi.load()
        p = C.c_void_p()
        capi.check(self.lib.ljx_device_alloc(device, nbytes, C.byref(p)), self.lib)
        self.ptr = p
        self.nbytes = nbytes

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        capi.check(self.lib.ljx_copy_to_device(self.ptr, arr.ctypes.data_as(C.c_void_p), arr.nbytes), self.lib)

    def download(self, arr):
        capi.check(self.lib.ljx_copy_to_host(arr.ctypes.data_as(C.c_void_p), self.ptr, arr.nbytes), self.lib)
        return arr

    def free(self):
        if self.ptr:
            self.lib.ljx_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Guide:
    """Learned path-guiding network on the device (no reference counterpart:
    an extension of the render path).  Parameters are FP32 in the canonical
    layout of include/lajolla_amd.h; the device computes in BF16 on the matrix
    cores with FP32 accumulation."""

    NUM_PARAMS = capi.GUIDE_NUM_PARAMS

    def __init__(self, device=0, learning_rate=1e-3, batch_size=1 << 20, seed=1, beta1=0.9, beta2=0.999,
                 epsilon=1e-8):
        self.lib = capi.load()
        cfg = capi.GuideConfig()
        cfg.learning_rate = learning_rate
        cfg.beta1 = beta1
        cfg.beta2 = beta2
        cfg.epsilon = epsilon
        cfg.batch_size = batch_size
        cfg.seed = seed
        h = C.c_void_p()
        capi.check(self.lib.ljx_guide_create(device, C.byref(cfg), C.byref(h)), self.lib)
        self.h = h
        self.device = device
        self.batch_size = batch_size

    def close(self):
        if getattr(self, "h", None):
            self.lib.ljx_guide_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(self.NUM_PARAMS, np.float32)
        capi.check(self.lib.ljx_guide_get_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float)), p.size), self.lib)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        capi.check(self.lib.ljx_guide_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float)), p.size), self.lib)

    def _upload(self, arr):
        arr = np.ascontiguousarray(arr, np.float32)
        b = DeviceBuffer(max(arr.nbytes, 16), self.device)
        b.upload(arr)
        return b

    def forward(self, feats):
        """Raw outputs (n, 40) for feature rows (n, 8)."""
        n = len(feats)
        f = self._upload(feats)
        out = DeviceBuffer(n * 40 * 4, self.device)
        capi.check(self.lib.ljx_guide_forward_device(self.h, f.ptr, n, out.ptr), self.lib)
        return out.download(np.zeros((n, 40), np.float32))

    def gradient(self, feats, dirs):
        """(normalised loss, gradient) for a batch; the parameters do not change."""
        n = len(feats)
        f, d = self._upload(feats), self._upload(dirs)
        g = DeviceBuffer(self.NUM_PARAMS * 4, self.device)
        loss = C.c_double()
        capi.check(self.lib.ljx_guide_gradient_device(self.h, f.ptr, d.ptr, n, g.ptr, C.byref(loss)), self.lib)
        return loss.value, g.download(np.zeros(self.NUM_PARAMS, np.float32))

    def train_step(self, feats, dirs):
        n = len(feats)
        f, d = self._upload(feats), self._upload(dirs)
        loss = C.c_double()
        capi.check(self.lib.ljx_guide_train_step_device(self.h, f.ptr, d.ptr, n, C.byref(loss)), self.lib)
        return loss.value


def guide_options(mix_prob=0.5, train_spp=0, train_spp_per_pass=1, steps_per_pass=4, max_record_depth=0):
    o = capi.GuideOptions()
    o.mix_prob = mix_prob
    o.train_spp = train_spp
    o.train_spp_per_pass = train_spp_per_pass
    o.steps_per_pass = steps_per_pass
    o.max_record_depth = max_record_depth
    return o


def render_guided(scene: Scene, guide: Guide, options=None, exec_options=None, gopts=None, stats=False):
    """Guided render into host memory (the image is produced in HBM and copied once)."""
    W, H = scene.width, scene.height
    out = DeviceBuffer(W * H * 3 * 4, guide.device)
    st = capi.RenderStats()
    g = gopts if gopts is not None else guide_options()
    capi.check(scene.lib.ljx_render_guided_device(scene.h, guide.h, C.byref(options) if options is not None else None,
                                                  C.byref(exec_options) if exec_options is not None else None,
                                                  C.byref(g), out.ptr, C.byref(st)), scene.lib)
    img = out.download(np.zeros((H, W, 3), np.float32))
    if stats:
        return img, stats_dict(st)
    return img


def render_batch(jobs, exec_options=None):
    """Batched render: jobs is a list of dicts with scene (Scene), camera
    (camera dict), spp, max_depth, seed.  Returns (images, per-job completion
    ms from the launch start, stats dict)."""
    from .scenes import camera_desc
    lib = capi.load()
    n = len(jobs)
    arr = (capi.BatchJob * n)()
    outs = []
    for i, j in enumerate(jobs):
        cam = j["camera"]
        buf = DeviceBuffer(cam["width"] * cam["height"] * 3 * 4, j["scene"].device)
        outs.append(buf)
        a = arr[i]
        a.scene = j["scene"].h
        a.camera = camera_desc(cam)
        a.samples_per_pixel = j["spp"]
        a.max_depth = j.get("max_depth", 8)
        a.seed = j.get("seed", capi.DEFAULT_SEED)
        a.rgb_out_device = outs[-1].ptr
    ms = (C.c_double * n)()
    st = capi.BatchStats()
    capi.check(lib.ljx_render_batch_device(arr, n, C.byref(exec_options) if exec_options is not None else None, ms,
                                           C.byref(st)), lib)
    imgs = [b.download(np.zeros((j["camera"]["height"], j["camera"]["width"], 3), np.float32))
            for b, j in zip(outs, jobs)]
    return imgs, list(ms), {k: getattr(st, k) for k, _ in capi.BatchStats._fields_}


def autotune(scene: Scene, options=None):
    """Runs the launch-settings autotuner for a scene; returns the chosen ExecOptions."""
    best = capi.ExecOptions()
    capi.check(scene.lib.ljx_autotune(scene.h, C.byref(options) if options is not None else None, C.byref(best)),
               scene.lib)
    return best
