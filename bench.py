"""Throughput benchmark of the wavefront path tracer on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one full render of the large-mesh workload: a seeded fBm
heightfield (1000x500 vertices) with 64 icospheres (subdivision 5) and 1024
emissive triangles (2,308,746 triangles), a constant environment,
Lambertian / RoughPlastic / DisneyBSDF materials, 1920x1080 at 64 samples per
pixel, max depth 8.  The scene and BVH are resident in HBM before timing; the
image stays in HBM.  Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=0, help="0: the workload's own spp")
    ap.add_argument("--workload", choices=["terrain_spheres", "nested_media", "hetvol_smoke"],
                    default="terrain_spheres",
                    help="terrain_spheres (default): the headline 1080p large-mesh workload; nested_media: the "
                         "1024x1024, 128 spp, depth 16 volumetric workload; hetvol_smoke: a 128^3 smoke grid "
                         "(heterogeneous medium), 768x576, 256 spp, unbounded depth with RR")
    ap.add_argument("--cpu-spp", type=int, default=0, help="spp of the bounded CPU baseline sample (0: auto)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--max-paths", type=int, default=0)
    ap.add_argument("--no-guide", action="store_true", help="skip the guided-vs-plain equal-time comparison")
    ap.add_argument("--no-batch", action="store_true", help="skip the 96-job batched-rendering run")
    ap.add_argument("--no-vol", action="store_true",
                    help="skip the volumetric workloads (nested media for each phase g, the smoke grid)")
    ap.add_argument("--guide-ref-spp", type=int, default=2048,
                    help="spp of the plain reference image the relMSE figures are measured against")
    ap.add_argument("--mode", choices=["render", "guided", "copycal"], default="render",
                    help="guided: guided renders of the guiding workload (terrain_gap 1080p, 64 spp, an untrained "
                         "network per step), with the matrix-core roofline of the inference kernel; copycal: one "
                         "plain FP32 elementwise kernel over 1 GiB (read) + 1 GiB (write), to scale the FETCH_SIZE / "
                         "WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); every box here has one GPU.",
              file=sys.stderr)
        sys.exit(2)
    if args.mode == "copycal":
        import torch
        n = (1 << 30) // 4
        a = torch.ones(n, device="cuda", dtype=torch.float32)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            torch.mul(a, 2.0, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"copycal_bytes_read": n * 4, "copycal_bytes_written": n * 4, "launches": 3}))
        return

    import numpy as np
    from lajolla_public_amd import _capi as capi
    from lajolla_public_amd import api, scenes

    if args.mode == "guided":
        guided_mode(args)
        return

    t0 = time.time()
    if args.workload == "terrain_spheres":
        desc = scenes.terrain_spheres(seed=7, spp=args.spp or 64)
        workload = ("terrain_spheres: fBm heightfield 1000x500 + 64 icospheres (subdiv 5) + 1024 emissive "
                    "triangles = 2,308,746 triangles, constant environment, 1920x1080, 64 spp, max depth 8")
        metric = "Msamples/s (1080p, depth 8)"
    elif args.workload == "hetvol_smoke":
        desc = scenes.hetvol_smoke(seed=3, spp=args.spp or 256, max_depth=-1, res=128, scale=40.0)
        workload = ("hetvol_smoke: 128^3 smoke density grid x 40 (heterogeneous medium, albedo 0.9, HG g=0) in an "
                    "index-matched box over a diffuse plane, sphere emitter, 768x576, 256 spp, unbounded depth, RR "
                    "from depth 5, volpath (delta and ratio tracking)")
        metric = "Msamples/s (768x576, volpath, heterogeneous)"
    else:
        desc = scenes.nested_media(seed=11, g=0.7, spp=args.spp or 128)
        workload = ("nested_media: 100,352-triangle index-matched sphere (homogeneous medium, HG g=0.7) inside a "
                    "RoughDielectric sphere with a second medium, 2 area lights, 1024x1024, 128 spp, max depth 16, "
                    "RR from depth 5, volpath")
        metric = "Msamples/s (1024x1024, depth 16, volpath)"
    print(f"[bench] scene generated in {time.time() - t0:.1f}s: {desc.num_triangles} triangles", file=sys.stderr,
          flush=True)
    sc = api.Scene(desc)
    bvh = sc.bvh_info()
    W, H, spp = desc.camera["width"], desc.camera["height"], desc.options["samples_per_pixel"]
    lib = sc.lib
    out = api.DeviceBuffer(W * H * 3 * 4)
    opts = scenes.render_options(**desc.options)

    def run(ex):
        st = capi.RenderStats()
        capi.check(lib.ljx_render_device(sc.h, C.byref(opts), C.byref(ex), C.cast(out.ptr, C.c_void_p),
                                         C.byref(st)), lib)
        return api.stats_dict(st)

    # instrumented replay: node visits / primitive tests per ray (not timed)
    inst = run(api.exec_options(collect_traversal=1, max_paths_in_flight=args.max_paths))
    print(f"[bench] instrumented: {json.dumps({k: inst[k] for k in ('extension_rays', 'node_visits', 'tri_tests', 'shadow_rays', 'shadow_node_visits', 'shadow_tri_tests', 'bvh_depth', 'build_ms')})}",
          file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        run(api.exec_options(max_paths_in_flight=args.max_paths))
    ex = api.exec_options(collect_timing=1, max_paths_in_flight=args.max_paths)
    stats = []
    t_start = time.perf_counter()
    for k in range(args.steps):
        stats.append(run(ex))
        print(f"[bench] step {k}: {stats[-1]['total_ms']:.1f} ms", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t_start
    ms_per_step = 1000 * elapsed / args.steps
    samples = W * H * spp
    value = samples * args.steps / elapsed / 1e6
    ext_rays = sum(s["extension_rays"] for s in stats)
    sh_rays = sum(s["shadow_rays"] for s in stats)
    grays = (ext_rays + sh_rays) / elapsed / 1e9

    # roofline of the dominant kernel (closest-hit traversal): algorithmic
    # bytes = 48 B per ray (32 B ray in, 16 B hit out) + 112 B per visited
    # 4-wide node (four child boxes and their references) + 36 B per
    # primitive test (three FP32 vertices)
    ext_ms = sum(s["extend_ms"] for s in stats)
    ext_launches = sum(s["extend_launches"] for s in stats)
    nodes_per_ray = inst["node_visits"] / max(1, inst["extension_rays"])
    tris_per_ray = inst["tri_tests"] / max(1, inst["extension_rays"])
    node_bytes = inst["node_bytes"]
    bytes_per_ray = 48 + node_bytes * nodes_per_ray + 36 * tris_per_ray
    alg_bytes = ext_rays * bytes_per_ray
    achieved_alg = alg_bytes / (ext_ms / 1000) / 1e9 if ext_ms > 0 else 0.0
    kernel_ms = {k: sum(s[k] for s in stats) / args.steps for k in ("extend_ms", "shadow_ms", "shade_ms", "other_ms")}
    alg_per_launch = alg_bytes / max(1, ext_launches)
    avg_launch_ms = ext_ms / max(1, ext_launches)
    traffic = measured_traffic(args.workload)
    hbm_gbs = traffic / (avg_launch_ms / 1000) / 1e9 if traffic and avg_launch_ms > 0 else None
    pipeline = measured_pipeline(args.workload)
    # the vector-memory pipeline's rate for incoherent fetches over this tree:
    # whole nodes at random indices, every lane its own node
    # (ljx_scene_node_fetch_rate); the traversal's node and primitive fetch
    # rate beside it, in bytes
    random_fetch_gbs = sc.node_fetch_rate() * node_bytes / 1e9
    fetch_gbs = (ext_rays * (node_bytes * nodes_per_ray + 48 * tris_per_ray)) / (ext_ms / 1000) / 1e9 \
        if ext_ms > 0 else 0.0

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(desc, args.cpu_spp)
    guiding = batch = volumetric = None
    if args.workload == "terrain_spheres":
        if not args.no_batch:
            batch = batched_jobs(sc, desc)
        del sc
        if not args.no_vol:
            volumetric = volumetric_workloads()
        if not args.no_guide:
            guiding = guided_vs_plain(args.guide_ref_spp)

    line = {
        "metric": metric,
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 traversal and path state between kernels (throughput, radiance); f64 shading and per-pixel sums",
        "data": "synthetic (seeded scene generator)",
        "config": {
            "workload": workload,
            "width": W, "height": H, "spp": spp, "max_depth": desc.options["max_depth"],
            "triangles": desc.num_triangles, "seed": desc.meta["seed"],
        },
        "grays_per_s": round(grays, 4),
        "scene_build": {"builder": "device" if bvh["builder"] == capi.BVH_BUILDER_GPU else "host",
                        "scene_build_ms": round(bvh["build_ms"], 2), "bvh_build_ms": round(bvh["bvh_ms"], 2),
                        "bvh_nodes": int(bvh["num_nodes"]), "bvh_depth": int(bvh["depth"]),
                        "sah_cost": round(bvh["sah_cost"], 4),
                        "note": "first scene creation in the process (includes one-time kernel loading); not in "
                                "the timed region"},
        "rays_per_step": {"extension": ext_rays // args.steps, "shadow": sh_rays // args.steps},
        "kernel_ms_per_step": {k: round(v, 2) for k, v in kernel_ms.items()},
        "roofline": {
            "kernel": "k_extend (closest-hit traversal)",
            "bound": "hbm",
            "achieved": round(hbm_gbs, 1) if hbm_gbs else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(hbm_gbs / HBM_PEAK_GBS, 4) if hbm_gbs else None,
            "traffic": traffic,
            "traffic_source": f"profiles/pmc_k_extend_{args.workload}.md (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE "
                              "x1.0 per dispatch, scaled by a copy kernel of known size)" if traffic else None,
            "avg_launch_ms": round(avg_launch_ms, 4),
            "launches_per_step": ext_launches // args.steps,
            "algorithmic_gbs": round(achieved_alg, 1),
            "algorithmic_bytes_per_launch": round(alg_per_launch),
            "traffic_over_algorithmic": round(traffic / alg_per_launch, 3) if traffic else None,
            "bytes_per_ray": round(bytes_per_ray, 1),
            "nodes_per_ray": round(nodes_per_ray, 2),
            "prims_per_ray": round(tris_per_ray, 2),
            "pipeline": pipeline,
            "fetch_pipe": {
                "traversal_fetch_gbs": round(fetch_gbs, 1),
                "random_fetch_gbs": round(random_fetch_gbs, 1),
                "ratio": round(fetch_gbs / random_fetch_gbs, 4) if random_fetch_gbs > 0 else None,
                "unit": "GB/s of node (node_bytes each) and primitive (48 B each) fetches",
                "note": "random_fetch_gbs: ljx_scene_node_fetch_rate on this scene's BVH, every lane of every wave "
                        "fetching whole nodes at random indices (no coherence), measured live. The traversal's "
                        "fetch rate is set against it: a ratio near or above 1 says the traversal moves node data "
                        "as fast as the vector-memory pipeline serves incoherent fetches (above 1: the shared upper "
                        "levels of the tree hit in cache)",
            },
            "note": "achieved and frac are HBM bytes measured by counters (traffic, per launch) over the live "
                    "HIP-event launch time, against the 8 TB/s HBM peak. algorithmic_gbs counts every node and "
                    "primitive fetch of every ray (48 B + node_bytes x nodes + 36 B x primitives per ray); the "
                    "caches serve most of those, so it is not an HBM figure. The unit that bounds the kernel is "
                    "the vector-memory pipeline (pipeline: texture address / data units busy fractions); fetch_pipe sets the "
                    "traversal's fetch rate beside that pipeline's measured rate for random node fetches.",
        },
        "cpu_baseline": cpu,
        "guiding": guiding,
        "batch": batch,
        "volumetric": volumetric,
    }
    if cpu and cpu.get("value"):
        line["gpu_over_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(line), flush=True)


MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16
GUIDE_FLOP_PER_QUERY = 2 * (160 * 64 + 3 * 64 * 64 + 64 * 40)  # 50,176
GUIDE_TRAIN_FLOP_PER_SAMPLE = 3 * GUIDE_FLOP_PER_QUERY  # forward, input deltas, weight gradients


def guided_mode(args):
    """Guided renders of terrain_gap at 1080p, 64 spp, depth 8, each step from
    an untrained network (8 recording passes, 6 Adam steps on 2^20-sample
    batches after each).  Roofline of k_guide_infer: 50,176 FLOP per query,
    one query per path that hit a guided material (the guided types'
    material buckets; the kernel counts them)."""
    from lajolla_public_amd import api, scenes
    desc = scenes.terrain_gap(seed=7, spp=args.spp or 64)
    sc = api.Scene(desc)
    W, H, spp = desc.camera["width"], desc.camera["height"], desc.options["samples_per_pixel"]
    opts = scenes.render_options(**desc.options)
    ex = api.exec_options(collect_timing=1)
    go = api.guide_options(mix_prob=0.5, train_spp=8, steps_per_pass=6)

    def step(seed):
        g = api.Guide(seed=seed)
        t = time.perf_counter()
        _, st = api.render_guided(sc, g, opts, exec_options=ex, gopts=go, stats=True)
        ms = (time.perf_counter() - t) * 1000
        g.close()
        return ms, st

    for k in range(args.warmup):
        step(1000 + k)
    runs = [step(k + 1) for k in range(args.steps)]
    ms = sum(r[0] for r in runs) / len(runs)
    st = [r[1] for r in runs]
    queries = sum(s["guide_queries"] for s in st)
    g_ms = sum(s["guide_ms"] for s in st)
    t_ms = sum(s["train_ms"] for s in st)
    steps = sum(s["train_steps"] for s in st)
    achieved = queries * GUIDE_FLOP_PER_QUERY / (g_ms / 1000) / 1e12
    line = {
        "metric": "Msamples/s (1080p, depth 8, guided, training included)",
        "value": round(W * H * spp / (ms / 1000) / 1e6, 3),
        "unit": "Msamples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 MFMA (fp32 accumulate) network; f32 traversal / f64 shading",
        "data": "synthetic (seeded scene generator, random-init network)",
        "config": {"workload": "terrain_gap 1920x1080, 64 spp, depth 8, guided (mix 0.5), 8 training passes x 6 "
                               "Adam steps", "width": W, "height": H, "spp": spp},
        "kernel_ms_per_step": {k: round(sum(s[k] for s in st) / len(st), 2)
                               for k in ("extend_ms", "shadow_ms", "shade_ms", "guide_ms", "train_ms", "other_ms")},
        "roofline": {
            "kernel": "k_guide_infer (MLP forward + lobes + sampling)", "bound": "mfma",
            "achieved": round(achieved, 1), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / MFMA_BF16_PEAK_TFLOPS, 4), "traffic": None,
            "flop_per_query": GUIDE_FLOP_PER_QUERY, "queries_per_step": queries // len(st),
            "extension_rays_per_step": sum(s["extension_rays"] for s in st) // len(st),
        },
        "training": {"adam_steps_per_step": steps // len(st),
                     "samples_per_step": sum(s["train_step_samples"] for s in st) // len(st),
                     "achieved_tflops_incl_collect": round(
                         sum(s["train_step_samples"] for s in st) * GUIDE_TRAIN_FLOP_PER_SAMPLE / (t_ms / 1000) / 1e12,
                         1) if t_ms > 0 else None},
    }
    print(json.dumps(line), flush=True)


def volumetric_workloads(steps=2):
    """The volumetric workloads beside the headline line: nested_media at its
    full size (1024x1024, 128 spp, depth 16, the 100,352-triangle medium
    sphere) for each Henyey-Greenstein g of the config, and hetvol_smoke
    (768x576, 256 spp, unbounded depth, 128^3 grid).  One warm-up render,
    then the best of `steps` renders (wall time of the render call)."""
    from lajolla_public_amd import api, scenes
    out = {}
    jobs = [(f"nested_media_g{g:+.1f}", scenes.nested_media(seed=11, g=g)) for g in (-0.3, 0.0, 0.7)]
    jobs.append(("hetvol_smoke", scenes.hetvol_smoke(seed=3, spp=256, max_depth=-1, res=128, scale=40.0)))
    for name, desc in jobs:
        sc = api.Scene(desc)
        ex = api.exec_options(collect_timing=1)
        sc.render(exec_options=ex)
        best = None
        for _ in range(steps):
            t = time.perf_counter()
            _, st = sc.render(exec_options=ex, stats=True)
            ms = (time.perf_counter() - t) * 1000
            if best is None or ms < best[0]:
                best = (ms, st)
        ms, st = best
        W, H, spp = desc.camera["width"], desc.camera["height"], desc.options["samples_per_pixel"]
        out[name] = {"msamples_per_s": round(W * H * spp / ms / 1e3, 2), "ms": round(ms, 1),
                     "width": W, "height": H, "spp": spp, "max_depth": desc.options["max_depth"],
                     "kernel_ms": {k: round(st[k], 1) for k in ("extend_ms", "shadow_ms", "shade_ms", "other_ms")}}
        print(f"[bench] {name}: {out[name]['msamples_per_s']} Msamples/s ({ms:.0f} ms)", file=sys.stderr, flush=True)
        sc.close()
    return out


def batched_jobs(sc_c1, desc_c1, n_jobs=96):
    """96 jobs over the large-mesh scene and the volumetric scene (depth 8),
    seeded orbit poses, resolutions log-uniform in [64, 1024] per side, spp
    from {4, 16, 64, 256}, all in one captured HIP graph with autotuned
    launch settings per scene."""
    import numpy as np
    from lajolla_public_amd import api, scenes
    t0 = time.time()
    desc_c2 = scenes.nested_media(seed=11, g=0.7, max_depth=8)
    sc_c2 = api.Scene(desc_c2)
    jobs = scenes.batch_jobs([(sc_c1, desc_c1), (sc_c2, desc_c2)], n_jobs=n_jobs, seed=3)
    t1 = time.time()
    tuned = {name: api.autotune(s) for name, s in (("terrain_spheres", sc_c1), ("nested_media", sc_c2))}
    t_tune = time.time() - t1
    print(f"[bench] batch: scenes ready in {t1 - t0:.1f}s, autotuned in {t_tune:.1f}s", file=sys.stderr, flush=True)
    api.render_batch(jobs[:4])  # warm-up of the batched kernels
    imgs, done_ms, st = api.render_batch(jobs)
    samples = st["samples"]
    done = np.array(done_ms)
    print(f"[bench] batch: {n_jobs} jobs, {samples / 1e6:.1f} Msamples in {st['launch_ms']:.0f} ms launch "
          f"(+{st['build_ms']:.0f} ms capture), {st['graph_nodes']} graph nodes", file=sys.stderr, flush=True)
    return {
        "workload": f"{n_jobs} jobs alternating terrain_spheres (2,308,746 triangles) and nested_media (volpath), "
                    "seeded orbit poses, W and H log-uniform in [64, 1024], spp in {4, 16, 64, 256}, depth 8",
        "value": round(samples / (st["launch_ms"] / 1000) / 1e6, 3),
        "unit": "Msamples/s (aggregate, graph launch time)",
        "samples": int(samples),
        "launch_ms": round(st["launch_ms"], 1),
        "graph_build_ms": round(st["build_ms"], 1),
        "graph_nodes": int(st["graph_nodes"]),
        "passes": int(st["passes"]),
        "autotune_s": round(t_tune, 2),
        "autotuned": {k: {"max_paths_in_flight": v.max_paths_in_flight, "block_size": v.block_size,
                          "sort_threshold": v.sort_threshold} for k, v in tuned.items()},
        "job_completion_ms": {"min": round(float(done.min()), 2), "median": round(float(np.median(done)), 2),
                              "max": round(float(done.max()), 2)},
        "per_job": [{"W": j["camera"]["width"], "H": j["camera"]["height"], "spp": j["spp"],
                     "scene": "terrain_spheres" if k % 2 == 0 else "nested_media", "done_ms": round(float(d), 2)}
                    for k, (j, d) in enumerate(zip(jobs, done))],
    }


def rel_mse(img, ref, trim=0.0):
    """mean((x - ref)^2 / (ref^2 + 0.01)) over pixels and channels; with trim,
    the largest per-pixel errors (that fraction of the pixels) are left out."""
    import numpy as np
    e = ((img.astype(np.float64) - ref) ** 2 / (ref.astype(np.float64) ** 2 + 1e-2)).mean(axis=2).ravel()
    if trim > 0:
        e = np.sort(e)[: int(round(len(e) * (1 - trim)))]
    return float(e.mean())


TRIM = 1e-4  # the 0.01% largest per-pixel errors (207 pixels at 1080p)


def guided_vs_plain(ref_spp, plain_spp=64, train_spp=4, steps_per_pass=3, plain_seeds=8):
    """relMSE at equal wall time, learned guiding vs the plain tracer, on the
    guiding workload (terrain_gap: the large-mesh scene under a ceiling with a
    5% gap, 90% of the light power behind it), 1920x1080, depth 8.  Each
    guided render starts from an untrained network: its training passes and
    Adam steps are inside the timed wall time, and their samples stay in the
    image.  relMSE = mean((x - ref)^2 / (ref^2 + 0.01)) against a plain
    render at ref_spp; the plain figure is the mean over plain_seeds renders.
    Both are also given with the 0.01% largest per-pixel errors left out,
    which shows how much of the standard figure is rare outliers."""
    import numpy as np
    from lajolla_public_amd import api, scenes
    t0 = time.time()
    desc = scenes.terrain_gap(seed=7)
    sc = api.Scene(desc)
    log = lambda m: print(f"[bench] guiding: {m}", file=sys.stderr, flush=True)
    log(f"scene ready in {time.time() - t0:.1f}s ({desc.num_triangles} triangles)")

    def opts(spp, seed):
        return scenes.render_options(**dict(desc.options, samples_per_pixel=spp, seed=seed))

    t0 = time.time()
    ref = sc.render(options=opts(ref_spp, 12345)).astype(np.float64)
    log(f"reference {ref_spp} spp in {time.time() - t0:.1f}s")
    sc.render(options=opts(4, 1))  # warm-up

    def guided(spp, seed):
        g = api.Guide(seed=seed)
        go = api.guide_options(mix_prob=0.5, train_spp=min(train_spp, spp), steps_per_pass=steps_per_pass)
        t = time.perf_counter()
        img, st = api.render_guided(sc, g, opts(spp, seed), gopts=go, stats=True)
        ms = (time.perf_counter() - t) * 1000
        g.close()
        return ms, img, st

    guided(16, 7)  # warm-up (first launch of the guided kernels)
    plain = []
    for k in range(plain_seeds):
        t = time.perf_counter()
        img = sc.render(options=opts(plain_spp, 1 + k))
        ms = (time.perf_counter() - t) * 1000
        plain.append({"seed": 1 + k, "ms": round(ms, 1), "relmse": round(rel_mse(img, ref), 6),
                      "relmse_trimmed": round(rel_mse(img, ref, TRIM), 6)})
        log(f"plain {plain_spp} spp seed {1 + k}: {ms:.0f} ms, relMSE {plain[-1]['relmse']:.5f} "
            f"(trimmed {plain[-1]['relmse_trimmed']:.5f})")
    p_ms = float(np.mean([p["ms"] for p in plain]))
    p_err = float(np.mean([p["relmse"] for p in plain]))
    p_trim = float(np.mean([p["relmse_trimmed"] for p in plain]))
    p_med = float(np.median([p["relmse"] for p in plain]))
    curve = []
    for spp in (8, 16, 24, 32, 48, 64):
        ms, img, st = guided(spp, 100 + spp)
        curve.append({"spp": spp, "ms": round(ms, 1), "relmse": round(rel_mse(img, ref), 6),
                      "relmse_trimmed": round(rel_mse(img, ref, TRIM), 6), "train_samples": st["train_samples"],
                      "train_steps": st["train_steps"]})
        log(f"guided {spp} spp: {ms:.0f} ms, relMSE {curve[-1]['relmse']:.5f} (trimmed "
            f"{curve[-1]['relmse_trimmed']:.5f}), {st['train_steps']} Adam steps on {st['train_samples']} samples")
    t = np.array([c["ms"] for c in curve])

    def at_time(key):
        e = np.array([c[key] for c in curve])
        return float(np.exp(np.interp(np.log(p_ms), np.log(t), np.log(e))))

    def time_to(key, target):
        e = np.array([c[key] for c in curve])
        if e.min() > target:
            return None
        k = int(np.argmax(e <= target))
        if k == 0:
            return float(t[0])
        f = (np.log(target) - np.log(e[k - 1])) / (np.log(e[k]) - np.log(e[k - 1]))
        return float(np.exp(np.log(t[k - 1]) + f * (np.log(t[k]) - np.log(t[k - 1]))))

    eq, eq_trim = at_time("relmse"), at_time("relmse_trimmed")
    t_match, t_match_trim = time_to("relmse", p_err), time_to("relmse_trimmed", p_trim)
    t_match_med = time_to("relmse", p_med)
    return {
        "workload": "terrain_gap: terrain_spheres (2,308,746 triangles) under a ceiling with a 5% gap; a 1x1 "
                    "emitter above the gap carries 90% of the light power; 1920x1080, depth 8",
        "reference_spp": ref_spp,
        "plain": {"spp": plain_spp, "ms": round(p_ms, 1), "relmse": round(p_err, 6),
                  "relmse_median": round(p_med, 6), "relmse_trimmed": round(p_trim, 6), "seeds": plain},
        "guided_curve": curve,
        "guided_relmse_at_equal_time": round(eq, 6),
        "relmse_ratio_plain_over_guided_equal_time": round(p_err / eq, 3),
        "guided_time_to_plain_relmse_ms": round(t_match, 1) if t_match else None,
        "time_ratio_guided_over_plain": round(t_match / p_ms, 3) if t_match else None,
        "time_ratio_guided_over_plain_median": round(t_match_med / p_ms, 3) if t_match_med else None,
        "trimmed": {"trim_fraction": TRIM, "guided_relmse_at_equal_time": round(eq_trim, 6),
                    "relmse_ratio_plain_over_guided_equal_time": round(p_trim / eq_trim, 3),
                    "time_ratio_guided_over_plain": round(t_match_trim / p_ms, 3) if t_match_trim else None},
        "mix_prob": 0.5,
        "train": {"train_spp": train_spp, "steps_per_pass": steps_per_pass, "batch": 1 << 20, "optimizer": "adam",
                  "lr": 1e-3},
    }


def measured_pipeline(workload):
    """Busy fractions of the vector-memory pipeline in k_extend from the
    committed rocprofv3 PMC summary of this workload, or None."""
    path = os.path.join(ROOT, "profiles", f"pmc_pipeline_{workload}.md")
    if not os.path.exists(path):
        return None
    out = {}
    for line in open(path):
        if line.startswith("busy k_extend "):
            k, v = line[len("busy k_extend "):].split(":")
            out[k.strip()] = float(v)
    return out or None


def measured_traffic(workload):
    """HBM bytes per k_extend launch from the committed rocprofv3 PMC summary
    of this workload, or None if there is none."""
    path = os.path.join(ROOT, "profiles", f"pmc_k_extend_{workload}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch k_extend:"):
            return float(line.split(":")[1])
    return None


def host_cpus():
    """CPUs this process may run on: the affinity mask, capped by a cgroup v2
    CPU quota where one is set; and nproc as the OS reports it."""
    usable = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    if quota:
        usable = max(1, min(usable, int(quota)))
    return usable, os.cpu_count(), quota


def cpu_baseline(desc, cpu_spp):
    """The oracle (lajolla's path_render / path_tracing restated in FP64 with
    the reference's tile RNG plan, CPU SAH BVH in place of Embree) on every
    host CPU this process may use, over every tile of the same image at
    reduced spp."""
    from lajolla_public_amd import scenes
    from oracle import pyoracle as po
    threads, nproc, quota = host_cpus()
    t0 = time.time()
    orc = po.OracleScene(desc)
    build_s = time.time() - t0
    spp = cpu_spp
    if spp <= 0:
        # ~0.25 Msamples/s per host thread on this scene: aim for ~15 s
        spp = max(1, min(desc.options["samples_per_pixel"], int(round(15 * 0.25e6 * threads /
                                                                          (desc.camera["width"] * desc.camera["height"])))))
    o = dict(desc.options)
    o["samples_per_pixel"] = spp
    t0 = time.time()
    img, cnt = orc.render(options=scenes.render_options(**o), plan=0, threads=threads)
    dt = time.time() - t0
    return {
        "value": round(cnt["paths"] / dt / 1e6, 4),
        "unit": "Msamples/s",
        "cores": threads,
        "nproc": nproc,
        "cgroup_cpu_quota": quota,
        "kind": "port",
        "sample": f"every 16x16 tile of the {desc.camera['width']}x{desc.camera['height']} image at {spp} spp "
                  f"({cnt['paths']} camera paths, "
                  f"{dt:.1f} s; BVH build {build_s:.1f} s not timed); lajolla-algorithm CPU baseline "
                  f"(project BVH, not Embree)",
        "grays_per_s": round((cnt["ext_rays"] + cnt["shadow_rays"]) / dt / 1e9, 5),
    }


if __name__ == "__main__":
    main()
