"""The BVH built on the device equals the host builder's, byte for byte.

Both builders take their split decisions from the same code
(csrc/ljx_bvh_sah.h, no FP contraction) and partition stably, so on the same
primitive boxes the 4-wide nodes (in depth-first order) and the leaf-slot
table must be identical.  ljx_scene_bvh_info fingerprints both (FNV-1a) and
reports the SAH cost, and checks the traversal's quantised nodes against the
FP32 nodes they are made from.  Cases: the Cornell box, spheres with meshes, the
2,308,746-triangle terrain scene, the 100,352-triangle sphere of the nested
media scene, degenerate input (many triangles with one centroid, which has no
binned split), a single triangle (the root is a leaf) and a range just above
one wave's share.
"""
import numpy as np
import pytest

from lajolla_public_amd import _capi as capi
from lajolla_public_amd import api, scenes

pytestmark = pytest.mark.gpu


def both(desc):
    g = api.Scene(desc, bvh_builder=capi.BVH_BUILDER_GPU)
    gi = g.bvh_info()
    g.close()
    h = api.Scene(desc, bvh_builder=capi.BVH_BUILDER_HOST)
    hi = h.bvh_info()
    h.close()
    return gi, hi


def check_same(desc, label):
    gi, hi = both(desc)
    print(f"{label}: {gi['num_slots']} prims, {gi['num_nodes']} nodes, depth {gi['depth']}, stack {gi['stack_need']}, "
          f"SAH {gi['sah_cost']:.3f}; device build {gi['bvh_ms']:.1f} ms (scene {gi['build_ms']:.1f} ms), "
          f"host build {hi['bvh_ms']:.1f} ms (scene {hi['build_ms']:.1f} ms)", flush=True)
    assert gi["builder"] == capi.BVH_BUILDER_GPU and hi["builder"] == capi.BVH_BUILDER_HOST
    for k in ("num_nodes", "num_slots", "depth", "stack_need", "binary_depth", "node_hash", "slot_hash"):
        assert gi[k] == hi[k], (label, k, gi[k], hi[k])
    assert gi["sah_cost"] == hi["sah_cost"]
    # the traversal's quantised copy (made on the device after either build):
    # the same bytes, and every half-precision box plane outside or on the
    # FP32 plane it stands for, at most ~2^-10 of its node's extent away
    assert gi["qnode_hash"] == hi["qnode_hash"], label
    assert gi["qnode_bad_planes"] == 0 and hi["qnode_bad_planes"] == 0, (label, gi["qnode_bad_planes"])
    assert 0 <= gi["qnode_max_loose"] <= 2.0 ** -9, (label, gi["qnode_max_loose"])
    return gi, hi


def test_cornell_box():
    check_same(scenes.cornell_box(seed=3), "cbox")


def test_spheres_and_meshes():
    rng = np.random.default_rng(2)
    s = scenes.SceneDesc()
    m = s.add_material(0, (0.5, 0.5, 0.5))
    for k in range(40):
        s.add_sphere(rng.uniform(-3, 3, 3), rng.uniform(0.2, 0.8), m)
    P, I = scenes.icosphere(3)
    for k in range(6):
        s.add_mesh(P * 0.7 + rng.uniform(-3, 3, 3), I, m, normals=P)
    check_same(s, "spheres+meshes")


def test_terrain_spheres_full():
    gi, hi = check_same(scenes.terrain_spheres(seed=7), "terrain_spheres")
    assert gi["num_slots"] == 2_308_746
    # the device build is the faster one at this size
    assert gi["bvh_ms"] < hi["bvh_ms"]


def test_nested_media_sphere():
    check_same(scenes.nested_media(seed=11, g=0.7, width=64, height=64, spp=4), "nested_media")


def test_degenerate_centroids():
    # 2000 tiny triangles around one point: every centroid coincides, so
    # ranges are halved in order down to the leaf size
    s = scenes.SceneDesc()
    m = s.add_material(0, (0.5, 0.5, 0.5))
    n = 2000
    P = np.tile(np.array([[-1e-3, 0, 0], [1e-3, 0, 0], [0, 0, 0]], np.float32), (n, 1))
    I = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    s.add_mesh(P, I, m)
    gi, _ = check_same(s, "degenerate")
    assert gi["num_slots"] == n


def test_single_triangle_and_one_wave_ranges():
    s = scenes.SceneDesc()
    m = s.add_material(0, (0.5, 0.5, 0.5))
    s.add_mesh(np.array([[-1, -1, -1], [1, -1, -1], [0, 1, -1]], np.float32), np.array([[0, 1, 2]]), m)
    gi, _ = check_same(s, "one triangle")
    assert gi["num_nodes"] == 1
    rng = np.random.default_rng(5)
    for n in (8, 9, 511, 512, 513, 2049):
        s = scenes.SceneDesc()
        m = s.add_material(0, (0.5, 0.5, 0.5))
        c = rng.uniform(-5, 5, (n, 1, 3))
        P = (c + rng.normal(scale=0.1, size=(n, 3, 3))).reshape(-1, 3).astype(np.float32)
        s.add_mesh(P, np.arange(3 * n, dtype=np.int32).reshape(n, 3), m)
        check_same(s, f"{n} triangles")


def test_node_fetch_rate():
    """The random node-fetch rate: positive, higher with every lane of a
    wave fetching than with eight, and argument errors cross the boundary."""
    sc = api.Scene(scenes.terrain_spheres(seed=7, width=64, height=64, spp=1, grid=(200, 100), n_spheres=8,
                                          sphere_subdiv=3, n_lights=64))
    r64, r8 = sc.node_fetch_rate(64), sc.node_fetch_rate(8)
    print(f"node fetches/s: 64 lanes {r64:.4g}, 8 lanes {r8:.4g}", flush=True)
    assert np.isfinite(r64) and r64 > r8 > 0
    for bad in (0, 65):
        with pytest.raises(capi.LjxError):
            sc.node_fetch_rate(bad)
