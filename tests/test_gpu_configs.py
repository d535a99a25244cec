"""The configs at their real geometry, sample-matched to the oracle.

Each render uses per-path PCG streams on both sides (the stream of path id
sample * W * H + pixel), so a pixel's device estimate equals the oracle's up to FP32/FP64
rounding except where rounding flips a discrete decision:

  * the large-mesh scene with every one of its 2,308,746 triangles, its
    camera at 96x54, 4 spp, depth 8;
  * the volumetric scene with its full 100,352-triangle medium sphere, for
    each Henyey-Greenstein g of the config (-0.3, 0, 0.7), 32x32, 4 spp;
  * the tent and Gaussian pixel filters (src/filters/tent.inl,
    gaussian.inl) on the Cornell box;
  * batched jobs on both config scenes (full geometry) equal their single
    renders.
"""
import copy

import numpy as np
import pytest

from lajolla_public_amd import _capi as capi
from lajolla_public_amd import api, scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def matched_fraction(img_g, img_o):
    diff = np.abs(img_g.astype(np.float64) - img_o)
    return float((diff <= 1e-3 * np.abs(img_o) + 1e-4).all(axis=2).mean())


def test_terrain_full_geometry_sample_matched():
    desc = scenes.terrain_spheres(seed=7, width=96, height=54, spp=4)
    assert desc.num_triangles == 2_308_746
    g = api.Scene(desc).render()
    o, _ = po.OracleScene(desc).render(plan=1)
    assert np.isfinite(g).all()
    f = matched_fraction(g, o)
    print(f"terrain (full geometry) sample-matched pixels {f:.4f}; means {g.mean():.6g} {o.mean():.6g}", flush=True)
    assert f >= 0.95
    assert abs(g.mean() - o.mean()) < 0.02 * o.mean()


@pytest.mark.parametrize("g", [-0.3, 0.0, 0.7])
def test_nested_media_full_sphere_sample_matched(g):
    desc = scenes.nested_media(seed=11, g=g, width=32, height=32, spp=4)
    assert desc.shapes[0]["indices"].shape[0] == 100_352
    img = api.Scene(desc).render()
    o, _ = po.OracleScene(desc).render(plan=1)
    assert np.isfinite(img).all()
    f = matched_fraction(img, o)
    print(f"nested media g={g}: sample-matched pixels {f:.4f}; means {img.mean():.6g} {o.mean():.6g}", flush=True)
    assert f >= 0.90
    assert abs(img.mean() - o.mean()) < 0.03 * o.mean()


@pytest.mark.parametrize("ftype,param", [(capi.FILTER_TENT, 2.0), (capi.FILTER_TENT, 1.0),
                                         (capi.FILTER_GAUSSIAN, 0.5)])
def test_pixel_filters_sample_matched(ftype, param):
    desc = scenes.cornell_box(seed=1, width=48, height=48, spp=8, max_depth=6)
    desc.camera.update(filter_type=ftype, filter_param=param)
    g = api.Scene(desc).render()
    o, _ = po.OracleScene(desc).render(plan=1)
    f = matched_fraction(g, o)
    print(f"filter {ftype} ({param}): sample-matched pixels {f:.4f}", flush=True)
    assert f >= 0.97
    # the filter is in effect: the box-filtered image differs
    desc_box = copy.deepcopy(desc)
    desc_box.camera.update(filter_type=capi.FILTER_BOX, filter_param=1.0)
    b = api.Scene(desc_box).render()
    assert np.abs(b - g).max() > 1e-3


def test_batch_jobs_on_config_scenes():
    c1 = scenes.terrain_spheres(seed=7, width=64, height=64, spp=4)
    c2 = scenes.nested_media(seed=11, g=-0.3, width=64, height=64, spp=4, max_depth=8)
    s1, s2 = api.Scene(c1), api.Scene(c2)
    jobs = scenes.batch_jobs([(s1, c1), (s2, c2)], n_jobs=8, seed=5, res_range=(24, 96), spps=(1, 2, 4), max_depth=8)
    imgs, ms, st = api.render_batch(jobs, api.exec_options(max_paths_in_flight=30000))
    print(st, flush=True)
    assert st["passes"] >= 2
    for k, (j, img) in enumerate(zip(jobs, imgs)):
        desc = copy.deepcopy(c1 if k % 2 == 0 else c2)
        desc.camera = j["camera"]
        desc.options.update(samples_per_pixel=j["spp"], max_depth=j["max_depth"], seed=j["seed"])
        sc = api.Scene(desc)
        single = sc.render()
        sc.close()
        assert img.shape == single.shape
        np.testing.assert_allclose(img, single, rtol=2e-5, atol=1e-6)
