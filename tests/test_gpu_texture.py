"""Textures, mipmaps, camera-ray footprints and image environment maps on the
device, against the oracle under the same per-path PCG streams:

  * the textured scene (checkerboard and image colour textures, a roughness
    texture, Disney metallic and base-colour textures, a coloured dielectric
    transmittance, an image environment map with a sun): at least 97% of
    pixels within 1e-3, as for the untextured path tracer;
  * the same scene without the area light (lit by the environment map alone,
    so its TableDist2D sampling and MIS carry the image);
  * the mipmap-level view of the textured scene (camera-ray footprints) pixel
    for pixel; a striped plane whose camera-ray footprints select levels
    around 2 (trilinear) or below 0; a textured scene through the volumetric
    path tracer;
  * a batched textured job equals its single render.
"""
import numpy as np
import pytest

from lajolla_public_amd import _capi as capi
from lajolla_public_amd import api, scenes
from oracle import pyoracle as po
from test_gpu_render import sample_matched
from test_oracle_texture import striped_plane

pytestmark = pytest.mark.gpu


def test_textured_scene_sample_matched():
    desc = scenes.textured_showcase(width=96, height=72, spp=16)
    g = api.Scene(desc).render()
    o, _ = po.OracleScene(desc).render(plan=1)
    assert np.isfinite(g).all()
    sample_matched(g, o, frac=0.97)
    assert abs(g.mean() - o.mean()) < 0.02 * o.mean()


def test_envmap_only_sample_matched():
    desc = scenes.textured_showcase(width=80, height=60, spp=16)
    # drop the area light: the image environment map is the only light
    desc.lights = [l for l in desc.lights if l["type"] == capi.LIGHT_ENVMAP]
    desc.envmap_light_id = 0
    for sh in desc.shapes:
        if sh["area_light_id"] >= 0:
            sh["area_light_id"] = -1
    g = api.Scene(desc).render()
    o, _ = po.OracleScene(desc).render(plan=1)
    sample_matched(g, o, frac=0.97)
    assert abs(g.mean() - o.mean()) < 0.02 * o.mean()


def test_mipmap_level_view_matches_oracle():
    desc = scenes.textured_showcase(width=96, height=72, spp=1)
    opts = dict(desc.options, integrator=capi.INTEGRATOR_MIPMAP_LEVEL, samples_per_pixel=1)
    g = api.Scene(desc).render(options=scenes.render_options(**opts)).astype(np.float64)
    desc.options.update(opts)
    o = po.OracleScene(desc).render()[0]
    ok = (np.abs(g - o) <= 1e-5 * np.abs(o) + 1e-5).all(axis=2)
    print(f"mipmap-level view: {ok.mean():.4f} of pixels match; levels {o.min():.2f} .. {o.max():.2f}", flush=True)
    assert ok.mean() > 0.995
    assert (o != 0).mean() > 0.05  # the image-textured spheres are in view


@pytest.mark.parametrize("depth,S", [(3.0, 0.25), (0.2, 8.0)])
def test_camera_ray_footprint_on_device(depth, S):
    """Levels around 2 (far: trilinear between levels 1, 2 and 3) and below 0
    (near: level 0) for the first vertex of each camera ray."""
    desc = striped_plane(depth, True, S=S)
    g = api.Scene(desc).render()
    o, _ = po.OracleScene(desc).render(plan=1)
    sample_matched(g, o, frac=0.99)


def test_textured_volpath_sample_matched():
    desc = scenes.textured_showcase(width=64, height=48, spp=8, envmap=False)
    med = desc.add_medium((0.05, 0.05, 0.05), (0.2, 0.25, 0.3), capi.PHASE_HENYEY_GREENSTEIN, 0.3)
    desc.camera["medium_id"] = med
    desc.options.update(integrator=capi.INTEGRATOR_VOLPATH)
    g = api.Scene(desc).render()
    o, _ = po.OracleScene(desc).render(plan=1)
    sample_matched(g, o, frac=0.90)
    assert abs(g.mean() - o.mean()) < 0.03 * o.mean()


def test_batched_textured_job_equals_single_render():
    desc = scenes.textured_showcase(width=64, height=48, spp=8)
    sc = api.Scene(desc)
    single = sc.render()
    imgs, _, _ = api.render_batch([dict(scene=sc, camera=dict(desc.camera), spp=8,
                                        max_depth=desc.options["max_depth"], seed=desc.options["seed"])])
    np.testing.assert_allclose(imgs[0], single, rtol=2e-5, atol=1e-6)
