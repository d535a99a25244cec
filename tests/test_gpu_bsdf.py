"""BSDF eval / pdf / sample on the device against the reference (north star
(b): 1e-5 relative on seeded inputs).

Lambertian, RoughPlastic and RoughDielectric are checked against the
reference's own outputs (tests/golden/ref_materials.json).  The Disney
materials, stubs in the reference, are checked against the oracle's
restatement of the handout at the same tolerance.
"""
import numpy as np
import pytest

from helpers import bsdf_queries, expected_to_numpy, golden, material_scene, rel_err, results_to_numpy
from lajolla_public_amd import _capi as capi
from lajolla_public_amd import api, scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
TOL = 1e-5


def check(got, want):
    assert np.array_equal(got["valid"], want["valid"])
    ef = rel_err(got["f"], want["f"], 1e-300).max(axis=1)
    ep = rel_err(got["pdf"], want["pdf"], 1e-300)
    print(f"max rel err f={ef.max():.3g} pdf={ep.max():.3g}", flush=True)
    assert ef.max() < TOL and ep.max() < TOL
    v = want["valid"] == 1
    assert np.abs(got["sdir"][v] - want["sdir"][v]).max(initial=0) < TOL
    assert np.array_equal(got["seta"][v], want["seta"][v])


def test_reference_materials_on_device():
    mats = golden("materials")["materials"]
    sc = api.Scene(material_scene(mats))
    q, exp = bsdf_queries(mats)
    check(results_to_numpy(sc.bsdf(q)), expected_to_numpy(exp))


def disney_scene_and_queries(seed=4, n=400):
    rng = np.random.default_rng(seed)
    s = scenes.SceneDesc()
    types = [capi.MAT_DISNEY_DIFFUSE, capi.MAT_DISNEY_METAL, capi.MAT_DISNEY_GLASS, capi.MAT_DISNEY_CLEARCOAT,
             capi.MAT_DISNEY_SHEEN, capi.MAT_DISNEY_BSDF, capi.MAT_DISNEY_BSDF]
    for t in types:
        p = rng.uniform(0, 1, 12)
        s.add_material(t, tuple(rng.uniform(0, 1, 3)), eta=float(rng.uniform(1.1, 2.0)), roughness=p[0],
                       specular_transmission=p[1], metallic=p[2], subsurface=p[3], specular=p[4], specular_tint=p[5],
                       anisotropic=p[6], sheen=p[7], sheen_tint=p[8], clearcoat=p[9], clearcoat_gloss=p[10])
    P, I, N = scenes.quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0))
    s.add_mesh(P, I, 0, normals=N)
    arr = (capi.BSDFQuery * (n * len(types)))()
    k = 0
    for mi in range(len(types)):
        for j in range(n):
            ng = rng.normal(size=3)
            ng /= np.linalg.norm(ng)
            sn = ng + (0.3 * rng.normal(size=3) if j % 3 == 1 else 0)
            sn /= np.linalg.norm(sn)
            # Frisvad basis of sn
            if sn[2] < -1 + 1e-6:
                fx, fy = np.array([0, -1, 0.0]), np.array([-1, 0, 0.0])
            else:
                a = 1 / (1 + sn[2])
                b = -sn[0] * sn[1] * a
                fx = np.array([1 - sn[0] ** 2 * a, b, -sn[0]])
                fy = np.array([b, 1 - sn[1] ** 2 * a, -sn[1]])
            din = rng.normal(size=3)
            din /= np.linalg.norm(din)
            dout = rng.normal(size=3)
            dout /= np.linalg.norm(dout)
            if j % 2 == 0:
                din = din if din @ ng > 0 else -din
                dout = dout if dout @ ng > 0 else -dout
            q = arr[k]
            q.material_id = mi
            q.transport_to_view = 0
            q.dir_in = capi.D3(*din)
            q.dir_out = capi.D3(*dout)
            q.geometric_normal = capi.D3(*ng)
            q.frame_x = capi.D3(*fx)
            q.frame_y = capi.D3(*fy)
            q.frame_n = capi.D3(*sn)
            q.rnd_uv = (np.ctypeslib.ctypes.c_double * 2)(*rng.uniform(0, 1, 2))
            q.rnd_w = rng.uniform(0, 1)
            k += 1
    return s, arr


def test_disney_materials_on_device_vs_oracle():
    s, q = disney_scene_and_queries()
    sc = api.Scene(s)
    orc = po.OracleScene(s)
    check(results_to_numpy(sc.bsdf(q)), results_to_numpy(orc.bsdf(q)))
