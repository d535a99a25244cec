"""C-ABI library checks that need no GPU: it loads, it exports every entry
point include/lajolla_amd.h declares, and the ctypes mirror has the header's
struct layout."""
import ctypes as C
import os
import re
import subprocess

import pytest

from lajolla_public_amd import _capi as capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lajolla_amd.h")


def header_symbols():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^int (ljx_\w+)\(", src, re.M)))


def test_header_and_mirror_agree():
    assert header_symbols() == sorted(capi.SIGNATURES)


def test_library_exports_every_symbol():
    lib = capi.load()
    for name in header_symbols():
        assert hasattr(lib, name), name
    assert lib.ljx_abi_version() == 1


STRUCTS = ["ljx_shape_desc", "ljx_material_desc", "ljx_light_desc", "ljx_medium_desc", "ljx_volume_desc",
           "ljx_texture_desc", "ljx_camera_desc",
           "ljx_render_options", "ljx_scene_desc", "ljx_ray", "ljx_hit", "ljx_bsdf_query", "ljx_bsdf_result",
           "ljx_render_stats", "ljx_exec_options", "ljx_guide_config", "ljx_guide_options", "ljx_batch_job",
           "ljx_batch_stats", "ljx_build_options", "ljx_bvh_info"]
MIRROR = {"ljx_shape_desc": capi.ShapeDesc, "ljx_material_desc": capi.MaterialDesc, "ljx_light_desc": capi.LightDesc,
          "ljx_medium_desc": capi.MediumDesc, "ljx_volume_desc": capi.VolumeDesc,
          "ljx_texture_desc": capi.TextureDesc, "ljx_camera_desc": capi.CameraDesc,
          "ljx_render_options": capi.RenderOptions, "ljx_scene_desc": capi.SceneDescC, "ljx_ray": capi.Ray,
          "ljx_hit": capi.Hit, "ljx_bsdf_query": capi.BSDFQuery, "ljx_bsdf_result": capi.BSDFResult,
          "ljx_render_stats": capi.RenderStats, "ljx_exec_options": capi.ExecOptions,
          "ljx_guide_config": capi.GuideConfig, "ljx_guide_options": capi.GuideOptions,
          "ljx_batch_job": capi.BatchJob, "ljx_batch_stats": capi.BatchStats,
          "ljx_build_options": capi.BuildOptions, "ljx_bvh_info": capi.BvhInfo}


def test_struct_sizes_match_c(tmp_path):
    prog = tmp_path / "sizes.c"
    prog.write_text('#include <stdio.h>\n#include "lajolla_amd.h"\nint main(void){\n' +
                    "".join(f'printf("%zu\\n", sizeof({s}));\n' for s in STRUCTS) + "return 0;}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I" + os.path.join(ROOT, "include"), str(prog), "-o", str(exe)], check=True)
    sizes = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    for s, n in zip(STRUCTS, sizes):
        assert C.sizeof(MIRROR[s]) == n, s


def test_errors_cross_the_boundary_as_status_codes():
    lib = capi.load()
    # NULL descriptor / handle: an error status and a message, no exception
    rc = lib.ljx_scene_create(None, 0, None)
    assert rc != 0
    buf = C.create_string_buffer(256)
    assert lib.ljx_last_error(buf, 256) == 0
    assert len(buf.value) > 0
    assert lib.ljx_scene_bounds(None, None) != 0
    with pytest.raises(capi.LjxError):
        capi.check(lib.ljx_render(None, None, None, None, None), lib)
