"""The C ABI without a GPU: libdrawhip.so loads, exports every function that
include/drawhip.h declares, and its host-side uniforms (Camera::new +
gen_transformation_matrix + gen_view_planes) equal the oracle's bit for bit."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from draw_amd import _lib, workloads
from draw_amd.render import frame_uniforms
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "drawhip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dh_[a-z_0-9]+)\s*\(", text)))


def test_header_symbols_exported():
    L = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), f"{n} declared in drawhip.h but not exported"
    assert set(names) <= set(_lib.SIGNATURES), "ctypes signatures missing for some declared symbols"


def test_structs_match_header_sizes():
    assert C.sizeof(_lib.DhFrameParams) == 208
    assert C.sizeof(_lib.DhFrameParams) == C.sizeof(O.OrFrame)
    assert C.sizeof(_lib.DhCamera) == 36


def test_struct_layouts_match_c_compiler(tmp_path):
    """sizeof/offsetof of every ABI struct as gcc lays out include/drawhip.h
    equal the ctypes mirrors."""
    structs = {"dh_frame_params": _lib.DhFrameParams, "dh_camera": _lib.DhCamera,
               "dh_material_desc": _lib.DhMaterialDesc, "dh_stats": _lib.DhStats, "dh_view": _lib.DhView,
               "dh_server_config": _lib.DhServerConfig, "dh_job": _lib.DhJob, "dh_job_info": _lib.DhJobInfo,
               "dh_server_stats": _lib.DhServerStats, "dh_obj_summary": _lib.DhObjSummary,
               "dh_obj_mesh_desc": _lib.DhObjMeshDesc, "dh_obj_texture_desc": _lib.DhObjTextureDesc,
               "dh_gui_vertex": _lib.DhGuiVertex, "dh_gui_cmd": _lib.DhGuiCmd}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "drawhip.h"', "int main(void) {"]
    for cname, ct in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in ct._fields_:
            cf = {"near_": "near_", "far_": "far_"}.get(fname, fname)
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {cf}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, ct in structs.items():
        assert got[(cname, "size")] == C.sizeof(ct), cname
        for fname, _ in ct._fields_:
            assert got[(cname, fname)] == getattr(ct, fname).offset, (cname, fname)


def test_frame_uniforms_bit_exact_vs_oracle():
    rng = np.random.default_rng(11)
    scenes = [workloads.torus_scene(1), workloads.soup_scene(2, n_tri=10)]
    for _ in range(30):
        s = workloads.torus_scene(1, width=int(rng.integers(8, 4000)), height=int(rng.integers(8, 2200)))
        s.camera = workloads.random_camera(rng, dist=(1.0, 400.0))
        s.camera.fov_x = float(rng.uniform(5, 170))
        s.camera.near = -float(rng.uniform(0.01, 20))
        s.camera.far = s.camera.near - float(rng.uniform(1, 1000))
        s.light_source = tuple(rng.standard_normal(3) * 100)
        scenes.append(s)
    for s in scenes:
        assert bytes(frame_uniforms(s)) == bytes(O.frame_uniforms(s))


def test_frame_uniforms_reference_defaults():
    """The reference's own constants: camera (0,0,150), near -10, far -510,
    fov_x 135 (scene/mod.rs:308-310, 763-766)."""
    L = _lib.lib()
    cam = _lib.DhCamera((C.c_float * 3)(0, 0, 150), (C.c_float * 3)(0, 0, -150), 135.0, -10.0, -510.0)
    fr = _lib.DhFrameParams()
    assert L.dh_frame_uniforms(C.byref(cam), 800, 600, None, C.byref(fr)) == 0
    assert list(fr.light) == [0.0, 300.0, 300.0]
    assert fr.clear_depth == 100000.0 and list(fr.clear_bgra) == [255, 186, 155, 255]


def test_frame_uniforms_errors():
    L = _lib.lib()
    fr = _lib.DhFrameParams()
    bad = _lib.DhCamera((C.c_float * 3)(0, 0, 3), (C.c_float * 3)(0, 0, -3), 60.0, 0.5, -10.0)   # near >= 0
    assert L.dh_frame_uniforms(C.byref(bad), 64, 64, None, C.byref(fr)) == -4
    assert L.dh_frame_uniforms(None, 64, 64, None, C.byref(fr)) == -1


def test_context_without_device_is_unsupported():
    """No GPU in this container: dh_ctx_create reports DH_E_UNSUPPORTED instead
    of falling back to anything."""
    from draw_amd._lib import lib
    L = lib()
    ctx = C.c_void_p()
    rc = L.dh_ctx_create(0, C.byref(ctx))
    if rc == 0:       # a GPU is present (this test also runs on a GPU box)
        assert L.dh_ctx_destroy(ctx) == 0
    else:
        assert rc == -9


def test_renderer_fails_loudly_without_device():
    from draw_amd._lib import DrawHipError, lib
    from draw_amd.render import Renderer
    ctx = C.c_void_p()
    if lib().dh_ctx_create(0, C.byref(ctx)) == 0:
        lib().dh_ctx_destroy(ctx)
        pytest.skip("a GPU is present")
    with pytest.raises(DrawHipError):
        Renderer(0)


def test_frame_uniforms_canvas_bit_exact_vs_oracle():
    """dh_frame_uniforms_canvas (screen size for the camera, a canvas of its
    own size and offset) equals the oracle's restatement bit for bit."""
    import copy
    rng = np.random.default_rng(12)
    for _ in range(30):
        sw, sh = int(rng.integers(64, 4000)), int(rng.integers(64, 2200))
        s = workloads.torus_scene(1, width=sw, height=sh)
        s.camera = workloads.random_camera(rng, dist=(1.0, 400.0))
        s.canvas = (int(rng.integers(1, sw + 1)), int(rng.integers(1, sh + 1)))
        s.canvas_offset = (int(rng.integers(-100, sw)), int(rng.integers(-100, sh)))
        assert bytes(frame_uniforms(s)) == bytes(O.frame_uniforms(s))
    app = workloads.app_scene(6, n_tri=10)
    fr = frame_uniforms(app)
    assert (fr.W, fr.H, fr.off_x, fr.off_y) == (800, 600, 560.0, 240.0)
    plain = copy.copy(app)
    plain.canvas = None
    assert bytes(frame_uniforms(plain))[:184] == bytes(fr)[:184]
