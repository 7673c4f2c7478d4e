"""The C ABI library: it loads, exports every function include/sharpray.h declares, and its
host-side error paths behave like the reference's exceptions. No GPU is used here: the checks run
on a host-only context (device -1)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "sharpray.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(sr_\w+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from sharpray_amd import _native as N
    handle = ctypes.CDLL(N.LIB_PATH)
    names = _declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(handle, name), name
    assert set(names) == set(N.SIGNATURES), "ctypes signatures cover exactly the declared ABI"


def test_gpu_entry_points_fail_loudly_without_a_device():
    from sharpray_amd import _native as N
    L = N.lib()
    ctx = ctypes.c_void_p()
    assert L.sr_create(-1, ctypes.byref(ctx)) == N.SR_OK
    try:
        v = N.View_()
        p = N.RenderParams_(8, 8, 32, 1, 1, 0.0, 1, 10.0, 0, 0, 1)
        assert L.sr_render_begin(ctx, ctypes.byref(v), ctypes.byref(p)) == -2
        assert b"host-only" in L.sr_last_error()
        s, t = ctypes.c_uint32(), ctypes.c_uint32()
        assert L.sr_render_tick(ctx, ctypes.byref(s), ctypes.byref(t)) == -2
    finally:
        L.sr_destroy(ctx)


def test_scene_errors_match_reference_exceptions():
    """scene.cs:525-526 ("Scene sky is required"), :543-571 ("Scene already built"), idempotent Build."""
    from sharpray_amd import _native as N
    L = N.lib()
    ctx = ctypes.c_void_p()
    assert L.sr_create(-1, ctypes.byref(ctx)) == N.SR_OK
    try:
        assert L.sr_scene_build(ctx) == -2
        assert b"sky is required" in L.sr_last_error()
        tex = np.ones((4, 8, 3), np.float32)
        tid = ctypes.c_int32()
        assert L.sr_texture_create(ctx, N.fptr(tex), 8, 4, 0, 1.0, 1.0, ctypes.byref(tid)) == 0
        assert L.sr_texture_create(ctx, N.fptr(tex), 0, 4, 0, 1.0, 1.0, ctypes.byref(tid)) == -1
        assert L.sr_scene_set_sky_texture(ctx, 5, 1.0, 0.0) == -1
        assert L.sr_scene_set_sky_texture(ctx, tid.value, 1.0, 0.0) == 0
        tr = N.Transform_((0, 0, 0), (0, 0, 0, 1), (1, 1, 1))
        m = N.Material_((0.5, 0.5, 0.5), 1.0, 0.0, 0.0, 1.5, (0, 0, 0), -1, -1, -1, -1)
        sphere = np.array([0, 0, 0, 1], np.float32)
        assert L.sr_scene_add_object(ctx, ctypes.byref(tr), ctypes.byref(m), 1, -1, N.fptr(sphere), None) == 0
        assert L.sr_scene_add_object(ctx, ctypes.byref(tr), ctypes.byref(m), 0, 3, None, None) == -1
        assert L.sr_scene_build(ctx) == 0
        assert L.sr_scene_build(ctx) == 0  # Build is idempotent (renderer.cs:77 calls it again)
        assert L.sr_scene_set_sky_texture(ctx, tid.value, 1.0, 0.0) == -2
        assert b"already built" in L.sr_last_error()
        assert L.sr_scene_add_object(ctx, ctypes.byref(tr), ctypes.byref(m), 1, -1, N.fptr(sphere), None) == -2
        nc, ic = ctypes.c_uint32(), ctypes.c_uint32()
        assert L.sr_bvh_export(ctx, -1, None, None, ctypes.byref(nc), ctypes.byref(ic)) == 0
        assert nc.value == 1 and ic.value == 1
    finally:
        L.sr_destroy(ctx)


def test_mesh_rejects_bad_arguments():
    from sharpray_amd import _native as N
    L = N.lib()
    ctx = ctypes.c_void_p()
    L.sr_create(-1, ctypes.byref(ctx))
    try:
        out = ctypes.c_int32()
        assert L.sr_mesh_create(ctx, None, 5, 8, ctypes.byref(out)) == -1
        tris = np.zeros((1, 36), np.float32)
        assert L.sr_mesh_create(ctx, N.fptr(tris), 1, 1, ctypes.byref(out)) == -1  # needs >= 2 bins
        assert L.sr_mesh_create(ctx, N.fptr(tris), 0, 8, ctypes.byref(out)) == 0   # empty mesh is allowed
    finally:
        L.sr_destroy(ctx)


def test_python_host_api_errors():
    import sharpray_amd as sr
    with pytest.raises(ValueError):
        sr.Transform(scale=(1, 0, 1))
    with pytest.raises(ValueError):
        sr.View(fov=4.0)
    with pytest.raises(ValueError):
        sr.Sphere((0, 0, 0), 0.0)
