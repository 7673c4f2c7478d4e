"""ctypes wrapper of the CPU oracle (oracle/sr_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by bench.py's
cpu_baseline leg, as the checker. The product (sharpray_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
COUNTERS = ["Pixel", "Sample", "SampleBounce", "SampleHit", "SampleMiss", "SampleTerminate", "SampleFogScatter",
            "SampleFogEscape", "SamplePixelConverged", "SamplePixelMaxed", "SceneTrace", "SceneOcclude",
            "MeshIntersect", "BvhIntersectNode", "BvhIntersectShape", "FilterSample", "DenoiseEarlyOut",
            "SampleCountMin", "SampleCountMax", "SampleVarianceMax", "DenoiseRejectMissing", "DenoiseRejectNormal",
            "DenoiseRejectDepth", "DenoiseRejectLum", "DenoiseWeightMax", "DenoiseMaxLum", "DenoiseMaxLumBoost"]
# Initial values of the min- and max-reduced counters (counters.cs:221-226): such a counter still at
# its initial value was never bumped.
COUNTER_UNSET = {"SampleCountMin": 2 ** 63 - 1, "SampleCountMax": -2 ** 63, "SampleVarianceMax": -2 ** 63,
                 "DenoiseWeightMax": -2 ** 63, "DenoiseMaxLum": -2 ** 63, "DenoiseMaxLumBoost": -2 ** 63}

_F = ctypes.POINTER(ctypes.c_float)
_I = ctypes.POINTER(ctypes.c_int32)
_U = ctypes.POINTER(ctypes.c_uint32)
_L = ctypes.POINTER(ctypes.c_int64)
_V = ctypes.c_void_p

_SIGS = {
    "oc_real_size": (ctypes.c_int, []),
    "oc_rng_uints": (None, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _U]),
    "oc_rng_floats": (None, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _F]),
    "oc_splitmix64": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_uint32]),
    "oc_philox": (None, [_U, ctypes.c_uint32, ctypes.c_uint32, _U]),
    "oc_rng_sphere": (None, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _F]),
    "oc_tri_intersect": (None, [_F, _F, ctypes.c_uint32, _F]),
    "oc_box_intersect": (None, [_F, _F, ctypes.c_uint32, _F]),
    "oc_mesh_intersect": (None, [_V, ctypes.c_int, _F, ctypes.c_uint32, _I, _F]),
    "oc_brdf_scalar": (None, [ctypes.c_int, _F, _F, ctypes.c_uint32, _F]),
    "oc_ggx_sample": (None, [ctypes.c_float, _F, ctypes.c_uint32, _F]),
    "oc_hg_scatter": (None, [_F, ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _F]),
    "oc_cdf1": (ctypes.c_float, [_F, ctypes.c_uint32, _F, _F, ctypes.c_uint32, _U]),
    "oc_triangles_make": (None, [_F, _F, _F, ctypes.c_uint32, ctypes.c_int, _F]),
    "oc_scene_new": (_V, []),
    "oc_scene_free": (None, [_V]),
    "oc_scene_add_texture": (ctypes.c_int, [_V, _F, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_float,
                                            ctypes.c_float]),
    "oc_scene_add_mesh": (ctypes.c_int, [_V, _F, ctypes.c_uint32, ctypes.c_int]),
    "oc_scene_add_object": (ctypes.c_int, [_V, _F, _F, _I, ctypes.c_int, ctypes.c_int, _F]),
    "oc_scene_set_sky_texture": (None, [_V, ctypes.c_int, ctypes.c_float, ctypes.c_float]),
    "oc_scene_set_sky_procedural": (None, [_V, _F]),
    "oc_scene_set_fog": (None, [_V, ctypes.c_float, _F, ctypes.c_float, ctypes.c_float]),
    "oc_scene_build": (ctypes.c_int, [_V]),
    "oc_param_count": (ctypes.c_uint32, [_V]),
    "oc_params_get": (None, [_V, _F]),
    "oc_params_set": (None, [_V, _F]),
    "oc_adjoint": (ctypes.c_int, [_V, _F, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                  ctypes.c_float, ctypes.c_uint32, _F, ctypes.c_int, _F]),
    "oc_scene_set_fog_volume": (None, [_V, ctypes.c_int, ctypes.c_float, _F, ctypes.c_float, _F, _F, _U,
                                       ctypes.POINTER(ctypes.c_uint16)]),
    "oc_medium_optical_depth": (None, [_V, _F, _F, ctypes.c_uint32, _F]),
    "oc_medium_scatter": (None, [_V, _F, ctypes.c_float, ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.c_uint32, _F]),
    "oc_bvh_export": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_int, _F, _I]),
    "oc_bvh_stats": (ctypes.c_float, [_V, ctypes.c_int, ctypes.c_int, _I]),
    "oc_trace_primary": (None, [_V, _F, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _I, _I,
                                _F, _F]),
    "oc_render": (ctypes.c_int, [_V, _F, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.c_float, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.c_int, _F, _F, _F, _F, _F, _U, _F, _L, ctypes.c_int, ctypes.c_uint32,
                                 ctypes.c_int, _F]),
    "oc_filter": (None, [_F, _F, _F, _F, _F, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _F, _L, _F,
                          ctypes.c_float]),
    "oc_filter_fog": (None, [ctypes.c_float, ctypes.c_float, _F, ctypes.c_uint32, ctypes.c_uint32, _F]),
    "oc_tonemap": (None, [_F, _F, ctypes.c_uint32, ctypes.c_int, ctypes.c_float, ctypes.c_char_p]),
    "oc_scene_trace": (None, [_V, _F, ctypes.c_uint32, _F]),
    "oc_shadow_transmittance": (None, [_V, _F, ctypes.c_uint32, _F]),
    "oc_refract": (None, [_F, _F, ctypes.c_float, ctypes.c_uint32, _F]),
    "oc_russian_roulette": (None, [_F, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _F]),
    "oc_clamp_luminance": (None, [_F, ctypes.c_float, ctypes.c_uint32, _F]),
    "oc_beer": (None, [_F, ctypes.c_float, _F]),
    "oc_relative_std_err": (ctypes.c_float, [ctypes.c_float, ctypes.c_float, ctypes.c_uint32]),
    "oc_sky_eval": (None, [_V, _F, ctypes.c_uint32, _F, _F]),
    "oc_sky_sample": (None, [_V, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _F]),
    "oc_surface_eval": (None, [_F, _F, _F, _F, _F, ctypes.c_uint32, _F]),
    "oc_surface_eval_grad": (None, [_F, _F, _F, _F, _F, ctypes.c_uint32, _F]),
    "oc_surface_scatter": (None, [_F, _F, _F, _F, _F, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _F]),
    "oc_fog_optical_depth": (None, [_F, _F, _F, ctypes.c_uint32, _F]),
    "oc_fog_scatter": (None, [_F, _F, ctypes.c_float, ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32,
                              ctypes.c_uint32, _F]),
}

_libs = {}


def build():
    """Compiles both oracle libraries (float and FP64) with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load(key, path):
    h = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(h, name)
        fn.restype = res
        fn.argtypes = args
    _libs[key] = h
    return h


def lib(precision="f32"):
    if precision not in _libs:
        path = os.path.join(HERE, f"liboracle_{precision}.so")
        if not os.path.exists(path):
            build()
        _load(precision, path)
    return _libs[precision]


def build_native(out_dir):
    """Builds the float oracle with -O3 -march=native -ffp-contract=off for the host it runs on into
    out_dir and makes it available as precision "f32_native" (used to time the CPU baseline)."""
    path = os.path.join(out_dir, "liboracle_f32_native.so")
    subprocess.run(["make", "-s", "-C", HERE, "native", f"OUT={path}"], check=True)
    return _load("f32_native", path)


def f(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32).ctypes.data_as(_F)


def _keep(a, dtype):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


class OracleScene:
    """A scene inside the oracle, built from a sharpray_amd.scenes.SceneSpec."""

    def __init__(self, spec, precision="f32", rebuild_triangles=None):
        self.L = lib(precision)
        self.spec = spec
        self.h = self.L.oc_scene_new()
        self._refs = []
        modes = getattr(spec, "texture_modes", {})
        for i, t in enumerate(spec.textures):
            rgb = _keep(t, np.float32)
            self._refs.append(rgb)
            flt, tx, ty = modes.get(i, (0, 1.0, 1.0))
            self.L.oc_scene_add_texture(self.h, f(rgb), rgb.shape[1], rgb.shape[0], int(flt), float(tx), float(ty))
        self.mesh_ids = []
        for m, bins in zip(spec.meshes, spec.mesh_bins):
            m = _keep(m, np.float32)
            self._refs.append(m)
            mid = self.L.oc_scene_add_mesh(self.h, f(m), m.shape[0], bins)
            if mid < 0:
                raise RuntimeError("oracle: degenerate BVH split")
            self.mesh_ids.append(mid)
        for o in spec.objects:
            tex = np.array(o.textures, dtype=np.int32)
            params = None if o.params is None else _keep(o.params, np.float32)
            self._refs += [tex, params]
            self.L.oc_scene_add_object(self.h, f(o.transform), f(o.material), tex.ctypes.data_as(_I), o.kind,
                                       o.mesh, f(params))
        if spec.sky_texture >= 0:
            self.L.oc_scene_set_sky_texture(self.h, spec.sky_texture, spec.sky_multiplier, spec.sky_yaw)
        else:
            self.L.oc_scene_set_sky_procedural(self.h, f(spec.sky_procedural))
        if spec.fog is not None:
            d, c, g, fo = spec.fog
            self.L.oc_scene_set_fog(self.h, d, f(np.asarray(c, np.float32)), g, fo)
        vol = getattr(spec, "fog_volume", None)
        if vol is not None:
            grid = None
            if vol.density is not None:
                grid = np.ascontiguousarray(vol.density, np.float16).view(np.uint16)
                self._refs.append(grid)
            res = np.array(vol.res, np.uint32)
            self._refs.append(res)
            self.L.oc_scene_set_fog_volume(self.h, vol.kind, vol.sigma, f(np.asarray(vol.color, np.float32)),
                                           vol.anisotropy, f(np.asarray(vol.bounds_min, np.float32)),
                                           f(np.asarray(vol.bounds_max, np.float32)), res.ctypes.data_as(_U),
                                           None if grid is None else grid.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)))
        if self.L.oc_scene_build(self.h) != 0:
            raise RuntimeError("oracle: scene build failed (sky required)")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.oc_scene_free(self.h)
            self.h = None

    def bvh(self, mesh=None):
        level, mid = (0, 0) if mesh is None else (1, mesh)
        n = self.L.oc_bvh_export(self.h, level, mid, None, None)
        nitems = len(self.spec.objects) if mesh is None else self.spec.meshes[mesh].shape[0]
        nodes = np.empty((n, 8), np.float32)
        items = np.empty(nitems, np.int32)
        self.L.oc_bvh_export(self.h, level, mid, nodes.ctypes.data_as(_F), items.ctypes.data_as(_I))
        return nodes, items

    def mesh_intersect(self, mesh, rays):
        rays = _keep(rays, np.float32).reshape(-1, 6)
        n = rays.shape[0]
        idx = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        self.L.oc_mesh_intersect(self.h, mesh, f(rays), n, idx.ctypes.data_as(_I), t.ctypes.data_as(_F))
        return idx, t

    def trace(self, rays):
        """Scene.Trace of rays (n, 6): dict of object, prim, dist, normal, normal_geo, tangent, uv,
        albedo, roughness, metallic (object -1 on a miss)."""
        rays = _keep(rays, np.float32).reshape(-1, 6)
        out = np.empty((rays.shape[0], 20), np.float32)
        self.L.oc_scene_trace(self.h, f(rays), rays.shape[0], out.ctypes.data_as(_F))
        return dict(object=out[:, 0].astype(np.int32), prim=out[:, 1].astype(np.int32), dist=out[:, 2],
                    normal=out[:, 3:6], normal_geo=out[:, 6:9], tangent=out[:, 9:13], uv=out[:, 13:15],
                    albedo=out[:, 15:18], roughness=out[:, 18], metallic=out[:, 19])

    def shadow_transmittance(self, rays):
        rays = _keep(rays, np.float32).reshape(-1, 6)
        out = np.empty((rays.shape[0], 3), np.float32)
        self.L.oc_shadow_transmittance(self.h, f(rays), rays.shape[0], out.ctypes.data_as(_F))
        return out

    def bvh_stats(self, mesh=None):
        level, mid = (0, 0) if mesh is None else (1, mesh)
        ints = np.zeros(7, np.int32)
        sah = self.L.oc_bvh_stats(self.h, level, mid, ints.ctypes.data_as(_I))
        keys = ["nodeCount", "depthMax", "leafCount", "leafSizeMin", "leafSizeMax", "shapeCount", "depthWeighted"]
        d = dict(zip(keys, ints.tolist()))
        d["sah"] = sah
        return d

    def trace_primary(self, y0=0, y1=None):
        s = self.spec
        y1 = s.height if y1 is None else y1
        n = (y1 - y0) * s.width
        obj = np.empty(n, np.int32)
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        nrm = np.empty((n, 3), np.float32)
        self.L.oc_trace_primary(self.h, f(s.view), s.width, s.height, y0, y1, obj.ctypes.data_as(_I),
                                prim.ctypes.data_as(_I), t.ctypes.data_as(_F), nrm.ctypes.data_as(_F))
        shp = (y1 - y0, s.width)
        return obj.reshape(shp), prim.reshape(shp), t.reshape(shp), nrm.reshape(shp + (3,))

    def params(self):
        p = np.empty(self.L.oc_param_count(self.h), np.float32)
        self.L.oc_params_get(self.h, p.ctypes.data_as(_F))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.L.oc_param_count(self.h)
        self.L.oc_params_set(self.h, p.ctypes.data_as(_F))

    def adjoint(self, d_image, seed, threads=None):
        """Gradient of sum(d_image * Radiance) over the parameters (fixed spp, counter RNG)."""
        s = self.spec
        d = np.ascontiguousarray(d_image, np.float32).reshape(-1)
        g = np.empty(self.L.oc_param_count(self.h), np.float32)
        rc = self.L.oc_adjoint(self.h, f(s.view), s.width, s.height, s.max_samples, s.bounces, s.indirect_clamp,
                               int(seed), d.ctypes.data_as(_F), threads or os.cpu_count() or 1, g.ctypes.data_as(_F))
        if rc != 0:
            raise RuntimeError(f"oracle: adjoint failed ({rc})")
        return g

    def render(self, y0=0, y1=None, threads=None, counters=False, rng_policy=0, rng_seed=0, light_sampler=0):
        s = self.spec
        y1 = s.height if y1 is None else y1
        threads = threads or os.cpu_count() or 1
        n = (y1 - y0) * s.width
        out = dict(radiance=np.empty((n, 3), np.float32), radiance_fog=np.empty((n, 3), np.float32),
                   normals=np.empty((n, 3), np.float32), uv=np.empty((n, 2), np.float32),
                   depth=np.empty(n, np.float32), samples=np.empty(n, np.uint32), variance=np.empty(n, np.float32),
                   albedo=np.empty((n, 3), np.float32))
        ctr = np.zeros(len(COUNTERS), np.int64)
        rc = self.L.oc_render(self.h, f(s.view), s.width, s.height, s.min_samples, s.max_samples,
                              s.variance_threshold, s.bounces, s.indirect_clamp, y0, y1, threads,
                              out["radiance"].ctypes.data_as(_F), out["radiance_fog"].ctypes.data_as(_F),
                              out["normals"].ctypes.data_as(_F), out["uv"].ctypes.data_as(_F),
                              out["depth"].ctypes.data_as(_F), out["samples"].ctypes.data_as(_U),
                              out["variance"].ctypes.data_as(_F), ctr.ctypes.data_as(_L) if counters else None,
                              int(rng_policy), int(rng_seed), int(light_sampler),
                              out["albedo"].ctypes.data_as(_F))
        if rc != 0:
            raise RuntimeError("oracle: render on an unbuilt scene")
        if counters:
            out["counters"] = dict(zip(COUNTERS, ctr.tolist()))
        return out


def triangles_make(pos, normals=None, uv=None, smooth=False, precision="f32"):
    pos = _keep(pos, np.float32).reshape(-1, 9)
    n = pos.shape[0]
    out = np.empty((n, 36), np.float32)
    nrm = _keep(normals, np.float32)
    uvs = _keep(uv, np.float32)
    lib(precision).oc_triangles_make(f(pos), f(nrm), f(uvs), n, 1 if smooth else 0, out.ctypes.data_as(_F))
    return out


def filter_fog(radius, strength, fog, width, height):
    """Compositor.FilterFog (compose.cs:206-245) of a w*h*3 fog radiance image."""
    out = np.empty((width * height, 3), np.float32)
    lib().oc_filter_fog(float(radius), float(strength), f(fog), width, height, out.ctypes.data_as(_F))
    return out


def tonemap(main, fog, tonemapper, exposure):
    """Tonemap (compose.cs:247-266) + ToPixel (math.cs:127-136) of main + fog: uint8 (n, 3)."""
    main = np.ascontiguousarray(main, np.float32).reshape(-1, 3)
    fog = np.ascontiguousarray(fog, np.float32).reshape(-1, 3)
    out = ctypes.create_string_buffer(main.shape[0] * 3)
    lib().oc_tonemap(f(main), f(fog), main.shape[0], int(tonemapper), float(exposure), out)
    return np.frombuffer(out.raw, np.uint8).reshape(-1, 3).copy()


def encode_bmp(rgb, width, height):
    """Image.WriteBmp (image.cs:147-191): 24-bit BGR, top-down (negative height), rows padded to 4."""
    import struct
    px = np.asarray(rgb, np.uint8).reshape(height, width, 3)
    stride = (width * 3 + 3) & ~3
    data = bytearray()
    data += b"BM" + struct.pack("<III", 14 + 40 + stride * height, 0, 54)
    data += struct.pack("<IiiHHII", 40, width, -height, 1, 24, 0, stride * height) + bytes(16)
    for y in range(height):
        data += px[y, :, ::-1].tobytes() + bytes(stride - width * 3)
    return bytes(data)


def bilateral(params7, radiance, normals, depth, variance, width, height, threads=None, albedo=None,
              albedo_limit=0.0):
    """Compositor.Filter of a whole image: (filtered (n, 3), the filter's counters by name)."""
    out = np.empty((width * height, 3), np.float32)
    ctr = np.zeros(len(COUNTERS), np.int64)
    lib().oc_filter(f(params7), f(radiance), f(normals), f(depth), f(variance), width, height,
                    threads or os.cpu_count() or 1, out.ctypes.data_as(_F), ctr.ctypes.data_as(_L), f(albedo),
                    float(albedo_limit))
    names = ("FilterSample", "DenoiseEarlyOut", "DenoiseRejectMissing", "DenoiseRejectNormal", "DenoiseRejectDepth",
             "DenoiseRejectLum", "DenoiseWeightMax", "DenoiseMaxLum", "DenoiseMaxLumBoost")
    full = dict(zip(COUNTERS, ctr.tolist()))
    return out, {k: full[k] for k in names}


def refract(inc, normal, eta):
    """Brdf.Refract of (n, 3) directions: (dirs (n, 3), refracted (n,) bool)."""
    inc = _keep(inc, np.float32).reshape(-1, 3)
    nrm = np.broadcast_to(_keep(normal, np.float32).reshape(-1, 3), inc.shape).copy()
    out = np.empty((inc.shape[0], 4), np.float32)
    lib().oc_refract(f(inc), f(nrm), float(eta), inc.shape[0], out.ctypes.data_as(_F))
    return out[:, :3], out[:, 3] > 0


def russian_roulette(energy, bounce, n, a=1, b=2):
    """Scene.RussianRoulette over n draws of Rng(a, b): (energies (n, 3), terminated (n,) bool)."""
    out = np.empty((n, 4), np.float32)
    lib().oc_russian_roulette(f(np.asarray(energy, np.float32)), bounce, a, b, n, out.ctypes.data_as(_F))
    return out[:, :3], out[:, 3] > 0


def clamp_luminance(colors, mx):
    c = _keep(colors, np.float32).reshape(-1, 3)
    out = np.empty_like(c)
    lib().oc_clamp_luminance(f(c), float(mx), c.shape[0], out.ctypes.data_as(_F))
    return out


def beer(albedo, d):
    out = np.empty(3, np.float32)
    lib().oc_beer(f(np.asarray(albedo, np.float32)), float(d), out.ctypes.data_as(_F))
    return out


def relative_std_err(total, total_sqr, n, precision="f32"):
    return float(lib(precision).oc_relative_std_err(float(total), float(total_sqr), int(n)))
