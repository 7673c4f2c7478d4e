"""Host-side mirror of SharpRay's Scene / Renderer / Compositor API over the gfx950 C ABI.

Names, argument meaning and error behaviour follow the reference's C# classes
(BastianBlokland/sharpray src/scene.cs, renderer.cs, compose.cs, math.cs, mesh.cs, texture.cs);
every call goes through libsharpray_hip.so (include/sharpray.h). Arrays are numpy float32 in the
reference's blittable layouts (Color/Vec3 = 3 floats, Vec2 = 2, Triangle = 36).
"""
import ctypes
import math
import os

import numpy as np

from . import _native as N
from ._native import SharpRayError, check, fptr, iptr, uptr

__all__ = [
    "SharpRayError", "Quat", "Transform", "View", "Texture", "Mesh", "Sphere", "AABox", "Material", "Object",
    "SkyTexture", "SkyProcedural", "Fog", "Scene", "Renderer", "Compositor", "triangles_build", "RNG_REFERENCE",
    "ObjConfig", "ObjLoader", "load_hdr", "decode_hdr", "decode_tga", "Image",
    "RNG_COUNTER",
]

RNG_REFERENCE = N.SR_RNG_REFERENCE
RNG_COUNTER = N.SR_RNG_COUNTER
LIGHT_CDF = N.SR_LIGHT_CDF
LIGHT_ALIAS = N.SR_LIGHT_ALIAS


def _f32(a, shape_tail=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape_tail is not None and a.shape[-len(shape_tail):] != tuple(shape_tail):
        raise ValueError(f"expected trailing shape {shape_tail}, got {a.shape}")
    return a


class Quat:
    """Quat (math.cs:564-659): x, y, z, w."""

    def __init__(self, x=0.0, y=0.0, z=0.0, w=1.0):
        self.xyzw = np.array([x, y, z, w], dtype=np.float32)

    @staticmethod
    def identity():
        return Quat()

    @staticmethod
    def angle_axis(angle, axis):
        """Quat.AngleAxis (math.cs:621-626), evaluated in float32."""
        ax = np.asarray(axis, dtype=np.float32)
        h = np.float32(angle) * np.float32(0.5)
        s = np.float32(math.sin(float(h)))
        return Quat(*(ax * s).tolist(), float(np.float32(math.cos(float(h)))))

    def __mul__(self, o):  # Quat * Quat (math.cs:603-607)
        a, b = self.xyzw.astype(np.float64), o.xyzw.astype(np.float64)
        return Quat(
            a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1],
            a[3] * b[1] + a[1] * b[3] + a[2] * b[0] - a[0] * b[2],
            a[3] * b[2] + a[2] * b[3] + a[0] * b[1] - a[1] * b[0],
            a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2],
        )


class Transform:
    """Transform (math.cs:773-840): position, rotation, scale."""

    def __init__(self, pos=(0, 0, 0), rot=None, scale=(1, 1, 1)):
        self.pos = np.asarray(pos, dtype=np.float32)
        self.rot = rot if rot is not None else Quat.identity()
        self.scale = np.asarray(scale, dtype=np.float32)
        if np.any(self.scale == 0):
            raise ValueError("Scale cannot be zero")

    @staticmethod
    def identity():
        return Transform()

    def _c(self):
        t = N.Transform_()
        t.pos[:] = self.pos.tolist()
        t.rot[:] = self.rot.xyzw.tolist()
        t.scale[:] = self.scale.tolist()
        return t

    def as_array(self):
        return np.concatenate([self.pos, self.rot.xyzw, self.scale]).astype(np.float32)


class View:
    """View (math.cs:1582-1663): unscaled transform, field of view (radians), near plane."""

    def __init__(self, trans=None, fov=math.radians(60.0), near=0.1):
        self.trans = trans or Transform()
        self.fov = float(np.float32(fov))
        self.near = float(near)
        if not (0.0 < self.fov < math.pi):
            raise ValueError("Invalid fov")

    def _c(self):
        v = N.View_()
        v.pos[:] = self.trans.pos.tolist()
        v.rot[:] = self.trans.rot.xyzw.tolist()
        v.fov = self.fov
        v.near_plane = self.near
        return v

    def as_array(self):
        return np.concatenate([self.trans.pos, self.trans.rot.xyzw, [self.fov, self.near]]).astype(np.float32)


class Texture:
    """Texture (texture.cs:6-179): linear RGB texels (h, w, 3), bilinear or point filter, tiling."""

    BILINEAR, POINT = 0, 1

    def __init__(self, rgb, filter=0, tiling=(1.0, 1.0)):
        self.rgb = _f32(rgb)
        if self.rgb.ndim != 3 or self.rgb.shape[2] != 3:
            raise ValueError("texture must be (h, w, 3)")
        self.filter = int(filter)
        self.tiling = (float(tiling[0]), float(tiling[1]))

    @property
    def width(self):
        return self.rgb.shape[1]

    @property
    def height(self):
        return self.rgb.shape[0]

    @staticmethod
    def from_hdr(rgbe, max_luminance=None):
        """Texture.FromHdr (texture.cs:154-163) of RGBE pixels (h, w, 4) uint8 (math.cs:196-202)."""
        rgbe = np.asarray(rgbe, dtype=np.uint8)
        e = rgbe[..., 3].astype(np.int32)
        scale = np.ldexp(np.float32(1.0), e - 128).astype(np.float32) / np.float32(256.0)
        rgb = rgbe[..., :3].astype(np.float32) * scale[..., None]
        rgb[e == 0] = 0.0
        if max_luminance is not None:
            lum = (np.float32(0.2126) * rgb[..., 0] + np.float32(0.7152) * rgb[..., 1]) + np.float32(0.0722) * rgb[..., 2]
            over = lum > np.float32(max_luminance)
            rgb[over] = rgb[over] * (np.float32(max_luminance) / lum[over])[:, None]
        return Texture(rgb)


def triangles_build(pos, normals=None, uv=None, smooth=False):
    """Triangle records (n, 36) from positions (n, 9), optional normals (n, 9) and uvs (n, 6): the
    Triangle ctors (math.cs:1503-1531) plus Mesh.ComputeSmoothNormals (mesh.cs:99-123)."""
    pos = _f32(pos).reshape(-1, 9)
    n = pos.shape[0]
    nrm = _f32(normals).reshape(-1, 9) if normals is not None else None
    uvs = _f32(uv).reshape(-1, 6) if uv is not None else None
    out = np.empty((n, 36), dtype=np.float32)
    check(N.lib().sr_triangles_build(fptr(pos), fptr(nrm), fptr(uvs), n, 1 if smooth else 0, fptr(out)),
          "triangles_build")
    return out


class Mesh:
    """Mesh (mesh.cs:13-172) from Triangle records (n, 36); the BVH is built by the library."""

    def __init__(self, triangles, split_bins=8):
        self.triangles = _f32(triangles).reshape(-1, 36)
        self.split_bins = int(split_bins)


class Sphere:
    """Sphere (math.cs:1171-1234)."""

    def __init__(self, center, radius):
        if radius <= 0:
            raise ValueError("Radius must be positive")
        self.params = np.array(list(center) + [radius, 0, 0], dtype=np.float32)


class AABox:
    """AABox as a shape (math.cs:936-1059)."""

    def __init__(self, mn, mx):
        self.params = np.array(list(mn) + list(mx), dtype=np.float32)


class Material:
    """Material (scene.cs:161-202)."""

    def __init__(self, albedo, roughness, metallic=0.0, transparency=0.0, ior=1.5, radiance=(0.0, 0.0, 0.0),
                 color_texture=None, roughness_texture=None, metallic_texture=None, normal_texture=None):
        self.albedo = np.asarray(albedo if np.ndim(albedo) else (albedo,) * 3, dtype=np.float32)
        self.roughness = float(roughness)
        self.metallic = float(metallic)
        self.transparency = float(transparency)
        self.ior = float(ior)
        self.radiance = np.asarray(radiance, dtype=np.float32)
        self.textures = [color_texture, roughness_texture, metallic_texture, normal_texture]

    def as_array(self):
        return np.array(list(self.albedo) + [self.roughness, self.metallic, self.transparency, self.ior] +
                        list(self.radiance), dtype=np.float32)


class Object:
    """Object (scene.cs:204-270): name, transform, material, shape (Mesh, Sphere or AABox)."""

    def __init__(self, name, trans, material, shape):
        self.name = name
        self.trans = trans
        self.material = material
        self.shape = shape


class SkyTexture:
    """SkyTexture (scene.cs:379-435)."""

    def __init__(self, texture, multiplier=1.0, rot_yaw=0.0):
        self.texture = texture
        self.multiplier = float(multiplier)
        self.rot_yaw = float(rot_yaw)


class SkyProcedural:
    """SkyProcedural (scene.cs:334-377) with its SunProcedural (scene.cs:281-332)."""

    def __init__(self, radiance_top, radiance_middle, radiance_bottom, sun_dir, sun_radiance, sun_angle):
        self.values = [radiance_top, radiance_middle, radiance_bottom, sun_dir, sun_radiance]
        self.sun_angle = float(sun_angle)

    def as_array(self):
        return np.array([c for v in self.values for c in v] + [self.sun_angle], dtype=np.float32)


class Fog:
    """Fog (scene.cs:438-505): exponential height fog."""

    def __init__(self, density, color, anisotropy, height_falloff):
        self.density = float(density)
        self.color = np.asarray(color, dtype=np.float32)
        self.anisotropy = float(anisotropy)
        self.height_falloff = float(height_falloff)


MEDIUM_HOMOGENEOUS = 1
MEDIUM_GRID = 2


class FogVolume:
    """A bounded participating medium in place of Fog (no reference counterpart): a homogeneous box
    (kind MEDIUM_HOMOGENEOUS, extinction sigma) or a half-precision density grid spanning the box
    (kind MEDIUM_GRID, extinction sigma * voxel value; density shaped (nz, ny, nx)). color and
    anisotropy mean what they mean for Fog."""

    def __init__(self, kind, sigma, color, anisotropy, bounds_min, bounds_max, density=None):
        self.kind = int(kind)
        self.sigma = float(sigma)
        self.color = np.asarray(color, dtype=np.float32)
        self.anisotropy = float(anisotropy)
        self.bounds_min = np.asarray(bounds_min, dtype=np.float32)
        self.bounds_max = np.asarray(bounds_max, dtype=np.float32)
        self.density = None if density is None else np.ascontiguousarray(density, dtype=np.float16)
        self.res = (0, 0, 0) if density is None else tuple(int(x) for x in self.density.shape[::-1])


class Scene:
    """Scene (scene.cs:507-881) backed by one device context."""

    def __init__(self, device=0):
        self._ctx = ctypes.c_void_p()
        check(N.lib().sr_create(device, ctypes.byref(self._ctx)), "sr_create")
        self._built = False
        self._sky = None
        self.objects = []
        self._tex_ids = {}
        self._mesh_ids = {}

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx:
            N.lib().sr_destroy(ctx)
            self._ctx = None

    @property
    def ctx(self):
        return self._ctx

    def _guard(self):
        if self._built:
            raise SharpRayError("Scene already built")

    def texture_id(self, tex):
        if tex is None:
            return -1
        key = id(tex)
        if key not in self._tex_ids:
            out = ctypes.c_int32()
            check(N.lib().sr_texture_create(self._ctx, fptr(tex.rgb), tex.width, tex.height, tex.filter,
                                            tex.tiling[0], tex.tiling[1], ctypes.byref(out)), "texture")
            self._tex_ids[key] = (out.value, tex)
        return self._tex_ids[key][0]

    def mesh_id(self, mesh):
        key = id(mesh)
        if key not in self._mesh_ids:
            out = ctypes.c_int32()
            tris = mesh.triangles
            check(N.lib().sr_mesh_create(self._ctx, fptr(tris), tris.shape[0], mesh.split_bins, ctypes.byref(out)),
                  "mesh")
            self._mesh_ids[key] = (out.value, mesh)
        return self._mesh_ids[key][0]

    def set_sky(self, sky):
        self._guard()
        if isinstance(sky, SkyTexture):
            check(N.lib().sr_scene_set_sky_texture(self._ctx, self.texture_id(sky.texture), sky.multiplier,
                                                   sky.rot_yaw), "set_sky")
        elif isinstance(sky, SkyProcedural):
            p = N.SkyProcedural_()
            arr = sky.as_array()
            for i, name in enumerate(["radiance_top", "radiance_middle", "radiance_bottom", "sun_dir", "sun_radiance"]):
                getattr(p, name)[:] = arr[i * 3:i * 3 + 3].tolist()
            p.sun_angle = sky.sun_angle
            check(N.lib().sr_scene_set_sky_procedural(self._ctx, ctypes.byref(p)), "set_sky")
        else:
            raise TypeError("sky must be SkyTexture or SkyProcedural")
        self._sky = sky

    def set_fog(self, fog):
        self._guard()
        if isinstance(fog, FogVolume):
            v = N.FogVolume_()
            v.kind, v.sigma, v.anisotropy = fog.kind, fog.sigma, fog.anisotropy
            v.color[:] = fog.color.tolist()
            v.bounds_min[:] = fog.bounds_min.tolist()
            v.bounds_max[:] = fog.bounds_max.tolist()
            v.res[:] = list(fog.res)
            if fog.density is not None:
                v.density = fog.density.view(np.uint16).ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))
            check(N.lib().sr_scene_set_fog_volume(self._ctx, ctypes.byref(v)), "set_fog_volume")
            self.fog = fog
            return
        f = N.Fog_()
        f.density = fog.density
        f.color[:] = fog.color.tolist()
        f.anisotropy = fog.anisotropy
        f.height_falloff = fog.height_falloff
        check(N.lib().sr_scene_set_fog(self._ctx, ctypes.byref(f)), "set_fog")
        self.fog = fog

    def set_tail_segment(self, segment):
        """Renders finish every path from this segment on in one fused launch (results unchanged)."""
        check(N.lib().sr_set_tail_segment(self._ctx, int(segment)), "tail segment")

    def set_slices(self, slices):
        """Renders run each band as this many concurrent row slices on their own streams (results
        unchanged)."""
        check(N.lib().sr_set_slices(self._ctx, int(slices)), "slices")

    def set_samples_in_flight(self, samples):
        """Renders with independent samples (counter RNG, fixed sample count, counters off) keep
        this many samples of every pixel in flight at once (results unchanged)."""
        check(N.lib().sr_set_samples_in_flight(self._ctx, int(samples)), "samples in flight")

    def set_bvh_builder(self, device):
        """Mesh BVHs of later meshes built on the GPU (True) or the host (False); both give the
        reference's tree bit for bit."""
        check(N.lib().sr_set_bvh_builder(self._ctx, N.SR_BVH_BUILDER_DEVICE if device else N.SR_BVH_BUILDER_HOST),
              "bvh builder")

    def _material_struct(self, m):
        cm = N.Material_()
        cm.albedo[:] = m.albedo.tolist()
        cm.roughness, cm.metallic, cm.transparency, cm.ior = m.roughness, m.metallic, m.transparency, m.ior
        cm.radiance[:] = m.radiance.tolist()
        ids = [self.texture_id(t) for t in m.textures]
        cm.color_texture, cm.roughness_texture, cm.metallic_texture, cm.normal_texture = ids
        return cm

    # ---- what the scene holds (also objects added natively, e.g. by ObjLoader) ----------------
    def counts(self):
        """(textures, meshes, objects)."""
        t, m, o = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        check(N.lib().sr_scene_counts(self._ctx, ctypes.byref(t), ctypes.byref(m), ctypes.byref(o)), "counts")
        return t.value, m.value, o.value

    def mesh_triangles(self, mesh):
        """Triangle records (n, 36) of a mesh: lean (A, B-A, C-A), normals, tangents, uvs."""
        n = ctypes.c_uint32()
        check(N.lib().sr_mesh_read(self._ctx, mesh, None, ctypes.byref(n)), "mesh read")
        out = np.empty((n.value, 36), np.float32)
        check(N.lib().sr_mesh_read(self._ctx, mesh, fptr(out), ctypes.byref(n)), "mesh read")
        return out

    def texture_texels(self, tex):
        w, h = ctypes.c_uint32(), ctypes.c_uint32()
        check(N.lib().sr_texture_read(self._ctx, tex, None, ctypes.byref(w), ctypes.byref(h)), "texture read")
        out = np.empty((h.value, w.value, 3), np.float32)
        check(N.lib().sr_texture_read(self._ctx, tex, fptr(out), ctypes.byref(w), ctypes.byref(h)), "texture read")
        return out

    def object_record(self, index):
        """dict(transform, material, shape_kind, mesh) of an object as stored."""
        t, m = N.Transform_(), N.Material_()
        kind, mesh = ctypes.c_int32(), ctypes.c_int32()
        check(N.lib().sr_scene_object_read(self._ctx, index, ctypes.byref(t), ctypes.byref(m), ctypes.byref(kind),
                                           ctypes.byref(mesh)), "object read")
        mat = {k: (list(getattr(m, k)) if k in ("albedo", "radiance") else getattr(m, k)) for k, _ in m._fields_}
        return {"pos": list(t.pos), "rot": list(t.rot), "scale": list(t.scale), "material": mat,
                "shape_kind": kind.value, "mesh": mesh.value}

    def add_object(self, obj):
        self._guard()
        cm = self._material_struct(obj.material)
        params = None
        if isinstance(obj.shape, Mesh):
            kind, mesh = N.SR_SHAPE_MESH, self.mesh_id(obj.shape)
        elif isinstance(obj.shape, Sphere):
            kind, mesh, params = N.SR_SHAPE_SPHERE, -1, obj.shape.params
        elif isinstance(obj.shape, AABox):
            kind, mesh, params = N.SR_SHAPE_BOX, -1, obj.shape.params
        else:
            raise TypeError("unsupported shape")
        tr = obj.trans._c()
        out = ctypes.c_int32()
        check(N.lib().sr_scene_add_object(self._ctx, ctypes.byref(tr), ctypes.byref(cm), kind, mesh, fptr(params),
                                          ctypes.byref(out)), "add_object")
        self.objects.append(obj)
        return out.value

    def build(self):
        """Scene.Build (scene.cs:517-541); idempotent."""
        check(N.lib().sr_scene_build(self._ctx), "build")
        self._built = True

    def bvh(self, mesh=None):
        """(nodes (n, 8) float32 with child/count as int32 bit patterns, items) of a mesh or the scene."""
        if mesh is None:
            mid = -1
        elif isinstance(mesh, int):
            mid = mesh
        else:
            mid = self.mesh_id(mesh)
        nc, ic = ctypes.c_uint32(), ctypes.c_uint32()
        check(N.lib().sr_bvh_export(self._ctx, mid, None, None, ctypes.byref(nc), ctypes.byref(ic)), "bvh")
        nodes = np.empty((nc.value, 8), dtype=np.float32)
        items = np.empty(ic.value, dtype=np.int32)
        check(N.lib().sr_bvh_export(self._ctx, mid, fptr(nodes), iptr(items), None, None), "bvh")
        return nodes, items

    def bvh_stats(self, mesh=None):
        """Bvh.GetStats (bvh.cs:79-125) of a mesh's BVH or (mesh None) of the built scene's: dict with
        the BvhStats record's fields (bvh.cs:5-14)."""
        mid = -1 if mesh is None else (mesh if isinstance(mesh, int) else self.mesh_id(mesh))
        st = N.BvhStats_()
        check(N.lib().sr_bvh_stats_get(self._ctx, mid, ctypes.byref(st)), "bvh stats")
        return {"NodeCount": st.node_count, "DepthMax": st.depth_max, "LeafCount": st.leaf_count,
                "LeafSizeMin": st.leaf_size_min, "LeafSizeMax": st.leaf_size_max, "LeafSizeAvg": st.leaf_size_avg,
                "LeafDepthAvg": st.leaf_depth_avg, "SahCost": st.sah_cost}

    def counters_enable(self, enable=True):
        """Collect counters from now on (mesh and scene builds, renders, filters)."""
        check(N.lib().sr_counters_enable(self._ctx, 1 if enable else 0), "counters")

    def counters_all(self, include_unset=False):
        """The context's Counters.Type book (counters.cs:24-79) by name; entries still at their
        reference initial value are left out unless include_unset."""
        out = (ctypes.c_int64 * N.SR_COUNTER_TYPE_COUNT)()
        check(N.lib().sr_counters_read_all(self._ctx, out, N.SR_COUNTER_TYPE_COUNT), "counters")
        d = dict(zip(N.COUNTER_TYPE_NAMES, list(out)))
        if include_unset:
            return d
        return {k: v for k, v in d.items() if v != N.COUNTER_TYPE_UNSET.get(k, 0)}

    def counters_reset(self):
        check(N.lib().sr_counters_reset(self._ctx), "counters")

    def trace_primary(self, view, width, height):
        """Sample-0 primary hits: (object, triangle, dist) arrays of shape (height, width)."""
        obj = np.empty(width * height, dtype=np.int32)
        tri = np.empty(width * height, dtype=np.int32)
        dist = np.empty(width * height, dtype=np.float32)
        v = view._c()
        check(N.lib().sr_trace_primary(self._ctx, ctypes.byref(v), width, height, iptr(obj), iptr(tri), fptr(dist)),
              "trace_primary")
        return obj.reshape(height, width), tri.reshape(height, width), dist.reshape(height, width)


class Renderer:
    """Renderer (renderer.cs:5-238). Construction starts the render asynchronously on the GPU."""

    def __init__(self, scene, view, width, height, block_size, min_samples, max_samples, variance_threshold, bounces,
                 indirect_clamp, rng_policy=RNG_REFERENCE, rng_seed=0, tick_bands=1, counters=False,
                 kernel_timing=False, light_sampler=LIGHT_CDF):
        if width <= 0 or height <= 0 or block_size <= 0:
            raise ValueError("size must be positive")
        if not (0 < min_samples <= max_samples):
            raise ValueError("need 0 < minSamples <= maxSamples")
        self.scene = scene
        self.view = view
        self.width, self.height = int(width), int(height)
        p = N.RenderParams_()
        p.width, p.height, p.block_size = self.width, self.height, int(block_size)
        p.min_samples, p.max_samples = int(min_samples), int(max_samples)
        p.variance_threshold = float(variance_threshold)
        p.bounces, p.indirect_clamp = int(bounces), float(indirect_clamp)
        p.rng_policy, p.rng_seed, p.tick_bands = int(rng_policy), int(rng_seed), int(tick_bands)
        p.light_sampler = int(light_sampler)
        self.params = p
        check(N.lib().sr_counters_enable(scene.ctx, 1 if counters else 0), "counters")
        check(N.lib().sr_profile_enable(scene.ctx, 1 if kernel_timing else 0), "kernel timing")
        v = view._c()
        check(N.lib().sr_render_begin(scene.ctx, ctypes.byref(v), ctypes.byref(p)), "render_begin")
        scene._built = True
        self._aovs = None

    def tick(self):
        """Renderer.Tick (renderer.cs:98-105): (completed blocks, total blocks)."""
        s, t = ctypes.c_uint32(), ctypes.c_uint32()
        check(N.lib().sr_render_tick(self.scene.ctx, ctypes.byref(s), ctypes.byref(t)), "tick")
        return s.value, t.value

    def wait(self):
        while True:
            s, t = self.tick()
            if s == t:
                return

    def _read(self):
        if self._aovs is None:
            n = self.width * self.height
            a = dict(
                radiance=np.empty((n, 3), np.float32), radiance_fog=np.empty((n, 3), np.float32),
                normals=np.empty((n, 3), np.float32), uv=np.empty((n, 2), np.float32),
                depth=np.empty(n, np.float32), samples=np.empty(n, np.uint32), variance=np.empty(n, np.float32))
            check(N.lib().sr_render_read(self.scene.ctx, fptr(a["radiance"]), fptr(a["radiance_fog"]),
                                         fptr(a["normals"]), fptr(a["uv"]), fptr(a["depth"]), uptr(a["samples"]),
                                         fptr(a["variance"])), "render_read")
            a["albedo"] = np.empty((n, 3), np.float32)
            check(N.lib().sr_render_read_albedo(self.scene.ctx, fptr(a["albedo"])), "render_read_albedo")
            self._aovs = a
        return self._aovs

    radiance = property(lambda self: self._read()["radiance"])
    radiance_fog = property(lambda self: self._read()["radiance_fog"])
    normals = property(lambda self: self._read()["normals"])
    uv = property(lambda self: self._read()["uv"])
    depth = property(lambda self: self._read()["depth"])
    samples = property(lambda self: self._read()["samples"])

    @property
    def sample_mean(self):
        """Radiance * Samples / min(Samples, max_samples): the per-pixel mean of the samples taken
        (the reference divides pixels that reach the cap by M + 1, renderer.cs:205-217)."""
        c = self.samples.astype(np.float32)
        m = np.float32(self.params.max_samples)
        f = np.where(c > m, c / m, np.float32(1.0)).astype(np.float32)
        return (self.radiance * f[:, None]).astype(np.float32)
    variance = property(lambda self: self._read()["variance"])
    albedo = property(lambda self: self._read()["albedo"])  # primary-hit albedo guide (no reference counterpart)

    def gpu_seconds(self):
        out = ctypes.c_double()
        check(N.lib().sr_render_gpu_seconds(self.scene.ctx, ctypes.byref(out)), "gpu_seconds")
        return out.value

    def kernel_times(self):
        """{kind: (total ms, launches)} of this render: "trace" (closest hit of a segment's rays plus
        the previous segment's shadow rays), "shade", "trace_tail" (the fused launch that finishes
        the paths from the tail segment on, or the trace kernel on a sample's last shadow rays when
        there is none) and "other" (camera rays, accumulation)."""
        ms = (ctypes.c_double * 4)()
        n = (ctypes.c_uint32 * 4)()
        check(N.lib().sr_profile_read(self.scene.ctx, ms, n), "kernel times")
        return {k: (ms[i], n[i]) for i, k in enumerate(["trace", "shade", "trace_tail", "other"])}

    def kernel_busy(self):
        """{kind: ms} of this render during which at least one launch of that kind ran on any of
        the concurrent slices (kinds as in kernel_times)."""
        ms = (ctypes.c_double * 4)()
        check(N.lib().sr_profile_busy(self.scene.ctx, ms), "kernel busy")
        return {k: ms[i] for i, k in enumerate(["trace", "shade", "trace_tail", "other"])}

    def counters(self):
        out = (ctypes.c_int64 * N.SR_COUNTER_COUNT)()
        check(N.lib().sr_counters_read(self.scene.ctx, out), "counters")
        return dict(zip(N.COUNTER_NAMES, list(out)))

    def counters_fused(self):
        """The fused tail launch's share of the traversal counters."""
        out = (ctypes.c_int64 * 4)()
        check(N.lib().sr_counters_read_fused(self.scene.ctx, out), "counters")
        return dict(zip(["BvhIntersectNode", "BvhIntersectShape", "SceneTrace", "SceneOcclude"], list(out)))

    def counters_reference_only(self):
        """The per-segment traversal launches' share of the counters that only the reference's way of
        tracing shadow rays incurs (rays whose light lies behind the surface, scene.cs:774-783, and
        closest-hit instead of any-hit walks, scene.cs:576-600): traced only while counting, so the
        work of a render without counters excludes this share."""
        out = (ctypes.c_int64 * 4)()
        check(N.lib().sr_counters_read_reference_only(self.scene.ctx, out), "counters")
        return dict(zip(["BvhIntersectNode", "BvhIntersectShape", "MeshIntersect", "SceneOcclude"], list(out)))


class Adjoint:
    """Adjoint rendering of a built scene (no reference counterpart): parameters are each object's
    albedo rgb and roughness, then the texture sky's texels; backward(d_image) returns
    d sum(d_image * Radiance) / d parameters for the last forward render (counter RNG policy,
    fixed sample count)."""

    def __init__(self, scene):
        self.scene = scene
        scene.build()
        n = ctypes.c_uint32()
        check(N.lib().sr_adjoint_param_count(scene.ctx, ctypes.byref(n)), "adjoint param count")
        self.param_count = n.value
        self._wh = None

    def params(self):
        p = np.empty(self.param_count, np.float32)
        check(N.lib().sr_adjoint_params_get(self.scene.ctx, fptr(p)), "adjoint params get")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.param_count:
            raise ValueError(f"expected {self.param_count} parameters")
        check(N.lib().sr_adjoint_params_set(self.scene.ctx, fptr(p)), "adjoint params set")

    def forward(self, view, width, height, spp, bounces, indirect_clamp, seed):
        rp = N.RenderParams_()
        rp.width, rp.height, rp.block_size = int(width), int(height), 32
        rp.min_samples = rp.max_samples = int(spp)
        rp.variance_threshold = 0.0
        rp.bounces, rp.indirect_clamp = int(bounces), float(indirect_clamp)
        rp.rng_policy, rp.rng_seed, rp.tick_bands, rp.light_sampler = RNG_COUNTER, int(seed), 1, LIGHT_CDF
        img = np.empty((int(width) * int(height), 3), np.float32)
        v = view._c()
        check(N.lib().sr_adjoint_forward(self.scene.ctx, ctypes.byref(v), ctypes.byref(rp), fptr(img)),
              "adjoint forward")
        self._wh = (int(width), int(height))
        return img

    def backward(self, d_image):
        if self._wh is None:
            raise RuntimeError("forward first")
        d = np.ascontiguousarray(d_image, np.float32).reshape(-1)
        if d.size != self._wh[0] * self._wh[1] * 3:
            raise ValueError("d_image must be width * height * 3 floats")
        g = np.empty(self.param_count, np.float32)
        check(N.lib().sr_adjoint_backward(self.scene.ctx, fptr(d), fptr(g)), "adjoint backward")
        return g


# Kernel-predicting denoiser topology (mirrors sr_cnn.hip): (cinA, cinB, cout, mode, srcA, srcB, level,
# relu); mode 0 same resolution, 1 2x2 average pool of srcA, 2 concat(nearest-up(srcA), srcB); -1 = the
# 10 input features.
CNN_LAYERS = [
    (10, 0, 32, 0, -1, -1, 0, True), (32, 0, 32, 0, 0, -1, 0, True),
    (32, 0, 48, 1, 1, -1, 1, True), (48, 0, 48, 0, 2, -1, 1, True),
    (48, 0, 64, 1, 3, -1, 2, True), (64, 0, 64, 0, 4, -1, 2, True),
    (64, 0, 96, 1, 5, -1, 3, True), (96, 0, 96, 0, 6, -1, 3, True),
    (96, 0, 128, 1, 7, -1, 4, True), (128, 0, 128, 0, 8, -1, 4, True),
    (128, 96, 96, 2, 9, 7, 3, True), (96, 0, 96, 0, 10, -1, 3, True),
    (96, 64, 64, 2, 11, 5, 2, True), (64, 0, 64, 0, 12, -1, 2, True),
    (64, 48, 48, 2, 13, 3, 1, True), (48, 0, 48, 0, 14, -1, 1, True),
    (48, 32, 32, 2, 15, 1, 0, True), (32, 0, 32, 0, 16, -1, 0, True),
    (32, 0, 100, 0, 17, -1, 0, False),
]


def cnn_weight_count():
    n = ctypes.c_uint64()
    check(N.lib().sr_cnn_weight_count(ctypes.byref(n)), "cnn weight count")
    return n.value


class KpnDenoiser:
    """Kernel-predicting U-Net denoiser of a scene's renders (no reference counterpart; the
    reference's Compositor.Filter is the bilateral alternative)."""

    def __init__(self, scene):
        self.scene = scene

    @staticmethod
    def random_weights(seed=0):
        """He-initialised weights (zero biases) in the blob layout sr_cnn_load expects."""
        rng = np.random.default_rng(seed)
        parts = []
        for cin_a, cin_b, cout, *_ in CNN_LAYERS:
            cin = cin_a + cin_b
            parts.append(rng.normal(0.0, np.sqrt(2.0 / (9 * cin)), size=cout * cin * 9).astype(np.float32))
            parts.append(np.zeros(cout, np.float32))
        return np.concatenate(parts)

    WEIGHTS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "kpn_unet.npz")

    @staticmethod
    def _bias_mask():
        parts = []
        for cin_a, cin_b, cout, *_ in CNN_LAYERS:
            parts += [np.zeros(cout * (cin_a + cin_b) * 9, bool), np.ones(cout, bool)]
        return np.concatenate(parts)

    @staticmethod
    def save_weights(path, weights):
        """Stores a blob compactly: conv weights as their bf16 bits (the precision the device
        multiplies with, round to nearest even), biases as float32 (the device adds them in fp32)."""
        w = np.ascontiguousarray(weights, np.float32)
        m = KpnDenoiser._bias_mask()
        if w.size != m.size:
            raise ValueError(f"expected {m.size} floats, got {w.size}")
        u = w[~m].view(np.uint32).astype(np.uint64)
        bits = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
        np.savez(path, wbits=bits, bias=w[m])

    @staticmethod
    def pretrained_weights(path=None):
        """The shipped weights, trained on tracer-rendered pairs (1 and 4 spp inputs, 1024 spp
        targets) by sharpray_amd.train_denoiser."""
        with np.load(path or KpnDenoiser.WEIGHTS_FILE, allow_pickle=False) as z:
            bits, bias = z["wbits"], z["bias"]
        m = KpnDenoiser._bias_mask()
        if bits.size + bias.size != m.size:
            raise ValueError("weight file does not match the network topology")
        w = np.empty(m.size, np.float32)
        w[~m] = (bits.astype(np.uint32) << 16).view(np.float32)
        w[m] = bias
        return w

    def load(self, weights):
        w = np.ascontiguousarray(weights, np.float32)
        check(N.lib().sr_cnn_load(self.scene.ctx, fptr(w), w.size), "cnn load")

    def denoise(self, renderer):
        out = np.empty((renderer.width * renderer.height, 3), np.float32)
        check(N.lib().sr_cnn_denoise(renderer.scene.ctx, fptr(out)), "cnn denoise")
        return out

    def layer(self, index):
        """Activations of layer `index` (-1: input features) of the last denoise, shape (h, w, c)."""
        c, w, h = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        L = N.lib()
        check(L.sr_cnn_layer_read(self.scene.ctx, index, None, ctypes.byref(c), ctypes.byref(w), ctypes.byref(h)),
              "cnn layer")
        out = np.empty((h.value, w.value, c.value), np.float32)
        check(L.sr_cnn_layer_read(self.scene.ctx, index, fptr(out), ctypes.byref(c), ctypes.byref(w), ctypes.byref(h)),
              "cnn layer")
        return out

    # ---- training ------------------------------------------------------------------------------
    def train_begin(self):
        check(N.lib().sr_cnn_train_begin(self.scene.ctx), "cnn train begin")

    def train_add(self, radiance, albedo, normals, depth, target, width, height):
        a = [np.ascontiguousarray(x, np.float32) for x in (radiance, albedo, normals, depth, target)]
        check(N.lib().sr_cnn_train_add(self.scene.ctx, *[fptr(x) for x in a], int(width), int(height)), "cnn train add")

    def train(self, steps, learning_rate=1e-3, seed=0):
        loss = ctypes.c_float()
        check(N.lib().sr_cnn_train_run(self.scene.ctx, int(steps), float(learning_rate), int(seed), ctypes.byref(loss)),
              "cnn train")
        return loss.value

    def gradient(self, example=0):
        """(loss, gradient blob) of a stored example without an update."""
        g = np.empty(cnn_weight_count(), np.float32)
        loss = ctypes.c_float()
        check(N.lib().sr_cnn_train_gradient(self.scene.ctx, int(example), fptr(g), ctypes.byref(loss)), "cnn gradient")
        return loss.value, g

    def export(self):
        w = np.empty(cnn_weight_count(), np.float32)
        check(N.lib().sr_cnn_train_export(self.scene.ctx, fptr(w)), "cnn export")
        return w

    def benchmark(self, renderer, iterations=10):
        """Milliseconds per device-resident denoise of the renderer's AOVs."""
        ms = ctypes.c_double()
        check(N.lib().sr_cnn_benchmark(renderer.scene.ctx, iterations, ctypes.byref(ms)), "cnn benchmark")
        return ms.value

    @staticmethod
    def flops(width, height):
        f = ctypes.c_double()
        check(N.lib().sr_cnn_flops(width, height, ctypes.byref(f)), "cnn flops")
        return f.value


class ObjConfig:
    """ObjConfig (obj.cs:127-134): transform of every object, material for objects without an MTL
    material (None: Material(White, roughness 1)), smooth normals for files without vn."""

    def __init__(self, transform=None, material=None, smooth_normals=True):
        self.transform = transform
        self.material = material
        self.smooth_normals = bool(smooth_normals)


class ObjLoader:
    """ObjLoader (obj.cs:136-512), natively: OBJ + MTL (+ TGA maps) into a scene."""

    @staticmethod
    def load(path, scene, config=None):
        """Adds the file's objects to `scene`; returns how many were added."""
        config = config or ObjConfig()
        scene._guard()
        t = config.transform._c() if config.transform is not None else None
        m = None
        if config.material is not None:
            m = scene._material_struct(config.material)
        n = ctypes.c_uint32()
        check(N.lib().sr_obj_load(scene.ctx, str(path).encode(), ctypes.byref(t) if t is not None else None,
                                  ctypes.byref(m) if m is not None else None, 1 if config.smooth_normals else 0,
                                  ctypes.byref(n)), "obj load")
        return n.value


def load_hdr(path, max_luminance=None):
    """Texture.FromHdr(ImageHdr.Load(path), maxLuminance) (image.cs:307-382, texture.cs:154-163):
    (h, w, 3) float32 texels."""
    w, h = ctypes.c_uint32(), ctypes.c_uint32()
    lim = np.float32(np.finfo(np.float32).max if max_luminance is None else max_luminance)
    check(N.lib().sr_hdr_load(str(path).encode(), lim, None, ctypes.byref(w), ctypes.byref(h)), "hdr load")
    out = np.empty((h.value, w.value, 3), np.float32)
    check(N.lib().sr_hdr_load(str(path).encode(), lim, fptr(out), ctypes.byref(w), ctypes.byref(h)), "hdr load")
    return out


def decode_hdr(data, max_luminance=None):
    """The same decoding from the bytes of a .hdr file."""
    buf = np.frombuffer(bytes(data), np.uint8)
    w, h = ctypes.c_uint32(), ctypes.c_uint32()
    lim = np.float32(np.finfo(np.float32).max if max_luminance is None else max_luminance)
    p = buf.ctypes.data_as(N.c_uint8_p)
    check(N.lib().sr_hdr_decode(p, buf.size, lim, None, ctypes.byref(w), ctypes.byref(h)), "hdr decode")
    out = np.empty((h.value, w.value, 3), np.float32)
    check(N.lib().sr_hdr_decode(p, buf.size, lim, fptr(out), ctypes.byref(w), ctypes.byref(h)), "hdr decode")
    return out


def decode_tga(data):
    """Image.LoadTga (image.cs:217-266) of the bytes of a .tga file: (h, w, 3) uint8, top row first."""
    buf = np.frombuffer(bytes(data), np.uint8)
    w, h = ctypes.c_uint32(), ctypes.c_uint32()
    p = buf.ctypes.data_as(N.c_uint8_p)
    check(N.lib().sr_tga_decode(p, buf.size, None, ctypes.byref(w), ctypes.byref(h)), "tga decode")
    out = np.empty((h.value, w.value, 3), np.uint8)
    check(N.lib().sr_tga_decode(p, buf.size, out.ctypes.data_as(N.c_uint8_p), ctypes.byref(w), ctypes.byref(h)),
          "tga decode")
    return out


TONEMAP_REINHARD = 0
TONEMAP_LINEAR_SMOOTH = 1


class Image:
    """Image (image.cs): width x height RGB bytes, row-major, top row first."""

    def __init__(self, width, height, pixels=None):
        self.width, self.height = int(width), int(height)
        self.pixels = (np.zeros((self.height, self.width, 3), np.uint8) if pixels is None
                       else np.ascontiguousarray(pixels, np.uint8).reshape(self.height, self.width, 3))

    def encode_bmp(self):
        """The bytes Image.WriteBmp (image.cs:147-191) writes."""
        size = ctypes.c_uint64()
        px = self.pixels.ctypes.data_as(N.c_uint8_p)
        check(N.lib().sr_image_encode_bmp(px, self.width, self.height, None, ctypes.byref(size)), "bmp")
        out = np.empty(size.value, np.uint8)
        check(N.lib().sr_image_encode_bmp(px, self.width, self.height, out.ctypes.data_as(N.c_uint8_p),
                                          ctypes.byref(size)), "bmp")
        return out.tobytes()

    def save(self, path):
        """Image.Save (image.cs) for .bmp paths."""
        if not str(path).lower().endswith(".bmp"):
            raise ValueError("only .bmp output is supported")
        check(N.lib().sr_image_write_bmp(str(path).encode(), self.pixels.ctypes.data_as(N.c_uint8_p), self.width,
                                         self.height), "bmp write")


class Compositor:
    """Compositor (compose.cs:7-267): Filter (compose.cs:101-204), FilterFog (206-245), Tonemap
    (247-266), Preview and Compose (57-99; the overlay is the C# host's and is not drawn)."""

    def __init__(self, denoise_radius, denoise_strength, denoise_strength_max, denoise_luminance_boost,
                 denoise_luminance_limit, denoise_normal_limit, denoise_depth_limit, denoise_fog_radius=0.0,
                 denoise_fog_strength=0.0, tonemapper=TONEMAP_REINHARD, exposure=1.0):
        p = N.DenoiseParams_()
        p.radius, p.strength, p.strength_max = denoise_radius, denoise_strength, denoise_strength_max
        p.luminance_boost, p.luminance_limit = denoise_luminance_boost, denoise_luminance_limit
        p.normal_limit, p.depth_limit = denoise_normal_limit, denoise_depth_limit
        self.params = p
        c = N.ComposeParams_()
        c.denoise = p
        c.tonemapper, c.exposure = int(tonemapper), float(exposure)
        c.fog_radius, c.fog_strength = float(denoise_fog_radius), float(denoise_fog_strength)
        self.compose_params = c

    def filter_fog(self, renderer):
        """FilterFog (compose.cs:206-245) of the renderer's fog radiance: (w*h, 3) floats."""
        out = np.empty((renderer.width * renderer.height, 3), np.float32)
        c = self.compose_params
        check(N.lib().sr_filter_fog(renderer.scene.ctx, c.fog_radius, c.fog_strength, fptr(out)), "fog filter")
        return out

    def preview(self, renderer):
        img = Image(renderer.width, renderer.height)
        c = self.compose_params
        check(N.lib().sr_preview(renderer.scene.ctx, c.tonemapper, c.exposure, img.pixels.ctypes.data_as(N.c_uint8_p)),
              "preview")
        return img

    def compose(self, renderer):
        img = Image(renderer.width, renderer.height)
        check(N.lib().sr_compose(renderer.scene.ctx, ctypes.byref(self.compose_params),
                                 img.pixels.ctypes.data_as(N.c_uint8_p)), "compose")
        return img

    def as_array(self):
        p = self.params
        return np.array([p.radius, p.strength, p.strength_max, p.luminance_boost, p.luminance_limit, p.normal_limit,
                         p.depth_limit], dtype=np.float32)

    def filter(self, renderer, albedo_limit=None, sample_mean=False):
        """Filter (compose.cs:101-204) of the renderer's surface radiance; with albedo_limit the
        albedo AOV joins normal and depth as a guide; with sample_mean the filter sees
        Radiance * Samples / min(Samples, max_samples), the unbiased per-pixel mean (neither has a
        reference counterpart)."""
        out = np.empty((renderer.width * renderer.height, 3), np.float32)
        if sample_mean:
            if albedo_limit is not None:
                raise ValueError("sample_mean and albedo_limit cannot be combined")
            check(N.lib().sr_denoise_bilateral_mean(renderer.scene.ctx, ctypes.byref(self.params), fptr(out)),
                  "denoise")
        elif albedo_limit is None:
            check(N.lib().sr_denoise_bilateral(renderer.scene.ctx, ctypes.byref(self.params), fptr(out)), "denoise")
        else:
            check(N.lib().sr_denoise_bilateral_guided(renderer.scene.ctx, ctypes.byref(self.params),
                                                      float(albedo_limit), fptr(out)), "denoise")
        return out
