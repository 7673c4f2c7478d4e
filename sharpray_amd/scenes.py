"""Seeded procedural workloads for the tracer (the reference's assets are not available).

A `SceneSpec` is a plain description (triangle arrays, objects, sky, fog, view, render settings)
that can be instantiated on the GPU through the public API (`instantiate`) and handed unchanged
to any other consumer, so that two renderers see bit-identical inputs.

Workloads:
  - `config_plain_cpu`: 256x256, 16 spp, 6 bounces; a seeded level-5 icosphere (20,480 triangles)
    over a 2-triangle ground quad, under a 512x256 procedural HDR sky (gradient + Gaussian sun
    lobe of peak radiance 50) quantised through RGBE like an .hdr file; Lambertian albedo 0.7
    and GGX roughness 0.3; mesh BVH with 16 bins.
  - `config_plain_gpu`: 1920x1080, 64 spp, 8 bounces; a 1,048,576-triangle sphere (512 rings x 1024
    segments) displaced by 6-octave fBm noise (amplitude 0.15) plus a ground plane; eight
    materials (albedo U[0.05, 0.95], roughness U[0.02, 1], metallic in {0, 1}), one per latitude
    band of the sphere; a 2048x1024 float32 environment map with log-normal texel radiance
    (sigma 1.5), importance sampled through its CDF with MIS.
Both take a `scale` so tests can run the same scene structure at reduced size.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import (MEDIUM_GRID, AABox, Fog, FogVolume, Material, Mesh, Object, Quat, Renderer, Scene, SkyProcedural,
               SkyTexture, Sphere, Texture, Transform, View, triangles_build)

SEED_BASE = 0x5EED0000


@dataclass
class ObjectSpec:
    name: str
    transform: np.ndarray          # pos3 quat4 scale3
    material: np.ndarray           # albedo3 roughness metallic transparency ior radiance3
    kind: int = 0                  # 0 mesh, 1 sphere, 2 box
    mesh: int = -1                 # index into SceneSpec.meshes
    params: np.ndarray = None      # sphere (c3, r) / box (min3, max3)
    textures: tuple = (-1, -1, -1, -1)


@dataclass
class SceneSpec:
    meshes: list = field(default_factory=list)       # (n, 36) float32 Triangle records
    mesh_bins: list = field(default_factory=list)
    objects: list = field(default_factory=list)
    textures: list = field(default_factory=list)     # (h, w, 3) float32
    texture_modes: dict = field(default_factory=dict)  # texture index -> (filter, tiling x, tiling y)
    sky_texture: int = -1
    sky_multiplier: float = 1.0
    sky_yaw: float = 0.0
    sky_procedural: np.ndarray = None               # 16 floats
    fog: tuple = None                                # (density, color3, anisotropy, falloff)
    fog_volume: object = None                        # sharpray_amd.FogVolume (bounded medium)
    view: np.ndarray = None                          # pos3 quat4 fov near
    width: int = 0
    height: int = 0
    min_samples: int = 1
    max_samples: int = 1
    variance_threshold: float = 0.0
    bounces: int = 8
    indirect_clamp: float = 10.0
    name: str = ""

    @property
    def triangle_count(self):
        return sum(int(m.shape[0]) for m in self.meshes)


def instantiate(spec, device=0, counters=False):
    """Builds a GPU Scene (and its View) from a spec through the public API; counters: collect the
    Counters.Type book from the first mesh on."""
    scene = Scene(device)
    if counters:
        scene.counters_enable(True)
    textures = [Texture(t, *_texture_mode(spec, i)) for i, t in enumerate(spec.textures)]
    meshes = [Mesh(m, split_bins=b) for m, b in zip(spec.meshes, spec.mesh_bins)]
    for m in meshes:  # registration order = spec index = device mesh id
        scene.mesh_id(m)
    if spec.sky_texture >= 0:
        scene.set_sky(SkyTexture(textures[spec.sky_texture], spec.sky_multiplier, spec.sky_yaw))
    else:
        p = spec.sky_procedural
        scene.set_sky(SkyProcedural(p[0:3], p[3:6], p[6:9], p[9:12], p[12:15], p[15]))
    if spec.fog is not None:
        d, c, g, f = spec.fog
        scene.set_fog(Fog(d, c, g, f))
    if spec.fog_volume is not None:
        scene.set_fog(spec.fog_volume)
    for o in spec.objects:
        t = o.transform
        tr = Transform(t[0:3], Quat(*t[3:7]), t[7:10])
        m = o.material
        mat = Material(m[0:3], m[3], m[4], m[5], m[6], m[7:10],
                       *[textures[i] if i >= 0 else None for i in o.textures])
        if o.kind == 0:
            shape = meshes[o.mesh]
        elif o.kind == 1:
            shape = Sphere(o.params[0:3], o.params[3])
        else:
            shape = AABox(o.params[0:3], o.params[3:6])
        scene.add_object(Object(o.name, tr, mat, shape))
    v = spec.view
    view = View(Transform(v[0:3], Quat(*v[3:7])), float(v[7]), float(v[8]))
    return scene, view


def _texture_mode(spec, i):
    f, tx, ty = getattr(spec, "texture_modes", {}).get(i, (0, 1.0, 1.0))
    return f, (tx, ty)


def render(spec, device=0, **kw):
    scene, view = instantiate(spec, device)
    r = Renderer(scene, view, spec.width, spec.height, 32, spec.min_samples, spec.max_samples,
                 spec.variance_threshold, spec.bounces, spec.indirect_clamp, **kw)
    return scene, view, r


# ---- geometry -------------------------------------------------------------------------------------
def _rotation_matrix(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def icosphere(level, seed=0, radius=1.0, center=(0.0, 0.0, 0.0)):
    """Positions (20 * 4^level, 9) of a unit icosphere with a seeded orientation, CCW from outside."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    v = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in verts]
    cache = {}

    def mid(a, b):
        key = (min(a, b), max(a, b))
        if key not in cache:
            m = v[a] + v[b]
            v.append(m / np.linalg.norm(m))
            cache[key] = len(v) - 1
        return cache[key]

    for _ in range(level):
        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(v) @ _rotation_matrix(np.random.default_rng(seed)).T * radius + np.asarray(center)
    F = np.array(faces)
    # Wind so the face normal points away from the centre (CCW seen from outside, right-handed).
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    n = np.cross(b - a, c - a)
    flip = np.einsum("ij,ij->i", n, a - np.asarray(center)) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return V[F].reshape(-1, 9).astype(np.float32)


def quad(y, half, z0=None, z1=None):
    """Two upward-facing triangles spanning [-half, half]^2 (or [z0, z1] in z) at height y."""
    z0 = -half if z0 is None else z0
    z1 = half if z1 is None else z1
    p00, p10 = (-half, y, z0), (half, y, z0)
    p01, p11 = (-half, y, z1), (half, y, z1)
    pos = np.array([p00 + p01 + p11, p00 + p11 + p10], dtype=np.float32)
    uv = np.array([[0, 0, 0, 1, 1, 1], [0, 0, 1, 1, 1, 0]], dtype=np.float32)
    return pos, uv


def _hash3(ix, iy, iz, seed):
    salt = np.uint64((int(seed) * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF)
    h = (ix.astype(np.uint64) * np.uint64(0x8DA6B343) ^ iy.astype(np.uint64) * np.uint64(0xD8163841)
         ^ iz.astype(np.uint64) * np.uint64(0xCB1AB31F) ^ salt)
    h ^= h >> np.uint64(33)
    h *= np.uint64(0xFF51AFD7ED558CCD)
    h ^= h >> np.uint64(33)
    h *= np.uint64(0xC4CEB9FE1A85EC53)
    h ^= h >> np.uint64(33)
    return (h >> np.uint64(11)).astype(np.float64) / float(1 << 53) * 2.0 - 1.0


def value_noise(p, seed):
    """Trilinearly interpolated lattice noise in [-1, 1] with a quintic fade, points (n, 3)."""
    f = np.floor(p)
    t = p - f
    i = f.astype(np.int64) + (1 << 20)
    w = t * t * t * (t * (t * 6.0 - 15.0) + 10.0)
    acc = np.zeros(p.shape[0])
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                h = _hash3(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed)
                wx = w[:, 0] if dx else 1.0 - w[:, 0]
                wy = w[:, 1] if dy else 1.0 - w[:, 1]
                wz = w[:, 2] if dz else 1.0 - w[:, 2]
                acc += h * wx * wy * wz
    return acc


def fbm(p, octaves, seed, base_freq=2.0):
    total = np.zeros(p.shape[0])
    amp, norm, freq = 1.0, 0.0, base_freq
    for o in range(octaves):
        total += amp * value_noise(p * freq, seed + o)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def fbm_sphere(rings, segments, amplitude, octaves, seed, radius=1.0):
    """CreateSphere topology (mesh.cs:125-165): 2 * rings * segments triangles on a UV sphere whose
    vertices are displaced along the radius by fBm. Returns (pos (n, 9), uv (n, 6), ring index)."""
    theta = np.pi * np.arange(rings + 1) / rings
    phi = np.pi + 2.0 * np.pi * np.arange(segments) / segments
    st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
    dirs = np.stack([st * np.cos(phi)[None, :], np.broadcast_to(ct, (rings + 1, segments)),
                     st * np.sin(phi)[None, :]], axis=-1)
    dirs[0, :] = (0.0, 1.0, 0.0)
    dirs[-1, :] = (0.0, -1.0, 0.0)
    d = fbm(dirs.reshape(-1, 3), octaves, seed).reshape(rings + 1, segments)
    d[0, :] = d[0, 0]
    d[-1, :] = d[-1, 0]
    P = dirs * (radius * (1.0 + amplitude * d))[..., None]
    r = np.arange(rings)[:, None]
    s = np.arange(segments)[None, :]
    s1 = (s + 1) % segments
    p00, p01 = P[r, s], P[r, s1]
    p10, p11 = P[r + 1, s], P[r + 1, s1]
    u0 = np.broadcast_to(s / segments, (rings, segments))
    u1 = np.broadcast_to((s + 1) / segments, (rings, segments))
    v0 = np.broadcast_to(r / rings, (rings, segments))
    v1 = np.broadcast_to((r + 1) / rings, (rings, segments))
    t1 = np.concatenate([p00, p11, p10], axis=-1)
    t2 = np.concatenate([p00, p01, p11], axis=-1)
    pos = np.stack([t1, t2], axis=2).reshape(-1, 9)
    uv1 = np.stack([u0, v0, u1, v1, u0, v1], axis=-1)
    uv2 = np.stack([u0, v0, u1, v0, u1, v1], axis=-1)
    uv = np.stack([uv1, uv2], axis=2).reshape(-1, 6)
    ring = np.repeat(np.arange(rings), segments * 2)
    return pos.astype(np.float32), uv.astype(np.float32), ring


# ---- skies ----------------------------------------------------------------------------------------
def rgbe_encode(rgb):
    """Float RGB (h, w, 3) -> RGBE bytes (h, w, 4) as written to Radiance .hdr files (Ward's rule)."""
    rgb = np.asarray(rgb, dtype=np.float32)
    m = rgb.max(axis=-1).astype(np.float64)
    out = np.zeros(rgb.shape[:-1] + (4,), dtype=np.uint8)
    ok = m > 1e-32
    mant, expo = np.frexp(m[ok])
    scale = mant * 256.0 / m[ok]
    out[ok, :3] = np.clip(np.floor(rgb[ok].astype(np.float64) * scale[:, None]), 0, 255).astype(np.uint8)
    out[ok, 3] = (expo + 128).astype(np.uint8)
    return out


def write_hdr(path, rgbe):
    """Writes RGBE bytes (h, w, 4) as a Radiance .hdr file with flat (uncompressed) scanlines."""
    rgbe = np.ascontiguousarray(rgbe, dtype=np.uint8)
    h, w, _ = rgbe.shape
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode() + rgbe.tobytes())


def write_obj(path, pos, uv=None, name="mesh"):
    """Writes triangles (n, 9) (and per-corner uvs (n, 6), stored with V flipped back) as an OBJ
    without vn, three vertices per face, every float with 9 significant digits (exact for float32)."""
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    lines = [f"o {name}"]
    lines += ["v %.9g %.9g %.9g" % tuple(float(c) for c in p) for p in pos]
    if uv is not None:
        t = np.asarray(uv, np.float32).reshape(-1, 2)
        lines += ["vt %.9g %.9g" % (float(a), float(np.float32(1.0) - b)) for a, b in t]
        lines += [f"f {3 * i + 1}/{3 * i + 1} {3 * i + 2}/{3 * i + 2} {3 * i + 3}/{3 * i + 3}" for i in range(len(pos) // 3)]
    else:
        lines += [f"f {3 * i + 1} {3 * i + 2} {3 * i + 3}" for i in range(len(pos) // 3)]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def procedural_sky(w, h, peak=50.0, sun_dir=(0.35, 0.6, 0.45), sun_width=0.05, seed=0):
    """Equirectangular float sky: zenith-to-horizon gradient, dark ground, Gaussian sun lobe."""
    u = (np.arange(w) + 0.5) / w
    v = (np.arange(h) + 0.5) / h
    phi = (u - 0.5) * 2.0 * np.pi
    theta = np.pi * v
    d = np.stack([np.sin(theta)[:, None] * np.cos(phi)[None, :], np.broadcast_to(np.cos(theta)[:, None], (h, w)),
                  np.sin(theta)[:, None] * np.sin(phi)[None, :]], axis=-1)
    y = d[..., 1]
    zen = np.array([0.25, 0.45, 0.95])
    hor = np.array([0.9, 0.9, 0.85])
    gnd = np.array([0.12, 0.1, 0.08])
    up = np.clip(y, 0.0, 1.0)[..., None]
    sky = np.where(y[..., None] >= 0.0, hor + (zen - hor) * up ** 0.6, gnd + (hor - gnd) * np.exp(8.0 * y)[..., None])
    s = np.asarray(sun_dir, dtype=np.float64)
    s /= np.linalg.norm(s)
    ang = np.arccos(np.clip(d @ s, -1.0, 1.0))
    sun = peak * np.exp(-0.5 * (ang / sun_width) ** 2)
    rgb = sky + sun[..., None] * np.array([1.0, 0.95, 0.85])
    return rgb.astype(np.float32)


def lognormal_env(w, h, sigma, seed):
    """Float32 environment map with independent log-normal texel radiance (mean 1) and a warm tint."""
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((h, w))
    lum = np.exp(sigma * z - 0.5 * sigma * sigma)
    tint = np.array([1.0, 0.93, 0.85])
    return (lum[..., None] * tint).astype(np.float32)


def _view(pos, pitch_deg, fov_deg, near=0.1):
    q = Quat.angle_axis(math.radians(pitch_deg), (1.0, 0.0, 0.0)).xyzw
    return np.array(list(pos) + q.tolist() + [np.float32(math.radians(fov_deg)), near], dtype=np.float32)


def _identity():
    return np.array([0, 0, 0, 0, 0, 0, 1, 1, 1, 1], dtype=np.float32)


def _mat(albedo, roughness, metallic=0.0, transparency=0.0, ior=1.5, radiance=(0, 0, 0)):
    a = albedo if np.ndim(albedo) else (albedo,) * 3
    return np.array(list(a) + [roughness, metallic, transparency, ior] + list(radiance), dtype=np.float32)


def config_plain_cpu(scale=1.0, spp=16, level=5, seed=SEED_BASE + 0):
    """The reference's CPU-runnable configuration (see module docstring)."""
    s = SceneSpec(name="plain_cpu_256px_16spp_ico5")
    s.width = max(8, int(round(256 * scale)))
    s.height = max(8, int(round(256 * scale)))
    s.min_samples = s.max_samples = spp
    s.bounces = 6
    s.indirect_clamp = 10.0
    sky = procedural_sky(512, 256, peak=50.0)
    s.textures.append(Texture.from_hdr(rgbe_encode(sky)).rgb)
    s.sky_texture = 0
    ico = icosphere(level, seed=seed, radius=1.0, center=(0.0, 0.0, 0.0))
    s.meshes.append(triangles_build(ico, None, None, smooth=True))
    s.mesh_bins.append(16)
    gp, guv = quad(-1.0, 6.0)
    s.meshes.append(triangles_build(gp, None, guv))
    s.mesh_bins.append(16)
    s.objects.append(ObjectSpec("icosphere", _identity(), _mat(0.7, 0.3, metallic=1.0), 0, 0))
    s.objects.append(ObjectSpec("ground", _identity(), _mat(0.7, 1.0), 0, 1))
    s.view = _view((0.0, 0.6, -3.2), 8.0, 55.0)
    return s


def config_plain_gpu(scale=1.0, spp=64, rings=512, segments=1024, env=(2048, 1024), seed=SEED_BASE + 1):
    """The 1080p / 1M-triangle forward configuration (see module docstring)."""
    s = SceneSpec(name="plain_gpu_1080p_64spp_1Mtri")
    s.width = max(8, int(round(1920 * scale)))
    s.height = max(8, int(round(1080 * scale)))
    s.min_samples = s.max_samples = spp
    s.bounces = 8
    s.indirect_clamp = 10.0
    rng = np.random.default_rng(seed)
    s.textures.append(lognormal_env(env[0], env[1], 1.5, seed + 100))
    s.sky_texture = 0
    pos, uv, ring = fbm_sphere(rings, segments, 0.15, 6, seed + 200)
    tris = triangles_build(pos, None, uv, smooth=True)
    bands = 8
    per = rings // bands
    for b in range(bands):
        sel = (ring >= b * per) & (ring < (b + 1) * per) if b < bands - 1 else ring >= b * per
        s.meshes.append(np.ascontiguousarray(tris[sel]))
        s.mesh_bins.append(8)
    gp, guv = quad(-1.25, 12.0)
    s.meshes.append(triangles_build(gp, None, guv))
    s.mesh_bins.append(8)
    for b in range(bands):
        albedo = rng.uniform(0.05, 0.95, size=3)
        rough = rng.uniform(0.02, 1.0)
        metal = float(rng.integers(0, 2))
        s.objects.append(ObjectSpec(f"band_{b}", _identity(), _mat(albedo, rough, metal), 0, b))
    s.objects.append(ObjectSpec("ground", _identity(), _mat(0.5, 1.0), 0, bands))
    s.view = _view((0.0, 0.35, -3.4), 4.0, 60.0)
    return s


def fbm_density_grid(res, seed, octaves=4, freq=0.6, box=((-6.0, -1.25, -3.0), (6.0, 3.0, 9.0))):
    """A res^3 half-precision density grid in [0, 1] (shape (nz, ny, nx)): fBm value noise sampled at
    the voxel centres of `box`, remapped by clip(0.5 + fbm, 0, 1)."""
    lo, hi = np.array(box[0]), np.array(box[1])
    ax = [lo[k] + (np.arange(res) + 0.5) * (hi[k] - lo[k]) / res for k in range(3)]
    z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    d = np.clip(0.5 + fbm(p, octaves, seed, base_freq=freq), 0.0, 1.0)
    return d.reshape(res, res, res).astype(np.float16), box


def config_fog(scale=1.0, spp=4, res=128, seed=SEED_BASE + 2):
    """The forward scene with a res^3 fp16 fBm density grid (extinction 0 to 0.2 per unit), HG
    anisotropy 0.3, delta-tracked free flight."""
    s = config_plain_gpu(scale, spp)
    s.name = f"fog_grid_1080p_{spp}spp_1Mtri"
    grid, box = fbm_density_grid(res, seed)
    s.fog_volume = FogVolume(MEDIUM_GRID, 0.2, (0.95, 0.95, 0.95), 0.3, box[0], box[1], grid)
    return s


def config_adjoint(scale=1.0, spp=16, bounces=4, rings=224, segments=224, materials=16, seed=SEED_BASE + 3):
    """The adjoint workload: 512x512, a ~100k-triangle fBm noise sphere split into `materials`
    latitude bands (one material each: albedo, roughness, metallic 0 or 1) over a ground quad,
    under a 256x128 log-normal environment map. Parameters: 4 per object + 256*128*3 texels."""
    s = SceneSpec(name=f"adjoint_512px_{spp}spp_100ktri")
    s.width = max(8, int(round(512 * scale)))
    s.height = max(8, int(round(512 * scale)))
    s.min_samples = s.max_samples = spp
    s.bounces = bounces
    s.indirect_clamp = 10.0
    rng = np.random.default_rng(seed)
    s.textures.append(lognormal_env(256, 128, 1.0, seed + 100))
    s.sky_texture = 0
    pos, uv, ring = fbm_sphere(rings, segments, 0.15, 5, seed + 200)
    tris = triangles_build(pos, None, uv, smooth=True)
    per = rings // materials
    for b in range(materials):
        sel = (ring >= b * per) & (ring < (b + 1) * per) if b < materials - 1 else ring >= b * per
        s.meshes.append(np.ascontiguousarray(tris[sel]))
        s.mesh_bins.append(8)
    gp, guv = quad(-1.25, 12.0)
    s.meshes.append(triangles_build(gp, None, guv))
    s.mesh_bins.append(8)
    for b in range(materials):
        s.objects.append(ObjectSpec(f"band_{b}", _identity(),
                                    _mat(rng.uniform(0.1, 0.9, size=3), rng.uniform(0.2, 0.9),
                                         float(rng.integers(0, 2))), 0, b))
    s.objects.append(ObjectSpec("ground", _identity(), _mat(0.5, 0.8), 0, materials))
    s.view = _view((0.0, 0.3, -3.0), 4.0, 50.0)
    return s


def config_adjoint_small(width=32, height=24, spp=16, bounces=2):
    """A small analytic scene for gradient checks: a box floor and four spheres (dielectric and
    metal, rough and glossy) under a 32x16 log-normal sky; bounces 2 keeps Russian roulette off."""
    s = SceneSpec(name="adjoint_small")
    s.textures.append(lognormal_env(32, 16, 1.0, 123))
    s.sky_texture = 0
    s.width, s.height = width, height
    s.min_samples = s.max_samples = spp
    s.bounces = bounces
    s.indirect_clamp = 1e30
    s.objects.append(ObjectSpec("floor", _identity(), _mat((0.6, 0.5, 0.4), 0.8, 0.0), 2, -1,
                                np.array([-5, -1.2, -3, 5, -1.0, 8], np.float32)))
    mats = [((0.8, 0.2, 0.2), 0.5, 0.0), ((0.2, 0.7, 0.3), 0.3, 1.0), ((0.3, 0.3, 0.9), 0.7, 0.0),
            ((0.9, 0.8, 0.2), 0.25, 1.0)]
    for i, (a, r, m) in enumerate(mats):
        c = (-1.5 + i * 1.0, -0.55 + 0.1 * i, 2.5 + 0.4 * (i % 2))
        s.objects.append(ObjectSpec(f"s{i}", _identity(), _mat(a, r, m), 1, -1,
                                    np.array(list(c) + [0.45, 0, 0], np.float32)))
    s.view = _view((0.0, 0.2, -1.0), 5.0, 60.0)
    return s


def config_analytic(scale=1.0, spp=4):
    """program.cs-like scene (spheres and a box floor under a procedural sky with sun) exercising
    the analytic shapes and the procedural sky."""
    s = SceneSpec(name="analytic_spheres")
    s.width = max(8, int(round(160 * scale)))
    s.height = max(8, int(round(90 * scale)))
    s.min_samples = s.max_samples = spp
    s.bounces = 8
    s.sky_procedural = np.array([0.08, 0.17, 0.70, 0.50, 0.65, 0.90, 0.12, 0.09, 0.07,
                                 *(np.array([0.5, 1.0, -0.5]) / np.linalg.norm([0.5, 1.0, -0.5])),
                                 100000.0, 90000.0, 65000.0, math.radians(0.53)], dtype=np.float32)
    s.objects.append(ObjectSpec("floor", _identity(), _mat(1.0, 1.0), 2, -1,
                                np.array([-30, -0.2, -2, 30, 0, 15], dtype=np.float32)))
    for i in range(10):
        ang = i * (math.pi * 2.0 / 10)
        rough = 1.0 - i / 9.0
        c = (1.0 + math.cos(ang) * 3.5, 0.75, 4.0 + math.sin(ang) * 3.5)
        s.objects.append(ObjectSpec(f"sphere_{i}", _identity(), _mat((1.0, 0.35, 0.05), rough, 1.0), 1, -1,
                                    np.array(list(c) + [0.75, 0, 0], dtype=np.float32)))
    s.view = _view((1.0, 5.0, -2.0), 35.0, 75.0)
    return s


def _normal_bumps(w, h, seed, strength=0.6):
    """A tangent-space normal map (texels in [0, 1]) of random smooth bumps: texel = (n + 1) / 2 with
    the G channel flipped as Texture.SampleNormal (texture.cs:71-75) reads it back."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    hgt = np.zeros((h, w))
    for _ in range(12):
        cx, cy, r = rng.uniform(0, w), rng.uniform(0, h), rng.uniform(2, w / 4)
        dx = (xx - cx + w / 2) % w - w / 2
        dy = (yy - cy + h / 2) % h - h / 2
        hgt += np.exp(-(dx * dx + dy * dy) / (2 * r * r)) * rng.uniform(-1, 1)
    gx = (np.roll(hgt, -1, 1) - np.roll(hgt, 1, 1)) * strength
    gy = (np.roll(hgt, -1, 0) - np.roll(hgt, 1, 0)) * strength
    n = np.stack([-gx, -gy, np.ones_like(hgt)], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    return np.stack([(n[..., 0] + 1) / 2, (1 - n[..., 1]) / 2, (n[..., 2] + 1) / 2], -1).astype(np.float32)


def _unit_axis(v):
    v = np.asarray(v, np.float64)
    return tuple((v / np.linalg.norm(v)).tolist())


def config_textured_transformed(scale=1.0, spp=4, seed=SEED_BASE + 5):
    """Object transforms and textured materials (program.cs:95-130 uses both): a box floor with
    tiled color, roughness and normal-map textures (Material.Sample*, scene.cs:173-187); an
    icosphere mesh rotated, translated and non-uniformly scaled, with a metallic texture; an
    ellipsoid (scaled, rotated sphere); a rotated box; a glass ellipsoid; under a log-normal sky."""
    s = SceneSpec(name="textured_transformed")
    s.width = max(8, int(round(160 * scale)))
    s.height = max(8, int(round(96 * scale)))
    s.min_samples = s.max_samples = spp
    s.bounces = 6
    s.indirect_clamp = 10.0
    rng = np.random.default_rng(seed)
    s.textures.append(lognormal_env(64, 32, 1.0, seed + 100))
    s.sky_texture = 0
    yy, xx = np.mgrid[0:32, 0:32]
    checker = ((xx // 8 + yy // 8) % 2)[..., None]
    s.textures.append((0.25 + 0.6 * checker * np.array([1.0, 0.8, 0.6]) + rng.uniform(0, 0.1, (32, 32, 3))
                       ).astype(np.float32))                                   # 1 floor color
    s.textures.append(np.repeat(rng.uniform(0.2, 1.0, (16, 16, 1)), 3, 2).astype(np.float32))  # 2 roughness
    s.textures.append(_normal_bumps(32, 32, seed + 7))                          # 3 normal map
    s.textures.append(np.repeat(rng.uniform(0.0, 1.0, (8, 8, 1)), 3, 2).astype(np.float32))   # 4 metallic
    s.texture_modes = {1: (0, 6.0, 1.7), 2: (0, 6.0, 1.7), 3: (0, 6.0, 1.7), 4: (1, 2.0, 2.0)}
    s.objects.append(ObjectSpec("floor", _identity(), _mat(1.0, 1.0), 2, -1,
                                np.array([-30, -0.2, -2, 30, 0, 15], np.float32), textures=(1, 2, -1, 3)))
    ico = icosphere(3, seed=seed)
    s.meshes.append(triangles_build(ico, None, None, smooth=True))
    s.mesh_bins.append(8)
    q = Quat.angle_axis(math.radians(-40.0), _unit_axis((0.3, 1.0, 0.1))).xyzw
    s.objects.append(ObjectSpec("ico", np.array([-1.2, 0.9, 4.5] + list(q) + [0.8, 1.1, 0.6], np.float32),
                                _mat((0.8, 0.5, 0.3), 0.35, 0.9), 0, 0, textures=(-1, -1, 4, -1)))
    q = Quat.angle_axis(math.radians(25.0), (0.0, 0.0, 1.0)).xyzw
    s.objects.append(ObjectSpec("ellipsoid", np.array([1.4, 0.7, 4.0] + list(q) + [1.6, 0.7, 0.9], np.float32),
                                _mat((0.2, 0.5, 0.9), 0.5, 0.0), 1, -1, np.array([0, 0, 0, 0.8, 0, 0], np.float32)))
    q = Quat.angle_axis(math.radians(35.0), _unit_axis((1.0, 1.0, 0.0))).xyzw
    s.objects.append(ObjectSpec("box", np.array([0.2, 0.5, 6.0] + list(q) + [1.0, 1.0, 1.0], np.float32),
                                _mat((0.9, 0.9, 0.9), 0.15, 1.0), 2, -1,
                                np.array([-0.5, -0.5, -0.5, 0.5, 0.5, 0.5], np.float32)))
    q = Quat.angle_axis(math.radians(60.0), (0.0, 1.0, 0.0)).xyzw
    s.objects.append(ObjectSpec("glass", np.array([0.0, 0.45, 3.0] + list(q) + [1.2, 0.9, 0.8], np.float32),
                                _mat((0.95, 0.97, 1.0), 0.05, 0.0, 1.0, 1.5), 1, -1,
                                np.array([0, 0, 0, 0.5, 0, 0], np.float32)))
    s.view = _view((0.0, 1.6, -1.5), 15.0, 70.0)
    return s
