"""Seeded benchmark scenes (the SceneFactory / DataSetFactory side).

A :class:`Scene` holds everything one frame needs: the scalar field in FP64 (the
exact values) and in its storage type, rectilinear coordinates of the unit cube,
the transfer function, one light, the map size, the camera and image size.
The recipes follow the benchmark configurations:

* ``plain_small``   — 64^3 FP32 field of 16 blobs, 256-entry TF from 6 control
  points, one directional light, 64^3-cell map, 512x512, step 0.5 voxel.
* ``precision``     — 256^3 field of 64 blobs + 3-octave value noise (amp 0.2),
  point light outside the box, 256^3-cell map, 1024x1024, step 0.5 voxel.
* ``learned`` / ``time_series`` — 512^3 / 1024^3 blob+noise fields (device-side).

Random draws for light and camera use numpy's PCG64 seeded with the scene seed;
the field and TF come from the library's SplitMix64 generators.
"""

from dataclasses import dataclass, field as dc_field

import numpy as np

from . import beams as B


@dataclass
class Scene:
    name: str
    dims: tuple
    coords: list            # three float32 coordinate arrays
    coords64: list          # the same coordinates in FP64 (i / (n-1))
    field_store: object     # what the renderer receives (float32 / float16 / uint16 bf16)
    field_dtype: int
    field32: np.ndarray     # storage values decoded to FP32 (oracle input of the FP32 twin)
    field64: np.ndarray     # exact field (FP64 oracle input)
    tf32: np.ndarray
    tf64: np.ndarray
    scalar_range: tuple
    light_kind: int
    light_pos: np.ndarray   # float32 position or direction
    light_color: np.ndarray # float32 color * intensity
    map_cells: tuple
    sample_distance: float
    alpha_stop: float
    camera: object
    width: int
    height: int
    map_dtype: int = B.F32
    extras: dict = dc_field(default_factory=dict)

    @property
    def bounds(self):
        b = []
        for c in self.coords:
            b += [float(c[0]), float(c[-1])]
        return b


def unit_coords(n):
    """x_i = float(0) + float(1/(n-1)) * i in FP32 (the ArrayHandleCounting values)."""
    sp = np.float32(1.0 / (n - 1))
    c32 = (np.float32(0.0) + sp * np.arange(n, dtype=np.float32)).astype(np.float32)
    c64 = np.arange(n, dtype=np.float64) / (n - 1)
    return c32, c64


def _decode(store, dtype):
    if dtype == B.F32:
        return np.asarray(store, np.float32)
    if dtype == B.F16:
        return np.asarray(store, np.float16).astype(np.float32)
    if dtype == B.BF16:
        return B.bf16_to_f32(store)
    raise ValueError(dtype)


def _light_outside(rng):
    v = rng.normal(size=3)
    v /= np.linalg.norm(v)
    r = rng.uniform(1.5, 2.5)
    return (np.array([0.5, 0.5, 0.5]) + r * v).astype(np.float32)


def make_scene(name, n, seed, nblobs, noise_octaves=0, noise_amp=0.0, light="point",
               map_cells=None, width=512, height=512, field_dtype=B.F32, map_dtype=B.F32,
               alpha_stop=1.0, time=0.0, drift_speed=0.0, noise_phase_rate=0.0, fov=60.0,
               device=False):
    """device=True generates the field straight into HBM (a devmem.DeviceArray) and keeps
    no host copy (field32/field64 are None)."""
    dims = (n, n, n)
    rng = np.random.default_rng(seed)
    gen = dict(seed=seed, nblobs=nblobs, noise_octaves=noise_octaves, noise_amp=noise_amp,
               noise_freq=4.0, time=time, drift_speed=drift_speed, noise_phase_rate=noise_phase_rate)
    if device:
        from .devmem import DeviceArray

        npdt = {B.F32: np.float32, B.F16: np.float16, B.BF16: np.uint16}[field_dtype]
        store = B.gen_field(dims, dtype=field_dtype, out=DeviceArray(n * n * n, npdt), **gen)
        f64 = None
    else:
        f64 = B.gen_field(dims, dtype=B.F64, **gen)
        store = B.gen_field(dims, dtype=field_dtype, **gen)
    tf32, tf64 = B.gen_colormap(seed + 1, 6, 256)
    c32, c64 = unit_coords(n)
    if light == "directional":
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        kind, pos = B.LIGHT_DIRECTIONAL, v.astype(np.float32)
    else:
        kind, pos = B.LIGHT_POINT, _light_outside(rng)
    az = float(rng.uniform(0.0, 360.0))
    el = float(rng.uniform(-30.0, 30.0))
    bounds = [float(c32[0]), float(c32[-1])] * 3
    cam = B.orbit_camera(bounds, az, el, fov)
    mc = tuple(map_cells) if map_cells is not None else (n - 1, n - 1, n - 1)
    return Scene(name=name, dims=dims, coords=[c32, c32, c32], coords64=[c64, c64, c64],
                 field_store=store, field_dtype=field_dtype,
                 field32=None if device else _decode(store, field_dtype),
                 field64=f64, tf32=tf32, tf64=tf64, scalar_range=(0.0, 1.0), light_kind=kind,
                 light_pos=pos, light_color=np.float32([2.0, 2.0, 2.0]), map_cells=mc,
                 sample_distance=float(np.float32(0.5 / (n - 1))), alpha_stop=alpha_stop,
                 camera=cam, width=width, height=height, map_dtype=map_dtype,
                 extras={"azimuth": az, "elevation": el, "seed": seed})


def plain_small(seed=1, **kw):
    """64^3 FP32, 16 blobs, directional light, 64^3-cell map, 512x512, step 0.5 voxel."""
    kw.setdefault("light", "directional")
    kw.setdefault("map_cells", (64, 64, 64))
    return make_scene("plain_small", 64, seed, 16, width=512, height=512, **kw)


def precision(seed=2, field_dtype=B.F32, map_dtype=B.F32, **kw):
    """256^3 blobs (64) + 3-octave noise at 0.2, point light, 256^3-cell map, 1024x1024."""
    kw.setdefault("map_cells", (256, 256, 256))
    return make_scene("precision", 256, seed, 64, noise_octaves=3, noise_amp=0.2,
                      light="point", width=kw.pop("width", 1024), height=kw.pop("height", 1024),
                      field_dtype=field_dtype, map_dtype=map_dtype, **kw)


def configure(renderer, scene, shadow=True):
    """Scene::Ready-style set-up of a LightedVolumeRenderer (SpheresScene.cxx:181-211)."""
    renderer.SetData(scene.field_store, scene.dims, coords=scene.coords,
                     scalar_range=scene.scalar_range, dtype=scene.field_dtype)
    renderer.SetColorMap(scene.tf32)
    renderer.AddLight(scene.light_pos, scene.light_color, 1.0, kind=scene.light_kind)
    renderer.SetShadow(shadow, scene.map_cells, scene.map_dtype)
    renderer.SetSampleDistance(scene.sample_distance)
    renderer.SetAlphaStop(scene.alpha_stop)
    return renderer
