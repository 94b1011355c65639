"""ctypes binding of the C ABI in include/beams_mi355x.h.

The classes mirror the reference's C++ surface so that tests read like calls into
Beams: :class:`LightedVolumeRenderer` follows
beams/rendering/LightedVolumeRenderer.h:20-101 and :class:`Rays` carries the
fields of vtkm::rendering::raytracing::Ray<Float32> that the renderer reads and
writes.  Every method raises :class:`BeamsError` on a non-zero status.
"""

import ctypes as C
import os

import numpy as np

from . import LIB_PATH

F32, F16, BF16, F64 = 0, 1, 2, 3
MEM_HOST, MEM_DEVICE = 0, 1
ASSOC_POINTS, ASSOC_CELLS = 0, 1
LIGHT_POINT, LIGHT_DIRECTIONAL = 0, 1
SHADOW_REF_MARCH = 0
SHADOW_SWEEP = 1
SHADOW_LEARNED = 2
ORDER_CENTER_DISTANCE, ORDER_KD_FRONT_TO_BACK = 0, 1
MLP_FEATURES = 32
PHASES = ("map", "exchange", "remap", "march", "composite", "frame")

STATUS = {
    0: "BEAMS_OK",
    1: "BEAMS_ERR_INVALID",
    2: "BEAMS_ERR_UNSUPPORTED",
    3: "BEAMS_ERR_STATE",
    4: "BEAMS_ERR_HIP",
    5: "BEAMS_ERR_NO_DEVICE",
}


class BeamsError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


class Camera(C.Structure):
    _fields_ = [
        ("eye", C.c_double * 3),
        ("look_at", C.c_double * 3),
        ("up", C.c_double * 3),
        ("fov_y_deg", C.c_double),
        ("near_dist", C.c_double),
        ("far_dist", C.c_double),
    ]


class FieldRecipe(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("nblobs", C.c_int32),
        ("noise_octaves", C.c_int32),
        ("noise_amp", C.c_double),
        ("noise_freq", C.c_double),
        ("time", C.c_double),
        ("drift_speed", C.c_double),
        ("noise_phase_rate", C.c_double),
    ]


_P = C.c_void_p
_I32, _I64, _F, _D = C.c_int32, C.c_int64, C.c_float, C.c_double
_PF = C.POINTER(C.c_float)
_PD = C.POINTER(C.c_double)
_PI64 = C.POINTER(C.c_int64)

# symbol -> (restype, argtypes); kept in the order of include/beams_mi355x.h
SIGNATURES = {
    "beams_create": (C.c_int, [C.POINTER(_P)]),
    "beams_destroy": (C.c_int, [_P]),
    "beams_last_error": (C.c_char_p, [_P]),
    "beams_synchronize": (C.c_int, [_P]),
    "beams_set_field": (C.c_int, [_P, _P, _I32, _I32, _PI64, C.POINTER(_PF), _PD, _PD, _I32]),
    "beams_set_bricks": (C.c_int, [_P, _I32, _PI64, _I32, _PI64]),
    "beams_get_brick_map": (C.c_int, [_P, _I32, _I32, _PF]),
    "beams_set_persistent_groups": (C.c_int, [_P, _I32]),
    "beams_set_graph": (C.c_int, [_P, _I32]),
    "beams_set_composite_order": (C.c_int, [_P, _I32]),
    "beams_autotune": (C.c_int, [_P, C.POINTER(Camera), _I32, _I32, _I32, C.POINTER(C.c_int32), _PF]),
    "beams_set_scalar_range": (C.c_int, [_P, _D, _D]),
    "beams_set_colormap": (C.c_int, [_P, _PF, _I64]),
    "beams_set_light": (C.c_int, [_P, _I32, _PF, _PF, _F]),
    "beams_set_shadow": (C.c_int, [_P, _I32, _PI64, _I32, _I32]),
    "beams_set_sample_distance": (C.c_int, [_P, _F]),
    "beams_set_alpha_stop": (C.c_int, [_P, _F]),
    "beams_render_rays": (C.c_int, [_P, _I64, C.POINTER(_P), C.POINTER(_P), _P, _P, _P, _P, _I32]),
    "beams_render_image": (C.c_int, [_P, C.POINTER(Camera), _I32, _I32, _PF, _P, _P, _I32]),
    "beams_get_map": (C.c_int, [_P, _P, _I32]),
    "beams_get_timings": (C.c_int, [_P, _PF, _I32]),
    "beams_get_timing_stats": (C.c_int, [_P, _PF, _I32, C.POINTER(C.c_int32)]),
    "beams_reset_timings": (C.c_int, [_P]),
    "beams_set_sample_counting": (C.c_int, [_P, _I32]),
    "beams_get_sample_counts": (C.c_int, [_P, _PI64, _PI64]),
    "beams_gen_field": (C.c_int, [C.POINTER(FieldRecipe), _PI64, _I32, _P, _I32]),
    "beams_gen_colormap": (C.c_int, [C.c_uint64, _I32, _I64, _PF, _PD]),
    "beams_gen_kd_bricks": (C.c_int, [C.c_uint64, _PI64, _I32, _I64, _I64, _PI64]),
    "beams_gen_orbit_camera": (C.c_int, [_PD, _D, _D, _D, C.POINTER(Camera)]),
    "beams_gen_camera_rays": (C.c_int, [C.POINTER(Camera), _I32, _I32, C.POINTER(_PF), C.POINTER(_PF),
                                        C.POINTER(_PD), C.POINTER(_PD)]),
    "beams_mlp_num_params": (C.c_int, []),
    "beams_mlp_create": (C.c_int, [C.POINTER(_P), C.c_uint64]),
    "beams_mlp_destroy": (C.c_int, [_P]),
    "beams_mlp_last_error": (C.c_char_p, [_P]),
    "beams_mlp_get_params": (C.c_int, [_P, _PF]),
    "beams_mlp_set_params": (C.c_int, [_P, _PF]),
    "beams_mlp_forward": (C.c_int, [_P, _P, _I64, _P, _I32]),
    "beams_mlp_gradients": (C.c_int, [_P, _P, _P, _I64, _PF, _PF, _I32]),
    "beams_mlp_train": (C.c_int, [_P, _P, _P, _I64, _I32, _I32, _F, _PF, _I32]),
    "beams_set_learned": (C.c_int, [_P, _P, _I32, _F, _I32]),
    "beams_learned_samples": (C.c_int, [_P, _I32, _I64, C.c_uint64, _P, _P, _I32]),
    "beams_get_learned_stats": (C.c_int, [_P, _PF, C.POINTER(C.c_int32)]),
    "beams_render_rays_noalloc": (C.c_int, [_P, _I64, C.POINTER(_P), C.POINTER(_P), _P, _P, _P, _P]),
    "beams_set_schedule": (C.c_int, [_P, C.c_void_p]),
    "beams_get_schedule": (C.c_int, [_P, C.c_void_p]),
}

_LIB = None


def lib():
    """Load libbeams_mi355x.so; raises ImportError when it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


def _f32p(a):
    return a.ctypes.data_as(_PF)


def _f64p(a):
    return a.ctypes.data_as(_PD)


def _i64(vals):
    return (C.c_int64 * 3)(*[int(v) for v in vals])


def _ptr(x):
    """Raw address of a numpy array or a torch tensor."""
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    return x.ctypes.data


def _mem_of(x):
    return MEM_DEVICE if getattr(x, "is_cuda", False) else MEM_HOST


def _check(status, handle=None):
    if status != 0:
        msg = lib().beams_last_error(handle).decode() if handle else "(no handle)"
        raise BeamsError(status, msg)


def bf16_bits(a):
    """float32 -> bfloat16 bit patterns (round to nearest even), as uint16."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_to_f32(bits):
    return (np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


class Rays:
    """Host SoA of Ray<Float32> (origin, dir, MinDistance, MaxDistance, Distance, Buffers[0])."""

    def __init__(self, origin, direction, min_d=None, max_d=None):
        # copies: Render() updates min/max/dist/rgba in place
        self.origin = np.array(origin, dtype=np.float32).reshape(3, -1)
        self.dir = np.array(direction, dtype=np.float32).reshape(3, -1)
        n = self.origin.shape[1]
        self.min_d = np.zeros(n, np.float32) if min_d is None else np.array(min_d, np.float32)
        self.max_d = np.full(n, np.inf, np.float32) if max_d is None else np.array(max_d, np.float32)
        self.dist = np.zeros(n, np.float32)
        self.rgba = np.zeros((n, 4), np.float32)

    @property
    def n(self):
        return self.origin.shape[1]

    def copy(self):
        r = Rays(self.origin.copy(), self.dir.copy(), self.min_d.copy(), self.max_d.copy())
        r.dist = self.dist.copy()
        r.rgba = self.rgba.copy()
        return r


class LightedVolumeRenderer:
    """beams::rendering::LightedVolumeRenderer over the HIP library (one handle, one stream)."""

    def __init__(self):
        h = C.c_void_p()
        st = lib().beams_create(C.byref(h))
        if st != 0:
            raise BeamsError(st, "beams_create failed (no gfx950 device?)")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().beams_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _c(self, status):
        _check(status, self._h)

    # ---- LightedVolumeRenderer::SetData (LightedVolumeRenderer.cxx:766-779)
    def SetData(self, values, dims, coords=None, origin=None, spacing=None,
                scalar_range=(0.0, 1.0), dtype=F32, assoc=ASSOC_POINTS):
        if coords is not None:
            cs = [np.ascontiguousarray(c, dtype=np.float32) for c in coords]
            self._coords_keep = cs
            carr = (_PF * 3)(*[_f32p(c) for c in cs])
            o = s = None
        else:
            carr = None
            o = (C.c_double * 3)(*origin)
            s = (C.c_double * 3)(*spacing)
        if not hasattr(values, "data_ptr"):
            values = np.ascontiguousarray(values)
            self._values_keep = values
        self._c(lib().beams_set_field(self._h, C.c_void_p(_ptr(values)), dtype, assoc, _i64(dims),
                                      carr, o, s, _mem_of(values)))
        self._c(lib().beams_set_scalar_range(self._h, float(scalar_range[0]), float(scalar_range[1])))

    def SetColorMap(self, rgba):
        t = np.ascontiguousarray(rgba, dtype=np.float32).reshape(-1, 4)
        self._c(lib().beams_set_colormap(self._h, _f32p(t), t.shape[0]))

    def AddLight(self, position, color=(1.0, 1.0, 1.0), intensity=1.0, kind=LIGHT_POINT):
        p = np.ascontiguousarray(position, dtype=np.float32)
        c = np.ascontiguousarray(color, dtype=np.float32)
        self._c(lib().beams_set_light(self._h, kind, _f32p(p), _f32p(c), float(intensity)))

    def SetShadow(self, enabled=True, map_cells=(16, 16, 16), map_dtype=F32, mode=SHADOW_REF_MARCH):
        self._c(lib().beams_set_shadow(self._h, int(bool(enabled)), _i64(map_cells), mode, map_dtype))

    def SetUseShadowMap(self, enabled):
        self._use_shadow = bool(enabled)
        self.SetShadow(enabled, getattr(self, "_map_cells", (16, 16, 16)), getattr(self, "_map_dtype", F32))

    def SetShadowMapSize(self, cells, map_dtype=F32):
        self._map_cells = tuple(int(c) for c in cells)
        self._map_dtype = map_dtype
        self.SetShadow(getattr(self, "_use_shadow", True), self._map_cells, map_dtype)

    def SetSampleDistance(self, d):
        self._c(lib().beams_set_sample_distance(self._h, float(d)))

    def SetAlphaStop(self, a):
        self._c(lib().beams_set_alpha_stop(self._h, float(a)))

    # ---- LightedVolumeRenderer::Render(Ray<Float32>&)
    def Render(self, rays):
        """Renders host Rays in place (map build + CalcRayStart + Sampler)."""
        org = (_P * 3)(*[C.c_void_p(rays.origin[i].ctypes.data) for i in range(3)])
        dr = (_P * 3)(*[C.c_void_p(rays.dir[i].ctypes.data) for i in range(3)])
        self._c(lib().beams_render_rays(self._h, rays.n, org, dr, C.c_void_p(rays.min_d.ctypes.data),
                                        C.c_void_p(rays.max_d.ctypes.data), C.c_void_p(rays.dist.ctypes.data),
                                        C.c_void_p(rays.rgba.ctypes.data), MEM_HOST))
        return rays

    def RenderDevice(self, n, origin, direction, min_d, max_d, dist, rgba):
        """Render with device tensors (torch, on the handle's device) used in place."""
        org = (_P * 3)(*[C.c_void_p(_ptr(t)) for t in origin])
        dr = (_P * 3)(*[C.c_void_p(_ptr(t)) for t in direction])
        self._c(lib().beams_render_rays(self._h, n, org, dr, C.c_void_p(_ptr(min_d)), C.c_void_p(_ptr(max_d)),
                                        C.c_void_p(_ptr(dist)), C.c_void_p(_ptr(rgba)), MEM_DEVICE))

    # ---- MapperLightedVolume::RenderCells
    def RenderImage(self, camera, width, height, background=None, rgba=None, depth=None):
        bg = None if background is None else (C.c_float * 4)(*background)
        if rgba is not None and hasattr(rgba, "data_ptr"):
            self._c(lib().beams_render_image(self._h, C.byref(camera), width, height, bg,
                                             C.c_void_p(_ptr(rgba)),
                                             C.c_void_p(_ptr(depth)) if depth is not None else None,
                                             MEM_DEVICE))
            return rgba, depth
        out = np.zeros((height, width, 4), np.float32)
        dep = np.zeros((height, width), np.float32)
        self._c(lib().beams_render_image(self._h, C.byref(camera), width, height, bg,
                                         C.c_void_p(out.ctypes.data), C.c_void_p(dep.ctypes.data), MEM_HOST))
        return out, dep

    def GetMap(self, map_cells):
        n = int(np.prod([c + 1 for c in map_cells]))
        out = np.zeros(n, np.float32)
        self._c(lib().beams_get_map(self._h, C.c_void_p(out.ctypes.data), MEM_HOST))
        return out

    def GetTimings(self):
        t = np.zeros(len(PHASES), np.float32)
        self._c(lib().beams_get_timings(self._h, _f32p(t), len(PHASES)))
        return dict(zip(PHASES, t.tolist()))

    # ---- BoundsMap -> bricks
    def SetBricks(self, point_ranges, ghost=1, map_cells=None):
        r = np.ascontiguousarray(point_ranges, dtype=np.int64).reshape(-1, 6)
        mc = None if map_cells is None else np.ascontiguousarray(map_cells, dtype=np.int64).reshape(-1, 3)
        self._c(lib().beams_set_bricks(self._h, r.shape[0], r.ctypes.data_as(_PI64), int(ghost),
                                       None if mc is None else mc.ctypes.data_as(_PI64)))

    def ClearBricks(self):
        self._c(lib().beams_set_bricks(self._h, 0, None, 0, None))

    def GetBrickMap(self, brick, nverts, final=True):
        out = np.zeros(nverts, np.float32)
        self._c(lib().beams_get_brick_map(self._h, int(brick), 1 if final else 0, _f32p(out)))
        return out

    def SetPersistentGroups(self, groups):
        self._c(lib().beams_set_persistent_groups(self._h, int(groups)))

    SCHEDULE_KNOBS = ("persistent_groups", "xcd_ranges", "light_rows", "march_xcd_tiles",
                      "map_fast_fetch", "march_fast_fetch", "map_pipeline", "sweep_block",
                      "march_groups_per_cu", "march_blocked")

    def SetSchedule(self, **knobs):
        """Brick-launch schedule knobs (SCHEDULE_KNOBS; results are the same under every
        schedule); a knob not given, or None, keeps its current value."""
        unknown = set(knobs) - set(self.SCHEDULE_KNOBS)
        if unknown:
            raise ValueError(f"unknown schedule knobs {sorted(unknown)}")
        cur = self.GetSchedule()
        arr = (C.c_int32 * len(self.SCHEDULE_KNOBS))(
            *[cur[k] if knobs.get(k) is None else int(knobs[k]) for k in self.SCHEDULE_KNOBS])
        self._c(lib().beams_set_schedule(self._h, C.cast(arr, C.c_void_p)))

    def GetSchedule(self):
        arr = (C.c_int32 * len(self.SCHEDULE_KNOBS))()
        self._c(lib().beams_get_schedule(self._h, C.cast(arr, C.c_void_p)))
        return dict(zip(self.SCHEDULE_KNOBS, list(arr)))

    def SetGraph(self, enabled):
        self._c(lib().beams_set_graph(self._h, int(bool(enabled))))

    def SetCompositeOrder(self, order):
        """ORDER_CENTER_DISTANCE (the reference's) or ORDER_KD_FRONT_TO_BACK (an extension)."""
        self._c(lib().beams_set_composite_order(self._h, int(order)))

    def Autotune(self, camera, width, height, frames=3):
        g = C.c_int32()
        ms = C.c_float()
        self._c(lib().beams_autotune(self._h, C.byref(camera), width, height, frames, C.byref(g),
                                     C.byref(ms)))
        return g.value, ms.value

    def GetTimingStats(self):
        t = np.zeros(len(PHASES), np.float32)
        n = C.c_int32()
        self._c(lib().beams_get_timing_stats(self._h, _f32p(t), len(PHASES), C.byref(n)))
        return dict(zip(PHASES, t.tolist())), n.value

    def ResetTimings(self):
        self._c(lib().beams_reset_timings(self._h))

    def SetSampleCounting(self, enabled):
        self._c(lib().beams_set_sample_counting(self._h, int(bool(enabled))))

    def GetSampleCounts(self):
        a = C.c_int64()
        b = C.c_int64()
        self._c(lib().beams_get_sample_counts(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def Synchronize(self):
        self._c(lib().beams_synchronize(self._h))

    # ---- learned shadow closure
    def SetLearned(self, net=None, fallback=SHADOW_REF_MARCH, error_budget=2e-3, check_stride=8):
        """Network for SHADOW_LEARNED (None: the coarse optical depth alone)."""
        self._net = net  # the renderer keeps the pointer
        self._c(lib().beams_set_learned(self._h, net._h if net is not None else None, fallback,
                                        float(error_budget), int(check_stride)))

    def LearnedSamples(self, n, seed, mode=SHADOW_SWEEP, features=None, targets=None):
        """Features (n, 32) and targets (n,) from the exact map of `mode`; device outputs when
        features/targets are device arrays."""
        if features is not None:
            self._c(lib().beams_learned_samples(self._h, mode, int(n), int(seed), C.c_void_p(_ptr(features)),
                                                C.c_void_p(_ptr(targets)), _mem_of(features)))
            return features, targets
        x = np.zeros((n, MLP_FEATURES), np.float32)
        t = np.zeros(n, np.float32)
        self._c(lib().beams_learned_samples(self._h, mode, int(n), int(seed), C.c_void_p(x.ctypes.data),
                                            C.c_void_p(t.ctypes.data), MEM_HOST))
        return x, t

    def GetLearnedStats(self):
        e = np.zeros(2, np.float32)
        fb = C.c_int32()
        self._c(lib().beams_get_learned_stats(self._h, _f32p(e), C.byref(fb)))
        return float(e[0]), float(e[1]), int(fb.value)


class ShadowMLP:
    """The learned closure's network (beams_mlp): parameters, forward, gradients, training."""

    def __init__(self, seed=1):
        h = C.c_void_p()
        st = lib().beams_mlp_create(C.byref(h), int(seed))
        if st != 0:
            raise BeamsError(st, "beams_mlp_create failed (no gfx950 device?)")
        self._h = h
        self.nparams = lib().beams_mlp_num_params()

    def close(self):
        if getattr(self, "_h", None):
            lib().beams_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _c(self, status):
        if status != 0:
            raise BeamsError(status, lib().beams_mlp_last_error(self._h).decode())

    def get_params(self):
        p = np.zeros(self.nparams, np.float32)
        self._c(lib().beams_mlp_get_params(self._h, _f32p(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        self._c(lib().beams_mlp_set_params(self._h, _f32p(p)))

    def forward(self, x):
        if hasattr(x, "data_ptr"):
            raise TypeError("forward takes host arrays; use the device entry point directly")
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_FEATURES)
        y = np.zeros(x.shape[0], np.float32)
        self._c(lib().beams_mlp_forward(self._h, C.c_void_p(x.ctypes.data), x.shape[0],
                                        C.c_void_p(y.ctypes.data), MEM_HOST))
        return y

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_FEATURES)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1)
        g = np.zeros(self.nparams, np.float32)
        loss = C.c_float()
        self._c(lib().beams_mlp_gradients(self._h, C.c_void_p(x.ctypes.data), C.c_void_p(t.ctypes.data),
                                          x.shape[0], _f32p(g), C.byref(loss), MEM_HOST))
        return g, loss.value

    def train(self, x, t, epochs=1, batch=4096, lr=1e-3):
        """Adam over consecutive minibatches; x/t host arrays or device arrays."""
        loss = C.c_float()
        if hasattr(x, "data_ptr"):
            n = int(t.shape[0])
            self._c(lib().beams_mlp_train(self._h, C.c_void_p(_ptr(x)), C.c_void_p(_ptr(t)), n, int(epochs),
                                          int(batch), float(lr), C.byref(loss), MEM_DEVICE))
            return loss.value
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_FEATURES)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1)
        self._c(lib().beams_mlp_train(self._h, C.c_void_p(x.ctypes.data), C.c_void_p(t.ctypes.data),
                                      x.shape[0], int(epochs), int(batch), float(lr), C.byref(loss), MEM_HOST))
        return loss.value


# ---- scene inputs (DataSetFactory / SceneFactory side) -----------------------------------------

def gen_field(dims, seed=1, nblobs=16, noise_octaves=0, noise_amp=0.0, noise_freq=4.0,
              time=0.0, drift_speed=0.0, noise_phase_rate=0.0, dtype=F32, out=None):
    """Blob + value-noise field normalised to [0,1].  out may be a device tensor."""
    r = FieldRecipe(seed, nblobs, noise_octaves, noise_amp, noise_freq, time, drift_speed,
                    noise_phase_rate)
    if out is None:
        npdt = {F32: np.float32, F16: np.float16, BF16: np.uint16, F64: np.float64}[dtype]
        out = np.zeros(int(np.prod(dims)), npdt)
    st = lib().beams_gen_field(C.byref(r), _i64(dims), dtype, C.c_void_p(_ptr(out)), _mem_of(out))
    _check(st)
    return out


def gen_colormap(seed, npoints=6, n=256):
    t32 = np.zeros((n, 4), np.float32)
    t64 = np.zeros((n, 4), np.float64)
    _check(lib().beams_gen_colormap(seed, npoints, n, _f32p(t32), _f64p(t64)))
    return t32, t64


def kd_bricks(dims, nbricks, seed, min_cells=32, max_cells=160):
    """Seeded kd bricks: (nbricks, 6) inclusive point ranges (i0,i1,j0,j1,k0,k1)."""
    out = np.zeros((nbricks, 6), np.int64)
    _check(lib().beams_gen_kd_bricks(seed, _i64(dims), nbricks, min_cells, max_cells,
                                     out.ctypes.data_as(_PI64)))
    return out


def orbit_camera(bounds, azimuth, elevation, fov=60.0):
    cam = Camera()
    b = (C.c_double * 6)(*bounds)
    _check(lib().beams_gen_orbit_camera(b, float(azimuth), float(elevation), float(fov), C.byref(cam)))
    return cam


def camera_rays(cam, width, height, want64=False):
    n = width * height
    o = np.zeros((3, n), np.float32)
    d = np.zeros((3, n), np.float32)
    oa = (_PF * 3)(*[_f32p(o[i]) for i in range(3)])
    da = (_PF * 3)(*[_f32p(d[i]) for i in range(3)])
    if want64:
        o64 = np.zeros((3, n), np.float64)
        d64 = np.zeros((3, n), np.float64)
        oa64 = (_PD * 3)(*[_f64p(o64[i]) for i in range(3)])
        da64 = (_PD * 3)(*[_f64p(d64[i]) for i in range(3)])
    else:
        o64 = d64 = oa64 = da64 = None
    _check(lib().beams_gen_camera_rays(C.byref(cam), width, height, oa, da, oa64, da64))
    return (o, d, o64, d64) if want64 else (o, d)
