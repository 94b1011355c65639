#!/usr/bin/env python3
"""Lit-frame benchmark of the MI355X Beams renderer on one GPU.

A step is one lit frame of the time-series workload with exact shadows: a 1024^3 FP16
blob+noise field (one of 8 drifting timesteps, resident in HBM) cut into 128 uneven kd bricks
that stand in for the reference's 128 MPI ranks; each frame rebuilds every brick's
transmittance map (local light march, non-local exchange, seeded re-march), marches every
brick's camera rays at 2048x2048 and composites them in depth order, as one hipGraph launch.

Prints one JSON line: ray samples/s (composited camera-ray samples, counted on the GPU) and
frame ms, per-phase kernel times from HIP events on the renderer's stream, the dominant
kernel's roofline entry, and the CPU oracle timed (Serial and OpenMP) on a bounded sample of
the same frame, plus the other workloads (precision-sweep frame, batched bricks, learned
closure, learned time series, RGBA error against the FP64 oracle).
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "ray samples/s & lit frame ms, exact vs learned shadows, at max RGBA err ≤2e-3"


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=256, help="points per axis of the field")
    ap.add_argument("--image", type=int, default=1024, help="image width = height")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "0") or 0))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-variants", action="store_true", help="skip the FP16/BF16 storage variants")
    ap.add_argument("--no-bricks", action="store_true", help="skip the batched-bricks workload")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-closure workloads")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the RGBA error report against the FP64 oracle")
    ap.add_argument("--learned-size", type=int, default=512)
    ap.add_argument("--learned-samples", type=int, default=4 << 20)
    ap.add_argument("--learned-epochs", type=int, default=12)
    ap.add_argument("--learned-eval-seeds", type=int, default=4, help="held-out fields evaluated (1..4)")
    ap.add_argument("--series-timesteps", type=int, default=8, help="timesteps of the time series (1..8)")
    ap.add_argument("--series-size", type=int, default=1024, help="points per axis of the time series")
    ap.add_argument("--series-image", type=int, default=2048, help="time-series image width = height")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed runs of each CPU mode (median)")
    ap.add_argument("--schedule", default=None,
                    help="time-series brick schedule as JSON (the knobs of SetSchedule): no trials or "
                         "autotuning, so that separate profiling passes run the same kernels")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device-to-device 5 times (the byte-counter calibration run)")
    return ap.parse_args()


def frame_loop(r, sc, rgba, depth, steps, warmup):
    from pilot_amd import beams as B

    for _ in range(warmup):
        r.RenderImage(sc.camera, sc.width, sc.height, rgba=rgba, depth=depth)
    r.SetSampleCounting(True)
    r.RenderImage(sc.camera, sc.width, sc.height, rgba=rgba, depth=depth)
    cam_samples, light_steps = r.GetSampleCounts()
    r.SetSampleCounting(False)
    r.Synchronize()
    r.ResetTimings()
    t0 = time.perf_counter()
    for _ in range(steps):
        r.RenderImage(sc.camera, sc.width, sc.height, rgba=rgba, depth=depth)
    r.Synchronize()
    t1 = time.perf_counter()
    stats, nframes = r.GetTimingStats()
    assert nframes == steps, (nframes, steps)
    del B
    return (t1 - t0) / steps, stats, cam_samples, light_steps


def graph_frame_s(r, sc, rgba, depth, steps):
    """Seconds per frame replayed from the frame's hipGraph (the whole frame: map build,
    with the learned closure's check and device-side fallback, march and composite)."""
    r.SetGraph(True)
    r.RenderImage(sc.camera, sc.width, sc.height, rgba=rgba, depth=depth)  # capture
    r.Synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r.RenderImage(sc.camera, sc.width, sc.height, rgba=rgba, depth=depth)
    r.Synchronize()
    sec = (time.perf_counter() - t0) / steps
    r.SetGraph(False)
    return sec


def bricks_workload(steps, warmup, n=512, nbricks=96, img=1024):
    """Batched bricks: n^3 FP16 blob+noise field cut into nbricks seeded kd bricks (cell
    extents 32..160, one ghost layer), FP16 sub-maps (one cell per voxel), 1024^2, autotuned
    persistent schedule; timed eagerly (per-phase events) and as hipGraph replays."""
    import numpy as np

    from pilot_amd import beams as B
    from pilot_amd import scenes
    from pilot_amd.devmem import DeviceArray

    sc = scenes.make_scene("bricks", n, seed=3, nblobs=64, noise_octaves=3, noise_amp=0.2,
                           light="point", width=img, height=img, field_dtype=B.F16,
                           map_dtype=B.F16, device=True)
    ranges = B.kd_bricks(sc.dims, nbricks, 3, 32, 160)
    rgba = DeviceArray((img, img, 4), np.float32)
    depth = DeviceArray((img, img), np.float32)
    with B.LightedVolumeRenderer() as r:
        scenes.configure(r, sc)
        r.SetBricks(ranges, ghost=1)
        groups, tuned_ms = r.Autotune(sc.camera, img, img, frames=2)
        sec, stats, cam_s, light_s = frame_loop(r, sc, rgba, depth, steps, warmup)
        r.SetGraph(True)
        r.RenderImage(sc.camera, img, img, rgba=rgba, depth=depth)  # capture
        r.Synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            r.RenderImage(sc.camera, img, img, rgba=rgba, depth=depth)
        r.Synchronize()
        gsec = (time.perf_counter() - t0) / steps
    ext = (ranges[:, 1::2] - ranges[:, 0::2])
    return {
        "workload": (f"batched bricks: {n}^3 FP16 field in {nbricks} kd bricks (cell extents "
                     f"{int(ext.min())}..{int(ext.max())}, 1 ghost layer), FP16 sub-maps, {img}x{img}, "
                     "exact shadows with the non-local exchange, depth-ordered composite"),
        "frame_ms_eager": sec * 1e3,
        "frame_ms_graph": gsec * 1e3,
        "phase_ms": stats,
        "samples_per_frame": cam_s,
        "light_steps_per_frame": light_s,
        "ray_samples_per_s": cam_s / gsec,
        "autotuned_persistent_groups": groups,
        "autotune_ms": tuned_ms,
    }


class DevView:
    """A device address (an allocation's base plus a byte offset) for device-memory calls."""

    is_cuda = True

    def __init__(self, addr):
        self.addr = int(addr)

    def data_ptr(self):
        return self.addr


def image_error(a, b):
    d = abs(a.astype("float64") - b.astype("float64"))
    return {"max": float(d.max()), "mean": float(d.mean())}


def learned_scene(seed, n, img, **kw):
    from pilot_amd import beams as B
    from pilot_amd import scenes

    kw.setdefault("map_cells", (n, n, n))
    kw.setdefault("device", True)
    return scenes.make_scene("learned", n, seed=seed, nblobs=64, noise_octaves=3, noise_amp=0.2,
                             light="point", width=img, height=img, field_dtype=B.F16,
                             map_dtype=B.F16, **kw)


def train_closure(n, seeds, per_field, epochs, batch=4096, lr=1e-3):
    """The network trained on per_field map vertices of each seeded field's exact map (the
    reference march, the map the check and the FP64 oracle follow; features and targets drawn
    on the GPU, shuffled once on the host, Adam on the GPU)."""
    import numpy as np

    from pilot_amd import beams as B
    from pilot_amd import scenes
    from pilot_amd.devmem import DeviceArray

    total = per_field * len(seeds)
    x = DeviceArray((total, B.MLP_FEATURES), np.float32)
    t = DeviceArray((total,), np.float32)
    t0 = time.perf_counter()
    for i, s in enumerate(seeds):
        sc = learned_scene(s, n, 64)
        with B.LightedVolumeRenderer() as r:
            scenes.configure(r, sc)
            r.LearnedSamples(per_field, seed=s, mode=B.SHADOW_REF_MARCH,
                             features=DevView(x.data_ptr() + i * per_field * B.MLP_FEATURES * 4),
                             targets=DevView(t.data_ptr() + i * per_field * 4))
        sc.field_store.free()
    t_samples = time.perf_counter() - t0
    perm = np.random.default_rng(0).permutation(total)
    xs = DeviceArray.from_host(x.numpy()[perm])
    ts = DeviceArray.from_host(t.numpy()[perm])
    x.free()
    t.free()
    net = B.ShadowMLP(seed=1)
    t0 = time.perf_counter()
    losses = []
    for stage_lr in (lr, lr, lr * 0.3):
        losses.append(net.train(xs, ts, epochs=max(1, epochs // 3), batch=batch, lr=stage_lr))
    t_train = time.perf_counter() - t0
    xs.free()
    ts.free()
    steps = 3 * max(1, epochs // 3) * ((total + batch - 1) // batch)
    return net, {"samples": total, "fields": len(seeds), "sample_s": t_samples, "train_s": t_train,
                 "optimizer_steps": steps, "batch": batch, "loss_per_stage": losses}


def learned_workload(steps, warmup, net, n=512, img=2048, seeds=(2001, 2002, 2003, 2004),
                     budget=None):
    """Held-out fields: the exact paths (reference per-vertex march and the exact sweep) and
    the learned closure (with and without its fallback) on the same frame."""
    import numpy as np

    budget = LEARNED_BUDGET if budget is None else budget

    from pilot_amd import beams as B
    from pilot_amd import scenes
    from pilot_amd.devmem import DeviceArray

    rgba = DeviceArray((img, img, 4), np.float32)
    depth = DeviceArray((img, img), np.float32)
    out = []
    for s in seeds:
        sc = learned_scene(s, n, img)
        row = {"seed": s}
        images = {}
        for label, mode, budget_ in (("exact_ref_march", B.SHADOW_REF_MARCH, None),
                                     ("exact_sweep", B.SHADOW_SWEEP, None),
                                     ("learned_no_fallback", B.SHADOW_LEARNED, 1e30),
                                     ("learned", B.SHADOW_LEARNED, budget)):
            with B.LightedVolumeRenderer() as r:
                scenes.configure(r, sc)
                r.SetShadow(True, sc.map_cells, B.F16, mode=mode)
                if mode == B.SHADOW_LEARNED:
                    r.SetLearned(net, fallback=B.SHADOW_REF_MARCH, error_budget=budget_, check_stride=8)
                sec, stats, cam_s, _ = frame_loop(r, sc, rgba, depth, steps, warmup)
                gsec = graph_frame_s(r, sc, rgba, depth, steps)
                images[label] = rgba.numpy()
                row[label] = {"frame_ms": gsec * 1e3, "frame_ms_eager": sec * 1e3, "map_ms": stats["map"],
                              "march_ms": stats["march"], "samples_per_frame": cam_s,
                              "ray_samples_per_s": cam_s / gsec}
                if mode == B.SHADOW_LEARNED:
                    e_l, e_c, fb = r.GetLearnedStats()
                    row[label].update({"check_max_opacity_err": e_l, "coarse_only_check_err": e_c,
                                       "fell_back": fb})
            print(f"[bench] learned seed {s} {label}: {sec * 1e3:.2f} ms", file=sys.stderr, flush=True)
        for label in ("exact_sweep", "learned_no_fallback", "learned"):
            row[label]["rgba_err_vs_exact_ref_march"] = image_error(images[label], images["exact_ref_march"])
        out.append(row)
        sc.field_store.free()
    rgba.free()
    depth.free()

    def mean(key, sub):
        return float(np.mean([r[key][sub] for r in out]))

    return {
        "workload": (f"learned shadow closure: {n}^3 FP16 blob+noise fields, {n}^3-cell FP16 maps, "
                     f"{img}x{img}, point light, step 0.5 voxel; held-out seeds {list(seeds)}; "
                     f"trained on 8 fields' reference-march maps; the check tiles (every 8th vertex per "
                     f"axis) whose own or neighbouring corners exceed a {budget:g} opacity budget "
                     f"re-marched by the reference march; one hipGraph per frame"),
        "frame_ms": {k: mean(k, "frame_ms") for k in ("exact_ref_march", "exact_sweep",
                                                      "learned_no_fallback", "learned")},
        "speedup_learned_no_fallback_vs_ref_march": mean("exact_ref_march", "frame_ms") /
        mean("learned_no_fallback", "frame_ms"),
        "speedup_learned_vs_ref_march": mean("exact_ref_march", "frame_ms") / mean("learned", "frame_ms"),
        "max_rgba_err_learned_vs_ref_march": max(r["learned"]["rgba_err_vs_exact_ref_march"]["max"] for r in out),
        "speedup_learned_no_fallback_vs_sweep": mean("exact_sweep", "frame_ms") /
        mean("learned_no_fallback", "frame_ms"),
        "per_seed": out,
    }


# ---- the headline: the time series (1024^3 FP16, 8 timesteps, 128 kd bricks, 2048^2) ------------

SERIES = dict(n=1024, img=2048, nbricks=128, timesteps=8, seed=9, brick_seed=5, min_cells=96,
              max_cells=320)


def series_scene(ts, n, img, device=True):
    """Timestep ts of the drifting field: 64 blobs moving 0.01 box/step + 3-octave value noise
    (amplitude 0.2) whose phase advances 0.25/step, generated in HBM as FP16 (device) or on the
    host; FP16 maps."""
    from pilot_amd import beams as B
    from pilot_amd import scenes

    return scenes.make_scene("series", n, seed=SERIES["seed"], nblobs=64, noise_octaves=3,
                             noise_amp=0.2, light="point", width=img, height=img,
                             field_dtype=B.F16, map_dtype=B.F16, device=device, time=float(ts),
                             drift_speed=0.01, noise_phase_rate=0.25)


def series_bricks(n):
    from pilot_amd import beams as B

    ranges = B.kd_bricks((n, n, n), SERIES["nbricks"], SERIES["brick_seed"],
                         SERIES["min_cells"] * n // 1024, SERIES["max_cells"] * n // 1024)
    return ranges, ranges[:, 1::2] - ranges[:, 0::2]  # a map cell per field cell


def series_exact(steps, warmup, n, img, timesteps, keep_first=True, schedule=None):
    """The exact-shadow frames of the time series.  One renderer per timestep holds that
    timestep's field in its bricks (inputs resident in HBM before timing); every frame rebuilds
    the maps (phase-1 local maps, the non-local exchange, the seeded re-march), marches every
    brick and composites, as one hipGraph launch.  Untimed: per-timestep sample counts (one
    counting frame) and per-phase HIP-event times (eager frames).  Timed: `steps` graph
    launches cycling through the timesteps, bracketed by stream synchronisation."""
    import numpy as np

    from pilot_amd import beams as B
    from pilot_amd import scenes
    from pilot_amd.devmem import DeviceArray, synchronize

    ranges, cells = series_bricks(n)
    rgba = DeviceArray((img, img, 4), np.float32)
    depth = DeviceArray((img, img), np.float32)
    rs, samples, light_steps, phase_rows, first = [], [], [], [], None
    sched_ms = {}
    for ts in range(timesteps):
        sc = series_scene(ts, n, img)
        r = B.LightedVolumeRenderer()
        scenes.configure(r, sc)
        r.SetBricks(ranges, ghost=1, map_cells=cells)
        if schedule is None:
            # the schedule of the brick launches: the frame with every switch off and with every
            # switch on, then the autotuner's pick (coordinate descent), used for every timestep
            knobs = [k for k in B.LightedVolumeRenderer.SCHEDULE_KNOBS
                     if k not in ("persistent_groups", "sweep_block", "march_groups_per_cu")]
            trials = [("all switches off", {k: 0 for k in knobs}), ("all switches on", {k: 1 for k in knobs})]
            trials += [(f"all on but {k}", {q: int(q != k) for q in knobs}) for k in knobs]
            sched_of = {}
            for label, sw in trials:
                r.SetSchedule(persistent_groups=2048, **sw)
                sched_of[label] = r.GetSchedule()
                r.RenderImage(sc.camera, img, img, rgba=rgba, depth=depth)
                synchronize()
                t0 = time.perf_counter()
                for _ in range(2):
                    r.RenderImage(sc.camera, img, img, rgba=rgba, depth=depth)
                synchronize()
                sched_ms[label] = (time.perf_counter() - t0) / 2 * 1e3
            # the autotuner's descent starts from the fastest trial; the fastest of all is kept
            best_trial = min(sched_of, key=lambda k: sched_ms[k])
            r.SetSchedule(**sched_of[best_trial])
            _, sched_ms["autotuned"] = r.Autotune(sc.camera, img, img, frames=2)
            sched_of["autotuned"] = r.GetSchedule()
            schedule = sched_of[min(sched_of, key=lambda k: sched_ms[k])]
            print(f"[bench] series schedules: {sched_ms}, picked {schedule}", file=sys.stderr, flush=True)
        r.SetSchedule(**schedule)
        # eager frames with per-phase events (the first one also warms up)
        r.RenderImage(sc.camera, img, img, rgba=rgba, depth=depth)
        r.ResetTimings()
        for _ in range(2):
            r.RenderImage(sc.camera, img, img, rgba=rgba, depth=depth)
        stats, _ = r.GetTimingStats()
        phase_rows.append(stats)
        r.SetSampleCounting(True)
        r.RenderImage(sc.camera, img, img, rgba=rgba, depth=depth)
        cs, ls = r.GetSampleCounts()
        r.SetSampleCounting(False)
        samples.append(cs)
        light_steps.append(ls)
        r.SetGraph(True)
        r.RenderImage(sc.camera, img, img, rgba=rgba, depth=depth)  # capture
        if ts == 0 and keep_first:
            first = sc  # its device field stays for the CPU sample
        else:
            sc.field_store.free()
        rs.append((r, sc.camera))
        print(f"[bench] series t={ts}: eager {stats['frame']:.1f} ms (map {stats['map']:.1f}, exchange "
              f"{stats['exchange']:.1f}, remap {stats['remap']:.1f}, march {stats['march']:.1f}), "
              f"{cs} samples", file=sys.stderr, flush=True)
    for k in range(warmup):
        r, cam = rs[k % timesteps]
        r.RenderImage(cam, img, img, rgba=rgba, depth=depth)
    synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        r, cam = rs[k % timesteps]
        r.RenderImage(cam, img, img, rgba=rgba, depth=depth)
    synchronize()
    sec = (time.perf_counter() - t0) / steps
    timed_samples = sum(samples[k % timesteps] for k in range(steps))
    images = []
    for ts, (r, cam) in enumerate(rs):
        r.SetGraph(False)
        im, _ = r.RenderImage(cam, img, img)
        images.append(im.reshape(-1, 4))
    phase = {k: float(np.mean([p[k] for p in phase_rows])) for k in phase_rows[0]}
    out = dict(sec=sec, samples=samples, light_steps=light_steps, timed_samples=timed_samples,
               phase=phase, phase_per_timestep=phase_rows, ranges=ranges, cells=cells,
               images=images, first_scene=first, renderer0=rs[0][0], schedule=schedule,
               schedule_ms=sched_ms)
    for r, _ in rs[1:]:
        r.close()
    rgba.free()
    depth.free()
    return out


def series_map_bytes(ranges, cells, ghost=1, n=1024):
    """Algorithmic bytes of one k_bricks_map launch: every brick's ghosted FP16 field read
    once, its FP16 map written once, the 256-entry TF read."""
    import numpy as np

    lo = np.maximum(ranges[:, 0::2] - ghost, 0)
    hi = np.minimum(ranges[:, 1::2] + ghost, n - 1)
    field = int(np.prod(hi - lo + 1, axis=1).sum()) * 2
    maps = int(np.prod(cells + 1, axis=1).sum()) * 2
    return field + maps + 256 * 16, field, maps


def series_cpu_baseline(ex, n, img, threads, runs, per_brick=4, tile=256):
    """The oracle (FP32 restatement of the reference) on a bounded sample of one timestep of
    the time series, Serial (1 thread) and OpenMP (`threads`), median of `runs` after one
    warm-up run each:
      map:      `per_brick` ranges of 256 map vertices in every brick (phase-1 local march);
      exchange: the same vertices' non-local seeds (hits, fetch from the other bricks' phase-1
                maps, composite), from the GPU's phase-1 maps;
      remap:    the same vertices' seeded re-march;
      march:    every brick's CalcRayStart + Sampler over a tile x tile block of camera rays at
                the image centre against the GPU's final maps, then the depth-ordered composite.
    Every sampled value is compared with the GPU's (local and final map vertices, the tile's
    RGBA): a full-size parity check.  The frame time is extrapolated: each map phase by the
    frame's map vertices over the sampled ones, the march by the frame's camera samples over
    the tile's."""
    import dataclasses

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from pilot_amd import beams as B

    sc = ex["first_scene"]
    r0 = ex["renderer0"]
    ranges, cells = ex["ranges"], ex["cells"]
    f16 = sc.field_store.numpy()
    sc.field_store.free()
    field32 = f16.astype(np.float32)
    del f16
    scf = dataclasses.replace(sc, field32=field32)
    vols = O.brick_volumes(scf, ranges, twin="f32", map_cells=cells)
    del field32, scf
    nb = len(vols)
    nverts = [v.nverts for v in vols]
    local = [r0.GetBrickMap(b, nverts[b], final=False) for b in range(nb)]
    final = [r0.GetBrickMap(b, nverts[b], final=True) for b in range(nb)]
    sel = []
    for b in range(nb):
        for q in range(per_brick):
            lo = (nverts[b] - 256) * (2 * q + 1) // (2 * per_brick)
            sel.append((b, lo, lo + 256))
    sampled = 256 * len(sel)
    o, d = B.camera_rays(sc.camera, img, img)
    y0 = x0 = (img - tile) // 2
    pix = ((np.arange(tile)[:, None] + y0) * img + (np.arange(tile)[None, :] + x0)).ravel()
    to, td = np.ascontiguousarray(o[:, pix]), np.ascontiguousarray(d[:, pix])
    bounds = []
    for v in vols:
        bounds += [float(v.coords[0][0]), float(v.coords[0][-1]), float(v.coords[1][0]),
                   float(v.coords[1][-1]), float(v.coords[2][0]), float(v.coords[2][-1])]
    bounds = np.array(bounds)
    gpu_tile = ex["images"][0][pix]
    r16 = lambda m: m.astype(np.float16).astype(np.float32)  # noqa: E731
    scratch = [np.zeros(nv, np.float32) for nv in nverts]
    seeds = [np.zeros(nv, np.float32) for nv in nverts]

    def one_run(nthr):
        t = {}
        ok = {"local": True, "final": True, "tile": True}
        t0 = time.perf_counter()
        for b, lo, hi in sel:
            O.tmap_build_range(vols[b], scratch[b], lo, hi, nthr)
        t["map"] = time.perf_counter() - t0
        for b, lo, hi in sel:
            ok["local"] &= bool(np.array_equal(r16(scratch[b][lo:hi]), local[b][lo:hi]))
        t0 = time.perf_counter()
        for b, lo, hi in sel:
            O.nonlocal_seed_range(vols, local, b, lo, hi, seeds[b], nthr)
        t["exchange"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        for b, lo, hi in sel:
            O.tmap_build_range(vols[b], scratch[b], lo, hi, nthr, seed=seeds[b])
        t["remap"] = time.perf_counter() - t0
        for b, lo, hi in sel:
            ok["final"] &= bool(np.array_equal(r16(scratch[b][lo:hi]), final[b][lo:hi]))
        t0 = time.perf_counter()
        partial, nsamp = [], 0
        for b in range(nb):
            rb = O.RayBuf(to, td, "f32")
            O.sample_rays(vols[b], final[b], rb, nthr)
            partial.append(rb.rgba)
            nsamp += int(rb.nsamples.sum())
        timg = O.composite(bounds, np.asarray(sc.camera.eye, np.float64), partial, "f32")
        t["march"] = time.perf_counter() - t0
        ok["tile"] = bool(np.array_equal(timg, gpu_tile))
        return t, ok, nsamp

    frame_samples = ex["samples"][0]
    total_verts = int(sum(nverts))
    modes = {}
    for label, nthr in (("serial", 1), ("openmp", threads)):
        one_run(nthr)  # warm-up
        rows = [one_run(nthr) for _ in range(runs)]
        med = {k: float(np.median([r[0][k] for r in rows])) for k in rows[0][0]}
        tile_samples = rows[0][2]
        frame_s = (med["map"] + med["exchange"] + med["remap"]) * total_verts / sampled + \
            med["march"] * frame_samples / max(tile_samples, 1)
        modes[label] = {"threads": nthr, "median_s": med, "tile_samples": tile_samples,
                        "frame_s_extrapolated": frame_s, "ray_samples_per_s": frame_samples / frame_s,
                        "equal_gpu": {k: all(r[1][k] for r in rows) for k in rows[0][1]}}
        print(f"[bench] cpu {label}: {med} -> frame {frame_s:.1f} s (extrapolated)", file=sys.stderr,
              flush=True)
    omp = modes["openmp"]
    return {
        "value": omp["ray_samples_per_s"],
        "unit": "ray samples/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"FP32 oracle on timestep 0, OpenMP {threads} threads, median of {runs}: "
                   f"{per_brick} ranges of 256 map vertices in each of the {nb} bricks ({sampled} of "
                   f"{total_verts} vertices: local march, exchange, seeded re-march) and the "
                   f"{tile}x{tile} centre tile ({omp['tile_samples']} of {frame_samples} camera samples); "
                   f"frame time extrapolated per vertex and per sample to "
                   f"{omp['frame_s_extrapolated']:.1f} s"),
        "modes": modes,
        "equal_gpu": {k: modes["serial"]["equal_gpu"][k] and omp["equal_gpu"][k] for k in omp["equal_gpu"]},
    }


# Opacity budget of the learned closure's check: a quarter of the 2e-3 RGBA bar, since a tile's
# interior can differ more than its (and its neighbours') check vertices.
LEARNED_BUDGET = 5e-4


def series_learned(steps, net, ex, n, img, timesteps):
    """The learned-shadow frames of the same timesteps, one captured graph per frame (map,
    per-brick check, device-side fallback list, fallback, march, composite).  The learned maps
    approximate the single-block map at the bricks' vertices; their exact counterpart is the
    same frame with every brick on the fallback (a negative budget), and the RGBA error is
    stated against it, and against the multi-rank exact frame (per-brick light step and
    exchange) for information."""
    import numpy as np

    from pilot_amd import beams as B
    from pilot_amd import scenes
    from pilot_amd.devmem import DeviceArray

    ranges, cells = ex["ranges"], ex["cells"]
    rgba = DeviceArray((img, img, 4), np.float32)
    depth = DeviceArray((img, img), np.float32)
    rows = []
    for ts in range(timesteps):
        sc = series_scene(ts, n, img)
        row = {"timestep": ts}
        images = {}
        with B.LightedVolumeRenderer() as r:
            scenes.configure(r, sc)
            r.SetShadow(True, sc.map_cells, B.F16, mode=B.SHADOW_LEARNED)
            r.SetBricks(ranges, ghost=1, map_cells=cells)
            r.SetGraph(True)
            for label, budget in (("learned", LEARNED_BUDGET), ("single_block_map_exact", -1.0)):
                r.SetLearned(net, fallback=B.SHADOW_REF_MARCH, error_budget=budget, check_stride=8)
                sec, stats, cam_s, _ = frame_loop(r, sc, rgba, depth, steps, 1)
                images[label] = rgba.numpy().reshape(-1, 4)
                e_l, _, fb = r.GetLearnedStats()
                row[label] = {"frame_ms": sec * 1e3, "phase_ms": stats, "samples_per_frame": cam_s,
                              "ray_samples_per_s": cam_s / sec, "check_max_opacity_err": e_l,
                              "bricks_fell_back": fb}
                print(f"[bench] series {label} t={ts}: {sec * 1e3:.1f} ms, fell back {fb}", file=sys.stderr,
                      flush=True)
        sc.field_store.free()
        row["learned"]["rgba_err_vs_single_block_map_exact"] = image_error(images["learned"],
                                                                           images["single_block_map_exact"])
        row["learned"]["rgba_err_vs_multirank_exact"] = image_error(images["learned"], ex["images"][ts])
        row["single_block_map_exact"]["rgba_err_vs_multirank_exact"] = image_error(
            images["single_block_map_exact"], ex["images"][ts])
        rows.append(row)
    rgba.free()
    depth.free()

    def mean(label, key):
        return float(np.mean([r[label][key] for r in rows]))

    worst = max(r["learned"]["rgba_err_vs_single_block_map_exact"]["max"] for r in rows)
    return {"workload": (f"learned shadows on the time series: the same bricks and frames, per-brick "
                         f"learned maps checked at every 8th vertex per axis against the whole volume's "
                         f"light march, tiles over a {LEARNED_BUDGET} opacity budget (own or "
                         f"neighbouring corners) re-marched; one hipGraph per frame"),
            "check_budget": LEARNED_BUDGET,
            "frame_ms": mean("learned", "frame_ms"),
            "single_block_map_exact_frame_ms": mean("single_block_map_exact", "frame_ms"),
            "speedup_vs_single_block_map_exact": mean("single_block_map_exact", "frame_ms") /
            mean("learned", "frame_ms"),
            "max_rgba_err_vs_single_block_map_exact": worst,
            "per_timestep": rows}


def precision_report(n, img, threads, net=None):
    """Every storage and shadow variant of the precision workload against the FP64 oracle:
    the reference algorithm in FP64 on the exact (FP64) field with FP64 camera rays."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from pilot_amd import beams as B
    from pilot_amd import scenes

    def scene(fdt, mdt):
        return scenes.make_scene("precision", n, seed=2, nblobs=64, noise_octaves=3, noise_amp=0.2,
                                 light="point", width=img, height=img, field_dtype=fdt, map_dtype=mdt,
                                 map_cells=(n, n, n))

    sc = scene(B.F32, B.F32)
    o, d, o64, d64 = B.camera_rays(sc.camera, img, img, want64=True)
    t0 = time.perf_counter()
    rb = O.RayBuf(o64, d64, "f64")
    O.render_rays(O.volume_from_scene(sc, "f64"), rb, threads)
    t_oracle = time.perf_counter() - t0
    ref = rb.rgba
    variants = [("f32 field / f32 map, reference march", B.F32, B.F32, B.SHADOW_REF_MARCH),
                ("f16 field / f16 map, reference march", B.F16, B.F16, B.SHADOW_REF_MARCH),
                ("bf16 field / f16 map, reference march", B.BF16, B.F16, B.SHADOW_REF_MARCH),
                ("f32 field / f32 map, exact sweep", B.F32, B.F32, B.SHADOW_SWEEP),
                ("f16 field / f16 map, exact sweep", B.F16, B.F16, B.SHADOW_SWEEP)]
    if net is not None:
        variants.append(("f16 field / f16 map, learned (no fallback)", B.F16, B.F16, B.SHADOW_LEARNED))
    rows = {}
    for label, fdt, mdt, mode in variants:
        scv = sc if (fdt, mdt) == (B.F32, B.F32) else scene(fdt, mdt)
        with B.LightedVolumeRenderer() as r:
            scenes.configure(r, scv)
            r.SetShadow(True, scv.map_cells, mdt, mode=mode)
            if mode == B.SHADOW_LEARNED:
                r.SetLearned(net, error_budget=1e30)
            out, _ = r.RenderImage(scv.camera, img, img)
        rows[label] = image_error(out.reshape(-1, 4), ref)
        print(f"[bench] precision {label}: {rows[label]}", file=sys.stderr, flush=True)
    return {
        "workload": (f"precision sweep: {n}^3 blob+noise field, {n}^3-cell map, point light, "
                     f"{img}x{img}; per-pixel RGBA error against the FP64 oracle "
                     f"({threads} threads, {t_oracle:.1f} s)"),
        "rgba_err_vs_fp64_oracle": rows,
    }


def fp64_report(threads, net, learned_n=512, learned_img=2048, tile=256, small=161):
    """RGBA error against the FP64 oracle (the reference algorithm in FP64 on the exact field,
    FP64 rays) for the workloads beyond the precision frame:
    - the 512^3 single-block learned workload (exact reference march; learned with and without
      its fallback) on a 256^2 tile of its 2048^2 frame (the one with the most partly
      transparent, lit pixels): the FP64 map is built at the vertices inside the bounding box of
      the tile rays' entry and exit points (the rays' samples look up no other vertex);
    - the brick workloads (time-series recipe, 128 bricks; batched-bricks recipe, 96 bricks) at
      161^3 with 256^2 frames: the FP64 multi-rank pipeline, and for the learned bricks their
      FP64 counterpart (the single-block map at the bricks' vertices)."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from pilot_amd import beams as B
    from pilot_amd import scenes

    out = {}
    # ---- single block, tile
    sc = learned_scene(2001, learned_n, learned_img, device=False)
    vol = O.volume_from_scene(sc, "f64")
    o, d, o64, d64 = B.camera_rays(sc.camera, learned_img, learned_img, want64=True)
    modes = [("exact reference march", B.SHADOW_REF_MARCH, None)]
    if net is not None:
        modes += [("learned, no fallback", B.SHADOW_LEARNED, 1e30),
                  (f"learned, budget {LEARNED_BUDGET:g}", B.SHADOW_LEARNED, LEARNED_BUDGET)]
    images = {}
    for label, mode, budget in modes:
        with B.LightedVolumeRenderer() as r:
            scenes.configure(r, sc)
            r.SetShadow(True, sc.map_cells, B.F16, mode=mode)
            if budget is not None:
                r.SetLearned(net, fallback=B.SHADOW_REF_MARCH, error_budget=budget, check_stride=8)
            img, _ = r.RenderImage(sc.camera, learned_img, learned_img)
            images[label] = img.reshape(-1, 4)
    # the tile with the most partly transparent, lit pixels of the exact frame (much of the
    # frame is background or the fully shadowed core, where every path agrees)
    ex = images["exact reference march"].reshape(learned_img, learned_img, 4)
    nt = learned_img // tile
    score = [(((ex[ty * tile:(ty + 1) * tile, tx * tile:(tx + 1) * tile, 3] > 0.05) &
               (ex[ty * tile:(ty + 1) * tile, tx * tile:(tx + 1) * tile, :3].max(axis=2) > 0.05)).sum(), ty, tx)
             for ty in range(nt) for tx in range(nt)]
    _, ty, tx = max(score)
    idx = ((np.arange(ty * tile, (ty + 1) * tile)[:, None] * learned_img) +
           np.arange(tx * tile, (tx + 1) * tile)[None, :]).ravel()
    to, td = o64[:, idx], d64[:, idx]
    lo = np.array([c[0] for c in sc.coords64]); hi = np.array([c[-1] for c in sc.coords64])
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = (lo[:, None] - to) / td
        t1 = (hi[:, None] - to) / td
    tin = np.maximum(np.nanmax(np.minimum(t0, t1), axis=0), 0.0)
    tout = np.nanmin(np.maximum(t0, t1), axis=0)
    hit = tout > tin
    pts = np.concatenate([to[:, hit] + td[:, hit] * tin[hit], to[:, hit] + td[:, hit] * tout[hit]], axis=1)
    cells = np.array(sc.map_cells)
    sp = (hi - lo) / cells
    vlo = np.clip(np.floor((pts.min(axis=1) - lo) / sp).astype(int) - 1, 0, cells)
    vhi = np.clip(np.ceil((pts.max(axis=1) - lo) / sp).astype(int) + 1, 0, cells)
    ax = [lo[a] + sp[a] * np.arange(vlo[a], vhi[a] + 1) for a in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    t_map = time.perf_counter()
    vals = O.march_points(vol, [X.ravel(), Y.ravel(), Z.ravel()], threads)
    mp = cells + 1
    full = np.full(int(np.prod(mp)), np.nan)
    ii, jj, kk = np.meshgrid(np.arange(vlo[2], vhi[2] + 1), np.arange(vlo[1], vhi[1] + 1),
                             np.arange(vlo[0], vhi[0] + 1), indexing="ij")
    full[(kk + mp[0] * (jj + mp[1] * ii)).ravel()] = vals
    rb = O.RayBuf(to, td, "f64")
    O.sample_rays(vol, full, rb, threads)
    ref = rb.rgba.reshape(-1, 4)
    assert np.isfinite(ref).all(), "a tile ray looked up a map vertex outside the built box"
    t_map = time.perf_counter() - t_map
    rows = {}
    for label, _, _ in modes:
        rows[label] = image_error(images[label][idx], ref)
        print(f"[bench] fp64 {learned_n}^3 tile {label}: {rows[label]}", file=sys.stderr, flush=True)
    lit = int(((ref[:, 3] > 0.05) & (ref[:, :3].max(axis=1) > 0.05)).sum())
    out[f"{learned_n}^3 FP16 single block, {learned_img}^2, {tile}^2 tile"] = {
        "tile_origin_px": [int(tx * tile), int(ty * tile)], "tile_lit_px": lit,
        "rgba_err_vs_fp64_oracle": rows, "fp64_map_vertices": int(vals.size), "oracle_s": t_map}
    # ---- bricks at a reduced size
    recipes = [("time-series recipe, 128 bricks", series_scene(0, small, tile, device=False),
                B.kd_bricks((small,) * 3, 128, SERIES["brick_seed"], SERIES["min_cells"] * small // 1024,
                            SERIES["max_cells"] * small // 1024)),
               ("batched-bricks recipe, 96 bricks",
                scenes.make_scene("bricks", small, seed=3, nblobs=64, noise_octaves=3, noise_amp=0.2,
                                  light="point", width=tile, height=tile, field_dtype=B.F16, map_dtype=B.F16),
                B.kd_bricks((small,) * 3, 96, 3, 32 * small // 512, 160 * small // 512))]
    for label, scb, ranges in recipes:
        cells_b = ranges[:, 1::2] - ranges[:, 0::2]
        o, d, o64, d64 = B.camera_rays(scb.camera, tile, tile, want64=True)
        vols64 = O.brick_volumes(scb, ranges, twin="f64", map_cells=cells_b)
        t_o = time.perf_counter()
        ref, _, _, _ = O.render_bricks(vols64, o64, d64, list(scb.camera.eye), nthreads=threads)
        rows = {}
        with B.LightedVolumeRenderer() as r:
            scenes.configure(r, scb)
            r.SetBricks(ranges, ghost=1, map_cells=cells_b)
            img, _ = r.RenderImage(scb.camera, tile, tile)
        rows["exact (multi-rank)"] = image_error(img.reshape(-1, 4), ref)
        if net is not None:
            whole = O.volume_from_scene(scb, "f64")
            maps = [O.march_points(whole, O.map_points(v), threads) for v in vols64]
            partial, bounds = [], []
            for v, m in zip(vols64, maps):
                rb = O.RayBuf(o64, d64, "f64")
                O.sample_rays(v, m, rb, threads)
                partial.append(rb.rgba)
                bounds += [float(v.coords[a][k]) for a in range(3) for k in (0, -1)]
            lref = O.composite(np.array(bounds), np.asarray(scb.camera.eye, np.float64), partial, "f64")
            with B.LightedVolumeRenderer() as r:
                scenes.configure(r, scb)
                r.SetShadow(True, scb.map_cells, B.F16, mode=B.SHADOW_LEARNED)
                r.SetLearned(net, fallback=B.SHADOW_REF_MARCH, error_budget=LEARNED_BUDGET, check_stride=8)
                r.SetBricks(ranges, ghost=1, map_cells=cells_b)
                img_l, _ = r.RenderImage(scb.camera, tile, tile)
            rows["learned (budget %g) vs its FP64 counterpart (single-block map at the bricks)"
                 % LEARNED_BUDGET] = image_error(img_l.reshape(-1, 4), lref)
            rows["learned vs the FP64 multi-rank frame"] = image_error(img_l.reshape(-1, 4), ref)
        t_o = time.perf_counter() - t_o
        for k, v in rows.items():
            print(f"[bench] fp64 {small}^3 {label} {k}: {v}", file=sys.stderr, flush=True)
        out[f"{small}^3 FP16 {label}, {tile}^2 frame"] = {"rgba_err_vs_fp64_oracle": rows, "oracle_s": t_o}
    return out


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the byte-counter runs summarised in
    profiles/traffic.json (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the counters' reading
    of a copy of known size), or None when that workload was not counted."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        entry = json.load(f).get(workload, {}).get(kernel)
    return None if entry is None else entry["bytes_per_launch"]


# MI355X vector issue: 256 CUs x 4 SIMDs, a wave64 VALU instruction occupies a SIMD for 4
# cycles (64 FP32 FMA lanes per CU per clock), 2.4 GHz peak engine clock.
SIMDS, VALU_CYCLES, CLOCK_GHZ = 1024, 4, 2.4


def valu_issue(kernel, workload, kernel_ms):
    """The kernel's VALU-issue bound from its instruction counters (profiles/instructions.json:
    SQ_INSTS_VALU and SQ_WAVES per dispatch): the time the SIMDs need just to issue those
    instructions, and its fraction of the measured kernel time; None when not counted."""
    path = os.path.join(ROOT, "profiles", "instructions.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        e = json.load(f).get(workload, {}).get(kernel)
    if e is None:
        return None
    issue_ms = e["sq_insts_valu"] * VALU_CYCLES / (SIMDS * CLOCK_GHZ * 1e9) * 1e3
    return {"valu_per_wave": e["valu_per_wave"], "waves": e["sq_waves"], "issue_ms": issue_ms,
            "kernel_ms": kernel_ms, "frac": issue_ms / kernel_ms}


def c2_cpu_check(sc, field32, gpu_map, threads, samples_per_frame, gpu_rgba=None):
    """Oracle (FP32 restatement of the reference) over the whole frame on the host cores:
    every map vertex (checked against the GPU map) and every camera ray, shaded against the
    bit-identical GPU map and compared with the GPU frame (gpu_rgba, W*H x 4).  The frame is
    a few seconds of work on 16 threads."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from pilot_amd import beams as B

    vol = O.Volume(sc.dims, sc.coords, field32, sc.tf32, sc.scalar_range, sc.light_kind,
                   [float(v) for v in sc.light_pos], [float(v) for v in sc.light_color],
                   sc.map_cells, True, sc.sample_distance, sc.alpha_stop, twin="f32")
    nv = vol.nverts
    cpu_map = np.zeros(nv, np.float32)
    t0 = time.perf_counter()
    O.tmap_build_range(vol, cpu_map, 0, nv, threads)
    t_map = time.perf_counter() - t0
    assert np.array_equal(cpu_map, gpu_map), "GPU map != oracle map"
    del cpu_map
    o, d = B.camera_rays(sc.camera, sc.width, sc.height)
    rb = O.RayBuf(o, d, "f32")
    t0 = time.perf_counter()
    O.sample_rays(vol, gpu_map, rb, threads)
    t_march = time.perf_counter() - t0
    frame_s = t_map + t_march
    nsamples = int(rb.nsamples.sum())
    return {
        "value": nsamples / frame_s,
        "unit": "ray samples/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"FP32 oracle, {threads} threads, the whole frame: map of {nv} vertices "
                   f"({t_map:.2f} s, equal to the GPU map) + {sc.width * sc.height} camera rays "
                   f"({t_march:.2f} s, {nsamples} samples) = {frame_s:.2f} s"),
        "frame_s": frame_s,
        "samples_equal_gpu": nsamples == samples_per_frame,
        "image_equal_gpu": (None if gpu_rgba is None
                            else bool(np.array_equal(rb.rgba, gpu_rgba.reshape(-1, 4)))),
    }


def c2_workload(args):
    """The precision-sweep frame (256^3 FP32 field, 256^3-cell FP32 map, 1024^2, reference
    march): eager frames with phase events and hipGraph replays; FP16/BF16 storage, cell-field
    and rectilinear variants; the FP32 oracle over the whole frame on the host cores (map and
    image compared with the GPU's)."""
    import numpy as np

    from pilot_amd import beams as B
    from pilot_amd import scenes
    from pilot_amd.devmem import DeviceArray

    n, img = args.size, args.image

    def scene(fdt, mdt):
        return scenes.make_scene("precision", n, seed=2, nblobs=64, noise_octaves=3, noise_amp=0.2,
                                 light="point", width=img, height=img, field_dtype=fdt,
                                 map_dtype=mdt, map_cells=(n, n, n), device=True)

    results = {}
    main_state = None
    for label, fdt, mdt in [("f32/f32", B.F32, B.F32), ("f16/f16", B.F16, B.F16),
                            ("bf16/f16", B.BF16, B.F16)]:
        if label != "f32/f32" and args.no_variants:
            continue
        sc = scene(fdt, mdt)
        rgba = DeviceArray((img, img, 4), np.float32)
        depth = DeviceArray((img, img), np.float32)
        r = B.LightedVolumeRenderer()
        scenes.configure(r, sc)
        sec, stats, cam_s, light_s = frame_loop(r, sc, rgba, depth, args.steps, args.warmup)
        results[label] = dict(frame_ms=sec * 1e3, map_ms=stats["map"], march_ms=stats["march"],
                              samples_per_frame=cam_s, light_steps_per_frame=light_s)
        if label == "f32/f32":
            gsec = graph_frame_s(r, sc, rgba, depth, args.steps)
            results[label].update(frame_ms_graph=gsec * 1e3, ray_samples_per_s=cam_s / gsec)
            main_state = (sc, r, stats, cam_s)
        else:
            r.close()
        print(f"[bench] precision frame {label}: {sec * 1e3:.3f} ms/frame eager", file=sys.stderr, flush=True)
    sc, r, stats, cam_s = main_state
    if not args.no_variants:
        cells = np.ascontiguousarray(sc.field_store.numpy().reshape(n, n, n)[:-1, :-1, :-1]).ravel()
        rgba = DeviceArray((img, img, 4), np.float32)
        depth = DeviceArray((img, img), np.float32)
        with B.LightedVolumeRenderer() as rc:
            rc.SetData(cells, sc.dims, coords=sc.coords, dtype=B.F32, assoc=B.ASSOC_CELLS)
            rc.SetColorMap(sc.tf32)
            rc.AddLight(sc.light_pos, sc.light_color, 1.0, kind=sc.light_kind)
            rc.SetShadow(False, sc.map_cells, B.F32)
            rc.SetSampleDistance(sc.sample_distance)
            csec, cstats, ccam, _ = frame_loop(rc, sc, rgba, depth, args.steps, args.warmup)
        results["f32 cells, no shadow"] = dict(frame_ms=csec * 1e3, march_ms=cstats["march"],
                                               samples_per_frame=ccam)
        rng = np.random.default_rng(7)
        rcoords = []
        for _ in range(3):
            w = rng.uniform(0.5, 1.5, size=n - 1)
            rcoords.append(np.concatenate([[0.0], np.cumsum(w) / w.sum()]).astype(np.float32))
        with B.LightedVolumeRenderer() as rr:
            rr.SetData(sc.field_store, sc.dims, coords=rcoords, dtype=B.F32)
            rr.SetColorMap(sc.tf32)
            rr.AddLight(sc.light_pos, sc.light_color, 1.0, kind=sc.light_kind)
            rr.SetShadow(True, sc.map_cells, B.F32)
            rr.SetSampleDistance(sc.sample_distance)
            rsec, rstats, rcam, rlight = frame_loop(rr, sc, rgba, depth, args.steps, args.warmup)
        results["f32 rectilinear coordinates"] = dict(frame_ms=rsec * 1e3, map_ms=rstats["map"],
                                                      march_ms=rstats["march"], samples_per_frame=rcam,
                                                      light_steps_per_frame=rlight)
        rgba.free()
        depth.free()
    npts = n ** 3
    nmap = int(np.prod([c + 1 for c in sc.map_cells]))
    bytes_map_kernel = npts * 4 + nmap * 4 + 256 * 16
    out = {
        "workload": (f"precision sweep, exact shadows: {n}^3 FP32 field, {n}^3-cell FP32 map (reference "
                     f"per-vertex march), point light, {img}x{img}, step 0.5 voxel"),
        "frame_ms_graph": results["f32/f32"]["frame_ms_graph"],
        "ray_samples_per_s": results["f32/f32"]["ray_samples_per_s"],
        "phase_ms": {"map": stats["map"], "march": stats["march"]},
        "k_tmap_ref": {"algorithmic_bytes": bytes_map_kernel, "ms": stats["map"],
                       "achieved_gbs": bytes_map_kernel / (stats["map"] * 1e-3) / 1e9,
                       "traffic": measured_traffic("k_tmap_ref", f"{n}^3 f32/f32, {n}^3-cell map, {img}x{img}"),
                       "valu_issue": valu_issue("k_tmap_ref", f"{n}^3 f32/f32, {n}^3-cell map, {img}x{img}",
                                                stats["map"])},
        "variants": results,
    }
    if not args.no_cpu:
        field32 = sc.field_store.numpy().astype(np.float32)
        gimg, _ = r.RenderImage(sc.camera, sc.width, sc.height)
        gmap = r.GetMap(sc.map_cells)
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_full_frame"] = c2_cpu_check(sc, field32, gmap, threads, cam_s, gimg)
    r.close()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this renderer runs on one GPU; bench.py measures --gpus 1 only",
                          "n_gpus_requested": args.gpus}))
        sys.exit(2)
    import numpy as np

    from pilot_amd.devmem import D2D, DeviceArray, hip

    if args.copy_calibration:
        a = DeviceArray((1 << 28,), np.float32).zero()
        b = DeviceArray((1 << 28,), np.float32)
        for _ in range(5):
            hip().hipMemcpy(b.data_ptr(), a.data_ptr(), a.nbytes, D2D)
        hip().hipDeviceSynchronize()
        print(json.dumps({"copy_calibration_bytes": a.nbytes, "copies": 5}))
        return

    n, img, T = args.series_size, args.series_image, max(1, min(8, args.series_timesteps))
    nb = SERIES["nbricks"]
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    ex = series_exact(args.steps, args.warmup, n, img, T, schedule=json.loads(args.schedule) if args.schedule else None)
    sec = ex["sec"]
    map_bytes, field_bytes, maps_bytes = series_map_bytes(ex["ranges"], ex["cells"], 1, n)
    map_ms = ex["phase"]["map"]
    achieved = map_bytes / (map_ms * 1e-3) / 1e9
    counted = f"series {n}^3 f16/f16, {nb} bricks, {img}x{img}"  # counter-run key
    ext = ex["ranges"][:, 1::2] - ex["ranges"][:, 0::2]
    line = {
        "metric": METRIC,
        "value": ex["timed_samples"] / (sec * args.steps),
        "unit": "ray samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": ("synthetic: seeded 64-blob + 3-octave value-noise fields (blobs drifting 0.01 box/step, "
                 "noise phase +0.25/step) generated in HBM, seeded TF/light/camera/kd bricks"),
        "config": {
            "workload": (f"time series, exact shadows: {n}^3 FP16 field (FP32 arithmetic), {T} timesteps "
                         f"cycled, {nb} kd bricks (cell extents {int(ext.min())}..{int(ext.max())}, 1 ghost "
                         f"layer, FP16 maps with a cell per field cell), maps rebuilt every frame (local "
                         f"march, non-local exchange, seeded re-march), per-brick march, depth-ordered "
                         f"composite, {img}x{img}, step 0.5 voxel, one hipGraph per frame"),
            "field": f"{n}^3", "bricks": nb, "timesteps": T, "image": f"{img}x{img}",
            "map_vertices": int(np.prod(ex["cells"] + 1, axis=1).sum()),
            "shadow": "exact (reference per-vertex march per brick + exchange)",
            "storage": "f16 field / f16 maps",
        },
        "frame_ms": sec * 1e3,
        "timing": ("value and ms_per_step: K hipGraph launches of whole frames cycling through the "
                   "timesteps between stream synchronisations; value = their camera-ray samples "
                   "(counted on the GPU per timestep) / their time; phase_ms: HIP events around each "
                   "phase of 2 eager frames per timestep on the renderer's stream"),
        "phase_ms": ex["phase"],
        "schedule": ex["schedule"],
        "schedule_frame_ms": ex["schedule_ms"],
        "samples_per_frame": ex["samples"],
        "light_steps_per_frame": ex["light_steps"],
        "roofline": {
            "kernel": "k_bricks_map (phase-1 local maps)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic("k_bricks_map", counted),
            "algorithmic_bytes": map_bytes,
            "algorithmic_bytes_parts": {"brick_fields_f16": field_bytes, "maps_f16": maps_bytes,
                                        "tf": 256 * 16},
            "launch_ms": map_ms,
            # the bound that applies: instruction issue, not HBM
            "valu_issue": valu_issue("k_bricks_map", counted, map_ms),
        },
        "cpu_baseline": None,
        "workloads": {},
    }
    print(f"[bench] headline: {sec * 1e3:.1f} ms/frame, {line['value']:.4g} samples/s", file=sys.stderr,
          flush=True)
    if not args.no_cpu:
        line["cpu_baseline"] = series_cpu_baseline(ex, n, img, threads, args.cpu_runs)
    else:
        ex["first_scene"].field_store.free()
    ex["renderer0"].close()
    exact_images = ex["images"]
    del ex
    if not args.no_variants:
        line["workloads"]["precision_frame"] = c2_workload(args)
    if not args.no_bricks:
        line["workloads"]["bricks_512_96"] = bricks_workload(max(3, args.steps // 2), 1)
        print(f"[bench] bricks: {line['workloads']['bricks_512_96']['frame_ms_graph']:.3f} ms/frame",
              file=sys.stderr, flush=True)
    net = None
    if not args.no_learned:
        seeds = tuple(1001 + i for i in range(8))
        net, train = train_closure(args.learned_size, seeds, args.learned_samples // len(seeds),
                                   args.learned_epochs)
        print(f"[bench] closure trained: {train}", file=sys.stderr, flush=True)
        wl = learned_workload(max(2, args.steps // 4), 1, net, n=args.learned_size,
                              seeds=(2001, 2002, 2003, 2004)[:max(1, min(4, args.learned_eval_seeds))])
        wl["training"] = train
        line["workloads"]["learned"] = wl
        sl = series_learned(2, net, {"ranges": series_bricks(n)[0], "cells": series_bricks(n)[1],
                                     "images": exact_images}, n, img, T)
        sl["speedup_vs_exact"] = line["frame_ms"] / sl["frame_ms"]
        line["workloads"]["series_learned"] = sl
    if not args.no_precision:
        line["precision"] = precision_report(args.size, args.image, threads, net)
        line["precision"]["other_workloads"] = fp64_report(threads, net)
    if net is not None:
        net.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
