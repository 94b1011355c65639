#!/usr/bin/env python3
"""Lit-frame benchmark of the MI355X Beams renderer on one GPU.

A step is one lit frame of the exact-shadow path: transmittance-map build (the
reference's per-vertex light march) followed by the camera-ray march with shadow
lookups, rendered into HBM-resident RGBA + depth buffers.  The workload is the
precision-sweep configuration: a 256^3 field (64 Gaussian blobs + 3-octave value
noise at 0.2) generated directly in HBM, a 256^3-cell (257^3-vertex) map, a point
light outside the box, 1024x1024 pixels, step 0.5 voxel.

Prints one JSON line: ray samples/s (composited camera-ray samples, counted on the
GPU) with frame ms, per-phase kernel times from HIP events on the renderer's
stream, the dominant kernel's roofline entry, and the CPU oracle timed on a bounded
sample of the same frame.
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
    ap.add_argument("--no-series", action="store_true", help="skip the 1024^3 time-series workload")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the RGBA error report against the FP64 oracle")
    ap.add_argument("--learned-size", type=int, default=512)
    ap.add_argument("--learned-samples", type=int, default=4 << 20)
    ap.add_argument("--learned-epochs", type=int, default=12)
    ap.add_argument("--learned-eval-seeds", type=int, default=4, help="held-out fields evaluated (1..4)")
    ap.add_argument("--series-timesteps", type=int, default=8, help="timesteps of the time series (1..8)")
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
    """Seconds per frame replayed from the frame's hipGraph (exact modes: the whole frame;
    the learned closure: its march and composite, its map build stays eager)."""
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
    return scenes.make_scene("learned", n, seed=seed, nblobs=64, noise_octaves=3, noise_amp=0.2,
                             light="point", width=img, height=img, field_dtype=B.F16,
                             map_dtype=B.F16, device=True, **kw)


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
                     budget=2e-3):
    """Held-out fields: the exact paths (reference per-vertex march and the exact sweep) and
    the learned closure (with and without its fallback) on the same frame."""
    import numpy as np

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
                     f"trained on 8 fields' reference-march maps; fallback to the reference march above "
                     f"a {budget} opacity budget at every 8th vertex"),
        "frame_ms": {k: mean(k, "frame_ms") for k in ("exact_ref_march", "exact_sweep",
                                                      "learned_no_fallback", "learned")},
        "speedup_learned_no_fallback_vs_ref_march": mean("exact_ref_march", "frame_ms") /
        mean("learned_no_fallback", "frame_ms"),
        "speedup_learned_no_fallback_vs_sweep": mean("exact_sweep", "frame_ms") /
        mean("learned_no_fallback", "frame_ms"),
        "per_seed": out,
    }


def time_series_workload(steps, warmup, net, n=1024, img=2048, timesteps=8, nbricks=128):
    """Drifting 1024^3 FP16 field over 8 timesteps in 128 kd bricks (one ghost layer, maps at
    half the voxel resolution), the shadow maps rebuilt every timestep, one hipGraph per frame
    for the exact path; exact and learned side by side."""
    import numpy as np

    from pilot_amd import beams as B
    from pilot_amd import scenes
    from pilot_amd.devmem import DeviceArray

    ranges = B.kd_bricks((n, n, n), nbricks, 5, 96, 320)
    cells = ranges[:, 1::2] - ranges[:, 0::2]  # one map cell per field cell
    rgba = DeviceArray((img, img, 4), np.float32)
    depth = DeviceArray((img, img), np.float32)
    rows = []
    for ts in range(timesteps):
        sc = scenes.make_scene("series", n, seed=9, nblobs=64, noise_octaves=3, noise_amp=0.2,
                               light="point", width=img, height=img, field_dtype=B.F16,
                               map_dtype=B.F16, device=True, time=float(ts), drift_speed=0.01,
                               noise_phase_rate=0.25)
        row = {"timestep": ts}
        imgs = {}
        for label, mode in (("exact", B.SHADOW_REF_MARCH), ("learned", B.SHADOW_LEARNED)):
            with B.LightedVolumeRenderer() as r:
                scenes.configure(r, sc)
                r.SetShadow(True, sc.map_cells, B.F16, mode=mode)
                if mode == B.SHADOW_LEARNED:
                    r.SetLearned(net, fallback=B.SHADOW_REF_MARCH, error_budget=2e-3, check_stride=8)
                r.SetBricks(ranges, ghost=1, map_cells=cells)
                r.SetGraph(True)
                sec, stats, cam_s, _ = frame_loop(r, sc, rgba, depth, steps, warmup)
                imgs[label] = rgba.numpy()
                row[label] = {"frame_ms": sec * 1e3, "phase_ms": stats, "samples_per_frame": cam_s,
                              "ray_samples_per_s": cam_s / sec}
                if mode == B.SHADOW_LEARNED:
                    e_l, _, fb = r.GetLearnedStats()
                    row[label].update({"check_max_opacity_err": e_l, "bricks_fell_back": fb})
            print(f"[bench] series t={ts} {label}: {sec * 1e3:.2f} ms", file=sys.stderr, flush=True)
        row["learned"]["rgba_err_vs_exact"] = image_error(imgs["learned"], imgs["exact"])
        rows.append(row)
        sc.field_store.free()
    rgba.free()
    depth.free()
    ext = ranges[:, 1::2] - ranges[:, 0::2]
    return {
        "workload": (f"time series: {n}^3 FP16 blob+noise field, {timesteps} timesteps (blob drift 0.01 "
                     f"box/step, noise phase +0.25/step), {nbricks} kd bricks (cell extents "
                     f"{int(ext.min())}..{int(ext.max())}, 1 ghost layer, a map cell per field cell), "
                     f"{img}x{img}, maps rebuilt every timestep; exact = multi-rank exchange in one "
                     "hipGraph per frame, learned = per-brick closure with per-brick fallback"),
        "frame_ms_exact": float(np.mean([r["exact"]["frame_ms"] for r in rows])),
        "frame_ms_learned": float(np.mean([r["learned"]["frame_ms"] for r in rows])),
        "per_timestep": rows,
    }


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


def cpu_baseline(sc, field32, gpu_map, threads, samples_per_frame, gpu_rgba=None):
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


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this renderer runs on one GPU; bench.py measures --gpus 1 only",
                          "n_gpus_requested": args.gpus}))
        sys.exit(2)
    import numpy as np

    from pilot_amd import beams as B
    from pilot_amd import scenes
    from pilot_amd.devmem import D2D, DeviceArray, hip

    if args.copy_calibration:
        a = DeviceArray((1 << 28,), np.float32).zero()
        b = DeviceArray((1 << 28,), np.float32)
        for _ in range(5):
            hip().hipMemcpy(b.data_ptr(), a.data_ptr(), a.nbytes, D2D)
        hip().hipDeviceSynchronize()
        print(json.dumps({"copy_calibration_bytes": a.nbytes, "copies": 5}))
        return

    n, img = args.size, args.image

    def scene(fdt, mdt):
        # a map of n cells per axis: (n+1)^3 vertices (SetShadowMapSize(n, n, n))
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
        print(f"[bench] {label}: {sec * 1e3:.3f} ms/frame, map {stats['map']:.3f} ms, "
              f"march {stats['march']:.3f} ms, {cam_s} samples", file=sys.stderr, flush=True)
        if label == "f32/f32":
            # the headline frame as the renderer replays it: one hipGraph of map + march
            gsec = graph_frame_s(r, sc, rgba, depth, args.steps)
            print(f"[bench] {label}: {gsec * 1e3:.3f} ms/frame as hipGraph replays", file=sys.stderr,
                  flush=True)
            main_state = (sc, r, sec, gsec, stats, cam_s, light_s)
        else:
            r.close()

    sc, r, sec_eager, sec, stats, cam_s, light_s = main_state
    if not args.no_variants:
        # SamplerCellAssoc: the same field as (n-1)^3 cell values, no shadow term
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
        print(f"[bench] f32 cell field, no shadow: {csec * 1e3:.3f} ms/frame, {ccam} samples",
              file=sys.stderr, flush=True)
        # RectilinearLocator: the same values on unevenly spaced coordinates (cell widths
        # varying 3x along each axis), exact shadows
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
        print(f"[bench] f32 rectilinear coordinates: {rsec * 1e3:.3f} ms/frame, map "
              f"{rstats['map']:.3f} ms, march {rstats['march']:.3f} ms, {rcam} samples",
              file=sys.stderr, flush=True)
        rgba.free()
        depth.free()
    npts = n ** 3
    nmap = int(np.prod([c + 1 for c in sc.map_cells]))
    npix = img * img
    bytes_map_kernel = npts * 4 + nmap * 4 + 256 * 16
    bytes_march_kernel = npts * 4 + nmap * 4 + npix * 20 + 256 * 16
    if stats["map"] >= stats["march"]:
        dom, dom_ms, dom_bytes = "k_tmap_ref", stats["map"], bytes_map_kernel
    else:
        dom, dom_ms, dom_bytes = "k_march_image", stats["march"], bytes_march_kernel
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    counted = f"{n}^3 f32/f32, {sc.map_cells[0]}^3-cell map, {img}x{img}"  # counter-run key
    line = {
        "metric": METRIC,
        "value": cam_s / sec,
        "unit": "ray samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded 64-blob + 3-octave value-noise field generated in HBM, seeded TF/light/camera",
        "config": {
            "workload": (f"precision sweep, exact shadows: {n}^3 FP32 field, {n}^3-cell FP32 "
                         f"transmittance map (reference per-vertex march), point light, "
                         f"{img}x{img} image, step 0.5 voxel"),
            "field": f"{n}^3", "map_vertices": nmap, "image": f"{img}x{img}",
            "shadow": "exact (reference per-vertex march)", "storage": "f32 field / f32 map",
        },
        "frame_ms": sec * 1e3,
        "frame_ms_eager": sec_eager * 1e3,
        "timing": ("value and ms_per_step: K hipGraph replays of the whole frame (map + march) "
                   "between stream synchronisations; phase_ms: HIP events around each phase "
                   "of K eager frames"),
        "phase_ms": {"map": stats["map"], "march": stats["march"]},
        "samples_per_frame": cam_s,
        "light_steps_per_frame": light_s,
        "variants": results,
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(dom, counted),
            "algorithmic_bytes": dom_bytes,
            # the bound that applies: instruction issue, not HBM
            "valu_issue": valu_issue(dom, counted, dom_ms),
        },
        "cpu_baseline": None,
    }
    if not args.no_bricks:
        line["workloads"] = {"bricks_512_96": bricks_workload(max(3, args.steps // 2), 1)}
        print(f"[bench] bricks: {line['workloads']['bricks_512_96']['frame_ms_graph']:.3f} ms/frame",
              file=sys.stderr, flush=True)
    net = None
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    if not args.no_learned:
        seeds = tuple(1001 + i for i in range(8))
        net, train = train_closure(args.learned_size, seeds, args.learned_samples // len(seeds),
                                   args.learned_epochs)
        print(f"[bench] closure trained: {train}", file=sys.stderr, flush=True)
        wl = learned_workload(max(2, args.steps // 4), 1, net, n=args.learned_size,
                              seeds=(2001, 2002, 2003, 2004)[:max(1, min(4, args.learned_eval_seeds))])
        wl["training"] = train
        line.setdefault("workloads", {})["learned"] = wl
        if not args.no_series:
            line["workloads"]["time_series"] = time_series_workload(
                2, 1, net, timesteps=max(1, min(8, args.series_timesteps)))
    if not args.no_precision:
        line["precision"] = precision_report(n, img, threads, net)
    if net is not None:
        net.close()
    if not args.no_cpu:
        field32 = sc.field_store.numpy().astype(np.float32)
        gimg, _ = r.RenderImage(sc.camera, sc.width, sc.height)  # this frame's RGBA on the host
        gmap = r.GetMap(sc.map_cells)
        line["cpu_baseline"] = cpu_baseline(sc, field32, gmap, threads, cam_s, gimg)
    r.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
