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


def cpu_baseline(sc, field32, gpu_map, threads, samples_per_frame):
    """Oracle (FP32 restatement of the reference) on a bounded sample of the same frame:
    a z-slab of the map and a centred ray tile, extrapolated to the full frame."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from pilot_amd import beams as B

    vol = O.Volume(sc.dims, sc.coords, field32, sc.tf32, sc.scalar_range, sc.light_kind,
                   [float(v) for v in sc.light_pos], [float(v) for v in sc.light_color],
                   sc.map_cells, True, sc.sample_distance, sc.alpha_stop, twin="f32")
    nv = vol.nverts
    mp = [c + 1 for c in sc.map_cells]
    slab_k = max(1, mp[2] // 16)
    k0 = mp[2] // 2 - slab_k // 2
    begin, end = k0 * mp[0] * mp[1], (k0 + slab_k) * mp[0] * mp[1]
    scratch = np.zeros(nv, np.float32)
    t0 = time.perf_counter()
    O.tmap_build_range(vol, scratch, begin, end, threads)
    t_map = time.perf_counter() - t0
    assert np.array_equal(scratch[begin:end], gpu_map[begin:end]), "GPU map slab != oracle"
    # centred tile of rays, shaded against the (bit-identical) GPU map
    o, d = B.camera_rays(sc.camera, sc.width, sc.height)
    tile = min(256, sc.width)
    ys, xs = np.mgrid[0:tile, 0:tile]
    idx = ((ys + (sc.height - tile) // 2) * sc.width + xs + (sc.width - tile) // 2).ravel()
    rb = O.RayBuf(o[:, idx], d[:, idx], "f32")
    t0 = time.perf_counter()
    O.sample_rays(vol, gpu_map, rb, threads)
    t_march = time.perf_counter() - t0
    frame_s = t_map * nv / (end - begin) + t_march * (sc.width * sc.height) / idx.size
    return {
        "value": samples_per_frame / frame_s,
        "unit": "ray samples/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"FP32 oracle, {threads} threads: map z-slab of {end - begin} of {nv} vertices "
                   f"({t_map:.2f} s) + {idx.size} centred rays ({t_march:.2f} s, "
                   f"{int(rb.nsamples.sum())} samples), extrapolated to the full frame "
                   f"= {frame_s:.1f} s"),
        "frame_s_extrapolated": frame_s,
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
    from pilot_amd.devmem import DeviceArray

    n, img = args.size, args.image

    def scene(fdt, mdt):
        return scenes.make_scene("precision", n, seed=2, nblobs=64, noise_octaves=3, noise_amp=0.2,
                                 light="point", width=img, height=img, field_dtype=fdt,
                                 map_dtype=mdt, device=True)

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
            main_state = (sc, r, sec, stats, cam_s, light_s)
        else:
            r.close()

    sc, r, sec, stats, cam_s, light_s = main_state
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
            "workload": (f"precision sweep, exact shadows: {n}^3 FP32 field, {n - 1}^3-cell FP32 "
                         f"transmittance map (reference per-vertex march), point light, "
                         f"{img}x{img} image, step 0.5 voxel"),
            "field": f"{n}^3", "map_vertices": nmap, "image": f"{img}x{img}",
            "shadow": "exact (reference per-vertex march)", "storage": "f32 field / f32 map",
        },
        "frame_ms": sec * 1e3,
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
            "traffic": None,
            "algorithmic_bytes": dom_bytes,
        },
        "cpu_baseline": None,
    }
    if not args.no_bricks:
        line["workloads"] = {"bricks_512_96": bricks_workload(max(3, args.steps // 2), 1)}
        print(f"[bench] bricks: {line['workloads']['bricks_512_96']['frame_ms_graph']:.3f} ms/frame",
              file=sys.stderr, flush=True)
    if not args.no_cpu:
        field32 = sc.field_store.numpy().astype(np.float32)
        gmap = r.GetMap(sc.map_cells)
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(sc, field32, gmap, threads, cam_s)
    r.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
