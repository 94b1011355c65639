"""Writes tests/golden/*.npz: small seeded scenes with the oracle's FP32 and FP64 outputs.

The reference ships no golden vectors for this path; these fixtures pin the oracle
(and the GPU path) against regressions.  Inputs come from the library's seeded
generators; outputs from oracle/liboracle.so.  Run from the repository root:
    python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import oracle as O  # noqa: E402
from pilot_amd import beams as B  # noqa: E402
from pilot_amd import scenes  # noqa: E402

CASES = {
    # name: (n, seed, nblobs, octaves, light, map cells, width, height, ntf)
    "tiny_point": (8, 101, 4, 0, "point", (5, 6, 7), 16, 12, 16),
    "tiny_directional_noise": (10, 202, 6, 3, "directional", (9, 9, 9), 20, 14, 256),
}


def build(name):
    n, seed, nblobs, octv, light, mc, w, h, ntf = CASES[name]
    sc = scenes.make_scene(name, n, seed, nblobs, noise_octaves=octv, noise_amp=0.2 if octv else 0.0,
                           light=light, map_cells=mc, width=w, height=h)
    tf32, tf64 = B.gen_colormap(seed + 1, 6, ntf)
    sc.tf32, sc.tf64 = tf32, tf64
    o, d, o64, d64 = B.camera_rays(sc.camera, w, h, want64=True)
    out = dict(field32=sc.field32, field64=sc.field64, coords32=np.stack(sc.coords),
               coords64=np.stack(sc.coords64), tf32=tf32, tf64=tf64,
               light_kind=np.int32(sc.light_kind), light_pos=sc.light_pos, light_color=sc.light_color,
               map_cells=np.array(mc), sample_distance=np.float32(sc.sample_distance),
               origin=o, dir=d, origin64=o64, dir64=d64)
    for tw in ("f32", "f64"):
        vol = O.volume_from_scene(sc, tw)
        rb = O.RayBuf(o if tw == "f32" else o64, d if tw == "f32" else d64, tw)
        m = O.render_rays(vol, rb)
        out[f"map_{tw}"] = m
        out[f"rgba_{tw}"] = rb.rgba
        out[f"min_{tw}"] = rb.min_d
        out[f"max_{tw}"] = rb.max_d
        out[f"nsamples_{tw}"] = rb.nsamples
    return sc, out


def main():
    for name in CASES:
        _, out = build(name)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print(name, {k: v.shape for k, v in out.items() if hasattr(v, "shape")})


if __name__ == "__main__":
    main()
