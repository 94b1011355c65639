"""hipGraph replay of a frame and the schedule autotuner: every replay, with eager work in
between, equals the eager frame; the chosen schedule does not change the image."""

import numpy as np
import pytest

from pilot_amd import beams as B
from pilot_amd import scenes
from pilot_amd.devmem import DeviceArray

pytestmark = pytest.mark.gpu


def _device_frame(r, sc, rgba, depth):
    r.RenderImage(sc.camera, sc.width, sc.height, rgba=rgba, depth=depth)
    r.Synchronize()
    return rgba.numpy().copy(), depth.numpy().copy()


@pytest.mark.parametrize("bricks", [False, True])
def test_graph_replays_equal_eager(bricks):
    sc = scenes.make_scene("graph", 41, seed=23, nblobs=10, width=96, height=80)
    rgba = DeviceArray((sc.height, sc.width, 4), np.float32)
    depth = DeviceArray((sc.height, sc.width), np.float32)
    other = DeviceArray(40 * 40 * 40, np.float32)
    with B.LightedVolumeRenderer() as r:
        scenes.configure(r, sc)
        if bricks:
            r.SetBricks(B.kd_bricks(sc.dims, 8, 3, 8, 24), ghost=1)
        ref_img, ref_depth = _device_frame(r, sc, rgba, depth)
        r.SetGraph(True)
        for k in range(4):
            rgba.zero()
            B.gen_field((40, 40, 40), seed=k, nblobs=3, out=other)  # eager work between replays
            img, dep = _device_frame(r, sc, rgba, depth)
            assert np.array_equal(img, ref_img), f"replay {k}"
            assert np.array_equal(dep, ref_depth), f"replay {k}"
        t = r.GetTimings()
        assert t["frame"] > 0
        # a changed camera is captured again
        cam2 = B.orbit_camera(sc.bounds, sc.extras["azimuth"] + 20.0, 5.0)
        r.RenderImage(cam2, sc.width, sc.height, rgba=rgba, depth=depth)
        r.SetGraph(False)
        eager2, _ = r.RenderImage(cam2, sc.width, sc.height)
        r.Synchronize()
        assert np.array_equal(rgba.numpy(), eager2)


def test_autotune_keeps_image():
    sc = scenes.make_scene("tune", 41, seed=29, nblobs=10, width=96, height=80)
    with B.LightedVolumeRenderer() as r:
        scenes.configure(r, sc)
        r.SetBricks(B.kd_bricks(sc.dims, 8, 5, 8, 24), ghost=1)
        before, _ = r.RenderImage(sc.camera, sc.width, sc.height)
        groups, ms = r.Autotune(sc.camera, sc.width, sc.height, frames=2)
        assert groups in (256, 512, 1024, 2048, 4096, 8192) and ms > 0
        after, _ = r.RenderImage(sc.camera, sc.width, sc.height)
    assert np.array_equal(before, after)
