"""The exact sweep (BEAMS_SHADOW_SWEEP): closed forms pin its FP64 restatement
(oracle/sweep_ref.py) on the CPU; on the GPU the kernel matches the restatement and the
march over a sweep map matches the oracle's sampler over the same map bit for bit."""

import numpy as np
import pytest

from pilot_amd import beams as B
from pilot_amd import scenes


def _unit(n):
    c = np.arange(n, dtype=np.float64) / (n - 1)
    return [c, c, c]


def _sigma(alpha_ratio, h):
    return -np.log1p(-min(alpha_ratio, 1 - 2.0 ** -24)) / h


def test_sweep_axis_aligned_light_is_exact():
    import sweep_ref as S

    n, cells = 17, 12
    tf = np.tile([1, 1, 1, 0.9], (256, 1))
    tf[int(0.4 * 255)] = [1, 1, 1, 0.9 * 0.3]
    m = S.sweep_map(np.full(n ** 3, 0.4), (n, n, n), _unit(n), tf, (0, 1), 1, (0, 0, 1.0),
                    (cells, cells, cells)).reshape(cells + 1, cells + 1, cells + 1)
    h = np.sqrt(3) / 128
    sig = _sigma(0.3, h)
    z = np.arange(cells + 1) / cells
    expect = 1 - np.exp(-sig * (z + 0.5 * h))
    assert np.allclose(m, expect[:, None, None], atol=1e-12, rtol=0)


@pytest.mark.parametrize("kind,light", [(1, (0.3, -0.45, 0.84)), (0, (1.7, 2.4, -1.6))])
def test_sweep_constant_field_matches_path_length(kind, light):
    """Constant extinction: opacity = 1 - exp(-sigma (L + h/2)) with L the light path
    inside the box; exact for a directional light away from side-entry seams, within the
    bilinear error elsewhere."""
    import sweep_ref as S

    n, cells = 17, 16
    tf = np.tile([1, 1, 1, 0.8], (256, 1))
    tf[int(0.6 * 255)] = [1, 1, 1, 0.8 * 0.05]
    m = S.sweep_map(np.full(n ** 3, 0.6), (n, n, n), _unit(n), tf, (0, 1), kind, light,
                    (cells, cells, cells))
    h = np.sqrt(3) / 128
    sig = _sigma(0.05, h)
    g = np.arange(cells + 1) / cells
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([X, Y, Z], -1).reshape(-1, 3)
    L = np.asarray(light, np.float64)
    if kind == 1:
        d = np.broadcast_to(L / np.linalg.norm(L), P.shape)
    else:
        d = (P - L) / np.linalg.norm(P - L, axis=1, keepdims=True)
    with np.errstate(divide="ignore"):
        t0 = np.where(d != 0, (0 - P) / d, np.inf)
        t1 = np.where(d != 0, (1 - P) / d, -np.inf)
    # distance back from P along -d to the box boundary
    back = np.min(np.where(d > 0, (P - 0) / np.where(d > 0, d, 1), np.where(d < 0, (1 - P) / np.where(d < 0, -d, 1), np.inf)), axis=1)
    del t0, t1
    expect = 1 - np.exp(-sig * (back + 0.5 * h))
    err = np.abs(m - expect)
    # bilinear interpolation is exact where the path length is linear across a cell; where a
    # cell straddles the line at which rays switch from a side face to the near face, the
    # path is underestimated by up to ~half a cell (here 1/16: ~0.12 in opacity at worst)
    # (a point light's path length is also curved across a cell)
    assert np.median(err) < (1e-4 if kind == 1 else 1e-3)
    assert np.quantile(err, 0.9) < 0.03
    assert err.max() < 0.15


def test_sweep_transparent_is_zero():
    import sweep_ref as S

    tf = np.tile([1, 1, 1, 0.0], (256, 1))
    tf[-1] = [1, 1, 1, 1.0]  # only scalars at 1.0 are dense
    m = S.sweep_map(np.full(9 ** 3, 0.5), (9, 9, 9), _unit(9), tf, (0, 1), 0, (2, 0.5, 0.5), (8, 8, 8))
    assert np.all(m == 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("light", ["point", "directional"])
def test_gpu_sweep_matches_restatement(oracle_mod, light):
    import sweep_ref as S

    O = oracle_mod
    sc = scenes.make_scene("sweep", 40, seed=37, nblobs=12, noise_octaves=3, noise_amp=0.2,
                           light=light, width=128, height=96)
    o, d = B.camera_rays(sc.camera, sc.width, sc.height)
    with B.LightedVolumeRenderer() as r:
        scenes.configure(r, sc)
        r.SetShadow(True, sc.map_cells, B.F32, mode=B.SHADOW_SWEEP)
        rays = r.Render(B.Rays(o, d))
        gmap = r.GetMap(sc.map_cells)
    ref = S.sweep_map(sc.field32, sc.dims, sc.coords, sc.tf32, sc.scalar_range, sc.light_kind,
                      sc.light_pos, sc.map_cells)
    err = np.abs(gmap - ref)
    print(f"sweep vs FP64 restatement: median {np.median(err):.2e} p99 {np.quantile(err, 0.99):.2e} max {err.max():.2e}")
    assert np.median(err) < 1e-5
    assert np.quantile(err, 0.99) < 1e-3
    # the march over this map is the exact path: bit-equal to the oracle's sampler
    rb = O.RayBuf(o, d, "f32")
    O.sample_rays(O.volume_from_scene(sc, "f32"), gmap, rb)
    assert np.array_equal(rays.rgba, rb.rgba)
