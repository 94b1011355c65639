"""Committed fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the oracle reproduces them exactly (CPU), the library's generators reproduce their
inputs (CPU), and the GPU path reproduces the FP32 outputs exactly (GPU)."""

import glob
import os

import numpy as np
import pytest

from pilot_amd import beams as B

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(glob.glob(os.path.join(HERE, "*.npz")))


def _load(path):
    return dict(np.load(path, allow_pickle=False))


def _vol(O, fx, twin):
    return O.Volume([fx["coords32"].shape[1]] * 3,
                    list(fx["coords32"] if twin == "f32" else fx["coords64"]),
                    fx["field32"] if twin == "f32" else fx["field64"],
                    fx["tf32"] if twin == "f32" else fx["tf64"],
                    (0.0, 1.0), int(fx["light_kind"]), [float(v) for v in fx["light_pos"]],
                    [float(v) for v in fx["light_color"]], tuple(int(c) for c in fx["map_cells"]),
                    True,
                    float(fx["sample_distance"]) if twin == "f32"
                    else 0.5 / (fx["coords32"].shape[1] - 1),
                    1.0, twin=twin)


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
@pytest.mark.parametrize("twin", ["f32", "f64"])
def test_oracle_reproduces_fixture(oracle_mod, path, twin):
    O = oracle_mod
    fx = _load(path)
    vol = _vol(O, fx, twin)
    o = fx["origin"] if twin == "f32" else fx["origin64"]
    d = fx["dir"] if twin == "f32" else fx["dir64"]
    rb = O.RayBuf(o, d, twin)
    m = O.render_rays(vol, rb)
    assert np.array_equal(m, fx[f"map_{twin}"])
    assert np.array_equal(rb.rgba, fx[f"rgba_{twin}"])
    assert np.array_equal(rb.nsamples, fx[f"nsamples_{twin}"])
    assert np.array_equal(rb.min_d, fx[f"min_{twin}"])


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_fixture_twins_consistent(path):
    """FP32 and FP64 outputs of a fixture agree to FP32 rounding except near TF-index
    boundaries: a check that neither twin's plumbing drifted."""
    fx = _load(path)
    err = np.abs(fx["rgba_f32"] - fx["rgba_f64"])
    assert np.median(err) < 1e-6
    assert np.quantile(err, 0.95) < 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_gpu_reproduces_fixture(path):
    fx = _load(path)
    n = fx["coords32"].shape[1]
    with B.LightedVolumeRenderer() as r:
        r.SetData(fx["field32"], (n, n, n), coords=list(fx["coords32"]), dtype=B.F32)
        r.SetColorMap(fx["tf32"])
        r.AddLight(fx["light_pos"], fx["light_color"], 1.0, kind=int(fx["light_kind"]))
        r.SetShadow(True, tuple(int(c) for c in fx["map_cells"]))
        r.SetSampleDistance(float(fx["sample_distance"]))
        rays = r.Render(B.Rays(fx["origin"], fx["dir"]))
        m = r.GetMap(tuple(int(c) for c in fx["map_cells"]))
    assert np.array_equal(m, fx["map_f32"])
    assert np.array_equal(rays.rgba, fx["rgba_f32"])
    assert np.array_equal(rays.min_d, fx["min_f32"])
