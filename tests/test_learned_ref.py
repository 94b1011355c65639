"""Pins of the learned closure's CPU restatement (oracle/learned_ref.py): the backward pass
against finite differences, the coarse optical depth against closed-form path lengths, and
the target/opacity round trip."""

import numpy as np
import pytest

import learned_ref as L


def _params(rng, scale=0.3):
    p = rng.normal(scale=scale, size=L.NPARAMS)
    p[L.OFF_W1 + 31 * L.HID:L.OFF_B1] = 0.0
    return p


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(3)
    p = _params(rng)
    x = rng.uniform(-1, 1, size=(24, L.FEAT))
    x[:, 31] = 0.0
    t = rng.normal(size=24)
    g, _ = L.gradients(p, x, t, emulate=False)

    def loss(q):
        y, _ = L.forward(q, x, emulate=False)
        return np.mean((y - t) ** 2)

    eps = 1e-6
    idx = np.concatenate([rng.choice(L.NPARAMS, 300, replace=False),
                          [L.OFF_B1, L.OFF_W2, L.OFF_B2, L.OFF_W3, L.OFF_B3]])
    fd = np.array([(loss(p + eps * np.eye(1, L.NPARAMS, i)[0]) - loss(p - eps * np.eye(1, L.NPARAMS, i)[0]))
                   / (2 * eps) for i in idx])
    assert np.allclose(g[idx], fd, rtol=1e-5, atol=1e-8)


def test_emulated_gradients_follow_rounding_points():
    """With every operand already a BF16 value and activations kept exact by construction
    (identity-sized weights), the emulated and plain gradients agree."""
    rng = np.random.default_rng(4)
    p = L.bf16(_params(rng))
    x = L.bf16(rng.integers(-4, 5, size=(8, L.FEAT)) / 4.0)
    x[:, 31] = 0
    t = rng.integers(-4, 5, size=8) / 4.0
    y0, _ = L.forward(p, x, emulate=False)
    y1, _ = L.forward(p, x, emulate=True)
    # only the hidden activations are rounded: the outputs differ by at most BF16 rounding
    assert np.allclose(y0, y1, rtol=2e-2, atol=2e-2 * np.abs(y0).max())
    g0, _ = L.gradients(p, x, t, emulate=False)
    g1, _ = L.gradients(p, x, t, emulate=True)
    assert np.linalg.norm(g1 - g0) <= 3e-2 * np.linalg.norm(g0)


@pytest.mark.parametrize("kind,light", [(1, (0.3, -0.5, 0.81)), (0, (1.8, 0.4, -1.1))])
def test_coarse_tau_constant_field_is_path_length(kind, light):
    n = 17
    c = (np.arange(n) / (n - 1)).astype(np.float32)
    tf = np.tile([1, 1, 1, 0.5], (256, 1)).astype(np.float32)
    sc = L.Scene(np.full(n ** 3, 0.5), (n, n, n), [c, c, c], tf, (0, 1), kind, light, (8, 8, 8))
    csig = sc.coarse_sigma()
    sig0 = -np.log1p(-(1.0 - 2.0 ** -24)) / sc.h  # alpha ratio 1 clamped below 1
    assert np.allclose(csig, sig0, rtol=1e-12)
    tau = sc.coarse_tau(csig)
    cn = L.COARSE_CELLS + 1
    g = np.arange(cn) / L.COARSE_CELLS
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    v = np.stack([X, Y, Z], -1)
    d = sc.light_dir(v)
    with np.errstate(divide="ignore", invalid="ignore"):
        back = np.where(-d > 0, (1 - v) / -d, np.where(-d < 0, (0 - v) / -d, np.inf)).min(-1)
    assert np.allclose(tau, sig0 * back, rtol=1e-9, atol=1e-9)


def test_features_of_constant_field():
    n = 9
    c = (np.arange(n) / (n - 1)).astype(np.float32)
    tf = np.tile([1, 1, 1, 0.8], (256, 1)).astype(np.float32)
    tf[int(0.25 * 255)] = [1, 1, 1, 0.2]
    sc = L.Scene(np.full(n ** 3, 0.25), (n, n, n), [c, c, c], tf, (0, 1), 1, (0, 0, 1), (4, 4, 4))
    ctau = sc.coarse_tau(sc.coarse_sigma())
    f, tc = sc.features(np.arange(125), ctau)
    assert np.allclose(f[:, 1:28], -np.log1p(-0.25) / 4, rtol=1e-6)
    assert np.allclose(f[:, 28:31], [0, 0, 1])
    assert np.all(f[:, 31] == 0)
    # directional light along +z: path from the z = 0 face to the vertex
    z = sc.map_vertices(np.arange(125))[:, 2]
    sig = -np.log1p(-0.25) / sc.h
    assert np.allclose(tc, np.minimum(sig * z, L.TAU_MAX), rtol=1e-9, atol=1e-9)
    assert np.allclose(f[:, 0], tc / 4)


def test_target_round_trip():
    rng = np.random.default_rng(5)
    op = rng.uniform(0, 1 - np.exp(-L.TAU_MAX) - 1e-6, 1000)
    tc = rng.uniform(0, L.TAU_MAX, 1000)
    y = L.target(tc, op)
    assert np.allclose(L.learned_opacity(tc, y), op, atol=1e-12)


def test_sample_ids_splitmix_reference_values():
    # SplitMix64 reference outputs for state 0 (the generator's published test vector)
    assert L.splitmix64(0) == 0xE220A8397B1DCDAF
    ids = L.sample_ids(7, 100, 1000)
    assert ids.min() >= 0 and ids.max() < 1000 and len(set(ids.tolist())) > 80
