"""Test oracle — NOT product code.

numpy restatement of the learned shadow closure (the product's BEAMS_SHADOW_LEARNED), which
has no counterpart in the reference.  It restates:

* the network 32 -> 64 -> 64 -> 1 (ReLU), forward and the mean-squared-error gradients,
  with the product's rounding points emulated (BF16 operands of the three products and of
  the two backward products, FP32-or-wider accumulation) or, with ``emulate=False``, in
  plain FP64 for the finite-difference pin;
* the coarse extinction mip, its optical depth and the per-vertex features and targets.

Only tests import this module.
"""

import numpy as np

FEAT, HID = 32, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + FEAT * HID
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID
NPARAMS = OFF_B3 + 1
COARSE_CELLS = 32
TAU_MAX = 8.0


def bf16(a):
    """Round to bfloat16 (nearest even) through float32, returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(FEAT, HID), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(HID, HID),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def forward(p, x, emulate=True):
    """Returns (y, cache)."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, w3, b3 = unpack(p)
    X = r(x)
    h1p = X @ r(W1) + b1
    h1 = np.maximum(h1p, 0.0)
    h1b = r(h1)
    h2p = h1b @ r(W2) + b2
    h2 = np.maximum(h2p, 0.0)
    y = h2 @ r(w3) + b3
    return y, (X, h1p, h1b, h2p, h2)


def gradients(p, x, t, emulate=True):
    """Gradients of mean((y - t)^2) with respect to the (rounded, when emulating) parameters,
    and the loss."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    W1, b1, W2, b2, w3, b3 = unpack(p)
    y, (X, h1p, h1b, h2p, h2) = forward(p, x, emulate)
    n = x.shape[0]
    err = y - np.asarray(t, np.float64)
    dy = 2.0 * err / n
    g = np.zeros(NPARAMS)
    g[OFF_B3] = dy.sum()
    g[OFF_W3:OFF_B3] = h2.T @ dy
    dh2 = np.where(h2p > 0, dy[:, None] * r(w3)[None, :], 0.0)
    g[OFF_B2:OFF_W3] = dh2.sum(0)
    dh2b = r(dh2)
    g[OFF_W2:OFF_B2] = (h1b.T @ dh2b).ravel()
    dh1 = np.where(h1p > 0, dh2b @ r(W2).T, 0.0)
    g[OFF_B1:OFF_W2] = dh1.sum(0)
    g[OFF_W1:OFF_B1] = (X.T @ r(dh1)).ravel()
    return g, float(np.mean(err ** 2))


# ---- features ---------------------------------------------------------------------------------

def splitmix64(seed):
    """One SplitMix64 output for the state `seed` (Python ints, mod 2^64)."""
    m = (1 << 64) - 1
    z = (seed + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def sample_ids(seed, n, nv):
    """The map vertices the product's sampler draws: SplitMix64(seed ^ (A (k + 1))) mod nv."""
    m = (1 << 64) - 1
    return np.array([splitmix64((seed ^ ((0x5851F42D4C957F2D * (k + 1)) & m)) & m) % nv for k in range(n)],
                    np.int64)


class Scene:
    """Inputs of the closure: point field (x fastest), coordinates, TF, range, light."""

    def __init__(self, field, dims, coords, tf, scalar_range, light_kind, light, map_cells):
        self.dims = tuple(int(d) for d in dims)
        self.f3 = np.asarray(field, np.float32).reshape(self.dims[2], self.dims[1], self.dims[0])
        self.coords = [np.asarray(c, np.float32) for c in coords]
        self.tf = np.asarray(tf, np.float32).reshape(-1, 4)
        self.smin, self.smax = (np.float32(scalar_range[0]), np.float32(scalar_range[1]))
        self.light_kind = light_kind
        L = np.asarray(light, np.float32)
        if light_kind == 1:
            L = (L / np.float32(np.sqrt(np.float32(np.dot(L, L))))).astype(np.float32)
        self.light = L
        self.map_cells = tuple(int(c) for c in map_cells)
        self.lo = np.array([c[0] for c in self.coords], np.float64)
        self.hi = np.array([c[-1] for c in self.coords], np.float64)
        size = (self.hi - self.lo).astype(np.float32)
        self.h = float(np.float32(np.sqrt(np.float32(np.sum(size * size)))) / np.float32(128.0))
        amax = np.float32(self.tf[:, 3].max())
        self.a_tab = (self.tf[:, 3] / amax).astype(np.float32)

    def tf_index(self, val):
        d = self.smax - self.smin
        inv = np.float32(1.0) / d if d != 0 else self.smin
        sc = ((np.asarray(val, np.float32) - self.smin) * np.float32(inv)).astype(np.float32)
        idx = (sc * np.float32(len(self.tf) - 1)).astype(np.float32)
        idx = np.trunc(idx).astype(np.int64)
        return np.clip(idx, 0, len(self.tf) - 1)

    def alpha(self, val):
        return self.a_tab[self.tf_index(val)].astype(np.float64)

    def patch(self, val):
        """Density-patch feature: -ln(1 - min(a, 1 - 2^-24)) / 4, in FP32 like the product."""
        a = np.minimum(self.a_tab[self.tf_index(val)], np.float32(1.0 - 2.0 ** -24))
        return (-np.log1p(-a) * np.float32(0.25)).astype(np.float32).astype(np.float64)

    def sigma(self, val):
        a = np.minimum(self.alpha(val), 1.0 - 2.0 ** -24)
        return -np.log1p(-a) / self.h

    # ---- coarse mip
    def coarse_sigma(self):
        cn = COARSE_CELLS + 1
        csp = (self.hi - self.lo) / COARSE_CELLS
        ranges = []
        for d in range(3):
            x = self.coords[d].astype(np.float64)
            rd = []
            for i in range(cn):
                c = self.lo[d] + csp[d] * i
                lo = int(np.searchsorted(x, c - 0.5 * csp[d], side="left"))
                hi = len(x) if i == cn - 1 else int(np.searchsorted(x, c + 0.5 * csp[d], side="left"))
                if hi <= lo:
                    lo, hi = self.nearest(d, c), self.nearest(d, c) + 1
                rd.append((lo, hi))
            ranges.append(rd)
        sig = self.sigma(self.f3)
        out = np.zeros((cn, cn, cn))
        for k, (z0, z1) in enumerate(ranges[2]):
            for j, (y0, y1) in enumerate(ranges[1]):
                for i, (x0, x1) in enumerate(ranges[0]):
                    out[k, j, i] = sig[z0:z1, y0:y1, x0:x1].mean()
        return out

    def nearest(self, d, v):
        x = self.coords[d].astype(np.float64)
        i = int(np.clip(np.searchsorted(x, v, side="right") - 1, 0, len(x) - 2))
        return i if v - x[i] <= x[i + 1] - v else i + 1

    def light_dir(self, v):
        v = np.asarray(v, np.float64)
        if self.light_kind == 1:
            return np.broadcast_to(self.light.astype(np.float64), v.shape).copy()
        d = v - self.light.astype(np.float64)
        return d / np.linalg.norm(d, axis=-1, keepdims=True)

    def coarse_lerp(self, grid, v):
        cn = grid.shape[0]
        csp = (self.hi - self.lo) / COARSE_CELLS
        u = np.clip((np.asarray(v, np.float64) - self.lo) / csp, 0, cn - 1)
        i = np.minimum(np.floor(u).astype(np.int64), cn - 2)
        t = u - i
        out = 0.0
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    w = ((t[..., 0] if dx else 1 - t[..., 0]) * (t[..., 1] if dy else 1 - t[..., 1]) *
                         (t[..., 2] if dz else 1 - t[..., 2]))
                    out = out + w * grid[i[..., 2] + dz, i[..., 1] + dy, i[..., 0] + dx]
        return out

    def coarse_tau(self, csig):
        """Optical depth from each coarse vertex to the light (midpoint rule, half the
        smallest coarse spacing)."""
        cn = COARSE_CELLS + 1
        csp = (self.hi - self.lo) / COARSE_CELLS
        g = np.arange(cn)
        Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
        v = self.lo + csp * np.stack([X, Y, Z], -1)
        d = self.light_dir(v)
        u = -d
        with np.errstate(divide="ignore", invalid="ignore"):
            tt = np.where(u > 0, (self.hi - v) / u, np.where(u < 0, (self.lo - v) / u, np.inf))
        L = tt.min(-1)
        if self.light_kind == 0:
            L = np.minimum(L, np.linalg.norm(self.light.astype(np.float64) - v, axis=-1))
        hstep = 0.5 * csp.min()
        n = np.clip(np.ceil(L / hstep), 1, 512).astype(np.int64)
        tau = np.zeros(L.shape)
        for idx in np.ndindex(L.shape):
            if not (L[idx] > 0 and np.isfinite(L[idx])):
                continue
            dt = L[idx] / n[idx]
            ts = (np.arange(n[idx]) + 0.5) * dt
            q = v[idx] - ts[:, None] * d[idx]
            tau[idx] = (self.coarse_lerp(csig, q) * dt).sum()
        return tau

    def map_vertices(self, ids):
        mp = [c + 1 for c in self.map_cells]
        ids = np.asarray(ids, np.int64)
        ijk = np.stack([ids % mp[0], (ids // mp[0]) % mp[1], ids // (mp[0] * mp[1])], -1)
        size = (self.hi - self.lo).astype(np.float32)
        msp = (size / np.array(self.map_cells, np.float32)).astype(np.float32)
        return (self.lo.astype(np.float32) + msp * ijk.astype(np.float32)).astype(np.float64)

    def features(self, ids, ctau):
        v = self.map_vertices(ids)
        f = np.zeros((len(ids), FEAT))
        tc = np.minimum(self.coarse_lerp(ctau, v), TAU_MAX)
        f[:, 0] = tc / 4
        c = np.array([[self.nearest(d, vv[d]) for d in range(3)] for vv in v], np.int64)
        q = 1
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    i = np.clip(c[:, 0] + dx, 0, self.dims[0] - 1)
                    j = np.clip(c[:, 1] + dy, 0, self.dims[1] - 1)
                    k = np.clip(c[:, 2] + dz, 0, self.dims[2] - 1)
                    f[:, q] = self.patch(self.f3[k, j, i])
                    q += 1
        f[:, 28:31] = self.light_dir(v)
        return f, tc


def target(tc, exact_opacity):
    with np.errstate(divide="ignore"):
        te = np.minimum(-np.log(1.0 - np.asarray(exact_opacity, np.float64)), TAU_MAX)
    return tc - te


def learned_opacity(tc, y):
    return 1.0 - np.exp(-np.maximum(tc - y, 0.0))
