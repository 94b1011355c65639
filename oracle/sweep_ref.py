"""Test oracle — NOT product code.

FP64 numpy restatement of the exact-sweep transmittance map (the product's
BEAMS_SHADOW_SWEEP).  It has no counterpart in the reference: the sweep replaces the
reference's per-vertex light march (TransmittanceMap.h:239-340) by one pass over
light-aligned slices, with the extinction sigma = -ln(1 - alpha/alpha_max) / h that makes
one reference step h = |size|/128 composite alpha/alpha_max, and the map opacity
1 - exp(-(tau + sigma h / 2)).
"""

import numpy as np


def _field_at(field3d, coords, p):
    """Trilinear field values at points p (n, 3) inside the grid (x fastest storage)."""
    out_idx = []
    ts = []
    for d in range(3):
        c = coords[d]
        i = np.searchsorted(c, p[:, d], side="right") - 1
        i = np.clip(i, 0, len(c) - 2)
        t = np.clip((p[:, d] - c[i]) / (c[i + 1] - c[i]), 0.0, 1.0)
        out_idx.append(i)
        ts.append(t)
    i, j, k = out_idx
    tx, ty, tz = ts
    f = field3d  # [z, y, x]

    def g(dz, dy, dx):
        return f[k + dz, j + dy, i + dx]

    x0 = g(0, 0, 0) + tx * (g(0, 0, 1) - g(0, 0, 0))
    x1 = g(0, 1, 0) + tx * (g(0, 1, 1) - g(0, 1, 0))
    x2 = g(1, 0, 0) + tx * (g(1, 0, 1) - g(1, 0, 0))
    x3 = g(1, 1, 0) + tx * (g(1, 1, 1) - g(1, 1, 0))
    y0 = x0 + ty * (x1 - x0)
    y1 = x2 + ty * (x3 - x2)
    return y0 + tz * (y1 - y0)


def sweep_axis(bounds, light_kind, light):
    light = np.asarray(light, np.float64)
    if light_kind == 1:
        axis = int(np.argmax(np.abs(light)))
        step = 1 if light[axis] > 0 else -1
    else:
        outs = [max(bounds[2 * d] - light[d], light[d] - bounds[2 * d + 1]) for d in range(3)]
        axis = int(np.argmax(outs))
        if outs[axis] <= 0:
            raise ValueError("the sweep needs the light outside the bounds")
        step = 1 if light[axis] < bounds[2 * axis] else -1
    a1 = 1 if axis == 0 else 0
    a2 = 1 if axis == 2 else 2
    return axis, a1, a2, step


def sweep_map(field, dims, coords, tf, scalar_range, light_kind, light, map_cells):
    """Opacity map (x fastest, (cells+1)^3) of the exact sweep, in FP64."""
    nx, ny, nz = dims
    f3 = np.asarray(field, np.float64).reshape(nz, ny, nx)
    coords = [np.asarray(c, np.float64) for c in coords]
    bounds = [coords[0][0], coords[0][-1], coords[1][0], coords[1][-1], coords[2][0], coords[2][-1]]
    tf = np.asarray(tf, np.float64).reshape(-1, 4)
    ntf = tf.shape[0]
    amax = tf[:, 3].max()
    a = np.minimum(tf[:, 3] / amax, 1.0 - 2.0 ** -24)
    size = np.array([bounds[1] - bounds[0], bounds[3] - bounds[2], bounds[5] - bounds[4]])
    h = np.linalg.norm(size) / 128.0
    sig_tab = -np.log1p(-a) / h
    mp = [c + 1 for c in map_cells]
    org = np.array([bounds[0], bounds[2], bounds[4]])
    sp = size / np.array(map_cells, np.float64)
    light = np.asarray(light, np.float64)
    if light_kind == 1:
        light = light / np.linalg.norm(light)
    axis, a1, a2, step = sweep_axis(bounds, light_kind, light)
    n1, n2 = mp[a1], mp[a2]
    I1, I2 = np.meshgrid(np.arange(n1), np.arange(n2), indexing="xy")  # [i2, i1]
    out = np.zeros((mp[2], mp[1], mp[0]))
    smin, smax = scalar_range
    inv_delta = 1.0 / (smax - smin) if smax != smin else smin
    prev = None
    order = range(mp[axis]) if step > 0 else range(mp[axis] - 1, -1, -1)
    lo = org
    hi = org + sp * (np.array(mp) - 1)
    for s, k in enumerate(order):
        v = np.zeros((n2, n1, 3))
        v[..., axis] = org[axis] + sp[axis] * k
        v[..., a1] = org[a1] + sp[a1] * I1
        v[..., a2] = org[a2] + sp[a2] * I2
        fv = _field_at(f3, coords, v.reshape(-1, 3)).reshape(n2, n1)
        idx = np.clip(((fv - smin) * inv_delta * (ntf - 1)).astype(np.int64), 0, ntf - 1)
        sig = sig_tab[idx]
        if s == 0:
            tau = np.zeros_like(sig)
        else:
            if light_kind == 1:
                d = np.broadcast_to(light, v.shape)
            else:
                d = v - light
                d = d / np.linalg.norm(d, axis=-1, keepdims=True)
            tb = sp[axis] / np.abs(d[..., axis])
            u1 = I1 - tb * d[..., a1] / sp[a1]
            u2 = I2 - tb * d[..., a2] / sp[a2]
            inside = (u1 >= 0) & (u2 >= 0) & (u1 <= n1 - 1) & (u2 <= n2 - 1)
            c1 = np.clip(np.floor(np.clip(u1, 0, n1 - 1)).astype(np.int64), 0, n1 - 2)
            c2 = np.clip(np.floor(np.clip(u2, 0, n2 - 1)).astype(np.int64), 0, n2 - 2)
            f1 = np.clip(u1, 0, n1 - 1) - c1
            f2 = np.clip(u2, 0, n2 - 1) - c2

            def bil(q):
                a0 = q[c2, c1] + f1 * (q[c2, c1 + 1] - q[c2, c1])
                b0 = q[c2 + 1, c1] + f1 * (q[c2 + 1, c1 + 1] - q[c2 + 1, c1])
                return a0 + f2 * (b0 - a0)

            tau_in = bil(prev[0]) + 0.5 * (bil(prev[1]) + sig) * tb
            te = tb.copy()
            for a in (a1, a2):
                da = d[..., a]
                va = v[..., a]
                with np.errstate(divide="ignore", invalid="ignore"):
                    te = np.where(da > 0, np.minimum(te, (va - lo[a]) / da), te)
                    te = np.where(da < 0, np.minimum(te, (hi[a] - va) / -da), te)
            tau = np.where(inside, tau_in, sig * np.maximum(te, 0.0))
        prev = (tau, sig)
        op = 1.0 - np.exp(-(tau + 0.5 * h * sig))
        # scatter the slice into the x-fastest map
        sl = [slice(None)] * 3
        sl[2 - axis] = k
        dst = out[tuple(sl)]  # remaining axes in (z, y, x) order without `axis`
        # dst axes are the two transverse axes in descending axis order: (a2, a1)
        dst[...] = op if a2 > a1 else op.T
    return out.ravel()
