"""Brute-force numpy voxelizer used to pin the oracle on tiny scenes.

Independent restatement of the per-voxel rule of fastpli's
FillVoxelsAroundFiberSegment (src/simulation/generator.cpp:321-378) that
tests EVERY voxel against EVERY segment (no bounding boxes, no culling,
no tiling), in global (bundle, fiber, segment) order, with numpy's IEEE
elementwise arithmetic (one rounding per operation, no FMA):

    closest point on the segment (FP64), dist^2 rounded to float32,
    r(t)^2 rounded to float32, reject if dist^2 > float32(scale_last^2) * r^2,
    update if stored >= dist^2, layer = searchsorted(scale^2, float32 ratio).

Fibers that the reference culls by their AABB can never reach a voxel of the
volume, and every voxel a segment can reach lies inside its integer bbox, so
this gives the same labels as the bbox-limited reference.  It does not model
the translate-by-origin step (scenes use origin 0) nor vector rounding of
'r' layers beyond what unit() does in FP64.
"""

import numpy as np


def voxelize(fiber_bundles, dim, voxel_size):
    nx, ny, nz = dim
    gx, gy, gz = np.meshgrid(np.arange(nx) + 0.5, np.arange(ny) + 0.5, np.arange(nz) + 0.5,
                             indexing='ij')
    pts = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)
    n = pts.shape[0]
    dist = np.full(n, np.inf, dtype=np.float32)
    label = np.zeros(n, dtype=np.int32)
    vec = np.zeros((n, 3), dtype=np.float32)
    layers = [sorted([tuple(ly) for ly in fb.layers], key=lambda l: l[0])
              for fb in fiber_bundles]
    max_layer = max(len(l) for l in layers)
    f = 1.0 / voxel_size
    for b, fb in enumerate(fiber_bundles):
        scale_sq = np.array([l[0] * l[0] for l in layers[b]], dtype=np.float64)
        f_last = np.float32(scale_sq[-1])
        for fiber in fb:
            data = np.asarray(fiber, dtype=np.float64)
            P = data[:, :3] * f
            R = data[:, 3] * f
            for s in range(len(P) - 1):
                p, q = P[s], P[s + 1]
                v = q - p
                w = pts - p
                c1 = ((0.0 + w[:, 0] * v[0]) + w[:, 1] * v[1]) + w[:, 2] * v[2]
                c2 = ((0.0 + v[0] * v[0]) + v[1] * v[1]) + v[2] * v[2]
                t = np.where(c1 < 0, 0.0, np.where(c2 <= c1, 1.0, c1 / c2))
                inner = (c1 >= 0) & (c2 > c1)
                cp = np.where((c1 < 0)[:, None], p[None, :],
                              np.where((c2 <= c1)[:, None], q[None, :],
                                       p[None, :] + v[None, :] * t[:, None]))
                cp = np.where(inner[:, None], p[None, :] + v[None, :] * t[:, None], cp)
                d = cp - pts
                d2 = (((0.0 + d[:, 0] * d[:, 0]) + d[:, 1] * d[:, 1]) +
                      d[:, 2] * d[:, 2]).astype(np.float32)
                rr = R[s] * (1 - t) + R[s + 1] * t
                r2 = (rr * rr).astype(np.float32)
                ok = ~(d2 > f_last * r2) & (dist >= d2)
                if not np.any(ok):
                    continue
                idx = np.nonzero(ok)[0]
                ratio = (d2[idx] / r2[idx]).astype(np.float64)
                ly = np.searchsorted(scale_sq, ratio, side='left')
                dist[idx] = d2[idx]
                label[idx] = ly + 1 + b * max_layer
                lyo = np.minimum(ly, len(scale_sq) - 1)
                for j, l in zip(idx, lyo):
                    model = layers[b][l][3]
                    if model == 'p':
                        u = v
                    elif model == 'r':
                        u = pts[j] - cp[j]
                    else:
                        vec[j] = 0
                        continue
                    n2 = ((0.0 + u[0] * u[0]) + u[1] * u[1]) + u[2] * u[2]
                    if n2 > 0:
                        u = u * (1.0 / np.sqrt(n2))
                    vec[j] = u.astype(np.float32)
    return label.reshape(nx, ny, nz), vec.reshape(nx, ny, nz, 3)


def random_scene(seed, n_bundles=2, n_fibers=3, n_points=4, size=6.0):
    """Tiny random scene of crossing polylines (coordinates in um, voxel 1 um)."""
    from fastpli_amd.objects import FiberBundles
    rng = np.random.default_rng(seed)
    data, layers = [], []
    models = ['p', 'r', 'b']
    for b in range(n_bundles):
        fibers = []
        for _ in range(n_fibers):
            start = rng.uniform(-1, size + 1, size=3)
            steps = rng.normal(scale=1.5, size=(n_points - 1, 3))
            pts = np.vstack([start, start + np.cumsum(steps, axis=0)])
            r = rng.uniform(0.4, 1.6, size=(n_points, 1))
            fibers.append(np.hstack([pts, r]))
        data.append(fibers)
        n_l = int(rng.integers(1, 4))
        scales = np.sort(rng.uniform(0.2, 1.0, size=n_l))
        scales[-1] = 1.0 if rng.uniform() < 0.5 else scales[-1]
        scales = np.unique(scales)
        lys = []
        for s in scales:
            m = models[int(rng.integers(0, 3))]
            dn = 0.0 if m == 'b' else (-0.004 if m == 'p' else 0.004)
            lys.append((float(s), dn, float(rng.uniform(0, 10)), m))
        layers.append(lys)
    return FiberBundles(data, layers)
