"""Seeded synthetic fiber models of the benchmark workloads.

``parallel_fibers``   64 straight parallel fibers along x (2-point polylines,
                      radius 1.0 um) on a jittered 8 x 8 hexagonal grid spaced
                      2.2 um, in a 30 um cube at 0.25 um voxels (120^3).
``crossing_bundles``  two bundles crossing at 90 deg in the xy-plane, 2000 fibers
                      each, radii lognormal (mean 1.0 um, sigma 0.2), 0.5 um
                      segments with a sinusoidal out-of-plane undulation
                      (amplitude 0.5 um, period 20 um, random phase), in a
                      120 x 120 x 60 um volume at 0.2 um voxels (600 x 600 x 300).
                      The two bundles interpenetrate: both fill the whole volume
                      (a crossing region throughout), centres on a jittered
                      50 x 40 grid of the bundle's cross-section.
``single_bundle``     one bundle with per-fiber directions drawn from a von
                      Mises-Fisher distribution about a random axis.

Layers of every model: [(0.75, 0, 5, 'b'), (1.0, -0.004, 1, 'p')] (myelin
model of the benchmark), unless given.
"""

import numpy as np

from . import objects

MYELIN = [(0.75, 0.0, 5.0, 'b'), (1.0, -0.004, 1.0, 'p')]


def parallel_fibers(seed=0, n_side=8, spacing=2.2, radius=1.0, size=30.0, jitter=0.05,
                    layers=MYELIN, size_z=None, n_rows=None):
    """Straight fibers along x on a jittered hexagonal grid (config 1: 8 x 8 in a
    30 um cube).  ``size_z`` / ``n_rows`` give a taller stack of rows."""
    rng = np.random.default_rng(seed)
    pitch = spacing * np.sqrt(3.0) / 2.0
    size_z = size if size_z is None else size_z
    n_rows = n_side if n_rows is None else n_rows
    ys, zs = [], []
    for row in range(n_rows):
        for col in range(n_side):
            ys.append(col * spacing + (spacing / 2.0 if row % 2 else 0.0))
            zs.append(row * pitch)
    ys = np.array(ys)
    zs = np.array(zs)
    ys += size / 2.0 - (ys.max() + ys.min()) / 2.0
    zs += size_z / 2.0 - (zs.max() + zs.min()) / 2.0
    fibers = []
    for y, z in zip(ys, zs):
        j = rng.uniform(-jitter, jitter, size=(2, 3))
        fibers.append(
            np.array([[-1.0 + j[0, 0], y + j[0, 1], z + j[0, 2], radius],
                      [size + 1.0 + j[1, 0], y + j[1, 1], z + j[1, 2], radius]]))
    return objects.FiberBundles([fibers], [list(layers)])


def _undulating_bundle(rng, axis, n_u, n_v, extent_u, extent_v, length, seg, amp, period,
                       sigma, jitter, layers):
    """Fibers along `axis` (0 = x, 1 = y), undulating in z, centres on a jittered grid."""
    t = np.arange(-2.0, length + 2.0 + 1e-9, seg)
    fibers = []
    du, dv = extent_u / n_u, extent_v / n_v
    for a in range(n_u):
        for b in range(n_v):
            cu = (a + 0.5) * du + rng.uniform(-jitter, jitter)
            cv = (b + 0.5) * dv + rng.uniform(-jitter, jitter)
            phase = rng.uniform(0, 2 * np.pi)
            r = np.exp(rng.normal(np.log(1.0) - sigma**2 / 2.0, sigma))
            pts = np.empty((t.size, 4))
            pts[:, axis] = t
            pts[:, 1 - axis] = cu
            pts[:, 2] = cv + amp * np.sin(2 * np.pi * t / period + phase)
            pts[:, 3] = r
            fibers.append(pts)
    return fibers


def crossing_bundles(seed=0, n_fibers=2000, size_xy=120.0, size_z=60.0, seg=0.5, amp=0.5,
                     period=20.0, sigma=0.2, jitter=0.2, layers=MYELIN):
    """Two interpenetrating bundles along x and y (config 2)."""
    rng = np.random.default_rng(seed)
    n_u = int(round(np.sqrt(n_fibers * size_xy / size_z)))
    n_v = int(np.ceil(n_fibers / n_u))
    a = _undulating_bundle(rng, 0, n_u, n_v, size_xy, size_z, size_xy, seg, amp, period, sigma,
                           jitter, layers)[:n_fibers]
    b = _undulating_bundle(rng, 1, n_u, n_v, size_xy, size_z, size_xy, seg, amp, period, sigma,
                           jitter, layers)[:n_fibers]
    return objects.FiberBundles([a, b], [list(layers), list(layers)])


def single_bundle(seed=0, n_fibers=400, size=(60.0, 60.0, 30.0), kappa=10.0, seg=0.5,
                  radius_sigma=0.2, layers=MYELIN, axis=None):
    """One bundle of straight fibers with vMF-distributed directions about a
    seeded random axis (or the given unit `axis`; the random draws are the same)."""
    rng = np.random.default_rng(seed)
    drawn = rng.normal(size=3)
    axis = drawn if axis is None else np.asarray(axis, dtype=float)
    axis = axis / np.linalg.norm(axis)
    size = np.asarray(size, dtype=float)
    centre = size / 2.0
    fibers = []
    for _ in range(n_fibers):
        d = _vmf_sample(rng, axis, kappa)
        c = centre + (rng.uniform(-0.5, 0.5, size=3) * size)
        half = np.linalg.norm(size)
        t = np.arange(-half, half + 1e-9, seg)
        pts = np.empty((t.size, 4))
        pts[:, :3] = c[None, :] + t[:, None] * d[None, :]
        pts[:, 3] = np.exp(rng.normal(-radius_sigma**2 / 2.0, radius_sigma))
        fibers.append(pts)
    return objects.FiberBundles([fibers], [list(layers)])


def _vmf_sample(rng, mu, kappa):
    """One unit vector from a von Mises-Fisher distribution on S^2 (Wood 1994)."""
    u = rng.uniform()
    w = 1.0 + np.log(u + (1.0 - u) * np.exp(-2.0 * kappa)) / kappa
    v = rng.normal(size=3)
    v -= v.dot(mu) * mu
    v /= np.linalg.norm(v)
    return w * mu + np.sqrt(max(0.0, 1.0 - w * w)) * v
