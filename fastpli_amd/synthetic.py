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
``serving_model``     a request of the serving stream: 1-3 bundles of straight
                      fibers, each bundle along a random direction (azimuth
                      uniform, inclination within +-30 deg), radii lognormal
                      (mean 1.0 um, sigma 0.2), its 'p' layer's dn uniform in
                      [-0.006, -0.001] and both layers' mu uniform in [1, 10].

Layers of every model: [(0.75, 0, 5, 'b'), (1.0, -0.004, 1, 'p')] (myelin
model of the benchmark), unless given.
"""

import numpy as np

from . import objects

MYELIN = [(0.75, 0.0, 5.0, 'b'), (1.0, -0.004, 1.0, 'p')]


def parallel_fibers(seed=0, n_side=8, spacing=2.2, radius=1.0, size=30.0, jitter=0.05,
                    layers=MYELIN, size_z=None, n_rows=None):
    """Straight fibers along x on a jittered hexagonal grid (default: 8 x 8 in a
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
    """Two interpenetrating bundles along x and y (default: 2 x 2000 undulating
    fibers in 120 x 120 x 60 um)."""
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


SERVE_SIZES = ((64, 64, 60), (128, 128, 120), (256, 256, 200), (512, 512, 300))


def serving_model(seed, dim, voxel_size=0.2, density=0.28, seg=0.5, sigma=0.2):
    """Fiber model of one serving request in a volume of `dim` voxels.

    Every bundle gets ``density`` fibers per um^2 of the volume's cross-section
    normal to its direction (the crossing bundles' 2000 fibers over
    120 x 60 um^2), centres uniform in the volume, straight polylines with
    `seg` um segments reaching 2 um beyond the volume.  Returns FiberBundles
    with their layers [(0.75, 0, mu_b, 'b'), (1.0, dn, mu_p, 'p')].
    """
    rng = np.random.default_rng(seed)
    size = np.asarray(dim, dtype=float) * voxel_size
    n_bundles = int(rng.integers(1, 4))
    data, layers = [], []
    half = 0.5 * np.linalg.norm(size) + 2.0
    t = np.arange(-half, half + 1e-9, seg)
    for _ in range(n_bundles):
        az = rng.uniform(0, 2 * np.pi)
        inc = np.deg2rad(rng.uniform(-30.0, 30.0))
        d = np.array([np.cos(inc) * np.cos(az), np.cos(inc) * np.sin(az), np.sin(inc)])
        # area of the box's projection along d
        area = abs(d[0]) * size[1] * size[2] + abs(d[1]) * size[0] * size[2] + \
            abs(d[2]) * size[0] * size[1]
        n_f = max(1, int(round(density * area)))
        c = rng.uniform(0, 1, size=(n_f, 3)) * size
        r = np.exp(rng.normal(np.log(1.0) - sigma**2 / 2.0, sigma, size=n_f))
        pts = c[:, None, :] + t[None, :, None] * d[None, None, :]
        # the part of every fiber within 2 um of the volume (a convex region:
        # one run of points per fiber)
        inside = np.all((pts > -2.0) & (pts < size + 2.0), axis=2)
        first = np.argmax(inside, axis=1)
        last = t.size - 1 - np.argmax(inside[:, ::-1], axis=1)
        fibers = []
        for k in range(n_f):
            if inside[k, first[k]] and last[k] > first[k]:
                f = np.empty((last[k] + 1 - first[k], 4))
                f[:, :3] = pts[k, first[k]:last[k] + 1]
                f[:, 3] = r[k]
                fibers.append(f)
        dn = rng.uniform(-0.006, -0.001)
        mu_b, mu_p = rng.uniform(1.0, 10.0, size=2)
        data.append(fibers)
        layers.append([(0.75, 0.0, float(mu_b), 'b'), (1.0, float(dn), float(mu_p), 'p')])
    return objects.FiberBundles(data, layers)
