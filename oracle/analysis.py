"""CPU oracle of the image steps after the ray march -- TEST INFRASTRUCTURE.

* ``filter_resample``: fastpli's ``optic.filter_resample``
  (src/fastpli/simulation/optic.py:36-55, 85-120, 169-187) -- the reference's
  own call ``scipy.ndimage.gaussian_filter(image, sigma / scale)`` followed by a
  restatement of its numba ``resample`` loop (same accumulation order).
* ``epa``: fastpli's ``analysis.epa.epa`` (src/fastpli/analysis/epa.py:9-45),
  restated in numpy.

Pinned by tests/test_oracle_analysis.py: with numpy's legacy RNG seeded 42 and
the reference's noise model, the oracle reproduces the golden
``/simulation/optic/{0..4}`` of tests/simulation/simpli_rep.h5 bit for bit,
and ``epa`` of those images reproduces ``/analysis/epa/*``.
"""

import numpy as np
import scipy.ndimage


def resample(image, scale):
    """Box resampling to round(n * scale) pixels (optic.py:85-120)."""
    sx = int(round(image.shape[0] * scale))
    sy = int(round(image.shape[1] * scale))
    data = np.zeros((sx, sy), dtype=image.dtype)
    divisor = np.zeros((sx, sy), dtype=np.int64)
    ii = np.floor((np.arange(image.shape[0]) + 0.5) * scale).astype(np.int64)
    jj = np.floor((np.arange(image.shape[1]) + 0.5) * scale).astype(np.int64)
    for i in range(image.shape[0]):
        if ii[i] >= sx:
            continue
        for j in range(image.shape[1]):
            if jj[j] >= sy:
                continue
            divisor[ii[i], jj[j]] += 1
            data[ii[i], jj[j]] += image[i, j]
    with np.errstate(invalid='ignore', divide='ignore'):
        return np.divide(data, divisor)


def filter_resample(image, sigma, scale):
    """Gaussian filter (sigma / scale) then resample (optic.py:169-187)."""
    if sigma != 0:
        image = scipy.ndimage.gaussian_filter(image, sigma / scale)
    return resample(image, scale)


def optic_resample(images, optical_sigma, voxel_size, pixel_size):
    """Simpli.apply_optic_resample for an (nx, ny, n_rho) stack (__simpli.py:491-549)."""
    images = np.atleast_3d(np.asarray(images, dtype=np.float64))
    scale = voxel_size / pixel_size
    out = [filter_resample(images[:, :, k], optical_sigma, scale)
           for k in range(images.shape[2])]
    return np.squeeze(np.stack(out, axis=-1))


def epa(data):
    """Transmittance, direction, retardation (epa.py:9-45)."""
    data = np.asarray(data)
    dtype = np.float32 if data.itemsize <= 4 else np.float64
    data = data.astype(dtype)
    n = data.shape[-1]
    rho_2 = 2 * np.linspace(0, np.pi, n, False, dtype=dtype)
    a0 = np.sum(data, -1) / n
    a1 = 2 * np.sum(data * np.sin(rho_2), -1) / n
    b1 = 2 * np.sum(data * np.cos(rho_2), -1) / n
    t = 2 * a0
    d = 0.5 * np.arctan2(-b1, a1) + np.pi
    a0[a0 < np.finfo(dtype).eps * 100] = 1
    r = np.sqrt(a1 * a1 + b1 * b1) / a0
    d = d % np.pi
    return t, d, r
