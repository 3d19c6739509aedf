"""Fiber models in text files (.dat / .txt).

The format is the reference's (src/fastpli/io/fiber_bundles.py:43-78,
save_dat :140-160): one point per line ``x y z r`` (further numbers on a line
are ignored); an empty line ends a fiber, two or more end a fiber bundle (an
empty bundle is not started twice).  The HDF5 variant needs h5py, which this
package does not use.

The reader parses all point lines with one numpy call and finds fibers and
bundles from the runs of empty lines, so large models (e.g. the reference's
optic_chiasm.dat, 6.6k lines) load in milliseconds.  The writer prints every
number with 17 significant digits, so a save / load round trip is exact.
"""

import os

import numpy as np

from .. import objects


def load(file_name, group_name='/'):
    """Load fiber bundles from a .dat / .txt file."""
    _, ext = os.path.splitext(file_name)
    if ext in ('.dat', '.txt'):
        with open(file_name, 'r') as f:
            return load_dat(f)
    if ext == '.h5':
        raise TypeError('.h5 fiber models need h5py, which is not available')
    raise TypeError(ext + ' is not implemented yet')


def _points(lines):
    """(n, 4) float64 of the first four numbers of every line."""
    if not lines:
        return np.zeros((0, 4))
    flat = np.array(' '.join(lines).split(), dtype=np.float64)
    if flat.size == 4 * len(lines):
        return flat.reshape(-1, 4)
    # lines of other lengths: per line
    return np.array([[float(v) for v in ln.split()[:4]] for ln in lines], dtype=np.float64)


def load_dat(file):
    """Parse an opened text file into :class:`FiberBundles`."""
    lines = file.read().splitlines()
    blank = np.fromiter((not ln.strip() for ln in lines), dtype=bool, count=len(lines))
    pts = _points([ln for ln, b in zip(lines, blank) if not b])
    # runs of equal `blank`: starts, lengths and kinds
    edges = np.flatnonzero(np.diff(blank.astype(np.int8))) + 1
    starts = np.concatenate([[0], edges]) if len(lines) else np.zeros(0, int)
    lengths = np.diff(np.concatenate([starts, [len(lines)]]))
    bundles = [[]]
    row = 0  # next row of pts
    for s, n in zip(starts, lengths):
        if not blank[s]:
            bundles[-1].append(pts[row:row + n])
            row += n
        elif n >= 2 and bundles[-1]:
            bundles.append([])
    return objects.FiberBundles(bundles)


def save(file_name, fiber_bundles, group_name='/', mode='w-'):
    """Write fiber bundles to a .dat / .txt file (mode 'w-': the file must not exist)."""
    _, ext = os.path.splitext(file_name)
    if ext not in ('.dat', '.txt'):
        raise TypeError(ext + ' is not implemented yet')
    with open(file_name, 'x' if mode == 'w-' else mode) as f:
        save_dat(f, fiber_bundles)


def save_dat(file, fiber_bundles):
    """Write fiber bundles as text in the format :func:`load_dat` reads."""
    fiber_bundles = objects.FiberBundles(fiber_bundles)
    for k, fb in enumerate(fiber_bundles):
        for fiber in fb:
            np.savetxt(file, np.asarray(fiber[:, :4]), fmt='%.17g')
            file.write('\n')
        if k != len(fiber_bundles) - 1:
            file.write('\n')
