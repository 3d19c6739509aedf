"""Text (.dat / .txt) fiber model loader and saver.

Format, as read by fastpli's ``src/fastpli/io/fiber_bundles.py:43-78``: one
point per line ``x y z r``; one empty line ends a fiber, two consecutive empty
lines end a fiber bundle.  The HDF5 variant of the reference needs h5py and
is not provided.
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
    raise TypeError(ext + ' is not implemented yet')


def load_dat(file):
    """Parse an opened text file into :class:`FiberBundles`."""
    fiber = []
    bundles = [[]]
    empty_lines = 0
    for line in file:
        if line.strip():
            empty_lines = 0
            numbers = [float(s) for s in line.split()]
            fiber.append(numbers[0:4])
        else:
            empty_lines += 1
            if empty_lines == 1:
                if fiber:
                    bundles[-1].append(np.array(fiber, float))
                    fiber = []
            elif empty_lines == 2:
                if bundles[-1]:
                    bundles.append([])
    if fiber:
        bundles[-1].append(np.array(fiber, float))
    return objects.FiberBundles(bundles)


def save(file_name, fiber_bundles, mode='w'):
    """Write fiber bundles as text in the format :func:`load_dat` reads."""
    with open(file_name, mode) as f:
        for fb in fiber_bundles:
            for fiber in fb:
                for p in np.asarray(fiber):
                    f.write(f'{p[0]:.6f} {p[1]:.6f} {p[2]:.6f} {p[3]:.6f}\n')
                f.write('\n')
            f.write('\n')
