"""The reference-side binding: the pybind11 modules' classes over the C ABI.

fastpli's ``Simpli`` reaches its native core through two classes,
``_Generator`` (src/simulation/bindings/generator_module.cpp:26-139) and
``_Simulator`` (src/simulation/bindings/simulator_module.cpp:21-94), imported
at src/fastpli/simulation/simpli.py:11-12.  This module provides both with the
same methods, arguments and return values, backed by libfastpli_amd.so, so a
maintainer can keep the reference's own ``Simpli`` and swap only that import.
Arrays cross as numpy, as in the reference (the device copies are the
shim's own).

``install_alias()`` registers this package under the reference's module paths
(``fastpli``, ``fastpli.simulation``, ``fastpli.objects``, ``fastpli.io``), for
user code that imports them by name; it refuses when a real ``fastpli`` is
already imported.
"""

import sys

import numpy as np

from . import _lib


def _torch():
    import torch
    return torch


class _Generator:
    """generator_module.cpp:26-139 over fpl_generate."""

    def __init__(self):
        self._vol = None
        self._fbs = None
        self._layers = None

    def set_volume(self, global_dim, origin, voxel_size):
        self._vol = ([int(d) for d in global_dim], [float(o) for o in origin], float(voxel_size))

    def set_fiber_bundles(self, fbs, prs):
        """fbs: [[(n, 4) arrays]] per bundle; prs: [[(scale, dn, mu, model)]] per bundle."""
        if len(fbs) != len(prs):
            raise ValueError('fbs and prs not the same size')  # generator_module.cpp:45-46
        self._fbs = fbs
        self._layers = [[tuple(ly) for ly in lys] for lys in prs]

    def set_omp_num_threads(self, num_threads):
        return int(num_threads)

    def dim_local(self):
        return list(self._vol[0])

    def dim_offset(self):
        return [0, 0, 0]

    def run_generation(self, only_label=False):
        from .objects import FiberBundles
        torch = _torch()
        dim, origin, voxel_size = self._vol
        pts, fptr, bptr, lptr, table = FiberBundles(self._fbs, self._layers).to_flat()
        dev = torch.device('cuda', torch.cuda.current_device())
        label = torch.empty(dim, dtype=torch.int32, device=dev)
        vector = None if only_label else torch.empty(dim + [3], dtype=torch.float32, device=dev)
        props, warn = _lib.generate(dim, origin, voxel_size, torch.from_numpy(pts).to(dev),
                                    torch.from_numpy(fptr).to(dev), bptr, lptr, table,
                                    only_label, label, vector)
        _lib.raise_warnings(warn)
        return (label.cpu().numpy(),
                np.empty((0,), np.float32) if vector is None else vector.cpu().numpy(), props)


class _Simulator:
    """simulator_module.cpp:21-94 over fpl_simulate."""

    def __init__(self):
        self._setup = None

    def set_pli_setup(self, step_size, light_intensity, voxel_size, wavelength,
                      tissue_refrection, interpolate, untilt_sensor_view, flip_z,
                      filter_rotations):
        self._setup = _lib.make_setup(step_size, light_intensity, voxel_size, wavelength,
                                      tissue_refrection, interpolate, untilt_sensor_view, flip_z,
                                      filter_rotations)

    def set_omp_num_threads(self, num_threads):
        return int(num_threads)

    def run_simulation(self, dim, label, vector, props, theta=0, phi=0):
        torch = _torch()
        setup, _keep = self._setup
        dim = [int(d) for d in dim]
        dev = torch.device('cuda', torch.cuda.current_device())
        lab = torch.from_numpy(np.ascontiguousarray(label, dtype=np.int32)).to(dev)
        vec = torch.from_numpy(np.ascontiguousarray(vector, dtype=np.float32)).to(dev)
        images = torch.empty((1, dim[0], dim[1], setup.n_rho), dtype=torch.float64, device=dev)
        _lib.raise_warnings(_lib.simulate(setup, dim, lab, vec, props, [[theta, phi]], images))
        return images[0].cpu().numpy()


def install_alias():
    """Registers this package as ``fastpli`` (with .simulation, .objects, .io)."""
    real = sys.modules.get('fastpli')
    pkg = sys.modules[__name__.rsplit('.', 1)[0]]
    if real is not None and real is not pkg:
        raise RuntimeError('a module named fastpli is imported already')
    from . import io, objects, simulation
    sys.modules['fastpli'] = pkg
    sys.modules['fastpli.simulation'] = simulation
    sys.modules['fastpli.objects'] = objects
    sys.modules['fastpli.io'] = io
    return pkg
