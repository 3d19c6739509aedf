"""FP64 CPU oracle of the fastpli simulation hot path -- TEST INFRASTRUCTURE.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module, and only as the checker / the timed CPU
baseline.  The product package ``fastpli_amd`` never imports it.

``liboracle.so`` is built from ``fpl_oracle.cpp`` (``make -C oracle``), a plain
C++17 + OpenMP restatement of fastpli's ``PliGenerator::RunTissueGeneration``
(src/simulation/generator.cpp:165-383,475-526) and
``PliSimulator::RunSimulation`` (src/simulation/simulator.cpp:182-773) in the
reference's own precision, compiled without FMA contraction.

Parity pins (see tests/test_oracle_*.py):
  * generate_tissue: bit-exact labels and vectors against the reference's
    golden ``simpli_rep.h5`` (cube.dat, 24^3); closed-form voxel tests of
    ``tests/simulation/simpli_test.py:65-218``; brute force.
  * run_simulation: golden images of all 5 tilts (Slerp) at 1e-5 relative;
    closed forms for uniform columns; rho-rotation identity.
  * Lerp, flip_z, step_size != 1, tilted sensor (untilt_sensor_view=False)
    and refraction have no reference vectors; they are pinned by closed
    forms, invariants, a library routine (Lerp on same-hemisphere fields =
    scipy's trilinear ``RegularGridInterpolator``) and geometry (each tilted
    entry point, moved up the light direction and turned by scipy's ZYZ
    rotation, lies on its pixel's vertical line): ``light_starts`` and
    ``get_vec`` expose those two pieces of the march.

``oracle.adjoint`` (torch float64) restates the NN march differentiably: the
FP64 reference gradient of the adjoint kernel, its forward pinned to
``simulate`` here and its gradient to central differences of itself.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

WARN_OVERLAP = 1 << 0
WARN_SINGLE_POINT = 1 << 1
WARN_ZERO_VEC = 1 << 2
WARN_NOT_NORM = 1 << 3
WARN_NO_LIGHT = 1 << 4

INTERP = {'NN': 0, 'Lerp': 1, 'Slerp': 2}

_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_d = ctypes.c_double
_int = ctypes.c_int


def build():
    """Compile liboracle.so with make (g++, no FMA contraction)."""
    import subprocess
    subprocess.run(['make', '-s', '-C', _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, 'liboracle.so')
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.oracle_generate.restype = _int
        L.oracle_generate.argtypes = [
            _i64, _i64, _i64, _p, _d, _i64, _p, _p, _p, _p, _p, _p, _p, _p,
            _int, _int, _p, _p, _p, _p
        ]
        L.oracle_simulate.restype = _int
        L.oracle_simulate.argtypes = [
            _i64, _i64, _i64, _p, _p, _p, _i64, _d, _d, _d, _d, _d, _int, _int,
            _int, _p, _i64, _d, _d, _int, _p, _p
        ]
        L.oracle_count_steps.restype = _i64
        L.oracle_count_steps.argtypes = [_i64, _i64, _i64, _d, _d, _d]
        L.oracle_light_starts.restype = None
        L.oracle_light_starts.argtypes = [_i64, _i64, _i64, _d, _d, _d, _int, _p]
        L.oracle_get_vec.restype = None
        L.oracle_get_vec.argtypes = [_i64, _i64, _i64, _p, _p, _int, _int, _d, _d, _d, _p]
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def generate(fiber_bundles, dim, origin, voxel_size, only_label=False,
             n_threads=1, out=None):
    """Voxelize ``FiberBundles`` -> (label int32, vector f32 or (0,), props f64, warn).

    ``fiber_bundles`` may also be the tuple ``FiberBundles.to_flat()`` returns;
    ``out`` = (label, vector) C-contiguous arrays to write into (the oracle
    clears them itself).  ``props[0, 1]`` is left 0 (the Simpli layer sets the
    background absorption).
    """
    flat = fiber_bundles if isinstance(fiber_bundles, tuple) else fiber_bundles.to_flat()
    points, fiber_ptr, bundle_ptr, layer_ptr, table = flat
    scale, dn, mu, model = table
    n_bundles = len(bundle_ptr) - 1
    max_layer = int(np.max(np.diff(layer_ptr))) if n_bundles else 0
    n_props = n_bundles * max_layer + 1
    nx, ny, nz = (int(d) for d in dim)
    if out is not None:
        label, vector = out
        assert label.shape == (nx, ny, nz) and label.dtype == np.int32
        assert label.flags.c_contiguous and (only_label or vector.flags.c_contiguous)
        assert only_label or (vector.shape == (nx, ny, nz, 3) and vector.dtype == np.float32)
    else:
        label = np.zeros((nx, ny, nz), dtype=np.int32)
        vector = (np.zeros((0,), dtype=np.float32) if only_label else np.zeros(
            (nx, ny, nz, 3), dtype=np.float32))
    props = np.zeros((n_props, 2), dtype=np.float64)
    warn = np.zeros(1, dtype=np.uint32)
    origin = np.ascontiguousarray(origin, dtype=np.float64)
    orient = np.ascontiguousarray(model.astype(np.uint8))
    rc = lib().oracle_generate(nx, ny, nz, _ptr(origin), float(voxel_size),
                               n_bundles, _ptr(bundle_ptr), _ptr(fiber_ptr),
                               _ptr(points), _ptr(layer_ptr), _ptr(scale),
                               _ptr(dn), _ptr(mu), _ptr(orient),
                               int(bool(only_label)), int(n_threads),
                               _ptr(label), _ptr(vector) if not only_label else None,
                               _ptr(props), _ptr(warn))
    if rc != 0:
        raise ValueError(f'oracle_generate failed: {rc}')
    return label, vector, props, int(warn[0])


def simulate(label, vector, props, theta, phi, *, step_size=1.0,
             light_intensity=26000.0, voxel_size=1.0, wavelength=525.0,
             tissue_refrection=1.0, interpolate='Slerp',
             untilt_sensor_view=True, flip_z=False, filter_rotations=None,
             n_threads=1, out=None):
    """March one tilt -> (images f64 (nx, ny, n_rho), warn); ``out``: the images array."""
    label = np.ascontiguousarray(label, dtype=np.int32)
    vector = np.ascontiguousarray(vector, dtype=np.float32)
    props = np.ascontiguousarray(props, dtype=np.float64)
    rho = np.ascontiguousarray(filter_rotations, dtype=np.float64)
    nx, ny, nz = label.shape
    if out is None:
        out = np.empty((nx, ny, rho.size), dtype=np.float64)
    assert out.shape == (nx, ny, rho.size) and out.dtype == np.float64 and out.flags.c_contiguous
    warn = np.zeros(1, dtype=np.uint32)
    rc = lib().oracle_simulate(nx, ny, nz, _ptr(label), _ptr(vector),
                               _ptr(props), props.shape[0], float(step_size),
                               float(light_intensity), float(voxel_size),
                               float(wavelength), float(tissue_refrection),
                               INTERP[interpolate], int(bool(untilt_sensor_view)),
                               int(bool(flip_z)), _ptr(rho), rho.size,
                               float(theta), float(phi), int(n_threads),
                               _ptr(out), _ptr(warn))
    if rc != 0:
        raise ValueError(f'oracle_simulate failed: {rc}')
    return out, int(warn[0])


def count_steps(dim, step_size, theta, phi):
    """Ray-steps of one tilt (rays x z-march iterations)."""
    nx, ny, nz = (int(d) for d in dim)
    return int(lib().oracle_count_steps(nx, ny, nz, float(step_size),
                                        float(theta), float(phi)))


def light_starts(dim, theta, phi, tissue_refrection=1.0, untilt_sensor_view=False):
    """Each pixel's ray entry point on the bottom plane, (nx, ny, 2) x, y (before the half step)."""
    nx, ny, nz = (int(d) for d in dim)
    out = np.empty((nx, ny, 2), dtype=np.float64)
    lib().oracle_light_starts(nx, ny, nz, float(theta), float(phi), float(tissue_refrection),
                              int(bool(untilt_sensor_view)), _ptr(out))
    return out


def get_vec(label, vector, pos, interpolate='Lerp', flip_z=False):
    """GetVec at one position (x, y, z) -> float64 (3,)."""
    label = np.ascontiguousarray(label, dtype=np.int32)
    vector = np.ascontiguousarray(vector, dtype=np.float32)
    nx, ny, nz = label.shape
    out = np.empty(3, dtype=np.float64)
    lib().oracle_get_vec(nx, ny, nz, _ptr(label), _ptr(vector), int(bool(flip_z)),
                         INTERP[interpolate], float(pos[0]), float(pos[1]), float(pos[2]),
                         _ptr(out))
    return out
