"""GPU paths for large property / layer tables.

The march keeps the property table in LDS when it has at most 128 rows, the
adjoint at most 256 rows, and the voxelizer keeps the layer tables in LDS up
to 64 rows; larger tables are read from HBM by separate kernel builds.  These
tests drive the large-table builds: a 90-bundle scene (up to 271 labels) vs
the oracle, and padded property tables whose extra rows are never used,
which must leave images bit-identical and gradients equal (up to the order
of atomic additions) to the compact table's.
"""

import sys
import warnings

import numpy as np
import pytest

from tests import bruteforce
from tests.conftest import GOLDEN_ROT, GOLDEN_TILTS
from tests.test_gpu_adjoint import TILTS5, _adjoint, _small_tissue
from tests.test_gpu_generate import _compare
from tests.test_gpu_simulate import _oracle, _setup, assert_intensity_close

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pad(props, n, seed=0):
    rng = np.random.default_rng(seed)
    extra = np.stack([rng.uniform(-0.01, 0.01, n - len(props)), rng.uniform(0, 10, n - len(props))], 1)
    return np.vstack([props, extra])


def test_many_bundles_generate_and_simulate():
    _say('large tables: 90 bundles, generate + simulate vs oracle')
    fbs = bruteforce.random_scene(11, n_bundles=90, n_fibers=1, n_points=3, size=14.0)
    dim = (16, 16, 16)
    label, vector, _ = _compare(fbs, dim, [0.0, 0.0, 0.0], 1.0)
    s = _setup(dim, 1.0, interpolate='Slerp', rho=GOLDEN_ROT)
    s.fiber_bundles = fbs
    with warnings.catch_warnings():
        warnings.simplefilter('ignore')
        _, _, props = s.generate_tissue()
    assert props.shape[0] > 256, props.shape
    assert label.max() > 128
    for interp in ('Slerp', 'NN'):
        s.interpolate = interp
        with warnings.catch_warnings():
            warnings.simplefilter('ignore')
            imgs = s.run_simulation_tilts(label, vector, props, GOLDEN_TILTS)
        for t, (theta, phi) in enumerate(GOLDEN_TILTS):
            ref, _ = _oracle(s, label, vector, props, theta, phi)
            assert_intensity_close(imgs[t], ref, s.light_intensity)


def test_padded_property_table_simulate_identical(golden):
    _say('large tables: padded property table, images bit-identical')
    props = golden['tissue_properties']
    for interp in ('Slerp', 'NN'):
        s = _setup((24, 24, 24), 5.0, interpolate=interp, untilt_sensor_view=True)
        a = s.run_simulation_tilts(golden['tissue'], golden['optical_axis'], props, GOLDEN_TILTS)
        b = s.run_simulation_tilts(golden['tissue'], golden['optical_axis'], _pad(props, 300),
                                   GOLDEN_TILTS)
        assert np.array_equal(a, b)


def test_padded_property_table_adjoint_unchanged():
    _say('large tables: padded property table, adjoint unchanged')
    s, label, vector, props = _small_tissue(seed=3)
    rng = np.random.default_rng(4)
    w = rng.standard_normal((len(TILTS5), label.shape[0], label.shape[1], 9))
    gp, gv = _adjoint(s, label, vector, props, w)
    gp2, gv2 = _adjoint(s, label, vector, _pad(props, 300), w)
    n = len(props)
    # the per-label sums and the axis gradient are accumulated with atomics,
    # so their rounding depends on the order of the additions: FP64 sums to
    # 1e-12 of the largest entry, FP32 axis sums to 1e-5 of the largest one
    assert np.max(np.abs(gp2[:n] - gp)) <= 1e-12 * np.max(np.abs(gp))
    assert not np.any(gp2[n:])
    assert np.max(np.abs(gv2 - gv)) <= 1e-5 * np.max(np.abs(gv))


def test_labels_beyond_15_bits():
    _say('large tables: 33000 layers (labels beyond 15 bits), generate vs oracle')
    from fastpli_amd.objects import FiberBundles
    rng = np.random.default_rng(9)
    fibers = []
    for _ in range(3):
        start = rng.uniform(0, 12, 3)
        pts = start + np.cumsum(rng.normal(scale=2.0, size=(4, 3)), axis=0)
        fibers.append(np.hstack([pts, rng.uniform(1.5, 3.0, (4, 1))]))
    scales = np.linspace(0.3, 1.0, 33000)
    layers = [(float(sc), (-0.004 if k % 2 else 0.0), float(k % 7), 'p' if k % 2 else 'b')
              for k, sc in enumerate(scales)]
    fbs = FiberBundles([fibers], [layers])
    label, vector, _ = _compare(fbs, (12, 12, 12), [0.0, 0.0, 0.0], 1.0)
    assert label.max() > 0x7FFF
