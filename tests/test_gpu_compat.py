"""The reference-side binding (fastpli_amd/compat.py) driven as the reference's
Simpli drives its pybind11 modules (src/fastpli/simulation/simpli.py:82-114,
182-183), on the golden cube: labels, axes and the property table equal the
reference's golden fixture; the five tilted images match it to 1e-4 relative."""

import sys

import numpy as np
import pytest

from fastpli_amd.compat import _Generator, _Simulator
from tests.conftest import GOLDEN_ROT, GOLDEN_TILTS
from tests.test_gpu_simulate import assert_intensity_close

pytestmark = pytest.mark.gpu


def test_compat_modules_on_the_golden_cube(golden, cube_bundles):
    print('compat: _Generator / _Simulator on the golden cube', flush=True)
    sys.stdout.flush()
    gen = _Generator()
    assert gen.set_omp_num_threads(2) == 2
    gen.set_volume([24, 24, 24], [-60.0, -60.0, -60.0], 5.0)
    gen.set_fiber_bundles(cube_bundles.as_array(), cube_bundles.layers)
    assert gen.dim_local() == [24, 24, 24] and gen.dim_offset() == [0, 0, 0]
    label, vector, props = gen.run_generation()
    props[0, 1] = 0.0  # Simpli's background_absorption
    assert np.array_equal(label, golden['tissue'])
    assert np.array_equal(vector, golden['optical_axis'])
    assert np.array_equal(props, golden['tissue_properties'])
    lab_only, empty, _ = gen.run_generation(only_label=True)
    assert np.array_equal(lab_only, label) and empty.shape == (0,)
    sim = _Simulator()
    sim.set_pli_setup(1.0, 26000, 5.0, 525, 1.0, 'Slerp', True, False, GOLDEN_ROT)
    for t, (theta, phi) in enumerate(GOLDEN_TILTS):
        img = sim.run_simulation([24, 24, 24], label, vector, props, theta, phi)
        assert img.shape == (24, 24, 6) and img.dtype == np.float64
        assert_intensity_close(img, golden['data'][t], 26000)
    with pytest.raises(ValueError, match='not the same size'):
        gen.set_fiber_bundles(cube_bundles.as_array(), [])
