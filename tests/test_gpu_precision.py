"""Precision storages of the march (uint16 labels + fp16 / 16-bit angle axes).

Reports the intensity error and the EPA transmittance / direction /
retardation errors against the FP64 oracle for the 64-parallel-fiber volume
(120^3 and a 300-voxel-high variant) and checks them against stated bounds:

  storage   max rel. intensity   max |direction| [rad]   max |retardation|
  f32       1e-4                 1e-4                    1e-4
  f16       2e-3                 2e-3                    1e-3
  angle16   5e-4                 5e-4                    3e-4
"""

import sys

import numpy as np
import pytest

import oracle
from oracle import analysis
from fastpli_amd import synthetic
from fastpli_amd.simulation import Simpli
from tests.conftest import GOLDEN_ROT, GOLDEN_TILTS

pytestmark = pytest.mark.gpu

BOUNDS = {'f32': (1e-4, 1e-4, 1e-4), 'f16': (2e-3, 2e-3, 1e-3), 'angle16': (5e-4, 5e-4, 3e-4)}
ROT9 = np.linspace(0, np.pi, 9, endpoint=False)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _errors(img, ref, I0):
    rel = np.abs(img - ref) / np.maximum(np.abs(ref), 1e-8 * I0)
    te, de, re = analysis.epa(img)
    tr, dr, rr = analysis.epa(ref)
    dd = np.abs(np.angle(np.exp(2j * (de - dr)))) / 2  # direction is mod pi
    return rel.max(), dd.max(), np.abs(re - rr).max(), np.abs(te - tr).max() / I0


@pytest.fixture(scope='module')
def volumes():
    out = {}
    for name, fbs, dim in [('120', synthetic.parallel_fibers(), (120, 120, 120)),
                           ('300', synthetic.parallel_fibers(size_z=75.0, n_rows=36),
                            (120, 120, 300))]:
        s = Simpli()
        s.voxel_size = 0.25
        s.dim = np.array(dim)
        s.fiber_bundles = fbs
        out[name] = s.generate_tissue()
    return out


@pytest.mark.parametrize('storage', ['f32', 'f16', 'angle16'])
@pytest.mark.parametrize('vol', ['120', '300'])
def test_precision_parallel_fibers(volumes, storage, vol):
    tissue, axis, props = volumes[vol]
    s = Simpli()
    s.voxel_size = 0.25
    s.dim = np.array(tissue.shape)
    s.filter_rotations = ROT9
    s.light_intensity = 26000
    s.wavelength = 525
    s.interpolate = 'NN'
    s.field_storage = storage
    for theta, phi in [(0.0, 0.0), (np.deg2rad(5.5), np.deg2rad(90))]:
        img = s.run_simulation(tissue, axis, props, theta, phi)
        ref, _ = oracle.simulate(tissue, axis, props, theta, phi, voxel_size=0.25,
                                 filter_rotations=ROT9, interpolate='NN', n_threads=8)
        ie, de, re, te = _errors(img, ref, 26000)
        _say(f'precision {storage} nz={tissue.shape[2]} tilt=({theta:.3f},{phi:.3f}): '
             f'intensity {ie:.2e} direction {de:.2e} rad retardation {re:.2e} '
             f'transmittance {te:.2e}')
        bi, bd, br = BOUNDS[storage]
        assert ie <= bi and de <= bd and re <= br


@pytest.mark.parametrize('storage', ['f16', 'angle16'])
def test_precision_golden_slerp(golden, storage):
    s = Simpli()
    s.voxel_size = 5.0
    s.dim = np.array([24, 24, 24])
    s.filter_rotations = GOLDEN_ROT
    s.light_intensity = 26000
    s.wavelength = 525
    s.field_storage = storage
    imgs = s.run_simulation_tilts(golden['tissue'], golden['optical_axis'],
                                  golden['tissue_properties'], GOLDEN_TILTS)
    for t in range(5):
        ie, de, re, te = _errors(imgs[t], golden['data'][t], 26000)
        _say(f'precision golden {storage} tilt {t}: intensity {ie:.2e} direction {de:.2e} '
             f'retardation {re:.2e}')
        assert ie <= 2e-3


def test_precision_label_limits(golden):
    s = Simpli()
    s.voxel_size = 5.0
    s.dim = np.array([24, 24, 24])
    s.filter_rotations = GOLDEN_ROT
    s.light_intensity = 26000
    s.wavelength = 525
    tissue = golden['tissue'].copy()
    tissue[tissue == 3] = 40000
    props = np.zeros((40001, 2))
    props[1:4] = golden['tissue_properties'][1:4]
    props[40000] = golden['tissue_properties'][3]
    s.field_storage = 'angle16'
    with pytest.raises(ValueError, match='does not fit'):
        s.run_simulation(tissue, golden['optical_axis'], props, 0, 0)
    s.field_storage = 'f16'
    img = s.run_simulation(tissue, golden['optical_axis'], props, 0, 0)
    assert img.shape == (24, 24, 6)
