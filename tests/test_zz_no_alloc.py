"""fpl_simulate_noalloc, the entry point that allocates nothing (its scratch is
a static device array), against fpl_simulate: identical images and warnings
on two calls in a row, the second on a smaller volume (the shared array keeps
the first call's contents), in the f32 and f16 storages; volumes beyond the
static scratch are rejected with an error."""

import sys
import warnings

import numpy as np
import pytest
import torch

from fastpli_amd import _lib, synthetic
from fastpli_amd.simulation import Simpli
from tests.conftest import GOLDEN_ROT, GOLDEN_TILTS

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _setup(voxel, storage, rot=GOLDEN_ROT):
    s = Simpli()
    s.voxel_size = voxel
    s.filter_rotations = rot
    s.light_intensity = 26000
    s.wavelength = 525
    s.interpolate = 'Slerp'
    s.field_storage = storage
    return s._setup()


def _both(setup, dim, label, vector, props, tilts):
    images = torch.empty((len(tilts), dim[0], dim[1], setup.n_rho), dtype=torch.float64,
                         device='cuda')
    ref = torch.empty_like(images)
    w_ref = _lib.simulate(setup, dim, label, vector, props, tilts, ref)
    images.fill_(float('nan'))
    w = _lib.simulate_noalloc(setup, dim, label, vector, props, tilts, images)
    torch.cuda.synchronize()
    assert w == w_ref
    assert torch.equal(images, ref)


@pytest.mark.parametrize('storage', ['f32', 'f16'])
def test_noalloc_equals_simulate_twice(golden, storage):
    # first call: 64 parallel fibers in 120^3 voxels; second: the 24^3 golden cube
    _say(f'noalloc: 120^3 then 24^3 ({storage})')
    s = Simpli()
    s.voxel_size = 0.25
    s.dim = [120, 120, 120]
    s.fiber_bundles = synthetic.parallel_fibers()
    with warnings.catch_warnings():
        warnings.simplefilter('ignore')
        label, vector, props = s.generate_tissue_device()
    setup, _keep = _setup(0.25, storage, np.linspace(0, np.pi, 9, endpoint=False))
    tilts = np.deg2rad([(0, 0), (5.5, 0), (5.5, 90)])
    _both(setup, [120, 120, 120], label, vector, props, tilts)
    lab = torch.from_numpy(golden['tissue']).cuda()
    vec = torch.from_numpy(golden['optical_axis']).cuda()
    setup, _keep = _setup(5.0, storage)
    _both(setup, [24, 24, 24], lab, vec, golden['tissue_properties'], GOLDEN_TILTS)


def test_noalloc_rejects_volumes_beyond_its_scratch():
    _say('noalloc: a volume beyond the static scratch')
    dim = [256, 256, 80]  # 5.2 M voxels x 16 B > 64 MiB
    label = torch.zeros(dim, dtype=torch.int32, device='cuda')
    vector = torch.zeros(dim + [3], dtype=torch.float32, device='cuda')
    images = torch.empty((1, 256, 256, 6), dtype=torch.float64, device='cuda')
    setup, _keep = _setup(1.0, 'f32')
    with pytest.raises(ValueError, match='static scratch'):
        _lib.simulate_noalloc(setup, dim, label, vector, np.zeros((1, 2)), [(0.0, 0.0)], images)
    # the f16 storage (8 B per voxel) fits the same volume
    setup, _keep = _setup(1.0, 'f16')
    _lib.simulate_noalloc(setup, dim, label, vector, np.zeros((1, 2)), [(0.0, 0.0)], images)
    torch.cuda.synchronize()
    assert torch.all(images == 26000 / 4)
