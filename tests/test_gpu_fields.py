"""March fields written by the voxelizer (generate_tissue_fields) against the
pack of the same C-order fields (pack_fields): the record buffers are equal
byte for byte in every storage, and run_simulation_fields gives the images of
run_simulation_tilts bit for bit.  Also the checks run_simulation_fields makes
before the march (CheckInput's label range, the volume shape)."""

import sys
import warnings

import numpy as np
import pytest
import torch

from fastpli_amd import _lib, objects, synthetic
from fastpli_amd.simulation import MarchFields, Simpli
from tests import bruteforce

pytestmark = pytest.mark.gpu

ROT9 = np.linspace(0, np.pi, 9, endpoint=False)
TILTS5 = np.deg2rad([(0, 0), (5.5, 0), (5.5, 90), (5.5, 180), (5.5, 270)])


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _simpli(fbs, dim, origin, voxel_size, interp='Slerp'):
    s = Simpli()
    s.fiber_bundles = fbs
    s.voxel_size = voxel_size
    s.dim = np.array(dim, dtype=int)
    s.dim_origin = origin
    s.filter_rotations = ROT9
    s.light_intensity = 26000
    s.wavelength = 525
    s.interpolate = interp
    return s


def _written(f):
    """The bytes of a MarchFields buffer the calls write (records, then codes)."""
    n = int(np.prod(f.dim))
    rb = _lib.record_bytes(f.storage)
    parts = [f.buffer[:n * rb]]
    if f.storage == 'angle16':
        _, off = _lib.fields_layout(f.dim, f.storage)
        parts.append(f.buffer[off:off + 4 * n])
    return torch.cat(parts)


def _check(s, storage, tilts=TILTS5):
    with warnings.catch_warnings():
        warnings.simplefilter('ignore')
        tissue, axis, props = s.generate_tissue_device()
        t2, a2, p2, fields = s.generate_tissue_fields(storage)
        _, none_t, p3, fields_only = s.generate_tissue_fields(storage, with_arrays=False)
    assert none_t is None
    assert torch.equal(tissue, t2) and torch.equal(axis, a2)
    assert np.array_equal(props, p2) and np.array_equal(props, p3)
    packed = s.pack_fields(tissue, axis, storage)
    assert fields.storage == storage and fields.dim == tuple(tissue.shape)
    for a, b in ((fields, packed), (fields_only, packed)):
        assert torch.equal(_written(a), _written(b)), 'voxelizer records != packed records'

    # the generator's label bound covers the packed range
    assert fields.label_min <= packed.label_min and fields.label_max >= packed.label_max
    s.field_storage = storage
    ref = s.run_simulation_tilts(tissue, axis, props, tilts)
    got = s.run_simulation_fields(fields, props, tilts)
    assert torch.equal(got, ref), 'images of the fields differ from run_simulation_tilts'
    return fields


@pytest.mark.parametrize('storage', ['f32', 'f16', 'angle16'])
def test_fields_golden_cube(cube_bundles, storage):
    _say(f'fields: golden cube {storage}')
    _check(_simpli(cube_bundles, (24, 24, 24), [-60.0] * 3, 5.0), storage)


@pytest.mark.parametrize('seed', range(4))
def test_fields_random_scenes(seed):
    _say(f'fields: random scene {seed}')
    fbs = bruteforce.random_scene(seed, n_fibers=4, n_points=6)
    for storage in ('f32', 'f16', 'angle16'):
        _check(_simpli(fbs, (13, 11, 37), [0.0] * 3, 1.0), storage)
        _check(_simpli(fbs, (9, 10, 70), [-1.5, 0.5, 0.25], 0.5, 'Lerp'), storage)


def test_fields_parallel_fibers_nn():
    _say('fields: 64 parallel fibers 120^3, NN')
    _check(_simpli(synthetic.parallel_fibers(), (120, 120, 120), [0.0] * 3, 0.25, 'NN'), 'f32',
           tilts=[(0.0, 0.0)])


def test_fields_crossing_full_size():
    _say('fields: crossing bundles 600x600x300, f32')
    fields = _check(_simpli(synthetic.crossing_bundles(seed=0), (600, 600, 300), [0.0] * 3, 0.2),
                    'f32')
    assert fields.nbytes == 16 * 600 * 600 * 300


def test_fields_label_range_is_checked(cube_bundles):
    _say('fields: label range against the property table')
    s = _simpli(cube_bundles, (24, 24, 24), [-60.0] * 3, 5.0)
    _, _, props, fields = s.generate_tissue_fields('f32', with_arrays=False)
    with pytest.raises(ValueError, match='label exceed properties.size()'):
        s.run_simulation_fields(fields, props[:2], TILTS5[:1])
    # a packed field with a negative label
    tissue = torch.zeros((24, 24, 24), dtype=torch.int32, device='cuda')
    tissue[3, 4, 5] = -1
    axis = torch.zeros((24, 24, 24, 3), dtype=torch.float32, device='cuda')
    packed = s.pack_fields(tissue, axis, 'f32')
    assert packed.label_min == -1 and packed.label_max == 0
    with pytest.raises(ValueError, match='label < 0 detected'):
        s.run_simulation_fields(packed, props, TILTS5[:1])


def test_fields_shape_and_type_are_checked(cube_bundles):
    _say('fields: shape and type checks')
    s = _simpli(cube_bundles, (24, 24, 24), [-60.0] * 3, 5.0)
    _, _, props, fields = s.generate_tissue_fields('f32', with_arrays=False)
    s.dim = np.array([24, 24, 23])
    with pytest.raises(ValueError):
        s.run_simulation_fields(fields, props, TILTS5[:1])
    with pytest.raises(TypeError):
        s.run_simulation_fields(fields.buffer, props, TILTS5[:1])
    with pytest.raises(ValueError):
        s.generate_tissue_fields('f8')


def test_fields_all_zero_warning():
    _say('fields: all labels 0')
    fbs = objects.FiberBundles([[np.array([[100.0, 100, 100, 1], [110.0, 100, 100, 1]])]])
    fbs.layers = [[(1.0, -0.004, 1, 'p')]]
    s = _simpli(fbs, (8, 8, 8), [0.0] * 3, 1.0)
    with pytest.warns(UserWarning, match='All labels are 0'):
        _, _, _, fields = s.generate_tissue_fields('f32', with_arrays=False)
    assert fields.label_max == 0
    assert isinstance(fields, MarchFields)
