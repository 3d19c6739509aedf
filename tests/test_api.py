"""CPU tests of the host-side API: Simpli parameters, containers, the C ABI library."""

import ctypes
import os
import re
import warnings

import numpy as np
import pytest

from fastpli_amd import _lib, objects
from fastpli_amd.io import fiber_bundles as fio
from fastpli_amd.simulation import Simpli
from tests.conftest import GOLDEN, ROOT


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, 'include', 'fastpli_amd.h')).read()
    declared = set(re.findall(r'\b(fpl_[a-z_]+)\s*\(', header))
    assert declared == set(_lib.EXPORTS)
    L = _lib.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert L.fpl_version().startswith(b'fastpli_amd')


def test_count_voxel_steps_host_only():
    s = Simpli()
    s.voxel_size = 0.2
    s.dim = [600, 600, 300]
    s.light_intensity = 26000
    s.wavelength = 525
    s.filter_rotations = np.linspace(0, np.pi, 9, endpoint=False)
    assert s.count_voxel_steps(0, 0) == 9 * 600 * 600 * 300
    assert s.count_voxel_steps(np.deg2rad(5.5), 0) == 9 * 600 * 600 * 301


def test_frozen_and_dict():
    s = Simpli()
    with pytest.raises(TypeError, match='frozen'):
        s.not_a_member = 1
    d = s.get_dict()
    assert 'interpolate' in d and d['interpolate'] == 'Slerp'
    assert d['step_size'] == 1.0 and d['untilt_sensor_view'] is True
    assert not any(k.startswith('Simpli') or k.startswith('_') for k in d)
    with pytest.warns(UserWarning):
        s.set_dict(d)  # None values warn (simpli_test.py:27-29)
    with pytest.raises(ValueError):
        s.set_dict({'_dim': 1})


def test_setters_validate():
    s = Simpli()
    with pytest.raises(TypeError, match='voxel_size is not set yet'):
        s.set_voi([0, 0, 0], [12, 12, 12])
    with pytest.raises(ValueError):
        s.voxel_size = 0
    with pytest.raises(ValueError, match='Only'):
        s.interpolate = 'Cubic'
    with pytest.raises(TypeError):
        s.dim = [1.5, 2, 3]
    with pytest.raises(TypeError):
        s.dim = [1, 2]
    with pytest.raises(TypeError):
        s.filter_rotations = []
    with pytest.raises(TypeError):
        s.tilts = 5
    with pytest.raises(TypeError):
        s.tilts = [[1, 2, 3]]
    with pytest.raises(TypeError):
        s.omp_num_threads = 1.5
    s.voxel_size = 0.5
    s.set_voi([-1, -2, -3], [1, 2, 3])
    assert np.array_equal(s.dim, [4, 8, 12])
    v_min, v_max = s.get_voi()
    assert np.allclose(v_min, [-1, -2, -3]) and np.allclose(v_max, [1, 2, 3])
    s.voxel_size = 0.25  # recalculates dim for the same VOI
    assert np.array_equal(s.dim, [8, 16, 24])
    assert s.memory_usage('MB') == pytest.approx(8 * 16 * 24 * 20 / 1024**2)


def test_crop_tilt():
    # test_crop_tilt (simpli_test.py:300-325)
    s = Simpli()
    s.voxel_size = 0.51
    s.pixel_size = 10
    s.set_voi([-100] * 3, [100] * 3)
    s.tilts = np.deg2rad([(15.5, 0)])
    dim_org = s.dim.copy()
    s.add_crop_tilt_halo()
    dim_crop = s.dim.copy()
    dim_crop[:2] -= 2 * s.crop_tilt_voxel()
    assert np.array_equal(dim_org, dim_crop)
    image_crop = s.rm_crop_tilt_halo(np.empty(s.dim))
    assert np.array_equal(dim_org, image_crop.shape)


def test_generation_input_checks():
    s = Simpli()
    s.voxel_size = 1
    s.dim = [4, 4, 4]
    with pytest.raises(ValueError, match='fiber_bundles is not set'):
        s.generate_tissue()
    s.fiber_bundles = [[[[0, 0, 0, 1], [1, 1, 1, 1]]]]
    with pytest.raises(ValueError, match='None'):
        s.generate_tissue()


def test_layers_validation():
    with pytest.raises(ValueError, match='same radii'):
        objects.Layers([(0.5, 0, 1, 'b'), (0.5, -0.004, 1, 'p')])
    with pytest.raises(ValueError, match='background'):
        objects.Layers([(1.0, 0.1, 1, 'b')])
    with pytest.warns(UserWarning, match='sorted'):
        lys = objects.Layers([(1.0, -0.004, 1, 'p'), (0.5, 0, 1, 'b')])
    assert [ly.scale for ly in lys] == [0.5, 1.0]
    with pytest.raises(TypeError):
        objects.Layer((1.0, 0, 1))


def test_fiber_containers_and_flat():
    with pytest.raises(TypeError):
        objects.Fiber([[1, 2, 3]])
    fbs = objects.FiberBundles([[[[0, 0, 0, 1], [1, 0, 0, 1]], [[0, 1, 0, 1]]],
                                [[[0, 0, 1, 2], [0, 0, 2, 2], [0, 0, 3, 2]]]],
                               [[(1.0, -0.004, 1, 'p')], [(0.5, 0, 1, 'b'), (1.0, 0.004, 1, 'r')]])
    points, fiber_ptr, bundle_ptr, layer_ptr, table = fbs.to_flat()
    assert points.shape == (6, 4) and points.flags.c_contiguous
    assert list(fiber_ptr) == [0, 2, 3, 6]
    assert list(bundle_ptr) == [0, 2, 3]
    assert list(layer_ptr) == [0, 1, 3]
    assert list(table[3]) == [ord('p'), ord('b'), ord('r')]
    with pytest.raises(TypeError):
        objects.FiberBundles([[[[0, 0, 0, 1]]]], [[(1, 0, 0, 'b')], [(1, 0, 0, 'b')]])


def _dat_line_by_line(text):
    """The .dat grammar read one line at a time (empty line: end of fiber;
    a second one in a row: end of bundle), to check the run-based reader."""
    fiber, bundles, empty = [], [[]], 0
    for line in text.splitlines():
        if line.strip():
            empty = 0
            fiber.append([float(v) for v in line.split()[:4]])
            continue
        empty += 1
        if empty == 1 and fiber:
            bundles[-1].append(np.array(fiber))
            fiber = []
        elif empty == 2 and bundles[-1]:
            bundles.append([])
    if fiber:
        bundles[-1].append(np.array(fiber))
    return bundles


@pytest.mark.parametrize('text', [
    '1 2 3 4\n5 6 7 8\n\n9 9 9 9\n\n\n1 1 1 1\n',        # two bundles
    '\n\n\n1 2 3 4\n\n\n\n\n5 6 7 8 9 10\n  \n',        # leading / long blank runs, 6 numbers
    '1 2 3 4\n\n\n',                                    # trailing empty bundle
    '1 2 3 4\n2 3 4 5',                                   # no final newline
    '',
])
def test_dat_reader_grammar(tmp_path, text):
    p = tmp_path / 'g.dat'
    p.write_text(text)
    got = fio.load(str(p))
    ref = _dat_line_by_line(text)
    assert len(got) == len(ref)
    for gb, rb in zip(got, ref):
        assert len(gb) == len(rb)
        for gf, rf in zip(gb, rb):
            assert np.array_equal(gf.as_array(), rf)


@pytest.mark.parametrize('name, bundles, n_points', [('cube.dat', [100], 1994),
                                                     ('optic_chiasm.dat', [61, 61, 19, 19],
                                                      6465)])
def test_dat_models_and_exact_round_trip(tmp_path, name, bundles, n_points):
    path = os.path.join(GOLDEN, name)
    fbs = fio.load(path)
    assert [len(b) for b in fbs] == bundles
    assert sum(f.shape[0] for b in fbs for f in b) == n_points
    ref = _dat_line_by_line(open(path).read())
    assert all(np.array_equal(f.as_array(), r) for b, rb in zip(fbs, ref) for f, r in zip(b, rb))
    p = tmp_path / 'x.dat'
    fio.save(str(p), fbs)
    with pytest.raises(FileExistsError):
        fio.save(str(p), fbs)  # mode 'w-'
    again = fio.load(str(p))
    assert [len(b) for b in again] == bundles
    for b, bb in zip(fbs, again):
        for f, g in zip(b, bb):
            assert np.array_equal(f.as_array(), g.as_array())


def test_field_storage_property():
    s = Simpli()
    assert s.field_storage == 'f32'
    s.field_storage = 'angle16'
    assert s.get_dict()['field_storage'] == 'angle16'
    with pytest.raises(ValueError):
        s.field_storage = 'bf16'


def test_adjoint_binding_checks_array_shapes():
    """The adjoint's raw pointers are checked against setup and dim before the call."""
    import torch
    setup, _keep = _lib.make_setup(1.0, 26000, 1.0, 525, 1.0, 'NN', True, False,
                                   np.linspace(0, np.pi, 6, endpoint=False))
    dim = [4, 5, 6]
    lab = torch.zeros(dim, dtype=torch.int32)
    vec = torch.zeros(dim + [3], dtype=torch.float32)
    props = np.zeros((2, 2))
    tilts = [(0.0, 0.0)]
    good = torch.zeros((1, 4, 5, 6), dtype=torch.float64)
    with pytest.raises(ValueError, match='grad_images has shape'):
        _lib.simulate_adjoint(setup, dim, lab, vec, props, tilts, torch.zeros((1, 4, 5, 9),
                                                                              dtype=torch.float64))
    with pytest.raises(ValueError, match='grad_images has dtype'):
        _lib.simulate_adjoint(setup, dim, lab, vec, props, tilts, good.float())
    with pytest.raises(ValueError, match='grad_vector has shape'):
        _lib.simulate_adjoint(setup, dim, lab, vec, props, tilts, good,
                              torch.zeros((4, 5, 5, 3), dtype=torch.float32))
    with pytest.raises(ValueError, match='optical_axis has dtype'):
        _lib.simulate_adjoint(setup, dim, lab, vec.double(), props, tilts, good)
    with pytest.raises(ValueError, match='must be contiguous'):
        _lib.simulate_adjoint(setup, dim, lab, vec, props, tilts,
                              torch.zeros((1, 5, 4, 6), dtype=torch.float64).transpose(1, 2))


def test_near_parallel_tilt_is_rejected_on_the_host():
    """A tilt just short of 90 degrees has a z-step of ~1e-8 voxel: rejected
    (by trip count) before any device work, like the exactly parallel one."""
    s = Simpli()
    s.voxel_size = 1.0
    s.dim = [4, 4, 300]
    s.light_intensity = 26000
    s.wavelength = 525
    s.filter_rotations = np.linspace(0, np.pi, 6, endpoint=False)
    with pytest.raises(ValueError, match='z-steps per ray'):
        s.count_voxel_steps(1.5707963, 0.0)
    assert s.count_voxel_steps(np.deg2rad(60), 0.0) > 0


def test_fiber_model_edits_are_counted():
    """Every edit of a fiber model object bumps the counter the device copy of
    a model is keyed on; arrays handed out are read-only snapshots."""
    from fastpli_amd.objects._sealed import edit_count
    fbs = objects.FiberBundles([[np.array([[0.0, 0, 0, 1], [1.0, 0, 0, 1]])]],
                               [[(1.0, -0.004, 1, 'p')]])
    flat0 = fbs.to_flat()
    edits = [
        lambda: fbs[0][0].__setitem__((0, 0), 5.0),
        lambda: setattr(fbs[0][0], 'radii', [2.0, 2.0]),
        lambda: fbs[0].append(np.array([[0.0, 1, 0, 1], [1.0, 1, 0, 1]])),
        lambda: fbs[0].layers.__setitem__(0, (1.0, -0.002, 1, 'p')),
        lambda: setattr(fbs[0], 'layers', [(1.0, -0.001, 1, 'p')]),
        lambda: fbs.append(fbs[0]),
        lambda: fbs.__delitem__(1),
    ]
    for edit in edits:
        n = edit_count()
        edit()
        assert edit_count() == n + 1
    pts = fbs[0][0].points
    with pytest.raises(ValueError):
        pts[0, 0] = 7.0  # read-only snapshot
    assert fbs.to_flat()[0][0, 0] == 5.0 and flat0[0][0, 0] == 0.0
    assert fbs.to_flat()[4][1][0] == -0.001


def test_install_alias_gives_the_reference_module_paths():
    """fastpli_amd.compat.install_alias() makes `from fastpli.simulation import
    Simpli` (and .objects, .io) this package's classes; run in a fresh process."""
    import subprocess
    import sys
    code = ('import fastpli_amd, fastpli_amd.compat as c; c.install_alias();'
            'from fastpli.simulation import Simpli; from fastpli.objects import FiberBundles;'
            'import fastpli.io;'
            'import fastpli_amd.simulation as s, fastpli_amd.objects as o;'
            'assert Simpli is s.Simpli and FiberBundles is o.FiberBundles;'
            'assert fastpli.io.fiber_bundles is fastpli_amd.io.fiber_bundles;'
            'print("ok")')
    out = subprocess.run([sys.executable, '-c', code], cwd=ROOT, capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == 'ok', out.stderr


def test_flag_bits_match_the_header_and_negative_images_raise():
    """The warn-mask bits of the Python layer are the header's; the march's
    FPL_FLAG_NEGATIVE is no warning but the reference's ValueError
    "intensity < 0 detected" (simpli.py:184-185)."""
    header = open(os.path.join(ROOT, 'include', 'fastpli_amd.h')).read()
    bits = {name: 1 << int(sh) for name, sh in
            re.findall(r'#define FPL_((?:WARN|FLAG)_[A-Z_]+) \(1u << (\d+)\)', header)}
    assert bits == {'WARN_OVERLAP': _lib.WARN_OVERLAP, 'WARN_SINGLE_POINT': _lib.WARN_SINGLE_POINT,
                    'WARN_ZERO_VEC': _lib.WARN_ZERO_VEC, 'WARN_NOT_NORM': _lib.WARN_NOT_NORM,
                    'WARN_NO_LIGHT': _lib.WARN_NO_LIGHT, 'FLAG_NEGATIVE': _lib.FLAG_NEGATIVE}
    with warnings.catch_warnings():
        warnings.simplefilter('error')
        _lib.raise_warnings(_lib.FLAG_NEGATIVE)  # no UserWarning for it
        _lib.check_images(_lib.WARN_NOT_NORM | _lib.WARN_ZERO_VEC)
    with pytest.raises(ValueError, match='intensity < 0 detected'):
        _lib.check_images(_lib.FLAG_NEGATIVE | _lib.WARN_NOT_NORM)
