"""GPU parity of run_simulation (FP32 path) vs the FP64 oracle.

Tolerance (FP32 mode), for every pixel and rotation:
|I_gpu - I_oracle| <= max(1e-4 * |I_oracle|, 1e-8 * I0).  The relative bound
governs wherever I_oracle >= 1e-4 * I0 (2.6 for I0 = 26000); below that the
absolute floor 1e-8 * I0 does (intensities that small come from near-total
extinction, where a relative bound would ask for more digits than FP32
attenuation sums carry).
"""

import sys

import numpy as np
import pytest

import oracle
from fastpli_amd import objects, synthetic
from fastpli_amd.simulation import Simpli
from tests.conftest import GOLDEN_ROT, GOLDEN_TILTS

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def assert_intensity_close(img, ref, I0, rtol=RTOL):
    err = np.abs(img - ref)
    bound = np.maximum(rtol * np.abs(ref), 1e-8 * I0)
    bad = err > bound
    assert not np.any(bad), (f'{np.count_nonzero(bad)} of {bad.size} values off; '
                             f'max rel {np.max(err / np.maximum(np.abs(ref), 1e-300)):.3e}')


def _setup(dim, voxel_size, interpolate='Slerp', rho=GOLDEN_ROT, **kw):
    s = Simpli()
    s.voxel_size = voxel_size
    s.dim = np.array(dim, dtype=int)
    s.filter_rotations = rho
    s.light_intensity = kw.pop('light_intensity', 26000)
    s.wavelength = kw.pop('wavelength', 525)
    s.interpolate = interpolate
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def _oracle(s, tissue, axis, props, theta, phi):
    img, w = oracle.simulate(tissue, axis, props, theta, phi, step_size=s.step_size,
                             light_intensity=s.light_intensity, voxel_size=s.voxel_size,
                             wavelength=s.wavelength, tissue_refrection=s.tissue_refrection,
                             interpolate=s.interpolate, untilt_sensor_view=s.untilt_sensor_view,
                             flip_z=s.flip_z_beam, filter_rotations=s.filter_rotations,
                             n_threads=8)
    return img, w


@pytest.mark.parametrize('t', range(5))
def test_simulate_golden(golden, t):
    _say(f'simulate: golden tilt {t}')
    s = _setup((24, 24, 24), 5.0)
    theta, phi = GOLDEN_TILTS[t]
    img = s.run_simulation(golden['tissue'], golden['optical_axis'], golden['tissue_properties'],
                           theta, phi)
    assert img.shape == (24, 24, 6) and img.dtype == np.float64
    assert_intensity_close(img, golden['data'][t], 26000)


def test_simulate_golden_all_tilts_one_call(golden):
    _say('simulate: golden, 5 tilts fused')
    s = _setup((24, 24, 24), 5.0)
    imgs = s.run_simulation_tilts(golden['tissue'], golden['optical_axis'],
                                  golden['tissue_properties'], GOLDEN_TILTS)
    for t in range(5):
        assert_intensity_close(imgs[t], golden['data'][t], 26000)


@pytest.mark.parametrize('interp', ['NN', 'Lerp', 'Slerp'])
@pytest.mark.parametrize('tilt', [(0.0, 0.0), (5.5, 0.0), (5.5, 90.0), (8.0, 42.0),
                                  (5.5, 270.0)])
def test_simulate_cube_modes(golden, interp, tilt):
    _say(f'simulate: cube {interp} tilt {tilt}')
    s = _setup((24, 24, 24), 5.0, interpolate=interp)
    theta, phi = np.deg2rad(tilt)
    img = s.run_simulation(golden['tissue'], golden['optical_axis'], golden['tissue_properties'],
                           theta, phi)
    ref, _ = _oracle(s, golden['tissue'], golden['optical_axis'], golden['tissue_properties'],
                     theta, phi)
    assert_intensity_close(img, ref, 26000)


@pytest.mark.parametrize('opts', [
    dict(flip_z_beam=True),
    dict(step_size=0.5),
    dict(step_size=1.7),
    dict(untilt_sensor_view=False),
    dict(untilt_sensor_view=False, tissue_refrection=1.3),
    dict(background_absorption=2.0),
])
def test_simulate_options(golden, opts):
    _say(f'simulate: options {opts}')
    props = golden['tissue_properties'].copy()
    bg = opts.pop('background_absorption', None)
    if bg is not None:
        props[0, 1] = bg
    s = _setup((24, 24, 24), 5.0, **opts)
    for tilt in [(0.0, 0.0), (5.5, 180.0), (10.0, 33.0)]:
        theta, phi = np.deg2rad(tilt)
        img = s.run_simulation(golden['tissue'], golden['optical_axis'], props, theta, phi)
        ref, _ = _oracle(s, golden['tissue'], golden['optical_axis'], props, theta, phi)
        assert_intensity_close(img, ref, 26000)


def test_simulate_parallel_fibers_config():
    _say('simulate: 64 parallel fibers 120^3, NN, 9 rotations')
    s = _setup((120, 120, 120), 0.25, interpolate='NN',
               rho=np.linspace(0, np.pi, 9, endpoint=False))
    s.fiber_bundles = synthetic.parallel_fibers()
    tissue, axis, props = s.generate_tissue()
    img = s.run_simulation(tissue, axis, props, 0.0, 0.0)
    ref, _ = _oracle(s, tissue, axis, props, 0.0, 0.0)
    assert_intensity_close(img, ref, 26000)


def test_simulate_uniform_column_closed_form():
    _say('simulate: uniform column closed form')
    v = np.array([0.6, 0.48, 0.64], dtype=np.float32)
    v /= np.linalg.norm(v)
    dim = (5, 7, 300)
    tissue = np.ones(dim, np.int32)
    axis = np.empty(dim + (3,), np.float32)
    axis[:] = v
    props = np.array([[0.0, 0.0], [-0.004, 1.0]])
    rho = np.linspace(0, np.pi, 9, endpoint=False)
    s = _setup(dim, 0.2, interpolate='NN', rho=rho)
    img = s.run_simulation(tissue, axis, props, 0.0, 0.0)
    t = 0.2e-6
    vd = v.astype(np.float64)
    cos2a = 1 - (vd[2] / np.linalg.norm(vd))**2
    delta = np.pi / 2 * (-0.004) * 4 * t / 525e-9 * np.linalg.norm(vd) * cos2a
    phi = np.arctan2(vd[1], vd[0])
    expect = 26000 / 4 * np.exp(-1e3 * t * 300) * (1 - np.sin(300 * delta) *
                                                   np.sin(2 * (rho - phi)))
    assert_intensity_close(img, np.broadcast_to(expect, img.shape), 26000)


def test_simulate_warnings_length_vec():
    # test_length_vec (simpli_test.py:327-358)
    _say('simulate: warnings')
    import warnings
    s = Simpli()
    s.voxel_size = 1
    s.set_voi([-10] * 3, [10] * 3)
    s.fiber_bundles = objects.FiberBundles([[[[-100, 0, 0, 10], [100, 0, 0, 10]]]],
                                           [[(0.333, -0.004, 10, 'p'), (0.666, 0, 5, 'b'),
                                             (1.0, 0.004, 1, 'r')]])
    tissue, optical_axis, tissue_properties = s.generate_tissue()
    s.filter_rotations = np.deg2rad([0, 30, 60, 90, 120, 150])
    s.light_intensity = 26000
    s.wavelength = 525
    s.untilt_sensor_view = False
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter('always')
        img = s.run_simulation(tissue, optical_axis, tissue_properties, 0, 0)
        assert len(w) == 0, [str(x.message) for x in w]
    assert not np.any(np.isnan(img))
    assert np.any(img != s.light_intensity / 4)
    optical_axis[tissue > 0, :] *= 1.0001
    with pytest.warns(UserWarning, match='not normalized'):
        s.run_simulation(tissue, optical_axis, tissue_properties, 0, 0)
    optical_axis[:] = 0
    with pytest.warns(UserWarning, match='optical_axis is 0'):
        s.run_simulation(tissue, optical_axis, tissue_properties, 0, 0)


def test_simulate_errors(golden):
    _say('simulate: error paths')
    s = _setup((24, 24, 24), 5.0)
    tissue = golden['tissue'].copy()
    props = golden['tissue_properties']
    tissue[3, 4, 5] = -1
    with pytest.raises(ValueError, match='label < 0'):
        s.run_simulation(tissue, golden['optical_axis'], props, 0, 0)
    tissue[3, 4, 5] = 4
    with pytest.raises(ValueError, match='label exceed'):
        s.run_simulation(tissue, golden['optical_axis'], props, 0, 0)
    bad = props.copy()
    bad[0, 0] = 0.1
    with pytest.raises(ValueError, match='background birefringence'):
        s.run_simulation(golden['tissue'], golden['optical_axis'], bad, 0, 0)
    bad = props.copy()
    bad[2, 1] = -1
    with pytest.raises(ValueError, match='mu < 0'):
        s.run_simulation(golden['tissue'], golden['optical_axis'], bad, 0, 0)
    with pytest.raises(ValueError, match='tissue_properties.shape'):
        s.run_simulation(golden['tissue'], golden['optical_axis'], props[:, :1], 0, 0)
    s.step_size = -1
    with pytest.raises(ValueError, match='step_size'):
        s.run_simulation(golden['tissue'], golden['optical_axis'], props, 0, 0)
    # the call after errors still works
    s.step_size = 1.0
    img = s.run_simulation(golden['tissue'], golden['optical_axis'], props, 0, 0)
    assert_intensity_close(img, golden['data'][0], 26000)


def test_simulate_device_tensors(golden):
    import torch
    _say('simulate: device tensors')
    s = _setup((24, 24, 24), 5.0)
    tissue = torch.from_numpy(golden['tissue']).cuda()
    axis = torch.from_numpy(golden['optical_axis']).cuda()
    imgs = s.run_simulation_tilts(tissue, axis, golden['tissue_properties'], GOLDEN_TILTS)
    assert imgs.is_cuda and imgs.shape == (5, 24, 24, 6)
    for t in range(5):
        assert_intensity_close(imgs[t].cpu().numpy(), golden['data'][t], 26000)


@pytest.mark.parametrize('interp', ['Slerp', 'Lerp'])
def test_simulate_dense_birefringence_queue(interp):
    """Interpolating views of a volume that is birefringent almost everywhere:
    every ray reaches its limit of queued steps, so the wave evaluates partial
    batches; rays of unequal length (tilted sensor) and a ray grid that is not a
    multiple of the wave width (23 x 29 pixels) end at different times."""
    _say(f'simulate: dense birefringent volume, {interp}, queued steps')
    rng = np.random.default_rng(5)
    dim = (23, 29, 40)
    tissue = np.where(rng.random(dim) < 0.93, 1 + (rng.random(dim) < 0.3), 0).astype(np.int32)
    # a smoothly varying field with noise, plus a few zero and non-unit axes
    x, y, z = np.meshgrid(*[np.arange(n) for n in dim], indexing='ij')
    a = 0.15 * x + 0.1 * z + 0.3 * rng.standard_normal(dim)
    b = 0.4 + 0.02 * y + 0.2 * rng.standard_normal(dim)
    axis = np.stack([np.cos(a) * np.sin(b), np.sin(a) * np.sin(b), np.cos(b)], -1)
    axis = axis.astype(np.float32)
    axis[rng.random(dim) < 0.01] = 0
    axis[tissue == 0] = 0
    props = np.array([[0.0, 0.5], [-0.004, 1.0], [0.006, 3.0]])
    for opts in (dict(), dict(untilt_sensor_view=False, tissue_refrection=1.2), dict(flip_z_beam=True)):
        s = _setup(dim, 0.5, interpolate=interp, rho=np.linspace(0, np.pi, 9, endpoint=False),
                   **opts)
        tilts = np.deg2rad([(0, 0), (5.5, 0), (8.0, 90), (12.0, 200), (30.0, 45)])
        with pytest.warns(UserWarning, match='optical_axis is 0'):
            imgs = s.run_simulation_tilts(tissue, axis, props, tilts)
        for t, (theta, phi) in enumerate(tilts):
            ref, _ = _oracle(s, tissue, axis, props, theta, phi)
            assert_intensity_close(imgs[t], ref, 26000)


@pytest.mark.parametrize('interp', ['NN', 'Slerp'])
def test_simulate_flip_z_is_an_index_mirror(golden, interp):
    """flip_z mirrors every read in z: equal, bit for bit, to the z-mirrored
    volume without it (simulator.cpp:427-428, 559-560)."""
    _say(f'simulate: flip_z mirror, {interp}')
    T, V, P = golden['tissue'], golden['optical_axis'], golden['tissue_properties']
    Tm, Vm = np.ascontiguousarray(T[:, :, ::-1]), np.ascontiguousarray(V[:, :, ::-1])
    a = _setup((24, 24, 24), 5.0, interpolate=interp, flip_z_beam=True).run_simulation_tilts(
        T, V, P, GOLDEN_TILTS)
    b = _setup((24, 24, 24), 5.0, interpolate=interp).run_simulation_tilts(Tm, Vm, P, GOLDEN_TILTS)
    assert np.array_equal(a, b)
