"""GPU ROFL (Simpli.apply_rofl) vs the oracle and the golden maps.

The GPU fit is the reference's: MINPACK lmdif with a forward-difference
Jacobian, or lmder with the reference's Dfun in grad mode, as
scipy.optimize.leastsq runs them (fpl_lmdif.h), from the same 6 x 6
brute-force start.  So the maps match the reference's golden maps at its own
regression bar (1e-5 relative, simpli_test.py:368-373), itermap equals the
oracle's nfev, and the angles agree with the oracle to 1e-6 (what remains is
the last bits of libm and of summation orders).  Parameter deviations
(numpy's BLAS products in the oracle) to 1e-6 relative.
"""

import sys

import numpy as np
import pytest

from oracle import rofl
from fastpli_amd.simulation import Simpli
from tests.conftest import GOLDEN
from tests.test_oracle_rofl import TILTS, golden_stack, single_voxel_stack

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _simpli():
    s = Simpli()
    s.tilts = TILTS
    return s


def test_rofl_golden_and_oracle():
    import os
    _say('rofl: golden stack')
    data = golden_stack()
    d, i, t, (dd, di, dt, f, n) = _simpli().apply_rofl(data)
    od, oi, ot, (odd, odi, odt, of, on) = rofl.rofl_map(data, TILTS[1, 0])
    _say(f'rofl: max |GPU - oracle| direction {np.max(np.abs(d - od)):.2e} inclination '
         f'{np.max(np.abs(i - oi)):.2e} t_rel {np.max(np.abs(t - ot)):.2e}; chi^2 GPU - oracle '
         f'max {np.max(f - of):.2e}')
    np.testing.assert_allclose(d, od, atol=1e-6, rtol=0)
    np.testing.assert_allclose(i, oi, atol=1e-6, rtol=0)
    np.testing.assert_allclose(t, ot, atol=1e-6, rtol=0)
    np.testing.assert_allclose(f, of, rtol=1e-6)
    for g, o in [(dd, odd), (di, odi), (dt, odt)]:
        np.testing.assert_allclose(g, o, rtol=1e-6)
    assert np.array_equal(n, on), 'itermap differs from the oracle nfev'
    for name, got in [('direction', np.rad2deg(d)), ('inclination', np.rad2deg(i)), ('trel', t)]:
        gold = np.load(os.path.join(GOLDEN, 'simpli_rep', f'rofl_{name}.npy'))
        np.testing.assert_allclose(got, gold, rtol=1e-5)


def test_rofl_grad_mode_golden_and_oracle():
    # grad_mode: lmder with the reference's Dfun, against scipy's
    _say('rofl: golden stack, grad_mode')
    data = golden_stack()
    d, i, t, (dd, di, dt, f, n) = _simpli().apply_rofl(data, grad_mode=True)
    od, oi, ot, (odd, odi, odt, of, on) = rofl.rofl_map(data, TILTS[1, 0], grad=True)
    for g, o in [(d, od), (i, oi), (t, ot)]:
        np.testing.assert_allclose(g, o, atol=1e-6, rtol=0)
    np.testing.assert_allclose(f, of, rtol=1e-6)
    for g, o in [(dd, odd), (di, odi), (dt, odt)]:
        np.testing.assert_allclose(g, o, rtol=1e-6)
    assert np.array_equal(n, on), 'itermap differs from the oracle nfev'


def test_rofl_single_voxel_closed_form():
    _say('rofl: single voxel closed form (GPU simulation + GPU fit)')
    s = Simpli()
    s.voxel_size = 60
    s.dim = [1, 1, 1]
    tissue = np.ones((1, 1, 1), dtype=np.int32)
    axis = np.zeros((1, 1, 1, 3), dtype=np.float32)
    axis[:] = [np.cos(np.deg2rad(45)), 0, np.sin(np.deg2rad(45))]
    props = np.array([[0, 0], [-0.001, 0]])
    s.filter_rotations = np.deg2rad([0, 30, 60, 90, 120, 150])
    s.light_intensity = 26000
    s.interpolate = 'Slerp'
    s.untilt_sensor_view = False
    s.wavelength = 525
    s.tilts = TILTS
    stack = []
    for theta, phi in TILTS:
        s.step_size = 1 / np.cos(theta)
        stack.append(s.run_simulation(tissue, axis, props, theta, phi))
    t_rel = 4 * s.voxel_size * abs(props[1][0]) / (s.wavelength / 1e3)
    # the reference's tolerances (rofl_test.py:55-61) on the FP64 stack
    d, i, t, _ = s.apply_rofl(single_voxel_stack())
    assert abs(d[0, 0] - np.pi) < 1e-10 or abs(d[0, 0]) < 1e-10
    assert abs(abs(i[0, 0]) - np.deg2rad(45)) < 1e-9
    assert abs(t[0, 0] - t_rel) < 1e-8
    # on the GPU's FP32-intensity stack the fit inherits the intensities'
    # 1e-5 relative tolerance
    d, i, t, _ = s.apply_rofl(np.array(stack))
    assert abs(d[0, 0] - np.pi) < 1e-6 or abs(d[0, 0]) < 1e-6
    assert abs(abs(i[0, 0]) - np.deg2rad(45)) < 1e-6
    assert abs(t[0, 0] - t_rel) < 1e-6


def test_rofl_many_pixels_mask_and_dtypes():
    _say('rofl: 4096 model pixels, mask, float32')
    rng = np.random.default_rng(11)
    mats = rofl.tilt_matrices(4, TILTS[1, 0])
    n = 64
    data = np.empty((5, n, n, 9))
    truth = np.empty((3, n, n))
    for a in range(n):
        for b in range(n):
            phi, alpha, tr = rng.uniform(0.05, 3.1), rng.uniform(-1.3, 1.3), rng.uniform(0.05, 0.85)
            d, al, tt = rofl.rotated_params(phi, alpha, tr, mats, TILTS[1, 0])
            data[:, a, b] = 1000.0 * (1.0 + rofl.intensities(d, al, tt, 9, 0).reshape(5, 9))
            truth[:, a, b] = phi, alpha, tr
    noisy = data + rng.normal(0, 2.0, data.shape)
    s = _simpli()
    d, i, t, (_, _, _, _, it) = s.apply_rofl(noisy)
    for a in range(0, n, 7):
        for b in range(0, n, 9):
            od, oi, ot, _, _, _, _, onf = rofl.fit(noisy[:, a, b], TILTS[1, 0])
            assert abs(d[a, b] - od) < 1e-6 and abs(i[a, b] - oi) < 1e-6
            assert abs(t[a, b] - ot) < 1e-6 and it[a, b] == onf
    # noise-free model pixels: the generating parameters come back
    d0, i0, t0, (_, _, _, f0, _) = s.apply_rofl(data[:, :8, :8])
    found = 0
    for a in range(8):
        for b in range(8):
            od, oi, ot, *_ = rofl.fit(data[:, a, b], TILTS[1, 0])
            assert abs(d0[a, b] - od) < 1e-9 and abs(i0[a, b] - oi) < 1e-9
            assert abs(t0[a, b] - ot) < 1e-9
            if f0[a, b] < 1e-16:
                found += 1
                assert abs(d0[a, b] - truth[0, a, b]) < 1e-9
                assert abs(abs(i0[a, b]) - abs(truth[1, a, b])) < 1e-9
                assert abs(t0[a, b] - truth[2, a, b]) < 1e-9
    assert found >= 60
    mask = np.zeros((n, n), bool)
    mask[3, 4] = True
    dm, _, _, _ = s.apply_rofl(noisy, mask=mask)
    assert np.isnan(dm[0, 0]) and dm[3, 4] == d[3, 4]
    d32, _, _, _ = s.apply_rofl(noisy.astype(np.float32))
    assert d32.dtype == np.float32


def test_rofl_errors():
    _say('rofl: errors')
    s = Simpli()
    with pytest.raises(ValueError, match='tilts not set'):
        s.apply_rofl(np.ones((5, 2, 2, 6)))
    s.tilts = np.deg2rad([(0, 0), (5.5, 0), (5.5, 90), (5.5, 180), (5.5, 90)])
    with pytest.raises(ValueError, match='not suitable'):
        s.apply_rofl(np.ones((5, 2, 2, 6)))
    s.tilts = TILTS
    with pytest.raises(ValueError, match='1 \\+ 4'):
        s.apply_rofl(np.ones((4, 2, 2, 6)))
    with pytest.raises(ValueError, match='at least 3'):
        s.apply_rofl(np.ones((5, 2, 2, 3)))
    with pytest.raises(TypeError):
        s.apply_rofl(np.ones((5, 2, 6)))
