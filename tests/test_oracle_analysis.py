"""Pins of the optic / EPA oracle against the reference's golden vectors (CPU)."""

import numpy as np

from oracle import analysis

# golden run: voxel 5 um, pixel 20 um, optical_sigma 0.71, noise model
# negative_binomial(x / (3 - 1), 1 / 3) after np.random.seed(42)
# (tests/simulation/simpli_rep.py:10, 56-62)


def test_optic_and_noise_reproduce_golden(golden):
    import os
    from tests.conftest import GOLDEN
    np.random.seed(42)
    for t in range(5):
        res = analysis.optic_resample(golden['data'][t], 0.71, 5.0, 20.0)
        assert res.shape == (6, 6, 6)
        noisy = np.random.negative_binomial(res / (3 - 1), 1 / 3)
        gold = np.load(os.path.join(GOLDEN, 'simpli_rep', f'optic_{t}.npy'))
        assert np.array_equal(noisy, gold)


def test_epa_reproduces_golden():
    import os
    from tests.conftest import GOLDEN
    for t in range(5):
        gold = lambda m: np.load(os.path.join(GOLDEN, 'simpli_rep', f'epa_{t}_{m}.npy'))
        optic = np.load(os.path.join(GOLDEN, 'simpli_rep', f'optic_{t}.npy'))
        tr, d, r = analysis.epa(optic)
        np.testing.assert_allclose(tr, gold('transmittance'), rtol=1e-12)
        np.testing.assert_allclose(np.rad2deg(d), gold('direction'), rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(r, gold('retardation'), rtol=1e-12)


def test_epa_closed_form():
    # I(rho) = T/2 (1 + R sin(2 (rho - phi))): transmittance T, retardation R,
    # direction phi (mod pi)
    rho = np.linspace(0, np.pi, 9, endpoint=False)
    T, R, phi = 1000.0, 0.3, 0.4
    img = T / 2 * (1 + R * np.sin(2 * (rho - phi)))
    t, d, r = analysis.epa(img[None, :])
    assert np.allclose(t, T) and np.allclose(r, R) and np.allclose(d, phi)


def test_resample_closed_form():
    # constant image stays constant; 4x4 -> 2x2 averages blocks
    img = np.arange(16, dtype=float).reshape(4, 4)
    out = analysis.resample(img, 0.5)
    assert np.allclose(out, [[2.5, 4.5], [10.5, 12.5]])
