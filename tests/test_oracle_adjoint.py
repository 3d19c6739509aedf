"""The FP64 reference gradient (oracle/adjoint.py): its forward against the C++
oracle, its autograd gradient against central differences of itself."""

import numpy as np
import pytest
import torch

import oracle
from oracle import adjoint

ROT = np.linspace(0, np.pi, 9, endpoint=False)
TILTS = [(0.0, 0.0), (np.deg2rad(5.5), 0.0), (np.deg2rad(5.5), np.pi / 2), (0.3, 2.0)]
SETUP = dict(voxel_size=0.5, wavelength=525.0, light_intensity=26000.0, filter_rotations=ROT)


def _fields(seed, dim=(7, 6, 9)):
    """Labels 0..3 (label 2 on the attenuation-only branch, dn = 0), unit axes
    with a few zero ones, a few non-unit ones."""
    rng = np.random.default_rng(seed)
    label = rng.integers(0, 4, size=dim).astype(np.int32)
    v = rng.normal(size=dim + (3,))
    v /= np.linalg.norm(v, axis=-1, keepdims=True)
    v[rng.uniform(size=dim) < 0.05] = 0.0
    v[rng.uniform(size=dim) < 0.05] *= 1.2
    props = np.array([[0.0, 0.5], [-0.004, 2.0], [0.0, 5.0], [0.006, 1.0]])
    return label, v.astype(np.float32), props


@pytest.mark.parametrize('interp', ['NN', 'Lerp', 'Slerp'])
@pytest.mark.parametrize('seed', [0, 1])
def test_forward_matches_cpp_oracle(seed, interp):
    label, vec, props = _fields(seed)
    got = adjoint.simulate(label, torch.tensor(vec, dtype=torch.float64),
                           torch.tensor(props), TILTS, interpolate=interp, **SETUP).numpy()
    for t, (theta, phi) in enumerate(TILTS):
        ref, _ = oracle.simulate(label, vec, props, theta, phi, voxel_size=0.5,
                                 interpolate=interp, filter_rotations=ROT)
        assert np.allclose(got[t], ref, rtol=1e-12, atol=1e-9), (t, np.abs(got[t] - ref).max())


@pytest.mark.parametrize('interp', ['NN', 'Lerp', 'Slerp'])
def test_gradient_matches_central_differences(interp):
    label, vec, props = _fields(3)
    rng = np.random.default_rng(4)
    w = rng.normal(size=(len(TILTS),) + label.shape[:2] + (ROT.size,))
    gp, gv, _ = adjoint.gradient(label, vec, props, TILTS, w, interpolate=interp, **SETUP)

    def loss(p, v):
        with torch.no_grad():
            img = adjoint.simulate(label, torch.tensor(v), torch.tensor(p), TILTS,
                                   interpolate=interp, **SETUP)
        return float((img.numpy() * w).sum())

    vec64 = vec.astype(np.float64)
    # dn of the birefringent labels, mu of every label
    for (l, c) in [(1, 0), (3, 0), (0, 1), (1, 1), (2, 1), (3, 1)]:
        h = 1e-6 * max(abs(props[l, c]), 1e-3)
        pp, pm = props.copy(), props.copy()
        pp[l, c] += h
        pm[l, c] -= h
        fd = (loss(pp, vec64) - loss(pm, vec64)) / (2 * h)
        assert np.isclose(gp[l, c], fd, rtol=1e-5, atol=1e-6 * np.abs(gp).max()), (l, c)
    # dn of the attenuation-only label does not enter
    assert gp[2, 0] == 0.0 and gp[0, 0] == 0.0
    # axis components of voxels on the retarder branch
    idx = np.argwhere((label == 1) | (label == 3))
    for x, y, z in idx[rng.choice(len(idx), 6, replace=False)]:
        if not np.any(vec[x, y, z]):
            continue
        for c in range(3):
            vp, vm = vec64.copy(), vec64.copy()
            vp[x, y, z, c] += 1e-6
            vm[x, y, z, c] -= 1e-6
            fd = (loss(props, vp) - loss(props, vm)) / 2e-6
            assert np.isclose(gv[x, y, z, c], fd, rtol=1e-5,
                              atol=1e-6 * np.abs(gv).max()), (x, y, z, c)
