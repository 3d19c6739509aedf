"""GPU adjoint of the ray march (fpl_simulate_adjoint, simulate_stack).

Checked against the FP64 reference gradient (torch autograd of the FP64
restatement in oracle/adjoint.py) for NN, Lerp and Slerp views, against
central differences of the GPU forward itself (h = 1e-3 relative), and
through torch.autograd.
"""

import sys
import warnings

import numpy as np
import pytest
import torch

from fastpli_amd import _lib, synthetic
from fastpli_amd.simulation import Simpli, simulate_stack
from oracle import adjoint as ref

pytestmark = pytest.mark.gpu

ROT9 = np.linspace(0, np.pi, 9, endpoint=False)
TILTS5 = np.deg2rad([(0, 0), (5.5, 0), (5.5, 90), (5.5, 180), (5.5, 270)])


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _small_tissue(seed=0, dim=(24, 20, 16), voxel=0.5, interp='NN'):
    """A vMF bundle (kappa = 10) in a small volume, generated on the GPU."""
    s = Simpli()
    s.voxel_size = voxel
    s.dim = np.array(dim)
    s.fiber_bundles = synthetic.single_bundle(seed=seed, n_fibers=40,
                                              size=tuple(np.array(dim) * voxel))
    s.filter_rotations = ROT9
    s.light_intensity = 26000
    s.wavelength = 525
    s.interpolate = interp
    with warnings.catch_warnings():
        warnings.simplefilter('ignore')
        label, vector, props = s.generate_tissue()
    return s, label, vector, props


def _adjoint(s, label, vector, props, w, with_axis=True):
    dev = torch.device('cuda', 0)
    lab = torch.from_numpy(label).to(dev)
    vec = torch.from_numpy(vector).to(dev)
    g = torch.from_numpy(np.ascontiguousarray(w)).to(dev)
    gv = torch.empty_like(vec) if with_axis else None
    setup, _keep = s._setup()
    gp = _lib.simulate_adjoint(setup, s.dim, lab, vec, props, TILTS5, g, gv)
    return gp, (gv.cpu().numpy() if with_axis else None)


@pytest.mark.parametrize('interp', ['NN', 'Lerp', 'Slerp'])
def test_adjoint_matches_fp64_reference_gradient(interp):
    _say(f'adjoint: vs FP64 reference gradient, vMF bundle 24 x 20 x 16, {interp}, '
         '5 tilts x 9 rot')
    s, label, vector, props = _small_tissue(interp=interp)
    props = props.copy()
    props[0, 1] = 0.3  # background absorption, so that label 0 has a mu-gradient too
    rng = np.random.default_rng(1)
    w = rng.normal(size=(5,) + label.shape[:2] + (9,))
    gp, gv = _adjoint(s, label, vector, props, w)
    rp, rv, _ = ref.gradient(label, vector, props, TILTS5, w, voxel_size=0.5,
                             wavelength=525.0, light_intensity=26000.0, filter_rotations=ROT9,
                             interpolate=interp)
    scale_p = np.abs(rp).max()
    for l in range(props.shape[0]):
        for c in range(2):
            # FP32 march vs FP64: 1e-3 of the largest table entry
            assert abs(gp[l, c] - rp[l, c]) <= 1e-3 * scale_p + 1e-3 * abs(rp[l, c]), \
                (l, c, gp[l, c], rp[l, c])
    # the attenuation-only labels (dn == 0) have no dn-gradient
    for l in range(props.shape[0]):
        if props[l, 0] == 0:
            assert gp[l, 0] == 0.0
    err = np.linalg.norm(gv - rv) / np.linalg.norm(rv)
    assert err < 1e-3, err
    assert np.abs(gv - rv).max() < 1e-2 * np.abs(rv).max()


@pytest.mark.parametrize('interp', ['NN', 'Slerp'])
def test_adjoint_matches_central_differences(interp):
    _say(f'adjoint: vs central differences (h = 1e-3) of the GPU forward, {interp}')
    s, label, vector, props = _small_tissue(seed=2, interp=interp)
    props = props.copy()
    props[0, 1] = 0.3
    # positive weights that vary over pixels, tilts and rotations (a sum over
    # the equidistant rotations alone sees the attenuation only)
    ii, jj = np.meshgrid(np.arange(label.shape[0]), np.arange(label.shape[1]), indexing='ij')
    w = ((1.0 + 0.5 * np.cos(0.4 * ii + 0.3 * jj))[None, :, :, None] *
         (1.0 + 0.5 * np.cos(2 * ROT9 + 0.7))[None, None, None, :] *
         (1.0 + 0.2 * np.arange(5))[:, None, None, None])
    gp, gv = _adjoint(s, label, vector, props, w)

    def loss(p, vec):
        with warnings.catch_warnings():
            warnings.simplefilter('ignore')
            img = s.run_simulation_tilts(label, vec, p, TILTS5)
        return float((img * w).sum())

    for l in range(props.shape[0]):
        for c in range(2):
            if props[l, c] == 0:
                continue
            h = 1e-3 * abs(props[l, c])
            pp, pm = props.copy(), props.copy()
            pp[l, c] += h
            pm[l, c] -= h
            fd = (loss(pp, vector) - loss(pm, vector)) / (2 * h)
            assert np.isclose(gp[l, c], fd, rtol=5e-3, atol=1e-4 * np.abs(gp).max()), \
                (l, c, gp[l, c], fd)
    # the axis field rotated by eps about a fixed direction: dL/d eps = sum gv . (w x v)
    om = np.array([0.3, -0.5, 0.8])
    om /= np.linalg.norm(om)
    eps = 1e-3

    def rotated(e):
        k = np.array([[0, -om[2], om[1]], [om[2], 0, -om[0]], [-om[1], om[0], 0]])
        R = np.eye(3) + np.sin(e) * k + (1 - np.cos(e)) * (k @ k)
        return np.ascontiguousarray((vector.astype(np.float64) @ R.T).astype(np.float32))

    fd = (loss(props, rotated(eps)) - loss(props, rotated(-eps))) / (2 * eps)
    an = float((gv * np.cross(om, vector.astype(np.float64))).sum())
    assert np.isclose(an, fd, rtol=5e-3), (an, fd)


def test_simulate_stack_autograd():
    _say('adjoint: simulate_stack through torch.autograd')
    s, label, vector, props = _small_tissue(seed=3)
    dev = torch.device('cuda', 0)
    lab = torch.from_numpy(label).to(dev)
    axis = torch.from_numpy(vector).to(dev).requires_grad_(True)
    pr = torch.tensor(props, dtype=torch.float64, requires_grad=True)
    images = simulate_stack(s, lab, axis, pr, TILTS5)
    with warnings.catch_warnings():
        warnings.simplefilter('ignore')
        plain = s.run_simulation_tilts(label, vector, props, TILTS5)
    assert np.array_equal(images.detach().cpu().numpy(), plain)
    rng = np.random.default_rng(5)
    w = torch.from_numpy(rng.normal(size=images.shape)).to(dev)
    (images * w).sum().backward()
    gp, gv = _adjoint(s, label, vector, props, w.cpu().numpy())
    # the same kernel; only the order of its atomic sums differs between calls
    assert np.allclose(pr.grad.numpy(), gp, rtol=1e-9, atol=0)
    assert np.allclose(axis.grad.cpu().numpy(), gv, rtol=1e-5, atol=1e-6 * np.abs(gv).max())
    # Slerp views: axis gradients through the interpolation, as the C entry gives them
    s.interpolate = 'Slerp'
    pr2 = torch.tensor(props, dtype=torch.float64, requires_grad=True)
    axis2 = torch.from_numpy(vector).to(dev).requires_grad_(True)
    img2 = simulate_stack(s, lab, axis2, pr2, TILTS5)
    (img2 * w).sum().backward()
    gp2, gv2 = _adjoint(s, label, vector, props, w.cpu().numpy())
    assert torch.isfinite(pr2.grad).all() and torch.isfinite(axis2.grad).all()
    assert np.allclose(pr2.grad.numpy(), gp2, rtol=1e-9, atol=0)
    assert np.allclose(axis2.grad.cpu().numpy(), gv2, rtol=1e-5, atol=1e-6 * np.abs(gv2).max())


@pytest.mark.parametrize('interp', ['NN', 'Slerp'])
def test_adjoint_central_differences_full_size(interp):
    # one vMF bundle of 400 fibers, 300 x 300 x 150 voxels @ 0.2 um,
    # 5 tilts x 9 rotations: dL/d dn and dL/d mu of the myelin label and the
    # derivative along a rotation of the whole axis field, against central
    # differences of the GPU forward
    _say(f'adjoint: vMF bundle 300 x 300 x 150 central differences, {interp}')
    dim = (300, 300, 150)
    s = Simpli()
    s.voxel_size = 0.2
    s.dim = np.array(dim)
    s.fiber_bundles = synthetic.single_bundle(seed=0, size=tuple(np.array(dim) * 0.2))
    s.filter_rotations = ROT9
    s.light_intensity = 26000
    s.wavelength = 525
    s.interpolate = interp
    with warnings.catch_warnings():
        warnings.simplefilter('ignore')
        label, vector, props = s.generate_tissue()
    ii, jj = np.meshgrid(np.arange(dim[0]), np.arange(dim[1]), indexing='ij')
    w = ((1.0 + 0.5 * np.cos(0.05 * ii + 0.03 * jj))[None, :, :, None] *
         (1.0 + 0.5 * np.cos(2 * ROT9 + 0.7))[None, None, None, :] *
         (1.0 + 0.2 * np.arange(5))[:, None, None, None])
    gp, gv = _adjoint(s, label, vector, props, w)
    dev = torch.device('cuda', 0)
    lab_d = torch.from_numpy(label).to(dev)
    w_d = torch.from_numpy(w).to(dev)

    def loss(p, vec):
        with warnings.catch_warnings():
            warnings.simplefilter('ignore')
            img = s.run_simulation_tilts(lab_d, torch.from_numpy(vec).to(dev), p, TILTS5)
        return float((img * w_d).sum())

    p_lab = 2  # the bundle's myelin ('p') layer: label 1 + layer 1
    assert props[p_lab, 0] != 0
    for (l, c) in [(p_lab, 0), (p_lab, 1), (1, 1)]:
        h = 1e-3 * abs(props[l, c])
        pp, pm = props.copy(), props.copy()
        pp[l, c] += h
        pm[l, c] -= h
        fd = (loss(pp, vector) - loss(pm, vector)) / (2 * h)
        assert np.isclose(gp[l, c], fd, rtol=5e-3), (l, c, gp[l, c], fd)
    om = np.array([0.3, -0.5, 0.8])
    om /= np.linalg.norm(om)
    k = np.array([[0, -om[2], om[1]], [om[2], 0, -om[0]], [-om[1], om[0], 0]])

    def rotated(e):
        R = np.eye(3) + np.sin(e) * k + (1 - np.cos(e)) * (k @ k)
        return np.ascontiguousarray((vector.astype(np.float64) @ R.T).astype(np.float32))

    eps = 1e-3
    fd = (loss(props, rotated(eps)) - loss(props, rotated(-eps))) / (2 * eps)
    an = float((gv.astype(np.float64) * np.cross(om, vector.astype(np.float64))).sum())
    assert np.isclose(an, fd, rtol=5e-3), (an, fd)
