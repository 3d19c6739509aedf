"""GPU adjoint of the ray march (fpl_simulate_adjoint, simulate_stack).

Checked against the FP64 reference gradient (torch autograd of the FP64
restatement in oracle/adjoint.py), against central differences of the GPU
forward itself (h = 1e-3 relative), and through torch.autograd.
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


def _small_tissue(seed=0, dim=(24, 20, 16), voxel=0.5):
    """A vMF bundle (kappa = 10) in a small volume, generated on the GPU."""
    s = Simpli()
    s.voxel_size = voxel
    s.dim = np.array(dim)
    s.fiber_bundles = synthetic.single_bundle(seed=seed, n_fibers=40,
                                              size=tuple(np.array(dim) * voxel))
    s.filter_rotations = ROT9
    s.light_intensity = 26000
    s.wavelength = 525
    s.interpolate = 'NN'
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


def test_adjoint_matches_fp64_reference_gradient():
    _say('adjoint: vs FP64 reference gradient, vMF bundle 24 x 20 x 16, NN, 5 tilts x 9 rot')
    s, label, vector, props = _small_tissue()
    props = props.copy()
    props[0, 1] = 0.3  # background absorption, so that label 0 has a mu-gradient too
    rng = np.random.default_rng(1)
    w = rng.normal(size=(5,) + label.shape[:2] + (9,))
    gp, gv = _adjoint(s, label, vector, props, w)
    rp, rv, _ = ref.gradient(label, vector, props, TILTS5, w, voxel_size=0.5,
                             wavelength=525.0, light_intensity=26000.0, filter_rotations=ROT9)
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


def test_adjoint_matches_central_differences():
    _say('adjoint: vs central differences (h = 1e-3) of the GPU forward')
    s, label, vector, props = _small_tissue(seed=2)
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
    # props-only gradients also work with Slerp views
    s.interpolate = 'Slerp'
    pr2 = torch.tensor(props, dtype=torch.float64, requires_grad=True)
    img2 = simulate_stack(s, lab, torch.from_numpy(vector).to(dev), pr2, TILTS5)
    img2.sum().backward()
    assert torch.isfinite(pr2.grad).all()
    with pytest.raises(ValueError, match='NN'):
        axis2 = torch.from_numpy(vector).to(dev).requires_grad_(True)
        simulate_stack(s, lab, axis2, pr2.detach(), TILTS5).sum().backward()
