"""Differentiable run_simulation: the ray march's adjoint behind torch.autograd.

``simulate_stack(simpli, tissue, optical_axis, tissue_properties, tilts)``
returns the image stack of ``Simpli.run_simulation_tilts`` (n_tilts, nx, ny,
n_rho), float64 on the GPU, as a node of the autograd graph.  Its backward is
the hand-written adjoint kernel (``fpl_simulate_adjoint``): gradients with
respect to ``tissue_properties`` (the (dn, mu) table, any float tensor) and
``optical_axis`` (float32 on the GPU).  The reference has no derivatives; the
function differentiated is its forward (src/simulation/simulator.cpp:182-411).
Labels whose dn is 0 take the reference's attenuation-only branch (:280-288)
and have a zero dn-gradient.  With NN a step's axis is its voxel's own; with
Lerp / Slerp (GetVec, :449-555) the gradient of a step's interpolated axis is
taken back through the interpolation tree to the corners of the ray's label
(the branches of the interpolation, such as the hemisphere flip, are held
fixed: the derivative is that of the branch the forward took).
"""

import numpy as np

from .. import _lib


def _torch():
    import torch
    return torch


class _March:
    """Namespace of the autograd.Function (defined lazily: torch is optional at import)."""
    fn = None


def _function():
    if _March.fn is not None:
        return _March.fn
    torch = _torch()

    class MarchFunction(torch.autograd.Function):

        @staticmethod
        def forward(ctx, props, optical_axis, tissue, simpli, tilts):
            setup, keep = simpli._setup()
            if setup.field_storage != 0:
                raise ValueError("simulate_stack needs field_storage='f32'")
            dim = simpli._dim_checked()
            props_np = props.detach().to('cpu', torch.float64).numpy()
            images = torch.empty((tilts.shape[0], dim[0], dim[1], setup.n_rho),
                                 dtype=torch.float64, device=tissue.device)
            warn = _lib.simulate(setup, dim, tissue, optical_axis.detach(), props_np, tilts,
                                 images)
            _lib.raise_warnings(warn)
            _lib.check_images(warn)
            ctx.save_for_backward(props, optical_axis, tissue)
            # the backward differentiates this forward: its setup (and the
            # rotations it points to), not the simpli's state at backward time
            ctx.setup, ctx.keep, ctx.tilts, ctx.dim = setup, keep, tilts, dim
            return images

        @staticmethod
        def backward(ctx, grad_images):
            props, axis, tissue = ctx.saved_tensors
            g = grad_images.to(torch.float64).contiguous()
            gvec = torch.empty_like(axis) if ctx.needs_input_grad[1] else None
            gp = _lib.simulate_adjoint(ctx.setup, ctx.dim, tissue, axis.detach(),
                                       props.detach().to('cpu', torch.float64).numpy(),
                                       ctx.tilts, g, gvec)
            gprops = None
            if ctx.needs_input_grad[0]:
                gprops = torch.from_numpy(gp).to(device=props.device, dtype=props.dtype)
            return gprops, gvec, None, None, None

    _March.fn = MarchFunction
    return MarchFunction


def simulate_stack(simpli, tissue, optical_axis, tissue_properties, tilts):
    """Differentiable image stack (n_tilts, nx, ny, n_rho) of ``simpli``'s setup.

    tissue: int32 tensor (nx, ny, nz) on the GPU; optical_axis: float32 tensor
    (nx, ny, nz, 3) on the GPU (may require grad);
    tissue_properties: float tensor (n, 2) = (dn, mu) (may require grad);
    tilts: (n_tilts, 2) = (theta, phi) in radian.
    """
    torch = _torch()
    simpli._check_volume_input()
    if not (torch.is_tensor(tissue) and torch.is_tensor(optical_axis)):
        raise TypeError('simulate_stack takes torch tensors on the GPU')
    if tissue.dtype != torch.int32 or optical_axis.dtype != torch.float32:
        raise TypeError('tissue must be int32 and optical_axis float32')
    if tissue.device.type != 'cuda' or optical_axis.device.type != 'cuda':
        raise TypeError('tissue and optical_axis must be on the GPU')
    if tuple(optical_axis.shape) != tuple(tissue.shape) + (3,):
        raise ValueError('not np.equal(tissue.shape, optical_axis.shape[:-1])')
    if tuple(tissue.shape) != tuple(int(d) for d in simpli._dim_checked()):
        raise ValueError('tissue.shape does not match dim')
    props = tissue_properties
    if not torch.is_tensor(props):
        props = torch.as_tensor(np.asarray(props, dtype=np.float64))
    if props.dim() != 2 or props.shape[1] != 2:
        raise ValueError('tissue_properties.shape[1] != 2')
    tilts = np.array(tilts, dtype=np.float64, ndmin=2).reshape(-1, 2)
    return _function().apply(props, optical_axis.contiguous(), tissue.contiguous(), simpli,
                             tilts)
