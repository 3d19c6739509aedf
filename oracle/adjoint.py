"""FP64 reference gradient of the ray march -- TEST INFRASTRUCTURE.

A torch float64 (CPU) restatement of ``PliSimulator::RunSimulation``
(src/simulation/simulator.cpp:182-411) for views with the untilted-sensor
start, written as the reference writes it and vectorized over the rays of a
tilt:

  * starts (i + 1/2, j + 1/2) - shift / 2 with shift = nz dir / cos(theta), then
    + step / 2, and `pos += step` while floor(z) lies in [0, nz)
    (:227-231, :259-261, :721-773);
  * GetLabel / GetVec NN: floor of the position, label 0 outside (:420-443,
    :557-579); GetVec Lerp / Slerp (:449-555): the eight corners around
    position - 1/2 (clamped), those of another label zeroed, seven
    interpolations x, y, z with LinearOrientationInterpolation /
    SphericalInterpolation (zero inputs, hemisphere flip, the 1e-12 identity
    test, acos and sin);
  * the attenuation-only branch for dn == 0 or label 0 (:280-288), the skipped
    zero axis (:292-297), and otherwise the per-rho Stokes update with asin,
    atan2, sin and cos (:309-343) and exp(-mu 1e3 t) attenuation (:276-278);
  * the output row of PolY(1): (s0 - s1) / 2 (:355-360); the input state
    RetarderMatrix(pi/2, -pi/2) PolX(1) (1, 0, 0, 0) I0 = (I0/2, 0, 0, -I0/2)
    (:207-214).

torch autograd differentiates it exactly, which gives the FP64 gradient of
L = sum(g * images) with respect to the (dn, mu) table and the optical-axis
field.  Only tests import this module, as the checker of the adjoint kernel;
its forward is pinned against the C++ oracle (``oracle.simulate``).
"""

import math

import numpy as np
import torch


def _pi_sin(v):
    """vm::rot_pi_cases::sin (vemath.hpp:499-508)."""
    v = math.fmod(v, 2 * math.pi)
    if v in (0.0, math.pi, -math.pi):
        return 0.0
    if v in (math.pi / 2, -(math.pi + math.pi / 2)):
        return 1.0
    if v in (math.pi + math.pi / 2, -math.pi / 2):
        return -1.0
    return math.sin(v)


def _pi_cos(v):
    """vm::rot_pi_cases::cos (vemath.hpp:510-520)."""
    v = math.fmod(v, 2 * math.pi)
    if v == 0.0:
        return 1.0
    if v in (math.pi / 2, math.pi + math.pi / 2, -math.pi / 2, -(math.pi + math.pi / 2)):
        return 0.0
    if v in (math.pi, -math.pi):
        return -1.0
    return math.cos(v)


def _rot_zyz(phi, theta):
    """vm::rot_pi_cases::Mat3RotZYZ(phi, theta, -phi) (vemath.hpp:522-544)."""
    def rz(a):
        s, c = _pi_sin(a), _pi_cos(a)
        return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])

    def ry(a):
        s, c = _pi_sin(a), _pi_cos(a)
        return np.array([[c, 0, s], [0, 1.0, 0], [-s, 0, c]])

    return rz(phi) @ (ry(theta) @ rz(-phi))


def _dot(a, b):
    return (a * b).sum(dim=1)


def _lerp(v, u, t):
    """LinearOrientationInterpolation (simulator.cpp:449-471) over rays."""
    vz, uz = (v == 0).all(dim=1), (u == 0).all(dim=1)
    a, b = v * (1 - t)[:, None], u * t[:, None]
    s = torch.where((_dot(a, b) >= 0)[:, None], a + b, a - b)
    s = torch.where(uz[:, None], v, s)
    return torch.where(vz[:, None], u, s)


def _slerp(v, u, t):
    """SphericalInterpolation (simulator.cpp:473-494) over rays.  The acos
    branch gets fixed orthogonal inputs where it is not taken, so that its
    unused gradients stay finite."""
    vz, uz = (v == 0).all(dim=1), (u == 0).all(dim=1)
    u = torch.where((_dot(v, u) < 0)[:, None], -u, u)
    same = ((v - u) ** 2).sum(dim=1) < 1e-12
    take = (~vz & ~uz & ~same)[:, None]
    vs = torch.where(take, v, torch.tensor([1.0, 0.0, 0.0], dtype=v.dtype))
    us = torch.where(take, u, torch.tensor([0.0, 1.0, 0.0], dtype=v.dtype))
    r = _dot(vs, us) / (torch.sqrt(_dot(vs, vs)) * torch.sqrt(_dot(us, us)))
    th = torch.acos(torch.clamp(r, -1.0, 1.0))
    st = torch.sin(th)
    s = vs * (torch.sin((1 - t) * th) / st)[:, None] + us * (torch.sin(t * th) / st)[:, None]
    s = torch.where(take, s, v)
    s = torch.where(uz[:, None], v, s)
    return torch.where(vz[:, None], u, s)


def _get_vec(lab, axis, l, px, py, pz, interpolate):
    """GetVec (simulator.cpp:496-579) for every ray at (px, py, pz); l = the
    rays' own labels."""
    nx, ny, nz = lab.shape
    ix, iy = np.floor(px).astype(np.int64), np.floor(py).astype(np.int64)
    iz = int(math.floor(pz))
    cx, cy = np.clip(ix, 0, nx - 1), np.clip(iy, 0, ny - 1)
    if interpolate == 'NN':
        return axis[cx, cy, iz]

    def corners(p, n):
        lo = np.clip(np.floor(p - 0.5), 0, n - 1).astype(np.int64)
        hi = np.clip(np.ceil(p - 0.5), 0, n - 1).astype(np.int64)
        d = np.where(lo == hi, 0.0, (p - lo - 0.5) / np.where(lo == hi, 1, hi - lo))
        return lo, hi, d

    x0, x1, xd = corners(px, nx)
    y0, y1, yd = corners(py, ny)
    z0, z1, zd = corners(np.full(px.shape, pz), nz)
    xd, yd, zd = (torch.as_tensor(a) for a in (xd, yd, zd))

    def c(x, y, z):
        return axis[x, y, z] * (lab[x, y, z] == l)[:, None].to(axis.dtype)

    ip = _slerp if interpolate == 'Slerp' else _lerp
    c00 = ip(c(x0, y0, z0), c(x1, y0, z0), xd)
    c01 = ip(c(x0, y0, z1), c(x1, y0, z1), xd)
    c10 = ip(c(x0, y1, z0), c(x1, y1, z0), xd)
    c11 = ip(c(x0, y1, z1), c(x1, y1, z1), xd)
    v = ip(ip(c00, c10, yd), ip(c01, c11, yd), zd)
    same = torch.as_tensor((x0 == x1) & (y0 == y1) & (z0 == z1))
    v = torch.where(same[:, None], axis[x0, y0, z0], v)
    return torch.where((l == 0)[:, None], torch.zeros_like(v), v)


def simulate(tissue, optical_axis, props, tilts, *, voxel_size, wavelength=525.0,
             light_intensity=26000.0, filter_rotations=None, step_size=1.0,
             interpolate='NN'):
    """Images (n_tilts, nx, ny, n_rho) float64 as a differentiable torch tensor.

    tissue: integer array (nx, ny, nz); optical_axis: torch float64
    (nx, ny, nz, 3); props: torch float64 (n, 2) = (dn, mu).  Either tensor
    may require grad.
    """
    lab = torch.as_tensor(np.asarray(tissue), dtype=torch.long)
    nx, ny, nz = lab.shape
    rho = torch.as_tensor(np.asarray(filter_rotations, dtype=np.float64))
    n_rho = rho.numel()
    lam = wavelength * 1e-9
    thick = voxel_size * 1e-6 * step_size
    ii, jj = np.meshgrid(np.arange(nx, dtype=np.float64), np.arange(ny, dtype=np.float64),
                         indexing='ij')
    ii, jj = ii.ravel(), jj.ravel()
    n_rays = ii.size
    out = []
    for theta, phi in tilts:
        theta, phi = float(theta), float(phi)
        rot = torch.as_tensor(_rot_zyz(phi, theta))
        d = np.array([_pi_cos(phi) * _pi_sin(theta), _pi_sin(phi) * _pi_sin(theta),
                      _pi_cos(theta)])
        step = d * step_size
        shift = d / math.cos(theta) * nz
        px = (ii + 0.5) - 0.5 * shift[0]
        py = (jj + 0.5) - 0.5 * shift[1]
        pz = np.zeros(n_rays)
        px, py, pz = px + 0.5 * step[0], py + 0.5 * step[1], pz + 0.5 * step[2]
        s = torch.zeros((n_rays, n_rho, 4), dtype=torch.float64)
        s[:, :, 0] = light_intensity / 2
        s[:, :, 3] = -light_intensity / 2
        while 0 <= math.floor(pz[0]) < nz:
            ix = np.floor(px).astype(np.int64)
            iy = np.floor(py).astype(np.int64)
            iz = int(math.floor(pz[0]))
            inside = (ix >= 0) & (ix < nx) & (iy >= 0) & (iy < ny)
            cx, cy = np.clip(ix, 0, nx - 1), np.clip(iy, 0, ny - 1)
            l = torch.where(torch.as_tensor(inside), lab[cx, cy, iz], 0)
            dn = props[l, 0]
            mu = props[l, 1] * 1e3
            att = torch.exp(-mu * thick)
            bg = (dn.detach() == 0) | (l == 0)
            v = _get_vec(lab, optical_axis, l, px, py, pz[0], interpolate)
            zero = (v.detach() == 0).all(dim=1)
            act = ~bg & ~zero
            # the retarder branch on a safe axis where it is not taken, so that
            # its unused gradients stay finite
            vs = torch.where(act[:, None], v, torch.tensor([1.0, 0.0, 0.0], dtype=v.dtype))
            d_rel = dn * 4.0 * thick / lam * torch.sqrt((vs * vs).sum(dim=1))
            if theta != 0:
                vs = vs @ rot.T
            alpha = torch.asin(vs[:, 2] / torch.sqrt((vs * vs).sum(dim=1)))
            ret = math.pi / 2 * d_rel * torch.cos(alpha) ** 2
            sin_r, cos_r = torch.sin(ret)[:, None], torch.cos(ret)[:, None]
            ph = torch.atan2(vs[:, 1], vs[:, 0])
            beta = 2 * (rho[None, :] - ph[:, None])
            sb, cb = torch.sin(-beta), torch.cos(-beta)
            a1 = s[:, :, 1] * cb - s[:, :, 2] * sb
            c1 = s[:, :, 1] * sb + s[:, :, 2] * cb
            b1 = c1 * cos_r + s[:, :, 3] * sin_r
            ret_state = torch.stack([s[:, :, 0], a1 * cb + b1 * sb, -a1 * sb + b1 * cb,
                                     -c1 * sin_r + s[:, :, 3] * cos_r], dim=2)
            a = att[:, None, None]
            s = torch.where(act[:, None, None], ret_state * a,
                            torch.where(bg[:, None, None], s * a, s))
            px, py, pz = px + step[0], py + step[1], pz + step[2]
        out.append((0.5 * (s[:, :, 0] - s[:, :, 1])).reshape(nx, ny, n_rho))
    return torch.stack(out)


def gradient(tissue, optical_axis, props, tilts, weights, **setup):
    """(dL/d props (n, 2), dL/d optical_axis (nx, ny, nz, 3)) of L = sum(weights * images).

    numpy float64 in and out; weights has the shape of the images.
    """
    axis = torch.tensor(np.asarray(optical_axis, dtype=np.float64), requires_grad=True)
    pr = torch.tensor(np.asarray(props, dtype=np.float64), requires_grad=True)
    images = simulate(tissue, axis, pr, tilts, **setup)
    loss = (images * torch.as_tensor(np.asarray(weights, dtype=np.float64))).sum()
    loss.backward()
    return pr.grad.numpy(), axis.grad.numpy(), images.detach().numpy()
