"""ROFL tilt analysis, restated in numpy + scipy (oracle: test infrastructure only).

Follows src/fastpli/analysis/rofl.py:17-60 and the local-fit path of
src/fastpli/analysis/_ROFL_with_jacobi.py (Schmitz et al., Front. Neuroanat.
2018, doi:10.3389/fnana.2018.00075): centered intensities and weights
(:47-57), a 6 x 6 brute-force start over inclination and t_rel (:313-370),
scipy.optimize.leastsq (MINPACK lmdif with finite differences, or lmder with
the analytic Jacobian in grad mode; maxfev 150) on the weighted residual (:383-400, :433-475), angle symmetrization (:60-69) and
the covariance from the Jacobian at the minimum (:489-522).  The reference
wraps the model in numba; the arithmetic here is the same numpy arithmetic.

Only tests import this module, as the checker of the GPU kernel.
"""

import itertools

import numpy as np
from scipy import optimize


def epa_direction(data):
    """epa.direction (src/fastpli/analysis/epa.py:48-75) of one (n_rho,) curve."""
    data = np.asarray(data, dtype=np.float64)
    n = data.shape[-1]
    rho_2 = 2 * np.linspace(0, np.pi, n, False)
    a1 = 2 * np.sum(data * np.sin(rho_2), -1) / n
    b1 = 2 * np.sum(data * np.cos(rho_2), -1) / n
    d = 0.5 * np.arctan2(-b1, a1) + np.pi
    return d % np.pi


def centered_values(stack, gain):
    """_calc_centered_vals (_ROFL_with_jacobi.py:47-57)."""
    _, n_rho = stack.shape
    trans = np.tile(np.mean(stack, 1), (n_rho, 1)).T
    centered = stack / trans - np.ones_like(stack)
    sigma = np.sqrt(gain * (stack / trans**2 + stack**2 / (trans**3 * n_rho)))
    return centered, 1 / sigma


def symmetrize(direction, inclination):
    """_symmetrize_angles (_ROFL_with_jacobi.py:60-69)."""
    inclination = (((inclination + np.pi / 2.) % np.pi) - np.pi / 2.) * np.sign(
        -(np.floor(direction / np.pi) % 2) + 0.5)
    return direction % np.pi, inclination


def rotation(psi, tau):
    """_create_rotation_matrix (_ROFL_with_jacobi.py:72-100)."""
    c, s, ct, st = np.cos(psi), np.sin(psi), np.cos(tau), np.sin(tau)
    return np.array([[ct * c**2 + s**2, (ct - 1) * s * c, c * st],
                     [(ct - 1) * s * c, ct * s**2 + c**2, s * st],
                     [-c * st, -s * st, ct]])


def tilt_matrices(n_tilts, tau):
    """_list_all_tilt_matrices (_ROFL_with_jacobi.py:124-142)."""
    psi = np.linspace(0, 2 * np.pi, n_tilts + 1)[:-1]
    return np.array([rotation(p, tau) for p in psi])


def rotated_params(phi, alpha, t_rel, mats, tau):
    """_calculate_rotated_fiber_params (_ROFL_with_jacobi.py:164-193)."""
    v = np.array([np.cos(alpha) * np.cos(phi), np.cos(alpha) * np.sin(phi), np.sin(alpha)])
    n = mats.shape[0]
    out = np.zeros((n + 1, 2))
    for k in range(n):
        w = np.dot(mats[k], v)
        out[k + 1] = [np.arctan(w[1] / float(w[0])), np.arcsin(w[2])]
    out[0] = [phi, alpha]
    t = np.empty(n + 1)
    t[1:] = t_rel / np.cos(tau)
    t[0] = t_rel
    return out[:, 0], out[:, 1], t


def intensities(phi, alpha, t_rel, n_rho, dir_offset):
    """_calc_Int_single_fiber_fitted (_ROFL_with_jacobi.py:196-223), flattened."""
    phi, alpha = symmetrize(phi, alpha)
    t_rel = np.abs(t_rel)
    rho = np.linspace(0, np.pi, n_rho + 1)[:-1] + dir_offset
    out = np.empty((len(phi), n_rho))
    for j in range(len(phi)):
        out[j] = np.sin(np.pi / 2 * t_rel[j] * np.cos(alpha[j])**2) * np.sin(2 * (rho - phi[j]))
    return out.ravel()


def jacobi(phi, alpha, t_rel, n_rho, dir_offset):
    """_calc_Jacobi (_ROFL_with_jacobi.py:226-270): (3, n_tilts * n_rho)."""
    phi, alpha = symmetrize(phi, alpha)
    t_rel = np.abs(t_rel)
    rho = np.linspace(0, np.pi, n_rho + 1)[:-1] + dir_offset
    n = len(phi)
    J = np.empty((3, n, n_rho))
    for j in range(n):
        arg = np.pi / 2 * t_rel[j] * np.cos(alpha[j])**2
        s2 = np.sin(2 * (rho - phi[j]))
        J[0, j] = -2 * np.cos(2 * (rho - phi[j])) * np.sin(arg)
        J[1, j] = -np.pi * t_rel[j] * np.cos(alpha[j]) * np.sin(alpha[j]) * np.cos(arg) * s2
        J[2, j] = np.pi / 2 * np.cos(alpha[j])**2 * s2 * np.cos(arg)
    return J.reshape(3, -1)


def invert3(m):
    """_invertmatrix (_ROFL_with_jacobi.py:14-40) as the reference computes it.

    Its determinant's third term is m[2,0] * m[1,2] (the factor m[0,1] of the
    cofactor expansion is absent); the deviations it yields are reproduced.
    """
    det = (m[0, 0] * m[1, 1] * m[2, 2] + m[1, 0] * m[2, 1] * m[0, 2] + m[2, 0] * m[1, 2] -
           m[0, 0] * m[2, 1] * m[1, 2] - m[2, 0] * m[1, 1] * m[0, 2] - m[1, 0] * m[0, 1] * m[2, 2])
    inv = np.empty((3, 3))
    inv[0, 0] = m[1, 1] * m[2, 2] - m[1, 2] * m[2, 1]
    inv[0, 1] = m[0, 2] * m[2, 1] - m[0, 1] * m[2, 2]
    inv[0, 2] = m[0, 1] * m[1, 2] - m[0, 2] * m[1, 1]
    inv[1, 0] = m[1, 2] * m[2, 0] - m[1, 0] * m[2, 2]
    inv[1, 1] = m[0, 0] * m[2, 2] - m[0, 2] * m[2, 0]
    inv[1, 2] = m[0, 2] * m[1, 0] - m[0, 0] * m[1, 2]
    inv[2, 0] = m[1, 0] * m[2, 1] - m[1, 1] * m[2, 0]
    inv[2, 1] = m[0, 1] * m[2, 0] - m[0, 0] * m[2, 1]
    inv[2, 2] = m[0, 0] * m[1, 1] - m[0, 1] * m[1, 0]
    return inv / det


def _model(p, mats, tau, n_rho, dir_offset):
    d, a, t = rotated_params(p[0], p[1], p[2], mats, tau)
    return intensities(d, a, t, n_rho, dir_offset)


def _residuum(p, mats, tau, n_rho, dir_offset, measures, weights):
    return weights * (_model(p, mats, tau, n_rho, dir_offset) - measures)


def brute_force(phi, n_incl, n_trel, n_rho, mats, dir_offset, tau, measures, weights):
    """_brute_force_grid (_ROFL_with_jacobi.py:313-370): first grid minimum."""
    incl = np.linspace(-np.pi / 2, np.pi / 2, n_incl + 2)[1:-1]
    trel = np.linspace(0, 0.9, n_trel + 2)[1:-1]
    grid = list(itertools.product(incl, trel))
    cost = [np.sum(np.square(weights * (_model((phi, a, t), mats, tau, n_rho, dir_offset) -
                                         measures))) for a, t in grid]
    k = int(np.argmin(cost))
    return phi, grid[k][0], grid[k][1]


def _weighted_jacobi(p, mats, tau, n_rho, dir_offset, _, weights):
    """_weighted_jacobi (_ROFL_with_jacobi.py:383-393): the Dfun of grad mode."""
    d, a, t = rotated_params(p[0], p[1], p[2], mats, tau)
    return jacobi(d, a, t, n_rho, dir_offset) * weights


def fit(stack, tilt_angle=np.deg2rad(5.5), gain=3, dir_offset=0, grad=False):
    """rofl.rofl (rofl.py:17-60) on one (5, n_rho) stack, local fit.

    grad: leastsq with Dfun = the weighted analytic Jacobian (MINPACK lmder),
    as _ROFL_with_jacobi.py:467-475; else finite differences (lmdif, :476-480).

    Returns (direction, inclination, t_rel, dir_dev, incl_dev, trel_dev, chi^2,
    function evaluations).
    """
    stack = np.asarray(stack, dtype=np.float64)
    if stack.ndim != 2:
        raise TypeError('data: np.array([tilts,stack])')
    if stack.shape[0] != 5:
        raise ValueError('data need 1 + 4 measurements')
    if stack.shape[1] <= 3:
        raise ValueError('data needs at least 3 equidistant rotations')
    if gain <= 1:
        raise ValueError('rofl gain <= 1')
    phi0 = epa_direction(stack[0]) + dir_offset
    centered, weights = centered_values(stack, gain)
    n_tilt_steps, n_rho = centered.shape
    mats = tilt_matrices(n_tilt_steps - 1, tilt_angle)
    centered = centered.ravel()
    weights = weights.ravel()
    args = (mats, tilt_angle, n_rho, dir_offset, centered, weights)
    start = brute_force(phi0, 6, 6, n_rho, mats, dir_offset, tilt_angle, centered, weights)
    if grad:
        x, _, info, _, _ = optimize.leastsq(_residuum, start, args=args, full_output=1,
                                            maxfev=150, Dfun=_weighted_jacobi, col_deriv=1)
    else:
        x, _, info, _, _ = optimize.leastsq(_residuum, start, args=args, full_output=1,
                                            maxfev=150)
    fvalue = np.sum(np.square(info['fvec']))
    d, a = symmetrize(x[0], x[1])
    t = np.abs(x[2])
    sigma_sq = fvalue / float(len(centered) - 3)
    dd, aa, tt = rotated_params(d, a, t, mats, tilt_angle)
    wj = np.tile(weights, (3, 1)) * jacobi(dd, aa, tt, n_rho, dir_offset)
    cov = invert3(np.dot(wj, wj.T)) * sigma_sq
    dev = np.diag(np.abs(cov)**0.5)
    return d, a, t, dev[0], dev[1], dev[2], fvalue, info['nfev']


def rofl_map(data, tilt_angle=np.deg2rad(5.5), gain=3, dir_offset=0, mask=None, grad=False):
    """Simpli.apply_rofl (__simpli.py:623-706) without the pool: data (5, nx, ny, n_rho).

    Returns (direction, inclination, t_rel, (dir_dev, incl_dev, trel_dev, chi^2, nfev)).
    """
    data = np.asarray(data, dtype=np.float64)
    nx, ny = data.shape[1:3]
    if mask is None:
        mask = np.ones((nx, ny), bool)
    out = np.full((8, nx, ny), np.nan)
    for i in range(nx):
        for j in range(ny):
            if mask[i, j]:
                out[:, i, j] = fit(data[:, i, j, :], tilt_angle, gain, dir_offset, grad)
    return out[0], out[1], out[2], tuple(out[3:])
