"""Pins of the FP64 CPU oracle (CPU only).

Golden vectors: the reference's tests/simulation/simpli_rep.h5 (labels,
optical axes and the images of 5 tilts with Slerp).  Closed forms: the voxel
tests of tests/simulation/simpli_test.py:65-218, uniform columns, the
rho-rotation identity and a single tilted voxel.  Brute force: every voxel
against every segment on tiny random scenes.
"""

import numpy as np
import pytest

import oracle
from fastpli_amd import objects
from tests import bruteforce
from tests.conftest import GOLDEN_ROT, GOLDEN_TILTS, rel_err


# ----------------------------------------------------------------- golden
def test_oracle_generation_matches_golden(golden, cube_bundles):
    label, vector, props, warn = oracle.generate(cube_bundles, (24, 24, 24),
                                                 np.array([-60.0] * 3), 5.0)
    assert np.array_equal(label, golden['tissue'])
    assert np.array_equal(vector, golden['optical_axis'])
    assert np.array_equal(props, golden['tissue_properties'])
    assert warn == 0


@pytest.mark.parametrize('t', range(5))
def test_oracle_simulation_matches_golden(golden, t):
    theta, phi = GOLDEN_TILTS[t]
    img, warn = oracle.simulate(golden['tissue'], golden['optical_axis'],
                                golden['tissue_properties'], theta, phi, voxel_size=5.0,
                                filter_rotations=GOLDEN_ROT, interpolate='Slerp', n_threads=4)
    # the reference's own bar: h5diff --relative=0.00001
    assert rel_err(img, golden['data'][t]).max() < 1e-5
    assert warn == 0


def test_golden_background_pixel_closed_form(golden):
    # pixel (1,1) of the untilted golden image crosses two 'b' voxels (mu = 5/mm,
    # t = 5 um) and background only: I = I0/4 * exp(-2 * 5e3 * 5e-6)
    assert np.allclose(golden['data'][0][1, 1, :], 26000 / 4 * np.exp(-0.05), rtol=1e-12)
    col = golden['tissue'][1, 1, :]
    assert np.count_nonzero(col == 2) == 2 and np.count_nonzero(col) == 2


# ------------------------------------------------------------ closed forms
def _single_fiber(layers):
    fbs = objects.FiberBundles([[[[0, 0.25, -500, 1], [0, 0.25, 500, 1]]]], [layers])
    return fbs


def test_oracle_tissue_radial():
    fbs = _single_fiber([(1.0, 0.004, 1, 'r')])
    label, vector, _, _ = oracle.generate(fbs, (15, 15, 15), np.zeros(3), 0.1)
    x, y, r = 0 / 0.1, 0.25 / 0.1, 1 / 0.1
    for i in range(15):
        for j in range(15):
            t = 0 if (x - (i + 0.5))**2 + (y - (j + 0.5))**2 > r**2 else 1
            assert np.all(label[i, j, :] == t)
            if t:
                phi = np.arctan2((j + 0.5) - y, (i + 0.5) - x)
                np.testing.assert_array_almost_equal(vector[i, j, 0, :],
                                                     [np.cos(phi), np.sin(phi), 0], decimal=6)


def test_oracle_tissue_layered():
    fbs = _single_fiber([(0.3, -0.004, 1, 'p'), (0.6, 0, 1, 'b'), (1.0, 0.004, 1, 'r')])
    label, vector, props, _ = oracle.generate(fbs, (15, 15, 15), np.zeros(3), 0.1)
    x, y, r = 0.0, 2.5, 10.0
    for i in range(15):
        for j in range(15):
            rr = (x - (i + 0.5))**2 + (y - (j + 0.5))**2
            if rr > r**2:
                assert label[i, j, 0] == 0
            elif np.sqrt(rr) < 0.3 * r:
                assert label[i, j, 0] == 1
                np.testing.assert_array_almost_equal(vector[i, j, 0], [0, 0, 1], decimal=6)
            elif np.sqrt(rr) < 0.6 * r:
                assert label[i, j, 0] == 2
                assert np.array_equal(vector[i, j, 0], [0, 0, 0])
            else:
                assert label[i, j, 0] == 3
    assert np.array_equal(props, [[0, 0], [-0.004, 1], [0, 1], [0.004, 1]])


@pytest.mark.parametrize('seed', range(6))
def test_oracle_matches_bruteforce(seed):
    fbs = bruteforce.random_scene(seed)
    dim = (7, 6, 8)
    label, vector, _, _ = oracle.generate(fbs, dim, np.zeros(3), 1.0)
    bl, bv = bruteforce.voxelize(fbs, dim, 1.0)
    assert np.array_equal(label, bl)
    assert np.array_equal(vector, bv)


def _uniform_volume(dim, vec, dn, mu):
    label = np.ones(dim, dtype=np.int32)
    vector = np.empty(tuple(dim) + (3,), dtype=np.float32)
    vector[:] = vec
    props = np.array([[0.0, 0.0], [dn, mu]])
    return label, vector, props


@pytest.mark.parametrize('incl', [0.0, 0.3, 1.1])
def test_oracle_uniform_column_closed_form(incl):
    phi_v = 0.7
    v = np.array([np.cos(incl) * np.cos(phi_v), np.cos(incl) * np.sin(phi_v), np.sin(incl)])
    v = (v / np.linalg.norm(v.astype(np.float32))).astype(np.float32)
    dim = (3, 2, 40)
    dn, mu, vs, wl, I0 = -0.004, 3.0, 0.5, 525.0, 26000.0
    label, vector, props = _uniform_volume(dim, v, dn, mu)
    rho = np.linspace(0, np.pi, 9, endpoint=False)
    img, _ = oracle.simulate(label, vector, props, 0.0, 0.0, voxel_size=vs, wavelength=wl,
                             light_intensity=I0, filter_rotations=rho, interpolate='NN')
    t = vs * 1e-6
    vd = v.astype(np.float64)
    cos2a = 1.0 - (vd[2] / np.linalg.norm(vd))**2
    delta = np.pi / 2 * dn * 4 * t / (wl * 1e-9) * np.linalg.norm(vd) * cos2a
    phi = np.arctan2(vd[1], vd[0])
    expect = I0 / 4 * np.exp(-mu * 1e3 * t * dim[2]) * (
        1 - np.sin(dim[2] * delta) * np.sin(2 * (rho - phi)))
    assert rel_err(img, expect[None, None, :]).max() < 1e-10


def test_oracle_rho_rotation_identity(golden):
    # every pixel is T (1 + a cos 2rho + b sin 2rho): three free numbers
    rho = np.deg2rad(np.arange(0, 180, 7.5))
    img, _ = oracle.simulate(golden['tissue'], golden['optical_axis'],
                             golden['tissue_properties'], GOLDEN_TILTS[1][0],
                             GOLDEN_TILTS[1][1], voxel_size=5.0, filter_rotations=rho,
                             interpolate='Slerp', n_threads=4)
    A = np.stack([np.ones_like(rho), np.cos(2 * rho), np.sin(2 * rho)], axis=1)
    y = img.reshape(-1, rho.size).T
    coef, *_ = np.linalg.lstsq(A, y, rcond=None)
    assert np.abs(A @ coef - y).max() < 1e-9 * np.abs(y).max()


def test_oracle_single_tilted_voxel():
    # one voxel, tilt (theta, phi), step_size = 1/cos(theta): the ray crosses it
    # once; closed form with the tilt rotation built by scipy
    from scipy.spatial.transform import Rotation
    v = np.array([np.cos(np.pi / 4), 0, np.sin(np.pi / 4)], dtype=np.float32)
    label, vector, props = _uniform_volume((1, 1, 1), v, -0.001, 2.0)
    rho = np.linspace(0, np.pi, 9, endpoint=False)
    for theta, phi in [(0.0, 0.0), (np.deg2rad(5.5), 0.0), (np.deg2rad(5.5), np.pi / 2),
                       (np.deg2rad(5.5), np.pi), (np.deg2rad(5.5), 3 * np.pi / 2)]:
        step = 1.0 / np.cos(theta)
        img, _ = oracle.simulate(label, vector, props, theta, phi, step_size=step,
                                 voxel_size=60.0, filter_rotations=rho, interpolate='NN')
        R = Rotation.from_euler('ZYZ', [phi, theta, -phi]).as_matrix()
        w = R @ v.astype(np.float64)
        t = 60e-6 * step
        alpha = np.arcsin(w[2] / np.linalg.norm(w))
        delta = np.pi / 2 * (-0.001) * 4 * t / 525e-9 * np.linalg.norm(w) * np.cos(alpha)**2
        ph = np.arctan2(w[1], w[0])
        expect = 26000 / 4 * np.exp(-2e3 * t) * (1 - np.sin(delta) * np.sin(2 * (rho - ph)))
        assert rel_err(img[0, 0], expect).max() < 1e-9, (theta, phi)


def test_oracle_count_steps():
    # untilted: nz steps; 5.5 deg at nz=300: 301 (floor((nz - cos/2)/cos) + 1)
    assert oracle.count_steps((2, 3, 300), 1.0, 0.0, 0.0) == 2 * 3 * 300
    assert oracle.count_steps((2, 3, 300), 1.0, np.deg2rad(5.5), 0.0) == 2 * 3 * 301


def test_oracle_errors():
    label = np.zeros((2, 2, 2), np.int32)
    vec = np.zeros((2, 2, 2, 3), np.float32)
    props = np.array([[0.0, 0.0], [0.1, 1.0]])
    rho = [0.0]
    label[0, 0, 0] = -1
    with pytest.raises(ValueError):
        oracle.simulate(label, vec, props, 0, 0, filter_rotations=rho)
    label[0, 0, 0] = 2
    with pytest.raises(ValueError):
        oracle.simulate(label, vec, props, 0, 0, filter_rotations=rho)


# ------------------------------------------------- setup branches (simulator.cpp)
@pytest.mark.parametrize('interp', ['NN', 'Lerp', 'Slerp'])
def test_oracle_flip_z_is_an_index_mirror(golden, interp):
    # GetLabel / GetVec map z -> nz - 1 - z before every read (simulator.cpp:
    # 427-428, 559-560) and nothing else changes: flip_z on a volume equals the
    # z-mirrored volume without it, bit for bit
    T, V, P = golden['tissue'], golden['optical_axis'], golden['tissue_properties']
    Tm, Vm = np.ascontiguousarray(T[:, :, ::-1]), np.ascontiguousarray(V[:, :, ::-1])
    for theta, phi in GOLDEN_TILTS[:3]:
        a, _ = oracle.simulate(T, V, P, theta, phi, voxel_size=5.0, interpolate=interp,
                               flip_z=True, filter_rotations=GOLDEN_ROT, n_threads=4)
        b, _ = oracle.simulate(Tm, Vm, P, theta, phi, voxel_size=5.0, interpolate=interp,
                               flip_z=False, filter_rotations=GOLDEN_ROT, n_threads=4)
        assert np.array_equal(a, b)


@pytest.mark.parametrize('step', [0.5, 1.7, 2.0])
def test_oracle_step_size_uniform_column(step):
    # uniform column, untilted: n steps of thickness step * voxel_size each
    v = np.array([0.48, 0.6, 0.64], dtype=np.float32)
    v = (v / np.linalg.norm(v)).astype(np.float32)
    dim = (2, 3, 37)
    dn, mu, vs, wl, I0 = -0.004, 2.0, 0.5, 525.0, 26000.0
    label, vector, props = _uniform_volume(dim, v, dn, mu)
    rho = np.linspace(0, np.pi, 9, endpoint=False)
    img, _ = oracle.simulate(label, vector, props, 0.0, 0.0, step_size=step, voxel_size=vs,
                             wavelength=wl, light_intensity=I0, filter_rotations=rho,
                             interpolate='NN')
    n = oracle.count_steps(dim, step, 0.0, 0.0) // (dim[0] * dim[1])
    # positions step/2 + k step with floor inside [0, nz)
    assert n == int(np.ceil((dim[2] - step / 2) / step))
    t = vs * step * 1e-6
    vd = v.astype(np.float64)
    cos2a = 1.0 - (vd[2] / np.linalg.norm(vd))**2
    delta = np.pi / 2 * dn * 4 * t / (wl * 1e-9) * np.linalg.norm(vd) * cos2a
    phi = np.arctan2(vd[1], vd[0])
    expect = I0 / 4 * np.exp(-mu * 1e3 * t * n) * (1 - np.sin(n * delta) * np.sin(2 * (rho - phi)))
    assert rel_err(img, expect[None, None, :]).max() < 1e-10


@pytest.mark.parametrize('opts', [dict(), dict(untilt_sensor_view=False, tissue_refrection=1.3),
                                  dict(flip_z=True), dict(step_size=0.7)])
def test_oracle_background_absorption_closed_form(opts):
    # background only (label 0 inside and outside the volume): every step of
    # every ray attenuates by exp(-mu_bg 1e3 t); the state stays (1, 0, 0, -1)
    dim = (6, 5, 23)
    label = np.zeros(dim, np.int32)
    vector = np.zeros(dim + (3,), np.float32)
    props = np.array([[0.0, 3.5], [-0.004, 1.0]])
    rho = np.linspace(0, np.pi, 9, endpoint=False)
    step = opts.get('step_size', 1.0)
    for theta, phi in [(0.0, 0.0), (np.deg2rad(5.5), 0.3), (np.deg2rad(20.0), 4.0)]:
        img, _ = oracle.simulate(label, vector, props, theta, phi, voxel_size=0.5,
                                 filter_rotations=rho, interpolate='Slerp', **opts)
        n = oracle.count_steps(dim, step, theta, phi) // (dim[0] * dim[1])
        expect = 26000 / 4 * np.exp(-3.5e3 * 0.5e-6 * step * n)
        assert rel_err(img, np.full_like(img, expect)).max() < 1e-12, (theta, phi, opts)


@pytest.mark.parametrize('interp', ['Lerp', 'Slerp'])
def test_oracle_interpolation_of_a_uniform_field(interp):
    # all corners of every GetVec carry the same label and axis: Lerp and
    # Slerp return that axis, so every tilted view equals its NN view
    v = np.array([0.36, -0.48, 0.8], dtype=np.float32)
    label, vector, props = _uniform_volume((9, 8, 30), v, -0.004, 1.5)
    rho = np.linspace(0, np.pi, 9, endpoint=False)
    for theta, phi in [(np.deg2rad(5.5), 0.0), (np.deg2rad(12.0), 2.2)]:
        a, _ = oracle.simulate(label, vector, props, theta, phi, voxel_size=0.4,
                               filter_rotations=rho, interpolate=interp)
        b, _ = oracle.simulate(label, vector, props, theta, phi, voxel_size=0.4,
                               filter_rotations=rho, interpolate='NN')
        assert rel_err(a, b).max() < 1e-12


def test_oracle_tilted_sensor_at_normal_incidence():
    # theta = 0: the refracted angle is 0 and the tilted-sensor start positions
    # (simulator.cpp:614-719) are the pixel centres of the untilted view
    rng = np.random.default_rng(2)
    dim = (7, 6, 9)
    T = rng.integers(0, 2, dim).astype(np.int32)
    V = rng.standard_normal(dim + (3,)).astype(np.float32)
    V /= np.linalg.norm(V, axis=-1, keepdims=True)
    P = np.array([[0.0, 0.5], [-0.004, 1.0]])
    rho = np.linspace(0, np.pi, 9, endpoint=False)
    for interp in ('NN', 'Slerp'):
        a, _ = oracle.simulate(T, V, P, 0.0, 0.0, interpolate=interp, filter_rotations=rho,
                               untilt_sensor_view=False, tissue_refrection=1.4)
        b, _ = oracle.simulate(T, V, P, 0.0, 0.0, interpolate=interp, filter_rotations=rho)
        assert rel_err(a, b).max() < 1e-12


def test_oracle_tilted_sensor_start_geometry():
    """The tilted sensor's ray entry points (CalcStartingLightPositionsTilted,
    simulator.cpp:618-703) by their geometric definition: a pixel's vertical
    line in the laboratory meets the tissue's top plane turned by
    R = Rz(phi) Ry(a) Rz(-phi), a = asin(sin(-theta) n), about the volume
    centre; R^-1 of that point, moved down along the light direction to
    z = 0, is the entry point.  Checked backwards with scipy's rotation: from
    each entry point up the light direction to z = nz, then R, the point lies
    on its pixel's vertical line (this fails for R and R^-1 swapped, a missing
    refraction, a sign of theta or phi, or a transposed pixel grid)."""
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(3)
    cases = [(np.deg2rad(5.5), p, 1.0) for p in (0.0, np.pi / 2, np.pi, -np.pi / 2)]
    cases += [(rng.uniform(0.01, 0.5), rng.uniform(-np.pi, np.pi), n) for n in (1.0, 1.3, 1.0)]
    for dim in [(10, 8, 6), (33, 17, 40)]:
        nx, ny, nz = dim
        pc = np.array(dim, dtype=float) / 2
        for theta, phi, n in cases:
            st = oracle.light_starts(dim, theta, phi, n)
            d = np.array([np.cos(phi) * np.sin(theta), np.sin(phi) * np.sin(theta),
                          np.cos(theta)])
            bottom = np.concatenate([st, np.zeros(st.shape[:2] + (1,))], axis=-1)
            top = bottom + d / np.cos(theta) * nz
            R = Rotation.from_euler('ZYZ', [phi, np.arcsin(np.sin(-theta) * n), -phi])
            lab = (top - pc) @ R.as_matrix().T + pc
            ii, jj = np.meshgrid(np.arange(nx) + 0.5, np.arange(ny) + 0.5, indexing='ij')
            assert np.allclose(lab[..., 0], ii, rtol=0, atol=1e-9), (dim, theta, phi, n)
            assert np.allclose(lab[..., 1], jj, rtol=0, atol=1e-9), (dim, theta, phi, n)


def _cone_field(rng, shape, half_angle_deg=40.0):
    """Unit axes within a cone about a random axis (every pair, and every
    positive combination of them, has a positive dot product)."""
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    v = rng.normal(size=shape + (3,))
    v -= (v @ axis)[..., None] * axis
    v /= np.linalg.norm(v, axis=-1, keepdims=True)
    ang = np.deg2rad(rng.uniform(0, half_angle_deg, size=shape))[..., None]
    return (np.cos(ang) * axis + np.sin(ang) * v).astype(np.float32)


def test_oracle_lerp_blend_is_trilinear_interpolation():
    """GetVec with Lerp (LinearOrientationInterpolation, simulator.cpp:457-471)
    on one label, where every blend has a non-negative dot product: the x, y, z
    cascade of (1 - t) a + t b is trilinear interpolation of the components,
    checked against scipy's RegularGridInterpolator (voxel centres at i + 0.5;
    beyond the outer centres the clamped corner indices equal interpolation at
    the clipped position).  Wrong weights, a transposed corner or a missing
    clamp fail it."""
    from scipy.interpolate import RegularGridInterpolator
    rng = np.random.default_rng(11)
    shape = (5, 4, 6)
    vector = _cone_field(rng, shape)
    label = np.ones(shape, dtype=np.int32)
    grid = tuple(np.arange(n) + 0.5 for n in shape)
    ref = RegularGridInterpolator(grid, vector.astype(np.float64), method='linear')
    hi = np.array(shape, dtype=float)
    for _ in range(400):
        pos = rng.uniform(0.0, hi)
        got = oracle.get_vec(label, vector, pos, 'Lerp')
        want = ref(np.clip(pos, 0.5, hi - 0.5)[None])[0]
        assert np.allclose(got, want, rtol=1e-12, atol=1e-12), (pos, got, want)


@pytest.mark.parametrize('interp', ['Lerp', 'Slerp'])
def test_oracle_interpolation_ignores_axis_sign(interp):
    """An axis and its negative are one fiber orientation: negating any corners
    leaves the interpolated axis unchanged up to its sign (the Lerp and Slerp
    hemisphere rules, simulator.cpp:449-494); without them the blend of a
    flipped corner would shrink or turn the result."""
    rng = np.random.default_rng(12)
    shape = (4, 5, 3)
    vector = _cone_field(rng, shape, 60.0)
    label = np.ones(shape, dtype=np.int32)
    flipped = vector * rng.choice([-1.0, 1.0], size=shape + (1,)).astype(np.float32)
    for _ in range(300):
        pos = rng.uniform(0.0, np.array(shape, dtype=float))
        a = oracle.get_vec(label, vector, pos, interp)
        b = oracle.get_vec(label, flipped, pos, interp)
        s = 1.0 if a @ b >= 0 else -1.0
        assert np.allclose(b, s * a, rtol=1e-9, atol=1e-12), (pos, a, b)


def test_oracle_lerp_corners_of_other_labels():
    """Corners whose label differs from the sample's voxel count as zero, and a
    zero corner leaves the blend to the other one (v == 0 -> u,
    simulator.cpp:461-464, 533-540): between a label-1 voxel with axis a and a
    label-2 voxel with axis b, every sample in the first voxel returns a
    exactly and every sample in the second b; a label-0 voxel returns zero.
    Known answer of the subtraction branch: a = (1, 0, 0), b = (-0.6, 0.8, 0)
    at weight 1/4 give 3/4 a - 1/4 b = (0.9, -0.2, 0)."""
    a = np.array([0.6, 0.0, 0.8], dtype=np.float32)
    b = np.array([0.0, 1.0, 0.0], dtype=np.float32)
    vector = np.stack([a, b, a]).reshape(3, 1, 1, 3)
    label = np.array([1, 2, 0], dtype=np.int32).reshape(3, 1, 1)
    for x in np.linspace(0.05, 0.95, 7):
        assert np.array_equal(oracle.get_vec(label, vector, (x, 0.5, 0.5)), a.astype(float))
        assert np.array_equal(oracle.get_vec(label, vector, (1 + x, 0.5, 0.5)),
                              b.astype(float))
        assert np.array_equal(oracle.get_vec(label, vector, (2 + x, 0.5, 0.5)), np.zeros(3))
    vector = np.array([[1.0, 0.0, 0.0], [-0.6, 0.8, 0.0]], dtype=np.float32).reshape(2, 1, 1, 3)
    label = np.ones((2, 1, 1), dtype=np.int32)
    got = oracle.get_vec(label, vector, (0.75, 0.5, 0.5), 'Lerp')
    assert np.allclose(got, [0.9, -0.2, 0.0], rtol=0, atol=1e-7), got


def test_oracle_past_the_end_label_overlap_known_answer():
    """Hand-derived case: with r = 1 and outer scale 0.7, a voxel centre at
    distance 0.7 has float(d^2) / float(r^2) = float(0.49) > 0.49 (double), so
    lower_bound passes the last layer: label 1 + 0 * 2 + 2 = 3 (generator.cpp:
    345-351).  A second fiber of the bundle at distance 0.1 gives label 1 and
    replaces it (0.01 <= 0.49), |3 - 1| = 2 >= max_layer: "objects overlap"
    (generator.cpp:353-359).  Listed the other way round, nothing is replaced."""
    a = np.array([[-1.0, 1.2, 0.5, 1.0], [9.0, 1.2, 0.5, 1.0]])
    b = np.array([[-1.0, 0.5, 0.6, 1.0], [9.0, 0.5, 0.6, 1.0]])
    layers = [[(0.45, -0.004, 1, 'p'), (0.7, 0.004, 2, 'r')]]
    assert float(np.float32(0.7 * 0.7)) > 0.49
    only_a = objects.FiberBundles([[a]], layers)
    la, _, pa, _ = oracle.generate(only_a, (8, 4, 4), np.zeros(3), 1.0, n_threads=1)
    assert np.all(la[:, 0, 0] == 3) and pa.shape[0] == 3
    for order, warns in (([a, b], True), ([b, a], False)):
        fbs = objects.FiberBundles([order], layers)
        lab, _, _, w = oracle.generate(fbs, (8, 4, 4), np.zeros(3), 1.0, n_threads=1)
        assert np.all(lab[:, 0, 0] == 1)
        assert bool(w & oracle.WARN_OVERLAP) == warns
