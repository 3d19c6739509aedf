"""ctypes binding of libfastpli_amd.so (the C ABI declared in include/fastpli_amd.h).

The library is required: there is no CPU fallback.  Importing this module
without the built library, or calling into it without a HIP device, raises.
"""

import ctypes
import os
import warnings

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'lib', 'libfastpli_amd.so')

FPL_OK = 0
FPL_E_INVALID = -1
FPL_E_LABEL_NEG = -2
FPL_E_LABEL_MAX = -3
FPL_E_BG_DN = -4
FPL_E_DEVICE = -5
FPL_E_NOMEM = -6
FPL_E_PENDING = -7

WARN_OVERLAP = 1 << 0
WARN_SINGLE_POINT = 1 << 1
WARN_ZERO_VEC = 1 << 2
WARN_NOT_NORM = 1 << 3
WARN_NO_LIGHT = 1 << 4
# not a warning: the march wrote an image value < 0 (raised as ValueError)
FLAG_NEGATIVE = 1 << 8

# messages of the reference (generator.cpp:215-216,357; simulator.cpp:294-301,715)
WARN_MESSAGES = {
    WARN_OVERLAP: 'objects overlap',
    WARN_SINGLE_POINT: 'fiber object with only single point',
    WARN_ZERO_VEC: 'optical_axis is 0 for tissue > 0',
    WARN_NOT_NORM: 'optical_axis not normalized',
    WARN_NO_LIGHT: 'light_positions is empty',
}

INTERP = {'NN': 0, 'Lerp': 1, 'Slerp': 2}
STORAGE = {'f32': 0, 'f16': 1, 'angle16': 2}

_c_i64 = ctypes.c_int64
_c_p = ctypes.c_void_p


class FplVolume(ctypes.Structure):
    _fields_ = [('dim', _c_i64 * 3), ('origin', ctypes.c_double * 3),
                ('voxel_size', ctypes.c_double)]


class FplFiberBundles(ctypes.Structure):
    _fields_ = [('n_bundles', _c_i64), ('n_fibers', _c_i64), ('n_points', _c_i64),
                ('points', _c_p), ('fiber_ptr', _c_p), ('bundle_ptr', _c_p),
                ('layer_ptr', _c_p), ('layer_scale', _c_p), ('layer_dn', _c_p),
                ('layer_mu', _c_p), ('layer_model', _c_p)]


class FplSetup(ctypes.Structure):
    _fields_ = [('step_size', ctypes.c_double), ('light_intensity', ctypes.c_double),
                ('voxel_size', ctypes.c_double), ('wavelength', ctypes.c_double),
                ('tissue_refrection', ctypes.c_double), ('interpolate', ctypes.c_int32),
                ('untilt_sensor_view', ctypes.c_int32), ('flip_z', ctypes.c_int32),
                ('n_rho', ctypes.c_int32), ('filter_rotations', _c_p),
                ('field_storage', ctypes.c_int32)]


class FplFields(ctypes.Structure):
    _fields_ = [('dim', _c_i64 * 3), ('storage', ctypes.c_int32), ('label_min', ctypes.c_int32),
                ('label_max', ctypes.c_int32), ('pad_', ctypes.c_int32), ('records', _c_p),
                ('codes', _c_p)]


class FplRequest(ctypes.Structure):
    _fields_ = [('tissue', _c_i64), ('setup', ctypes.POINTER(FplSetup)), ('props', _c_p),
                ('n_props', _c_i64), ('tilts', _c_p), ('n_tilts', _c_i64), ('images', _c_p),
                ('stream', _c_p)]


class FplModelRequest(ctypes.Structure):
    _fields_ = [('volume', FplVolume), ('bundles', FplFiberBundles),
                ('background_absorption', ctypes.c_double), ('setup', ctypes.POINTER(FplSetup)),
                ('tilts', _c_p), ('n_tilts', _c_i64), ('images', _c_p), ('stream', _c_p)]


_LIB = None

# every symbol include/fastpli_amd.h declares
EXPORTS = ('fpl_num_properties', 'fpl_generate', 'fpl_simulate', 'fpl_optic_resample', 'fpl_epa',
           'fpl_count_voxel_steps', 'fpl_profile_enable', 'fpl_profile_read',
           'fpl_release_workspace', 'fpl_version', 'fpl_server_create', 'fpl_server_add_tissue',
           'fpl_server_remove_tissue', 'fpl_server_submit', 'fpl_server_wait',
           'fpl_server_stats', 'fpl_server_destroy', 'fpl_rofl', 'fpl_simulate_adjoint',
           'fpl_record_bytes', 'fpl_fields_bytes', 'fpl_fields_codes_offset',
           'fpl_generate_fields', 'fpl_pack_fields', 'fpl_simulate_fields',
           'fpl_server_submit_model', 'fpl_simulate_noalloc')


def lib():
    """Load the shared library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f'{LIB_PATH} is missing: build it with '
                              '`python -m fastpli_amd.csrc.build`')
        L = ctypes.CDLL(LIB_PATH)
        L.fpl_num_properties.restype = _c_i64
        L.fpl_num_properties.argtypes = [ctypes.POINTER(FplFiberBundles)]
        L.fpl_generate.restype = ctypes.c_int
        L.fpl_generate.argtypes = [
            ctypes.POINTER(FplVolume), ctypes.POINTER(FplFiberBundles), ctypes.c_int, _c_p,
            _c_p, _c_p, ctypes.POINTER(ctypes.c_uint32), _c_p, ctypes.c_char_p, ctypes.c_size_t
        ]
        L.fpl_simulate.restype = ctypes.c_int
        L.fpl_simulate.argtypes = [
            ctypes.POINTER(FplSetup), _c_p, _c_p, _c_p, _c_p, _c_i64, _c_p, _c_i64, _c_p,
            ctypes.POINTER(ctypes.c_uint32), _c_p, ctypes.c_char_p, ctypes.c_size_t
        ]
        L.fpl_record_bytes.restype = ctypes.c_size_t
        L.fpl_record_bytes.argtypes = [ctypes.c_int32]
        L.fpl_fields_bytes.restype = ctypes.c_size_t
        L.fpl_fields_bytes.argtypes = [_c_p, ctypes.c_int32]
        L.fpl_fields_codes_offset.restype = ctypes.c_size_t
        L.fpl_fields_codes_offset.argtypes = [_c_p, ctypes.c_int32]
        L.fpl_generate_fields.restype = ctypes.c_int
        L.fpl_generate_fields.argtypes = [
            ctypes.POINTER(FplVolume), ctypes.POINTER(FplFiberBundles), _c_p, _c_p, _c_p,
            ctypes.POINTER(FplFields), ctypes.POINTER(ctypes.c_uint32), _c_p, ctypes.c_char_p,
            ctypes.c_size_t
        ]
        L.fpl_pack_fields.restype = ctypes.c_int
        L.fpl_pack_fields.argtypes = [_c_p, _c_p, _c_p, ctypes.POINTER(FplFields), _c_p,
                                      ctypes.c_char_p, ctypes.c_size_t]
        L.fpl_simulate_fields.restype = ctypes.c_int
        L.fpl_simulate_fields.argtypes = [
            ctypes.POINTER(FplSetup), ctypes.POINTER(FplFields), _c_p, _c_i64, _c_p, _c_i64,
            _c_p, ctypes.POINTER(ctypes.c_uint32), _c_p, ctypes.c_char_p, ctypes.c_size_t
        ]
        L.fpl_simulate_noalloc.restype = ctypes.c_int
        L.fpl_simulate_noalloc.argtypes = L.fpl_simulate.argtypes
        L.fpl_simulate_adjoint.restype = ctypes.c_int
        L.fpl_simulate_adjoint.argtypes = [
            ctypes.POINTER(FplSetup), _c_p, _c_p, _c_p, _c_p, _c_i64, _c_p, _c_i64, _c_p, _c_p,
            _c_p, _c_p, ctypes.c_char_p, ctypes.c_size_t
        ]
        L.fpl_optic_resample.restype = ctypes.c_int
        L.fpl_optic_resample.argtypes = [
            _c_p, _c_i64, _c_i64, _c_i64, ctypes.c_double, ctypes.c_double, _c_p, _c_p,
            ctypes.c_char_p, ctypes.c_size_t
        ]
        L.fpl_epa.restype = ctypes.c_int
        L.fpl_epa.argtypes = [
            _c_p, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_p, ctypes.c_char_p, ctypes.c_size_t
        ]
        L.fpl_count_voxel_steps.restype = _c_i64
        L.fpl_count_voxel_steps.argtypes = [
            ctypes.POINTER(FplSetup), _c_p, ctypes.c_double, ctypes.c_double
        ]
        L.fpl_profile_enable.restype = None
        L.fpl_profile_enable.argtypes = [ctypes.c_int]
        L.fpl_profile_read.restype = ctypes.c_int
        L.fpl_profile_read.argtypes = [
            ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
            ctypes.POINTER(ctypes.c_int64)
        ]
        L.fpl_release_workspace.restype = None
        L.fpl_release_workspace.argtypes = []
        L.fpl_version.restype = ctypes.c_char_p
        L.fpl_version.argtypes = []
        msgargs = [ctypes.c_char_p, ctypes.c_size_t]
        L.fpl_rofl.restype = ctypes.c_int
        L.fpl_rofl.argtypes = [_c_p, _c_i64, _c_i64, ctypes.c_double, ctypes.c_double,
                               ctypes.c_double, ctypes.c_int, _c_p, _c_p, _c_p] + msgargs
        L.fpl_server_create.restype = ctypes.c_int
        L.fpl_server_create.argtypes = [ctypes.c_int, _c_i64, ctypes.c_int32,
                                        ctypes.POINTER(_c_p)] + msgargs
        L.fpl_server_add_tissue.restype = ctypes.c_int
        L.fpl_server_add_tissue.argtypes = [_c_p, _c_p, _c_p, _c_p, ctypes.c_int32,
                                            ctypes.POINTER(_c_i64)] + msgargs
        L.fpl_server_remove_tissue.restype = ctypes.c_int
        L.fpl_server_remove_tissue.argtypes = [_c_p, _c_i64] + msgargs
        L.fpl_server_submit.restype = ctypes.c_int
        L.fpl_server_submit.argtypes = [_c_p, ctypes.POINTER(FplRequest),
                                        ctypes.POINTER(_c_i64)] + msgargs
        L.fpl_server_submit_model.restype = ctypes.c_int
        L.fpl_server_submit_model.argtypes = [_c_p, ctypes.POINTER(FplModelRequest),
                                              ctypes.POINTER(_c_i64)] + msgargs
        L.fpl_server_wait.restype = ctypes.c_int
        L.fpl_server_wait.argtypes = [_c_p, _c_i64, ctypes.c_double,
                                      ctypes.POINTER(ctypes.c_uint32),
                                      ctypes.POINTER(ctypes.c_double)] + msgargs
        L.fpl_server_stats.restype = ctypes.c_int
        L.fpl_server_stats.argtypes = [_c_p, ctypes.POINTER(_c_i64)]
        L.fpl_server_destroy.restype = ctypes.c_int
        L.fpl_server_destroy.argtypes = [_c_p]
        _LIB = L
    return _LIB


def raise_warnings(mask, stacklevel=3):
    for bit, msg in WARN_MESSAGES.items():
        if mask & bit:
            warnings.warn(msg, UserWarning, stacklevel=stacklevel)


def check_images(mask):
    """The reference's check of the images after the native call
    (simpli.py:184-185), from the march's flag instead of a pass over them."""
    if mask & FLAG_NEGATIVE:
        raise ValueError('intensity < 0 detected')


def check(rc, msg):
    if rc == FPL_OK:
        return
    text = msg.value.decode(errors='replace') if msg.value else f'error {rc}'
    if rc in (FPL_E_DEVICE, FPL_E_NOMEM):
        raise RuntimeError(text)
    raise ValueError(text)


def _ptr(a):
    """Raw address of a numpy array or torch tensor."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


def _stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


def make_setup(step_size, light_intensity, voxel_size, wavelength, tissue_refrection,
               interpolate, untilt_sensor_view, flip_z, filter_rotations, field_storage='f32'):
    """Returns (FplSetup, keep-alive rotations array)."""
    rho = np.ascontiguousarray(filter_rotations, dtype=np.float64).ravel()
    s = FplSetup(float(step_size), float(light_intensity), float(voxel_size), float(wavelength),
                 float(tissue_refrection), INTERP[interpolate], int(bool(untilt_sensor_view)),
                 int(bool(flip_z)), int(rho.size), rho.ctypes.data if rho.size else None,
                 STORAGE[field_storage])
    return s, rho


def _bundles(points, fiber_ptr, bundle_ptr, layer_ptr, layer_table):
    """FplFiberBundles over device points / fiber offsets and host tables; returns
    (struct, keep-alive arrays)."""
    scale, dn, mu, model = (np.ascontiguousarray(a) for a in layer_table)
    bundle_ptr = np.ascontiguousarray(bundle_ptr, dtype=np.int64)
    layer_ptr = np.ascontiguousarray(layer_ptr, dtype=np.int64)
    fbs = FplFiberBundles(
        len(bundle_ptr) - 1,
        int(fiber_ptr.shape[0]) - 1, int(points.shape[0]), _ptr(points), _ptr(fiber_ptr),
        bundle_ptr.ctypes.data, layer_ptr.ctypes.data,
        scale.ctypes.data if scale.size else None, dn.ctypes.data if dn.size else None,
        mu.ctypes.data if mu.size else None, model.ctypes.data if model.size else None)
    return fbs, (scale, dn, mu, model, bundle_ptr, layer_ptr)


def _volume(dim, origin, voxel_size):
    return FplVolume((_c_i64 * 3)(*[int(d) for d in dim]),
                     (ctypes.c_double * 3)(*[float(o) for o in origin]), float(voxel_size))


def record_bytes(storage):
    """Bytes per voxel of the record array of a field storage ('f32', 'f16', 'angle16')."""
    return int(lib().fpl_record_bytes(STORAGE[storage]))


def fields_layout(dim, storage):
    """(total bytes, byte offset of the angle codes) of one allocation of march fields."""
    d = np.ascontiguousarray(dim, dtype=np.int64)
    L = lib()
    return (int(L.fpl_fields_bytes(d.ctypes.data, STORAGE[storage])),
            int(L.fpl_fields_codes_offset(d.ctypes.data, STORAGE[storage])))


def fields_struct(buffer, dim, storage, label_min=0, label_max=0):
    """FplFields over a device byte buffer laid out as fields_layout says."""
    _, off = fields_layout(dim, storage)
    base = _ptr(buffer)
    codes = base + off if storage == 'angle16' else None
    return FplFields((_c_i64 * 3)(*[int(d) for d in dim]), STORAGE[storage], int(label_min),
                     int(label_max), 0, base, codes)


def generate_fields(dim, origin, voxel_size, points, fiber_ptr, bundle_ptr, layer_ptr,
                    layer_table, label, vector, fields):
    """fpl_generate_fields: label / vector (device tensors or None) and the march
    fields (FplFields, filled in place); returns (props ndarray, warn mask)."""
    fbs, _keep = _bundles(points, fiber_ptr, bundle_ptr, layer_ptr, layer_table)
    vol = _volume(dim, origin, voxel_size)
    L = lib()
    n_props = L.fpl_num_properties(ctypes.byref(fbs))
    props = np.zeros((n_props, 2), dtype=np.float64)
    warn = ctypes.c_uint32(0)
    msg = ctypes.create_string_buffer(512)
    rc = L.fpl_generate_fields(ctypes.byref(vol), ctypes.byref(fbs), _ptr(label), _ptr(vector),
                               props.ctypes.data, ctypes.byref(fields), ctypes.byref(warn),
                               _stream(), msg, 512)
    check(rc, msg)
    return props, warn.value


def pack_fields(dim, label, vector, fields):
    """fpl_pack_fields: C-order device fields -> march fields (filled in place)."""
    d = np.ascontiguousarray(dim, dtype=np.int64)
    msg = ctypes.create_string_buffer(512)
    rc = lib().fpl_pack_fields(d.ctypes.data, _ptr(label), _ptr(vector), ctypes.byref(fields),
                               _stream(), msg, 512)
    check(rc, msg)


def simulate_fields(setup, fields, props, tilts, images):
    """fpl_simulate_fields on march fields; returns the warn mask."""
    props = np.ascontiguousarray(props, dtype=np.float64)
    tilts = np.ascontiguousarray(tilts, dtype=np.float64).reshape(-1, 2)
    warn = ctypes.c_uint32(0)
    msg = ctypes.create_string_buffer(512)
    rc = lib().fpl_simulate_fields(ctypes.byref(setup), ctypes.byref(fields), props.ctypes.data,
                                   props.shape[0], tilts.ctypes.data, tilts.shape[0],
                                   _ptr(images), ctypes.byref(warn), _stream(), msg, 512)
    check(rc, msg)
    return warn.value


def generate(dim, origin, voxel_size, points, fiber_ptr, bundle_ptr, layer_ptr, layer_table,
             only_label, label, vector):
    """Call fpl_generate on device tensors; returns (props ndarray, warn mask)."""
    scale, dn, mu, model = (np.ascontiguousarray(a) for a in layer_table)
    bundle_ptr = np.ascontiguousarray(bundle_ptr, dtype=np.int64)
    layer_ptr = np.ascontiguousarray(layer_ptr, dtype=np.int64)
    vol = FplVolume((_c_i64 * 3)(*[int(d) for d in dim]),
                    (ctypes.c_double * 3)(*[float(o) for o in origin]), float(voxel_size))
    fbs = FplFiberBundles(
        len(bundle_ptr) - 1,
        int(fiber_ptr.shape[0]) - 1, int(points.shape[0]), _ptr(points), _ptr(fiber_ptr),
        bundle_ptr.ctypes.data, layer_ptr.ctypes.data,
        scale.ctypes.data if scale.size else None, dn.ctypes.data if dn.size else None,
        mu.ctypes.data if mu.size else None, model.ctypes.data if model.size else None)
    L = lib()
    n_props = L.fpl_num_properties(ctypes.byref(fbs))
    props = np.zeros((n_props, 2), dtype=np.float64)
    warn = ctypes.c_uint32(0)
    msg = ctypes.create_string_buffer(512)
    rc = L.fpl_generate(ctypes.byref(vol), ctypes.byref(fbs), int(bool(only_label)),
                        _ptr(label), _ptr(vector), props.ctypes.data, ctypes.byref(warn),
                        _stream(), msg, 512)
    check(rc, msg)
    return props, warn.value


def simulate(setup, dim, label, vector, props, tilts, images):
    """Call fpl_simulate on device tensors; returns the warn mask."""
    props = np.ascontiguousarray(props, dtype=np.float64)
    tilts = np.ascontiguousarray(tilts, dtype=np.float64).reshape(-1, 2)
    dim_arr = np.ascontiguousarray(dim, dtype=np.int64)
    warn = ctypes.c_uint32(0)
    msg = ctypes.create_string_buffer(512)
    rc = lib().fpl_simulate(ctypes.byref(setup), dim_arr.ctypes.data, _ptr(label), _ptr(vector),
                            props.ctypes.data, props.shape[0], tilts.ctypes.data,
                            tilts.shape[0], _ptr(images), ctypes.byref(warn), _stream(), msg,
                            512)
    check(rc, msg)
    return warn.value


def simulate_noalloc(setup, dim, label, vector, props, tilts, images):
    """fpl_simulate_noalloc (static device scratch, no allocation); returns the warn mask."""
    props = np.ascontiguousarray(props, dtype=np.float64)
    tilts = np.ascontiguousarray(tilts, dtype=np.float64).reshape(-1, 2)
    dim_arr = np.ascontiguousarray(dim, dtype=np.int64)
    warn = ctypes.c_uint32(0)
    msg = ctypes.create_string_buffer(512)
    rc = lib().fpl_simulate_noalloc(ctypes.byref(setup), dim_arr.ctypes.data, _ptr(label),
                                    _ptr(vector), props.ctypes.data, props.shape[0],
                                    tilts.ctypes.data, tilts.shape[0], _ptr(images),
                                    ctypes.byref(warn), _stream(), msg, 512)
    check(rc, msg)
    return warn.value


def _check_array(name, a, shape, dtype):
    """The shape, element type and contiguity a raw-pointer argument must have."""
    if tuple(a.shape) != tuple(int(d) for d in shape):
        raise ValueError(f'{name} has shape {tuple(a.shape)}, expected {tuple(shape)}')
    if str(a.dtype).split('.')[-1] != dtype:
        raise ValueError(f'{name} has dtype {a.dtype}, expected {dtype}')
    contiguous = a.is_contiguous() if hasattr(a, 'is_contiguous') else a.flags['C_CONTIGUOUS']
    if not contiguous:
        raise ValueError(f'{name} must be contiguous')


def simulate_adjoint(setup, dim, label, vector, props, tilts, grad_images, grad_vector=None):
    """fpl_simulate_adjoint on device tensors; returns dL/d props as (n_props, 2) float64.

    The kernel reads and writes through raw pointers: every array's shape,
    element type and contiguity is checked against the setup and dim first."""
    props = np.ascontiguousarray(props, dtype=np.float64)
    tilts = np.ascontiguousarray(tilts, dtype=np.float64).reshape(-1, 2)
    dim_arr = np.ascontiguousarray(dim, dtype=np.int64)
    nx, ny, nz = (int(d) for d in dim_arr)
    _check_array('tissue', label, (nx, ny, nz), 'int32')
    _check_array('optical_axis', vector, (nx, ny, nz, 3), 'float32')
    _check_array('grad_images', grad_images, (tilts.shape[0], nx, ny, setup.n_rho), 'float64')
    if grad_vector is not None:
        _check_array('grad_vector', grad_vector, (nx, ny, nz, 3), 'float32')
    grad_props = np.zeros(props.shape, dtype=np.float64)
    msg = ctypes.create_string_buffer(512)
    rc = lib().fpl_simulate_adjoint(ctypes.byref(setup), dim_arr.ctypes.data, _ptr(label),
                                    _ptr(vector), props.ctypes.data, props.shape[0],
                                    tilts.ctypes.data, tilts.shape[0], _ptr(grad_images),
                                    grad_props.ctypes.data, _ptr(grad_vector), _stream(), msg,
                                    512)
    check(rc, msg)
    return grad_props


def count_voxel_steps(setup, dim, theta, phi):
    dim_arr = np.ascontiguousarray(dim, dtype=np.int64)
    n = int(lib().fpl_count_voxel_steps(ctypes.byref(setup), dim_arr.ctypes.data, float(theta),
                                        float(phi)))
    if n < 0:
        raise ValueError(f'tilt theta = {theta} needs more than 2^24 z-steps per ray')
    return n


def profile_enable(on=True):
    lib().fpl_profile_enable(int(bool(on)))


def profile_read(name):
    """(total milliseconds, launches) of a kernel since profile_enable()."""
    ms = ctypes.c_double(0)
    n = ctypes.c_int64(0)
    lib().fpl_profile_read(name.encode(), ctypes.byref(ms), ctypes.byref(n))
    return ms.value, n.value


def optic_resample(images, sigma, scale, out):
    """fpl_optic_resample on float64 device tensors (nx, ny, n_rho) -> out."""
    nx, ny, nr = images.shape
    msg = ctypes.create_string_buffer(512)
    rc = lib().fpl_optic_resample(_ptr(images), nx, ny, nr, float(sigma), float(scale),
                                  _ptr(out), _stream(), msg, 512)
    check(rc, msg)


def epa(images, t, d, r):
    """fpl_epa on a float64 device tensor (..., n_rho) -> t, d, r (...)."""
    nr = images.shape[-1]
    n_pix = images.numel() // nr
    msg = ctypes.create_string_buffer(512)
    rc = lib().fpl_epa(_ptr(images), n_pix, nr, _ptr(t), _ptr(d), _ptr(r), _stream(), msg, 512)
    check(rc, msg)


def rofl(images, tilt_angle, gain, dir_offset, mask, out, grad_mode=False):
    """fpl_rofl on a float64 device tensor (5, ..., n_rho) -> out (8, ...)."""
    nr = images.shape[-1]
    n_pix = images.numel() // (5 * nr)
    msg = ctypes.create_string_buffer(512)
    rc = lib().fpl_rofl(_ptr(images), n_pix, nr, float(tilt_angle), float(gain),
                        float(dir_offset), int(bool(grad_mode)), _ptr(mask), _ptr(out),
                        _stream(), msg, 512)
    check(rc, msg)
