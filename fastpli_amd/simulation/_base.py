"""Parameter surface of :class:`Simpli`, declared as a table.

Contract kept from the reference (src/fastpli/simulation/__simpli.py:35-416,
simpli.py:44-60, 518-535): the property names and defaults, what each setter
accepts and the errors it raises (type and message, in the same order), the
"frozen class" behaviour, and ``get_dict`` / ``set_dict`` (property name ->
value, arrays as lists, without the fiber bundles).

Design: every property is one ``Param`` entry of the class body (default,
checks on the raw value, a conversion, checks on the converted value, an
optional custom assignment); each entry is a descriptor that stores the value
in the instance's ``_<name>`` slot, and ``get_dict`` walks the same entries.
"""

import warnings

import numpy as np

from .. import objects
from ..objects._sealed import Sealed
from ..objects.layers import apply_rules


class Param:
    """One Simpli property (a data descriptor over the attribute `_<name>`)."""

    def __init__(self, default, before=(), convert=None, after=(), assign=None, doc=''):
        self.default = default
        self.before = before    # rules on the value as given
        self.convert = convert  # value -> stored value
        self.after = after      # rules on the converted value
        self.assign = assign    # (instance, value) -> None instead of storing it
        self.__doc__ = doc

    def __set_name__(self, owner, name):
        self.name = name
        self.slot = '_' + name

    def __get__(self, obj, owner=None):
        if obj is None:
            return self
        try:
            return obj.__dict__[self.slot]
        except KeyError:
            raise AttributeError(self.name) from None

    def __set__(self, obj, value):
        apply_rules(self.before, value)
        if self.convert is not None:
            value = self.convert(value)
        apply_rules(self.after, value)
        if self.assign is not None:
            self.assign(obj, value)
        else:
            obj.__dict__[self.slot] = value


def _voxel_size(obj, value):
    """A new voxel size keeps the volume of interest: dim follows it."""
    voi = obj.get_voi() if obj._voxel_size is not None and obj._dim is not None else None
    obj.__dict__['_voxel_size'] = value
    if voi is not None:
        obj.set_voi(*voi)
        obj._print('voxel_size: recalculated dimensions')


def _bundles(value):
    return None if value is None else objects.FiberBundles(value)


_positive = lambda name: ((lambda v: v <= 0, ValueError, f'{name} <= 0'),)  # noqa: E731


class _SimpliBase(Sealed):
    """Frozen parameter container shared by :class:`Simpli` (order = get_dict order)."""

    # ------------------------------------------------------------ volume
    dim = Param(None, convert=np.array,
                after=((lambda d: d.dtype != int, TypeError, 'dim is not np.ndarray(int)'),
                       (lambda d: d.size != 3, TypeError, 'dim.size != 3'),
                       (lambda d: d.ndim != 1, TypeError, 'dim.ndim != 1')),
                doc='dimension of the volume in voxels: (3)-np.ndarray(int)')
    dim_origin = Param(np.zeros(3), convert=lambda v: np.array(v, dtype=float),
                       doc='origin of the volume in micro meter: (3)-np.ndarray')
    voxel_size = Param(None, before=_positive('voxel_size'), convert=float, assign=_voxel_size,
                       doc='side length of a voxel in micro meter')
    # ------------------------------------------------------------ generation
    fiber_bundles = Param(None, convert=_bundles, doc='fiber model: FiberBundles')
    # ------------------------------------------------------------ simulation
    filter_rotations = Param(
        None, convert=lambda v: np.array(v, dtype=float),
        after=((lambda r: r.size == 0 or r.ndim != 1, TypeError, 'filter_rotations : nx1'),),
        doc='polarization filter rotations in radian: (n)-np.ndarray')
    flip_z_beam = Param(False, convert=bool, doc='light travels along -z')
    interpolate = Param(
        'Slerp',
        before=((lambda v: v not in ('NN', 'Lerp', 'Slerp'), ValueError,
                 'Only \'NN\', \'Lerp\' or \'Slerp\' are supported'),),
        doc="optical axis interpolation along tilted rays: 'NN', 'Lerp' or 'Slerp'")
    light_intensity = Param(None, convert=float, doc='initial light intensity [a.u.]')
    pixel_size = Param(None, before=_positive('pixel_size'), convert=float,
                       doc='side length of a pixel of the optical image in micro meter')
    optical_sigma = Param(None,
                          before=((lambda v: not isinstance(v, (int, float)), TypeError,
                                   'optical_sigma is not a number'),
                                  (lambda v: v < 0, ValueError, 'optical_sigma is < 0')),
                          doc='standard deviation of the optical blur in pixels')
    step_size = Param(1.0, convert=float, doc='light step along a ray in voxels')
    tilts = Param(None,
                  before=((lambda v: not isinstance(v, (list, tuple, np.ndarray)), TypeError,
                           'tilts is not a list or array'),),
                  convert=lambda v: np.array(v, ndmin=2),
                  after=((lambda t: t.ndim != 2 or t.shape[-1] != 2, TypeError,
                          'tilts.shape != nx2'),),
                  doc='tilt angles [[theta, phi], ...] in radian')
    tissue_refrection = Param(1, convert=float, doc='refractive index of the tissue')
    background_absorption = Param(0, convert=float, doc='absorption of the background [1/mm]')
    untilt_sensor_view = Param(True, convert=bool, doc='untilt the images of tilted views')
    wavelength = Param(None, convert=float, doc='wavelength of the light in nano meter')
    # storage of label / axis inside the ray march: this package's precision
    # modes, no counterpart in the reference
    field_storage = Param(
        'f32',
        before=((lambda v: v not in ('f32', 'f16', 'angle16'), ValueError,
                 'field_storage must be \'f32\', \'f16\' or \'angle16\''),),
        doc="storage of the fields inside the march: 'f32' (int32 + float32, exact), "
        "'f16' (uint16 + fp16, 8 B/voxel) or 'angle16' (uint16 + 16-bit angles, 6 B/voxel)")
    # ------------------------------------------------------------ analysis
    noise_model = Param(None,
                        before=((lambda v: not callable(v), TypeError,
                                 'noise model is not callable'),),
                        doc='noise applied to the resampled image: callable')
    # ------------------------------------------------------------ other
    omp_num_threads = Param(
        1,
        before=((lambda v: not isinstance(v, int), TypeError, 'num_threads != int'),
                (lambda v: v <= 0, TypeError, 'num_threads <= 0')),
        doc='kept for API compatibility; the GPU path uses no OpenMP')
    verbose = Param(False, convert=bool, doc='print progress')

    def __init__(self):
        for p in self._params():
            default = p.default
            self.__dict__[p.slot] = default.copy() if isinstance(default, np.ndarray) else default

    @classmethod
    def _params(cls):
        seen, out = set(), []
        for klass in reversed(cls.__mro__):
            for name, v in vars(klass).items():
                if isinstance(v, Param) and name not in seen:
                    seen.add(name)
                    out.append(v)
        return out

    def _frozen_label(self):
        return self.__class__.__name__

    def _print(self, msg):
        if self._verbose:
            print(msg)

    def get_dict(self):
        """The properties as a dictionary (arrays as lists, without fiber_bundles)."""
        out = {}
        for p in self._params():
            if p.name == 'fiber_bundles':
                continue
            v = self.__dict__[p.slot]
            out[p.name] = v.tolist() if isinstance(v, np.ndarray) else v
        return out

    def set_dict(self, data):
        """Sets the properties of a dictionary as get_dict returns it."""
        for key, value in data.items():
            if key.startswith('_'):
                raise ValueError('member variable cant be set directly')
            if value is None:
                warnings.warn('None value in dict detected')
                continue
            setattr(self, key, value)

    # ------------------------------------------------------- volume of interest
    def get_voi(self):
        """(v_min, v_max) of the volume in micro meter."""
        for value, name in ((self._voxel_size, 'voxel_size'), (self._dim, 'dim'),
                            (self._dim_origin, 'dim_origin')):
            if value is None:
                raise ValueError(f'{name} is not set, voi can\'t be calculated')
        v_min = np.array(self._dim_origin, dtype=float)
        return v_min, v_min + self._dim * self._voxel_size

    def set_voi(self, v_min, v_max):
        """Sets origin and dim from the volume [v_min, v_max] in micro meter."""
        v_min = np.array(v_min, dtype=float)
        v_max = np.array(v_max, dtype=float)
        apply_rules(((lambda _: v_min.ndim != 1 or v_max.ndim != 1, TypeError,
                      'v_min,v_max : ndim != 1'),
                     (lambda _: v_min.size != 3 or v_max.size != 3, TypeError,
                      'v_min,v_max : size != 1'),
                     (lambda _: np.any(v_min >= v_max), ValueError, 'v_min >= v_max'),
                     (lambda _: self._voxel_size is None, TypeError, 'voxel_size is not set yet')),
                    None)
        self.__dict__['_dim'] = np.array(np.round((v_max - v_min) / self._voxel_size), dtype=int)
        self.__dict__['_dim_origin'] = v_min

    # ------------------------------------------------------------ checks
    def _check_volume_input(self):
        for value, name in ((self._dim, 'dim'), (self._dim_origin, 'dim_origin'),
                            (self._voxel_size, 'voxel_size')):
            if value is None:
                raise ValueError(f'{name} not set')
        self.get_voi()

    def _check_generation_input(self):
        if not self._fiber_bundles:
            raise ValueError('fiber_bundles is not set')
        for lys in self._fiber_bundles.layers:
            if lys is None or lys.is_none() or any(ly is None for ly in lys):
                raise ValueError('fiber_bundle properties are "None"')

    def _check_simulation_input(self):
        if self._step_size <= 0:
            raise ValueError('step_size <= 0')
        for value, name in ((self._light_intensity, 'light_intensity'),
                            (self._voxel_size, 'voxel_size'), (self._wavelength, 'wavelength'),
                            (self._filter_rotations, 'filter_rotations')):
            if value is None:
                raise ValueError(f'{name} not set')
