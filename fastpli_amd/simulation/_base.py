"""Parameter surface of :class:`Simpli`: properties, validation, dictionaries.

Mirrors fastpli's ``src/fastpli/simulation/__simpli.py:20-448``: the same
attribute names and defaults (:35-64), the same setter validation and error
messages (:97-416), the frozen-class behaviour (unknown attributes raise
``TypeError``, :27-33) and ``get_dict`` / ``set_dict`` (:70-95), which skip
private members whose mangled names start with ``_Simpli``.
"""

import warnings

import numpy as np

from .. import objects


class _SimpliBase:
    """Frozen parameter container shared by :class:`Simpli`."""

    __is_frozen = False

    def __setattr__(self, key, value):
        if self.__is_frozen and not hasattr(self, key):
            raise TypeError(f'{self.__class__.__name__} is a frozen class')
        object.__setattr__(self, key, value)

    def _freeze(self):
        self.__is_frozen = True

    def __init__(self):
        # DIMENSIONS
        self._dim = None
        self._dim_origin = np.array([0, 0, 0], dtype=float)
        self._voxel_size = None
        # GENERATION
        self._fiber_bundles = None
        # SIMULATION
        self._filter_rotations = None
        self._flip_z_beam = False
        self._interpolate = 'Slerp'
        self._light_intensity = None
        self._pixel_size = None
        self._optical_sigma = None
        self._step_size = 1.0
        self._tilts = None
        self._tissue_refrection = 1
        self._background_absorption = 0
        self._untilt_sensor_view = True
        self._wavelength = None
        # storage of the fields inside the march: 'f32' (exact), 'f16', 'angle16'
        self._field_storage = 'f32'
        # ANALYSIS
        self._noise_model = None
        # OTHER
        self._omp_num_threads = 1
        self._verbose = False

    def _print(self, msg):
        if self._verbose:
            print(msg)

    def get_dict(self):
        """All property members as a dictionary (without fiber_bundles)."""
        members = dict()
        for key, value in self.__dict__.items():
            if key == '_fiber_bundles':
                continue
            if key.startswith('_') and not key.startswith('__') and not key.startswith(
                    '_Simpli'):
                if isinstance(value, np.ndarray):
                    members[key[1:]] = value.tolist()
                else:
                    members[key[1:]] = value
        return members

    def set_dict(self, data):
        """Set properties from a dictionary produced by :meth:`get_dict`."""
        for key, value in data.items():
            if key.startswith('_'):
                raise ValueError('member variable cant be set directly')
            if value is not None:
                setattr(self, key, value)
            else:
                warnings.warn('None value in dict detected')

    # ------------------------------------------------------------ volume
    @property
    def dim(self):
        """dimension of volume in voxel: (3)-np.ndarray"""
        return self._dim

    @dim.setter
    def dim(self, dim):
        dim = np.array(dim)
        if dim.dtype != int:
            raise TypeError('dim is not np.ndarray(int)')
        if dim.size != 3:
            raise TypeError('dim.size != 3')
        if dim.ndim != 1:
            raise TypeError('dim.ndim != 1')
        self._dim = dim

    @property
    def dim_origin(self):
        """origin of volume in micro meter: (3)-np.ndarray"""
        return self._dim_origin

    @dim_origin.setter
    def dim_origin(self, dim_origin):
        self._dim_origin = np.array(dim_origin, dtype=float)

    @property
    def voxel_size(self):
        """side length of a voxel in micro meter"""
        return self._voxel_size

    @voxel_size.setter
    def voxel_size(self, voxel_size):
        if voxel_size <= 0:
            raise ValueError('voxel_size <= 0')
        flag = False
        if self._voxel_size is not None and self._dim is not None and \
                self._dim_origin is not None:
            v_min, v_max = self.get_voi()
            flag = True
        self._voxel_size = float(voxel_size)
        if flag and self._dim is not None and self._dim_origin is not None:
            self.set_voi(v_min, v_max)
            self._print('voxel_size: recalculated dimensions')

    @property
    def pixel_size(self):
        """side length of a pixel of the optical image in micro meter"""
        return self._pixel_size

    @pixel_size.setter
    def pixel_size(self, pixel_size):
        if pixel_size <= 0:
            raise ValueError('pixel_size <= 0')
        self._pixel_size = float(pixel_size)

    def get_voi(self):
        """volume of interest: (v_min, v_max) in micro meter"""
        if self._voxel_size is None:
            raise ValueError('voxel_size is not set, voi can\'t be calculated')
        if self._dim is None:
            raise ValueError('dim is not set, voi can\'t be calculated')
        if self._dim_origin is None:
            raise ValueError('dim_origin is not set, voi can\'t be calculated')
        voi = np.zeros((6,))
        voi[::2] = self._dim_origin
        voi[1::2] = voi[::2] + self._dim * self._voxel_size
        return np.array(voi[0::2]), np.array(voi[1::2])

    def set_voi(self, v_min, v_max):
        """set the volume of interest [v_min, v_max] in micro meter"""
        v_min = np.array(v_min, dtype=float)
        v_max = np.array(v_max, dtype=float)
        if v_min.ndim != 1 or v_max.ndim != 1:
            raise TypeError('v_min,v_max : ndim != 1')
        if v_min.size != 3 or v_max.size != 3:
            raise TypeError('v_min,v_max : size != 1')
        if np.any(v_min >= v_max):
            raise ValueError('v_min >= v_max')
        if self._voxel_size is None:
            raise TypeError('voxel_size is not set yet')
        self._dim = np.array(np.round((v_max - v_min) / self._voxel_size), dtype=int)
        self._dim_origin = v_min

    # -------------------------------------------------------- simulation
    @property
    def filter_rotations(self):
        """filter rotation angles in radian: (n)-np.ndarray"""
        return self._filter_rotations

    @filter_rotations.setter
    def filter_rotations(self, filter_rotations):
        filter_rotations = np.array(filter_rotations, dtype=float)
        if filter_rotations.size == 0 or filter_rotations.ndim != 1:
            raise TypeError('filter_rotations : nx1')
        self._filter_rotations = filter_rotations

    @property
    def light_intensity(self):
        """initial light intensity in a.u."""
        return self._light_intensity

    @light_intensity.setter
    def light_intensity(self, light_intensity):
        self._light_intensity = float(light_intensity)

    @property
    def tilts(self):
        """tilting angles [[theta, phi], ...] in radian"""
        return self._tilts

    @tilts.setter
    def tilts(self, tilts):
        if not isinstance(tilts, (list, tuple, np.ndarray)):
            raise TypeError('tilts is not a list or array')
        tilts = np.array(tilts, ndmin=2)
        if tilts.ndim != 2:
            raise TypeError('tilts.shape != nx2')
        if tilts.shape[-1] != 2:
            raise TypeError('tilts.shape != nx2')
        self._tilts = tilts

    @property
    def noise_model(self):
        """noise model applied to the resampled image: callable"""
        return self._noise_model

    @noise_model.setter
    def noise_model(self, noise_model):
        if not callable(noise_model):
            raise TypeError('noise model is not callable')
        self._noise_model = noise_model

    @property
    def optical_sigma(self):
        """sigma of the optical blur in pixel size"""
        return self._optical_sigma

    @optical_sigma.setter
    def optical_sigma(self, optical_sigma):
        if not isinstance(optical_sigma, (int, float)):
            raise TypeError('optical_sigma is not a number')
        if optical_sigma < 0:
            raise ValueError('optical_sigma is < 0')
        self._optical_sigma = optical_sigma

    @property
    def wavelength(self):
        """wavelength of the light in nano meter"""
        return self._wavelength

    @wavelength.setter
    def wavelength(self, wavelength):
        self._wavelength = float(wavelength)

    @property
    def background_absorption(self):
        """absorption coefficient of the background"""
        return self._background_absorption

    @background_absorption.setter
    def background_absorption(self, background_absorption):
        self._background_absorption = float(background_absorption)

    @property
    def tissue_refrection(self):
        """tissue refraction index used for tilted light paths"""
        return self._tissue_refrection

    @tissue_refrection.setter
    def tissue_refrection(self, tissue_refrection):
        self._tissue_refrection = float(tissue_refrection)

    @property
    def step_size(self):
        """light step inside the tissue in voxel"""
        return self._step_size

    @step_size.setter
    def step_size(self, step_size):
        self._step_size = float(step_size)

    @property
    def field_storage(self):
        """storage of label / optical axis inside the ray march (this package's
        precision modes, no counterpart in the reference): 'f32' (int32 +
        float32, exact input), 'f16' (uint16 + fp16, 8 B/voxel) or 'angle16'
        (uint16 + 16-bit polar/azimuth codes, 6 B/voxel)"""
        return self._field_storage

    @field_storage.setter
    def field_storage(self, field_storage):
        if field_storage not in ('f32', 'f16', 'angle16'):
            raise ValueError('field_storage must be \'f32\', \'f16\' or \'angle16\'')
        self._field_storage = field_storage

    @property
    def verbose(self):
        return self._verbose

    @verbose.setter
    def verbose(self, verbose):
        self._verbose = bool(verbose)

    @property
    def untilt_sensor_view(self):
        """automatic untilting of the images"""
        return self._untilt_sensor_view

    @untilt_sensor_view.setter
    def untilt_sensor_view(self, untilt_sensor_view):
        self._untilt_sensor_view = bool(untilt_sensor_view)

    @property
    def flip_z_beam(self):
        """flip the light's z direction"""
        return self._flip_z_beam

    @flip_z_beam.setter
    def flip_z_beam(self, flip_z_beam):
        self._flip_z_beam = bool(flip_z_beam)

    @property
    def fiber_bundles(self):
        """fiber bundles: FiberBundles"""
        return self._fiber_bundles

    @fiber_bundles.setter
    def fiber_bundles(self, fbs):
        if fbs is None:
            self._fiber_bundles = None
        else:
            self._fiber_bundles = objects.FiberBundles(fbs)

    # ------------------------------------------------------------ checks
    def _check_volume_input(self):
        if self._dim is None:
            raise ValueError('dim not set')
        if self._dim_origin is None:
            raise ValueError('dim_origin not set')
        if self._voxel_size is None:
            raise ValueError('voxel_size not set')
        self.get_voi()

    def _check_generation_input(self):
        if not self._fiber_bundles:
            raise ValueError('fiber_bundles is not set')
        for lys in self._fiber_bundles.layers:
            if lys is None or lys.is_none():
                raise ValueError('fiber_bundle properties are "None"')
            for ly in lys:
                if ly is None:
                    raise ValueError('fiber_bundle properties are "None"')

    def _check_simulation_input(self):
        if self._step_size <= 0:
            raise ValueError('step_size <= 0')
        if self._light_intensity is None:
            raise ValueError('light_intensity not set')
        if self._voxel_size is None:
            raise ValueError('voxel_size not set')
        if self._wavelength is None:
            raise ValueError('wavelength not set')
        if self._filter_rotations is None:
            raise ValueError('filter_rotations not set')
