"""Simpli: the fastpli simulation API on MI355X.

Drop-in counterpart of fastpli's ``src/fastpli/simulation/simpli.py`` for the
simulation hot path: ``generate_tissue`` (:61-104) and ``run_simulation``
(:125-187) keep their signatures, validation, warnings, output dtypes and
layouts.  The work runs in libfastpli_amd.so (HIP kernels for gfx950) instead
of the pybind11/OpenMP core; there is no CPU fallback.

Extensions of this package (not in the reference):
  * ``generate_tissue_device`` / ``run_simulation`` with torch tensors on the
    GPU keep the fields resident in HBM (no host round trip);
  * ``generate_tissue_fields`` / ``pack_fields`` give ``MarchFields``, the
    fields in the march's own record layout, and ``run_simulation_fields``
    simulates them without repacking (generation writes the records itself);
  * ``run_simulation_tilts`` runs several tilts in one device call;
  * ``count_voxel_steps`` gives the exact Mueller voxel-step count of a run.
"""

import warnings

import numpy as np

from .. import _lib
from ..objects._sealed import edit_count
from ._base import _SimpliBase


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('fastpli_amd needs a HIP device (torch.cuda.is_available() is False)')
    return torch


def _is_tensor(x):
    try:
        import torch
    except ImportError:  # pragma: no cover
        return False
    return isinstance(x, torch.Tensor)


class MarchFields:
    """Label and optical-axis fields in the ray march's record layout, on the GPU.

    Made by ``Simpli.generate_tissue_fields`` (the voxelizer writes the records
    as it writes the tissue) or ``Simpli.pack_fields`` (from C-order fields),
    consumed by ``Simpli.run_simulation_fields``.  A snapshot: later changes of
    the arrays it was made from do not reach it.

    storage: 'f32' (16 B/voxel, exact), 'f16' (8 B) or 'angle16' (6 B).
    label_min / label_max: bounds of the labels it holds, checked against the
    property table before every simulation (the reference's CheckInput).
    """

    def __init__(self, dim, storage, buffer):
        self.dim = tuple(int(d) for d in dim)
        self.storage = storage
        self.buffer = buffer
        self.label_min = 0
        self.label_max = 0

    @classmethod
    def empty(cls, dim, storage, device):
        torch = _torch()
        nbytes, _ = _lib.fields_layout(dim, storage)
        return cls(dim, storage, torch.empty(nbytes, dtype=torch.uint8, device=device))

    def struct(self):
        return _lib.fields_struct(self.buffer, self.dim, self.storage, self.label_min,
                                  self.label_max)

    def update(self, st):
        self.label_min = int(st.label_min)
        self.label_max = int(st.label_max)

    @property
    def nbytes(self):
        return int(self.buffer.numel())


class Simpli(_SimpliBase):
    """Simulation of 3D-PLI images (tissue generation + light propagation)."""

    def __init__(self, mpi_comm=None):
        super().__init__()
        if mpi_comm is not None:
            raise NotImplementedError(
                'MPI domain decomposition is not supported: one GPU holds the whole volume')
        self.mpi = None
        self._model_cache = None  # (fiber bundles, edit count, device, flattened model)
        self._seal()

    # ------------------------------------------------------------ generation
    def _dim_checked(self):
        dim = [int(d) for d in self._dim]
        if any(d <= 0 for d in dim):
            raise ValueError(f'dim.global[any] <= 0: [{dim[0]},{dim[1]},{dim[2]}]')
        return dim

    def _device_fiber_model(self, dev):
        """The fiber model flattened and its points on the device, kept while the
        model object is the same and unedited (objects' edit counter)."""
        torch = _torch()
        c = self._model_cache
        if (c is not None and c[0] is self._fiber_bundles and c[1] == edit_count() and
                c[2] == dev):
            return c[3]
        points, fiber_ptr, bundle_ptr, layer_ptr, table = self._fiber_bundles.to_flat()
        points_d = torch.from_numpy(points).to(dev)
        fiber_ptr_d = torch.from_numpy(fiber_ptr).to(dev)
        model = (points_d, fiber_ptr_d, bundle_ptr, layer_ptr, table)
        self._model_cache = (self._fiber_bundles, edit_count(), dev, model)
        return model

    def _generate(self, only_tissue, fields, with_arrays=True):
        self._check_volume_input()
        self._check_generation_input()
        dim = self._dim_checked()
        torch = _torch()
        dev = torch.device('cuda', torch.cuda.current_device())
        model = self._device_fiber_model(dev)
        tissue = optical_axis = None
        if with_arrays:
            tissue = torch.empty(dim, dtype=torch.int32, device=dev)
            if not only_tissue:
                optical_axis = torch.empty(dim + [3], dtype=torch.float32, device=dev)
        st = fields.struct() if fields is not None else _lib.FplFields()
        props, warn = _lib.generate_fields(dim, self._dim_origin, self._voxel_size, *model,
                                           tissue, optical_axis, st)
        if fields is not None:
            fields.update(st)
        props[0, 1] = self._background_absorption
        _lib.raise_warnings(warn)
        if st.label_max == 0:
            warnings.warn(
                'All labels are 0. Usually this means, that the VOI contains ' +
                'no fiber_bundles or that the voxel_size is to large.', UserWarning)
        return tissue, optical_axis, props

    def generate_tissue_device(self, only_tissue=False):
        """Voxelize the fiber bundles into device tensors.

        Returns (tissue int32 (nx, ny, nz), optical_axis float32 (nx, ny, nz, 3)
        or None, tissue_properties float64 ndarray (n, 2)).
        """
        return self._generate(only_tissue, None)

    def generate_tissue_fields(self, field_storage=None, with_arrays=True):
        """Voxelize the fiber bundles into march fields (and the device arrays).

        The voxelizer writes every voxel's march record as it writes the tissue,
        so ``run_simulation_fields`` needs no repacking pass.  field_storage
        defaults to the ``field_storage`` property.  Returns (tissue, optical_axis,
        tissue_properties, fields); tissue and optical_axis are None when
        with_arrays is False (only the records are written).
        """
        self._check_volume_input()
        self._check_generation_input()
        storage = self._field_storage if field_storage is None else field_storage
        if storage not in _lib.STORAGE:
            raise ValueError('field_storage must be \'f32\', \'f16\' or \'angle16\'')
        torch = _torch()
        dev = torch.device('cuda', torch.cuda.current_device())
        fields = MarchFields.empty(self._dim_checked(), storage, dev)
        tissue, optical_axis, props = self._generate(False, fields, with_arrays)
        return tissue, optical_axis, props, fields

    def pack_fields(self, tissue, optical_axis, field_storage=None):
        """March fields from C-order fields (numpy arrays or device tensors)."""
        storage = self._field_storage if field_storage is None else field_storage
        if storage not in _lib.STORAGE:
            raise ValueError('field_storage must be \'f32\', \'f16\' or \'angle16\'')
        dim, tissue_, axis_, _ = self._prepare_fields(tissue, optical_axis, np.zeros((1, 2)))
        fields = MarchFields.empty(dim, storage, tissue_.device)
        st = fields.struct()
        _lib.pack_fields(dim, tissue_, axis_, st)
        fields.update(st)
        return fields

    def generate_tissue(self, only_tissue=False):
        """Generation of the discretized tissue.

        Returns (tissue int32 (nx, ny, nz), optical_axis float32 (nx, ny, nz, 3)
        or shape (0,) if only_tissue, tissue_properties float64 (n, 2)).
        """
        self._print('Generate Tissue')
        if self._dim is not None:
            self._print(f'Memory needed: ~{np.ceil(self.memory_usage()):.0f} MB')
        tissue, optical_axis, props = self.generate_tissue_device(only_tissue)
        tissue = tissue.cpu().numpy()
        if optical_axis is None:
            optical_axis = np.empty((0,), dtype=np.float32)
        else:
            optical_axis = optical_axis.cpu().numpy()
        return tissue, optical_axis, props

    # ------------------------------------------------------------ simulation
    def _setup(self):
        self._check_simulation_input()
        return _lib.make_setup(self._step_size, self._light_intensity, self._voxel_size,
                               self._wavelength, self._tissue_refrection, self._interpolate,
                               self._untilt_sensor_view, self._flip_z_beam,
                               self._filter_rotations, self._field_storage)

    def _prepare_fields(self, tissue, optical_axis, tissue_properties):
        """Validation of simpli.py:154-180; returns device tensors + props."""
        tissue_properties = np.array(tissue_properties)
        if tissue_properties.ndim != 2:
            raise ValueError('tissue_properties.ndim !=2')
        if tissue_properties.shape[1] != 2:
            raise ValueError('tissue_properties.shape[1] != 2')
        torch = _torch()
        dev = torch.device('cuda', torch.cuda.current_device())
        if _is_tensor(tissue) or _is_tensor(optical_axis):
            tissue_ = torch.as_tensor(tissue, device=dev)
            optical_axis_ = torch.as_tensor(optical_axis, device=dev)
            if tissue_.dtype != torch.int32:
                warnings.warn('tissue is copied', UserWarning)
                tissue_ = tissue_.to(torch.int32)
            if optical_axis_.dtype != torch.float32:
                warnings.warn('optical_axis is copied', UserWarning)
                optical_axis_ = optical_axis_.to(torch.float32)
            tissue_ = tissue_.contiguous()
            optical_axis_ = optical_axis_.contiguous()
            t_shape, o_shape = tuple(tissue_.shape), tuple(optical_axis_.shape)
        else:
            tissue_np = np.asarray(tissue, dtype=np.int32)
            axis_np = np.asarray(optical_axis, dtype=np.float32)
            if tissue_np is not tissue:
                warnings.warn('tissue is copied', UserWarning)
            if axis_np is not optical_axis:
                warnings.warn('optical_axis is copied', UserWarning)
            t_shape, o_shape = tissue_np.shape, axis_np.shape
            tissue_ = None
            optical_axis_ = None
        if len(t_shape) != 3:
            raise ValueError('tissue.ndim != 3')
        if len(o_shape) != 4:
            raise ValueError('optical_axis.ndim != 4')
        if o_shape[3] != 3:
            raise ValueError('optical_axis.shape[3] != 3')
        if tuple(t_shape) != tuple(o_shape[:-1]):
            raise ValueError('not np.equal(tissue.shape, optical_axis.shape[:-1])')
        dim = self._dim_checked()
        if int(np.prod(dim)) != int(np.prod(t_shape)):
            raise ValueError('dim_.local.x() * dim_.local.y() * dim_.local.z() != '
                             'label_field_.size()')
        if tissue_ is None:
            tissue_ = torch.from_numpy(np.ascontiguousarray(tissue_np)).to(dev)
            optical_axis_ = torch.from_numpy(np.ascontiguousarray(axis_np)).to(dev)
        return dim, tissue_, optical_axis_, tissue_properties

    def run_simulation_tilts(self, tissue, optical_axis, tissue_properties, tilts):
        """Simulate several tilts in one device call.

        Returns images float64 (n_tilts, nx, ny, n_rho): a torch tensor on the
        GPU when the fields are torch tensors, else a numpy array.
        """
        self._print('Run simulation')
        self._check_volume_input()
        setup, _keep = self._setup()
        as_tensor = _is_tensor(tissue) or _is_tensor(optical_axis)
        dim, tissue_, axis_, props = self._prepare_fields(tissue, optical_axis,
                                                          tissue_properties)
        torch = _torch()
        tilts = np.array(tilts, dtype=np.float64, ndmin=2).reshape(-1, 2)
        images = torch.empty((tilts.shape[0], dim[0], dim[1], setup.n_rho),
                             dtype=torch.float64,
                             device=tissue_.device)
        warn = _lib.simulate(setup, dim, tissue_, axis_, props, tilts, images)
        _lib.raise_warnings(warn)
        _lib.check_images(warn)
        if as_tensor:
            return images
        return images.cpu().numpy()

    def run_simulation_fields(self, fields, tissue_properties, tilts):
        """Simulate several tilts of march fields in one device call.

        Returns images float64 (n_tilts, nx, ny, n_rho) as a torch tensor on the
        GPU.  The fields' storage applies (the field_storage property does not).
        """
        self._print('Run simulation')
        self._check_volume_input()
        if not isinstance(fields, MarchFields):
            raise TypeError('fields must be MarchFields')
        if tuple(self._dim_checked()) != fields.dim:
            raise ValueError('dim_.local.x() * dim_.local.y() * dim_.local.z() != '
                             'label_field_.size()')
        tissue_properties = np.array(tissue_properties)
        if tissue_properties.ndim != 2:
            raise ValueError('tissue_properties.ndim !=2')
        if tissue_properties.shape[1] != 2:
            raise ValueError('tissue_properties.shape[1] != 2')
        setup, _keep = self._setup()
        torch = _torch()
        tilts = np.array(tilts, dtype=np.float64, ndmin=2).reshape(-1, 2)
        dim = fields.dim
        images = torch.empty((tilts.shape[0], dim[0], dim[1], setup.n_rho), dtype=torch.float64,
                             device=fields.buffer.device)
        warn = _lib.simulate_fields(setup, fields.struct(), tissue_properties, tilts, images)
        _lib.raise_warnings(warn)
        _lib.check_images(warn)
        return images

    def run_simulation(self, tissue, optical_axis, tissue_properties, theta, phi):
        """Simulate the PLI images of one tilt (theta, phi).

        Returns images float64 (nx, ny, n_rho) with a pixel size of voxel_size.
        """
        images = self.run_simulation_tilts(tissue, optical_axis, tissue_properties,
                                           [[theta, phi]])
        return images[0]

    def count_voxel_steps(self, theta=0.0, phi=0.0):
        """Mueller voxel-steps of one tilt: n_rho x rays x z-march iterations."""
        self._check_volume_input()
        setup, _keep = self._setup()
        return _lib.count_voxel_steps(setup, self._dim_checked(), theta, phi)

    # ------------------------------------------------------------ optic / analysis
    def apply_optic_resample(self, data, shift=(0, 0), mp_pool=None):
        """Gaussian blur + CCD resampling of an image stack (__simpli.py:491-549).

        The blur has standard deviation optical_sigma / scale voxels with
        scale = voxel_size / pixel_size.  Runs on the GPU; returns a numpy array
        (squeezed as the reference) or a torch tensor for tensor input.
        ``mp_pool`` is accepted for API compatibility and ignored.
        """
        self._print('Apply optic resample')
        if self._optical_sigma is None:
            raise ValueError('optical_sigma is None')
        if self._voxel_size is None:
            raise ValueError('voxel_size not set')
        if self._pixel_size is None:
            raise ValueError('pixel_size not set')
        as_tensor = _is_tensor(data)
        torch = _torch()
        dev = torch.device('cuda', torch.cuda.current_device())
        if as_tensor:
            x = data.to(device=dev)
            while x.dim() < 3:
                x = x.unsqueeze(-1)
        else:
            x = torch.from_numpy(np.ascontiguousarray(np.atleast_3d(np.array(data)))).to(dev)
        if x.dim() > 3:
            raise TypeError('data can be 1d, 2d or 3d')
        shift = np.array(shift, int)
        if shift.size != 2 or shift.ndim != 1:
            raise TypeError('shift has to be (x,y)')
        if shift[0] > 0 and shift[1] > 0:
            x = x[shift[0]:, shift[1]:, ...]
        scale = self._voxel_size / self._pixel_size
        size = np.array(np.round(np.array(x.shape[0:2]) * scale), dtype=int)
        if np.amin(size) == 0:
            raise ValueError(f'voxel_size {self._voxel_size}, ' +
                             f'pixel_size {self._pixel_size} ' +
                             f'and data shape {tuple(x.shape[0:2])} ' +
                             f'result in optical image size of {size}')
        x = x.to(torch.float64).contiguous()
        out = torch.empty((int(size[0]), int(size[1]), x.shape[2]), dtype=torch.float64,
                          device=dev)
        _lib.optic_resample(x, self._optical_sigma / scale, scale, out)
        if as_tensor:
            return out.squeeze()
        return np.squeeze(out.cpu().numpy())

    def apply_optic(self, data, mp_pool=None):
        """Optical resampling then the noise model (__simpli.py:462-489).

        The noise model is the user's Python callable and runs on the host, as
        in the reference.  Returns (resampled, noisy) numpy arrays.
        """
        self._print('Apply optic')
        if self._noise_model is None:
            raise ValueError('noise_model not set')
        resampled = self.apply_optic_resample(data, mp_pool=mp_pool)
        if _is_tensor(resampled):
            resampled = resampled.cpu().numpy()
        if np.amin(resampled) < 0:
            raise AssertionError('intensity < 0 detected')
        noisy = self._noise_model(np.array(resampled, copy=False))
        return resampled, noisy

    def apply_epa(self, data, mask=None):
        """Transmittance, direction and retardation (__simpli.py:597-621, epa.py:9-45).

        Computed on the GPU in float64; float32 input gives float32 results
        (the reference computes those in float32).
        """
        self._print('Apply epa')
        torch = _torch()
        dev = torch.device('cuda', torch.cuda.current_device())
        as_tensor = _is_tensor(data)
        if as_tensor:
            x = data.to(device=dev)
            itemsize = x.element_size()
        else:
            arr = np.asarray(data)
            itemsize = arr.dtype.itemsize
            x = torch.from_numpy(np.ascontiguousarray(arr)).to(dev)
        x = x.to(torch.float64).contiguous()
        shape = tuple(x.shape[:-1])
        t = torch.empty(shape, dtype=torch.float64, device=dev)
        d = torch.empty_like(t)
        r = torch.empty_like(t)
        _lib.epa(x, t, d, r)
        out = [t, d, r]
        if itemsize <= 4:
            out = [o.to(torch.float32) for o in out]
        if not as_tensor:
            out = [o.cpu().numpy() for o in out]
        if mask is not None:
            m = np.asarray(mask) if not as_tensor else mask
            for o in out:
                o[~m] = float('nan')
        return tuple(out)

    def apply_rofl(self, data, mask=None, mp_pool=None, grad_mode=False):
        """Tilt analysis: direction, inclination and t_rel per pixel (__simpli.py:623-706).

        data: (5, x, y, rho) images of the tilts (0, 0), (tau, 0), (tau, 90 deg),
        (tau, 180 deg), (tau, 270 deg).  Returns (directionmap, inclmap, trelmap,
        (dirdevmap, incldevmap, treldevmap, funcmap, itermap)).  Runs one GPU
        thread per pixel in float64: the reference's fit as
        scipy.optimize.leastsq runs it (maxfev 150), MINPACK lmdif with forward
        differences, or with ``grad_mode=True`` lmder with the reference's
        analytic per-view Jacobian; itermap is its nfev.  ``mp_pool`` is
        accepted for API compatibility.
        """
        self._print('Apply rofl')
        if self._tilts is None:
            raise ValueError('tilts not set')
        tilts = np.asarray(self._tilts)
        if (tilts.shape != (5, 2) or np.any(tilts[:, 1] != np.deg2rad([0, 0, 90, 180, 270])) or
                tilts[0, 0] != 0 or np.any(tilts[1:, 0] != tilts[1, 0])):
            raise ValueError('tilts not suitable for ROFL')
        tilt_angle = float(tilts[1, 0])
        torch = _torch()
        dev = torch.device('cuda', torch.cuda.current_device())
        as_tensor = _is_tensor(data)
        if as_tensor:
            x = data.to(device=dev)
            itemsize = x.element_size()
        else:
            arr = np.asarray(data)
            itemsize = arr.dtype.itemsize
            x = torch.from_numpy(np.ascontiguousarray(arr)).to(dev)
        if x.dim() != 4:
            raise TypeError('data: np.array([tilts,x,y,stack])')
        if x.shape[0] != 5:
            raise ValueError('data need 1 + 4 measurements')
        if x.shape[-1] <= 3:
            raise ValueError('data needs at least 3 equidistant rotations')
        x = x.to(torch.float64).contiguous()
        nx, ny = int(x.shape[1]), int(x.shape[2])
        m = None
        if mask is not None:
            m = torch.as_tensor(np.asarray(mask.cpu() if _is_tensor(mask) else mask, dtype=bool)
                                .astype(np.uint8), device=dev).contiguous()
            if tuple(m.shape) != (nx, ny):
                raise ValueError('mask shape does not match the images')
        out = torch.empty((8, nx, ny), dtype=torch.float64, device=dev)
        _lib.rofl(x, tilt_angle, 3.0, 0.0, m, out, grad_mode)
        if itemsize <= 4:
            out = out.to(torch.float32)
        if not as_tensor:
            out = out.cpu().numpy()
        return out[0], out[1], out[2], (out[3], out[4], out[5], out[6], out[7])

    # ------------------------------------------------------------ helpers
    def crop_tilt_pixel(self):
        """number of boundary pixels affected by tilting (simpli.py:246-262)"""
        if not self.untilt_sensor_view:
            raise ValueError('currently only for untilt_sensor_view=True')
        if self._voxel_size is None:
            raise ValueError('voxel_size is not set')
        if self._pixel_size is None:
            raise ValueError('pixel_size is not set')
        if self._tilts is None:
            raise ValueError('tilts are not set')
        return int(
            np.ceil(
                np.tan(np.max(np.abs(self._tilts[:, 0]))) * self._dim[-1] / 2 *
                self._voxel_size / self._pixel_size))

    def crop_tilt_voxel(self):
        """number of boundary voxels affected by tilting"""
        return int(np.round(self.crop_tilt_pixel() * self._pixel_size / self._voxel_size))

    def add_crop_tilt_halo(self):
        """enlarge the volume by the tilt halo"""
        self.dim_origin[:2] -= self.crop_tilt_voxel() * self.voxel_size
        self.dim[:2] += 2 * self.crop_tilt_voxel()

    def rm_crop_tilt_halo(self, data):
        """remove the tilt halo from images"""
        delta_voxel = self.crop_tilt_voxel()
        if delta_voxel == 0:
            return data
        return data[delta_voxel:-delta_voxel, delta_voxel:-delta_voxel, :]

    def memory_usage(self, unit='MB', item='all'):
        """expected memory usage (simpli.py:537-571)"""
        if unit == 'MB':
            div = 1024**2
        elif unit == 'GB':
            div = 1024**3
        else:
            raise ValueError('allowed is only \'MB\', \'GB\'')
        if self._dim is None:
            raise TypeError('dimension not set yet')
        if item == 'tissue':
            usage = np.prod(self._dim) * (32 + 32) / 8 / div
        elif item == 'all':
            usage = np.prod(self._dim) * (32 + 32 + 3 * 32) / 8 / div
        else:
            raise ValueError('allowed is only \'tissue\' or \'all\'')
        return usage
