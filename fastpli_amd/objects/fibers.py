"""Fiber containers: a fiber is an (n, 4) float array of points (x, y, z, r).

Conversion and validation follow fastpli's ``src/fastpli/objects/fibers.py``
(``_convert_to_fiber`` :19-35, ``FiberBundle`` :401-520, ``FiberBundles``
:709-830).  The geometry helpers of the reference (cut, rotate, ...) are
model-editing tools, not part of the simulation path, and are not provided
except ``translate`` / ``scale`` which the benchmark workloads use.

``FiberBundles.to_flat()`` is this package's own addition: it packs all
points into one contiguous (n_points, 4) float64 array with fiber and bundle
offsets, the layout the native library takes.
"""

import copy

import numpy as np

from .layers import Layers


def _convert_to_fiber(data, dtype):
    if data is None:
        return np.empty((0, 4), dtype)
    if isinstance(data, Fiber):
        return np.array(data._data, dtype=dtype, copy=True)
    data = np.atleast_2d(np.array(data, dtype=dtype, copy=True))
    if data.ndim != 2:
        raise TypeError('Wrong shape: expected (n,4)')
    if data.shape[1] != 4:
        raise TypeError('Wrong shape: expected (n,4)')
    if not np.issubdtype(data.dtype, np.floating):
        raise TypeError('Wrong type: has to be floating')
    return data


class Fiber:
    """(n, 4) array of points and radii."""

    __is_frozen = False

    def __setattr__(self, key, value):
        if self.__is_frozen and not hasattr(self, key):
            raise TypeError(f'{self!r} is a frozen class')
        object.__setattr__(self, key, value)

    def __freeze(self):
        self.__is_frozen = True

    def __init__(self, data, dtype=float):
        self._data = _convert_to_fiber(data, dtype)
        self.__freeze()

    def __getitem__(self, item):
        return self._data[item]

    def __setitem__(self, item, value):
        self._data[item] = value

    def __str__(self):
        return str(self._data)

    def __repr__(self):
        return repr(self._data)

    def __iter__(self):
        return iter(self._data)

    def __len__(self):
        return self._data.shape[-1]

    def copy(self):
        return copy.deepcopy(self)

    @property
    def shape(self):
        return self._data.shape

    @property
    def dtype(self):
        return self._data.dtype

    @property
    def points(self):
        return self._data[:, :-1]

    @property
    def radii(self):
        return self._data[:, -1]

    def as_array(self):
        return self._data.copy()

    def cast(self, dtype):
        return Fiber(self, dtype)

    def translate(self, offset):
        data = self._data.copy()
        data[:, :3] += np.asarray(offset, dtype=data.dtype)
        return Fiber(data)

    def scale(self, scale, mode='all'):
        data = self._data.copy()
        if mode in ('all', 'points'):
            data[:, :3] *= scale
        if mode in ('all', 'radii'):
            data[:, -1] *= scale
        return Fiber(data)


def _convert_to_fiber_bundle(data, layers, dtype):
    if data is None:
        return [], Layers(layers)
    if isinstance(data, Fiber):
        return [data], Layers(layers)
    if isinstance(data, FiberBundle):
        return data[:], Layers(layers)
    if not isinstance(data, (list, tuple)):
        raise TypeError(f'data is not a list: {type(data)}')
    return [Fiber(f) for f in data], Layers(layers)


class FiberBundle:
    """List of :class:`Fiber` sharing one :class:`Layers` description."""

    __is_frozen = False

    def __setattr__(self, key, value):
        if self.__is_frozen and not hasattr(self, key):
            raise TypeError(f'{self!r} is a frozen class')
        object.__setattr__(self, key, value)

    def __freeze(self):
        self.__is_frozen = True

    def __init__(self, data=None, layers=None, dtype=None):
        if isinstance(data, FiberBundle) and layers is None:
            layers = data.layers
        self._data, self._layers = _convert_to_fiber_bundle(data, layers, dtype)
        self.__freeze()

    def __getitem__(self, item):
        return self._data[item]

    def __setitem__(self, item, value):
        self._data[item] = Fiber(value)

    def __delitem__(self, item):
        del self._data[item]

    def __str__(self):
        return str(self._data)

    def __repr__(self):
        return repr(self._data)

    def __iter__(self):
        return iter(self._data)

    def __len__(self):
        return len(self._data)

    def copy(self):
        return copy.deepcopy(self)

    @property
    def dtype(self):
        return self._data[0].dtype if len(self) > 1 else None

    def as_array(self):
        return [f.as_array() for f in self]

    @property
    def layers(self):
        return self._layers

    @layers.setter
    def layers(self, value):
        self._layers = Layers(value)

    def append(self, fiber):
        self._data.append(Fiber(fiber))

    def extend(self, fibers):
        for f in fibers:
            self._data.append(Fiber(f))


def _convert_to_fiber_bundles(data, layers, dtype):
    if data is None:
        return []
    if isinstance(data, Fiber):
        return [FiberBundle(data, layers)]
    if isinstance(data, FiberBundle):
        return [data]
    if isinstance(data, FiberBundles):
        return data[:]
    if not isinstance(data, (list, tuple)):
        raise TypeError('data is not a list')
    if layers is not None:
        if len(data) != len(layers):
            raise TypeError('[FiberBundle] and [Layers] differ in length')
    else:
        layers = [None] * len(data)
    return [FiberBundle(fb, lys) for fb, lys in zip(data, layers)]


class FiberBundles:
    """List of :class:`FiberBundle`."""

    __is_frozen = False

    def __setattr__(self, key, value):
        if self.__is_frozen and not hasattr(self, key):
            raise TypeError(f'{self!r} is a frozen class')
        object.__setattr__(self, key, value)

    def __freeze(self):
        self.__is_frozen = True

    def __init__(self, data=None, layers=None, dtype=None):
        self._data = _convert_to_fiber_bundles(data, layers, dtype)
        self.__freeze()

    def __getitem__(self, item):
        return self._data[item]

    def __setitem__(self, item, value):
        self._data[item] = FiberBundle(value)

    def __delitem__(self, item):
        del self._data[item]

    def __str__(self):
        return str(self._data)

    def __repr__(self):
        return repr(self._data)

    def __iter__(self):
        return iter(self._data)

    def __len__(self):
        return len(self._data)

    def copy(self):
        return copy.deepcopy(self)

    @property
    def dtype(self):
        return self[0].dtype if len(self) > 1 else None

    def as_array(self):
        return [fb.as_array() for fb in self]

    @property
    def layers(self):
        return [fb.layers for fb in self]

    @layers.setter
    def layers(self, value):
        if len(value) != len(self):
            raise ValueError('Wrong number of [layers]')
        for fb, lys in zip(self, value):
            fb.layers = lys

    def append(self, fiber_bundle):
        self._data.append(FiberBundle(fiber_bundle))

    def extend(self, fiber_bundles):
        for fb in fiber_bundles:
            self._data.append(FiberBundle(fb))

    def to_flat(self):
        """Pack all fibers into contiguous arrays.

        Returns ``(points, fiber_ptr, bundle_ptr, layer_ptr, layer_table)``:
        points (n_points, 4) float64 C-contiguous in bundle -> fiber -> point
        order; ``fiber_ptr`` (n_fibers + 1) int64 point offsets; ``bundle_ptr``
        (n_bundles + 1) int64 fiber offsets; ``layer_ptr`` (n_bundles + 1)
        int64 offsets into ``layer_table`` = (scale f64[], dn f64[], mu f64[],
        model uint8[]) with layers in their sorted order.
        """
        fibers = [f._data for fb in self for f in fb]
        counts = np.array([f.shape[0] for f in fibers], dtype=np.int64)
        fiber_ptr = np.zeros(len(fibers) + 1, dtype=np.int64)
        np.cumsum(counts, out=fiber_ptr[1:])
        if fibers:
            points = np.ascontiguousarray(np.concatenate(fibers, axis=0),
                                          dtype=np.float64)
        else:
            points = np.zeros((0, 4), dtype=np.float64)
        bundle_ptr = np.zeros(len(self) + 1, dtype=np.int64)
        np.cumsum([len(fb) for fb in self], out=bundle_ptr[1:])
        scale, dn, mu, model = [], [], [], []
        layer_ptr = np.zeros(len(self) + 1, dtype=np.int64)
        for i, fb in enumerate(self):
            for ly in fb.layers:
                scale.append(ly[0])
                dn.append(ly[1])
                mu.append(ly[2])
                model.append(ord(ly[3]))
            layer_ptr[i + 1] = len(scale)
        table = (np.array(scale, dtype=np.float64), np.array(dn, dtype=np.float64),
                 np.array(mu, dtype=np.float64), np.array(model, dtype=np.uint8))
        return points, fiber_ptr, bundle_ptr, layer_ptr, table
