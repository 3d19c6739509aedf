"""Fiber models: a fiber is an (n, 4) float array of points and radii (x, y, z, r).

Contract kept from the reference (src/fastpli/objects/fibers.py): the classes
``Fiber`` / ``FiberBundle`` / ``FiberBundles``, their constructors' accepted
inputs and error messages (:19-35, :401-420, :709-737), indexing, ``layers``,
``as_array``, ``cast``, ``copy`` and the per-fiber transforms ``scale``,
``rotate``, ``translate``, ``apply``, ``apply_to_points``, ``apply_to_radii``,
which return new objects.  (The model-editing ``cut`` / ``cut_sphere`` are not
on the simulation path and are not provided.)

Design:
  * a fiber's array is read-only and replaced on every edit (copy on write):
    arrays handed out earlier are snapshots, and every change goes through a
    method, which bumps one module-wide edit counter (``edit_count``).  The
    simulation keys its device copy of a fiber model on it, so an unchanged
    model is flattened and uploaded once.  (Writing through ``fiber.points``
    as through a view raises; assign ``fiber.points = ...`` instead.)
  * the bundle-level transforms are the fiber-level ones mapped over the
    fibers (``_lifted``), not separate code.
  * ``FiberBundles.to_flat()`` packs a model into the contiguous arrays the
    native library takes.
"""

import numpy as np

from ._sealed import Sealed, edit_count, edited
from .layers import Layers

__all__ = ['Fiber', 'FiberBundle', 'FiberBundles', 'edit_count']


def _frozen_array(a):
    a.flags.writeable = False
    return a


def _fiber_array(data, dtype):
    """A read-only (n, 4) floating copy of `data` (fibers.py:19-35)."""
    if data is None:
        return _frozen_array(np.empty((0, 4), dtype))
    src = data._data if isinstance(data, Fiber) else data
    a = np.atleast_2d(np.array(src, dtype=dtype, copy=True))
    if a.ndim != 2 or a.shape[1] != 4:
        raise TypeError('Wrong shape: expected (n,4)')
    if not np.issubdtype(a.dtype, np.floating):
        raise TypeError('Wrong type: has to be floating')
    return _frozen_array(a)


class Fiber(Sealed):
    """Points and radii of one fiber, an (n, 4) array."""

    def __init__(self, data, dtype=float):
        self._data = _fiber_array(data, dtype)
        self._seal()

    # every in-place change: on a fresh copy, which then replaces the array
    def _edit(self, change):
        a = self._data.copy()
        change(a)
        self._data = _frozen_array(a)
        edited()

    def __getitem__(self, item):
        return self._data[item]

    def __setitem__(self, item, value):
        self._edit(lambda a: a.__setitem__(item, value))

    def __iter__(self):
        return iter(self._data)

    def __len__(self):
        return self._data.shape[-1]  # as the reference (fibers.py:75-76)

    def __str__(self):
        return str(self._data)

    def __repr__(self):
        return repr(self._data)

    def copy(self):
        return Fiber(self, self.dtype)

    @property
    def shape(self):
        return self._data.shape

    @property
    def dtype(self):
        return self._data.dtype

    def as_array(self):
        return np.array(self._data)

    @property
    def points(self):
        return self._data[:, :-1]

    @points.setter
    def points(self, value):
        self._edit(lambda a: a.__setitem__((slice(None), slice(0, 3)), value))

    @property
    def radii(self):
        return self._data[:, -1]

    @radii.setter
    def radii(self, value):
        self._edit(lambda a: a.__setitem__((slice(None), -1), value))

    # ---------------------------------------------------------- transforms
    def _mapped(self, cols, fun):
        a = np.array(self._data)
        a[cols] = fun(a[cols])
        return Fiber(a, a.dtype)

    def cast(self, dtype):
        return Fiber(self, dtype)

    def scale(self, scale, mode='all'):
        cols = {'all': np.s_[:], 'points': np.s_[:, :3], 'radii': np.s_[:, -1]}
        if mode not in cols:
            raise ValueError('mode = [all, points, radii]')
        return self._mapped(cols[mode], lambda v: v * scale)

    def rotate(self, rot, offset=None):
        rot = np.asarray(rot)
        o = np.zeros(3) if offset is None else np.asarray(offset)
        return self._mapped(np.s_[:, :3], lambda p: (p - o) @ rot.T + o)

    def translate(self, offset):
        return self._mapped(np.s_[:, :3], lambda p: p + np.asarray(offset))

    def apply(self, fun):
        return self._mapped(np.s_[:], fun)

    def apply_to_points(self, fun):
        return self._mapped(np.s_[:, :-1], fun)

    def apply_to_radii(self, fun):
        return self._mapped(np.s_[:, -1], fun)


_TRANSFORMS = ('cast', 'scale', 'rotate', 'translate', 'apply', 'apply_to_points',
               'apply_to_radii')


def _lifted(name, container):
    """The Fiber transform `name` mapped over a container's items."""

    def method(self, *args, **kwargs):
        return container._rebuilt(self, [getattr(it, name)(*args, **kwargs) for it in self])

    method.__name__ = name
    method.__doc__ = f'Fiber.{name} applied to every fiber; returns a new {container.__name__}.'
    return method


class _Items(Sealed):
    """A list of `_item` objects (every stored element converted to it)."""

    _item = None

    def __getitem__(self, item):
        return self._data[item]

    def __setitem__(self, item, value):
        self._data[item] = self._item(value)
        edited()

    def __delitem__(self, item):
        del self._data[item]
        edited()

    def __iter__(self):
        return iter(self._data)

    def __len__(self):
        return len(self._data)

    def __str__(self):
        return str(self._data)

    def __repr__(self):
        return repr(self._data)

    def append(self, value):
        self._data.append(self._item(value))
        edited()

    def extend(self, values):
        self._data.extend(self._item(v) for v in values)
        edited()

    @property
    def dtype(self):
        return self[0].dtype if len(self) > 1 else None  # as the reference

    def as_array(self):
        return [it.as_array() for it in self]


class FiberBundle(_Items):
    """Fibers sharing one layer description."""

    _item = Fiber

    def __init__(self, data=None, layers=None, dtype=None):
        if isinstance(data, FiberBundle) and layers is None:
            layers = data.layers  # a bundle keeps its layers unless given others
        if data is None:
            fibers = []
        elif isinstance(data, Fiber):
            fibers = [data]
        elif isinstance(data, FiberBundle):
            fibers = list(data)
        elif isinstance(data, (list, tuple)):
            fibers = [Fiber(f) if dtype is None else Fiber(f, dtype) for f in data]
        else:
            raise TypeError(f'data is not a list: {type(data)}')
        self._data = fibers
        self._layers = Layers(layers)
        self._seal()

    @staticmethod
    def _rebuilt(old, fibers):
        return FiberBundle(fibers, old.layers)

    def copy(self):
        return FiberBundle([f.copy() for f in self], self.layers.copy())

    @property
    def layers(self):
        return self._layers

    @layers.setter
    def layers(self, value):
        self._layers = Layers(value)
        edited()


class FiberBundles(_Items):
    """The fiber bundles of a tissue."""

    _item = FiberBundle

    def __init__(self, data=None, layers=None, dtype=None):
        if data is None:
            bundles = []
        elif isinstance(data, Fiber):
            bundles = [FiberBundle(data, layers)]
        elif isinstance(data, FiberBundle):
            bundles = [data]
        elif isinstance(data, FiberBundles):
            bundles = list(data)
        elif isinstance(data, (list, tuple)):
            if layers is not None and len(data) != len(layers):
                raise TypeError('[FiberBundle] and [Layers] differ in length')
            lys = [None] * len(data) if layers is None else layers
            bundles = [FiberBundle(b, ly, dtype) for b, ly in zip(data, lys)]
        else:
            raise TypeError('data is not a list')
        self._data = bundles
        self._seal()

    @staticmethod
    def _rebuilt(old, bundles):
        return FiberBundles(bundles)

    def copy(self):
        return FiberBundles([fb.copy() for fb in self])

    @property
    def layers(self):
        return [fb.layers for fb in self]

    @layers.setter
    def layers(self, value):
        if len(value) != len(self):
            raise ValueError('Wrong number of [layers]')
        for fb, lys in zip(self, value):
            fb.layers = lys

    def to_flat(self):
        """Pack all fibers into contiguous arrays.

        Returns ``(points, fiber_ptr, bundle_ptr, layer_ptr, layer_table)``:
        points (n_points, 4) float64 C-contiguous in bundle -> fiber -> point
        order; ``fiber_ptr`` (n_fibers + 1) int64 point offsets; ``bundle_ptr``
        (n_bundles + 1) int64 fiber offsets; ``layer_ptr`` (n_bundles + 1)
        int64 offsets into ``layer_table`` = (scale f64[], dn f64[], mu f64[],
        model uint8[]) with layers in their sorted order.
        """
        arrays = [f._data for fb in self for f in fb]
        counts = [a.shape[0] for a in arrays]
        fiber_ptr = np.concatenate([[0], np.cumsum(counts, dtype=np.int64)]).astype(np.int64)
        points = (np.concatenate(arrays, axis=0).astype(np.float64, copy=False)
                  if arrays else np.zeros((0, 4)))
        points = np.ascontiguousarray(points, dtype=np.float64)
        bundle_ptr = np.concatenate([[0], np.cumsum([len(fb) for fb in self],
                                                    dtype=np.int64)]).astype(np.int64)
        rows = [ly for fb in self for ly in fb.layers]
        layer_ptr = np.concatenate([[0], np.cumsum([len(fb.layers) for fb in self],
                                                   dtype=np.int64)]).astype(np.int64)
        table = (np.array([ly[0] for ly in rows], dtype=np.float64),
                 np.array([ly[1] for ly in rows], dtype=np.float64),
                 np.array([ly[2] for ly in rows], dtype=np.float64),
                 np.array([ord(ly[3]) for ly in rows], dtype=np.uint8))
        return points, fiber_ptr, bundle_ptr, layer_ptr, table


for _name in _TRANSFORMS:
    setattr(FiberBundle, _name, _lifted(_name, FiberBundle))
    setattr(FiberBundles, _name, _lifted(_name, FiberBundles))
