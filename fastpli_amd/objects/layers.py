"""Layer containers of a fiber bundle: (scale, birefringence dn, absorption mu, model).

Behaviour follows fastpli's ``src/fastpli/objects/layers.py:9-153``: a layer is
a 4-tuple ``(float, float, float, 'b'|'p'|'r')``; a list of layers is sorted by
scale (warning 'layers had to be sorted'), duplicate scales raise
``ValueError('detected layers of same radii')``, and a background layer ('b')
with non-zero birefringence raises ``ValueError``.
"""

import copy
import warnings

_MODELS = ('b', 'p', 'r')


def _check_layer(ly):
    """Validate one layer tuple (layers.py:9-18)."""
    if len(ly) != 4:
        raise TypeError('layer != (float, float, float, char)')
    if ly[1] < 0 and ly[-1] == 'r':
        warnings.warn('birefringence negative and radial')
    if ly[1] > 0 and ly[-1] == 'p':
        warnings.warn('birefringence positive and parallel')
    if ly[1] != 0 and ly[-1] == 'b':
        raise ValueError('birefringence != 0 for background.')
    return (float(ly[0]), float(ly[1]), float(ly[2]), ly[3])


class Layer:
    """One layer: scale relative to the fiber radius, dn, mu [1/mm], model."""

    __is_frozen = False

    def __setattr__(self, key, value):
        if self.__is_frozen and not hasattr(self, key):
            raise TypeError(f'{self!r} is a frozen class')
        object.__setattr__(self, key, value)

    def __freeze(self):
        self.__is_frozen = True

    def __init__(self, *args):
        if len(args) == 1:
            args = args[0]
        self._data = _check_layer(args)
        self.__freeze()

    def __getitem__(self, item):
        return self._data[item]

    def __str__(self):
        return str(self._data)

    def __repr__(self):
        return repr(self._data)

    def copy(self):
        return copy.deepcopy(self)

    def data(self):
        return self._data

    @property
    def scale(self):
        return self._data[0]

    @property
    def birefringence(self):
        return self._data[1]

    @property
    def absorption(self):
        return self._data[2]

    @property
    def model(self):
        return self._data[3]


def _sort_and_check(lys):
    keys = sorted(range(len(lys)), key=lambda i: lys[i][0])
    if any(keys[i] > keys[i + 1] for i in range(len(keys) - 1)):
        warnings.warn('layers had to be sorted')
        lys = [lys[i] for i in keys]
    if any(lys[i][0] == lys[i + 1][0] for i in range(len(lys) - 1)):
        raise ValueError('detected layers of same radii')
    return lys


def _check_layers(lys):
    """Validate and sort a layer list (layers.py:76-99)."""
    if lys is None:
        return None
    if isinstance(lys, Layers):
        lys = lys.data()
    if all(isinstance(ly, (list, tuple, Layer)) for ly in lys):
        lys = [Layer(ly) for ly in lys]
    elif isinstance(lys, (list, tuple, Layer)):
        lys = [Layer(lys)]
    else:
        raise TypeError('Expected "Layer" or "[Layer]"')
    if any(ly.scale > 1 for ly in lys):
        warnings.warn('layer scale > 1 detected')
    return _sort_and_check(lys)


class Layers:
    """Sorted list of :class:`Layer` of one fiber bundle."""

    __is_frozen = False

    def __setattr__(self, key, value):
        if self.__is_frozen and not hasattr(self, key):
            raise TypeError(f'{self!r} is a frozen class')
        object.__setattr__(self, key, value)

    def __freeze(self):
        self.__is_frozen = True

    def __init__(self, data):
        self._data = _check_layers(data)
        self.__freeze()

    def __getitem__(self, item):
        return self._data[item]

    def __setitem__(self, item, value):
        lys = list(self._data)
        lys[item] = Layer(value)
        self._data = _sort_and_check(lys)

    def __str__(self):
        return str(self._data)

    def __repr__(self):
        return repr(self._data)

    def copy(self):
        return copy.deepcopy(self)

    def __iter__(self):
        if self._data is None:
            return iter(())
        return iter(self._data)

    def __len__(self):
        return 0 if self._data is None else len(self._data)

    def is_none(self):
        return self._data is None

    def data(self):
        return [ly.data() for ly in self]
