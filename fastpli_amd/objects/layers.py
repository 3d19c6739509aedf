"""Layers of a fiber bundle: (scale, birefringence dn, absorption mu [1/mm], model).

Contract kept from the reference (src/fastpli/objects/layers.py): a layer is
four values ``(scale, dn, mu, 'b'|'p'|'r')``; the messages and the order of the
checks below; a bundle's layers are held sorted by scale.

Design: a ``Layer`` is an immutable tuple subclass (nothing to copy or
freeze), and the checks are data, in the order the reference applies them.
"""

import warnings

from ._sealed import Sealed, edited

# What one layer must satisfy, in the reference's order (layers.py:9-18):
# (violated?, TypeError / ValueError raises, a Warning class warns, message)
LAYER_RULES = (
    (lambda ly: len(ly) != 4, TypeError, 'layer != (float, float, float, char)'),
    (lambda ly: ly[1] < 0 and ly[3] == 'r', UserWarning, 'birefringence negative and radial'),
    (lambda ly: ly[1] > 0 and ly[3] == 'p', UserWarning, 'birefringence positive and parallel'),
    (lambda ly: ly[1] != 0 and ly[3] == 'b', ValueError, 'birefringence != 0 for background.'),
)


def apply_rules(rules, value):
    """Raises or warns for every violated rule of `rules`, in order."""
    for violated, kind, message in rules:
        if violated(value):
            if issubclass(kind, Warning):
                warnings.warn(message, kind, stacklevel=3)
            else:
                raise kind(message)


class Layer(tuple):
    """One layer: ``Layer(scale, dn, mu, model)`` or ``Layer((scale, dn, mu, model))``."""

    __slots__ = ()

    def __new__(cls, *args):
        raw = args[0] if len(args) == 1 else args
        if isinstance(raw, Layer):
            return raw
        raw = tuple(raw)
        apply_rules(LAYER_RULES, raw)
        return tuple.__new__(cls, (float(raw[0]), float(raw[1]), float(raw[2]), raw[3]))

    def __setattr__(self, name, value):
        if not hasattr(self, name):
            raise TypeError(f'{self!r} is a frozen class')
        object.__setattr__(self, name, value)  # read-only properties raise here

    def copy(self):
        return self  # immutable

    def data(self):
        return tuple(self)

    scale = property(lambda self: self[0], doc='radius scale relative to the fiber radius')
    birefringence = property(lambda self: self[1], doc='birefringence dn')
    absorption = property(lambda self: self[2], doc='absorption coefficient mu [1/mm]')
    model = property(lambda self: self[3], doc="'b' background, 'p' parallel, 'r' radial")


def _ordered(layers):
    """`layers` sorted by scale (stable; warns when the order changes) with
    distinct scales (layers.py:91-99)."""
    order = sorted(range(len(layers)), key=lambda k: layers[k].scale)
    if order != list(range(len(layers))):
        warnings.warn('layers had to be sorted', stacklevel=3)
        layers = [layers[k] for k in order]
    scales = [ly.scale for ly in layers]
    if any(a == b for a, b in zip(scales, scales[1:])):
        raise ValueError('detected layers of same radii')
    return list(layers)


def as_layer_list(value):
    """A list of Layer from a layer, a sequence of layers or Layers (layers.py:76-99);
    None stays None."""
    if value is None:
        return None
    items = list(value)
    if all(isinstance(it, (list, tuple, Layer)) for it in items):
        layers = [Layer(it) for it in items]
    elif isinstance(value, (list, tuple, Layer)):
        layers = [Layer(value)]
    else:
        raise TypeError('Expected "Layer" or "[Layer]"')
    if any(ly.scale > 1 for ly in layers):
        warnings.warn('layer scale > 1 detected', stacklevel=2)
    return _ordered(layers)


class Layers(Sealed):
    """The layers of one fiber bundle, sorted by scale (None: not set)."""

    def __init__(self, data):
        self._layers = as_layer_list(data)
        self._seal()

    def __getitem__(self, item):
        return self._layers[item]

    def __setitem__(self, item, value):
        layers = list(self._layers)
        layers[item] = Layer(value)
        self._layers = _ordered(layers)
        edited()

    def __iter__(self):
        return iter(self._layers if self._layers is not None else ())

    def __len__(self):
        return len(self._layers) if self._layers is not None else 0

    def __str__(self):
        return str(self._layers)

    def __repr__(self):
        return repr(self._layers)

    def copy(self):
        new = Layers(None)
        object.__setattr__(new, '_layers', None if self._layers is None else list(self._layers))
        return new

    def is_none(self):
        return self._layers is None

    def data(self):
        return [ly.data() for ly in self]
