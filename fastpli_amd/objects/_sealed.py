"""The one base class of the parameter and fiber-model containers.

The reference's containers refuse attributes they were not built with
("<x> is a frozen class", e.g. src/fastpli/objects/layers.py:22-29,
src/fastpli/simulation/__simpli.py:25-33); user code relies on that to catch
misspelt property names.  Here every container derives from ``Sealed``:
attributes may be added while ``__init__`` runs, and ``_seal()`` ends that.
"""


class Sealed:
    """Refuses new attributes once ``_seal()`` has been called."""

    _sealed = False

    def _frozen_label(self):
        """How the TypeError names the instance (its repr by default)."""
        return repr(self)

    def __setattr__(self, name, value):
        if self._sealed and not hasattr(self, name):
            raise TypeError(f'{self._frozen_label()} is a frozen class')
        object.__setattr__(self, name, value)

    def _seal(self):
        object.__setattr__(self, '_sealed', True)


# One counter of edits for all fiber-model objects (fibers, bundles, layers):
# a cached copy of a model (flattened, on the device) stays valid while it
# does not change.
_EDITS = [0]


def edit_count():
    """Number of edits of any fiber model object so far."""
    return _EDITS[0]


def edited():
    _EDITS[0] += 1
