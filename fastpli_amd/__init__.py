"""fastpli_amd -- MI355X-native simulation hot path of fastpli (3D-PLI).

Subpackages mirror the reference's module layout for the simulation path:
``fastpli_amd.simulation.Simpli``, ``fastpli_amd.objects`` and
``fastpli_amd.io``.  The compute runs in libfastpli_amd.so (hand-written HIP
kernels for gfx950) behind the C ABI of ``include/fastpli_amd.h``.
"""

from . import io, objects, simulation

__version__ = '0.1.0'

__all__ = ['io', 'objects', 'simulation', '__version__']
