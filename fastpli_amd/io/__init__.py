from . import fiber_bundles

__all__ = ['fiber_bundles']
