from .fibers import Fiber, FiberBundle, FiberBundles
from .layers import Layer, Layers

__all__ = ['Fiber', 'FiberBundle', 'FiberBundles', 'Layer', 'Layers']
