"""MI355X-native Lorenz-Mie hologram model and Levenberg-Marquardt fitter.

Drop-in for pylorenzmie's ``LorenzMie(...).field()/hologram()`` and
``Optimizer.optimize()``; the arithmetic runs in the gfx950 kernels of
liblmz.so (include/lmz.h), reached through ctypes.
"""
__version__ = '0.1.0'

from . import _lmz

__all__ = ['theory', 'analysis', 'lib', '__version__']
