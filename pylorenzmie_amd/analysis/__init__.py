"""Fitting: holograms, masks and the GPU fitters (pylorenzmie.analysis)."""
from .Hologram import Hologram
from .Mask import Mask
from .Optimizer import Optimizer, hipOptimizer
from .BatchOptimizer import BatchOptimizer
from .lm import LMBatch, LMSettings

__all__ = ['Hologram', 'Mask', 'Optimizer', 'hipOptimizer', 'BatchOptimizer',
           'LMBatch', 'LMSettings']
