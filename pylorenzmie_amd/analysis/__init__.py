"""Fitting: holograms, masks and the GPU fitters (pylorenzmie.analysis)."""
from .Hologram import Hologram
from .Mask import Mask
from .RadialMask import RadialMask
from .Optimizer import Optimizer, hipOptimizer
from .BatchOptimizer import BatchOptimizer
from .BaseEstimator import BaseEstimator
from .DEEstimator import DEEstimator
from .RadialEstimator import RadialEstimator
from .Feature import Feature
from .Frame import Frame
from .lm import LMBatch, LMSettings

__all__ = ['Hologram', 'Mask', 'RadialMask', 'Optimizer', 'hipOptimizer',
           'BatchOptimizer',
           'BaseEstimator', 'DEEstimator', 'RadialEstimator', 'Feature', 'Frame',
           'LMBatch', 'LMSettings']
