"""Frame pre-processing on the GPU, mirroring utilities/__init__.py:
background normalization (Normalizer) and the running median of a video
(VMedian)."""
from .normalizer import Normalizer
from .vmedian import VMedian, vmedian

__all__ = ['Normalizer', 'VMedian', 'vmedian']
