"""Forward model: instrument, particles and the GPU Lorenz-Mie calculator.

Mirrors pylorenzmie.theory (theory/__init__.py:1-44).  There is one
backend, so ``LorenzMie`` is the HIP calculator itself; ``hipLorenzMie``
and ``LMHologram`` are aliases.
"""
from .Instrument import Instrument
from .Particle import Particle
from .Sphere import Sphere
from .LayeredSphere import LayeredSphere
from .LorenzMie import LorenzMie, PARAMS
from .Cluster import Cluster
from .Dimer import Dimer
from .Aberrated import Aberrated, AberratedLorenzMie
from .LorenzMieBatch import LorenzMieBatch

hipLorenzMie = LorenzMie
LMHologram = LorenzMie
hipLorenzMieBatch = LorenzMieBatch     # theory/torchLorenzMieBatch.py:218

__all__ = ['Instrument', 'Particle', 'Sphere', 'LayeredSphere', 'Cluster',
           'Dimer', 'Aberrated', 'AberratedLorenzMie', 'LorenzMie', 'hipLorenzMie',
           'LMHologram', 'LorenzMieBatch', 'hipLorenzMieBatch', 'PARAMS']
