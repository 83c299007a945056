"""Golden vectors for superposed and aberrated holograms, from the
reference numpy path (theory/Aberrated.py, theory/Cluster.py,
theory/Dimer.py on theory/LorenzMie.py).

Runs ONLY in the development container with the reference checkout at
/root/reference; the fixture it writes (tests/golden/superposed.npz) is
data.  Imported because nothing else pins them: the spherical-aberration
phase mask (theory/Aberrated.py:61-84), the cluster origin propagation
(theory/Cluster.py:50-59) and the dimer geometry (theory/Dimer.py:64-78).

Usage:  python tests/golden/make_golden_superposed.py
"""
import importlib.util
import os
import sys

import numpy as np

sys.dont_write_bytecode = True
REF = '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))
INSTRUMENT = dict(wavelength=0.447, magnification=0.048, n_m=1.340)


def _load():
    spec = importlib.util.spec_from_file_location(
        'pylorenzmie', os.path.join(REF, '__init__.py'),
        submodule_search_locations=[REF])
    pkg = importlib.util.module_from_spec(spec)
    sys.modules['pylorenzmie'] = pkg
    spec.loader.exec_module(pkg)
    from pylorenzmie.theory.LorenzMie import LorenzMie
    from pylorenzmie.theory.Sphere import Sphere
    from pylorenzmie.theory.Instrument import Instrument
    from pylorenzmie.theory.Aberrated import Aberrated
    from pylorenzmie.theory.Cluster import Cluster
    from pylorenzmie.theory.Dimer import Dimer
    return LorenzMie, Sphere, Instrument, Aberrated, Cluster, Dimer


def main():
    LorenzMie, Sphere, Instrument, Aberrated, Cluster, Dimer = _load()
    out = {}
    c = LorenzMie.meshgrid((64, 64), corner=(10., 20.))
    out['coords'] = c
    inst = Instrument(**INSTRUMENT)
    # one aberrated sphere, two signs of the coefficient
    s = Sphere(x_p=41.3, y_p=52.6, z_p=120., a_p=0.9, n_p=1.5, k_p=0.002)
    AL = Aberrated(LorenzMie)
    for tag, sph in (('s09', 0.9), ('sm04', -0.4)):
        m = AL(coordinates=c, particle=s, instrument=inst, spherical=sph)
        out[f'aberrated_{tag}'] = m.hologram()
    out['aberrated_particle'] = np.array([41.3, 52.6, 120., 0.9, 1.5, 0.002])
    # a cluster of two spheres moved as one (r_0 propagation), aberrated
    p1 = Sphere(x_p=-6.0, y_p=2.0, z_p=5.0, a_p=0.6, n_p=1.45)
    p2 = Sphere(x_p=7.0, y_p=-3.0, z_p=-4.0, a_p=0.8, n_p=1.55, k_p=0.001)
    cl = Cluster(x_p=40.0, y_p=50.0, z_p=140.0, particles=[p1, p2])
    m = AL(coordinates=c, particle=cl, instrument=inst, spherical=0.5)
    out['cluster_holo'] = m.hologram()
    m = LorenzMie(coordinates=c, particle=cl, instrument=inst)
    out['cluster_holo_plain'] = m.hologram()
    out['cluster_centre'] = np.array([40.0, 50.0, 140.0])
    out['cluster_members'] = np.array([[-6.0, 2.0, 5.0, 0.6, 1.45, 0.0],
                                       [7.0, -3.0, -4.0, 0.8, 1.55, 0.001]])
    # a dimer
    d = Dimer(a_p=0.5, n_p=1.42, k_p=0.0, theta=0.6, phi=1.1,
              magnification=INSTRUMENT['magnification'])
    d.x_p, d.y_p, d.z_p = 42.0, 51.5, 130.0
    m = LorenzMie(coordinates=c, particle=d, instrument=inst)
    out['dimer_holo'] = m.hologram()
    out['dimer_params'] = np.array([42.0, 51.5, 130.0, 0.5, 1.42, 0.0, 0.6, 1.1])
    out['dimer_member_rp'] = np.array([p.r_p for p in d.particles])
    np.savez_compressed(os.path.join(HERE, 'superposed.npz'), **out)
    print('wrote', {k: v.shape for k, v in out.items()})


if __name__ == '__main__':
    main()
