"""Write golden input/output vectors from the reference numpy path.

Runs ONLY in the development container where the reference checkout is
mounted at /root/reference; the fixtures it writes (tests/golden/*.npz,
*.json) are data and travel with the repository, the reference does not.

What is imported from the reference and why (nothing else pins it):
  * theory/LorenzMie.py field()/hologram() -- the sign, bohren/cartesian,
    on-axis and phase conventions of the full forward model;
  * theory/Sphere.py mie_coefficients -- the effect of starting D1 at
    n_max (theory/Sphere.py:190) on the top orders;
  * analysis/Optimizer.py optimize() -- the fitted values and SVD
    uncertainties at the reference's default MINPACK settings.

The reference's analysis/__init__.py imports trackpy (absent here) through
Localizer; the analysis modules used here do not need it, so the package
object is created without running that __init__.  No bytecode is written.

Usage:  python tests/golden/make_golden.py
"""
import importlib.util
import json
import os
import platform
import subprocess
import sys
import time
import types

import numpy as np

sys.dont_write_bytecode = True
REF = '/root/reference'
HERE = os.path.dirname(os.path.abspath(__file__))


def _load_reference():
    spec = importlib.util.spec_from_file_location(
        'pylorenzmie', os.path.join(REF, '__init__.py'),
        submodule_search_locations=[REF])
    pkg = importlib.util.module_from_spec(spec)
    sys.modules['pylorenzmie'] = pkg
    spec.loader.exec_module(pkg)
    # bare package object for analysis/ (skips its trackpy-importing __init__)
    ana = types.ModuleType('pylorenzmie.analysis')
    ana.__path__ = [os.path.join(REF, 'analysis')]
    sys.modules['pylorenzmie.analysis'] = ana
    from pylorenzmie.theory.LorenzMie import LorenzMie
    from pylorenzmie.theory.Sphere import Sphere
    from pylorenzmie.theory.Instrument import Instrument
    from pylorenzmie.analysis.Hologram import Hologram
    from pylorenzmie.analysis.Optimizer import Optimizer
    return LorenzMie, Sphere, Instrument, Hologram, Optimizer


INSTRUMENT = dict(wavelength=0.447, magnification=0.048, n_m=1.340)

AB_CASES = [  # (a_p, n_p, k_p)
    (0.75, 1.45, 0.0), (1.0, 1.5, 0.0), (1.0, 1.4, 0.0), (0.4, 1.38, 0.0),
    (3.0, 1.65, 0.01), (1.5, 1.59, 0.001), (0.2, 1.33, 0.0),
    (2.0, 2.0, 0.0), (5.0, 1.45, 0.05), (0.05, 1.5, 0.0),
    (10.0, 1.33, 0.0), (10.0, 2.0, 0.0), (1.1, 1.4, 0.0),
]


def make_model(LorenzMie, Sphere, Instrument, coords, p):
    inst = Instrument(**INSTRUMENT)
    s = Sphere(x_p=p[0], y_p=p[1], z_p=p[2], a_p=p[3], n_p=p[4], k_p=p[5])
    return LorenzMie(coordinates=coords, particle=s, instrument=inst)


def main():
    LorenzMie, Sphere, Instrument, Hologram, Optimizer = _load_reference()
    out = {}

    # --- Mie coefficients -------------------------------------------------
    abs_, offs = [], [0]
    for a_p, n_p, k_p in AB_CASES:
        ab = Sphere(a_p=a_p, n_p=n_p, k_p=k_p).ab(INSTRUMENT['n_m'],
                                                   INSTRUMENT['wavelength'])
        abs_.append(ab)
        offs.append(offs[-1] + ab.shape[0])
    out['ab_cases'] = np.array(AB_CASES)
    out['ab_offsets'] = np.array(offs)
    out['ab_values'] = np.concatenate(abs_)

    # --- fields on 64x64 for every flag combination --------------------------
    c64 = LorenzMie.meshgrid((64, 64))
    p64 = (32.3, 31.7, 150.0, 0.75, 1.45, 0.0)
    m = make_model(LorenzMie, Sphere, Instrument, c64, p64)
    out['field64_params'] = np.array(p64)
    for cart in (True, False):
        for boh in (True, False):
            key = f'field64_c{int(cart)}_b{int(boh)}'
            out[key] = m.field(cartesian=cart, bohren=boh).copy()

    # --- config-0 hologram, 201x201, particle on a pixel (on-axis pixel) -----
    c201 = LorenzMie.meshgrid((201, 201))
    p0 = (100.0, 100.0, 150.0, 0.75, 1.45, 0.0)
    m = make_model(LorenzMie, Sphere, Instrument, c201, p0)
    out['cfg0_params'] = np.array(p0)
    out['cfg0_holo'] = m.hologram()

    # --- assorted holograms on explicit coordinate lists ---------------------
    rng = np.random.default_rng(1)
    cases, coords_all, holos, offs = [], [], [], [0]
    for i in range(8):
        s = int(rng.integers(64, 401))
        a_p = rng.uniform(0.4, 3.0)
        n_p = rng.uniform(1.38, 1.65)
        k_p = rng.uniform(0., 0.01)
        z_p = rng.uniform(50., 400.)
        x_p = (s - 1) / 2 + rng.uniform(-5, 5)
        y_p = (s - 1) / 2 + rng.uniform(-5, 5)
        corner = (float(rng.integers(0, 900)), float(rng.integers(0, 700)))
        c = LorenzMie.meshgrid((s, s), corner=corner)[:, ::7]
        p = (x_p + corner[0], y_p + corner[1], z_p, a_p, n_p, k_p)
        m = make_model(LorenzMie, Sphere, Instrument, c, p)
        cases.append(p)
        coords_all.append(c)
        holos.append(m.hologram())
        offs.append(offs[-1] + c.shape[1])
    # particle below the focal plane, pixels off z = 0, and a tiny sphere
    extra = [((20.4, 19.6, -60.0, 0.9, 1.5, 0.0), 0.0),
             ((20.0, 20.0, 80.0, 0.6, 1.59, 0.002), 12.5),
             ((19.5, 20.5, 40.0, 0.05, 1.5, 0.0), 0.0)]
    for p, zc in extra:
        c = LorenzMie.meshgrid((41, 41))
        c = np.vstack([c, np.full((1, c.shape[1]), zc)])
        m = make_model(LorenzMie, Sphere, Instrument, c, p)
        cases.append(p)
        coords_all.append(c[:2] if zc == 0. else c)
        holos.append(m.hologram())
        offs.append(offs[-1] + c.shape[1])
    out['holo_cases'] = np.array(cases)
    out['holo_offsets'] = np.array(offs)
    out['holo_coords'] = np.concatenate(
        [np.vstack([c, np.zeros((3 - c.shape[0], c.shape[1]))])
         for c in coords_all], axis=1)
    out['holo_values'] = np.concatenate(holos)

    # --- two-sphere superposition --------------------------------------------
    c48 = LorenzMie.meshgrid((48, 48))
    inst = Instrument(**INSTRUMENT)
    pa = Sphere(x_p=15.2, y_p=20.0, z_p=120., a_p=0.6, n_p=1.45)
    pb = Sphere(x_p=33.0, y_p=27.5, z_p=90., a_p=0.9, n_p=1.59, k_p=0.001)
    m = LorenzMie(coordinates=c48, particle=[pa, pb], instrument=inst)
    out['pair_params'] = np.array([[15.2, 20.0, 120., 0.6, 1.45, 0.],
                                   [33.0, 27.5, 90., 0.9, 1.59, 0.001]])
    out['pair_holo'] = m.hologram()

    np.savez_compressed(os.path.join(HERE, 'forward.npz'), **out)

    # --- reference fits -------------------------------------------------------
    fits = {}
    t_fit = {}
    variables = ['x_p', 'y_p', 'z_p', 'a_p', 'n_p', 'k_p']
    tight = {'ftol': 1e-10, 'xtol': 1e-10, 'gtol': 1e-10}
    fit_inputs = []
    # config 0: 201x201, truth p0, sigma 0.05, seed 0, perturbed start
    rng = np.random.default_rng(0)
    m = make_model(LorenzMie, Sphere, Instrument, c201, p0)
    data = (m.hologram() + 0.05 * rng.standard_normal(c201.shape[1]))
    data = data.astype(np.float32).reshape(201, 201)
    start = (102.0, 98.0, 152.0, 0.7875, 1.47, 0.0)
    fit_inputs.append(('cfg0', data, p0, start, 0.05))
    # config-3-like crops, side log-uniform [48, 160]
    rng = np.random.default_rng(3)
    for i in range(12):
        s = int(round(np.exp(rng.uniform(np.log(48), np.log(160)))))
        a_p = rng.uniform(0.4, 3.0)
        n_p = rng.uniform(1.38, 1.65)
        k_p = rng.uniform(0., 0.01)
        z_p = rng.uniform(50., 400.)
        x_p = (s - 1) / 2 + rng.uniform(-5, 5)
        y_p = (s - 1) / 2 + rng.uniform(-5, 5)
        sigma = float(rng.choice([0.02, 0.05, 0.1]))
        truth = (x_p, y_p, z_p, a_p, n_p, k_p)
        sg = rng.choice([-1., 1.], size=5)
        start = (x_p + 2 * sg[0], y_p + 2 * sg[1], z_p + 2 * sg[2],
                 a_p * (1 + 0.05 * sg[3]), n_p + 0.02 * sg[4], k_p)
        cs = LorenzMie.meshgrid((s, s))
        m = make_model(LorenzMie, Sphere, Instrument, cs, truth)
        data = m.hologram() + sigma * rng.standard_normal(cs.shape[1])
        fit_inputs.append((f'crop{i:02d}', data.astype(np.float32).reshape(s, s),
                           truth, start, sigma))
    fixt = {}
    for name, data, truth, start, sigma in fit_inputs:
        fixt[f'{name}_data'] = data
        fixt[f'{name}_truth'] = np.array(truth)
        fixt[f'{name}_start'] = np.array(start)
        fixt[f'{name}_sigma'] = np.array(sigma)
        for mode, extra_settings in (('default', {}), ('tight', tight)):
            ny, nx = data.shape
            model = make_model(LorenzMie, Sphere, Instrument,
                               LorenzMie.meshgrid((ny, nx)), start)
            model.instrument.noise = sigma
            opt = Optimizer(model=model)
            opt.variables = variables
            opt.settings.update(extra_settings)
            t0 = time.perf_counter()
            res = opt.optimize(Hologram(data.astype(float)))
            t_fit[f'{name}_{mode}'] = time.perf_counter() - t0
            fits[f'{name}_{mode}'] = {k: (bool(v) if k == 'success' else
                                          float(v)) for k, v in res.items()}
            fits[f'{name}_{mode}']['nfev'] = int(opt._result.nfev)
            print(name, mode, fits[f'{name}_{mode}'], flush=True)
    np.savez_compressed(os.path.join(HERE, 'fits.npz'), **fixt)
    cpu = subprocess.run(['lscpu'], capture_output=True, text=True).stdout
    model_name = [ln.split(':', 1)[1].strip() for ln in cpu.splitlines()
                  if ln.startswith('Model name')]
    meta = dict(variables=variables, tight_settings=tight, fits=fits,
                wall_seconds=t_fit,
                host=dict(cpu=model_name[0] if model_name else '',
                          python=platform.python_version(),
                          numpy=np.__version__,
                          scipy=__import__('scipy').__version__,
                          omp_threads=os.environ.get('OMP_NUM_THREADS')))
    with open(os.path.join(HERE, 'fits.json'), 'w') as f:
        json.dump(meta, f, indent=1)


if __name__ == '__main__':
    main()
