"""Write tests/golden/headline.npz and headline.json: fit-parity fixtures on
the headline workload (7-parameter fits of 201 x 201 crops, sigma = 0.05).

Inputs
  * 32 features drawn exactly as the first 32 of the benchmark's headline
    batch (seed 3: a_p U[0.4, 3.0] um, n_p U[1.38, 1.65], k_p U[0, 0.01],
    z_p U[50, 400] px, x_p, y_p at the crop centre +-5 px, alpha
    U[0.8, 1.2]; starts +-2 px, x(1 +- 0.05) a_p, +-0.02 n_p);
  * 25 hard small crops of the uneven-crop workload (the first features of
    its seed-3 draw with side <= 80 px and a_p > 2.4 um: holograms larger
    than their crops, multimodal chi2), sigma as drawn.
  The data are the FP64 oracle's holograms plus numpy Gaussian noise
  (seeded), stored as float32.

Fits (all from the same starts, tolerances 1e-12)
  * oracle 7:   oracle.fit with all 7 parameters (x_p y_p z_p a_p n_p k_p
                alpha) -- the reference's MINPACK fitter restated in FP64
                (analysis/Optimizer.py:196-278), with alpha free;
  * oracle 6:   oracle.fit of x_p .. k_p with alpha held at 1 (the
                reference's variables and model);
  * reference:  the reference's own analysis/Optimizer.py optimize() with
                the same 6 variables and tolerances (imported from
                /root/reference as tests/golden/make_golden.py does; nothing
                else pins its end-to-end result on this workload).  Only for
                the 32 headline features.
  * exact J (--exact-jacobian, added to an existing headline.npz): the same
                MINPACK (scipy least_squares, method 'lm', x_scale 'jac')
                with the oracle's 5-point central-difference Jacobian
                (oracle.lorenzmie.jacobian, errors ~1e-10 of scale) in place
                of lmdif's forward differences, 7 parameters, for the
                headline features and the small crops: the minimum MINPACK
                reaches when its Jacobian is exact, as the device's
                analytic Jacobian is.

Runs ONLY in the development container (the reference is read there); the
fixtures are data and travel with the repository.  About 3 minutes on 8
CPUs.

Usage:  python tests/golden/make_golden_headline.py [--exact-jacobian]
"""
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

sys.dont_write_bytecode = True
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

INSTRUMENT = dict(wavelength=0.447, magnification=0.048, n_m=1.340)
PARAMS = ('x_p', 'y_p', 'z_p', 'a_p', 'n_p', 'k_p', 'alpha')
TIGHT = dict(ftol=1e-12, xtol=1e-12, gtol=1e-12, max_nfev=4000)
SIDE = 201
SIGMA = 0.05
NHEAD = 32
NSMALL = 25


def headline_draw(n, side=SIDE, seed=3):
    """The benchmark's headline draw, feature by feature in its order."""
    rng = np.random.default_rng(seed)
    truth = np.zeros((n, 7))
    start = np.zeros((n, 7))
    c = (side - 1) / 2
    for i in range(n):
        a = rng.uniform(0.4, 3.0)
        npart = rng.uniform(1.38, 1.65)
        k = rng.uniform(0., 0.01)
        z = rng.uniform(50., 400.)
        x = c + rng.uniform(-5, 5)
        y = c + rng.uniform(-5, 5)
        al = rng.uniform(0.8, 1.2)
        truth[i] = (x, y, z, a, npart, k, al)
        sg = rng.choice([-1., 1.], size=5)
        start[i] = (x + 2 * sg[0], y + 2 * sg[1], z + 2 * sg[2],
                    a * (1 + 0.05 * sg[3]), npart + 0.02 * sg[4], k, al)
    return truth, start


def small_draw(nkeep, seed=3, n=4096):
    """The uneven-crop draw (side log-U[48, 512], sigma in {0.02, 0.05,
    0.1}), keeping the first features with side <= 80 and a_p > 2.4."""
    rng = np.random.default_rng(seed)
    keep = []
    for i in range(n):
        s = int(round(np.exp(rng.uniform(np.log(48.), np.log(512.)))))
        c = (s - 1) / 2
        t = np.array((c + rng.uniform(-5, 5), c + rng.uniform(-5, 5),
                      rng.uniform(50., 400.), rng.uniform(0.4, 3.0),
                      rng.uniform(1.38, 1.65), rng.uniform(0., 0.01),
                      rng.uniform(0.8, 1.2)))
        sigma = float(rng.choice([0.02, 0.05, 0.1]))
        sg = rng.choice([-1., 1.], size=5)
        st = t + (2 * sg[0], 2 * sg[1], 2 * sg[2], 0, 0.02 * sg[4], 0, 0)
        st[3] = t[3] * (1 + 0.05 * sg[3])
        if s <= 80 and t[3] > 2.4:
            keep.append((i, s, t, st, sigma))
            if len(keep) == nkeep:
                break
    return keep


def _oracle_fit(job):
    from oracle import fit, lorenzmie as lm
    data, side, start, sigma, nvar = job
    c = lm.meshgrid((side, side))
    t0 = time.perf_counter()
    r = fit.fit(data.ravel(), c, dict(zip(PARAMS, start)),
                variables=PARAMS[:nvar], instrument=INSTRUMENT, noise=sigma,
                settings=TIGHT)
    r['seconds'] = time.perf_counter() - t0
    r['chi2'] = r['redchi'] * (r['npix'] - nvar)
    return r


_REF = []


def _reference_fit(job):
    from make_golden import _load_reference, make_model
    if not _REF:
        _REF.append(_load_reference())
    LorenzMie, Sphere, Instrument, Hologram, Optimizer = _REF[0]
    data, side, start, sigma = job
    model = make_model(LorenzMie, Sphere, Instrument,
                       LorenzMie.meshgrid((side, side)), start[:6])
    model.instrument.noise = sigma
    opt = Optimizer(model=model)
    opt.variables = list(PARAMS[:6])
    opt.settings.update(TIGHT)
    t0 = time.perf_counter()
    res = opt.optimize(Hologram(data.reshape(side, side).astype(float)))
    r = {k: (bool(v) if k == 'success' else float(v)) for k, v in res.items()}
    r['nfev'] = int(opt._result.nfev)
    r['seconds'] = time.perf_counter() - t0
    r['chi2'] = r['redchi'] * (r['npix'] - 6)
    return r


def _exact_j_fit(job):
    from scipy.optimize import least_squares
    from oracle import fit, lorenzmie as lm
    data, side, start, sigma = job
    c = lm.meshgrid((side, side))
    data = data.ravel().astype(float)

    def fun(x):
        return fit.residuals(x, PARAMS, {}, data, c, INSTRUMENT, sigma)

    def jac(x):
        return lm.jacobian(c, dict(zip(PARAMS, x)), INSTRUMENT, order=4).T / sigma
    t0 = time.perf_counter()
    r = least_squares(fun, np.asarray(start, float), jac=jac, method='lm',
                      x_scale='jac', ftol=TIGHT['ftol'], xtol=TIGHT['xtol'],
                      gtol=TIGHT['gtol'], max_nfev=TIGHT['max_nfev'])
    redchi, unc = fit.statistics(r.jac, r.cost, data.size, len(PARAMS))
    return dict(values=r.x, uncertainties=unc, chi2=2. * r.cost,
                nfev=r.nfev, success=bool(r.success),
                seconds=time.perf_counter() - t0)


def exact_jacobian():
    path = os.path.join(HERE, 'headline.npz')
    out = dict(np.load(path))
    with open(os.path.join(HERE, 'headline.json')) as f:
        meta = json.load(f)
    jobs = [(out['head_data'][i], SIDE, out['head_start'][i], SIGMA)
            for i in range(len(out['head_start']))]
    off = out['small_offsets']
    jobs_s = [(out['small_data'][off[k]:off[k + 1]], int(out['small_side'][k]),
               out['small_start'][k], float(out['small_sigma'][k]))
              for k in range(len(out['small_side']))]
    os.environ.update(OMP_NUM_THREADS='1', MKL_NUM_THREADS='1',
                      OPENBLAS_NUM_THREADS='1')
    t0 = time.perf_counter()
    with mp.get_context('spawn').Pool(7) as pool:
        ah = pool.map_async(_exact_j_fit, jobs)
        asm = pool.map_async(_exact_j_fit, jobs_s)
        fits = dict(oracle7j=ah.get(), small7j=asm.get())
    print(f'exact-Jacobian fits done in {time.perf_counter() - t0:.0f} s')
    for key, rows in fits.items():
        for f in ('values', 'uncertainties', 'chi2', 'nfev', 'success'):
            out[f'{key}_{f}'] = np.array([r[f] for r in rows])
        meta[f'{key}_seconds'] = [r['seconds'] for r in rows]
    np.savez_compressed(path, **out)
    with open(os.path.join(HERE, 'headline.json'), 'w') as f:
        json.dump(meta, f, indent=1)


def main():
    from oracle import lorenzmie as lm
    out, meta = {}, dict(instrument=INSTRUMENT, tolerances=TIGHT,
                         params=list(PARAMS))
    truth, start = headline_draw(NHEAD)
    c = lm.meshgrid((SIDE, SIDE))
    rng = np.random.default_rng(1003)
    data = np.zeros((NHEAD, SIDE, SIDE), np.float32)
    for i in range(NHEAD):
        h = lm.hologram(c, dict(zip(PARAMS[:6], truth[i, :6])),
                        alpha=truth[i, 6], **INSTRUMENT)
        data[i] = (h + SIGMA * rng.standard_normal(h.size)).reshape(SIDE, SIDE)
    out.update(head_truth=truth, head_start=start, head_data=data,
               head_sigma=np.array(SIGMA))

    small = small_draw(NSMALL)
    sides = np.array([s for _, s, _, _, _ in small])
    s_truth = np.array([t for _, _, t, _, _ in small])
    s_start = np.array([st for _, _, _, st, _ in small])
    s_sigma = np.array([sg for _, _, _, _, sg in small])
    rng = np.random.default_rng(1004)
    s_data = []
    for (_, s, t, _, sg) in small:
        h = lm.hologram(lm.meshgrid((s, s)), dict(zip(PARAMS[:6], t[:6])),
                        alpha=t[6], **INSTRUMENT)
        s_data.append((h + sg * rng.standard_normal(h.size)).astype(np.float32))
    out.update(small_index=np.array([i for i, *_ in small]), small_side=sides,
               small_truth=s_truth, small_start=s_start, small_sigma=s_sigma,
               small_offsets=np.concatenate([[0], np.cumsum(sides ** 2)]),
               small_data=np.concatenate(s_data))

    jobs7 = [(data[i], SIDE, start[i], SIGMA, 7) for i in range(NHEAD)]
    start1 = start.copy()
    start1[:, 6] = 1.                 # alpha held at 1, as the reference's
    jobs6 = [(data[i], SIDE, start1[i], SIGMA, 6) for i in range(NHEAD)]
    jobs_s = [(s_data[k], int(sides[k]), s_start[k], float(s_sigma[k]), 7)
              for k in range(len(small))]
    jobs_r = [(data[i], SIDE, start[i], SIGMA) for i in range(NHEAD)]
    env = dict(OMP_NUM_THREADS='1', MKL_NUM_THREADS='1',
               OPENBLAS_NUM_THREADS='1')
    os.environ.update(env)
    t0 = time.perf_counter()
    with mp.get_context('spawn').Pool(7) as pool:
        a7 = pool.map_async(_oracle_fit, jobs7)
        a6 = pool.map_async(_oracle_fit, jobs6)
        asm = pool.map_async(_oracle_fit, jobs_s)
        ar = pool.map_async(_reference_fit, jobs_r)
        fits = dict(oracle7=a7.get(), oracle6=a6.get(), small7=asm.get(),
                    reference6=ar.get())
    print(f'fits done in {time.perf_counter() - t0:.0f} s', flush=True)
    for key, rows in fits.items():
        nvar = 6 if key.endswith('6') else 7
        names = PARAMS[:nvar]
        out[f'{key}_values'] = np.array([[r[v] for v in names] for r in rows])
        out[f'{key}_uncertainties'] = np.array([[r['d' + v] for v in names]
                                                for r in rows])
        out[f'{key}_chi2'] = np.array([r['chi2'] for r in rows])
        out[f'{key}_nfev'] = np.array([r['nfev'] for r in rows])
        out[f'{key}_success'] = np.array([r['success'] for r in rows])
        meta[f'{key}_seconds'] = [r['seconds'] for r in rows]
    d = np.abs(out['oracle6_values'] - out['reference6_values'])
    rel = d / np.abs(out['reference6_values'])
    rel[:, 5] = d[:, 5]
    meta['oracle6_vs_reference6_max'] = float(rel.max())
    print('oracle (6 parameters) against the reference: max difference '
          f'{rel.max():.2e}')
    np.savez_compressed(os.path.join(HERE, 'headline.npz'), **out)
    with open(os.path.join(HERE, 'headline.json'), 'w') as f:
        json.dump(meta, f, indent=1)


if __name__ == '__main__':
    if '--exact-jacobian' in sys.argv[1:]:
        exact_jacobian()
    else:
        main()
