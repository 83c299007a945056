"""Headline benchmark: 7-parameter sphere fits/s at 201x201 crops on one GPU.

Workload (one "step"): 4096 features, every crop 201x201, parameters drawn
as the config-3 distribution (a_p U[0.4, 3.0] um, n_p U[1.38, 1.65], k_p
U[0, 0.01], z_p U[50, 400] px, x/y at the crop centre +-5 px, alpha
U[0.8, 1.2]), Gaussian noise sigma = 0.05, starts perturbed by +-2 px,
x(1 +- 0.05) a_p, +-0.02 n_p.  One step fits all 4096 with the batched
device-resident LM (forward-mode FP32 Jacobian and FP64 trial forward pass,
both from radial tables of the multipole sums, FP64 solve, HIP graph per
iteration, <= 50 iterations), from the starting parameters already in HBM to
the results copied back to the host.

Also measured:
  fd      the plain GPU finite-difference form (hipOptimizer: one feature at
          a time, host MINPACK LM with forward differences, residuals copied
          to the host per evaluation) on a sample of features -> speedup;
  fd_batched  the same batched LM with a forward-difference Jacobian of the
          FP64 hologram (lmz_lm_settings.fd_step = 1e-5, the reference's
          diff_step) in place of the analytic one -> the adjoint's share of
          the speedup;
  roofline  the LM's trial forward pass (lmz_residual_tabulated: FP64
          radial tables of the multipole sums, k_fwd_ftab, and their
          12-point interpolation with the residual and chi2 per pixel,
          k_forward_ftab), timed with HIP events on the stream it runs on,
          against the FP64 vector peak; algorithmic FLOP: 64 N per table
          sample (N the order count) and 280 per pixel; and against the
          attainable FP64 FMA rate of this build, measured in the same run
          by the FMA-chain probe (tests/native/lmz_valu_peak.hip): the
          counted FP64 FLOP and VALU instructions per launch (from the PMC
          summary committed for these sources) over the probe's.  For
          comparison the order loop at every pixel (k_forward<2>, 64 N + 70
          FLOP per pixel);
  jacobian  the LM's Jacobian pass (lmz_jacobian_tabulated: FP32 radial
          tables and their 6-point interpolation, the J^T J partials;
          296 N per sample, 954 per pixel) against the FP32 peak, and the
          order loop at every pixel (k_jacobian, 296 N + 480 per pixel);
  cpu_baseline  the FP64 oracle (numpy model + scipy MINPACK LM) on the
          host's cores (one process per core, two features each, at most
          16 cores), started as a child process that never touches the
          GPU; cpu_baseline_1core: the same fits one after another on one
          core.
  fit_agreement  the timed fits of the first 32 features against the
          FP64 oracle's fits of the same data (tests/golden/headline.npz,
          whose data replace those 32 synthetic crops).

--extra superposed | frame | de | batch | normalize measures a path around
the fitter instead:
  the five-sphere hologram of the reference's MultiSphere example, a
  frame's features fitted as one batch from device memory, the DE
  estimator's GPU population, LorenzMieBatch.batch_hologram of 256
  particle lists, and the Normalizer's three methods on 1280 x 1024 frames
  of four pixel types (one JSON line each).

--config 1..4 measures the other configurations instead (one JSON line):
  1  batched forward holograms (1024 crops, side U{64..400}): holograms/s,
     FP64 TFLOP/s, error against the oracle at sampled pixels;
  2  forward + adjoint of sum (I - d)^2 (256 crops of 256^2): time, error
     of the gradient against FP64 central differences;
  3  batched LM on uneven crops (side log-U[48, 512]): fits/s, the
     Jacobian's FP32 fraction, the J^T J reduction's HBM rate;
  4  hard regime (512 crops of 512^2, layered and large spheres): hologram
     time and error, 7-parameter fits/s of both halves.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

INSTRUMENT = dict(wavelength=0.447, magnification=0.048, n_m=1.340)
PEAK_FP32 = 157.3        # TFLOP/s, MI355X FP32 vector (packed)
PEAK_FP32_UNPACKED = 78.6
PEAK_FP64 = 78.6         # TFLOP/s, MI355X FP64 vector
PEAK_HBM = 8000.0        # GB/s
PARAMS = ('x_p', 'y_p', 'z_p', 'a_p', 'n_p', 'k_p', 'alpha')


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def draw_features(n: int, side: int, seed: int):
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


def nmax_of(a_p, n_p, k_p):
    from pylorenzmie_amd._lmz import wiscombe_yang
    k = 2 * math.pi / INSTRUMENT['wavelength'] * INSTRUMENT['n_m']
    return wiscombe_yang(k * a_p, (n_p + 1j * k_p) / INSTRUMENT['n_m'])


def make_data(truth, side, sigma, seed):
    """Noisy synthetic holograms on the device (FP32 kernel + torch noise)."""
    import torch
    from pylorenzmie_amd.theory import Instrument
    from pylorenzmie_amd.theory.batch import HologramBatch
    from pylorenzmie_amd import _lmz
    inst = Instrument(**INSTRUMENT)
    caps = [_lmz.nmax_cap(p[3], p[4], p[5], inst.n_m, inst.wavelength, 0.)
            for p in truth]
    hb = HologramBatch([(side, side)] * len(truth), caps, instrument=inst)
    hb.load(truth)
    clean = hb.compute().clone()
    g = torch.Generator(device='cuda')
    g.manual_seed(seed)
    noise = torch.randn(clean.shape, generator=g, device='cuda',
                        dtype=torch.float32)
    return clean + sigma * noise


def run_batched(args, truth, start, data_dev):
    import torch
    from pylorenzmie_amd import _lmz
    from pylorenzmie_amd.theory import Instrument
    from pylorenzmie_amd.analysis.lm import LMBatch, LMSettings
    inst = Instrument(noise=args.sigma, **INSTRUMENT)
    side = args.side
    crops = [dict(data=np.zeros((side, side), np.float32), width=side,
                  origin=(0., 0., 0.),
                  nmax_cap=_lmz.nmax_cap(p[3], p[4], p[5], inst.n_m,
                                         inst.wavelength))
             for p in start]
    lm = LMBatch(crops, inst)
    lm.data.copy_(data_dev)
    settings = LMSettings(max_iter=50, ftol=args.ftol, xtol=args.xtol,
                          gtol=args.gtol)
    p0 = torch.zeros_like(lm.params)
    p0[:, :7] = torch.from_numpy(start).to(p0.device)
    # launch shapes of the pixel kernels, timed on this batch (not in the
    # timed region; cached per process)
    lm.params.copy_(p0)
    args.schedule = lm.autotune() if not args.no_tune else lm.batch.schedule.as_dict()
    log(f'bench: schedule {args.schedule}')

    def step():
        lm.params.copy_(p0)
        lm.run(settings, PARAMS)
        return lm.fetch()

    mark = roctx_marker() if args.mark else None
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        if mark:
            mark.push(f'lm step {i}')
        params, res = step()
        if mark:
            mark.pop()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return lm, dt, params, res, settings, p0


class roctx_marker:
    """roctx ranges (librocprofiler-sdk-roctx) around the timed steps, so
    that a marker trace shows which kernels each step ran and its wall
    time."""

    def __init__(self):
        import ctypes
        self._lib = ctypes.CDLL('librocprofiler-sdk-roctx.so.1')
        self._lib.roctxRangePushA.argtypes = [ctypes.c_char_p]
        self._lib.roctxRangePushA.restype = ctypes.c_int
        self._lib.roctxRangePop.restype = ctypes.c_int

    def push(self, name: str) -> None:
        self._lib.roctxRangePushA(name.encode())

    def pop(self) -> None:
        self._lib.roctxRangePop()


def valu_probe(reps: int = 5) -> dict:
    """Attainable vector-ALU issue rates of code built with the library's
    flags (tests/native/lmz_valu_peak.hip): 16 independent chains per
    thread of FMA with uniform operands (VOP3 encoding), FMA with per-lane
    operands (the two-operand VOP2 v_fmac form), multiply and add, in FP32
    and FP64, over 8 workgroups of 256 threads per CU; the best of reps
    launches of each."""
    import ctypes
    import torch
    lib = ctypes.CDLL(os.path.join(ROOT, 'tests', 'native', 'liblmz_valu_peak.so'))
    lib.valu_peak.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int, ctypes.c_void_p,
                              ctypes.POINTER(ctypes.c_float)]
    props = torch.cuda.get_device_properties(0)
    blocks, iters = props.multi_processor_count * 8, 4096
    chains, nt = lib.valu_peak_chains(), lib.valu_peak_threads()
    rates = {}
    ops = (('fma_uniform', 2.), ('fma_lane', 2.), ('mul', 1.), ('add', 1.))
    for dtype, name, tdt in ((0, 'fp32', torch.float32), (1, 'fp64', torch.float64)):
        out = torch.empty(blocks * nt, dtype=tdt, device='cuda')
        for op, (opname, fl) in enumerate(ops):
            best = None
            for _ in range(reps):
                ms = ctypes.c_float()
                if lib.valu_peak(dtype, op, blocks, iters,
                                 ctypes.c_void_p(out.data_ptr()), ctypes.byref(ms)):
                    raise RuntimeError('valu_peak launch failed')
                best = ms.value if best is None else min(best, ms.value)
            inst = float(blocks * nt) * iters * chains
            rates[f'{name}_{opname}'] = {
                'lane_instructions_per_s': inst / (best * 1e-3),
                'tflops': fl * inst / (best * 1e-3) / 1e12, 'ms': best}
    return {'workload': f'{blocks} workgroups of {nt} threads, {chains} '
                        f'independent FMA chains x {iters} per thread',
            'rates': rates}


def kernel_times(lm, settings, args):
    """Average duration of the Jacobian and trial-forward kernels over the
    whole batch, launched on torch's current stream and timed with HIP
    events on that stream."""
    import torch
    from pylorenzmie_amd import _lmz
    lib = _lmz.load_library()
    b = lm.batch
    dev = lm.data.device
    coef = torch.zeros(b.coef_floats(), dtype=torch.float32, device=dev)
    dcoef = torch.zeros(b.dcoef_floats(), dtype=torch.float32, device=dev)
    nmax = torch.zeros(2 * b.nfeat, dtype=torch.int32, device=dev)
    work = torch.zeros(b.work_doubles, dtype=torch.float64, device=dev)
    partials = torch.zeros(36 * b.ntiles, dtype=torch.float64, device=dev)
    chi2 = torch.zeros(b.ntiles, dtype=torch.float64, device=dev)
    s = _lmz.stream_handle()
    _lmz.check(lib.lmz_mie_ab(b.handle, _lmz.ptr(lm.params), None, _lmz.ptr(nmax),
                              _lmz.ptr(coef), _lmz.ptr(dcoef), None, _lmz.ptr(work), s))
    flags = 3
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]

    def jac():
        _lmz.check(lib.lmz_jacobian(b.handle, _lmz.ptr(lm.params), None, _lmz.ptr(nmax),
                                    _lmz.ptr(coef), _lmz.ptr(dcoef), None, flags,
                                    _lmz.ptr(lm.data), None, None, _lmz.ptr(partials),
                                    None, s))

    def fwd():
        _lmz.check(lib.lmz_hologram(b.handle, _lmz.ptr(lm.params), None, _lmz.ptr(nmax),
                                    _lmz.ptr(coef), None, flags, _lmz.ptr(lm.data),
                                    None, _lmz.ptr(chi2), None, s))
    # the LM's Jacobian pass on dense crops: radial tables of the sums,
    # interpolated per pixel (k_jac_rtab, k_jacobian_rtab and their tile
    # lists), the same partials
    import ctypes
    nb = ctypes.c_size_t()
    _lmz.check(lib.lmz_jacobian_tabulated_bytes(b.handle, ctypes.byref(nb)))
    tabwork = torch.empty(nb.value // 8 + 64, dtype=torch.float64, device=dev)

    def jtab():
        _lmz.check(lib.lmz_jacobian_tabulated(
            b.handle, _lmz.ptr(lm.params), _lmz.ptr(nmax), _lmz.ptr(coef),
            _lmz.ptr(dcoef), flags, _lmz.ptr(lm.data), _lmz.ptr(partials),
            _lmz.ptr(tabwork), s))
    # the LM's trial forward pass on dense crops: FP64 radial tables of the
    # sums, interpolated per pixel (k_fwd_ftab, k_forward_ftab)
    _lmz.check(lib.lmz_residual_tabulated_bytes(b.handle, ctypes.byref(nb)))
    fwork = torch.empty(nb.value // 8 + 64, dtype=torch.float64, device=dev)
    # the residual written as the LM writes it (FP32, one per pixel)
    fres = torch.empty_like(lm.data)

    def ftab():
        _lmz.check(lib.lmz_residual_tabulated(
            b.handle, _lmz.ptr(lm.params), _lmz.ptr(nmax), _lmz.ptr(coef),
            flags, _lmz.ptr(lm.data), _lmz.ptr(fres), _lmz.ptr(chi2),
            _lmz.ptr(fwork), s))
    jac()
    fwd()
    jtab()
    ftab()
    reps = args.kernel_reps
    et = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ev[0].record()
    for _ in range(reps):
        jac()
    ev[1].record()
    for _ in range(reps):
        fwd()
    ev[2].record()
    et[0].record()
    for _ in range(reps):
        jtab()
    et[1].record()
    for _ in range(reps):
        ftab()
    et[2].record()
    torch.cuda.synchronize()
    tj = ev[0].elapsed_time(ev[1]) / reps * 1e-3
    tf = ev[1].elapsed_time(ev[2]) / reps * 1e-3
    tt = et[0].elapsed_time(et[1]) / reps * 1e-3
    tft = et[1].elapsed_time(et[2]) / reps * 1e-3
    del tabwork, fwork, fres
    nm = nmax.cpu().numpy()[0::2].astype(float)
    P = np.array(lm.npix, float)
    # radial table samples per launch, as lmz_pixel.h's rtab_params and
    # ftab_params place them (dense crops; the particle's axis from the
    # parameters, the crop from the feature table)
    pr = lm.params.cpu().numpy()
    kpx = 2. * np.pi / INSTRUMENT['wavelength'] * INSTRUMENT['n_m'] * \
        INSTRUMENT['magnification']
    side = np.sqrt(np.asarray(lm.npix, float))
    px, py, Zk = pr[:, 0], pr[:, 1], kpx * np.abs(pr[:, 2])
    nx = np.maximum(np.maximum(-px, px - (side - 1)), 0.)
    ny = np.maximum(np.maximum(-py, py - (side - 1)), 0.)
    fx = np.maximum(px, side - 1 - px)
    fy = np.maximum(py, side - 1 - py)
    span = kpx * (np.hypot(fx, fy) - np.hypot(nx, ny))
    m_j = np.floor(span / np.minimum(0.35 * Zk / nm, 1.) + 12.)
    m_f = np.floor(span / np.minimum(0.25 * Zk / nm, 1.) + 24.)
    tables = {'jacobian_samples': float(m_j.sum()), 'forward_samples': float(m_f.sum())}
    flop_j = float(np.sum(P * (296. * nm + 480.)))
    flop_f = float(np.sum(P * (64. * nm + 70.)))
    # the standalone J^T J reduction over a Jacobian written to HBM
    # (lmz_normal_eq: 7 J columns + the residual, 32 B per pixel)
    npx = int(b.npix_total)
    J = torch.empty(7 * npx, dtype=torch.float32, device=dev)
    r = torch.empty(npx, dtype=torch.float32, device=dev)
    neq = torch.empty(36 * b.nfeat, dtype=torch.float64, device=dev)
    _lmz.check(lib.lmz_jacobian(b.handle, _lmz.ptr(lm.params), None, _lmz.ptr(nmax),
                                _lmz.ptr(coef), _lmz.ptr(dcoef), None, flags,
                                _lmz.ptr(lm.data), _lmz.ptr(J), _lmz.ptr(r), None,
                                None, s))

    def normal_eq():
        _lmz.check(lib.lmz_normal_eq(b.handle, _lmz.ptr(J), _lmz.ptr(r),
                                     _lmz.ptr(neq), s))
    normal_eq()
    ev[2].record()
    for _ in range(max(reps, 5)):
        normal_eq()
    ev[3].record()
    torch.cuda.synchronize()
    tn = ev[2].elapsed_time(ev[3]) / max(reps, 5) * 1e-3
    del J, r
    # algorithmic FLOP of the tabulated passes: the order loop per table
    # sample (296 N, 64 N) and per pixel the interpolation (2 x 30 x 6 + 4 x
    # 6 weights, 2 x 6 x 12 + 4 x 12 weights), the turn by e^{i psi} (15
    # and 3 complex products) and the pixel's geometry and projection (480,
    # 70 as the loop's)
    flop_jt = float(np.sum(296. * nm * m_j) + np.sum(P) * (360. + 24. + 90. + 480.))
    flop_ft = float(np.sum(64. * nm * m_f) + np.sum(P) * (144. + 48. + 18. + 70.))
    return (tj, tf, tn, tt, tft, flop_j, flop_f, flop_jt, flop_ft,
            float(np.sum(P)), float(np.mean(nm)), tables)


HEADLINE_FIXTURE = os.path.join(ROOT, 'tests', 'golden', 'headline.npz')


def load_parity_sample(args, truth, start, data_dev):
    """The oracle's fits of the first 32 headline features
    (tests/golden/headline.npz: the same draws, FP64 holograms plus seeded
    noise, fitted by the FP64 MINPACK oracle to 1e-12): their data replace
    the first 32 synthetic crops, so that the timed run fits them.  None
    for other workloads."""
    import torch
    if not (args.side == 201 and args.seed == 3 and args.sigma == 0.05 and
            args.nfeat >= 32 and os.path.exists(HEADLINE_FIXTURE)):
        return None
    fx = np.load(HEADLINE_FIXTURE)
    n = len(fx['head_truth'])
    if not (np.array_equal(fx['head_truth'], truth[:n]) and
            np.array_equal(fx['head_start'], start[:n])):
        return None
    px = args.side * args.side
    data_dev[: n * px] = torch.from_numpy(
        fx['head_data'].reshape(-1)).to(data_dev.device)
    return fx


def oracle_parity(fx, params, res):
    """The timed fits of the fixture's features against the oracle's: the
    same minimum where chi2 agree to 1e-7 relative (parameters then within
    1e-4 relative, k_p 1e-5 absolute), and whether the device's chi2 is
    higher than the oracle's by more than 1e-9 relative."""
    n = len(fx['oracle7_chi2'])
    ref_p, ref_c = fx['oracle7_values'], fx['oracle7_chi2']
    chi2 = res[:n, 35]
    rel_c = (chi2 - ref_c) / ref_c
    d = np.abs(params[:n, :7] - ref_p)
    rel = d / np.abs(ref_p)
    rel[:, 5] = d[:, 5]
    same = np.abs(rel_c) <= 1e-7
    ok = (np.delete(rel, 5, axis=1).max(1) <= 1e-4) & (rel[:, 5] <= 1e-5)
    capped = (res[:n, 38].astype(int) & 8) != 0
    # the same MINPACK given an exact Jacobian (the oracle's 5-point
    # differences), the method the device's analytic Jacobian stands for
    dj = np.abs(params[:n, :7] - fx['oracle7j_values'])
    relj = dj / np.abs(fx['oracle7j_values'])
    relj[:, 5] = dj[:, 5]
    okj = (np.delete(relj, 5, axis=1).max(1) <= 1e-4) & (relj[:, 5] <= 1e-5)
    return {'reference': 'FP64 oracle (numpy model + scipy MINPACK, 1e-12 '
                         'tolerances) fits of the first 32 features, '
                         'tests/golden/headline.npz',
            'within_1e-4_of_exact_jacobian_minpack': int(okj.sum()),
            'features': int(n),
            'same_minimum': int(same.sum()),
            'same_minimum_within_1e-4': int((same & ok).sum()),
            'higher_chi2': int((rel_c > 1e-9).sum()),
            'lower_chi2': int((rel_c < -1e-7).sum()),
            'capped_at_50': int(capped.sum()),
            'max_param_rel_diff_same_minimum': (float(np.delete(rel[same], 5, axis=1).max())
                                                if same.any() else None),
            'max_kp_abs_diff_same_minimum': (float(rel[same, 5].max())
                                             if same.any() else None),
            'chi2_rel_diff': [float(v) for v in rel_c]}


def convergence_check(lm, params, res, p0):
    """Distance of the timed fits from the minimum each one is heading for:
    the same LM rerun from the same starts to 1e-13 tolerances and 200
    iterations (not timed).  Features capped in either run are left out.
    Relative differences for all parameters but k_p, whose value can be 0:
    its difference is absolute."""
    from pylorenzmie_amd.analysis.lm import LMSettings
    tight = LMSettings(max_iter=200, ftol=1e-13, xtol=1e-13, gtol=1e-15)
    lm.params.copy_(p0)
    lm.run(tight, PARAMS)
    pref, rref = lm.fetch()
    keep = (((rref[:, 38].astype(int) | res[:, 38].astype(int)) & 8) == 0)
    rel = np.abs(params[:, :7] - pref[:, :7]) / np.abs(pref[:, :7])
    rel[:, 5] = np.abs(params[:, 5] - pref[:, 5])
    rel = rel[keep]
    ok = (np.delete(rel, 5, axis=1).max(1) <= 1e-4) & (rel[:, 5] <= 1e-5)
    return {'reference': 'the same LM from the same starts to 1e-13 '
                         'tolerances, <=200 iterations (self-consistency)',
            'features_compared': int(keep.sum()),
            'frac_within_1e-4_rel_and_kp_1e-5_abs': float(ok.mean()),
            'p99': {k: float(np.percentile(rel[:, i], 99))
                    for i, k in enumerate(PARAMS)},
            'median': {k: float(np.median(rel[:, i]))
                       for i, k in enumerate(PARAMS)}}


def kernel_source_hash() -> str:
    """First 16 hex digits of the SHA-256 of the library's sources (the
    HIP kernels, their headers and include/lmz.h), in name order."""
    import glob
    import hashlib
    h = hashlib.sha256()
    files = sorted(glob.glob(os.path.join(ROOT, 'pylorenzmie_amd', 'csrc', 'lmz_*')))
    files.append(os.path.join(ROOT, 'include', 'lmz.h'))
    for path in files:
        h.update(os.path.basename(path).encode())
        with open(path, 'rb') as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def committed_figure(name, key, nfeat, side):
    """A per-launch figure of a kernel from a PMC summary committed under
    profiles/ for this workload, or None for other workloads: HBM bytes
    (FETCH_SIZE + WRITE_SIZE, calibrated against a copy kernel) or counted
    FP32 FLOP (SQ_INSTS_VALU_FLOPS_FP32).  The summary names the source
    hash (kernel_source_hash) of the build it measured; a figure measured
    on other sources is None, never a stale number."""
    import re
    path = os.path.join(ROOT, 'profiles', f'{name}_{nfeat}x{side}x{side}.md')
    if not os.path.exists(path):
        return None
    text = open(path).read()
    stamp = re.search(r'source_sha256_16\s*=\s*([0-9a-f]{16})', text)
    if not stamp or stamp.group(1) != kernel_source_hash():
        return None
    m = re.search(key + r'\s*=\s*([0-9.eE+]+)', text)
    return float(m.group(1)) if m else None


def run_fd_batched(lm, p0, settings, params_adjoint, res_adjoint):
    """The same batched LM with a forward-difference Jacobian of the FP64
    hologram in place of the analytic one (steps 1e-5 |p_j|, the
    reference's diff_step): one warm and one timed run of all features.
    Returns fits/s, mean iterations, and the fraction of features whose
    fit agrees with the analytic-Jacobian fit to 1e-4 (relative; k_p 1e-5
    absolute) among those both runs converged."""
    import dataclasses
    import torch
    fd = dataclasses.replace(settings, fd_step=1e-5)
    lm.params.copy_(p0)
    lm.run(fd, PARAMS)
    torch.cuda.synchronize()
    lm.params.copy_(p0)
    t0 = time.perf_counter()
    lm.run(fd, PARAMS)
    params, res = lm.fetch()
    dt = time.perf_counter() - t0
    ok = lambda r: ((r[:, 38].astype(int) & 7) != 0) & ((r[:, 38].astype(int) & 56) == 0)
    both = ok(res) & ok(res_adjoint)
    d = np.abs(params[:, :7] - params_adjoint[:, :7])
    rel = d / np.maximum(np.abs(params_adjoint[:, :7]), 1e-300)
    rel[:, 5] = d[:, 5] * 1e-4 / 1e-5           # k_p: absolute 1e-5
    within = np.all(rel <= 1e-4, axis=1)
    return dict(value=lm.nfeat / dt, unit='fits/s', fd_step=1e-5,
                iterations_mean=float(res[:, 37].mean()),
                agree_with_adjoint=float(within[both].mean()) if both.any() else None,
                features_compared=int(both.sum()), seconds=dt)


def run_fd(args, truth, start, data_host):
    """Plain GPU FD form on a sample of features, sequentially."""
    from pylorenzmie_amd.theory import LorenzMie, Sphere, Instrument
    from pylorenzmie_amd.analysis import hipOptimizer, Hologram
    side = args.side
    n = min(args.fd_sample, len(start))
    t0 = time.perf_counter()
    for i in range(n):
        inst = Instrument(noise=args.sigma, **INSTRUMENT)
        s = Sphere(**dict(zip(PARAMS[:6], start[i, :6])))
        m = LorenzMie(coordinates=LorenzMie.meshgrid((side, side)), particle=s,
                      instrument=inst, alpha=start[i, 6])
        opt = hipOptimizer(model=m)
        opt.variables = list(PARAMS)
        opt.optimize(Hologram(data_host[i].astype(float)))
    dt = time.perf_counter() - t0
    return n / dt, n, dt


def run_cpu(args, start, data_host):
    """FP64 oracle: numpy model + scipy MINPACK LM at the reference's
    default settings with its SVD uncertainties (oracle/fit.py, as
    analysis/Optimizer.py:196-278 drives it), one feature at a time."""
    from oracle import fit, lorenzmie as ref
    side = args.side
    c = ref.meshgrid((side, side))
    n = 0
    t0 = time.perf_counter()
    while True:
        i = n % len(start)
        fit.fit(data_host[i], c, dict(zip(PARAMS, start[i])), variables=PARAMS,
                instrument=INSTRUMENT, noise=args.sigma)
        n += 1
        if time.perf_counter() - t0 > args.cpu_seconds or n >= args.cpu_max_fits:
            break
    dt = time.perf_counter() - t0
    return n / dt, n, dt


def run_cpu_pool(args, start, data_dev):
    """The same oracle fits on args.cpu_workers host cores (oracle/cpu_pool.py,
    one process per core, two features each), started as a child process
    that never touches the GPU.  The sample goes through a file."""
    import subprocess
    import tempfile
    n = 2 * args.cpu_workers
    data = data_dev[: args.side * args.side * n].cpu().numpy()
    data = data.reshape(n, args.side, args.side)
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, 'sample.npz')
        np.savez(path, data=data, start=start[:n], sigma=np.array(args.sigma))
        env = dict(os.environ, OMP_NUM_THREADS='1', MKL_NUM_THREADS='1',
                   OPENBLAS_NUM_THREADS='1')
        out = subprocess.run([sys.executable, '-m', 'oracle.cpu_pool', path,
                              str(args.cpu_workers)], capture_output=True,
                             text=True, env=env, cwd=ROOT, timeout=900,
                             check=True)
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return dict(value=r['value'], unit='fits/s', cores=r['workers'],
                kind='port', sample=f"{r['fits']} 7-parameter fits of "
                f"{args.side}^2 crops, {r['workers']} processes, "
                f"{r['seconds']:.1f} s wall")


# ---------------------------------------------------------------------------
# The other configurations (--config N): one JSON line each with the
# quantities the measurement plan asks of them.  Seeds are the config index.
def _timed(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps * 1e-3


def _draw_cfg(rng, sides, alpha):
    """Parameters as configs 1-3: centre +-5 px, z U[50, 400], a_p
    U[0.4, 3.0], n_p U[1.38, 1.65], k_p U[0, 0.01], alpha U[0.8, 1.2] or 1."""
    p = np.zeros((len(sides), 7))
    for i, s in enumerate(sides):
        c = (s - 1) / 2
        p[i] = (c + rng.uniform(-5, 5), c + rng.uniform(-5, 5),
                rng.uniform(50., 400.), rng.uniform(0.4, 3.0),
                rng.uniform(1.38, 1.65), rng.uniform(0., 0.01),
                rng.uniform(0.8, 1.2) if alpha else 1.0)
    return p


def _oracle_error(hb, params, picks, rng, layers=None):
    """(max |I - I_oracle|, max of the same less the float32 rounding of
    the stored value 2^-23 I) over 2000 sampled pixels of each picked
    feature."""
    from oracle import lorenzmie as ref
    host = hb.out.cpu().numpy()
    worst, excess = 0., -np.inf
    for i in picks:
        ny, nx = hb.shapes[i]
        idx = np.unique(rng.choice(ny * nx, min(2000, ny * nx), replace=False))
        c = ref.meshgrid((ny, nx))[:, idx]
        if layers is None:
            want = ref.hologram(c, dict(zip(ref.PARAMS, params[i])), **INSTRUMENT)
        else:
            lay = np.asarray(layers[i])
            ab = (ref.mie_coefficients(lay[0, 0], lay[0, 1], lay[0, 2], 1.34, 0.447)
                  if len(lay) == 1 else
                  ref.mie_coefficients_layered(lay[:, 0], lay[:, 1], lay[:, 2],
                                               1.34, 0.447))
            want = ref.hologram(c, dict(x_p=params[i, 0], y_p=params[i, 1],
                                        z_p=params[i, 2], ab=ab),
                                alpha=params[i, 6], **INSTRUMENT)
        err = np.abs(host[hb.offsets[i] + idx] - want)
        worst = max(worst, float(err.max()))
        excess = max(excess, float((err - 2. ** -23 * np.abs(want)).max()))
    return worst, excess


def run_config(args):
    import torch
    from pylorenzmie_amd import _lmz
    from pylorenzmie_amd.theory import Instrument
    from pylorenzmie_amd.theory.batch import HologramBatch
    cfg = args.config
    rng = np.random.default_rng(cfg)
    inst = Instrument(**INSTRUMENT)
    line = {'config': cfg}
    if cfg == 1:
        sides = [int(rng.integers(64, 401)) for _ in range(1024)]
        params = _draw_cfg(rng, sides, alpha=False)
        caps = [_lmz.nmax_cap(p[3], p[4], p[5], inst.n_m, inst.wavelength)
                for p in params]
        hb = HologramBatch([(s, s) for s in sides], caps, instrument=inst)
        hb.load(params)
        t = _timed(hb.compute, 5)
        nm = hb.nmax.cpu().numpy().astype(float)
        P = np.array([s * s for s in sides], float)
        flop = float(np.sum(P * (64. * nm + 70.)))
        line.update(workload='batched forward holograms, 1024 crops of side '
                             'U{64..400}, Mie + FP64 pixel sum',
                    pixels=float(P.sum()), ms=t * 1e3, holograms_per_s=1024 / t,
                    achieved_tflops=flop / t / 1e12,
                    frac_fp64_peak=flop / t / 1e12 / 78.6)
        e, x = _oracle_error(hb, params, range(0, 1024, 256), rng)
        line.update(max_abs_err_vs_oracle=e, max_err_beyond_float32_rounding=x)
    elif cfg == 2:
        from pylorenzmie_amd.theory.autograd import hologram
        from oracle import lorenzmie as ref
        n, s = 256, 256
        truth = _draw_cfg(rng, [s] * n, alpha=True)
        q = truth.copy()
        q[:, :3] += (2., -2., 2.)
        q[:, 3] *= 1.05
        q[:, 4] -= 0.02
        caps = [_lmz.nmax_cap(p[3], p[4], p[5], inst.n_m, inst.wavelength)
                for p in np.maximum(truth, q)]
        hb = HologramBatch([(s, s)] * n, caps, instrument=inst)
        hb.load(truth)
        g = torch.Generator(device='cuda')
        g.manual_seed(cfg)
        data = hb.compute().clone() + 0.05 * torch.randn(
            hb.npix_total, generator=g, device='cuda', dtype=torch.float32)
        pq = torch.tensor(q, device='cuda', requires_grad=True)

        def step():
            pq.grad = None
            I = hologram(pq, hb)
            ((I.double() - data.double()) ** 2).sum().backward()
        t = _timed(step, 5)
        grad = pq.grad.cpu().numpy()
        # FP64 central differences of each feature's loss, h = 1e-6 relative
        d_host = data.cpu().numpy()
        c = ref.meshgrid((s, s))
        worst = 0.
        for i in (0, 1):
            nmax = ref.wiscombe_yang(ref.size_parameter(q[i, 3]),
                                     (q[i, 4] + 1j * q[i, 5]) / 1.34)
            d = d_host[hb.offsets[i]:hb.offsets[i] + s * s].astype(float)

            def L(v):
                I = ref.hologram(c, dict(x_p=v[0], y_p=v[1], z_p=v[2], a_p=v[3],
                                         n_p=v[4], k_p=v[5], nmax=nmax),
                                 alpha=v[6], **INSTRUMENT)
                return np.sum((I - d) ** 2)
            for j in range(7):
                h = 1e-6 * max(abs(q[i, j]), 1e-3)
                vp, vm = q[i].copy(), q[i].copy()
                vp[j] += h
                vm[j] -= h
                fd = (L(vp) - L(vm)) / (2 * h)
                worst = max(worst, abs(grad[i, j] - fd) / max(abs(fd), 1e-6))
        line.update(workload='forward + adjoint (VJP) of L = sum (I - d)^2, '
                             '256 crops of 256x256, alpha U[0.8, 1.2]',
                    pixels=float(hb.npix_total), ms_forward_and_backward=t * 1e3,
                    max_rel_err_vs_fp64_central_differences=worst)
    elif cfg == 3:
        from pylorenzmie_amd.analysis.lm import LMBatch, LMSettings
        n = args.nfeat
        sides = [int(round(np.exp(rng.uniform(np.log(48.), np.log(512.)))))
                 for _ in range(n)]
        truth = _draw_cfg(rng, sides, alpha=True)
        sigma = rng.choice([0.02, 0.05, 0.1], n)
        sg = rng.choice([-1., 1.], size=(n, 5))
        start = truth.copy()
        start[:, :3] += 2 * sg[:, :3]
        start[:, 3] *= 1 + 0.05 * sg[:, 3]
        start[:, 4] += 0.02 * sg[:, 4]
        caps = [_lmz.nmax_cap(p[3], p[4], p[5], inst.n_m, inst.wavelength)
                for p in np.maximum(truth, start)]
        hb = HologramBatch([(x, x) for x in sides], caps, instrument=inst)
        hb.load(truth)
        g = torch.Generator(device='cuda')
        g.manual_seed(cfg)
        sig = torch.from_numpy(np.repeat(sigma, [x * x for x in sides]).astype(np.float32))
        data = hb.compute().clone() + sig.cuda() * torch.randn(
            hb.npix_total, generator=g, device='cuda', dtype=torch.float32)
        crops = [dict(data=np.zeros((x, x), np.float32), width=x, origin=(0., 0., 0.),
                      noise=float(sigma[i]), nmax_cap=caps[i])
                 for i, x in enumerate(sides)]
        lm = LMBatch(crops, inst)
        lm.data.copy_(data)
        settings = LMSettings(max_iter=50, ftol=args.ftol, xtol=args.xtol,
                              gtol=args.gtol)
        lm.load(start)
        lm.autotune()
        lm.load(start)
        lm.run(settings, PARAMS)
        lm.load(start)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lm.run(settings, PARAMS)
        params, res = lm.fetch()
        dt = time.perf_counter() - t0
        st = res[:, 38].astype(int)
        ok = ((st & 7) != 0) & ((st & 56) == 0)
        lm.params.copy_(torch.from_numpy(np.hstack([start, np.zeros((n, 1))])).cuda())
        (tj, tf, tn, tt, tft, flop_j, flop_f, flop_jt, flop_ft, npx, nmean,
         _) = kernel_times(lm, settings, args)
        bytes_n = 32. * npx + 36 * 8. * n
        line.update(workload=f'batched 7-parameter LM, {n} crops of side '
                             'log-U[48, 512], sigma in {0.02, 0.05, 0.1}',
                    pixels=float(npx), fits_per_s=n / dt, seconds=dt,
                    converged_frac=float(ok.mean()),
                    iterations_mean=float(res[:, 37].mean()),
                    forward_tabulated_ms=tft * 1e3,
                    jacobian_tabulated_ms=tt * 1e3,
                    jacobian_loop_ms=tj * 1e3,
                    jacobian_loop_tflops=flop_j / tj / 1e12,
                    jacobian_loop_frac_fp32_peak=flop_j / tj / 1e12 / PEAK_FP32,
                    forward_ms=tf * 1e3, forward_tflops=flop_f / tf / 1e12,
                    forward_frac_fp64_peak=flop_f / tf / 1e12 / PEAK_FP64,
                    normal_eq_gbs=bytes_n / tn / 1e9,
                    normal_eq_frac_hbm_peak=bytes_n / tn / 1e9 / PEAK_HBM)
    elif cfg == 4:
        from pylorenzmie_amd.analysis.lm import LMBatch, LMSettings
        from pylorenzmie_amd.theory.LayeredSphere import layered_cap
        n, side, mag = 512, 512, INSTRUMENT['magnification']
        params, layers = np.zeros((n, 7)), []
        for i in range(n):
            a_p = rng.uniform(0.5, 10.)
            if rng.random() < 0.5:
                L = int(rng.integers(2, 6))
                radii = np.sort(rng.uniform(0.05, 1., L)) * a_p
                radii[-1] = a_p
                nl = rng.uniform(1.33, 2.0, L)
            else:
                radii, nl = np.array([a_p]), rng.uniform(1.33, 2.0, 1)
            layers.append(np.stack([radii, nl, np.zeros(len(radii))], 1))
            c = (side - 1) / 2
            params[i] = (c + rng.uniform(-5, 5), c + rng.uniform(-5, 5),
                         a_p / mag + 30. + rng.uniform(0., 250.), a_p, nl[-1], 0.,
                         rng.uniform(0.8, 1.2))
        caps = [layered_cap(lay, 1.34, 0.447) for lay in layers]
        hb = HologramBatch([(side, side)] * n, caps, instrument=inst)
        hb.load(params)
        hb.load_layers(layers)
        t = _timed(hb.compute, 3)
        nm = hb.nmax.cpu().numpy()
        lay_idx = [i for i in range(n) if len(layers[i]) > 1]
        err, excess = _oracle_error(hb, params, [int(np.argmax(nm))] + lay_idx[:3],
                                    rng, layers)
        g = torch.Generator(device='cuda')
        g.manual_seed(cfg)
        data = hb.compute().clone() + 0.05 * torch.randn(
            hb.npix_total, generator=g, device='cuda', dtype=torch.float32)
        inst_n = Instrument(noise=0.05, **INSTRUMENT)
        out = {}
        # both halves fit all seven parameters from the uneven-crop
        # workload's starts; a layered sphere's a_p scales every radius and
        # n_p + i k_p is its outer layer's index (lmz_lm_run_layered)
        for name, idx in (('layered', lay_idx),
                          ('homogeneous', [i for i in range(n) if len(layers[i]) == 1])):
            vary = list(PARAMS)
            sub = data.view(n, side * side)[idx].reshape(-1)
            crops = [dict(data=np.zeros((side, side), np.float32), width=side,
                          origin=(0., 0., 0.),
                          nmax_cap=(layered_cap(layers[i], 1.34, 0.447, headroom=0.25)
                                    if name == 'layered' else
                                    _lmz.nmax_cap(1.05 * params[i, 3], params[i, 4],
                                                  0., 1.34, 0.447)))
                     for i in idx]
            lm = LMBatch(crops, inst_n)
            lm.data.copy_(sub)
            start = params[idx].copy()
            start[:, :3] += (2., -2., 2.)
            start[:, 3] *= 1.05
            start[:, 4] -= 0.02
            if name == 'layered':
                lm.fit_layers([layers[i] for i in idx])
            settings = LMSettings(max_iter=50)
            lm.load(start)
            lm.run(settings, vary)
            lm.load(start)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            lm.run(settings, vary)
            p, res = lm.fetch()
            dt = time.perf_counter() - t0
            st = res[:, 38].astype(int)
            ok = ((st & 7) != 0) & ((st & 56) == 0)
            redchi = res[:, 35] / (side * side - len(vary))
            good = ok & (redchi < 1.02)
            out[name] = dict(features=len(idx), fits_per_s=len(idx) / dt,
                             varied=vary, converged_frac=float(ok.mean()),
                             at_noise_level_frac=float(good.mean()),
                             median_abs_err_xyz=float(np.median(
                                 np.abs(p[good, :3] - params[idx][good, :3]))),
                             median_rel_err_a_n=float(np.median(
                                 np.abs(p[good, 3:5] / params[idx][good, 3:5] - 1.))))
            lm.close()
        line.update(workload='hard regime: 512 crops of 512x512, half layered '
                             '(2-5 layers; a_p scales the radii, outer n_p, k_p), '
                             'half homogeneous a_p U[0.5, 10] um; 7 parameters '
                             'from config-3 starts',
                    nmax_max=int(nm.max()), holograms_ms=t * 1e3,
                    max_abs_err_vs_oracle=err,
                    max_err_beyond_float32_rounding=excess, **out)
    else:
        raise SystemExit('--config must be 1, 2, 3 or 4')
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------------------
# Rows around the path (--extra NAME): one JSON line each, with the FP64
# oracle on one host core beside the GPU figure where the reference's own
# path is a CPU one.
def run_extra(args):
    import torch
    from pylorenzmie_amd.theory import Instrument, LorenzMie, Sphere
    from oracle import lorenzmie as ref
    line = {'extra': args.extra}
    if args.extra == 'superposed':
        # docs/MultiSphere.ipynb: five default spheres on a 480 x 640 grid
        # (4.43 s per hologram() there, hardware unstated)
        shape = (480, 640)
        parts = [Sphere(x_p=100. * (i + 1), y_p=240., z_p=100. + 50. * i)
                 for i in range(5)]
        m = LorenzMie(coordinates=LorenzMie.meshgrid(shape), particle=parts,
                      instrument=Instrument())
        h = m.hologram()
        torch.cuda.synchronize()
        reps = 20
        t0 = time.perf_counter()
        for _ in range(reps):
            h = m.hologram()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / reps
        c = ref.meshgrid(shape)
        dicts = [dict(x_p=p.x_p, y_p=p.y_p, z_p=p.z_p, a_p=p.a_p, n_p=p.n_p,
                      k_p=p.k_p) for p in parts]
        t0 = time.perf_counter()
        want = ref.hologram(c, dicts, **INSTRUMENT)
        tc = time.perf_counter() - t0
        line.update(workload='hologram of 5 spheres (a_p 1 um, n_p 1.5) on a '
                             '480 x 640 grid, one fused pass',
                    ms_per_hologram=t * 1e3, max_abs_err_vs_oracle=float(
                        np.abs(h - want).max()),
                    cpu_baseline={'value': tc * 1e3, 'unit': 'ms per hologram',
                                  'cores': 1, 'kind': 'port'},
                    reference_notebook_ms=4430.)
    elif args.extra == 'frame':
        from pylorenzmie_amd.analysis import Frame
        rng = np.random.default_rng(11)
        ny, nx, side, k = 1024, 1280, 201, 16
        inst = Instrument(noise=0.05, **INSTRUMENT)
        truth, boxes = [], []
        for i in range(k):
            x0, y0 = 40 + 300 * (i % 4), 20 + 250 * (i // 4)
            truth.append((x0 + 100 + rng.uniform(-2, 2), y0 + 100 + rng.uniform(-2, 2),
                          rng.uniform(100., 250.), rng.uniform(0.5, 1.5),
                          rng.uniform(1.4, 1.6), 0.))
            boxes.append(((x0, y0), side, side))
        frame = np.ones((ny, nx))
        for q, ((x0, y0), _, _) in zip(truth, boxes):
            cc = ref.meshgrid((side, side), corner=(x0, y0))
            frame[y0:y0 + side, x0:x0 + side] = ref.hologram(
                cc, dict(zip(ref.PARAMS[:6], q)), **INSTRUMENT).reshape(side, side)
        frame += 0.05 * rng.standard_normal(frame.shape)
        dev = torch.from_numpy(frame.astype(np.float32)).cuda()

        def setup(data):
            f = Frame(data, instrument=inst)
            f.bboxes = boxes
            for ft, q in zip(f.features, truth):
                ft.particle.x_p += 1.5
                ft.particle.z_p, ft.particle.a_p, ft.particle.n_p = \
                    q[2] * 1.02, q[3] * 1.03, q[4] - 0.01
                ft.optimizer.settings.update(ftol=1e-9, xtol=1e-9, gtol=1e-10)
            return f
        f = setup(dev)
        f.optimize()
        torch.cuda.synchronize()
        f = setup(dev)
        np.random.seed(1)
        t0 = time.perf_counter()
        batched = f.optimize()
        torch.cuda.synchronize()
        tb = time.perf_counter() - t0
        f2 = setup(frame.astype(np.float32))
        np.random.seed(1)
        t0 = time.perf_counter()
        seq = [ft.optimize() for ft in f2.features]
        torch.cuda.synchronize()
        ts = time.perf_counter() - t0
        agree = np.mean([all(abs(batched.iloc[i][v] - seq[i][v]) <=
                             1e-4 * max(abs(seq[i][v]), 1e-3)
                             for v in ('x_p', 'y_p', 'z_p', 'a_p', 'n_p'))
                         for i in range(k)])
        # the DE estimates of all features, one search after another
        fe = setup(dev)
        for i, ft in enumerate(fe.features):
            ft.estimator.seed = i
        np.random.seed(2)
        t0 = time.perf_counter()
        fe.estimate()
        torch.cuda.synchronize()
        est_ms = (time.perf_counter() - t0) * 1e3
        fits = fe.optimize()
        seq_ok = np.mean([all(abs(fits.iloc[i][v] - q[j]) <= 6 * fits.iloc[i]['d' + v]
                              for j, v in enumerate(('x_p', 'y_p', 'z_p', 'a_p', 'n_p')))
                          for i, q in enumerate(truth)])
        line.update(fits_from_sequential_estimates_at_truth=float(seq_ok))
        # the same searches advanced together, then the fits from them
        fb = setup(dev)
        for ft in fb.features:
            ft.estimator.seed = 0
        np.random.seed(2)
        t0 = time.perf_counter()
        fb.estimate(batched=True)
        torch.cuda.synchronize()
        estb_ms = (time.perf_counter() - t0) * 1e3
        fitb = fb.optimize()
        truth_ok = np.mean([all(abs(fitb.iloc[i][v] - q[j]) <= 6 * fitb.iloc[i]['d' + v]
                                for j, v in enumerate(('x_p', 'y_p', 'z_p', 'a_p', 'n_p')))
                            for i, q in enumerate(truth)])
        line.update(frame_estimate_batched_ms=estb_ms,
                    fits_from_batched_estimates_at_truth=float(truth_ok))
        line.update(workload=f'{k} features of 201x201 (10% masks) in a '
                             '1280x1024 frame in device memory',
                    frame_optimize_ms=tb * 1e3, fits_per_s=k / tb,
                    sequential_feature_optimize_ms=ts * 1e3,
                    speedup_vs_sequential=ts / tb, agreement=float(agree),
                    success=float(batched['success'].mean()),
                    frame_estimate_ms=est_ms)
    elif args.extra == 'de':
        from pylorenzmie_amd.analysis import DEEstimator, Hologram, Mask
        from pylorenzmie_amd.analysis.DEEstimator import PopulationChi2
        q = (100.3, 99.6, 150., 0.75, 1.45, 0.)
        c = ref.meshgrid((201, 201))
        rng = np.random.default_rng(2)
        d = (ref.hologram(c, dict(zip(ref.PARAMS[:6], q)), **INSTRUMENT) +
             0.05 * rng.standard_normal(201 * 201)).reshape(201, 201)
        inst = Instrument(noise=0.05, **INSTRUMENT)
        est = DEEstimator(model=LorenzMie(particle=Sphere(), instrument=inst),
                          seed=1)
        np.random.seed(1)
        est.estimate(Hologram(d))
        torch.cuda.synchronize()
        np.random.seed(1)
        t0 = time.perf_counter()
        g = est.estimate(Hologram(d))
        te = time.perf_counter() - t0
        # one generation's objective: the population on the GPU vs the oracle
        np.random.seed(3)
        data, coords = Mask(fraction=0.02).apply(Hologram(d))
        pop = np.zeros((30, 7))
        pop[:, :2] = q[:2]
        pop[:, 2] = rng.uniform(10., 600., 30)
        pop[:, 3] = rng.uniform(0.25, 10., 30)
        pop[:, 4] = rng.uniform(1.0, 3.0, 30)
        pop[:, 6] = 1.
        chi2 = PopulationChi2(coords, data, inst, pop[0],
                              cap=_lmz_cap(10., 3.0))
        chi2(pop)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(10):
            chi2(pop)
        tg = (time.perf_counter() - t0) / 10
        t0 = time.perf_counter()
        for p in pop[:6]:
            np.sum(((ref.hologram(coords, dict(zip(ref.PARAMS[:6], p)),
                                  **INSTRUMENT) - data) / 0.05) ** 2)
        tc = (time.perf_counter() - t0) / 6
        line.update(workload='DE estimate of z_p, a_p, n_p on a 201x201 crop '
                             '(2% mask, popsize 10, default bounds)',
                    estimate_ms=te * 1e3,
                    estimate={k2: float(g[k2]) for k2 in ('z_p', 'a_p', 'n_p')},
                    generation_ms=tg * 1e3, candidates_per_s=30 / tg,
                    cpu_baseline={'value': 1. / tc, 'unit': 'candidates/s',
                                  'cores': 1, 'kind': 'port'})
    elif args.extra == 'batch':
        # TorchLorenzMieBatch.batch_hologram (theory/torchLorenzMieBatch.py:
        # 258-283): 256 holograms on one 201 x 201 grid, 1-3 spheres each
        from pylorenzmie_amd.theory import LorenzMieBatch
        rng = np.random.default_rng(4)
        B, side = 256, 201
        lists = []
        for _ in range(B):
            lists.append([Sphere(x_p=rng.uniform(60., 140.), y_p=rng.uniform(60., 140.),
                                 z_p=rng.uniform(80., 300.), a_p=rng.uniform(0.4, 1.5),
                                 n_p=rng.uniform(1.38, 1.65))
                          for _ in range(int(rng.integers(1, 4)))])
        m = LorenzMieBatch(coordinates=LorenzMie.meshgrid((side, side)),
                           instrument=Instrument())
        h = m.batch_hologram(lists)
        t0 = time.perf_counter()
        reps = 5
        for _ in range(reps):
            h = m.batch_hologram(lists)
        tb = (time.perf_counter() - t0) / reps
        m.batch_hologram(lists, device=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            m.batch_hologram(lists, device=True)
        torch.cuda.synchronize()
        td = (time.perf_counter() - t0) / reps
        # the same holograms one LorenzMie.hologram() call at a time
        m1 = LorenzMie(coordinates=LorenzMie.meshgrid((side, side)),
                       instrument=Instrument())
        t0 = time.perf_counter()
        for plist in lists[:32]:
            m1.particle = plist
            m1.hologram()
        ts = (time.perf_counter() - t0) / 32
        c = ref.meshgrid((side, side))
        err, tc = 0., 0.
        for i in range(0, B, 64):
            dicts = [dict(x_p=p.x_p, y_p=p.y_p, z_p=p.z_p, a_p=p.a_p,
                          n_p=p.n_p, k_p=p.k_p) for p in lists[i]]
            t0 = time.perf_counter()
            want = ref.hologram(c, dicts, **INSTRUMENT)
            tc += time.perf_counter() - t0
            err = max(err, float(np.abs(h[i] - want).max()))
        tc /= len(range(0, B, 64))
        line.update(workload=f'LorenzMieBatch.batch_hologram: {B} holograms '
                             f'of 1-3 spheres (a_p 0.4-1.5 um) on one '
                             f'{side} x {side} grid, one pass',
                    ms_per_batch=tb * 1e3, holograms_per_s=B / tb,
                    ms_per_batch_on_device=td * 1e3,
                    holograms_per_s_on_device=B / td,
                    sequential_ms_per_hologram=ts * 1e3,
                    speedup_vs_sequential=ts * B / tb,
                    max_abs_err_vs_oracle=err,
                    cpu_baseline={'value': 1. / tc, 'unit': 'holograms/s',
                                  'cores': 1, 'kind': 'port'})
    elif args.extra == 'normalize':
        # utilities/normalizer.py:11-76: Normalizer('median'), ('filter',
        # size 51) and ('reference') of a 1280 x 1024 camera frame already
        # in device memory, as 8-bit, 16-bit, float32 and float64 pixels
        from pylorenzmie_amd.utilities import Normalizer
        ny, nx, size, reps = 1024, 1280, 51, 20
        raw = _synthetic_frame(ny, nx)
        frames = {'u8': raw, 'u16': raw.astype(np.uint16),
                  'f32': raw.astype(np.float32) / np.float32(100.),
                  'f64': raw.astype(np.float64) / 100.}
        background = torch.from_numpy(_synthetic_frame(ny, nx, seed=1)).cuda()
        cases = []
        for name, arr in frames.items():
            d = torch.from_numpy(arr).cuda()
            for method, norm in (('median', Normalizer('median')),
                                 ('filter', Normalizer('filter', size=size)),
                                 ('reference', Normalizer('reference',
                                                          reference=background))):
                ms = _timed(lambda: norm(d), reps) * 1e3
                cases.append(dict(dtype=name, method=method,
                                  ms_per_frame=round(ms, 4),
                                  mpix_per_s=round(ny * nx / ms / 1e3, 1)))
        line.update(workload=f'frame normalization, {nx} x {ny} frames in '
                             f'device memory (median filter window {size})',
                    cases=cases)
    elif args.extra == 'valu_peak':
        # the attainable FP32 and FP64 FMA rates of code built with the
        # library's flags: independent FMA chains in registers
        props = torch.cuda.get_device_properties(0)
        probe = valu_probe()
        line.update(workload=probe.pop('workload'),
                    cus=props.multi_processor_count, rates=probe['rates'],
                    peak_published={'fp32': PEAK_FP32, 'fp32_unpacked': PEAK_FP32_UNPACKED,
                                    'fp64': 78.6})
    elif args.extra == 'stages':
        # the forward's per-pixel stages in FP32 or FP64 (variant bits as in
        # tests/native/lmz_stages.hip): time per launch of one sphere over a
        # 4096 x 4096 grid, one thread per pixel
        import ctypes
        from pylorenzmie_amd import _lmz
        P = ctypes.c_void_p
        lib = ctypes.CDLL(os.path.join(ROOT, 'tests', 'native', 'liblmz_stages.so'))
        lib.stage_gpu.argtypes = [ctypes.c_int, P, P, P, ctypes.c_int, P,
                                  ctypes.c_int, ctypes.c_int, P,
                                  ctypes.POINTER(ctypes.c_float)]
        chk = ctypes.CDLL(os.path.join(ROOT, 'tests', 'native', 'liblmz_hostcheck.so'))
        chk.chk_mie.restype = ctypes.c_int
        chk.chk_mie.argtypes = [ctypes.c_double] * 5 + [ctypes.c_int, P, P, P]
        names = {0: 'fp64', 1: 'geometry', 2: 'angular', 4: 'xi', 8: 'products',
                 16: 'accumulate', 32: 'epilogue', 24: 'products+accumulate',
                 56: 'products+accumulate+epilogue', 60: 'all but geometry, angular',
                 62: 'all but geometry', 63: 'fp32'}
        side = 4096
        yy, xx = np.mgrid[0:side, 0:side]
        coords = np.ascontiguousarray(np.stack([xx.ravel(), yy.ravel(),
                                                np.zeros(side * side)]), np.float32)
        npix = side * side
        out = np.zeros(npix)
        cases = []
        inst = Instrument(**INSTRUMENT)
        for a_p in (0.75, 1.5, 3.0):
            coef = np.zeros((400, 40), np.float32)
            dcoef = np.zeros((400, 24), np.float32)
            ab = np.zeros((401, 4))
            N = chk.chk_mie(a_p, 1.5, 0.001, 1.34, 0.447, 400,
                            coef.ctypes.data_as(P), dcoef.ctypes.data_as(P),
                            ab.ctypes.data_as(P))
            f = _lmz.make_feature(npix, 0, (0., 0., 0.), inst, 400)
            p = np.array([2047.6, 2048.2, 200., a_p, 1.5, 0.001, 1.0, 0.])
            ms = {}
            for v, name in names.items():
                t = ctypes.c_float()
                if lib.stage_gpu(v, ctypes.byref(f), p.ctypes.data_as(P),
                                 coef.ctypes.data_as(P), N, coords.ctypes.data_as(P),
                                 npix, 5, out.ctypes.data_as(P), ctypes.byref(t)):
                    raise RuntimeError('stage_gpu failed')
                ms[name] = round(t.value, 4)
            cases.append(dict(a_p=a_p, nmax=N, ms=ms,
                              speed_vs_fp64={k: round(ms['fp64'] / v, 3)
                                             for k, v in ms.items()}))
        line.update(workload=f'one sphere (n_p 1.5, z_p 200 px) over a {side} x '
                             f'{side} grid, one thread per pixel; the stages named '
                             'run in FP32, the rest in FP64', cases=cases)
    else:
        raise SystemExit('--extra must be superposed, frame, de, batch, '
                         'normalize, valu_peak or stages')
    print(json.dumps(line), flush=True)


def _synthetic_frame(ny: int, nx: int, seed: int = 0) -> np.ndarray:
    """An 8-bit camera frame: fringes of two features on a sloped
    background, with shot noise."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[:ny, :nx].astype(float)
    v = 90. + 20. * x / nx + 10. * y / ny
    for cx, cy, k in ((0.3 * nx, 0.5 * ny, 0.35), (0.7 * nx, 0.3 * ny, 0.25)):
        r = np.hypot(x - cx, y - cy)
        v += 40. * np.cos(k * r) * np.exp(-r / 150.)
    v += rng.normal(0., 3., v.shape)
    return np.clip(v, 0, 255).astype(np.uint8)


def _lmz_cap(a_p, n_p):
    from pylorenzmie_amd import _lmz
    return _lmz.nmax_cap(a_p, n_p, 0., INSTRUMENT['n_m'], INSTRUMENT['wavelength'])


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--nfeat', type=int, default=4096)
    ap.add_argument('--side', type=int, default=201)
    ap.add_argument('--sigma', type=float, default=0.05)
    ap.add_argument('--ftol', type=float, default=1e-9)
    ap.add_argument('--xtol', type=float, default=1e-9)
    ap.add_argument('--gtol', type=float, default=1e-10)
    ap.add_argument('--fd-sample', type=int, default=32)
    ap.add_argument('--cpu-seconds', type=float, default=15.)
    ap.add_argument('--cpu-max-fits', type=int, default=4)
    ap.add_argument('--cpu-workers', type=int, default=-1,
                    help='time the oracle on this many host cores (one '
                         'process each, 2 features each; -1: the CPUs this '
                         'process may use, at most 16; 0: skip)')
    ap.add_argument('--kernel-reps', type=int, default=3)
    ap.add_argument('--calibrate-copy', type=int, default=0, metavar='MIB',
                    help='also run one device copy of this many MiB (PMC '
                         'traffic calibration) and exit')
    ap.add_argument('--dump', default='', help='write per-feature truth, '
                    'start, fitted parameters and LM results to this .npz')
    ap.add_argument('--no-tune', action='store_true',
                    help='keep the default launch schedule (no autotuner)')
    ap.add_argument('--mark', action='store_true',
                    help='roctx range around each timed step (for '
                         'rocprofv3 --marker-trace)')
    ap.add_argument('--no-fd', action='store_true')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--seed', type=int, default=3)
    ap.add_argument('--extra', default='',
                    help='superposed | frame | de | batch | normalize | valu_peak | stages: measure that path instead '
                         'of the headline (one JSON line)')
    ap.add_argument('--config', type=int, default=0,
                    help='1-4: measure that configuration instead of the '
                         'headline (one JSON line)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py runs on one GPU: there is no multi-GPU path '
              '(features would shard across GPUs with no communication).  '
              'Use --gpus 1.', file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), 'bench.py needs a GPU'
    if args.calibrate_copy:
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f'bench: copied {args.calibrate_copy} MiB device to device')
        return
    if args.config:
        run_config(args)
        return
    if args.extra:
        run_extra(args)
        return

    truth, start = draw_features(args.nfeat, args.side, args.seed)
    log(f'bench: generating {args.nfeat} synthetic {args.side}^2 holograms')
    data_dev = make_data(truth, args.side, args.sigma, args.seed)
    parity_fx = load_parity_sample(args, truth, start, data_dev)
    torch.cuda.synchronize()
    log('bench: batched LM warmup + timed steps')
    lm, dt, params, res, settings, p0 = run_batched(args, truth, start, data_dev)
    ms_per_step = dt / args.steps * 1e3
    fits_per_s = args.nfeat * args.steps / dt
    status = res[:, 38].astype(int)
    iters = res[:, 37]
    ok = ((status & 7) != 0) & ((status & 48) == 0)
    err = np.abs(params[:, :7] - truth)
    npix = args.side * args.side
    redchi = res[:, 35] / (npix - 7)
    capped = (status & 8) != 0
    far = (err[:, 0] > 0.5) | (err[:, 2] > 5.) | (err[:, 3] > 0.05)
    log(f'bench: {fits_per_s:.1f} fits/s, {ms_per_step:.1f} ms/step, '
        f'iters mean {iters.mean():.1f} max {iters.max():.0f}, '
        f'converged {ok.mean():.3f}')

    if args.dump:
        np.savez_compressed(args.dump, truth=truth, start=start, params=params,
                            result=res)
    parity = oracle_parity(parity_fx, params, res) if parity_fx is not None else None
    if parity:
        log(f"bench: oracle parity: same minimum {parity['same_minimum']}/"
            f"{parity['features']}, higher chi2 {parity['higher_chi2']}")
    log('bench: agreement with fully converged fits')
    agree = convergence_check(lm, params, res, p0)
    log(f"bench: {agree['frac_within_1e-4_rel_and_kp_1e-5_abs']:.4f} of "
        f"{agree['features_compared']} fits within 1e-4 of the converged minimum")
    log('bench: kernel timing')
    (tj, tf, tn, tt, tft, flop_j, flop_f, flop_jt, flop_ft, npx, nmean,
     tables) = kernel_times(lm, settings, args)
    ach_j = flop_j / tj / 1e12
    ach_f = flop_f / tf / 1e12
    ach_jt = flop_jt / tt / 1e12
    ach_ft = flop_ft / tft / 1e12
    log(f'bench: tabulated forward {tft * 1e3:.2f} ms {ach_ft:.1f} TFLOP/s, '
        f'jacobian {tt * 1e3:.2f} ms {ach_jt:.1f} TFLOP/s; order loops '
        f'{tf * 1e3:.2f} ms, {tj * 1e3:.2f} ms')

    # J reads the data (4 B / pixel) and writes 36 FP64 partials per tile
    bytes_j = 4. * npx + 36 * 8. * lm.batch.ntiles
    # reductions: J^T J from J in HBM (32 B / pixel in, 288 B / feature
    # out); the residual and chi2 (4 B / pixel in, the FP32 residual and
    # one FP64 per tile out)
    bytes_n = 32. * npx + 36 * 8. * lm.batch.nfeat
    bytes_r = 8. * npx + 8. * lm.batch.ntiles
    reductions = {
        'normal_eq': {'kernel': 'k_normal_eq', 'bound': 'hbm',
                      'achieved': bytes_n / tn / 1e9, 'peak': PEAK_HBM,
                      'unit': 'GB/s', 'frac': bytes_n / tn / 1e9 / PEAK_HBM,
                      'avg_launch_ms': tn * 1e3, 'bytes_per_launch': bytes_n,
                      'traffic': committed_figure(
                          'k_normal_eq_traffic', 'traffic_bytes_per_launch',
                          args.nfeat, args.side)},
        'residual_chi2': {'kernel': 'k_forward_ftab', 'bound': 'valu (FP64)',
                          'achieved': bytes_r / tft / 1e9, 'unit': 'GB/s',
                          'avg_launch_ms': tft * 1e3, 'bytes_per_launch': bytes_r}}
    log(f'bench: normal_eq {tn * 1e3:.3f} ms {bytes_n / tn / 1e9:.0f} GB/s')
    traffic = committed_figure('k_jacobian_traffic', 'traffic_bytes_per_launch',
                               args.nfeat, args.side)
    counted = committed_figure('valu_counters', 'counted_fp32_flop_per_launch',
                               args.nfeat, args.side)
    lane_inst = committed_figure('valu_counters', 'valu_lane_instructions_per_launch',
                                 args.nfeat, args.side)
    traffic_ft = committed_figure('tabulated_passes', 'forward_hbm_bytes',
                                  args.nfeat, args.side)
    counted_ft = committed_figure('tabulated_passes', 'forward_counted_fp64_flop',
                                  args.nfeat, args.side)
    lane_ft = committed_figure('tabulated_passes', 'forward_lane_instructions',
                               args.nfeat, args.side)
    traffic_jt = committed_figure('tabulated_passes', 'jacobian_hbm_bytes',
                                  args.nfeat, args.side)
    lane_jt = committed_figure('tabulated_passes', 'jacobian_lane_instructions',
                               args.nfeat, args.side)
    log('bench: VALU issue-rate probe')
    probe = valu_probe()['rates']
    vop2 = probe['fp32_fma_lane']
    attainable = {
        'source': 'tests/native/lmz_valu_peak.hip, measured in this run',
        'fp32_fma_vop2_tflops': vop2['tflops'],
        'fp32_fma_vop3_tflops': probe['fp32_fma_uniform']['tflops'],
        'fp64_fma_tflops': probe['fp64_fma_lane']['tflops'],
        'fp32_vop2_lane_instructions_per_s': vop2['lane_instructions_per_s']}
    issue = (lane_inst / tj) if lane_inst else None
    issue_ft = (lane_ft / tft) if lane_ft else None
    fp64_lane = probe['fp64_fma_lane']['lane_instructions_per_s']
    attainable['fp64_lane_instructions_per_s'] = fp64_lane
    # the tabulated forward reads the data and writes the FP32 residual
    # (4 + 4 B / pixel), one FP64 chi2 partial per tile and its tables (48 B
    # a sample, read back once a feature's tile run); the tabulated Jacobian
    # reads the data and writes 36 FP64 partials per tile and its tables
    # (128 B a sample)
    bytes_ft = 8. * npx + 8. * lm.batch.ntiles + 2 * 48. * tables['forward_samples']
    bytes_jt = 4. * npx + 36 * 8. * lm.batch.ntiles + 2 * 128. * tables['jacobian_samples']
    fd = None
    fdb = None
    if not args.no_fd:
        log('bench: batched LM with a forward-difference Jacobian')
        fdb = run_fd_batched(lm, p0, settings, params, res)
        log(f"bench: batched FD {fdb['value']:.1f} fits/s, "
            f"{fdb['iterations_mean']:.1f} iterations, "
            f"{fdb['agree_with_adjoint']:.4f} agree with the adjoint fits")
        log('bench: plain GPU FD form sample')
        data_host = data_dev[: args.side * args.side * args.fd_sample].cpu().numpy()
        data_host = data_host.reshape(-1, args.side, args.side)
        rate, n, t = run_fd(args, truth, start, data_host)
        fd = dict(value=rate, unit='fits/s', sample=f'{n} features, sequential',
                  seconds=t)
        log(f'bench: FD form {rate:.2f} fits/s')
    cpu = None
    if not args.no_cpu:
        log('bench: CPU oracle sample')
        data_host = data_dev[: args.side * args.side * 4].cpu().numpy()
        data_host = data_host.reshape(-1, args.side, args.side)
        os.environ.setdefault('OMP_NUM_THREADS', '1')
        rate, n, t = run_cpu(args, start, data_host)
        cpu = dict(value=rate, unit='fits/s', cores=1, kind='port',
                   sample=f'{n} 7-parameter fit(s) of {args.side}^2 crops, '
                          f'FP64 numpy oracle + scipy MINPACK LM (forward '
                          f'differences), {t:.1f} s')
        log(f'bench: CPU oracle {rate:.4f} fits/s')
    cpu_pool = None
    workers = args.cpu_workers
    if workers < 0:
        workers = min(16, len(os.sched_getaffinity(0)))
    if workers > 0 and not args.no_cpu:
        args.cpu_workers = workers
        cpu_pool = run_cpu_pool(args, start, data_dev)
        log(f"bench: CPU oracle on {cpu_pool['cores']} cores "
            f"{cpu_pool['value']:.4f} fits/s")

    line = {
        'metric': '7-param sphere fits/sec at 201² crops, adjoint vs FD-Jacobian, 1 GPU',
        'value': fits_per_s,
        'unit': 'fits/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms_per_step,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64 forward/residual/chi2, f32 Jacobian with f64 tile sums, '
                 'f64 Mie and 7x7 solve',
        'data': 'synthetic (seeded config-3 parameter draws, sigma=0.05)',
        'config': {'workload': f'batched 7-parameter LM, {args.nfeat} features, '
                               f'{args.side}x{args.side} crops, <=50 iterations',
                   'features': args.nfeat, 'side': args.side,
                   'mean_nmax': nmean, 'ftol': args.ftol, 'xtol': args.xtol,
                   'gtol': args.gtol, 'schedule': args.schedule},
        'roofline': {'bound': 'valu', 'kernels': ['k_fwd_ftab', 'k_forward_ftab'],
                     'pass': 'trial forward (FP64 residual and chi2) from radial '
                             'tables of the sums',
                     'dtype': 'f64',
                     'achieved': ach_ft, 'peak': PEAK_FP64, 'unit': 'TFLOP/s',
                     'frac': ach_ft / PEAK_FP64,
                     'avg_launch_ms': tft * 1e3,
                     'algorithmic_flop_per_launch': flop_ft,
                     'pixels_per_launch': npx,
                     'table_samples_per_launch': tables['forward_samples'],
                     'traffic': traffic_ft,
                     'algorithmic_bytes_per_launch': bytes_ft,
                     'counted_fp64_flop_per_launch': counted_ft,
                     'counted_achieved': (counted_ft / tft / 1e12) if counted_ft else None,
                     'attainable': attainable,
                     'frac_of_attainable_counted': ((counted_ft / tft / 1e12) /
                                                    attainable['fp64_fma_tflops']
                                                    if counted_ft else None),
                     'valu_lane_instructions_per_launch': lane_ft,
                     'frac_of_attainable_issue': ((issue_ft / fp64_lane)
                                                  if issue_ft else None),
                     'order_loop': {'kernel': 'k_forward<2>',
                                    'avg_launch_ms': tf * 1e3,
                                    'achieved': ach_f, 'frac': ach_f / PEAK_FP64,
                                    'algorithmic_flop_per_launch': flop_f},
                     'order_loop_over_tabulated': tf / tft},
        'jacobian': {'kernels': ['k_jac_rtab', 'k_jacobian_rtab'],
                     'pass': 'Jacobian J^T J, J^T r from radial tables of the sums',
                     'dtype': 'f32 (f64 tile sums)', 'bound': 'valu',
                     'achieved': ach_jt, 'peak': PEAK_FP32, 'unit': 'TFLOP/s',
                     'frac': ach_jt / PEAK_FP32,
                     'avg_launch_ms': tt * 1e3,
                     'algorithmic_flop_per_launch': flop_jt,
                     'pixels_per_launch': npx,
                     'table_samples_per_launch': tables['jacobian_samples'],
                     'traffic': traffic_jt,
                     'algorithmic_bytes_per_launch': bytes_jt,
                     'valu_lane_instructions_per_launch': lane_jt,
                     'frac_of_attainable_issue': (
                         (lane_jt / tt / attainable['fp32_vop2_lane_instructions_per_s'])
                         if lane_jt else None),
                     'order_loop': {
                         'kernel': 'k_jacobian', 'bound': 'valu',
                         'avg_launch_ms': tj * 1e3,
                         'achieved': ach_j, 'peak': PEAK_FP32, 'unit': 'TFLOP/s',
                         'frac': ach_j / PEAK_FP32,
                         'algorithmic_flop_per_launch': flop_j,
                         'traffic': traffic,
                         'algorithmic_bytes_per_launch': bytes_j,
                         'counted_fp32_flop_per_launch': counted,
                         'valu_lane_instructions_per_launch': lane_inst,
                         'frac_of_attainable_issue': (
                             (issue / attainable['fp32_vop2_lane_instructions_per_s'])
                             if issue else None)},
                     'order_loop_over_tabulated': tj / tt},
        'reductions': reductions,
        'fit_quality': {'converged_frac': float(ok.mean()),
                        'iterations_mean': float(iters.mean()),
                        'iterations_max': float(iters.max()),
                        'iterations_p90': float(np.percentile(iters, 90)),
                        'jacobians_mean': float(res[:, 39].mean()),
                        'status_counts': {name: int(((status & bit) != 0).sum())
                                          for name, bit in (('ftol', 1), ('xtol', 2),
                                                            ('gtol', 4), ('maxiter', 8),
                                                            ('stalled', 16),
                                                            ('nonfinite', 32))},
                        'capped_redchi_median': (float(np.median(redchi[capped]))
                                                 if capped.any() else None),
                        'capped_far_from_truth': int((capped & far).sum()),
                        'converged_redchi_median': float(np.median(redchi[ok])),
                        'median_abs_err': {k: float(np.median(err[:, i]))
                                           for i, k in enumerate(PARAMS)}},
        'fit_agreement': parity,
        'convergence_check': agree,
        'fd_plain_gpu': fd,
        'speedup_vs_fd': (fits_per_s / fd['value']) if fd else None,
        'fd_batched': fdb,
        'speedup_vs_fd_batched': (fits_per_s / fdb['value']) if fdb else None,
        'cpu_baseline': cpu_pool or cpu,
        'cpu_baseline_1core': cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
