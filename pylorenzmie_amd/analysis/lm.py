"""Device-resident batched Levenberg-Marquardt (lmz_lm_run).

``LMBatch`` owns everything one batch of features needs on the GPU: the
feature table (dense crops or explicit coordinates), packed data, the LM
workspace and plan with its captured HIP graph.  ``run()`` fits every
feature from the given starting parameters; repeated runs reuse the graph.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from .. import _lmz
from ..theory.LorenzMie import PARAMS


@dataclass
class LMSettings:
    """Stopping rules of the on-device LM (see lmz_lm_settings)."""
    max_iter: int = 50
    ftol: float = 1e-9
    xtol: float = 1e-9
    gtol: float = 1e-10
    lambda0: float = 1e-3
    check_every: int = 4
    use_graph: bool = True
    fd_step: float = 0.0      # > 0: forward-difference Jacobian (transparency)
    damping: str = 'minpack'  # MINPACK's trust region, or 'nielsen'


def stats_from_jtj(jtj_upper: np.ndarray, chi2: float, npix: int,
                   vary: list[int]):
    """Reduced chi2 and uncertainties, analysis/Optimizer.py:262-278.

    The reference takes the SVD of J; J^T J = V S^2 V^T gives the same
    V and singular values, with the same cut s > eps max(P, n) s_0.
    """
    full = np.zeros((7, 7))
    iu = np.triu_indices(7)
    full[iu] = jtj_upper
    full = full + full.T - np.diag(np.diag(full))
    a = full[np.ix_(vary, vary)]
    nvar = len(vary)
    ndeg = npix - nvar
    redchi = chi2 / ndeg if ndeg > 0 else np.nan
    w, V = np.linalg.eigh(a)
    order = np.argsort(w)[::-1]
    w, V = w[order], V[:, order]
    s = np.sqrt(np.clip(w, 0., None))
    if s.size == 0 or s[0] == 0.:
        return redchi, np.full(nvar, np.nan)
    keep = s > np.finfo(float).eps * max(npix, nvar) * s[0]
    s, V = s[keep], V[:, keep]
    pcov = (V / s ** 2) @ V.T
    return redchi, np.sqrt(redchi * np.diag(pcov))


class LMBatch:
    """Features (crops or coordinate lists) fitted together on the GPU.

    ``crops``: list of dicts with keys
        data (ny, nx) or (npts,) float array,
        width (0 for explicit coordinates), origin (x, y, z),
        coords (3, npts) float array relative to origin when width == 0,
        noise, nmax_cap.
    ``frame``: (tensor, xs, ys) -- a float32 (ny, nx) frame in device memory
        with its column and row coordinates; the crops then give
        box (x0, y0, w, h), npix and, for width == 0, index (the crop-local
        row-major indices of their pixels) instead of data and coords, and
        lmz_gather_frame fills the batch on the device.
    """

    def __init__(self, crops: list[dict], instrument, flags: int = 3,
                 frame=None):
        t = _lmz.torch()
        dev = _lmz.device()
        feats = []
        off = 0
        datas, coords = [], []
        self.npix = []
        for c in crops:
            if frame is None:
                data = np.asarray(c['data'], np.float32).ravel()
                n = data.size
            else:
                data, n = None, int(c['npix'])
            feats.append(_lmz.make_feature(n, c.get('width', 0),
                                           c.get('origin', (0., 0., 0.)),
                                           instrument, c['nmax_cap'], off,
                                           noise=c.get('noise')))
            datas.append(data)
            if c.get('width', 0) == 0 and frame is None:
                coords.append(np.asarray(c['coords'], np.float32).reshape(3, n))
            else:
                coords.append(None)
            self.npix.append(n)
            off += n
        self.batch = _lmz.Batch(feats, 2)
        self.nfeat = len(crops)
        self.flags = flags
        if frame is not None:
            self._gather(crops, frame, off)
        else:
            self.data = t.from_numpy(np.concatenate(datas)).to(dev)
            if any(c is not None for c in coords):
                allc = np.zeros((3, off), np.float32)
                o = 0
                for c, n in zip(coords, self.npix):
                    if c is not None:
                        allc[:, o:o + n] = c
                    o += n
                self.coords = t.from_numpy(allc).to(dev)
            else:
                self.coords = None
        nbytes = self.batch.lm_workspace_bytes()
        self.workspace = t.empty(nbytes + 256, dtype=t.uint8, device=dev)
        base = self.workspace.data_ptr()
        self._ws_ptr = (base + 255) & ~255
        self.params = t.zeros((self.nfeat, _lmz.NPARAM), dtype=t.float64,
                              device=dev)
        self.result = t.zeros((self.nfeat, _lmz.NRES), dtype=t.float64,
                              device=dev)
        plan = ctypes.c_void_p()
        _lmz.check(_lmz.load_library().lmz_lm_plan_create(
            self.batch.handle, ctypes.byref(plan)))
        self.plan = plan

    def _gather(self, crops, frame, npix_total: int) -> None:
        """Data (and explicit coordinates) of every feature read from the
        device frame by lmz_gather_frame; only the boxes and the picked
        pixel indices travel from the host."""
        t = _lmz.torch()
        image, xs, ys = frame
        if image.dtype != t.float32 or image.dim() != 2 or not image.is_cuda:
            raise TypeError('frame must be a float32 (ny, nx) device tensor')
        image = image.contiguous()
        dev = image.device
        ny, nx = image.shape
        if len(xs) != nx or len(ys) != ny:
            raise ValueError('xs and ys must hold the frame\'s nx column and '
                             'ny row coordinates')
        box = np.zeros((self.nfeat, 4), np.int32)
        index = None
        explicit = any(c.get('width', 0) == 0 for c in crops)
        if explicit:
            index = np.zeros(npix_total, np.int32)
        o = 0
        for i, (c, n) in enumerate(zip(crops, self.npix)):
            x0, y0, w, h = (int(v) for v in c['box'])
            if not (0 <= x0 and x0 + w <= nx and 0 <= y0 and y0 + h <= ny):
                raise ValueError(f'box {c["box"]} is outside the {nx}x{ny} frame')
            box[i] = (x0, y0, w, h)
            if c.get('width', 0) == 0:
                idx = np.asarray(c['index'], np.int64)
                if idx.size != n or (n and (idx.min() < 0 or idx.max() >= w * h)):
                    raise ValueError('index must hold npix crop-local indices')
                index[o:o + n] = idx
            elif int(c['width']) != w or n != w * h:
                raise ValueError('a dense crop must cover its box')
            o += n
        self.data = t.empty(npix_total, dtype=t.float32, device=dev)
        self.coords = (t.empty((3, npix_total), dtype=t.float32, device=dev)
                       if explicit else None)
        self._box = t.from_numpy(box).to(dev)
        self._index = t.from_numpy(index).to(dev) if explicit else None
        self._xs = t.from_numpy(np.ascontiguousarray(xs, np.float64)).to(dev)
        self._ys = t.from_numpy(np.ascontiguousarray(ys, np.float64)).to(dev)
        _lmz.check(_lmz.load_library().lmz_gather_frame(
            self.batch.handle, _lmz.ptr(image), ny, nx, _lmz.ptr(self._xs),
            _lmz.ptr(self._ys), _lmz.ptr(self._box), _lmz.ptr(self._index),
            _lmz.ptr(self.data), _lmz.ptr(self.coords), _lmz.stream_handle()))

    def close(self):
        if getattr(self, 'plan', None):
            _lmz.load_library().lmz_lm_plan_destroy(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def lm_settings(settings: LMSettings, vary) -> _lmz.LMSettings:
        s = _lmz.LMSettings()
        s.max_iter = int(settings.max_iter)
        s.check_every = int(settings.check_every)
        s.ftol = float(settings.ftol)
        s.xtol = float(settings.xtol)
        s.gtol = float(settings.gtol)
        s.lambda0 = float(settings.lambda0)
        for i in range(_lmz.NJAC):
            s.vary[i] = 1 if (PARAMS[i] in vary) else 0
        s.use_graph = 1 if settings.use_graph else 0
        s.fd_step = float(settings.fd_step)
        rules = {'minpack': _lmz.DAMPING_MINPACK, 'nielsen': _lmz.DAMPING_NIELSEN}
        if settings.damping not in rules:
            raise ValueError(f"damping must be one of {sorted(rules)}")
        s.damping = rules[settings.damping]
        return s

    def load(self, params0: np.ndarray) -> None:
        """Starting parameters, (nfeat, 7 or 8) rows of PARAMS order."""
        t = _lmz.torch()
        p = np.zeros((self.nfeat, _lmz.NPARAM))
        p0 = np.asarray(params0, float)
        p[:, :p0.shape[1]] = p0
        self.params.copy_(t.from_numpy(p))

    def load_layers(self, layers_list) -> None:
        """Fixed Mie coefficients for every feature from layer tables
        ((L, 3) rows of outer radius [um], n, k per feature; one row for a
        homogeneous sphere), computed once by lmz_mie_ab_layered.  Runs
        then move x_p, y_p, z_p and alpha only (lmz_lm_run_fixed_ab)."""
        from ..theory.LayeredSphere import layer_table
        t = _lmz.torch()
        dev = self.params.device
        b = self.batch
        table, counts = layer_table(layers_list)
        lt = t.from_numpy(table).to(dev)
        nl = t.from_numpy(counts).to(dev)
        self.fixed_nmax = t.zeros(2 * self.nfeat, dtype=t.int32, device=dev)
        self.fixed_ab = t.zeros(b.coef_slots * 4, dtype=t.float64, device=dev)
        work = t.zeros(b.work_doubles, dtype=t.float64, device=dev)
        _lmz.check(_lmz.load_library().lmz_mie_ab_layered(
            b.handle, _lmz.ptr(lt), _lmz.ptr(nl), _lmz.ptr(self.fixed_nmax),
            None, _lmz.ptr(self.fixed_ab), _lmz.ptr(work), _lmz.stream_handle()))
        if int(self.fixed_nmax[0::2].min()) < 1:
            raise ValueError('invalid layer table')
        self.layer_tables = None

    def fit_layers(self, layers_list) -> None:
        """Layered particles whose outer radius a_p and outer index n_p +
        i k_p may be fitted (lmz_lm_run_layered): per feature an (L, 3)
        table of outer radius [um], n, k, innermost first.  At every trial
        the radii are the table's scaled by a_p / (its outer radius) --
        fixed radius fractions -- and the outer layer's index is n_p + i
        k_p; the inner indices are held.  Runs then vary any of the seven
        parameters, with the coefficient tangents from the device."""
        from ..theory.LayeredSphere import layer_table
        t = _lmz.torch()
        dev = self.params.device
        table, counts = layer_table(layers_list)
        if len(counts) != self.nfeat:
            raise ValueError(f'need {self.nfeat} layer tables')
        if np.any(counts < 1) or np.any(counts > _lmz.MAX_LAYERS):
            raise ValueError(f'1..{_lmz.MAX_LAYERS} layers per particle')
        self.layer_tables = (t.from_numpy(table).to(dev),
                             t.from_numpy(counts).to(dev))
        self.fixed_ab = None

    def autotune(self, force: bool = False):
        """Time the schedules of the pixel kernels on this batch at the
        loaded parameters and keep the fastest (lmz_autotune); returns it
        as a dict."""
        t = _lmz.torch()
        b = self.batch
        dev = self.params.device
        nmax = t.zeros(2 * self.nfeat, dtype=t.int32, device=dev)
        coef = t.zeros(b.coef_floats(), dtype=t.float32, device=dev)
        dcoef = t.zeros(b.dcoef_floats(), dtype=t.float32, device=dev)
        work = t.zeros(b.work_doubles, dtype=t.float64, device=dev)
        _lmz.check(_lmz.load_library().lmz_mie_ab(
            b.handle, _lmz.ptr(self.params), None, _lmz.ptr(nmax),
            _lmz.ptr(coef), _lmz.ptr(dcoef), None, _lmz.ptr(work),
            _lmz.stream_handle()))
        return b.autotune(self.params, nmax, coef, dcoef, self.data,
                          self.coords, self.flags, force).as_dict()

    def run(self, settings: LMSettings, vary, stream=None) -> None:
        """Fit on the device, stream-ordered; results stay in HBM."""
        s = self.lm_settings(settings, vary)
        lib = _lmz.load_library()
        layers = getattr(self, 'layer_tables', None)
        if layers is not None:
            _lmz.check(lib.lmz_lm_run_layered(
                self.plan, _lmz.ptr(self.coords), self.flags, _lmz.ptr(self.data),
                ctypes.byref(s), _lmz.ptr(layers[0]), _lmz.ptr(layers[1]),
                _lmz.ptr(self.params), _lmz.ptr(self.result), self._ws_ptr,
                _lmz.stream_handle(stream)))
            return
        if getattr(self, 'fixed_ab', None) is not None:
            _lmz.check(lib.lmz_lm_run_fixed_ab(
                self.plan, _lmz.ptr(self.coords), self.flags, _lmz.ptr(self.data),
                ctypes.byref(s), _lmz.ptr(self.fixed_ab), _lmz.ptr(self.fixed_nmax),
                _lmz.ptr(self.params), _lmz.ptr(self.result), self._ws_ptr,
                _lmz.stream_handle(stream)))
            return
        _lmz.check(lib.lmz_lm_run(
            self.plan, _lmz.ptr(self.coords), self.flags, _lmz.ptr(self.data),
            ctypes.byref(s), _lmz.ptr(self.params), _lmz.ptr(self.result),
            self._ws_ptr, _lmz.stream_handle(stream)))

    def fetch(self):
        """(params (nfeat, 8), result (nfeat, NRES)) as numpy."""
        return self.params.cpu().numpy(), self.result.cpu().numpy()
