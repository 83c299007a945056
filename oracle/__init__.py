"""FP64 CPU oracle for the Lorenz-Mie hologram and its Levenberg-Marquardt fit.

TEST INFRASTRUCTURE ONLY.  This package restates, in float64 numpy, the
algorithms of pylorenzmie's numpy path so that the HIP kernels of
``pylorenzmie_amd`` can be checked against it.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import it, and only as the checker: the product path never calls it.

Modules
-------
lorenzmie
    wavenumber, n_max, Mie coefficients (homogeneous and layered), the
    per-pixel multipole sum, hologram with the scattered-field amplitude
    alpha, and a high-order finite-difference Jacobian.
fit
    The reference's fitter semantics (scipy MINPACK LM with the
    reference's default settings, residual scaling and SVD uncertainties).
cpu_pool
    ``fit`` over a sample of features on several host processes, the
    multi-core CPU baseline (started by bench.py as a child process).
normalize
    Background normalization of a frame (np.median, scipy's median filter,
    a reference image) and the running median of a video.
independent
    Textbook restatements that share no code path with ``lorenzmie``
    (Bessel functions from scipy.special, Legendre functions from
    scipy.special.lpmv, mpmath Mie coefficients) used to pin it.

Parity pins (what checks each function; see tests/test_oracle.py):

* wavenumber, wiscombe_yang, neves_pisignano -- the reference's own
  known-answer tests (tests/test_instrument.py:81-103,
  tests/test_sphere.py:169-185) plus golden vectors.
* mie_coefficients -- mpmath textbook Bohren-Huffman Eq. 4.53 at 30 digits,
  the lossless invariant |a_n - 1/2| = 1/2, spot values printed in the
  reference's docs/tutorials/Theory.ipynb, and golden vectors.
* mie_coefficients_layered -- reduces to the homogeneous result when all
  layers share one index; 2-layer transfer-matrix brute force in mpmath.
* lorenzmie / hologram -- an independent B&H Eq. 4.45-4.50 sum built from
  scipy.special.spherical_jn/yn and lpmv, mirror symmetries, a = b = 0
  and alpha = 0 giving I == 1, the alpha closed form, golden vectors.
* jacobian -- agreement of 5-point and 7-point stencils, and
  directional derivatives of the independent model.
* fit -- the reference Optimizer's golden fits (values, SVD
  uncertainties, reduced chi2 at default and tight tolerances) to 1e-9
  relative (identical in practice); a finite-difference step other than
  1e-5 or a reduced chi2 over npix misses them.
* normalize -- the filter against its definition (explicit reflected
  windows and an exact rank, median_filter_brute) at odd, even and
  frame-exceeding sizes; the result types against numpy's promotion of
  the reference's expression; VMedian against np.median of the frames it
  has seen (order 0) and the median of medians (order 1).
"""
