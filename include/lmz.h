/*
 * lmz.h -- C ABI of the MI355X (gfx950) Lorenz-Mie hologram and
 * Levenberg-Marquardt library, liblmz.so.
 *
 * Conventions
 *   - Every compute entry point is stream-ordered on the hipStream_t it is
 *     given and takes DEVICE pointers for per-pixel and per-feature arrays.
 *     The caller owns every buffer; the library allocates only the device
 *     copy of a batch's feature table and tile map (lmz_batch_create) and a
 *     capture stream/graph inside an lmz_lm_plan.
 *   - Return value: 0 on success, LMZ_EINVAL for a bad argument (checked on
 *     the host before any launch), LMZ_EHIP for a HIP runtime error.  The
 *     message of the last failure on the calling thread is lmz_last_error().
 *   - No exceptions cross the ABI; calls on distinct streams are independent.
 *
 * Data layout in HBM
 *   params   double [nfeat][LMZ_NPARAM]  x_p y_p z_p [px], a_p [um], n_p, k_p,
 *                                         alpha, (unused)
 *   data     float  [npix_total]          normalized hologram pixels, features
 *                                         packed back to back (pix_offset)
 *   coords   float  [3][npix_total]       explicit coordinates x, y, z of
 *                                         features with width == 0, relative
 *                                         to the feature's origin; may be NULL
 *                                         when every feature is a dense crop
 *   field    float2 [3][npix_total]       complex field components
 *   jac      float  [LMZ_NJAC][npix_total] dI/d(param), same parameter order
 *   coef     float  [slots][coef_stride]  per order n: P = i E_n a_n and
 *                                         -Q = -E_n b_n (complex, with -Im P
 *                                         and Im Q beside them), n, n(n+1),
 *                                         the recurrence constants, and as
 *                                         doubles i^(n+1)(2n+1) a_n, P, Q
 *                                         (lmz_batch_sizes gives the strides;
 *                                         the layout is the library's own)
 *   dcoef    float  [slots][dcoef_stride] the tangents of those entries by
 *                                         a_p, n_p, k_p
 *   partials double [ntiles][36]          per pixel tile: upper triangle of
 *                                         J^T J (28, row major), J^T r (7),
 *                                         r^T r (1), with J and r scaled by
 *                                         1/noise
 */
#ifndef LMZ_H
#define LMZ_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LMZ_VERSION 104
#define LMZ_NPARAM 8       /* stride of params rows */
#define LMZ_NJAC 7         /* x_p y_p z_p a_p n_p k_p alpha */
#define LMZ_NRED 36        /* 28 + 7 + 1 */
#define LMZ_TILE 1024      /* pixels per tile (256 threads x 4 pixels) */
#define LMZ_MAX_WIDTH 8192 /* widest dense crop */

#define LMZ_OK 0
#define LMZ_EINVAL 1
#define LMZ_EHIP 2

/* flags */
#define LMZ_CARTESIAN 1    /* project onto x, y, z (theory/LorenzMie.py:331) */
#define LMZ_BOHREN 2       /* +z along propagation (theory/LorenzMie.py:249) */
#define LMZ_ACCUMULATE 4   /* field += instead of field = */
#define LMZ_FP32 8         /* lmz_hologram (hologram, or residuals with
                              data) in FP32 arithmetic: 1.6x faster,
                              errors up to 1e-4 + 2e-5 I (Cartesian only;
                              not for the LM; cupyLorenzMie's
                              double_precision = False,
                              theory/cupyLorenzMie.py:46-65) */

/* One particle and the pixels it is evaluated on.  Fields marked (out) are
 * written by lmz_batch_create into the caller's array. */
typedef struct lmz_feature {
    int64_t pix_offset;    /* first pixel in the packed pixel arrays */
    int32_t npix;          /* number of pixels (> 0) */
    int32_t width;         /* > 0: dense row-major crop of this width, pixel
                              (col, row) at origin + (col, row, 0);
                              0: explicit coordinates in coords[] */
    double origin[3];      /* dense: coordinates of pixel (0, 0);
                              explicit: offset added to coords[] */
    double n_m;            /* medium refractive index */
    double wavelength;     /* vacuum wavelength [um] */
    double magnification;  /* [um / px] */
    double noise;          /* residual scale (analysis/Optimizer.py:260) */
    int32_t nmax_cap;      /* orders reserved per coefficient buffer */
    int32_t coef_offset;   /* (out) first coefficient slot */
    int32_t tile_offset;   /* (out) first tile */
    int32_t ntiles;        /* (out) ceil(npix / LMZ_TILE) */
} lmz_feature;

typedef struct lmz_batch lmz_batch;

/* Launch shape of the pixel kernels of a batch.  A tile of LMZ_TILE pixels
 * is worked by jac_threads (fwd_threads) threads: 64 (one wavefront per
 * tile), 128 or 256; the LM's persistent launches use *_blocks_per_cu
 * workgroups per CU (0: as many as are resident at once).  Default
 * {256, 256, 0, 0}; lmz_autotune picks one by timing. */
typedef struct lmz_schedule {
    int32_t jac_threads;
    int32_t fwd_threads;
    int32_t jac_blocks_per_cu;
    int32_t fwd_blocks_per_cu;
} lmz_schedule;

typedef struct lmz_batch_sizes {
    int32_t nfeat;
    int32_t ntiles;
    int64_t npix_total;    /* max(pix_offset + npix) */
    int32_t coef_buffers;  /* 1, or 2 for double-buffered LM trial tables */
    int32_t coef_slots;    /* rows of coef / dcoef / ab_out */
    int32_t cap_max;       /* largest nmax_cap */
    int64_t work_doubles;  /* scratch of lmz_mie_ab */
    int32_t coef_stride;   /* floats per coef row */
    int32_t dcoef_stride;  /* floats per dcoef row */
} lmz_batch_sizes;

/* Validate a feature table on the host, fill its (out) fields, and upload it
 * with the tile -> feature map.  coef_buffers is 1 for plain evaluation or 2
 * for the double-buffered LM tables. */
int lmz_batch_create(int32_t nfeat, lmz_feature *feat, int32_t coef_buffers,
                     lmz_batch **out);
int lmz_batch_destroy(lmz_batch *batch);
int lmz_batch_query(const lmz_batch *batch, lmz_batch_sizes *sizes);

/* Mie coefficients a_n, b_n of homogeneous spheres in FP64 and their FP32
 * per-order tables for the pixel kernels.
 * Replaces Sphere.ab / Sphere.mie_coefficients (theory/Sphere.py:70-89,
 * 147-226) with n_max from Sphere.wiscombe_yang (theory/Sphere.py:91-123),
 * and folds in the per-order prefactors of theory/LorenzMie.py:297-303.
 *   sel      per-feature buffer index (0/1) or NULL for buffer 0
 *   nmax_out int32 [nfeat * coef_buffers], entry f*coef_buffers + buffer
 *   dcoef    NULL skips the a_p, n_p, k_p tangents
 *   ab_out   NULL or double [slots][4] = a_n.re a_n.im b_n.re b_n.im, n >= 1
 *   work     double [work_doubles] scratch */
int lmz_mie_ab(const lmz_batch *batch, const double *params,
               const int32_t *sel, int32_t *nmax_out, float *coef,
               float *dcoef, double *ab_out, double *work,
               hipStream_t stream);

/* Mie coefficients of layered spheres in FP64 (Yang's recursion) and the
 * FP32 per-order tables of buffer 0, no tangents.
 * Replaces the multilayer recursion of lib/obsolete/sphere_coefficients.py:
 * 126-215 (Yang, Appl. Opt. 42, 1710 (2003)) with n_max from Yang's Eq. 30
 * (theory/Sphere.py:91-123).
 *   layers   double [nfeat][LMZ_MAX_LAYERS][3]: outer radius [um], n, k of
 *            each layer, innermost first, radii increasing
 *   nlayers  int32 [nfeat], 1..LMZ_MAX_LAYERS
 *   nmax_out int32 [nfeat * coef_buffers], entry f*coef_buffers; -1 marks a
 *            feature whose layer table is invalid
 *   coef, ab_out  as lmz_mie_ab (either may be NULL, not both)
 *   work     double [work_doubles] scratch */
#define LMZ_MAX_LAYERS 8
int lmz_mie_ab_layered(const lmz_batch *batch, const double *layers,
                       const int32_t *nlayers, int32_t *nmax_out, float *coef,
                       double *ab_out, double *work, hipStream_t stream);

/* Layered spheres in the fitting parametrization: each feature's layer
 * table (as lmz_mie_ab_layered) with its radii scaled by params a_p / (the
 * table's outer radius) -- fixed radius fractions -- and the outer layer's
 * index replaced by params n_p + i k_p; inner indices held.  Writes the
 * coefficient tables of buffer 0 and, unless dcoef is NULL, the tangents
 * by a_p, n_p, k_p (forward-mode through Yang's recursion; the k_p tangent
 * is i times the n_p one, the coefficients being holomorphic in the outer
 * index), which lmz_jacobian and lmz_hologram_vjp use as for homogeneous
 * spheres.  nmax_out as lmz_mie_ab; 0 for a parameter row outside the
 * domain (a_p <= 0, n_p <= 0).  The layered extension of the coefficient
 * tangents Optimizer needs to fit a layered particle's size and index;
 * the reference has no working layered recursion
 * (lib/obsolete/sphere_coefficients.py:126-215 fails for layered input). */
int lmz_mie_ab_layered_fit(const lmz_batch *batch, const double *params,
                           const double *layers, const int32_t *nlayers,
                           int32_t *nmax_out, float *coef, float *dcoef,
                           double *work, hipStream_t stream);

/* Per-order tables from coefficients computed elsewhere (any particle
 * whose ab() is not a homogeneous sphere): ab_in double [slots][4] rows
 * n = 1..nmax_in[f] of each feature, as ab_out of lmz_mie_ab.
 * Replaces the per-order prefactors of theory/LorenzMie.py:297-303 applied
 * to Particle.ab() (theory/Particle.py:97-119). */
int lmz_coef_from_ab(const lmz_batch *batch, const double *ab_in,
                     const int32_t *nmax_in, int32_t *nmax_out, float *coef,
                     hipStream_t stream);

/* Scattered field of every feature's particle at its pixels, FP32.
 * Replaces LorenzMie.field / scattered_field / lorenzmie
 * (theory/LorenzMie.py:135-181, 183-342). */
int lmz_field(const lmz_batch *batch, const double *params,
              const int32_t *sel, const int32_t *nmax, const float *coef,
              const float *coords, int32_t flags, float *field,
              hipStream_t stream);

/* Hologram |x + alpha E_s|^2, or the residual (I - d) / noise when data is
 * not NULL, with per-tile r^T r in chi2_tiles (double [ntiles], may be NULL).
 * Replaces LorenzMie.hologram (theory/LorenzMie.py:113-133) and
 * Optimizer._residuals (analysis/Optimizer.py:257-260).
 * active: uint8 [nfeat] or NULL; inactive features are skipped. */
int lmz_hologram(const lmz_batch *batch, const double *params,
                 const int32_t *sel, const int32_t *nmax, const float *coef,
                 const float *coords, int32_t flags, const float *data,
                 float *out, double *chi2_tiles, const uint8_t *active,
                 hipStream_t stream);

/* Superposed holograms (or fields) of several particles per feature in
 * one pass over the pixels, FP64, with an optional spherical-aberration
 * mask per particle.
 *   part_offset  int32 [nfeat + 1]: particles of feature f are
 *                part_offset[f] .. part_offset[f + 1] - 1
 *   pparams      double [nparticles][LMZ_NPARAM]: x_p y_p z_p (absolute,
 *                any cluster origin added), a_p, n_p, k_p (unused here)
 *   pnmax        int32 [nparticles], orders of each particle's table
 *   pcoef        float [nparticles * pcap][coef_stride] per-order tables
 *                (particle i at rows i * pcap), as lmz_mie_ab or
 *                lmz_coef_from_ab write them for a batch of the particles
 *   spherical    aberration coefficient [wavelengths], 0 for none
 *   alpha        double [nfeat] or NULL (1)
 *   out          float [npix_total] hologram, or field float2 [3][npix_total]
 *                (alpha E): exactly one of the two
 * Replaces the particle loop of LorenzMie.field (theory/LorenzMie.py:
 * 160-181) with Cluster / Dimer particles (theory/Cluster.py:7-59,
 * theory/Dimer.py:7-78), and the mask of theory/Aberrated.py:61-84. */
int lmz_hologram_superposed(const lmz_batch *batch, int32_t nparticles,
                            const int32_t *part_offset, const double *pparams,
                            const int32_t *pnmax, const float *pcoef,
                            int32_t pcap, const float *coords, int32_t flags,
                            double spherical, const double *alpha, float *out,
                            float *field, hipStream_t stream);

/* Intensity of an accumulated field: out = |x + E|^2 per pixel, where the
 * field already carries alpha (lmz_field writes alpha E).  Used for
 * superposed particles (theory/LorenzMie.py:131-133 after :178-181). */
int lmz_intensity(const lmz_batch *batch, const double *params,
                  const float *field, int32_t flags, float *out,
                  hipStream_t stream);

/* Forward-mode Jacobian dI/d(x_p, y_p, z_p, a_p, n_p, k_p, alpha), FP32,
 * recomputing the recurrences and their tangents per pixel.  Any of jac,
 * resid, partials may be NULL (resid and partials need data).  Needs the
 * dcoef tangent tables of lmz_mie_ab.
 *   jac       float [LMZ_NJAC][npix_total], column c at c * npix_total
 *   resid     float [npix_total], (I - data) / noise
 *   partials  double [ntiles][LMZ_NRED], the normal equations of each tile
 *             (J scaled by 1/noise)
 * Replaces jaxLorenzMie.jac (theory/jaxLorenzMie.py:354-395) and
 * Optimizer._analytical_jac (analysis/Optimizer.py:251-255), extended to
 * k_p and alpha. */
int lmz_jacobian(const lmz_batch *batch, const double *params,
                 const int32_t *sel, const int32_t *nmax, const float *coef,
                 const float *dcoef, const float *coords, int32_t flags,
                 const float *data, float *jac, float *resid,
                 double *partials, const uint8_t *active,
                 hipStream_t stream);

/* Normal-equation partials as lmz_jacobian's (J scaled by 1/noise,
 * residuals from data), the way the batched LM forms them: on dense crops
 * (width > 0, every pixel in the plane z = 0) the multipole sums and their
 * tangents depend on the pixel only through its distance from the
 * particle's axis, so the order loop runs once per sample of a radial grid
 * and each pixel interpolates them; crops with explicit coordinates take
 * the order loop at every pixel.  Agrees with lmz_jacobian to FP32
 * rounding.  work: lmz_jacobian_tabulated_bytes, 256-byte aligned.
 * No counterpart in the reference: the Jacobian pass of lmz_lm_run, for
 * callers that want J^T J of dense crops without J in memory. */
int lmz_jacobian_tabulated_bytes(const lmz_batch *batch, size_t *bytes);
int lmz_jacobian_tabulated(const lmz_batch *batch, const double *params,
                           const int32_t *nmax, const float *coef,
                           const float *dcoef, int32_t flags,
                           const float *data, double *partials, void *work,
                           hipStream_t stream);

/* The FP64 residual (I - data) / noise and per-tile r^T r as lmz_hologram
 * with data, the way the batched LM forms them: on dense crops from FP64
 * radial tables of the multipole sums (one order loop per sample of the
 * distance from the particle's axis, 12-point interpolation per pixel),
 * elsewhere by the order loop at every pixel.  Agrees with lmz_hologram to
 * FP64 rounding.  resid may be NULL.  work: lmz_residual_tabulated_bytes,
 * 256-byte aligned.  No counterpart in the reference: the trial forward
 * pass of lmz_lm_run, for callers that evaluate chi2 of dense crops. */
int lmz_residual_tabulated_bytes(const lmz_batch *batch, size_t *bytes);
int lmz_residual_tabulated(const lmz_batch *batch, const double *params,
                           const int32_t *nmax, const float *coef,
                           int32_t flags, const float *data, float *resid,
                           double *chi2_tiles, void *work, hipStream_t stream);

/* Vector-Jacobian product of the hologram: grad_params[f][c] = sum over
 * feature f's pixels of grad_out[px] dI/d(param c), for the seven
 * parameters x_p y_p z_p a_p n_p k_p alpha, accumulated in FP64.  J is
 * formed per pixel in registers (the recurrences and their tangents are
 * recomputed, nothing is stored) and contracted at once.  The backward of
 * lmz_hologram for reverse-mode callers (a loss over many holograms);
 * coefficient buffer 0 is used.
 * Replaces the hologram derivatives that theory/jaxLorenzMie.py:280-286
 * and :354-395 take by jax.jacfwd, contracted with an upstream gradient.
 *   grad_out     float  [npix_total]      dL/dI
 *   tile_work    double [ntiles][LMZ_NJAC] scratch
 *   grad_params  double [nfeat][LMZ_NJAC] */
int lmz_hologram_vjp(const lmz_batch *batch, const double *params,
                     const int32_t *nmax, const float *coef,
                     const float *dcoef, const float *coords, int32_t flags,
                     const float *grad_out, double *tile_work,
                     double *grad_params, hipStream_t stream);

/* lmz_hologram_vjp without tile_work: the per-tile partial gradients go to
 * a fixed device array of LMZ_VJP_STATIC_MAX_TILES rows (3.5 MiB, one copy
 * per GPU, held until the process ends), so the call allocates nothing and
 * calls no allocating routine.  Same kernels and results as
 * lmz_hologram_vjp; batches with more tiles are rejected (LMZ_EINVAL).
 * Every call on a GPU shares that GPU's array: calls that overlap on one
 * GPU (two streams, two threads) race and must be serialized by the caller.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc). */
#define LMZ_VJP_STATIC_MAX_TILES 65536
int lmz_hologram_vjp_static(const lmz_batch *batch, const double *params,
                            const int32_t *nmax, const float *coef,
                            const float *dcoef, const float *coords,
                            int32_t flags, const float *grad_out,
                            double *grad_params, hipStream_t stream);

/* Launch shape of a batch's pixel kernels (see lmz_schedule). */
int lmz_batch_set_schedule(lmz_batch *batch, const lmz_schedule *schedule);
int lmz_batch_get_schedule(const lmz_batch *batch, lmz_schedule *schedule);

/* Autotuner: times the Jacobian kernel (with its J^T J reduction) and the
 * residual kernel as the LM launches them (persistent grid over a
 * device-built tile list), over every tile and over the tiles of every
 * 8th feature (the LM's later iterations; weighted 8x), for 64, 128 and
 * 256 threads per tile and 1, 2 and 4 times the resident workgroups per
 * CU; keeps the fastest of each in the batch's schedule and returns it.
 * Results are cached per process (behind a mutex) by device, tile count,
 * mean order count and coordinate layout; LMZ_TUNE_FORCE re-times.
 * Synchronizes the stream.  Inputs as lmz_jacobian with data; scratch is a
 * 256-byte aligned device buffer of lmz_autotune_scratch_bytes.
 * Has no counterpart in the reference (its kernels have fixed shapes,
 * theory/cupyLorenzMie.py:73-75); it serves every caller of the batched
 * kernels. */
#define LMZ_TUNE_FORCE 1
int lmz_autotune_scratch_bytes(const lmz_batch *batch, size_t *bytes);
int lmz_autotune(lmz_batch *batch, const double *params, const int32_t *nmax,
                 const float *coef, const float *dcoef, const float *coords,
                 int32_t flags, const float *data, void *scratch,
                 int32_t options, lmz_schedule *chosen, hipStream_t stream);

/* Normal equations from a Jacobian and residual already in HBM:
 * out double [nfeat][36] (J^T J upper triangle, J^T r, r^T r), with J
 * scaled by 1/noise.  The reduction MINPACK performs inside lmder
 * (scipy least_squares, called at analysis/Optimizer.py:221). */
int lmz_normal_eq(const lmz_batch *batch, const float *jac,
                  const float *resid, double *out, hipStream_t stream);

/* Batched Levenberg-Marquardt over every feature of the batch
 * (coef_buffers must be 2).  Replaces Optimizer.optimize
 * (analysis/Optimizer.py:196-222) -> scipy least_squares(method='lm')
 * with an on-device Marquardt iteration: forward-mode Jacobian, FP64 7x7
 * Cholesky solve, trial forward, accept/reject, one HIP graph per
 * iteration replayed up to max_iter times. */
typedef struct lmz_lm_settings {
    int32_t max_iter;      /* iteration cap (accepted + rejected steps) */
    int32_t check_every;   /* replays between host checks for completion */
    double ftol;           /* stop: actual and predicted relative chi2
                              reduction both <= ftol */
    double xtol;           /* stop: ||D dp|| <= xtol (||D p|| + xtol) */
    double gtol;           /* stop: max_j |J_j^T r| / (|J_j| |r|) <= gtol */
    double lambda0;        /* initial damping relative to diag(J^T J) */
    uint8_t vary[LMZ_NJAC];/* 1 = fitted, 0 = held fixed */
    uint8_t use_graph;     /* capture the iteration in a HIP graph */
    double fd_step;        /* 0: analytic Jacobian (lmz_jacobian's kernel);
                              > 0: forward differences of the FP64 hologram
                              with steps fd_step |p_j| (fd_step if p_j = 0),
                              MINPACK lmdif's rule for scipy's diff_step
                              (analysis/Optimizer.py:111) */
    int32_t damping;       /* LMZ_DAMPING_MINPACK or LMZ_DAMPING_NIELSEN */
    int32_t reserved;      /* 0 */
} lmz_lm_settings;

/* Damping rules of lmz_lm_settings.damping.
 * LMZ_DAMPING_MINPACK: the step rule of MINPACK lmder/lmdif, which scipy
 *   least_squares(method='lm') runs for Optimizer.optimize
 *   (analysis/Optimizer.py:221): a trust region of radius delta in the
 *   scaled norm ||D dp|| (D from the running maximum of the Jacobian's
 *   column norms, x_scale='jac'), delta = 100 ||D p0|| at the start, the
 *   Levenberg parameter found by lmpar's Newton iteration, radius updates
 *   from the ratio of actual to predicted reduction, steps accepted when
 *   that ratio is >= 1e-4; ftol, xtol and gtol as MINPACK tests them
 *   (relative chi2 reductions <= ftol, delta <= xtol ||D p||).  lambda0 is
 *   not used.
 * LMZ_DAMPING_NIELSEN: Marquardt damping lambda diag(D^2) from lambda0,
 *   Nielsen's update on accepted steps and doubling growth on rejections;
 *   ftol on the chi2 reductions of accepted steps, xtol on ||D dp|| <=
 *   xtol (||D p|| + xtol). */
#define LMZ_DAMPING_MINPACK 0
#define LMZ_DAMPING_NIELSEN 1

/* per-feature LM result, double [nfeat][LMZ_NRES] */
#define LMZ_NRES 40
/* 0..27 J^T J at the solution (upper triangle), 28..34 J^T r, 35 chi2,
   36 lambda, 37 iterations, 38 status, 39 Jacobian evaluations */

/* status codes */
#define LMZ_LM_RUNNING 0
#define LMZ_LM_FTOL 1
#define LMZ_LM_XTOL 2
#define LMZ_LM_GTOL 4
#define LMZ_LM_MAXITER 8
#define LMZ_LM_STALLED 16
#define LMZ_LM_NONFINITE 32

typedef struct lmz_lm_plan lmz_lm_plan;

/* Bytes of device workspace lmz_lm_run needs for this batch. */
int lmz_lm_workspace_bytes(const lmz_batch *batch, size_t *bytes);
int lmz_lm_plan_create(const lmz_batch *batch, lmz_lm_plan **out);
int lmz_lm_plan_destroy(lmz_lm_plan *plan);

/* params_io: double [nfeat][LMZ_NPARAM] device, start in / solution out.
 * result:    double [nfeat][LMZ_NRES] device.
 * workspace: device buffer of lmz_lm_workspace_bytes. */
int lmz_lm_run(lmz_lm_plan *plan, const float *coords, int32_t flags,
               const float *data, const lmz_lm_settings *settings,
               double *params_io, double *result, void *workspace,
               hipStream_t stream);

/* As lmz_lm_run for particles whose Mie coefficients do not move during
 * the fit: layered spheres (lmz_mie_ab_layered), or any particle with a_p,
 * n_p, k_p held.  ab: double [coef_slots][4], rows n = 1..nmax of each
 * feature at its buffer-0 slots, as lmz_mie_ab / lmz_mie_ab_layered write
 * them for this batch; nmax: int32 [nfeat * coef_buffers], entry
 * f*coef_buffers.  settings->vary must hold a_p, n_p and k_p fixed.
 * Replaces Optimizer.optimize (analysis/Optimizer.py:196-222) for such
 * particles, whose ab() the reference re-evaluates on every residual
 * (theory/LorenzMie.py:156). */
int lmz_lm_run_fixed_ab(lmz_lm_plan *plan, const float *coords, int32_t flags,
                        const float *data, const lmz_lm_settings *settings,
                        const double *ab, const int32_t *nmax,
                        double *params_io, double *result, void *workspace,
                        hipStream_t stream);

/* As lmz_lm_run for layered spheres whose outer radius a_p and outer index
 * n_p + i k_p may be fitted, in lmz_mie_ab_layered_fit's parametrization
 * (radius fractions and inner indices from layers / nlayers, device
 * arrays as lmz_mie_ab_layered takes them): the coefficients and their
 * tangents are recomputed at every trial; with settings->fd_step > 0 the
 * a_p, n_p, k_p columns difference the layered coefficients instead.
 * Replaces Optimizer.optimize (analysis/Optimizer.py:196-222) for layered
 * particles, whose ab() the reference re-evaluates on every residual
 * (theory/LorenzMie.py:156). */
int lmz_lm_run_layered(lmz_lm_plan *plan, const float *coords, int32_t flags,
                       const float *data, const lmz_lm_settings *settings,
                       const double *layers, const int32_t *nlayers,
                       double *params_io, double *result, void *workspace,
                       hipStream_t stream);

/* ------------------------------------------------------------------ *
 * Differential-evolution estimates of many features, on the device.
 * ------------------------------------------------------------------ */

#define LMZ_DE_MAX_DIM 7
#define LMZ_DE_MAX_POP 1024

typedef struct lmz_de_settings {
    int32_t ndim;           /* searched parameters, 1..LMZ_DE_MAX_DIM */
    int32_t cols[LMZ_DE_MAX_DIM]; /* their columns in a params row (0..6) */
    int32_t npop;           /* members per search, 5..LMZ_DE_MAX_POP
                               (scipy: max(5, popsize * ndim)) */
    int32_t maxiter;        /* generations after the initial population */
    int32_t check_every;    /* generations between host reads of the
                               number of converged searches */
    int32_t reserved;
    double tol, atol;       /* stop a search when std(E) <= atol + tol |mean(E)| */
    double mutation_lo, mutation_hi;  /* dither range of the scale */
    double recombination;   /* crossover probability */
    uint64_t seed;
} lmz_de_settings;

/* Device workspace of lmz_de_run (256-byte aligned). */
int lmz_de_workspace_bytes(const lmz_batch *batch,
                           const lmz_de_settings *settings, size_t *bytes);

/* nsearch = nfeat / npop differential-evolution searches advanced together,
 * each the search scipy's differential_evolution runs for one feature in
 * the reference's DEEstimator.estimate (analysis/DEEstimator.py:119-173:
 * strategy best1bin, deferred updating, Latin-hypercube start, no polish):
 * every generation is one trial kernel, the Mie and residual launches of
 * all members of all unconverged searches, and one selection kernel; the
 * host only reads the count of converged searches every check_every
 * generations.  The batch holds the searches feature-major: features
 * s*npop .. s*npop + npop - 1 are search s's members, each a copy of that
 * search's pixels.  base: double [nsearch][LMZ_NPARAM], the parameters a
 * search does not set; bounds: double [nsearch][ndim][2] (lo, hi); the
 * energy is chi2 = sum ((I - d) / noise)^2 over the member's pixels
 * (flags as lmz_hologram's; LMZ_FP32 allowed, the search only ranks).
 * Outputs (device): best double [nsearch][ndim] (the member of lowest
 * energy, in parameter units), energy double [nsearch], generations
 * int32 [nsearch] (the generation whose selection met the stopping test,
 * else maxiter).  Random draws are counter-based: a seed gives the same
 * searches on any device and batch composition. */
int lmz_de_run(const lmz_batch *batch, const lmz_de_settings *settings,
               const double *base, const double *bounds, const float *coords,
               int32_t flags, const float *data, double *best, double *energy,
               int32_t *generations, void *workspace, hipStream_t stream);

/* ------------------------------------------------------------------ *
 * Frames resident in device memory: background normalization and the
 * crops that feed a batch.
 * ------------------------------------------------------------------ */

/* pixel types of frames, backgrounds and normalized output */
#define LMZ_U8 0
#define LMZ_U16 1
#define LMZ_F32 2
#define LMZ_F64 3
#define LMZ_MAX_FILTER_SIZE 255   /* widest median-filter window */

/* Bytes of scratch lmz_median needs (8-byte aligned device memory). */
size_t lmz_median_scratch_bytes(void);

/* numpy.median of n pixels of type dtype, written as a double to the device
 * scalar *median: the mean of the two middle values for even n (float32
 * arithmetic for LMZ_F32, float64 otherwise), NaN when a pixel is NaN or
 * n == 0.  Exact radix select in device memory; no host synchronization.
 * Replaces np.median(image) of utilities/normalizer.py:74. */
int lmz_median(const void *image, int32_t dtype, int64_t n, double *median,
               void *scratch, hipStream_t stream);

/* scipy.ndimage.median_filter(image, size=size) with its default 'reflect'
 * boundary on an (ny, nx) row-major frame: the element of rank
 * size*size // 2 of the size x size window whose offsets run from -(size//2)
 * to size - size//2 - 1.  Writes the background (type dtype) and/or
 * out = image / where(background > 0, background, 1.) as LMZ_F32 or
 * LMZ_F64 (the exact quotient rounded once).  1 <= size <=
 * LMZ_MAX_FILTER_SIZE.  Replaces utilities/normalizer.py:71-72 and 75-76. */
int lmz_median_filter(const void *image, int32_t dtype, int32_t ny, int32_t nx,
                      int32_t size, void *background, void *out,
                      int32_t out_dtype, hipStream_t stream);

/* out = image / where(b > 0, b, 1.) over n pixels, b being the device scalar
 * *background_scalar (from lmz_median), else the image reference[] of type
 * ref_dtype, else the host scalar ref_scalar.  The quotient of the exact
 * operands rounded once to out_dtype, which equals numpy's result in its
 * result type.  Replaces utilities/normalizer.py:64-76 ('median' and
 * 'reference'). */
int lmz_normalize(const void *image, int32_t dtype, int64_t n,
                  const double *background_scalar, const void *reference,
                  int32_t ref_dtype, double ref_scalar, void *out,
                  int32_t out_dtype, hipStream_t stream);

/* Pixelwise median of three 8-bit frames frames[3][n] into out[n]: one
 * level of the running background median of utilities/vmedian.py:61-75. */
int lmz_median3_u8(const uint8_t *frames, int64_t n, uint8_t *out,
                   hipStream_t stream);

/* Fill a batch's packed data (and explicit coordinates) from an (ny, nx)
 * float32 frame in device memory, without host copies.
 *   xs, ys double [nx], [ny] device: the frame's column and row
 *          coordinates (lib/meshgrid.py:10-40); needed only for
 *          explicit-coordinate features
 *   box    int32 [nfeat][4] device: x0, y0, w, h of each feature's crop in
 *          the frame (the Hologram slice of analysis/Hologram.py:92-108)
 *   index  int32 [npix_total] device, for explicit-coordinate features:
 *          the crop-local row-major index of each pixel (the pixels a Mask
 *          picks, analysis/Mask.py:82-105), NULL if every feature is dense
 *   coords float [3][npix_total] (out) for explicit-coordinate features:
 *          (xs[x], ys[y], 0) - origin, rounded once to float
 * A dense feature's width must equal its box width.  Pixels outside the
 * frame or the box read as NaN. */
int lmz_gather_frame(const lmz_batch *batch, const float *frame, int32_t ny,
                     int32_t nx, const double *xs, const double *ys,
                     const int32_t *box, const int32_t *index, float *data,
                     float *coords, hipStream_t stream);

const char *lmz_last_error(void);
int lmz_version(void);

#ifdef __cplusplus
}
#endif
#endif /* LMZ_H */
