// Batched Levenberg-Marquardt on gfx950: every feature of a batch is
// fitted at once; one iteration is six launches captured in a HIP graph.
//
// Replaces Optimizer.optimize (analysis/Optimizer.py:196-222), which hands
// one feature at a time to scipy least_squares(method='lm') (MINPACK lmdif,
// with a forward-difference Jacobian), by a Marquardt iteration whose
// Jacobian, trial forward pass and normal-equation reductions stay on the
// device.  Per iteration and feature:
//   J^T J, J^T r   from the forward-mode Jacobian (skipped after a rejection)
//   solve          (J^T J + par D^2) dp = -J^T r, FP64 Cholesky, par from
//                  MINPACK's trust-region rule (or Marquardt-Nielsen)
//   trial          a_n, b_n and chi2 at p + dp
//   accept/reject  radius / damping update; ftol / xtol / gtol tests
// The per-feature step logic is in lmz_lmstep.h.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "lmz_device.h"
#include "lmz_lmstep.h"

namespace lmz {

struct LMLayout {
    size_t coef, dcoef, nmax, sel, work, partials, chi2_tiles, state, ptrial,
        mask_j, mask_f, count, resid, list_j, list_f, tile_start, fd_h,
        fd_params, fd_nmax,
        fd_coef, rpar, rtab, mask_jt, mask_jl, list_jl, fpar, ftab, mask_ft,
        mask_fl, list_fl, total;
};

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

static LMLayout lm_layout(const lmz_batch *b) {
    LMLayout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off = align256(off + bytes); return o; };
    L.coef = take(sizeof(float) * LMZ_COEF_STRIDE * (size_t)b->coef_slots);
    L.dcoef = take(sizeof(float) * LMZ_DCOEF_STRIDE * (size_t)b->coef_slots);
    L.nmax = take(sizeof(int32_t) * 2 * (size_t)b->nfeat);
    L.sel = take(sizeof(int32_t) * (size_t)b->nfeat);
    L.work = take(sizeof(double) * LMZ_WORK_PER_ORDER * LMZ_WORK_PASSES *
                  (size_t)(b->cap_max + 1) *
                  b->nfeat);
    L.partials = take(sizeof(double) * LMZ_NRED * (size_t)b->ntiles);
    L.chi2_tiles = take(sizeof(double) * (size_t)b->ntiles);
    L.state = take(sizeof(LMState) * (size_t)b->nfeat);
    L.ptrial = take(sizeof(double) * LMZ_NPARAM * (size_t)b->nfeat);
    L.mask_j = take((size_t)b->nfeat);
    L.mask_f = take((size_t)b->nfeat);
    L.count = take(sizeof(int32_t) * 80);   // active count, four TileLists
    L.resid = take(sizeof(float) * 2 * (size_t)b->npix_total);
    L.list_j = take(sizeof(int32_t) * (size_t)b->ntiles);
    L.list_f = take(sizeof(int32_t) * (size_t)b->ntiles);
    L.tile_start = take(sizeof(int32_t) * (size_t)b->nfeat);   // both lists
    // forward-difference Jacobian (settings.fd_step > 0)
    L.fd_h = take(sizeof(double) * LMZ_NPARAM * (size_t)b->nfeat);
    L.fd_params = take(sizeof(double) * 3 * LMZ_NPARAM * (size_t)b->nfeat);
    L.fd_nmax = take(sizeof(int32_t) * 3 * 2 * (size_t)b->nfeat);
    L.fd_coef = take(sizeof(float) * 3 * LMZ_COEF_STRIDE * (size_t)b->coef_slots);
    // radial tables of the Jacobian's sums (dense crops)
    L.rpar = take(sizeof(RTabPar) * (size_t)b->nfeat);
    L.rtab = take(sizeof(float) * LMZ_RTAB_CAP * LMZ_RTAB_STRIDE * (size_t)b->nfeat);
    L.mask_jt = take((size_t)b->nfeat);    // tabulated / order-loop features
    L.mask_jl = take((size_t)b->nfeat);
    L.list_jl = take(sizeof(int32_t) * (size_t)b->ntiles);
    // FP64 radial tables of the trial forward pass (dense crops)
    L.fpar = take(sizeof(FTabPar) * (size_t)b->nfeat);
    L.ftab = take(sizeof(double) * LMZ_FTAB_CAP * LMZ_FTAB_STRIDE * (size_t)b->nfeat);
    L.mask_ft = take((size_t)b->nfeat);
    L.mask_fl = take((size_t)b->nfeat);
    L.list_fl = take(sizeof(int32_t) * (size_t)b->ntiles);
    L.total = off;
    return L;
}

__global__ __launch_bounds__(64) void k_lm_init(
    int nfeat, LMState *__restrict__ st, const double *__restrict__ params,
    double *__restrict__ ptrial, int32_t *__restrict__ sel,
    uint8_t *__restrict__ mask_j, uint8_t *__restrict__ mask_f,
    double lambda0) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfeat) return;
    LMState s;
    lm_init(s, lambda0);
    st[f] = s;
    for (int i = 0; i < LMZ_NPARAM; ++i)
        ptrial[(long long)LMZ_NPARAM * f + i] = params[(long long)LMZ_NPARAM * f + i];
    sel[f] = 0;
    mask_j[f] = 1;
    mask_f[f] = 1;
}

// Sum a feature's per-tile doubles (stride NV) over its tiles: one wave.
template <int NV>
__device__ __forceinline__ void feature_tile_sum(const lmz_feature &F,
                                                 const double *__restrict__ tiles,
                                                 double (&out)[NV]) {
    const int lane = threadIdx.x & 63;
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = 0.0;
    for (int t = lane; t < F.ntiles; t += 64) {
        const double *row = tiles + (long long)NV * (F.tile_offset + t);
#pragma unroll
        for (int i = 0; i < NV; ++i) acc[i] += row[i];
    }
#pragma unroll
    for (int i = 0; i < NV; ++i) out[i] = wave_sum_d(acc[i]);
}

__global__ __launch_bounds__(64) void k_lm_init_chi2(
    int nfeat, const lmz_feature *__restrict__ feat,
    const double *__restrict__ chi2_tiles, LMState *__restrict__ st) {
    const int f = blockIdx.x;
    double c[1];
    feature_tile_sum<1>(feat[f], chi2_tiles, c);
    if (threadIdx.x == 0) {
        st[f].chi2 = c[0];
        if (!isfinite(c[0])) st[f].status = LMZ_LM_NONFINITE;
    }
}

__global__ __launch_bounds__(64) void k_lm_mask_init(
    int nfeat, const LMState *__restrict__ st, uint8_t *__restrict__ mask_j,
    uint8_t *__restrict__ mask_f) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfeat) return;
    const bool running = st[f].status == LMZ_LM_RUNNING;
    mask_j[f] = running ? 1 : 0;
    mask_f[f] = running ? 1 : 0;
}

// Reduce the Jacobian partials (if fresh), then solve for the trial step.
__global__ __launch_bounds__(64) void k_lm_solve(
    const lmz_feature *__restrict__ feat, const double *__restrict__ partials,
    LMState *__restrict__ st, const double *__restrict__ params,
    double *__restrict__ ptrial, uint8_t *__restrict__ mask_j,
    uint8_t *__restrict__ mask_f, lmz_lm_settings set) {
    const int f = blockIdx.x;
    if (!mask_f[f]) return;
    const lmz_feature &F = feat[f];
    const double *p = params + (long long)LMZ_NPARAM * f;
    const bool fresh = mask_j[f];
    double red[LMZ_NRED];
    if (fresh) feature_tile_sum<LMZ_NRED>(F, partials, red);
    if (threadIdx.x != 0) return;
    LMState S = st[f];                  // registers; written back once
    if (fresh) lm_take_jacobian(S, red, set, p);
    // gradient test (analysis/Optimizer.py:108 gtol), on a fresh Jacobian
    if (fresh && set.gtol > 0.0 && lm_gtol_met(S, set)) {
        S.status |= LMZ_LM_GTOL;
        mask_f[f] = 0;
        mask_j[f] = 0;
    } else if (!lm_solve(S, set, p, ptrial + (long long)LMZ_NPARAM * f)) {
        S.status |= LMZ_LM_STALLED;
        mask_f[f] = 0;
        mask_j[f] = 0;
    }
    st[f] = S;
}

// Compare the trial chi2 with the current one; update damping and status.
__global__ __launch_bounds__(64) void k_lm_accept(
    const lmz_feature *__restrict__ feat,
    const double *__restrict__ chi2_tiles, LMState *__restrict__ st,
    double *__restrict__ params, const double *__restrict__ ptrial,
    int32_t *__restrict__ sel, uint8_t *__restrict__ mask_j,
    uint8_t *__restrict__ mask_f, lmz_lm_settings set) {
    const int f = blockIdx.x;
    if (!mask_f[f]) return;
    double c[1];
    feature_tile_sum<1>(feat[f], chi2_tiles, c);
    if (threadIdx.x != 0) return;
    LMState S = st[f];
    const double *pt = ptrial + (long long)LMZ_NPARAM * f;
    if (lm_accept(S, set, c[0], pt)) {
        for (int i = 0; i < LMZ_NPARAM; ++i)
            params[(long long)LMZ_NPARAM * f + i] = pt[i];
        sel[f] ^= 1;
    }
    st[f] = S;
    const bool running = (S.status == LMZ_LM_RUNNING);
    mask_f[f] = running ? 1 : 0;
    mask_j[f] = (running && S.need_j) ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_lm_count(int nfeat,
                                                  const uint8_t *__restrict__ mask_f,
                                                  int32_t *__restrict__ count) {
    __shared__ int part[256];
    int s = 0;
    for (int f = threadIdx.x; f < nfeat; f += 256) s += mask_f[f];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) count[0] = part[0];
}

// Results: J^T J and J^T r of the last Jacobian evaluation (at the solution,
// or one converged step before it), chi2 at the solution, and the status.
__global__ __launch_bounds__(64) void k_lm_finalize(
    int nfeat, const LMState *__restrict__ st, double *__restrict__ result) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfeat) return;
    const LMState &S = st[f];
    double *r = result + (long long)LMZ_NRES * f;
    for (int i = 0; i < 28; ++i) r[i] = S.jtj[i];
    for (int i = 0; i < 7; ++i) r[28 + i] = S.g[i];
    r[35] = S.chi2;
    r[36] = S.lambda;
    r[37] = (double)S.iter;
    r[38] = (double)S.status;
    r[39] = (double)S.njev;
}

}  // namespace lmz

using namespace lmz;

struct lmz_lm_plan {
    const lmz_batch *batch = nullptr;
    int grid_j = 0, grid_f = 0;       // persistent grids of the pixel kernels
    int grid_t = 0, grid_ft = 0;      // and of k_jacobian_rtab, k_forward_ftab
    hipStream_t s = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    int32_t *h_count = nullptr;       // pinned
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // arguments the captured graph was built for
    const void *key_ptrs[7] = {nullptr, nullptr, nullptr, nullptr, nullptr,
                               nullptr, nullptr};
    int32_t key_flags = -1;
    int key_fixed = -1;
    lmz_lm_settings key_set{};
    lmz_schedule key_sched{};
};

static void plan_release_graph(lmz_lm_plan *p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
}

// One LM iteration: the six launches a graph replays.
// Coefficient tables of both buffers from caller-supplied FP64 a_n, b_n
// (rows at each feature's buffer-0 slots), with zero tangents: the
// coefficients stay fixed during the fit.  One lane per feature.
__global__ __launch_bounds__(64) void k_lm_load_ab(
    const lmz_feature *__restrict__ feat, int nfeat,
    const double *__restrict__ ab, const int32_t *__restrict__ nmax_in,
    int32_t *__restrict__ nmax, float *__restrict__ coef,
    float *__restrict__ dcoef) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfeat) return;
    const lmz_feature F = feat[f];
    int N = nmax_in[2LL * f];
    N = N < 0 ? 0 : (N > F.nmax_cap ? F.nmax_cap : N);
    for (int b = 0; b < 2; ++b) {
        const long long slot0 = F.coef_offset + (long long)b * F.nmax_cap;
        for (int n = 1; n <= N; ++n) {
            const double *r = ab + 4LL * (F.coef_offset + n - 1);
            write_order(n, cd{r[0], r[1]}, cd{r[2], r[3]},
                        coef + LMZ_COEF_STRIDE * (slot0 + n - 1));
            float *drow = dcoef + LMZ_DCOEF_STRIDE * (slot0 + n - 1);
            for (int i = 0; i < LMZ_DCOEF_STRIDE; ++i) drow[i] = 0.f;
        }
        nmax[2LL * f + b] = N;
    }
}

// Where a trial's Mie coefficients come from: recomputed for homogeneous
// spheres (k_mie), held (fixed_ab), or recomputed for layered spheres in
// their fitting parametrization (layers, nlayers).
struct MieSource {
    bool fixed_ab;
    const double *layers;
    const int32_t *nlayers;
};

// The FP64 residual pass at params (coefficient buffer sel[f] ^ flip) over
// the features in mask (all when null): residuals into resid's buffer, r^T r
// per tile into chi2_tiles.  Dense crops from FP64 radial tables of the sums
// (k_forward_ftab), the rest by the order loop at every pixel (k_forward<2>);
// all by the order loop with FP32 residuals, and with a forward-difference
// Jacobian, whose differences take r at the base point from the same
// kernel as at the displaced ones.
static void enqueue_forward(const lmz_lm_plan *plan, const LMLayout &L,
                            char *ws, const float *coords, int32_t flags,
                            const float *data, const lmz_lm_settings &set,
                            const double *params, int flip,
                            const uint8_t *mask, hipStream_t s) {
    const lmz_batch *b = plan->batch;
    const int32_t *nmax = reinterpret_cast<const int32_t *>(ws + L.nmax);
    const int32_t *sel = reinterpret_cast<const int32_t *>(ws + L.sel);
    const float *coef = reinterpret_cast<const float *>(ws + L.coef);
    double *chi2_tiles = reinterpret_cast<double *>(ws + L.chi2_tiles);
    int32_t *count = reinterpret_cast<int32_t *>(ws + L.count);
    float *resid = reinterpret_cast<float *>(ws + L.resid);
    int32_t *tile_start = reinterpret_cast<int32_t *>(ws + L.tile_start);
    const long long np = b->npix_total;
    const TileList tl_f{reinterpret_cast<int32_t *>(ws + L.list_f), count + 32,
                        plan->grid_f, tile_start};
    if ((flags & LMZ_FP32) || set.fd_step > 0.0) {
        if (mask) launch_tile_list(b, mask, tl_f, s);
        launch_residual(b, params, sel, flip, nmax, coef, coords, flags, data,
                        resid, np, chi2_tiles, nullptr, mask ? &tl_f : nullptr,
                        s);
        return;
    }
    FTabPar *fpar = reinterpret_cast<FTabPar *>(ws + L.fpar);
    double *ftab = reinterpret_cast<double *>(ws + L.ftab);
    uint8_t *mask_ft = reinterpret_cast<uint8_t *>(ws + L.mask_ft);
    uint8_t *mask_fl = reinterpret_cast<uint8_t *>(ws + L.mask_fl);
    const TileList tl_ft{tl_f.list, tl_f.count, plan->grid_ft, tile_start};
    const TileList tl_fl{reinterpret_cast<int32_t *>(ws + L.list_fl),
                         count + 64, plan->grid_f, tile_start};
    launch_fwd_ftab(b, params, sel, flip, nmax, coef, flags, mask, fpar, ftab,
                    mask_ft, mask_fl, s);
    launch_tile_list(b, mask_ft, tl_ft, s, LMZ_FTAB_GROUP);
    launch_forward_ftab(b, params, sel, flip, nmax, flags, data, resid, np,
                        chi2_tiles, fpar, ftab, &tl_ft, s);
    launch_tile_list(b, mask_fl, tl_fl, s);
    launch_residual(b, params, sel, flip, nmax, coef, coords, flags, data,
                    resid, np, chi2_tiles, nullptr, &tl_fl, s);
}

static void enqueue_iteration(const lmz_lm_plan *plan, const LMLayout &L,
                              char *ws, const float *coords, int32_t flags,
                              const float *data, const lmz_lm_settings &set,
                              const MieSource &src, double *params,
                              hipStream_t s) {
    const lmz_batch *b = plan->batch;
    float *coef = reinterpret_cast<float *>(ws + L.coef);
    float *dcoef = reinterpret_cast<float *>(ws + L.dcoef);
    int32_t *nmax = reinterpret_cast<int32_t *>(ws + L.nmax);
    int32_t *sel = reinterpret_cast<int32_t *>(ws + L.sel);
    double *work = reinterpret_cast<double *>(ws + L.work);
    double *partials = reinterpret_cast<double *>(ws + L.partials);
    double *chi2_tiles = reinterpret_cast<double *>(ws + L.chi2_tiles);
    LMState *st = reinterpret_cast<LMState *>(ws + L.state);
    double *ptrial = reinterpret_cast<double *>(ws + L.ptrial);
    uint8_t *mask_j = reinterpret_cast<uint8_t *>(ws + L.mask_j);
    uint8_t *mask_f = reinterpret_cast<uint8_t *>(ws + L.mask_f);
    int32_t *count = reinterpret_cast<int32_t *>(ws + L.count);
    float *resid = reinterpret_cast<float *>(ws + L.resid);
    const long long np = b->npix_total;
    int32_t *tile_start = reinterpret_cast<int32_t *>(ws + L.tile_start);
    const TileList tl_j{reinterpret_cast<int32_t *>(ws + L.list_j), count + 16,
                        plan->grid_j, tile_start};
    // J^T J and J^T r at the current point, r from its FP64 forward pass,
    // over the tiles of the features that accepted their last step
    if (set.fd_step > 0.0) {
        launch_tile_list(b, mask_j, tl_j, s);
        FdTables fd;
        fd.eps = set.fd_step;
        fd.h = reinterpret_cast<double *>(ws + L.fd_h);
        fd.params = reinterpret_cast<double *>(ws + L.fd_params);
        fd.nmax = reinterpret_cast<int32_t *>(ws + L.fd_nmax);
        fd.coef = reinterpret_cast<float *>(ws + L.fd_coef);
        fd.coef_stride = (long long)LMZ_COEF_STRIDE * b->coef_slots;
        fd.layers = src.layers;
        fd.nlayers = src.nlayers;
        int vary = 0;
        for (int i = 0; i < LMZ_NJAC; ++i) vary |= set.vary[i] ? (1 << i) : 0;
        launch_fd_jacobian(b, params, sel, 0, nmax, coef, fd, vary, coords,
                           flags, resid, np, partials, mask_j, &tl_j, work, s);
    } else {
        // dense crops from radial tables of the sums (k_jacobian_rtab), the
        // rest by the order loop at every pixel (k_jacobian)
        RTabPar *rpar = reinterpret_cast<RTabPar *>(ws + L.rpar);
        float *rtab = reinterpret_cast<float *>(ws + L.rtab);
        uint8_t *mask_jt = reinterpret_cast<uint8_t *>(ws + L.mask_jt);
        uint8_t *mask_jl = reinterpret_cast<uint8_t *>(ws + L.mask_jl);
        const TileList tl_t{tl_j.list, tl_j.count, plan->grid_t, tile_start};
        const TileList tl_l{reinterpret_cast<int32_t *>(ws + L.list_jl),
                            count + 48, plan->grid_j, tile_start};
        launch_jac_rtab(b, params, sel, 0, nmax, coef, dcoef, flags, mask_j,
                        rpar, rtab, mask_jt, mask_jl, s);
        launch_tile_list(b, mask_jt, tl_t, s, LMZ_RTAB_GROUP);
        launch_jacobian_rtab(b, params, sel, 0, nmax, coords, flags, data,
                             resid, np, partials, rpar, rtab, &tl_t, s);
        launch_tile_list(b, mask_jl, tl_l, s);
        launch_jacobian(b, params, sel, 0, nmax, coef, dcoef, coords, flags,
                        data, resid, np, nullptr, nullptr, partials, nullptr,
                        &tl_l, s);
    }
    hipLaunchKernelGGL(k_lm_solve, dim3(b->nfeat), dim3(64), 0, s, b->d_feat,
                       partials, st, params, ptrial, mask_j, mask_f, set);
    if (src.layers)
        launch_mie_layered(b, ptrial, sel, 1, src.layers, src.nlayers, nmax,
                           coef, dcoef, work, mask_f, s);
    else if (!src.fixed_ab)
        launch_mie(b, ptrial, sel, 1, nmax, coef, dcoef, nullptr, work, mask_f, s);
    enqueue_forward(plan, L, ws, coords, flags, data, set, ptrial, 1, mask_f,
                    s);
    hipLaunchKernelGGL(k_lm_accept, dim3(b->nfeat), dim3(64), 0, s, b->d_feat,
                       chi2_tiles, st, params, ptrial, sel, mask_j, mask_f, set);
    hipLaunchKernelGGL(k_lm_count, dim3(1), dim3(256), 0, s, b->nfeat, mask_f,
                       count);
}

extern "C" {

int lmz_lm_workspace_bytes(const lmz_batch *b, size_t *bytes) {
    if (int rc = check_batch(b)) return rc;
    LMZ_REQUIRE(bytes != nullptr, "bytes is NULL");
    *bytes = lm_layout(b).total;
    return LMZ_OK;
}

int lmz_lm_plan_create(const lmz_batch *b, lmz_lm_plan **out) {
    if (int rc = check_batch(b)) return rc;
    LMZ_REQUIRE(out != nullptr, "out is NULL");
    LMZ_REQUIRE(b->coef_buffers == 2,
                "the LM needs a batch created with coef_buffers = 2");
    *out = nullptr;
    lmz_lm_plan *p = new lmz_lm_plan;
    p->batch = b;
    hipError_t e = hipStreamCreateWithFlags(&p->s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_in, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_out, hipEventDisableTiming);
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void **>(&p->h_count), sizeof(int32_t) * 4,
                          hipHostMallocDefault);
    if (e != hipSuccess) {
        lmz_lm_plan_destroy(p);
        return hip_fail(e, "lmz_lm_plan_create");
    }
    *out = p;
    return LMZ_OK;
}

int lmz_lm_plan_destroy(lmz_lm_plan *p) {
    if (!p) return LMZ_OK;
    if (p->s) (void)hipStreamSynchronize(p->s);
    plan_release_graph(p);
    if (p->h_count) (void)hipHostFree(p->h_count);
    if (p->ev_in) (void)hipEventDestroy(p->ev_in);
    if (p->ev_out) (void)hipEventDestroy(p->ev_out);
    if (p->s) (void)hipStreamDestroy(p->s);
    delete p;
    return LMZ_OK;
}

static int lm_run(lmz_lm_plan *plan, const float *coords, int32_t flags,
                  const float *data, const lmz_lm_settings *settings,
                  const double *ab_fixed, const int32_t *nmax_fixed,
                  const double *layers, const int32_t *nlayers,
                  double *params_io, double *result, void *workspace,
                  hipStream_t stream) {
    LMZ_REQUIRE(plan != nullptr && settings != nullptr, "NULL plan or settings");
    const lmz_batch *b = plan->batch;
    if (int rc = check_batch(b)) return rc;
    LMZ_REQUIRE(data && params_io && result && workspace,
                "data, params_io, result and workspace are required");
    LMZ_REQUIRE(coords || b->all_dense,
                "coords is required when a feature has explicit coordinates");
    LMZ_REQUIRE(flags & LMZ_CARTESIAN, "the LM fits the Cartesian hologram");
    const lmz_lm_settings set = *settings;
    LMZ_REQUIRE(set.max_iter >= 1 && set.max_iter <= 100000, "max_iter out of range");
    LMZ_REQUIRE(set.check_every >= 1, "check_every must be >= 1");
    LMZ_REQUIRE(set.ftol >= 0.0 && set.xtol >= 0.0 && set.gtol >= 0.0,
                "tolerances must be >= 0");
    LMZ_REQUIRE(set.damping == LMZ_DAMPING_MINPACK ||
                    set.damping == LMZ_DAMPING_NIELSEN,
                "damping must be LMZ_DAMPING_MINPACK or LMZ_DAMPING_NIELSEN");
    LMZ_REQUIRE(set.damping != LMZ_DAMPING_NIELSEN || set.lambda0 > 0.0,
                "lambda0 must be > 0");
    LMZ_REQUIRE(set.fd_step >= 0.0 && set.fd_step < 1.0,
                "fd_step must be in [0, 1): 0 for the analytic Jacobian");
    LMZ_REQUIRE(!(flags & LMZ_FP32),
                "the LM's residuals are FP64 (its 1e-9 ftol test needs them)");
    int nv = 0;
    for (int i = 0; i < LMZ_NJAC; ++i) nv += set.vary[i] ? 1 : 0;
    LMZ_REQUIRE(nv >= 1, "no parameter is varied");
    const bool fixed_ab = ab_fixed != nullptr;
    LMZ_REQUIRE(!fixed_ab || (!set.vary[3] && !set.vary[4] && !set.vary[5]),
                "a_p, n_p and k_p must be held fixed with fixed coefficients");
    LMZ_REQUIRE(!fixed_ab || nmax_fixed, "nmax is required with ab");
    LMZ_REQUIRE(!layers || (nlayers && !fixed_ab),
                "nlayers is required with layers, and excludes ab");
    const MieSource src{fixed_ab, layers, nlayers};
    LMZ_REQUIRE((reinterpret_cast<uintptr_t>(workspace) & 255) == 0,
                "workspace must be 256-byte aligned");

    if (int rc = persistent_grids(b, &plan->grid_j, &plan->grid_f)) return rc;
    if (int rc = rtab_grid(&plan->grid_t, &plan->grid_ft)) return rc;
    const LMLayout L = lm_layout(b);
    char *ws = static_cast<char *>(workspace);
    LMState *st = reinterpret_cast<LMState *>(ws + L.state);
    int32_t *sel = reinterpret_cast<int32_t *>(ws + L.sel);
    float *coef = reinterpret_cast<float *>(ws + L.coef);
    float *dcoef = reinterpret_cast<float *>(ws + L.dcoef);
    int32_t *nmax = reinterpret_cast<int32_t *>(ws + L.nmax);
    double *work = reinterpret_cast<double *>(ws + L.work);
    double *partials = reinterpret_cast<double *>(ws + L.partials);
    double *chi2_tiles = reinterpret_cast<double *>(ws + L.chi2_tiles);
    double *ptrial = reinterpret_cast<double *>(ws + L.ptrial);
    uint8_t *mask_j = reinterpret_cast<uint8_t *>(ws + L.mask_j);
    uint8_t *mask_f = reinterpret_cast<uint8_t *>(ws + L.mask_f);
    int32_t *count = reinterpret_cast<int32_t *>(ws + L.count);
    float *resid = reinterpret_cast<float *>(ws + L.resid);
    const long long np = b->npix_total;
    hipStream_t s = plan->s;

    LMZ_HIP(hipEventRecord(plan->ev_in, stream));
    LMZ_HIP(hipStreamWaitEvent(s, plan->ev_in, 0));
    // initial state, coefficients with tangents, chi2 at the start
    hipLaunchKernelGGL(k_lm_init, dim3((b->nfeat + 63) / 64), dim3(64), 0, s,
                       b->nfeat, st, params_io, ptrial, sel, mask_j, mask_f,
                       set.lambda0);
    if (fixed_ab)
        hipLaunchKernelGGL(k_lm_load_ab, dim3((b->nfeat + 63) / 64), dim3(64), 0, s,
                           b->d_feat, b->nfeat, ab_fixed, nmax_fixed, nmax, coef,
                           dcoef);
    else if (layers)
        launch_mie_layered(b, params_io, sel, 0, layers, nlayers, nmax, coef,
                           dcoef, work, nullptr, s);
    else
        launch_mie(b, params_io, sel, 0, nmax, coef, dcoef, nullptr, work, nullptr, s);
    enqueue_forward(plan, L, ws, coords, flags, data, set, params_io, 0,
                    nullptr, s);
    hipLaunchKernelGGL(k_lm_init_chi2, dim3(b->nfeat), dim3(64), 0, s, b->nfeat,
                       b->d_feat, chi2_tiles, st);
    hipLaunchKernelGGL(k_lm_mask_init, dim3((b->nfeat + 63) / 64), dim3(64), 0, s,
                       b->nfeat, st, mask_j, mask_f);
    LMZ_HIP(hipGetLastError());

    if (set.use_graph) {
        const void *key[7] = {coords, data, params_io, workspace, result,
                              layers, nlayers};
        const bool same = plan->exec && std::memcmp(key, plan->key_ptrs, sizeof key) == 0 &&
                          plan->key_flags == flags &&
                          plan->key_fixed == (int)fixed_ab &&
                          std::memcmp(&b->sched, &plan->key_sched,
                                      sizeof b->sched) == 0 &&
                          std::memcmp(&set, &plan->key_set, sizeof set) == 0;
        if (!same) {
            plan_release_graph(plan);
            LMZ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            enqueue_iteration(plan, L, ws, coords, flags, data, set, src,
                              params_io, s);
            hipError_t e = hipStreamEndCapture(s, &plan->graph);
            if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
            LMZ_HIP(hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0));
            std::memcpy(plan->key_ptrs, key, sizeof key);
            plan->key_flags = flags;
            plan->key_fixed = (int)fixed_ab;
            plan->key_sched = b->sched;
            plan->key_set = set;
        }
    }
    for (int it = 0; it < set.max_iter; ++it) {
        if (set.use_graph) {
            LMZ_HIP(hipGraphLaunch(plan->exec, s));
        } else {
            enqueue_iteration(plan, L, ws, coords, flags, data, set, src,
                              params_io, s);
            LMZ_HIP(hipGetLastError());
        }
        if ((it + 1) % set.check_every == 0 && it + 1 < set.max_iter) {
            LMZ_HIP(hipMemcpyAsync(plan->h_count, count, sizeof(int32_t),
                                   hipMemcpyDeviceToHost, s));
            LMZ_HIP(hipStreamSynchronize(s));
            if (plan->h_count[0] == 0) break;
        }
    }
    hipLaunchKernelGGL(k_lm_finalize, dim3((b->nfeat + 63) / 64), dim3(64), 0, s,
                       b->nfeat, st, result);
    LMZ_HIP(hipGetLastError());
    LMZ_HIP(hipEventRecord(plan->ev_out, s));
    LMZ_HIP(hipStreamWaitEvent(stream, plan->ev_out, 0));
    return LMZ_OK;
}

int lmz_lm_run(lmz_lm_plan *plan, const float *coords, int32_t flags,
               const float *data, const lmz_lm_settings *settings,
               double *params_io, double *result, void *workspace,
               hipStream_t stream) {
    return lm_run(plan, coords, flags, data, settings, nullptr, nullptr,
                  nullptr, nullptr, params_io, result, workspace, stream);
}

int lmz_lm_run_fixed_ab(lmz_lm_plan *plan, const float *coords, int32_t flags,
                        const float *data, const lmz_lm_settings *settings,
                        const double *ab, const int32_t *nmax,
                        double *params_io, double *result, void *workspace,
                        hipStream_t stream) {
    LMZ_REQUIRE(ab != nullptr, "ab is NULL");
    return lm_run(plan, coords, flags, data, settings, ab, nmax, nullptr,
                  nullptr, params_io, result, workspace, stream);
}

int lmz_lm_run_layered(lmz_lm_plan *plan, const float *coords, int32_t flags,
                       const float *data, const lmz_lm_settings *settings,
                       const double *layers, const int32_t *nlayers,
                       double *params_io, double *result, void *workspace,
                       hipStream_t stream) {
    LMZ_REQUIRE(layers != nullptr, "layers is NULL");
    return lm_run(plan, coords, flags, data, settings, nullptr, nullptr,
                  layers, nlayers, params_io, result, workspace, stream);
}

}  // extern "C"
