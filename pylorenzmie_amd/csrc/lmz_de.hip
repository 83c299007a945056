// Differential-evolution estimates of many features on the device
// (lmz_de_run): the search scipy's differential_evolution runs for one
// feature in the reference's DEEstimator.estimate
// (analysis/DEEstimator.py:119-173), for every search of a batch at once.
//
// Per generation, four launches on the caller's stream:
//   k_de_trial   one lane per member: best1bin trial, its parameter row;
//   Mie + FP32/FP64 residual of every member of the unconverged searches
//                (the library's launches, with a per-feature active mask);
//   k_de_select  one wavefront per search: the members' chi2 from their
//                tile partials, deferred selection, the stopping test, the
//                best member and the outputs.
// The population lives in the workspace; nothing goes back to the host but
// the count of converged searches, read every check_every generations.
#include <hip/hip_runtime.h>

#include "lmz_device.h"
#include "lmz_de.h"

namespace lmz {
namespace {

struct DELayout {
    size_t unit, trial, energy, params, nmax, coef, work, chi2_tiles, active,
        best, done, count, total;
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

DELayout de_layout(const lmz_batch *b, const lmz_de_settings &set) {
    DELayout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off = align256(off + bytes); return o; };
    const size_t nf = (size_t)b->nfeat;
    const size_t ns = nf / (size_t)set.npop;
    L.unit = take(sizeof(double) * nf * set.ndim);
    L.trial = take(sizeof(double) * nf * set.ndim);
    L.energy = take(sizeof(double) * nf);
    L.params = take(sizeof(double) * LMZ_NPARAM * nf);
    L.nmax = take(sizeof(int32_t) * nf);
    L.coef = take(sizeof(float) * LMZ_COEF_STRIDE * (size_t)b->coef_slots);
    L.work = take(sizeof(double) * LMZ_WORK_PER_ORDER * (size_t)(b->cap_max + 1) * nf);
    L.chi2_tiles = take(sizeof(double) * (size_t)b->ntiles);
    L.active = take(nf);
    L.best = take(sizeof(int32_t) * ns);
    L.done = take(sizeof(int32_t) * ns);
    L.count = take(sizeof(int32_t) * 4);
    L.total = off;
    return L;
}

// Parameter row of member f: the search's base with the searched columns
// mapped from the unit cube.
__device__ __forceinline__ void de_write_row(const lmz_de_settings &set,
                                             const double *base,
                                             const double *bounds,
                                             const double *u, double *row) {
    for (int k = 0; k < LMZ_NPARAM; ++k) row[k] = base[k];
    for (int j = 0; j < set.ndim; ++j) {
        const double lo = bounds[2 * j], hi = bounds[2 * j + 1];
        row[set.cols[j]] = lo + u[j] * (hi - lo);
    }
}

// Latin-hypercube start: one workgroup per search.
__global__ __launch_bounds__(64) void k_de_init(
    int nsearch, lmz_de_settings set, const double *__restrict__ base,
    const double *__restrict__ bounds, double *__restrict__ unit,
    double *__restrict__ params, uint8_t *__restrict__ active,
    int32_t *__restrict__ done, int32_t *__restrict__ gens,
    int32_t *__restrict__ count) {
    __shared__ double keys[LMZ_DE_MAXPOP];
    const int s = blockIdx.x;
    if (s >= nsearch) return;
    const int np = set.npop;
    const long long f0 = (long long)s * np;
    for (int j = 0; j < set.ndim; ++j) {
        for (int i = threadIdx.x; i < np; i += blockDim.x)
            keys[i] = de_lhs_key(set.seed, (uint32_t)s, i, j);
        __syncthreads();
        for (int i = threadIdx.x; i < np; i += blockDim.x)
            unit[(f0 + i) * set.ndim + j] = de_lhs(set.seed, (uint32_t)s, i, j, np, keys);
        __syncthreads();
    }
    for (int i = threadIdx.x; i < np; i += blockDim.x) {
        de_write_row(set, base + (long long)LMZ_NPARAM * s,
                     bounds + 2LL * set.ndim * s, unit + (f0 + i) * set.ndim,
                     params + LMZ_NPARAM * (f0 + i));
        active[f0 + i] = 1;
    }
    if (threadIdx.x == 0) {
        done[s] = 0;
        gens[s] = 0;
        if (s == 0) count[0] = 0;
    }
}

// best1bin trials of the unconverged searches, one lane per member.
__global__ __launch_bounds__(64) void k_de_trial(
    int nfeat, int gen, lmz_de_settings set, const double *__restrict__ base,
    const double *__restrict__ bounds, const double *__restrict__ unit,
    const int32_t *__restrict__ best, const int32_t *__restrict__ done,
    double *__restrict__ trial, double *__restrict__ params) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfeat) return;
    const int np = set.npop, s = f / np, i = f - s * np;
    if (done[s]) return;
    const double F = de_scale(set.seed, (uint32_t)gen, (uint32_t)s,
                              set.mutation_lo, set.mutation_hi);
    const long long f0 = (long long)s * np;
    double t[LMZ_DE_MAXDIM];
    de_trial(set.seed, (uint32_t)gen, (uint32_t)s, i, np, set.ndim,
             unit + f0 * set.ndim, best[s], F, set.recombination, t);
    for (int j = 0; j < set.ndim; ++j) trial[(long long)f * set.ndim + j] = t[j];
    de_write_row(set, base + (long long)LMZ_NPARAM * s, bounds + 2LL * set.ndim * s,
                 t, params + (long long)LMZ_NPARAM * f);
}

// Deferred selection and the stopping test, one wavefront per search.
// gen 0 takes the initial population's energies.
__global__ __launch_bounds__(64) void k_de_select(
    int nsearch, int gen, lmz_de_settings set, const lmz_feature *__restrict__ feat,
    const double *__restrict__ chi2_tiles, const double *__restrict__ bounds,
    double *__restrict__ unit, const double *__restrict__ trial,
    double *__restrict__ energy, int32_t *__restrict__ best,
    int32_t *__restrict__ done, uint8_t *__restrict__ active,
    int32_t *__restrict__ count, double *__restrict__ best_out,
    double *__restrict__ energy_out, int32_t *__restrict__ gens) {
    const int s = blockIdx.x;
    if (s >= nsearch || done[s]) return;
    const int np = set.npop;
    const long long f0 = (long long)s * np;
    for (int i = threadIdx.x; i < np; i += blockDim.x) {
        const lmz_feature F = feat[f0 + i];
        double c = 0.0;
        for (int t = 0; t < F.ntiles; ++t) c += chi2_tiles[F.tile_offset + t];
        const double et = de_energy(c);
        if (gen == 0) {
            energy[f0 + i] = et;
        } else if (et < energy[f0 + i]) {
            energy[f0 + i] = et;
            for (int j = 0; j < set.ndim; ++j)
                unit[(f0 + i) * set.ndim + j] = trial[(f0 + i) * set.ndim + j];
        }
    }
    __syncthreads();
    __shared__ int s_stop;
    if (threadIdx.x == 0) {
        const double *E = energy + f0;
        const int b = de_argmin(E, np);
        best[s] = b;
        energy_out[s] = E[b];
        const double *bnd = bounds + 2LL * set.ndim * s;
        for (int j = 0; j < set.ndim; ++j)
            best_out[(long long)s * set.ndim + j] =
                bnd[2 * j] + unit[(f0 + b) * set.ndim + j] * (bnd[2 * j + 1] - bnd[2 * j]);
        const bool conv = gen > 0 && de_converged(E, np, set.tol, set.atol);
        const bool stop = conv || gen >= set.maxiter;
        gens[s] = gen;
        if (stop) {
            done[s] = 1;
            atomicAdd(count, 1);
        }
        s_stop = stop ? 1 : 0;
    }
    __syncthreads();
    if (s_stop)
        for (int i = threadIdx.x; i < np; i += blockDim.x) active[f0 + i] = 0;
}

}  // namespace
}  // namespace lmz

using namespace lmz;

static int de_check(const lmz_batch *b, const lmz_de_settings *set) {
    if (int rc = check_batch(b)) return rc;
    LMZ_REQUIRE(set != nullptr, "settings is NULL");
    LMZ_REQUIRE(set->ndim >= 1 && set->ndim <= LMZ_DE_MAX_DIM,
                "ndim must be in 1..LMZ_DE_MAX_DIM");
    for (int j = 0; j < set->ndim; ++j) {
        LMZ_REQUIRE(set->cols[j] >= 0 && set->cols[j] < LMZ_NJAC,
                    "cols must name parameters 0..6");
        for (int k = 0; k < j; ++k)
            LMZ_REQUIRE(set->cols[k] != set->cols[j], "cols must be distinct");
    }
    LMZ_REQUIRE(set->npop >= 5 && set->npop <= LMZ_DE_MAX_POP,
                "npop must be in 5..LMZ_DE_MAX_POP");
    LMZ_REQUIRE(b->nfeat % set->npop == 0,
                "the batch must hold npop members per search");
    LMZ_REQUIRE(set->maxiter >= 0 && set->maxiter <= 1000000, "maxiter out of range");
    LMZ_REQUIRE(set->check_every >= 1, "check_every must be >= 1");
    LMZ_REQUIRE(set->tol >= 0.0 && set->atol >= 0.0, "tol and atol must be >= 0");
    LMZ_REQUIRE(set->mutation_lo >= 0.0 && set->mutation_hi >= set->mutation_lo &&
                    set->mutation_hi <= 2.0,
                "the mutation range must satisfy 0 <= lo <= hi <= 2");
    LMZ_REQUIRE(set->recombination >= 0.0 && set->recombination <= 1.0,
                "recombination must be in [0, 1]");
    return LMZ_OK;
}

extern "C" {

int lmz_de_workspace_bytes(const lmz_batch *b, const lmz_de_settings *set,
                           size_t *bytes) {
    if (int rc = de_check(b, set)) return rc;
    LMZ_REQUIRE(bytes != nullptr, "bytes is NULL");
    *bytes = de_layout(b, *set).total;
    return LMZ_OK;
}

int lmz_de_run(const lmz_batch *b, const lmz_de_settings *settings,
               const double *base, const double *bounds, const float *coords,
               int32_t flags, const float *data, double *best_out,
               double *energy_out, int32_t *generations, void *workspace,
               hipStream_t s) {
    if (int rc = de_check(b, settings)) return rc;
    const lmz_de_settings set = *settings;
    LMZ_REQUIRE(base && bounds && data && best_out && energy_out && generations &&
                    workspace,
                "base, bounds, data, best, energy, generations and workspace "
                "are required");
    LMZ_REQUIRE(coords || b->all_dense,
                "coords is required when a feature has explicit coordinates");
    LMZ_REQUIRE(b->coef_buffers == 1, "the search needs a batch with coef_buffers = 1");
    LMZ_REQUIRE(!(flags & LMZ_FP32) || (flags & LMZ_CARTESIAN),
                "FP32 residuals are defined for the Cartesian hologram");
    LMZ_REQUIRE((reinterpret_cast<uintptr_t>(workspace) & 255) == 0,
                "workspace must be 256-byte aligned");
    const DELayout L = de_layout(b, set);
    char *ws = static_cast<char *>(workspace);
    double *unit = reinterpret_cast<double *>(ws + L.unit);
    double *trial = reinterpret_cast<double *>(ws + L.trial);
    double *energy = reinterpret_cast<double *>(ws + L.energy);
    double *params = reinterpret_cast<double *>(ws + L.params);
    int32_t *nmax = reinterpret_cast<int32_t *>(ws + L.nmax);
    float *coef = reinterpret_cast<float *>(ws + L.coef);
    double *work = reinterpret_cast<double *>(ws + L.work);
    double *chi2 = reinterpret_cast<double *>(ws + L.chi2_tiles);
    uint8_t *active = reinterpret_cast<uint8_t *>(ws + L.active);
    int32_t *best = reinterpret_cast<int32_t *>(ws + L.best);
    int32_t *done = reinterpret_cast<int32_t *>(ws + L.done);
    int32_t *count = reinterpret_cast<int32_t *>(ws + L.count);
    const int ns = b->nfeat / set.npop;

    auto evaluate = [&]() {
        launch_mie(b, params, nullptr, 0, nmax, coef, nullptr, nullptr, work,
                   active, s);
        launch_residual(b, params, nullptr, 0, nmax, coef, coords, flags, data,
                        nullptr, 0LL, chi2, active, nullptr, s);
    };
    hipLaunchKernelGGL(k_de_init, dim3(ns), dim3(64), 0, s, ns, set, base, bounds,
                       unit, params, active, done, generations, count);
    evaluate();
    hipLaunchKernelGGL(k_de_select, dim3(ns), dim3(64), 0, s, ns, 0, set, b->d_feat,
                       chi2, bounds, unit, trial, energy, best, done, active,
                       count, best_out, energy_out, generations);
    LMZ_HIP(hipGetLastError());
    for (int gen = 1; gen <= set.maxiter; ++gen) {
        hipLaunchKernelGGL(k_de_trial, dim3((b->nfeat + 63) / 64), dim3(64), 0, s,
                           b->nfeat, gen, set, base, bounds, unit, best, done,
                           trial, params);
        evaluate();
        hipLaunchKernelGGL(k_de_select, dim3(ns), dim3(64), 0, s, ns, gen, set,
                           b->d_feat, chi2, bounds, unit, trial, energy, best,
                           done, active, count, best_out, energy_out, generations);
        LMZ_HIP(hipGetLastError());
        if (gen % set.check_every == 0 && gen < set.maxiter) {
            int32_t h = 0;
            LMZ_HIP(hipMemcpyAsync(&h, count, sizeof h, hipMemcpyDeviceToHost, s));
            LMZ_HIP(hipStreamSynchronize(s));
            if (h >= ns) break;
        }
    }
    return LMZ_OK;
}

}  // extern "C"
