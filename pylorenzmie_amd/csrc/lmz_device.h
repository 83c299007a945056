// Library-internal definitions shared by the liblmz translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

#include "lmz_pixel.h"

#define LMZ_THREADS 256
#define LMZ_PPT (LMZ_TILE / LMZ_THREADS)   // pixels per thread

static_assert(LMZ_TILE % LMZ_THREADS == 0, "tile must be whole threads");

struct lmz_batch {
    int32_t nfeat = 0;
    int32_t ntiles = 0;
    int64_t npix_total = 0;
    int32_t coef_buffers = 1;
    int32_t coef_slots = 0;
    int32_t cap_max = 0;
    int32_t all_dense = 1;
    lmz_schedule sched{LMZ_THREADS, LMZ_THREADS, 0, 0};   // launch shapes
    lmz_feature *d_feat = nullptr;      // device copy of the feature table
    int32_t *d_tile_feature = nullptr;  // tile -> feature
};

namespace lmz {

void set_error(const std::string &msg);

// A device-built list of active tiles and the persistent grid that walks it.
struct TileList {
    int32_t *list;    // device, [ntiles]
    int32_t *count;   // device, [2 + LMZ_XCDS]: list length, unused,
                      // per-XCD fetch counters
    int grid;         // blocks of the persistent launch
    int32_t *start;   // device, [nfeat]: each feature's first list position
};

void launch_mie(const lmz_batch *b, const double *params, const int32_t *sel,
                int flip, int32_t *nmax, float *coef, float *dcoef,
                double *ab_out, double *work, const uint8_t *active,
                hipStream_t stream);
// layered spheres in the fitting parametrization (k_mie_layered_fit)
void launch_mie_layered(const lmz_batch *b, const double *params,
                        const int32_t *sel, int flip, const double *layers,
                        const int32_t *nlayers, int32_t *nmax, float *coef,
                        float *dcoef, double *work, const uint8_t *active,
                        hipStream_t stream);
// residual (I - d) / noise written to out + (sel[f] ^ flip) * out_stride
void launch_residual(const lmz_batch *b, const double *params,
                     const int32_t *sel, int flip, const int32_t *nmax,
                     const float *coef, const float *coords, int flags,
                     const float *data, float *out, long long out_stride,
                     double *chi2_tiles, const uint8_t *active,
                     const TileList *tl, hipStream_t stream);
// rin: residuals to use instead of the kernel's own, read from
// rin + (sel[f] ^ flip) * rin_stride (NULL: computed from data)
void launch_jacobian(const lmz_batch *b, const double *params,
                     const int32_t *sel, int flip, const int32_t *nmax,
                     const float *coef, const float *dcoef,
                     const float *coords, int flags, const float *data,
                     const float *rin, long long rin_stride, float *jac,
                     float *resid, double *partials, const uint8_t *active,
                     const TileList *tl, hipStream_t stream);
// Radial tables of the Jacobian's sums for the active features (rpar
// [nfeat], rtab [nfeat][LMZ_RTAB_CAP][LMZ_RTAB_STRIDE]); mask_tab marks the
// tabulated ones, mask_loop the rest (both [nfeat], 0 for inactive ones).
void launch_jac_rtab(const lmz_batch *b, const double *params,
                     const int32_t *sel, int flip, const int32_t *nmax,
                     const float *coef, const float *dcoef, int flags,
                     const uint8_t *active, RTabPar *rpar, float *rtab,
                     uint8_t *mask_tab, uint8_t *mask_loop, hipStream_t stream);
// FP64 radial tables of the forward sums (fpar [nfeat], ftab
// [nfeat][LMZ_FTAB_CAP][LMZ_FTAB_STRIDE]) and the residual pass over the
// tabulated features' tiles in tl, as launch_residual's.
void launch_fwd_ftab(const lmz_batch *b, const double *params,
                     const int32_t *sel, int flip, const int32_t *nmax,
                     const float *coef, int flags, const uint8_t *active,
                     FTabPar *fpar, double *ftab, uint8_t *mask_tab,
                     uint8_t *mask_loop, hipStream_t stream);
void launch_forward_ftab(const lmz_batch *b, const double *params,
                         const int32_t *sel, int flip, const int32_t *nmax,
                         int flags, const float *data, float *out,
                         long long out_stride, double *chi2_tiles,
                         const FTabPar *fpar, const double *ftab,
                         const TileList *tl, hipStream_t stream);
// Normal-equation partials of the tabulated features' tiles in tl.
void launch_jacobian_rtab(const lmz_batch *b, const double *params,
                          const int32_t *sel, int flip, const int32_t *nmax,
                          const float *coords, int flags, const float *data,
                          const float *rin, long long rin_stride,
                          double *partials, const RTabPar *rpar,
                          const float *rtab, const TileList *tl,
                          hipStream_t stream);
// Buffers of the forward-difference Jacobian: steps, the a_p / n_p / k_p
// parameter rows and their coefficient tables.
struct FdTables {
    double eps;            // relative step
    double *h;             // [nfeat][LMZ_NPARAM]
    double *params;        // [3][nfeat][LMZ_NPARAM]
    int32_t *nmax;         // [3][nfeat * coef_buffers]
    float *coef;           // 3 tables, coef_stride floats apart
    long long coef_stride;
    // layered spheres in their fitting parametrization (null: homogeneous)
    const double *layers = nullptr;
    const int32_t *nlayers = nullptr;
};
// Normal-equation partials of a forward-difference Jacobian of the FP64
// hologram, for the features in tl (or all); rin as in launch_jacobian.
void launch_fd_jacobian(const lmz_batch *b, const double *params,
                        const int32_t *sel, int flip, const int32_t *nmax,
                        const float *coef, const FdTables &fd, int vary,
                        const float *coords, int flags, const float *rin,
                        long long rin_stride, double *partials,
                        const uint8_t *active, const TileList *tl,
                        double *work, hipStream_t stream);
// The list of the masked features' tiles, or of every group-th of them
// (the first tile of each group of consecutive tiles of a feature).
void launch_tile_list(const lmz_batch *b, const uint8_t *mask,
                      const TileList &tl, hipStream_t stream, int group = 1);
// Persistent grids of the LM's pixel launches under the batch's schedule.
int persistent_grids(const lmz_batch *b, int *grid_jacobian, int *grid_forward);
int rtab_grid(int *grid, int *grid_fwd);
int hip_fail(hipError_t e, const char *what);
int check_batch(const lmz_batch *b);

#define LMZ_HIP(call)                                                    \
    do {                                                                 \
        hipError_t e_ = (call);                                          \
        if (e_ != hipSuccess) return ::lmz::hip_fail(e_, #call);         \
    } while (0)

#define LMZ_REQUIRE(cond, msg)                                           \
    do {                                                                 \
        if (!(cond)) { ::lmz::set_error(msg); return LMZ_EINVAL; }       \
    } while (0)

// Workgroups of a launch are dispatched round-robin over the 8 XCDs of an
// MI355X (workgroup b on XCD b % 8), each with its own L2.  Tile of
// workgroup b when XCD x takes one contiguous range of the G tiles: the
// tiles of a feature, which share its coefficient tables, then meet in one
// L2 instead of eight.  A bijection of [0, G).
#define LMZ_XCDS 8
__device__ __forceinline__ int xcd_tile(int b, int G) {
    const int q = G / LMZ_XCDS, r = G % LMZ_XCDS;
    const int x = b % LMZ_XCDS, i = b / LMZ_XCDS;
    return x * q + (x < r ? x : r) + i;
}

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block reduction of NV per-thread floats to NV doubles, written to
// out[0..NV) by the first NV threads.  Fixed order: deterministic.
template <int NV, int NT = LMZ_THREADS>
__device__ __forceinline__ void block_reduce_store(float (&v)[NV],
                                                   double *__restrict__ out) {
    __shared__ double red[NT / 64][NV];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const float s = wave_sum(v[i]);
        if (lane == 0) red[wave][i] = (double)s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NT / 64; ++w) s += red[w][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

// out[i] = block sum of v[i] for i < n (n <= NV); each sum as for NV = 1.
template <int NV, int NT = LMZ_THREADS>
__device__ __forceinline__ void block_reduce_store_d(double (&v)[NV],
                                                     double *__restrict__ out,
                                                     int n = NV) {
    __shared__ double red[NT / 64][NV];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const double s = wave_sum_d(v[i]);
        if (lane == 0) red[wave][i] = s;
    }
    __syncthreads();
    if (threadIdx.x < n) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NT / 64; ++w) s += red[w][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

}  // namespace lmz
