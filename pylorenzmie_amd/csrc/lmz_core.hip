// liblmz: batch tables, Mie coefficients, forward model, Jacobian and
// normal equations for gfx950.  C ABI in include/lmz.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "lmz_device.h"

namespace lmz {

static thread_local std::string g_error;

void set_error(const std::string &msg) { g_error = msg; }

int hip_fail(hipError_t e, const char *what) {
    g_error = std::string(what) + ": " + hipGetErrorString(e);
    return LMZ_EHIP;
}

int check_batch(const lmz_batch *b) {
    LMZ_REQUIRE(b != nullptr, "batch is NULL");
    LMZ_REQUIRE(b->d_feat != nullptr && b->d_tile_feature != nullptr,
                "batch was not created by lmz_batch_create");
    return LMZ_OK;
}

// ====================================================================
// kernels
// ====================================================================

// The Mie coefficients in two launches.  First the three sequential chains
// (D1(m x) downward; D1(x) downward, then the upward chain, which needs
// D1(x) only), one lane per feature and chain: blockIdx.y picks the chain
// group, so 64 features share each wavefront.  A wavefront per feature with
// one or two busy lanes would issue every FP64 instruction of the chains
// for the whole wavefront.  (The complex divisions branch on |Re| >= |Im|
// and diverge across features; a branch-free form changed the last bits of
// a_n, which moved fits on crops with neighbouring minima, so the chains
// keep the division the rest of the library uses.)  Then the per-order
// coefficients and tangents, one wavefront per feature with the orders
// across its lanes (lane + 1, lane + 65, ...).  The scratch of feature f is
// the contiguous block work + f * LMZ_WORK_PER_ORDER * (cap_max + 1).
__global__ __launch_bounds__(64) void k_mie_chains(
    const lmz_feature *__restrict__ feat, int nfeat, int cap_max,
    const double *__restrict__ params, double *__restrict__ work,
    const uint8_t *__restrict__ active) {
    const int f = blockIdx.x * 64 + threadIdx.x;
    if (f >= nfeat) return;
    if (active && !active[f]) return;
    const lmz_feature F = feat[f];
    const double *p = params + (long long)LMZ_NPARAM * f;
    const MieSetup ms = mie_setup(p[3], p[4], p[5], F.n_m, F.wavelength,
                                  F.nmax_cap);
    double *w = work + (long long)f * LMZ_WORK_PER_ORDER * (cap_max + 1);
    if (blockIdx.y == 0) {
        mie_d1_down(ms.x * ms.m, ms.N, w, 1, 0);
    } else {
        mie_d1_down(cd{ms.x, 0.0}, ms.N, w, 1, 4);
        mie_up_chain(ms.x, ms.N, w, 1);
    }
}

__global__ __launch_bounds__(64) void k_mie(
    const lmz_feature *__restrict__ feat, int nfeat, int cap_max,
    const double *__restrict__ params, const int32_t *__restrict__ sel,
    int flip, int nbuf, int32_t *__restrict__ nmax_out,
    float *__restrict__ coef, float *__restrict__ dcoef,
    double *__restrict__ ab_out, const double *__restrict__ work,
    const uint8_t *__restrict__ active) {
    const int f = blockIdx.x;
    if (f >= nfeat) return;
    if (active && !active[f]) return;
    const int lane = threadIdx.x;
    const lmz_feature F = feat[f];
    const double *p = params + (long long)LMZ_NPARAM * f;
    const int b = sel ? (sel[f] ^ flip) : 0;
    const long long slot0 = F.coef_offset + (long long)b * F.nmax_cap;
    const MieSetup ms = mie_setup(p[3], p[4], p[5], F.n_m, F.wavelength,
                                  F.nmax_cap);
    const double *w = work + (long long)f * LMZ_WORK_PER_ORDER * (cap_max + 1);
    for (int n = lane + 1; n <= ms.N; n += 64)
        mie_order(ms, n, w, 1, coef ? coef + LMZ_COEF_STRIDE * slot0 : nullptr,
                  dcoef ? dcoef + LMZ_DCOEF_STRIDE * slot0 : nullptr,
                  ab_out ? ab_out + 4 * slot0 : nullptr);
    if (lane == 0) nmax_out[(long long)f * nbuf + b] = ms.N;
}

// Layered spheres, one lane per feature (FP64, sequential in n).  A
// feature whose layer table is invalid (layer count outside 1..
// LMZ_MAX_LAYERS, radii not positive and increasing, or a non-finite
// index) gets n_max = -1 and no coefficients.
__global__ __launch_bounds__(64) void k_mie_layered(
    const lmz_feature *__restrict__ feat, int nfeat,
    const double *__restrict__ layers, const int32_t *__restrict__ nlayers,
    int nbuf, int32_t *__restrict__ nmax_out, float *__restrict__ coef,
    double *__restrict__ ab_out, double *__restrict__ work) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfeat) return;
    const lmz_feature F = feat[f];
    const int L = nlayers[f];
    double r[LMZ_MAX_LAYERS], nl[LMZ_MAX_LAYERS], kl[LMZ_MAX_LAYERS];
    bool ok = L >= 1 && L <= LMZ_MAX_LAYERS;
    for (int l = 0; ok && l < L; ++l) {
        const double *q = layers + ((long long)f * LMZ_MAX_LAYERS + l) * 3;
        r[l] = q[0];
        nl[l] = q[1];
        kl[l] = q[2];
        ok = r[l] > (l ? r[l - 1] : 0.0) && isfinite(r[l]) && isfinite(nl[l]) &&
             isfinite(kl[l]) && nl[l] > 0.0;
    }
    if (!ok) {
        nmax_out[(long long)f * nbuf] = -1;
        return;
    }
    const long long slot0 = F.coef_offset;
    nmax_out[(long long)f * nbuf] = mie_layered_feature(
        r, nl, kl, L, F.n_m, F.wavelength, F.nmax_cap, work + f, nfeat,
        coef ? coef + LMZ_COEF_STRIDE * slot0 : nullptr,
        ab_out ? ab_out + 4 * slot0 : nullptr);
}

// Layered spheres in the fitting parametrization: the table's radii scaled
// by params a_p / (the table's outer radius), the outer layer's index
// n_p + i k_p from params, inner indices from the table; coefficient tables
// of buffer (sel ^ flip) with the a_p, n_p, k_p tangents (dcoef, may be
// NULL).  One lane per feature and pass (blockIdx.y): 0 the coefficients
// and n_max, 1 and 2 the a_p and n_p (with k_p) tangents, each carried with
// its own values through the sequential recursion, in scratch column block
// pass * nfeat + f of the LMZ_WORK_PASSES interleaved blocks.  The three
// recursions of a feature run side by side instead of one after another.
// A trial point outside the domain (a_p <= 0, n_p <= 0, non-finite) gets no
// orders (n_max 0: no scattered field), which the LM rejects by its chi2.
__global__ __launch_bounds__(64) void k_mie_layered_fit(
    const lmz_feature *__restrict__ feat, int nfeat,
    const double *__restrict__ params, const int32_t *__restrict__ sel,
    int flip, int nbuf, const double *__restrict__ layers,
    const int32_t *__restrict__ nlayers, int32_t *__restrict__ nmax_out,
    float *__restrict__ coef, float *__restrict__ dcoef,
    double *__restrict__ work, const uint8_t *__restrict__ active) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    const int pass = blockIdx.y;
    if (f >= nfeat) return;
    if (active && !active[f]) return;
    if (pass > 0 && !dcoef) return;
    const lmz_feature F = feat[f];
    const int b = sel ? (sel[f] ^ flip) : 0;
    const double *p = params + (long long)LMZ_NPARAM * f;
    const int L = nlayers[f];
    double r[LMZ_MAX_LAYERS], nl[LMZ_MAX_LAYERS], kl[LMZ_MAX_LAYERS];
    bool ok = L >= 1 && L <= LMZ_MAX_LAYERS && p[3] > 0.0 && p[4] > 0.0 &&
              isfinite(p[3]) && isfinite(p[4]) && isfinite(p[5]);
    for (int l = 0; ok && l < L; ++l) {
        const double *q = layers + ((long long)f * LMZ_MAX_LAYERS + l) * 3;
        r[l] = q[0];
        nl[l] = q[1];
        kl[l] = q[2];
        ok = r[l] > (l ? r[l - 1] : 0.0) && isfinite(r[l]);
    }
    if (!ok) {
        if (pass == 0) nmax_out[(long long)f * nbuf + b] = 0;
        return;
    }
    const double scale = p[3] / r[L - 1];
    for (int l = 0; l < L - 1; ++l) r[l] *= scale;
    r[L - 1] = p[3];
    nl[L - 1] = p[4];
    kl[L - 1] = p[5];
    const long long slot0 = F.coef_offset + (long long)b * F.nmax_cap;
    double *w = work + (long long)pass * nfeat + f;
    const long long wstride = (long long)LMZ_WORK_PASSES * nfeat;
    if (pass == 0) {
        const int N = mie_layered_feature(r, nl, kl, L, F.n_m, F.wavelength,
                                          F.nmax_cap, w, wstride,
                                          coef + LMZ_COEF_STRIDE * slot0, nullptr);
        nmax_out[(long long)f * nbuf + b] = N;
    } else {
        const int N = layered_nmax(r, nl, kl, L, F.n_m, F.wavelength, F.nmax_cap);
        mie_layered_tangent(r, nl, kl, L, F.n_m, F.wavelength, pass - 1, N, w,
                            wstride, dcoef + LMZ_DCOEF_STRIDE * slot0);
    }
}

// Per-order tables from FP64 a_n, b_n supplied by the caller.
__global__ __launch_bounds__(64) void k_coef_from_ab(
    const lmz_feature *__restrict__ feat, int nfeat,
    const double *__restrict__ ab_in, const int32_t *__restrict__ nmax_in,
    int nbuf, int32_t *__restrict__ nmax_out, float *__restrict__ coef) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfeat) return;
    const lmz_feature F = feat[f];
    int N = nmax_in[f];
    N = N < 0 ? 0 : (N > F.nmax_cap ? F.nmax_cap : N);
    for (int n = 1; n <= N; ++n) {
        const double *r = ab_in + 4LL * (F.coef_offset + n - 1);
        write_order(n, cd{r[0], r[1]}, cd{r[2], r[3]},
                    coef + (long long)LMZ_COEF_STRIDE * (F.coef_offset + n - 1));
    }
    nmax_out[(long long)f * nbuf] = N;
}

// Next entry of a work list for the whole block.  The list (length
// count[0]) is split into LMZ_XCDS contiguous chunks, chunk x fetched
// through counter count[2 + x] (set to 0 by k_tile_starts).  A block starts on
// the chunk of its own XCD, so the tiles of one feature stay in one L2, and
// moves on to the other chunks when its own is exhausted.  Each fetch claims
// LMZ_CLAIM consecutive entries (fewer at a chunk's end), which the block
// then takes one by one from its cursor without a fetch or a barrier.
// `k` (thread 0's count of exhausted chunks, initially 0) only grows, so
// every block reaches the end value count[0] after at most count[0] +
// LMZ_XCDS fetches.
#ifndef LMZ_CLAIM
#define LMZ_CLAIM 2
#endif
struct TileCursor {
    int cur = 0, end = 0;   // claimed entries [cur, end), the same in every thread
    int k = 0;              // thread 0: chunks found exhausted
};

template <int CLAIM = LMZ_CLAIM>
__device__ __forceinline__ int next_tile(int32_t *count, int (&slot)[2],
                                         TileCursor &c) {
    if (c.cur < c.end) return c.cur++;
    if (threadIdx.x == 0) {
        const int n = count[0];
        const int q = n / LMZ_XCDS, r = n % LMZ_XCDS;
        const int home = (int)(blockIdx.x % LMZ_XCDS);
        int w = n, e = n;
        while (c.k < LMZ_XCDS) {
            const int x = (home + c.k) % LMZ_XCDS;
            const int len = q + (x < r ? 1 : 0);
            if (len == 0) {   // an empty chunk: nothing to claim
                ++c.k;
                continue;
            }
            const int i = atomicAdd(count + 2 + x, CLAIM);
            if (i < len) {
                const int base = x * q + (x < r ? x : r);
                w = base + i;
                e = base + (i + CLAIM < len ? i + CLAIM : len);
                break;
            }
            ++c.k;
        }
        slot[0] = w;
        slot[1] = e;
    }
    __syncthreads();
    const int w = slot[0], e = slot[1];
    __syncthreads();
    c.cur = w + 1;
    c.end = e;
    return w;
}

// Forward pass over tiles of LMZ_TILE pixels, in FP64.
// MODE 0: field, 1: hologram, 2: residual (+ per-tile r^T r), 3: the
// residual in FP32 arithmetic (Cartesian; chi2 partials still FP64), or
// without data the FP32 hologram into out.
// Without a work list each block takes tile xcd_tile(blockIdx.x); with one
// (the LM's compacted list of active tiles, length count[0]) the grid is
// persistent and blocks fetch list entries dynamically (next_tile) until
// the list is exhausted.
// NT threads share a tile, LMZ_TILE / NT pixels each (the schedule).
template <int MODE, int NT = LMZ_THREADS>
__global__ __launch_bounds__(NT) void k_forward(
    const lmz_feature *__restrict__ feat,
    const int32_t *__restrict__ tile_feature,
    const double *__restrict__ params, const int32_t *__restrict__ sel,
    int flip, int nbuf, const int32_t *__restrict__ nmax,
    const float *__restrict__ coef, const float *__restrict__ coords,
    long long npix_total, int flags, const float *__restrict__ data,
    float *__restrict__ out, long long out_stride, cf *__restrict__ field,
    double *__restrict__ chi2_tiles, const uint8_t *__restrict__ active,
    const int32_t *__restrict__ list, int32_t *count) {
  __shared__ int next_w[2];
  TileCursor cursor;
  const int nwork = count ? count[0] : (int)gridDim.x;
  for (int w = list ? next_tile(count, next_w, cursor)
                    : xcd_tile((int)blockIdx.x, (int)gridDim.x);
       w < nwork;
       w = list ? next_tile(count, next_w, cursor) : nwork) {
    const int t = list ? list[w] : w;
    const int f = tile_feature[t];
    if (active && !active[f]) continue;
    const lmz_feature &F = feat[f];
    const int b = sel ? (sel[f] ^ flip) : 0;
    const int N = nmax[(long long)f * nbuf + b];
    const FeatUniform fu = feature_uniform(
        F, params + (long long)LMZ_NPARAM * f, N,
        F.coef_offset + b * F.nmax_cap);
    const bool cart = flags & LMZ_CARTESIAN;
    const bool bohren = flags & LMZ_BOHREN;
    const float *cf0 = coef + (long long)LMZ_COEF_STRIDE * fu.slot0;
    float *outf = out ? out + b * out_stride : nullptr;
    double chi = 0.0;
#pragma unroll 1
    for (int q = 0; q < LMZ_TILE / NT; ++q) {
        const int j = (t - fu.tile0) * LMZ_TILE + q * NT + threadIdx.x;
        if (j >= fu.npix) break;
        float ax, ay, az;
        pixel_coords(fu, j, coords, npix_total, ax, ay, az);
        const long long gi = fu.pix_offset + j;
        if constexpr (MODE == 3) {   // FP32 arithmetic (Cartesian)
            PixGeom gf;
            pixel_geometry(fu, ax, ay, az, bohren, gf);
            const float I = intensity_f(gf, forward_sum_f(gf, cf0, fu.nmax),
                                        fu.alpha);
            if (data == nullptr) {   // the hologram itself
                outf[gi] = I;
                continue;
            }
            const double r = ((double)I - (double)data[gi]) / F.noise;
            if (outf) outf[gi] = (float)r;
            chi = fma(r, r, chi);
            continue;
        }
        PixGeomD g;
        pixel_geometry_d(fu, ax, ay, az, bohren, g);
        const SumsD S = forward_sum(g, cf0, fu.nmax);
        cd E[3];
        field_from_sums(g, S, cart, E);
        if (MODE == 0) {
            const double al = fu.alphad;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                cf e{(float)(al * E[c].re), (float)(al * E[c].im)};
                if (flags & LMZ_ACCUMULATE) e = e + field[c * npix_total + gi];
                field[c * npix_total + gi] = e;
            }
        } else {
            const double I = intensity(E, fu.alphad);
            if (MODE == 1) {
                out[gi] = (float)I;
            } else {
                const double r = (I - (double)data[gi]) / F.noise;
                if (outf) outf[gi] = (float)r;
                chi = fma(r, r, chi);
            }
        }
    }
    if ((MODE == 2 || MODE == 3) && chi2_tiles) {
        double v[1] = {chi};
        block_reduce_store_d<1, NT>(v, chi2_tiles + t);
        __syncthreads();   // the reduction buffer is reused by the next tile
    }
  }
}

// Forward-mode Jacobian over one tile; optional J, residual, and per-tile
// normal-equation partials.  NT threads per tile as k_forward.
template <int NT = LMZ_THREADS>
__global__ __launch_bounds__(NT) void k_jacobian(
    const lmz_feature *__restrict__ feat,
    const int32_t *__restrict__ tile_feature,
    const double *__restrict__ params, const int32_t *__restrict__ sel,
    int flip, int nbuf, const int32_t *__restrict__ nmax,
    const float *__restrict__ coef, const float *__restrict__ dcoef,
    const float *__restrict__ coords, long long npix_total, int flags,
    const float *__restrict__ data, const float *__restrict__ rin,
    long long rin_stride, float *__restrict__ jac,
    float *__restrict__ resid, double *__restrict__ partials,
    const uint8_t *__restrict__ active, const int32_t *__restrict__ list,
    int32_t *count) {
  __shared__ int next_w[2];
  // Each thread's normal-equation partials live in LDS (its own column of
  // acc_s), not in registers: the 36 registers they would hold keep the
  // kernel below the register budget of four waves per SIMD.
  __shared__ float acc_s[LMZ_NRED][NT];
  float *mine = &acc_s[0][threadIdx.x];
  TileCursor cursor;
  const int nwork = count ? count[0] : (int)gridDim.x;
  for (int w = list ? next_tile(count, next_w, cursor)
                    : xcd_tile((int)blockIdx.x, (int)gridDim.x);
       w < nwork;
       w = list ? next_tile(count, next_w, cursor) : nwork) {
    const int t = list ? list[w] : w;
    const int f = tile_feature[t];
    if (active && !active[f]) continue;
    const lmz_feature &F = feat[f];
    const int b = sel ? (sel[f] ^ flip) : 0;
    const int N = nmax[(long long)f * nbuf + b];
    const FeatUniform fu = feature_uniform(
        F, params + (long long)LMZ_NPARAM * f, N,
        F.coef_offset + b * F.nmax_cap);
    const bool bohren = flags & LMZ_BOHREN;
    const float *cf0 = coef + (long long)LMZ_COEF_STRIDE * fu.slot0;
    const float *df0 = dcoef + (long long)LMZ_DCOEF_STRIDE * fu.slot0;
    const float *rinf = rin ? rin + b * rin_stride : nullptr;
#pragma unroll
    for (int i = 0; i < LMZ_NRED; ++i) mine[i * NT] = 0.f;
#pragma unroll 1
    for (int q = 0; q < LMZ_TILE / NT; ++q) {
        // the partials stay in LDS across pixels (no register copies)
        asm volatile("" ::: "memory");
        const int j = (t - fu.tile0) * LMZ_TILE + q * NT + threadIdx.x;
        if (j >= fu.npix) break;
        float ax, ay, az;
        pixel_coords(fu, j, coords, npix_total, ax, ay, az);
        PixGeom g;
        pixel_geometry<true>(fu, ax, ay, az, bohren, g);
        JacPixel jp;
        jacobian_pixel(fu, g, cf0, df0, fu.nmax, jp);
        const long long gi = fu.pix_offset + j;
        if (jac) {
#pragma unroll
            for (int c = 0; c < LMZ_NJAC; ++c) jac[c * npix_total + gi] = jp.J[c];
        }
        if (data) {
            const float r = rinf ? rinf[gi] : (jp.I - data[gi]) * fu.inv_noise;
            if (resid) resid[gi] = r;
            if (partials) {
                float Jn[LMZ_NJAC];
#pragma unroll
                for (int c = 0; c < LMZ_NJAC; ++c) Jn[c] = jp.J[c] * fu.inv_noise;
                int idx = 0;
#pragma unroll
                for (int a = 0; a < LMZ_NJAC; ++a)
#pragma unroll
                    for (int c = a; c < LMZ_NJAC; ++c, ++idx)
                        mine[idx * NT] = fmaf(Jn[a], Jn[c], mine[idx * NT]);
#pragma unroll
                for (int a = 0; a < LMZ_NJAC; ++a)
                    mine[(28 + a) * NT] = fmaf(Jn[a], r, mine[(28 + a) * NT]);
                mine[35 * NT] = fmaf(r, r, mine[35 * NT]);
            }
        }
    }
    if (partials && data) {
        float acc[LMZ_NRED];
#pragma unroll
        for (int i = 0; i < LMZ_NRED; ++i) acc[i] = mine[i * NT];
        block_reduce_store<LMZ_NRED, NT>(acc, partials + (long long)LMZ_NRED * t);
        __syncthreads();   // the reduction buffer is reused by the next tile
    }
  }
}

// Orders whose coefficients the table builders stage in LDS (a feature of
// more orders reads them from global memory): every sample's order loop
// reads each order's coefficients, one dependent load per order.
#define LMZ_TAB_NSTAGE 128
#define LMZ_TAB_THREADS 256

// Copy n floats from global memory to LDS (the whole block).
__device__ __forceinline__ void stage_floats(float *dst, const float *src, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// Radial tables of the Jacobian's sums (lmz_pixel.h, rtab_params): one
// thread per sample, LMZ_TAB_THREADS samples per block, blockIdx.y the
// sample block, the feature's coefficients staged in LDS;
// thread 0 of the first block writes the feature's grid (m = 0 when the
// feature is not tabulated) and sorts the active features into mask_tab
// (k_jacobian_rtab) and mask_loop (k_jacobian).  The coefficient tables are
// those k_jacobian reads (buffer sel[f] ^ flip).
__global__ __launch_bounds__(LMZ_TAB_THREADS) void k_jac_rtab(
    const lmz_feature *__restrict__ feat, const double *__restrict__ params,
    const int32_t *__restrict__ sel, int flip, int nbuf,
    const int32_t *__restrict__ nmax, const float *__restrict__ coef,
    const float *__restrict__ dcoef, int flags,
    const uint8_t *__restrict__ active, RTabPar *__restrict__ rpar,
    float *__restrict__ rtab, uint8_t *__restrict__ mask_tab,
    uint8_t *__restrict__ mask_loop) {
    const int f = blockIdx.x;
    if (active && !active[f]) {
        if (blockIdx.y == 0 && threadIdx.x == 0) mask_tab[f] = mask_loop[f] = 0;
        return;
    }
    const lmz_feature &F = feat[f];
    const int b = sel ? (sel[f] ^ flip) : 0;
    const int N = nmax[(long long)f * nbuf + b];
    const FeatUniform fu = feature_uniform(
        F, params + (long long)LMZ_NPARAM * f, N,
        F.coef_offset + b * F.nmax_cap);
    const RTabPar rp = rtab_params(fu);
    if (blockIdx.y == 0 && threadIdx.x == 0) {
        rpar[f] = rp;
        mask_tab[f] = rp.m > 0 ? 1 : 0;
        mask_loop[f] = rp.m > 0 ? 0 : 1;
    }
    if ((int)blockIdx.y * LMZ_TAB_THREADS >= rp.m) return;   // the whole block
    __shared__ float sc[LMZ_TAB_NSTAGE * LMZ_COEF_STRIDE];
    __shared__ float sd[LMZ_TAB_NSTAGE * LMZ_DCOEF_STRIDE];
    const float *c = coef + (long long)LMZ_COEF_STRIDE * fu.slot0;
    const float *d = dcoef + (long long)LMZ_DCOEF_STRIDE * fu.slot0;
    const bool lds = fu.nmax <= LMZ_TAB_NSTAGE;
    if (lds) {
        stage_floats(sc, c, fu.nmax * LMZ_COEF_STRIDE);
        stage_floats(sd, d, fu.nmax * LMZ_DCOEF_STRIDE);
        __syncthreads();
    }
    const int j = blockIdx.y * LMZ_TAB_THREADS + threadIdx.x;
    if (j >= rp.m) return;
    float *out = rtab + ((long long)f * LMZ_RTAB_CAP + j) * LMZ_RTAB_STRIDE;
    if (lds)
        rtab_sample(fu, rp, flags & LMZ_BOHREN, j, sc, sd, out);
    else
        rtab_sample(fu, rp, flags & LMZ_BOHREN, j, c, d, out);
}

// Normal-equation partials of the tabulated features' tiles (the LM's
// Jacobian pass on dense crops): each pixel interpolates the radial table
// of its feature (k_jac_rtab) from LDS: the whole table when it fits in
// LMZ_RTAB_LDS rows, else the window of rows its group of tiles reads
// (table_rows), restaged only when a group needs rows the block does not
// hold; global memory when even the window is larger.  Projects as
// k_jacobian.  Persistent grid over a device-built tile list; residuals
// from rin (the FP64 forward pass) or data.  The partials stay in registers:
// without the order loop the kernel has them to spare.
template <int NT = LMZ_THREADS>
__global__ __launch_bounds__(NT) void k_jacobian_rtab(
    const lmz_feature *__restrict__ feat,
    const int32_t *__restrict__ tile_feature,
    const double *__restrict__ params, const int32_t *__restrict__ sel,
    int flip, int nbuf, const int32_t *__restrict__ nmax,
    const float *__restrict__ coords, long long npix_total, int flags,
    const float *__restrict__ data, const float *__restrict__ rin,
    long long rin_stride, double *__restrict__ partials,
    const RTabPar *__restrict__ rpar, const float *__restrict__ rtab,
    const int32_t *__restrict__ list, int32_t *count) {
  __shared__ int next_w[2];
  __shared__ float4 stab[LMZ_RTAB_LDS * LMZ_RTAB_LDS_STRIDE / 4];
  TileCursor cursor;
  int staged = -1, sb = 0, se = 0;   // stab holds rows sb .. se - 1 of staged's table
  // Each list entry is the first tile of a group of LMZ_RTAB_GROUP
  // consecutive tiles of one feature (group k: its tiles k G .. k G + G - 1);
  // the group's partials are summed in registers and reduced once, into the
  // group's first tile, its other tiles get zeros (the LM sums a feature's
  // tiles).  The grouping is a function of the tile's place in its feature,
  // so a feature's partials do not depend on the rest of the batch.
  const int nwork = count[0];
  for (int w = next_tile(count, next_w, cursor); w < nwork;
       w = next_tile(count, next_w, cursor)) {
    const int t = list[w];
    const int f = tile_feature[t];
    const lmz_feature &F = feat[f];
    const int b = sel ? (sel[f] ^ flip) : 0;
    const int N = nmax[(long long)f * nbuf + b];
    const FeatUniform fu = feature_uniform(
        F, params + (long long)LMZ_NPARAM * f, N,
        F.coef_offset + b * F.nmax_cap);
    const bool bohren = flags & LMZ_BOHREN;
    const RTabPar rp = rpar[f];
    const float *gtab = rtab + (long long)f * (LMZ_RTAB_CAP * LMZ_RTAB_STRIDE);
    const int ng = min(LMZ_RTAB_GROUP, F.ntiles - (t - fu.tile0));
    // the rows the group reads: the whole table, or, for a table larger
    // than LDS, the window of the group's pixels
    int lo = 0, hi = rp.m - 1;
    if (rp.m > LMZ_RTAB_LDS) {
        const int jp0 = (t - fu.tile0) * LMZ_TILE;
        table_rows<LMZ_RTAB_P, float>(fu.width, jp0,
                                      min(jp0 + ng * LMZ_TILE, fu.npix), fu.k,
                                      fu.ix + fu.fx, fu.iy + fu.fy, rp.rho0,
                                      rp.inv_h, rp.m, lo, hi);
    }
    const bool lds = hi - lo < LMZ_RTAB_LDS;
    if (lds && !(staged == f && lo >= sb && hi < se)) {
        // every thread has finished the previous group (the barriers of
        // the reduction and of next_tile)
        sb = rp.m > LMZ_RTAB_LDS ? table_window(lo, hi, rp.m, LMZ_RTAB_LDS) : 0;
        se = min(sb + LMZ_RTAB_LDS, rp.m);
        const float4 *src = reinterpret_cast<const float4 *>(
            gtab + (long long)sb * LMZ_RTAB_STRIDE);
        constexpr int G = LMZ_RTAB_STRIDE / 4, S = LMZ_RTAB_LDS_STRIDE / 4;
        for (int i = threadIdx.x; i < (se - sb) * G; i += NT)
            stab[(i / G) * S + i % G] = src[i];
        __syncthreads();
        staged = f;
    }
    const float *rinf = rin ? rin + b * rin_stride : nullptr;
    float acc[LMZ_NRED];
#pragma unroll
    for (int i = 0; i < LMZ_NRED; ++i) acc[i] = 0.f;
#pragma unroll 1
    for (int q = 0; q < ng * (LMZ_TILE / NT); ++q) {
        const int j = (t - fu.tile0) * LMZ_TILE + q * NT + threadIdx.x;
        if (j >= fu.npix) break;
        float ax, ay, az;
        pixel_coords(fu, j, coords, npix_total, ax, ay, az);
        PixGeom g;
        pixel_geometry<true>(fu, ax, ay, az, bohren, g);
        JacPixel jp;   // J scaled by 1 / noise
        if (lds)
            jacobian_pixel_rtab<LMZ_RTAB_LDS_STRIDE>(
                fu, g, rp, reinterpret_cast<const float *>(stab), fu.inv_noise,
                jp, sb);
        else
            jacobian_pixel_rtab<LMZ_RTAB_STRIDE>(fu, g, rp, gtab, fu.inv_noise, jp);
        const long long gi = fu.pix_offset + j;
        const float r = rinf ? rinf[gi] : (jp.I - data[gi]) * fu.inv_noise;
        const float *Jn = jp.J;
        int idx = 0;
#pragma unroll
        for (int a = 0; a < LMZ_NJAC; ++a)
#pragma unroll
            for (int c = a; c < LMZ_NJAC; ++c, ++idx)
                acc[idx] = fmaf(Jn[a], Jn[c], acc[idx]);
#pragma unroll
        for (int a = 0; a < LMZ_NJAC; ++a) acc[28 + a] = fmaf(Jn[a], r, acc[28 + a]);
        acc[35] = fmaf(r, r, acc[35]);
    }
    block_reduce_store<LMZ_NRED, NT>(acc, partials + (long long)LMZ_NRED * t);
    // after the group's loads (a load issued after a store waits for it)
    if (threadIdx.x < LMZ_NRED)
        for (int k = 1; k < ng; ++k)
            partials[(long long)LMZ_NRED * (t + k) + threadIdx.x] = 0.0;
    __syncthreads();   // the reduction buffer is reused by the next group
  }
}

// FP64 radial tables of the forward sums (ftab_params): one thread per
// sample, 64 per block, blockIdx.y the sample block; thread 0 of the first
// block writes the feature's grid and sorts the active features into
// mask_tab (k_forward_ftab) and mask_loop (k_forward).  Coefficient buffer
// sel[f] ^ flip, as the residual kernel reads.
__global__ __launch_bounds__(LMZ_TAB_THREADS) void k_fwd_ftab(
    const lmz_feature *__restrict__ feat, const double *__restrict__ params,
    const int32_t *__restrict__ sel, int flip, int nbuf,
    const int32_t *__restrict__ nmax, const float *__restrict__ coef,
    int flags, const uint8_t *__restrict__ active, FTabPar *__restrict__ fpar,
    double *__restrict__ ftab, uint8_t *__restrict__ mask_tab,
    uint8_t *__restrict__ mask_loop) {
    const int f = blockIdx.x;
    if (active && !active[f]) {
        if (blockIdx.y == 0 && threadIdx.x == 0) mask_tab[f] = mask_loop[f] = 0;
        return;
    }
    const lmz_feature &F = feat[f];
    const int b = sel ? (sel[f] ^ flip) : 0;
    const int N = nmax[(long long)f * nbuf + b];
    const FeatUniform fu = feature_uniform(
        F, params + (long long)LMZ_NPARAM * f, N,
        F.coef_offset + b * F.nmax_cap);
    const FTabPar fp = ftab_params(fu);
    if (blockIdx.y == 0 && threadIdx.x == 0) {
        fpar[f] = fp;
        mask_tab[f] = fp.m > 0 ? 1 : 0;
        mask_loop[f] = fp.m > 0 ? 0 : 1;
    }
    if ((int)blockIdx.y * LMZ_TAB_THREADS >= fp.m) return;   // the whole block
    __shared__ float sc[LMZ_TAB_NSTAGE * LMZ_COEF_STRIDE];
    const float *c = coef + (long long)LMZ_COEF_STRIDE * fu.slot0;
    const bool lds = fu.nmax <= LMZ_TAB_NSTAGE;
    if (lds) {
        stage_floats(sc, c, fu.nmax * LMZ_COEF_STRIDE);
        __syncthreads();
    }
    const int j = blockIdx.y * LMZ_TAB_THREADS + threadIdx.x;
    if (j >= fp.m) return;
    double *out = ftab + ((long long)f * LMZ_FTAB_CAP + j) * LMZ_FTAB_STRIDE;
    if (lds)
        ftab_sample(fu, fp, flags & LMZ_BOHREN, j, sc, out);
    else
        ftab_sample(fu, fp, flags & LMZ_BOHREN, j, c, out);
}

// The FP64 residual (I - d) / noise and per-tile r^T r of the tabulated
// features' tiles (the LM's trial forward pass on dense crops): each pixel
// interpolates its feature's table (k_fwd_ftab) from LDS as
// k_jacobian_rtab does (whole table, window, or global memory; LMZ_FTAB_LDS
// rows), and projects and squares as
// k_forward<2>.  Persistent grid over a list of groups of LMZ_FTAB_GROUP
// consecutive tiles of a feature (as k_jacobian_rtab's); each tile's r^T r
// is reduced on its own.  The residual goes to out + (sel[f] ^ flip) *
// out_stride.
template <int NT = LMZ_THREADS>
__global__ __launch_bounds__(NT) void k_forward_ftab(
    const lmz_feature *__restrict__ feat,
    const int32_t *__restrict__ tile_feature,
    const double *__restrict__ params, const int32_t *__restrict__ sel,
    int flip, int nbuf, const int32_t *__restrict__ nmax,
    const float *__restrict__ coords, long long npix_total, int flags,
    const float *__restrict__ data, float *__restrict__ out,
    long long out_stride, double *__restrict__ chi2_tiles,
    const FTabPar *__restrict__ fpar, const double *__restrict__ ftab,
    const int32_t *__restrict__ list, int32_t *count) {
  __shared__ int next_w[2];
  __shared__ double stab[LMZ_FTAB_LDS * LMZ_FTAB_STRIDE];
  TileCursor cursor;
  int staged = -1, sb = 0, se = 0;   // stab holds rows sb .. se - 1 of staged's table
  const int nwork = count[0];
  for (int w = next_tile(count, next_w, cursor); w < nwork;
       w = next_tile(count, next_w, cursor)) {
    const int t = list[w];
    const int f = tile_feature[t];
    const lmz_feature &F = feat[f];
    const int b = sel ? (sel[f] ^ flip) : 0;
    const int N = nmax[(long long)f * nbuf + b];
    const FeatUniform fu = feature_uniform(
        F, params + (long long)LMZ_NPARAM * f, N,
        F.coef_offset + b * F.nmax_cap);
    const bool bohren = flags & LMZ_BOHREN;
    const FTabPar fp = fpar[f];
    const double *gtab = ftab + (long long)f * (LMZ_FTAB_CAP * LMZ_FTAB_STRIDE);
    constexpr int PT = LMZ_TILE / NT;   // pixels per thread and tile
    const int ng = min(LMZ_FTAB_GROUP, F.ntiles - (t - fu.tile0));
    const int jp0 = (t - fu.tile0) * LMZ_TILE;
    // the data of a tile are loaded one tile ahead, before the previous
    // tile's residual stores: a load issued after a store waits for it
    // (one memory counter)
    float dcur[PT];
#pragma unroll
    for (int p = 0; p < PT; ++p) {
        const int j = jp0 + p * NT + threadIdx.x;
        dcur[p] = j < fu.npix ? data[fu.pix_offset + j] : 0.f;
    }
    // the rows the group reads: the whole table, or, for a table larger
    // than LDS, the window of the group's pixels
    int lo = 0, hi = fp.m - 1;
    if (fp.m > LMZ_FTAB_LDS) {
        table_rows<LMZ_FTAB_P, double>(fu.width, jp0,
                                       min(jp0 + ng * LMZ_TILE, fu.npix), fu.kd,
                                       (double)fu.ix + fu.fxd,
                                       (double)fu.iy + fu.fyd, fp.rho0,
                                       fp.inv_h, fp.m, lo, hi);
    }
    const bool lds = hi - lo < LMZ_FTAB_LDS;
    if (lds && !(staged == f && lo >= sb && hi < se)) {
        // every thread has finished the previous tile (the barriers of the
        // reduction and of next_tile)
        sb = fp.m > LMZ_FTAB_LDS ? table_window(lo, hi, fp.m, LMZ_FTAB_LDS) : 0;
        se = min(sb + LMZ_FTAB_LDS, fp.m);
        const double2 *src = reinterpret_cast<const double2 *>(
            gtab + (long long)sb * LMZ_FTAB_STRIDE);
        double2 *dst = reinterpret_cast<double2 *>(stab);
        for (int i = threadIdx.x; i < (se - sb) * (LMZ_FTAB_STRIDE / 2); i += NT)
            dst[i] = src[i];
        __syncthreads();
        staged = f;
    }
    float *outf = out ? out + b * out_stride : nullptr;
    // one pixel at a time (registers for more waves); each tile's r^T r
    // reduced at its end, as a tile of its own
#pragma unroll 1
    for (int q = 0; q < ng; ++q) {
        float dnext[PT];
#pragma unroll
        for (int p = 0; p < PT; ++p) {
            const int j = jp0 + (q + 1) * LMZ_TILE + p * NT + threadIdx.x;
            dnext[p] = (q + 1 < ng && j < fu.npix) ? data[fu.pix_offset + j] : 0.f;
        }
        double chi = 0.0;
#pragma unroll 1
        for (int p = 0; p < PT; ++p) {
            const int j = jp0 + q * LMZ_TILE + p * NT + threadIdx.x;
            if (j >= fu.npix) break;
            float dq = 0.f;   // dcur[p] by unrolled selects, not a run-time index
#pragma unroll
            for (int i = 0; i < PT; ++i) dq = (i == p) ? dcur[i] : dq;
            float ax, ay, az;
            pixel_coords(fu, j, coords, npix_total, ax, ay, az);
            const long long gi = fu.pix_offset + j;
            PixGeomD g;
            pixel_geometry_d(fu, ax, ay, az, bohren, g);
            const SumsD S = lds ? forward_sum_ftab<LMZ_FTAB_STRIDE>(g, fp, stab, sb)
                                : forward_sum_ftab<LMZ_FTAB_STRIDE>(g, fp, gtab);
            cd E[3];
            field_from_sums(g, S, true, E);
            const double I = intensity(E, fu.alphad);
            const double r = (I - (double)dq) / F.noise;
            if (outf) outf[gi] = (float)r;
            chi = fma(r, r, chi);
        }
        double v[1] = {chi};
        block_reduce_store_d<1, NT>(v, chi2_tiles + t + q);
        __syncthreads();   // the reduction buffer is reused by the next tile
#pragma unroll
        for (int p = 0; p < PT; ++p) dcur[p] = dnext[p];
    }
  }
}

// Hologram of one pixel in FP64 (the residual kernel's arithmetic) for the
// particle described by fu with coefficient table cf; E receives the field
// without alpha.
__device__ inline double hologram_d(const FeatUniform &fu, const float *cf,
                                    float ax, float ay, float az, bool bohren,
                                    bool cart, cd (&E)[3]) {
    PixGeomD g;
    pixel_geometry_d(fu, ax, ay, az, bohren, g);
    const SumsD S = forward_sum(g, cf, fu.nmax);
    field_from_sums(g, S, cart, E);
    return intensity(E, fu.alphad);
}

// Steps and perturbed parameter rows of a forward-difference Jacobian, as
// MINPACK's lmdif (scipy least_squares method 'lm' with diff_step, called
// at analysis/Optimizer.py:221) takes them: h_j = eps |p_j|, or eps when
// p_j = 0.  Rows fd_params[c][f] = p + h_{3+c} e_{3+c} for a_p, n_p, k_p,
// whose coefficient tables differ from the base point's.
__global__ __launch_bounds__(64) void k_fd_setup(
    int nfeat, const double *__restrict__ params, double eps,
    const uint8_t *__restrict__ active, double *__restrict__ fd_h,
    double *__restrict__ fd_params) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfeat) return;
    if (active && !active[f]) return;
    const double *p = params + (long long)LMZ_NPARAM * f;
    for (int c = 0; c < LMZ_NPARAM; ++c) {
        double h = eps * fabs(p[c]);
        if (h == 0.0) h = eps;
        fd_h[(long long)LMZ_NPARAM * f + c] = h;
    }
    for (int q = 0; q < 3; ++q) {
        double *row = fd_params + ((long long)q * nfeat + f) * LMZ_NPARAM;
        for (int c = 0; c < LMZ_NPARAM; ++c) row[c] = p[c];
        row[3 + q] += fd_h[(long long)LMZ_NPARAM * f + 3 + q];
    }
}

// Forward-difference Jacobian over one tile: the FP64 hologram at p and at
// p + h_j e_j for every varied j (x_p, y_p, z_p move the particle, a_p,
// n_p, k_p use the coefficient tables of their perturbed rows, alpha needs
// only the base field), J_j = (I_j - I) / h_j, reduced to the same per-tile
// normal-equation partials as k_jacobian.  The finite-difference form the
// analytic Jacobian replaces, inside the same batched LM.
template <int NT = LMZ_THREADS>
__global__ __launch_bounds__(NT) void k_fd_jacobian(
    const lmz_feature *__restrict__ feat,
    const int32_t *__restrict__ tile_feature, int nfeat,
    const double *__restrict__ params, const int32_t *__restrict__ sel,
    int flip, int nbuf, const int32_t *__restrict__ nmax,
    const float *__restrict__ coef, const double *__restrict__ fd_params,
    const int32_t *__restrict__ fd_nmax, const float *__restrict__ fd_coef,
    long long fd_coef_stride, const double *__restrict__ fd_h, int vary,
    const float *__restrict__ coords, long long npix_total, int flags,
    const float *__restrict__ rin, long long rin_stride,
    double *__restrict__ partials, const int32_t *__restrict__ list,
    int32_t *count) {
  __shared__ int next_w[2];
  TileCursor cursor;
  const int nwork = count ? count[0] : (int)gridDim.x;
  for (int w = list ? next_tile(count, next_w, cursor)
                    : xcd_tile((int)blockIdx.x, (int)gridDim.x);
       w < nwork;
       w = list ? next_tile(count, next_w, cursor) : nwork) {
    const int t = list ? list[w] : w;
    const int f = tile_feature[t];
    const lmz_feature &F = feat[f];
    const int b = sel ? (sel[f] ^ flip) : 0;
    const int N = nmax[(long long)f * nbuf + b];
    const double *p = params + (long long)LMZ_NPARAM * f;
    const double *h = fd_h + (long long)LMZ_NPARAM * f;
    const int slot = F.coef_offset + b * F.nmax_cap;
    const FeatUniform fu = feature_uniform(F, p, N, slot);
    const bool cart = flags & LMZ_CARTESIAN;
    const bool bohren = flags & LMZ_BOHREN;
    const float *cf0 = coef + (long long)LMZ_COEF_STRIDE * slot;
    const float *rinf = rin + b * rin_stride;
    float acc[LMZ_NRED];
#pragma unroll
    for (int i = 0; i < LMZ_NRED; ++i) acc[i] = 0.f;
#pragma unroll 1
    for (int q = 0; q < LMZ_TILE / NT; ++q) {
        const int j = (t - fu.tile0) * LMZ_TILE + q * NT + threadIdx.x;
        if (j >= fu.npix) break;
        float ax, ay, az;
        pixel_coords(fu, j, coords, npix_total, ax, ay, az);
        cd E[3], Ej[3];
        const double I0 = hologram_d(fu, cf0, ax, ay, az, bohren, cart, E);
        float Jn[LMZ_NJAC];
#pragma unroll 1
        for (int c = 0; c < 3; ++c) {          // x_p, y_p, z_p
            Jn[c] = 0.f;
            if (!(vary & (1 << c))) continue;
            double pc[LMZ_NPARAM];
            for (int i = 0; i < LMZ_NPARAM; ++i) pc[i] = p[i];
            pc[c] += h[c];
            const FeatUniform fc = feature_uniform(F, pc, N, slot);
            const double Ic = hologram_d(fc, cf0, ax, ay, az, bohren, cart, Ej);
            Jn[c] = (float)((Ic - I0) / h[c] * (double)fu.inv_noise);
        }
#pragma unroll 1
        for (int c = 3; c < 6; ++c) {          // a_p, n_p, k_p
            Jn[c] = 0.f;
            if (!(vary & (1 << c))) continue;
            const int qn = c - 3;
            const int Nc = fd_nmax[(long long)qn * nfeat * nbuf + (long long)f * nbuf];
            const FeatUniform fc = feature_uniform(
                F, fd_params + ((long long)qn * nfeat + f) * LMZ_NPARAM, Nc,
                F.coef_offset);
            const float *cfc = fd_coef + qn * fd_coef_stride +
                               (long long)LMZ_COEF_STRIDE * F.coef_offset;
            const double Ic = hologram_d(fc, cfc, ax, ay, az, bohren, cart, Ej);
            Jn[c] = (float)((Ic - I0) / h[c] * (double)fu.inv_noise);
        }
        Jn[6] = 0.f;
        if (vary & (1 << 6))                   // alpha: the same field
            Jn[6] = (float)((intensity(E, fu.alphad + h[6]) - I0) / h[6] *
                            (double)fu.inv_noise);
        const long long gi = fu.pix_offset + j;
        const float r = rinf[gi];
        int idx = 0;
#pragma unroll
        for (int a = 0; a < LMZ_NJAC; ++a)
#pragma unroll
            for (int c = a; c < LMZ_NJAC; ++c, ++idx)
                acc[idx] = fmaf(Jn[a], Jn[c], acc[idx]);
#pragma unroll
        for (int a = 0; a < LMZ_NJAC; ++a) acc[28 + a] = fmaf(Jn[a], r, acc[28 + a]);
        acc[35] = fmaf(r, r, acc[35]);
    }
    block_reduce_store<LMZ_NRED, NT>(acc, partials + (long long)LMZ_NRED * t);
    __syncthreads();   // the reduction buffer is reused by the next tile
  }
}

// Superposed field of several particles per feature, FP64, in one pass
// over the pixels (theory/LorenzMie.py:160-181 loops over particles with a
// full pass each): particles part_offset[f] .. part_offset[f + 1] - 1,
// rows of pparams (x_p y_p z_p absolute, a_p n_p k_p, -, -) with
// coefficient tables at pcoef + LMZ_COEF_STRIDE * i * pcap.  When spherical
// is not 0 each particle's field is multiplied by its aberration mask
// exp(i (4 n_m^3 / k) s r^4 / (r^2 + z_p^2)^2) (theory/Aberrated.py:61-84),
// r the in-plane pixel distance to the particle.
// MODE 0: field alpha E, 1: hologram |x + alpha E|^2.
template <int MODE>
__global__ __launch_bounds__(LMZ_THREADS) void k_superpose(
    const lmz_feature *__restrict__ feat,
    const int32_t *__restrict__ tile_feature,
    const int32_t *__restrict__ part_offset,
    const double *__restrict__ pparams, const int32_t *__restrict__ pnmax,
    const float *__restrict__ pcoef, int pcap,
    const float *__restrict__ coords, long long npix_total, int flags,
    double spherical, const double *__restrict__ alpha,
    float *__restrict__ out, cf *__restrict__ field) {
    const int t = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int f = tile_feature[t];
    const lmz_feature &F = feat[f];
    const int p0 = part_offset[f], p1 = part_offset[f + 1];
    const bool cart = flags & LMZ_CARTESIAN;
    const bool bohren = flags & LMZ_BOHREN;
    const double al = alpha ? alpha[f] : 1.0;
    const double kk = wavenumber_px(F);
    const double ab_scale = 4.0 * F.n_m * F.n_m * F.n_m / kk * spherical;
#pragma unroll 1
    for (int q = 0; q < LMZ_PPT; ++q) {
        const int j = (t - F.tile_offset) * LMZ_TILE + q * LMZ_THREADS + threadIdx.x;
        if (j >= F.npix) break;
        cd E[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
        float ax = 0.f, ay = 0.f, az = 0.f;
#pragma unroll 1
        for (int i = p0; i < p1; ++i) {
            const double *pp = pparams + (long long)LMZ_NPARAM * i;
            const FeatUniform fu = feature_uniform(F, pp, pnmax[i], 0);
            if (i == p0) pixel_coords(fu, j, coords, npix_total, ax, ay, az);
            PixGeomD g;
            pixel_geometry_d(fu, ax, ay, az, bohren, g);
            const SumsD S = forward_sum(
                g, pcoef + (long long)LMZ_COEF_STRIDE * i * pcap, fu.nmax);
            cd Ei[3];
            field_from_sums(g, S, cart, Ei);
            cd m{1.0, 0.0};
            if (spherical != 0.0) {
                const double dx = ((double)ax - (double)fu.ix) - fu.fxd;
                const double dy = ((double)ay - (double)fu.iy) - fu.fyd;
                const double rsq = fma(dx, dx, dy * dy);
                const double den = rsq + pp[2] * pp[2];
                double sn, cs;
                sincos(ab_scale * rsq * rsq / (den * den), &sn, &cs);
                m = cd{cs, sn};
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) E[c] = E[c] + m * Ei[c];
        }
        const long long gi = F.pix_offset + j;
        if (MODE == 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c)
                field[c * npix_total + gi] = cf{(float)(al * E[c].re),
                                                (float)(al * E[c].im)};
        } else {
            out[gi] = (float)intensity(E, al);
        }
    }
}

// Vector-Jacobian product over one tile: sum of g dI/d(param) for the
// seven parameters, J formed per pixel in registers and never written.
__global__ __launch_bounds__(LMZ_THREADS) void k_vjp(
    const lmz_feature *__restrict__ feat,
    const int32_t *__restrict__ tile_feature,
    const double *__restrict__ params, int nbuf,
    const int32_t *__restrict__ nmax, const float *__restrict__ coef,
    const float *__restrict__ dcoef, const float *__restrict__ coords,
    long long npix_total, int flags, const float *__restrict__ g,
    double *__restrict__ tile_grad) {
    const int t = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int f = tile_feature[t];
    const lmz_feature &F = feat[f];
    const FeatUniform fu = feature_uniform(
        F, params + (long long)LMZ_NPARAM * f, nmax[(long long)f * nbuf],
        F.coef_offset);
    const bool bohren = flags & LMZ_BOHREN;
    const float *cf0 = coef + (long long)LMZ_COEF_STRIDE * fu.slot0;
    const float *df0 = dcoef + (long long)LMZ_DCOEF_STRIDE * fu.slot0;
    float acc[LMZ_NJAC];
#pragma unroll
    for (int c = 0; c < LMZ_NJAC; ++c) acc[c] = 0.f;
#pragma unroll 1
    for (int q = 0; q < LMZ_PPT; ++q) {
        const int j = (t - fu.tile0) * LMZ_TILE + q * LMZ_THREADS + threadIdx.x;
        if (j >= fu.npix) break;
        float ax, ay, az;
        pixel_coords(fu, j, coords, npix_total, ax, ay, az);
        PixGeom pg;
        pixel_geometry<true>(fu, ax, ay, az, bohren, pg);
        JacPixel jp;
        jacobian_pixel(fu, pg, cf0, df0, fu.nmax, jp);
        const float gv = g[fu.pix_offset + j];
#pragma unroll
        for (int c = 0; c < LMZ_NJAC; ++c) acc[c] = fmaf(jp.J[c], gv, acc[c]);
    }
    block_reduce_store<LMZ_NJAC>(acc, tile_grad + (long long)LMZ_NJAC * t);
}

// Per-feature sums of per-tile rows of NV doubles, tiles in order (the
// same result on every run): one thread per (feature, entry).
template <int NV>
__global__ __launch_bounds__(LMZ_THREADS) void k_tile_sum(
    const lmz_feature *__restrict__ feat, int nfeat,
    const double *__restrict__ tiles, double *__restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nfeat * NV) return;
    const int f = (int)(i / NV), c = (int)(i % NV);
    const lmz_feature &F = feat[f];
    double s = 0.0;
    for (int k = 0; k < F.ntiles; ++k)
        s += tiles[(long long)NV * (F.tile_offset + k) + c];
    out[i] = s;
}

// Normal equations from J and r in HBM: one block per feature.
template <int NT>
__global__ __launch_bounds__(NT) void k_normal_eq(
    const lmz_feature *__restrict__ feat, const float *__restrict__ jac,
    const float *__restrict__ resid, long long npix_total,
    double *__restrict__ out) {
    const int f = blockIdx.x;
    const lmz_feature &F = feat[f];
    const float inv_noise = (float)(1.0 / F.noise);
    float acc[LMZ_NRED];
#pragma unroll
    for (int i = 0; i < LMZ_NRED; ++i) acc[i] = 0.f;
    for (int j = threadIdx.x; j < F.npix; j += NT) {
        const long long gi = F.pix_offset + j;
        float Jn[LMZ_NJAC];
#pragma unroll
        for (int c = 0; c < LMZ_NJAC; ++c) Jn[c] = jac[c * npix_total + gi] * inv_noise;
        const float r = resid[gi];
        int idx = 0;
#pragma unroll
        for (int a = 0; a < LMZ_NJAC; ++a)
#pragma unroll
            for (int c = a; c < LMZ_NJAC; ++c, ++idx)
                acc[idx] = fmaf(Jn[a], Jn[c], acc[idx]);
#pragma unroll
        for (int a = 0; a < LMZ_NJAC; ++a) acc[28 + a] = fmaf(Jn[a], r, acc[28 + a]);
        acc[35] = fmaf(r, r, acc[35]);
    }
    block_reduce_store<LMZ_NRED, NT>(acc, out + (long long)LMZ_NRED * f);
}

// |x + alpha E|^2 of an accumulated field, one thread per pixel.
__global__ __launch_bounds__(LMZ_THREADS) void k_intensity(
    const lmz_feature *__restrict__ feat,
    const int32_t *__restrict__ tile_feature,
    const double *__restrict__ params, const cf *__restrict__ field,
    long long npix_total, float *__restrict__ out) {
    const int t = blockIdx.x;
    const int f = tile_feature[t];
    const lmz_feature &F = feat[f];
    for (int q = 0; q < LMZ_PPT; ++q) {
        const int j = (t - F.tile_offset) * LMZ_TILE + q * LMZ_THREADS + threadIdx.x;
        if (j >= F.npix) break;
        const long long gi = F.pix_offset + j;
        // the field already carries alpha (lmz_field writes alpha E)
        cd E[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const cf e = field[c * npix_total + gi];
            E[c] = cd{(double)e.re, (double)e.im};
        }
        out[gi] = (float)intensity(E, 1.0);
    }
}

// ---------------------------------------------------------------- launchers
// Host-side launchers shared with the LM translation unit.
void launch_mie(const lmz_batch *b, const double *params, const int32_t *sel,
                int flip, int32_t *nmax, float *coef, float *dcoef,
                double *ab_out, double *work, const uint8_t *active,
                hipStream_t stream) {
    hipLaunchKernelGGL(k_mie_chains, dim3((b->nfeat + 63) / 64, 2), dim3(64), 0,
                       stream, b->d_feat, b->nfeat, b->cap_max, params, work,
                       active);
    hipLaunchKernelGGL(k_mie, dim3(b->nfeat), dim3(64), 0, stream, b->d_feat,
                       b->nfeat, b->cap_max, params, sel, flip, b->coef_buffers,
                       nmax, coef, dcoef, ab_out, work, active);
}

void launch_mie_layered(const lmz_batch *b, const double *params,
                        const int32_t *sel, int flip, const double *layers,
                        const int32_t *nlayers, int32_t *nmax, float *coef,
                        float *dcoef, double *work, const uint8_t *active,
                        hipStream_t stream) {
    hipLaunchKernelGGL(k_mie_layered_fit,
                       dim3((b->nfeat + 63) / 64, dcoef ? LMZ_WORK_PASSES : 1),
                       dim3(64), 0, stream, b->d_feat, b->nfeat, params, sel, flip,
                       b->coef_buffers, layers, nlayers, nmax, coef, dcoef,
                       work, active);
}

// Pixel-kernel launches at the batch's threads per tile (64, 128, 256).
template <int MODE, typename... Args>
static void launch_forward_nt(int nt, int grid, hipStream_t stream,
                              Args... args) {
    if (nt == 64)
        hipLaunchKernelGGL((k_forward<MODE, 64>), dim3(grid), dim3(64), 0,
                           stream, args...);
    else if (nt == 128)
        hipLaunchKernelGGL((k_forward<MODE, 128>), dim3(grid), dim3(128), 0,
                           stream, args...);
    else
        hipLaunchKernelGGL((k_forward<MODE, 256>), dim3(grid), dim3(256), 0,
                           stream, args...);
}

template <typename... Args>
static void launch_jacobian_nt(int nt, int grid, hipStream_t stream,
                               Args... args) {
    if (nt == 64)
        hipLaunchKernelGGL((k_jacobian<64>), dim3(grid), dim3(64), 0, stream,
                           args...);
    else if (nt == 128)
        hipLaunchKernelGGL((k_jacobian<128>), dim3(grid), dim3(128), 0,
                           stream, args...);
    else
        hipLaunchKernelGGL((k_jacobian<256>), dim3(grid), dim3(256), 0,
                           stream, args...);
}

void launch_residual(const lmz_batch *b, const double *params,
                     const int32_t *sel, int flip, const int32_t *nmax,
                     const float *coef, const float *coords, int flags,
                     const float *data, float *out, long long out_stride,
                     double *chi2_tiles, const uint8_t *active,
                     const TileList *tl, hipStream_t stream) {
    const int grid = tl ? tl->grid : b->ntiles;
    if (flags & LMZ_FP32) {
        launch_forward_nt<3>(b->sched.fwd_threads, grid, stream, b->d_feat,
                             b->d_tile_feature, params, sel, flip,
                             b->coef_buffers, nmax, coef, coords,
                             b->npix_total, flags, data, out, out_stride,
                             (cf *)nullptr, chi2_tiles, active,
                             tl ? (const int32_t *)tl->list : nullptr,
                             tl ? tl->count : nullptr);
        return;
    }
    launch_forward_nt<2>(b->sched.fwd_threads, grid, stream, b->d_feat,
                         b->d_tile_feature, params, sel, flip, b->coef_buffers,
                         nmax, coef, coords, b->npix_total, flags, data, out,
                         out_stride, (cf *)nullptr, chi2_tiles, active,
                         tl ? (const int32_t *)tl->list : nullptr,
                         tl ? tl->count : nullptr);
}

void launch_jacobian(const lmz_batch *b, const double *params,
                     const int32_t *sel, int flip, const int32_t *nmax,
                     const float *coef, const float *dcoef,
                     const float *coords, int flags, const float *data,
                     const float *rin, long long rin_stride, float *jac,
                     float *resid, double *partials, const uint8_t *active,
                     const TileList *tl, hipStream_t stream) {
    const int grid = tl ? tl->grid : b->ntiles;
    launch_jacobian_nt(b->sched.jac_threads, grid, stream, b->d_feat,
                       b->d_tile_feature, params, sel, flip, b->coef_buffers,
                       nmax, coef, dcoef, coords, b->npix_total, flags, data,
                       rin, rin_stride, jac, resid, partials, active,
                       tl ? (const int32_t *)tl->list : nullptr,
                       tl ? tl->count : nullptr);
}

void launch_jac_rtab(const lmz_batch *b, const double *params,
                     const int32_t *sel, int flip, const int32_t *nmax,
                     const float *coef, const float *dcoef, int flags,
                     const uint8_t *active, RTabPar *rpar, float *rtab,
                     uint8_t *mask_tab, uint8_t *mask_loop,
                     hipStream_t stream) {
    hipLaunchKernelGGL(k_jac_rtab, dim3(b->nfeat, LMZ_RTAB_CAP / LMZ_TAB_THREADS),
                       dim3(LMZ_TAB_THREADS),
                       0, stream, b->d_feat, params, sel, flip,
                       b->coef_buffers, nmax, coef, dcoef, flags, active, rpar,
                       rtab, mask_tab, mask_loop);
}

void launch_fwd_ftab(const lmz_batch *b, const double *params,
                     const int32_t *sel, int flip, const int32_t *nmax,
                     const float *coef, int flags, const uint8_t *active,
                     FTabPar *fpar, double *ftab, uint8_t *mask_tab,
                     uint8_t *mask_loop, hipStream_t stream) {
    hipLaunchKernelGGL(k_fwd_ftab, dim3(b->nfeat, LMZ_FTAB_CAP / LMZ_TAB_THREADS),
                       dim3(LMZ_TAB_THREADS),
                       0, stream, b->d_feat, params, sel, flip,
                       b->coef_buffers, nmax, coef, flags, active, fpar, ftab,
                       mask_tab, mask_loop);
}

void launch_forward_ftab(const lmz_batch *b, const double *params,
                         const int32_t *sel, int flip, const int32_t *nmax,
                         int flags, const float *data, float *out,
                         long long out_stride, double *chi2_tiles,
                         const FTabPar *fpar, const double *ftab,
                         const TileList *tl, hipStream_t stream) {
    hipLaunchKernelGGL((k_forward_ftab<LMZ_THREADS>), dim3(tl->grid),
                       dim3(LMZ_THREADS), 0, stream, b->d_feat,
                       b->d_tile_feature, params, sel, flip, b->coef_buffers,
                       nmax, (const float *)nullptr, b->npix_total, flags, data,
                       out, out_stride, chi2_tiles, fpar, ftab,
                       (const int32_t *)tl->list, tl->count);
}

void launch_jacobian_rtab(const lmz_batch *b, const double *params,
                          const int32_t *sel, int flip, const int32_t *nmax,
                          const float *coords, int flags, const float *data,
                          const float *rin, long long rin_stride,
                          double *partials, const RTabPar *rpar,
                          const float *rtab, const TileList *tl,
                          hipStream_t stream) {
    hipLaunchKernelGGL((k_jacobian_rtab<LMZ_THREADS>), dim3(tl->grid),
                       dim3(LMZ_THREADS), 0, stream, b->d_feat,
                       b->d_tile_feature, params, sel, flip, b->coef_buffers,
                       nmax, coords, b->npix_total, flags, data, rin,
                       rin_stride, partials, rpar, rtab,
                       (const int32_t *)tl->list, tl->count);
}

void launch_fd_jacobian(const lmz_batch *b, const double *params,
                        const int32_t *sel, int flip, const int32_t *nmax,
                        const float *coef, const FdTables &fd, int vary,
                        const float *coords, int flags, const float *rin,
                        long long rin_stride, double *partials,
                        const uint8_t *active, const TileList *tl,
                        double *work, hipStream_t stream) {
    hipLaunchKernelGGL(k_fd_setup, dim3((b->nfeat + 63) / 64), dim3(64), 0,
                       stream, b->nfeat, params, fd.eps, active, fd.h,
                       fd.params);
    // coefficient tables of the a_p, n_p, k_p rows (the params rows of a
    // feature are contiguous per set, so the Mie kernel reads set q as its
    // params array)
    for (int q = 0; q < 3; ++q) {
        if (!(vary & (1 << (3 + q)))) continue;
        const double *pq = fd.params + (long long)q * b->nfeat * LMZ_NPARAM;
        int32_t *nq = fd.nmax + (long long)q * b->nfeat * b->coef_buffers;
        if (fd.layers)
            launch_mie_layered(b, pq, nullptr, 0, fd.layers, fd.nlayers, nq,
                               fd.coef + q * fd.coef_stride, nullptr, work,
                               active, stream);
        else
            launch_mie(b, pq, nullptr, 0, nq, fd.coef + q * fd.coef_stride,
                       nullptr, nullptr, work, active, stream);
    }
    const int grid = tl ? tl->grid : b->ntiles;
    const int nt = b->sched.jac_threads;
    const int32_t *list = tl ? (const int32_t *)tl->list : nullptr;
    int32_t *count = tl ? tl->count : nullptr;
    auto go = [&](auto kern, int threads) {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), 0, stream, b->d_feat,
                           b->d_tile_feature, b->nfeat, params, sel, flip,
                           b->coef_buffers, nmax, coef, fd.params, fd.nmax,
                           fd.coef, fd.coef_stride, fd.h, vary, coords,
                           b->npix_total, flags, rin, rin_stride, partials,
                           list, count);
    };
    if (nt == 64) go(k_fd_jacobian<64>, 64);
    else if (nt == 128) go(k_fd_jacobian<128>, 128);
    else go(k_fd_jacobian<256>, 256);
}

// Compacted list of the tiles of the features whose mask is set, in feature
// order, and its length, in two launches.  k_tile_starts (one block):
// thread t counts the tiles of features [t per, (t + 1) per), a block scan
// gives each feature's first list position (start[f]) and the length.
// k_tile_fill: one wavefront per feature writes its tile range, the tiles
// across the lanes, so the stores are contiguous and every feature's range
// is written in parallel.
__global__ __launch_bounds__(1024) void k_tile_starts(
    const lmz_feature *__restrict__ feat, int nfeat,
    const uint8_t *__restrict__ mask, int32_t *__restrict__ start,
    int32_t *__restrict__ count, int group) {
    __shared__ int scan[1024];
    const int tid = threadIdx.x;
    const int per = (nfeat + 1023) / 1024;
    const int f0 = min(nfeat, tid * per), f1 = min(nfeat, f0 + per);
    int mine = 0;
    for (int f = f0; f < f1; ++f)
        if (mask[f]) mine += (feat[f].ntiles + group - 1) / group;
    scan[tid] = mine;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = tid >= off ? scan[tid - off] : 0;
        __syncthreads();
        scan[tid] += v;
        __syncthreads();
    }
    int pos = scan[tid] - mine;
    for (int f = f0; f < f1; ++f) {
        start[f] = pos;
        if (mask[f]) pos += (feat[f].ntiles + group - 1) / group;
    }
    if (tid == 1023) count[0] = scan[1023];
    if (tid < LMZ_XCDS) count[2 + tid] = 0;   // fetch counters (next_tile)
}

__global__ __launch_bounds__(256) void k_tile_fill(
    const lmz_feature *__restrict__ feat, int nfeat,
    const uint8_t *__restrict__ mask, const int32_t *__restrict__ start,
    int32_t *__restrict__ list, int group) {
    const int f = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (f >= nfeat || !mask[f]) return;
    const int s0 = start[f], t0 = feat[f].tile_offset;
    const int ng = (feat[f].ntiles + group - 1) / group;
    for (int i = threadIdx.x & 63; i < ng; i += 64) list[s0 + i] = t0 + i * group;
}

void launch_tile_list(const lmz_batch *b, const uint8_t *mask,
                      const TileList &tl, hipStream_t stream, int group) {
    hipLaunchKernelGGL(k_tile_starts, dim3(1), dim3(1024), 0, stream, b->d_feat,
                       b->nfeat, mask, tl.start, tl.count, group);
    hipLaunchKernelGGL(k_tile_fill, dim3((b->nfeat + 3) / 4), dim3(256), 0, stream,
                       b->d_feat, b->nfeat, mask, tl.start, tl.list, group);
}

// Workgroups of one pixel kernel resident per CU at once.
static int occupancy(bool jacobian, int nt) {
    const void *fn;
    if (jacobian)
        fn = nt == 64 ? reinterpret_cast<const void *>(&k_jacobian<64>)
           : nt == 128 ? reinterpret_cast<const void *>(&k_jacobian<128>)
                       : reinterpret_cast<const void *>(&k_jacobian<256>);
    else
        fn = nt == 64 ? reinterpret_cast<const void *>(&k_forward<2, 64>)
           : nt == 128 ? reinterpret_cast<const void *>(&k_forward<2, 128>)
                       : reinterpret_cast<const void *>(&k_forward<2, 256>);
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, nt, 0) != hipSuccess)
        n = 1;
    return n > 0 ? n : 1;
}

static int device_cus(int *cus) {
    int dev = 0;
    LMZ_HIP(hipGetDevice(&dev));
    LMZ_HIP(hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev));
    return LMZ_OK;
}

// Persistent grid sizes: the schedule's workgroups per CU, by default as
// many as are resident at once.
int persistent_grids(const lmz_batch *b, int *grid_jacobian, int *grid_forward) {
    int cus = 0;
    if (int rc = device_cus(&cus)) return rc;
    const lmz_schedule &sc = b->sched;
    const int bj = sc.jac_blocks_per_cu > 0 ? sc.jac_blocks_per_cu
                                            : occupancy(true, sc.jac_threads);
    const int bf = sc.fwd_blocks_per_cu > 0 ? sc.fwd_blocks_per_cu
                                            : occupancy(false, sc.fwd_threads);
    *grid_jacobian = cus * bj;
    *grid_forward = cus * bf;
    return LMZ_OK;
}

// Persistent grids of k_jacobian_rtab and k_forward_ftab: as many
// workgroups as are resident.
int rtab_grid(int *grid, int *grid_fwd) {
    int cus = 0;
    if (int rc = device_cus(&cus)) return rc;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n, reinterpret_cast<const void *>(&k_jacobian_rtab<LMZ_THREADS>),
            LMZ_THREADS, 0) != hipSuccess || n < 1)
        n = 1;
    *grid = cus * n;
    if (grid_fwd) {
        n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &n, reinterpret_cast<const void *>(&k_forward_ftab<LMZ_THREADS>),
                LMZ_THREADS, 0) != hipSuccess || n < 1)
            n = 1;
        *grid_fwd = cus * n;
    }
    return LMZ_OK;
}

// ---------------------------------------------------------------- autotuner
// mask[f] = 1 for every stride-th feature, else 0
__global__ void k_mask_every(uint8_t *__restrict__ mask, int n, int stride) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) mask[i] = (i % stride) == 0 ? 1 : 0;
}

struct TuneScratch {
    size_t mask, list, start, count, partials, chi2, total;
};

static TuneScratch tune_layout(const lmz_batch *b) {
    TuneScratch L{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = (off + bytes + 255) & ~size_t(255);
        return o;
    };
    L.mask = take((size_t)b->nfeat);
    L.list = take(sizeof(int32_t) * (size_t)b->ntiles);
    L.start = take(sizeof(int32_t) * (size_t)b->nfeat);
    L.count = take(sizeof(int32_t) * 16);
    L.partials = take(sizeof(double) * LMZ_NRED * (size_t)b->ntiles);
    L.chi2 = take(sizeof(double) * (size_t)b->ntiles);
    L.total = off;
    return L;
}

// Schedules already tuned in this process, by device and batch shape.
using TuneKey = std::tuple<int, int, int, int, int>;
static std::mutex g_tune_mutex;
static std::map<TuneKey, lmz_schedule> g_tune_cache;

}  // namespace lmz

using namespace lmz;

// ====================================================================
// C ABI
// ====================================================================
// Per-tile partial gradients of lmz_hologram_vjp_static: a fixed device
// array, so that the call allocates nothing.  Every GPU holds its own copy
// until the process ends; its address is looked up once per device.  Each
// call writes rows 0 .. ntiles - 1 before it reads them, so nothing is
// cleared; calls that overlap on one GPU race for it.
static __device__ __align__(256) double
    scratch[(size_t)LMZ_VJP_STATIC_MAX_TILES * LMZ_NJAC];

#define LMZ_MAX_DEVICES 64
static std::mutex g_scratch_mutex;
static double *g_scratch_ptr[LMZ_MAX_DEVICES];

// Device address of the current device's scratch (hipGetSymbolAddress on
// the first call on that device, with it current).
static int vjp_scratch(double **out) {
    int dev = 0;
    LMZ_HIP(hipGetDevice(&dev));
    LMZ_REQUIRE(dev >= 0 && dev < LMZ_MAX_DEVICES, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        LMZ_REQUIRE(p != nullptr, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return LMZ_OK;
}

// k_vjp over every tile into tile_work, then per-feature sums in order.
static int launch_vjp(const lmz_batch *b, const double *params,
                      const int32_t *nmax, const float *coef,
                      const float *dcoef, const float *coords, int32_t flags,
                      const float *grad_out, double *tile_work,
                      double *grad_params, hipStream_t stream) {
    hipLaunchKernelGGL(k_vjp, dim3(b->ntiles), dim3(LMZ_THREADS), 0, stream,
                       b->d_feat, b->d_tile_feature, params, b->coef_buffers,
                       nmax, coef, dcoef, coords, b->npix_total, flags,
                       grad_out, tile_work);
    LMZ_HIP(hipGetLastError());
    const long long n = (long long)b->nfeat * LMZ_NJAC;
    hipLaunchKernelGGL(k_tile_sum<LMZ_NJAC>,
                       dim3((unsigned)((n + LMZ_THREADS - 1) / LMZ_THREADS)),
                       dim3(LMZ_THREADS), 0, stream, b->d_feat, b->nfeat,
                       tile_work, grad_params);
    LMZ_HIP(hipGetLastError());
    return LMZ_OK;
}

extern "C" {

const char *lmz_last_error(void) { return g_error.c_str(); }
int lmz_version(void) { return LMZ_VERSION; }

int lmz_batch_create(int32_t nfeat, lmz_feature *feat, int32_t coef_buffers,
                     lmz_batch **out) {
    LMZ_REQUIRE(out != nullptr, "out is NULL");
    *out = nullptr;
    LMZ_REQUIRE(nfeat > 0, "nfeat must be positive");
    LMZ_REQUIRE(feat != nullptr, "feature table is NULL");
    LMZ_REQUIRE(coef_buffers == 1 || coef_buffers == 2,
                "coef_buffers must be 1 or 2");
    std::vector<int32_t> tile_feature;
    int64_t npix_total = 0, prev_end = 0;
    int32_t slots = 0, cap_max = 0, all_dense = 1;
    for (int32_t f = 0; f < nfeat; ++f) {
        lmz_feature &F = feat[f];
        char msg[160];
        if (F.npix <= 0 || F.pix_offset < prev_end || F.width < 0 ||
            F.width > LMZ_MAX_WIDTH || F.nmax_cap < 1 || F.nmax_cap > 4096 ||
            !(F.noise > 0.0) || !(F.wavelength > 0.0) || !(F.n_m > 0.0) ||
            !(F.magnification > 0.0) ||
            (F.width > 0 && F.npix >= (1 << 24))) {
            snprintf(msg, sizeof msg,
                     "invalid feature %d (npix %d, offset %lld, width %d, cap %d)",
                     f, F.npix, (long long)F.pix_offset, F.width, F.nmax_cap);
            set_error(msg);
            return LMZ_EINVAL;
        }
        if (F.width == 0) all_dense = 0;
        prev_end = F.pix_offset + F.npix;
        npix_total = std::max<int64_t>(npix_total, prev_end);
        F.coef_offset = slots;
        F.ntiles = (F.npix + LMZ_TILE - 1) / LMZ_TILE;
        F.tile_offset = (int32_t)tile_feature.size();
        for (int32_t i = 0; i < F.ntiles; ++i) tile_feature.push_back(f);
        const int64_t s = (int64_t)slots + (int64_t)F.nmax_cap * coef_buffers;
        LMZ_REQUIRE(s < (1LL << 30), "coefficient table too large");
        slots = (int32_t)s;
        cap_max = std::max(cap_max, F.nmax_cap);
        LMZ_REQUIRE(tile_feature.size() < (1ULL << 31), "too many tiles");
    }
    lmz_batch *b = new lmz_batch;
    b->nfeat = nfeat;
    b->ntiles = (int32_t)tile_feature.size();
    b->npix_total = npix_total;
    b->coef_buffers = coef_buffers;
    b->coef_slots = slots;
    b->cap_max = cap_max;
    b->all_dense = all_dense;
    hipError_t e = hipMalloc(&b->d_feat, sizeof(lmz_feature) * nfeat);
    if (e == hipSuccess)
        e = hipMalloc(&b->d_tile_feature, sizeof(int32_t) * tile_feature.size());
    if (e == hipSuccess)
        e = hipMemcpy(b->d_feat, feat, sizeof(lmz_feature) * nfeat,
                      hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_tile_feature, tile_feature.data(),
                      sizeof(int32_t) * tile_feature.size(),
                      hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (b->d_feat) (void)hipFree(b->d_feat);
        if (b->d_tile_feature) (void)hipFree(b->d_tile_feature);
        delete b;
        return hip_fail(e, "lmz_batch_create");
    }
    *out = b;
    return LMZ_OK;
}

int lmz_batch_destroy(lmz_batch *b) {
    if (!b) return LMZ_OK;
    if (b->d_feat) (void)hipFree(b->d_feat);
    if (b->d_tile_feature) (void)hipFree(b->d_tile_feature);
    delete b;
    return LMZ_OK;
}

int lmz_batch_query(const lmz_batch *b, lmz_batch_sizes *s) {
    LMZ_REQUIRE(b != nullptr && s != nullptr, "NULL argument");
    s->nfeat = b->nfeat;
    s->ntiles = b->ntiles;
    s->npix_total = b->npix_total;
    s->coef_buffers = b->coef_buffers;
    s->coef_slots = b->coef_slots;
    s->cap_max = b->cap_max;
    s->work_doubles = (long long)LMZ_WORK_PER_ORDER * LMZ_WORK_PASSES *
                      (b->cap_max + 1) * b->nfeat;
    s->coef_stride = LMZ_COEF_STRIDE;
    s->dcoef_stride = LMZ_DCOEF_STRIDE;
    return LMZ_OK;
}

int lmz_mie_ab(const lmz_batch *b, const double *params, const int32_t *sel,
               int32_t *nmax_out, float *coef, float *dcoef, double *ab_out,
               double *work, hipStream_t stream) {
    if (int rc = check_batch(b)) return rc;
    LMZ_REQUIRE(params && nmax_out && work, "params, nmax_out and work are required");
    LMZ_REQUIRE(coef || ab_out, "one of coef, ab_out is required");
    LMZ_REQUIRE(sel == nullptr || b->coef_buffers == 2,
                "sel needs a batch with two coefficient buffers");
    launch_mie(b, params, sel, 0, nmax_out, coef, dcoef, ab_out, work, nullptr,
               stream);
    LMZ_HIP(hipGetLastError());
    return LMZ_OK;
}

int lmz_mie_ab_layered(const lmz_batch *b, const double *layers,
                       const int32_t *nlayers, int32_t *nmax_out, float *coef,
                       double *ab_out, double *work, hipStream_t stream) {
    if (int rc = check_batch(b)) return rc;
    LMZ_REQUIRE(layers && nlayers && nmax_out && work,
                "layers, nlayers, nmax_out and work are required");
    LMZ_REQUIRE(coef || ab_out, "nothing to compute: coef and ab_out are NULL");
    hipLaunchKernelGGL(k_mie_layered, dim3((b->nfeat + 63) / 64), dim3(64), 0,
                       stream, b->d_feat, b->nfeat, layers, nlayers,
                       b->coef_buffers, nmax_out, coef, ab_out, work);
    LMZ_HIP(hipGetLastError());
    return LMZ_OK;
}

int lmz_mie_ab_layered_fit(const lmz_batch *b, const double *params,
                           const double *layers, const int32_t *nlayers,
                           int32_t *nmax_out, float *coef, float *dcoef,
                           double *work, hipStream_t stream) {
    if (int rc = check_batch(b)) return rc;
    LMZ_REQUIRE(params && layers && nlayers && nmax_out && coef && work,
                "params, layers, nlayers, nmax_out, coef and work are required");
    launch_mie_layered(b, params, nullptr, 0, layers, nlayers, nmax_out, coef,
                       dcoef, work, nullptr, stream);
    LMZ_HIP(hipGetLastError());
    return LMZ_OK;
}

int lmz_coef_from_ab(const lmz_batch *b, const double *ab_in,
                     const int32_t *nmax_in, int32_t *nmax_out, float *coef,
                     hipStream_t stream) {
    if (int rc = check_batch(b)) return rc;
    LMZ_REQUIRE(ab_in && nmax_in && nmax_out && coef,
                "ab_in, nmax_in, nmax_out and coef are required");
    const int grid = (b->nfeat + 63) / 64;
    hipLaunchKernelGGL(k_coef_from_ab, dim3(grid), dim3(64), 0, stream,
                       b->d_feat, b->nfeat, ab_in, nmax_in, b->coef_buffers,
                       nmax_out, coef);
    LMZ_HIP(hipGetLastError());
    return LMZ_OK;
}

static int forward_args(const lmz_batch *b, const double *params,
                        const int32_t *nmax, const float *coef,
                        const float *coords, const int32_t *sel) {
    if (int rc = check_batch(b)) return rc;
    LMZ_REQUIRE(params && nmax && coef, "params, nmax and coef are required");
    LMZ_REQUIRE(coords || b->all_dense,
                "coords is required when a feature has explicit coordinates");
    LMZ_REQUIRE(sel == nullptr || b->coef_buffers == 2,
                "sel needs a batch with two coefficient buffers");
    return LMZ_OK;
}

static int vjp_args(const lmz_batch *b, const double *params,
                    const int32_t *nmax, const float *coef, const float *dcoef,
                    const float *coords, int32_t flags) {
    if (int rc = forward_args(b, params, nmax, coef, coords, nullptr)) return rc;
    LMZ_REQUIRE(dcoef != nullptr, "dcoef (coefficient tangents) is required");
    LMZ_REQUIRE(flags & LMZ_CARTESIAN,
                "the Jacobian is defined for the Cartesian hologram");
    return LMZ_OK;
}

int lmz_field(const lmz_batch *b, const double *params, const int32_t *sel,
              const int32_t *nmax, const float *coef, const float *coords,
              int32_t flags, float *field, hipStream_t stream) {
    if (int rc = forward_args(b, params, nmax, coef, coords, sel)) return rc;
    LMZ_REQUIRE(field != nullptr, "field is NULL");
    launch_forward_nt<0>(b->sched.fwd_threads, b->ntiles, stream, b->d_feat,
                         b->d_tile_feature, params, sel, 0, b->coef_buffers,
                         nmax, coef, coords, b->npix_total, flags,
                         (const float *)nullptr, (float *)nullptr, 0LL,
                         reinterpret_cast<cf *>(field), (double *)nullptr,
                         (const uint8_t *)nullptr, (const int32_t *)nullptr,
                         (int32_t *)nullptr);
    LMZ_HIP(hipGetLastError());
    return LMZ_OK;
}

int lmz_hologram(const lmz_batch *b, const double *params, const int32_t *sel,
                 const int32_t *nmax, const float *coef, const float *coords,
                 int32_t flags, const float *data, float *out,
                 double *chi2_tiles, const uint8_t *active,
                 hipStream_t stream) {
    if (int rc = forward_args(b, params, nmax, coef, coords, sel)) return rc;
    if (data == nullptr && (flags & LMZ_FP32)) {
        LMZ_REQUIRE(out != nullptr, "out is NULL");
        LMZ_REQUIRE(flags & LMZ_CARTESIAN,
                    "the FP32 hologram is defined for the Cartesian field");
        launch_forward_nt<3>(b->sched.fwd_threads, b->ntiles, stream, b->d_feat,
                             b->d_tile_feature, params, sel, 0, b->coef_buffers,
                             nmax, coef, coords, b->npix_total, flags,
                             (const float *)nullptr, out, 0LL, (cf *)nullptr,
                             (double *)nullptr, active, (const int32_t *)nullptr,
                             (int32_t *)nullptr);
    } else if (data == nullptr) {
        LMZ_REQUIRE(out != nullptr, "out is NULL");
        launch_forward_nt<1>(b->sched.fwd_threads, b->ntiles, stream, b->d_feat,
                             b->d_tile_feature, params, sel, 0, b->coef_buffers,
                             nmax, coef, coords, b->npix_total, flags,
                             (const float *)nullptr, out, 0LL, (cf *)nullptr,
                             (double *)nullptr, active, (const int32_t *)nullptr,
                             (int32_t *)nullptr);
    } else {
        LMZ_REQUIRE(out || chi2_tiles, "one of out, chi2_tiles is required");
        LMZ_REQUIRE(!(flags & LMZ_FP32) || (flags & LMZ_CARTESIAN),
                    "FP32 residuals are defined for the Cartesian hologram");
        launch_residual(b, params, sel, 0, nmax, coef, coords, flags, data, out,
                        0LL, chi2_tiles, active, nullptr, stream);
    }
    LMZ_HIP(hipGetLastError());
    return LMZ_OK;
}

int lmz_hologram_superposed(const lmz_batch *b, int32_t nparticles,
                            const int32_t *part_offset, const double *pparams,
                            const int32_t *pnmax, const float *pcoef,
                            int32_t pcap, const float *coords, int32_t flags,
                            double spherical, const double *alpha, float *out,
                            float *field, hipStream_t stream) {
    if (int rc = check_batch(b)) return rc;
    LMZ_REQUIRE(nparticles >= 1 && pcap >= 1, "nparticles and pcap must be >= 1");
    LMZ_REQUIRE(part_offset && pparams && pnmax && pcoef,
                "part_offset, pparams, pnmax and pcoef are required");
    LMZ_REQUIRE(coords || b->all_dense,
                "coords is required when a feature has explicit coordinates");
    LMZ_REQUIRE((out != nullptr) != (field != nullptr),
                "exactly one of out (hologram) and field is required");
    LMZ_REQUIRE(std::isfinite(spherical), "spherical must be finite");
    if (out)
        hipLaunchKernelGGL(k_superpose<1>, dim3(b->ntiles), dim3(LMZ_THREADS), 0,
                           stream, b->d_feat, b->d_tile_feature, part_offset,
                           pparams, pnmax, pcoef, pcap, coords, b->npix_total,
                           flags, spherical, alpha, out, (cf *)nullptr);
    else
        hipLaunchKernelGGL(k_superpose<0>, dim3(b->ntiles), dim3(LMZ_THREADS), 0,
                           stream, b->d_feat, b->d_tile_feature, part_offset,
                           pparams, pnmax, pcoef, pcap, coords, b->npix_total,
                           flags, spherical, alpha, (float *)nullptr,
                           reinterpret_cast<cf *>(field));
    LMZ_HIP(hipGetLastError());
    return LMZ_OK;
}

int lmz_intensity(const lmz_batch *b, const double *params, const float *field,
                  int32_t flags, float *out, hipStream_t stream) {
    if (int rc = check_batch(b)) return rc;
    LMZ_REQUIRE(params && field && out, "params, field and out are required");
    (void)flags;
    hipLaunchKernelGGL(k_intensity, dim3(b->ntiles), dim3(LMZ_THREADS), 0,
                       stream, b->d_feat, b->d_tile_feature, params,
                       reinterpret_cast<const cf *>(field), b->npix_total, out);
    LMZ_HIP(hipGetLastError());
    return LMZ_OK;
}

int lmz_jacobian(const lmz_batch *b, const double *params, const int32_t *sel,
                 const int32_t *nmax, const float *coef, const float *dcoef,
                 const float *coords, int32_t flags, const float *data,
                 float *jac, float *resid, double *partials,
                 const uint8_t *active, hipStream_t stream) {
    if (int rc = forward_args(b, params, nmax, coef, coords, sel)) return rc;
    LMZ_REQUIRE(dcoef != nullptr, "dcoef (coefficient tangents) is required");
    LMZ_REQUIRE(flags & LMZ_CARTESIAN,
                "the Jacobian is defined for the Cartesian hologram");
    LMZ_REQUIRE(jac || ((resid || partials) && data),
                "nothing to compute: jac is NULL and no data for resid/partials");
    LMZ_REQUIRE(data || (!resid && !partials), "resid and partials need data");
    launch_jacobian(b, params, sel, 0, nmax, coef, dcoef, coords, flags, data,
                    nullptr, 0LL, jac, resid, partials, active, nullptr, stream);
    LMZ_HIP(hipGetLastError());
    return LMZ_OK;
}

// Workspace of lmz_jacobian_tabulated: tables, their grids, the two masks
// and tile lists, list counters, list starts.
struct TabLayout {
    size_t rpar, rtab, mask_t, mask_l, list_t, list_l, count, start, total;
};

static TabLayout tab_layout(const lmz_batch *b) {
    TabLayout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = (off + bytes + 255) & ~size_t(255);
        return o;
    };
    L.rpar = take(sizeof(RTabPar) * (size_t)b->nfeat);
    L.rtab = take(sizeof(float) * LMZ_RTAB_CAP * LMZ_RTAB_STRIDE * (size_t)b->nfeat);
    L.mask_t = take((size_t)b->nfeat);
    L.mask_l = take((size_t)b->nfeat);
    L.list_t = take(sizeof(int32_t) * (size_t)b->ntiles);
    L.list_l = take(sizeof(int32_t) * (size_t)b->ntiles);
    L.count = take(sizeof(int32_t) * 32);
    L.start = take(sizeof(int32_t) * (size_t)b->nfeat);
    L.total = off;
    return L;
}

int lmz_jacobian_tabulated_bytes(const lmz_batch *b, size_t *bytes) {
    LMZ_REQUIRE(bytes != nullptr, "bytes is NULL");
    if (int rc = check_batch(b)) return rc;
    *bytes = tab_layout(b).total;
    return LMZ_OK;
}

int lmz_jacobian_tabulated(const lmz_batch *b, const double *params,
                           const int32_t *nmax, const float *coef,
                           const float *dcoef, int32_t flags,
                           const float *data, double *partials, void *work,
                           hipStream_t stream) {
    if (int rc = forward_args(b, params, nmax, coef, nullptr, nullptr)) return rc;
    LMZ_REQUIRE(b->all_dense, "lmz_jacobian_tabulated needs dense crops");
    LMZ_REQUIRE(dcoef != nullptr, "dcoef (coefficient tangents) is required");
    LMZ_REQUIRE(flags & LMZ_CARTESIAN,
                "the Jacobian is defined for the Cartesian hologram");
    LMZ_REQUIRE(data && partials && work, "data, partials and work are required");
    LMZ_REQUIRE((reinterpret_cast<uintptr_t>(work) & 255) == 0,
                "work must be 256-byte aligned");
    int grid_j = 0, grid_f = 0, grid_t = 0;
    if (int rc = persistent_grids(b, &grid_j, &grid_f)) return rc;
    if (int rc = rtab_grid(&grid_t, nullptr)) return rc;
    const TabLayout L = tab_layout(b);
    char *ws = static_cast<char *>(work);
    RTabPar *rpar = reinterpret_cast<RTabPar *>(ws + L.rpar);
    float *rtab = reinterpret_cast<float *>(ws + L.rtab);
    uint8_t *mask_t = reinterpret_cast<uint8_t *>(ws + L.mask_t);
    uint8_t *mask_l = reinterpret_cast<uint8_t *>(ws + L.mask_l);
    int32_t *count = reinterpret_cast<int32_t *>(ws + L.count);
    int32_t *start = reinterpret_cast<int32_t *>(ws + L.start);
    const TileList tl_t{reinterpret_cast<int32_t *>(ws + L.list_t), count,
                        grid_t, start};
    const TileList tl_l{reinterpret_cast<int32_t *>(ws + L.list_l), count + 16,
                        grid_j, start};
    launch_jac_rtab(b, params, nullptr, 0, nmax, coef, dcoef, flags, nullptr,
                    rpar, rtab, mask_t, mask_l, stream);
    launch_tile_list(b, mask_t, tl_t, stream, LMZ_RTAB_GROUP);
    launch_jacobian_rtab(b, params, nullptr, 0, nmax, nullptr, flags, data,
                         nullptr, 0LL, partials, rpar, rtab, &tl_t, stream);
    launch_tile_list(b, mask_l, tl_l, stream);
    launch_jacobian(b, params, nullptr, 0, nmax, coef, dcoef, nullptr, flags,
                    data, nullptr, 0LL, nullptr, nullptr, partials, nullptr,
                    &tl_l, stream);
    LMZ_HIP(hipGetLastError());
    return LMZ_OK;
}

// Workspace of lmz_residual_tabulated: FP64 tables, grids, masks, lists.
struct FTabLayout {
    size_t fpar, ftab, mask_t, mask_l, list_t, list_l, count, start, total;
};

static FTabLayout ftab_layout(const lmz_batch *b) {
    FTabLayout L{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = (off + bytes + 255) & ~size_t(255);
        return o;
    };
    L.fpar = take(sizeof(FTabPar) * (size_t)b->nfeat);
    L.ftab = take(sizeof(double) * LMZ_FTAB_CAP * LMZ_FTAB_STRIDE * (size_t)b->nfeat);
    L.mask_t = take((size_t)b->nfeat);
    L.mask_l = take((size_t)b->nfeat);
    L.list_t = take(sizeof(int32_t) * (size_t)b->ntiles);
    L.list_l = take(sizeof(int32_t) * (size_t)b->ntiles);
    L.count = take(sizeof(int32_t) * 32);
    L.start = take(sizeof(int32_t) * (size_t)b->nfeat);
    L.total = off;
    return L;
}

int lmz_residual_tabulated_bytes(const lmz_batch *b, size_t *bytes) {
    LMZ_REQUIRE(bytes != nullptr, "bytes is NULL");
    if (int rc = check_batch(b)) return rc;
    *bytes = ftab_layout(b).total;
    return LMZ_OK;
}

int lmz_residual_tabulated(const lmz_batch *b, const double *params,
                           const int32_t *nmax, const float *coef,
                           int32_t flags, const float *data, float *resid,
                           double *chi2_tiles, void *work, hipStream_t stream) {
    if (int rc = forward_args(b, params, nmax, coef, nullptr, nullptr)) return rc;
    LMZ_REQUIRE(b->all_dense, "lmz_residual_tabulated needs dense crops");
    LMZ_REQUIRE(flags & LMZ_CARTESIAN, "the residual is of the Cartesian hologram");
    LMZ_REQUIRE(!(flags & LMZ_FP32), "the tabulated residual is FP64");
    LMZ_REQUIRE(data && chi2_tiles && work, "data, chi2_tiles and work are required");
    LMZ_REQUIRE((reinterpret_cast<uintptr_t>(work) & 255) == 0,
                "work must be 256-byte aligned");
    int grid_j = 0, grid_f = 0, grid_t = 0, grid_ft = 0;
    if (int rc = persistent_grids(b, &grid_j, &grid_f)) return rc;
    if (int rc = rtab_grid(&grid_t, &grid_ft)) return rc;
    const FTabLayout L = ftab_layout(b);
    char *ws = static_cast<char *>(work);
    FTabPar *fpar = reinterpret_cast<FTabPar *>(ws + L.fpar);
    double *ftab = reinterpret_cast<double *>(ws + L.ftab);
    uint8_t *mask_t = reinterpret_cast<uint8_t *>(ws + L.mask_t);
    uint8_t *mask_l = reinterpret_cast<uint8_t *>(ws + L.mask_l);
    int32_t *count = reinterpret_cast<int32_t *>(ws + L.count);
    int32_t *start = reinterpret_cast<int32_t *>(ws + L.start);
    const TileList tl_t{reinterpret_cast<int32_t *>(ws + L.list_t), count,
                        grid_ft, start};
    const TileList tl_l{reinterpret_cast<int32_t *>(ws + L.list_l), count + 16,
                        grid_f, start};
    launch_fwd_ftab(b, params, nullptr, 0, nmax, coef, flags, nullptr, fpar,
                    ftab, mask_t, mask_l, stream);
    launch_tile_list(b, mask_t, tl_t, stream, LMZ_FTAB_GROUP);
    launch_forward_ftab(b, params, nullptr, 0, nmax, flags, data, resid, 0LL,
                        chi2_tiles, fpar, ftab, &tl_t, stream);
    launch_tile_list(b, mask_l, tl_l, stream);
    launch_residual(b, params, nullptr, 0, nmax, coef, nullptr, flags, data,
                    resid, 0LL, chi2_tiles, nullptr, &tl_l, stream);
    LMZ_HIP(hipGetLastError());
    return LMZ_OK;
}

int lmz_hologram_vjp(const lmz_batch *b, const double *params,
                     const int32_t *nmax, const float *coef,
                     const float *dcoef, const float *coords, int32_t flags,
                     const float *grad_out, double *tile_work,
                     double *grad_params, hipStream_t stream) {
    if (int rc = vjp_args(b, params, nmax, coef, dcoef, coords, flags))
        return rc;
    LMZ_REQUIRE(grad_out && tile_work && grad_params,
                "grad_out, tile_work and grad_params are required");
    return launch_vjp(b, params, nmax, coef, dcoef, coords, flags, grad_out,
                      tile_work, grad_params, stream);
}

int lmz_hologram_vjp_static(const lmz_batch *b, const double *params,
                            const int32_t *nmax, const float *coef,
                            const float *dcoef, const float *coords,
                            int32_t flags, const float *grad_out,
                            double *grad_params, hipStream_t stream) {
    if (int rc = vjp_args(b, params, nmax, coef, dcoef, coords, flags))
        return rc;
    LMZ_REQUIRE(grad_out && grad_params, "grad_out and grad_params are required");
    LMZ_REQUIRE(b->ntiles <= LMZ_VJP_STATIC_MAX_TILES,
                "the batch has more tiles than lmz_hologram_vjp_static's "
                "scratch holds (LMZ_VJP_STATIC_MAX_TILES); use "
                "lmz_hologram_vjp with a tile_work buffer");
    double *work = nullptr;
    if (int rc = vjp_scratch(&work)) return rc;
    return launch_vjp(b, params, nmax, coef, dcoef, coords, flags, grad_out,
                      work, grad_params, stream);
}

int lmz_batch_set_schedule(lmz_batch *b, const lmz_schedule *sc) {
    if (int rc = check_batch(b)) return rc;
    LMZ_REQUIRE(sc != nullptr, "schedule is NULL");
    auto ok_nt = [](int nt) { return nt == 64 || nt == 128 || nt == 256; };
    LMZ_REQUIRE(ok_nt(sc->jac_threads) && ok_nt(sc->fwd_threads),
                "threads per tile must be 64, 128 or 256");
    LMZ_REQUIRE(sc->jac_blocks_per_cu >= 0 && sc->jac_blocks_per_cu <= 64 &&
                sc->fwd_blocks_per_cu >= 0 && sc->fwd_blocks_per_cu <= 64,
                "blocks per CU must be in 0..64");
    b->sched = *sc;
    return LMZ_OK;
}

int lmz_batch_get_schedule(const lmz_batch *b, lmz_schedule *sc) {
    if (int rc = check_batch(b)) return rc;
    LMZ_REQUIRE(sc != nullptr, "schedule is NULL");
    *sc = b->sched;
    return LMZ_OK;
}

int lmz_autotune_scratch_bytes(const lmz_batch *b, size_t *bytes) {
    if (int rc = check_batch(b)) return rc;
    LMZ_REQUIRE(bytes != nullptr, "bytes is NULL");
    *bytes = tune_layout(b).total;
    return LMZ_OK;
}

int lmz_autotune(lmz_batch *b, const double *params, const int32_t *nmax,
                 const float *coef, const float *dcoef, const float *coords,
                 int32_t flags, const float *data, void *scratch,
                 int32_t options, lmz_schedule *chosen, hipStream_t stream) {
    if (int rc = forward_args(b, params, nmax, coef, coords, nullptr)) return rc;
    LMZ_REQUIRE(dcoef && data && scratch && chosen,
                "dcoef, data, scratch and chosen are required");
    LMZ_REQUIRE((reinterpret_cast<uintptr_t>(scratch) & 255) == 0,
                "scratch must be 256-byte aligned");
    LMZ_REQUIRE(flags & LMZ_CARTESIAN, "the LM fits the Cartesian hologram");
    int dev = 0, cus = 0;
    LMZ_HIP(hipGetDevice(&dev));
    if (int rc = device_cus(&cus)) return rc;
    // the cache key: device, tiles (log2), mean order count (/8), layout
    std::vector<int32_t> hn((size_t)b->nfeat * b->coef_buffers);
    LMZ_HIP(hipMemcpyAsync(hn.data(), nmax, sizeof(int32_t) * hn.size(),
                           hipMemcpyDeviceToHost, stream));
    LMZ_HIP(hipStreamSynchronize(stream));
    double nsum = 0.0;
    for (int f = 0; f < b->nfeat; ++f) nsum += hn[(size_t)f * b->coef_buffers];
    const TuneKey key{dev, (int)std::lround(std::log2((double)b->ntiles)),
                      (int)std::lround(nsum / b->nfeat / 8.0), b->all_dense,
                      flags};
    if (!(options & LMZ_TUNE_FORCE)) {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            b->sched = it->second;
            *chosen = it->second;
            return LMZ_OK;
        }
    }
    const TuneScratch L = tune_layout(b);
    char *ws = static_cast<char *>(scratch);
    uint8_t *mask = reinterpret_cast<uint8_t *>(ws + L.mask);
    TileList tl{reinterpret_cast<int32_t *>(ws + L.list),
                reinterpret_cast<int32_t *>(ws + L.count), 0,
                reinterpret_cast<int32_t *>(ws + L.start)};
    double *partials = reinterpret_cast<double *>(ws + L.partials);
    double *chi2 = reinterpret_cast<double *>(ws + L.chi2);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    LMZ_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(hipErrorOutOfMemory, "hipEventCreate");
    }
    const int reps = 3;
    const lmz_schedule saved = b->sched;
    lmz_schedule best = saved;
    hipError_t err = hipSuccess;
    for (int kind = 0; kind < 2 && err == hipSuccess; ++kind) {
        float best_ms = 1e30f;
        for (int nt : {64, 128, 256}) {
            const int occ = occupancy(kind == 0, nt);
            int last = -1;
            for (int mult : {1, 2, 4}) {
                const int bpc = std::min(occ * mult, 64);
                if (bpc == last) continue;
                last = bpc;
                tl.grid = cus * bpc;
                lmz_schedule sc = b->sched;
                if (kind == 0) { sc.jac_threads = nt; sc.jac_blocks_per_cu = bpc; }
                else { sc.fwd_threads = nt; sc.fwd_blocks_per_cu = bpc; }
                b->sched = sc;
                // the LM launches over every tile in its first iterations and
                // over a thinning set of features later: time both, the
                // sparse one (every 8th feature) weighted to equal work
                float total = 0.f;
                for (int r = 0; r <= 2 * reps && err == hipSuccess; ++r) {
                    const int stride = (r & 1) ? 8 : 1;
                    hipLaunchKernelGGL(k_mask_every, dim3((b->nfeat + 255) / 256),
                                       dim3(256), 0, stream, mask, b->nfeat, stride);
                    launch_tile_list(b, mask, tl, stream);
                    (void)hipEventRecord(e0, stream);
                    if (kind == 0)
                        launch_jacobian(b, params, nullptr, 0, nmax, coef, dcoef,
                                        coords, flags, data, nullptr, 0LL,
                                        nullptr, nullptr, partials, nullptr,
                                        &tl, stream);
                    else
                        launch_residual(b, params, nullptr, 0, nmax, coef, coords,
                                        flags, data, nullptr, 0LL, chi2, nullptr,
                                        &tl, stream);
                    (void)hipEventRecord(e1, stream);
                    err = hipEventSynchronize(e1);
                    float ms = 0.f;
                    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
                    if (r > 0) total += ms * stride;   // the first launch warms up
                }
                if (err == hipSuccess && total / reps < best_ms) {
                    best_ms = total / reps;
                    if (kind == 0) {
                        best.jac_threads = nt;
                        best.jac_blocks_per_cu = bpc;
                    } else {
                        best.fwd_threads = nt;
                        best.fwd_blocks_per_cu = bpc;
                    }
                }
            }
        }
        b->sched = saved;
        if (kind == 0) { b->sched.jac_threads = best.jac_threads;
                         b->sched.jac_blocks_per_cu = best.jac_blocks_per_cu; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (err != hipSuccess) {
        b->sched = saved;
        return hip_fail(err, "lmz_autotune");
    }
    b->sched = best;
    *chosen = best;
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        g_tune_cache[key] = best;
    }
    return LMZ_OK;
}

int lmz_normal_eq(const lmz_batch *b, const float *jac, const float *resid,
                  double *out, hipStream_t stream) {
    if (int rc = check_batch(b)) return rc;
    LMZ_REQUIRE(jac && resid && out, "jac, resid and out are required");
    // one workgroup per feature, 1024 threads so that the largest crops do
    // not trail the rest
    hipLaunchKernelGGL(k_normal_eq<1024>, dim3(b->nfeat), dim3(1024), 0,
                       stream, b->d_feat, jac, resid, b->npix_total, out);
    LMZ_HIP(hipGetLastError());
    return LMZ_OK;
}

}  // extern "C"
