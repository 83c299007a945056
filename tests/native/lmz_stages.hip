// Precision stages of the forward hologram.  The per-pixel computation of
// pylorenzmie_amd/csrc/lmz_pixel.h (forward_sum and its epilogue) is split
// into six stages, each of which runs in FP64 (as the library does) or in
// FP32; bit V of the variant selects FP32 for
//   1 geometry (pixel_geometry in place of pixel_geometry_d),
//   2 angular recurrence (Angular32), 4 xi recurrence,
//   8 coefficient products, 16 accumulation of the sums, 32 epilogue.
// The host build gives each variant's error against the all-FP64 variant
// (tests/test_precision_stages.py); the gfx950 build its time per launch
// (bench.py --extra stages).  Test instrument only.
#include <hip/hip_runtime.h>

#include "../../pylorenzmie_amd/csrc/lmz_pixel.h"

using namespace lmz;

namespace {

template <int V>
LMZ_HD double stage_pixel(const FeatUniform &fu, float ax, float ay, float az,
                          const float *coef, int nmax) {
    constexpr bool G = V & 1, A = V & 2, X = V & 4, P = V & 8, S = V & 16,
                   E = V & 32;
    PixGeomD g;
    if constexpr (G) {
        PixGeom f;
        pixel_geometry(fu, ax, ay, az, true, f);
        g.R = f.R;
        g.invR = f.invR;
        g.rho = f.rho;
        g.c = f.c;
        g.s = f.s;
        g.u = f.u;
        g.v = f.v;
        g.w = f.w;
        g.invRl = f.invRl;
        g.mir = f.mir;
        g.xi_m1 = cd{f.xi_m1.re, f.xi_m1.im};
        g.xi_0 = cd{f.xi_0.re, f.xi_0.im};
    } else {
        pixel_geometry_d(fu, ax, ay, az, true, g);
    }
    Angular ang;
    ang.init();
    Angular32 ang32;
    ang32.init();
    cd x2 = g.xi_m1, x1 = g.xi_0;
    cf fx2{(float)x2.re, (float)x2.im}, fx1{(float)x1.re, (float)x1.im};
    const double invR = g.invRl, w = g.w;
    const float finvR = (float)invR, fw = (float)w;
    cd Sr{0, 0}, St{0, 0}, Sp{0, 0};
    cf fSr{0, 0}, fSt{0, 0}, fSp{0, 0};
    for (int n = 1; n <= nmax; ++n) {
        const float *cp = coef + LMZ_COEF_STRIDE * (n - 1);
        double pn, tau;
        if constexpr (A) {
            float a, b;
            ang32.step(fw, cp, a, b);
            pn = a;
            tau = b;
        } else {
            ang.step_d(w, cp, pn, tau);
        }
        cd xi, D;
        if constexpr (X) {
            const float ninv = cp[6] * finvR, cx = cp[31] * finvR;
            const cf fxi{fmaf(cx, fx1.re, -fx2.re), fmaf(cx, fx1.im, -fx2.im)};
            const cf fD{fmaf(ninv, fxi.re, -fx1.re), fmaf(ninv, fxi.im, -fx1.im)};
            fx2 = fx1;
            fx1 = fxi;
            xi = cd{fxi.re, fxi.im};
            D = cd{fD.re, fD.im};
        } else {
            const double ninv = row_d(cp, 14) * invR;
            const double cx = fma(2.0, ninv, -invR);
            xi = cd{fma(cx, x1.re, -x2.re), fma(cx, x1.im, -x2.im)};
            D = cd{fma(ninv, xi.re, -x1.re), fma(ninv, xi.im, -x1.im)};
            x2 = x1;
            x1 = xi;
        }
        cd PD, QX, PRX;
        if constexpr (P) {   // FP32 table entries, as forward_sum_f
            const cf fxi{(float)xi.re, (float)xi.im}, fD{(float)D.re, (float)D.im};
            const cf pd = cmul3(cp[2], cp[3], cp[0], fD);
            const cf px = cmul3(cp[2], cp[3], cp[0], fxi);
            const cf mqx = cmul3(cp[4], cp[5], cp[1], fxi);
            PD = cd{pd.re, pd.im};
            QX = cd{-mqx.re, -mqx.im};
            PRX = cd{(double)(cp[7] * px.re), (double)(cp[7] * px.im)};
        } else {
            const cd PRc{row_d(cp, 16), row_d(cp, 18)};
            const cd Pc{row_d(cp, 20), row_d(cp, 22)};
            const cd Qc{row_d(cp, 24), row_d(cp, 26)};
            PD = Pc * D;
            QX = Qc * xi;
            PRX = PRc * xi;
        }
        if constexpr (S) {
            const float fp = (float)pn, ft = (float)tau;
            const cf pd{(float)PD.re, (float)PD.im}, qx{(float)QX.re, (float)QX.im};
            const cf prx{(float)PRX.re, (float)PRX.im};
            cmac(fSr, fp, prx);
            cmac(fSt, ft, pd);
            cmac(fSt, -fp, qx);
            cmac(fSp, fp, pd);
            cmac(fSp, -ft, qx);
        } else {
            cdmac(Sr, pn, PRX);
            cdmac(St, tau, PD);
            cdmac(St, -pn, QX);
            cdmac(Sp, pn, PD);
            cdmac(Sp, -tau, QX);
        }
    }
    if constexpr (S) {
        Sr = cd{fSr.re, fSr.im};
        St = cd{fSt.re, fSt.im};
        Sp = cd{fSp.re, fSp.im};
    }
    Sp = g.mir * Sp;
    if constexpr (E) {
        PixGeom f;
        f.invR = (float)g.invR;
        f.s = (float)g.s;
        f.c = (float)g.c;
        f.u = (float)g.u;
        f.v = (float)g.v;
        const SumsF sf{cf{(float)Sr.re, (float)Sr.im}, cf{(float)St.re, (float)St.im},
                       cf{(float)Sp.re, (float)Sp.im}};
        return intensity_f(f, sf, fu.alpha);
    } else {
        cd Ef[3];
        field_from_sums(g, SumsD{Sr, St, Sp}, true, Ef);
        return intensity(Ef, fu.alphad);
    }
}

template <int V>
__global__ __launch_bounds__(256) void k_stage(FeatUniform fu, const float *coef,
                                               int nmax, const float *coords,
                                               int npix, double *out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npix) return;
    out[j] = stage_pixel<V>(fu, coords[j], coords[npix + j], coords[2 * npix + j],
                            coef, nmax);
}

}  // namespace

#define LMZ_STAGE_VARIANTS(X) \
    X(0) X(1) X(2) X(4) X(8) X(16) X(32) X(24) X(56) X(60) X(62) X(63)

extern "C" {

// Intensity of every pixel (explicit coordinates relative to feat->origin)
// under variant V on the host; 0, or -1 for an unknown variant.
int stage_host(int V, const lmz_feature *feat, const double *params,
               const float *coef, int nmax, const float *coords, int npix,
               double *out) {
    const FeatUniform fu = feature_uniform(*feat, params, nmax, 0);
#define LMZ_CASE(v)                                                        \
    case v:                                                                \
        for (int j = 0; j < npix; ++j)                                     \
            out[j] = stage_pixel<v>(fu, coords[j], coords[npix + j],       \
                                    coords[2 * npix + j], coef, nmax);     \
        return 0;
    switch (V) {
        LMZ_STAGE_VARIANTS(LMZ_CASE)
        default: return -1;
    }
#undef LMZ_CASE
}

// The same on the GPU, one thread per pixel, grid ceil(npix / 256): one
// warm launch, then the mean time of reps launches into *ms; out gets the
// values.  0, or a negative code (unknown variant, HIP error).
int stage_gpu(int V, const lmz_feature *feat, const double *params,
              const float *coef, int nmax, const float *coords, int npix,
              int reps, double *out, float *ms) {
    if (npix < 1 || nmax < 1 || reps < 1) return -1;
    const FeatUniform fu = feature_uniform(*feat, params, nmax, 0);
    float *dcoef = nullptr, *dcoords = nullptr;
    double *dout = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    int rc = 0;
    const size_t ncoef = (size_t)LMZ_COEF_STRIDE * nmax;
    if (hipMalloc(&dcoef, sizeof(float) * ncoef) != hipSuccess ||
        hipMalloc(&dcoords, sizeof(float) * 3 * (size_t)npix) != hipSuccess ||
        hipMalloc(&dout, sizeof(double) * (size_t)npix) != hipSuccess ||
        hipMemcpy(dcoef, coef, sizeof(float) * ncoef, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dcoords, coords, sizeof(float) * 3 * (size_t)npix,
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
        rc = -2;
    const dim3 grid((npix + 255) / 256), block(256);
#define LMZ_CASE(v)                                                          \
    case v:                                                                  \
        hipLaunchKernelGGL(k_stage<v>, grid, block, 0, 0, fu, dcoef, nmax,   \
                           dcoords, npix, dout);                             \
        (void)hipEventRecord(a, 0);                                          \
        for (int r = 0; r < reps; ++r)                                       \
            hipLaunchKernelGGL(k_stage<v>, grid, block, 0, 0, fu, dcoef,     \
                               nmax, dcoords, npix, dout);                   \
        (void)hipEventRecord(b, 0);                                          \
        break;
    if (rc == 0) {
        switch (V) {
            LMZ_STAGE_VARIANTS(LMZ_CASE)
            default: rc = -1;
        }
    }
#undef LMZ_CASE
    if (rc == 0) {
        float t = 0.f;
        if (hipEventSynchronize(b) != hipSuccess ||
            hipEventElapsedTime(&t, a, b) != hipSuccess ||
            hipMemcpy(out, dout, sizeof(double) * (size_t)npix,
                      hipMemcpyDeviceToHost) != hipSuccess)
            rc = -3;
        *ms = t / reps;
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (dcoef) (void)hipFree(dcoef);
    if (dcoords) (void)hipFree(dcoords);
    if (dout) (void)hipFree(dout);
    return rc;
}

}  // extern "C"
