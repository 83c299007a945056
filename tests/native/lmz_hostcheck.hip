// Host build of the per-pixel and per-feature arithmetic in
// pylorenzmie_amd/csrc/lmz_pixel.h, for CPU tests of the kernel formulas
// against the FP64 oracle.  Test instrument only: the library's product
// path is the gfx950 kernels in liblmz.so.
#include <cstdint>
#include <vector>

#include "../../pylorenzmie_amd/csrc/lmz_pixel.h"

using namespace lmz;

extern "C" {

// FP64 Mie coefficients with tangents; returns n_max.
int chk_mie(double a_p, double n_p, double k_p, double n_m, double wavelength,
            int cap, float *coef, float *dcoef, double *ab) {
    std::vector<double> work(LMZ_WORK_PER_ORDER * (cap + 1));
    return mie_feature(a_p, n_p, k_p, n_m, wavelength, cap, work.data(), 1,
                       coef, dcoef, ab);
}

// FP64 layered-sphere coefficients; returns n_max.
int chk_mie_layered(const double *radius, const double *nl, const double *kl,
                    int L, double n_m, double wavelength, int cap, double *ab) {
    std::vector<double> work(LMZ_WORK_PER_ORDER * (cap + 1));
    return mie_layered_feature(radius, nl, kl, L, n_m, wavelength, cap,
                               work.data(), 1, nullptr, ab);
}

// Field (flags as in lmz.h), hologram and Jacobian of one feature at
// explicit coordinates (x, y, z arrays relative to feat->origin).
void chk_pixels(const lmz_feature *feat, const double *params,
                const float *coef, const float *dcoef, int nmax,
                const float *coords, int npix, int flags, float *field,
                float *holo, float *jac) {
    const FeatUniform fu = feature_uniform(*feat, params, nmax, 0);
    const bool cart = flags & LMZ_CARTESIAN, bohren = flags & LMZ_BOHREN;
    for (int j = 0; j < npix; ++j) {
        const float ax = coords[j], ay = coords[npix + j], az = coords[2 * npix + j];
        if (field || holo) {
            PixGeomD gd;
            pixel_geometry_d(fu, ax, ay, az, bohren, gd);
            const SumsD S = forward_sum(gd, coef, nmax);
            cd E[3];
            field_from_sums(gd, S, cart, E);
            if (field)
                for (int c = 0; c < 3; ++c) {
                    field[2 * (c * npix + j)] = (float)(fu.alphad * E[c].re);
                    field[2 * (c * npix + j) + 1] = (float)(fu.alphad * E[c].im);
                }
            if (holo) holo[j] = (float)intensity(E, fu.alphad);
        }
        PixGeom g;
        pixel_geometry(fu, ax, ay, az, bohren, g);
        if (jac) {
            JacPixel jp;
            jacobian_pixel(fu, g, coef, dcoef, nmax, jp);
            for (int c = 0; c < LMZ_NJAC; ++c) jac[c * npix + j] = jp.J[c];
        }
    }
}

// FP32 hologram of one feature (the residual kernel's fast mode).
void chk_holo32(const lmz_feature *feat, const double *params,
                const float *coef, int nmax, const float *coords, int npix,
                int flags, float *holo) {
    const FeatUniform fu = feature_uniform(*feat, params, nmax, 0);
    const bool bohren = flags & LMZ_BOHREN;
    for (int j = 0; j < npix; ++j) {
        PixGeom g;
        pixel_geometry(fu, coords[j], coords[npix + j], coords[2 * npix + j],
                       bohren, g);
        holo[j] = intensity_f(g, forward_sum_f(g, coef, nmax), fu.alpha);
    }
}

// Dense-grid pixel coordinates as the kernels derive them.
void chk_dense_coords(int width, int npix, float *ax, float *ay) {
    lmz_feature F{};
    F.width = width;
    F.npix = npix;
    F.noise = 1.0; F.wavelength = 1.0; F.n_m = 1.0; F.magnification = 1.0;
    double p[8] = {0};
    const FeatUniform fu = feature_uniform(F, p, 1, 0);
    for (int j = 0; j < npix; ++j) {
        float z;
        pixel_coords(fu, j, nullptr, 0, ax[j], ay[j], z);
    }
}

}  // extern "C"
