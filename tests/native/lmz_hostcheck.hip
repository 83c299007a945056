// Host build of the per-pixel and per-feature arithmetic in
// pylorenzmie_amd/csrc/lmz_pixel.h, the LM step logic in
// pylorenzmie_amd/csrc/lmz_lmstep.h, the differential-evolution step in
// pylorenzmie_amd/csrc/lmz_de.h and the frame algorithms in
// pylorenzmie_amd/csrc/lmz_frame.h, for CPU tests of the kernel formulas
// against the FP64 oracle, numpy and scipy.  Test instrument only: the library's product
// path is the gfx950 kernels in liblmz.so.
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <vector>

#include "../../pylorenzmie_amd/csrc/lmz_de.h"
#include "../../pylorenzmie_amd/csrc/lmz_frame.h"
#include "../../pylorenzmie_amd/csrc/lmz_lmstep.h"
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

// Tangent rows (dcoef, LMZ_DCOEF_STRIDE floats per order) of a layered
// sphere's coefficients in the fitting parametrization (outer radius
// scaling every radius, outer index n + i k); returns n_max.
int chk_mie_layered_tangent(const double *radius, const double *nl,
                            const double *kl, int L, double n_m,
                            double wavelength, int cap, float *dcoef) {
    std::vector<double> work(LMZ_WORK_PER_ORDER * (cap + 1));
    std::vector<double> ab(4 * (cap + 1));
    const int N = mie_layered_feature(radius, nl, kl, L, n_m, wavelength, cap,
                                      work.data(), 1, nullptr, ab.data());
    for (int dir = 0; dir < 2; ++dir)
        mie_layered_tangent(radius, nl, kl, L, n_m, wavelength, dir, N,
                            work.data(), 1, dcoef);
    return N;
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
        pixel_geometry<true>(fu, ax, ay, az, bohren, g);
        if (jac) {
            JacPixel jp;
            jacobian_pixel(fu, g, coef, dcoef, nmax, jp);
            for (int c = 0; c < LMZ_NJAC; ++c) jac[c * npix + j] = jp.J[c];
        }
    }
}

}  // extern "C"

// The table rows a group of pixels j0 .. j1 - 1 reads as the persistent
// kernels give them: a table of more than cap rows (cap > 0) is read
// through a window of cap rows (table_rows, table_window) where the group's
// rows fit in one.  Returns the window's first row (the base the
// interpolation takes) and the table to read: a copy holding the window's
// rows and NaN in every other, so a read outside the window shows; win
// counts the groups read through a window.
template <int P, int RS, typename T, typename Par>
const T *chk_window(const FeatUniform &fu, const Par &p, T k, T px, T py,
                    int j0, int j1, int cap, const std::vector<T> &tab,
                    std::vector<T> &buf, int &base, int &win) {
    base = 0;
    if (p.m <= 0 || cap <= 0 || p.m <= cap) return tab.data();
    int lo, hi;
    table_rows<P, T>(fu.width, j0, j1, k, px, py, p.rho0, p.inv_h, p.m, lo, hi);
    if (hi - lo >= cap) return tab.data();
    const int sb = table_window(lo, hi, p.m, cap);
    const int se = std::min(sb + cap, p.m);
    buf.assign(tab.size(), std::numeric_limits<T>::quiet_NaN());
    std::copy(tab.begin() + (size_t)sb * RS, tab.begin() + (size_t)se * RS,
              buf.begin() + (size_t)sb * RS);
    base = sb;
    ++win;
    return buf.data() + (size_t)sb * RS;
}

extern "C" {

// Jacobian of every pixel of a dense crop (feat->width > 0) by the order
// loop (jac_loop) and by the radial table (jac_tab, when the feature is
// tabulated), the table read as chk_window gives it to groups of `group`
// pixels; returns the table's sample count (0: not tabulated), its spacing
// in k rho and the groups read through a window.
int chk_jac_rtab(const lmz_feature *feat, const double *params,
                 const float *coef, const float *dcoef, int nmax, int flags,
                 float *jac_loop, float *jac_tab, float *h_out, int cap,
                 int group, int *nwin) {
    const FeatUniform fu = feature_uniform(*feat, params, nmax, 0);
    const bool bohren = flags & LMZ_BOHREN;
    const RTabPar p = rtab_params(fu);
    std::vector<float> tab((size_t)LMZ_RTAB_STRIDE * (p.m > 0 ? p.m : 1)), buf;
    for (int j = 0; j < p.m; ++j)
        rtab_sample(fu, p, bohren, j, coef, dcoef, &tab[(size_t)j * LMZ_RTAB_STRIDE]);
    *nwin = 0;
    for (int j0 = 0; j0 < fu.npix; j0 += group) {
        const int j1 = std::min(j0 + group, fu.npix);
        int base;
        const float *src = chk_window<LMZ_RTAB_P, LMZ_RTAB_STRIDE, float>(
            fu, p, fu.k, fu.ix + fu.fx, fu.iy + fu.fy, j0, j1, cap, tab, buf,
            base, *nwin);
        for (int j = j0; j < j1; ++j) {
            float ax, ay, az;
            pixel_coords(fu, j, nullptr, fu.npix, ax, ay, az);
            PixGeom g;
            pixel_geometry<true>(fu, ax, ay, az, bohren, g);
            JacPixel a, b;
            jacobian_pixel(fu, g, coef, dcoef, nmax, a);
            if (p.m > 0) jacobian_pixel_rtab(fu, g, p, src, 1.f, b, base);
            for (int c = 0; c < LMZ_NJAC; ++c) {
                jac_loop[c * fu.npix + j] = a.J[c];
                jac_tab[c * fu.npix + j] = p.m > 0 ? b.J[c] : 0.f;
            }
        }
    }
    *h_out = p.h;
    return p.m;
}

// FP64 hologram of every pixel of a dense crop by the order loop (holo_loop)
// and by the FP64 radial table (holo_tab, when tabulated), the table read
// as chk_window gives it to groups of `group` pixels; returns the table's
// sample count (0: not tabulated), its spacing and the groups read through
// a window.
int chk_fwd_ftab(const lmz_feature *feat, const double *params,
                 const float *coef, int nmax, int flags, double *holo_loop,
                 double *holo_tab, double *h_out, int cap, int group,
                 int *nwin) {
    const FeatUniform fu = feature_uniform(*feat, params, nmax, 0);
    const bool bohren = flags & LMZ_BOHREN;
    const FTabPar p = ftab_params(fu);
    std::vector<double> tab((size_t)LMZ_FTAB_STRIDE * (p.m > 0 ? p.m : 1)), buf;
    for (int j = 0; j < p.m; ++j)
        ftab_sample(fu, p, bohren, j, coef, &tab[(size_t)j * LMZ_FTAB_STRIDE]);
    *nwin = 0;
    for (int j0 = 0; j0 < fu.npix; j0 += group) {
        const int j1 = std::min(j0 + group, fu.npix);
        int base;
        const double *src = chk_window<LMZ_FTAB_P, LMZ_FTAB_STRIDE, double>(
            fu, p, fu.kd, (double)fu.ix + fu.fxd, (double)fu.iy + fu.fyd, j0,
            j1, cap, tab, buf, base, *nwin);
        for (int j = j0; j < j1; ++j) {
            float ax, ay, az;
            pixel_coords(fu, j, nullptr, fu.npix, ax, ay, az);
            PixGeomD g;
            pixel_geometry_d(fu, ax, ay, az, bohren, g);
            cd E[3];
            field_from_sums(g, forward_sum(g, coef, nmax), true, E);
            holo_loop[j] = intensity(E, fu.alphad);
            if (p.m > 0) {
                field_from_sums(g, forward_sum_ftab(g, p, src, base), true, E);
                holo_tab[j] = intensity(E, fu.alphad);
            } else {
                holo_tab[j] = 0.0;
            }
        }
    }
    *h_out = p.h;
    return p.m;
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

// The LM's per-feature step logic (lmz_lmstep.h) as k_lm_solve and
// k_lm_accept run it, on an opaque LMState owned by the caller.
int chk_lm_state_bytes(void) { return (int)sizeof(LMState); }
void chk_lm_init(void *st, double lambda0) {
    lm_init(*static_cast<LMState *>(st), lambda0);
}
void chk_lm_set_chi2(void *st, double chi2) {
    static_cast<LMState *>(st)->chi2 = chi2;
}
// A fresh Jacobian's reduction (J^T J upper triangle, J^T r, r^T r) at p;
// returns 1 when the gradient test stops the fit (status GTOL).
int chk_lm_fresh(void *st, const double *red, const lmz_lm_settings *set,
                 const double *p) {
    LMState &S = *static_cast<LMState *>(st);
    lm_take_jacobian(S, red, *set, p);
    if (set->gtol > 0.0 && lm_gtol_met(S, *set)) {
        S.status |= LMZ_LM_GTOL;
        return 1;
    }
    return 0;
}
int chk_lm_solve(void *st, const lmz_lm_settings *set, const double *p,
                 double *pt) {
    return lm_solve(*static_cast<LMState *>(st), *set, p, pt) ? 1 : 0;
}
int chk_lm_accept(void *st, const lmz_lm_settings *set, double chi2t,
                  const double *pt) {
    return lm_accept(*static_cast<LMState *>(st), *set, chi2t, pt) ? 1 : 0;
}
// chi2, lambda (par), delta, status, iter, njev, need_j
void chk_lm_read(const void *st, double *out) {
    const LMState &S = *static_cast<const LMState *>(st);
    out[0] = S.chi2; out[1] = S.lambda; out[2] = S.delta;
    out[3] = S.status; out[4] = S.iter; out[5] = S.njev; out[6] = S.need_j;
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


// ---- differential evolution (lmz_de.h), driven in the order the kernels
// of lmz_de.hip run it: k_de_init, energies, k_de_select(0), then per
// generation k_de_trial for every member, energies, k_de_select.

// Latin-hypercube start of search s: unit[npop][ndim].
void chk_de_lhs(uint64_t seed, int s, int npop, int ndim, double *unit) {
    std::vector<double> keys(npop);
    for (int j = 0; j < ndim; ++j) {
        for (int i = 0; i < npop; ++i) keys[i] = de_lhs_key(seed, (uint32_t)s, i, j);
        for (int i = 0; i < npop; ++i)
            unit[i * ndim + j] = de_lhs(seed, (uint32_t)s, i, j, npop, keys.data());
    }
}

// One best1bin trial of member i; also returns the scale used.
double chk_de_trial(uint64_t seed, int gen, int s, int i, int npop, int ndim,
                    const double *u, int best, double mlo, double mhi,
                    double CR, double *trial) {
    const double F = de_scale(seed, (uint32_t)gen, (uint32_t)s, mlo, mhi);
    de_trial(seed, (uint32_t)gen, (uint32_t)s, i, npop, ndim, u, best, F, CR, trial);
    return F;
}

// sizeof(lmz_de_settings) and the offsets of npop, tol and seed
void chk_de_settings_layout(int *out) {
    out[0] = (int)sizeof(lmz_de_settings);
    out[1] = (int)offsetof(lmz_de_settings, npop);
    out[2] = (int)offsetof(lmz_de_settings, tol);
    out[3] = (int)offsetof(lmz_de_settings, seed);
}

int chk_de_converged(const double *E, int n, double tol, double atol) {
    return de_converged(E, n, tol, atol) ? 1 : 0;
}

// The whole search of nsearch independent problems whose energy is the
// bowl 1 + sum_j ((x_j - target[s][j]) / scale[s][j])^2 in parameter units
// (x_j = lo_j + u_j (hi_j - lo_j), bounds [s][j][2]), or, with
// flat != 0, a constant.  Outputs as lmz_de_run's.
void chk_de_bowl(uint64_t seed, int nsearch, int npop, int ndim, int maxiter,
                 double tol, double atol, double mlo, double mhi, double CR,
                 const double *bounds, const double *target, const double *scale,
                 int flat, double *best_out, double *energy_out, int *gens) {
    auto energy = [&](int s, const double *u) {
        if (flat) return 1.0;
        double e = 1.0;
        for (int j = 0; j < ndim; ++j) {
            const double *b = bounds + 2 * (s * ndim + j);
            const double x = b[0] + u[j] * (b[1] - b[0]);
            const double d = (x - target[s * ndim + j]) / scale[s * ndim + j];
            e += d * d;
        }
        return e;
    };
    for (int s = 0; s < nsearch; ++s) {
        std::vector<double> unit(npop * ndim), trial(npop * ndim), E(npop);
        chk_de_lhs(seed, s, npop, ndim, unit.data());
        for (int i = 0; i < npop; ++i) E[i] = de_energy(energy(s, &unit[i * ndim]));
        int best = de_argmin(E.data(), npop);
        gens[s] = 0;
        for (int gen = 1; gen <= maxiter; ++gen) {
            const double F = de_scale(seed, (uint32_t)gen, (uint32_t)s, mlo, mhi);
            for (int i = 0; i < npop; ++i)
                de_trial(seed, (uint32_t)gen, (uint32_t)s, i, npop, ndim,
                         unit.data(), best, F, CR, &trial[i * ndim]);
            for (int i = 0; i < npop; ++i) {
                const double et = de_energy(energy(s, &trial[i * ndim]));
                if (et < E[i]) {
                    E[i] = et;
                    for (int j = 0; j < ndim; ++j) unit[i * ndim + j] = trial[i * ndim + j];
                }
            }
            best = de_argmin(E.data(), npop);
            gens[s] = gen;
            if (de_converged(E.data(), npop, tol, atol)) break;
        }
        energy_out[s] = E[best];
        for (int j = 0; j < ndim; ++j) {
            const double *b = bounds + 2 * (s * ndim + j);
            best_out[s * ndim + j] = b[0] + unit[best * ndim + j] * (b[1] - b[0]);
        }
    }
}

}  // extern "C"

namespace {

struct HostHist {
    uint32_t w[128];
    void add(uint32_t i, uint32_t v) { w[i] += v; }
    uint32_t word(int i) const { return w[i]; }
    void clear() {
        for (int i = 0; i < 128; ++i) w[i] = 0u;
    }
};

// The radix-select median filter as the kernel runs it: 8 x 8 output tiles,
// selection from the first digit where the tile's keys differ.
template <int DT>
void medfilt_radix_host(const void *image, int ny, int nx, int size, void *bgv) {
    using P = Pix<DT>;
    using K = typename P::K;
    using T = typename P::T;
    const T *img = (const T *)image;
    T *bg = (T *)bgv;
    const int lo = size / 2, span = 7 + size;
    HostHist h;
    auto key_at = [&](int i, int j) {
        return P::key(img[(long long)reflect_index(i, ny) * nx + reflect_index(j, nx)]);
    };
    for (int by = 0; by < ny; by += 8)
        for (int bx = 0; bx < nx; bx += 8) {
            const K kref = key_at(by - lo, bx - lo);
            unsigned long long diff = 0ull;
            for (int i = 0; i < span; ++i)
                for (int j = 0; j < span; ++j)
                    diff |= (unsigned long long)(key_at(by - lo + i, bx - lo + j) ^ kref);
            for (int ty = 0; ty < 8; ++ty)
                for (int tx = 0; tx < 8; ++tx) {
                    const int x = bx + tx, y = by + ty;
                    if (x >= nx || y >= ny) continue;
                    K key = kref;
                    if (diff) {
                        const int hb = 63 - __builtin_clzll(diff);
                        const int top = (hb / 8) * 8;
                        const K prefix = (top + 8 >= (int)(8 * sizeof(K)))
                                             ? (K)0 : (K)(kref & (~(K)0 << (top + 8)));
                        auto fetch = [&](int r, int c) {
                            return key_at(y - lo + r, x - lo + c);
                        };
                        key = window_select<K>(fetch, size, size * size / 2,
                                               prefix, top, h);
                    }
                    bg[(long long)y * nx + x] = P::value(key);
                }
        }
}

}  // namespace

extern "C" {

// scipy.ndimage.median_filter by the device algorithms: path 0 the sliding
// histogram (8-bit frames), path 1 the radix select, path 2 the two-level
// sliding histogram (16-bit frames).
int chk_median_filter(const void *image, int dt, int ny, int nx, int size,
                      void *bg, int path) {
    if (path == 2) {
        if (dt != LMZ_U16) return 1;
        const uint16_t *img = (const uint16_t *)image;
        uint16_t *out = (uint16_t *)bg;
        const int lo = size / 2, rank = size * size / 2;
        HostHist H, F;
        for (int x = 0; x < nx; ++x)
            for (int y0 = 0; y0 < ny; y0 += 32) {
                auto px = [&](int yy, int c) {
                    return (uint32_t)img[(long long)reflect_index(yy, ny) * nx +
                                         reflect_index(x - lo + c, nx)];
                };
                int ycur = y0;
                auto refill = [&](int m) {
                    for (int r = 0; r < size; ++r)
                        for (int c = 0; c < size; ++c) {
                            const uint32_t v = px(ycur - lo + r, c);
                            if ((int)(v >> 8) == m) hist_add(F, v & 0xFFu);
                        }
                };
                H.clear();
                for (int r = 0; r < size; ++r)
                    for (int c = 0; c < size; ++c) hist_add(H, px(y0 - lo + r, c) >> 8);
                Slide16 st;
                st.m = hist_find(H, rank, &st.bc);
                F.clear();
                refill(st.m);
                st.mf = hist_find(F, rank - st.bc, &st.bf);
                out[(long long)y0 * nx + x] = (uint16_t)((st.m << 8) | st.mf);
                for (int y = y0 + 1; y < std::min(y0 + 32, ny); ++y) {
                    for (int c = 0; c < size; ++c) {
                        slide16_remove(H, F, st, px(y - 1 - lo, c));
                        slide16_add(H, F, st, px(y - 1 - lo + size, c));
                    }
                    ycur = y;
                    out[(long long)y * nx + x] =
                        (uint16_t)slide16_settle(H, F, st, rank, refill);
                }
            }
        return 0;
    }
    if (path == 0) {
        if (dt != LMZ_U8) return 1;
        const uint8_t *img = (const uint8_t *)image;
        uint8_t *out = (uint8_t *)bg;
        const int lo = size / 2, rank = size * size / 2;
        HostHist h;
        for (int x = 0; x < nx; ++x)
            for (int y0 = 0; y0 < ny; y0 += 32) {
                auto px = [&](int yy, int c) {
                    return img[(long long)reflect_index(yy, ny) * nx +
                               reflect_index(x - lo + c, nx)];
                };
                h.clear();
                for (int r = 0; r < size; ++r)
                    for (int c = 0; c < size; ++c) hist_add(h, px(y0 - lo + r, c));
                int below;
                int m = hist_find(h, rank, &below);
                out[(long long)y0 * nx + x] = (uint8_t)m;
                for (int y = y0 + 1; y < std::min(y0 + 32, ny); ++y) {
                    for (int c = 0; c < size; ++c) {
                        slide_remove(h, px(y - 1 - lo, c), m, below);
                        slide_add(h, px(y - 1 - lo + size, c), m, below);
                    }
                    slide_settle(h, rank, m, below);
                    out[(long long)y * nx + x] = (uint8_t)m;
                }
            }
        return 0;
    }
    switch (dt) {
    case LMZ_U8: medfilt_radix_host<LMZ_U8>(image, ny, nx, size, bg); break;
    case LMZ_U16: medfilt_radix_host<LMZ_U16>(image, ny, nx, size, bg); break;
    case LMZ_F32: medfilt_radix_host<LMZ_F32>(image, ny, nx, size, bg); break;
    default: medfilt_radix_host<LMZ_F64>(image, ny, nx, size, bg); break;
    }
    return 0;
}

// Order-preserving key of a value and back (lmz_frame.h's Pix<dt>).
unsigned long long chk_key(int dt, double v) {
    switch (dt) {
    case LMZ_U8: return Pix<LMZ_U8>::key((uint8_t)v);
    case LMZ_U16: return Pix<LMZ_U16>::key((uint16_t)v);
    case LMZ_F32: return Pix<LMZ_F32>::key((float)v);
    default: return Pix<LMZ_F64>::key(v);
    }
}

double chk_mean_of_two(int dt, double a, double b) {
    switch (dt) {
    case LMZ_U8: return mean_of_two<LMZ_U8>((uint8_t)a, (uint8_t)b);
    case LMZ_U16: return mean_of_two<LMZ_U16>((uint16_t)a, (uint16_t)b);
    case LMZ_F32: return mean_of_two<LMZ_F32>((float)a, (float)b);
    default: return mean_of_two<LMZ_F64>(a, b);
    }
}

int chk_reflect(int i, int n) { return reflect_index(i, n); }

}  // extern "C"
