// Per-pixel (FP32) and per-feature (FP64) arithmetic of the Lorenz-Mie
// kernels.  Every function here is __host__ __device__ so that the same
// source runs in the gfx950 kernels and in the host-side checker library
// the tests build from it.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

#include "../../include/lmz.h"

#define LMZ_HD __host__ __device__ __forceinline__

namespace lmz {

// ====================================================================
// complex helpers
// ====================================================================
struct cf { float re, im; };
struct cd { double re, im; };

LMZ_HD cf operator+(cf a, cf b) { return {a.re + b.re, a.im + b.im}; }
LMZ_HD cf operator-(cf a, cf b) { return {a.re - b.re, a.im - b.im}; }
LMZ_HD cf operator*(float s, cf a) { return {s * a.re, s * a.im}; }
LMZ_HD cf cmul(cf a, cf b) {
    return {fmaf(a.re, b.re, -a.im * b.im), fmaf(a.re, b.im, a.im * b.re)};
}
// acc += s * a
LMZ_HD void cmac(cf &acc, float s, cf a) {
    acc.re = fmaf(s, a.re, acc.re);
    acc.im = fmaf(s, a.im, acc.im);
}
// Re(conj(a) b)
LMZ_HD float credot(cf a, cf b) { return fmaf(a.re, b.re, a.im * b.im); }
// c z for a table coefficient given as (Re c, Im c, -Im c): each part is a
// product and a multiply-add into it
LMZ_HD cf cmul3(float cre, float cim, float cnim, cf z) {
    return {fmaf(cre, z.re, cnim * z.im), fmaf(cre, z.im, cim * z.re)};
}

LMZ_HD cd operator+(cd a, cd b) { return {a.re + b.re, a.im + b.im}; }
LMZ_HD cd operator-(cd a, cd b) { return {a.re - b.re, a.im - b.im}; }
LMZ_HD cd operator-(cd a) { return {-a.re, -a.im}; }
LMZ_HD cd operator*(cd a, cd b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
LMZ_HD cd operator*(double s, cd a) { return {s * a.re, s * a.im}; }
LMZ_HD cd operator/(cd a, cd b) {  // Smith's algorithm
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, d = b.re + b.im * r;
        return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    const double r = b.re / b.im, d = b.re * r + b.im;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
LMZ_HD cd cinv(cd b) { return cd{1.0, 0.0} / b; }
LMZ_HD cd mul_i(cd a) { return {-a.im, a.re}; }        // i a
// i^k a
LMZ_HD cd mul_ipow(int k, cd a) {
    switch (k & 3) {
        case 0: return a;
        case 1: return {-a.im, a.re};
        case 2: return {-a.re, -a.im};
        default: return {a.im, -a.re};
    }
}

// ====================================================================
// per-feature uniforms (FP64 in, FP32 out)
// ====================================================================
// x_p - origin = ix + fx with ix integral, so the per-pixel displacement
// (x_rel - ix) - fx loses no digits to the particle's absolute position.
struct FeatUniform {
    float k;           // wavenumber [rad / px]
    float ix, iy, iz;  // integral parts of the particle offset from origin
    float fx, fy, fz;  // fractional parts
    float kz0;         // k * origin z
    double kd, fxd, fyd, fzd, kz0d;   // the same in FP64
    float alpha;
    double alphad;
    float inv_noise;
    int npix;
    int width;
    float inv_width;
    long long pix_offset;
    int tile0;
    int nmax;
    int slot0;
};

// Instrument.wavenumber, theory/Instrument.py:46-70 (in medium, scaled)
LMZ_HD double wavenumber_px(const lmz_feature &F) {
    return 2.0 * M_PI / F.wavelength * F.n_m * F.magnification;
}

LMZ_HD FeatUniform feature_uniform(const lmz_feature &F, const double *p,
                                   int nmax, int slot0) {
    FeatUniform u;
    const double kk = wavenumber_px(F);
    u.k = (float)kk;
    const double dx = p[0] - F.origin[0];
    const double dy = p[1] - F.origin[1];
    const double dz = p[2] - F.origin[2];
    const double rx = rint(dx), ry = rint(dy), rz = rint(dz);
    u.ix = (float)rx;
    u.iy = (float)ry;
    u.iz = (float)rz;
    u.fx = (float)(dx - rx);
    u.fy = (float)(dy - ry);
    u.fz = (float)(dz - rz);
    u.kz0 = (float)(kk * F.origin[2]);
    u.kd = kk;
    u.fxd = dx - rx;
    u.fyd = dy - ry;
    u.fzd = dz - rz;
    u.kz0d = kk * F.origin[2];
    u.alpha = (float)p[6];
    u.alphad = p[6];
    u.inv_noise = (float)(1.0 / F.noise);
    u.npix = F.npix;
    u.width = F.width;
    u.inv_width = F.width > 0 ? 1.0f / (float)F.width : 0.f;
    u.pix_offset = F.pix_offset;
    u.tile0 = F.tile_offset;
    u.nmax = nmax;
    u.slot0 = slot0;
    return u;
}

// Pixel j of a feature: coordinates relative to the feature origin, from
// the implicit dense grid (x fastest, lib/meshgrid.py:34-40) or the
// explicit arrays.
LMZ_HD void pixel_coords(const FeatUniform &fu, int j, const float *coords,
                         long long npix_total, float &ax, float &ay,
                         float &az) {
    if (fu.width > 0) {
        int row = (int)(((float)j + 0.5f) * fu.inv_width);
        int col = j - row * fu.width;
        if (col < 0) { row -= 1; col += fu.width; }
        else if (col >= fu.width) { row += 1; col -= fu.width; }
        ax = (float)col;
        ay = (float)row;
        az = 0.f;
    } else {
        const long long gi = fu.pix_offset + j;
        ax = coords[gi];
        ay = coords[npix_total + gi];
        az = coords[2 * npix_total + gi];
    }
}

// ====================================================================
// pixel geometry, theory/LorenzMie.py:221-257
// ====================================================================
// The particle phase exp(-i k z_p) of theory/LorenzMie.py:157-158 is
// folded into the starting values of the Riccati-Bessel recurrence (it is
// linear with real coefficients, so every xi_n carries the same factor).
// For the outgoing wave the phase k (r - z_p) is formed as
// (rho^2 / (r + dz) - z) k, free of cancellation near the axis.
// 1 / sqrt(x) and 1 / x by the hardware approximations (within 1 ulp) in
// the kernels; the host build of the checker divides.  For the Jacobian's
// FP32 geometry, where a correctly rounded sqrtf or division costs ten or
// more instructions per pixel.
LMZ_HD float rsqrt_hw(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / std::sqrt(x);
#endif
}

LMZ_HD float rcp_hw(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

struct PixGeom {
    float X, Y, Z;      // k (x - x_p), k (y - y_p), k (z_p - z)
    float R, invR;      // kr, 1 / kr
    float rho, irho;    // k rho, 1 / (k rho) (0 on the axis)
    float c, s;         // cos theta, sin theta
    float u, v;         // cos phi, sin phi
    // Recurrence inputs.  For cos theta < 0 the recurrences run on the mirror
    // image: pi_n(-c) = (-1)^(n+1) pi_n(c), tau_n(-c) = (-1)^n tau_n(c), and
    // (-1)^(n+1) xi_n obeys the xi recurrence with -1/kr from (xi_{-1}, -xi_0);
    // the sums are then equal except for the sign of S_phi (mir = -1).
    float w;            // 1 - |cos theta|, without cancellation
    float cl;           // |cos theta|
    float invRl;        // +-1 / kr
    float mir;          // +1, or -1 for cos theta < 0
    cf xi_m1, xi_0;     // recurrence start, times exp(-i k z_p)
};

// HW: kr, k rho, their inverses and R - |Z| from the hardware reciprocal
// square root and reciprocal (the Jacobian, whose FP32 arithmetic is
// accurate to ~1e-5 anyway); otherwise from correctly rounded sqrtf and
// division (the FP32 hologram, whose phase k r carries the error of kr).
template <bool HW = false>
LMZ_HD void pixel_geometry(const FeatUniform &fu, float ax, float ay, float az,
                           bool bohren, PixGeom &g) {
    const float dx = (ax - fu.ix) - fu.fx;
    const float dy = (ay - fu.iy) - fu.fy;
    const float dz = (fu.iz - az) + fu.fz;
    g.X = fu.k * dx;
    g.Y = fu.k * dy;
    g.Z = fu.k * dz;
    const float rho2 = fmaf(g.X, g.X, g.Y * g.Y);
    const float R2 = fmaf(g.Z, g.Z, rho2);
    if constexpr (HW) {
        g.invR = rsqrt_hw(R2);
        g.R = R2 * g.invR;
        g.irho = rho2 > 0.f ? rsqrt_hw(rho2) : 0.f;
        g.rho = rho2 * g.irho;
    } else {
        g.R = sqrtf(R2);
        g.invR = 1.0f / g.R;
        g.rho = sqrtf(rho2);
        g.irho = rho2 > 0.f ? 1.0f / g.rho : 0.f;
    }
    g.c = g.Z * g.invR;
    g.s = g.rho * g.invR;
    if (rho2 > 0.f) {
        g.u = g.X * g.irho;
        g.v = g.Y * g.irho;
    } else {  // arctan2(0, 0) = 0
        g.u = 1.f;
        g.v = 0.f;
    }
    const float sz = (g.Z > 0.f) ? 1.f : ((g.Z < 0.f) ? -1.f : 0.f);
    const float sig = bohren ? sz : -sz;
    const float kzpix = fmaf(fu.k, az, fu.kz0);  // k * z of the pixel
    const float az_ = fabsf(g.Z);
    const float t = HW ? rho2 * rcp_hw(g.R + az_)      // R - |Z|
                       : rho2 / (g.R + az_);
    g.w = t * g.invR;
    g.cl = az_ * g.invR;
    g.mir = (g.Z < 0.f) ? -1.f : 1.f;
    g.invRl = g.mir * g.invR;
    if (sig != 0.f) {
        float psi;
        if (sig > 0.f) {  // k (r - z_p) = (R - Z) - k z_pix
            psi = ((g.Z > 0.f) ? t : g.R - g.Z) - kzpix;
        } else {          // -k (r + z_p) = -((R + Z) + k z_pix)
            psi = -(((g.Z < 0.f) ? t : g.R + g.Z) + kzpix);
        }
        float sn, cs;
        sincosf(psi, &sn, &cs);
        g.xi_m1 = {cs, sn};                 // exp(i psi)
        g.xi_0 = {sig * sn, -sig * cs};     // -i sig exp(i psi)
    } else {  // kz == 0: cos kr, sin kr, times exp(-i k z_pix)
        float sr, cr, sp, cp;
        sincosf(g.R, &sr, &cr);
        sincosf(kzpix, &sp, &cp);
        g.xi_m1 = {cr * cp, -cr * sp};
        g.xi_0 = {sr * cp, -sr * sp};
    }
    g.xi_0 = g.mir * g.xi_0;
}

// ====================================================================
// per-order table row (LMZ_COEF_STRIDE floats, scalar loads per order)
// ====================================================================
//   [0] -Im P   [1] Im Q    (P = i E_n a_n, Q = E_n b_n)
//   [2..3]   P              [4..5]  -Q                           (FP32)
//   [6]      n          [7]  n(n+1)
//   [8..9]   1 + 1/n    [10..11]  n(n+1)/2    [12..13]  (n+1)(n+2)/2
//   [14..15] n                                                   (FP64)
//   [16..27] PR = i^(n+1)(2n+1) a_n, P, Q                        (FP64)
//   [28] 1 + 1/n   [29] -n(n+1)/2   [30] (n+1)(n+2)/2   [31] 2n - 1
//   [32] -n                                                       (FP32)
//   [33..35] sigma_n P (re, im, -im)   [36..38] -sigma_n Q (re, im, -im)
//   [39] 1 - 2n                  (FP32, the Jacobian's signed loop, jac_order)
// E_n = i^n (2n+1)/(n(n+1)) as theory/LorenzMie.py:297-303.  The FP32
// entries carry the signs the per-pixel loops need, so that every complex
// product is a multiply and a multiply-add into its own result and every
// sum accumulates a product of positive factors: on gfx950 those are the
// two-operand (VOP2) instructions that issue two at a time, where forms
// with a negated operand (VOP3) issue one at a time.
#define LMZ_COEF_STRIDE 40
// dcoef row: for each of a_p and n_p (9 floats each, at 0 and 9): the
// tangents of PR (re, im, -im), of P (re, im, -im) and of -Q (re, im, -im),
// times sigma_n (order_sign);
// then the k_p tangents of PR, P, Q (re, im) at 18..23.
#define LMZ_DCOEF_STRIDE 24

LMZ_HD double row_d(const float *row, int i) {
    double v;
    __builtin_memcpy(&v, row + i, sizeof v);
    return v;
}

// ====================================================================
// Wiscombe's angular recurrence in deviation form (FP64)
// ====================================================================
// theory/LorenzMie.py:279-281, 311-312 compute pi_n(c) by
//   S = pi_n c,  T = S - pi_{n-1},  tau_n = pi_{n-1} - n T,
//   pi_{n+1} = S + (1 + 1/n) T.
// Near the axis pi_n is ~n^4/8 times as sensitive to c as it is large, and an
// FP32 cos(theta) has already lost the digits of 1 - c.  With a_n = n(n+1)/2
// (= pi_n(1)), w = 1 - c (from rho^2 / (R + Z)) and d_n = a_n - pi_n, the same
// recurrence becomes, exactly,
//   Td = (d_n - d_{n-1}) + w pi_n,   tau_n = n Td - (a_n + d_{n-1}),
//   d_{n+1} = d_n + w pi_n + (1 + 1/n) Td,   pi_{n+1} = a_{n+1} - d_{n+1},
// whose rounding errors scale with d_n, not a_n.  The forward model carries
// it in FP64 (Angular); the Jacobian in FP32 (Angular32).
struct Angular {
    double dm, dn, pn;
    LMZ_HD void init() { dm = 0.0; dn = 0.0; pn = 1.0; }
    // returns pi_n, tau_n of the current order and advances to n + 1
    LMZ_HD void step_d(double w, const float *row, double &pi, double &tau) {
        const double fn = row_d(row, 14);
        const double c1 = row_d(row, 8), an = row_d(row, 10), an1 = row_d(row, 12);
        const double wp = w * pn;
        const double Td = (dn - dm) + wp;
        pi = pn;
        tau = fma(fn, Td, -(an + dm));
        const double dnext = fma(c1, Td, dn + wp);
        dm = dn;
        dn = dnext;
        pn = an1 - dnext;
    }
};

// The same recurrence in FP32 for the Jacobian: its errors scale with d_n
// near the axis, and far from it (d_n ~ a_n) with a_n / pi_n ~ n, which
// keeps the Jacobian's columns within 2e-5 of their scale to n ~ 280.
// a_n = n(n+1)/2 and a_{n+1} = a_n + n + 1 are exact in FP32 for n < 4096.
struct Angular32 {
    float dm, dn, pn;
    LMZ_HD void init() { dm = 0.f; dn = 0.f; pn = 1.f; }
    LMZ_HD void step(float w, const float *row, float &pi, float &tau) {
        const float fn = row[6], c1 = row[28];
        const float man = row[29], an1 = row[30];    // -a_n, a_{n+1}
        const float wp = w * pn;
        const float Td = (dn - dm) + wp;
        pi = pn;
        tau = fmaf(fn, Td, man - dm);                // n Td - (a_n + d_{n-1})
        const float dnext = fmaf(c1, Td, dn + wp);
        dm = dn;
        dn = dnext;
        pn = an1 - dnext;
    }
};

// ====================================================================
// pixel geometry in FP64 for the forward model
// ====================================================================
// The hologram is formed in FP64 end to end (geometry, both recurrences,
// products, sums, projection): gfx950 issues FP64 FMA at the rate of
// unpacked FP32 FMA, and FP32 geometry or recurrences leave errors above
// 1e-5 on bright forward-scattering lobes (I ~ 40).  Same conventions as
// pixel_geometry (phase folded into the start, mirror for cos theta < 0).
struct PixGeomD {
    double R, invR, rho, c, s, u, v;
    double w, invRl, mir;
    cd xi_m1, xi_0;
};

LMZ_HD void pixel_geometry_d(const FeatUniform &fu, float ax, float ay,
                             float az, bool bohren, PixGeomD &g) {
    const double dx = ((double)ax - (double)fu.ix) - fu.fxd;
    const double dy = ((double)ay - (double)fu.iy) - fu.fyd;
    const double dz = ((double)fu.iz - (double)az) + fu.fzd;
    const double X = fu.kd * dx, Y = fu.kd * dy, Z = fu.kd * dz;
    const double rho2 = fma(X, X, Y * Y);
    g.rho = sqrt(rho2);
    g.R = sqrt(fma(Z, Z, rho2));
    g.invR = 1.0 / g.R;
    g.c = Z * g.invR;
    g.s = g.rho * g.invR;
    if (g.rho > 0.0) {
        const double ir = 1.0 / g.rho;
        g.u = X * ir;
        g.v = Y * ir;
    } else {  // arctan2(0, 0) = 0
        g.u = 1.0;
        g.v = 0.0;
    }
    const double sz = (Z > 0.0) ? 1.0 : ((Z < 0.0) ? -1.0 : 0.0);
    const double sig = bohren ? sz : -sz;
    const double kzpix = fma(fu.kd, (double)az, fu.kz0d);
    const double az_ = fabs(Z);
    const double t = rho2 / (g.R + az_);               // R - |Z|
    g.w = t * g.invR;
    g.mir = (Z < 0.0) ? -1.0 : 1.0;
    g.invRl = g.mir * g.invR;
    if (sig != 0.0) {
        double psi;
        if (sig > 0.0) psi = ((Z > 0.0) ? t : g.R - Z) - kzpix;
        else psi = -(((Z < 0.0) ? t : g.R + Z) + kzpix);
        double sn, cs;
        sincos(psi, &sn, &cs);
        g.xi_m1 = {cs, sn};
        g.xi_0 = {sig * sn, -sig * cs};
    } else {
        double sr, cr, sp, cp;
        sincos(g.R, &sr, &cr);
        sincos(kzpix, &sp, &cp);
        g.xi_m1 = {cr * cp, -cr * sp};
        g.xi_0 = {sr * cp, -sr * sp};
    }
    g.xi_0 = g.mir * g.xi_0;
}

// ====================================================================
// forward multipole sum, theory/LorenzMie.py:272-316 (FP64)
// ====================================================================
// With tau = -tau_n (the reference's sign),
//   S_r += PR pi xi,  S_t += P tau D - Q pi xi,  S_p += P pi D - Q tau xi
// and the geometric factors of :322-325 applied in field_from_sums.
struct SumsD { cd r, t, p; };

LMZ_HD void cdmac(cd &acc, double s, cd a) {
    acc.re = fma(s, a.re, acc.re);
    acc.im = fma(s, a.im, acc.im);
}

LMZ_HD SumsD forward_sum(const PixGeomD &g, const float *coef, int nmax) {
    cd Sr{0.0, 0.0}, St{0.0, 0.0}, Sp{0.0, 0.0};
    Angular ang;
    ang.init();
    cd x2 = g.xi_m1, x1 = g.xi_0;
    const double invR = g.invRl, w = g.w;
#pragma unroll 2
    for (int n = 1; n <= nmax; ++n) {
        const float *cp = coef + LMZ_COEF_STRIDE * (n - 1);
        const cd PR{row_d(cp, 16), row_d(cp, 18)};
        const cd P{row_d(cp, 20), row_d(cp, 22)};
        const cd Q{row_d(cp, 24), row_d(cp, 26)};
        const double fn = row_d(cp, 14);
        double pn, tau;
        ang.step_d(w, cp, pn, tau);
        const double ninv = fn * invR;
        const double cx = fma(2.0, ninv, -invR);          // (2n - 1) / kr
        const cd xi{fma(cx, x1.re, -x2.re), fma(cx, x1.im, -x2.im)};
        const cd D{fma(ninv, xi.re, -x1.re), fma(ninv, xi.im, -x1.im)};
        const cd PD = P * D, QX = Q * xi, PRX = PR * xi;
        cdmac(Sr, pn, PRX);
        cdmac(St, tau, PD);
        cdmac(St, -pn, QX);
        cdmac(Sp, pn, PD);
        cdmac(Sp, -tau, QX);
        x2 = x1;
        x1 = xi;
    }
    return SumsD{Sr, St, g.mir * Sp};
}

// The same sum in FP32 with the FP32 angular recurrence (the LM's early
// iterations; errors up to ~2e-5 in I on bright forward lobes).
struct SumsF { cf r, t, p; };

LMZ_HD SumsF forward_sum_f(const PixGeom &g, const float *coef, int nmax) {
    cf Sr{0.f, 0.f}, St{0.f, 0.f}, Sp{0.f, 0.f};
    Angular32 ang;
    ang.init();
    cf x2 = g.xi_m1, x1 = g.xi_0;
    const float invR = g.invRl, w = g.w;
#pragma unroll 2
    for (int n = 1; n <= nmax; ++n) {
        const float *cp = coef + LMZ_COEF_STRIDE * (n - 1);
        const float nn1 = cp[7];
        float pn, tau;
        ang.step(w, cp, pn, tau);
        const float ninv = cp[6] * invR;
        const float cx = cp[31] * invR;                  // (2n - 1) / kr
        const cf xi{fmaf(cx, x1.re, -x2.re), fmaf(cx, x1.im, -x2.im)};
        const cf D{fmaf(ninv, xi.re, -x1.re), fmaf(ninv, xi.im, -x1.im)};
        const cf PD = cmul3(cp[2], cp[3], cp[0], D);
        const cf PX = cmul3(cp[2], cp[3], cp[0], xi);
        const cf mQX = cmul3(cp[4], cp[5], cp[1], xi);   // -Q xi
        cmac(Sr, pn * nn1, PX);          // PR = n(n+1) P
        cmac(St, tau, PD);
        cmac(St, pn, mQX);
        cmac(Sp, pn, PD);
        cmac(Sp, tau, mQX);
        x2 = x1;
        x1 = xi;
    }
    return SumsF{Sr, St, g.mir * Sp};
}

// |x_hat + alpha E|^2 from the FP32 sums (Cartesian projection).
LMZ_HD float intensity_f(const PixGeom &g, const SumsF &S, float alpha) {
    const float invR2 = g.invR * g.invR;
    const cf A = (g.s * g.s * invR2) * S.r + (g.c * g.invR) * S.t;
    const cf B = g.invR * S.p;
    const cf C = (g.c * invR2) * S.r - g.invR * S.t;
    const cf E0 = (g.u * g.u) * A - (g.v * g.v) * B;
    const cf E1 = (g.u * g.v) * (A + B);
    const cf E2 = (g.u * g.s) * C;
    const cf F0{fmaf(alpha, E0.re, 1.f), alpha * E0.im};
    const float a2 = alpha * alpha;
    return credot(F0, F0) + a2 * (credot(E1, E1) + credot(E2, E2));
}

// Spherical -> field components.  Cartesian (theory/LorenzMie.py:331-340):
//   A = S_r s^2/R^2 + S_t c/R,  B = S_p / R,  C = S_r c/R^2 - S_t / R
//   E_x = u^2 A - v^2 B,  E_y = u v (A + B),  E_z = u s C
// Spherical (:322-325): (S_r u s / R^2, S_t u / R, S_p v / R).
LMZ_HD void field_from_sums(const PixGeomD &g, const SumsD &S, bool cart,
                            cd (&E)[3]) {
    const double invR2 = g.invR * g.invR;
    if (cart) {
        const double s2 = g.s * g.s;
        const cd A = (s2 * invR2) * S.r + (g.c * g.invR) * S.t;
        const cd B = g.invR * S.p;
        const cd C = (g.c * invR2) * S.r - g.invR * S.t;
        E[0] = (g.u * g.u) * A - (g.v * g.v) * B;
        E[1] = (g.u * g.v) * (A + B);
        E[2] = (g.u * g.s) * C;
    } else {
        E[0] = (g.u * g.s * invR2) * S.r;
        E[1] = (g.u * g.invR) * S.t;
        E[2] = (g.v * g.invR) * S.p;
    }
}

// |x_hat + alpha E|^2, theory/LorenzMie.py:131-133 with alpha
LMZ_HD double intensity(const cd (&E)[3], double alpha) {
    const double e0r = fma(alpha, E[0].re, 1.0), e0i = alpha * E[0].im;
    const double a2 = alpha * alpha;
    return fma(e0r, e0r, e0i * e0i) +
           a2 * (fma(E[1].re, E[1].re, E[1].im * E[1].im) +
                 fma(E[2].re, E[2].re, E[2].im * E[2].im));
}

// ====================================================================
// forward-mode Jacobian of one pixel
// ====================================================================
// Tangents are recomputed alongside the recurrences, never stored:
//   d xi_n / dR = -D_n,  d D_n / dR = (1 - n(n+1)/R^2) xi_n  (Riccati-Bessel
//   equation), d pi_n / dc and d tau_n / dc by differentiating Wiscombe's
//   recurrence, and the coefficient tangents dcoef for a_p, n_p, k_p.
// The geometric factors are then differentiated through
// (R, c, u, v, s)(X, Y, Z); the only 1/rho term, K = (A + B)/rho, vanishes
// on the axis (A + B = O(rho^2)), where dI/dX = dI/dY = 0 by the mirror
// symmetries of the hologram.  a_n and b_n are holomorphic in
// m = (n_p + i k_p) / n_m, so d/dk_p = i d/dn_p: the k_p column comes from
// the n_p sums and needs no sums of its own.
struct JacPixel {
    cf E[3];       // field without alpha
    float I;       // intensity
    float J[LMZ_NJAC];
};

#define LMZ_RHO_AXIS 3.0e-4f   // k rho below which K is taken as 0

// State of the Jacobian's order loop.
struct JacLoop {
    // Sums, each accumulating products of positive factors (see the table
    // layout): mSrR = -S_r,R, mStR = -S_t,R, mSpR = -S_p,R.
    cf Sr, St, Sp, mSrR, mStR, mSpR, Src, Stc, Spc;
    cf Tr[2], Tt[2], Tp[2];      // coefficient tangents: a_p, n_p
    Angular32 ang;
    float dpm, dpn;              // d pi_{n-1} / dc, d pi_n / dc
    cf x1, x2;                   // sigma xi of orders n - 1, n - 2
};

// One order n of the Jacobian's loop; ODD = n odd.  The Riccati-Bessel
// functions are carried as sigma_n xi_n and sigma_n D_n with the period-4
// sign sigma_n = +, +, -, -, ... for n = -1, 0, 1, 2, ... (order_sign), so
// that xi's three-term recurrence becomes sigma_n xi_n = s_n (2n-1)/kr
// sigma_{n-1} xi_{n-1} + sigma_{n-2} xi_{n-2} with s_n = sigma_n
// sigma_{n-1} = (-1)^n: a two-operand FMA into the register of the order
// that is no longer needed.  sigma_n D_n = n/kr sigma_n xi_n - s_n
// sigma_{n-1} xi_{n-1}.  The coefficient rows the products read (slots
// 33-38 and the dcoef triples) carry sigma_n, so every product is that of
// the unsigned functions.
template <bool ODD>
LMZ_HD void jac_order(JacLoop &L, int n, const float *coef, const float *dcoef,
                      float w, float cl, float invRl, float invR2) {
    const float *cp = coef + LMZ_COEF_STRIDE * (n - 1);
    const float *dp = dcoef + LMZ_DCOEF_STRIDE * (n - 1);
    const float nn1 = cp[7];
    const float c1 = cp[28];
    float pn, tau;
    L.ang.step(w, cp, pn, tau);
    // tangent of Wiscombe's recurrence with respect to c
    const float dSv = fmaf(L.dpn, cl, pn);
    const float dT = dSv - L.dpm;
    const float dtau = fmaf(cp[32], dT, L.dpm);     // dpm - n dT
    const float ninv = cp[6] * invRl;
    const float cx = (ODD ? cp[39] : cp[31]) * invRl;   // s_n (2n - 1) / kr
    const cf xi{fmaf(cx, L.x1.re, L.x2.re), fmaf(cx, L.x1.im, L.x2.im)};
    const cf D = ODD ? cf{fmaf(ninv, xi.re, L.x1.re), fmaf(ninv, xi.im, L.x1.im)}
                     : cf{fmaf(ninv, xi.re, -L.x1.re), fmaf(ninv, xi.im, -L.x1.im)};
    const float mgn = fmaf(nn1, invR2, -1.f);      // -(1 - n(n+1)/R^2)
    // PR = n(n+1) P: the S_r terms use P xi and P D with pi_n n(n+1)
    const cf PD = cmul3(cp[33], cp[34], cp[35], D);
    const cf PX = cmul3(cp[33], cp[34], cp[35], xi);
    const cf mQX = cmul3(cp[36], cp[37], cp[38], xi); // -Q xi
    const cf mQD = cmul3(cp[36], cp[37], cp[38], D);  // -Q D
    const float pnn = pn * nn1;
    // value
    cmac(L.Sr, pnn, PX);
    cmac(L.St, tau, PD);
    cmac(L.St, pn, mQX);
    cmac(L.Sp, pn, PD);
    cmac(L.Sp, tau, mQX);
    // d/dR (negated)
    cmac(L.mSrR, pnn, PD);
    const float mtg = tau * mgn, mpg = pn * mgn;
    cmac(L.mStR, mtg, PX);
    cmac(L.mStR, pn, mQD);
    cmac(L.mSpR, mpg, PX);
    cmac(L.mSpR, tau, mQD);
    // d/dc
    cmac(L.Src, L.dpn * nn1, PX);
    cmac(L.Stc, dtau, PD);
    cmac(L.Stc, L.dpn, mQX);
    cmac(L.Spc, L.dpn, PD);
    cmac(L.Spc, dtau, mQX);
    // coefficient tangents; dPR_q pi xi as dPR_q (pi xi), the product pi xi
    // shared by both directions (four FMAs a direction, not two multiplies
    // and four FMAs)
    const cf pxi{pn * xi.re, pn * xi.im};
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const float *d = dp + 9 * q;
        const cf a2 = cmul3(d[3], d[4], d[5], D);
        const cf b3 = cmul3(d[6], d[7], d[8], xi);   // -dQ xi
        L.Tr[q].re = fmaf(d[0], pxi.re, fmaf(d[2], pxi.im, L.Tr[q].re));
        L.Tr[q].im = fmaf(d[0], pxi.im, fmaf(d[1], pxi.re, L.Tr[q].im));
        cmac(L.Tt[q], tau, a2);
        cmac(L.Tt[q], pn, b3);
        cmac(L.Tp[q], pn, a2);
        cmac(L.Tp[q], tau, b3);
    }
    L.dpm = L.dpn;
    L.dpn = fmaf(c1, dT, dSv);
    L.x2 = L.x1;
    L.x1 = xi;
}

// The order loop of one pixel (or one radial table sample): the sums of
// JacLoop from the starting values (xi_m1, xi_0).
LMZ_HD void jacobian_loop(float w, float cl, float invRl, float invR,
                          cf xi_m1, cf xi_0, const float *coef,
                          const float *dcoef, int nmax, JacLoop &L) {
    L.Sr = L.St = L.Sp = L.mSrR = L.mStR = L.mSpR = cf{0.f, 0.f};
    L.Src = L.Stc = L.Spc = cf{0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 2; ++q) L.Tr[q] = L.Tt[q] = L.Tp[q] = cf{0.f, 0.f};
    L.ang.init();
    L.dpm = 0.f;
    L.dpn = 0.f;
    L.x2 = xi_m1;      // sigma_{-1} = sigma_0 = 1
    L.x1 = xi_0;
    const float invR2 = invR * invR;
    int n = 1;
    for (; n + 1 <= nmax; n += 2) {
        jac_order<true>(L, n, coef, dcoef, w, cl, invRl, invR2);
        jac_order<false>(L, n + 1, coef, dcoef, w, cl, invRl, invR2);
    }
    if (n <= nmax) jac_order<true>(L, n, coef, dcoef, w, cl, invRl, invR2);
}

// The intensity and its seven derivatives of one pixel from the sums of
// the order loop (L's sums; its loop state is not read).
// ROT: L's sums are e^{-i psi} times the pixel's (the radial tables'), with
// e^{i psi} = ph: the field is turned once, and the adjoint weights, formed
// from conj(ph) F, meet the unturned sums (Re[conj(a) ph b] =
// Re[conj(conj(ph) a) b]).  jscale multiplies every derivative.
template <bool ROT = false>
LMZ_HD void jacobian_project(const FeatUniform &fu, const PixGeom &g,
                             const JacLoop &L, cf ph, float jscale,
                             JacPixel &out) {
    const float invR2 = g.invR * g.invR;
    // ---- projection by adjoint weights.  The field is E = M S, S = (S_r,
    // S_t, S_p), with the real 3 x 3 geometry matrix of field_from_sums:
    //   E_x = u^2 A - v^2 B,  E_y = u v (A + B),  E_z = u s C,
    //   A = a_r S_r + a_t S_t,  B = S_p / R,  C = c_r S_r - S_t / R.
    // Every derivative of I = |F|^2, F = x_hat + alpha E, is 2 alpha
    // Re[conj(F) . dE], and Re[conj(F) . M X] = sum_k credot(g_k, X_k) with
    // g = M^T F:  g_r = a_r W + c_r U, g_t = a_t W - U / R, g_p = V / R for
    // W = u^2 F_x + u v F_y, V = u v F_y - v^2 F_x, U = u s F_z.  So the sums'
    // derivatives (d/dR, d/dc, the coefficient tangents) enter through three
    // complex dot products each, and the geometry's own derivatives through
    // a few scalars.  The loop's S_p, T_p and d/dc of S_r, S_t carry the
    // mirror sign m, its d/dR sums are negated (and those of S_r, S_t carry
    // m as well); both are folded into the weights.
    const cf Sr = L.Sr, St = L.St, SpL = L.Sp;
    const float m = g.mir;
    const float s = g.s, s2 = g.s * g.s, u = g.u, v = g.v;
    const float c = g.c, invR = g.invR;
    const float u2 = u * u, v2 = v * v, uv = u * v, us = u * s;
    const float a_r = s2 * invR2, a_t = c * invR;          // A = a_r Sr + a_t St
    const float c_r = c * invR2;                            // C = c_r Sr - invR St
    const cf A = a_r * Sr + a_t * St;
    const cf B = (m * invR) * SpL;
    const cf C = c_r * Sr - invR * St;
    const cf ApB = A + B;
    out.E[0] = u2 * A - v2 * B;
    out.E[1] = uv * ApB;
    out.E[2] = us * C;
    if constexpr (ROT) {
#pragma unroll
        for (int i = 0; i < 3; ++i) out.E[i] = cmul(ph, out.E[i]);
    }
    const float al = fu.alpha;
    cf F0{fmaf(al, out.E[0].re, 1.f), al * out.E[0].im};
    cf F1 = al * out.E[1], F2 = al * out.E[2];
    out.I = credot(F0, F0) + credot(F1, F1) + credot(F2, F2);
    if constexpr (ROT) {
        const cf pc{ph.re, -ph.im};
        F0 = cmul(pc, F0);
        F1 = cmul(pc, F1);
        F2 = cmul(pc, F2);
    }

    const cf W = u2 * F0 + uv * F1, V = uv * F1 - v2 * F0, U = us * F2;
    const cf gr = a_r * W + c_r * U, gt = a_t * W - invR * U, gp = invR * V;
    const cf gpm = m * gp;
    const float rr = credot(gr, Sr), rt = credot(gt, St), rp = credot(gpm, SpL);
    // d/dR: the sums' partials, and the explicit R^-2 (S_r) and R^-1 (S_t,
    // S_p) factors of M
    const float GR =
        -m * (credot(gr, L.mSrR) + credot(gt, L.mStR)) - credot(gp, L.mSpR) -
        invR * (2.f * rr + rt + rp);
    // d/dc: the sums' partials, and d a_r/dc = -2 c / R^2, d a_t/dc = 1/R,
    // d c_r/dc = 1/R^2 (through W and U)
    const cf er = invR2 * U - (2.f * c_r) * W, et = invR * W;
    const float Gc = m * (credot(gr, L.Src) + credot(gt, L.Stc)) +
                     credot(gp, L.Spc) + credot(er, Sr) + credot(et, St);
    // derivatives of the prefactors u^2, v^2, u v (K = (A + B) / rho, zero on
    // the axis, where dI/dX = dI/dY = 0) and u s
    const float irho = (g.rho > LMZ_RHO_AXIS) ? g.irho : 0.f;
    const float k0 = irho * credot(F0, ApB), k1 = irho * credot(F1, ApB);
    const float tz = invR * credot(F2, C);
    const float PX = (2.f * u * v2) * k0 + (v * (v2 - u2)) * k1 + (v2 + u2 * c * c) * tz;
    const float PY = (-2.f * u2 * v) * k0 + (u * (u2 - v2)) * k1 - (uv * s2) * tz;
    const float PZ = -(us * c) * tz;
    // chain rule to X, Y, Z: R_X = s u, R_Y = s v, R_Z = c; c_X = -c s u / R,
    // c_Y = -c s v / R, c_Z = s^2 / R
    const float csR = c * s * invR;
    const float DX = s * u * GR - csR * u * Gc + PX;
    const float DY = s * v * GR - csR * v * Gc + PY;
    const float DZ = c * GR + s2 * invR * Gc + PZ;
    const float two_al = 2.f * al * jscale, k2a = fu.k * two_al;
    out.J[0] = -k2a * DX;
    out.J[1] = -k2a * DY;
    out.J[2] = k2a * (out.E[0].im + DZ);
    // a_p, n_p tangents; k_p: dE/dk_p = i dE/dn_p (the n_p sums times i)
#pragma unroll
    for (int q = 0; q < 2; ++q)
        out.J[3 + q] = two_al * (credot(gr, L.Tr[q]) + credot(gt, L.Tt[q]) +
                                 credot(gpm, L.Tp[q]));
    {
        const cf iTr{-L.Tr[1].im, L.Tr[1].re}, iTt{-L.Tt[1].im, L.Tt[1].re},
            iTp{-L.Tp[1].im, L.Tp[1].re};
        out.J[5] = two_al * (credot(gr, iTr) + credot(gt, iTt) + credot(gpm, iTp));
    }
    out.J[6] = (2.f * jscale) * (rr + rt + rp);
}

LMZ_HD void jacobian_pixel(const FeatUniform &fu, const PixGeom &g,
                           const float *coef, const float *dcoef, int nmax,
                           JacPixel &out) {
    JacLoop L;
    jacobian_loop(g.w, g.cl, g.invRl, g.invR, g.xi_m1, g.xi_0, coef, dcoef,
                  nmax, L);
    jacobian_project(fu, g, L, cf{1.f, 0.f}, 1.f, out);
}

// ====================================================================
// radial tables of the Jacobian's sums (dense crops)
// ====================================================================
// On a dense crop every pixel lies in the plane z = 0, so the sums of the
// order loop depend on the pixel only through its distance rho from the
// particle's axis (R = sqrt(rho^2 + Z^2), cos theta = Z / R), and, since the
// loop is linear in its starting values (xi_-1, xi_0) = e^{i psi} (1,
// -i sig m), they are e^{i psi(rho)} times the sums started from (1, -i sig
// m), which vary slowly in rho: no faster than N / |Z| radians per unit of
// k rho (the angular functions of order N).  So the loop runs once per
// sample of a uniform grid in k rho, with spacing h = LMZ_RTAB_HW |Z| / N
// (at most 1), and each pixel takes its sums by LMZ_RTAB_P-point Lagrange
// interpolation times e^{i psi}; within 1e-6 of the column scale of each
// Jacobian column on the feature (measured against the loop at every pixel).
// The sums are even in rho, so samples at rho < 0 take |rho|.
#define LMZ_RTAB_CAP 1024    // samples per feature
#define LMZ_RTAB_STRIDE 32   // floats per sample: 15 complex sums, 2 spare
#define LMZ_RTAB_P 6         // interpolation points
#define LMZ_RTAB_HW 0.35f    // h N / |Z|
#define LMZ_RTAB_LDS 352      // rows of a table the Jacobian kernel keeps in LDS
#define LMZ_RTAB_GROUP 4      // tiles of a feature per work item of that kernel
#define LMZ_RTAB_LDS_STRIDE 36   // floats per LDS row: rows 4 banks apart

struct RTabPar {
    float rho0, h, inv_h;    // grid k rho_j = rho0 + j h
    int m;                   // samples; 0: the feature uses the order loop
};

// The table grid of a feature at its current parameters: every pixel of the
// crop (height rows of width) lies between the nearest and farthest points
// of the crop from the particle's axis, with LMZ_RTAB_P / 2 samples to
// spare either side.  m = 0 for explicit coordinates (not known to be
// planar), for Z = 0, or when the grid would need more than LMZ_RTAB_CAP
// samples.
LMZ_HD RTabPar rtab_params(const FeatUniform &fu) {
    RTabPar p{0.f, 0.f, 0.f, 0};
    if (fu.width <= 0 || fu.nmax < 1) return p;
    const float Z = fu.k * (fu.iz + fu.fz);     // pixel_geometry at z = 0
    if (!(fabsf(Z) > 0.f)) return p;
    const int height = (fu.npix + fu.width - 1) / fu.width;
    const float px = fu.ix + fu.fx, py = fu.iy + fu.fy;   // axis in the crop
    const float xm = (float)(fu.width - 1), ym = (float)(height - 1);
    const float nx = fmaxf(fmaxf(-px, px - xm), 0.f);
    const float ny = fmaxf(fmaxf(-py, py - ym), 0.f);
    const float fx = fmaxf(px, xm - px), fy = fmaxf(py, ym - py);
    const float rmin = fu.k * sqrtf(fmaf(nx, nx, ny * ny));
    const float rmax = fu.k * sqrtf(fmaf(fx, fx, fy * fy));
    const float h = fminf(LMZ_RTAB_HW * fabsf(Z) / (float)fu.nmax, 1.f);
    const float span = (rmax - rmin) / h + (float)(2 * LMZ_RTAB_P);
    if (!(span < (float)LMZ_RTAB_CAP)) return p;
    p.h = h;
    p.inv_h = 1.f / h;
    p.rho0 = rmin - (float)(LMZ_RTAB_P / 2 + 1) * h;
    p.m = (int)span;
    return p;
}

// Sample j of the table: the order loop at k rho = |rho0 + j h|, z = 0,
// started from (1, -i sig m); 30 floats in JacLoop's sum order.
LMZ_HD void rtab_sample(const FeatUniform &fu, const RTabPar &p, bool bohren,
                        int j, const float *coef, const float *dcoef,
                        float *out) {
    const float Z = fu.k * (fu.iz + fu.fz);
    const float rho = fabsf(fmaf((float)j, p.h, p.rho0));
    const float rho2 = rho * rho;
    const float R2 = fmaf(Z, Z, rho2);
    const float invR = rsqrt_hw(R2);
    const float R = R2 * invR;
    const float az_ = fabsf(Z);
    const float w = rho2 * rcp_hw(R + az_) * invR;   // 1 - |cos theta|
    const float cl = az_ * invR;
    const float mir = (Z < 0.f) ? -1.f : 1.f;
    const float sz = (Z > 0.f) ? 1.f : -1.f;
    const float sig = bohren ? sz : -sz;
    JacLoop L;
    jacobian_loop(w, cl, mir * invR, invR, cf{1.f, 0.f}, cf{0.f, -sig * mir},
                  coef, dcoef, fu.nmax, L);
    const cf v[15] = {L.Sr, L.St, L.Sp, L.mSrR, L.mStR, L.mSpR, L.Src, L.Stc,
                      L.Spc, L.Tr[0], L.Tr[1], L.Tt[0], L.Tt[1], L.Tp[0],
                      L.Tp[1]};
#pragma unroll
    for (int i = 0; i < 15; ++i) {
        out[2 * i] = v[i].re;
        out[2 * i + 1] = v[i].im;
    }
    out[30] = 0.f;
    out[31] = 0.f;
}

// The Jacobian of one pixel of a tabulated feature: its sums interpolated
// in k rho, then projected as the loop's with the turn by e^{i psi}
// (g.xi_m1); every derivative times jscale.  RS: floats from one table row
// to the next (LMZ_RTAB_STRIDE in global memory, LMZ_RTAB_LDS_STRIDE in
// LDS).
template <int RS = LMZ_RTAB_STRIDE>
LMZ_HD void jacobian_pixel_rtab(const FeatUniform &fu, const PixGeom &g,
                                const RTabPar &p, const float *tab,
                                float jscale, JacPixel &out, int base = 0) {
    const float t = (g.rho - p.rho0) * p.inv_h;
    int i0 = (int)floorf(t) - (LMZ_RTAB_P / 2 - 1);
    i0 = i0 < 0 ? 0 : (i0 > p.m - LMZ_RTAB_P ? p.m - LMZ_RTAB_P : i0);
    const float d = t - (float)i0;
    // Lagrange weights on nodes 0..5 at d: prefix and suffix products of
    // (d - k) over the constant denominators
    float e[LMZ_RTAB_P], pre[LMZ_RTAB_P], suf[LMZ_RTAB_P];
#pragma unroll
    for (int k = 0; k < LMZ_RTAB_P; ++k) e[k] = d - (float)k;
    pre[0] = 1.f;
#pragma unroll
    for (int k = 1; k < LMZ_RTAB_P; ++k) pre[k] = pre[k - 1] * e[k - 1];
    suf[LMZ_RTAB_P - 1] = 1.f;
#pragma unroll
    for (int k = LMZ_RTAB_P - 2; k >= 0; --k) suf[k] = suf[k + 1] * e[k + 1];
    const float den[LMZ_RTAB_P] = {-1.f / 120.f, 1.f / 24.f, -1.f / 12.f,
                                   1.f / 12.f, -1.f / 24.f, 1.f / 120.f};
    float wgt[LMZ_RTAB_P];
#pragma unroll
    for (int k = 0; k < LMZ_RTAB_P; ++k) wgt[k] = pre[k] * suf[k] * den[k];
    float acc[30];
    const float *row = tab + (i0 - base) * RS;   // tab holds rows base ..
#pragma unroll
    for (int c = 0; c < 30; ++c) acc[c] = wgt[0] * row[c];
#pragma unroll
    for (int k = 1; k < LMZ_RTAB_P; ++k) {
        const float *rk = row + k * RS;
#pragma unroll
        for (int c = 0; c < 30; ++c) acc[c] = fmaf(wgt[k], rk[c], acc[c]);
    }
    JacLoop L;
    cf *dst[15] = {&L.Sr, &L.St, &L.Sp, &L.mSrR, &L.mStR, &L.mSpR, &L.Src,
                   &L.Stc, &L.Spc, &L.Tr[0], &L.Tr[1], &L.Tt[0], &L.Tt[1],
                   &L.Tp[0], &L.Tp[1]};
#pragma unroll
    for (int i = 0; i < 15; ++i) *dst[i] = cf{acc[2 * i], acc[2 * i + 1]};
    jacobian_project<true>(fu, g, L, g.xi_m1, jscale, out);
}

// ====================================================================
// FP64 radial tables of the forward sums (dense crops)
// ====================================================================
// The forward's sums S_r, S_t, S_p on a dense crop are, as the Jacobian's
// (rtab_params), e^{i psi} times functions of k rho alone; in FP64 they are
// tabulated at spacing h = LMZ_FTAB_HW |Z| / N (at most 1) and each pixel
// takes them by LMZ_FTAB_P-point Lagrange interpolation: within a few ulp
// of the largest sum of the order loop at every pixel (measured from
// a_p 0.5 to 3 um, z_p 50 to 400 px), so the intensity agrees with the
// loop's to FP64 rounding.
#define LMZ_FTAB_CAP 2048    // samples per feature
#define LMZ_FTAB_STRIDE 6    // doubles per sample: S_r, S_t, S_p
#define LMZ_FTAB_P 12        // interpolation points
#define LMZ_FTAB_HW 0.25     // h N / |Z|
#define LMZ_FTAB_LDS 640     // rows of a table the forward kernel keeps in LDS
#define LMZ_FTAB_GROUP 8     // tiles of a feature per work item of that kernel

struct FTabPar {
    double rho0, h, inv_h;   // grid k rho_j = rho0 + j h
    int m;                   // samples; 0: the feature uses the order loop
};

// The rows lo .. hi of a radial table (grid k rho_i = rho0 + i / inv_h,
// m rows, P-point stencils placed as jacobian_pixel_rtab and
// forward_sum_ftab place them) that pixels j0 .. j1 - 1 of a dense crop
// read: the stencils of the nearest and the farthest point of the pixels'
// bounding rectangle from the particle's axis (px, py) (k: wavenumber per
// pixel), two rows wider on either side for the rounding of each pixel's
// own k rho.
template <int P, typename T>
LMZ_HD void table_rows(int width, int j0, int j1, T k, T px, T py, T rho0,
                       T inv_h, int m, int &lo, int &hi) {
    const int r0 = j0 / width, r1 = (j1 - 1) / width;
    const T c0 = (T)(r0 == r1 ? j0 - r0 * width : 0);
    const T c1 = (T)(r0 == r1 ? j1 - 1 - r1 * width : width - 1);
    const T y0 = (T)r0, y1 = (T)r1;
    const T nx = (px < c0 ? c0 : (px > c1 ? c1 : px)) - px;
    const T ny = (py < y0 ? y0 : (py > y1 ? y1 : py)) - py;
    const T fx = (px - c0 > c1 - px) ? px - c0 : c1 - px;
    const T fy = (py - y0 > y1 - py) ? py - y0 : y1 - py;
    const T tmin = (k * sqrt(nx * nx + ny * ny) - rho0) * inv_h;
    const T tmax = (k * sqrt(fx * fx + fy * fy) - rho0) * inv_h;
    lo = (int)floor(tmin) - (P / 2 - 1) - 2;
    hi = (int)floor(tmax) - (P / 2 - 1) + (P - 1) + 2;
    lo = lo < 0 ? 0 : (lo > m - P ? m - P : lo);
    hi = hi < P - 1 ? P - 1 : (hi > m - 1 ? m - 1 : hi);
}

// The first row of an LDS window of cap rows (cap < m) that holds rows
// lo .. hi (hi - lo < cap), centred on them where the table allows.
LMZ_HD int table_window(int lo, int hi, int m, int cap) {
    int sb = lo - (cap - (hi - lo + 1)) / 2;
    return sb < 0 ? 0 : (sb > m - cap ? m - cap : sb);
}

// The grid of a feature at its current parameters (as rtab_params, in
// FP64); m = 0 for explicit coordinates, Z = 0, or more than LMZ_FTAB_CAP
// samples.
LMZ_HD FTabPar ftab_params(const FeatUniform &fu) {
    FTabPar p{0.0, 0.0, 0.0, 0};
    if (fu.width <= 0 || fu.nmax < 1) return p;
    const double Z = fu.kd * ((double)fu.iz + fu.fzd);   // pixel_geometry_d, z = 0
    if (!(fabs(Z) > 0.0)) return p;
    const int height = (fu.npix + fu.width - 1) / fu.width;
    const double px = (double)fu.ix + fu.fxd, py = (double)fu.iy + fu.fyd;
    const double xm = (double)(fu.width - 1), ym = (double)(height - 1);
    const double nx = fmax(fmax(-px, px - xm), 0.0);
    const double ny = fmax(fmax(-py, py - ym), 0.0);
    const double fx = fmax(px, xm - px), fy = fmax(py, ym - py);
    const double rmin = fu.kd * sqrt(fma(nx, nx, ny * ny));
    const double rmax = fu.kd * sqrt(fma(fx, fx, fy * fy));
    const double h = fmin(LMZ_FTAB_HW * fabs(Z) / (double)fu.nmax, 1.0);
    const double span = (rmax - rmin) / h + (double)(2 * LMZ_FTAB_P);
    if (!(span < (double)LMZ_FTAB_CAP)) return p;
    p.h = h;
    p.inv_h = 1.0 / h;
    p.rho0 = rmin - (double)(LMZ_FTAB_P / 2 + 1) * h;
    p.m = (int)span;
    return p;
}

// Sample j: the FP64 order loop at k rho = |rho0 + j h|, z = 0, started
// from (1, -i sig m) (pixel_geometry_d with psi = 0).
LMZ_HD void ftab_sample(const FeatUniform &fu, const FTabPar &p, bool bohren,
                        int j, const float *coef, double *out) {
    const double Z = fu.kd * ((double)fu.iz + fu.fzd);
    const double rho = fabs(fma((double)j, p.h, p.rho0));
    const double rho2 = rho * rho;
    PixGeomD g{};
    g.R = sqrt(fma(Z, Z, rho2));
    g.invR = 1.0 / g.R;
    const double az_ = fabs(Z);
    g.w = rho2 / (g.R + az_) * g.invR;
    g.mir = (Z < 0.0) ? -1.0 : 1.0;
    g.invRl = g.mir * g.invR;
    const double sz = (Z > 0.0) ? 1.0 : -1.0;
    const double sig = bohren ? sz : -sz;
    g.xi_m1 = cd{1.0, 0.0};
    g.xi_0 = g.mir * cd{0.0, -sig};
    const SumsD S = forward_sum(g, coef, fu.nmax);
    out[0] = S.r.re; out[1] = S.r.im;
    out[2] = S.t.re; out[3] = S.t.im;
    out[4] = S.p.re; out[5] = S.p.im;
}

// The forward sums of one pixel of a tabulated feature: interpolated in
// k rho and turned by e^{i psi} (g.xi_m1).  RS: doubles per table row.
template <int RS = LMZ_FTAB_STRIDE>
LMZ_HD SumsD forward_sum_ftab(const PixGeomD &g, const FTabPar &p,
                              const double *tab, int base = 0) {
    constexpr int P = LMZ_FTAB_P;
    const double t = (g.rho - p.rho0) * p.inv_h;
    int i0 = (int)floor(t) - (P / 2 - 1);
    i0 = i0 < 0 ? 0 : (i0 > p.m - P ? p.m - P : i0);
    const double d = t - (double)i0;
    // weights on nodes 0..P-1: prefix and suffix products of (d - k) over
    // prod_{q != k} (k - q)
    constexpr double inv_den[P] = {
        1.0 / -39916800.0, 1.0 / 3628800.0, 1.0 / -725760.0, 1.0 / 241920.0,
        1.0 / -120960.0, 1.0 / 86400.0, 1.0 / -86400.0, 1.0 / 120960.0,
        1.0 / -241920.0, 1.0 / 725760.0, 1.0 / -3628800.0, 1.0 / 39916800.0};
    double pre[P], suf[P];
    pre[0] = 1.0;
#pragma unroll
    for (int k = 1; k < P; ++k) pre[k] = pre[k - 1] * (d - (double)(k - 1));
    suf[P - 1] = 1.0;
#pragma unroll
    for (int k = P - 2; k >= 0; --k) suf[k] = suf[k + 1] * (d - (double)(k + 1));
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const double *row = tab + (i0 - base) * RS;   // tab holds rows base ..
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const double wk = pre[k] * suf[k] * inv_den[k];
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[c] = fma(wk, row[k * RS + c], acc[c]);
    }
    const cd ph = g.xi_m1;   // e^{i psi}
    return SumsD{ph * cd{acc[0], acc[1]}, ph * cd{acc[2], acc[3]},
                 ph * cd{acc[4], acc[5]}};
}

// ====================================================================
// Mie coefficients in FP64 with a_p, n_p, k_p tangents
// ====================================================================
// doubles of scratch per order and feature and pass; the batch's scratch
// (lmz_batch_sizes.work_doubles) holds LMZ_WORK_PASSES passes, which the
// layered-sphere fit runs side by side (value, a_p and n_p tangents)
#define LMZ_WORK_PER_ORDER 16
#define LMZ_WORK_PASSES 3
// Sphere.wiscombe_yang, theory/Sphere.py:91-123 (one layer: the rolled
// term equals |x m|)
LMZ_HD int wiscombe_yang(double x, cd m) {
    const double xl = fabs(x);
    double ns;
    if (xl <= 8.0) ns = floor(xl + 4.0 * cbrt(xl) + 1.0);
    else if (xl <= 4200.0) ns = floor(xl + 4.05 * cbrt(xl) + 2.0);
    else ns = floor(xl + 4.0 * cbrt(xl) + 2.0);
    const double xm = hypot(x * m.re, x * m.im);
    return (int)fmax(ns, xm);
}

// sigma_n = +1, +1, -1, -1, ... for n = -1, 0, 1, 2, ... (period 4): the
// sign the Jacobian's Riccati-Bessel functions carry (jac_order).
LMZ_HD float order_sign(int n) { return ((n + 1) & 2) ? -1.f : 1.f; }

// Output of one order: FP32 table rows and optional FP64 a_n, b_n.
LMZ_HD void put_d(float *row, int i, double v) { __builtin_memcpy(row + i, &v, sizeof v); }

LMZ_HD void write_order(int n, cd a, cd b, float *crow) {
    const double nn1 = (double)n * (n + 1.0);
    const double en = (2.0 * n + 1.0) / nn1;
    const cd P = mul_ipow(n + 1, en * a);           // i E_n a_n
    const cd Q = mul_ipow(n, en * b);               // E_n b_n
    const cd PR = nn1 * P;                          // i^(n+1) (2n+1) a_n
    crow[0] = -(float)P.im; crow[1] = (float)Q.im;
    crow[2] = (float)P.re;  crow[3] = (float)P.im;
    crow[4] = -(float)Q.re; crow[5] = -(float)Q.im;
    crow[6] = (float)n;
    crow[7] = (float)nn1;
    put_d(crow, 8, 1.0 + 1.0 / n);
    put_d(crow, 10, 0.5 * nn1);
    put_d(crow, 12, 0.5 * (n + 1.0) * (n + 2.0));
    put_d(crow, 14, (double)n);
    put_d(crow, 16, PR.re); put_d(crow, 18, PR.im);
    put_d(crow, 20, P.re);  put_d(crow, 22, P.im);
    put_d(crow, 24, Q.re);  put_d(crow, 26, Q.im);
    crow[28] = (float)(1.0 + 1.0 / n);
    crow[29] = (float)(-0.5 * nn1);
    crow[30] = (float)(0.5 * (n + 1.0) * (n + 2.0));
    crow[31] = (float)(2.0 * n - 1.0);
    crow[32] = (float)(-n);
    // the P and -Q triples of slots 2, 3, 0 and 4, 5, 1 times sigma_n, for
    // the Jacobian's signed Riccati-Bessel functions (jac_order)
    const float sg = order_sign(n);
    crow[33] = sg * crow[2]; crow[34] = sg * crow[3]; crow[35] = sg * crow[0];
    crow[36] = sg * crow[4]; crow[37] = sg * crow[5]; crow[38] = sg * crow[1];
    crow[39] = (float)(1.0 - 2.0 * n);   // s_n (2n - 1) for odd n
    for (int i = 40; i < LMZ_COEF_STRIDE; ++i) crow[i] = 0.f;
}

// Tangent rows: q = 0 (a_p), 1 (n_p) as (Re, Im, -Im) triples of PR, P,
// -Q at 9 q, times sigma_n (order_sign: the Jacobian reads them with its
// signed Riccati-Bessel functions); write_dorder_k: the k_p tangents (Re,
// Im of PR, P, Q) at 18, unsigned.
LMZ_HD void write_dorder(int n, cd da, cd db, float *drow) {
    const double en = (2.0 * n + 1.0) / ((double)n * (n + 1.0));
    const double sg = order_sign(n);
    const cd P = mul_ipow(n + 1, (sg * en) * da);
    const cd Q = mul_ipow(n, (sg * en) * db);
    const cd PR = ((double)n * (n + 1.0)) * P;
    drow[0] = (float)PR.re; drow[1] = (float)PR.im; drow[2] = -(float)PR.im;
    drow[3] = (float)P.re;  drow[4] = (float)P.im;  drow[5] = -(float)P.im;
    drow[6] = -(float)Q.re; drow[7] = -(float)Q.im; drow[8] = (float)Q.im;
}
LMZ_HD void write_dorder_k(int n, cd da, cd db, float *drow) {
    const double en = (2.0 * n + 1.0) / ((double)n * (n + 1.0));
    const cd P = mul_ipow(n + 1, en * da);
    const cd Q = mul_ipow(n, en * db);
    const cd PR = ((double)n * (n + 1.0)) * P;
    drow[0] = (float)PR.re; drow[1] = (float)PR.im;
    drow[2] = (float)P.re;  drow[3] = (float)P.im;
    drow[4] = (float)Q.re;  drow[5] = (float)Q.im;
}

struct MieSetup {
    double kum, x, n_m;
    cd m;
    int N;
};

LMZ_HD MieSetup mie_setup(double a_p, double n_p, double k_p, double n_m,
                          double wavelength, int cap) {
    MieSetup ms;
    ms.kum = 2.0 * M_PI / wavelength * n_m;    // rad / um in the medium
    ms.x = ms.kum * a_p;
    ms.n_m = n_m;
    ms.m = cd{n_p / n_m, k_p / n_m};
    int N = wiscombe_yang(ms.x, ms.m);
    if (N > cap) N = cap;
    if (N < 1) N = 1;
    ms.N = N;
    return ms;
}

#define LMZ_W(n, j) work[((long long)(n) * LMZ_WORK_PER_ORDER + (j)) * wstride]

// Phase 1: D1_n(z) downward from D1_N = 0 (Eq. 16b) with dD1_n/dz, into
// columns col..col+3 (re, im, d re, d im) for n = 0..N.  n / z is a true
// division, as theory/Sphere.py:196, 203 evaluate it: for real z near the
// poles of D1 the recursion amplifies a last-bit difference in n / z to
// 1e-9 in a_n (a_p = 9.5 um), which n * (1 / z) would introduce.
LMZ_HD void mie_d1_down(cd z, int N, double *work, long long wstride, int col) {
    const cd iz = cinv(z);
    cd d1{0.0, 0.0}, dd1{0.0, 0.0};
    LMZ_W(N, col) = 0.0; LMZ_W(N, col + 1) = 0.0;
    LMZ_W(N, col + 2) = 0.0; LMZ_W(N, col + 3) = 0.0;
    for (int n = N; n >= 1; --n) {
        const cd nz = cd{(double)n, 0.0} / z;         // n / z
        const cd dnz = -(nz * iz);                    // -n / z^2
        const cd inv = cinv(d1 + nz);
        dd1 = dnz + (dd1 + dnz) * (inv * inv);
        d1 = nz - inv;
        LMZ_W(n - 1, col) = d1.re; LMZ_W(n - 1, col + 1) = d1.im;
        LMZ_W(n - 1, col + 2) = dd1.re; LMZ_W(n - 1, col + 3) = dd1.im;
    }
}

// Phase 2: Psi_n(x) (Eq. 20), zeta_n(x) (Eq. 21), Psi_n zeta_n and D3_n
// (Eqs. 18a-d) upward with d/dx; Psi, dPsi into columns 8, 9, zeta and
// dzeta into 10-13, for n = 0..N.  Needs D1(x) in column 4, dD1/dx in 6.
LMZ_HD void mie_up_chain(double x, int N, double *work, long long wstride) {
    double sx, cx;
    sincos(x, &sx, &cx);
    double s2x, c2x;
    sincos(2.0 * x, &s2x, &c2x);
    double psi = sx, dpsi = cx;                      // Eq. 20a
    cd zeta{sx, -cx}, dzeta{cx, sx};                 // -i e^{ix}, Eq. 21a
    cd pz{0.5 * (1.0 - c2x), -0.5 * s2x};            // Eq. 18a
    cd dpz{s2x, -c2x};                               // d/dx = -i e^{2ix}
    cd d3{0.0, 1.0}, dd3{0.0, 0.0};                  // Eq. 18b
    for (int n = 1; n <= N; ++n) {
        // D1 and dD1 of orders n - 1 and n are read before the stores of
        // order n - 1, which are held from the previous pass: a load issued
        // after stores waits for them to complete (one memory counter
        // covers both), which put two memory round trips on the chain's
        // critical path per order.
        const double d1m = LMZ_W(n - 1, 4), dd1m = LMZ_W(n - 1, 6);
        const double d1n = LMZ_W(n, 4), dd1n = LMZ_W(n, 6);
        LMZ_W(n - 1, 8) = psi; LMZ_W(n - 1, 9) = dpsi;
        LMZ_W(n - 1, 10) = zeta.re; LMZ_W(n - 1, 11) = zeta.im;
        LMZ_W(n - 1, 12) = dzeta.re; LMZ_W(n - 1, 13) = dzeta.im;
        const double nx = n / x, dnx = -n / (x * x);
        const double w1 = nx - d1m;
        const double dw1 = dnx - dd1m;
        const cd w3 = cd{nx, 0.0} - d3;
        const cd dw3 = cd{dnx, 0.0} - dd3;
        const double psi_m = psi;
        const cd zeta_m = zeta;
        psi = psi_m * w1;                              // Eq. 20b
        dpsi = dpsi * w1 + psi_m * dw1;
        zeta = zeta_m * w3;                            // Eq. 21b
        dzeta = dzeta * w3 + zeta_m * dw3;
        const cd pz_m = pz;
        pz = pz_m * (w1 * w3);                         // Eq. 18c
        dpz = dpz * (w1 * w3) + pz_m * (dw1 * w3 + w1 * dw3);
        const cd ipz = cinv(pz);
        d3 = cd{d1n, 0.0} + mul_i(ipz);                // Eq. 18d
        dd3 = cd{dd1n, 0.0} - mul_i(dpz * (ipz * ipz));
    }
    LMZ_W(N, 8) = psi; LMZ_W(N, 9) = dpsi;
    LMZ_W(N, 10) = zeta.re; LMZ_W(N, 11) = zeta.im;
    LMZ_W(N, 12) = dzeta.re; LMZ_W(N, 13) = dzeta.im;
}

// Phase 3: a_n, b_n (Eqs. 5 and 6 with H_a = H_b = D1(m x)) and their
// tangents for one order n >= 1, written as table rows.
LMZ_HD void mie_order(const MieSetup &ms, int n, const double *work,
                      long long wstride, float *coef, float *dcoef,
                      double *ab) {
    const double x = ms.x;
    const cd m = ms.m;
    const double nx = n / x, dnx = -n / (x * x);
    const double psi = LMZ_W(n, 8), dpsi = LMZ_W(n, 9);
    const double psi_m = LMZ_W(n - 1, 8), dpsi_m = LMZ_W(n - 1, 9);
    const cd zeta{LMZ_W(n, 10), LMZ_W(n, 11)}, dzeta{LMZ_W(n, 12), LMZ_W(n, 13)};
    const cd zeta_m{LMZ_W(n - 1, 10), LMZ_W(n - 1, 11)};
    const cd dzeta_m{LMZ_W(n - 1, 12), LMZ_W(n - 1, 13)};
    const cd H{LMZ_W(n, 0), LMZ_W(n, 1)};
    const cd dH{LMZ_W(n, 2), LMZ_W(n, 3)};                 // dD1/dz at z = m x
    const cd im = cinv(m);
    const cd fa = H * im + cd{nx, 0.0};
    const cd fa_x = dH + cd{dnx, 0.0};             // (dH m)/m - n/x^2
    const cd fa_m = x * (dH * im) - H * (im * im);
    const cd numA = fa * cd{psi, 0.0} - cd{psi_m, 0.0};
    const cd denA = fa * zeta - zeta_m;
    const cd idA = cinv(denA);
    const cd a = numA * idA;
    const cd a_x = (fa_x * cd{psi, 0.0} + fa * cd{dpsi, 0.0} -
                    cd{dpsi_m, 0.0} -
                    a * (fa_x * zeta + fa * dzeta - dzeta_m)) * idA;
    const cd a_m = fa_m * (cd{psi, 0.0} - a * zeta) * idA;
    const cd fb = m * H + cd{nx, 0.0};
    const cd fb_x = (m * m) * dH + cd{dnx, 0.0};
    const cd fb_m = H + x * (m * dH);
    const cd numB = fb * cd{psi, 0.0} - cd{psi_m, 0.0};
    const cd denB = fb * zeta - zeta_m;
    const cd idB = cinv(denB);
    const cd b = numB * idB;
    const cd b_x = (fb_x * cd{psi, 0.0} + fb * cd{dpsi, 0.0} -
                    cd{dpsi_m, 0.0} -
                    b * (fb_x * zeta + fb * dzeta - dzeta_m)) * idB;
    const cd b_m = fb_m * (cd{psi, 0.0} - b * zeta) * idB;
    if (coef) write_order(n, a, b, coef + LMZ_COEF_STRIDE * (n - 1));
    if (dcoef) {
        const double inm = 1.0 / ms.n_m;
        float *drow = dcoef + LMZ_DCOEF_STRIDE * (n - 1);
        write_dorder(n, ms.kum * a_x, ms.kum * b_x, drow);                 // a_p
        write_dorder(n, inm * a_m, inm * b_m, drow + 9);                  // n_p
        write_dorder_k(n, mul_i(inm * a_m), mul_i(inm * b_m), drow + 18); // k_p
    }
    if (ab) {
        double *r = ab + 4 * (n - 1);
        r[0] = a.re; r[1] = a.im; r[2] = b.re; r[3] = b.im;
    }
}

// All three phases in sequence on one thread.  Returns n_max (clamped to
// cap).
LMZ_HD int mie_feature(double a_p, double n_p, double k_p, double n_m,
                       double wavelength, int cap, double *work,
                       long long wstride, float *coef, float *dcoef,
                       double *ab) {
    const MieSetup ms = mie_setup(a_p, n_p, k_p, n_m, wavelength, cap);
    mie_d1_down(ms.x * ms.m, ms.N, work, wstride, 0);
    mie_d1_down(cd{ms.x, 0.0}, ms.N, work, wstride, 4);
    mie_up_chain(ms.x, ms.N, work, wstride);
    for (int n = 1; n <= ms.N; ++n)
        mie_order(ms, n, work, wstride, coef, dcoef, ab);
    return ms.N;
}
#undef LMZ_W

// Order count of a layered sphere: Wiscombe (1980) on the outer size
// parameter with Yang (2003) Eq. 30, N = max(n_s, |m_l x_l|, |m_l x_{l-1}|).
// theory/Sphere.py:91-123 states the same rule; its np.roll term pairs x_l
// with m_{l-1} and wraps around, which the reference never reaches (its
// scalar-only branch fails for arrays), so Yang's pairing is used.
LMZ_HD int wiscombe_yang_layers(const double *x, const cd *m, int L) {
    int N = wiscombe_yang(x[L - 1], cd{0.0, 0.0});
    for (int l = 0; l < L; ++l) {
        N = (int)fmax((double)N, hypot(x[l] * m[l].re, x[l] * m[l].im));
        if (l > 0)
            N = (int)fmax((double)N, hypot(x[l - 1] * m[l].re, x[l - 1] * m[l].im));
    }
    return N;
}

LMZ_HD cd cexp_i2(cd z) {          // exp(2 i z)
    const double r = exp(-2.0 * z.im);
    double sn, cs;
    sincos(2.0 * z.re, &sn, &cs);
    return cd{r * cs, r * sn};
}

// Layered sphere (L layers, innermost first, outer radii increasing):
// Yang, Appl. Opt. 42, 1710 (2003) Eqs. 7-8 and 19 with Pena & Pal (2009)
// Eqs. 5-6, the recursion lib/obsolete/sphere_coefficients.py:126-215
// describes.  H_a = H_b = D1(m_1 x_1) in the core; for each further layer
// l, D1 downward at z_a = m_l x_{l-1} and z_b = m_l x_l from D1[N] = 0 (as
// theory/Sphere.py:190), D3 upward through Psi zeta, the ratio
// Q_n = [Psi_n/zeta_n](z_a) / [Psi_n/zeta_n](z_b) upward, and the update of
// H_a, H_b; then Eqs. 5-6 on the outer boundary.  FP64, no tangents; one
// layer reproduces mie_feature's coefficients.  layer l: radius[l] [um],
// index n[l] + i k[l].  work as mie_feature.  Returns n_max (clamped).
// n_max of a layered sphere, clamped to [1, cap] (mie_layered_feature's).
LMZ_HD int layered_nmax(const double *radius, const double *nl,
                        const double *kl, int L, double n_m,
                        double wavelength, int cap) {
    const double kum = 2.0 * M_PI / wavelength * n_m;
    double x[LMZ_MAX_LAYERS];
    cd m[LMZ_MAX_LAYERS];
    for (int l = 0; l < L; ++l) {
        x[l] = kum * radius[l];
        m[l] = cd{nl[l] / n_m, kl[l] / n_m};
    }
    int N = wiscombe_yang_layers(x, m, L);
    if (N > cap) N = cap;
    if (N < 1) N = 1;
    return N;
}

LMZ_HD int mie_layered_feature(const double *radius, const double *nl,
                               const double *kl, int L, double n_m,
                               double wavelength, int cap, double *work,
                               long long wstride, float *coef, double *ab) {
    const double kum = 2.0 * M_PI / wavelength * n_m;
    double x[LMZ_MAX_LAYERS];
    cd m[LMZ_MAX_LAYERS];
    for (int l = 0; l < L; ++l) {
        x[l] = kum * radius[l];
        m[l] = cd{nl[l] / n_m, kl[l] / n_m};
    }
    const int N = layered_nmax(radius, nl, kl, L, n_m, wavelength, cap);
#define W(n, j) work[((long long)(n) * LMZ_WORK_PER_ORDER + (j)) * wstride]
    // D1 downward from D1[N] = 0 into work columns (c, c + 1), orders 0..N
    auto d1_down = [&](cd z, int c) {
        cd d1{0.0, 0.0};
        W(N, c) = 0.0;
        W(N, c + 1) = 0.0;
        for (int n = N; n >= 1; --n) {
            const cd nz = cd{(double)n, 0.0} / z;     // as mie_d1_down
            d1 = nz - cinv(d1 + nz);
            W(n - 1, c) = d1.re;
            W(n - 1, c + 1) = d1.im;
        }
    };
    // core: H_a = H_b = D1(m_1 x_1) in columns 0..3
    d1_down(x[0] * m[0], 0);
    for (int n = 0; n <= N; ++n) {
        W(n, 2) = W(n, 0);
        W(n, 3) = W(n, 1);
    }
    for (int l = 1; l < L; ++l) {
        const cd za = x[l - 1] * m[l], zb = x[l] * m[l];
        d1_down(za, 4);
        d1_down(zb, 6);
        const cd iza = cinv(za), izb = cinv(zb);
        const cd ea = cexp_i2(za), eb = cexp_i2(zb);
        cd pza = 0.5 * (cd{1.0, 0.0} - ea), pzb = 0.5 * (cd{1.0, 0.0} - eb);
        cd d3a{0.0, 1.0}, d3b{0.0, 1.0};
        // Q_0 = exp(2i(zb - za)) (exp(2i za) - 1) / (exp(2i zb) - 1)
        cd q = cexp_i2(zb - za) * (ea - cd{1.0, 0.0}) / (eb - cd{1.0, 0.0});
        const double r2 = (x[l - 1] / x[l]) * (x[l - 1] / x[l]);
        for (int n = 1; n <= N; ++n) {
            const cd d1a_m{W(n - 1, 4), W(n - 1, 5)}, d1b_m{W(n - 1, 6), W(n - 1, 7)};
            const cd d1a{W(n, 4), W(n, 5)}, d1b{W(n, 6), W(n, 7)};
            const cd nza = (double)n * iza, nzb = (double)n * izb;
            // Q_n with D3_{n-1}, Yang Eq. 19
            const cd num = (zb * d1b + cd{(double)n, 0.0}) *
                           (cd{(double)n, 0.0} - zb * d3b);
            const cd den = (za * d1a + cd{(double)n, 0.0}) *
                           (cd{(double)n, 0.0} - za * d3a);
            q = (r2 * q) * num / den;
            // D3_n through Psi_n zeta_n, Yang Eqs. 18c-d
            pza = pza * ((nza - d1a_m) * (nza - d3a));
            pzb = pzb * ((nzb - d1b_m) * (nzb - d3b));
            d3a = d1a + mul_i(cinv(pza));
            d3b = d1b + mul_i(cinv(pzb));
            // H_a, H_b of the layer, Yang Eqs. 7b and 8b
            const cd ha{W(n, 0), W(n, 1)}, hb{W(n, 2), W(n, 3)};
            cd g1 = m[l] * ha - m[l - 1] * d1a;
            cd g2 = m[l] * ha - m[l - 1] * d3a;
            const cd han = (g2 * d1b - q * g1 * d3b) / (g2 - q * g1);
            g1 = m[l - 1] * hb - m[l] * d1a;
            g2 = m[l - 1] * hb - m[l] * d3a;
            const cd hbn = (g2 * d1b - q * g1 * d3b) / (g2 - q * g1);
            W(n, 0) = han.re; W(n, 1) = han.im;
            W(n, 2) = hbn.re; W(n, 3) = hbn.im;
        }
    }
    // outer boundary: D1(x_L) downward (real) in column 4
    const double xL = x[L - 1];
    {
        double r1 = 0.0;
        W(N, 4) = 0.0;
        for (int n = N; n >= 1; --n) {
            const double nx = n / xL;
            r1 = nx - 1.0 / (r1 + nx);
            W(n - 1, 4) = r1;
        }
    }
    double sx, cx;
    sincos(xL, &sx, &cx);
    double s2x, c2x;
    sincos(2.0 * xL, &s2x, &c2x);
    double psi = sx;
    cd zeta{sx, -cx};
    cd pz{0.5 * (1.0 - c2x), -0.5 * s2x};
    cd d3{0.0, 1.0};
    const cd mL = m[L - 1], imL = cinv(mL);
    for (int n = 1; n <= N; ++n) {
        const double nx = n / xL;
        const double w1 = nx - W(n - 1, 4);
        const cd w3 = cd{nx, 0.0} - d3;
        const double psi_m = psi;
        const cd zeta_m = zeta;
        psi = psi_m * w1;
        zeta = zeta_m * w3;
        pz = pz * (w1 * w3);
        d3 = cd{W(n, 4), 0.0} + mul_i(cinv(pz));
        const cd ha{W(n, 0), W(n, 1)}, hb{W(n, 2), W(n, 3)};
        const cd fa = ha * imL + cd{nx, 0.0};
        const cd a = (fa * cd{psi, 0.0} - cd{psi_m, 0.0}) / (fa * zeta - zeta_m);
        const cd fb = mL * hb + cd{nx, 0.0};
        const cd b = (fb * cd{psi, 0.0} - cd{psi_m, 0.0}) / (fb * zeta - zeta_m);
        if (coef) write_order(n, a, b, coef + LMZ_COEF_STRIDE * (n - 1));
        if (ab) {
            double *r = ab + 4 * (n - 1);
            r[0] = a.re; r[1] = a.im; r[2] = b.re; r[3] = b.im;
        }
    }
#undef W
    return N;
}

// ====================================================================
// Layered spheres: tangents of a_n, b_n for fitting
// ====================================================================
// Parametrization: the outer radius a_p scales every radius (fixed radius
// fractions r_l / a_p), and n_p + i k_p is the outer layer's index; inner
// indices are held.  a_n, b_n are holomorphic in the outer m_L, so the k_p
// tangent is i times the n_p one.  The tangents are forward-mode
// derivatives carried through the same recursion (value and tangent of
// every complex quantity in one direction per pass).
struct cdl { cd v, t; };                  // complex value and tangent
struct rdl { double v, t; };              // real value and tangent
LMZ_HD cdl operator+(cdl a, cdl b) { return {a.v + b.v, a.t + b.t}; }
LMZ_HD cdl operator-(cdl a, cdl b) { return {a.v - b.v, a.t - b.t}; }
LMZ_HD cdl operator*(cdl a, cdl b) { return {a.v * b.v, a.t * b.v + a.v * b.t}; }
LMZ_HD cdl operator*(double s, cdl a) { return {s * a.v, s * a.t}; }
LMZ_HD cdl operator*(rdl s, cdl a) {
    return {s.v * a.v, s.t * a.v + s.v * a.t};
}
LMZ_HD cdl operator/(cdl a, cdl b) {
    const cd q = a.v / b.v;
    return {q, (a.t - q * b.t) / b.v};
}
LMZ_HD cdl cinv(cdl b) {
    const cd q = cinv(b.v);
    return {q, -((q * q) * b.t)};
}
LMZ_HD cdl mul_i(cdl a) { return {mul_i(a.v), mul_i(a.t)}; }
LMZ_HD cdl cexp_i2(cdl z) {               // exp(2 i z), tangent 2 i exp(2 i z) z'
    const cd e = cexp_i2(z.v);
    return {e, mul_i(2.0 * (e * z.t))};
}
LMZ_HD cdl ndiv(int n, cdl z) {           // n / z, as the value recursion divides
    const cd q = cd{(double)n, 0.0} / z.v;
    return {q, -((q / z.v) * z.t)};
}

// One tangent direction (dir 0: a_p, 1: n_p) of the layered coefficients
// at n_max N (frozen at the value pass's), into the dcoef rows of orders
// 1..N: direction 0 at offset 0, direction 1 at 9 and the k_p rows at 18.
// Work: 16 doubles per order (value and tangent of H_a, H_b, and the two
// D1 chains of a layer).
LMZ_HD void mie_layered_tangent(const double *radius, const double *nl,
                                const double *kl, int L, double n_m,
                                double wavelength, int dir, int N,
                                double *work, long long wstride,
                                float *dcoef) {
    const double kum = 2.0 * M_PI / wavelength * n_m;
    const double aL = radius[L - 1];
    rdl x[LMZ_MAX_LAYERS];
    cdl m[LMZ_MAX_LAYERS];
    for (int l = 0; l < L; ++l) {
        x[l] = rdl{kum * radius[l], dir == 0 ? kum * radius[l] / aL : 0.0};
        m[l] = cdl{cd{nl[l] / n_m, kl[l] / n_m},
                   cd{(dir == 1 && l == L - 1) ? 1.0 / n_m : 0.0, 0.0}};
    }
    auto xm = [](rdl a, cdl b) { return a * b; };
#define W(n, j) work[((long long)(n) * LMZ_WORK_PER_ORDER + (j)) * wstride]
    auto put = [&](int n, int c, cdl z) {
        W(n, c) = z.v.re; W(n, c + 1) = z.v.im;
        W(n, c + 2) = z.t.re; W(n, c + 3) = z.t.im;
    };
    auto get = [&](int n, int c) {
        return cdl{cd{W(n, c), W(n, c + 1)}, cd{W(n, c + 2), W(n, c + 3)}};
    };
    auto d1_down = [&](cdl z, int c) {
        cdl d1{cd{0.0, 0.0}, cd{0.0, 0.0}};
        put(N, c, d1);
        for (int n = N; n >= 1; --n) {
            const cdl nz = ndiv(n, z);
            d1 = nz - cinv(d1 + nz);
            put(n - 1, c, d1);
        }
    };
    // core: H_a = H_b = D1(m_1 x_1) in columns 0..3 and 4..7
    d1_down(xm(x[0], m[0]), 0);
    for (int n = 0; n <= N; ++n) put(n, 4, get(n, 0));
    const cdl one{cd{1.0, 0.0}, cd{0.0, 0.0}};
    for (int l = 1; l < L; ++l) {
        const cdl za = xm(x[l - 1], m[l]), zb = xm(x[l], m[l]);
        d1_down(za, 8);
        d1_down(zb, 12);
        const cdl iza = cinv(za), izb = cinv(zb);
        const cdl ea = cexp_i2(za), eb = cexp_i2(zb);
        cdl pza = 0.5 * (one - ea), pzb = 0.5 * (one - eb);
        cdl d3a{cd{0.0, 1.0}, cd{0.0, 0.0}}, d3b = d3a;
        cdl q = cexp_i2(zb - za) * (ea - one) / (eb - one);
        const double rr = x[l - 1].v / x[l].v;
        const rdl r2{rr * rr, 2.0 * rr * (x[l - 1].t * x[l].v - x[l - 1].v * x[l].t) /
                                   (x[l].v * x[l].v)};
        for (int n = 1; n <= N; ++n) {
            const cdl d1a_m = get(n - 1, 8), d1b_m = get(n - 1, 12);
            const cdl d1a = get(n, 8), d1b = get(n, 12);
            const cdl nza = (double)n * iza, nzb = (double)n * izb;
            const cdl nn{cd{(double)n, 0.0}, cd{0.0, 0.0}};
            const cdl num = (zb * d1b + nn) * (nn - zb * d3b);
            const cdl den = (za * d1a + nn) * (nn - za * d3a);
            q = (r2 * q) * num / den;
            pza = pza * ((nza - d1a_m) * (nza - d3a));
            pzb = pzb * ((nzb - d1b_m) * (nzb - d3b));
            d3a = d1a + mul_i(cinv(pza));
            d3b = d1b + mul_i(cinv(pzb));
            const cdl ha = get(n, 0), hb = get(n, 4);
            cdl g1 = m[l] * ha - m[l - 1] * d1a;
            cdl g2 = m[l] * ha - m[l - 1] * d3a;
            const cdl han = (g2 * d1b - q * g1 * d3b) / (g2 - q * g1);
            g1 = m[l - 1] * hb - m[l] * d1a;
            g2 = m[l - 1] * hb - m[l] * d3a;
            const cdl hbn = (g2 * d1b - q * g1 * d3b) / (g2 - q * g1);
            put(n, 0, han);
            put(n, 4, hbn);
        }
    }
    // outer boundary: D1(x_L) downward (real) in columns 8, 9
    const rdl xL = x[L - 1];
    {
        rdl r1{0.0, 0.0};
        W(N, 8) = 0.0; W(N, 9) = 0.0;
        for (int n = N; n >= 1; --n) {
            const rdl nx{n / xL.v, -n * xL.t / (xL.v * xL.v)};
            const double den = r1.v + nx.v;
            r1 = rdl{nx.v - 1.0 / den, nx.t + (r1.t + nx.t) / (den * den)};
            W(n - 1, 8) = r1.v; W(n - 1, 9) = r1.t;
        }
    }
    double sx, cx;
    sincos(xL.v, &sx, &cx);
    double s2x, c2x;
    sincos(2.0 * xL.v, &s2x, &c2x);
    rdl psi{sx, cx * xL.t};
    cdl zeta{cd{sx, -cx}, cd{cx * xL.t, sx * xL.t}};           // -i e^{ix}
    cdl pz{cd{0.5 * (1.0 - c2x), -0.5 * s2x}, cd{s2x * xL.t, -c2x * xL.t}};
    cdl d3{cd{0.0, 1.0}, cd{0.0, 0.0}};
    const cdl mL = m[L - 1], imL = cinv(mL);
    for (int n = 1; n <= N; ++n) {
        const rdl nx{n / xL.v, -n * xL.t / (xL.v * xL.v)};
        const rdl w1{nx.v - W(n - 1, 8), nx.t - W(n - 1, 9)};
        const cdl w3 = cdl{cd{nx.v, 0.0}, cd{nx.t, 0.0}} - d3;
        const rdl psi_m = psi;
        const cdl zeta_m = zeta;
        psi = rdl{psi_m.v * w1.v, psi_m.t * w1.v + psi_m.v * w1.t};
        zeta = zeta_m * w3;
        pz = pz * (w1 * w3);
        d3 = cdl{cd{W(n, 8), 0.0}, cd{W(n, 9), 0.0}} + mul_i(cinv(pz));
        const cdl ha = get(n, 0), hb = get(n, 4);
        const cdl cnx{cd{nx.v, 0.0}, cd{nx.t, 0.0}};
        const cdl cpsi{cd{psi.v, 0.0}, cd{psi.t, 0.0}};
        const cdl cpsi_m{cd{psi_m.v, 0.0}, cd{psi_m.t, 0.0}};
        const cdl fa = ha * imL + cnx;
        const cdl a = (fa * cpsi - cpsi_m) / (fa * zeta - zeta_m);
        const cdl fb = mL * hb + cnx;
        const cdl b = (fb * cpsi - cpsi_m) / (fb * zeta - zeta_m);
        float *drow = dcoef + LMZ_DCOEF_STRIDE * (n - 1);
        if (dir == 0) {
            write_dorder(n, a.t, b.t, drow);
        } else {
            write_dorder(n, a.t, b.t, drow + 9);
            write_dorder_k(n, mul_i(a.t), mul_i(b.t), drow + 18);
        }
    }
#undef W
}

}  // namespace lmz
