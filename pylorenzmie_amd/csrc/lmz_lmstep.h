// The per-feature step logic of the batched Levenberg-Marquardt fitter:
// the normal-equation solve, the damping rules and the stopping tests, on
// one feature's 7 x 7 FP64 normal equations.  __host__ __device__ so that
// the LM kernels (lmz_lm.hip) and the host-side checker the tests build
// run the same code.
//
// Two damping rules (lmz_lm_settings.damping):
//  * LMZ_DAMPING_MINPACK restates the algorithm of MINPACK lmder/lmdif
//    (Moré, "The Levenberg-Marquardt algorithm: implementation and
//    theory", Lecture Notes in Mathematics 630 (1978); MINPACK-1, Argonne
//    ANL-80-74), which scipy.optimize.least_squares(method='lm') runs for
//    the reference's Optimizer.optimize (analysis/Optimizer.py:221) with
//    x_scale='jac' (scaling mode 1) and factor 100.  MINPACK works on a QR
//    factorization of J; here the same quantities come from a Cholesky
//    factorization of J^T J + par D^2 (||J p||^2 = p^T J^T J p, and the
//    lmpar Newton correction uses ||S^-T D^2 p|| with S^T S = J^T J +
//    par D^2, as qrsolv's S gives it).
//  * LMZ_DAMPING_NIELSEN: Marquardt damping with Nielsen's update.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

#include "../../include/lmz.h"

#ifndef LMZ_HD
#define LMZ_HD __host__ __device__ __forceinline__
#endif

namespace lmz {

struct LMState {
    double jtj[28];      // J^T J upper triangle (row major), J scaled by 1/noise
    double g[7];         // J^T r
    double dscale[7];    // D: running maximum of the column norms of J
    double chi2;         // r^T r at the current parameters
    double lambda;       // Nielsen: damping; MINPACK: Levenberg parameter par
    double nu;           // Nielsen: growth factor on rejection
    double pred;         // predicted chi2 reduction of the trial step
    double dnorm;        // ||D dp||
    double pnorm;        // ||D p||
    double delta;        // MINPACK: trust-region radius
    double jpn;          // MINPACK: ||J dp||^2 (chi2 units)
    double xnorm;        // MINPACK: ||D p|| at the last accepted point
    int32_t iter;        // trial steps taken
    int32_t status;
    int32_t need_j;      // the next step needs a fresh Jacobian
    int32_t njev;        // Jacobian evaluations
};

// finite and not NaN, in the same code for host and device
LMZ_HD bool lm_finite(double v) { return v - v == 0.0; }

LMZ_HD int lm_tri(int i, int j) {  // i <= j, row-major upper triangle
    return i * LMZ_NJAC - (i * (i - 1)) / 2 + (j - i);
}

LMZ_HD double lm_jtj(const LMState &S, int i, int j) {
    return (i <= j) ? S.jtj[lm_tri(i, j)] : S.jtj[lm_tri(j, i)];
}

LMZ_HD void lm_init(LMState &S, double lambda0) {
    for (int i = 0; i < 28; ++i) S.jtj[i] = 0.0;
    for (int i = 0; i < 7; ++i) { S.g[i] = 0.0; S.dscale[i] = 0.0; }
    S.chi2 = 0.0;
    S.lambda = lambda0;
    S.nu = 2.0;
    S.pred = S.dnorm = S.pnorm = S.delta = S.jpn = S.xnorm = 0.0;
    S.iter = 0;
    S.status = LMZ_LM_RUNNING;
    S.need_j = 1;
    S.njev = 0;
}

// The normal equations over all seven parameters with the held ones
// decoupled: their rows and columns of J^T J become the identity and their
// J^T r entries 0, so every solve gives them a zero step and the free
// parameters exactly the solution of the reduced system.  Fixed sizes keep
// every array in registers once the loops are unrolled.
struct LMSys {
    double A[7][7];   // J^T J (free block), identity elsewhere
    double g[7];      // J^T r, 0 for held parameters
    double d[7];      // D (1 for held parameters and zero columns)
};

LMZ_HD void lm_system(const LMState &S, const lmz_lm_settings &set,
                      LMSys &Y) {
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        const bool fi = set.vary[i] != 0;
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            const bool fj = set.vary[j] != 0;
            Y.A[i][j] = (fi && fj) ? lm_jtj(S, i, j) : (i == j ? 1.0 : 0.0);
        }
        Y.g[i] = fi ? S.g[i] : 0.0;
        Y.d[i] = (fi && S.dscale[i] > 0.0) ? S.dscale[i] : 1.0;
    }
}

// Cholesky factor L (lower) of A + par diag(d^2); false if it is not
// positive definite in floating point.
LMZ_HD bool lm_cholesky(const LMSys &Y, double par, double (&L)[7][7]) {
#pragma unroll
    for (int i = 0; i < 7; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) L[i][j] = Y.A[i][j];
#pragma unroll
    for (int i = 0; i < 7; ++i) L[i][i] += par * (Y.d[i] * Y.d[i]);
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        double s = L[j][j];
#pragma unroll
        for (int k = 0; k < j; ++k) s -= L[j][k] * L[j][k];
        ok = ok && (s > 0.0) && lm_finite(s);
        const double ljj = sqrt(fmax(s, 0.0));
        L[j][j] = ljj;
#pragma unroll
        for (int i = j + 1; i < 7; ++i) {
            double t = L[i][j];
#pragma unroll
            for (int k = 0; k < j; ++k) t -= L[i][k] * L[j][k];
            L[i][j] = t / ljj;
        }
    }
    return ok;
}

// y = L^-1 b
LMZ_HD void lm_lower_solve(const double (&L)[7][7], const double (&b)[7],
                           double (&y)[7]) {
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        double t = b[i];
#pragma unroll
        for (int k = 0; k < i; ++k) t -= L[i][k] * y[k];
        y[i] = t / L[i][i];
    }
}

// x = -(L L^T)^-1 g
LMZ_HD void lm_step(const double (&L)[7][7], const LMSys &Y, double (&x)[7]) {
    double b[7], y[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) b[i] = -Y.g[i];
    lm_lower_solve(L, b, y);
#pragma unroll
    for (int i = 6; i >= 0; --i) {
        double t = y[i];
#pragma unroll
        for (int k = i + 1; k < 7; ++k) t -= L[k][i] * x[k];
        x[i] = t / L[i][i];
    }
}

LMZ_HD double lm_dnorm(const LMSys &Y, const double (&x)[7]) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        const double t = Y.d[i] * x[i];
        s += t * t;
    }
    return sqrt(s);
}

// ||L^-1 D^2 x / ||D x|| ||: the lmpar Newton correction's denominator
LMZ_HD double lm_newton_norm(const double (&L)[7][7], const LMSys &Y,
                             const double (&x)[7], double dxnorm) {
    double w[7], y[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) w[i] = Y.d[i] * Y.d[i] * x[i] / dxnorm;
    lm_lower_solve(L, w, y);
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 7; ++i) s += y[i] * y[i];
    return sqrt(s);
}

// MINPACK lmpar on the normal equations: the Levenberg parameter par >= 0
// and step x = -(J^T J + par D^2)^-1 J^T r with either par = 0 and
// ||D x|| <= 1.1 delta, or | ||D x|| - delta | <= 0.1 delta (at most 10
// Newton iterations on ||D x(par)|| = delta between the bounds parl, paru).
// par holds the previous parameter on entry (the starting guess).  Returns
// false if no factorization succeeded.
LMZ_HD bool lm_lmpar(const LMSys &Y, double delta, double &par,
                     double (&x)[7]) {
    const double dwarf = 2.2250738585072014e-308;
    double L[7][7];
    // Gauss-Newton direction
    const bool nonsing = lm_cholesky(Y, 0.0, L);
    double dxnorm = INFINITY, fp = INFINITY;
    if (nonsing) {
        lm_step(L, Y, x);
        dxnorm = lm_dnorm(Y, x);
        fp = dxnorm - delta;
        if (fp <= 0.1 * delta) {
            par = 0.0;
            return true;
        }
    }
    // lower bound from the Newton step at par = 0 (full rank only)
    double parl = 0.0;
    if (nonsing) {
        const double t = lm_newton_norm(L, Y, x, dxnorm);
        parl = ((fp / delta) / t) / t;
    }
    // upper bound: ||D^-1 J^T r|| / delta
    double gn = 0.0;
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        const double t = Y.g[i] / Y.d[i];
        gn += t * t;
    }
    gn = sqrt(gn);
    double paru = gn / delta;
    if (paru == 0.0) paru = dwarf / fmin(delta, 0.1);
    par = fmax(par, parl);
    par = fmin(par, paru);
    if (par == 0.0) par = nonsing ? gn / dxnorm : paru;
    bool ok = false;
    for (int iter = 1; iter <= 10; ++iter) {
        if (par == 0.0) par = fmax(dwarf, 0.001 * paru);
        int grow = 0;
        while (!lm_cholesky(Y, par, L)) {
            // J^T J + par D^2 is positive definite for par > 0; only
            // rounding at a tiny par can defeat the factorization
            if (++grow > 30) return ok;
            par = fmax(par * 10.0, 1e-300);
        }
        lm_step(L, Y, x);
        ok = true;
        dxnorm = lm_dnorm(Y, x);
        const double temp = fp;
        fp = dxnorm - delta;
        if (fabs(fp) <= 0.1 * delta || (parl == 0.0 && fp <= temp && temp < 0.0) ||
            iter == 10)
            break;
        const double t = lm_newton_norm(L, Y, x, dxnorm);
        const double parc = ((fp / delta) / t) / t;
        if (fp > 0.0) parl = fmax(parl, par);
        if (fp < 0.0) paru = fmin(paru, par);
        par = fmax(parl, par + parc);
    }
    return ok;
}

// Take a fresh Jacobian's reduction red[36] (J^T J, J^T r, r^T r unused:
// chi2 comes from the FP64 forward pass) into the state: update D, and on
// the first Jacobian set MINPACK's initial radius 100 ||D p||.
LMZ_HD void lm_take_jacobian(LMState &S, const double *red,
                             const lmz_lm_settings &set, const double *p) {
    for (int i = 0; i < 28; ++i) S.jtj[i] = red[i];
    for (int i = 0; i < 7; ++i) S.g[i] = red[28 + i];
    const bool first = S.njev == 0;
    S.njev += 1;
    for (int i = 0; i < 7; ++i) {
        const double cn = sqrt(fmax(S.jtj[lm_tri(i, i)], 0.0));
        if (set.damping == LMZ_DAMPING_MINPACK && first)
            S.dscale[i] = cn == 0.0 ? 1.0 : cn;
        else
            S.dscale[i] = fmax(S.dscale[i], cn);
    }
    if (set.damping == LMZ_DAMPING_MINPACK && first) {
        double xs = 0.0;
#pragma unroll
        for (int i = 0; i < 7; ++i) {
            const double t = set.vary[i] ? S.dscale[i] * p[i] : 0.0;
            xs += t * t;
        }
        S.xnorm = sqrt(xs);
        S.delta = S.xnorm > 0.0 ? 100.0 * S.xnorm : 100.0;
        S.lambda = 0.0;
    }
}

// MINPACK's gradient test: max_j |(J^T r)_j| / (||J_j|| ||r||) <= gtol
// over the free parameters with nonzero columns.
LMZ_HD bool lm_gtol_met(const LMState &S, const lmz_lm_settings &set) {
    double gmax = 0.0;
    const double rn = sqrt(fmax(S.chi2, 0.0));
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        const double jn = sqrt(fmax(S.jtj[lm_tri(i, i)], 0.0));
        if (set.vary[i] && jn > 0.0 && rn > 0.0)
            gmax = fmax(gmax, fabs(S.g[i]) / (jn * rn));
    }
    return gmax <= set.gtol;
}

// The trial step from the current normal equations: pt = p + dp (rows of
// LMZ_NPARAM), with S.pred, S.dnorm, S.pnorm (and for MINPACK S.jpn,
// S.delta on the first outer iteration) set.  Returns false when no step
// can be formed (the feature stalls).
LMZ_HD bool lm_solve(LMState &S, const lmz_lm_settings &set, const double *p,
                     double *pt) {
    LMSys Y;
    lm_system(S, set, Y);
    double x[7] = {0, 0, 0, 0, 0, 0, 0};
    if (set.damping == LMZ_DAMPING_MINPACK) {
        double par = S.lambda;
        if (!lm_lmpar(Y, S.delta, par, x)) return false;
        S.lambda = par;
    } else {
        bool ok = false;
        for (int attempt = 0; attempt < 12 && !ok; ++attempt) {
            double L[7][7];
            ok = lm_cholesky(Y, S.lambda, L);
            if (!ok) { S.lambda *= 10.0; continue; }
            lm_step(L, Y, x);
        }
        if (!ok) return false;
    }
    // predicted reduction of chi2 under the linear model:
    // -2 g.dp - dp.H.dp = dp.H.dp + 2 par ||D dp||^2 for the damped step
    double gd = 0.0, dhd = 0.0, dn = 0.0, pn = 0.0;
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        gd += Y.g[i] * x[i];
        double hd = 0.0;
#pragma unroll
        for (int j = 0; j < 7; ++j) hd += Y.A[i][j] * x[j];
        dhd += x[i] * hd;
        const double d = set.vary[i] ? S.dscale[i] : 0.0;
        dn += d * d * x[i] * x[i];
        pn += d * d * p[i] * p[i];
    }
    S.jpn = dhd;
    S.dnorm = sqrt(dn);
    S.pnorm = sqrt(pn);
    // MINPACK forms it as ||J dp||^2 + 2 par ||D dp||^2
    S.pred = set.damping == LMZ_DAMPING_MINPACK ? dhd + 2.0 * S.lambda * dn
                                                : -2.0 * gd - dhd;
    // MINPACK: on the first outer iteration the radius shrinks to the step
    if (set.damping == LMZ_DAMPING_MINPACK && S.njev == 1)
        S.delta = fmin(S.delta, S.dnorm);
#pragma unroll
    for (int i = 0; i < LMZ_NPARAM; ++i) pt[i] = p[i] + (i < 7 ? x[i] : 0.0);
    return true;
}

// Judge the trial whose chi2 is chi2t: returns true if it is accepted (the
// caller then moves to pt).  Updates damping, status and need_j.
LMZ_HD bool lm_accept(LMState &S, const lmz_lm_settings &set, double chi2t,
                      const double *pt) {
    const double chi2 = S.chi2;
    bool accept = false;
    S.iter += 1;
    if (set.damping == LMZ_DAMPING_MINPACK) {
        const double eps = 2.220446049250313e-16;
        // relative reductions (MINPACK's actred, prered, dirder in units of
        // fnorm^2 = chi2)
        const bool finite = lm_finite(chi2t);
        const double actred = (finite && chi2t < 100.0 * chi2) ? 1.0 - chi2t / chi2
                                                                : -1.0;
        const double par = S.lambda;
        const double prered = S.pred / chi2;
        const double dirder = -(S.jpn + par * S.dnorm * S.dnorm) / chi2;
        const double ratio = prered != 0.0 ? actred / prered : 0.0;
        if (ratio <= 0.25) {
            double temp = actred >= 0.0 ? 0.5 : 0.5 * dirder / (dirder + 0.5 * actred);
            if (!(chi2t < 100.0 * chi2) || temp < 0.1) temp = 0.1;
            S.delta = temp * fmin(S.delta, S.dnorm / 0.1);
            S.lambda = par / temp;
        } else if (par == 0.0 || ratio >= 0.75) {
            S.delta = S.dnorm / 0.5;
            S.lambda = 0.5 * par;
        }
        accept = ratio >= 1e-4;
        if (accept) {
            S.chi2 = chi2t;
            double xs = 0.0;
#pragma unroll
            for (int i = 0; i < 7; ++i) {
                const double t = set.vary[i] ? S.dscale[i] * pt[i] : 0.0;
                xs += t * t;
            }
            S.xnorm = sqrt(xs);
        }
        const double xnorm = S.xnorm;
        S.need_j = accept ? 1 : 0;
        if (fabs(actred) <= set.ftol && prered <= set.ftol && 0.5 * ratio <= 1.0)
            S.status |= LMZ_LM_FTOL;
        if (S.delta <= set.xtol * xnorm) S.status |= LMZ_LM_XTOL;
        // MINPACK's stringent tests (info 6 and 7): nothing more can be
        // gained at machine precision
        if (S.status == LMZ_LM_RUNNING) {
            if (fabs(actred) <= eps && prered <= eps && 0.5 * ratio <= 1.0)
                S.status |= LMZ_LM_FTOL;
            if (S.delta <= eps * xnorm) S.status |= LMZ_LM_XTOL;
        }
    } else {
        const double act = chi2 - chi2t;
        if (lm_finite(chi2t) && act > 0.0) {
            accept = true;
            const double rho = (S.pred > 0.0) ? act / S.pred : 0.0;
            S.chi2 = chi2t;
            const double t = 2.0 * rho - 1.0;
            S.lambda *= fmax(1.0 / 3.0, 1.0 - t * t * t);
            S.nu = 2.0;
            S.need_j = 1;
            if (act <= set.ftol * chi2 && S.pred <= set.ftol * chi2)
                S.status |= LMZ_LM_FTOL;
            if (S.dnorm <= set.xtol * (S.pnorm + set.xtol)) S.status |= LMZ_LM_XTOL;
        } else {
            S.lambda *= S.nu;
            S.nu *= 2.0;
            S.need_j = 0;
            // nothing left to gain at this resolution of chi2
            if (lm_finite(chi2t) && S.pred <= set.ftol * chi2) S.status |= LMZ_LM_FTOL;
            if (S.dnorm <= set.xtol * (S.pnorm + set.xtol)) S.status |= LMZ_LM_XTOL;
            if (S.lambda > 1e16) S.status |= LMZ_LM_STALLED;
        }
    }
    if (S.status == LMZ_LM_RUNNING && S.iter >= set.max_iter)
        S.status |= LMZ_LM_MAXITER;
    return accept;
}

}  // namespace lmz
