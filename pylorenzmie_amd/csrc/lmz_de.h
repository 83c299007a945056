// The per-member logic of the batched differential-evolution search: the
// random draws, the Latin-hypercube start, scipy's best1bin trial with
// binomial crossover, and the deferred selection with its stopping test.
// __host__ __device__ so that the kernels (lmz_de.hip) and the host-side
// checker the tests build run the same code.
//
// The algorithm is scipy.optimize.differential_evolution's with
// strategy='best1bin', updating='deferred', init='latinhypercube',
// polish=False, which the reference's DEEstimator runs once per feature
// (analysis/DEEstimator.py:119-173; scipy's _differentialevolution.py
// DifferentialEvolutionSolver):
//   * population of npop members in the unit cube, parameter j mapped to
//     lo_j + u_j (hi_j - lo_j);
//   * each generation: one dither scale F ~ U[mutation_lo, mutation_hi];
//     for every member i, two distinct others r0, r1 (not i), the mutant
//     best + F (u_r0 - u_r1), binomial crossover with probability CR and
//     one forced dimension, coordinates outside [0, 1] redrawn uniformly;
//   * deferred selection: every trial is evaluated, then each replaces its
//     member where its energy is lower;
//   * stop when std(E) <= atol + tol |mean(E)| (population standard
//     deviation), or after maxiter generations.
// The random numbers are not scipy's stream: every draw is a counter-based
// hash of (seed, generation, search, member, draw), so the searches are
// reproducible for a seed and independent of launch shape and of each
// other.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

#ifndef LMZ_HD
#define LMZ_HD __host__ __device__ __forceinline__
#endif

namespace lmz {

#define LMZ_DE_MAXDIM 7
#define LMZ_DE_MAXPOP 1024

// SplitMix64's finalizer (Steele, Lea and Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014): a bijective 64-bit mix.
LMZ_HD uint64_t de_mix(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Draw indices of one member in one generation.
enum : uint32_t {
    DE_R0 = 0,         // first difference member
    DE_R1 = 1,         // second difference member
    DE_JRAND = 2,      // the dimension crossed over in any case
    DE_CROSS = 3,      // + j: crossover test of dimension j
    DE_REDRAW = 10,    // + j: redraw of dimension j outside [0, 1]
    DE_LHS_KEY = 20,   // + j: permutation key of the Latin hypercube
    DE_LHS_OFF = 30,   // + j: position inside the hypercube's segment
};
#define LMZ_DE_SEARCH_DRAW 0xffffffffu   // member index of per-search draws

// Uniform double in [0, 1) for (seed, generation, search, member, draw).
LMZ_HD double de_uniform(uint64_t seed, uint32_t gen, uint32_t search,
                         uint32_t member, uint32_t draw) {
    uint64_t h = de_mix(seed ^ de_mix(((uint64_t)gen << 32) | search));
    h = de_mix(h ^ (((uint64_t)member << 32) | draw));
    return (double)(h >> 11) * 0x1.0p-53;
}

// Integer uniform in [0, n).
LMZ_HD int de_below(double u, int n) {
    const int k = (int)(u * (double)n);
    return k < n ? k : n - 1;
}

// The dither scale of a search in a generation.
LMZ_HD double de_scale(uint64_t seed, uint32_t gen, uint32_t search,
                       double lo, double hi) {
    return lo + de_uniform(seed, gen, search, LMZ_DE_SEARCH_DRAW, 0) * (hi - lo);
}

// Member i's coordinate j in the Latin-hypercube start: its rank among
// the npop members by a random key, plus a uniform offset inside that
// segment, over npop (the distribution of scipy's _latinhypercube_init).
// keys: the npop members' keys of dimension j.
LMZ_HD double de_lhs(uint64_t seed, uint32_t search, int i, int j, int npop,
                     const double *keys) {
    const double ki = keys[i];
    int rank = 0;
    for (int k = 0; k < npop; ++k)
        rank += (keys[k] < ki || (keys[k] == ki && k < i)) ? 1 : 0;
    return ((double)rank + de_uniform(seed, 0, search, (uint32_t)i,
                                      DE_LHS_OFF + (uint32_t)j)) / (double)npop;
}

LMZ_HD double de_lhs_key(uint64_t seed, uint32_t search, int i, int j) {
    return de_uniform(seed, 0, search, (uint32_t)i, DE_LHS_KEY + (uint32_t)j);
}

// best1bin trial of member i (npop >= 3): u holds the search's members
// (npop x ndim, row-major), best the index of its lowest energy, F the
// generation's scale.  Writes ndim coordinates to trial.
LMZ_HD void de_trial(uint64_t seed, uint32_t gen, uint32_t search, int i,
                     int npop, int ndim, const double *u, int best, double F,
                     double CR, double *trial) {
    // two distinct members other than i (scipy's _select_samples)
    int r0 = de_below(de_uniform(seed, gen, search, i, DE_R0), npop - 1);
    r0 += (r0 >= i) ? 1 : 0;
    const int a = i < r0 ? i : r0, b = i < r0 ? r0 : i;
    int r1 = de_below(de_uniform(seed, gen, search, i, DE_R1), npop - 2);
    r1 += (r1 >= a) ? 1 : 0;
    r1 += (r1 >= b) ? 1 : 0;
    const int jrand = de_below(de_uniform(seed, gen, search, i, DE_JRAND), ndim);
    for (int j = 0; j < ndim; ++j) {
        const bool cross =
            j == jrand ||
            de_uniform(seed, gen, search, i, DE_CROSS + (uint32_t)j) < CR;
        double t = cross ? u[best * ndim + j] +
                               F * (u[r0 * ndim + j] - u[r1 * ndim + j])
                         : u[i * ndim + j];
        if (!(t >= 0.0 && t <= 1.0))          // scipy's _ensure_constraint
            t = de_uniform(seed, gen, search, i, DE_REDRAW + (uint32_t)j);
        trial[j] = t;
    }
}

// Energy of a trial for selection: non-finite chi2 counts as +inf.
LMZ_HD double de_energy(double chi2) {
    return (chi2 - chi2 == 0.0) ? chi2 : INFINITY;
}

// scipy's convergence test on the population energies E[0..n): population
// standard deviation <= atol + tol |mean|; false while any energy is
// infinite (numpy's std is then nan).
LMZ_HD bool de_converged(const double *E, int n, double tol, double atol) {
    double s = 0.0;
    for (int k = 0; k < n; ++k) {
        if (!(E[k] - E[k] == 0.0)) return false;
        s += E[k];
    }
    const double mean = s / n;
    double v = 0.0;
    for (int k = 0; k < n; ++k) v += (E[k] - mean) * (E[k] - mean);
    return sqrt(v / n) <= atol + tol * fabs(mean);
}

// Index of the lowest energy (the first of equal ones, as np.argmin).
LMZ_HD int de_argmin(const double *E, int n) {
    int best = 0;
    for (int k = 1; k < n; ++k)
        if (E[k] < E[best]) best = k;
    return best;
}

}  // namespace lmz
