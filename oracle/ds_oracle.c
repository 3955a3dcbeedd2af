/*
 * ds_oracle.c -- FP64 CPU restatement of the reference direct-summation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: only the
 * test-suite, the smoke check and the CPU-baseline leg of the benchmark may
 * load it.  The product library (libnbody_mi355x.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the AstroNbodySim.jl repository root).  Arithmetic follows the reference's
 * evaluation order literally (left-to-right, no contraction: the build uses
 * -ffp-contract=off), including its decimal literals.
 *
 * Pins (what checks each function independently of this text) are listed in
 * the header comment of each function; functions without an independent pin
 * would say "parity unpinned" -- none do.
 *
 * Deviations from the reference that the oracle records on purpose:
 *   - h == 0 with r == 0 (the self pair when softening is zero) makes the
 *     reference compute 0*Inf = NaN (gravity.jl:16-17,27).  The pair
 *     functions below keep that behaviour; the j-sums take a `skip` index so
 *     a caller can drop the self pair, which equals the reference for h > 0.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define DS_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* a1: compute_force_from_particle  src/directsum/gravity.jl:8-28           */
/* Pins: Newtonian closed form for r >= h; continuity at u=0.5 and u=1;      */
/* antisymmetry; mpmath brute force; StaticTest binary closed form.          */
/* ------------------------------------------------------------------------ */
static inline void pair_force(const double p[3], double sx, double sy, double sz,
                              double mass, double G, double h, double out[3]) {
    double dx = sx - p[0], dy = sy - p[1], dz = sz - p[2];   /* dp = sourcePos - pos */
    double r2 = dx * dx + dy * dy + dz * dz;                   /* dp * dp */
    double r = sqrt(r2);
    double h_inv = 1.0 / h;
    double h3_inv = h_inv * h_inv * h_inv;                     /* h_inv^3 (literal_pow) */
    double fac;
    if (r >= h) {
        fac = mass / (r2 * r);
    } else {
        double u = r * h_inv;
        if (u < 0.5) {
            fac = mass * h3_inv * (10.666666666667 + u * u * (32.0 * u - 38.4));
        } else {
            fac = mass * h3_inv * (21.333333333333 - 48.0 * u + 38.4 * (u * u)
                                   - 10.666666666667 * (u * u * u) - 0.066666666667 / (u * u * u));
        }
    }
    double g = 1.0 * G;
    out[0] = (dx * fac) * g;   /* dp * fac * (1.0 * G) */
    out[1] = (dy * fac) * g;
    out[2] = (dz * fac) * g;
}

/* ------------------------------------------------------------------------ */
/* a2: compute_unit_potential_from_particle  src/directsum/potential.jl:2-27 */
/* Pins: -Gm/r for r >= h; continuity at u=0.5, u=1; StaticTest binary W.    */
/* ------------------------------------------------------------------------ */
static inline double pair_potential(const double p[3], double sx, double sy, double sz,
                                    double mass, double G, double h) {
    double dx = sx - p[0], dy = sy - p[1], dz = sz - p[2];
    double r2 = dx * dx + dy * dy + dz * dz;
    double r = sqrt(r2);
    double h_inv = 1.0 / h;
    double pot;
    if (r >= h) {
        pot = (-mass) / r * (1.0 * G);
    } else {
        double u = r * h_inv;
        double wp;
        if (u < 0.5) {
            wp = -2.8 + u * u * (5.333333333333 + u * u * (6.4 * u - 9.6));
        } else {
            wp = -3.2 + 0.066666666667 / u
                 + u * u * (10.666666666667 + u * (-16.0 + u * (9.6 - 2.133333333333 * u)));
        }
        pot = mass * h_inv * wp * (1.0 * G);
    }
    if (r == 0.0) pot = pot * 0.0;   /* iszero(r) */
    return pot;
}

DS_EXPORT void ds_pair_force(const double p[3], const double s[3], double mass, double G,
                             double h, double out[3]) {
    pair_force(p, s[0], s[1], s[2], mass, G, h, out);
}

DS_EXPORT double ds_pair_potential(const double p[3], const double s[3], double mass,
                                   double G, double h) {
    return pair_potential(p, s[0], s[1], s[2], mass, G, h);
}

/* ------------------------------------------------------------------------ */
/* Julia Base.sum over a vector: pairwise reduction (mapreduce_impl with     */
/* pairwise_blocksize 1024), used by sum(acc) in gravity.jl:45 and           */
/* potential.jl:40.  Rounding-level only; pinned against the long-double     */
/* variant to ~1e-15.                                                        */
/* ------------------------------------------------------------------------ */
static void pairwise_sum3(const double *a, int64_t lo, int64_t hi, double out[3]) {
    if (lo == hi) { out[0] = a[3 * lo]; out[1] = a[3 * lo + 1]; out[2] = a[3 * lo + 2]; return; }
    if (hi - lo < 1024) {
        double s0 = a[3 * lo] + a[3 * (lo + 1)];
        double s1 = a[3 * lo + 1] + a[3 * (lo + 1) + 1];
        double s2 = a[3 * lo + 2] + a[3 * (lo + 1) + 2];
        for (int64_t i = lo + 2; i <= hi; ++i) { s0 += a[3 * i]; s1 += a[3 * i + 1]; s2 += a[3 * i + 2]; }
        out[0] = s0; out[1] = s1; out[2] = s2;
        return;
    }
    int64_t mid = lo + ((hi - lo) >> 1);
    double l[3], r[3];
    pairwise_sum3(a, lo, mid, l);
    pairwise_sum3(a, mid + 1, hi, r);
    out[0] = l[0] + r[0]; out[1] = l[1] + r[1]; out[2] = l[2] + r[2];
}

static double pairwise_sum1(const double *a, int64_t lo, int64_t hi) {
    if (lo == hi) return a[lo];
    if (hi - lo < 1024) {
        double s = a[lo] + a[lo + 1];
        for (int64_t i = lo + 2; i <= hi; ++i) s += a[i];
        return s;
    }
    int64_t mid = lo + ((hi - lo) >> 1);
    return pairwise_sum1(a, lo, mid) + pairwise_sum1(a, mid + 1, hi);
}

/* Variants of the j-sum.  0 = faithful (per-target scratch array + pairwise
 * sum, like gravity.jl:40-45), 1 = tuned (register accumulation),
 * 2 = long-double pair math and accumulation (bounds the oracle's own error),
 * 3 = vectorised (SIMD lanes accumulate interleaved partial sums; same pair
 *     formula, different summation order -- the CPU timing baseline). */
enum { DS_FAITHFUL = 0, DS_TUNED = 1, DS_LONGDOUBLE = 2, DS_SIMD = 3 };

/* a1 written branch-free for SIMD lanes; same formula and literals. */
#pragma omp declare simd notinbranch
static inline double pair_fac_simd(double r2, double mass, double h, double h_inv, double h3_inv) {
    double r = sqrt(r2);
    double u = r * h_inv;
    double u2 = u * u, u3 = u2 * u;
    double fn = mass / (r2 * r);
    double f1 = mass * h3_inv * (10.666666666667 + u2 * (32.0 * u - 38.4));
    double f2 = mass * h3_inv * (21.333333333333 - 48.0 * u + 38.4 * u2 - 10.666666666667 * u3
                                 - 0.066666666667 / u3);
    return r >= h ? fn : (u < 0.5 ? f1 : f2);
}

static void pair_force_ld(const double p[3], double sx, double sy, double sz, double mass,
                          double G, double h, long double out[3]) {
    long double dx = (long double)sx - p[0], dy = (long double)sy - p[1], dz = (long double)sz - p[2];
    long double r2 = dx * dx + dy * dy + dz * dz;
    long double r = sqrtl(r2);
    long double h_inv = 1.0L / h;
    long double h3_inv = h_inv * h_inv * h_inv;
    long double fac;
    if (r >= h) {
        fac = mass / (r2 * r);
    } else {
        long double u = r * h_inv;
        if (u < 0.5L) fac = mass * h3_inv * (10.666666666667L + u * u * (32.0L * u - 38.4L));
        else fac = mass * h3_inv * (21.333333333333L - 48.0L * u + 38.4L * u * u
                                    - 10.666666666667L * u * u * u - 0.066666666667L / (u * u * u));
    }
    out[0] = dx * fac * G; out[1] = dy * fac * G; out[2] = dz * fac * G;
}

static long double pair_potential_ld(const double p[3], double sx, double sy, double sz,
                                     double mass, double G, double h) {
    long double dx = (long double)sx - p[0], dy = (long double)sy - p[1], dz = (long double)sz - p[2];
    long double r2 = dx * dx + dy * dy + dz * dz;
    long double r = sqrtl(r2);
    long double h_inv = 1.0L / h;
    long double pot;
    if (r >= h) pot = -(long double)mass / r * G;
    else {
        long double u = r * h_inv, wp;
        if (u < 0.5L) wp = -2.8L + u * u * (5.333333333333L + u * u * (6.4L * u - 9.6L));
        else wp = -3.2L + 0.066666666667L / u
                  + u * u * (10.666666666667L + u * (-16.0L + u * (9.6L - 2.133333333333L * u)));
        pot = mass * h_inv * wp * G;
    }
    if (r == 0.0L) pot = pot * 0.0L;
    return pot;
}

/* ------------------------------------------------------------------------ */
/* a3: compute_force_at_point(pos, data, G, h)  src/directsum/gravity.jl:37-46 */
/* Sum over ALL sources (self included).  `skip` >= 0 drops that source      */
/* index (used only for h == 0, where the reference yields NaN).             */
/* Pins: StaticTest binary |a| closed form, momentum conservation, mpmath.   */
/* ------------------------------------------------------------------------ */
static void force_at_point(const double p[3], int64_t n, const double *x, const double *y,
                           const double *z, const double *m, double G, double h, int64_t skip,
                           int variant, double *scratch, double out[3]) {
    if (n == 0) { out[0] = out[1] = out[2] = 0.0; return; }
    if (variant == DS_FAITHFUL) {
        for (int64_t j = 0; j < n; ++j) {
            if (j == skip) { scratch[3 * j] = scratch[3 * j + 1] = scratch[3 * j + 2] = 0.0; continue; }
            pair_force(p, x[j], y[j], z[j], m[j], G, h, scratch + 3 * j);
        }
        pairwise_sum3(scratch, 0, n - 1, out);
    } else if (variant == DS_TUNED) {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, t[3];
        for (int64_t j = 0; j < n; ++j) {
            if (j == skip) continue;
            pair_force(p, x[j], y[j], z[j], m[j], G, h, t);
            a0 += t[0]; a1 += t[1]; a2 += t[2];
        }
        out[0] = a0; out[1] = a1; out[2] = a2;
    } else if (variant == DS_SIMD) {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        const double h_inv = 1.0 / h, h3_inv = h_inv * h_inv * h_inv, g = 1.0 * G;
        const double px = p[0], py = p[1], pz = p[2];
#pragma omp simd reduction(+ : a0, a1, a2)
        for (int64_t j = 0; j < n; ++j) {
            double dx = x[j] - px, dy = y[j] - py, dz = z[j] - pz;
            double r2 = dx * dx + dy * dy + dz * dz;
            double fac = (j == skip) ? 0.0 : pair_fac_simd(r2, m[j], h, h_inv, h3_inv);
            a0 += dx * fac * g; a1 += dy * fac * g; a2 += dz * fac * g;
        }
        out[0] = a0; out[1] = a1; out[2] = a2;
    } else {
        long double a0 = 0.0L, a1 = 0.0L, a2 = 0.0L, t[3];
        for (int64_t j = 0; j < n; ++j) {
            if (j == skip) continue;
            pair_force_ld(p, x[j], y[j], z[j], m[j], G, h, t);
            a0 += t[0]; a1 += t[1]; a2 += t[2];
        }
        out[0] = (double)a0; out[1] = (double)a1; out[2] = (double)a2;
    }
}

/* a4: compute_unit_potential_at_point(pos, data, G, h)  src/directsum/potential.jl:33-41 */
static double potential_at_point(const double p[3], int64_t n, const double *x, const double *y,
                                 const double *z, const double *m, double G, double h,
                                 int64_t skip, int variant, double *scratch) {
    if (n == 0) return 0.0;
    if (variant == DS_FAITHFUL) {
        for (int64_t j = 0; j < n; ++j)
            scratch[j] = (j == skip) ? 0.0 : pair_potential(p, x[j], y[j], z[j], m[j], G, h);
        return pairwise_sum1(scratch, 0, n - 1);
    } else if (variant == DS_TUNED) {
        double s = 0.0;
        for (int64_t j = 0; j < n; ++j)
            if (j != skip) s += pair_potential(p, x[j], y[j], z[j], m[j], G, h);
        return s;
    } else {
        long double s = 0.0L;
        for (int64_t j = 0; j < n; ++j)
            if (j != skip) s += pair_potential_ld(p, x[j], y[j], z[j], m[j], G, h);
        return (double)s;
    }
}

static int set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    return nthreads;
#else
    (void)nthreads;
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* a5: compute_local_force + compute_force(sim, DirectSum, CPU)              */
/*     src/directsum/gravity.jl:80-101,160-169                               */
/* Targets [t0, t1) of the n particles; active iff active==NULL or           */
/* active[i] != 0 (Ti_endstep[i] == system_time_int, gravity.jl:95-96);      */
/* h = soft[coll[i]-1] (target-side ForceSofteningTable[Int(Collection)],    */
/* gravity.jl:97).  skip_self drops j == i (only meaningful for h == 0).     */
/* Threads split targets statically (split_block, gravity.jl:92-93).         */
/* Inactive targets keep whatever acc holds.                                 */
/* ------------------------------------------------------------------------ */
DS_EXPORT int ds_compute_force(int64_t n, const double *x, const double *y, const double *z,
                               const double *m, const uint8_t *coll, const double soft[6],
                               double G, const uint8_t *active, int64_t t0, int64_t t1,
                               int skip_self, int variant, int nthreads, double *acc /* 3n */) {
    if (t0 < 0 || t1 > n || t0 > t1) return -1;
    for (int64_t i = t0; i < t1; ++i) if (coll[i] < 1 || coll[i] > 6) return -2;
    int nt = set_threads(nthreads);
    int err = 0;
#pragma omp parallel num_threads(nt)
    {
        double *scratch = NULL;
        if (variant == DS_FAITHFUL) {
            scratch = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
            if (!scratch) {
#pragma omp atomic write
                err = -3;
            }
        }
#pragma omp for schedule(static)
        for (int64_t i = t0; i < t1; ++i) {
            if (err) continue;
            if (active && !active[i]) continue;
            double p[3] = {x[i], y[i], z[i]};
            double h = soft[coll[i] - 1];
            force_at_point(p, n, x, y, z, m, G, h, skip_self ? i : -1, variant, scratch, acc + 3 * i);
        }
        free(scratch);
    }
    return err;
}

/* ------------------------------------------------------------------------ */
/* a6: compute_local_potential / compute_potential(sim, DirectSum, CPU)      */
/*     src/directsum/potential.jl:67-79,142-148 (no active filter).          */
/* ------------------------------------------------------------------------ */
DS_EXPORT int ds_compute_potential(int64_t n, const double *x, const double *y, const double *z,
                                   const double *m, const uint8_t *coll, const double soft[6],
                                   double G, int64_t t0, int64_t t1, int skip_self, int variant,
                                   int nthreads, double *pot) {
    if (t0 < 0 || t1 > n || t0 > t1) return -1;
    for (int64_t i = t0; i < t1; ++i) if (coll[i] < 1 || coll[i] > 6) return -2;
    int nt = set_threads(nthreads);
    int err = 0;
#pragma omp parallel num_threads(nt)
    {
        double *scratch = NULL;
        if (variant == DS_FAITHFUL) {
            scratch = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
            if (!scratch) {
#pragma omp atomic write
                err = -3;
            }
        }
#pragma omp for schedule(static)
        for (int64_t i = t0; i < t1; ++i) {
            if (err) continue;
            double p[3] = {x[i], y[i], z[i]};
            double h = soft[coll[i] - 1];
            pot[i] = potential_at_point(p, n, x, y, z, m, G, h, skip_self ? i : -1, variant, scratch);
        }
        free(scratch);
    }
    return err;
}

/* a7: remove_self_potential_local  src/base/Potential.jl:5-14 -- the public  */
/* compute_potential(sim) adds G*m_i/(h_i/2.8) after the sum (Potential.jl:20-24). */
DS_EXPORT void ds_remove_self_potential(int64_t n, const double *m, const uint8_t *coll,
                                        const double soft[6], double G, double *pot) {
    for (int64_t i = 0; i < n; ++i) pot[i] += G * m[i] / (soft[coll[i] - 1] / 2.8);
}

/* ------------------------------------------------------------------------ */
/* a8: probe points with a caller-given h                                    */
/*     compute_force(sim, pos, h, DirectSum, CPU)  src/directsum/gravity.jl:171-194 */
/*     compute_potential(sim, pos, h, ...)         src/directsum/potential.jl:155-180 */
/* ------------------------------------------------------------------------ */
DS_EXPORT int ds_force_at_points(int64_t npts, const double *pts /* 3*npts */, double h, int64_t n,
                                 const double *x, const double *y, const double *z, const double *m,
                                 double G, int variant, int nthreads, double *acc /* 3*npts */) {
    int nt = set_threads(nthreads);
    int err = 0;
#pragma omp parallel num_threads(nt)
    {
        double *scratch = NULL;
        if (variant == DS_FAITHFUL) {
            scratch = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
            if (!scratch) {
#pragma omp atomic write
                err = -3;
            }
        }
#pragma omp for schedule(static)
        for (int64_t k = 0; k < npts; ++k) {
            if (err) continue;
            force_at_point(pts + 3 * k, n, x, y, z, m, G, h, -1, variant, scratch, acc + 3 * k);
        }
        free(scratch);
    }
    return err;
}

DS_EXPORT int ds_potential_at_points(int64_t npts, const double *pts, double h, int64_t n,
                                     const double *x, const double *y, const double *z,
                                     const double *m, double G, int variant, int nthreads,
                                     double *pot) {
    int nt = set_threads(nthreads);
    int err = 0;
#pragma omp parallel num_threads(nt)
    {
        double *scratch = NULL;
        if (variant == DS_FAITHFUL) {
            scratch = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
            if (!scratch) {
#pragma omp atomic write
                err = -3;
            }
        }
#pragma omp for schedule(static)
        for (int64_t k = 0; k < npts; ++k) {
            if (err) continue;
            pot[k] = potential_at_point(pts + 3 * k, n, x, y, z, m, G, h, -1, variant, scratch);
        }
        free(scratch);
    }
    return err;
}

/* ------------------------------------------------------------------------ */
/* a10: analytic point-mass background field (a central body as a field),    */
/* as used through apply_background_force  src/base/PostProcessing/Force.jl:7-20 */
/* with the closure of examples/01-binary/01-binary.jl:148-149:              */
/* acc += -G M (x - c) / |x - c|^3 ; potential -G M / |x - c|.               */
/* ------------------------------------------------------------------------ */
DS_EXPORT void ds_background_point_mass(int64_t n, const double *x, const double *y,
                                        const double *z, double G, double M, const double c[3],
                                        double *acc, double *pot) {
    for (int64_t i = 0; i < n; ++i) {
        double dx = x[i] - c[0], dy = y[i] - c[1], dz = z[i] - c[2];
        double r2 = dx * dx + dy * dy + dz * dz;
        double r = sqrt(r2);
        if (acc) {
            double f = -G * M / (r2 * r);
            acc[3 * i] += f * dx; acc[3 * i + 1] += f * dy; acc[3 * i + 2] += f * dz;
        }
        if (pot) pot[i] += -G * M / r;
    }
}

/* ------------------------------------------------------------------------ */
/* a11: constant-dt KDK pieces                                               */
/*   leapfrog_half_kick_local (const)  src/base/Timestep.jl:116-125          */
/*     Vel += Acc * forward * 0.5 * dt                                       */
/*   advance_and_find_timestep_local (const)  src/base/Timestep.jl:84-93     */
/*     Vel += Acc * dt                                                       */
/*   drift_particles_kernel (const)  src/base/Timestep.jl:181-191            */
/*     Pos += Vel * (t_new - t_old)                                          */
/* Pins: the hand integrator integrate_const_leapfrog                        */
/* (test/Local-Timestep.jl:54-91) restated independently in the tests,       */
/* time reversibility, second-order energy-error scaling.                    */
/* ------------------------------------------------------------------------ */
DS_EXPORT void ds_half_kick(int64_t n, double *vel, const double *acc, int forward, double dt) {
    for (int64_t i = 0; i < 3 * n; ++i) vel[i] += acc[i] * (double)forward * 0.5 * dt;
}

DS_EXPORT void ds_kick(int64_t n, double *vel, const double *acc, double dt) {
    for (int64_t i = 0; i < 3 * n; ++i) vel[i] += acc[i] * dt;
}

DS_EXPORT void ds_drift(int64_t n, double *x, double *y, double *z, const double *vel, double dt) {
    for (int64_t i = 0; i < n; ++i) {
        x[i] += vel[3 * i] * dt;
        y[i] += vel[3 * i + 1] * dt;
        z[i] += vel[3 * i + 2] * dt;
    }
}

/* ------------------------------------------------------------------------ */
/* a14: energy and momentum  src/base/Energy.jl:13-25,53-59,77-100           */
/*   K = sum 0.5 m v.v ; W = sum pot*m (the caller halves it, as             */
/*   total_potential(sim) does at src/directsum/potential.jl:152);           */
/*   P = sum m v ; L = sum m (x cross v).  Pairwise sums as Base.sum.        */
/* ------------------------------------------------------------------------ */
DS_EXPORT int ds_energy(int64_t n, const double *x, const double *y, const double *z,
                        const double *vel, const double *m, const double *pot, double *K,
                        double *Wsum, double P[3], double L[3]) {
    if (n <= 0) { *K = 0; *Wsum = 0; P[0] = P[1] = P[2] = 0; L[0] = L[1] = L[2] = 0; return 0; }
    double *s = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    if (!s) return -3;
    for (int64_t i = 0; i < n; ++i) {
        const double *v = vel + 3 * i;
        s[i] = 0.5 * m[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    *K = pairwise_sum1(s, 0, n - 1);
    if (pot) {
        for (int64_t i = 0; i < n; ++i) s[i] = pot[i] * m[i];
        *Wsum = pairwise_sum1(s, 0, n - 1);
    } else {
        *Wsum = 0.0;
    }
    for (int64_t i = 0; i < n; ++i) {
        const double *v = vel + 3 * i;
        s[3 * i] = v[0] * m[i]; s[3 * i + 1] = v[1] * m[i]; s[3 * i + 2] = v[2] * m[i];
    }
    pairwise_sum3(s, 0, n - 1, P);
    for (int64_t i = 0; i < n; ++i) {
        const double *v = vel + 3 * i;
        s[3 * i] = m[i] * (y[i] * v[2] - z[i] * v[1]);
        s[3 * i + 1] = m[i] * (z[i] * v[0] - x[i] * v[2]);
        s[3 * i + 2] = m[i] * (x[i] * v[1] - y[i] * v[0]);
    }
    pairwise_sum3(s, 0, n - 1, L);
    free(s);
    return 0;
}

DS_EXPORT int ds_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Ensemble of independent systems, each integrated with the constant-dt    */
/* KDK sequence of run(sim) (src/run.jl:76-112; src/base/Timestep.jl:84-93, */
/* :116-125, :181-191): force, +1/2 kick, then nsteps x {drift dt, force,   */
/* kick dt}.  No final -1/2 kick.  Per system s: G[s], softening soft[s]    */
/* for every particle, timestep dt[s]; the self pair is skipped (h may be   */
/* 0).  Massless sources are skipped: their terms are exactly +0.0.         */
/* E0[s] = K + W before the first kick; energy[s*slots + k] = K + W after   */
/* step (k+1)*every with velocities synchronised to the step (v - a dt/2).  */
/* Pinned against run_const_kdk (the run-loop restatement) in the tests.    */
/* ------------------------------------------------------------------------ */
static void sys_force(int64_t n, const double *x, const double *y, const double *z,
                      const double *m, double G, double h, double *acc, double *pot) {
    for (int64_t i = 0; i < n; ++i) {
        const double p[3] = {x[i], y[i], z[i]};
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, ph = 0.0, t[3];
        for (int64_t j = 0; j < n; ++j) {
            if (j == i || m[j] == 0.0) continue;
            pair_force(p, x[j], y[j], z[j], m[j], G, h, t);
            a0 += t[0]; a1 += t[1]; a2 += t[2];
            if (pot) ph += pair_potential(p, x[j], y[j], z[j], m[j], G, h);
        }
        acc[3 * i] = a0; acc[3 * i + 1] = a1; acc[3 * i + 2] = a2;
        if (pot) pot[i] = ph;
    }
}

static double sys_energy(int64_t n, const double *x, const double *y, const double *z,
                         const double *vel, const double *acc, const double *m, double G,
                         double h, double half_dt, double *scratch) {
    sys_force(n, x, y, z, m, G, h, scratch, scratch + 3 * n);
    double K = 0.0, W = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double ux = vel[3 * i] - half_dt * acc[3 * i];
        const double uy = vel[3 * i + 1] - half_dt * acc[3 * i + 1];
        const double uz = vel[3 * i + 2] - half_dt * acc[3 * i + 2];
        K += 0.5 * m[i] * (ux * ux + uy * uy + uz * uz);
        W += scratch[3 * n + i] * m[i];
    }
    return K + 0.5 * W;
}

DS_EXPORT int ds_run_ensemble_kdk(int32_t nsys, const int64_t *off, double *x, double *y,
                                  double *z, double *vel, const double *m, const double *G,
                                  const double *soft, const double *dt, int64_t nsteps,
                                  int32_t every, double *energy, double *E0, int nthreads) {
    int nt = set_threads(nthreads);
    const int64_t slots = every > 0 ? nsteps / every : 0;
    int err = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int32_t s = 0; s < nsys; ++s) {
        const int64_t o = off[s], n = off[s + 1] - off[s];
        double *X = x + o, *Y = y + o, *Z = z + o, *V = vel + 3 * o;
        const double *M = m + o;
        double *acc = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
        double *scr = (double *)malloc(sizeof(double) * 4 * (size_t)(n > 0 ? n : 1));
        if (!acc || !scr) {
            free(acc); free(scr);
#pragma omp atomic write
            err = -3;
            continue;
        }
        sys_force(n, X, Y, Z, M, G[s], soft[s], acc, NULL);
        if (E0) E0[s] = sys_energy(n, X, Y, Z, V, acc, M, G[s], soft[s], 0.0, scr);
        ds_half_kick(n, V, acc, +1, dt[s]);
        for (int64_t k = 1; k <= nsteps; ++k) {
            ds_drift(n, X, Y, Z, V, dt[s]);
            sys_force(n, X, Y, Z, M, G[s], soft[s], acc, NULL);
            ds_kick(n, V, acc, dt[s]);
            if (energy && every > 0 && k % every == 0 && k / every <= slots)
                energy[(int64_t)s * slots + (k / every - 1)] =
                    sys_energy(n, X, Y, Z, V, acc, M, G[s], soft[s], 0.5 * dt[s], scr);
        }
        free(acc);
        free(scr);
    }
    return err;
}
