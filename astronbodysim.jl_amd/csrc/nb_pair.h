// nb_pair.h -- device pair arithmetic for the Gadget-2 cubic-spline softened
// gravity of src/directsum/gravity.jl:8-28 (force) and
// src/directsum/potential.jl:2-27 (potential).
//
// Conventions used by every kernel:
//   fac  : force factor, so that acc_i += fac * (x_j - x_i)
//   q    : potential magnitude, so that phi_i -= q   (q = m/r outside h)
//   G is applied once per target after the sum.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>

namespace nb {

// Decimal literals as written in the reference (not 32/3 etc.).
#define NB_C1 10.666666666667
#define NB_C2 32.0
#define NB_C3 38.4
#define NB_C4 21.333333333333
#define NB_C5 48.0
#define NB_C6 0.066666666667
#define NB_P1 2.8
#define NB_P2 5.333333333333
#define NB_P3 6.4
#define NB_P4 9.6
#define NB_P5 3.2
#define NB_P6 16.0
#define NB_P7 2.133333333333

// ---------------------------------------------------------------------------
// FP32: softened / degenerate pairs.  Called only for r2 < max(h2, FLT_MIN):
// r2 == 0 (self pair or coincident particles) contributes nothing, which is
// the reference's value for h > 0 (dp == 0, and iszero(r) zeroes the
// potential); for h == 0 the reference yields NaN there and this returns 0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void spline_f32(float r2, float m, float h2, float hinv, float hinv3,
                                           float &fac, float &q) {
    if (r2 == 0.0f) {
        fac = 0.0f;
        q = 0.0f;
        return;
    }
    if (r2 >= h2) {  // only reachable with h == 0 and denormal r2
        float ri = __builtin_amdgcn_rsqf(r2);
        q = m * ri;
        fac = q * ri * ri;
        return;
    }
    float r = __builtin_sqrtf(r2);
    float u = r * hinv;
    float u2 = u * u;
    if (u < 0.5f) {
        fac = m * hinv3 * ((float)NB_C1 + u2 * ((float)NB_C2 * u - (float)NB_C3));
        q = -m * hinv * (-(float)NB_P1 + u2 * ((float)NB_P2 + u2 * ((float)NB_P3 * u - (float)NB_P4)));
    } else {
        float u3 = u2 * u;
        float rcpu = __builtin_amdgcn_rcpf(u);
        fac = m * hinv3 *
              ((float)NB_C4 - (float)NB_C5 * u + (float)NB_C3 * u2 - (float)NB_C1 * u3 -
               (float)NB_C6 * (rcpu * rcpu * rcpu));
        q = -m * hinv *
            (-(float)NB_P5 + (float)NB_C6 * rcpu +
             u2 * ((float)NB_C1 + u * (-(float)NB_P6 + u * ((float)NB_P4 - (float)NB_P7 * u))));
    }
}

// One FP32 pair.  CHECKED selects the per-pair softening test; the caller
// guarantees r >= h for every pair of an unchecked tile (and r > 0).
template <bool CHECKED, bool POT>
__device__ __forceinline__ void pair_f32(float sx, float sy, float sz, float sm, float xi, float yi,
                                         float zi, float h2c, float h2, float hinv, float hinv3,
                                         float &ax, float &ay, float &az, float &aq) {
    float dx = sx - xi;
    float dy = sy - yi;
    float dz = sz - zi;
    float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
    float ri = __builtin_amdgcn_rsqf(r2);
    float q = sm * ri;
    float fac = q * (ri * ri);
    if (CHECKED) {
        if (__builtin_expect(r2 < h2c, 0)) spline_f32(r2, sm, h2, hinv, hinv3, fac, q);
    }
    ax = __builtin_fmaf(fac, dx, ax);
    ay = __builtin_fmaf(fac, dy, ay);
    az = __builtin_fmaf(fac, dz, az);
    if (POT) aq += q;
}

// One FP32 pair with hi/lo coordinates: x = hi + lo (both FP32, relative to
// the block origin).  dx = (xj_hi - xi_hi) + (xj_lo - xi_lo) keeps the
// relative error of a close pair's separation at ~FP32 epsilon whatever the
// distance of the pair from the origin.
template <bool CHECKED, bool POT>
__device__ __forceinline__ void pair_f32_hilo(float sx, float sy, float sz, float sm, float sxl,
                                              float syl, float szl, float xi, float yi, float zi,
                                              float xil, float yil, float zil, float h2c,
                                              float h2, float hinv, float hinv3, float &ax,
                                              float &ay, float &az, float &aq) {
    float dx = (sx - xi) + (sxl - xil);
    float dy = (sy - yi) + (syl - yil);
    float dz = (sz - zi) + (szl - zil);
    float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
    float ri = __builtin_amdgcn_rsqf(r2);
    float q = sm * ri;
    float fac = q * (ri * ri);
    if (CHECKED) {
        if (__builtin_expect(r2 < h2c, 0)) spline_f32(r2, sm, h2, hinv, hinv3, fac, q);
    }
    ax = __builtin_fmaf(fac, dx, ax);
    ay = __builtin_fmaf(fac, dy, ay);
    az = __builtin_fmaf(fac, dz, az);
    if (POT) aq += q;
}

// One FP32 pair whose separation comes from FP64 coordinates (both relative
// to the block origin): dx = fl32(xj - xi) carries a relative error of at most
// 2^-24 (plus 2^-53 (|xi| + |xj|) / |dx| from the FP64 subtraction), as the
// hi/lo difference does, with one FP64 subtraction and one conversion per axis
// instead of three FP32 operations.
template <bool CHECKED, bool POT>
__device__ __forceinline__ void pair_f32_d(double sx, double sy, double sz, float sm, double xi,
                                           double yi, double zi, float h2c, float h2, float hinv,
                                           float hinv3, float &ax, float &ay, float &az, float &aq) {
    const float dx = (float)(sx - xi);
    const float dy = (float)(sy - yi);
    const float dz = (float)(sz - zi);
    float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
    float ri = __builtin_amdgcn_rsqf(r2);
    float q = sm * ri;
    float fac = q * (ri * ri);
    if (CHECKED) {
        if (__builtin_expect(r2 < h2c, 0)) spline_f32(r2, sm, h2, hinv, hinv3, fac, q);
    }
    ax = __builtin_fmaf(fac, dx, ax);
    ay = __builtin_fmaf(fac, dy, ay);
    az = __builtin_fmaf(fac, dz, az);
    if (POT) aq += q;
}

// ---------------------------------------------------------------------------
// FP64 pair (correctness mode): the reference's formula in double.
// ---------------------------------------------------------------------------
template <bool POT>
__device__ __forceinline__ void pair_f64(double sx, double sy, double sz, double sm, double xi,
                                         double yi, double zi, double h, double hinv, double hinv3,
                                         double &ax, double &ay, double &az, double &aq) {
    double dx = sx - xi, dy = sy - yi, dz = sz - zi;
    double r2 = dx * dx + dy * dy + dz * dz;
    if (r2 == 0.0) return;  // self pair / coincident: 0 for h > 0 (NaN in the reference for h == 0)
    double r = __builtin_sqrt(r2);
    double fac, q;
    if (r >= h) {
        fac = sm / (r2 * r);
        q = sm / r;
    } else {
        double u = r * hinv;
        double u2 = u * u;
        if (u < 0.5) {
            fac = sm * hinv3 * (NB_C1 + u2 * (NB_C2 * u - NB_C3));
            q = -(sm * hinv * (-NB_P1 + u2 * (NB_P2 + u2 * (NB_P3 * u - NB_P4))));
        } else {
            double u3 = u2 * u;
            fac = sm * hinv3 * (NB_C4 - NB_C5 * u + NB_C3 * u2 - NB_C1 * u3 - NB_C6 / u3);
            q = -(sm * hinv *
                  (-NB_P5 + NB_C6 / u + u2 * (NB_C1 + u * (-NB_P6 + u * (NB_P4 - NB_P7 * u)))));
        }
    }
    ax += dx * fac;
    ay += dy * fac;
    az += dz * fac;
    if (POT) aq += q;
}

}  // namespace nb
