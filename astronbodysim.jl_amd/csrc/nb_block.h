// nb_block.h -- the block-level all-pairs routine shared by the one-system
// kernel (nb_force.hip) and the ragged ensemble kernel (nb_batch.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "nb_internal.h"
#include "nb_pair.h"

namespace nb {

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).
// Map hardware block ids so that each XCD receives a contiguous range of
// logical blocks; logical blocks of one source segment then share an L2.
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
    constexpr int X = 8;
    int per = nblocks / X, rem = nblocks % X;
    int xcd = b % X, slot = b / X;
    return xcd * per + (xcd < rem ? xcd : rem) + slot;
}

template <typename V>
__device__ __forceinline__ V wave_sum(V v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
template <typename V>
__device__ __forceinline__ V wave_min(V v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}
template <typename V>
__device__ __forceinline__ V wave_max(V v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Block statistics of the targets: centroid, bounding box, largest h.
struct BlockStats {
    double ox, oy, oz;        // origin (centroid, or 0)
    double lo[3], hi[3];      // bounding box of the valid targets
    double hmax;
};

template <int TPT>
__device__ void block_stats(const double (&px)[TPT], const double (&py)[TPT], const double (&pz)[TPT],
                            const double (&ph)[TPT], const bool (&valid)[TPT], bool centroid,
                            double (*red)[NW], BlockStats &st) {
    double v[11];
    v[0] = v[1] = v[2] = v[3] = 0.0;
    v[4] = v[5] = v[6] = DBL_MAX;
    v[7] = v[8] = v[9] = -DBL_MAX;
    v[10] = 0.0;
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        if (valid[k]) {
            v[0] += px[k]; v[1] += py[k]; v[2] += pz[k]; v[3] += 1.0;
            v[4] = fmin(v[4], px[k]); v[5] = fmin(v[5], py[k]); v[6] = fmin(v[6], pz[k]);
            v[7] = fmax(v[7], px[k]); v[8] = fmax(v[8], py[k]); v[9] = fmax(v[9], pz[k]);
            v[10] = fmax(v[10], ph[k]);
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = wave_sum(v[q]);
#pragma unroll
    for (int q = 4; q < 7; ++q) v[q] = wave_min(v[q]);
#pragma unroll
    for (int q = 7; q < 11; ++q) v[q] = wave_max(v[q]);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 11; ++q) red[q][w] = v[q];
    }
    __syncthreads();
    double r[11];
#pragma unroll
    for (int q = 0; q < 11; ++q) r[q] = red[q][0];
#pragma unroll
    for (int ww = 1; ww < NW; ++ww) {
#pragma unroll
        for (int q = 0; q < 4; ++q) r[q] += red[q][ww];
#pragma unroll
        for (int q = 4; q < 7; ++q) r[q] = fmin(r[q], red[q][ww]);
#pragma unroll
        for (int q = 7; q < 11; ++q) r[q] = fmax(r[q], red[q][ww]);
    }
    __syncthreads();
    if (centroid && r[3] > 0.0) {
        st.ox = r[0] / r[3]; st.oy = r[1] / r[3]; st.oz = r[2] / r[3];
    } else {
        st.ox = st.oy = st.oz = 0.0;
    }
    st.lo[0] = r[4]; st.lo[1] = r[5]; st.lo[2] = r[6];
    st.hi[0] = r[7]; st.hi[1] = r[8]; st.hi[2] = r[9];
    st.hmax = r[10];
}

// Source staged for LDS: position relative to the block origin, FP32.
// Sources past the end of the range (or excluded) become massless particles
// placed far away, so both tile paths can run a fixed trip count.
constexpr float PAD_COORD = 1.0e18f;

// Softening-checked pairs, GR sources at a time: the Newtonian terms of all
// of them first, then one wave-level test for a pair inside some target's h
// (rare even in these tiles), then the sums in source order.  One branch per
// group instead of one per pair keeps the group's arithmetic in one block for
// the scheduler; the results are those of pair_f32 / pair_f32_d pair by pair.
template <int TPT>
struct CheckedGroup {
    static constexpr int GR = TPT >= 4 ? 1 : 4 / TPT;
};

template <int TPT, bool POT, int GR>
__device__ __forceinline__ void checked_group_sums(const float (&dx)[GR][TPT], const float (&dy)[GR][TPT],
                                                   const float (&dz)[GR][TPT], const float (&r2)[GR][TPT],
                                                   float (&fac)[GR][TPT], float (&q)[GR][TPT],
                                                   const float (&m)[GR], bool any, const float (&h2c)[TPT],
                                                   const float (&h2)[TPT], const float (&hinv)[TPT],
                                                   const float (&hinv3)[TPT], float (&ax)[TPT], float (&ay)[TPT],
                                                   float (&az)[TPT], float (&aq)[TPT]) {
    if (__builtin_expect(any, 0)) {
#pragma unroll
        for (int g = 0; g < GR; ++g)
#pragma unroll
            for (int k = 0; k < TPT; ++k)
                if (r2[g][k] < h2c[k]) spline_f32(r2[g][k], m[g], h2[k], hinv[k], hinv3[k], fac[g][k], q[g][k]);
    }
#pragma unroll
    for (int g = 0; g < GR; ++g)
#pragma unroll
        for (int k = 0; k < TPT; ++k) {
            ax[k] = __builtin_fmaf(fac[g][k], dx[g][k], ax[k]);
            ay[k] = __builtin_fmaf(fac[g][k], dy[g][k], ay[k]);
            az[k] = __builtin_fmaf(fac[g][k], dz[g][k], az[k]);
            if (POT) aq[k] += q[g][k];
        }
}

// NS sources from s: a whole tile (BS) or one 64-source quarter of it.
template <int TPT, bool CHECKED, bool POT, int NS = BS>
__device__ __forceinline__ void tile_f32(const float4 *__restrict__ s, const float (&xi)[TPT],
                                         const float (&yi)[TPT], const float (&zi)[TPT],
                                         const float (&h2c)[TPT], const float (&h2)[TPT],
                                         const float (&hinv)[TPT], const float (&hinv3)[TPT],
                                         float (&ax)[TPT], float (&ay)[TPT], float (&az)[TPT],
                                         float (&aq)[TPT]) {
    if constexpr (CHECKED && CheckedGroup<TPT>::GR > 1) {
        constexpr int GR = CheckedGroup<TPT>::GR;
#pragma unroll 1
        for (int jj = 0; jj < NS; jj += GR) {
            float dx[GR][TPT], dy[GR][TPT], dz[GR][TPT], r2[GR][TPT], fac[GR][TPT], q[GR][TPT], m[GR];
            bool any = false;
#pragma unroll
            for (int g = 0; g < GR; ++g) {
                const float4 v = s[jj + g];
                m[g] = v.w;
#pragma unroll
                for (int k = 0; k < TPT; ++k) {
                    dx[g][k] = v.x - xi[k];
                    dy[g][k] = v.y - yi[k];
                    dz[g][k] = v.z - zi[k];
                    r2[g][k] = __builtin_fmaf(dx[g][k], dx[g][k], __builtin_fmaf(dy[g][k], dy[g][k], dz[g][k] * dz[g][k]));
                    const float ri = __builtin_amdgcn_rsqf(r2[g][k]);
                    q[g][k] = v.w * ri;
                    fac[g][k] = q[g][k] * (ri * ri);
                    any = any || r2[g][k] < h2c[k];
                }
            }
            checked_group_sums<TPT, POT, GR>(dx, dy, dz, r2, fac, q, m, any, h2c, h2, hinv, hinv3, ax, ay, az, aq);
        }
        return;
    }
#pragma unroll 4
    for (int jj = 0; jj < NS; ++jj) {
        const float4 v = s[jj];
#pragma unroll
        for (int k = 0; k < TPT; ++k)
            pair_f32<CHECKED, POT>(v.x, v.y, v.z, v.w, xi[k], yi[k], zi[k], h2c[k], h2[k], hinv[k],
                                   hinv3[k], ax[k], ay[k], az[k], aq[k]);
    }
}

template <int TPT, bool CHECKED, bool POT, int NS = BS>
__device__ __forceinline__ void tile_f32_hilo(const float4 *__restrict__ s,
                                              const float4 *__restrict__ sl,
                                              const float (&xi)[TPT], const float (&yi)[TPT],
                                              const float (&zi)[TPT], const float (&xil)[TPT],
                                              const float (&yil)[TPT], const float (&zil)[TPT],
                                              const float (&h2c)[TPT], const float (&h2)[TPT],
                                              const float (&hinv)[TPT], const float (&hinv3)[TPT],
                                              float (&ax)[TPT], float (&ay)[TPT],
                                              float (&az)[TPT], float (&aq)[TPT]) {
#pragma unroll 2
    for (int jj = 0; jj < NS; ++jj) {
        const float4 v = s[jj];
        const float4 w = sl[jj];
#pragma unroll
        for (int k = 0; k < TPT; ++k)
            pair_f32_hilo<CHECKED, POT>(v.x, v.y, v.z, v.w, w.x, w.y, w.z, xi[k], yi[k], zi[k],
                                        xil[k], yil[k], zil[k], h2c[k], h2[k], hinv[k], hinv3[k],
                                        ax[k], ay[k], az[k], aq[k]);
    }
}

// A source for the near tiles of FP64-state blocks on far path 0: its FP64
// position relative to the block origin and its mass.
struct SrcF64 {
    double x, y, z;
    float m, pad;
};

template <int TPT, bool CHECKED, bool POT, int NS = BS>
__device__ __forceinline__ void tile_f32_d(const SrcF64 *__restrict__ s, const double (&xi)[TPT],
                                           const double (&yi)[TPT], const double (&zi)[TPT],
                                           const float (&h2c)[TPT], const float (&h2)[TPT],
                                           const float (&hinv)[TPT], const float (&hinv3)[TPT],
                                           float (&ax)[TPT], float (&ay)[TPT], float (&az)[TPT],
                                           float (&aq)[TPT]) {
    if constexpr (CHECKED && CheckedGroup<TPT>::GR > 1) {
        constexpr int GR = CheckedGroup<TPT>::GR;
#pragma unroll 1
        for (int jj = 0; jj < NS; jj += GR) {
            float dx[GR][TPT], dy[GR][TPT], dz[GR][TPT], r2[GR][TPT], fac[GR][TPT], q[GR][TPT], m[GR];
            bool any = false;
#pragma unroll
            for (int g = 0; g < GR; ++g) {
                const SrcF64 v = s[jj + g];
                m[g] = v.m;
#pragma unroll
                for (int k = 0; k < TPT; ++k) {
                    dx[g][k] = (float)(v.x - xi[k]);
                    dy[g][k] = (float)(v.y - yi[k]);
                    dz[g][k] = (float)(v.z - zi[k]);
                    r2[g][k] = __builtin_fmaf(dx[g][k], dx[g][k], __builtin_fmaf(dy[g][k], dy[g][k], dz[g][k] * dz[g][k]));
                    const float ri = __builtin_amdgcn_rsqf(r2[g][k]);
                    q[g][k] = v.m * ri;
                    fac[g][k] = q[g][k] * (ri * ri);
                    any = any || r2[g][k] < h2c[k];
                }
            }
            checked_group_sums<TPT, POT, GR>(dx, dy, dz, r2, fac, q, m, any, h2c, h2, hinv, hinv3, ax, ay, az, aq);
        }
        return;
    }
#pragma unroll 2
    for (int jj = 0; jj < NS; ++jj) {
        const SrcF64 v = s[jj];
#pragma unroll
        for (int k = 0; k < TPT; ++k)
            pair_f32_d<CHECKED, POT>(v.x, v.y, v.z, v.m, xi[k], yi[k], zi[k], h2c[k], h2[k], hinv[k],
                                     hinv3[k], ax[k], ay[k], az[k], aq[k]);
    }
}

// Tile classification (wave-uniform), from the Euclidean distance d between
// a source tile's box and the box of the wave's targets:
//   d <= h_max (+ a resolution margin)      -> "soft": softening-checked pairs
//                                              (hi/lo coordinates in SPLIT);
//   d <= R/2 (SPLIT only; R = largest         -> "close": hi/lo coordinates,
//     distance of the wave's targets from       Newtonian (every pair has r > h);
//     the block's FP64 origin)
//   otherwise                               -> "far": 12-VALU Newtonian loop on
//                                              hi coordinates.
// For a far pair the separation x_j - x_i, formed from two FP32 roundings of
// coordinates of size <= R + r, carries a relative error <= 2 eps (1 + R/r)
// <= 6 eps since r > R/2.
// With MF >= 1 (SPLIT only) the far tiles use the matrix cores for the sum:
//   a_i = sum_j m_j r_ij^-3 (x'_j - x'_i) = S_xyz - x'_i S_m,
//   S = sum_j r_ij^-3 (m_j x'_j, m_j y'_j, m_j z'_j, m_j),
// one v_mfma_f32_4x4x1f32 per source and wave: 16 blocks of 4 targets x 1
// source x 4 columns.  With MF = 1 the VALU keeps the 8-op r^-3 part of every
// pair.  In a far tile |x'_i| <= R < 2r and |x'_j| <= R + r < 3r, so each term
// of S is bounded by a few times the pair's force and the subtraction costs a
// few eps.
// With MF = 2 the squared distance of the farthest tiles comes from the matrix
// cores too:
//   r_ij^2 = |x'_i|^2 + |x'_j|^2 - 2 x'_i . x'_j,
// four v_mfma_f32_4x4x1f32 per 4 sources and wave (K = x, y, z, |x'_j|^2 with
// C = |x'_i|^2), so the VALU does only rsq and two multiplies per pair.  The
// expansion loses about 3 eps (|x'_i| + |x'_j|)^2 / r^2 relative; the tiles it
// serves have r >= MF2_REACH * Re, where Re is the wave's largest Euclidean
// distance from the block origin, so (|x'_i| + |x'_j|) <= r + 2 Re
// <= (1 + 2 / MF2_REACH) r and the loss stays below about 8 eps.  Tiles in
// between take the MF = 1 loop.
constexpr double MF2_REACH = 3.0;
// Far tiles with r >= FARW_REACH * Re in SPLIT mode ("weighted" far tiles):
//   f_ij = m_j r_ij^-3 = q^3,  q = rsq(w_j r_ij^2),  w_j = m_j^(-2/3),
//   w_j r_ij^2 = w_j |x'_j|^2 + w_j |x'_i|^2 - 2 (w_j x'_j) . x'_i,
// four FMAs from the source's (w x', w y', w z', w |x'|^2, w) and the target's
// (-2x'_i, -2y'_i, -2z'_i, |x'_i|^2), with the sum kept as
//   S = sum_j f_ij (x'_j, y'_j, z'_j, 1),   a_i = S_xyz - x'_i S_m:
// ten VALU and one v_rsq_f32 per pair against twelve and one.  The expansion
// loses about 3 eps (|x'_i| + |x'_j|)^2 / r^2 <= 3 eps (1 + 2 / 0.5)^2 = 75 eps
// relative in w r^2 for r >= 0.5 Re, since |x'_i| <= Re and |x'_j| <= r + Re;
// sampled pairs at that reach stay below 41 eps (median 1 eps), and the
// forces of the two-sphere benchmark keep a median error of 3.5e-8 against
// FP64 (3.2e-8 with the tiles between 0.5 Re and 1.5 Re on the exact loop).
// Tiles nearer than R/2 take the hi/lo loop anyway.  Massless sources get
// w = 0 and w |x'|^2 = 2^120, so their f is exactly zero.
constexpr double FARW_REACH = 0.5;

// With FAR_MFMA_BF16 the squared distances of the tiles past FARW_REACH * Re
// come from bf16 matrix cores (v_mfma_f32_16x16x32_bf16), which run alongside
// the VALU, and the sum stays on the VALU in the S form above.  As in the
// weighted path, the source rows carry w_j = m_j^(-2/3), so the product is
// w_j r_ij^2 and f = rsq(w r^2)^3: rsq, two multiplies, three FMAs and an
// add per pair.  Every FP32 operand is split into three bf16 (x = x0 + x1 +
// x2 exactly); the K = 32 product holds the cross terms of split orders
// a + b <= 2 of each axis (the dropped ones are below 2^-23 |x_i||x_j|),
// w |x'_j|^2 and w |x'_i|^2, accumulated in FP32, so w r^2 carries the
// few-eps (|x'_i| + |x'_j|)^2 error of the weighted expansion.

typedef __bf16 nb_bf8 __attribute__((ext_vector_type(8)));
typedef unsigned nb_u4v __attribute__((ext_vector_type(4)));
constexpr float FARW_MASSLESS = 1.329228e36f;   // 2^120

// Shared memory of one all-pairs block.
// HILO: positions relative to the block's FP64 origin, near tiles on hi/lo
// coordinates.  FP64 state always; FP32 state with the bf16 matrix-core far
// tiles, whose squared-distance expansion needs origin-relative coordinates.
// ABSNEAR (FP32 state with those far tiles): the near tiles take the FP32
// positions themselves instead of hi/lo, as the FP32-state loops do: two
// FP32 positions differ by an FP32 subtraction with a relative error of at
// most 2^-24, as small as the hi/lo difference's, for 12 VALU per pair
// instead of 18.  sjl then holds (x, y, z, m) of each source.
// D64NEAR (FP64 state on far path 0): the near tiles take the sources' FP64
// positions relative to the origin (sjd) and form each separation with one
// FP64 subtraction and a conversion per axis (pair_f32_d): 15 VALU per pair
// instead of the hi/lo loop's 18, at the same accuracy; sjl is then unused.
template <typename T, int MF = 0>
struct BlockShared {
    static constexpr bool HILO = sizeof(T) == 8 || MF == FAR_MFMA_BF16;
    static constexpr bool ABSNEAR = sizeof(T) == 4 && MF == FAR_MFMA_BF16;
    static constexpr bool D64NEAR = sizeof(T) == 8 && MF == FAR_VALU;
    float4 sj[2][BS];
    float4 sjl[HILO && !D64NEAR ? 2 : 1][HILO && !D64NEAR ? BS : 1];
    SrcF64 sjd[D64NEAR ? 2 : 1][D64NEAR ? BS : 1];
    static constexpr bool COLS = MF == 1 || MF == 2;   // the MFMA far paths
    float sB[COLS ? 2 : 1][COLS ? 4 : 1][COLS ? BS : 1];   // MFMA B columns, component-major
    float sQ[MF == 2 ? 2 : 1][MF == 2 ? BS : 1];     // |x'_j|^2 of each source
    static constexpr bool WFAR = HILO && MF == 3;
    float4 sW[WFAR ? 2 : 1][WFAR ? BS : 1];          // (w x', w y', w z', w |x'|^2)
    float sw[WFAR ? 2 : 1][WFAR ? BS : 1];           // w = m^(-2/3)
    static constexpr bool ROWS = HILO && MF == 4;
    nb_bf8 sA[1][ROWS ? BS : 1][4];                   // bf16 A rows (K = 32), single-buffered
    double red[11][NW];
};

typedef float nb_f4v __attribute__((ext_vector_type(4)));

// Far tile on the matrix cores.  Lane l = 4b + q supplies A = r^-3 of its own
// target and B = component q of the source; D[v] of lane 4b + q accumulates
// target 4b + v, component q.  Sources are taken four at a time: the VALU
// part of the four pairs first (independent, so it interleaves), then four
// MFMAs alternating between two accumulators (no back-to-back dependence).
__device__ __forceinline__ void tile_mfma(const float4 *__restrict__ s, const float *__restrict__ sB,
                                          int q, float xi, float yi, float zi, double (&S)[4]) {
    const float *bp = sB + q * BS;
    nb_f4v D0 = {0.0f, 0.0f, 0.0f, 0.0f}, D1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 2
    for (int jj = 0; jj < BS; jj += 4) {
        const float4 bq = *reinterpret_cast<const float4 *>(bp + jj);
        float r3[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float4 v = s[jj + u];
            const float dx = v.x - xi, dy = v.y - yi, dz = v.z - zi;
            const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
            const float ri = __builtin_amdgcn_rsqf(r2);
            r3[u] = (ri * ri) * ri;
        }
        D0 = __builtin_amdgcn_mfma_f32_4x4x1f32(r3[0], bq.x, D0, 0, 0, 0);
        D1 = __builtin_amdgcn_mfma_f32_4x4x1f32(r3[1], bq.y, D1, 0, 0, 0);
        D0 = __builtin_amdgcn_mfma_f32_4x4x1f32(r3[2], bq.z, D0, 0, 0, 0);
        D1 = __builtin_amdgcn_mfma_f32_4x4x1f32(r3[3], bq.w, D1, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) S[v] += (double)D0[v] + (double)D1[v];
}

// Farthest tiles with MF = 2.  Lane l = 4b + q supplies, for the four sources
// jj..jj+3, A = component of source jj + q and B = the matching component of
// its own target (-2x'_i, -2y'_i, -2z'_i, 1); R[v] of lane l then holds r^2 of
// (target l, source jj + v).  The r^-3 values feed the same column MFMAs as
// tile_mfma.  Two groups of four sources are interleaved so that no MFMA
// waits on the one just issued.
__device__ __forceinline__ void tile_mfma2(const float4 *__restrict__ s, const float *__restrict__ sq,
                                           const float *__restrict__ sB, int q, float bx, float by,
                                           float bz, nb_f4v ci, double (&S)[4]) {
    const float *bp = sB + q * BS;
    nb_f4v D0 = {0.0f, 0.0f, 0.0f, 0.0f}, D1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 1
    for (int jj = 0; jj < BS; jj += 8) {
        const float4 a0 = s[jj + q], a1 = s[jj + 4 + q];
        const float q0 = sq[jj + q], q1 = sq[jj + 4 + q];
        nb_f4v R0 = __builtin_amdgcn_mfma_f32_4x4x1f32(a0.x, bx, ci, 0, 0, 0);
        nb_f4v R1 = __builtin_amdgcn_mfma_f32_4x4x1f32(a1.x, bx, ci, 0, 0, 0);
        R0 = __builtin_amdgcn_mfma_f32_4x4x1f32(a0.y, by, R0, 0, 0, 0);
        R1 = __builtin_amdgcn_mfma_f32_4x4x1f32(a1.y, by, R1, 0, 0, 0);
        R0 = __builtin_amdgcn_mfma_f32_4x4x1f32(a0.z, bz, R0, 0, 0, 0);
        R1 = __builtin_amdgcn_mfma_f32_4x4x1f32(a1.z, bz, R1, 0, 0, 0);
        R0 = __builtin_amdgcn_mfma_f32_4x4x1f32(q0, 1.0f, R0, 0, 0, 0);
        R1 = __builtin_amdgcn_mfma_f32_4x4x1f32(q1, 1.0f, R1, 0, 0, 0);
        const float4 b0 = *reinterpret_cast<const float4 *>(bp + jj);
        const float4 b1 = *reinterpret_cast<const float4 *>(bp + jj + 4);
        float r3a[4], r3b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float ra = __builtin_amdgcn_rsqf(R0[u]);
            const float rb = __builtin_amdgcn_rsqf(R1[u]);
            r3a[u] = (ra * ra) * ra;
            r3b[u] = (rb * rb) * rb;
        }
        D0 = __builtin_amdgcn_mfma_f32_4x4x1f32(r3a[0], b0.x, D0, 0, 0, 0);
        D1 = __builtin_amdgcn_mfma_f32_4x4x1f32(r3a[1], b0.y, D1, 0, 0, 0);
        D0 = __builtin_amdgcn_mfma_f32_4x4x1f32(r3a[2], b0.z, D0, 0, 0, 0);
        D1 = __builtin_amdgcn_mfma_f32_4x4x1f32(r3a[3], b0.w, D1, 0, 0, 0);
        D0 = __builtin_amdgcn_mfma_f32_4x4x1f32(r3b[0], b1.x, D0, 0, 0, 0);
        D1 = __builtin_amdgcn_mfma_f32_4x4x1f32(r3b[1], b1.y, D1, 0, 0, 0);
        D0 = __builtin_amdgcn_mfma_f32_4x4x1f32(r3b[2], b1.z, D0, 0, 0, 0);
        D1 = __builtin_amdgcn_mfma_f32_4x4x1f32(r3b[3], b1.w, D1, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) S[v] += (double)D0[v] + (double)D1[v];
}

// Weighted far tile (see FARW_REACH): TPT targets per lane, S-form sums.
template <int TPT>
__device__ __forceinline__ void tile_w(const float4 *__restrict__ s, const float4 *__restrict__ sW,
                                       const float *__restrict__ sw, const float (&bx)[TPT],
                                       const float (&by)[TPT], const float (&bz)[TPT],
                                       const float (&ci)[TPT], float (&Sx)[TPT], float (&Sy)[TPT],
                                       float (&Sz)[TPT], float (&Sm)[TPT]) {
#pragma unroll 1
    for (int j = 0; j < BS; ++j) {
        const float4 W = sW[j];
        const float w = sw[j];
        const float4 p = s[j];
#pragma unroll
        for (int k = 0; k < TPT; ++k) {
            const float r2w = __builtin_fmaf(
                W.x, bx[k], __builtin_fmaf(W.y, by[k], __builtin_fmaf(W.z, bz[k], __builtin_fmaf(w, ci[k], W.w))));
            const float q = __builtin_amdgcn_rsqf(r2w);
            const float f = (q * q) * q;
            Sx[k] = __builtin_fmaf(f, p.x, Sx[k]);
            Sy[k] = __builtin_fmaf(f, p.y, Sy[k]);
            Sz[k] = __builtin_fmaf(f, p.z, Sz[k]);
            Sm[k] += f;
        }
    }
}

// A wave-uniform double moved to scalar registers.
__device__ __forceinline__ double uniform_f64(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// ---- bf16-split squared distances on the matrix cores (FAR_MFMA_BF16) ----

// K layout of the weighted squared-distance product (A = source row, B =
// target column), w_j = m_j^(-2/3):
//   k = 6d + c (d = x, y, z; c = 0..5): A = (w x'_j)_d^RA[c], B = -2 (x'_i)_d^RB[c]
//   k = 18..20:                         A = split of w |x'_j|^2, B = 1
//   k = 21..26:                         A = w^RA[c], B = (|x'_i|^2)^RB[c]
//   k = 27..31:                         0
// with (RA[c], RB[c]) = (0,0) (0,1) (1,0) (0,2) (2,0) (1,1): every cross term of
// split orders a + b <= 2, the dropped ones being below 2^-23 of the product.
// The product is w_j r_ij^2, so f_ij = m_j r_ij^-3 = rsq(w_j r_ij^2)^3.

typedef float nb_f2v __attribute__((ext_vector_type(2)));
typedef __bf16 nb_bf2 __attribute__((ext_vector_type(2)));

// Two FP32 values (each exactly a bf16) packed into one dword (v_cvt_pk_bf16_f32).
__device__ __forceinline__ unsigned pk_bf16(float a, float b) {
    const nb_f2v v = {a, b};
    return __builtin_bit_cast(unsigned, __builtin_convertvector(v, nb_bf2));
}

// Two FP32 values that are exactly bf16 (low 16 bits zero) packed into one
// dword by a byte permute: the same bits as pk_bf16, without the conversion.
__device__ __forceinline__ unsigned pk_exact(float a, float b) {
    return __builtin_amdgcn_perm(__builtin_bit_cast(unsigned, b), __builtin_bit_cast(unsigned, a), 0x07060302u);
}

// The three-bf16 splits of two values at once, as FP32 (each exactly a bf16).
__device__ __forceinline__ void split2(float x0, float x1, float (&s0)[3], float (&s1)[3]) {
    const nb_f2v v = {x0, x1};
    const nb_f2v h = __builtin_convertvector(__builtin_convertvector(v, nb_bf2), nb_f2v);
    const nb_f2v r = v - h;
    const nb_f2v m = __builtin_convertvector(__builtin_convertvector(r, nb_bf2), nb_f2v);
    const nb_f2v l = r - m;
    s0[0] = h.x; s0[1] = m.x; s0[2] = l.x;
    s1[0] = h.y; s1[1] = m.y; s1[2] = l.y;
}

__device__ __forceinline__ void store_words(const unsigned (&w)[16], nb_bf8 (&out)[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const nb_u4v u = {w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]};
        out[q] = __builtin_bit_cast(nb_bf8, u);
    }
}

// A row of a source with weight w > 0: (wx, wy, wz) = w x'_j, wq = w |x'_j|^2.
// Pairs of K entries are packed from their FP32 splits: entries 6d..6d+5 are
// (h, h, m, h, l, m) of axis d (RA order), 18..20 (h, m, l) of wq, 21..26 the
// RA order of w, then zeros.
__device__ __forceinline__ void bf16_source_row(float wx, float wy, float wz, float wq, float w,
                                                nb_bf8 (&out)[4]) {
    float sx[3], sy[3], sz[3], sq[3], sw[3], s0[3];
    split2(wx, wy, sx, sy);
    split2(wz, wq, sz, sq);
    split2(w, 0.0f, sw, s0);
    unsigned d[16];
    d[0] = pk_exact(sx[0], sx[0]); d[1] = pk_exact(sx[1], sx[0]); d[2] = pk_exact(sx[2], sx[1]);
    d[3] = pk_exact(sy[0], sy[0]); d[4] = pk_exact(sy[1], sy[0]); d[5] = pk_exact(sy[2], sy[1]);
    d[6] = pk_exact(sz[0], sz[0]); d[7] = pk_exact(sz[1], sz[0]); d[8] = pk_exact(sz[2], sz[1]);
    d[9] = pk_exact(sq[0], sq[1]); d[10] = pk_exact(sq[2], sw[0]); d[11] = pk_exact(sw[0], sw[1]);
    d[12] = pk_exact(sw[0], sw[2]); d[13] = pk_exact(sw[1], 0.0f); d[14] = 0u; d[15] = 0u;
    store_words(d, out);
}

// A row of a massless (or padding) source: w r^2 = 2^120, so f = 0 exactly.
__device__ __forceinline__ void bf16_massless_row(nb_bf8 (&out)[4]) {
    unsigned d[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) d[q] = 0u;
    d[9] = pk_bf16(FARW_MASSLESS, 0.0f);
    store_words(d, out);
}

// B column of a target at x' with c = |x'|^2: entries 6d..6d+5 are -2 times
// (h, m, h, l, h, m) of axis d (RB order), 18..20 ones, 21..26 the RB order
// of c, then zeros.
__device__ __forceinline__ void bf16_target_col(float x, float y, float z, float c,
                                                nb_bf8 (&out)[4]) {
    float tx[3], ty[3], tz[3], tc[3];
    split2(x, y, tx, ty);
    split2(z, c, tz, tc);
    float (*t[3])[3] = {&tx, &ty, &tz};
    unsigned d[16];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float h = -2.0f * (*t[a])[0], m = -2.0f * (*t[a])[1], l = -2.0f * (*t[a])[2];
        d[3 * a] = pk_bf16(h, m);
        d[3 * a + 1] = pk_bf16(h, l);
        d[3 * a + 2] = pk_bf16(h, m);
    }
    d[9] = pk_bf16(1.0f, 1.0f);
    d[10] = pk_bf16(1.0f, tc[0]);
    d[11] = pk_bf16(tc[1], tc[0]);
    d[12] = pk_bf16(tc[2], tc[0]);
    d[13] = pk_bf16(tc[1], 0.0f);
    d[14] = 0u;
    d[15] = 0u;
    store_words(d, out);
}

__device__ __forceinline__ nb_bf8 shfl_bf8(nb_bf8 v, int src) {
    nb_u4v u = __builtin_bit_cast(nb_u4v, v);
#pragma unroll
    for (int e = 0; e < 4; ++e) u[e] = (unsigned)__shfl((int)u[e], src, 64);
    return __builtin_bit_cast(nb_bf8, u);
}

// The B operands of lane `lane` for the four 16-target groups of its wave
// (FAR_MFMA_BF16): lane l holds, for group g, the K slice (l >> 4) of the B
// column of target 16g + (l & 15).  (dx, dy, dz): the lane's target
// relative to the block origin.  Evaluated without FMA contraction, so that
// the force kernel and k_target_geom produce the same operands.
__device__ __forceinline__ void target_b_operands(double dx, double dy, double dz, int lane,
                                                  nb_bf8 (&Bg)[4]) {
#pragma clang fp contract(off)
    nb_bf8 mine[4];
    bf16_target_col((float)dx, (float)dy, (float)dz, (float)(dx * dx + dy * dy + dz * dz), mine);
#pragma unroll
    for (int g = 0; g < 4; ++g) {
#pragma unroll
        for (int sl = 0; sl < 4; ++sl) {
            const nb_bf8 t = shfl_bf8(mine[sl], 16 * g + (lane & 15));
            if (sl == (lane >> 4)) Bg[g] = t;
        }
    }
}

// Farthest tiles with FAR_MFMA_BF16.  Chunk of 16 sources jc..jc+15: MFMA g gives
// lane l the w r^2 of targets 16g + (l & 15) (the B columns) against sources
// jc + 4(l >> 4) + v in element v (the A rows); the lane then accumulates
// S[g] for its four targets over its four sources.
template <bool FIRST>
__device__ __forceinline__ void mf3_chunk(const float4 *__restrict__ s, const nb_bf8 a, int grp,
                                          const nb_bf8 (&Bg)[4], float (&S)[4][4]) {
    const nb_f4v zero = {0.0f, 0.0f, 0.0f, 0.0f};
    nb_f4v D[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) D[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, Bg[g], zero, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const float4 p = s[4 * grp + v];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const float q = __builtin_amdgcn_rsqf(D[g][v]);
            const float f = (q * q) * q;
            if (FIRST && v == 0) {
                S[g][0] = f * p.x;
                S[g][1] = f * p.y;
                S[g][2] = f * p.z;
                S[g][3] = f;
            } else {
                S[g][0] = __builtin_fmaf(f, p.x, S[g][0]);
                S[g][1] = __builtin_fmaf(f, p.y, S[g][1]);
                S[g][2] = __builtin_fmaf(f, p.z, S[g][2]);
                S[g][3] += f;
            }
        }
    }
}

// A whole tile; S is written (not accumulated into), so it lives only
// within the tile.
__device__ __forceinline__ void tile_mf3(const float4 *__restrict__ s, const nb_bf8 (*__restrict__ sa)[4],
                                         int lane, const nb_bf8 (&Bg)[4], float (&S)[4][4]) {
    const int c16 = lane & 15, grp = lane >> 4;
    mf3_chunk<true>(s, sa[c16][grp], grp, Bg, S);
#pragma unroll 1
    for (int jc = 16; jc < BS; jc += 16) mf3_chunk<false>(s + jc, sa[jc + c16][grp], grp, Bg, S);
}

// One 64-source quarter of a tile (sources 64q..64q+63), added into S.
__device__ __forceinline__ void quarter_mf3(const float4 *__restrict__ s, const nb_bf8 (*__restrict__ sa)[4],
                                            int lane, int q, const nb_bf8 (&Bg)[4], float (&S)[4][4]) {
    const int c16 = lane & 15, grp = lane >> 4;
#pragma unroll 1
    for (int jc = SB * q; jc < SB * q + SB; jc += 16) mf3_chunk<false>(s + jc, sa[jc + c16][grp], grp, Bg, S);
}

// Potential only (no accelerations): with q = rsq(w_j r^2) = m_j^(1/3) / r
// from the same product, m_j / r = q c_j where c_j = m_j^(2/3) = 1 / w_j
// (0 for massless sources) rides in the w slot of the lo coordinates: rsq and
// one FMA per pair.
template <bool FIRST>
__device__ __forceinline__ void mf3_pot_chunk(const float4 *__restrict__ sl, const nb_bf8 a, int grp,
                                              const nb_bf8 (&Bg)[4], float (&S)[4]) {
    const nb_f4v zero = {0.0f, 0.0f, 0.0f, 0.0f};
    nb_f4v D[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) D[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, Bg[g], zero, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const float c = sl[4 * grp + v].w;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const float q = __builtin_amdgcn_rsqf(D[g][v]);
            S[g] = (FIRST && v == 0) ? q * c : __builtin_fmaf(q, c, S[g]);
        }
    }
}

__device__ __forceinline__ void tile_mf3_pot(const float4 *__restrict__ sl,
                                             const nb_bf8 (*__restrict__ sa)[4], int lane,
                                             const nb_bf8 (&Bg)[4], float (&S)[4]) {
    const int c16 = lane & 15, grp = lane >> 4;
    mf3_pot_chunk<true>(sl, sa[c16][grp], grp, Bg, S);
#pragma unroll 1
    for (int jc = 16; jc < BS; jc += 16) mf3_pot_chunk<false>(sl + jc, sa[jc + c16][grp], grp, Bg, S);
}

__device__ __forceinline__ void quarter_mf3_pot(const float4 *__restrict__ sl,
                                                const nb_bf8 (*__restrict__ sa)[4], int lane, int q,
                                                const nb_bf8 (&Bg)[4], float (&S)[4]) {
    const int c16 = lane & 15, grp = lane >> 4;
#pragma unroll 1
    for (int jc = SB * q; jc < SB * q + SB; jc += 16) mf3_pot_chunk<false>(sl + jc, sa[jc + c16][grp], grp, Bg, S);
}

// FAR_MFMA_BF16 fold: target 16g + c sits in lanes c, 16 + c, 32 + c, 48 + c
// (slot g) and lane l keeps slot l >> 4, its own target.  permlane32_swap of
// slots g and g + 2 followed by an add leaves lanes 0-31 with slot g and lanes
// 32-63 with slot g + 2, each summed over lane bit 5; permlane16_swap of the
// two results does the same for lane bit 4 and slot bit 0.  The lane's four
// sums (S_xyz, S_m) go to FP64 as S_xyz - x'_i S_m.
__device__ __forceinline__ void fold_mf3(const float (&S)[4][4], double xd, double yd, double zd,
                                         double &Ax, double &Ay, double &Az) {
    float col[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float t1[2];
#pragma unroll
        for (int g0 = 0; g0 < 2; ++g0) {
            const auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(unsigned, S[g0][c]),
                                                            __builtin_bit_cast(unsigned, S[g0 + 2][c]), false,
                                                            false);
            t1[g0] = __builtin_bit_cast(float, (unsigned)r[0]) + __builtin_bit_cast(float, (unsigned)r[1]);
        }
        const auto r = __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(unsigned, t1[0]),
                                                        __builtin_bit_cast(unsigned, t1[1]), false, false);
        col[c] = __builtin_bit_cast(float, (unsigned)r[0]) + __builtin_bit_cast(float, (unsigned)r[1]);
    }
    const double sm = (double)col[3];
    Ax += (double)col[0] - xd * sm;
    Ay += (double)col[1] - yd * sm;
    Az += (double)col[2] - zd * sm;
}

// The same lane fold for one sum per target (the potential-only tiles).
__device__ __forceinline__ float fold_mf3_one(const float (&S)[4]) {
    float t1[2];
#pragma unroll
    for (int g0 = 0; g0 < 2; ++g0) {
        const auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(unsigned, S[g0]),
                                                        __builtin_bit_cast(unsigned, S[g0 + 2]), false, false);
        t1[g0] = __builtin_bit_cast(float, (unsigned)r[0]) + __builtin_bit_cast(float, (unsigned)r[1]);
    }
    const auto r = __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(unsigned, t1[0]),
                                                    __builtin_bit_cast(unsigned, t1[1]), false, false);
    return __builtin_bit_cast(float, (unsigned)r[0]) + __builtin_bit_cast(float, (unsigned)r[1]);
}

// Per-wave tile classes.  Each wave uses the box of its own targets, their
// largest h, their largest distance R from the block origin per axis and
// their largest Euclidean distance Re from it.  A margin keeps every far pair
// at FP32-resolvable, nonzero separation.  The two functions are evaluated
// without FMA contraction, so the classes do not depend on how the compiler
// schedules them (one lane per tile, 64 tiles at a time, in force_block).
struct WaveGeom {
    double lo[3], hi[3];
    double soft_d2, close_d2, farw_d2;
};

template <int TPT>
__device__ __forceinline__ WaveGeom wave_geometry(const double (&px)[TPT], const double (&py)[TPT],
                                                  const double (&pz)[TPT], const double (&ph)[TPT],
                                                  const bool (&valid)[TPT], double ox, double oy,
                                                  double oz, bool hilo) {
#pragma clang fp contract(off)
    WaveGeom w;
    double whmax = 0.0, re2 = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        w.lo[d] = DBL_MAX;
        w.hi[d] = -DBL_MAX;
    }
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        if (!valid[k]) continue;
        w.lo[0] = fmin(w.lo[0], px[k]); w.hi[0] = fmax(w.hi[0], px[k]);
        w.lo[1] = fmin(w.lo[1], py[k]); w.hi[1] = fmax(w.hi[1], py[k]);
        w.lo[2] = fmin(w.lo[2], pz[k]); w.hi[2] = fmax(w.hi[2], pz[k]);
        whmax = fmax(whmax, ph[k]);
        const double dx = px[k] - ox, dy = py[k] - oy, dz = pz[k] - oz;
        re2 = fmax(re2, dx * dx + dy * dy + dz * dz);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        w.lo[d] = wave_min(w.lo[d]);
        w.hi[d] = wave_max(w.hi[d]);
    }
    whmax = wave_max(whmax);
    const double re = sqrt(wave_max(re2));
    double rb = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double o = d == 0 ? ox : (d == 1 ? oy : oz);
        if (w.hi[d] >= w.lo[d]) rb = fmax(rb, fmax(fabs(w.lo[d] - o), fabs(w.hi[d] - o)));
    }
    const double soft_d = whmax * (1.0 + 1e-6) + 1e-6 * rb + 1e-300;
    const double close_d = hilo ? fmax(soft_d, 0.5 * rb) : soft_d;
    const double farw_d = fmax(close_d, FARW_REACH * re);
    w.soft_d2 = uniform_f64(soft_d * soft_d);
    w.close_d2 = uniform_f64(close_d * close_d);
    w.farw_d2 = uniform_f64(farw_d * farw_d);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        w.lo[d] = uniform_f64(w.lo[d]);
        w.hi[d] = uniform_f64(w.hi[d]);
    }
    return w;
}

// Squared Euclidean gap between a source tile's box (tb: lo xyz, hi xyz) and
// the wave's target box.
__device__ __forceinline__ double tile_gap2(const double *__restrict__ tb, const WaveGeom &w) {
#pragma clang fp contract(off)
    const double gx = fmax(0.0, fmax(tb[0] - w.hi[0], w.lo[0] - tb[3]));
    const double gy = fmax(0.0, fmax(tb[1] - w.hi[1], w.lo[1] - tb[4]));
    const double gz = fmax(0.0, fmax(tb[2] - w.hi[2], w.lo[2] - tb[5]));
    return gx * gx + gy * gy + gz * gz;
}

// A target block's geometry as force_block derives it (block_stats and
// wave_geometry), stored once per force call when the block's targets meet
// several source segments: the origin, then per wave lo xyz, hi xyz and the
// three squared class distances (TG_STRIDE doubles per block).
static_assert(TG_STRIDE >= 3 + 9 * NW, "target geometry record");

__device__ __forceinline__ void store_target_geom(double *__restrict__ g, double ox, double oy, double oz,
                                                  const WaveGeom &w, int wave, int lane) {
    if (lane != 0) return;
    if (wave == 0) { g[0] = ox; g[1] = oy; g[2] = oz; }
    double *gw = g + 3 + 9 * wave;
#pragma unroll
    for (int d = 0; d < 3; ++d) { gw[d] = w.lo[d]; gw[3 + d] = w.hi[d]; }
    gw[6] = w.soft_d2; gw[7] = w.close_d2; gw[8] = w.farw_d2;
}

__device__ __forceinline__ WaveGeom load_target_geom(const double *__restrict__ g, int wave) {
    const double *gw = g + 3 + 9 * wave;
    WaveGeom w;
#pragma unroll
    for (int d = 0; d < 3; ++d) { w.lo[d] = uniform_f64(gw[d]); w.hi[d] = uniform_f64(gw[3 + d]); }
    w.soft_d2 = uniform_f64(gw[6]);
    w.close_d2 = uniform_f64(gw[7]);
    w.farw_d2 = uniform_f64(gw[8]);
    return w;
}

// Target k (< TPT) of thread tid in a block whose targets start at i0: each
// wave owns 64*TPT consecutive targets, so that a wave's targets (sorted in
// Morton order) stay compact and its tile classes stay tight.
template <int TPT>
__device__ __forceinline__ int64_t target_of(int64_t i0, int k, int tid) {
    return i0 + (int64_t)(tid >> 6) * (64 * TPT) + k * 64 + (tid & 63);
}

// One block's share of an all-pairs sum: targets target_of(i0, k, tid) (k < TPT,
// i < i_end) against the sources [j0, j1).  tile_box points at the box of
// the BS-source tile that starts at j0 (tiles are BS-aligned from j0);
// sub_box, when not null, at the boxes of that tile's NQ quarters (SB sources
// each, [tile][NQ][8]).  With far path 0 (and 4 on FP32 state) a wave then
// classifies each quarter on its own, so a tile that only touches the wave's
// targets with one corner runs the cheaper arithmetic on its other quarters.
// tgeo, when not null, holds this target block's origin and wave geometry
// (TG_STRIDE doubles written by k_target_geom from the same targets), and
// tbg, when not null, its lanes' bf16 B operands (16 dwords per lane).
// Returns FP64 sums (without G) in Ax..Aq and the target validity.  With
// ACC = false (and POT) only Aq is meaningful: the kernel writes the
// potential alone, so the acceleration arithmetic is dead and removed.
// PASS_FAR (the bf16 matrix-core path, forces only) sums the far tiles
// (quarters) only and returns whether the wave left near ones out, which a
// separate pass (near_wave) adds: without the near loops beside it the far
// loop runs faster.
template <typename T, int TPT, bool POT, int MF = 0, bool ACC = true, int PASS = PASS_ALL>
__device__ __forceinline__ bool force_block(
    BlockShared<T, MF> &sh, const T *__restrict__ tx, const T *__restrict__ ty,
    const T *__restrict__ tz, const double *__restrict__ th, double th_const, int64_t i0,
    int64_t i_end, const T *__restrict__ sx, const T *__restrict__ sy, const T *__restrict__ sz,
    const T *__restrict__ sm, const float *__restrict__ sw, int64_t j0, int64_t j1, const double *__restrict__ tile_box,
    const double *__restrict__ sub_box, const double *__restrict__ tgeo, const unsigned *__restrict__ tbg,
    const uint8_t *__restrict__ skip_src, bool use_origin, double (&Ax)[TPT], double (&Ay)[TPT],
    double (&Az)[TPT], double (&Aq)[TPT], bool (&valid)[TPT]) {
    constexpr bool HILO = BlockShared<T, MF>::HILO;
    constexpr bool ABSNEAR = BlockShared<T, MF>::ABSNEAR;
    constexpr bool D64NEAR = BlockShared<T, MF>::D64NEAR;
    static_assert(!(MF == FAR_MFMA_SUM || MF == FAR_MFMA_R2) || (HILO && TPT == 1 && !POT),
                  "MFMA far tiles: SPLIT, 1 target/lane, acc only");
    static_assert(MF != FAR_WEIGHTED || !POT, "weighted far tiles: acc only");
    static_assert(MF != FAR_MFMA_BF16 || (HILO && TPT == 1 && (!POT || !ACC)),
                  "bf16 matrix-core far tiles: SPLIT, 1 target/lane, acc or potential alone");
    static_assert(ACC || POT, "nothing to compute");
    static_assert(PASS == PASS_ALL || (PASS == PASS_FAR && MF == FAR_MFMA_BF16 && ACC && !POT),
                  "far pass: bf16 far path, forces");
    const int tid = threadIdx.x;
    bool had_near = false;   // PASS_FAR: near tiles of this wave left to the near pass
    __syncthreads();  // the previous call's readers of sh are done
    double S[4] = {0.0, 0.0, 0.0, 0.0};

    double px[TPT], py[TPT], pz[TPT], ph[TPT];
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int64_t i = target_of<TPT>(i0, k, tid);
        valid[k] = i < i_end;
        px[k] = valid[k] ? (double)tx[i] : 0.0;
        py[k] = valid[k] ? (double)ty[i] : 0.0;
        pz[k] = valid[k] ? (double)tz[i] : 0.0;
        ph[k] = valid[k] ? (th ? th[i] : th_const) : 0.0;
    }
    // The block's origin and its waves' geometry: read when stored for this
    // target block (tgeo), else derived here (block-uniform, kept in scalar
    // registers).
    double ox, oy, oz;
    if (tgeo) {
        ox = uniform_f64(tgeo[0]); oy = uniform_f64(tgeo[1]); oz = uniform_f64(tgeo[2]);
    } else {
        BlockStats st;
        block_stats<TPT>(px, py, pz, ph, valid, use_origin, sh.red, st);
        ox = uniform_f64(st.ox); oy = uniform_f64(st.oy); oz = uniform_f64(st.oz);
    }

    float xi[TPT], yi[TPT], zi[TPT], xil[TPT], yil[TPT], zil[TPT];
    float xa[TPT], ya[TPT], za[TPT];   // ABSNEAR: the FP32 positions themselves
    double xd[TPT], yd[TPT], zd[TPT];  // D64NEAR: FP64 positions relative to the origin
    float h2c[TPT], h2[TPT], hinv[TPT], hinv3[TPT];
    float ax[TPT], ay[TPT], az[TPT], aq[TPT];
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const double dx = px[k] - ox, dy = py[k] - oy, dz = pz[k] - oz;
        xi[k] = (float)dx;
        yi[k] = (float)dy;
        zi[k] = (float)dz;
        xil[k] = (float)(dx - (double)xi[k]);
        yil[k] = (float)(dy - (double)yi[k]);
        zil[k] = (float)(dz - (double)zi[k]);
        xa[k] = (float)px[k];
        ya[k] = (float)py[k];
        za[k] = (float)pz[k];
        xd[k] = dx;
        yd[k] = dy;
        zd[k] = dz;
        const float hf = (float)ph[k];
        h2[k] = hf * hf;
        h2c[k] = fmaxf(h2[k], FLT_MIN);
        hinv[k] = 1.0f / hf;
        hinv3[k] = hinv[k] * hinv[k] * hinv[k];
        ax[k] = ay[k] = az[k] = aq[k] = 0.0f;
        Ax[k] = Ay[k] = Az[k] = Aq[k] = 0.0;
    }
    // Tiles are classified per wave (the branch is wave-uniform).
    // (ABSNEAR has no close class: every near tile takes the FP32 positions)
    const WaveGeom wg = tgeo ? load_target_geom(tgeo, __builtin_amdgcn_readfirstlane(tid >> 6))
                             : wave_geometry<TPT>(px, py, pz, ph, valid, ox, oy, oz, HILO && !ABSNEAR);
    const double soft_d2 = wg.soft_d2, close_d2 = wg.close_d2;
    const double close_d = sqrt(close_d2);
    // FAR_MFMA_BF16: lane l holds, for target group g, the K slice (l >> 4) of
    // the B column of target 16g + (l & 15).
    constexpr bool ROWS = BlockShared<T, MF>::ROWS;
    nb_bf8 Bg[4];
    // Classes of the quarters of tiles 16k..16k+15: bit 4 (t - 16k) + q.  Far
    // path 0, and far path 4 with FP32 state, classify quarters; the others
    // classify whole tiles (the four bits of a tile agree).  (The FP64-state
    // path-4 kernel would spill inside its far loop with the quarter code, and
    // its near tiles are few: the far reach grows with the block's extent.)
    constexpr bool QUARTERS = MF == FAR_VALU || (MF == FAR_MFMA_BF16 && ABSNEAR);
    const bool by_quarter = QUARTERS && sub_box != nullptr;
    uint64_t m_soft = 0, m_close = 0, m_far = 0;
    if constexpr (ROWS) {
        if (tbg) {   // stored for this target block with its geometry
            const nb_u4v *b = reinterpret_cast<const nb_u4v *>(tbg) + (int64_t)tid * 4;
#pragma unroll
            for (int g = 0; g < 4; ++g) Bg[g] = __builtin_bit_cast(nb_bf8, b[g]);
        } else {
            target_b_operands(px[0] - ox, py[0] - oy, pz[0] - oz, tid & 63, Bg);
        }
    }
    // weighted far tiles: per-target operands and S-form sums
    constexpr bool WFAR = BlockShared<T, MF>::WFAR && !POT;
    float wbx[TPT], wby[TPT], wbz[TPT], wci[TPT], Sx[TPT], Sy[TPT], Sz[TPT], Sm[TPT];
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const double dx = px[k] - ox, dy = py[k] - oy, dz = pz[k] - oz;
        wbx[k] = -2.0f * xi[k];
        wby[k] = -2.0f * yi[k];
        wbz[k] = -2.0f * zi[k];
        wci[k] = (float)(dx * dx + dy * dy + dz * dz);
        Sx[k] = Sy[k] = Sz[k] = Sm[k] = 0.0f;
    }
    // MF = 2: the Euclidean reach of the wave's targets and their MFMA operands
    double mf2_d2 = DBL_MAX;
    float b2x = 0.0f, b2y = 0.0f, b2z = 0.0f;
    nb_f4v ci = {0.0f, 0.0f, 0.0f, 0.0f};
    if constexpr (MF == 2) {
        const double dx = px[0] - ox, dy = py[0] - oy, dz = pz[0] - oz;
        const double re2 = valid[0] ? dx * dx + dy * dy + dz * dz : 0.0;
        const double re = sqrt(wave_max(re2));
        const double d = fmax(close_d, MF2_REACH * re);
        mf2_d2 = d * d;
        b2x = -2.0f * xi[0];
        b2y = -2.0f * yi[0];
        b2z = -2.0f * zi[0];
        const float c = (float)re2;
        ci = nb_f4v{c, c, c, c};
    }
    // Register prefetch of the first tile (with the source weight when the
    // weighted far paths are compiled in).
    constexpr bool WEIGHTS = WFAR || ROWS;
    T rx = 0, ry = 0, rz = 0, rm = 0;
    float rw = 0.0f;
    bool rv = false;
    {
        const int64_t j = j0 + tid;
        rv = j < j1 && !(skip_src && skip_src[j]);
        if (rv) {
            rx = sx[j]; ry = sy[j]; rz = sz[j]; rm = sm[j];
            if (WEIGHTS) rw = sw ? sw[j] : (rm > 0 ? (float)(1.0 / cbrt((double)rm * (double)rm)) : 0.0f);
        }
    }
    int it = 0;
    for (int64_t base = j0; base < j1; base += BS, ++it) {
        const int buf = it & 1;
        // the single-buffered bf16 rows: the last tile's readers finish first
        if constexpr (ROWS) if (it > 0) __syncthreads();
        if (rv) {
            const double dx = (double)rx - ox, dy = (double)ry - oy, dz = (double)rz - oz;
            const float hx = (float)dx, hy = (float)dy, hz = (float)dz;
            // m^(2/3) of the potential-only far tiles rides in the w slot of
            // the lo record (ABSNEAR: of the hi record)
            const float cw = (ROWS && !ACC && rw > 0.0f) ? __builtin_amdgcn_rcpf(rw) : 0.0f;
            sh.sj[buf][tid] = make_float4(hx, hy, hz, (ABSNEAR && !ACC) ? cw : (float)rm);
            if (PASS == PASS_FAR) {
                // the far tiles read sj and the rows only
            } else if (ABSNEAR)
                sh.sjl[HILO ? buf : 0][HILO ? tid : 0] = make_float4((float)rx, (float)ry, (float)rz, (float)rm);
            else if (D64NEAR)
                sh.sjd[D64NEAR ? buf : 0][D64NEAR ? tid : 0] = SrcF64{dx, dy, dz, (float)rm, 0.0f};
            else if (HILO)
                sh.sjl[HILO ? buf : 0][HILO ? tid : 0] =
                    make_float4((float)(dx - (double)hx), (float)(dy - (double)hy),
                                (float)(dz - (double)hz), cw);
            if constexpr (MF == 1 || MF == 2) {
                const double md = (double)rm;
                sh.sB[buf][0][tid] = (float)(md * dx);
                sh.sB[buf][1][tid] = (float)(md * dy);
                sh.sB[buf][2][tid] = (float)(md * dz);
                sh.sB[buf][3][tid] = (float)md;
            }
            if constexpr (MF == 2) sh.sQ[MF == 2 ? buf : 0][MF == 2 ? tid : 0] = (float)(dx * dx + dy * dy + dz * dz);
            if constexpr (ROWS) {
                if (rw > 0.0f) {
                    const double q2 = (double)hx * hx + (double)hy * hy + (double)hz * hz;
                    bf16_source_row(rw * hx, rw * hy, rw * hz, (float)((double)rw * q2), rw,
                                    sh.sA[0][ROWS ? tid : 0]);
                } else {
                    bf16_massless_row(sh.sA[0][ROWS ? tid : 0]);
                }
            }
            if constexpr (WFAR) {
                const double q2 = (double)hx * hx + (double)hy * hy + (double)hz * hz;
                float4 W;
                float w;
                if (rw > 0.0f) {
                    const double wd = (double)rw;
                    W = make_float4((float)(wd * hx), (float)(wd * hy), (float)(wd * hz), (float)(wd * q2));
                    w = rw;
                } else {
                    W = make_float4(0.0f, 0.0f, 0.0f, FARW_MASSLESS);
                    w = 0.0f;
                }
                sh.sW[WFAR ? buf : 0][WFAR ? tid : 0] = W;
                sh.sw[WFAR ? buf : 0][WFAR ? tid : 0] = w;
            }
        } else {
            sh.sj[buf][tid] = make_float4(PAD_COORD, PAD_COORD, PAD_COORD, 0.0f);
            if (PASS == PASS_FAR) {
            } else if (ABSNEAR)
                sh.sjl[HILO ? buf : 0][HILO ? tid : 0] = make_float4(PAD_COORD, PAD_COORD, PAD_COORD, 0.0f);
            else if (D64NEAR)
                sh.sjd[D64NEAR ? buf : 0][D64NEAR ? tid : 0] =
                    SrcF64{(double)PAD_COORD, (double)PAD_COORD, (double)PAD_COORD, 0.0f, 0.0f};
            else if (HILO)
                sh.sjl[HILO ? buf : 0][HILO ? tid : 0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if constexpr (MF == 1 || MF == 2) {
                sh.sB[buf][0][tid] = 0.0f;
                sh.sB[buf][1][tid] = 0.0f;
                sh.sB[buf][2][tid] = 0.0f;
                sh.sB[buf][3][tid] = 0.0f;
            }
            // |PAD|^2: r^2 ~ 3e36 keeps r^-3 at zero (and finite) for padding
            if constexpr (MF == 2) sh.sQ[MF == 2 ? buf : 0][MF == 2 ? tid : 0] = 3.0f * PAD_COORD * PAD_COORD;
            if constexpr (ROWS) bf16_massless_row(sh.sA[0][ROWS ? tid : 0]);
            if constexpr (WFAR) {
                sh.sW[WFAR ? buf : 0][WFAR ? tid : 0] = make_float4(0.0f, 0.0f, 0.0f, FARW_MASSLESS);
                sh.sw[WFAR ? buf : 0][WFAR ? tid : 0] = 0.0f;
            }
        }
        __syncthreads();
        {
            const int64_t j = base + BS + tid;
            rv = j < j1 && !(skip_src && skip_src[j]);
            if (rv) {
                rx = sx[j]; ry = sy[j]; rz = sz[j]; rm = sm[j];
                if (WEIGHTS) rw = sw ? sw[j] : (rm > 0 ? (float)(1.0 / cbrt((double)rm * (double)rm)) : 0.0f);
            }
        }
        // Quarter classes of the next 16 tiles, one per lane, as wave-wide bit
        // masks (the per-tile test is then scalar).  A quarter past the end of
        // the sources holds padding only, which every path sums to zero.
        if ((it & 15) == 0) {
            const int lane = tid & 63;
            const int64_t t = it + (lane >> 2);
            const int q = lane & 3;
            bool c_soft = false, c_close = false, c_far = true;
            if (j0 + t * BS + (by_quarter ? q * SB : 0) < j1) {
                const double g2 = tile_gap2(by_quarter ? sub_box + (t * NQ + q) * 8 : tile_box + t * 8, wg);
                c_soft = g2 <= soft_d2;
                c_close = !c_soft && g2 <= close_d2;
                c_far = g2 > wg.farw_d2;
            }
            m_soft = __builtin_amdgcn_ballot_w64(c_soft);
            m_close = __builtin_amdgcn_ballot_w64(c_close);
            m_far = __builtin_amdgcn_ballot_w64(c_far);
        }
        const int tsh = 4 * (it & 15);
        const unsigned q_soft = (unsigned)(m_soft >> tsh) & 15u;
        const unsigned q_close = (unsigned)(m_close >> tsh) & 15u;
        const unsigned q_far = (unsigned)(m_far >> tsh) & 15u;
        // far quarters with arithmetic of their own (on path 0 far is mid)
        const unsigned q_fown = ROWS ? q_far : 0u;
        if constexpr (PASS == PASS_FAR) {
            if (q_far != 15u) had_near = true;
            if (q_far == 15u || (by_quarter && q_far != 0u)) {
                float S3[4][4];
                if (q_far == 15u) {
                    tile_mf3(sh.sj[buf], sh.sA[0], tid & 63, Bg, S3);
                } else {
#pragma unroll
                    for (int g = 0; g < 4; ++g)
#pragma unroll
                        for (int c = 0; c < 4; ++c) S3[g][c] = 0.0f;
#pragma unroll 1
                    for (int q = 0; q < NQ; ++q)
                        if (q_far & (1u << q)) quarter_mf3(sh.sj[buf], sh.sA[0], tid & 63, q, Bg, S3);
                }
                fold_mf3(S3, px[0] - ox, py[0] - oy, pz[0] - oz, Ax[0], Ay[0], Az[0]);
            }
            continue;
        }
        if (by_quarter && q_soft != 15u && q_close != 15u && q_fown != 15u && (q_soft | q_close | q_fown) != 0u) {
            // a tile of mixed quarters: its far quarters first (their sums
            // folded before the near quarters' FP32 sums start)
            if constexpr (QUARTERS) {
                if constexpr (ROWS) {
                    if (q_fown != 0u) {
                        float S3[4][4], Sq[4];
#pragma unroll
                        for (int g = 0; g < 4; ++g) {
                            Sq[g] = 0.0f;
#pragma unroll
                            for (int c = 0; c < 4; ++c) S3[g][c] = 0.0f;
                        }
#pragma unroll 1
                        for (int q = 0; q < NQ; ++q) {
                            if (!(q_fown & (1u << q))) continue;
                            if constexpr (!ACC)
                                quarter_mf3_pot(ABSNEAR ? sh.sj[buf] : sh.sjl[HILO ? buf : 0], sh.sA[0], tid & 63,
                                                q, Bg, Sq);
                            else
                                quarter_mf3(sh.sj[buf], sh.sA[0], tid & 63, q, Bg, S3);
                        }
                        if constexpr (!ACC)
                            Aq[0] += (double)fold_mf3_one(Sq);
                        else
                            fold_mf3(S3, px[0] - ox, py[0] - oy, pz[0] - oz, Ax[0], Ay[0], Az[0]);
                    }
                }
#pragma unroll 1
                for (int q = 0; q < NQ; ++q) {
                    const unsigned qb = 1u << q;
                    if (q_fown & qb) continue;
                    const int o = q * SB;
                    if (q_soft & qb) {
                        if (ABSNEAR)
                            tile_f32<TPT, true, POT, SB>(sh.sjl[HILO ? buf : 0] + (HILO ? o : 0), xa, ya, za, h2c,
                                                         h2, hinv, hinv3, ax, ay, az, aq);
                        else if (D64NEAR)
                            tile_f32_d<TPT, true, POT, SB>(sh.sjd[D64NEAR ? buf : 0] + (D64NEAR ? o : 0), xd, yd,
                                                           zd, h2c, h2, hinv, hinv3, ax, ay, az, aq);
                        else if (HILO)
                            tile_f32_hilo<TPT, true, POT, SB>(sh.sj[buf] + o, sh.sjl[HILO ? buf : 0] + (HILO ? o : 0),
                                                              xi, yi, zi, xil, yil, zil, h2c, h2, hinv, hinv3,
                                                              ax, ay, az, aq);
                        else
                            tile_f32<TPT, true, POT, SB>(sh.sj[buf] + o, xi, yi, zi, h2c, h2, hinv, hinv3, ax, ay,
                                                         az, aq);
                    } else if (D64NEAR && (q_close & qb)) {
                        tile_f32_d<TPT, false, POT, SB>(sh.sjd[D64NEAR ? buf : 0] + (D64NEAR ? o : 0), xd, yd, zd,
                                                        h2c, h2, hinv, hinv3, ax, ay, az, aq);
                    } else if (HILO && !ABSNEAR && (q_close & qb)) {
                        tile_f32_hilo<TPT, false, POT, SB>(sh.sj[buf] + o, sh.sjl[HILO ? buf : 0] + (HILO ? o : 0),
                                                           xi, yi, zi, xil, yil, zil, h2c, h2, hinv, hinv3, ax,
                                                           ay, az, aq);
                    } else if (ABSNEAR) {
                        tile_f32<TPT, false, POT, SB>(sh.sjl[HILO ? buf : 0] + (HILO ? o : 0), xa, ya, za, h2c, h2,
                                                      hinv, hinv3, ax, ay, az, aq);
                    } else {
                        tile_f32<TPT, false, POT, SB>(sh.sj[buf] + o, xi, yi, zi, h2c, h2, hinv, hinv3, ax, ay,
                                                      az, aq);
                    }
                }
            }
        } else if (q_soft == 15u) {
            if (ABSNEAR)
                tile_f32<TPT, true, POT>(sh.sjl[HILO ? buf : 0], xa, ya, za, h2c, h2, hinv, hinv3, ax, ay, az,
                                         aq);
            else if (D64NEAR)
                tile_f32_d<TPT, true, POT>(sh.sjd[D64NEAR ? buf : 0], xd, yd, zd, h2c, h2, hinv, hinv3, ax, ay,
                                           az, aq);
            else if (HILO)
                tile_f32_hilo<TPT, true, POT>(sh.sj[buf], sh.sjl[HILO ? buf : 0], xi, yi, zi, xil,
                                              yil, zil, h2c, h2, hinv, hinv3, ax, ay, az, aq);
            else
                tile_f32<TPT, true, POT>(sh.sj[buf], xi, yi, zi, h2c, h2, hinv, hinv3, ax, ay, az,
                                         aq);
        } else if (D64NEAR && q_close == 15u) {
            tile_f32_d<TPT, false, POT>(sh.sjd[D64NEAR ? buf : 0], xd, yd, zd, h2c, h2, hinv, hinv3, ax, ay, az,
                                        aq);
        } else if (HILO && !ABSNEAR && q_close == 15u) {
            tile_f32_hilo<TPT, false, POT>(sh.sj[buf], sh.sjl[HILO ? buf : 0], xi, yi, zi, xil,
                                           yil, zil, h2c, h2, hinv, hinv3, ax, ay, az, aq);
        } else if (ROWS && PASS == PASS_ALL && q_far == 15u) {
            if constexpr (!ACC) {
                float Sq[4];
                tile_mf3_pot(ABSNEAR ? sh.sj[buf] : sh.sjl[HILO ? buf : 0], sh.sA[0], tid & 63, Bg, Sq);
                Aq[0] += (double)fold_mf3_one(Sq);
            } else {
                float S3[4][4];
                tile_mf3(sh.sj[buf], sh.sA[0], tid & 63, Bg, S3);
                fold_mf3(S3, px[0] - ox, py[0] - oy, pz[0] - oz, Ax[0], Ay[0], Az[0]);
            }
            continue;   // no FP32 VALU partials to flush
        } else if (WFAR && q_far == 15u) {
            tile_w<TPT>(sh.sj[buf], sh.sW[WFAR ? buf : 0], sh.sw[WFAR ? buf : 0], wbx, wby, wbz, wci,
                        Sx, Sy, Sz, Sm);
#pragma unroll
            for (int k = 0; k < TPT; ++k) {
                const double sm = (double)Sm[k];
                Ax[k] += (double)Sx[k] - (px[k] - ox) * sm;
                Ay[k] += (double)Sy[k] - (py[k] - oy) * sm;
                Az[k] += (double)Sz[k] - (pz[k] - oz) * sm;
                Sx[k] = Sy[k] = Sz[k] = Sm[k] = 0.0f;
            }
        } else if constexpr (MF == 1 || MF == 2) {
            if (MF == 2 && tile_gap2(tile_box + ((base - j0) / BS) * 8, wg) > mf2_d2)
                tile_mfma2(sh.sj[buf], &sh.sQ[MF == 2 ? buf : 0][0], &sh.sB[buf][0][0], tid & 3, b2x,
                           b2y, b2z, ci, S);
            else
                tile_mfma(sh.sj[buf], &sh.sB[buf][0][0], tid & 3, xi[0], yi[0], zi[0], S);
        } else if (ABSNEAR) {
            tile_f32<TPT, false, POT>(sh.sjl[HILO ? buf : 0], xa, ya, za, h2c, h2, hinv, hinv3, ax, ay, az, aq);
        } else {
            tile_f32<TPT, false, POT>(sh.sj[buf], xi, yi, zi, h2c, h2, hinv, hinv3, ax, ay, az, aq);
        }
#pragma unroll
        for (int k = 0; k < TPT; ++k) {
            Ax[k] += (double)ax[k]; Ay[k] += (double)ay[k]; Az[k] += (double)az[k];
            ax[k] = ay[k] = az[k] = 0.0f;
            if (POT) { Aq[k] += (double)aq[k]; aq[k] = 0.0f; }
        }
    }
    if constexpr (MF == 1 || MF == 2) {
        // Lane t = 4b + i collects target t's four columns: column c sits in
        // S[i] of lane 4b + c.
        const int lane = tid & 63, qb = lane & ~3, i = lane & 3;
        double col[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int v = 0; v < 4; ++v) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const double t = __shfl(S[v], qb + c, 64);
                if (v == i) col[c] = t;
            }
        }
        const double xd = px[0] - ox, yd = py[0] - oy, zd = pz[0] - oz;
        Ax[0] += col[0] - xd * col[3];
        Ay[0] += col[1] - yd * col[3];
        Az[0] += col[2] - zd * col[3];
    }
    return had_near;
}

// ---------------------------------------------------------------------------
// Near pass of a split bf16 launch, one wavefront at a time.  Wave `wave` of
// a force block walks the block's source segment on its own: it classifies
// tile quarters as force_block does, from the block's stored geometry (tgeo,
// as the far pass read it, so both passes draw the same line between far
// and near), and, for each quarter that is not far for it, stages the
// quarter's 64 sources into its LDS slice (one per lane) and runs the near
// arithmetic of its class over them.  A wavefront's LDS instructions
// complete in issue order, so wavefront-scope fences (code-motion barriers)
// order the staging and the reads.
// ---------------------------------------------------------------------------
template <typename T, int MF>
struct NearShared {
    static constexpr bool LO = BlockShared<T, MF>::HILO && !BlockShared<T, MF>::ABSNEAR;
    float4 wj[SB];              // hi parts (or the FP32 positions themselves)
    float4 wl[LO ? SB : 1];     // lo parts
};

template <typename T, int MF>
__device__ __forceinline__ void near_wave(NearShared<T, MF> &sh, int wave, const T *__restrict__ tx,
                                          const T *__restrict__ ty, const T *__restrict__ tz,
                                          const double *__restrict__ th, double th_const, int64_t i0,
                                          int64_t i_end, const T *__restrict__ sx, const T *__restrict__ sy,
                                          const T *__restrict__ sz, const T *__restrict__ sm, int64_t j0,
                                          int64_t j1, const double *__restrict__ tile_box,
                                          const double *__restrict__ sub_box, const double *__restrict__ tgeo,
                                          const uint8_t *__restrict__ skip_src, double &Ax, double &Ay,
                                          double &Az, bool &valid) {
    constexpr bool ABSNEAR = BlockShared<T, MF>::ABSNEAR;
    constexpr bool LO = NearShared<T, MF>::LO;
    static_assert(MF == FAR_MFMA_BF16, "near pass of the bf16 far path");
    const int lane = threadIdx.x & 63;
    double px, py, pz, ph;
    {
        const int64_t i = i0 + wave * 64 + lane;
        valid = i < i_end;
        px = valid ? (double)tx[i] : 0.0;
        py = valid ? (double)ty[i] : 0.0;
        pz = valid ? (double)tz[i] : 0.0;
        ph = valid ? (th ? th[i] : th_const) : 0.0;
    }
    const double ox = uniform_f64(tgeo[0]), oy = uniform_f64(tgeo[1]), oz = uniform_f64(tgeo[2]);
    const WaveGeom wg = load_target_geom(tgeo, wave);
    const double dx = px - ox, dy = py - oy, dz = pz - oz;
    float xi[1] = {(float)dx}, yi[1] = {(float)dy}, zi[1] = {(float)dz};
    float xil[1] = {(float)(dx - (double)xi[0])}, yil[1] = {(float)(dy - (double)yi[0])},
          zil[1] = {(float)(dz - (double)zi[0])};
    float xa[1] = {(float)px}, ya[1] = {(float)py}, za[1] = {(float)pz};
    const float hf = (float)ph;
    float h2[1] = {hf * hf}, h2c[1] = {fmaxf(hf * hf, FLT_MIN)}, hinv[1] = {1.0f / hf};
    float hinv3[1] = {hinv[0] * hinv[0] * hinv[0]};
    float ax[1] = {0.0f}, ay[1] = {0.0f}, az[1] = {0.0f}, aq[1] = {0.0f};
    Ax = Ay = Az = 0.0;
    constexpr bool QUARTERS = ABSNEAR;
    const bool by_quarter = QUARTERS && sub_box != nullptr;
    float4 *wj = sh.wj;
    float4 *wl = sh.wl;
    // The classes of the 16 tiles of window `win` (4 quarter bits per tile).
    uint64_t m_soft = 0, m_close = 0, m_far = ~0ull;
    int win = -1;
    // Steps (t, q) on to the wave's next near quarter; false past the segment.
    auto advance = [&](int &t, int &q) -> bool {
        for (;;) {
            if (++q == NQ) { q = 0; ++t; }
            if (j0 + (int64_t)t * BS >= j1) return false;
            if ((t >> 4) != win) {
                win = t >> 4;
                const int64_t u = 16 * (int64_t)win + (lane >> 2);
                const int lq = lane & 3;
                bool c_soft = false, c_close = false, c_far = true;
                if (j0 + u * BS + (by_quarter ? lq * SB : 0) < j1) {
                    const double g2 = tile_gap2(by_quarter ? sub_box + (u * NQ + lq) * 8 : tile_box + u * 8, wg);
                    c_soft = g2 <= wg.soft_d2;
                    c_close = !c_soft && g2 <= wg.close_d2;
                    c_far = g2 > wg.farw_d2;
                }
                m_soft = __builtin_amdgcn_ballot_w64(c_soft);
                m_close = __builtin_amdgcn_ballot_w64(c_close);
                m_far = __builtin_amdgcn_ballot_w64(c_far);
            }
            const unsigned q_far = (unsigned)(m_far >> (4 * (t & 15))) & 15u;
            if (q_far == 15u) { q = NQ - 1; continue; }   // the whole tile is far
            if (!((q_far >> q) & 1u)) return true;
        }
    };
    // The next quarter's sources (this lane's one) are loaded while the
    // current quarter is computed: index clamped into the segment, validity
    // decided at staging.
    T rx, ry, rz, rm;
    int64_t rj = 0;
    unsigned rskip = 0;   // the byte itself: tested at staging, so no wait here
    auto fetch = [&](int t, int q) {
        rj = j0 + (int64_t)t * BS + q * SB + lane;
        const int64_t jc = rj < j1 ? rj : j1 - 1;
        rx = sx[jc]; ry = sy[jc]; rz = sz[jc]; rm = sm[jc];
        if (skip_src) rskip = skip_src[jc];
    };
    int t = 0, q = -1;
    bool have = advance(t, q);
    if (have) fetch(t, q);
    while (have) {
        {   // stage the quarter fetched last (the previous quarter's reads are done)
            const bool rv = rj < j1 && rskip == 0;
            if (ABSNEAR) {
                wj[lane] = rv ? make_float4((float)rx, (float)ry, (float)rz, (float)rm)
                              : make_float4(PAD_COORD, PAD_COORD, PAD_COORD, 0.0f);
            } else {
                if (rv) {
                    const double sdx = (double)rx - ox, sdy = (double)ry - oy, sdz = (double)rz - oz;
                    const float hx = (float)sdx, hy = (float)sdy, hz = (float)sdz;
                    wj[lane] = make_float4(hx, hy, hz, (float)rm);
                    if (LO)
                        wl[lane] = make_float4((float)(sdx - (double)hx), (float)(sdy - (double)hy),
                                               (float)(sdz - (double)hz), 0.0f);
                } else {
                    wj[lane] = make_float4(PAD_COORD, PAD_COORD, PAD_COORD, 0.0f);
                    if (LO) wl[lane] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const int ct = t;
        const int bit = 4 * (t & 15) + q;
        const bool is_soft = (m_soft >> bit) & 1u, is_close = (m_close >> bit) & 1u;
        have = advance(t, q);
        if (have) fetch(t, q);
        if (is_soft) {
            if (ABSNEAR)
                tile_f32<1, true, false, SB>(wj, xa, ya, za, h2c, h2, hinv, hinv3, ax, ay, az, aq);
            else
                tile_f32_hilo<1, true, false, SB>(wj, wl, xi, yi, zi, xil, yil, zil, h2c, h2, hinv, hinv3, ax, ay,
                                                  az, aq);
        } else if (LO && is_close) {
            tile_f32_hilo<1, false, false, SB>(wj, wl, xi, yi, zi, xil, yil, zil, h2c, h2, hinv, hinv3, ax, ay, az,
                                               aq);
        } else if (ABSNEAR) {
            tile_f32<1, false, false, SB>(wj, xa, ya, za, h2c, h2, hinv, hinv3, ax, ay, az, aq);
        } else {
            tile_f32<1, false, false, SB>(wj, xi, yi, zi, h2c, h2, hinv, hinv3, ax, ay, az, aq);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");   // reads done before the next staging
        if (!have || t != ct) {   // a tile's quarters are summed in FP32, tiles in FP64
            Ax += (double)ax[0]; Ay += (double)ay[0]; Az += (double)az[0];
            ax[0] = ay[0] = az[0] = 0.0f;
        }
    }
}

}  // namespace nb
