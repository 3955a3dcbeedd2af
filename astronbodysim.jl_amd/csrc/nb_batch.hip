// nb_batch.hip -- many independent systems of uneven size on one GPU.
//
// Two execution paths, both following the one-system semantics per system
// (src/directsum/gravity.jl:8-28 pair law with the system's softening and G;
// constant-dt KDK as src/directsumgpu/timestep.jl:156-224):
//
//  * ragged path: one persistent launch of the all-pairs kernel serves every
//    system.  The work items are (system, tile of 256*TPT targets), ordered by
//    cost (largest systems first) and pulled from an atomic queue; each item
//    is the same block routine as the one-system kernel (force_block), with
//    the system's own source range, per-system Morton order, FP64 origin and
//    hi/lo near tiles.  KDK steps are captured in HIP graphs.
//  * resident path (small systems): the particles stay in registers and many
//    KDK steps run in a single launch, the sources broadcast through LDS;
//    central bodies use FP64 pair terms, the other massive bodies FP32 pair
//    terms on separations taken from FP64 positions.  Several systems share a
//    workgroup (k_batch_resident_packed), or one system takes one to four
//    wavefronts (k_batch_resident).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <rocprim/device/device_radix_sort.hpp>
#include <string>
#include <vector>

#include "nb_batch.h"
#include "nb_block.h"

namespace nb {

static int launch_ok(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of %s failed: %s", what, hipGetErrorString(e));
        return NB_ERR_HIP;
    }
    return NB_OK;
}

static inline unsigned blocks_for(int64_t n, int bs = 256) {
    return (unsigned)((n + bs - 1) / bs);
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BS) void k_batch_tile_boxes(const double *__restrict__ x,
                                                        const double *__restrict__ y,
                                                        const double *__restrict__ z,
                                                        const int64_t *__restrict__ off,
                                                        const int64_t *__restrict__ tile_base,
                                                        const int32_t *__restrict__ tile_sys,
                                                        double *__restrict__ box, double *__restrict__ sub) {
    __shared__ double red[6][NW];
    const int64_t t = blockIdx.x;
    const int s = tile_sys[t];
    const int64_t start = off[s] + (t - tile_base[s]) * BS;
    const int64_t end = off[s + 1];
    const int64_t j = start + threadIdx.x;
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    if (j < end) {
        lo[0] = hi[0] = x[j];
        lo[1] = hi[1] = y[j];
        lo[2] = hi[2] = z[j];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        lo[d] = wave_min(lo[d]);
        hi[d] = wave_max(hi[d]);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            red[d][w] = lo[d]; red[3 + d][w] = hi[d];
            sub[(t * NQ + w) * 8 + d] = lo[d];   // the quarter's box (empty: lo > hi)
            sub[(t * NQ + w) * 8 + 3 + d] = hi[d];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int q = threadIdx.x;
        double r = red[q][0];
        for (int ww = 1; ww < NW; ++ww) r = q < 3 ? fmin(r, red[q][ww]) : fmax(r, red[q][ww]);
        box[t * 8 + q] = r;
    }
}

struct BatchArgs {
    const int64_t *off;
    const int64_t *tile_base;
    const double *tile_box;
    const double *sub_box;        // [ntiles][NQ][8]: boxes of the tile quarters
    const double *Gs;
    const double *x, *y, *z, *m, *h;
    const float *w;               // source weights (weighted far paths) or nullptr
    const uint8_t *skip_src;
    double *ax, *ay, *az, *pot;
    const int64_t *items;
    int64_t nitems;
    unsigned *counter;
};

// Persistent ragged all-pairs kernel: blocks pull (system, target tile) items
// from an atomic queue until it is empty; the last block to finish resets the
// queue for the next launch (so the launch can be replayed from a graph).
// The bf16 matrix-core far path is held to 4 waves per SIMD, as in the
// one-system kernel: at the 2 waves its 169 VGPRs would give, an ensemble of
// 8 Morton-sorted 131072-particle Plummer spheres took 48.4 ms per force;
// at 4 waves (a few values spilled outside the pair loops) 43.2 ms, against
// 44.3 ms for far path 0.
template <int TPT, bool POT, int MF = 0, bool ACC = true>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(MF == 4 ? 4 : 1))) void k_batch_force(
    BatchArgs a) {
    __shared__ BlockShared<double, MF> sh;
    __shared__ int64_t item_sh;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) {
            const unsigned q = atomicAdd(a.counter, 1u);
            item_sh = (int64_t)q < a.nitems ? a.items[q] : -1;
        }
        __syncthreads();
        const int64_t item = item_sh;
        if (item < 0) break;
        const int s = (int)(item >> 32);
        const int64_t k = item & 0xffffffffll;
        const int64_t o0 = a.off[s], o1 = a.off[s + 1];
        const int64_t i0 = o0 + k * (BS * TPT);
        double Ax[TPT], Ay[TPT], Az[TPT], Aq[TPT];
        bool valid[TPT];
        // force_block starts with a barrier, so item_sh is read by every
        // thread before thread 0 overwrites it in the next iteration.
        force_block<double, TPT, POT, MF, ACC>(sh, a.x, a.y, a.z, a.h, 0.0, i0, o1, a.x, a.y, a.z, a.m, a.w, o0,
                                      o1, a.tile_box + a.tile_base[s] * 8,
                                      a.sub_box + a.tile_base[s] * (NQ * 8), nullptr, nullptr, a.skip_src, true,
                                      Ax,
                                      Ay, Az, Aq, valid);
        const double G = a.Gs[s];
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            if (!valid[q]) continue;
            const int64_t i = target_of<TPT>(i0, q, tid);
            if (ACC) {
                a.ax[i] = G * Ax[q];
                a.ay[i] = G * Ay[q];
                a.az[i] = G * Az[q];
            }
            if (POT) a.pot[i] = -G * Aq[q];
        }
    }
    if (tid == 0) {
        __threadfence();
        const unsigned done = atomicAdd(a.counter + 1, 1u);
        if (done == gridDim.x - 1) {
            atomicExch(a.counter, 0u);
            atomicExch(a.counter + 1, 0u);
        }
    }
}

// FP64 pair arithmetic, one item per block (correctness mode).
template <bool POT>
__global__ __launch_bounds__(BS) void k_batch_force_f64(BatchArgs a) {
    __shared__ double4 sj[BS];
    const int64_t item = a.items[blockIdx.x];
    const int s = (int)(item >> 32);
    const int64_t k = item & 0xffffffffll;
    const int64_t o0 = a.off[s], o1 = a.off[s + 1];
    const int64_t i = o0 + k * BS + threadIdx.x;
    const bool valid = i < o1;
    const double xi = valid ? a.x[i] : 0.0, yi = valid ? a.y[i] : 0.0, zi = valid ? a.z[i] : 0.0;
    const double h = valid ? a.h[i] : 1.0;
    const double hinv = 1.0 / h, hinv3 = hinv * hinv * hinv;
    double Ax = 0.0, Ay = 0.0, Az = 0.0, Aq = 0.0;
    for (int64_t base = o0; base < o1; base += BS) {
        const int64_t j = base + threadIdx.x;
        __syncthreads();
        sj[threadIdx.x] = j < o1 ? make_double4(a.x[j], a.y[j], a.z[j], a.m[j])
                                 : make_double4(1e150, 1e150, 1e150, 0.0);
        __syncthreads();
        const int cnt = (o1 - base < BS) ? (int)(o1 - base) : BS;
        for (int jj = 0; jj < cnt; ++jj) {
            const double4 v = sj[jj];
            pair_f64<POT>(v.x, v.y, v.z, v.w, xi, yi, zi, h, hinv, hinv3, Ax, Ay, Az, Aq);
        }
    }
    if (!valid) return;
    const double G = a.Gs[s];
    a.ax[i] = G * Ax;
    a.ay[i] = G * Ay;
    a.az[i] = G * Az;
    if (POT) a.pot[i] = -G * Aq;
}

// FP64 terms of the central bodies of each system, added to acc (and pot).
// With POT, a central body's own potential is summed over every source of its
// system in FP64 (not only over the central ones), so that energies carry the
// central bodies' terms in FP64 whichever side of the pair they are on.
template <bool POT, bool ACC = true>
__global__ void k_batch_central(const int32_t *__restrict__ sys_of, const int64_t *__restrict__ cidx,
                                int ncentral, const double *__restrict__ Gs,
                                const double *__restrict__ x, const double *__restrict__ y,
                                const double *__restrict__ z, const double *__restrict__ m,
                                const double *__restrict__ h, int64_t n, double *__restrict__ ax,
                                double *__restrict__ ay, double *__restrict__ az,
                                double *__restrict__ pot, const int64_t *__restrict__ off,
                                const uint8_t *__restrict__ central) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = sys_of[i];
    const double xi = x[i], yi = y[i], zi = z[i], hi = h[i];
    const double hinv = 1.0 / hi, hinv3 = hinv * hinv * hinv;
    double Ax = 0.0, Ay = 0.0, Az = 0.0, Aq = 0.0;
    if (POT && central[i]) {
        double bx = 0.0, by = 0.0, bz = 0.0;
        for (int64_t j = off[s]; j < off[s + 1]; ++j) {
            if (j == i) continue;
            pair_f64<true>(x[j], y[j], z[j], m[j], xi, yi, zi, hi, hinv, hinv3, bx, by, bz, Aq);
        }
        pot[i] = -Gs[s] * Aq;
        Aq = 0.0;
        if (!ACC) return;
    }
    for (int c = 0; c < ncentral; ++c) {
        const int64_t j = cidx[(int64_t)s * ncentral + c];
        if (j == i) continue;
        pair_f64<POT>(x[j], y[j], z[j], m[j], xi, yi, zi, hi, hinv, hinv3, Ax, Ay, Az, Aq);
    }
    const double G = Gs[s];
    if (ACC) {
        ax[i] += G * Ax;
        ay[i] += G * Ay;
        az[i] += G * Az;
    }
    if (POT && !central[i]) pot[i] -= G * Aq;
}

// v += a * coef * dt_s ; optionally then x += v * dt_s (the next step's drift)
__global__ void k_batch_kick_drift(double *__restrict__ x, double *__restrict__ y,
                                   double *__restrict__ z, double *__restrict__ vx,
                                   double *__restrict__ vy, double *__restrict__ vz,
                                   const double *__restrict__ ax, const double *__restrict__ ay,
                                   const double *__restrict__ az,
                                   const int32_t *__restrict__ sys_of,
                                   const double *__restrict__ dts, int64_t n, double kick_coef,
                                   int do_drift) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double dt = dts[sys_of[i]];
    double VX = vx[i], VY = vy[i], VZ = vz[i];
    if (kick_coef != 0.0) {
        const double c = kick_coef * dt;
        VX += ax[i] * c;
        VY += ay[i] * c;
        VZ += az[i] * c;
        vx[i] = VX;
        vy[i] = VY;
        vz[i] = VZ;
    }
    if (do_drift) {
        x[i] += VX * dt;
        y[i] += VY * dt;
        z[i] += VZ * dt;
    }
}

// Per-system FP64 reductions, one block per system: K, sum m phi, P, L, and
// (sync != 0) the kinetic energy of velocities synchronised to the step,
// v - a dt/2, in slot 0 instead.
__global__ __launch_bounds__(256) void k_batch_energy(
    const int64_t *__restrict__ off, const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, const double *__restrict__ vx, const double *__restrict__ vy,
    const double *__restrict__ vz, const double *__restrict__ ax, const double *__restrict__ ay,
    const double *__restrict__ az, const double *__restrict__ m, const double *__restrict__ pot,
    const double *__restrict__ dts, int sync, double *__restrict__ out) {
    __shared__ double red[8][4];
    const int s = blockIdx.x;
    const int64_t o0 = off[s], o1 = off[s + 1];
    const double half = sync ? 0.5 * dts[s] : 0.0;
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = o0 + threadIdx.x; i < o1; i += 256) {
        const double mi = m[i];
        const double VX = vx[i] - half * ax[i], VY = vy[i] - half * ay[i], VZ = vz[i] - half * az[i];
        v[0] += 0.5 * mi * (VX * VX + VY * VY + VZ * VZ);
        if (pot) v[1] += pot[i] * mi;
        v[2] += VX * mi;
        v[3] += VY * mi;
        v[4] += VZ * mi;
        v[5] += mi * (y[i] * VZ - z[i] * VY);
        v[6] += mi * (z[i] * VX - x[i] * VZ);
        v[7] += mi * (x[i] * VY - y[i] * VX);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = wave_sum(v[q]);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) red[q][w] = v[q];
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        const int q = threadIdx.x;
        out[(int64_t)s * 8 + q] = red[q][0] + red[q][1] + red[q][2] + red[q][3];
    }
}

// ---- resident multi-step kernel (small systems, many steps per launch) ----
struct ResArgs {
    const int64_t *off;
    const double *Gs, *hs, *dts;
    double *x, *y, *z, *vx, *vy, *vz;
    const double *m;
    double *ax, *ay, *az;
    int ncentral;
    double *energy;      // [nsys][slots] or nullptr
    int energy_every;
    int energy_slots;
};

__device__ __forceinline__ float readlane_f32(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ double readlane_f64(double v, int l) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// 1/sqrt(x) in FP64 for x > 0: v_rsq_f64 (relative error up to about 2^-23)
// refined by two Newton steps y <- y + y (1 - x y^2) / 2, each of which
// squares the relative error (times 3/2), so the result is within a few ulp.
// Replaces the sqrt and divide of the reference's m / (r^2 r) for the
// central-body terms.
__device__ __forceinline__ double rsq_f64(double x) {
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double e = __builtin_fma(-x * y, y, 1.0);
        y = __builtin_fma(0.5 * y, e, y);
    }
    return y;
}

// FP32 pair of the resident kernel from FP64 positions: source (sx, sy, sz,
// sm), target (x, y, z); the separation is fl32 of the FP64 difference
// (pair_f32_d).  SOFT: the spline inside h.  Otherwise Newtonian, with
// r^2 + tiny in place of r^2 (tiny from resident_tiny): the self pair (r = 0)
// then has a finite factor times dx = 0, so it contributes nothing without a
// select, and every pair with r^2 above about 2^24 tiny keeps its FP32 r^2.
template <bool SOFT>
__device__ __forceinline__ void res_pair_f32(double sx, double sy, double sz, float smf, double x, double y,
                                             double z, float h2cf, float h2f, float hinvf, float hinv3f,
                                             float tiny, float &fx, float &fy, float &fz, float &fq) {
    if constexpr (SOFT) {
        pair_f32_d<true, false>(sx, sy, sz, smf, x, y, z, h2cf, h2f, hinvf, hinv3f, fx, fy, fz, fq);
    } else {
        const float dx = (float)(sx - x);
        const float dy = (float)(sy - y);
        const float dz = (float)(sz - z);
        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, tiny)));
        const float ri = __builtin_amdgcn_rsqf(r2);
        const float fac = smf * (ri * ri) * ri;
        fx = __builtin_fmaf(fac, dx, fx);
        fy = __builtin_fmaf(fac, dy, fy);
        fz = __builtin_fmaf(fac, dz, fz);
    }
}

// The r^2 offsets of the unsoftened resident pairs: tiny = (4 max(m_max, 1) /
// MAX)^(2/3) makes m r^-3 at r^2 = tiny at most MAX / 4 for every source mass
// m <= m_max (MAX = FLT_MAX or DBL_MAX), so a zero separation gives a finite
// factor times zero.  In FP32 it is at least 5.2e-26, in FP64 at least 1e-205.
// (The cube root comes first: (4 m / DBL_MAX)^2 itself underflows.)
__device__ __forceinline__ void resident_tiny(double m_max, float &tiny32, double &tiny64) {
    const double m = fmax(m_max, 1.0);
    const double c32 = cbrt(4.0 * m / (double)FLT_MAX), c64 = cbrt(4.0 * m / DBL_MAX);
    tiny32 = (float)(c32 * c32);
    tiny64 = c64 * c64;
}

// State of one resident wave: position p = lane + 64 k holds one particle
// (see k_batch_resident), the massive ones at positions 0 .. nm - 1.
template <int NSLOT>
struct ResState {
    double X[NSLOT], Y[NSLOT], Z[NSLOT], VX[NSLOT], VY[NSLOT], VZ[NSLOT], AX[NSLOT], AY[NSLOT],
        AZ[NSLOT], M[NSLOT];
    bool act[NSLOT];
    uint64_t mmask[NSLOT];   // massive particles (sources)
    int nm, nc;              // massive particles, of which the first nc are central (FP64 terms)
    float tiny32;            // r^2 offsets of the unsoftened pairs (resident_tiny)
    double tiny64;
};

// One source of a resident wave in LDS, 32 B: FP64 (x, y, z, m) for a central
// body; FP64 (x, y, z) and the FP32 mass for any other (src_f32).
struct ResSrc {
    float4 a, b;
};

__device__ __forceinline__ float4 f64x2_bits(double u, double v) {
    const unsigned long long bu = __builtin_bit_cast(unsigned long long, u);
    const unsigned long long bv = __builtin_bit_cast(unsigned long long, v);
    return make_float4(__builtin_bit_cast(float, (unsigned)bu), __builtin_bit_cast(float, (unsigned)(bu >> 32)),
                       __builtin_bit_cast(float, (unsigned)bv), __builtin_bit_cast(float, (unsigned)(bv >> 32)));
}
__device__ __forceinline__ void f64x2_of(float4 f, double &u, double &v) {
    u = __builtin_bit_cast(double, ((unsigned long long)__builtin_bit_cast(unsigned, f.y) << 32) |
                                       __builtin_bit_cast(unsigned, f.x));
    v = __builtin_bit_cast(double, ((unsigned long long)__builtin_bit_cast(unsigned, f.w) << 32) |
                                       __builtin_bit_cast(unsigned, f.z));
}
// The record of a source with FP32 pair terms, and its reading.
__device__ __forceinline__ ResSrc src_f32(double x, double y, double z, double m) {
    const float4 zb = f64x2_bits(z, 0.0);
    return ResSrc{f64x2_bits(x, y), make_float4(zb.x, zb.y, (float)m, 0.0f)};
}
__device__ __forceinline__ void src_f32_of(const ResSrc &r, double &x, double &y, double &z, float &m) {
    double unused;
    f64x2_of(r.a, x, y);
    f64x2_of(make_float4(r.b.x, r.b.y, 0.0f, 0.0f), z, unused);
    m = r.b.z;
}

// nsteps KDK steps of one wave's particles; SOFT (h > 0) is wave-uniform and
// chosen outside the step loop, so the source loops carry no branches.  Each
// step the massive lanes publish their source records to the wave's own LDS
// array src (positions 0 .. nm - 1), which every lane then reads as
// broadcasts.  A wavefront's LDS instructions complete in issue order, so the
// wavefront-scope fences (code-motion barriers here) are all the ordering
// the write-then-read and read-then-next-write sequences need.
template <int NSLOT, bool SOFT>
__device__ __forceinline__ void resident_run(const ResArgs &a, ResState<NSLOT> &r, ResSrc *src, int s,
                                             int lane, bool energy_wave, int nsteps, double G, double dt,
                                             double hsys) {
    const double hinv = SOFT ? 1.0 / hsys : 0.0, hinv3 = hinv * hinv * hinv;
    const float hf = (float)hsys, h2f = hf * hf, h2cf = fmaxf(h2f, FLT_MIN);
    const float hinvf = SOFT ? 1.0f / hf : 0.0f, hinv3f = hinvf * hinvf * hinvf;
    const int nm = r.nm, nc = r.nc;
    const float tiny32 = r.tiny32;
    const double tiny64 = r.tiny64;
    int e_slot = 0;
    for (int step = 0; step < nsteps; ++step) {
        // drift (FP64)
#pragma unroll
        for (int k = 0; k < NSLOT; ++k) {
            r.X[k] += r.VX[k] * dt; r.Y[k] += r.VY[k] * dt; r.Z[k] += r.VZ[k] * dt;
        }
        float fx[NSLOT], fy[NSLOT], fz[NSLOT], fq[NSLOT];
        double Ax[NSLOT], Ay[NSLOT], Az[NSLOT], Aq[NSLOT];
#pragma unroll
        for (int k = 0; k < NSLOT; ++k) {
            fx[k] = fy[k] = fz[k] = fq[k] = 0.0f;
            Ax[k] = Ay[k] = Az[k] = Aq[k] = 0.0;
            const int p = lane + 64 * k;
            if (p < nc) {
                src[p].a = f64x2_bits(r.X[k], r.Y[k]);
                src[p].b = f64x2_bits(r.Z[k], r.M[k]);
            } else if (p < nm) {
                src[p] = src_f32(r.X[k], r.Y[k], r.Z[k], r.M[k]);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        // central bodies: FP64 terms
        for (int j = 0; j < nc; ++j) {
            double sx, sy, sz, sm;
            f64x2_of(src[j].a, sx, sy);
            f64x2_of(src[j].b, sz, sm);
#pragma unroll
            for (int kt = 0; kt < NSLOT; ++kt) {
                if constexpr (SOFT) {
                    pair_f64<false>(sx, sy, sz, sm, r.X[kt], r.Y[kt], r.Z[kt], hsys, hinv, hinv3, Ax[kt],
                                    Ay[kt], Az[kt], Aq[kt]);
                } else {
                    const double dx = sx - r.X[kt], dy = sy - r.Y[kt], dz = sz - r.Z[kt];
                    const double r2 = dx * dx + dy * dy + (dz * dz + tiny64);
                    const double ri = rsq_f64(r2);
                    const double fac = sm * (ri * ri) * ri;
                    Ax[kt] += dx * fac; Ay[kt] += dy * fac; Az[kt] += dz * fac;
                }
            }
        }
        // the other massive bodies: FP32 terms, two sources per trip
        // (position nm holds a massless record for an odd count)
        for (int j = nc; j < nm; j += 2) {
            double ax_, ay_, az_, bx_, by_, bz_;
            float am, bm;
            src_f32_of(src[j], ax_, ay_, az_, am);
            src_f32_of(src[j + 1], bx_, by_, bz_, bm);
#pragma unroll
            for (int kt = 0; kt < NSLOT; ++kt) {
                res_pair_f32<SOFT>(ax_, ay_, az_, am, r.X[kt], r.Y[kt], r.Z[kt], h2cf, h2f, hinvf, hinv3f, tiny32,
                                   fx[kt], fy[kt], fz[kt], fq[kt]);
                res_pair_f32<SOFT>(bx_, by_, bz_, bm, r.X[kt], r.Y[kt], r.Z[kt], h2cf, h2f, hinvf, hinv3f, tiny32,
                                   fx[kt], fy[kt], fz[kt], fq[kt]);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        // kick (FP64)
#pragma unroll
        for (int k = 0; k < NSLOT; ++k) {
            r.AX[k] = G * (Ax[k] + (double)fx[k]);
            r.AY[k] = G * (Ay[k] + (double)fy[k]);
            r.AZ[k] = G * (Az[k] + (double)fz[k]);
            r.VX[k] += r.AX[k] * dt; r.VY[k] += r.AY[k] * dt; r.VZ[k] += r.AZ[k] * dt;
        }
        if (a.energy && energy_wave && (step + 1) % a.energy_every == 0) {
            // total energy with velocities synchronised to the step; the
            // potential in FP64 over the massive sources (the reference's
            // pair potential, src/directsum/potential.jl:2-27)
            double q[NSLOT];
#pragma unroll
            for (int k = 0; k < NSLOT; ++k) q[k] = 0.0;
#pragma unroll
            for (int ks = 0; ks < NSLOT; ++ks) {
                uint64_t mm = r.mmask[ks];
                while (mm) {
                    const int j = (int)__builtin_ctzll(mm);
                    mm &= mm - 1;
                    const double sx = readlane_f64(r.X[ks], j), sy = readlane_f64(r.Y[ks], j);
                    const double sz = readlane_f64(r.Z[ks], j), sm = readlane_f64(r.M[ks], j);
#pragma unroll
                    for (int kt = 0; kt < NSLOT; ++kt) {
                        double bx = 0, by = 0, bz = 0;
                        if (!(ks == kt && lane == j))
                            pair_f64<true>(sx, sy, sz, sm, r.X[kt], r.Y[kt], r.Z[kt], hsys, SOFT ? hinv : 0.0,
                                           SOFT ? hinv3 : 0.0, bx, by, bz, q[kt]);
                    }
                }
            }
            double e = 0.0;
#pragma unroll
            for (int k = 0; k < NSLOT; ++k) {
                if (!r.act[k]) continue;
                const double ux = r.VX[k] - 0.5 * r.AX[k] * dt, uy = r.VY[k] - 0.5 * r.AY[k] * dt,
                             uz = r.VZ[k] - 0.5 * r.AZ[k] * dt;
                e += 0.5 * r.M[k] * (ux * ux + uy * uy + uz * uz) - 0.5 * G * r.M[k] * q[k];
            }
            e = wave_sum(e);
            if (lane == 0 && e_slot < a.energy_slots)
                a.energy[(int64_t)s * a.energy_slots + e_slot] = e;
            ++e_slot;
        }
    }
}

// W wavefronts per system (block of 64 W threads, W <= 4), each holding the
// system's massive particles and a 1/W share of its massless ones in
// registers for the whole launch: position p = lane + 64 k (k < NSLOT) of
// wave w holds the p-th massive particle (p < nM, the same in every wave) or
// massless particle w * share + p - nM (share = ceil(nT / W)), particles
// counted in the caller's order.  Massless particles are no sources, so each
// wave integrates its own copy of the massive ones (identical instructions on
// identical data, hence identical values) and the waves never exchange data.
// Each massive particle is a source, published to LDS each step.  Pair terms
// of the central bodies (the first ncentral particles; ALL_F64: of every
// source) are FP64; the others are FP32 on the fl32 of the FP64 separation.
// Drift and kick are FP64.  Every `energy_every` steps wave 0, which holds every massive
// particle, writes the system's total energy (velocities synchronised to the
// step, FP64 potential over the massive sources; massless particles add
// nothing) to energy[s * energy_slots + k].
// The host picks NSLOT and W with nM + share <= 64 NSLOT for every system.
template <int NSLOT, bool ALL_F64>
__global__ __launch_bounds__(256) void k_batch_resident(ResArgs a, int nsteps) {
    const int s = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, W = blockDim.x >> 6;
    const int64_t o0 = a.off[s];
    const int ns = (int)(a.off[s + 1] - o0);
    if (ns <= 0) {   // an empty system has zero energy
        if (a.energy && wave == 0)
            for (int k = lane; k < a.energy_slots; k += 64) a.energy[(int64_t)s * a.energy_slots + k] = 0.0;
        return;
    }
    // particle of each (wave, position), -1 for none (systems of up to 1024
    // particles)
    __shared__ int16_t held[4][64 * NSLOT];
    if (wave == 0) {
#pragma unroll
        for (int w = 0; w < 4; ++w)
#pragma unroll
            for (int k = 0; k < NSLOT; ++k) held[w][lane + 64 * k] = -1;
        int nm = 0;
        for (int base = 0; base < ns; base += 64)
            nm += __builtin_popcountll(__builtin_amdgcn_ballot_w64(base + lane < ns && a.m[o0 + base + lane] != 0.0));
        const int share = (ns - nm + W - 1) / W;
        const uint64_t below = (1ull << lane) - 1ull;
        int before = 0;   // massive particles before this chunk of 64
        for (int base = 0; base < ns; base += 64) {
            const int p = base + lane;
            const bool massive = p < ns && a.m[o0 + p] != 0.0;
            const uint64_t mk = __builtin_amdgcn_ballot_w64(massive);
            const int rm = before + __builtin_popcountll(mk & below);
            before += __builtin_popcountll(mk);
            if (p >= ns) continue;
            if (massive) {
                if (rm < 64 * NSLOT)
                    for (int w = 0; w < W; ++w) held[w][rm] = (int16_t)p;
            } else {
                const int rt = p - rm, w = rt / share, pos = nm + rt % share;
                if (pos < 64 * NSLOT) held[w][pos] = (int16_t)p;
            }
        }
    }
    __syncthreads();
    ResState<NSLOT> r;
    int pid[NSLOT];
    const int ncentral = ALL_F64 ? 1 << 30 : a.ncentral;
    r.nm = r.nc = 0;
#pragma unroll
    for (int k = 0; k < NSLOT; ++k) {
        const int p = held[wave][lane + 64 * k];
        pid[k] = p;
        const bool act = p >= 0;
        r.act[k] = act;
        const int64_t i = o0 + (act ? p : 0);
        r.X[k] = act ? a.x[i] : 0.0; r.Y[k] = act ? a.y[i] : 0.0; r.Z[k] = act ? a.z[i] : 0.0;
        r.VX[k] = act ? a.vx[i] : 0.0; r.VY[k] = act ? a.vy[i] : 0.0; r.VZ[k] = act ? a.vz[i] : 0.0;
        r.AX[k] = act ? a.ax[i] : 0.0; r.AY[k] = act ? a.ay[i] : 0.0; r.AZ[k] = act ? a.az[i] : 0.0;
        r.M[k] = act ? a.m[i] : 0.0;
        r.mmask[k] = __builtin_amdgcn_ballot_w64(act && r.M[k] != 0.0);
        r.nm += __builtin_popcountll(r.mmask[k]);
        // the central bodies are the first massive positions
        r.nc += __builtin_popcountll(__builtin_amdgcn_ballot_w64(act && r.M[k] != 0.0 && p < ncentral));
    }
    double m_max = 0.0;
#pragma unroll
    for (int k = 0; k < NSLOT; ++k) m_max = fmax(m_max, r.act[k] ? fabs(r.M[k]) : 0.0);
    resident_tiny(wave_max(m_max), r.tiny32, r.tiny64);
    // 64 NSLOT + 1 records per wave: position nm is a massless zero record
    // (it pads an odd count of FP32 sources)
    extern __shared__ ResSrc res_src[];
    ResSrc *src = res_src + wave * (64 * NSLOT + 1);
    if (lane == 0) {
        src[r.nm].a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        src[r.nm].b = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    const double G = a.Gs[s], dt = a.dts[s], hsys = a.hs[s];
    const bool energy_wave = wave == 0;
    if (hsys > 0.0) resident_run<NSLOT, true>(a, r, src, s, lane, energy_wave, nsteps, G, dt, hsys);
    else resident_run<NSLOT, false>(a, r, src, s, lane, energy_wave, nsteps, G, dt, hsys);
#pragma unroll
    for (int k = 0; k < NSLOT; ++k) {
        // the massive particles are written by wave 0 only
        if (!r.act[k] || (wave > 0 && r.M[k] != 0.0)) continue;
        const int64_t i = o0 + pid[k];
        a.x[i] = r.X[k]; a.y[i] = r.Y[k]; a.z[i] = r.Z[k];
        a.vx[i] = r.VX[k]; a.vy[i] = r.VY[k]; a.vz[i] = r.VZ[k];
        a.ax[i] = r.AX[k]; a.ay[i] = r.AY[k]; a.az[i] = r.AZ[k];
    }
}

// ---- packed resident kernel -----------------------------------------------
// Several systems per workgroup, split by role: the systems' massive
// particles fill the first wavefronts (each system's within one wave), their
// massless particles the waves after, so a planetary system's 64 test
// particles take one full wave and the massive waves serve several systems.
// Each step every massive lane publishes its 32-byte source record (ResSrc)
// into one of two LDS buffers; one workgroup barrier later every lane reads
// its own system's records.  The buffer alternates with the step,
// so a step's writes never meet the previous step's reads (every lane has
// passed this step's barrier only after those reads) and one barrier per step
// suffices.  A system's records: slots 0 .. NCB - 1 for its central bodies,
// NCB .. NMB - 1 for the others, massless zero records where it has fewer
// (NCB, NMB: the block's largest counts), and one more zero record that pads
// an odd FP32 count.  Each particle's arithmetic is the one-system resident
// kernel's, so both layouts give identical positions.
struct PackArgs {
    const int32_t *lane, *meta, *blk;
    const double *mmax;   // largest |mass| of each block's systems
    int smax;        // systems per block at most (block table stride 5 + 2 smax)
    int recs;        // records per system
    const double *Gs, *hs, *dts;
    double *x, *y, *z, *vx, *vy, *vz;
    const double *m;
    double *ax, *ay, *az;
    double *energy;      // [nsys][slots] or nullptr (zeroed by the host)
    int energy_every;
    int energy_slots;
};

// Block table: [0] NCB, [1] NMB, [2] systems in the block, [3] softened,
// [4] unused, then smax system ids and smax massive counts.  meta: bits 16+ system in block, bit 15 first massive lane
// of its system, bits 0-14 record slot (0x7fff: massless).
constexpr int PK_NONE = 0x7fff;

template <bool SOFT>
__device__ __forceinline__ void packed_run(const PackArgs &a, const int32_t *bi, int p, int meta, int nsteps,
                                           ResSrc *rec, ResSrc *erec, double *es) {
    const int t = threadIdx.x;
    const int NCB = bi[0], NMB = bi[1], nsb = bi[2];
    const int sl = meta >> 16, slot = meta & PK_NONE;
    const bool first = (meta >> 15) & 1;
    const bool act = p >= 0, massive = act && slot != PK_NONE;
    const int s = bi[5 + sl];
    const int nm_s = bi[5 + a.smax + sl];
    const double G = a.Gs[s], dt = a.dts[s], hsys = a.hs[s];
    const double hinv = SOFT ? 1.0 / hsys : 0.0, hinv3 = hinv * hinv * hinv;
    const float hf = (float)hsys, h2f = hf * hf, h2cf = fmaxf(h2f, FLT_MIN);
    const float hinvf = SOFT ? 1.0f / hf : 0.0f, hinv3f = hinvf * hinvf * hinvf;
    float tiny32;
    double tiny64;
    resident_tiny(a.mmax[blockIdx.x], tiny32, tiny64);
    double X = act ? a.x[p] : 0.0, Y = act ? a.y[p] : 0.0, Z = act ? a.z[p] : 0.0;
    double VX = act ? a.vx[p] : 0.0, VY = act ? a.vy[p] : 0.0, VZ = act ? a.vz[p] : 0.0;
    double AX = act ? a.ax[p] : 0.0, AY = act ? a.ay[p] : 0.0, AZ = act ? a.az[p] : 0.0;
    const double M = act ? a.m[p] : 0.0;
    const int rb = sl * a.recs;
    const int nrec = nsb * a.recs;
    int e_slot = 0;
    for (int step = 0; step < nsteps; ++step) {
        ResSrc *src = rec + (step & 1) * nrec;
        X += VX * dt; Y += VY * dt; Z += VZ * dt;
        if (massive) {
            if (slot < NCB) {
                src[rb + slot].a = f64x2_bits(X, Y);
                src[rb + slot].b = f64x2_bits(Z, M);
            } else {
                src[rb + slot] = src_f32(X, Y, Z, M);
            }
        }
        __syncthreads();
        double Ax = 0.0, Ay = 0.0, Az = 0.0, Aq = 0.0;
        for (int j = 0; j < NCB; ++j) {
            double sx, sy, sz, sm;
            f64x2_of(src[rb + j].a, sx, sy);
            f64x2_of(src[rb + j].b, sz, sm);
            if constexpr (SOFT) {
                pair_f64<false>(sx, sy, sz, sm, X, Y, Z, hsys, hinv, hinv3, Ax, Ay, Az, Aq);
            } else {
                const double dx = sx - X, dy = sy - Y, dz = sz - Z;
                const double r2 = dx * dx + dy * dy + (dz * dz + tiny64);
                const double ri = rsq_f64(r2);
                const double fac = sm * (ri * ri) * ri;
                Ax += dx * fac; Ay += dy * fac; Az += dz * fac;
            }
        }
        float fx = 0.0f, fy = 0.0f, fz = 0.0f, fq = 0.0f;
        for (int j = NCB; j < NMB; j += 2) {
            double ax_, ay_, az_, bx_, by_, bz_;
            float am, bm;
            src_f32_of(src[rb + j], ax_, ay_, az_, am);
            src_f32_of(src[rb + j + 1], bx_, by_, bz_, bm);
            res_pair_f32<SOFT>(ax_, ay_, az_, am, X, Y, Z, h2cf, h2f, hinvf, hinv3f, tiny32, fx, fy, fz, fq);
            res_pair_f32<SOFT>(bx_, by_, bz_, bm, X, Y, Z, h2cf, h2f, hinvf, hinv3f, tiny32, fx, fy, fz, fq);
        }
        AX = G * (Ax + (double)fx);
        AY = G * (Ay + (double)fy);
        AZ = G * (Az + (double)fz);
        VX += AX * dt; VY += AY * dt; VZ += AZ * dt;
        if (a.energy && (step + 1) % a.energy_every == 0) {
            // total energy with velocities synchronised to the step: FP64
            // records of every massive particle, the FP64 potential over them
            // (src/directsum/potential.jl:2-27), a sum over the system's
            // massive lanes in lane order; massless particles add nothing
            if (massive) {
                erec[rb + slot].a = f64x2_bits(X, Y);
                erec[rb + slot].b = f64x2_bits(Z, M);
            }
            __syncthreads();
            double e = 0.0;
            if (massive) {
                double q = 0.0;
                for (int j = 0; j < NMB; ++j) {
                    double sx, sy, sz, sm;
                    f64x2_of(erec[rb + j].a, sx, sy);
                    f64x2_of(erec[rb + j].b, sz, sm);
                    double bx = 0.0, by = 0.0, bz = 0.0;
                    if (j != slot && sm != 0.0)
                        pair_f64<true>(sx, sy, sz, sm, X, Y, Z, hsys, SOFT ? hinv : 0.0, SOFT ? hinv3 : 0.0, bx,
                                       by, bz, q);
                }
                const double ux = VX - 0.5 * AX * dt, uy = VY - 0.5 * AY * dt, uz = VZ - 0.5 * AZ * dt;
                e = 0.5 * M * (ux * ux + uy * uy + uz * uz) - 0.5 * G * M * q;
            }
            es[t] = e;
            __syncthreads();
            if (first && e_slot < a.energy_slots) {
                double sum = 0.0;
                for (int k = 0; k < nm_s; ++k) sum += es[t + k];
                a.energy[(int64_t)s * a.energy_slots + e_slot] = sum;
            }
            ++e_slot;
        }
    }
    if (act) {
        a.x[p] = X; a.y[p] = Y; a.z[p] = Z;
        a.vx[p] = VX; a.vy[p] = VY; a.vz[p] = VZ;
        a.ax[p] = AX; a.ay[p] = AY; a.az[p] = AZ;
    }
}

// One workgroup per block of the packed layout (k_batch_resident_packed's
// host side, build_packed, makes the tables); the zero records are written
// first.  Every lane, idle ones included, takes every barrier.
__global__ __launch_bounds__(1024) void k_batch_resident_packed(PackArgs a, int nsteps) {
    extern __shared__ ResSrc pk_src[];
    const int32_t *bi = a.blk + (int64_t)blockIdx.x * (5 + 2 * a.smax);
    const int nrec = bi[2] * a.recs;
    ResSrc *rec = pk_src, *erec = pk_src + 2 * nrec;
    double *es = reinterpret_cast<double *>(pk_src + 3 * nrec);
    for (int i = threadIdx.x; i < 3 * nrec; i += blockDim.x) {
        pk_src[i].a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        pk_src[i].b = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    __syncthreads();
    const int64_t li = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int p = a.lane[li], meta = a.meta[li];
    if (bi[3]) packed_run<true>(a, bi, p, meta, nsteps, rec, erec, es);
    else packed_run<false>(a, bi, p, meta, nsteps, rec, erec, es);
}

// ---- Morton order within systems -----------------------------------------
__global__ __launch_bounds__(256) void k_sys_box(const int64_t *__restrict__ off,
                                                 const double *__restrict__ x,
                                                 const double *__restrict__ y,
                                                 const double *__restrict__ z,
                                                 double *__restrict__ box) {
    __shared__ double red[6][4];
    const int s = blockIdx.x;
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int64_t i = off[s] + threadIdx.x; i < off[s + 1]; i += 256) {
        lo[0] = fmin(lo[0], x[i]); hi[0] = fmax(hi[0], x[i]);
        lo[1] = fmin(lo[1], y[i]); hi[1] = fmax(hi[1], y[i]);
        lo[2] = fmin(lo[2], z[i]); hi[2] = fmax(hi[2], z[i]);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        lo[d] = wave_min(lo[d]);
        hi[d] = wave_max(hi[d]);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int d = 0; d < 3; ++d) { red[d][w] = lo[d]; red[3 + d][w] = hi[d]; }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int q = threadIdx.x;
        double r = red[q][0];
        for (int ww = 1; ww < 4; ++ww) r = q < 3 ? fmin(r, red[q][ww]) : fmax(r, red[q][ww]);
        box[(int64_t)s * 8 + q] = r;
    }
}

__device__ __forceinline__ uint64_t spread17(uint64_t v) {
    v &= 0x1ffff;
    v = (v | (v << 32)) & 0x1f00000000ffffULL;
    v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
    v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}

// Sort key: the system index above a Morton code of `mbits` bits per axis,
// with 3 mbits + (bits of the largest system index) <= 64, so that no two
// systems' keys interleave (mbits = 17 up to 8192 systems, fewer beyond).
__global__ void k_batch_morton(const int32_t *__restrict__ sys_of, const double *__restrict__ x,
                               const double *__restrict__ y, const double *__restrict__ z,
                               const double *__restrict__ box, int64_t n, int mbits,
                               uint64_t *__restrict__ key, int64_t *__restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = sys_of[i];
    const double *b = box + (int64_t)s * 8;
    const double p[3] = {x[i], y[i], z[i]};
    uint64_t c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double ext = b[3 + d] - b[d];
        double u = ext > 0.0 ? (p[d] - b[d]) / ext : 0.0;
        u = fmin(fmax(u, 0.0), 1.0);
        c[d] = (uint64_t)(u * (double)((1u << mbits) - 1u));
    }
    key[i] = ((uint64_t)s << (3 * mbits)) | spread17(c[0]) | (spread17(c[1]) << 1) |
             (spread17(c[2]) << 2);
    idx[i] = i;
}

template <typename V>
__global__ void k_bgather(const V *__restrict__ src, const int64_t *__restrict__ idx, int64_t n,
                          V *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[i] = src[idx[i]];
}

// dst[perm[i]] = src[i]: back to the caller's order.
template <typename V>
__global__ void k_bscatter(const V *__restrict__ src, const int64_t *__restrict__ perm, int64_t n,
                           V *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[perm[i]] = src[i];
}

__global__ void k_bcompose(const int64_t *__restrict__ old_perm, const int64_t *__restrict__ idx,
                           int64_t n, int64_t *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = old_perm ? old_perm[idx[i]] : idx[i];
}

__global__ void k_central_list(const int64_t *__restrict__ perm, const int32_t *__restrict__ sys_of,
                               const int64_t *__restrict__ off, int64_t n, int ncentral,
                               uint8_t *__restrict__ flags, int64_t *__restrict__ cidx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t u = perm ? perm[i] : i;   // caller index
    const int s = sys_of[i];
    const int64_t local = u - off[s];
    const bool c = local < ncentral;
    flags[i] = c ? 1 : 0;
    if (c) cidx[(int64_t)s * ncentral + local] = i;
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
static int bcheck(nb_batch *b, bool need_upload = true) {
    clear_error();
    if (!b) {
        set_error("null batch");
        return NB_ERR_INVALID;
    }
    hipError_t e = hipSetDevice(b->dev);
    if (e != hipSuccess) {
        set_error("hipSetDevice failed: %s", hipGetErrorString(e));
        return NB_ERR_HIP;
    }
    if (need_upload && !b->uploaded) {
        set_error("no particle data uploaded (call nb_batch_upload first)");
        return NB_ERR_STATE;
    }
    return NB_OK;
}

static void bdrop_graph(nb_batch *b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
    b->g_steps = 0;
}

static int64_t resident_lanes(const nb_batch *b, int W);

// The resident kernel runs when some wave count W <= 4 per system fits every
// system's massive particles plus a 1/W share of its massless ones into 4
// lane slots (256 lanes) per wave; before the masses are known, when every
// system has at most 256 particles.
static bool packed_ok(const nb_batch *b);

static bool resident_ok(const nb_batch *b) {
    if (!b->tun.resident || b->nsys <= 0) return false;
    if (b->h_nm.size() != (size_t)b->nsys) return b->max_ns <= 256;
    if (packed_ok(b)) return true;
    for (int W = 1; W <= 4; ++W)
        if (resident_lanes(b, W) <= 256) return true;
    return false;
}

// (system, target tile) items, the most expensive (largest system) first.
static int build_items(nb_batch *b, int tpt) {
    if (b->items && b->items_tpt == tpt) return NB_OK;
    const int64_t ti = (int64_t)BS * tpt;
    std::vector<int32_t> order(b->nsys);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t p, int32_t q) {
        return b->h_off[p + 1] - b->h_off[p] > b->h_off[q + 1] - b->h_off[q];
    });
    std::vector<int64_t> items;
    for (int32_t s : order) {
        const int64_t ns = b->h_off[s + 1] - b->h_off[s];
        for (int64_t k = 0; k * ti < ns; ++k) items.push_back(((int64_t)s << 32) | k);
    }
    dfree(b->items);
    b->nitems = (int64_t)items.size();
    NB_TRY(dalloc(&b->items, sizeof(int64_t) * std::max<size_t>(items.size(), 1)));
    if (!items.empty())
        NB_HIP(hipMemcpy(b->items, items.data(), sizeof(int64_t) * items.size(),
                         hipMemcpyHostToDevice));
    b->items_tpt = tpt;
    return NB_OK;
}

static int bpermute(double *arr, const int64_t *idx, int64_t n, void *spare, hipStream_t s) {
    k_bgather<double><<<blocks_for(n), 256, 0, s>>>(arr, idx, n, (double *)spare);
    NB_TRY(launch_ok("k_bgather"));
    NB_HIP(hipMemcpyAsync(arr, spare, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
    return NB_OK;
}

static int rebuild_central(nb_batch *b) {
    if (b->ncentral <= 0 || b->n == 0) return NB_OK;
    k_central_list<<<blocks_for(b->n), 256, 0, b->stream>>>(b->perm, b->sys_of, b->off, b->n,
                                                            b->ncentral, b->central, b->cidx);
    return launch_ok("k_central_list");
}

static int batch_sort(nb_batch *b) {
    const int64_t n = b->n;
    if (n < 2) return NB_OK;
    hipStream_t s = b->stream;
    if (!b->keys) {
        NB_TRY(dalloc(&b->keys, sizeof(uint64_t) * 2 * (size_t)n));
        NB_TRY(dalloc(&b->idx, sizeof(int64_t) * 2 * (size_t)n));
        NB_TRY(dalloc(&b->sysbox, sizeof(double) * 8 * (size_t)b->nsys));
        NB_TRY(dalloc(&b->spare, sizeof(double) * (size_t)n));
        size_t bytes = 0;
        NB_HIP(rocprim::radix_sort_pairs(nullptr, bytes, b->keys, b->keys + n, b->idx, b->idx + n,
                                         (size_t)n, 0, 64, s));
        NB_TRY(dalloc(&b->sort_tmp, bytes));
        b->sort_tmp_bytes = bytes;
    }
    k_sys_box<<<b->nsys, 256, 0, s>>>(b->off, b->x, b->y, b->z, b->sysbox);
    NB_TRY(launch_ok("k_sys_box"));
    int sys_bits = 0;
    while (sys_bits < 31 && ((int64_t)1 << sys_bits) < (int64_t)b->nsys) ++sys_bits;
    const int mbits = std::min(17, (64 - sys_bits) / 3);
    k_batch_morton<<<blocks_for(n), 256, 0, s>>>(b->sys_of, b->x, b->y, b->z, b->sysbox, n, mbits,
                                                 b->keys, b->idx);
    NB_TRY(launch_ok("k_batch_morton"));
    size_t bytes = b->sort_tmp_bytes;
    NB_HIP(rocprim::radix_sort_pairs(b->sort_tmp, bytes, b->keys, b->keys + n, b->idx, b->idx + n,
                                     (size_t)n, 0, 64, s));
    const int64_t *idx = b->idx + n;
    double *arrs[12] = {b->x, b->y, b->z, b->vx, b->vy, b->vz, b->m, b->h, b->ax, b->ay, b->az, b->pot};
    for (double *p : arrs) NB_TRY(bpermute(p, idx, n, b->spare, s));
    int64_t *np = (int64_t *)b->keys;
    k_bcompose<<<blocks_for(n), 256, 0, s>>>(b->perm, idx, n, np);
    NB_TRY(launch_ok("k_bcompose"));
    if (!b->perm) NB_TRY(dalloc(&b->perm, sizeof(int64_t) * (size_t)n));
    NB_HIP(hipMemcpyAsync(b->perm, np, sizeof(int64_t) * (size_t)n, hipMemcpyDeviceToDevice, s));
    NB_TRY(rebuild_central(b));
    b->sorted = true;
    b->calls_since_sort = 0;
    return NB_OK;
}

// Undo the Morton order: every per-particle array back in the caller's order.
// The resident kernel needs it (it takes the central bodies and the origin by
// position in the system), and an ensemble sorted while it ran the ragged
// path (resident = 0, or before nb_batch_set_tuning switched resident on)
// comes back here before its first resident launch.
static int batch_unsort(nb_batch *b) {
    if (!b->perm) {
        b->sorted = false;
        return NB_OK;
    }
    const int64_t n = b->n;
    hipStream_t s = b->stream;
    if (n > 0) {
        if (!b->spare) NB_TRY(dalloc(&b->spare, sizeof(double) * (size_t)n));
        double *arrs[12] = {b->x, b->y, b->z, b->vx, b->vy, b->vz, b->m, b->h, b->ax, b->ay, b->az, b->pot};
        for (double *p : arrs) {
            k_bscatter<double><<<blocks_for(n), 256, 0, s>>>(p, b->perm, n, (double *)b->spare);
            NB_TRY(launch_ok("k_bscatter"));
            NB_HIP(hipMemcpyAsync(p, b->spare, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
        }
        NB_HIP(hipStreamSynchronize(s));
    }
    dfree(b->perm);
    b->sorted = false;
    b->calls_since_sort = 0;
    return rebuild_central(b);
}

// Systems small enough for the resident kernel stay in caller order (the
// resident kernel identifies central bodies and the origin by position in
// the system); only the ragged path of larger systems is re-sorted.
static int bmaybe_sort(nb_batch *b, int64_t upcoming) {
    if (b->mode == NB_MODE_FP64 || b->tun.sort_interval <= 0 || resident_ok(b)) return NB_OK;
    if (!b->sorted || b->calls_since_sort + upcoming > b->tun.sort_interval) return batch_sort(b);
    return NB_OK;
}

// pot_only (with with_pot): the FP32-pair kernels and the central pass write
// the potential alone and leave the accelerations as they are.
static int bforce(nb_batch *b, bool with_pot, bool allow_timing, bool pot_only = false) {
    if (b->n == 0) return NB_OK;
    hipStream_t s = b->stream;
    const bool f64 = b->mode == NB_MODE_FP64;
    const int tpt = f64 ? 1 : b->tun.targets_per_thread;
    NB_TRY(build_items(b, tpt));
    b->calls_since_sort++;
    BatchArgs a{};
    a.off = b->off; a.tile_base = b->tile_base; a.tile_box = b->tile_box; a.sub_box = b->sub_box; a.Gs = b->Gs;
    a.x = b->x; a.y = b->y; a.z = b->z; a.m = b->m; a.h = b->h;
    a.skip_src = (b->mode == NB_MODE_SPLIT_CENTRAL && b->ncentral > 0) ? b->central : nullptr;
    a.ax = b->ax; a.ay = b->ay; a.az = b->az; a.pot = b->pot;
    a.items = b->items; a.nitems = b->nitems; a.counter = b->counter;
    hipEvent_t e0 = nullptr;
    if (allow_timing && b->timing) {
        NB_HIP(hipEventCreate(&e0));
        NB_HIP(hipEventRecord(e0, s));
    }
    if (f64) {
        if (with_pot) k_batch_force_f64<true><<<(unsigned)b->nitems, BS, 0, s>>>(a);
        else k_batch_force_f64<false><<<(unsigned)b->nitems, BS, 0, s>>>(a);
        NB_TRY(launch_ok("k_batch_force_f64"));
    } else {
        k_batch_tile_boxes<<<(unsigned)b->ntiles, BS, 0, s>>>(b->x, b->y, b->z, b->off,
                                                              b->tile_base, b->tile_sys, b->tile_box,
                                                              b->sub_box);
        NB_TRY(launch_ok("k_batch_tile_boxes"));
        pot_only = pot_only && with_pot;
        // potential alone on far path 4: m/r = rsq(w r^2) m^(2/3)
        const bool pot_mf4 = pot_only && tpt == 1 && b->tun.far_path == FAR_MFMA_BF16;
        if ((!with_pot && far_path_weighted(b->tun.far_path)) || pot_mf4) {
            NB_TRY(launch_source_weights(b->m, true, b->n, b->wsrc, s));
            a.w = b->wsrc;
        }
        int nblk = b->tun.blocks > 0 ? b->tun.blocks : 8 * b->cu_count;
        nblk = (int)std::max<int64_t>(1, std::min<int64_t>(nblk, b->nitems));
#define NB_BL(T_, P_) k_batch_force<T_, P_><<<nblk, BS, 0, s>>>(a)
        const int fp = with_pot ? FAR_VALU : b->tun.far_path;
        if (pot_mf4) k_batch_force<1, true, FAR_MFMA_BF16, false><<<nblk, BS, 0, s>>>(a);
        else if (pot_only && tpt == 1) k_batch_force<1, true, FAR_VALU, false><<<nblk, BS, 0, s>>>(a);
        else if (pot_only && tpt == 2) k_batch_force<2, true, FAR_VALU, false><<<nblk, BS, 0, s>>>(a);
        else if (pot_only) k_batch_force<4, true, FAR_VALU, false><<<nblk, BS, 0, s>>>(a);
        else if (tpt == 1 && fp == FAR_MFMA_R2) k_batch_force<1, false, FAR_MFMA_R2><<<nblk, BS, 0, s>>>(a);
        else if (tpt == 1 && fp == FAR_MFMA_SUM) k_batch_force<1, false, FAR_MFMA_SUM><<<nblk, BS, 0, s>>>(a);
        else if (tpt == 1 && fp == FAR_MFMA_BF16) k_batch_force<1, false, FAR_MFMA_BF16><<<nblk, BS, 0, s>>>(a);
        else if (fp == FAR_WEIGHTED && tpt == 1) k_batch_force<1, false, FAR_WEIGHTED><<<nblk, BS, 0, s>>>(a);
        else if (fp == FAR_WEIGHTED && tpt == 2) k_batch_force<2, false, FAR_WEIGHTED><<<nblk, BS, 0, s>>>(a);
        else if (fp == FAR_WEIGHTED && tpt == 4) k_batch_force<4, false, FAR_WEIGHTED><<<nblk, BS, 0, s>>>(a);
        else if (tpt == 1) { if (with_pot) NB_BL(1, true); else NB_BL(1, false); }
        else if (tpt == 2) { if (with_pot) NB_BL(2, true); else NB_BL(2, false); }
        else { if (with_pot) NB_BL(4, true); else NB_BL(4, false); }
#undef NB_BL
        NB_TRY(launch_ok("k_batch_force"));
    }
    if (e0) {
        hipEvent_t e1;
        NB_HIP(hipEventCreate(&e1));
        NB_HIP(hipEventRecord(e1, s));
        b->pending.emplace_back(e0, e1);
    }
    if (b->mode == NB_MODE_SPLIT_CENTRAL && b->ncentral > 0) {
        if (with_pot && pot_only && !f64)
            k_batch_central<true, false><<<blocks_for(b->n), 256, 0, s>>>(
                b->sys_of, b->cidx, b->ncentral, b->Gs, b->x, b->y, b->z, b->m, b->h, b->n, b->ax,
                b->ay, b->az, b->pot, b->off, b->central);
        else if (with_pot)
            k_batch_central<true><<<blocks_for(b->n), 256, 0, s>>>(
                b->sys_of, b->cidx, b->ncentral, b->Gs, b->x, b->y, b->z, b->m, b->h, b->n, b->ax,
                b->ay, b->az, b->pot, b->off, b->central);
        else
            k_batch_central<false><<<blocks_for(b->n), 256, 0, s>>>(
                b->sys_of, b->cidx, b->ncentral, b->Gs, b->x, b->y, b->z, b->m, b->h, b->n, b->ax,
                b->ay, b->az, b->pot, b->off, b->central);
        NB_TRY(launch_ok("k_batch_central"));
    }
    return NB_OK;
}

// Pinned host staging of at least n doubles (grown on demand).
static int ensure_pinned(nb_batch *b, size_t n) {
    if (b->h_pinned_n >= n) return NB_OK;
    if (b->h_pinned) NB_HIP(hipHostFree(b->h_pinned));
    b->h_pinned = nullptr;
    b->h_pinned_n = 0;
    hipError_t e = hipHostMalloc((void **)&b->h_pinned, sizeof(double) * n, hipHostMallocDefault);
    if (e != hipSuccess) {
        b->h_pinned = nullptr;
        set_error("hipHostMalloc(%zu bytes) failed: %s", sizeof(double) * n, hipGetErrorString(e));
        return NB_ERR_HIP;
    }
    b->h_pinned_n = n;
    return NB_OK;
}

// Device buffer of at least n tracked energies (grown on demand).
static int ensure_edev(nb_batch *b, size_t n) {
    if (b->e_dev_n >= n) return NB_OK;
    dfree(b->e_dev);
    b->e_dev_n = 0;
    NB_TRY(dalloc(&b->e_dev, sizeof(double) * n));
    b->e_dev_n = n;
    return NB_OK;
}

static int bkick_drift(nb_batch *b, double kick_coef, bool drift) {
    if (b->n == 0) return NB_OK;
    k_batch_kick_drift<<<blocks_for(b->n), 256, 0, b->stream>>>(
        b->x, b->y, b->z, b->vx, b->vy, b->vz, b->ax, b->ay, b->az, b->sys_of, b->dts, b->n,
        kick_coef, drift ? 1 : 0);
    return launch_ok("k_batch_kick_drift");
}

// S steps as one schedule: D, F, (KD, F) x (S-1), K
static int enqueue_steps(nb_batch *b, int64_t S, bool timing) {
    NB_TRY(bkick_drift(b, 0.0, true));
    for (int64_t k = 0; k < S; ++k) {
        if (k > 0) NB_TRY(bkick_drift(b, 1.0, true));
        NB_TRY(bforce(b, false, timing));
    }
    NB_TRY(bkick_drift(b, 1.0, false));
    return NB_OK;
}

// Lanes a resident wave needs when each system runs on W waves: its massive
// particles plus a 1/W share of the massless ones.
static int64_t resident_lanes(const nb_batch *b, int W) {
    int64_t need = 0;
    for (int32_t q = 0; q < b->nsys; ++q) {
        const int64_t ns = b->h_off[q + 1] - b->h_off[q], nm = b->h_nm[(size_t)q];
        need = std::max(need, nm + (ns - nm + W - 1) / W);
    }
    return need;
}

// Waves per system and lane slots per wave of the resident kernel: the
// fewest wave slots (waves x slots) per system, ties to fewer waves (the
// kernel is bound by VALU issue, and every extra wave repeats the massive
// particles' arithmetic and the per-step overhead: on 2048 planetary
// systems of 73 bodies one wave of two slots took 1.74 us per step, two
// waves of one slot 1.85 us); tun.resident_waves forces the wave count where
// its lanes fit 4 slots.
static void resident_shape(const nb_batch *b, int &waves, int &nslot) {
    int best = 1 << 30;
    for (int W = 1; W <= 4; ++W) {
        const int ns = (int)std::max<int64_t>(1, (resident_lanes(b, W) + 63) / 64);
        if (ns > 4) continue;
        if (b->tun.resident_waves == W) {
            waves = W;
            nslot = ns;
            return;
        }
        if (W * ns < best) {
            best = W * ns;
            waves = W;
            nslot = ns;
        }
    }
}

// The packed layout applies when no wave count is forced, the particle
// indices fit its 32-bit lane tables, and every system has at most 64 massive
// particles and fits one workgroup (one massive wave plus its massless
// particles in at most 15 more).
static bool packed_ok(const nb_batch *b) {
    if (b->tun.resident_waves != 0 || b->h_nm.size() != (size_t)b->nsys || b->n >= INT32_MAX) return false;
    for (int32_t q = 0; q < b->nsys; ++q) {
        const int64_t ns = b->h_off[q + 1] - b->h_off[q], nm = b->h_nm[(size_t)q];
        if (nm > 64 || 1 + (ns - nm + 63) / 64 > 16) return false;
    }
    return true;
}

// Builds the packed tables.  Systems are ordered by (softened, central count,
// massive count) and taken up to S at a time (S below), a block closing early
// when the next system is softened differently or would take it past 16
// waves.
// Massive particles are packed into the block's first waves without splitting
// a system; massless ones follow, consecutively.
static int build_packed(nb_batch *b) {
    const int32_t ns_sys = b->nsys;
    const bool f64 = b->mode == NB_MODE_FP64;
    const int ncentral = b->mode == NB_MODE_SPLIT_CENTRAL ? b->ncentral : 0;
    std::vector<int32_t> nc(ns_sys), order(ns_sys);
    for (int32_t q = 0; q < ns_sys; ++q) {
        int c = 0;
        for (int64_t i = b->h_off[q]; i < b->h_off[q + 1]; ++i)
            c += b->h_massive[(size_t)i] && (f64 || i - b->h_off[q] < ncentral);
        nc[q] = c;
        order[q] = q;
    }
    auto soft = [&](int32_t q) { return b->hh[(size_t)q] > 0.0; };
    std::stable_sort(order.begin(), order.end(), [&](int32_t p, int32_t q) {
        if (soft(p) != soft(q)) return soft(p) < soft(q);
        if (nc[p] != nc[q]) return nc[p] < nc[q];
        return b->h_nm[(size_t)p] < b->h_nm[(size_t)q];
    });
    // S also keeps the workgroup's LDS (three buffers of 32-byte records)
    // within 64 KiB.  Records per system: the largest central count plus the
    // largest other massive count plus one (at most 129), which may come from
    // different systems.
    int nc_max = 0, nf_max = 0;
    for (int32_t q = 0; q < ns_sys; ++q) {
        nc_max = std::max(nc_max, nc[q]);
        nf_max = std::max(nf_max, b->h_nm[(size_t)q] - nc[q]);
    }
    const int rmax = nc_max + nf_max + 1;
    const int s_lds = std::max(1, (65536 - 1024 * (int)sizeof(double)) / (3 * rmax * (int)sizeof(ResSrc)));
    // S minimises the estimated makespan ceil(workgroups / CUs) x waves per
    // workgroup (the kernel is bound by VALU issue, and a CU runs its
    // workgroups' waves together), with a workgroup's waves estimated from
    // the mean massive and massless counts; ties go to the larger S.  On the
    // 2048 planetary systems of 73 bodies that gives 8 (256 workgroups of
    // 10 waves: 1.32 us per step, against 1.37 for S = 4, 1.36 for 3, 1.43
    // for 2 and 1.49 for 6, whose 342 workgroups leave CUs one short).
    double nm_avg = 0.0, nt_avg = 0.0;
    for (int32_t q = 0; q < ns_sys; ++q) {
        nm_avg += b->h_nm[(size_t)q];
        nt_avg += (double)(b->h_off[q + 1] - b->h_off[q]) - b->h_nm[(size_t)q];
    }
    nm_avg /= std::max(1, ns_sys);
    nt_avg /= std::max(1, ns_sys);
    int S = 1;
    double best = 1e300;
    for (int k = 1; k <= std::min(16, s_lds); ++k) {
        const double waves = std::ceil(k * nm_avg / 64.0) + std::ceil(k * nt_avg / 64.0);
        if (waves > 16) break;
        const double blocks = std::ceil((double)ns_sys / k);
        const double span = std::ceil(blocks / b->cu_count) * std::max(1.0, waves);
        if (span <= best) {
            best = span;
            S = k;
        }
    }
    struct Blk {
        std::vector<int32_t> sys;
        int mw = 0, lanes_in_mw = 64, nt = 0;   // massive waves, lanes used in the last one, massless
    };
    std::vector<Blk> blks;
    auto waves = [](const Blk &k) { return k.mw + (k.nt + 63) / 64; };
    for (int32_t q : order) {
        const int nm = b->h_nm[(size_t)q];
        const int nt = (int)(b->h_off[q + 1] - b->h_off[q]) - nm;
        bool open = !blks.empty() && (int)blks.back().sys.size() < S && soft(blks.back().sys[0]) == soft(q);
        if (open) {
            Blk t = blks.back();
            if (nm > 0 && t.lanes_in_mw + nm > 64) { ++t.mw; t.lanes_in_mw = 0; }
            t.lanes_in_mw += nm;
            t.nt += nt;
            open = waves(t) <= 16;
        }
        if (!open) blks.push_back(Blk{});
        Blk &k = blks.back();
        if (nm > 0 && k.lanes_in_mw + nm > 64) { ++k.mw; k.lanes_in_mw = 0; }
        k.lanes_in_mw += nm;
        k.nt += nt;
        k.sys.push_back(q);
    }
    int maxw = 1, smax = 1, ncb_max = 0, nfb_max = 0;
    for (const Blk &k : blks) {
        maxw = std::max(maxw, waves(k));
        smax = std::max(smax, (int)k.sys.size());
        for (int32_t q : k.sys) {
            ncb_max = std::max(ncb_max, nc[q]);
            nfb_max = std::max(nfb_max, b->h_nm[(size_t)q] - nc[q]);
        }
    }
    const int threads = 64 * maxw, stride = 5 + 2 * smax, recs = ncb_max + nfb_max + 1;
    const size_t nb = blks.size();
    std::vector<int32_t> lane(nb * threads, -1), meta(nb * threads, (int32_t)PK_NONE), blk(nb * stride, -1);
    std::vector<double> mmax_b(nb);
    for (size_t kb = 0; kb < nb; ++kb) {
        const Blk &k = blks[kb];
        int ncb = 0, nfb = 0;
        double mmax = 0.0;
        for (int32_t q : k.sys) {
            ncb = std::max(ncb, nc[q]);
            nfb = std::max(nfb, b->h_nm[(size_t)q] - nc[q]);
            mmax = std::max(mmax, b->h_mmax[(size_t)q]);
        }
        int32_t *bi = &blk[kb * stride];
        bi[0] = ncb; bi[1] = ncb + nfb; bi[2] = (int32_t)k.sys.size(); bi[3] = soft(k.sys[0]); bi[4] = 0;
        mmax_b[kb] = mmax;
        // massive waves first (each system's massive particles in one wave),
        // then the massless particles
        int mw_used = 0, cur_w = 0, in_w = 0;
        for (size_t j = 0; j < k.sys.size(); ++j) {
            const int32_t q = k.sys[j];
            const int nm = b->h_nm[(size_t)q];
            bi[5 + j] = q;
            bi[5 + smax + j] = nm;
            if (nm == 0) continue;
            if (in_w + nm > 64) { ++cur_w; in_w = 0; }
            int mlane = 64 * cur_w + in_w;
            in_w += nm;
            mw_used = cur_w + 1;
            int ci = 0, fi = 0;
            bool firstm = true;
            for (int64_t i = b->h_off[q]; i < b->h_off[q + 1]; ++i) {
                if (!b->h_massive[(size_t)i]) continue;
                const bool central = f64 || i - b->h_off[q] < ncentral;
                const int slot = central ? ci++ : ncb + fi++;
                lane[kb * threads + mlane] = (int32_t)i;
                meta[kb * threads + mlane] = ((int32_t)j << 16) | (firstm ? 0x8000 : 0) | slot;
                firstm = false;
                ++mlane;
            }
        }
        int tlane = 64 * mw_used;
        for (size_t j = 0; j < k.sys.size(); ++j) {
            const int32_t q = k.sys[j];
            for (int64_t i = b->h_off[q]; i < b->h_off[q + 1]; ++i) {
                if (b->h_massive[(size_t)i]) continue;
                lane[kb * threads + tlane] = (int32_t)i;
                meta[kb * threads + tlane] = ((int32_t)j << 16) | PK_NONE;
                ++tlane;
            }
        }
    }
    dfree(b->pk_lane); dfree(b->pk_meta); dfree(b->pk_blk); dfree(b->pk_tiny);
    NB_TRY(dalloc(&b->pk_lane, sizeof(int32_t) * lane.size()));
    NB_TRY(dalloc(&b->pk_meta, sizeof(int32_t) * meta.size()));
    NB_TRY(dalloc(&b->pk_blk, sizeof(int32_t) * blk.size()));
    NB_TRY(dalloc(&b->pk_tiny, sizeof(double) * mmax_b.size()));
    NB_HIP(hipMemcpy(b->pk_lane, lane.data(), sizeof(int32_t) * lane.size(), hipMemcpyHostToDevice));
    NB_HIP(hipMemcpy(b->pk_meta, meta.data(), sizeof(int32_t) * meta.size(), hipMemcpyHostToDevice));
    NB_HIP(hipMemcpy(b->pk_blk, blk.data(), sizeof(int32_t) * blk.size(), hipMemcpyHostToDevice));
    NB_HIP(hipMemcpy(b->pk_tiny, mmax_b.data(), sizeof(double) * mmax_b.size(), hipMemcpyHostToDevice));
    b->pk_nblk = (int)nb;
    b->pk_threads = threads;
    b->pk_smax = smax;
    b->pk_recs = recs;
    b->pk_stale = false;
    return NB_OK;
}

// The packed tables are built by the set-up calls (upload, set_central,
// set_params, set_tuning) that change what they depend on, so that the step
// calls do not pay for them.
static int prepare_resident(nb_batch *b) {
    if (b->uploaded && b->pk_stale && resident_ok(b) && packed_ok(b)) NB_TRY(build_packed(b));
    return NB_OK;
}

static int resident_steps(nb_batch *b, int64_t nsteps, int every, double *energy_dev, int slots) {
    ResArgs a{};
    a.off = b->off; a.Gs = b->Gs; a.hs = b->hs; a.dts = b->dts;
    a.x = b->x; a.y = b->y; a.z = b->z; a.vx = b->vx; a.vy = b->vy; a.vz = b->vz;
    a.m = b->m; a.ax = b->ax; a.ay = b->ay; a.az = b->az;
    a.ncentral = b->mode == NB_MODE_SPLIT_CENTRAL ? b->ncentral : 0;
    a.energy = energy_dev; a.energy_every = every > 0 ? every : 1; a.energy_slots = slots;
    if (nsteps > INT32_MAX) {
        set_error("too many steps for one resident launch");
        return NB_ERR_INVALID;
    }
    if (b->nsys <= 0) return NB_OK;
    if (packed_ok(b)) {
        if (b->pk_stale) NB_TRY(build_packed(b));
        PackArgs pa{};
        pa.lane = b->pk_lane; pa.meta = b->pk_meta; pa.blk = b->pk_blk; pa.mmax = b->pk_tiny;
        pa.smax = b->pk_smax; pa.recs = b->pk_recs;
        pa.Gs = b->Gs; pa.hs = b->hs; pa.dts = b->dts;
        pa.x = b->x; pa.y = b->y; pa.z = b->z; pa.vx = b->vx; pa.vy = b->vy; pa.vz = b->vz;
        pa.m = b->m; pa.ax = b->ax; pa.ay = b->ay; pa.az = b->az;
        pa.energy = energy_dev; pa.energy_every = a.energy_every; pa.energy_slots = slots;
        // systems without massive particles have zero energy and no lane that writes it
        if (energy_dev)
            NB_HIP(hipMemsetAsync(energy_dev, 0, sizeof(double) * (size_t)b->nsys * (size_t)slots, b->stream));
        const size_t lds = 3 * (size_t)b->pk_smax * b->pk_recs * sizeof(ResSrc) + b->pk_threads * sizeof(double);
        k_batch_resident_packed<<<b->pk_nblk, b->pk_threads, lds, b->stream>>>(pa, (int)nsteps);
        return launch_ok("k_batch_resident_packed");
    }
    int waves = 1, nslot = 1;
    resident_shape(b, waves, nslot);
    const bool f64 = b->mode == NB_MODE_FP64;
    const dim3 block(64 * waves);
    const size_t lds = sizeof(ResSrc) * (64 * (size_t)nslot + 1) * (size_t)waves;
#define NB_RES(NS_)                                                                                \
    do {                                                                                           \
        if (f64) k_batch_resident<NS_, true><<<b->nsys, block, lds, b->stream>>>(a, (int)nsteps);  \
        else k_batch_resident<NS_, false><<<b->nsys, block, lds, b->stream>>>(a, (int)nsteps);     \
    } while (0)
    if (nslot == 1) NB_RES(1);
    else if (nslot == 2) NB_RES(2);
    else if (nslot == 3) NB_RES(3);
    else NB_RES(4);
#undef NB_RES
    return launch_ok("k_batch_resident");
}

static int batch_steps(nb_batch *b, int64_t nsteps, bool use_graph) {
    if (nsteps <= 0 || b->n == 0) return NB_OK;
    if (resident_ok(b) && use_graph) {
        // the resident kernel runs in the caller's order
        NB_TRY(batch_unsort(b));
        return resident_steps(b, nsteps, 0, nullptr, 0);
    }
    int64_t k = 0;
    const int64_t S = std::max(2, b->tun.steps_per_graph & ~1);
    if (use_graph && nsteps >= S) {
        if (!b->gexec || b->g_steps != S) {
            bdrop_graph(b);
            NB_TRY(build_items(b, b->mode == NB_MODE_FP64 ? 1 : b->tun.targets_per_thread));
            NB_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            const int64_t keep = b->calls_since_sort;
            int rc = enqueue_steps(b, S, false);
            b->calls_since_sort = keep;
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc != NB_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) {
                set_error("hipStreamEndCapture failed: %s", hipGetErrorString(e));
                if (g) (void)hipGraphDestroy(g);
                return NB_ERR_HIP;
            }
            e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                b->gexec = nullptr;
                set_error("hipGraphInstantiate failed: %s", hipGetErrorString(e));
                return NB_ERR_HIP;
            }
            b->g_steps = S;
        }
        for (; k + S <= nsteps; k += S) {
            NB_TRY(bmaybe_sort(b, S));
            NB_HIP(hipGraphLaunch(b->gexec, b->stream));
            b->calls_since_sort += S;
        }
    }
    for (; k < nsteps; ++k) {
        NB_TRY(bmaybe_sort(b, 1));
        NB_TRY(enqueue_steps(b, 1, true));
    }
    return NB_OK;
}

static int collect(nb_batch *b) {
    for (auto &p : b->pending) {
        NB_HIP(hipEventSynchronize(p.second));
        float ms = 0.f;
        NB_HIP(hipEventElapsedTime(&ms, p.first, p.second));
        b->tms += ms;
        b->tcount++;
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    b->pending.clear();
    return NB_OK;
}

}  // namespace nb

using namespace nb;

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

int nb_batch_create(nb_batch **out, int32_t nsys, const int64_t *offsets, int32_t precision_mode) {
    clear_error();
    if (!out || !offsets || nsys < 0) {
        set_error("invalid batch arguments");
        return NB_ERR_INVALID;
    }
    *out = nullptr;
    if (precision_mode != NB_MODE_FP64 && precision_mode != NB_MODE_SPLIT &&
        precision_mode != NB_MODE_SPLIT_CENTRAL) {
        set_error("batch precision mode must be FP64, SPLIT or SPLIT_CENTRAL (got %d)",
                  precision_mode);
        return NB_ERR_INVALID;
    }
    if (offsets[0] != 0) {
        set_error("offsets[0] must be 0");
        return NB_ERR_INVALID;
    }
    int64_t max_ns = 0;
    for (int32_t s = 0; s < nsys; ++s) {
        const int64_t ns = offsets[s + 1] - offsets[s];
        if (ns < 0) {
            set_error("offsets must be non-decreasing (system %d)", s);
            return NB_ERR_INVALID;
        }
        max_ns = std::max(max_ns, ns);
    }
    if (offsets[nsys] > ((int64_t)1 << 31)) {
        set_error("too many particles");
        return NB_ERR_INVALID;
    }
    nb_batch *b = new (std::nothrow) nb_batch();
    if (!b) {
        set_error("host allocation failed");
        return NB_ERR_NOMEM;
    }
    b->nsys = nsys;
    b->mode = precision_mode;
    b->n = offsets[nsys];
    b->max_ns = max_ns;
    b->h_off.assign(offsets, offsets + nsys + 1);
    b->hG.assign(nsys, 1.0);
    b->hh.assign(nsys, 1e-4);
    b->hdt.assign(nsys, 1e-3);
    b->tun.targets_per_thread = 1;
    b->tun.blocks = 0;
    b->tun.steps_per_graph = 16;
    b->tun.sort_interval = precision_mode == NB_MODE_FP64 ? 0 : 16;
    b->tun.resident = 1;
    int rc = NB_OK;
    do {
        hipError_t e = hipGetDevice(&b->dev);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            set_error("device/stream setup failed: %s", hipGetErrorString(e));
            rc = NB_ERR_HIP;
            break;
        }
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, b->dev) == hipSuccess) b->cu_count = p.multiProcessorCount;
        const size_t n = (size_t)b->n, ns = (size_t)nsys;
        // tiles
        std::vector<int64_t> tb(ns + 1, 0);
        for (size_t s = 0; s < ns; ++s) tb[s + 1] = tb[s] + (b->h_off[s + 1] - b->h_off[s] + BS - 1) / BS;
        b->ntiles = tb[ns];
        std::vector<int32_t> tsys((size_t)b->ntiles);
        for (size_t s = 0; s < ns; ++s)
            for (int64_t t = tb[s]; t < tb[s + 1]; ++t) tsys[(size_t)t] = (int32_t)s;
        std::vector<int32_t> sys_of(n);
        for (size_t s = 0; s < ns; ++s)
            for (int64_t i = b->h_off[s]; i < b->h_off[s + 1]; ++i) sys_of[(size_t)i] = (int32_t)s;
        if ((rc = dalloc(&b->off, sizeof(int64_t) * (ns + 1))) ||
            (rc = dalloc(&b->sys_of, sizeof(int32_t) * n)) ||
            (rc = dalloc(&b->Gs, sizeof(double) * ns)) || (rc = dalloc(&b->hs, sizeof(double) * ns)) ||
            (rc = dalloc(&b->dts, sizeof(double) * ns)) || (rc = dalloc(&b->x, sizeof(double) * n)) ||
            (rc = dalloc(&b->y, sizeof(double) * n)) || (rc = dalloc(&b->z, sizeof(double) * n)) ||
            (rc = dalloc(&b->vx, sizeof(double) * n)) || (rc = dalloc(&b->vy, sizeof(double) * n)) ||
            (rc = dalloc(&b->vz, sizeof(double) * n)) || (rc = dalloc(&b->m, sizeof(double) * n)) ||
            (rc = dalloc(&b->h, sizeof(double) * n)) || (rc = dalloc(&b->ax, sizeof(double) * n)) ||
            (rc = dalloc(&b->ay, sizeof(double) * n)) || (rc = dalloc(&b->az, sizeof(double) * n)) ||
            (rc = dalloc(&b->pot, sizeof(double) * n)) || (rc = dalloc(&b->central, n)) ||
            (rc = dalloc(&b->tile_base, sizeof(int64_t) * (ns + 1))) ||
            (rc = dalloc(&b->tile_sys, sizeof(int32_t) * (size_t)std::max<int64_t>(b->ntiles, 1))) ||
            (rc = dalloc(&b->tile_box, sizeof(double) * 8 * (size_t)std::max<int64_t>(b->ntiles, 1))) ||
            (rc = dalloc(&b->sub_box, sizeof(double) * 8 * NQ * (size_t)std::max<int64_t>(b->ntiles, 1))) ||
            (rc = dalloc(&b->counter, sizeof(unsigned) * 2)) ||
            (rc = dalloc(&b->stage, sizeof(double) * 3 * n)) ||
            (rc = dalloc(&b->wsrc, sizeof(float) * std::max<int64_t>(n, 1))) ||
            (rc = dalloc(&b->red, sizeof(double) * 8 * std::max<size_t>(ns, 1))))
            break;
        e = hipMemcpy(b->off, offsets, sizeof(int64_t) * (ns + 1), hipMemcpyHostToDevice);
        if (e == hipSuccess && n) e = hipMemcpy(b->sys_of, sys_of.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(b->tile_base, tb.data(), sizeof(int64_t) * (ns + 1), hipMemcpyHostToDevice);
        if (e == hipSuccess && b->ntiles) e = hipMemcpy(b->tile_sys, tsys.data(), sizeof(int32_t) * tsys.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemset(b->counter, 0, sizeof(unsigned) * 2);
        if (e == hipSuccess && n) e = hipMemset(b->central, 0, n);
        if (e == hipSuccess && n) e = hipMemset(b->ax, 0, sizeof(double) * n);
        if (e == hipSuccess && n) e = hipMemset(b->ay, 0, sizeof(double) * n);
        if (e == hipSuccess && n) e = hipMemset(b->az, 0, sizeof(double) * n);
        if (e == hipSuccess && n) e = hipMemset(b->pot, 0, sizeof(double) * n);
        if (e != hipSuccess) {
            set_error("batch initialisation failed: %s", hipGetErrorString(e));
            rc = NB_ERR_HIP;
            break;
        }
        rc = nb_batch_set_params(b, b->hG.data(), b->hh.data(), b->hdt.data());
    } while (0);
    if (rc != NB_OK) {
        std::string keep = nb_last_error();
        nb_batch_destroy(b);
        set_error("%s", keep.c_str());
        return rc;
    }
    *out = b;
    return NB_OK;
}

int nb_batch_destroy(nb_batch *b) {
    clear_error();
    if (!b) return NB_OK;
    (void)hipSetDevice(b->dev);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto &p : b->pending) {
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    bdrop_graph(b);
    dfree(b->off); dfree(b->sys_of); dfree(b->Gs); dfree(b->hs); dfree(b->dts);
    dfree(b->x); dfree(b->y); dfree(b->z); dfree(b->vx); dfree(b->vy); dfree(b->vz);
    dfree(b->m); dfree(b->h); dfree(b->ax); dfree(b->ay); dfree(b->az); dfree(b->pot);
    dfree(b->perm); dfree(b->central); dfree(b->cidx);
    dfree(b->tile_base); dfree(b->tile_sys); dfree(b->tile_box); dfree(b->sub_box);
    dfree(b->items); dfree(b->counter); dfree(b->stage); dfree(b->red); dfree(b->wsrc);
    dfree(b->sort_tmp); dfree(b->keys); dfree(b->idx); dfree(b->sysbox); dfree(b->spare);
    dfree(b->pk_lane); dfree(b->pk_meta); dfree(b->pk_blk); dfree(b->pk_tiny);
    if (b->h_pinned) (void)hipHostFree(b->h_pinned);
    dfree(b->e_dev);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return NB_OK;
}

int nb_batch_set_params(nb_batch *b, const double *G, const double *soft, const double *dt) {
    NB_TRY(bcheck(b, false));
    if (!G || !soft || !dt) {
        set_error("null parameter array");
        return NB_ERR_INVALID;
    }
    for (int32_t s = 0; s < b->nsys; ++s) {
        if (!isfinite(G[s]) || !(soft[s] >= 0.0) || !isfinite(soft[s]) || !isfinite(dt[s])) {
            set_error("invalid parameters of system %d", s);
            return NB_ERR_INVALID;
        }
    }
    b->pk_stale = true;
    if (G != b->hG.data()) b->hG.assign(G, G + b->nsys);
    if (soft != b->hh.data()) b->hh.assign(soft, soft + b->nsys);
    if (dt != b->hdt.data()) b->hdt.assign(dt, dt + b->nsys);
    bdrop_graph(b);
    const size_t ns = (size_t)b->nsys;
    if (ns) {
        NB_HIP(hipMemcpy(b->Gs, b->hG.data(), sizeof(double) * ns, hipMemcpyHostToDevice));
        NB_HIP(hipMemcpy(b->hs, b->hh.data(), sizeof(double) * ns, hipMemcpyHostToDevice));
        NB_HIP(hipMemcpy(b->dts, b->hdt.data(), sizeof(double) * ns, hipMemcpyHostToDevice));
    }
    if (b->uploaded && b->n) {
        // per-particle h from the system table, in the current internal order
        std::vector<double> hh((size_t)b->n);
        std::vector<int32_t> so((size_t)b->n);
        NB_HIP(hipMemcpy(so.data(), b->sys_of, sizeof(int32_t) * so.size(), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < hh.size(); ++i) hh[i] = b->hh[(size_t)so[i]];
        NB_HIP(hipMemcpy(b->h, hh.data(), sizeof(double) * hh.size(), hipMemcpyHostToDevice));
    }
    return prepare_resident(b);
}

int nb_batch_upload(nb_batch *b, const double *pos, const double *vel, const double *mass) {
    NB_TRY(bcheck(b, false));
    const int64_t n = b->n;
    if (n > 0 && (!pos || !mass)) {
        set_error("pos and mass are required");
        return NB_ERR_INVALID;
    }
    bdrop_graph(b);
    dfree(b->perm);
    b->sorted = false;
    hipStream_t s = b->stream;
    if (n > 0) {
        NB_HIP(hipMemcpyAsync(b->stage, pos, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, s));
        NB_TRY(to_soa(b->stage, nullptr, n, b->x, b->y, b->z, true, s));
        NB_HIP(hipStreamSynchronize(s));
        if (vel) {
            NB_HIP(hipMemcpyAsync(b->stage, vel, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, s));
            NB_TRY(to_soa(b->stage, nullptr, n, b->vx, b->vy, b->vz, true, s));
        } else {
            NB_HIP(hipMemsetAsync(b->vx, 0, sizeof(double) * (size_t)n, s));
            NB_HIP(hipMemsetAsync(b->vy, 0, sizeof(double) * (size_t)n, s));
            NB_HIP(hipMemsetAsync(b->vz, 0, sizeof(double) * (size_t)n, s));
        }
        NB_HIP(hipMemcpyAsync(b->m, mass, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, s));
        NB_HIP(hipStreamSynchronize(s));
    }
    b->h_nm.assign((size_t)b->nsys, 0);
    b->h_mmax.assign((size_t)b->nsys, 0.0);
    b->h_massive.assign((size_t)n, 0);
    for (int32_t q = 0; q < b->nsys; ++q)
        for (int64_t i = b->h_off[q]; i < b->h_off[q + 1]; ++i) {
            b->h_massive[(size_t)i] = mass[i] != 0.0;
            b->h_nm[(size_t)q] += mass[i] != 0.0;
            b->h_mmax[(size_t)q] = std::max(b->h_mmax[(size_t)q], std::fabs(mass[i]));
        }
    b->pk_stale = true;
    NB_TRY(ensure_pinned(b, (size_t)std::max(b->nsys, 1) * 128));
    NB_TRY(ensure_edev(b, (size_t)std::max(b->nsys, 1) * 128));
    b->uploaded = true;
    NB_TRY(nb_batch_set_params(b, b->hG.data(), b->hh.data(), b->hdt.data()));
    NB_TRY(rebuild_central(b));
    NB_HIP(hipStreamSynchronize(s));
    return NB_OK;
}

int nb_batch_download(nb_batch *b, double *pos, double *vel, double *acc, double *pot) {
    NB_TRY(bcheck(b));
    const int64_t n = b->n;
    if (n == 0) return NB_OK;
    hipStream_t s = b->stream;
    const size_t b3 = sizeof(double) * 3 * (size_t)n;
    if (pos) {
        NB_TRY(to_aos(b->x, b->y, b->z, b->perm, n, true, b->stage, s));
        NB_HIP(hipMemcpyAsync(pos, b->stage, b3, hipMemcpyDeviceToHost, s));
        NB_HIP(hipStreamSynchronize(s));
    }
    if (vel) {
        NB_TRY(to_aos(b->vx, b->vy, b->vz, b->perm, n, true, b->stage, s));
        NB_HIP(hipMemcpyAsync(vel, b->stage, b3, hipMemcpyDeviceToHost, s));
        NB_HIP(hipStreamSynchronize(s));
    }
    if (acc) {
        NB_TRY(to_aos(b->ax, b->ay, b->az, b->perm, n, true, b->stage, s));
        NB_HIP(hipMemcpyAsync(acc, b->stage, b3, hipMemcpyDeviceToHost, s));
        NB_HIP(hipStreamSynchronize(s));
    }
    if (pot) {
        NB_TRY(scalar_out(b->pot, b->perm, n, b->stage, s));
        NB_HIP(hipMemcpyAsync(pot, b->stage, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, s));
        NB_HIP(hipStreamSynchronize(s));
    }
    return NB_OK;
}

int nb_batch_set_central(nb_batch *b, int32_t count) {
    NB_TRY(bcheck(b, false));
    if (count < 0 || (count > 0 && b->mode != NB_MODE_SPLIT_CENTRAL)) {
        set_error("central bodies need NB_MODE_SPLIT_CENTRAL and count >= 0");
        return NB_ERR_INVALID;
    }
    for (int32_t s = 0; s < b->nsys; ++s) {
        if (b->h_off[s + 1] - b->h_off[s] < count) {
            set_error("system %d has fewer than %d particles", s, count);
            return NB_ERR_INVALID;
        }
    }
    bdrop_graph(b);
    dfree(b->cidx);
    b->ncentral = count;
    b->pk_stale = true;
    NB_HIP(hipMemset(b->central, 0, (size_t)std::max<int64_t>(b->n, 1)));
    if (count > 0) {
        NB_TRY(dalloc(&b->cidx, sizeof(int64_t) * (size_t)b->nsys * (size_t)count));
        NB_TRY(rebuild_central(b));
        NB_HIP(hipStreamSynchronize(b->stream));
    }
    return prepare_resident(b);
}

int nb_batch_compute_force(nb_batch *b) {
    NB_TRY(bcheck(b));
    NB_TRY(bmaybe_sort(b, 1));
    NB_TRY(bforce(b, false, true));
    return NB_OK;
}

int nb_batch_compute_potential(nb_batch *b) {
    NB_TRY(bcheck(b));
    NB_TRY(bmaybe_sort(b, 1));
    NB_TRY(bforce(b, true, false, true));
    NB_HIP(hipStreamSynchronize(b->stream));
    return NB_OK;
}

int nb_batch_half_kick(nb_batch *b, int32_t forward) {
    NB_TRY(bcheck(b));
    if (forward != 1 && forward != -1) {
        set_error("forward must be +1 or -1");
        return NB_ERR_INVALID;
    }
    return bkick_drift(b, (double)forward * 0.5, false);
}

int nb_batch_kdk_steps(nb_batch *b, int64_t nsteps, int32_t use_graph) {
    NB_TRY(bcheck(b));
    if (nsteps < 0) {
        set_error("nsteps must be >= 0");
        return NB_ERR_INVALID;
    }
    return batch_steps(b, nsteps, use_graph != 0);
}

int nb_batch_kdk_steps_tracked(nb_batch *b, int64_t nsteps, int32_t every, double *energy) {
    NB_TRY(bcheck(b));
    if (nsteps < 0 || every <= 0 || !energy) {
        set_error("invalid tracked-step arguments");
        return NB_ERR_INVALID;
    }
    const int64_t slots = nsteps / every;
    const size_t ns = (size_t)b->nsys;
    if (slots == 0) return batch_steps(b, nsteps, true);
    if (resident_ok(b)) NB_TRY(batch_unsort(b));
    NB_TRY(ensure_edev(b, ns * (size_t)slots));
    double *dev = b->e_dev;
    int rc = NB_OK;
    if (resident_ok(b)) {
        rc = resident_steps(b, nsteps, every, dev, (int)slots);
    } else {
        std::vector<double> host(ns * 8);
        for (int64_t k = 0; k < slots && rc == NB_OK; ++k) {
            rc = batch_steps(b, every, true);
            if (rc == NB_OK) rc = bforce(b, true, false, true);   // potential alone
            if (rc == NB_OK) {
                k_batch_energy<<<b->nsys, 256, 0, b->stream>>>(
                    b->off, b->x, b->y, b->z, b->vx, b->vy, b->vz, b->ax, b->ay, b->az, b->m,
                    b->pot, b->dts, 1, b->red);
                rc = launch_ok("k_batch_energy");
            }
            if (rc == NB_OK) {
                hipError_t e = hipMemcpyAsync(host.data(), b->red, sizeof(double) * 8 * ns,
                                              hipMemcpyDeviceToHost, b->stream);
                if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
                if (e != hipSuccess) {
                    set_error("energy readback: %s", hipGetErrorString(e));
                    rc = NB_ERR_HIP;
                }
            }
            if (rc == NB_OK)
                for (size_t s = 0; s < ns; ++s)
                    energy[s * (size_t)slots + (size_t)k] = host[s * 8] + 0.5 * host[s * 8 + 1];
        }
        if (rc == NB_OK && nsteps % every) rc = batch_steps(b, nsteps % every, true);
        return rc;
    }
    if (rc == NB_OK) rc = ensure_pinned(b, ns * (size_t)slots);
    if (rc == NB_OK) {
        hipError_t e = hipMemcpyAsync(b->h_pinned, dev, sizeof(double) * ns * (size_t)slots,
                                      hipMemcpyDeviceToHost, b->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess) {
            set_error("energy readback: %s", hipGetErrorString(e));
            rc = NB_ERR_HIP;
        } else {
            std::memcpy(energy, b->h_pinned, sizeof(double) * ns * (size_t)slots);
        }
    }
    if (rc == NB_OK && nsteps % every) rc = resident_steps(b, nsteps % every, 0, nullptr, 0);
    (void)hipStreamSynchronize(b->stream);
    return rc;
}

int nb_batch_energy(nb_batch *b, double *K, double *W, double *P, double *L) {
    NB_TRY(bcheck(b));
    if (W) {
        NB_TRY(bmaybe_sort(b, 1));
        NB_TRY(bforce(b, true, false, true));
    }
    k_batch_energy<<<b->nsys, 256, 0, b->stream>>>(b->off, b->x, b->y, b->z, b->vx, b->vy, b->vz,
                                                   b->ax, b->ay, b->az, b->m, W ? b->pot : nullptr,
                                                   b->dts, 0, b->red);
    NB_TRY(launch_ok("k_batch_energy"));
    std::vector<double> host((size_t)b->nsys * 8);
    NB_HIP(hipMemcpyAsync(host.data(), b->red, sizeof(double) * host.size(), hipMemcpyDeviceToHost,
                          b->stream));
    NB_HIP(hipStreamSynchronize(b->stream));
    for (int32_t s = 0; s < b->nsys; ++s) {
        const double *r = host.data() + (size_t)s * 8;
        if (K) K[s] = r[0];
        if (W) W[s] = 0.5 * r[1];
        if (P) { P[3 * s] = r[2]; P[3 * s + 1] = r[3]; P[3 * s + 2] = r[4]; }
        if (L) { L[3 * s] = r[5]; L[3 * s + 1] = r[6]; L[3 * s + 2] = r[7]; }
    }
    // the outputs stay written; the first system with a NaN or Inf among the
    // requested quantities is reported as NB_ERR_NONFINITE
    for (int32_t s = 0; s < b->nsys; ++s) {
        const double *r = host.data() + (size_t)s * 8;
        const bool bad = (K && !std::isfinite(r[0])) || (W && !std::isfinite(r[1])) ||
                         (P && !(std::isfinite(r[2]) && std::isfinite(r[3]) && std::isfinite(r[4]))) ||
                         (L && !(std::isfinite(r[5]) && std::isfinite(r[6]) && std::isfinite(r[7])));
        if (bad) {
            set_error("non-finite energy or momentum in system %d (NaN or Inf in its state)", s);
            return NB_ERR_NONFINITE;
        }
    }
    return NB_OK;
}

int nb_batch_sync(nb_batch *b) {
    NB_TRY(bcheck(b, false));
    NB_HIP(hipStreamSynchronize(b->stream));
    return NB_OK;
}

int nb_batch_set_tuning(nb_batch *b, const nb_batch_tuning *t) {
    NB_TRY(bcheck(b, false));
    if (!t || (t->targets_per_thread != 1 && t->targets_per_thread != 2 &&
               t->targets_per_thread != 4) ||
        t->blocks < 0 || t->blocks > (1 << 20) || t->steps_per_graph < 2 ||
        t->steps_per_graph > 4096 || t->sort_interval < 0 || t->far_path < 0 || t->far_path > 4 ||
        t->resident_waves < 0 || t->resident_waves > 4) {
        set_error("invalid batch tuning");
        return NB_ERR_INVALID;
    }
    if ((t->far_path == FAR_MFMA_SUM || t->far_path == FAR_MFMA_R2 ||
         t->far_path == FAR_MFMA_BF16) && t->targets_per_thread != 1) {
        set_error("far_path %d needs targets_per_thread 1 (got %d)", t->far_path,
                  t->targets_per_thread);
        return NB_ERR_INVALID;
    }
    bdrop_graph(b);
    b->tun = *t;
    return prepare_resident(b);
}

int nb_batch_get_tuning(nb_batch *b, nb_batch_tuning *t) {
    NB_TRY(bcheck(b, false));
    if (!t) {
        set_error("null output");
        return NB_ERR_INVALID;
    }
    *t = b->tun;
    return NB_OK;
}

int nb_batch_autotune(nb_batch *b) {
    NB_TRY(bcheck(b));
    if (b->mode == NB_MODE_FP64 || b->n == 0) return NB_OK;
    NB_TRY(bmaybe_sort(b, 1));
    const nb_batch_tuning keep = b->tun;
    nb_batch_tuning best = keep;
    float best_ms = 1e30f;
    const int mult[4] = {2, 4, 8, 16};
    hipEvent_t e0, e1;
    NB_HIP(hipEventCreate(&e0));
    NB_HIP(hipEventCreate(&e1));
    int rc = NB_OK;
    // (targets per thread, far path) as in the one-system autotuner
    const int cand[9][2] = {{1, 0}, {2, 0}, {4, 0}, {1, 3}, {2, 3}, {4, 3}, {1, 4}, {1, 1}, {1, 2}};
    for (int ti = 0; ti < 9 && rc == NB_OK; ++ti) {
        for (int mi = 0; mi < 4 && rc == NB_OK; ++mi) {
            b->tun.targets_per_thread = cand[ti][0];
            b->tun.far_path = cand[ti][1];
            b->tun.blocks = mult[mi] * b->cu_count;
            rc = bforce(b, false, false);
            if (rc) break;
            (void)hipEventRecord(e0, b->stream);
            for (int r = 0; r < 3 && rc == NB_OK; ++r) rc = bforce(b, false, false);
            (void)hipEventRecord(e1, b->stream);
            (void)hipEventSynchronize(e1);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < best_ms) {
                best_ms = ms;
                best = b->tun;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->tun = best;
    bdrop_graph(b);
    if (rc == NB_OK) rc = bforce(b, false, false);
    if (rc == NB_OK) {
        hipError_t e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess) {
            set_error("autotune sync: %s", hipGetErrorString(e));
            rc = NB_ERR_HIP;
        }
    }
    return rc;
}

int nb_batch_set_timing(nb_batch *b, int32_t enable) {
    NB_TRY(bcheck(b, false));
    NB_TRY(collect(b));
    b->timing = enable != 0;
    b->tcount = 0;
    b->tms = 0.0;
    return NB_OK;
}

int nb_batch_get_timing(nb_batch *b, int64_t *count, double *total_ms) {
    NB_TRY(bcheck(b, false));
    NB_TRY(collect(b));
    if (count) *count = b->tcount;
    if (total_ms) *total_ms = b->tms;
    return NB_OK;
}

}  // extern "C"
