// nb_force.hip -- all-pairs gravity kernels for one system (gfx950).
//
// Replaces compute_force_gpu_kernel  src/directsumgpu/gravity.jl:22-34 (one
// thread per target, sources re-read from global memory) with an LDS-tiled
// kernel: each block owns BS*TPT targets held in registers and streams the
// sources through a double-buffered LDS stage of BS particles; every lane
// reads the same source (LDS broadcast).  Pair arithmetic is FP32 (v_rsq_f32,
// no MFMA); each tile's FP32 partial sums are folded into FP64 accumulators.
//
// Precision of close pairs (SPLIT): positions are FP64 in HBM; the block
// picks an FP64 origin (the centroid of its targets) and both targets and the
// staged sources are converted to FP32 *relative to that origin*, so the
// rounding error of a difference x_j - x_i is set by the block's extent, not
// by the absolute coordinate.
//
// Softening test: tiles whose bounding box is separated from the block's
// target box by more than the largest target h are evaluated without the
// per-pair r < h test (pure Newtonian); all other tiles take the checked path.
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

constexpr int NW = BS / 64;

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

template <int TPT, bool CHECKED, bool POT>
__device__ __forceinline__ void tile_f32(const float4 *__restrict__ s, const float (&xi)[TPT],
                                         const float (&yi)[TPT], const float (&zi)[TPT],
                                         const float (&h2c)[TPT], const float (&h2)[TPT],
                                         const float (&hinv)[TPT], const float (&hinv3)[TPT],
                                         float (&ax)[TPT], float (&ay)[TPT], float (&az)[TPT],
                                         float (&aq)[TPT]) {
#pragma unroll 4
    for (int jj = 0; jj < BS; ++jj) {
        const float4 v = s[jj];
#pragma unroll
        for (int k = 0; k < TPT; ++k)
            pair_f32<CHECKED, POT>(v.x, v.y, v.z, v.w, xi[k], yi[k], zi[k], h2c[k], h2[k], hinv[k],
                                   hinv3[k], ax[k], ay[k], az[k], aq[k]);
    }
}

// FP32-pair kernel.  T = double (SPLIT, FP64 state) or float (FP32 state).
template <typename T, int TPT, bool POT>
__global__ __launch_bounds__(BS) void k_force_f32(ForceArgs<T> a) {
    __shared__ float4 sj[2][BS];
    __shared__ double red[11][NW];

    const int tid = threadIdx.x;
    const int logical = xcd_remap(blockIdx.x, gridDim.x);
    const int seg = logical / a.nblk_i;
    const int ib = logical - seg * a.nblk_i;
    const int64_t i0 = (int64_t)ib * (BS * TPT);

    double px[TPT], py[TPT], pz[TPT], ph[TPT];
    bool valid[TPT];
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int64_t i = i0 + k * BS + tid;
        valid[k] = i < a.nt;
        px[k] = valid[k] ? (double)a.tx[i] : 0.0;
        py[k] = valid[k] ? (double)a.ty[i] : 0.0;
        pz[k] = valid[k] ? (double)a.tz[i] : 0.0;
        ph[k] = valid[k] ? (a.th ? a.th[i] : a.th_const) : 0.0;
    }
    BlockStats st;
    block_stats<TPT>(px, py, pz, ph, valid, a.use_origin != 0, red, st);
    const double ox = st.ox, oy = st.oy, oz = st.oz;

    float xi[TPT], yi[TPT], zi[TPT], h2c[TPT], h2[TPT], hinv[TPT], hinv3[TPT];
    float ax[TPT], ay[TPT], az[TPT], aq[TPT];
    double Ax[TPT], Ay[TPT], Az[TPT], Aq[TPT];
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        xi[k] = (float)(px[k] - ox);
        yi[k] = (float)(py[k] - oy);
        zi[k] = (float)(pz[k] - oz);
        const float hf = (float)ph[k];
        h2[k] = hf * hf;
        h2c[k] = fmaxf(h2[k], FLT_MIN);
        hinv[k] = 1.0f / hf;
        hinv3[k] = hinv[k] * hinv[k] * hinv[k];
        ax[k] = ay[k] = az[k] = aq[k] = 0.0f;
        Ax[k] = Ay[k] = Az[k] = Aq[k] = 0.0;
    }
    // Target box grown by the largest h plus a margin that keeps every pair
    // of an unchecked tile at FP32-resolvable, nonzero separation.
    double rb = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double o = d == 0 ? ox : (d == 1 ? oy : oz);
        rb = fmax(rb, fmax(fabs(st.lo[d] - o), fabs(st.hi[d] - o)));
    }
    const double grow = st.hmax * (1.0 + 1e-6) + 1e-6 * rb + 1e-300;
    const double glo0 = st.lo[0] - grow, glo1 = st.lo[1] - grow, glo2 = st.lo[2] - grow;
    const double ghi0 = st.hi[0] + grow, ghi1 = st.hi[1] + grow, ghi2 = st.hi[2] + grow;

    const int64_t j0 = (int64_t)seg * a.seg_len;
    const int64_t j1 = (j0 + a.seg_len < a.ns) ? j0 + a.seg_len : a.ns;

    // Register prefetch of the first tile.
    T rx = 0, ry = 0, rz = 0, rm = 0;
    bool rv = false;
    {
        const int64_t j = j0 + tid;
        rv = j < j1 && !(a.skip_src && a.skip_src[j]);
        if (rv) { rx = a.sx[j]; ry = a.sy[j]; rz = a.sz[j]; rm = a.sm[j]; }
    }
    int it = 0;
    for (int64_t base = j0; base < j1; base += BS, ++it) {
        const int buf = it & 1;
        sj[buf][tid] = rv ? make_float4((float)((double)rx - ox), (float)((double)ry - oy),
                                        (float)((double)rz - oz), (float)rm)
                          : make_float4(PAD_COORD, PAD_COORD, PAD_COORD, 0.0f);
        __syncthreads();
        {
            const int64_t j = base + BS + tid;
            rv = j < j1 && !(a.skip_src && a.skip_src[j]);
            if (rv) { rx = a.sx[j]; ry = a.sy[j]; rz = a.sz[j]; rm = a.sm[j]; }
        }
        const double *tb = a.tile_box + (base / BS) * 8;
        const bool near = !(tb[0] > ghi0 || tb[3] < glo0 || tb[1] > ghi1 || tb[4] < glo1 ||
                            tb[2] > ghi2 || tb[5] < glo2);
        if (near)
            tile_f32<TPT, true, POT>(sj[buf], xi, yi, zi, h2c, h2, hinv, hinv3, ax, ay, az, aq);
        else
            tile_f32<TPT, false, POT>(sj[buf], xi, yi, zi, h2c, h2, hinv, hinv3, ax, ay, az, aq);
#pragma unroll
        for (int k = 0; k < TPT; ++k) {
            Ax[k] += (double)ax[k]; Ay[k] += (double)ay[k]; Az[k] += (double)az[k];
            ax[k] = ay[k] = az[k] = 0.0f;
            if (POT) { Aq[k] += (double)aq[k]; aq[k] = 0.0f; }
        }
    }
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        if (!valid[k]) continue;
        const int64_t i = i0 + k * BS + tid;
        if (a.nseg == 1) {
            a.ax[i] = a.G * Ax[k];
            a.ay[i] = a.G * Ay[k];
            a.az[i] = a.G * Az[k];
            if (POT) a.pot[i] = -a.G * Aq[k];
        } else {
            double *p = a.part + (int64_t)seg * 4 * a.nt;
            p[i] = Ax[k];
            p[a.nt + i] = Ay[k];
            p[2 * a.nt + i] = Az[k];
            if (POT) p[3 * a.nt + i] = Aq[k];
        }
    }
}

// FP64-pair kernel (correctness mode, and the FP64 central-body terms):
// same tiling, double arithmetic, loop bounded by the valid sources.
template <bool POT>
__global__ __launch_bounds__(BS) void k_force_f64(ForceArgs<double> a) {
    __shared__ double4 sj[2][BS];
    const int tid = threadIdx.x;
    const int logical = xcd_remap(blockIdx.x, gridDim.x);
    const int seg = logical / a.nblk_i;
    const int ib = logical - seg * a.nblk_i;
    const int64_t i = (int64_t)ib * BS + tid;
    const bool valid = i < a.nt;
    const double xi = valid ? a.tx[i] : 0.0, yi = valid ? a.ty[i] : 0.0, zi = valid ? a.tz[i] : 0.0;
    const double h = valid ? (a.th ? a.th[i] : a.th_const) : 1.0;
    const double hinv = 1.0 / h, hinv3 = hinv * hinv * hinv;
    double Ax = 0.0, Ay = 0.0, Az = 0.0, Aq = 0.0;
    const int64_t j0 = (int64_t)seg * a.seg_len;
    const int64_t j1 = (j0 + a.seg_len < a.ns) ? j0 + a.seg_len : a.ns;
    int it = 0;
    for (int64_t base = j0; base < j1; base += BS, ++it) {
        const int buf = it & 1;
        const int64_t j = base + tid;
        const bool rv = j < j1 && !(a.skip_src && a.skip_src[j]);
        sj[buf][tid] = rv ? make_double4(a.sx[j], a.sy[j], a.sz[j], a.sm[j])
                          : make_double4(1.0e150, 1.0e150, 1.0e150, 0.0);
        __syncthreads();
        const int cnt = (j1 - base < BS) ? (int)(j1 - base) : BS;
#pragma unroll 2
        for (int jj = 0; jj < cnt; ++jj) {
            const double4 v = sj[buf][jj];
            pair_f64<POT>(v.x, v.y, v.z, v.w, xi, yi, zi, h, hinv, hinv3, Ax, Ay, Az, Aq);
        }
    }
    if (!valid) return;
    if (a.nseg == 1) {
        if (a.accumulate) {
            a.ax[i] += Ax * a.G;
            a.ay[i] += Ay * a.G;
            a.az[i] += Az * a.G;
            if (POT) a.pot[i] -= Aq * a.G;
        } else {
            a.ax[i] = Ax * a.G;
            a.ay[i] = Ay * a.G;
            a.az[i] = Az * a.G;
            if (POT) a.pot[i] = -Aq * a.G;
        }
    } else {
        double *p = a.part + (int64_t)seg * 4 * a.nt;
        p[i] = Ax;
        p[a.nt + i] = Ay;
        p[2 * a.nt + i] = Az;
        if (POT) p[3 * a.nt + i] = Aq;
    }
}

// Sum the per-segment partials in fixed segment order and apply G.
__global__ void k_finalize(const double *__restrict__ part, int nseg, int64_t n, double G,
                           double *__restrict__ ax, double *__restrict__ ay,
                           double *__restrict__ az, double *__restrict__ pot, int with_pot) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sx = 0.0, sy = 0.0, sz = 0.0, sq = 0.0;
    for (int s = 0; s < nseg; ++s) {
        const double *p = part + (int64_t)s * 4 * n;
        sx += p[i];
        sy += p[n + i];
        sz += p[2 * n + i];
        if (with_pot) sq += p[3 * n + i];
    }
    ax[i] = G * sx;
    ay[i] = G * sy;
    az[i] = G * sz;
    if (with_pot) pot[i] = -G * sq;
}

// Bounding box of each BS-source tile (lo xyz, hi xyz).
template <typename T>
__global__ __launch_bounds__(BS) void k_tile_boxes(const T *__restrict__ x, const T *__restrict__ y,
                                                   const T *__restrict__ z, int64_t n,
                                                   double *__restrict__ box) {
    __shared__ double red[6][NW];
    const int64_t j = (int64_t)blockIdx.x * BS + threadIdx.x;
    const bool v = j < n;
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    if (v) {
        lo[0] = hi[0] = (double)x[j];
        lo[1] = hi[1] = (double)y[j];
        lo[2] = hi[2] = (double)z[j];
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
        for (int ww = 1; ww < NW; ++ww) r = q < 3 ? fmin(r, red[q][ww]) : fmax(r, red[q][ww]);
        box[(int64_t)blockIdx.x * 8 + q] = r;
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of %s failed: %s", what, hipGetErrorString(e));
        return NB_ERR_HIP;
    }
    return NB_OK;
}

int launch_tile_boxes(const void *x, const void *y, const void *z, int is_double, int64_t n,
                      double *box, hipStream_t s) {
    if (n <= 0) return NB_OK;
    const int64_t ntiles = (n + BS - 1) / BS;
    if (is_double)
        k_tile_boxes<double><<<(unsigned)ntiles, BS, 0, s>>>((const double *)x, (const double *)y,
                                                            (const double *)z, n, box);
    else
        k_tile_boxes<float><<<(unsigned)ntiles, BS, 0, s>>>((const float *)x, (const float *)y,
                                                           (const float *)z, n, box);
    return check_launch("k_tile_boxes");
}

template <typename T>
static int launch_f32_impl(const ForceArgs<T> &a, int tpt, bool pot, hipStream_t s) {
    const int64_t nb = (int64_t)a.nblk_i * a.nseg;
    if (nb <= 0) return NB_OK;
    if (nb > INT32_MAX) {
        set_error("force grid too large (%lld blocks)", (long long)nb);
        return NB_ERR_INVALID;
    }
    dim3 g((unsigned)nb), b(BS);
#define NB_L(TPT_, POT_) k_force_f32<T, TPT_, POT_><<<g, b, 0, s>>>(a)
    if (tpt == 1) { if (pot) NB_L(1, true); else NB_L(1, false); }
    else if (tpt == 2) { if (pot) NB_L(2, true); else NB_L(2, false); }
    else if (tpt == 4) { if (pot) NB_L(4, true); else NB_L(4, false); }
    else {
        set_error("unsupported targets_per_thread %d", tpt);
        return NB_ERR_INVALID;
    }
#undef NB_L
    return check_launch("k_force_f32");
}

int launch_force_f32(const ForceArgs<float> &a, int tpt, bool pot, hipStream_t s) {
    return launch_f32_impl<float>(a, tpt, pot, s);
}
int launch_force_split(const ForceArgs<double> &a, int tpt, bool pot, hipStream_t s) {
    return launch_f32_impl<double>(a, tpt, pot, s);
}

int launch_force_f64(const ForceArgs<double> &a, bool pot, hipStream_t s) {
    const int64_t nb = (int64_t)a.nblk_i * a.nseg;
    if (nb <= 0) return NB_OK;
    if (pot) k_force_f64<true><<<(unsigned)nb, BS, 0, s>>>(a);
    else k_force_f64<false><<<(unsigned)nb, BS, 0, s>>>(a);
    return check_launch("k_force_f64");
}

int launch_finalize(const double *part, int nseg, int64_t n, double G, double *ax, double *ay,
                    double *az, double *pot, bool with_pot, hipStream_t s) {
    if (n <= 0) return NB_OK;
    const unsigned nb = (unsigned)((n + 255) / 256);
    k_finalize<<<nb, 256, 0, s>>>(part, nseg, n, G, ax, ay, az, pot, with_pot ? 1 : 0);
    return check_launch("k_finalize");
}

}  // namespace nb
