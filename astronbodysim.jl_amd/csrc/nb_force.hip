// nb_force.hip -- all-pairs gravity kernels for one system (gfx950).
//
// Replaces compute_force_gpu_kernel  src/directsumgpu/gravity.jl:22-34 (one
// thread per target, sources re-read from global memory) with an LDS-tiled
// kernel: each block owns BS*TPT targets held in registers and streams the
// sources through a double-buffered LDS stage of BS particles; every lane
// reads the same source (LDS broadcast).  Pair arithmetic is FP32 (v_rsq_f32);
// each tile's FP32 partial sums are folded into FP64 accumulators.  The block
// routine itself, with its per-wave tile classes (softening-checked, hi/lo
// Newtonian, far) and the optional matrix-core far tiles, is force_block in
// nb_block.h.
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "nb_block.h"
#include "nb_internal.h"
#include "nb_pair.h"

namespace nb {

// FP32-pair kernel.  T = double (SPLIT: FP64 state, per-block origin, hi/lo
// coordinates on near tiles) or float (FP32 state: absolute coordinates, or,
// with the bf16 far tiles, the SPLIT layout of the FP32 positions).  MF
// selects the matrix-core far tiles (TPT 1, forces or potentials alone).
// ACC = false writes the potential only.
// PASS: every tile (PASS_ALL), or the far half of a split bf16 launch
// (PASS_FAR): the far tiles' sums, with the waves that left near tiles out
// flagged in a.nflag (which holds zeros between launches) for k_force_near.
template <typename T, int TPT, bool POT, int MF = 0, bool ACC = true, int PASS = PASS_ALL>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(MF == 4 ? 4 : 3))) void k_force_f32(ForceArgs<T> a) {
    __shared__ BlockShared<T, MF> sh;
    const int tid = threadIdx.x;
    const int logical = xcd_remap(blockIdx.x, gridDim.x);
    const int seg = logical / a.nblk_i;
    const int ib = logical - seg * a.nblk_i;
    const int64_t i0 = (int64_t)ib * (BS * TPT);
    const int64_t j0 = (int64_t)seg * a.seg_len;
    const int64_t j1 = (j0 + a.seg_len < a.ns) ? j0 + a.seg_len : a.ns;
    double Ax[TPT], Ay[TPT], Az[TPT], Aq[TPT];
    bool valid[TPT];
    const bool had_near = force_block<T, TPT, POT, MF, ACC, PASS>(sh, a.tx, a.ty, a.tz, a.th, a.th_const, i0, a.nt, a.sx, a.sy, a.sz,
                             a.sm, a.sw, j0, j1, a.tile_box + (j0 / BS) * 8,
                             a.sub_box ? a.sub_box + (j0 / BS) * (NQ * 8) : nullptr,
                             a.tgeo ? a.tgeo + (int64_t)ib * TG_STRIDE : nullptr,
                             a.tbg ? a.tbg + (int64_t)ib * BS * 16 : nullptr, a.skip_src,
                             a.use_origin != 0, Ax, Ay, Az, Aq, valid);
    // (one plain byte store per wave that left near tiles out)
    if (PASS == PASS_FAR && had_near && (tid & 63) == 0) a.nflag[(int64_t)logical * NW + (tid >> 6)] = 1;
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        if (!valid[k]) continue;
        const int64_t i = target_of<TPT>(i0, k, tid);
        if (a.nseg == 1) {
            if (ACC) {
                a.ax[i] = a.G * Ax[k];
                a.ay[i] = a.G * Ay[k];
                a.az[i] = a.G * Az[k];
            }
            if (POT) a.pot[i] = -a.G * Aq[k];
        } else {
            double *p = a.part + (int64_t)seg * 4 * a.nt;
            if (ACC) {
                p[i] = Ax[k];
                p[a.nt + i] = Ay[k];
                p[2 * a.nt + i] = Az[k];
            }
            if (POT) p[3 * a.nt + i] = Aq[k];
        }
    }
}

// Near pass of a split bf16 launch (forces only): one wavefront per
// workgroup, one workgroup per (force block, wave) of the far pass.  The
// waves the far pass flagged add their near quarters (near_wave) to the same
// outputs and clear their flags; the rest return at once.  Each wave runs on
// its own, so a wave with many near tiles holds no other wave's resources.
template <typename T>
__global__ __launch_bounds__(64) void k_force_near(ForceArgs<T> a) {
    __shared__ NearShared<T, FAR_MFMA_BF16> sh;
    const int lw = xcd_remap(blockIdx.x, gridDim.x);
    if (!a.nflag[lw]) return;   // wave-uniform
    if (threadIdx.x == 0) a.nflag[lw] = 0;
    const int logical = lw / NW, wave = lw - logical * NW;
    const int seg = logical / a.nblk_i;
    const int ib = logical - seg * a.nblk_i;
    const int64_t i0 = (int64_t)ib * BS;
    const int64_t j0 = (int64_t)seg * a.seg_len;
    const int64_t j1 = (j0 + a.seg_len < a.ns) ? j0 + a.seg_len : a.ns;
    double Ax, Ay, Az;
    bool valid;
    near_wave<T, FAR_MFMA_BF16>(sh, wave, a.tx, a.ty, a.tz, a.th, a.th_const, i0, a.nt, a.sx, a.sy, a.sz, a.sm,
                                j0, j1, a.tile_box + (j0 / BS) * 8,
                                a.sub_box ? a.sub_box + (j0 / BS) * (NQ * 8) : nullptr,
                                a.tgeo + (int64_t)ib * TG_STRIDE, a.skip_src, Ax, Ay, Az, valid);
    if (!valid) return;
    const int64_t i = i0 + wave * 64 + (threadIdx.x & 63);
    if (a.nseg == 1) {
        a.ax[i] += a.G * Ax;
        a.ay[i] += a.G * Ay;
        a.az[i] += a.G * Az;
    } else {
        double *p = a.part + (int64_t)seg * 4 * a.nt;
        p[i] += Ax;
        p[a.nt + i] += Ay;
        p[2 * a.nt + i] += Az;
    }
}

// The bf16 matrix-core force launch (forces only, 1 target per lane): one
// kernel, or the far and near passes when the caller provides the flags
// and the target blocks' stored geometry (several source segments).
// The split pays where near tiles are few (FP64 state: the collision at 2^20
// has 1.8 % of its pairs in them, and its far loop alone runs about 10 %
// faster); with FP32 state (the FP32 Plummer sphere at 2^16: 18 %) the near
// pass cost more than it saved, so that launch stays in one kernel.
template <typename T>
static void launch_bf16_forces(const ForceArgs<T> &a, int64_t nb, hipStream_t s) {
    if constexpr (sizeof(T) == 8) {
        if (a.nflag && a.tgeo) {
            k_force_f32<T, 1, false, FAR_MFMA_BF16, true, PASS_FAR><<<(unsigned)nb, BS, 0, s>>>(a);
            k_force_near<T><<<(unsigned)(nb * NW), 64, 0, s>>>(a);
            return;
        }
    }
    k_force_f32<T, 1, false, FAR_MFMA_BF16><<<(unsigned)nb, BS, 0, s>>>(a);
}

// The origin and wave geometry of each target block (force_block's first
// steps), stored once per force call for launches whose target blocks meet
// several source segments.  hilo: the force kernel's near tiles have a
// close class (FP64 state).
template <typename T, int TPT, bool ROWS>
__global__ __launch_bounds__(BS) void k_target_geom(ForceArgs<T> a, bool hilo) {
    __shared__ double red[11][NW];
    const int tid = threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.x * (BS * TPT);
    double px[TPT], py[TPT], pz[TPT], ph[TPT];
    bool valid[TPT];
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int64_t i = target_of<TPT>(i0, k, tid);
        valid[k] = i < a.nt;
        px[k] = valid[k] ? (double)a.tx[i] : 0.0;
        py[k] = valid[k] ? (double)a.ty[i] : 0.0;
        pz[k] = valid[k] ? (double)a.tz[i] : 0.0;
        ph[k] = valid[k] ? (a.th ? a.th[i] : a.th_const) : 0.0;
    }
    BlockStats st;
    block_stats<TPT>(px, py, pz, ph, valid, a.use_origin != 0, red, st);
    const double ox = uniform_f64(st.ox), oy = uniform_f64(st.oy), oz = uniform_f64(st.oz);
    const WaveGeom wg = wave_geometry<TPT>(px, py, pz, ph, valid, ox, oy, oz, hilo);
    store_target_geom(a.tgeo + (int64_t)blockIdx.x * TG_STRIDE, ox, oy, oz, wg, tid >> 6, tid & 63);
    if constexpr (ROWS) {   // the lanes' bf16 B operands (TPT 1)
        nb_bf8 Bg[4];
        target_b_operands(px[0] - ox, py[0] - oy, pz[0] - oz, tid & 63, Bg);
        nb_u4v *b = reinterpret_cast<nb_u4v *>(a.tbg + ((int64_t)blockIdx.x * BS + tid) * 16);
#pragma unroll
        for (int g = 0; g < 4; ++g) b[g] = __builtin_bit_cast(nb_u4v, Bg[g]);
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
                           double *__restrict__ az, double *__restrict__ pot, int with_pot,
                           int with_acc) {
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
    if (with_acc) {
        ax[i] = G * sx;
        ay[i] = G * sy;
        az[i] = G * sz;
    }
    if (with_pot) pot[i] = -G * sq;
}

// Bounding box of each BS-source tile (lo xyz, hi xyz) and, when sub is not
// null, of each of its NQ quarters (the sources of one wave; an empty quarter
// gets lo = DBL_MAX, hi = -DBL_MAX).  One block per tile; thread t holds the
// position of source BS * blockIdx.x + t (v: the source exists).
__device__ __forceinline__ void tile_box_block(double px, double py, double pz, bool v,
                                               double *__restrict__ box, double *__restrict__ sub,
                                               double (&red)[6][NW]) {
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    if (v) {
        lo[0] = hi[0] = px;
        lo[1] = hi[1] = py;
        lo[2] = hi[2] = pz;
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
    if (sub && lane == 0) {
        double *o = sub + ((int64_t)blockIdx.x * NQ + w) * 8;
#pragma unroll
        for (int d = 0; d < 3; ++d) { o[d] = lo[d]; o[3 + d] = hi[d]; }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int q = threadIdx.x;
        double r = red[q][0];
        for (int ww = 1; ww < NW; ++ww) r = q < 3 ? fmin(r, red[q][ww]) : fmax(r, red[q][ww]);
        box[(int64_t)blockIdx.x * 8 + q] = r;
    }
}

template <typename T>
__global__ __launch_bounds__(BS) void k_tile_boxes(const T *__restrict__ x, const T *__restrict__ y,
                                                   const T *__restrict__ z, int64_t n,
                                                   double *__restrict__ box, double *__restrict__ sub) {
    __shared__ double red[6][NW];
    const int64_t j = (int64_t)blockIdx.x * BS + threadIdx.x;
    const bool v = j < n;
    tile_box_block(v ? (double)x[j] : 0.0, v ? (double)y[j] : 0.0, v ? (double)z[j] : 0.0, v, box, sub, red);
}

// A step's drift (Pos += Vel * (t_new - t_old), the operations of k_drift,
// with block 0 publishing t_new = time[rd] + dt to time[wr]) fused with the
// tile boxes of the drifted positions, which the next force needs.
// GEO: the block's target geometry as k_target_geom<T, 1, ROWS> stores it
// (the force call that follows has one target per lane and several source
// segments), from the drifted positions.
template <typename T, bool GEO, bool ROWS>
__global__ __launch_bounds__(BS) void k_drift_boxes(T *__restrict__ x, T *__restrict__ y, T *__restrict__ z,
                                                    const T *__restrict__ vx, const T *__restrict__ vy,
                                                    const T *__restrict__ vz, int64_t n, double dt,
                                                    double *time, int rd, int wr, double *__restrict__ box,
                                                    double *__restrict__ sub, const double *__restrict__ h,
                                                    double *__restrict__ tgeo, unsigned *__restrict__ tbg,
                                                    int use_origin, bool hilo) {
    __shared__ double red[6][NW];
    __shared__ double gred[GEO ? 11 : 1][NW];
    const int64_t j = (int64_t)blockIdx.x * BS + threadIdx.x;
    const double t_old = time[rd];
    const double t_new = t_old + dt;
    if (j == 0) time[wr] = t_new;
    const bool v = j < n;
    T px = 0, py = 0, pz = 0;
    if (v) {
        const T d = (T)(t_new - t_old);
        px = x[j]; py = y[j]; pz = z[j];
        px += vx[j] * d;
        py += vy[j] * d;
        pz += vz[j] * d;
        x[j] = px; y[j] = py; z[j] = pz;
    }
    tile_box_block((double)px, (double)py, (double)pz, v, box, sub, red);
    if constexpr (GEO) {
        const int tid = threadIdx.x;
        double gx[1] = {v ? (double)px : 0.0}, gy[1] = {v ? (double)py : 0.0}, gz[1] = {v ? (double)pz : 0.0};
        double gh[1] = {v ? h[j] : 0.0};
        bool gv[1] = {v};
        BlockStats st;
        block_stats<1>(gx, gy, gz, gh, gv, use_origin != 0, gred, st);
        const double ox = uniform_f64(st.ox), oy = uniform_f64(st.oy), oz = uniform_f64(st.oz);
        const WaveGeom wg = wave_geometry<1>(gx, gy, gz, gh, gv, ox, oy, oz, hilo);
        store_target_geom(tgeo + (int64_t)blockIdx.x * TG_STRIDE, ox, oy, oz, wg, tid >> 6, tid & 63);
        if constexpr (ROWS) {
            nb_bf8 Bg[4];
            target_b_operands(gx[0] - ox, gy[0] - oy, gz[0] - oz, tid & 63, Bg);
            nb_u4v *b = reinterpret_cast<nb_u4v *>(tbg + ((int64_t)blockIdx.x * BS + tid) * 16);
#pragma unroll
            for (int g = 0; g < 4; ++g) b[g] = __builtin_bit_cast(nb_u4v, Bg[g]);
        }
    }
}

// k_finalize followed by the step's kick (Vel += Acc * coef, the operations
// of k_kick) on the accelerations it has just written.
template <typename T>
__global__ void k_finalize_kick(const double *__restrict__ part, int nseg, int64_t n, double G,
                                double *__restrict__ ax, double *__restrict__ ay, double *__restrict__ az,
                                T *__restrict__ vx, T *__restrict__ vy, T *__restrict__ vz, double coef) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int s = 0; s < nseg; ++s) {
        const double *p = part + (int64_t)s * 4 * n;
        sx += p[i];
        sy += p[n + i];
        sz += p[2 * n + i];
    }
    const double fx = G * sx, fy = G * sy, fz = G * sz;
    ax[i] = fx;
    ay[i] = fy;
    az[i] = fz;
    const T c = (T)coef;
    vx[i] += (T)fx * c;
    vy[i] += (T)fy * c;
    vz[i] += (T)fz * c;
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
                      double *box, double *sub_box, hipStream_t s) {
    if (n <= 0) return NB_OK;
    const int64_t ntiles = (n + BS - 1) / BS;
    if (is_double)
        k_tile_boxes<double><<<(unsigned)ntiles, BS, 0, s>>>((const double *)x, (const double *)y,
                                                            (const double *)z, n, box, sub_box);
    else
        k_tile_boxes<float><<<(unsigned)ntiles, BS, 0, s>>>((const float *)x, (const float *)y,
                                                           (const float *)z, n, box, sub_box);
    return check_launch("k_tile_boxes");
}

// With several source segments, store each target block's geometry first
// (a.tgeo); otherwise the force kernel derives it (tgeo = nullptr).
// rows: the force kernel takes bf16 matrix-core far tiles (TPT 1), whose
// B operands are stored as well (a.tbg).
template <typename T>
static int prepare_target_geom(ForceArgs<T> &a, int tpt, hipStream_t s, bool rows = false) {
    if (!a.tgeo || a.nseg <= 1 || a.nblk_i <= 0) {
        a.tgeo = nullptr;
        a.tbg = nullptr;
        return NB_OK;
    }
    // already written by the step's drift (one target per lane), with the B
    // operands when this launch takes them
    if (tpt == 1 && a.tgeo_ready && (a.tgeo_ready == 2 || !rows || !a.tbg)) {
        if (!rows || !a.tbg) a.tbg = nullptr;
        return NB_OK;
    }
    const bool hilo = sizeof(T) == 8;
    if (!rows || tpt != 1 || !a.tbg) a.tbg = nullptr;
    if (tpt == 1 && a.tbg) k_target_geom<T, 1, true><<<(unsigned)a.nblk_i, BS, 0, s>>>(a, hilo);
    else if (tpt == 1) k_target_geom<T, 1, false><<<(unsigned)a.nblk_i, BS, 0, s>>>(a, hilo);
    else if (tpt == 2) k_target_geom<T, 2, false><<<(unsigned)a.nblk_i, BS, 0, s>>>(a, hilo);
    else if (tpt == 4) k_target_geom<T, 4, false><<<(unsigned)a.nblk_i, BS, 0, s>>>(a, hilo);
    else {
        a.tgeo = nullptr;
        a.tbg = nullptr;
        return NB_OK;
    }
    return check_launch("k_target_geom");
}

template <typename T, int MF>
static int launch_f32_impl(ForceArgs<T> a, int tpt, bool pot, hipStream_t s, bool acc) {
    const int64_t nb = (int64_t)a.nblk_i * a.nseg;
    if (nb <= 0) return NB_OK;
    if (nb > INT32_MAX) {
        set_error("force grid too large (%lld blocks)", (long long)nb);
        return NB_ERR_INVALID;
    }
    NB_TRY(prepare_target_geom(a, tpt, s));
    dim3 g((unsigned)nb), b(BS);
#define NB_L(TPT_, POT_) k_force_f32<T, TPT_, POT_, (POT_ ? 0 : MF)><<<g, b, 0, s>>>(a)
#define NB_LP(TPT_) k_force_f32<T, TPT_, true, 0, false><<<g, b, 0, s>>>(a)
    if (tpt == 1) { if (pot && !acc) NB_LP(1); else if (pot) NB_L(1, true); else NB_L(1, false); }
    else if (tpt == 2) { if (pot && !acc) NB_LP(2); else if (pot) NB_L(2, true); else NB_L(2, false); }
    else if (tpt == 4) { if (pot && !acc) NB_LP(4); else if (pot) NB_L(4, true); else NB_L(4, false); }
    else {
        set_error("unsupported targets_per_thread %d", tpt);
        return NB_ERR_INVALID;
    }
#undef NB_L
#undef NB_LP
    return check_launch("k_force_f32");
}

int launch_force_f32(const ForceArgs<float> &a, int tpt, bool pot, hipStream_t s, bool acc,
                     int far_path) {
    if (!acc && !pot) acc = true;
    if (far_path == FAR_MFMA_BF16 && tpt == 1 && (!pot || !acc)) {
        // FP32 state through the per-block FP64 origin: hi/lo near tiles, bf16
        // matrix-core squared distances on the far tiles (forces or potential alone)
        const int64_t nb = (int64_t)a.nblk_i * a.nseg;
        if (nb <= 0) return NB_OK;
        if (nb > INT32_MAX) {
            set_error("force grid too large (%lld blocks)", (long long)nb);
            return NB_ERR_INVALID;
        }
        ForceArgs<float> b = a;
        b.use_origin = 1;
        NB_TRY(prepare_target_geom(b, 1, s, true));
        if (pot) k_force_f32<float, 1, true, FAR_MFMA_BF16, false><<<(unsigned)nb, BS, 0, s>>>(b);
        else launch_bf16_forces(b, nb, s);
        return check_launch("k_force_f32<float, bf16 mfma>");
    }
    return launch_f32_impl<float, FAR_VALU>(a, tpt, pot, s, acc);
}
int launch_force_split(const ForceArgs<double> &a_in, int tpt, bool pot, hipStream_t s, int far_path,
                       bool acc) {
    if (!acc && !pot) acc = true;
    ForceArgs<double> a = a_in;
    if (far_path == FAR_MFMA_BF16 && tpt == 1 && pot && !acc) {
        const int64_t nb = (int64_t)a.nblk_i * a.nseg;
        if (nb <= 0) return NB_OK;
        NB_TRY(prepare_target_geom(a, tpt, s, true));
        k_force_f32<double, 1, true, FAR_MFMA_BF16, false><<<(unsigned)nb, BS, 0, s>>>(a);
        return check_launch("k_force_f32<bf16 mfma, potential>");
    }
    if ((far_path == FAR_MFMA_SUM || far_path == FAR_MFMA_R2) && tpt == 1 && !pot) {
        const int64_t nb = (int64_t)a.nblk_i * a.nseg;
        if (nb <= 0) return NB_OK;
        NB_TRY(prepare_target_geom(a, tpt, s));
        if (far_path == FAR_MFMA_R2) k_force_f32<double, 1, false, FAR_MFMA_R2><<<(unsigned)nb, BS, 0, s>>>(a);
        else k_force_f32<double, 1, false, FAR_MFMA_SUM><<<(unsigned)nb, BS, 0, s>>>(a);
        return check_launch("k_force_f32<mfma>");
    }
    if (far_path == FAR_MFMA_BF16 && tpt == 1 && !pot) {
        const int64_t nb = (int64_t)a.nblk_i * a.nseg;
        if (nb <= 0) return NB_OK;
        NB_TRY(prepare_target_geom(a, tpt, s, true));
        launch_bf16_forces(a, nb, s);
        return check_launch("k_force_f32<bf16 mfma>");
    }
    if (far_path == FAR_WEIGHTED) return launch_f32_impl<double, FAR_WEIGHTED>(a, tpt, pot, s, acc);
    return launch_f32_impl<double, FAR_VALU>(a, tpt, pot, s, acc);
}

template <typename T>
__global__ void k_source_weights(const T *__restrict__ m, int64_t n, float *__restrict__ w) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double mi = (double)m[i];
    w[i] = mi > 0.0 ? (float)(1.0 / cbrt(mi * mi)) : 0.0f;
}

int launch_source_weights(const void *m, bool is_double, int64_t n, float *w, hipStream_t s) {
    if (n <= 0) return NB_OK;
    const unsigned g = (unsigned)((n + 255) / 256);
    if (is_double) k_source_weights<double><<<g, 256, 0, s>>>((const double *)m, n, w);
    else k_source_weights<float><<<g, 256, 0, s>>>((const float *)m, n, w);
    return check_launch("k_source_weights");
}

int launch_force_f64(const ForceArgs<double> &a, bool pot, hipStream_t s) {
    const int64_t nb = (int64_t)a.nblk_i * a.nseg;
    if (nb <= 0) return NB_OK;
    if (pot) k_force_f64<true><<<(unsigned)nb, BS, 0, s>>>(a);
    else k_force_f64<false><<<(unsigned)nb, BS, 0, s>>>(a);
    return check_launch("k_force_f64");
}

int launch_drift_boxes(void *x, void *y, void *z, const void *vx, const void *vy, const void *vz, int64_t n,
                       double dt, double *time, int rd, int wr, int is_double, double *box, double *sub_box,
                       hipStream_t s, const StepGeometry *geo) {
    if (n <= 0) return NB_OK;
    const unsigned nt = (unsigned)((n + BS - 1) / BS);
    const bool g = geo && geo->tgeo, r = g && geo->tbg;
    const double *h = g ? geo->h : nullptr;
    double *tg = g ? geo->tgeo : nullptr;
    unsigned *tb = r ? geo->tbg : nullptr;
    const int uo = g ? geo->use_origin : 0;
#define NB_DB(T_, G_, R_)                                                                                       \
    k_drift_boxes<T_, G_, R_><<<nt, BS, 0, s>>>((T_ *)x, (T_ *)y, (T_ *)z, (const T_ *)vx, (const T_ *)vy,     \
                                                (const T_ *)vz, n, dt, time, rd, wr, box, sub_box, h, tg, tb, uo, \
                                                sizeof(T_) == 8)
    if (is_double) {
        if (r) NB_DB(double, true, true);
        else if (g) NB_DB(double, true, false);
        else NB_DB(double, false, false);
    } else {
        if (r) NB_DB(float, true, true);
        else if (g) NB_DB(float, true, false);
        else NB_DB(float, false, false);
    }
#undef NB_DB
    return check_launch("k_drift_boxes");
}

int launch_finalize_kick(const double *part, int nseg, int64_t n, double G, double *ax, double *ay, double *az,
                         void *vx, void *vy, void *vz, double coef, int is_double, hipStream_t s) {
    if (n <= 0) return NB_OK;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (is_double)
        k_finalize_kick<double><<<nb, 256, 0, s>>>(part, nseg, n, G, ax, ay, az, (double *)vx, (double *)vy,
                                                   (double *)vz, coef);
    else
        k_finalize_kick<float><<<nb, 256, 0, s>>>(part, nseg, n, G, ax, ay, az, (float *)vx, (float *)vy,
                                                  (float *)vz, coef);
    return check_launch("k_finalize_kick");
}

int launch_finalize(const double *part, int nseg, int64_t n, double G, double *ax, double *ay,
                    double *az, double *pot, bool with_pot, hipStream_t s, bool with_acc) {
    if (n <= 0) return NB_OK;
    const unsigned nb = (unsigned)((n + 255) / 256);
    k_finalize<<<nb, 256, 0, s>>>(part, nseg, n, G, ax, ay, az, pot, with_pot ? 1 : 0,
                                  with_acc ? 1 : 0);
    return check_launch("k_finalize");
}

}  // namespace nb
