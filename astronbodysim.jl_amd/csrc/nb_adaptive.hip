// nb_adaptive.hip -- adaptive (block, power-of-two) timesteps on the GPU.
//
// Follows the reference's CPU semantics (src/base/Timestep.jl), which its GPU
// kernels restate (src/directsumgpu/timestep.jl:1-47,103-154,178-192):
//   get_timestep      dt = sqrt(2 eta h / |a|), capped by MaxStep; errors when
//                     dt < MinStep or the integer step is <= 0; the integer
//                     step is the largest TimeBase / 2^k not above dt/TimeInterval
//                     (src/base/Timestep.jl:1-23)
//   init_timesteps    Ti_endstep = get_timestep, Ti_begstep = 0  (:29-39)
//   half kick         Vel += Acc * forward * 0.5 * (Ti_end - Ti_beg) * TimeInterval (:105-114)
//   advance           active particles (Ti_endstep == system time): new step,
//                     kick from the middle of the old step to the middle of the
//                     new one, Ti_beg = Ti_end, Ti_end += step  (:47-82)
//   min end step      minimum over Ti_endstep and TimeBase  (:135-144)
//   drift             every particle by (t_int - last_t_int) * TimeInterval (:169-179)
// The force is evaluated for the active particles only; inactive particles
// keep their accelerations (the reference CPU path does; its GPU path zeroes
// them, src/directsumgpu/gravity.jl:43).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstring>
#include <math.h>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "nb_ctx.h"

namespace nb {

static int launch_ok(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of %s failed: %s", what, hipGetErrorString(e));
        return NB_ERR_HIP;
    }
    return NB_OK;
}

static inline unsigned gridn(int64_t n) { return (unsigned)((n + 255) / 256); }

struct AdaptParams {
    double time_interval;
    int64_t time_base;
    double err_tol;
    double min_step;
    double max_step;
};

// Error flags: 1 = dt < MinStep, 2 = integer step <= 0.
__device__ __forceinline__ int64_t get_timestep_dev(double ax, double ay, double az, double h,
                                                    const AdaptParams &p, unsigned *flag) {
    const double ac = sqrt(ax * ax + ay * ay + az * az);
    double dt = sqrt(2.0 * p.err_tol * h / ac);
    if (dt > p.max_step) dt = p.max_step;
    if (dt < p.min_step) atomicOr(flag, 1u);
    const double ti_step = dt / p.time_interval;
    if (!(ti_step > 0.0)) atomicOr(flag, 2u);
    int64_t ti_min = p.time_base;
    while ((double)ti_min > ti_step && ti_min > 0) ti_min >>= 1;
    return ti_min;
}

template <typename T>
__global__ void k_init_timesteps(const double *__restrict__ ax, const double *__restrict__ ay,
                                 const double *__restrict__ az, const double *__restrict__ h,
                                 int64_t n, AdaptParams p, int64_t *__restrict__ te,
                                 int64_t *__restrict__ tb, unsigned *flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    te[i] = get_timestep_dev(ax[i], ay[i], az[i], h[i], p, flag);
    tb[i] = 0;
}

template <typename T>
__global__ void k_adaptive_half_kick(T *__restrict__ vx, T *__restrict__ vy, T *__restrict__ vz,
                                     const double *__restrict__ ax, const double *__restrict__ ay,
                                     const double *__restrict__ az, const int64_t *__restrict__ te,
                                     const int64_t *__restrict__ tb, int64_t n, double forward,
                                     double interval) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double c = forward * 0.5 * (double)(te[i] - tb[i]) * interval;
    vx[i] += (T)(ax[i] * c);
    vy[i] += (T)(ay[i] * c);
    vz[i] += (T)(az[i] * c);
}

template <typename T>
__global__ void k_adaptive_advance(T *__restrict__ vx, T *__restrict__ vy, T *__restrict__ vz,
                                   const double *__restrict__ ax, const double *__restrict__ ay,
                                   const double *__restrict__ az, const double *__restrict__ h,
                                   int64_t *__restrict__ te, int64_t *__restrict__ tb, int64_t n,
                                   int64_t t_int, AdaptParams p, unsigned *flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || te[i] != t_int) return;
    int64_t ti_step = get_timestep_dev(ax[i], ay[i], az[i], h[i], p, flag);
    if (t_int == p.time_base) ti_step = 0;
    if (p.time_base - t_int < ti_step) ti_step = p.time_base - t_int;
    const double tstart = (double)(tb[i] + te[i]) / 2.0;
    const double tend = (double)te[i] + 0.5 * (double)ti_step;
    const double dt = (tend - tstart) * p.time_interval;
    vx[i] += (T)(ax[i] * dt);
    vy[i] += (T)(ay[i] * dt);
    vz[i] += (T)(az[i] * dt);
    const int64_t beg = te[i];
    tb[i] = beg;
    te[i] = beg + ti_step;
}

template <typename T>
__global__ void k_adaptive_drift(T *__restrict__ x, T *__restrict__ y, T *__restrict__ z,
                                 const T *__restrict__ vx, const T *__restrict__ vy,
                                 const T *__restrict__ vz, int64_t n, double dt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T d = (T)dt;
    x[i] += vx[i] * d;
    y[i] += vy[i] * d;
    z[i] += vz[i] * d;
}

// Minimum end step and the number of particles at it: (m, k) pairs combine as
// (min, sum of the counts of the operands at that minimum).  Start value
// (time_base, 0), so an empty range leaves (time_base, 0).
constexpr int MINRED_BLOCKS = 256;

__device__ __forceinline__ void min_count_block(int64_t m, int64_t k, int64_t *sm, int64_t *sk) {
    const int t = threadIdx.x;
    sm[t] = m;
    sk[t] = k;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) {
            const int64_t m1 = sm[t], m2 = sm[t + off];
            const int64_t mm = m2 < m1 ? m2 : m1;
            sk[t] = (m1 == mm ? sk[t] : 0) + (m2 == mm ? sk[t + off] : 0);
            sm[t] = mm;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_min_count_part(const int64_t *__restrict__ te, int64_t n,
                                                        int64_t base, int64_t *__restrict__ part) {
    __shared__ int64_t sm[256], sk[256];
    int64_t m = base, k = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int64_t v = te[i];
        if (v < m) {
            m = v;
            k = 1;
        } else if (v == m) {
            ++k;
        }
    }
    min_count_block(m, k, sm, sk);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sm[0];
        part[2 * blockIdx.x + 1] = sk[0];
    }
}

// out = (minimum, count, error flag of the last kernel that set it)
__global__ __launch_bounds__(256) void k_min_count_final(const int64_t *__restrict__ part, int nb,
                                                         int64_t base,
                                                         const unsigned *__restrict__ flag,
                                                         int64_t *__restrict__ out) {
    __shared__ int64_t sm[256], sk[256];
    const int t = threadIdx.x;
    const bool in = t < nb;
    min_count_block(in ? part[2 * t] : base, in ? part[2 * t + 1] : 0, sm, sk);
    if (t == 0) {
        out[0] = sm[0];
        out[1] = sk[0];
        out[2] = (int64_t)*flag;
    }
}

__global__ void k_active_flags(const int64_t *__restrict__ te, int64_t n, int64_t t_int,
                               uint8_t *__restrict__ flags) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flags[i] = te[i] == t_int ? 1 : 0;
}

template <typename T>
__global__ void k_gather_targets(const T *__restrict__ x, const T *__restrict__ y,
                                 const T *__restrict__ z, const double *__restrict__ h,
                                 const int64_t *__restrict__ idx, const int64_t *__restrict__ cnt,
                                 T *__restrict__ gx, T *__restrict__ gy, T *__restrict__ gz,
                                 double *__restrict__ gh) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *cnt) return;
    const int64_t i = idx[k];
    gx[k] = x[i];
    gy[k] = y[i];
    gz[k] = z[i];
    gh[k] = h[i];
}

__global__ void k_scatter_acc(const double *__restrict__ gax, const double *__restrict__ gay,
                              const double *__restrict__ gaz, const int64_t *__restrict__ idx,
                              const int64_t *__restrict__ cnt, double *__restrict__ ax,
                              double *__restrict__ ay, double *__restrict__ az) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *cnt) return;
    const int64_t i = idx[k];
    ax[i] = gax[k];
    ay[i] = gay[k];
    az[i] = gaz[k];
}

__global__ void k_steps_in(const int64_t *__restrict__ e, const int64_t *__restrict__ b,
                           const int64_t *__restrict__ perm, int64_t n, int64_t *__restrict__ te,
                           int64_t *__restrict__ tb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t u = perm ? perm[i] : i;
    te[i] = e[u];
    tb[i] = b[u];
}

}  // namespace nb

using namespace nb;

namespace {

AdaptParams params_of(const nb_ctx *c) {
    AdaptParams p;
    p.time_interval = c->ad.time_interval;
    p.time_base = c->ad.time_base;
    p.err_tol = c->ad.err_tol;
    p.min_step = c->ad.min_step;
    p.max_step = c->ad.max_step;
    return p;
}

int check_adaptive(nb_ctx *c) {
    clear_error();
    if (!c) {
        set_error("null context");
        return NB_ERR_INVALID;
    }
    NB_HIP(hipSetDevice(c->dev));
    if (!c->uploaded) {
        set_error("no particle data uploaded");
        return NB_ERR_STATE;
    }
    if (!c->ti_end) {
        set_error("adaptive timesteps not initialised (call nb_adaptive_init)");
        return NB_ERR_STATE;
    }
    return NB_OK;
}

int flag_error(nb_ctx *c, unsigned f) {
    if (f) {
        NB_HIP(hipMemsetAsync(c->flag, 0, sizeof(unsigned), c->stream));
        NB_HIP(hipStreamSynchronize(c->stream));
        set_error(f & 1 ? "Timestep < MinStep" : "Wrong integer timestep");
        return NB_ERR_TIMESTEP;
    }
    return NB_OK;
}

int read_flag(nb_ctx *c) {
    unsigned f = 0;
    NB_HIP(hipMemcpyAsync(&f, c->flag, sizeof f, hipMemcpyDeviceToHost, c->stream));
    NB_HIP(hipStreamSynchronize(c->stream));
    return flag_error(c, f);
}

// (minimum end step, particles at it, error flag) of the current timeline, in
// one copy and one wait.
int reduce_steps(nb_ctx *c, int64_t (&r)[3]) {
    hipStream_t s = c->stream;
    const int nb = (int)std::min<int64_t>(MINRED_BLOCKS, std::max<int64_t>(1, (c->n + 255) / 256));
    k_min_count_part<<<nb, 256, 0, s>>>(c->ti_end, c->n, c->ad.time_base, c->ad.red);
    NB_TRY(launch_ok("k_min_count_part"));
    k_min_count_final<<<1, 256, 0, s>>>(c->ad.red, nb, c->ad.time_base, c->flag,
                                        c->ad.red + 2 * MINRED_BLOCKS);
    NB_TRY(launch_ok("k_min_count_final"));
    NB_HIP(hipMemcpyAsync(r, c->ad.red + 2 * MINRED_BLOCKS, sizeof r, hipMemcpyDeviceToHost, s));
    NB_HIP(hipStreamSynchronize(s));
    return NB_OK;
}

}  // namespace

extern "C" {

int nb_adaptive_init(nb_ctx *c, double time_interval, int64_t time_base, double err_tol,
                     double min_step, double max_step) {
    clear_error();
    if (!c) {
        set_error("null context");
        return NB_ERR_INVALID;
    }
    NB_HIP(hipSetDevice(c->dev));
    if (!c->uploaded) {
        set_error("no particle data uploaded");
        return NB_ERR_STATE;
    }
    if (!(time_interval > 0.0) || time_base <= 0 || !(err_tol > 0.0) || !(min_step >= 0.0) ||
        !(max_step > 0.0)) {
        set_error("invalid adaptive timestep parameters");
        return NB_ERR_INVALID;
    }
    c->ad.next_valid = false;
    c->ad.time_interval = time_interval;
    c->ad.time_base = time_base;
    c->ad.err_tol = err_tol;
    c->ad.min_step = min_step;
    c->ad.max_step = max_step;
    const size_t n = (size_t)std::max<int64_t>(c->n, 1);
    if (!c->ti_end) {
        NB_TRY(dalloc(&c->ti_end, sizeof(int64_t) * n));
        NB_TRY(dalloc(&c->ti_beg, sizeof(int64_t) * n));
        NB_TRY(dalloc(&c->ad.flags, n));
        NB_TRY(dalloc(&c->ad.idx, sizeof(int64_t) * n));
        NB_TRY(dalloc(&c->ad.cnt, sizeof(int64_t) * 2));
        NB_TRY(dalloc(&c->ad.red, sizeof(int64_t) * (2 * MINRED_BLOCKS + 3)));
        NB_TRY(dalloc(&c->ad.buf, sizeof(double) * 7 * n));
        size_t bytes = 0;
        rocprim::counting_iterator<int64_t> it(0);
        NB_HIP(rocprim::select(nullptr, bytes, it, c->ad.flags, c->ad.idx, c->ad.cnt, (size_t)n,
                               c->stream));
        NB_TRY(dalloc(&c->ad.tmp, bytes));
        c->ad.tmp_bytes = bytes;
    }
    NB_HIP(hipMemsetAsync(c->flag, 0, sizeof(unsigned), c->stream));
    if (c->n > 0) {
        k_init_timesteps<double><<<gridn(c->n), 256, 0, c->stream>>>(
            c->ax, c->ay, c->az, c->h, c->n, params_of(c), c->ti_end, c->ti_beg, c->flag);
        NB_TRY(launch_ok("k_init_timesteps"));
    }
    return read_flag(c);
}

int nb_adaptive_half_kick(nb_ctx *c, int32_t forward) {
    NB_TRY(check_adaptive(c));
    if (forward != 1 && forward != -1) {
        set_error("forward must be +1 or -1");
        return NB_ERR_INVALID;
    }
    if (c->n == 0) return NB_OK;
    if (c->is_double)
        k_adaptive_half_kick<double><<<gridn(c->n), 256, 0, c->stream>>>(
            (double *)c->vx, (double *)c->vy, (double *)c->vz, c->ax, c->ay, c->az, c->ti_end,
            c->ti_beg, c->n, (double)forward, c->ad.time_interval);
    else
        k_adaptive_half_kick<float><<<gridn(c->n), 256, 0, c->stream>>>(
            (float *)c->vx, (float *)c->vy, (float *)c->vz, c->ax, c->ay, c->az, c->ti_end,
            c->ti_beg, c->n, (double)forward, c->ad.time_interval);
    return launch_ok("k_adaptive_half_kick");
}

int nb_adaptive_min_endstep(nb_ctx *c, int64_t *min_endstep) {
    NB_TRY(check_adaptive(c));
    if (!min_endstep) {
        set_error("null output");
        return NB_ERR_INVALID;
    }
    // after a kick the reduction has already been read back with its flag
    if (!c->ad.next_valid) {
        int64_t r[3];
        NB_TRY(reduce_steps(c, r));
        c->ad.next_min = r[0];
        c->ad.next_count = r[1];
        c->ad.next_valid = true;
    }
    *min_endstep = c->ad.next_min;
    return NB_OK;
}

int nb_adaptive_drift(nb_ctx *c, int64_t from_int, int64_t to_int) {
    NB_TRY(check_adaptive(c));
    const double dt = (double)(to_int - from_int) * c->ad.time_interval;
    if (c->n == 0) return NB_OK;
    if (c->is_double)
        k_adaptive_drift<double><<<gridn(c->n), 256, 0, c->stream>>>(
            (double *)c->x, (double *)c->y, (double *)c->z, (const double *)c->vx,
            (const double *)c->vy, (const double *)c->vz, c->n, dt);
    else
        k_adaptive_drift<float><<<gridn(c->n), 256, 0, c->stream>>>(
            (float *)c->x, (float *)c->y, (float *)c->z, (const float *)c->vx,
            (const float *)c->vy, (const float *)c->vz, c->n, dt);
    return launch_ok("k_adaptive_drift");
}

int nb_adaptive_compute_force(nb_ctx *c, int64_t t_int, int64_t *nactive) {
    NB_TRY(check_adaptive(c));
    if (c->n == 0) {
        if (nactive) *nactive = 0;
        return NB_OK;
    }
    hipStream_t s = c->stream;
    const int64_t n = c->n;
    // ordered compaction of the active particles (deterministic)
    k_active_flags<<<gridn(n), 256, 0, s>>>(c->ti_end, n, t_int, c->ad.flags);
    NB_TRY(launch_ok("k_active_flags"));
    size_t bytes = c->ad.tmp_bytes;
    rocprim::counting_iterator<int64_t> it(0);
    NB_HIP(rocprim::select(c->ad.tmp, bytes, it, c->ad.flags, c->ad.idx, c->ad.cnt, (size_t)n, s));
    // The count is known without a wait when t_int is the minimum end step
    // reduced since Ti_endstep last changed (the usual block step); the
    // kernels below read the device count, which equals it.
    int64_t na = 0;
    if (c->ad.next_valid && t_int == c->ad.next_min) {
        na = c->ad.next_count;
    } else {
        NB_HIP(hipMemcpyAsync(&na, c->ad.cnt, sizeof na, hipMemcpyDeviceToHost, s));
        NB_HIP(hipStreamSynchronize(s));
    }
    if (nactive) *nactive = na;
    if (na == 0) return NB_OK;
    if (na == n) return ctx_run_force(c, false, true);   // everyone active: the plain path
    double *gh = c->ad.buf;
    double *gax = gh + n, *gay = gax + n, *gaz = gay + n;
    void *gx = gaz + n;
    void *gy = (char *)gx + (c->is_double ? 8 : 4) * (size_t)n;
    void *gz = (char *)gy + (c->is_double ? 8 : 4) * (size_t)n;
    NB_TRY(launch_tile_boxes(c->x, c->y, c->z, c->is_double, n, c->tile_box, c->sub_box, s));
    if (c->is_double)
        k_gather_targets<double><<<gridn(na), 256, 0, s>>>(
            (const double *)c->x, (const double *)c->y, (const double *)c->z, c->h, c->ad.idx,
            c->ad.cnt, (double *)gx, (double *)gy, (double *)gz, gh);
    else
        k_gather_targets<float><<<gridn(na), 256, 0, s>>>(
            (const float *)c->x, (const float *)c->y, (const float *)c->z, c->h, c->ad.idx,
            c->ad.cnt, (float *)gx, (float *)gy, (float *)gz, gh);
    NB_TRY(launch_ok("k_gather_targets"));
    // Few active targets still fill the GPU: split the sources into segments
    // (partials reduced in a fixed order) for about 4096 blocks.
    const int64_t ntiles = (n + BS - 1) / BS;
    const int64_t nblk = (na + BS - 1) / BS;
    int64_t want = (4096 + nblk - 1) / nblk;
    want = std::max<int64_t>(1, std::min<int64_t>(want, std::min<int64_t>(ntiles, 64)));
    const int64_t tiles_per_seg = (ntiles + want - 1) / want;
    const int nseg = (int)((ntiles + tiles_per_seg - 1) / tiles_per_seg);
    NB_TRY(ctx_ensure_part(c, nseg));
    if (c->is_double) {
        ForceArgs<double> a{};
        a.tx = (const double *)gx; a.ty = (const double *)gy; a.tz = (const double *)gz;
        a.th = gh; a.nt = na;
        a.sx = (const double *)c->x; a.sy = (const double *)c->y; a.sz = (const double *)c->z;
        a.sm = (const double *)c->m; a.ns = n;
        a.tile_box = c->tile_box; a.sub_box = c->sub_box; a.use_origin = 1;
        a.nseg = nseg; a.seg_len = tiles_per_seg * BS; a.nblk_i = (int)nblk; a.G = c->G;
        a.ax = gax; a.ay = gay; a.az = gaz; a.pot = nullptr; a.part = c->part;
        a.skip_src = (c->mode == NB_MODE_SPLIT_CENTRAL && c->ncentral > 0) ? c->central : nullptr;
        if (c->mode == NB_MODE_FP64) {
            NB_TRY(launch_force_f64(a, false, s));
        } else {
            // the context's far path (one target per lane here), with the
            // source weights that the weighted and matrix-core paths read
            const int fp = c->tun.far_path;
            if (far_path_weighted(fp)) {
                NB_TRY(launch_source_weights(c->m, true, n, c->wsrc, s));
                a.sw = c->wsrc;
            }
            NB_TRY(launch_force_split(a, 1, false, s, fp));
        }
        if (nseg > 1) NB_TRY(launch_finalize(c->part, nseg, na, c->G, gax, gay, gaz, nullptr, false, s));
        if (c->mode == NB_MODE_SPLIT_CENTRAL && c->ncentral > 0) {
            // FP64 central terms for the active targets
            ForceArgs<double> b = a;
            b.sx = c->cx; b.sy = c->cy; b.sz = c->cz; b.sm = c->cm; b.ns = c->ncentral;
            b.skip_src = nullptr; b.tile_box = nullptr; b.sub_box = nullptr; b.accumulate = 1;
            b.nseg = 1; b.part = nullptr;
            b.seg_len = ((c->ncentral + BS - 1) / BS) * BS;
            NB_TRY(ctx_gather_central(c));
            NB_TRY(launch_force_f64(b, false, s));
        }
    } else {
        ForceArgs<float> a{};
        a.tx = (const float *)gx; a.ty = (const float *)gy; a.tz = (const float *)gz;
        a.th = gh; a.nt = na;
        a.sx = (const float *)c->x; a.sy = (const float *)c->y; a.sz = (const float *)c->z;
        a.sm = (const float *)c->m; a.ns = n;
        a.tile_box = c->tile_box; a.sub_box = c->sub_box; a.use_origin = 0;
        a.nseg = nseg; a.seg_len = tiles_per_seg * BS; a.nblk_i = (int)nblk; a.G = c->G;
        a.ax = gax; a.ay = gay; a.az = gaz; a.pot = nullptr; a.part = c->part;
        const int fp = c->tun.far_path == FAR_MFMA_BF16 ? FAR_MFMA_BF16 : FAR_VALU;
        if (fp == FAR_MFMA_BF16) {
            NB_TRY(launch_source_weights(c->m, false, n, c->wsrc, s));
            a.sw = c->wsrc;
        }
        NB_TRY(launch_force_f32(a, 1, false, s, true, fp));
        if (nseg > 1) NB_TRY(launch_finalize(c->part, nseg, na, c->G, gax, gay, gaz, nullptr, false, s));
    }
    // post-processing of the active targets, in the order of
    // src/base/PostProcessing/Force.jl:33-46: background fields, then MOND
    for (const BgField &f : c->bg)
        NB_TRY(background(gx, gy, gz, na, c->is_double, f, gax, gay, gaz, nullptr, s));
    if (c->mond_acc0 > 0.0) NB_TRY(mond(gax, gay, gaz, na, c->mond_acc0, c->mond_index, s));
    k_scatter_acc<<<gridn(na), 256, 0, s>>>(gax, gay, gaz, c->ad.idx, c->ad.cnt, c->ax, c->ay, c->az);
    NB_TRY(launch_ok("k_scatter_acc"));
    return NB_OK;
}

int nb_adaptive_kick(nb_ctx *c, int64_t t_int) {
    NB_TRY(check_adaptive(c));
    if (c->n == 0) return NB_OK;
    c->ad.next_valid = false;
    if (c->is_double)
        k_adaptive_advance<double><<<gridn(c->n), 256, 0, c->stream>>>(
            (double *)c->vx, (double *)c->vy, (double *)c->vz, c->ax, c->ay, c->az, c->h, c->ti_end,
            c->ti_beg, c->n, t_int, params_of(c), c->flag);
    else
        k_adaptive_advance<float><<<gridn(c->n), 256, 0, c->stream>>>(
            (float *)c->vx, (float *)c->vy, (float *)c->vz, c->ax, c->ay, c->az, c->h, c->ti_end,
            c->ti_beg, c->n, t_int, params_of(c), c->flag);
    NB_TRY(launch_ok("k_adaptive_advance"));
    // the new timeline's next sync point comes back with the error flag
    int64_t r[3];
    NB_TRY(reduce_steps(c, r));
    NB_TRY(flag_error(c, (unsigned)r[2]));
    c->ad.next_min = r[0];
    c->ad.next_count = r[1];
    c->ad.next_valid = true;
    return NB_OK;
}

int nb_adaptive_set_steps(nb_ctx *c, const int64_t *endstep, const int64_t *begstep) {
    NB_TRY(check_adaptive(c));
    c->ad.next_valid = false;
    const int64_t n = c->n;
    if (n == 0) return NB_OK;
    if (!endstep || !begstep) {
        set_error("endstep and begstep are required");
        return NB_ERR_INVALID;
    }
    for (int64_t i = 0; i < n; ++i) {
        if (begstep[i] < 0 || endstep[i] < begstep[i] || endstep[i] > c->ad.time_base) {
            set_error("step %lld: begstep %lld / endstep %lld outside 0 <= beg <= end <= TimeBase",
                      (long long)i, (long long)begstep[i], (long long)endstep[i]);
            return NB_ERR_INVALID;
        }
    }
    // staging: the gathered-target buffer holds 7n doubles, 2n int64 fit
    int64_t *se = (int64_t *)c->ad.buf, *sb = se + n;
    NB_HIP(hipMemcpyAsync(se, endstep, sizeof(int64_t) * n, hipMemcpyHostToDevice, c->stream));
    NB_HIP(hipMemcpyAsync(sb, begstep, sizeof(int64_t) * n, hipMemcpyHostToDevice, c->stream));
    k_steps_in<<<gridn(n), 256, 0, c->stream>>>(se, sb, c->perm, n, c->ti_end, c->ti_beg);
    NB_TRY(launch_ok("k_steps_in"));
    NB_HIP(hipStreamSynchronize(c->stream));
    return NB_OK;
}

int nb_adaptive_get_steps(nb_ctx *c, int64_t *endstep, int64_t *begstep) {
    NB_TRY(check_adaptive(c));
    const int64_t n = c->n;
    if (n == 0) return NB_OK;
    std::vector<int64_t> e((size_t)n), b((size_t)n), perm;
    NB_HIP(hipMemcpyAsync(e.data(), c->ti_end, sizeof(int64_t) * n, hipMemcpyDeviceToHost, c->stream));
    NB_HIP(hipMemcpyAsync(b.data(), c->ti_beg, sizeof(int64_t) * n, hipMemcpyDeviceToHost, c->stream));
    if (c->perm) {
        perm.resize((size_t)n);
        NB_HIP(hipMemcpyAsync(perm.data(), c->perm, sizeof(int64_t) * n, hipMemcpyDeviceToHost,
                              c->stream));
    }
    NB_HIP(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < n; ++i) {
        const int64_t u = c->perm ? perm[(size_t)i] : i;
        if (endstep) endstep[u] = e[(size_t)i];
        if (begstep) begstep[u] = b[(size_t)i];
    }
    return NB_OK;
}

}  // extern "C"
