// nb_api.hip -- the C-ABI (include/nbody_mi355x.h) for one-system contexts.
//
// Device state is SoA in an internal particle order (perm maps internal ->
// caller index; nullptr = identity).  Positions/velocities/masses are FP64
// except in NB_MODE_FP32, where they are FP32 as in the reference's
// floattype = Float32 GPU simulations.  Accelerations and potentials are FP64.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <utility>
#include <vector>

#include "nb_ctx.h"
#include "nb_integrate.h"
#include "nb_internal.h"
#include "nb_pair.h"

#define NB_VERSION_MAJOR 1
#define NB_VERSION_MINOR 0
#define NB_VERSION_PATCH 0

namespace nb {

static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}
void clear_error() { g_err.clear(); }
const char *last_error() { return g_err.c_str(); }

}  // namespace nb

using namespace nb;

namespace {

// Defaults before any autotuned entry: one target per lane; in the SPLIT
// modes the bf16 matrix-core far path from 2^16 particles on (where most
// tiles are far), with source segments for about 16k blocks, else 8k.
int default_tpt(int64_t) { return 1; }

int default_far_path(int mode, int64_t n) {
    return mode != NB_MODE_FP64 && n >= (1 << 16) ? FAR_MFMA_BF16 : FAR_VALU;
}

int default_segments(int64_t n, int tpt, int far_path) {
    const int64_t nblk = (n + (int64_t)BS * tpt - 1) / ((int64_t)BS * tpt);
    const int64_t ntiles = (n + BS - 1) / BS;
    const int64_t want = far_path == FAR_MFMA_BF16 ? 16384 : 8192;
    int64_t nseg = (want + nblk - 1) / std::max<int64_t>(nblk, 1);
    // far path 4 keeps at least 16 segments: measured on the two-sphere
    // collision at 2^21 / 2^22 particles, 16 segments ran 6 % / 5 % faster
    // than the 2 / 1 that the block count alone gives.  32 gained only
    // 0.4 % / 0.2 % more there and doubles the partial sums (4 GB at 2^22).
    if (far_path == FAR_MFMA_BF16) nseg = std::max<int64_t>(nseg, 16);
    nseg = std::min<int64_t>(std::max<int64_t>(nseg, 1), std::max<int64_t>(ntiles, 1));
    nseg = std::min<int64_t>(nseg, 64);
    return (int)nseg;
}

int check_ctx(nb_ctx *c, bool need_upload = true) {
    clear_error();
    if (!c) {
        set_error("null context");
        return NB_ERR_INVALID;
    }
    hipError_t e = hipSetDevice(c->dev);
    if (e != hipSuccess) {
        set_error("hipSetDevice(%d) failed: %s", c->dev, hipGetErrorString(e));
        return NB_ERR_HIP;
    }
    if (need_upload && !c->uploaded) {
        set_error("no particle data uploaded (call nb_upload first)");
        return NB_ERR_STATE;
    }
    return NB_OK;
}

void drop_graph(nb_ctx *c) {
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->g_steps = 0;
}

int ensure_part(nb_ctx *c, int nseg) {
    if (nseg <= 1 || nseg <= c->part_cap) return NB_OK;
    drop_graph(c);
    dfree(c->part);
    c->part_cap = 0;
    NB_TRY(dalloc(&c->part, sizeof(double) * 4 * (size_t)c->n * (size_t)nseg));
    c->part_cap = nseg;
    return NB_OK;
}

struct Geometry {
    int tpt, nseg, nblk_i;
    int64_t seg_len;
};

// Workspaces of the matrix-core far path, allocated when a launch first needs
// them: the flags of the split far/near passes (one byte per wave) and, with
// several source segments, the bf16 B operands (16 dwords per target).
int ensure_tbg(nb_ctx *c, const Geometry &g) {
    if (g.tpt != 1 || c->tun.far_path != FAR_MFMA_BF16 || c->mode == NB_MODE_FP64) return NB_OK;
    const size_t nf = (size_t)g.nblk_i * (size_t)g.nseg * NW;
    if (nf > c->nflag_cap) {
        drop_graph(c);
        dfree(c->nflag);
        c->nflag = nullptr;
        c->nflag_cap = 0;
        NB_TRY(dalloc(&c->nflag, nf));
        NB_HIP(hipMemsetAsync(c->nflag, 0, nf, c->stream));   // before any capture
        c->nflag_cap = nf;
    }
    if (c->tbg || g.nseg <= 1) return NB_OK;
    drop_graph(c);
    return dalloc(&c->tbg, sizeof(unsigned) * 16 * (size_t)BS * (size_t)((c->n + BS - 1) / BS + 1));
}

Geometry geometry(int64_t nt, int64_t ns, int tpt, int want_seg) {
    Geometry g;
    g.tpt = tpt;
    g.nblk_i = (int)((nt + (int64_t)BS * tpt - 1) / ((int64_t)BS * tpt));
    const int64_t ntiles = std::max<int64_t>((ns + BS - 1) / BS, 1);
    int64_t nseg = std::min<int64_t>(std::max(want_seg, 1), ntiles);
    const int64_t tiles_per_seg = (ntiles + nseg - 1) / nseg;
    g.seg_len = tiles_per_seg * BS;
    g.nseg = (int)((ntiles + tiles_per_seg - 1) / tiles_per_seg);
    return g;
}

int record_start(nb_ctx *c, hipEvent_t *e0) {
    *e0 = nullptr;
    if (!c->timing) return NB_OK;
    NB_HIP(hipEventCreate(e0));
    NB_HIP(hipEventRecord(*e0, c->stream));
    return NB_OK;
}

int record_stop(nb_ctx *c, hipEvent_t e0) {
    if (!c->timing || !e0) return NB_OK;
    hipEvent_t e1;
    NB_HIP(hipEventCreate(&e1));
    NB_HIP(hipEventRecord(e1, c->stream));
    c->pending.emplace_back(e0, e1);
    return NB_OK;
}

int collect_timing(nb_ctx *c) {
    for (auto &p : c->pending) {
        NB_HIP(hipEventSynchronize(p.second));
        float ms = 0.f;
        NB_HIP(hipEventElapsedTime(&ms, p.first, p.second));
        c->tms += ms;
        c->tcount += 1;
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    c->pending.clear();
    return NB_OK;
}

// Refresh the compacted FP64 copies of the central sources.
__global__ void k_gather_central(const double *__restrict__ x, const double *__restrict__ y,
                                 const double *__restrict__ z, const double *__restrict__ m,
                                 const int64_t *__restrict__ idx, int64_t k, double *cx,
                                 double *cy, double *cz, double *cm) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= k) return;
    const int64_t i = idx[t];
    cx[t] = x[i];
    cy[t] = y[i];
    cz[t] = z[i];
    cm[t] = m[i];
}

// The potential of each central body summed over every source in FP64 (the
// FP32 kernel gave it the non-central sources' terms in FP32): one block per
// central body, a fixed-order block reduction, the result overwrites pot.
__global__ __launch_bounds__(256) void k_central_target_pot(const double *__restrict__ x,
                                                            const double *__restrict__ y,
                                                            const double *__restrict__ z,
                                                            const double *__restrict__ m,
                                                            const double *__restrict__ h, int64_t n,
                                                            const int64_t *__restrict__ cidx,
                                                            double G, double *__restrict__ pot) {
    __shared__ double red[4];
    const int64_t i = cidx[blockIdx.x];
    const double xi = x[i], yi = y[i], zi = z[i], hi = h[i];
    const double hinv = 1.0 / hi, hinv3 = hinv * hinv * hinv;
    double bx = 0.0, by = 0.0, bz = 0.0, q = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += 256) {
        if (j == i) continue;
        pair_f64<true>(x[j], y[j], z[j], m[j], xi, yi, zi, hi, hinv, hinv3, bx, by, bz, q);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) q += __shfl_xor(q, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = q;
    __syncthreads();
    if (threadIdx.x == 0) pot[i] = -G * (((red[0] + red[1]) + red[2]) + red[3]);
}

int gather_central(nb_ctx *c) {
    const int64_t k = c->ncentral;
    if (k <= 0) return NB_OK;
    k_gather_central<<<(unsigned)((k + 255) / 256), 256, 0, c->stream>>>(
        (const double *)c->x, (const double *)c->y, (const double *)c->z, (const double *)c->m,
        c->cidx, k, c->cx, c->cy, c->cz, c->cm);
    NB_HIP(hipGetLastError());
    return NB_OK;
}

// The all-pairs evaluation for every particle (acc, and potential when asked).
// Re-sort into Morton order when the order is stale.
static bool sort_due(const nb_ctx *c, int64_t upcoming) {
    return c->sort.interval > 0 && c->n >= 2 &&
           (!c->sort.valid || c->calls_since_sort + upcoming > c->sort.interval);
}

int maybe_sort(nb_ctx *c, int64_t upcoming) {
    if (sort_due(c, upcoming)) {
        NB_TRY(morton_reorder(c));
        c->calls_since_sort = 0;
    }
    return NB_OK;
}

// The source weights change only with the masses (upload) and their order
// (a re-sort): computed when a force needs them and they are stale, never
// inside a captured graph (nb_kdk_steps refreshes them before capturing and
// before each graph launch).
static int ensure_source_weights(nb_ctx *c) {
    if (c->wsrc_valid || c->n <= 0) return NB_OK;
    NB_TRY(launch_source_weights(c->m, c->is_double, c->n, c->wsrc, c->stream));
    c->wsrc_valid = true;
    return NB_OK;
}

// Work of a KDK step that rides on run_force's kernels: the drift has written
// the tile boxes already (boxes_ready), and the kick of `kick` (x dt) may be
// done by the finalize (kicked reports whether it was).
struct StepFusion {
    bool boxes_ready = false;
    int geo_ready = 0;     // the drift stored the target geometry (1), with the B operands (2)
    double kick = 0.0;
    bool kicked = false;
};

// pot_only: the FP32-pair kernels write the potential alone (the caller keeps
// the accelerations; the FP64 kernels and passes still write both).
int run_force(nb_ctx *c, bool with_pot, bool allow_timing, bool allow_sort = true,
              bool pot_only = false, StepFusion *fuse = nullptr) {
    hipStream_t s = c->stream;
    const int64_t n = c->n;
    if (n == 0) return NB_OK;
    const bool resorted = allow_sort && sort_due(c, 1);   // the boxes of a drift no longer hold
    if (allow_sort) NB_TRY(maybe_sort(c, 1));
    c->force_calls++;
    c->calls_since_sort++;
    if (!(fuse && fuse->boxes_ready && !resorted))
        NB_TRY(launch_tile_boxes(c->x, c->y, c->z, c->is_double, n, c->tile_box, c->sub_box, s));
    const bool f64 = c->mode == NB_MODE_FP64;
    const int tpt = f64 ? 1 : c->tun.targets_per_thread;
    Geometry g = geometry(n, n, tpt, c->tun.j_segments);
    NB_TRY(ensure_part(c, g.nseg));
    NB_TRY(ensure_tbg(c, g));
    hipEvent_t e0 = nullptr;
    if (allow_timing) NB_TRY(record_start(c, &e0));
    if (c->is_double) {
        ForceArgs<double> a{};
        a.tx = (const double *)c->x; a.ty = (const double *)c->y; a.tz = (const double *)c->z;
        a.th = c->h; a.th_const = 0.0; a.nt = n;
        a.sx = (const double *)c->x; a.sy = (const double *)c->y; a.sz = (const double *)c->z;
        a.sm = (const double *)c->m; a.ns = n;
        a.tile_box = c->tile_box;
        a.sub_box = c->sub_box;
        a.tgeo = c->tgeo;
        a.tbg = c->tbg;
        a.tgeo_ready = fuse && !resorted ? fuse->geo_ready : 0;
        a.nflag = c->nflag_cap >= (size_t)g.nblk_i * (size_t)g.nseg * NW ? c->nflag : nullptr;
        a.use_origin = 1;
        a.nseg = g.nseg; a.seg_len = g.seg_len; a.nblk_i = g.nblk_i;
        a.G = c->G;
        a.ax = c->ax; a.ay = c->ay; a.az = c->az; a.pot = c->pot;
        a.part = c->part;
        a.skip_src = (c->mode == NB_MODE_SPLIT_CENTRAL && c->ncentral > 0) ? c->central : nullptr;
        a.accumulate = 0;
        if (!f64 && (!with_pot || pot_only) && far_path_weighted(c->tun.far_path)) {
            NB_TRY(ensure_source_weights(c));
            a.sw = c->wsrc;
        }
        if (f64) NB_TRY(launch_force_f64(a, with_pot, s));
        else NB_TRY(launch_force_split(a, tpt, with_pot, s, c->tun.far_path, !(with_pot && pot_only)));
    } else {
        ForceArgs<float> a{};
        a.tx = (const float *)c->x; a.ty = (const float *)c->y; a.tz = (const float *)c->z;
        a.th = c->h; a.th_const = 0.0; a.nt = n;
        a.sx = (const float *)c->x; a.sy = (const float *)c->y; a.sz = (const float *)c->z;
        a.sm = (const float *)c->m; a.ns = n;
        a.tile_box = c->tile_box;
        a.sub_box = c->sub_box;
        a.tgeo = c->tgeo;
        a.tbg = c->tbg;
        a.tgeo_ready = fuse && !resorted ? fuse->geo_ready : 0;
        a.nflag = c->nflag_cap >= (size_t)g.nblk_i * (size_t)g.nseg * NW ? c->nflag : nullptr;
        a.use_origin = 0;
        a.nseg = g.nseg; a.seg_len = g.seg_len; a.nblk_i = g.nblk_i;
        a.G = c->G;
        a.ax = c->ax; a.ay = c->ay; a.az = c->az; a.pot = c->pot;
        a.part = c->part;
        a.skip_src = nullptr;
        a.accumulate = 0;
        const bool acc_out = !(with_pot && pot_only);
        if ((!with_pot || !acc_out) && c->tun.far_path == FAR_MFMA_BF16) {
            NB_TRY(ensure_source_weights(c));
            a.sw = c->wsrc;
        }
        NB_TRY(launch_force_f32(a, tpt, with_pot, s, acc_out, c->tun.far_path));
    }
    if (allow_timing) NB_TRY(record_stop(c, e0));
    // the kick rides on the finalize when nothing else adds to the accelerations
    const bool kick_here = fuse && !with_pot && g.nseg > 1 && c->bg.empty() && !(c->mond_acc0 > 0.0) &&
                           !(c->mode == NB_MODE_SPLIT_CENTRAL && c->ncentral > 0);
    if (kick_here) {
        NB_TRY(launch_finalize_kick(c->part, g.nseg, n, c->G, c->ax, c->ay, c->az, c->vx, c->vy, c->vz,
                                    fuse->kick, c->is_double, s));
        fuse->kicked = true;
    } else if (g.nseg > 1) {
        NB_TRY(launch_finalize(c->part, g.nseg, n, c->G, c->ax, c->ay, c->az, c->pot, with_pot, s,
                               !(with_pot && pot_only) || f64));
    }
    // FP64 central-body terms: every target, sources = the central particles.
    if (c->mode == NB_MODE_SPLIT_CENTRAL && c->ncentral > 0) {
        const int64_t k = c->ncentral;
        NB_TRY(gather_central(c));
        Geometry gc = geometry(n, k, 1, 1);
        ForceArgs<double> a{};
        a.tx = (const double *)c->x; a.ty = (const double *)c->y; a.tz = (const double *)c->z;
        a.th = c->h; a.nt = n;
        a.sx = c->cx; a.sy = c->cy; a.sz = c->cz; a.sm = c->cm; a.ns = k;
        a.tile_box = nullptr;
        a.nseg = 1; a.seg_len = gc.seg_len; a.nblk_i = gc.nblk_i;
        a.G = c->G;
        a.ax = c->ax; a.ay = c->ay; a.az = c->az; a.pot = c->pot;
        a.accumulate = 1;
        NB_TRY(launch_force_f64(a, with_pot, s));
        if (with_pot) {
            k_central_target_pot<<<(unsigned)k, 256, 0, s>>>((const double *)c->x, (const double *)c->y,
                                                            (const double *)c->z, (const double *)c->m,
                                                            c->h, n, c->cidx, c->G, c->pot);
            NB_HIP(hipGetLastError());
        }
    }
    for (const BgField &f : c->bg)
        NB_TRY(background(c->x, c->y, c->z, n, c->is_double, f, c->ax, c->ay, c->az,
                          with_pot ? c->pot : nullptr, s));
    if (c->mond_acc0 > 0.0) NB_TRY(mond(c->ax, c->ay, c->az, n, c->mond_acc0, c->mond_index, s));
    return NB_OK;
}

int upload_scalar_h(nb_ctx *c) { return fill_h(c->coll, c->n, c->soft, c->h, c->stream); }

__global__ void k_mark_central(const int64_t *__restrict__ perm, int64_t n, int64_t k,
                               uint8_t *__restrict__ flags) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t u = perm ? perm[i] : i;
    flags[i] = u < k ? 1 : 0;
}
__global__ void k_central_index(const int64_t *__restrict__ perm, int64_t n, int64_t k,
                                int64_t *__restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t u = perm ? perm[i] : i;
    if (u < k) idx[u] = i;
}

int rebuild_central(nb_ctx *c) {
    if (c->ncentral <= 0 || c->n == 0) return NB_OK;
    const unsigned g = (unsigned)((c->n + 255) / 256);
    k_mark_central<<<g, 256, 0, c->stream>>>(c->perm, c->n, c->ncentral, c->central);
    NB_HIP(hipGetLastError());
    k_central_index<<<g, 256, 0, c->stream>>>(c->perm, c->n, c->ncentral, c->cidx);
    NB_HIP(hipGetLastError());
    return NB_OK;
}
}  // namespace

namespace nb {
int ctx_run_force(nb_ctx *c, bool with_pot, bool allow_timing) {
    return run_force(c, with_pot, allow_timing);
}
int ctx_rebuild_central(nb_ctx *c) { return rebuild_central(c); }
int ctx_gather_central(nb_ctx *c) { return gather_central(c); }
void ctx_drop_graph(nb_ctx *c) { drop_graph(c); }
int ctx_ensure_part(nb_ctx *c, int nseg) { return ensure_part(c, nseg); }
}  // namespace nb

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

int nb_version(void) {
    return 10000 * NB_VERSION_MAJOR + 100 * NB_VERSION_MINOR + NB_VERSION_PATCH;
}

const char *nb_last_error(void) { return nb::last_error(); }

int nb_device_info(nb_devinfo *out) {
    clear_error();
    if (!out) {
        set_error("null output");
        return NB_ERR_INVALID;
    }
    int dev = 0;
    NB_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    NB_HIP(hipGetDeviceProperties(&p, dev));
    memset(out, 0, sizeof *out);
    snprintf(out->name, sizeof out->name, "%s", p.name);
    snprintf(out->arch, sizeof out->arch, "%s", p.gcnArchName);
    out->device = dev;
    out->compute_units = p.multiProcessorCount;
    out->clock_khz = p.clockRate;
    out->lds_per_block = (int32_t)p.sharedMemPerBlock;
    out->warp_size = p.warpSize;
    out->max_threads_per_block = p.maxThreadsPerBlock;
    out->global_mem_bytes = (int64_t)p.totalGlobalMem;
    out->l2_bytes = (int64_t)p.l2CacheSize;
    return NB_OK;
}

int nb_create(nb_ctx **out, int64_t n, int32_t precision_mode) {
    clear_error();
    if (!out) {
        set_error("null output handle");
        return NB_ERR_INVALID;
    }
    *out = nullptr;
    if (n < 0 || n > ((int64_t)1 << 31)) {
        set_error("particle count %lld out of range [0, 2^31]", (long long)n);
        return NB_ERR_INVALID;
    }
    if (precision_mode < NB_MODE_FP64 || precision_mode > NB_MODE_SPLIT_CENTRAL) {
        set_error("unknown precision mode %d", precision_mode);
        return NB_ERR_INVALID;
    }
    nb_ctx *c = new (std::nothrow) nb_ctx();
    if (!c) {
        set_error("host allocation failed");
        return NB_ERR_NOMEM;
    }
    c->n = n;
    c->mode = precision_mode;
    c->is_double = precision_mode != NB_MODE_FP32;
    int rc = NB_OK;
    do {
        hipError_t e = hipGetDevice(&c->dev);
        if (e != hipSuccess) {
            set_error("hipGetDevice failed: %s", hipGetErrorString(e));
            rc = NB_ERR_HIP;
            break;
        }
        e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            set_error("hipStreamCreate failed: %s", hipGetErrorString(e));
            rc = NB_ERR_HIP;
            break;
        }
        const size_t es = c->is_double ? sizeof(double) : sizeof(float);
        const size_t nn = (size_t)n;
        if ((rc = dalloc(&c->x, es * nn)) || (rc = dalloc(&c->y, es * nn)) ||
            (rc = dalloc(&c->z, es * nn)) || (rc = dalloc(&c->vx, es * nn)) ||
            (rc = dalloc(&c->vy, es * nn)) || (rc = dalloc(&c->vz, es * nn)) ||
            (rc = dalloc(&c->m, es * nn)) || (rc = dalloc(&c->h, sizeof(double) * nn)) ||
            (rc = dalloc(&c->coll, nn)) || (rc = dalloc(&c->ax, sizeof(double) * nn)) ||
            (rc = dalloc(&c->ay, sizeof(double) * nn)) ||
            (rc = dalloc(&c->az, sizeof(double) * nn)) ||
            (rc = dalloc(&c->pot, sizeof(double) * nn)) ||
            (rc = dalloc(&c->tile_box, sizeof(double) * 8 * (size_t)((n + BS - 1) / BS + 1))) ||
            (rc = dalloc(&c->sub_box, sizeof(double) * 8 * NQ * (size_t)((n + BS - 1) / BS + 1))) ||
            (rc = dalloc(&c->tgeo, sizeof(double) * TG_STRIDE * (size_t)((n + BS - 1) / BS + 1))) ||
            (rc = dalloc(&c->stage, sizeof(double) * 3 * nn)) ||
            (rc = dalloc(&c->wsrc, sizeof(float) * nn)) ||
            (rc = dalloc(&c->red, sizeof(double) * (ENERGY_BLOCKS * 8 + 8))) ||
            (rc = dalloc(&c->time, sizeof(double) * 2)) || (rc = dalloc(&c->flag, sizeof(unsigned))) ||
            (rc = dalloc(&c->central, nn)))
            break;
        e = hipMemsetAsync(c->ax, 0, sizeof(double) * nn, c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(c->ay, 0, sizeof(double) * nn, c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(c->az, 0, sizeof(double) * nn, c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(c->pot, 0, sizeof(double) * nn, c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(c->central, 0, nn, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            set_error("initialising device buffers failed: %s", hipGetErrorString(e));
            rc = NB_ERR_HIP;
            break;
        }
        c->tun.targets_per_thread = default_tpt(n);
        c->tun.far_path = default_far_path(precision_mode, n);
        c->tun.j_segments = default_segments(n, c->tun.targets_per_thread, c->tun.far_path);
        c->tun.sort_interval = precision_mode == NB_MODE_FP64 ? 0 : 16;
        tuner_lookup(c->mode, n, &c->tun);
        c->sort.interval = c->tun.sort_interval;
    } while (0);
    if (rc != NB_OK) {
        std::string keep = nb::last_error();
        nb_destroy(c);
        set_error("%s", keep.c_str());
        return rc;
    }
    *out = c;
    return NB_OK;
}

int nb_destroy(nb_ctx *c) {
    clear_error();
    if (!c) return NB_OK;
    (void)hipSetDevice(c->dev);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &p : c->pending) {
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    drop_graph(c);
    dfree(c->x); dfree(c->y); dfree(c->z);
    dfree(c->vx); dfree(c->vy); dfree(c->vz);
    dfree(c->m); dfree(c->h); dfree(c->coll);
    dfree(c->ax); dfree(c->ay); dfree(c->az); dfree(c->pot);
    dfree(c->perm);
    dfree(c->tile_box); dfree(c->sub_box); dfree(c->tgeo); dfree(c->tbg); dfree(c->nflag); dfree(c->part); dfree(c->stage); dfree(c->red); dfree(c->wsrc);
    dfree(c->time); dfree(c->flag); dfree(c->central);
    dfree(c->cx); dfree(c->cy); dfree(c->cz); dfree(c->cm); dfree(c->cidx);
    dfree(c->ti_end); dfree(c->ti_beg);
    dfree(c->ad.flags); dfree(c->ad.idx); dfree(c->ad.cnt); dfree(c->ad.red); dfree(c->ad.buf);
    dfree(c->ad.tmp);
    sort_free(c);
    snapshot_free(c);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return NB_OK;
}

int nb_set_params(nb_ctx *c, double G, const double soft[6]) {
    NB_TRY(check_ctx(c, false));
    if (!soft || !isfinite(G)) {
        set_error("invalid G or softening table");
        return NB_ERR_INVALID;
    }
    for (int k = 0; k < 6; ++k) {
        if (!(soft[k] >= 0.0) || !isfinite(soft[k])) {
            set_error("softening length %d must be finite and >= 0 (got %g)", k + 1, soft[k]);
            return NB_ERR_INVALID;
        }
    }
    drop_graph(c);
    c->G = G;
    for (int k = 0; k < 6; ++k) c->soft.h[k] = soft[k];
    if (c->uploaded) {
        NB_TRY(upload_scalar_h(c));
        NB_HIP(hipStreamSynchronize(c->stream));
    }
    return NB_OK;
}

int nb_upload(nb_ctx *c, const double *pos, const double *vel, const double *mass,
              const uint8_t *collection) {
    NB_TRY(check_ctx(c, false));
    const int64_t n = c->n;
    if (n > 0 && (!pos || !mass || !collection)) {
        set_error("pos, mass and collection are required");
        return NB_ERR_INVALID;
    }
    for (int64_t i = 0; i < n; ++i) {
        if (collection[i] < 1 || collection[i] > 6) {
            set_error("collection[%lld] = %d outside 1..6", (long long)i, (int)collection[i]);
            return NB_ERR_INVALID;
        }
    }
    drop_graph(c);
    // Upload resets the internal order to the caller's order.
    dfree(c->perm);
    c->sort.valid = false;
    hipStream_t s = c->stream;
    const size_t nn = (size_t)n;
    if (n > 0) {
        NB_HIP(hipMemcpyAsync(c->stage, pos, sizeof(double) * 3 * nn, hipMemcpyHostToDevice, s));
        NB_TRY(to_soa(c->stage, nullptr, n, c->x, c->y, c->z, c->is_double, s));
        NB_HIP(hipStreamSynchronize(s));
        if (vel) {
            NB_HIP(hipMemcpyAsync(c->stage, vel, sizeof(double) * 3 * nn, hipMemcpyHostToDevice, s));
            NB_TRY(to_soa(c->stage, nullptr, n, c->vx, c->vy, c->vz, c->is_double, s));
        } else {
            const size_t es = c->is_double ? sizeof(double) : sizeof(float);
            NB_HIP(hipMemsetAsync(c->vx, 0, es * nn, s));
            NB_HIP(hipMemsetAsync(c->vy, 0, es * nn, s));
            NB_HIP(hipMemsetAsync(c->vz, 0, es * nn, s));
        }
        NB_HIP(hipStreamSynchronize(s));
        NB_HIP(hipMemcpyAsync(c->stage, mass, sizeof(double) * nn, hipMemcpyHostToDevice, s));
        NB_TRY(scalar_in(c->stage, nullptr, n, c->m, c->is_double, s));
        c->wsrc_valid = false;
        NB_HIP(hipMemcpyAsync(c->coll, collection, nn, hipMemcpyHostToDevice, s));
        NB_TRY(upload_scalar_h(c));
        NB_HIP(hipStreamSynchronize(s));
    }
    c->uploaded = true;
    c->force_calls = 0;
    c->calls_since_sort = 0;
    NB_TRY(rebuild_central(c));
    NB_HIP(hipStreamSynchronize(s));
    return NB_OK;
}

int nb_upload_acc(nb_ctx *c, const double *acc, const double *pot) {
    NB_TRY(check_ctx(c));
    const int64_t n = c->n;
    if (n == 0) return NB_OK;
    hipStream_t s = c->stream;
    const size_t nn = (size_t)n;
    // caller order -> internal order: internal i takes caller perm[i]
    if (acc) {
        NB_HIP(hipMemcpyAsync(c->stage, acc, sizeof(double) * 3 * nn, hipMemcpyHostToDevice, s));
        NB_TRY(to_soa(c->stage, c->perm, n, c->ax, c->ay, c->az, true, s));
        NB_HIP(hipStreamSynchronize(s));
    }
    if (pot) {
        NB_HIP(hipMemcpyAsync(c->stage, pot, sizeof(double) * nn, hipMemcpyHostToDevice, s));
        NB_TRY(scalar_in(c->stage, c->perm, n, c->pot, true, s));
        NB_HIP(hipStreamSynchronize(s));
    }
    return NB_OK;
}

int nb_download(nb_ctx *c, double *pos, double *vel, double *acc, double *pot) {
    NB_TRY(check_ctx(c));
    const int64_t n = c->n;
    if (n == 0) return NB_OK;
    hipStream_t s = c->stream;
    const size_t nn = (size_t)n;
    if (pos) {
        NB_TRY(to_aos(c->x, c->y, c->z, c->perm, n, c->is_double, c->stage, s));
        NB_HIP(hipMemcpyAsync(pos, c->stage, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, s));
        NB_HIP(hipStreamSynchronize(s));
    }
    if (vel) {
        NB_TRY(to_aos(c->vx, c->vy, c->vz, c->perm, n, c->is_double, c->stage, s));
        NB_HIP(hipMemcpyAsync(vel, c->stage, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, s));
        NB_HIP(hipStreamSynchronize(s));
    }
    if (acc) {
        NB_TRY(to_aos(c->ax, c->ay, c->az, c->perm, n, true, c->stage, s));
        NB_HIP(hipMemcpyAsync(acc, c->stage, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, s));
        NB_HIP(hipStreamSynchronize(s));
    }
    if (pot) {
        NB_TRY(scalar_out(c->pot, c->perm, n, c->stage, s));
        NB_HIP(hipMemcpyAsync(pot, c->stage, sizeof(double) * nn, hipMemcpyDeviceToHost, s));
        NB_HIP(hipStreamSynchronize(s));
    }
    return NB_OK;
}

int nb_compute_force(nb_ctx *c) {
    NB_TRY(check_ctx(c));
    NB_TRY(run_force(c, false, true));
    return NB_OK;
}

int nb_compute_potential(nb_ctx *c, int32_t add_self_term) {
    NB_TRY(check_ctx(c));
    // The potential is produced together with the acceleration by the same
    // kernel; keep the caller's accelerations by saving them around it.
    hipStream_t s = c->stream;
    const size_t nb3 = sizeof(double) * (size_t)c->n;
    NB_TRY(maybe_sort(c, 1));   // reorder first, so the saved accelerations stay aligned
    if (c->n > 0) {
        NB_HIP(hipMemcpyAsync(c->stage, c->ax, nb3, hipMemcpyDeviceToDevice, s));
        NB_HIP(hipMemcpyAsync(c->stage + c->n, c->ay, nb3, hipMemcpyDeviceToDevice, s));
        NB_HIP(hipMemcpyAsync(c->stage + 2 * c->n, c->az, nb3, hipMemcpyDeviceToDevice, s));
    }
    NB_TRY(run_force(c, true, false, false, true));
    if (c->n > 0) {
        NB_HIP(hipMemcpyAsync(c->ax, c->stage, nb3, hipMemcpyDeviceToDevice, s));
        NB_HIP(hipMemcpyAsync(c->ay, c->stage + c->n, nb3, hipMemcpyDeviceToDevice, s));
        NB_HIP(hipMemcpyAsync(c->az, c->stage + 2 * c->n, nb3, hipMemcpyDeviceToDevice, s));
    }
    if (add_self_term) NB_TRY(self_term(c->m, c->h, c->n, c->G, c->is_double, c->pot, s));
    NB_HIP(hipStreamSynchronize(s));
    return NB_OK;
}

static int probe_points(nb_ctx *c, int64_t npts, const double *pts, double h, double *acc,
                        double *pot) {
    NB_TRY(check_ctx(c));
    if (npts < 0 || (npts > 0 && !pts) || !(h >= 0.0) || !isfinite(h)) {
        set_error("invalid probe points or softening");
        return NB_ERR_INVALID;
    }
    if (npts == 0) return NB_OK;
    hipStream_t s = c->stream;
    const bool with_pot = pot != nullptr;
    double *buf = nullptr;
    // target xyz (double or float), outputs ax ay az pot, packed xyz staging
    NB_TRY(dalloc(&buf, sizeof(double) * (size_t)npts * 10));
    double *tx = buf, *ty = buf + npts, *tz = buf + 2 * npts;
    double *oax = buf + 3 * npts, *oay = buf + 4 * npts, *oaz = buf + 5 * npts, *opot = buf + 6 * npts;
    double *aos = buf + 7 * npts;
    int rc = NB_OK;
    do {
        hipError_t e = hipMemcpyAsync(aos, pts, sizeof(double) * 3 * (size_t)npts,
                                      hipMemcpyHostToDevice, s);
        if (e != hipSuccess) { set_error("probe upload: %s", hipGetErrorString(e)); rc = NB_ERR_HIP; break; }
        if ((rc = launch_tile_boxes(c->x, c->y, c->z, c->is_double, c->n, c->tile_box, c->sub_box, s))) break;
        Geometry g = geometry(npts, c->n, 1, 1);
        if (c->is_double) {
            if ((rc = to_soa(aos, nullptr, npts, tx, ty, tz, true, s))) break;
            ForceArgs<double> a{};
            a.tx = tx; a.ty = ty; a.tz = tz; a.th = nullptr; a.th_const = h; a.nt = npts;
            a.sx = (const double *)c->x; a.sy = (const double *)c->y; a.sz = (const double *)c->z;
            a.sm = (const double *)c->m; a.ns = c->n;
            a.tile_box = c->tile_box; a.sub_box = c->sub_box; a.use_origin = 1;
            a.nseg = 1; a.seg_len = g.seg_len; a.nblk_i = g.nblk_i; a.G = c->G;
            a.ax = oax; a.ay = oay; a.az = oaz; a.pot = opot; a.accumulate = 0;
            rc = c->mode == NB_MODE_FP64 ? launch_force_f64(a, with_pot, s)
                                         : launch_force_split(a, 1, with_pot, s);
            if (rc) break;
        } else {
            // FP32 state: probe coordinates are rounded to FP32 like the particles.
            float *fx = (float *)tx, *fy = (float *)ty, *fz = (float *)tz;
            if ((rc = to_soa(aos, nullptr, npts, fx, fy, fz, false, s))) break;
            ForceArgs<float> a{};
            a.tx = fx; a.ty = fy; a.tz = fz; a.th = nullptr; a.th_const = h; a.nt = npts;
            a.sx = (const float *)c->x; a.sy = (const float *)c->y; a.sz = (const float *)c->z;
            a.sm = (const float *)c->m; a.ns = c->n;
            a.tile_box = c->tile_box; a.sub_box = c->sub_box; a.use_origin = 0;
            a.nseg = 1; a.seg_len = g.seg_len; a.nblk_i = g.nblk_i; a.G = c->G;
            a.ax = oax; a.ay = oay; a.az = oaz; a.pot = opot; a.accumulate = 0;
            if ((rc = launch_force_f32(a, 1, with_pot, s))) break;
            if ((rc = to_soa(aos, nullptr, npts, tx, ty, tz, true, s))) break;
        }
        for (const BgField &f : c->bg)
            if ((rc = background(tx, ty, tz, npts, true, f, oax, oay, oaz, with_pot ? opot : nullptr, s))) break;
        if (rc) break;
        if (acc) {
            if ((rc = to_aos(oax, oay, oaz, nullptr, npts, true, aos, s))) break;
            e = hipMemcpyAsync(acc, aos, sizeof(double) * 3 * (size_t)npts, hipMemcpyDeviceToHost, s);
            if (e != hipSuccess) { set_error("probe download: %s", hipGetErrorString(e)); rc = NB_ERR_HIP; break; }
        }
        if (pot) {
            e = hipMemcpyAsync(pot, opot, sizeof(double) * (size_t)npts, hipMemcpyDeviceToHost, s);
            if (e != hipSuccess) { set_error("probe download: %s", hipGetErrorString(e)); rc = NB_ERR_HIP; break; }
        }
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) { set_error("probe sync: %s", hipGetErrorString(e)); rc = NB_ERR_HIP; break; }
    } while (0);
    (void)hipStreamSynchronize(s);
    dfree(buf);
    return rc;
}

int nb_force_at_points(nb_ctx *c, int64_t npts, const double *pts, double h, double *acc) {
    if (!acc && npts > 0) {
        set_error("null output");
        return NB_ERR_INVALID;
    }
    return probe_points(c, npts, pts, h, acc, nullptr);
}

int nb_potential_at_points(nb_ctx *c, int64_t npts, const double *pts, double h, double *pot) {
    if (!pot && npts > 0) {
        set_error("null output");
        return NB_ERR_INVALID;
    }
    return probe_points(c, npts, pts, h, nullptr, pot);
}

int nb_half_kick(nb_ctx *c, int32_t forward, double dt) {
    NB_TRY(check_ctx(c));
    if (!isfinite(dt) || (forward != 1 && forward != -1)) {
        set_error("invalid half kick (forward=%d, dt=%g)", forward, dt);
        return NB_ERR_INVALID;
    }
    // Vel += Acc * forward * 0.5 * dt
    NB_TRY(kick(c->vx, c->vy, c->vz, c->ax, c->ay, c->az, c->n, (double)forward * 0.5 * dt,
                c->is_double, c->stream));
    return NB_OK;
}

int nb_drift(nb_ctx *c, double dt) {
    NB_TRY(check_ctx(c));
    if (!isfinite(dt)) {
        set_error("invalid drift dt %g", dt);
        return NB_ERR_INVALID;
    }
    // t_new - t_old == dt exactly when t_old == 0.
    NB_TRY(set_time(c->time, 0, 0.0, c->stream));
    NB_TRY(drift(c->x, c->y, c->z, c->vx, c->vy, c->vz, c->n, dt, c->time, 0, 1, c->is_double,
                 c->stream));
    return NB_OK;
}

int nb_kick(nb_ctx *c, double dt) {
    NB_TRY(check_ctx(c));
    if (!isfinite(dt)) {
        set_error("invalid kick dt %g", dt);
        return NB_ERR_INVALID;
    }
    NB_TRY(kick(c->vx, c->vy, c->vz, c->ax, c->ay, c->az, c->n, dt, c->is_double, c->stream));
    return NB_OK;
}

// Drift (with the tile boxes of the new positions), force, kick (with the
// finalize of the accelerations where it can).
static int enqueue_step(nb_ctx *c, int64_t k, double dt, bool allow_timing, bool allow_sort) {
    const int rd = (int)(k & 1), wr = 1 - rd;
    StepFusion fuse;
    if (c->n > 0) {
        // The target geometry of a force call with one target per lane and
        // several source segments rides on the drift as well, stored as that
        // call's launch would store it (origin mode, B operands of far path 4).
        StepGeometry sg{};
        const StepGeometry *gp = nullptr;
        if (c->mode != NB_MODE_FP64 && c->tun.targets_per_thread == 1 && c->tgeo) {
            Geometry g = geometry(c->n, c->n, 1, c->tun.j_segments);
            if (g.nseg > 1) {
                NB_TRY(ensure_part(c, g.nseg));
                NB_TRY(ensure_tbg(c, g));
                const bool bf16 = c->tun.far_path == FAR_MFMA_BF16;
                sg.h = c->h;
                sg.tgeo = c->tgeo;
                sg.tbg = bf16 ? c->tbg : nullptr;
                sg.use_origin = (c->is_double || bf16) ? 1 : 0;
                gp = &sg;
                fuse.geo_ready = sg.tbg ? 2 : 1;
            }
        }
        NB_TRY(launch_drift_boxes(c->x, c->y, c->z, c->vx, c->vy, c->vz, c->n, dt, c->time, rd, wr, c->is_double,
                                  c->tile_box, c->sub_box, c->stream, gp));
        fuse.boxes_ready = true;
    } else {
        NB_TRY(drift(c->x, c->y, c->z, c->vx, c->vy, c->vz, c->n, dt, c->time, rd, wr, c->is_double,
                     c->stream));
    }
    fuse.kick = dt;
    NB_TRY(run_force(c, false, allow_timing, allow_sort, false, &fuse));
    if (!fuse.kicked)
        NB_TRY(kick(c->vx, c->vy, c->vz, c->ax, c->ay, c->az, c->n, dt, c->is_double, c->stream));
    return NB_OK;
}

int nb_kdk_steps(nb_ctx *c, int64_t nsteps, double dt, double t_begin, int32_t use_graph) {
    NB_TRY(check_ctx(c));
    if (nsteps < 0 || !isfinite(dt) || !isfinite(t_begin)) {
        set_error("invalid kdk arguments (nsteps=%lld, dt=%g)", (long long)nsteps, dt);
        return NB_ERR_INVALID;
    }
    hipStream_t s = c->stream;
    NB_TRY(set_time(c->time, 0, t_begin, s));
    // Graph replay needs a static schedule: re-sorting happens between
    // graph launches, never inside one.
    const bool graph_ok = use_graph && nsteps >= 2;
    int64_t k = 0;
    if (graph_ok) {
        const int64_t chunk = std::min<int64_t>(nsteps & ~(int64_t)1, 16);
        const bool reuse = c->gexec && c->g_steps == chunk && c->g_dt == dt &&
                           c->g_tun.targets_per_thread == c->tun.targets_per_thread &&
                           c->g_tun.j_segments == c->tun.j_segments &&
                           c->g_tun.far_path == c->tun.far_path && c->g_nbg == c->bg.size();
        if (!reuse) {
            drop_graph(c);
            // Make sure partial buffers exist before capture (no allocation inside).
            Geometry g = geometry(c->n, c->n, c->mode == NB_MODE_FP64 ? 1 : c->tun.targets_per_thread,
                                  c->tun.j_segments);
            NB_TRY(ensure_part(c, g.nseg));
            NB_TRY(ensure_tbg(c, g));
            NB_TRY(ensure_source_weights(c));
            NB_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int rc = NB_OK;
            for (int64_t q = 0; q < chunk && rc == NB_OK; ++q)
                rc = enqueue_step(c, q, dt, false, false);
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamEndCapture(s, &graph);
            if (rc != NB_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            if (e != hipSuccess) {
                set_error("hipStreamEndCapture failed: %s", hipGetErrorString(e));
                if (graph) (void)hipGraphDestroy(graph);
                return NB_ERR_HIP;
            }
            e = hipGraphInstantiate(&c->gexec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                c->gexec = nullptr;
                set_error("hipGraphInstantiate failed: %s", hipGetErrorString(e));
                return NB_ERR_HIP;
            }
            c->g_steps = chunk;
            c->g_dt = dt;
            c->g_tun = c->tun;
            c->g_nbg = c->bg.size();
            // Force calls made inside the capture were counted; undo.
            c->force_calls -= chunk;
            c->calls_since_sort -= chunk;
        }
        for (; k + c->g_steps <= nsteps; k += c->g_steps) {
            NB_TRY(maybe_sort(c, c->g_steps));
            NB_TRY(ensure_source_weights(c));   // after a re-sort the graph's forces need them again
            NB_HIP(hipGraphLaunch(c->gexec, s));
            c->force_calls += c->g_steps;
            c->calls_since_sort += c->g_steps;
        }
    }
    for (; k < nsteps; ++k) NB_TRY(enqueue_step(c, k, dt, true, true));
    return NB_OK;
}

int nb_energy(nb_ctx *c, double *K, double *W, double P[3], double L[3]) {
    NB_TRY(check_ctx(c));
    hipStream_t s = c->stream;
    if (W) NB_TRY(nb_compute_potential(c, 0));
    double host[8];
    NB_TRY(energy_reduce(c->x, c->y, c->z, c->vx, c->vy, c->vz, c->m, W ? c->pot : nullptr, c->n,
                         c->is_double, c->red, s));
    NB_HIP(hipMemcpyAsync(host, c->red + (size_t)ENERGY_BLOCKS * 8, sizeof host,
                          hipMemcpyDeviceToHost, s));
    NB_HIP(hipStreamSynchronize(s));
    if (K) *K = host[0];
    if (W) *W = 0.5 * host[1];
    if (P) { P[0] = host[2]; P[1] = host[3]; P[2] = host[4]; }
    if (L) { L[0] = host[5]; L[1] = host[6]; L[2] = host[7]; }
    // the outputs stay written; a NaN or Inf among the requested ones is
    // reported as NB_ERR_NONFINITE
    const char *bad = nullptr;
    if (K && !std::isfinite(host[0])) bad = "kinetic energy";
    else if (W && !std::isfinite(host[1])) bad = "potential energy";
    else if (P && !(std::isfinite(host[2]) && std::isfinite(host[3]) && std::isfinite(host[4])))
        bad = "momentum";
    else if (L && !(std::isfinite(host[5]) && std::isfinite(host[6]) && std::isfinite(host[7])))
        bad = "angular momentum";
    if (bad) {
        set_error("non-finite %s (NaN or Inf in the particle state)", bad);
        return NB_ERR_NONFINITE;
    }
    return NB_OK;
}

int nb_set_central(nb_ctx *c, int64_t count) {
    NB_TRY(check_ctx(c, false));
    if (count < 0 || count > c->n) {
        set_error("central count %lld out of range", (long long)count);
        return NB_ERR_INVALID;
    }
    if (c->mode != NB_MODE_SPLIT_CENTRAL && count > 0) {
        set_error("central bodies need NB_MODE_SPLIT_CENTRAL");
        return NB_ERR_STATE;
    }
    drop_graph(c);
    dfree(c->cx); dfree(c->cy); dfree(c->cz); dfree(c->cm); dfree(c->cidx);
    c->ncentral = count;
    if (count > 0) {
        NB_TRY(dalloc(&c->cx, sizeof(double) * count));
        NB_TRY(dalloc(&c->cy, sizeof(double) * count));
        NB_TRY(dalloc(&c->cz, sizeof(double) * count));
        NB_TRY(dalloc(&c->cm, sizeof(double) * count));
        NB_TRY(dalloc(&c->cidx, sizeof(int64_t) * count));
    }
    NB_HIP(hipMemsetAsync(c->central, 0, (size_t)c->n, c->stream));
    NB_TRY(rebuild_central(c));
    NB_HIP(hipStreamSynchronize(c->stream));
    return NB_OK;
}

int nb_add_background(nb_ctx *c, int32_t kind, const double *params) {
    NB_TRY(check_ctx(c, false));
    if (kind == 0) {
        drop_graph(c);
        c->bg.clear();
        return NB_OK;
    }
    if (kind < 1 || kind > 4 || !params) {
        set_error("unknown background field kind %d", kind);
        return NB_ERR_INVALID;
    }
    if (c->bg.size() >= 8) {
        set_error("at most 8 background fields");
        return NB_ERR_INVALID;
    }
    BgField f{};
    f.kind = kind;
    const int np = kind == 1 ? 4 : kind == 2 ? 5 : kind == 3 ? 7 : 3;
    for (int k = 0; k < np; ++k) {
        if (!isfinite(params[k])) {
            set_error("non-finite background parameter %d", k);
            return NB_ERR_INVALID;
        }
        f.p[k] = params[k];
    }
    drop_graph(c);
    c->bg.push_back(f);
    return NB_OK;
}

int nb_set_mond(nb_ctx *c, double acc0, double nu_index) {
    NB_TRY(check_ctx(c, false));
    if (!isfinite(acc0) || !(nu_index > 0.0) || !isfinite(nu_index)) {
        set_error("invalid MOND parameters (ACC0=%g, nu index=%g)", acc0, nu_index);
        return NB_ERR_INVALID;
    }
    drop_graph(c);
    c->mond_acc0 = acc0 > 0.0 ? acc0 : 0.0;
    c->mond_index = nu_index;
    return NB_OK;
}

int nb_sync(nb_ctx *c) {
    NB_TRY(check_ctx(c, false));
    NB_HIP(hipStreamSynchronize(c->stream));
    return NB_OK;
}

int nb_set_tuning(nb_ctx *c, const nb_tuning *t) {
    NB_TRY(check_ctx(c, false));
    if (!t || (t->targets_per_thread != 1 && t->targets_per_thread != 2 &&
               t->targets_per_thread != 4) ||
        t->j_segments < 1 || t->j_segments > 256 || t->sort_interval < 0 || t->far_path < 0 ||
        t->far_path > 4) {
        set_error("invalid tuning");
        return NB_ERR_INVALID;
    }
    if (c->mode == NB_MODE_FP32 && t->far_path != FAR_VALU && t->far_path != FAR_MFMA_BF16) {
        set_error("the FP32 mode takes far_path 0 or 4 (got %d)", t->far_path);
        return NB_ERR_INVALID;
    }
    // the matrix-core far paths hold one target per lane (as nb_tuner_load requires)
    if ((t->far_path == FAR_MFMA_SUM || t->far_path == FAR_MFMA_R2 ||
         t->far_path == FAR_MFMA_BF16) && t->targets_per_thread != 1) {
        set_error("far_path %d needs targets_per_thread 1 (got %d)", t->far_path,
                  t->targets_per_thread);
        return NB_ERR_INVALID;
    }
    drop_graph(c);
    c->tun = *t;
    c->sort.interval = t->sort_interval;
    return NB_OK;
}

int nb_get_tuning(nb_ctx *c, nb_tuning *t) {
    NB_TRY(check_ctx(c, false));
    if (!t) {
        set_error("null output");
        return NB_ERR_INVALID;
    }
    *t = c->tun;
    return NB_OK;
}

int nb_autotune(nb_ctx *c) {
    NB_TRY(check_ctx(c));
    return autotune_ctx(c);
}

int nb_set_timing(nb_ctx *c, int32_t enable) {
    NB_TRY(check_ctx(c, false));
    NB_TRY(collect_timing(c));
    c->timing = enable != 0;
    c->tcount = 0;
    c->tms = 0.0;
    return NB_OK;
}

int nb_get_timing(nb_ctx *c, int64_t *count, double *total_ms) {
    NB_TRY(check_ctx(c, false));
    NB_TRY(collect_timing(c));
    if (count) *count = c->tcount;
    if (total_ms) *total_ms = c->tms;
    return NB_OK;
}

}  // extern "C"
