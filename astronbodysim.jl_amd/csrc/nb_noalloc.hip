// nb_noalloc.hip -- nb_compute_force_noalloc: the all-pairs force on particle
// arrays the caller already holds in device memory, with no device allocation.
//
// The kernels are the ones nb_compute_force launches (k_tile_boxes,
// k_force_f32 / k_force_f64, k_finalize).  Their scratch -- the bounding box
// of every 256-source tile and, with several source segments, the
// per-segment partial sums -- lives in one file-scope device array instead of
// context buffers.  Each GPU holds its own copy of that array for the life of
// the process; its address is looked up once per device with
// hipGetSymbolAddress, and only that address is handed to kernels.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>

#include "nb_internal.h"

namespace nb {
namespace {

constexpr size_t SCRATCH_BYTES = (size_t)NB_NOALLOC_SCRATCH_BYTES;
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

std::mutex g_mu;
void *g_ptr[MAX_DEVICES] = {};   // device address of `scratch`, per device

// The scratch address on the current device (looked up on the first call
// there, with that device current).
int scratch_ptr(int dev, void **out) {
    if (dev < 0 || dev >= MAX_DEVICES) {
        set_error("device %d beyond the %d devices the no-alloc scratch tracks", dev, MAX_DEVICES);
        return NB_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                      hipGetErrorString(e));
            return NB_ERR_HIP;
        }
        g_ptr[dev] = p;
    }
    *out = g_ptr[dev];
    return NB_OK;
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace
}  // namespace nb

using namespace nb;

extern "C" int nb_compute_force_noalloc(int32_t precision_mode, int64_t n, const double *x,
                                        const double *y, const double *z, const double *m,
                                        const double *h, double h_const, double G, double *ax,
                                        double *ay, double *az, double *pot, void *stream) {
    clear_error();
    if (precision_mode != NB_MODE_SPLIT && precision_mode != NB_MODE_FP64) {
        set_error("nb_compute_force_noalloc: precision mode must be NB_MODE_SPLIT or NB_MODE_FP64 (got %d)",
                  precision_mode);
        return NB_ERR_INVALID;
    }
    if (n < 0 || n > NB_NOALLOC_MAX_N) {
        set_error("nb_compute_force_noalloc: n = %lld outside [0, %lld] (the no-alloc scratch holds "
                  "%lld bytes)", (long long)n, (long long)NB_NOALLOC_MAX_N,
                  (long long)NB_NOALLOC_SCRATCH_BYTES);
        return NB_ERR_INVALID;
    }
    if (n == 0) return NB_OK;
    if (!x || !y || !z || !m || !ax || !ay || !az) {
        set_error("nb_compute_force_noalloc: x, y, z, m, ax, ay and az are required");
        return NB_ERR_INVALID;
    }
    if (!h && !(h_const >= 0.0)) {
        set_error("nb_compute_force_noalloc: h_const must be >= 0");
        return NB_ERR_INVALID;
    }
    int dev = 0;
    NB_HIP(hipGetDevice(&dev));
    void *base = nullptr;
    NB_TRY(scratch_ptr(dev, &base));
    hipStream_t s = (hipStream_t)stream;
    const bool f64 = precision_mode == NB_MODE_FP64;

    // scratch layout: tile boxes, then the per-segment partials
    const int64_t ntiles = (n + BS - 1) / BS;
    const size_t box_bytes = align256(sizeof(double) * 8 * (size_t)ntiles);
    const size_t seg_bytes = sizeof(double) * 4 * (size_t)n;
    const int64_t nblk = (n + BS - 1) / BS;
    // about 8k blocks to fill the 256 CUs, as many segments as the scratch holds
    int64_t want = std::min<int64_t>(std::max<int64_t>((8192 + nblk - 1) / nblk, 1), 64);
    want = std::min<int64_t>(want, ntiles);
    const int64_t fit = (int64_t)((SCRATCH_BYTES - box_bytes) / seg_bytes);
    if (want > fit) want = fit;
    if (want < 2) want = 1;
    const int64_t tiles_per_seg = (ntiles + want - 1) / want;
    const int nseg = (int)((ntiles + tiles_per_seg - 1) / tiles_per_seg);
    double *box = (double *)base;
    double *part = nseg > 1 ? (double *)((unsigned char *)base + box_bytes) : nullptr;

    ForceArgs<double> a{};
    a.tx = x; a.ty = y; a.tz = z;
    a.th = h; a.th_const = h_const; a.nt = n;
    a.sx = x; a.sy = y; a.sz = z; a.sm = m; a.ns = n;
    a.tile_box = box;
    a.use_origin = 1;
    a.nseg = nseg; a.seg_len = tiles_per_seg * BS; a.nblk_i = (int)nblk;
    a.G = G;
    a.ax = ax; a.ay = ay; a.az = az; a.pot = pot;
    a.part = part;
    a.skip_src = nullptr;
    a.accumulate = 0;
    const bool with_pot = pot != nullptr;
    if (f64) {
        NB_TRY(launch_force_f64(a, with_pot, s));
    } else {
        NB_TRY(launch_tile_boxes(x, y, z, 1, n, box, nullptr, s));
        NB_TRY(launch_force_split(a, 1, with_pot, s, FAR_VALU, true));
    }
    if (nseg > 1) NB_TRY(launch_finalize(part, nseg, n, G, ax, ay, az, pot, with_pot, s, true));
    return NB_OK;
}
