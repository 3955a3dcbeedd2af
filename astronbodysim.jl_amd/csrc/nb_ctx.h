// nb_ctx.h -- the one-system context behind the opaque nb_ctx handle.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>
#include <vector>

#include "nb_integrate.h"
#include "nb_internal.h"

namespace nb {

// Morton-order reordering state (nb_sort.hip).
struct SortState {
    int interval = 0;          // re-sort every `interval` force calls (0: never)
    bool valid = false;        // the internal order is a Morton order
    void *tmp = nullptr;       // rocPRIM temporary storage
    size_t tmp_bytes = 0;
    uint64_t *keys = nullptr;  // [2][n]
    int64_t *idx = nullptr;    // [2][n]
    double *box = nullptr;     // global bounding box (8 doubles)
    void *spare = nullptr;     // n * 8 bytes gather target
};

// Adaptive-timestep state (nb_adaptive.hip).
struct AdaptState {
    double time_interval = 0.0;
    int64_t time_base = 0;
    double err_tol = 0.0, min_step = 0.0, max_step = 0.0;
    uint8_t *flags = nullptr;  // active flags
    int64_t *idx = nullptr;    // compacted active indices
    int64_t *cnt = nullptr;    // [0] active count
    int64_t *red = nullptr;    // end-step reduction: block partials, then (min, count, flag)
    // (minimum end step, particles at it) of the current Ti_endstep, when
    // next_valid; read back together with the error flag after each kick
    bool next_valid = false;
    int64_t next_min = 0, next_count = 0;
    double *buf = nullptr;     // gathered targets and their accelerations (7n doubles)
    void *tmp = nullptr;       // rocPRIM select storage
    size_t tmp_bytes = 0;
};

// One pinned host slot of the asynchronous snapshot download (nb_snapshot.hip).
constexpr int SNAP_SLOTS = 2;
struct SnapSlot {
    double *host = nullptr;    // pinned: pos 3n | vel 3n | acc 3n | pot n (caller order)
    size_t cap = 0;            // doubles
    hipEvent_t ev = nullptr;   // recorded after the slot's copies
    int what = 0;
    int64_t n = 0;
    bool pending = false;
};

}  // namespace nb

struct nb_ctx {
    int dev = 0;
    int mode = NB_MODE_SPLIT;
    int64_t n = 0;
    bool is_double = true;
    hipStream_t stream = nullptr;
    double G = 1.0;
    nb::SoftTable soft{{1e-4, 1e-4, 1e-4, 1e-4, 1e-4, 1e-4}};
    void *x = nullptr, *y = nullptr, *z = nullptr;
    void *vx = nullptr, *vy = nullptr, *vz = nullptr;
    void *m = nullptr;
    double *h = nullptr;
    uint8_t *coll = nullptr;
    double *ax = nullptr, *ay = nullptr, *az = nullptr, *pot = nullptr;
    int64_t *perm = nullptr;      // internal -> caller index (nullptr: identity)
    double *tile_box = nullptr;
    double *sub_box = nullptr;    // [ntiles][NQ][8] boxes of the tile quarters
    double *tgeo = nullptr;       // [target blocks][TG_STRIDE] geometry workspace of the force kernel
    unsigned *tbg = nullptr;      // [target blocks][BS][16] bf16 B operands (far path 4), on first need
    uint8_t *nflag = nullptr;     // one byte per force wave: near tiles left to the near pass
    size_t nflag_cap = 0;
    double *part = nullptr;
    int part_cap = 0;             // segments the partial buffer holds
    double *stage = nullptr;      // 3n doubles
    float *wsrc = nullptr;        // source weights m^(-2/3) (weighted far paths), n floats
    bool wsrc_valid = false;      // wsrc holds the weights of the current masses in the current order
    double *red = nullptr;        // energy scratch
    double *time = nullptr;       // two FP64 time slots for device-side drift
    unsigned *flag = nullptr;
    uint8_t *central = nullptr;   // internal flags: source evaluated in FP64
    int64_t ncentral = 0;
    double *cx = nullptr, *cy = nullptr, *cz = nullptr, *cm = nullptr;  // compacted central sources
    int64_t *cidx = nullptr;      // internal indices of the central sources
    nb::SortState sort;
    int64_t *ti_end = nullptr, *ti_beg = nullptr;   // adaptive integer timeline (internal order)
    nb::AdaptState ad;
    int64_t force_calls = 0;
    int64_t calls_since_sort = 0;
    nb_tuning tun{2, 1, 0, 0};
    bool uploaded = false;
    std::vector<nb::BgField> bg;
    double mond_acc0 = 0.0;       // > 0: Milgrom-1983 rescaling after the background fields
    double mond_index = 2.0;
    nb::SnapSlot snap[nb::SNAP_SLOTS];
    // timing
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    int64_t tcount = 0;
    double tms = 0.0;
    // graph cache for nb_kdk_steps
    hipGraphExec_t gexec = nullptr;
    int64_t g_steps = 0;
    double g_dt = 0.0;
    nb_tuning g_tun{0, 0, 0, 0};
    size_t g_nbg = 0;
};

namespace nb {
template <typename P>
inline int dalloc(P **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc((void **)p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
        return NB_ERR_NOMEM;
    }
    return NB_OK;
}

template <typename P>
inline void dfree(P *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

// nb_sort.hip
int morton_reorder(nb_ctx *c);
void sort_free(nb_ctx *c);
// nb_tune.hip
void tuner_lookup(int mode, int64_t n, nb_tuning *t);
int autotune_ctx(nb_ctx *c);
// nb_api.hip
int ctx_run_force(nb_ctx *c, bool with_pot, bool allow_timing);
int ctx_rebuild_central(nb_ctx *c);
int ctx_gather_central(nb_ctx *c);
void ctx_drop_graph(nb_ctx *c);
int ctx_ensure_part(nb_ctx *c, int nseg);
// nb_snapshot.hip
void snapshot_free(nb_ctx *c);
}  // namespace nb
