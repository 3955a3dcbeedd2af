// nb_snapshot.hip -- asynchronous snapshot download.
//
// The reference writes a snapshot by copying every particle to the host
// synchronously and then writing the file while the simulation waits
// (src/base/Output.jl:36-50, src/run.jl:90,116-131).  Here a snapshot is a
// stream-ordered device-to-host copy into one of two pinned host slots,
// marked by an event: the integration continues on the stream at once, and a
// writer (another host thread) waits for the slot's event, takes the data in
// the caller's particle order and writes the file while later steps run.
#include <hip/hip_runtime.h>
#include <cstring>

#include "nb_ctx.h"
#include "nb_integrate.h"

using namespace nb;

namespace {

constexpr int SNAP_POS = 1, SNAP_VEL = 2, SNAP_ACC = 4, SNAP_POT = 8;

int check(nb_ctx *c, int32_t slot) {
    clear_error();
    if (!c) {
        set_error("null context");
        return NB_ERR_INVALID;
    }
    if (slot < 0 || slot >= SNAP_SLOTS) {
        set_error("snapshot slot %d out of range [0, %d)", slot, SNAP_SLOTS);
        return NB_ERR_INVALID;
    }
    NB_HIP(hipSetDevice(c->dev));
    return NB_OK;
}

}  // namespace

namespace nb {

void snapshot_free(nb_ctx *c) {
    for (SnapSlot &s : c->snap) {
        if (s.ev) (void)hipEventSynchronize(s.ev);
        if (s.host) (void)hipHostFree(s.host);
        if (s.ev) (void)hipEventDestroy(s.ev);
        s = SnapSlot{};
    }
}

}  // namespace nb

extern "C" {

int nb_snapshot_begin(nb_ctx *c, int32_t slot, int32_t what) {
    NB_TRY(check(c, slot));
    if (!c->uploaded) {
        set_error("no particle data uploaded (call nb_upload first)");
        return NB_ERR_STATE;
    }
    if (what <= 0 || what > (SNAP_POS | SNAP_VEL | SNAP_ACC | SNAP_POT)) {
        set_error("invalid snapshot field mask %d", what);
        return NB_ERR_INVALID;
    }
    SnapSlot &S = c->snap[slot];
    if (S.pending) {
        set_error("snapshot slot %d still holds an unfetched snapshot", slot);
        return NB_ERR_STATE;
    }
    const int64_t n = c->n;
    const size_t need = (size_t)10 * (size_t)(n > 0 ? n : 1);   // pos, vel, acc: 3n; pot: n
    if (S.cap < need) {
        if (S.host) {
            (void)hipEventSynchronize(S.ev);
            NB_HIP(hipHostFree(S.host));
            S.host = nullptr;
            S.cap = 0;
        }
        hipError_t e = hipHostMalloc((void **)&S.host, sizeof(double) * need, hipHostMallocDefault);
        if (e != hipSuccess) {
            S.host = nullptr;
            set_error("hipHostMalloc(%zu bytes) failed: %s", sizeof(double) * need,
                      hipGetErrorString(e));
            return NB_ERR_NOMEM;
        }
        S.cap = need;
    }
    if (!S.ev) NB_HIP(hipEventCreateWithFlags(&S.ev, hipEventDisableTiming));
    hipStream_t s = c->stream;
    const size_t nn = (size_t)n;
    if (n > 0) {
        // One device staging buffer serves all fields: the stream orders each
        // copy-out before the next conversion overwrites it.
        if (what & SNAP_POS) {
            NB_TRY(to_aos(c->x, c->y, c->z, c->perm, n, c->is_double, c->stage, s));
            NB_HIP(hipMemcpyAsync(S.host, c->stage, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, s));
        }
        if (what & SNAP_VEL) {
            NB_TRY(to_aos(c->vx, c->vy, c->vz, c->perm, n, c->is_double, c->stage, s));
            NB_HIP(hipMemcpyAsync(S.host + 3 * nn, c->stage, sizeof(double) * 3 * nn,
                                  hipMemcpyDeviceToHost, s));
        }
        if (what & SNAP_ACC) {
            NB_TRY(to_aos(c->ax, c->ay, c->az, c->perm, n, true, c->stage, s));
            NB_HIP(hipMemcpyAsync(S.host + 6 * nn, c->stage, sizeof(double) * 3 * nn,
                                  hipMemcpyDeviceToHost, s));
        }
        if (what & SNAP_POT) {
            NB_TRY(scalar_out(c->pot, c->perm, n, c->stage, s));
            NB_HIP(hipMemcpyAsync(S.host + 9 * nn, c->stage, sizeof(double) * nn,
                                  hipMemcpyDeviceToHost, s));
        }
    }
    NB_HIP(hipEventRecord(S.ev, s));
    S.what = what;
    S.n = n;
    S.pending = true;
    return NB_OK;
}

int nb_snapshot_fetch(nb_ctx *c, int32_t slot, double *pos, double *vel, double *acc, double *pot) {
    NB_TRY(check(c, slot));
    SnapSlot &S = c->snap[slot];
    if (!S.pending) {
        set_error("snapshot slot %d holds no snapshot (call nb_snapshot_begin first)", slot);
        return NB_ERR_STATE;
    }
    const int want = (pos ? SNAP_POS : 0) | (vel ? SNAP_VEL : 0) | (acc ? SNAP_ACC : 0) |
                     (pot ? SNAP_POT : 0);
    if (want & ~S.what) {
        set_error("snapshot slot %d was taken with field mask %d; mask %d requested", slot, S.what,
                  want);
        return NB_ERR_INVALID;
    }
    NB_HIP(hipEventSynchronize(S.ev));
    const size_t nn = (size_t)S.n;
    if (nn > 0) {
        if (pos) std::memcpy(pos, S.host, sizeof(double) * 3 * nn);
        if (vel) std::memcpy(vel, S.host + 3 * nn, sizeof(double) * 3 * nn);
        if (acc) std::memcpy(acc, S.host + 6 * nn, sizeof(double) * 3 * nn);
        if (pot) std::memcpy(pot, S.host + 9 * nn, sizeof(double) * nn);
    }
    S.pending = false;
    return NB_OK;
}

}  // extern "C"
