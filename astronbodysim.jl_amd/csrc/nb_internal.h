// nb_internal.h -- host-side internals shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/nbody_mi355x.h"

namespace nb {

// Source tile = one LDS stage = one block of BS threads (NW wavefronts).
constexpr int BS = 256;
constexpr int NW = BS / 64;
// Tile quarters: the SB sources one wave stages.  Their boxes let a wave
// classify a tile quarter by quarter.
constexpr int SB = 64;
constexpr int NQ = BS / SB;
// Doubles per target block of the force kernel's geometry workspace (the
// block origin, then nine per wave).
constexpr int TG_STRIDE = 3 + 9 * (BS / 64) + 1;

void set_error(const char *fmt, ...);
void clear_error();

#define NB_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            ::nb::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                            __LINE__);                                                     \
            return NB_ERR_HIP;                                                             \
        }                                                                                  \
    } while (0)

#define NB_TRY(expr)              \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != NB_OK) return rc_; \
    } while (0)

// Arguments of the FP32-pair all-pairs kernel (modes FP32 / SPLIT).
// Targets and sources may be different arrays (probe points).
template <typename T>
struct ForceArgs {
    const T *tx, *ty, *tz;   // targets (internal order)
    const double *th;        // per-target softening h (nullptr => th_const)
    double th_const;
    int64_t nt;
    const T *sx, *sy, *sz, *sm;  // sources
    int64_t ns;
    const double *tile_box;  // [ntiles][8]: lo xyz, hi xyz, (unused) of each BS-source tile
    const double *sub_box;   // [ntiles][NQ][8]: the same for each tile's quarters, or nullptr
    double *tgeo;            // [nblk_i][TG_STRIDE] workspace for the target blocks' geometry, or nullptr
    int tgeo_ready;          // 1: tgeo holds this call's geometry (written by the step's drift), 2: tbg too
    unsigned *tbg;           // [nblk_i][BS][16] workspace for their bf16 B operands, or nullptr
    uint8_t *nflag;          // [nblk_i * nseg * NW]: waves with near tiles (split passes), or nullptr
    int use_origin;          // 1: per-block FP64 origin (SPLIT), 0: absolute FP32 coordinates
    int nseg;                // source segments (grid = nblk_i * nseg)
    int64_t seg_len;         // sources per segment (multiple of BS)
    int nblk_i;              // target blocks
    double G;
    double *ax, *ay, *az, *pot;  // final outputs (nseg == 1)
    double *part;                // partials [nseg][4][nt] (nseg > 1)
    const uint8_t *skip_src;     // sources excluded from the FP32 sum (central bodies) or nullptr
    int accumulate;              // 1: add G*sum to the outputs instead of storing (nseg == 1)
    const float *sw;             // source weights m^(-2/3) (far paths 3 and 4) or nullptr
};

// Far-tile paths (the far_path tuning value; nb_block.h describes each).
constexpr int FAR_VALU = 0;       // 12 VALU + rsq per pair
constexpr int FAR_MFMA_SUM = 1;   // sum on the matrix cores
constexpr int FAR_MFMA_R2 = 2;    // r^2 and sum on the matrix cores (farthest tiles)
constexpr int FAR_WEIGHTED = 3;   // weighted far tiles: 10 VALU + rsq per pair
constexpr int FAR_MFMA_BF16 = 4;  // bf16-split r^2 on the matrix cores, sum on the VALU

// Passes of the force kernel (force_block): every tile, or the far tiles
// only (the split bf16 matrix-core launch, whose near tiles k_force_near adds).
constexpr int PASS_ALL = 0;
constexpr int PASS_FAR = 1;

// Launchers (nb_force.hip)
int launch_tile_boxes(const void *x, const void *y, const void *z, int is_double, int64_t n,
                      double *box, double *sub_box, hipStream_t s);
// acc = false (with pot) writes the potential alone.
// far_path FAR_MFMA_BF16 (tpt 1, forces or potential alone) runs the FP32
// state relative to a per-block FP64 origin; any other value is far path 0.
int launch_force_f32(const ForceArgs<float> &a, int tpt, bool pot, hipStream_t s, bool acc = true,
                     int far_path = FAR_VALU);
int launch_force_split(const ForceArgs<double> &a, int tpt, bool pot, hipStream_t s,
                       int far_path = 0, bool acc = true);
int launch_force_f64(const ForceArgs<double> &a, bool pot, hipStream_t s);
// w_i = m_i^(-2/3) (0 for massless particles) for the weighted far paths.
int launch_source_weights(const void *m, bool is_double, int64_t n, float *w, hipStream_t s);
inline bool far_path_weighted(int far_path) { return far_path == FAR_WEIGHTED || far_path == FAR_MFMA_BF16; }
int launch_finalize(const double *part, int nseg, int64_t n, double G, double *ax, double *ay,
                    double *az, double *pot, bool with_pot, hipStream_t s,
                    bool with_acc = true);
// The KDK step's drift fused with the tile boxes of the drifted positions, and
// its kick fused with the finalize of the accelerations (same arithmetic as
// the separate kernels).
// What the drift may also store for the force call that follows: the target
// blocks' geometry (tgeo non-null; one target per lane, several segments),
// with the bf16 B operands when tbg is non-null.
struct StepGeometry {
    const double *h;     // target softening lengths
    double *tgeo;
    unsigned *tbg;
    int use_origin;
};
int launch_drift_boxes(void *x, void *y, void *z, const void *vx, const void *vy, const void *vz, int64_t n,
                       double dt, double *time, int rd, int wr, int is_double, double *box, double *sub_box,
                       hipStream_t s, const StepGeometry *geo = nullptr);
int launch_finalize_kick(const double *part, int nseg, int64_t n, double G, double *ax, double *ay, double *az,
                         void *vx, void *vy, void *vz, double coef, int is_double, hipStream_t s);

}  // namespace nb
