/*
 * nbody_mi355x.h -- C-ABI of the MI355X direct-summation gravity engine.
 *
 * Drop-in for the (DirectSum, GPU) methods of AstroNbodySim.jl.  Every entry
 * point below cites the reference method it replaces (paths relative to the
 * AstroNbodySim.jl repository root).  Plain pointers and sizes only; no C++
 * exceptions cross this boundary; every call returns an int status
 * (NB_OK == 0, negative on error) and nb_last_error() describes the last
 * failure of the calling host thread.
 *
 * Host arrays are caller-owned and copied synchronously.  Vectors are passed
 * as packed xyz triples (3*n doubles), the memory layout of the reference's
 * Vector{PVector{Float64}} columns.  Collections are the integers Int(c)
 * that index the 6-entry ForceSofteningTable (1..6).
 *
 * Precision modes (nb_create):
 *   NB_MODE_FP64  : FP64 pair arithmetic and state (correctness mode).
 *   NB_MODE_FP32  : FP32 state (positions, velocities, masses), FP32 pair
 *                   arithmetic; the reference's floattype = Float32 on GPU.
 *   NB_MODE_SPLIT : FP64 state; FP32 pair arithmetic on positions taken
 *                   relative to a per-block FP64 origin; per-tile FP32 sums
 *                   folded into FP64 accumulators; FP64 kick/drift/energy.
 *   NB_MODE_SPLIT_CENTRAL : NB_MODE_SPLIT plus, in FP64, the pair terms whose
 *                   source is a particle marked central (nb_set_central), for
 *                   every target, and the potential of each central particle
 *                   over every source (so energies carry every term of a
 *                   central particle in FP64).
 *
 * One context owns one HIP stream on the device that was current when it was
 * created.  Calls on one context must not overlap; separate contexts are
 * independent.
 */
#ifndef NBODY_MI355X_H
#define NBODY_MI355X_H

#include <stdint.h>

#if defined(__GNUC__)
#define NB_API __attribute__((visibility("default")))
#else
#define NB_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define NB_OK 0
#define NB_ERR_INVALID (-1)   /* bad argument                                  */
#define NB_ERR_HIP (-2)       /* HIP runtime error                             */
#define NB_ERR_NOMEM (-3)     /* device or host allocation failed              */
#define NB_ERR_STATE (-4)     /* call out of order (e.g. no upload yet)        */
#define NB_ERR_NONFINITE (-5) /* NaN/Inf found in state or result              */
#define NB_ERR_TIMESTEP (-6)  /* adaptive timestep below MinStep / bad integer step */

#define NB_MODE_FP64 0
#define NB_MODE_FP32 1
#define NB_MODE_SPLIT 2
#define NB_MODE_SPLIT_CENTRAL 3

typedef struct nb_ctx nb_ctx;
typedef struct nb_batch nb_batch;

typedef struct nb_devinfo {
    char name[128];
    char arch[64];
    int32_t device;
    int32_t compute_units;
    int32_t clock_khz;
    int32_t lds_per_block;
    int32_t warp_size;
    int32_t max_threads_per_block;
    int64_t global_mem_bytes;
    int64_t l2_bytes;
} nb_devinfo;

/* Launch configuration of the all-pairs kernel (chosen by the autotuner). */
typedef struct nb_tuning {
    int32_t targets_per_thread; /* 1, 2 or 4                                   */
    int32_t j_segments;         /* source-range split (>=1; partials reduced in fixed order) */
    int32_t sort_interval;      /* re-sort into Morton order every k force calls (0 = never) */
    int32_t far_path;           /* far-tile arithmetic: 0 FP32 VALU, 12 ops + rsq per pair;
                                   3 weighted VALU, 10 ops + rsq on the farthest tiles; 1 sum
                                   on the matrix cores, 2 also r^2 there; 4 r^2 from
                                   bf16-split products on the matrix cores beside a VALU sum
                                   (also for potentials alone).  1, 2 and 4 need
                                   targets_per_thread 1; the FP32 mode takes 0 or 4, the
                                   SPLIT modes any; forces with potentials take 0 */
} nb_tuning;

/* ---- library / device ------------------------------------------------- */

/* Library version: 10000*major + 100*minor + patch. */
NB_API int nb_version(void);
/* Message of the last failed call on this host thread ("" if none). */
NB_API const char *nb_last_error(void);
/* Device properties; replaces gpuinfo()  src/base/UtilsGPU.jl:3-21. */
NB_API int nb_device_info(nb_devinfo *out);

/* ---- one-system context: Simulation(d; device = GPU()) ---------------- */

/* Allocate device state for n particles; replaces the DirectSum & GPU branch
 * of Simulation  src/base/Config.jl:1072-1089. */
NB_API int nb_create(nb_ctx **out, int64_t n, int32_t precision_mode);
NB_API int nb_destroy(nb_ctx *ctx);
/* Gravitational constant and ForceSofteningTable (6 entries, indexed by
 * collection-1); replaces sim.config.constants.G / GravityConfig
 * src/base/Config.jl:137-146 and set_softlen!  src/tools/SofteningLength.jl:79-93. */
NB_API int nb_set_params(nb_ctx *ctx, double G, const double soft[6]);
/* Copy particle state host->device (pos/vel 3n, mass n, collection n);
 * replaces cu(StructArray(d))  src/base/Config.jl:1078.  vel may be NULL (zero). */
NB_API int nb_upload(nb_ctx *ctx, const double *pos, const double *vel, const double *mass,
              const uint8_t *collection);
/* Copy state device->host in the caller's particle order; any pointer may be
 * NULL.  Replaces get_all_data / get_local_data  src/base/Config.jl:1199-1201,1228-1230. */
NB_API int nb_download(nb_ctx *ctx, double *pos, double *vel, double *acc, double *pot);
/* Set Acc and/or Potential (caller order; NULL skips).  A resumed run needs the
 * accelerations of the last force evaluation: with adaptive steps those of
 * inactive particles were taken at earlier positions and cannot be recomputed
 * (restart state, src/restart.jl:14-19,127-136). */
NB_API int nb_upload_acc(nb_ctx *ctx, const double *acc, const double *pot);

/* Asynchronous snapshot download, the device half of output(sim, gadget2)
 * src/base/Output.jl:36-50 (called from src/run.jl:90,116-131), which copies
 * synchronously.  nb_snapshot_begin enqueues, on the context's stream, the
 * copies of the fields in `what` (1 Pos, 2 Vel, 4 Acc, 8 Potential; caller
 * order) into pinned host slot `slot` (0 or 1) and returns at once.
 * nb_snapshot_fetch waits for that slot's copies and moves them into the
 * caller's arrays (NULL skips a field; asking for a field the slot did not
 * take is an error).  A slot must be fetched before it is begun again.
 * fetch may run on another host thread while the context keeps working, but
 * not at the same time as begin of the same slot or nb_destroy. */
NB_API int nb_snapshot_begin(nb_ctx *ctx, int32_t slot, int32_t what);
NB_API int nb_snapshot_fetch(nb_ctx *ctx, int32_t slot, double *pos, double *vel, double *acc,
                             double *pot);

/* ---- forces and potentials -------------------------------------------- */

/* Acc[i] = sum_j spline pair force (self pair contributes 0), every particle;
 * replaces compute_force(sim, DirectSum, GPU)  src/directsumgpu/gravity.jl:36-45,81-85
 * (kernel :22-34, pair function src/directsum/gravity.jl:8-28). */
NB_API int nb_compute_force(nb_ctx *ctx);
/* Potential[i] = sum_j spline pair potential; replaces
 * compute_potential(sim, DirectSum, CPU)  src/directsum/potential.jl:142-148
 * (no GPU method exists in the reference).  add_self_term != 0 also applies
 * the public compute_potential(sim) self term  src/base/Potential.jl:5-24.
 * The accelerations are left as they were. */
NB_API int nb_compute_potential(nb_ctx *ctx, int32_t add_self_term);
/* Force / potential at npts probe points with a caller-given softening h;
 * replaces compute_force(sim, pos, h, DirectSum, ...)  src/directsum/gravity.jl:171-194
 * and compute_potential(sim, pos, h, ...)  src/directsum/potential.jl:155-180. */
NB_API int nb_force_at_points(nb_ctx *ctx, int64_t npts, const double *pts, double h, double *acc);
NB_API int nb_potential_at_points(nb_ctx *ctx, int64_t npts, const double *pts, double h, double *pot);

/* Allocation-free force evaluation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory -- code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc.  It computes what
 * nb_compute_force computes (the same kernels; FP64 state with FP32 pair
 * arithmetic in NB_MODE_SPLIT, all FP64 in NB_MODE_FP64) on n particles
 * that the caller keeps in DEVICE memory as SoA arrays x, y, z, m, with the
 * per-target softening h[i] (h == NULL: h_const for every particle), and
 * writes ax, ay, az (and pot unless NULL: the unbiased potential) there.
 * Particles are taken in the order given (no Morton sort).  The launches are
 * enqueued on `stream` (a hipStream_t; NULL = the default stream) and the
 * call returns without waiting.  It calls no hipMalloc / hipFree: the tile
 * boxes and per-segment partial sums use one fixed device array of
 * NB_NOALLOC_SCRATCH_BYTES, which every GPU the library runs on holds until
 * the process ends; n above NB_NOALLOC_MAX_N is rejected (NB_ERR_INVALID).
 * All calls on one GPU share that array, so calls that overlap on one GPU
 * (other streams, other host threads) race: serialise them. */
#define NB_NOALLOC_SCRATCH_BYTES (64ll << 20)
#define NB_NOALLOC_MAX_N (1ll << 28)
NB_API int nb_compute_force_noalloc(int32_t precision_mode, int64_t n, const double *x,
                                    const double *y, const double *z, const double *m,
                                    const double *h, double h_const, double G, double *ax,
                                    double *ay, double *az, double *pot, void *stream);

/* ---- constant-timestep leapfrog (KDK) --------------------------------- */

/* Vel += Acc * forward * 0.5 * dt; replaces leapfrog_half_kick (ConstantTimestep)
 * src/directsumgpu/timestep.jl:194-214. */
NB_API int nb_half_kick(nb_ctx *ctx, int32_t forward, double dt);
/* Pos += Vel * dt; replaces drift_particles_kernel  src/base/Timestep.jl:181-191
 * (the caller passes t_new - t_old, the CPU semantics). */
NB_API int nb_drift(nb_ctx *ctx, double dt);
/* Vel += Acc * dt; replaces advance_and_find_timestep (ConstantTimestep)
 * src/directsumgpu/timestep.jl:156-176. */
NB_API int nb_kick(nb_ctx *ctx, double dt);
/* nsteps x step(sim, DirectSum, GPU)  src/directsumgpu/timestep.jl:216-224:
 * drift by (t+dt)-t with t accumulated from t_begin as the reference does
 * (src/base/Timestep.jl:198, :183), compute_force, kick by dt.  With
 * use_graph != 0 the steps are replayed from a captured HIP graph. */
NB_API int nb_kdk_steps(nb_ctx *ctx, int64_t nsteps, double dt, double t_begin, int32_t use_graph);

/* ---- reductions --------------------------------------------------------- */

/* K = sum 1/2 m v.v, W = 1/2 sum m phi (phi recomputed, unbiased, as
 * total_potential(sim, DirectSum, CPU)  src/directsum/potential.jl:150-153),
 * P = sum m v, L = sum m x cross v; FP64, fixed-order reductions.  Replaces
 * total_kinetic / total_potential / total_momentum / total_angular_momentum
 * src/base/Energy.jl:13-100.  Any output pointer may be NULL; W == NULL skips
 * the potential evaluation.  The outputs are always written; when one of the
 * requested ones is NaN or Inf the call returns NB_ERR_NONFINITE (the
 * reference returns such values silently). */
NB_API int nb_energy(nb_ctx *ctx, double *K, double *W, double P[3], double L[3]);

/* ---- hierarchical systems, background fields ---------------------------- */

/* Mark the first `count` particles (caller order) as central bodies: in
 * NB_MODE_SPLIT_CENTRAL their pair terms run in FP64.  No reference
 * counterpart: the reference evaluates every pair in one precision. */
NB_API int nb_set_central(nb_ctx *ctx, int64_t count);
/* Analytic FP64 background field added after the pair sum; replaces the
 * GPU bgforce closures  src/directsumgpu/gravity.jl:47-53,64-79 and
 * apply_background_potential  src/base/PostProcessing/Potential.jl:1-7 with a
 * fixed menu.  kind 0 clears all fields; kind 1 = point mass
 * params = {GM, cx, cy, cz}; kind 2 = Paczynski-Wiita pseudo-Newtonian mass
 * params = {GM, cx, cy, cz, r_g}  (src/tools/BlackHole.jl:48-60); kind 3 =
 * radial polynomial about a centre, |a| = c0 + c1 r + c2 r^2 + c3 r^3 toward
 * it, phi = c0 r + c1 r^2/2 + c2 r^3/3 + c3 r^4/4, params = {c0, cx, cy, cz,
 * c1, c2, c3}; kind 4 = uniform field, params = {gx, gy, gz}, phi = -g.x. */
NB_API int nb_add_background(nb_ctx *ctx, int32_t kind, const double *params);

/* MOND: after the background fields, Acc <- Acc * nu(|Acc|/acc0, nu_index) with
 * nu(y, n) = (0.5 (1 + sqrt(1 + 4 y^-n)))^(1/n); replaces mond_Milgrom1983
 * src/mond/milgrom1983.jl:6-8,33-47 and its GPU kernel
 * src/directsumgpu/gravity.jl:55-62,74-78.  acc0 <= 0 switches it off. */
NB_API int nb_set_mond(nb_ctx *ctx, double acc0, double nu_index);

/* ---- synchronisation, tuning, timing ------------------------------------ */

NB_API int nb_sync(nb_ctx *ctx);
NB_API int nb_set_tuning(nb_ctx *ctx, const nb_tuning *t);
NB_API int nb_get_tuning(nb_ctx *ctx, nb_tuning *t);
/* Time the candidate launch configurations on this context's data and keep
 * the fastest (state is restored afterwards).  No reference counterpart
 * (GPU_NumThreads is a fixed keyword there, src/base/Config.jl:295-305). */
NB_API int nb_autotune(nb_ctx *ctx);
/* Force-kernel timing with HIP events on the context's stream: enable, then
 * read the number of timed force evaluations and their summed duration. */
NB_API int nb_set_timing(nb_ctx *ctx, int32_t enable);
NB_API int nb_get_timing(nb_ctx *ctx, int64_t *count, double *total_ms);
/* Tuning table persistence: entries keyed by (precision mode, ceil(log2 N)),
 * consulted by nb_create / nb_batch_create. */
NB_API int nb_tuner_save(const char *path);
NB_API int nb_tuner_load(const char *path);
NB_API int nb_tuner_clear(void);


/* ---- adaptive (block, power-of-two) timesteps ----------------------------- */
/* The integer timeline of the reference's AdaptiveTimestep: TimeInterval =
 * (End - Begin) / TimeBase; each particle's step is TimeBase / 2^k. */

/* Ti_endstep = get_timestep(Acc, h), Ti_begstep = 0 for every particle (needs
 * a force evaluation first); replaces init_timesteps  src/base/Timestep.jl:29-44,
 * src/directsumgpu/timestep.jl:32-47.  NB_ERR_TIMESTEP when dt < MinStep or an
 * integer step is <= 0 (the reference's error(), src/base/Timestep.jl:9-16). */
NB_API int nb_adaptive_init(nb_ctx *ctx, double time_interval, int64_t time_base,
                            double err_tol, double min_step, double max_step);
/* Vel += Acc * forward * 0.5 * (Ti_endstep - Ti_begstep) * TimeInterval;
 * replaces leapfrog_half_kick (AdaptiveTimestep)  src/base/Timestep.jl:105-114. */
NB_API int nb_adaptive_half_kick(nb_ctx *ctx, int32_t forward);
/* min(TimeBase, min_i Ti_endstep); replaces find_min_endstep
 * src/base/Timestep.jl:135-144, src/directsumgpu/timestep.jl:103-114. */
NB_API int nb_adaptive_min_endstep(nb_ctx *ctx, int64_t *min_endstep);
/* Every particle: Pos += Vel * (to - from) * TimeInterval; replaces
 * drift_particles_kernel (AdaptiveTimestep)  src/base/Timestep.jl:169-179. */
NB_API int nb_adaptive_drift(nb_ctx *ctx, int64_t from_int, int64_t to_int);
/* Accelerations of the active particles (Ti_endstep == t_int) from every
 * source; others keep theirs.  Replaces compute_local_force's active filter
 * src/directsum/gravity.jl:94-98. */
NB_API int nb_adaptive_compute_force(nb_ctx *ctx, int64_t t_int, int64_t *nactive);
/* Active particles: new integer step, kick from the old mid-step to the new
 * one, Ti_begstep = Ti_endstep, Ti_endstep += step; replaces
 * advance_and_find_timestep (AdaptiveTimestep)  src/base/Timestep.jl:47-82,
 * src/directsumgpu/timestep.jl:122-154. */
NB_API int nb_adaptive_kick(nb_ctx *ctx, int64_t t_int);
/* Ti_endstep / Ti_begstep in the caller's order (either may be NULL). */
NB_API int nb_adaptive_get_steps(nb_ctx *ctx, int64_t *endstep, int64_t *begstep);
/* Restore Ti_endstep / Ti_begstep (caller order) after nb_adaptive_init, for a
 * run resumed from a restart file (the reference keeps them inside the
 * serialised Simulation, src/restart.jl:14-19,127-136). */
NB_API int nb_adaptive_set_steps(nb_ctx *ctx, const int64_t *endstep, const int64_t *begstep);

/* ---- batched ensembles: many independent systems in one launch ----------- */
/* No reference counterpart (the reference runs one system per Simulation);
 * each system follows the same per-system semantics as the one-system API. */

typedef struct nb_batch_tuning {
    int32_t targets_per_thread; /* 1, 2 or 4 (targets of a work item = 256 x this)   */
    int32_t blocks;             /* persistent blocks of the ragged kernel (0 = 8 per CU) */
    int32_t steps_per_graph;    /* KDK steps per captured HIP graph (even, >= 2)     */
    int32_t sort_interval;      /* re-sort every system into Morton order every k force calls (0 = never) */
    int32_t resident;           /* 1: multi-step resident kernel when every system fits it: its
                                   massive particles plus a 1/W share of its massless ones in
                                   <= 256 lanes for some W <= 4 (at most 256 particles, or
                                   more with massless test particles) */
    int32_t far_path;           /* as nb_tuning.far_path */
    int32_t resident_waves;     /* wavefronts per system of the resident kernel, 1-4; each holds
                                   every massive particle and a share of the massless ones
                                   (0 = fewest lane slots, ties to fewer waves) */
    int32_t reserved;
} nb_batch_tuning;

/* nsys systems; system s owns particles [offsets[s], offsets[s+1]) of the
 * packed arrays.  precision_mode: NB_MODE_FP64, NB_MODE_SPLIT or
 * NB_MODE_SPLIT_CENTRAL. */
NB_API int nb_batch_create(nb_batch **out, int32_t nsys, const int64_t *offsets,
                           int32_t precision_mode);
NB_API int nb_batch_destroy(nb_batch *b);
/* Per-system G, softening length (all collections) and timestep. */
NB_API int nb_batch_set_params(nb_batch *b, const double *G, const double *soft, const double *dt);
NB_API int nb_batch_upload(nb_batch *b, const double *pos, const double *vel, const double *mass);
NB_API int nb_batch_download(nb_batch *b, double *pos, double *vel, double *acc, double *pot);
/* First `count` particles of every system are central bodies (FP64 terms in
 * NB_MODE_SPLIT_CENTRAL). */
NB_API int nb_batch_set_central(nb_batch *b, int32_t count);
NB_API int nb_batch_compute_force(nb_batch *b);
/* Potential of every particle within its own system; the accelerations are
 * left as they were (nb_batch_energy's W and the tracked-energy steps run the
 * same evaluation). */
NB_API int nb_batch_compute_potential(nb_batch *b);
/* Vel += Acc * forward * 0.5 * dt_s for every system s. */
NB_API int nb_batch_half_kick(nb_batch *b, int32_t forward);
/* nsteps x (drift dt_s, force, kick dt_s) for every system.  use_graph != 0
 * replays captured graphs (or the resident kernel, see nb_batch_tuning). */
NB_API int nb_batch_kdk_steps(nb_batch *b, int64_t nsteps, int32_t use_graph);
/* As nb_batch_kdk_steps, recording every system's total energy (FP64, with
 * velocities synchronised to the step: v - a dt/2) after every
 * `every`-th step into energy[s * (nsteps / every) + k]. */
NB_API int nb_batch_kdk_steps_tracked(nb_batch *b, int64_t nsteps, int32_t every, double *energy);
/* Per-system K, W (= 1/2 sum m phi), P[3], L[3]; arrays of nsys / 3*nsys.
 * The outputs are always written; NB_ERR_NONFINITE names the first system
 * with a NaN or Inf among the requested ones. */
NB_API int nb_batch_energy(nb_batch *b, double *K, double *W, double *P, double *L);
NB_API int nb_batch_sync(nb_batch *b);
NB_API int nb_batch_set_tuning(nb_batch *b, const nb_batch_tuning *t);
NB_API int nb_batch_get_tuning(nb_batch *b, nb_batch_tuning *t);
/* Time candidate launch configurations of the ragged force kernel on this
 * batch and keep the fastest. */
NB_API int nb_batch_autotune(nb_batch *b);
NB_API int nb_batch_set_timing(nb_batch *b, int32_t enable);
NB_API int nb_batch_get_timing(nb_batch *b, int64_t *count, double *total_ms);

#ifdef __cplusplus
}
#endif
#endif /* NBODY_MI355X_H */
