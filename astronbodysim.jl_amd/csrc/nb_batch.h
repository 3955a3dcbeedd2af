// nb_batch.h -- the ensemble object behind the opaque nb_batch handle.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>
#include <vector>

#include "nb_ctx.h"

struct nb_batch {
    int dev = 0;
    int mode = NB_MODE_SPLIT;
    hipStream_t stream = nullptr;
    int32_t nsys = 0;
    int64_t n = 0;
    int64_t max_ns = 0;
    std::vector<int64_t> h_off;
    std::vector<int32_t> h_nm;    // massive particles of each system (set by upload)
    std::vector<uint8_t> h_massive;   // per particle in the caller's order: mass != 0
    std::vector<double> h_mmax;   // largest |mass| of each system
    // packed resident layout (k_batch_resident_packed), rebuilt when stale
    int32_t *pk_lane = nullptr;   // [blocks][threads] particle, -1 for an idle lane
    int32_t *pk_meta = nullptr;   // [blocks][threads] (system in block << 16) | first flag | record slot
    int32_t *pk_blk = nullptr;    // [blocks][stride] block table
    double *pk_tiny = nullptr;    // [blocks] largest |mass| (sets the r^2 offsets)
    int pk_nblk = 0, pk_threads = 0, pk_smax = 0, pk_recs = 0;
    bool pk_stale = true;
    // pinned host staging of the tracked energies (a pageable device-to-host
    // copy costs about 12 ms the first time in a process)
    double *h_pinned = nullptr;
    size_t h_pinned_n = 0;
    double *e_dev = nullptr;      // device buffer of the tracked energies
    size_t e_dev_n = 0;
    std::vector<double> hG, hh, hdt;
    int64_t *off = nullptr;       // [nsys+1]
    int32_t *sys_of = nullptr;    // system of each particle
    double *Gs = nullptr, *hs = nullptr, *dts = nullptr;  // per system
    double *x = nullptr, *y = nullptr, *z = nullptr;
    double *vx = nullptr, *vy = nullptr, *vz = nullptr;
    double *m = nullptr, *h = nullptr;
    double *ax = nullptr, *ay = nullptr, *az = nullptr, *pot = nullptr;
    int64_t *perm = nullptr;      // internal -> caller index (nullptr: identity)
    uint8_t *central = nullptr;   // per particle: FP64 source
    int32_t ncentral = 0;
    int64_t *cidx = nullptr;      // [nsys][ncentral] internal indices of the central bodies
    // source tiles: BS-aligned from each system's first particle
    int64_t ntiles = 0;
    int64_t *tile_base = nullptr; // [nsys+1]
    int32_t *tile_sys = nullptr;  // [ntiles]
    double *tile_box = nullptr;   // [ntiles][8]
    double *sub_box = nullptr;    // [ntiles][NQ][8] boxes of the tile quarters
    // work items of the ragged kernel: (system << 32) | target tile, by cost
    int64_t nitems = 0;
    int64_t *items = nullptr;
    int items_tpt = 0;
    unsigned *counter = nullptr;  // [2]: next item, finished blocks
    double *stage = nullptr;      // 3n doubles
    float *wsrc = nullptr;        // source weights m^(-2/3) (weighted far paths)
    double *red = nullptr;        // per-system reductions [nsys][8]
    // Morton order within systems
    void *sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    uint64_t *keys = nullptr;
    int64_t *idx = nullptr;
    double *sysbox = nullptr;     // [nsys][8]
    void *spare = nullptr;
    bool sorted = false;
    int64_t calls_since_sort = 0;
    nb_batch_tuning tun{};
    bool uploaded = false;
    int cu_count = 256;
    hipGraphExec_t gexec = nullptr;
    int64_t g_steps = 0;
    // timing of the ragged force kernel
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    int64_t tcount = 0;
    double tms = 0.0;
};
