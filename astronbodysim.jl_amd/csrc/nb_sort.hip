// nb_sort.hip -- Morton (Z-order) reordering of the internal particle order.
//
// Spatially sorted particles make each target block and each source tile
// compact: the per-block FP32 origin then resolves close pairs to the block's
// extent, and most (block, tile) pairs are separated by more than the
// softening length, so they take the unchecked Newtonian path.  The caller's
// order is kept in `perm` (internal -> caller index); downloads scatter back.
// No reference counterpart: the reference evaluates particles in input order.
#include <hip/hip_runtime.h>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include "nb_ctx.h"

namespace nb {

static int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of %s failed: %s", what, hipGetErrorString(e));
        return NB_ERR_HIP;
    }
    return NB_OK;
}

// Global box from the per-tile boxes (one block, fixed order).
__global__ __launch_bounds__(256) void k_global_box(const double *__restrict__ tb, int64_t ntiles,
                                                    double *__restrict__ box) {
    __shared__ double red[6][256];
    double v[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
    for (int64_t t = threadIdx.x; t < ntiles; t += 256) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            v[d] = fmin(v[d], tb[t * 8 + d]);
            v[3 + d] = fmax(v[3 + d], tb[t * 8 + 3 + d]);
        }
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
#pragma unroll
            for (int q = 0; q < 3; ++q) red[q][threadIdx.x] = fmin(red[q][threadIdx.x], red[q][threadIdx.x + off]);
#pragma unroll
            for (int q = 3; q < 6; ++q) red[q][threadIdx.x] = fmax(red[q][threadIdx.x], red[q][threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) box[threadIdx.x] = red[threadIdx.x][0];
}

__device__ __forceinline__ uint64_t spread21(uint64_t v) {
    v &= 0x1fffff;
    v = (v | (v << 32)) & 0x1f00000000ffffULL;
    v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
    v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}

template <typename T>
__global__ void k_morton(const T *__restrict__ x, const T *__restrict__ y, const T *__restrict__ z,
                         int64_t n, const double *__restrict__ box, uint64_t *__restrict__ key,
                         int64_t *__restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double scale = 2097151.0;  // 2^21 - 1
    uint64_t c[3];
    const T *p[3] = {x, y, z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double lo = box[d], ext = box[3 + d] - box[d];
        double u = ext > 0.0 ? ((double)p[d][i] - lo) / ext : 0.0;
        u = fmin(fmax(u, 0.0), 1.0);
        c[d] = (uint64_t)(u * scale);
    }
    key[i] = spread21(c[0]) | (spread21(c[1]) << 1) | (spread21(c[2]) << 2);
    idx[i] = i;
}

__global__ void k_compose_perm(const int64_t *__restrict__ old_perm, const int64_t *__restrict__ idx,
                               int64_t n, int64_t *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = old_perm ? old_perm[idx[i]] : idx[i];
}

// The particle arrays that follow a re-sort, gathered by one permutation in
// one pass (elements of 1, 4 or 8 bytes) into their own areas of the spare
// buffer, then copied back in a second pass: two launches for all of them.
constexpr int MAX_SORTED_ARRAYS = 16;
struct GatherSet {
    void *arr[MAX_SORTED_ARRAYS];
    int64_t off[MAX_SORTED_ARRAYS];   // byte offset of the array's area in the spare buffer
    int size[MAX_SORTED_ARRAYS];      // bytes per element
    int count;
};

__global__ void k_gather_set(GatherSet g, const int64_t *__restrict__ idx, int64_t n,
                             char *__restrict__ spare) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t k = idx[i];
    for (int a = 0; a < g.count; ++a) {
        char *dst = spare + g.off[a];
        if (g.size[a] == 8) ((uint64_t *)dst)[i] = ((const uint64_t *)g.arr[a])[k];
        else if (g.size[a] == 4) ((uint32_t *)dst)[i] = ((const uint32_t *)g.arr[a])[k];
        else ((uint8_t *)dst)[i] = ((const uint8_t *)g.arr[a])[k];
    }
}

__global__ void k_copy_back_set(GatherSet g, int64_t n, const char *__restrict__ spare) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int a = 0; a < g.count; ++a) {
        const char *src = spare + g.off[a];
        if (g.size[a] == 8) ((uint64_t *)g.arr[a])[i] = ((const uint64_t *)src)[i];
        else if (g.size[a] == 4) ((uint32_t *)g.arr[a])[i] = ((const uint32_t *)src)[i];
        else ((uint8_t *)g.arr[a])[i] = ((const uint8_t *)src)[i];
    }
}

// Bytes of spare buffer for the largest set: 16 arrays of 8-byte elements.
static size_t spare_bytes(int64_t n) { return (size_t)MAX_SORTED_ARRAYS * 8 * (size_t)n; }

static void add_array(GatherSet &g, void *arr, int size, int64_t n) {
    const int64_t off = g.count ? g.off[g.count - 1] + ((g.size[g.count - 1] * n + 7) / 8) * 8 : 0;
    g.arr[g.count] = arr;
    g.off[g.count] = off;
    g.size[g.count] = size;
    ++g.count;
}

int morton_reorder(nb_ctx *c) {
    const int64_t n = c->n;
    if (n < 2) return NB_OK;
    hipStream_t s = c->stream;
    SortState &S = c->sort;
    if (!S.keys) {
        NB_TRY(dalloc(&S.keys, sizeof(uint64_t) * 2 * (size_t)n));
        NB_TRY(dalloc(&S.idx, sizeof(int64_t) * 2 * (size_t)n));
        NB_TRY(dalloc(&S.box, sizeof(double) * 8));
        NB_TRY(dalloc(&S.spare, spare_bytes(n)));
        size_t bytes = 0;
        NB_HIP(rocprim::radix_sort_pairs(nullptr, bytes, S.keys, S.keys + n, S.idx, S.idx + n,
                                         (size_t)n, 0, 63, s));
        NB_TRY(dalloc(&S.tmp, bytes));
        S.tmp_bytes = bytes;
    }
    const int64_t ntiles = (n + BS - 1) / BS;
    NB_TRY(launch_tile_boxes(c->x, c->y, c->z, c->is_double, n, c->tile_box, c->sub_box, s));
    k_global_box<<<1, 256, 0, s>>>(c->tile_box, ntiles, S.box);
    NB_TRY(check_launch("k_global_box"));
    const unsigned g = (unsigned)((n + 255) / 256);
    if (c->is_double)
        k_morton<double><<<g, 256, 0, s>>>((const double *)c->x, (const double *)c->y,
                                           (const double *)c->z, n, S.box, S.keys, S.idx);
    else
        k_morton<float><<<g, 256, 0, s>>>((const float *)c->x, (const float *)c->y,
                                          (const float *)c->z, n, S.box, S.keys, S.idx);
    NB_TRY(check_launch("k_morton"));
    size_t bytes = S.tmp_bytes;
    NB_HIP(rocprim::radix_sort_pairs(S.tmp, bytes, S.keys, S.keys + n, S.idx, S.idx + n, (size_t)n,
                                     0, 63, s));
    const int64_t *idx = S.idx + n;  // sorted: internal position i takes old position idx[i]
    GatherSet set{};
    const int es = c->is_double ? 8 : 4;
    for (void *a : {c->x, c->y, c->z, c->vx, c->vy, c->vz, c->m}) add_array(set, a, es, n);
    for (void *a : {(void *)c->h, (void *)c->ax, (void *)c->ay, (void *)c->az, (void *)c->pot}) add_array(set, a, 8, n);
    if (c->coll) add_array(set, c->coll, 1, n);
    if (c->central) add_array(set, c->central, 1, n);
    if (c->ti_end) {
        add_array(set, c->ti_end, 8, n);
        add_array(set, c->ti_beg, 8, n);
    }
    k_gather_set<<<g, 256, 0, s>>>(set, idx, n, (char *)S.spare);
    NB_TRY(check_launch("k_gather_set"));
    k_copy_back_set<<<g, 256, 0, s>>>(set, n, (const char *)S.spare);
    NB_TRY(check_launch("k_copy_back_set"));
    // perm_new[i] = perm_old[idx[i]]  (written to the free key buffer half)
    int64_t *np = (int64_t *)S.keys;
    k_compose_perm<<<g, 256, 0, s>>>(c->perm, idx, n, np);
    NB_TRY(check_launch("k_compose_perm"));
    if (!c->perm) NB_TRY(dalloc(&c->perm, sizeof(int64_t) * (size_t)n));
    NB_HIP(hipMemcpyAsync(c->perm, np, sizeof(int64_t) * (size_t)n, hipMemcpyDeviceToDevice, s));
    NB_TRY(ctx_rebuild_central(c));
    S.valid = true;
    c->wsrc_valid = false;   // the source weights follow the masses' order
    return NB_OK;
}

void sort_free(nb_ctx *c) {
    SortState &S = c->sort;
    dfree(S.tmp);
    dfree(S.keys);
    dfree(S.idx);
    dfree(S.box);
    dfree(S.spare);
    S.tmp_bytes = 0;
    S.valid = false;
}

}  // namespace nb
