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

template <typename V>
__global__ void k_gather(const V *__restrict__ src, const int64_t *__restrict__ idx, int64_t n,
                         V *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dst[i] = src[idx[i]];
}

__global__ void k_compose_perm(const int64_t *__restrict__ old_perm, const int64_t *__restrict__ idx,
                               int64_t n, int64_t *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = old_perm ? old_perm[idx[i]] : idx[i];
}

template <typename V>
static int permute(V *arr, const int64_t *idx, int64_t n, void *spare, hipStream_t s) {
    const unsigned g = (unsigned)((n + 255) / 256);
    k_gather<V><<<g, 256, 0, s>>>(arr, idx, n, (V *)spare);
    NB_TRY(check_launch("k_gather"));
    NB_HIP(hipMemcpyAsync(arr, spare, sizeof(V) * (size_t)n, hipMemcpyDeviceToDevice, s));
    return NB_OK;
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
        NB_TRY(dalloc(&S.spare, sizeof(double) * (size_t)n));
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
    if (c->is_double) {
        NB_TRY(permute((double *)c->x, idx, n, S.spare, s));
        NB_TRY(permute((double *)c->y, idx, n, S.spare, s));
        NB_TRY(permute((double *)c->z, idx, n, S.spare, s));
        NB_TRY(permute((double *)c->vx, idx, n, S.spare, s));
        NB_TRY(permute((double *)c->vy, idx, n, S.spare, s));
        NB_TRY(permute((double *)c->vz, idx, n, S.spare, s));
        NB_TRY(permute((double *)c->m, idx, n, S.spare, s));
    } else {
        NB_TRY(permute((float *)c->x, idx, n, S.spare, s));
        NB_TRY(permute((float *)c->y, idx, n, S.spare, s));
        NB_TRY(permute((float *)c->z, idx, n, S.spare, s));
        NB_TRY(permute((float *)c->vx, idx, n, S.spare, s));
        NB_TRY(permute((float *)c->vy, idx, n, S.spare, s));
        NB_TRY(permute((float *)c->vz, idx, n, S.spare, s));
        NB_TRY(permute((float *)c->m, idx, n, S.spare, s));
    }
    NB_TRY(permute(c->h, idx, n, S.spare, s));
    NB_TRY(permute(c->ax, idx, n, S.spare, s));
    NB_TRY(permute(c->ay, idx, n, S.spare, s));
    NB_TRY(permute(c->az, idx, n, S.spare, s));
    NB_TRY(permute(c->pot, idx, n, S.spare, s));
    NB_TRY(permute(c->coll, idx, n, S.spare, s));
    NB_TRY(permute(c->central, idx, n, S.spare, s));
    if (c->ti_end) {
        NB_TRY(permute(c->ti_end, idx, n, S.spare, s));
        NB_TRY(permute(c->ti_beg, idx, n, S.spare, s));
    }
    // perm_new[i] = perm_old[idx[i]]  (written to the free key buffer half)
    int64_t *np = (int64_t *)S.keys;
    k_compose_perm<<<g, 256, 0, s>>>(c->perm, idx, n, np);
    NB_TRY(check_launch("k_compose_perm"));
    if (!c->perm) NB_TRY(dalloc(&c->perm, sizeof(int64_t) * (size_t)n));
    NB_HIP(hipMemcpyAsync(c->perm, np, sizeof(int64_t) * (size_t)n, hipMemcpyDeviceToDevice, s));
    NB_TRY(ctx_rebuild_central(c));
    S.valid = true;
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
