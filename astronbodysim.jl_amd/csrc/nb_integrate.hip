// nb_integrate.hip -- per-particle kernels around the pair sum: layout
// conversion, constant-timestep kick/drift, background fields and the FP64
// energy / momentum reductions.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nb_integrate.h"
#include "nb_internal.h"

namespace nb {

static int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of %s failed: %s", what, hipGetErrorString(e));
        return NB_ERR_HIP;
    }
    return NB_OK;
}

static inline unsigned grid_for(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// ---- layout: packed xyz (caller order) <-> SoA (internal order) ----------
template <typename T>
__global__ void k_from_aos(const double *__restrict__ aos, const int64_t *__restrict__ perm,
                           int64_t n, T *__restrict__ x, T *__restrict__ y, T *__restrict__ z) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t u = perm ? perm[i] : i;
    x[i] = (T)aos[3 * u];
    y[i] = (T)aos[3 * u + 1];
    z[i] = (T)aos[3 * u + 2];
}

template <typename T>
__global__ void k_to_aos(const T *__restrict__ x, const T *__restrict__ y, const T *__restrict__ z,
                         const int64_t *__restrict__ perm, int64_t n, double *__restrict__ aos) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t u = perm ? perm[i] : i;
    aos[3 * u] = (double)x[i];
    aos[3 * u + 1] = (double)y[i];
    aos[3 * u + 2] = (double)z[i];
}

template <typename T>
__global__ void k_from_scalar(const double *__restrict__ s, const int64_t *__restrict__ perm,
                              int64_t n, T *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = (T)s[perm ? perm[i] : i];
}

__global__ void k_to_scalar(const double *__restrict__ v, const int64_t *__restrict__ perm,
                            int64_t n, double *__restrict__ s) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    s[perm ? perm[i] : i] = v[i];
}

// h_i = ForceSofteningTable[collection_i]  (src/directsum/gravity.jl:97)
__global__ void k_fill_h(const uint8_t *__restrict__ coll, int64_t n, SoftTable t,
                         double *__restrict__ h) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    h[i] = t.h[coll[i] - 1];
}

int to_soa(const double *aos, const int64_t *perm, int64_t n, void *x, void *y, void *z,
           bool is_double, hipStream_t s) {
    if (n <= 0) return NB_OK;
    if (is_double)
        k_from_aos<double><<<grid_for(n), 256, 0, s>>>(aos, perm, n, (double *)x, (double *)y,
                                                       (double *)z);
    else
        k_from_aos<float><<<grid_for(n), 256, 0, s>>>(aos, perm, n, (float *)x, (float *)y,
                                                      (float *)z);
    return check_launch("k_from_aos");
}

int to_aos(const void *x, const void *y, const void *z, const int64_t *perm, int64_t n,
           bool is_double, double *aos, hipStream_t s) {
    if (n <= 0) return NB_OK;
    if (is_double)
        k_to_aos<double><<<grid_for(n), 256, 0, s>>>((const double *)x, (const double *)y,
                                                     (const double *)z, perm, n, aos);
    else
        k_to_aos<float><<<grid_for(n), 256, 0, s>>>((const float *)x, (const float *)y,
                                                    (const float *)z, perm, n, aos);
    return check_launch("k_to_aos");
}

int scalar_in(const double *s_, const int64_t *perm, int64_t n, void *out, bool is_double,
              hipStream_t s) {
    if (n <= 0) return NB_OK;
    if (is_double)
        k_from_scalar<double><<<grid_for(n), 256, 0, s>>>(s_, perm, n, (double *)out);
    else
        k_from_scalar<float><<<grid_for(n), 256, 0, s>>>(s_, perm, n, (float *)out);
    return check_launch("k_from_scalar");
}

int scalar_out(const double *v, const int64_t *perm, int64_t n, double *out, hipStream_t s) {
    if (n <= 0) return NB_OK;
    k_to_scalar<<<grid_for(n), 256, 0, s>>>(v, perm, n, out);
    return check_launch("k_to_scalar");
}

int fill_h(const uint8_t *coll, int64_t n, const SoftTable &t, double *h, hipStream_t s) {
    if (n <= 0) return NB_OK;
    k_fill_h<<<grid_for(n), 256, 0, s>>>(coll, n, t, h);
    return check_launch("k_fill_h");
}

// ---- constant-timestep KDK pieces ---------------------------------------
// Vel += Acc * coef   (coef = dt, or forward*0.5*dt for the half kicks)
template <typename T>
__global__ void k_kick(T *__restrict__ vx, T *__restrict__ vy, T *__restrict__ vz,
                       const double *__restrict__ ax, const double *__restrict__ ay,
                       const double *__restrict__ az, int64_t n, double coef) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T c = (T)coef;
    vx[i] += (T)ax[i] * c;
    vy[i] += (T)ay[i] * c;
    vz[i] += (T)az[i] * c;
}

// Pos += Vel * (t_new - t_old) where t_new = t_old + dt is accumulated on the
// device in FP64 exactly as src/base/Timestep.jl:198,183 does on the host.
// time[rd] holds t_old; block 0 publishes t_new to time[wr] (rd != wr).
template <typename T>
__global__ void k_drift(T *__restrict__ x, T *__restrict__ y, T *__restrict__ z,
                        const T *__restrict__ vx, const T *__restrict__ vy,
                        const T *__restrict__ vz, int64_t n, double dt, double *time, int rd,
                        int wr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double t_old = time[rd];
    const double t_new = t_old + dt;
    if (i == 0) time[wr] = t_new;
    if (i >= n) return;
    const T d = (T)(t_new - t_old);
    x[i] += vx[i] * d;
    y[i] += vy[i] * d;
    z[i] += vz[i] * d;
}

__global__ void k_set_time(double *time, int slot, double t) { time[slot] = t; }

int kick(void *vx, void *vy, void *vz, const double *ax, const double *ay, const double *az,
         int64_t n, double coef, bool is_double, hipStream_t s) {
    if (n <= 0) return NB_OK;
    if (is_double)
        k_kick<double><<<grid_for(n), 256, 0, s>>>((double *)vx, (double *)vy, (double *)vz, ax,
                                                   ay, az, n, coef);
    else
        k_kick<float><<<grid_for(n), 256, 0, s>>>((float *)vx, (float *)vy, (float *)vz, ax, ay,
                                                  az, n, coef);
    return check_launch("k_kick");
}

int drift(void *x, void *y, void *z, const void *vx, const void *vy, const void *vz, int64_t n,
          double dt, double *time, int rd, int wr, bool is_double, hipStream_t s) {
    const int64_t nn = n > 0 ? n : 1;
    if (is_double)
        k_drift<double><<<grid_for(nn), 256, 0, s>>>((double *)x, (double *)y, (double *)z,
                                                     (const double *)vx, (const double *)vy,
                                                     (const double *)vz, n, dt, time, rd, wr);
    else
        k_drift<float><<<grid_for(nn), 256, 0, s>>>((float *)x, (float *)y, (float *)z,
                                                    (const float *)vx, (const float *)vy,
                                                    (const float *)vz, n, dt, time, rd, wr);
    return check_launch("k_drift");
}

int set_time(double *time, int slot, double t, hipStream_t s) {
    k_set_time<<<1, 1, 0, s>>>(time, slot, t);
    return check_launch("k_set_time");
}

// ---- analytic background fields (FP64) -----------------------------------
// kind 1: point mass  a += -GM (x-c)/|x-c|^3,  phi += -GM/|x-c|
// kind 2: Paczynski-Wiita  a += -GM (x-c)/(|x-c| (|x-c|-r_g)^2),  phi += -GM/(|x-c|-r_g)
//         (src/tools/BlackHole.jl:48-60)
template <typename T>
__global__ void k_background(const T *__restrict__ x, const T *__restrict__ y,
                             const T *__restrict__ z, int64_t n, BgField f,
                             double *__restrict__ ax, double *__restrict__ ay,
                             double *__restrict__ az, double *__restrict__ pot) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (f.kind == 4) {
        // uniform field g: a = g, phi = -g . x
        if (ax) {
            ax[i] += f.p[0];
            ay[i] += f.p[1];
            az[i] += f.p[2];
        }
        if (pot) pot[i] -= f.p[0] * (double)x[i] + f.p[1] * (double)y[i] + f.p[2] * (double)z[i];
        return;
    }
    const double dx = (double)x[i] - f.p[1], dy = (double)y[i] - f.p[2], dz = (double)z[i] - f.p[3];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r = sqrt(r2);
    const double GM = f.p[0];
    double fac, phi;
    if (f.kind == 1) {
        fac = -GM / (r2 * r);
        phi = -GM / r;
    } else if (f.kind == 2) {
        const double rr = r - f.p[4];
        fac = -GM / (r * rr * rr);
        phi = -GM / rr;
    } else {
        // radial polynomial: |a| = c0 + c1 r + c2 r^2 + c3 r^3 toward the
        // centre, phi = c0 r + c1 r^2/2 + c2 r^3/3 + c3 r^4/4 (zero at r = 0)
        const double c0 = f.p[0], c1 = f.p[4], c2 = f.p[5], c3 = f.p[6];
        const double mag = c0 + r * (c1 + r * (c2 + r * c3));
        fac = r > 0.0 ? -mag / r : 0.0;
        phi = r * (c0 + r * (0.5 * c1 + r * (c2 / 3.0 + r * 0.25 * c3)));
    }
    if (ax) {
        ax[i] += fac * dx;
        ay[i] += fac * dy;
        az[i] += fac * dz;
    }
    if (pot) pot[i] += phi;
}

int background(const void *x, const void *y, const void *z, int64_t n, bool is_double,
               const BgField &f, double *ax, double *ay, double *az, double *pot, hipStream_t s) {
    if (n <= 0) return NB_OK;
    if (is_double)
        k_background<double><<<grid_for(n), 256, 0, s>>>((const double *)x, (const double *)y,
                                                         (const double *)z, n, f, ax, ay, az, pot);
    else
        k_background<float><<<grid_for(n), 256, 0, s>>>((const float *)x, (const float *)y,
                                                        (const float *)z, n, f, ax, ay, az, pot);
    return check_launch("k_background");
}

// ---- MOND (Milgrom 1983) rescaling --------------------------------------------
// Acc <- Acc * nu(|Acc| / ACC0, n), nu(y, n) = (0.5 (1 + sqrt(1 + 4 y^-n)))^(1/n)
// (src/mond/milgrom1983.jl:6-8,33-47).  A zero acceleration stays zero (the
// reference's formula gives 0 * Inf there).
__global__ void k_mond(double *__restrict__ ax, double *__restrict__ ay, double *__restrict__ az,
                       int64_t n, double acc0, double nidx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = sqrt(ax[i] * ax[i] + ay[i] * ay[i] + az[i] * az[i]);
    if (a == 0.0) return;
    const double y = a / acc0;
    const double nu = pow(0.5 * (1.0 + sqrt(1.0 + 4.0 * pow(y, -nidx))), 1.0 / nidx);
    ax[i] *= nu;
    ay[i] *= nu;
    az[i] *= nu;
}

int mond(double *ax, double *ay, double *az, int64_t n, double acc0, double nidx, hipStream_t s) {
    if (n <= 0) return NB_OK;
    k_mond<<<grid_for(n), 256, 0, s>>>(ax, ay, az, n, acc0, nidx);
    return check_launch("k_mond");
}

// ---- self term of the public compute_potential(sim) -----------------------
// Potential[i] += G * m_i / (h_i / 2.8)   (src/base/Potential.jl:5-14)
template <typename T>
__global__ void k_self_term(const T *__restrict__ m, const double *__restrict__ h, int64_t n,
                            double G, double *__restrict__ pot) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pot[i] += G * (double)m[i] / (h[i] / 2.8);
}

int self_term(const void *m, const double *h, int64_t n, double G, bool is_double, double *pot,
              hipStream_t s) {
    if (n <= 0) return NB_OK;
    if (is_double)
        k_self_term<double><<<grid_for(n), 256, 0, s>>>((const double *)m, h, n, G, pot);
    else
        k_self_term<float><<<grid_for(n), 256, 0, s>>>((const float *)m, h, n, G, pot);
    return check_launch("k_self_term");
}

// ---- FP64 energy / momentum reductions -----------------------------------
// Per block: K = sum 1/2 m v.v, Wm = sum m phi, P = sum m v, L = sum m x cross v.
// Block partials are reduced by one block in fixed order: deterministic.
constexpr int RED_BS = 256;
constexpr int RED_VALS = 8;

template <typename T>
__global__ __launch_bounds__(RED_BS) void k_energy_partials(
    const T *__restrict__ x, const T *__restrict__ y, const T *__restrict__ z,
    const T *__restrict__ vx, const T *__restrict__ vy, const T *__restrict__ vz,
    const T *__restrict__ m, const double *__restrict__ pot, int64_t n, double *__restrict__ out) {
    __shared__ double red[RED_VALS][RED_BS];
    double v[RED_VALS] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = (int64_t)blockIdx.x * RED_BS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * RED_BS) {
        const double mi = (double)m[i];
        const double X = (double)x[i], Y = (double)y[i], Z = (double)z[i];
        const double VX = (double)vx[i], VY = (double)vy[i], VZ = (double)vz[i];
        v[0] += 0.5 * mi * (VX * VX + VY * VY + VZ * VZ);
        if (pot) v[1] += pot[i] * mi;
        v[2] += VX * mi;
        v[3] += VY * mi;
        v[4] += VZ * mi;
        v[5] += mi * (Y * VZ - Z * VY);
        v[6] += mi * (Z * VX - X * VZ);
        v[7] += mi * (X * VY - Y * VX);
    }
#pragma unroll
    for (int q = 0; q < RED_VALS; ++q) red[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int off = RED_BS / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
#pragma unroll
            for (int q = 0; q < RED_VALS; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x < RED_VALS) out[(int64_t)blockIdx.x * RED_VALS + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(RED_BS) void k_reduce_partials(const double *__restrict__ part,
                                                            int nblocks, double *__restrict__ out) {
    __shared__ double red[RED_VALS][RED_BS];
    double v[RED_VALS] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nblocks; b += RED_BS) {
#pragma unroll
        for (int q = 0; q < RED_VALS; ++q) v[q] += part[(int64_t)b * RED_VALS + q];
    }
#pragma unroll
    for (int q = 0; q < RED_VALS; ++q) red[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int off = RED_BS / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
#pragma unroll
            for (int q = 0; q < RED_VALS; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x < RED_VALS) out[threadIdx.x] = red[threadIdx.x][0];
}

int energy_reduce(const void *x, const void *y, const void *z, const void *vx, const void *vy,
                  const void *vz, const void *m, const double *pot, int64_t n, bool is_double,
                  double *scratch /* ENERGY_BLOCKS*8 + 8 */, hipStream_t s) {
    const int nb = ENERGY_BLOCKS;
    if (is_double)
        k_energy_partials<double><<<nb, RED_BS, 0, s>>>(
            (const double *)x, (const double *)y, (const double *)z, (const double *)vx,
            (const double *)vy, (const double *)vz, (const double *)m, pot, n, scratch);
    else
        k_energy_partials<float><<<nb, RED_BS, 0, s>>>(
            (const float *)x, (const float *)y, (const float *)z, (const float *)vx,
            (const float *)vy, (const float *)vz, (const float *)m, pot, n, scratch);
    NB_TRY(check_launch("k_energy_partials"));
    k_reduce_partials<<<1, RED_BS, 0, s>>>(scratch, nb, scratch + (int64_t)nb * RED_VALS);
    return check_launch("k_reduce_partials");
}

// ---- non-finite detection --------------------------------------------------
__global__ void k_count_nonfinite(const double *__restrict__ a, int64_t n, unsigned *flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!isfinite(a[i])) atomicOr(flag, 1u);
}

int count_nonfinite(const double *a, int64_t n, unsigned *flag, hipStream_t s) {
    if (n <= 0) return NB_OK;
    k_count_nonfinite<<<grid_for(n), 256, 0, s>>>(a, n, flag);
    return check_launch("k_count_nonfinite");
}

}  // namespace nb
