// nb_integrate.h -- launchers of the per-particle kernels (nb_integrate.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nb {

struct SoftTable {
    double h[6];
};

struct BgField {
    int kind;      // 1 point mass, 2 Paczynski-Wiita, 3 radial polynomial, 4 uniform
    double p[8];   // see nb_add_background
};

constexpr int ENERGY_BLOCKS = 512;

int to_soa(const double *aos, const int64_t *perm, int64_t n, void *x, void *y, void *z,
           bool is_double, hipStream_t s);
int to_aos(const void *x, const void *y, const void *z, const int64_t *perm, int64_t n,
           bool is_double, double *aos, hipStream_t s);
int scalar_in(const double *src, const int64_t *perm, int64_t n, void *out, bool is_double,
              hipStream_t s);
int scalar_out(const double *v, const int64_t *perm, int64_t n, double *out, hipStream_t s);
int fill_h(const uint8_t *coll, int64_t n, const SoftTable &t, double *h, hipStream_t s);
int kick(void *vx, void *vy, void *vz, const double *ax, const double *ay, const double *az,
         int64_t n, double coef, bool is_double, hipStream_t s);
int drift(void *x, void *y, void *z, const void *vx, const void *vy, const void *vz, int64_t n,
          double dt, double *time, int rd, int wr, bool is_double, hipStream_t s);
int set_time(double *time, int slot, double t, hipStream_t s);
int background(const void *x, const void *y, const void *z, int64_t n, bool is_double,
               const BgField &f, double *ax, double *ay, double *az, double *pot, hipStream_t s);
int mond(double *ax, double *ay, double *az, int64_t n, double acc0, double nidx, hipStream_t s);
int self_term(const void *m, const double *h, int64_t n, double G, bool is_double, double *pot,
              hipStream_t s);
int energy_reduce(const void *x, const void *y, const void *z, const void *vx, const void *vy,
                  const void *vz, const void *m, const double *pot, int64_t n, bool is_double,
                  double *scratch, hipStream_t s);
int count_nonfinite(const double *a, int64_t n, unsigned *flag, hipStream_t s);

}  // namespace nb
