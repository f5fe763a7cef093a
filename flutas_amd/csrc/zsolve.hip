// Batched z tridiagonal solves, one lane per (i,j) column, lanes adjacent in
// i so every k-row access is coalesced.
//
// The arithmetic replays the reference's Thomas algorithm operation for
// operation (gaussel / dgtsv_homebrewed, src/solver_cpu.f90:117-223):
// bb = b + lambda, z = 1/(bb - a d), d = c z, p = (p - a p_prev) z, the
// zero-pivot rule on the last row, then back substitution.  FMA contraction
// is disabled so the data-independent pivots are bit-identical to the FP64
// CPU path; only the data recurrences see different (FFT) rounding.  The
// elimination always runs in FP64, whatever the storage type T.
#include "internal.hpp"

namespace flutas {

// Non-periodic z (src/solver_cpu.f90:117-145, :187-223).  D is an
// nx*ny*nz FP64 scratch for the elimination factors d(k).
template <typename T>
__global__ void __launch_bounds__(256)
    k_thomas(T *__restrict__ W, double *__restrict__ D,
             const double *__restrict__ lamx, const double *__restrict__ lamy,
             const double *__restrict__ a, const double *__restrict__ b,
             const double *__restrict__ c, int nx, int ny, int n, int pin) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nx) return;
  const double lam = lamx[i] + lamy[j];
  const long long st = (long long)nx * ny;
  T *p = W + i + (long long)nx * j;
  double *d = D + i + (long long)nx * j;
  double z = 1.0 / (b[0] + lam);
  double dprev = c[0] * z;
  d[0] = dprev;
  double pprev = (double)p[0] * z;
  p[0] = (T)pprev;
  for (int l = 1; l < n - 1; ++l) {
    const double bb = b[l] + lam;
    z = 1.0 / (bb - a[l] * dprev);
    dprev = c[l] * z;
    d[l * st] = dprev;
    pprev = ((double)p[l * st] - a[l] * pprev) * z;
    p[l * st] = (T)pprev;
  }
  z = (b[n - 1] + lam) - a[n - 1] * dprev;
  // the singular column of an NN z-problem (lambda == 0) is pinned to
  // p(n) = 0, the reference's zero-pivot branch (src/solver_cpu.f90:209-214)
  if (z != 0.0 && !(pin && lam == 0.0))
    pprev = ((double)p[(n - 1) * st] - a[n - 1] * pprev) / z;
  else
    pprev = 0.0;
  p[(n - 1) * st] = (T)pprev;
  for (int l = n - 2; l >= 0; --l) {
    pprev = (double)p[l * st] - d[l * st] * pprev;
    p[l * st] = (T)pprev;
  }
}

// Periodic z (src/solver_cpu.f90:147-185): two Thomas solves of order n-1
// sharing the factors d, then the rank-one closure for p(n).  P2 holds the
// second solution.  pin: the lambda == 0 column (singular) takes p(n) = 0.
template <typename T>
__global__ void __launch_bounds__(256)
    k_thomas_periodic(T *__restrict__ W, double *__restrict__ D,
                      double *__restrict__ P2, const double *__restrict__ lamx,
                      const double *__restrict__ lamy,
                      const double *__restrict__ a,
                      const double *__restrict__ b,
                      const double *__restrict__ c, int nx, int ny, int n,
                      int pin) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nx) return;
  const double lam = lamx[i] + lamy[j];
  const long long st = (long long)nx * ny;
  T *p = W + i + (long long)nx * j;
  double *d = D + i + (long long)nx * j;
  double *q = P2 + i + (long long)nx * j;
  const int m = n - 1;  // order of the two reduced systems
  // forward elimination of p1 (in W) and p2 (in P2)
  double z = 1.0 / (b[0] + lam);
  double dprev = c[0] * z;
  d[0] = dprev;
  double p1 = (double)p[0] * z;
  p[0] = (T)p1;
  // p2 starts as (-a(1), 0, ..., 0, -c(n-1))
  double p2 = (m == 1 ? -c[0] : -a[0]) * z;
  q[0] = p2;
  for (int l = 1; l < m - 1; ++l) {
    const double bb = b[l] + lam;
    z = 1.0 / (bb - a[l] * dprev);
    dprev = c[l] * z;
    d[l * st] = dprev;
    p1 = ((double)p[l * st] - a[l] * p1) * z;
    p[l * st] = (T)p1;
    p2 = (0.0 - a[l] * p2) * z;
    q[l * st] = p2;
  }
  z = (b[m - 1] + lam) - a[m - 1] * dprev;
  const double rhs2 = -c[m - 1];
  if (z != 0.0) {
    p1 = ((double)p[(m - 1) * st] - a[m - 1] * p1) / z;
    p2 = (rhs2 - a[m - 1] * p2) / z;
  } else {
    p1 = 0.0;
    p2 = 0.0;
  }
  p[(m - 1) * st] = (T)p1;
  q[(m - 1) * st] = p2;
  for (int l = m - 2; l >= 0; --l) {
    const double dl = d[l * st];
    p1 = (double)p[l * st] - dl * p1;
    p[l * st] = (T)p1;
    p2 = q[l * st] - dl * p2;
    q[l * st] = p2;
  }
  // closure (p1, p2 now hold element 1 of each solution)
  const double p1n1 = (double)p[(m - 1) * st], p2n1 = q[(m - 1) * st];
  double pn;
  if (pin && lam == 0.0)
    pn = 0.0;
  else
    pn = ((double)p[(n - 1) * st] - c[n - 1] * p1 - a[n - 1] * p1n1) /
         ((b[n - 1] + lam) + c[n - 1] * p2 + a[n - 1] * p2n1);
  p[(n - 1) * st] = (T)pn;
  for (int l = 0; l < m; ++l)
    p[l * st] = (T)((double)p[l * st] + q[l * st] * pn);
}

template <typename T>
int launch_thomas(T *W, double *D, const double *lamx, const double *lamy,
                  const double *a, const double *b, const double *c, int nx,
                  int ny, int nz, int pin, hipStream_t s) {
  if (nz < 2) return fail(FLUTAS_EINVAL, "nz must be >= 2");
  dim3 blk(std::min(256, ((nx + 63) / 64) * 64));
  dim3 grid((nx + blk.x - 1) / blk.x, ny);
  hipLaunchKernelGGL(k_thomas<T>, grid, blk, 0, s, W, D, lamx, lamy, a, b, c,
                     nx, ny, nz, pin);
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

template <typename T>
int launch_thomas_periodic(T *W, double *D, double *P2, const double *lamx,
                           const double *lamy, const double *a,
                           const double *b, const double *c, int nx, int ny,
                           int nz, int pin, hipStream_t s) {
  if (nz < 3) return fail(FLUTAS_EINVAL, "periodic z needs nz >= 3");
  dim3 blk(std::min(256, ((nx + 63) / 64) * 64));
  dim3 grid((nx + blk.x - 1) / blk.x, ny);
  hipLaunchKernelGGL(k_thomas_periodic<T>, grid, blk, 0, s, W, D, P2, lamx,
                     lamy, a, b, c, nx, ny, nz, pin);
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

template int launch_thomas<double>(double *, double *, const double *, const double *, const double *, const double *, const double *, int, int, int, int, hipStream_t);
template int launch_thomas<float>(float *, double *, const double *, const double *, const double *, const double *, const double *, int, int, int, int, hipStream_t);
template int launch_thomas_periodic<double>(double *, double *, double *, const double *, const double *, const double *, const double *, const double *, int, int, int, int, hipStream_t);
template int launch_thomas_periodic<float>(float *, double *, double *, const double *, const double *, const double *, const double *, const double *, int, int, int, int, hipStream_t);

}  // namespace flutas
