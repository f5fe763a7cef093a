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

// The factors d(k) are not stored for every plane: the forward sweep
// keeps one checkpoint (the d before the group) every kZsGroup planes and
// the backward sweep recomputes each group's factors from it with the same
// operations, so they are the forward sweep's bit for bit.  ck holds
// nx*ny*zs_groups(n) values (twice that for periodic z, whose second
// right-hand side keeps checkpoints of its own).
constexpr int kZsGroup = 8;
__host__ __device__ constexpr int zs_groups(int n) { return (n + kZsGroup - 1) / kZsGroup; }

// Non-periodic z (src/solver_cpu.f90:117-145, :187-223).
template <typename T>
__global__ void __launch_bounds__(256)
    k_thomas(T *__restrict__ W, double *__restrict__ CK,
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
  double *ck = CK + i + (long long)nx * j;
  double dprev = 0.0, pprev = 0.0;
  for (int l = 0; l < n; ++l) {
    if (l % kZsGroup == 0) ck[(l / kZsGroup) * st] = dprev;
    const double y = (double)p[l * st];
    if (l == 0) {
      const double z = 1.0 / (b[0] + lam);
      dprev = c[0] * z;
      pprev = y * z;
    } else if (l < n - 1) {
      const double bb = b[l] + lam;
      const double z = 1.0 / (bb - a[l] * dprev);
      dprev = c[l] * z;
      pprev = (y - a[l] * pprev) * z;
    } else {
      const double z = (b[n - 1] + lam) - a[n - 1] * dprev;
      // the singular column of an NN z-problem (lambda == 0) is pinned to
      // p(n) = 0, the reference's zero-pivot branch
      // (src/solver_cpu.f90:209-214)
      if (z != 0.0 && !(pin && lam == 0.0))
        pprev = (y - a[n - 1] * pprev) / z;
      else
        pprev = 0.0;
    }
    p[l * st] = (T)pprev;
  }
  // back substitution, group by group from the top
  for (int g = zs_groups(n) - 1; g >= 0; --g) {
    const int k0 = g * kZsGroup;
    double d[kZsGroup];
    double dp = ck[g * st];
#pragma unroll
    for (int l = 0; l < kZsGroup; ++l) {
      const int k = k0 + l;
      if (k < n - 1) {
        if (k == 0) {
          dp = c[0] * (1.0 / (b[0] + lam));
        } else {
          const double bb = b[k] + lam;
          dp = c[k] * (1.0 / (bb - a[k] * dp));
        }
      }
      d[l] = dp;
    }
#pragma unroll
    for (int l = kZsGroup - 1; l >= 0; --l) {
      const int k = k0 + l;
      if (k >= n - 1) continue;
      pprev = (double)p[k * st] - d[l] * pprev;
      p[k * st] = (T)pprev;
    }
  }
}

// Periodic z (src/solver_cpu.f90:147-185): the first m = n-1 rows are
// eliminated for the field (p1) and for r2 = (-a(1), 0, ..., 0, -c(m)) (p2)
// with the same factors; p1(1) and p2(1) come out of the forward sweep as
// sums of g(k) p'(k), g(k) = prod_{k'<k} (-d(k')), so the closing row
// p(n) is formed there, and one backward sweep writes p1 + p2 p(n), with
// p2's forward values recomputed from their own checkpoints.  pin: the
// lambda == 0 column (singular) takes p(n) = 0.
template <typename T>
__global__ void __launch_bounds__(256)
    k_thomas_periodic(T *__restrict__ W, double *__restrict__ CK,
                      const double *__restrict__ lamx,
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
  double *ckd = CK + i + (long long)nx * j;
  double *ckq = ckd + st * zs_groups(n);
  const int m = n - 1;  // order of the two reduced systems
  double dprev = 0.0, p1 = 0.0, q = 0.0, s1 = 0.0, s2 = 0.0, g = 1.0;
  for (int l = 0; l < m; ++l) {
    if (l % kZsGroup == 0) {
      ckd[(l / kZsGroup) * st] = dprev;
      ckq[(l / kZsGroup) * st] = q;
    }
    const double y = (double)p[l * st];
    // r2 = (-a(1), 0, ..., 0, -c(m))
    const double r2 = (l == m - 1) ? -c[m - 1] : ((l == 0) ? -a[0] : 0.0);
    if (l == 0) {
      const double z = 1.0 / (b[0] + lam);
      dprev = c[0] * z;
      p1 = y * z;
      q = r2 * z;
    } else if (l < m - 1) {
      const double bb = b[l] + lam;
      const double z = 1.0 / (bb - a[l] * dprev);
      dprev = c[l] * z;
      p1 = (y - a[l] * p1) * z;
      q = (r2 - a[l] * q) * z;
    } else {
      const double z = (b[m - 1] + lam) - a[m - 1] * dprev;
      if (z != 0.0) {
        p1 = (y - a[m - 1] * p1) / z;
        q = (r2 - a[m - 1] * q) / z;
      } else {
        p1 = 0.0;
        q = 0.0;
      }
    }
    s1 = s1 + g * p1;
    s2 = s2 + g * q;
    g = -(dprev * g);
    p[l * st] = (T)p1;
  }
  // closure with p1(1) = s1, p2(1) = s2, p1(m) = p1, p2(m) = q
  double pn;
  if (pin && lam == 0.0)
    pn = 0.0;
  else
    pn = ((double)p[m * st] - c[n - 1] * s1 - a[n - 1] * p1) /
         ((b[n - 1] + lam) + c[n - 1] * s2 + a[n - 1] * q);
  p[m * st] = (T)pn;
  double p1n = 0.0, p2n = 0.0;
  for (int gr = zs_groups(m) - 1; gr >= 0; --gr) {
    const int k0 = gr * kZsGroup;
    double d[kZsGroup], qq[kZsGroup];
    double dp = ckd[gr * st], qp = ckq[gr * st];
#pragma unroll
    for (int l = 0; l < kZsGroup; ++l) {
      const int k = k0 + l;
      if (k < m) {
        const double r2 = (k == m - 1) ? -c[m - 1] : ((k == 0) ? -a[0] : 0.0);
        if (k == 0) {
          const double z = 1.0 / (b[0] + lam);
          dp = c[0] * z;
          qp = r2 * z;
        } else if (k < m - 1) {
          const double bb = b[k] + lam;
          const double z = 1.0 / (bb - a[k] * dp);
          dp = c[k] * z;
          qp = (r2 - a[k] * qp) * z;
        } else {
          const double z = (b[m - 1] + lam) - a[m - 1] * dp;
          qp = (z != 0.0) ? (r2 - a[m - 1] * qp) / z : 0.0;
        }
      }
      d[l] = dp;
      qq[l] = qp;
    }
#pragma unroll
    for (int l = kZsGroup - 1; l >= 0; --l) {
      const int k = k0 + l;
      if (k >= m) continue;
      const double pk = (double)p[k * st];
      double x1, x2;
      if (k == m - 1) {
        x1 = pk;
        x2 = qq[l];
      } else {
        x1 = pk - d[l] * p1n;
        x2 = qq[l] - d[l] * p2n;
      }
      p1n = x1;
      p2n = x2;
      p[k * st] = (T)(x1 + x2 * pn);
    }
  }
}

template <typename T>
int launch_thomas(T *W, double *CK, const double *lamx, const double *lamy,
                  const double *a, const double *b, const double *c, int nx,
                  int ny, int nz, int pin, hipStream_t s) {
  if (nz < 2) return fail(FLUTAS_EINVAL, "nz must be >= 2");
  dim3 blk(std::min(256, ((nx + 63) / 64) * 64));
  dim3 grid((nx + blk.x - 1) / blk.x, ny);
  hipLaunchKernelGGL(k_thomas<T>, grid, blk, 0, s, W, CK, lamx, lamy, a, b, c,
                     nx, ny, nz, pin);
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

template <typename T>
int launch_thomas_periodic(T *W, double *CK, const double *lamx,
                           const double *lamy, const double *a,
                           const double *b, const double *c, int nx, int ny,
                           int nz, int pin, hipStream_t s) {
  if (nz < 3) return fail(FLUTAS_EINVAL, "periodic z needs nz >= 3");
  dim3 blk(std::min(256, ((nx + 63) / 64) * 64));
  dim3 grid((nx + blk.x - 1) / blk.x, ny);
  hipLaunchKernelGGL(k_thomas_periodic<T>, grid, blk, 0, s, W, CK, lamx,
                     lamy, a, b, c, nx, ny, nz, pin);
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

template int launch_thomas<double>(double *, double *, const double *, const double *, const double *, const double *, const double *, int, int, int, int, hipStream_t);
template int launch_thomas<float>(float *, double *, const double *, const double *, const double *, const double *, const double *, int, int, int, int, hipStream_t);
template int launch_thomas_periodic<double>(double *, double *, const double *, const double *, const double *, const double *, const double *, int, int, int, int, hipStream_t);
template int launch_thomas_periodic<float>(float *, double *, const double *, const double *, const double *, const double *, const double *, int, int, int, int, hipStream_t);

size_t thomas_checkpoint_doubles(int nx, int ny, int nz, bool periodic) {
  return (size_t)nx * ny * zs_groups(nz) * (periodic ? 2 : 1);
}

}  // namespace flutas
