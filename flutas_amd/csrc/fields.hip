// Streaming kernels of the projection step around the Poisson solve:
// fillps, pres_tw_src, pres_sp_src, correc, chkdiv, boundp, updt_rhs_b, the
// increment bookkeeping, and the FP64 residual of the refined solver.
// One lane per grid point, lanes adjacent in i (coalesced); every kernel is
// a single HBM pass.  The arithmetic keeps the reference's operation order
// (contraction off) so results match the FP64 CPU path to rounding.
#include "internal.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <vector>

namespace flutas {

// index of (i,j,k) in a halo-h block whose first element is (1-h,1-h,1-h)
__device__ __forceinline__ long long ixh(int i, int j, int k, int h, int nx,
                                         int ny) {
  return (long long)(i - 1 + h) +
         (long long)(nx + 2 * h) *
             ((long long)(j - 1 + h) + (long long)(ny + 2 * h) * (k - 1 + h));
}

// Point kernels: one workgroup per (j, k) row of up to kRowThreads points
// (one lane per point; a 512-point row is one 512-lane workgroup, which
// streams faster than two 256-lane halves).
constexpr int kRowThreads = 1024;
static inline dim3 grid3(int nx, int ny, int nz, int bx) {
  return dim3((nx + bx - 1) / bx, ny, nz);
}

// ---- fillps (src/fillps.f90:42-61, _CONSTANT_COEFFS_POISSON) -------------
template <typename T>
__global__ void __launch_bounds__(kRowThreads)
    k_fillps(int nx, int ny, int nz, int nh_d, int nh_u, T dxi, T dyi,
             const T *__restrict__ dzfi, T dti, T rho0,
             const T *__restrict__ u, const T *__restrict__ v,
             const T *__restrict__ w, T *__restrict__ p) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > nx) return;
  const T dtidxi = dti * dxi, dtidyi = dti * dyi;
  const long long c = ixh(i, j, k, nh_u, nx, ny);
  const T val =
      ((w[c] - w[ixh(i, j, k - 1, nh_u, nx, ny)]) * dti * dzfi[k + nh_d - 1] +
       (v[c] - v[ixh(i, j - 1, k, nh_u, nx, ny)]) * dtidyi) +
      (u[c] - u[c - 1]) * dtidxi;
  p[ixh(i, j, k, 1, nx, ny)] = val * rho0;
}

// ---- pres_tw_src (src/source.f90:272-306) -------------------------------
template <typename T>
__global__ void __launch_bounds__(kRowThreads)
    k_pres_tw_src(int nx, int ny, int nz, T dxi, T dyi, int nh_d,
                  int nh_u, const T *__restrict__ dzci, T rho0i,
                  T f_t12, T f1, T f2,
                  const T *__restrict__ p, const T *__restrict__ pold,
                  const T *__restrict__ rho, T *__restrict__ u,
                  T *__restrict__ v, T *__restrict__ w, int cc) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > nx) return;
  const long long c1 = ixh(i, j, k, 1, nx, ny), cu = ixh(i, j, k, nh_u, nx, ny);
  const long long ip = c1 + 1, jp = c1 + (nx + 2),
                  kp = c1 + (long long)(nx + 2) * (ny + 2);
  const T rc = rho[c1];
  const T rhoxi = T(1.0) / (T(0.5) * (rho[ip] + rc));
  const T rhoyi = T(1.0) / (T(0.5) * (rho[jp] + rc));
  const T rhozi = T(1.0) / (T(0.5) * (rho[kp] + rc));
  const T dzk = dzci[k + nh_d - 1];
  const T pc = p[c1];
  if (cc) {
    const T hc = f1 * pc - f2 * pold[c1];
    u[cu] = u[cu] + f_t12 * ((-((p[ip] - pc) * dxi)) * rho0i -
                             (rhoxi - rho0i) * ((f1 * p[ip] - f2 * pold[ip]) - hc) * dxi);
    v[cu] = v[cu] + f_t12 * ((-((p[jp] - pc) * dyi)) * rho0i -
                             (rhoyi - rho0i) * ((f1 * p[jp] - f2 * pold[jp]) - hc) * dyi);
    w[cu] = w[cu] + f_t12 * ((-((p[kp] - pc) * dzk)) * rho0i -
                             (rhozi - rho0i) * ((f1 * p[kp] - f2 * pold[kp]) - hc) * dzk);
  } else {
    u[cu] = u[cu] + f_t12 * (-((p[ip] - pc) * dxi)) * rhoxi;
    v[cu] = v[cu] + f_t12 * (-((p[jp] - pc) * dyi)) * rhoyi;
    w[cu] = w[cu] + f_t12 * (-((p[kp] - pc) * dzk)) * rhozi;
  }
}

// ---- pres_sp_src (src/source.f90:326-343) -------------------------------
template <typename T>
__global__ void __launch_bounds__(kRowThreads)
    k_pres_sp_src(int nx, int ny, int nz, T f_t12, T dxi, T dyi,
                  int nh_d, int nh_u, const T *__restrict__ dzci,
                  T rho0i, const T *__restrict__ pold,
                  T *__restrict__ u, T *__restrict__ v,
                  T *__restrict__ w) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > nx) return;
  const long long c1 = ixh(i, j, k, 1, nx, ny), cu = ixh(i, j, k, nh_u, nx, ny);
  const T pc = pold[c1];
  u[cu] = u[cu] + f_t12 * (-((pold[c1 + 1] - pc) * dxi)) * rho0i;
  v[cu] = v[cu] + f_t12 * (-((pold[c1 + (nx + 2)] - pc) * dyi)) * rho0i;
  w[cu] = w[cu] + f_t12 * (-((pold[c1 + (long long)(nx + 2) * (ny + 2)] - pc) *
                             dzci[k + nh_d - 1])) * rho0i;
}

// ---- correc (src/correc.f90:120-144) ------------------------------------
template <typename T>
__global__ void __launch_bounds__(kRowThreads)
    k_correc(int nx, int ny, int nz, int nh_d, int nh_u, T factori,
             T factorj, const T *__restrict__ dzci, T dt,
             T rho0i, const T *__restrict__ p, T *__restrict__ u,
             T *__restrict__ v, T *__restrict__ w,
             const T *__restrict__ rho, int cc) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > nx) return;
  const long long c1 = ixh(i, j, k, 1, nx, ny), cu = ixh(i, j, k, nh_u, nx, ny);
  const long long ip = c1 + 1, jp = c1 + (nx + 2),
                  kp = c1 + (long long)(nx + 2) * (ny + 2);
  const T pc = p[c1];
  const T fz = dt * dzci[k + nh_d - 1];
  if (cc) {
    u[cu] = u[cu] - factori * (p[ip] - pc) * rho0i;
    v[cu] = v[cu] - factorj * (p[jp] - pc) * rho0i;
    w[cu] = w[cu] - fz * (p[kp] - pc) * rho0i;
  } else {
    const T rc = rho[c1];
    const T rhox = T(0.5) * (rho[ip] + rc), rhoy = T(0.5) * (rho[jp] + rc),
                 rhoz = T(0.5) * (rho[kp] + rc);
    u[cu] = u[cu] - factori * (p[ip] - pc) / rhox;
    v[cu] = v[cu] - factorj * (p[jp] - pc) / rhoy;
    w[cu] = w[cu] - fz * (p[kp] - pc) / rhoz;
  }
}

// ---- fused prologue: fillps then updt_rhs_b in one pass -----------------
// The face terms are added in the order updt_rhs_b adds them (x side 0,
// x side 1, y, z), so the result equals fillps followed by updt_rhs_b bit
// for bit.  q1..q3 are the 'f' offsets of updt_rhs_b (0 here).
__global__ void __launch_bounds__(kRowThreads)
    k_fillps_rhs(int nx, int ny, int nz, int nh_d, int nh_u, double dxi, double dyi,
                 const double *__restrict__ dzfi, double dti, double rho0,
                 const double *__restrict__ u, const double *__restrict__ v,
                 const double *__restrict__ w, int q1, int q2, int q3,
                 const double *__restrict__ rx, const double *__restrict__ ry,
                 const double *__restrict__ rz, double *__restrict__ p) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > nx) return;
  const double dtidxi = dti * dxi, dtidyi = dti * dyi;
  const long long c = ixh(i, j, k, nh_u, nx, ny);
  double val =
      (((w[c] - w[ixh(i, j, k - 1, nh_u, nx, ny)]) * dti * dzfi[k + nh_d - 1] +
        (v[c] - v[ixh(i, j - 1, k, nh_u, nx, ny)]) * dtidyi) +
       (u[c] - u[c - 1]) * dtidxi) *
      rho0;
  const long long mx = (j - 1) + (long long)ny * (k - 1);
  const long long my = (i - 1) + (long long)nx * (k - 1);
  const long long mz = (i - 1) + (long long)nx * (j - 1);
  if (i == 1) val += rx[mx];
  if (i == nx - q1) val += rx[mx + (long long)ny * nz];
  if (j == 1) val += ry[my];
  if (j == ny - q2) val += ry[my + (long long)nx * nz];
  if (k == 1) val += rz[mz];
  if (k == nz - q3) val += rz[mz + (long long)nx * ny];
  p[ixh(i, j, k, 1, nx, ny)] = val;
}

// ---- fused epilogue: correc, then pnew = pold + phi ---------------------
// phi is the solver's increment and is only read, so the pass has no
// neighbour hazards; pnew may be pold (in place) but not phi.  Same
// arithmetic as correc and the increment update.
//
// BC = true also stands in for the two boundp calls around it
// (main__single_phase.f90:727-747: boundp(phi) before correc, boundp(p)
// after the update): the halo values of phi that correc reads (the upper
// faces i = nx+1, j = ny+1, k = nz+1 at interior tangential positions,
// each written by one face of set_bc) are computed from phi's interior
// with set_bc's rule instead of read, and every halo point of pnew is
// written by the thread that owns its interior source point, as the
// composition of set_bc's face rules in boundp's face order (periodic y,
// periodic z, x, then the other y and z faces), which is the value the
// sequence of face passes leaves there.  Bit for bit the three calls.
struct HaloBc {
  char t[3][2];      // BC type per direction and side
  double val[3][2];  // BC value
  double dr[3][2];   // spacing of the N rule (dl in x and y, dzc(0)/dzc(nz) in z)
  int rank[3];       // position of the direction's faces in boundp's order
};

// set_bc's rule for one face (k_set_bc): the ghost value from the value
// one cell inside.
__device__ __forceinline__ double face_rule(char t, int side, double value, double dr, double v) {
  if (t == 'P') return v;
  double factor = value, sgn = 0.0;
  if (t == 'D') { factor = 2.0 * factor; sgn = -1.0; }
  if (t == 'N') { factor = (side == 0) ? -dr * factor : dr * factor; sgn = 1.0; }
  return factor + sgn * v;
}

template <bool BC>
__global__ void __launch_bounds__(kRowThreads)
    k_correc_update(int nx, int ny, int nz, int nh_d, int nh_u, double factori,
                    double factorj, const double *__restrict__ dzci, double dt,
                    double rho0i, const double *__restrict__ phi,
                    const double *pold, double *pnew, double *__restrict__ u,
                    double *__restrict__ v, double *__restrict__ w,
                    const double *__restrict__ rho, int cc, HaloBc hb) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > nx) return;
  const long long sy = nx + 2, sz = (long long)(nx + 2) * (ny + 2);
  const long long c1 = ixh(i, j, k, 1, nx, ny), cu = ixh(i, j, k, nh_u, nx, ny);
  const long long ip = c1 + 1, jp = c1 + sy, kp = c1 + sz;
  const double pc = phi[c1];
  double phx, phy, phz;
  if constexpr (BC) {
    // upper-face ghosts of phi: periodic copies of the first interior
    // plane, else set_bc's rule on the last one
    phx = (i < nx) ? phi[ip]
          : (hb.t[0][1] == 'P' ? phi[c1 - (nx - 1)]
                               : face_rule(hb.t[0][1], 1, hb.val[0][1], hb.dr[0][1], pc));
    phy = (j < ny) ? phi[jp]
          : (hb.t[1][1] == 'P' ? phi[c1 - sy * (ny - 1)]
                               : face_rule(hb.t[1][1], 1, hb.val[1][1], hb.dr[1][1], pc));
    phz = (k < nz) ? phi[kp]
          : (hb.t[2][1] == 'P' ? phi[c1 - sz * (nz - 1)]
                               : face_rule(hb.t[2][1], 1, hb.val[2][1], hb.dr[2][1], pc));
  } else {
    phx = phi[ip];
    phy = phi[jp];
    phz = phi[kp];
  }
  const double fz = dt * dzci[k + nh_d - 1];
  const double po = pold[c1];
  if (cc) {
    u[cu] = u[cu] - factori * (phx - pc) * rho0i;
    v[cu] = v[cu] - factorj * (phy - pc) * rho0i;
    w[cu] = w[cu] - fz * (phz - pc) * rho0i;
  } else {
    const double rc = rho[c1];
    const double rhox = 0.5 * (rho[ip] + rc), rhoy = 0.5 * (rho[jp] + rc),
                 rhoz = 0.5 * (rho[kp] + rc);
    u[cu] = u[cu] - factori * (phx - pc) / rhox;
    v[cu] = v[cu] - factorj * (phy - pc) / rhoy;
    w[cu] = w[cu] - fz * (phz - pc) / rhoz;
  }
  const double pv = po + pc;
  pnew[c1] = pv;
  if constexpr (BC) {
    // the halo points whose source is this interior point: per direction,
    // P: halo 0 from the last plane, halo n+1 from the first; D, N: halo
    // 0 from the first plane, halo n+1 from the last
    const int c[3] = {i, j, k}, n[3] = {nx, ny, nz};
    const long long st[3] = {1, sy, sz};
    int hc[3], side[3];
    bool has[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      has[d] = false;
      hc[d] = c[d];
      side[d] = 0;
      const bool per = hb.t[d][0] == 'P';
      if (c[d] == 1) { has[d] = true; side[d] = per ? 1 : 0; hc[d] = per ? n[d] + 1 : 0; }
      else if (c[d] == n[d]) { has[d] = true; side[d] = per ? 0 : 1; hc[d] = per ? 0 : n[d] + 1; }
    }
    if (has[0] || has[1] || has[2]) {
      for (int mask = 1; mask < 8; ++mask) {
        bool ok = true;
#pragma unroll
        for (int d = 0; d < 3; ++d)
          if ((mask >> d) & 1) ok = ok && has[d];
        if (!ok) continue;
        // apply the face rules of the masked directions in boundp's order
        double val = pv;
        for (int r = 0; r < 5; ++r) {
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            if (!((mask >> d) & 1) || hb.rank[d] != r) continue;
            const int sd = side[d];
            val = face_rule(hb.t[d][sd], sd, hb.val[d][sd], hb.dr[d][sd], val);
          }
        }
        long long off = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) off += (long long)(((mask >> d) & 1) ? hc[d] : c[d]) * st[d];
        pnew[off] = val;
      }
    }
  }
}

// ---- chkdiv (src/chkdiv.f90:38-60): per-block partial sum and max -------
constexpr int kRedThreads = 256;

template <typename T>
__global__ void __launch_bounds__(kRedThreads)
    k_chkdiv(int nx, int ny, int nz, int nh_d, int nh_u, T dxi, T dyi,
             const T *__restrict__ dzfi, const T *__restrict__ u,
             const T *__restrict__ v, const T *__restrict__ w,
             double *__restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double ssum[kRedThreads], smax[kRedThreads];
  const long long total = (long long)nx * ny * nz;
  double s = 0.0, m = 0.0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % nx) + 1;
    const long long r = t / nx;
    const int j = (int)(r % ny) + 1, k = (int)(r / ny) + 1;
    const long long c = ixh(i, j, k, nh_u, nx, ny);
    // the divergence in the fields' precision, the sums in FP64
    const T div =
        ((w[c] - w[ixh(i, j, k - 1, nh_u, nx, ny)]) * dzfi[k + nh_d - 1] +
         (v[c] - v[ixh(i, j - 1, k, nh_u, nx, ny)]) * dyi) +
        (u[c] - u[c - 1]) * dxi;
    s += (double)div;
    const double ad = fabs((double)div);
    m = (ad > m || ad != ad) ? ad : m;
  }
  ssum[threadIdx.x] = s;
  smax[threadIdx.x] = m;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      ssum[threadIdx.x] += ssum[threadIdx.x + off];
      const double o = smax[threadIdx.x + off];
      if (o > smax[threadIdx.x] || o != o) smax[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = ssum[0];
    part[2 * blockIdx.x + 1] = smax[0];
  }
}

// Final reduction of the nb block partials of k_chkdiv (one workgroup, a
// fixed order, so repeated calls on the same field give the same bits):
// out[0] = divtot, out[1] = divmax (NaN propagating).
__global__ void __launch_bounds__(kRedThreads)
    k_chkdiv_final(const double *__restrict__ part, int nb, double *__restrict__ out) {
#pragma clang fp contract(off)
  __shared__ double ssum[kRedThreads], smax[kRedThreads];
  double s = 0.0, m = 0.0;
  for (int b = threadIdx.x; b < nb; b += kRedThreads) {
    s += part[2 * b];
    const double o = part[2 * b + 1];
    m = (o > m || o != o) ? o : m;
  }
  ssum[threadIdx.x] = s;
  smax[threadIdx.x] = m;
  __syncthreads();
  for (int off = kRedThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      ssum[threadIdx.x] += ssum[threadIdx.x + off];
      const double o = smax[threadIdx.x + off];
      if (o > smax[threadIdx.x] || o != o) smax[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = ssum[0];
    out[1] = smax[0];
  }
}

// Partial count of a chkdiv launch: one per 256 points, at most 2048 (the
// kernel strides over the grid, so any grid fits).
constexpr int kChkdivBlocks = 2048;
static int chkdiv_blocks(long long total) {
  return (int)std::min<long long>(kChkdivBlocks, (total + kRedThreads - 1) / kRedThreads);
}

// Enqueues chkdiv on stream s with the partials and the two results in
// `work` (2 * chkdiv_blocks + 2 doubles, results last) and copies the
// results to the host values; the caller synchronises.
template <typename T>
static int chkdiv_enqueue(int nx, int ny, int nz, int nh_d, int nh_u, T dxi, T dyi,
                          const T *dzfi, const T *u, const T *v,
                          const T *w, double *work, double *res_host, hipStream_t s) {
  const int nb = chkdiv_blocks((long long)nx * ny * nz);
  hipLaunchKernelGGL(k_chkdiv<T>, dim3(nb), dim3(kRedThreads), 0, s, nx, ny, nz, nh_d, nh_u, dxi,
                     dyi, dzfi, u, v, w, work);
  FL_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_chkdiv_final, dim3(1), dim3(kRedThreads), 0, s, work, nb, work + 2 * nb);
  FL_HIP(hipGetLastError());
  FL_HIP(hipMemcpyAsync(res_host, work + 2 * nb, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  return FLUTAS_OK;
}

// ---- scratch of flutas_chkdiv_noalloc -------------------------------------
// The block partials and the two results of one chkdiv live in this
// file-scope device array: each GPU the library runs on holds its own copy
// (32 KiB) until the process ends.  In host code `scratch` is only a
// stand-in for the symbol; kernels and copies get the device address that
// hipGetSymbolAddress returns for the current device, resolved once per
// device on its first call and kept in chkdiv_scratch_ptr.
constexpr int kNoAllocDoubles = 2 * kChkdivBlocks + 2;
static __device__ __align__(256) double scratch[kNoAllocDoubles];
constexpr int kMaxDevices = 64;
static std::atomic<double *> g_scratch_dev[kMaxDevices];
static std::mutex g_scratch_mu;

static int chkdiv_scratch_ptr(double **ptr) {
  int dev = 0;
  FL_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(FLUTAS_EDEVICE, "device index beyond the scratch table");
  double *p = g_scratch_dev[dev].load(std::memory_order_acquire);
  if (!p) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    p = g_scratch_dev[dev].load(std::memory_order_relaxed);
    if (!p) {
      void *addr = nullptr;
      FL_HIP(hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)));
      if (!addr) return fail(FLUTAS_EDEVICE, "no device address for the chkdiv scratch");
      p = static_cast<double *>(addr);
      g_scratch_dev[dev].store(p, std::memory_order_release);
    }
  }
  *ptr = p;
  return FLUTAS_OK;
}

// ---- boundp / set_bc, nh_p = 1 (src/bound.f90:146-646) -------------------
// One face: dir 0/1/2, side 0/1; type 'P' periodic copy, 'D' centred
// Dirichlet p0 = 2v - p1, 'N' p0 = p1 -+ dr v.  Loops over the full
// tangential extent including the halo, as set_bc does.
template <typename T>
__global__ void k_set_bc(int nx, int ny, int nz, char ctype, int side, int dir,
                         T value, T dr, T *__restrict__ p) {
  const int n[3] = {nx, ny, nz};
  const long long st[3] = {1, nx + 2, (long long)(nx + 2) * (ny + 2)};
  const int t1 = (dir == 0) ? 1 : 0, t2 = (dir == 2) ? 1 : 2;
  const int b1 = blockIdx.x * blockDim.x + threadIdx.x, b2 = blockIdx.y;
  if (b1 > n[t1] + 1 || b2 > n[t2] + 1) return;
  const long long base = b1 * st[t1] + b2 * st[t2], s = st[dir];
  const int nd = n[dir];
  if (ctype == 'P') {
    if (side == 0) p[base] = p[base + nd * s];
    else p[base + (nd + 1) * s] = p[base + s];
    return;
  }
  T factor = value, sgn = T(0.0);
  if (ctype == 'D') { factor = T(2.0) * factor; sgn = -T(1.0); }
  if (ctype == 'N') { factor = (side == 0) ? -dr * factor : dr * factor; sgn = T(1.0); }
  if (side == 0) p[base] = factor + sgn * p[base + s];
  else p[base + (nd + 1) * s] = factor + sgn * p[base + nd * s];
}

// ---- bounduvw: set_bc on u, v, w with halo nh, and outflow ---------------
// (src/bound.f90:17-144, :227-646, :649-773)
constexpr int kMaxHalo = 8;
template <typename T>
struct FaceField {
  T *p;
  T dr[kMaxHalo];  // spacing of layer q (Neumann factor)
  T value;
  char ctype;           // 'P', 'D', 'N'; 0: field not on this face
  int centered;         // cell-centred component (else face-normal)
};
template <typename T>
struct FaceArgs {
  FaceField<T> f[3];
  int n[3];
  int nh, dir, side;
};

// One face of set_bc on up to three fields: one lane per tangential
// position, halos included (set_bc's loop bounds), each lane running the
// reference's layer loop in order along its own line, so that a later layer
// overwrites an earlier one exactly where the reference's does (the upper
// side of the face-normal kinds writes positions nd + q and nd + 1 + q).
// Reads are interior positions of the line only, so lanes are independent.
template <typename T>
__global__ void k_set_bc_face(FaceArgs<T> a) {
#pragma clang fp contract(off)
  const int d = a.dir, t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2;
  const int nh = a.nh, nd = a.n[d];
  const long long ext[3] = {a.n[0] + 2LL * nh, a.n[1] + 2LL * nh, a.n[2] + 2LL * nh};
  const long long st[3] = {1, ext[0], ext[0] * ext[1]};
  const long long b1 = blockIdx.x * (long long)blockDim.x + threadIdx.x, b2 = blockIdx.y;
  if (b1 >= ext[t1] || b2 >= ext[t2]) return;
  const long long base = b1 * st[t1] + b2 * st[t2], s = st[d];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const char t = a.f[c].ctype;
    if (!t) continue;
    T *p = a.f[c].p;
    auto L = [&](int pos) -> T & { return p[base + (long long)(pos - 1 + nh) * s]; };
    const bool cen = a.f[c].centered != 0;
    T sgn = T(0.0);
    if (t == 'D' && cen) sgn = -T(1.0);
    if (t == 'N') sgn = +T(1.0);
    for (int q = 0; q < nh; ++q) {
      T f = a.f[c].value;
      if (t == 'D' && cen) f = +T(2.0) * f;
      if (t == 'N') f = (a.side == 0) ? -a.f[c].dr[q] * f : a.f[c].dr[q] * f;
      if (t == 'P') {
        L(0 - q) = L(nd - q);
        L(nd + 1 + q) = L(1 + q);
      } else if (cen) {
        if (a.side == 0) L(0 - q) = f + sgn * L(1 + q);
        else L(nd + 1 + q) = f + sgn * L(nd - q);
      } else if (t == 'D') {
        if (a.side == 0) {
          L(0 - q) = f;
        } else {
          L(nd + q) = f;
          L(nd + 1 + q) = L(nd - 1 - q);
        }
      } else {
        if (a.side == 0) {
          L(0 - q) = T(1.0) * f + L(1 + q);
        } else {
          L(nd + q) = T(1.0) * f + L(nd - 1 - q);
          L(nd + 1 + q) = T(2.0) * f + L(nd - 1 - q);
        }
      }
    }
  }
}

// outflow (src/bound.f90:649-773): the velocity normal to an outflow face
// from zero divergence, layers q = 0..nh on the upper side and 0..nh-1 on
// the lower side; one lane per tangential interior position.  dzf: device
// (1-nh_d : nz+nh_d) for the x and y faces; dzq: dzf at the z-face layers
// (upper: dzf(nz+q); lower: dzf(-q)).
template <typename T>
struct OutflowArgs {
  T *u, *v, *w;
  const T *dzf;
  T dzq[kMaxHalo + 1];
  T dx, dy, dxi, dyi;
  int n[3];
  int nh, nh_d, dir, side;
};
template <typename T>
__global__ void k_outflow(OutflowArgs<T> a) {
#pragma clang fp contract(off)
  const int nx = a.n[0], ny = a.n[1], nz = a.n[2], nh = a.nh;
  const long long sx = nx + 2LL * nh, sy = sx * (ny + 2LL * nh);
  auto F = [&](T *f, int i, int j, int k) -> T & {
    return f[(i - 1 + nh) + sx * (j - 1 + nh) + sy * (k - 1 + nh)];
  };
  const int t1 = blockIdx.x * blockDim.x + threadIdx.x + 1, t2 = blockIdx.y + 1;
  const int qmin = nh - 1;
  T *u = a.u, *v = a.v, *w = a.w;
  if (a.dir == 0) {
    const int j = t1, k = t2;
    if (j > ny || k > nz) return;
    const T dzfi = T(1.0) / a.dzf[k + a.nh_d - 1];
    if (a.side == 1) {
      const int i = nx;
      for (int q = 0; q <= qmin + 1; ++q)
        F(u, i + q, j, k) = F(u, i - 1 - q, j, k) -
            a.dx * ((F(v, i + q, j, k) - F(v, i + q, j - 1, k)) * a.dyi +
                    (F(w, i + q, j, k) - F(w, i + q, j, k - 1)) * dzfi);
    } else {
      const int i = 0;
      for (int q = 0; q <= qmin; ++q)
        F(u, i - q, j, k) = F(u, i + 1 + q, j, k) +
            a.dx * ((F(v, i + 1 + q, j, k) - F(v, i + 1 + q, j - 1, k)) * a.dyi +
                    (F(w, i + 1 + q, j, k) - F(w, i + 1 + q, j, k - 1)) * dzfi);
    }
  } else if (a.dir == 1) {
    const int i = t1, k = t2;
    if (i > nx || k > nz) return;
    const T dzfi = T(1.0) / a.dzf[k + a.nh_d - 1];
    if (a.side == 1) {
      const int j = ny;
      for (int q = 0; q <= qmin + 1; ++q)
        F(v, i, j + q, k) = F(v, i, j - 1 - q, k) -
            a.dy * ((F(u, i, j + q, k) - F(u, i - 1, j + q, k)) * a.dxi +
                    (F(w, i, j + q, k) - F(w, i, j + q, k - 1)) * dzfi);
    } else {
      const int j = 0;
      for (int q = 0; q <= qmin; ++q)
        F(v, i, j - q, k) = F(v, i, j + 1 + q, k) +
            a.dy * ((F(u, i, j + 1 + q, k) - F(u, i - 1, j + 1 + q, k)) * a.dxi +
                    (F(w, i, j + 1 + q, k) - F(w, i, j + 1 + q, k - 1)) * dzfi);
    }
  } else {
    const int i = t1, j = t2;
    if (i > nx || j > ny) return;
    if (a.side == 1) {
      const int k = nz;
      for (int q = 0; q <= qmin + 1; ++q)
        F(w, i, j, k + q) = F(w, i, j, k - 1 - q) -
            a.dzq[q] * ((F(u, i, j, k + q) - F(u, i - 1, j, k + q)) * a.dxi +
                        (F(v, i, j, k + q) - F(v, i, j - 1, k + q)) * a.dyi);
    } else {
      const int k = 0;
      for (int q = 0; q <= qmin; ++q)
        F(w, i, j, k - q) = F(w, i, j, k + 1 + q) +
            a.dzq[q] * ((F(u, i, j, k + 1 + q) - F(u, i - 1, j, k + 1 + q)) * a.dxi +
                        (F(v, i, j, k + 1 + q) - F(v, i, j - 1, k + 1 + q)) * a.dyi);
    }
  }
}

// ---- updt_rhs_b (src/bound.f90:829-944) ------------------------------------
template <typename T>
__global__ void k_add_face(int nx, int ny, int nz, int dir, int q,
                           const T *__restrict__ rhs,
                           T *__restrict__ p) {
  const int n[3] = {nx, ny, nz};
  const long long st[3] = {1, nx + 2, (long long)(nx + 2) * (ny + 2)};
  const int t1 = (dir == 0) ? 1 : 0, t2 = (dir == 2) ? 1 : 2;
  const int a = blockIdx.x * blockDim.x + threadIdx.x, bb = blockIdx.y;
  if (a >= n[t1] || bb >= n[t2]) return;
  const long long base = (a + 1) * st[t1] + (bb + 1) * st[t2];
  const long long m = a + (long long)n[t1] * bb, plane = (long long)n[t1] * n[t2];
  p[base + 1 * st[dir]] += rhs[m];
  p[base + (n[dir] - q) * st[dir]] += rhs[m + plane];
}

// ---- interior copies ------------------------------------------------------
__global__ void __launch_bounds__(kRowThreads)
    k_interior(int nx, int ny, int nz, const double *__restrict__ src,
               double *__restrict__ dst, int add) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > nx) return;
  const long long c = ixh(i, j, k, 1, nx, ny);
  dst[c] = add ? src[c] + dst[c] : src[c];
}

// ---- residual r = f - L p (refined solver) ------------------------------
// L is the operator the spectral solve inverts: second differences in x, y
// with homogeneous ghost rules (P wrap, N p0 = p1, D p0 = -p1) and the
// tridiagonal rows a, b, c in z (periodic wrap for PP).
// One thread walks one (i, j) column through kResidualPlanes planes,
// keeping the z-neighbours in registers and loading plane k+1 while plane k
// is evaluated; one partial max per workgroup.
//
// The grid is one-dimensional and its workgroups are renumbered so that
// consecutive rows j of one plane chunk run on the same XCD (the hardware
// deals workgroups to the 8 XCDs round robin): the rows j-1 and j+1 that a
// row reads as neighbours are then the rows its XCD's L2 is already
// streaming.
constexpr int kResidualPlanes = 32;
constexpr int kXcds = 8;

struct BlockIdx3 {
  int ib, j, kc;
};
__device__ __forceinline__ BlockIdx3 residual_block(int nbx, int ny) {
  const unsigned w = blockIdx.x, nb = gridDim.x;
  const unsigned per = nb / kXcds;
  const unsigned L = (w < per * kXcds) ? (w % kXcds) * per + w / kXcds : w;
  BlockIdx3 b;
  b.ib = (int)(L % (unsigned)nbx);
  const unsigned r = L / (unsigned)nbx;
  b.j = (int)(r % (unsigned)ny);
  b.kc = (int)(r / (unsigned)ny);
  return b;
}

// The values of one plane that the residual of cell (i, j, k) reads.
struct Plane7 {
  double f, c, xm, xp, ym, yp;
};

// STORE: r receives the residual (held in registers through the walk);
// without it only the maximum is formed, with no registers held.
template <typename Tr, bool STORE>
__global__ void __launch_bounds__(256)
    k_residual(OperatorDesc op, const double *__restrict__ f,
               const double *__restrict__ p, Tr *__restrict__ r,
               double *__restrict__ part, int nbx) {
#pragma clang fp contract(off)
  __shared__ double smax[256];
  const int nx = op.nx, ny = op.ny, nz = op.nz;
  const BlockIdx3 bi = residual_block(nbx, ny);
  const int i = bi.ib * blockDim.x + threadIdx.x;
  const int j = bi.j;
  const int k0 = bi.kc * kResidualPlanes;
  const int k1 = min(nz, k0 + kResidualPlanes);
  const bool zper = op.cbc[4] == 'P';
  double am = 0.0;
  if (i < nx) {
    const long long sy = nx + 2, sz = (long long)(nx + 2) * (ny + 2);
    const long long c0 = (i + 1) + sy * (j + 1);  // plane-0 offset minus sz
    // x and y neighbours with the homogeneous ghost rules
    auto gather = [&](int k) {
      const long long c = c0 + sz * (k + 1);
      Plane7 v;
      // unconditional loads at selected addresses (a non-periodic boundary
      // neighbour re-reads the centre; ghost() replaces it)
      const long long ixm = (i > 0) ? c - 1 : ((op.cbc[0] == 'P') ? c + (nx - 1) : c);
      const long long ixp = (i < nx - 1) ? c + 1 : ((op.cbc[1] == 'P') ? c - (nx - 1) : c);
      const long long iym = (j > 0) ? c - sy : ((op.cbc[2] == 'P') ? c + sy * (ny - 1) : c);
      const long long iyp = (j < ny - 1) ? c + sy : ((op.cbc[3] == 'P') ? c - sy * (ny - 1) : c);
      v.f = f[c];
      v.c = p[c];
      v.xm = p[ixm];
      v.xp = p[ixp];
      v.ym = p[iym];
      v.yp = p[iyp];
      return v;
    };
    auto ghost = [&](double nb, double pc, char bc, bool inside) {
      if (inside || bc == 'P') return nb;
      return bc == 'N' ? pc : -pc;
    };
    // plane index of the walk's neighbour after plane k (periodic wrap; a
    // non-periodic last plane re-reads itself, unused)
    auto knext = [&](int k) { return (k + 1 < nz) ? k + 1 : (zper ? k + 1 - nz : nz - 1); };
    double pzm = (k0 > 0) ? p[c0 + sz * k0] : (zper ? p[c0 + sz * nz] : 0.0);
    // two planes in flight: plane k+1 is loaded while plane k is evaluated
    // and plane k+2 is requested; plane k's residual is stored after plane
    // k+2's loads are issued (the shared load/store counter completes in
    // issue order, so no later wait for loads waits for it)
    Plane7 cur = gather(k0);
    Plane7 nxt = gather(knext(k0));
#pragma unroll
    for (int q = 0; q < kResidualPlanes; ++q) {
      const int k = k0 + q;
      if (k >= k1) continue;
      Plane7 nxt2 = nxt;
      if (k + 1 < k1) nxt2 = gather(knext(k + 1));
      const double pc = cur.c;
      const double lx = (ghost(cur.xp, pc, op.cbc[1], i < nx - 1) - 2.0 * pc +
                         ghost(cur.xm, pc, op.cbc[0], i > 0)) * op.dx2;
      const double ly = (ghost(cur.yp, pc, op.cbc[3], j < ny - 1) - 2.0 * pc +
                         ghost(cur.ym, pc, op.cbc[2], j > 0)) * op.dy2;
      const double pzp = (k < nz - 1 || zper) ? nxt.c : 0.0;
      double lz = op.b[k] * pc;
      if (k > 0 || zper) lz += op.a[k] * pzm;
      if (k < nz - 1 || zper) lz += op.c[k] * pzp;
      const double res = cur.f - ((lx + ly) + lz);
      // the store follows the loads of plane k+2 in issue order, so the
      // waits for later planes' loads do not wait for it
      if constexpr (STORE) r[i + (long long)nx * (j + (long long)ny * k)] = (Tr)res;
      const double ar = fabs(res);
      if (ar > am || res != res) am = (res != res) ? res : ar;
      pzm = pc;
      cur = nxt;
      nxt = nxt2;
    }
  }
  smax[threadIdx.x] = am;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      const double o = smax[threadIdx.x + off];
      if (o > smax[threadIdx.x] || o != o) smax[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = smax[0];
}

__global__ void __launch_bounds__(256)
    k_absmax(const double *__restrict__ f, int nx, int ny, int nz,
             double *__restrict__ part) {
  __shared__ double smax[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int k0 = blockIdx.z * kResidualPlanes;
  const int k1 = min(nz, k0 + kResidualPlanes);
  double am = 0.0;
  if (i < nx) {
    for (int k = k0; k < k1; ++k) {
      const double v = f[(i + 1) + (long long)(nx + 2) * ((j + 1) + (long long)(ny + 2) * (k + 1))];
      const double av = fabs(v);
      if (av > am || v != v) am = (v != v) ? v : av;
    }
  }
  smax[threadIdx.x] = am;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      const double o = smax[threadIdx.x + off];
      if (o > smax[threadIdx.x] || o != o) smax[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    part[blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z)] = smax[0];
}

// flag = seq where the interior of the halo-1 field f holds a NaN or an
// infinity (the five-pass line pipeline, whose x transform has no check of
// its own)
__global__ void __launch_bounds__(256)
    k_flag_nonfinite(const double *__restrict__ f, int nx, int ny, int *flag, int seq) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= nx) return;
  const double v = f[(i + 1) + (long long)(nx + 2) * ((j + 1) + (long long)(ny + 2) * (k + 1))];
  if (!(fabs(v) <= 1.7976931348623157e308)) *flag = seq;
}

__global__ void __launch_bounds__(256)
    k_accumulate(const float *__restrict__ r, double *__restrict__ p, int nx,
                 int ny) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= nx) return;
  const long long c = (i + 1) + (long long)(nx + 2) * ((j + 1) + (long long)(ny + 2) * (k + 1));
  p[c] = p[c] + (double)r[i + (long long)nx * (j + (long long)ny * k)];
}

int launch_accumulate(const float *r, double *p, int nx, int ny, int nz,
                      hipStream_t s) {
  const int bx = 256;
  hipLaunchKernelGGL(k_accumulate, dim3((nx + bx - 1) / bx, ny, nz), dim3(bx), 0, s, r, p, nx, ny);
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

// max over n partial maxima (NaN propagates), one workgroup
__global__ void __launch_bounds__(1024)
    k_reduce_max(const double *__restrict__ part, int n, double *__restrict__ out) {
  __shared__ double wmax[16];
  double m[4] = {0.0, 0.0, 0.0, 0.0};
  auto take = [](double &acc, double o) {
    if (o > acc || o != o) acc = o;
  };
  // four independent loads in flight per step
  int t = threadIdx.x;
  for (; t + 3 * 1024 < n; t += 4 * 1024) {
    const double o0 = part[t], o1 = part[t + 1024], o2 = part[t + 2048], o3 = part[t + 3072];
    take(m[0], o0);
    take(m[1], o1);
    take(m[2], o2);
    take(m[3], o3);
  }
  for (; t < n; t += 1024) take(m[0], part[t]);
  take(m[0], m[1]);
  take(m[2], m[3]);
  take(m[0], m[2]);
  double am = m[0];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) take(am, __shfl_xor(am, off));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = am;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = wmax[0];
    for (int w = 1; w < 16; ++w) take(r, wmax[w]);
    out[0] = r;
  }
}

int launch_reduce_max(const double *part, int n, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, s, part, n, out);
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

static int xblock(int nx) { return std::min(kRowThreads, ((nx + 63) / 64) * 64); }

template <typename Tr>
int launch_residual(const OperatorDesc &op, const double *f, const double *p,
                    Tr *r, double *partial, int *nparts, hipStream_t s) {
  const int bx = 256;
  const int nbx = (op.nx + bx - 1) / bx;
  const long long nb = (long long)nbx * op.ny * ((op.nz + kResidualPlanes - 1) / kResidualPlanes);
  if (nb > 0x7fffffffLL) return fail(FLUTAS_EINVAL, "residual grid too large");
  if (r)
    hipLaunchKernelGGL((k_residual<Tr, true>), dim3((unsigned)nb), dim3(bx), 0, s, op, f, p, r,
                       partial, nbx);
  else
    hipLaunchKernelGGL((k_residual<Tr, false>), dim3((unsigned)nb), dim3(bx), 0, s, op, f, p, r,
                       partial, nbx);
  FL_HIP(hipGetLastError());
  *nparts = (int)nb;
  return FLUTAS_OK;
}
template int launch_residual<float>(const OperatorDesc &, const double *, const double *, float *, double *, int *, hipStream_t);
template int launch_residual<double>(const OperatorDesc &, const double *, const double *, double *, double *, int *, hipStream_t);

size_t residual_blocks(int nx, int ny, int nz) {
  return (size_t)((nx + 255) / 256) * ny * ((nz + kResidualPlanes - 1) / kResidualPlanes);
}

int launch_absmax(const double *f, int nx, int ny, int nz, double *partial,
                  int *nparts, hipStream_t s) {
  const int bx = 256;
  dim3 grid((nx + bx - 1) / bx, ny, (nz + kResidualPlanes - 1) / kResidualPlanes);
  hipLaunchKernelGGL(k_absmax, grid, dim3(bx), 0, s, f, nx, ny, nz, partial);
  FL_HIP(hipGetLastError());
  *nparts = (int)(grid.x * grid.y * grid.z);
  return FLUTAS_OK;
}

int launch_flag_nonfinite(const double *f, int nx, int ny, int nz, int *flag, int seq,
                          hipStream_t s) {
  const int bx = 256;
  hipLaunchKernelGGL(k_flag_nonfinite, dim3((nx + bx - 1) / bx, ny, nz), dim3(bx), 0, s, f, nx,
                     ny, flag, seq);
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

}  // namespace flutas

using namespace flutas;

static int check_dims(int nx, int ny, int nz) {
  if (nx < 1 || ny < 1 || nz < 1) return fail(FLUTAS_EINVAL, "grid sizes must be positive");
  return FLUTAS_OK;
}

// ---- host sides of the field entry points, for FP64 and FP32 fields ----
template <typename T>
static int chkdiv_impl(int nx, int ny, int nz, T dxi, T dyi, T dzi, int nh_d, int nh_u,
                       const T *dzfi, const T *u, const T *v, const T *w, double res[2]) {
  (void)dzi;
  FL_TRY(check_dims(nx, ny, nz));
  if (!dzfi || !u || !v || !w || !res) return fail(FLUTAS_EINVAL, "null pointer");
  hipStream_t s = current_stream();
  const int nb = chkdiv_blocks((long long)nx * ny * nz);
  double *work = nullptr;
  FL_HIP(hipMallocAsync((void **)&work, sizeof(double) * (2 * nb + 2), s));
  res[0] = res[1] = 0.0;
  const int rc = chkdiv_enqueue<T>(nx, ny, nz, nh_d, nh_u, dxi, dyi, dzfi, u, v, w, work, res, s);
  FL_HIP(hipFreeAsync(work, s));
  FL_HIP(hipStreamSynchronize(s));
  return rc;
}
template <typename T>
static int fillps_impl(int nx, int ny, int nz, int nh_d, int nh_u, T dxi,
                  T dyi, T dzi, const T *dzfi, T dti,
                  T rho0, const T *u, const T *v,
                  const T *w, T *p) {
  (void)dzi;
  FL_TRY(check_dims(nx, ny, nz));
  if (nh_u < 1 || nh_d < 1) return fail(FLUTAS_EINVAL, "halo widths must be >= 1");
  if (!dzfi || !u || !v || !w || !p) return fail(FLUTAS_EINVAL, "null pointer");
  const int bx = xblock(nx);
  hipLaunchKernelGGL(k_fillps<T>, grid3(nx, ny, nz, bx), dim3(bx), 0,
                     current_stream(), nx, ny, nz, nh_d, nh_u, dxi, dyi, dzfi,
                     dti, rho0, u, v, w, p);
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}
template <typename T>
static int pres_tw_src_impl(int nx, int ny, int nz, T dxi, T dyi,
                       T dzi, int nh_d, int nh_u, const T *dzci,
                       T rho0i, T f_t12, T f_t12_o,
                       const T *p, const T *pold, const T *rho,
                       T *u, T *v, T *w, int const_coeff) {
  (void)dzi;
  FL_TRY(check_dims(nx, ny, nz));
  if (nh_u < 1 || nh_d < 1) return fail(FLUTAS_EINVAL, "halo widths must be >= 1");
  if (!dzci || !p || !pold || !rho || !u || !v || !w) return fail(FLUTAS_EINVAL, "null pointer");
  const T f1 = T(1.0) + (f_t12 / f_t12_o), f2 = (f_t12 / f_t12_o);
  const int bx = xblock(nx);
  hipLaunchKernelGGL(k_pres_tw_src<T>, grid3(nx, ny, nz, bx), dim3(bx), 0,
                     current_stream(), nx, ny, nz, dxi, dyi, nh_d, nh_u, dzci,
                     rho0i, f_t12, f1, f2, p, pold, rho, u, v, w,
                     const_coeff ? 1 : 0);
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}
template <typename T>
static int pres_sp_src_impl(int nx, int ny, int nz, T f_t12, T dxi,
                       T dyi, T dzi, int nh_d, int nh_u,
                       const T *dzci, T rho0i, const T *pold,
                       T *u, T *v, T *w) {
  (void)dzi;
  FL_TRY(check_dims(nx, ny, nz));
  if (nh_u < 1 || nh_d < 1) return fail(FLUTAS_EINVAL, "halo widths must be >= 1");
  if (!dzci || !pold || !u || !v || !w) return fail(FLUTAS_EINVAL, "null pointer");
  const int bx = xblock(nx);
  hipLaunchKernelGGL(k_pres_sp_src<T>, grid3(nx, ny, nz, bx), dim3(bx), 0,
                     current_stream(), nx, ny, nz, f_t12, dxi, dyi, nh_d, nh_u,
                     dzci, rho0i, pold, u, v, w);
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}
template <typename T>
static int correc_impl(int nx, int ny, int nz, int nh_d, int nh_u, T dxi,
                  T dyi, T dzi, const T *dzci, T dt,
                  T rho0, const T *p, T *u, T *v,
                  T *w, const T *rho, int const_coeff) {
  (void)dzi;
  FL_TRY(check_dims(nx, ny, nz));
  if (nh_u < 1 || nh_d < 1) return fail(FLUTAS_EINVAL, "halo widths must be >= 1");
  if (!dzci || !p || !u || !v || !w || (!const_coeff && !rho))
    return fail(FLUTAS_EINVAL, "null pointer");
  const T rho0i = T(1.0) / rho0, factori = dt * dxi, factorj = dt * dyi;
  const int bx = xblock(nx);
  hipLaunchKernelGGL(k_correc<T>, grid3(nx, ny, nz, bx), dim3(bx), 0,
                     current_stream(), nx, ny, nz, nh_d, nh_u, factori, factorj,
                     dzci, dt, rho0i, p, u, v, w, rho, const_coeff ? 1 : 0);
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}
template <typename T>
static int boundp_impl(const char cbc[6], const int n[3], const T bc[6],
                  const T dl[3], T dzc_0, T dzc_n, T *p) {
  if (!cbc || !n || !bc || !dl || !p) return fail(FLUTAS_EINVAL, "null pointer");
  FL_TRY(check_dims(n[0], n[1], n[2]));
  for (int q = 0; q < 6; ++q)
    if (cbc[q] != 'P' && cbc[q] != 'D' && cbc[q] != 'N')
      return fail(FLUTAS_EINVAL, "BC must be one of P, D, N");
  hipStream_t s = current_stream();
  const int nx = n[0], ny = n[1], nz = n[2];
  auto face = [&](char t, int side, int dir, T value, T dr) -> int {
    const int ext[3] = {nx + 2, ny + 2, nz + 2};
    const int t1 = (dir == 0) ? 1 : 0, t2 = (dir == 2) ? 1 : 2;
    dim3 grid((ext[t1] + 255) / 256, ext[t2]);
    hipLaunchKernelGGL(k_set_bc<T>, grid, dim3(256), 0, s, nx, ny, nz, t, side,
                       dir, value, dr, p);
    FL_HIP(hipGetLastError());
    return FLUTAS_OK;
  };
  // periodic y and z halos first (the self-exchange of updthalo at one
  // rank, src/bound.f90:182-183), then set_bc for x, y, z
  if (cbc[2] == 'P') { FL_TRY(face('P', 0, 1, 0, 0)); FL_TRY(face('P', 1, 1, 0, 0)); }
  if (cbc[4] == 'P') { FL_TRY(face('P', 0, 2, 0, 0)); FL_TRY(face('P', 1, 2, 0, 0)); }
  FL_TRY(face(cbc[0], 0, 0, bc[0], dl[0]));
  FL_TRY(face(cbc[1], 1, 0, bc[1], dl[0]));
  if (cbc[2] != 'P') { FL_TRY(face(cbc[2], 0, 1, bc[2], dl[1])); FL_TRY(face(cbc[3], 1, 1, bc[3], dl[1])); }
  if (cbc[4] != 'P') { FL_TRY(face(cbc[4], 0, 2, bc[4], dzc_0)); FL_TRY(face(cbc[5], 1, 2, bc[5], dzc_n)); }
  return FLUTAS_OK;
}
// ---- bounduvw in one launch ----------------------------------------------
// The face launches of bounduvw (updthalo in y and z, then the x, y, z
// faces, each side in turn) as one kernel.  A face stage writes a halo
// position from one value of its own line (a copy, a constant, or
// A + sgn * value, sgn = +-1), so the final value of a point is a chain: its
// last writer's rule applied to the source point's value at the stage
// before, and so on, ending at a constant or at a point no stage writes.
// Each thread walks that chain for its own point and applies the rules
// from the innermost out, the order and operations of the face launches,
// so the result is theirs bit for bit.  A chain only moves a coordinate
// from the halo into the line's interior and, because a stage's write set
// depends only on the coordinate along its direction, every stage writing
// a halo coordinate of the end point also wrote the start point earlier in
// the chain: the end point is never written by any stage.  So each thread
// writes only its own point and reads only points nobody writes.  Used
// when every n > 2 nh (so that a source never lands in the opposite halo).
constexpr int kUvwStages = 8;
template <typename T>
struct UvwStage {
  int dir, side;
  char t[3];  // rule per field ('P' both sides, 'D', 'N'); 0: not touched
  T f0[3];
  T dr[3][kMaxHalo];
};
template <typename T>
struct UvwArgs {
  T *p[3];
  UvwStage<T> st[kUvwStages];
  int nst, nh;
  int n[3];
  long long cnt[3];  // points of the direction slabs of the shell
};

// Rule of stage S for field c at position pos along S.dir: 0 not written,
// 1 constant A, 2 copy of src, 3 A + sg * (value at src).
template <typename T>
__device__ __forceinline__ int uvw_rule(const UvwStage<T> &S, int c, int pos, int nd, int nh,
                                        T *A, T *sg, int *src) {
#pragma clang fp contract(off)
  const char t = S.t[c];
  if (!t) return 0;
  if (t == 'P') {
    if (pos <= 0 && pos >= 1 - nh) {
      *src = nd + pos;
      return 2;
    }
    if (pos >= nd + 1 && pos <= nd + nh) {
      *src = pos - nd;
      return 2;
    }
    return 0;
  }
  const T f0 = S.f0[c];
  if (c != S.dir) {
    // cell-centred component: f + sgn * L, f = 2 value (D) or -+dr value (N)
    int q;
    if (S.side == 0) {
      if (pos > 0 || pos < 1 - nh) return 0;
      q = -pos;
      *src = 1 + q;
      *A = (t == 'D') ? T(2.0) * f0 : -S.dr[c][q] * f0;
    } else {
      if (pos < nd + 1 || pos > nd + nh) return 0;
      q = pos - nd - 1;
      *src = nd - q;
      *A = (t == 'D') ? T(2.0) * f0 : S.dr[c][q] * f0;
    }
    *sg = (t == 'D') ? -T(1.0) : T(1.0);
    return 3;
  }
  if (t == 'D') {
    // face-normal wall: the value; the outer upper layer copies nd - nh
    if (S.side == 0) {
      if (pos > 0 || pos < 1 - nh) return 0;
      *A = f0;
      return 1;
    }
    if (pos >= nd && pos <= nd + nh - 1) {
      *A = f0;
      return 1;
    }
    if (pos == nd + nh) {
      *src = nd - nh;
      return 2;
    }
    return 0;
  }
  // face-normal N: 1.0 f + L (2.0 f + L on the outer upper layer)
  *sg = T(1.0);
  if (S.side == 0) {
    if (pos > 0 || pos < 1 - nh) return 0;
    const int q = -pos;
    *A = T(1.0) * (-S.dr[c][q] * f0);
    *src = 1 + q;
    return 3;
  }
  if (pos >= nd && pos <= nd + nh - 1) {
    const int r = pos - nd;
    *A = T(1.0) * (S.dr[c][r] * f0);
    *src = nd - 1 - r;
    return 3;
  }
  if (pos == nd + nh) {
    *A = T(2.0) * (S.dr[c][nh - 1] * f0);
    *src = nd - nh;
    return 3;
  }
  return 0;
}

template <typename T>
__global__ void __launch_bounds__(256) k_bounduvw_one(UvwArgs<T> a) {
#pragma clang fp contract(off)
  long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  int d = 0;
  while (d < 3 && t >= a.cnt[d]) {
    t -= a.cnt[d];
    ++d;
  }
  if (d == 3) return;
  const int nh = a.nh;
  // slab d: positions 1-nh..0 and n_d..n_d+nh along d; along e < d the
  // interior positions 1..n_e-1 (the rest belongs to slab e), along e > d
  // every position
  int pos[3];
  const int L = (int)(t % (2 * nh + 1));
  t /= (2 * nh + 1);
  pos[d] = (L < nh) ? (1 - nh + L) : (a.n[d] + (L - nh));
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    if (e == d) continue;
    const int len = (e < d) ? a.n[e] - 1 : a.n[e] + 2 * nh;
    pos[e] = ((e < d) ? 1 : 1 - nh) + (int)(t % len);
    t /= len;
  }
  const long long ex = a.n[0] + 2LL * nh, exy = ex * (a.n[1] + 2LL * nh);
  auto at = [&](const int *q) {
    return (q[0] - 1 + nh) + ex * (q[1] - 1 + nh) + exy * (q[2] - 1 + nh);
  };
#pragma unroll 1
  for (int c = 0; c < 3; ++c) {
    int cur[3] = {pos[0], pos[1], pos[2]};
    T As[kUvwStages], Ss[kUvwStages];
    int nr = 0, s = a.nst - 1;
    bool written = false;
    T v = T(0.0);
    for (;;) {
      int kind = 0, src = 0, sw = -1;
      T A = T(0.0), sg = T(0.0);
      for (; s >= 0; --s) {
        const int dd = a.st[s].dir;
        kind = uvw_rule(a.st[s], c, cur[dd], a.n[dd], nh, &A, &sg, &src);
        if (kind) {
          sw = s;
          break;
        }
      }
      if (sw < 0) {
        v = a.p[c][at(cur)];
        break;
      }
      written = true;
      if (kind == 1) {
        v = A;
        break;
      }
      if (kind == 3) {
        As[nr] = A;
        Ss[nr] = sg;
        ++nr;
      }
      cur[a.st[sw].dir] = src;
      s = sw - 1;
    }
    if (!written) continue;
    for (int r = nr - 1; r >= 0; --r) v = As[r] + Ss[r] * v;
    a.p[c][at(pos)] = v;
  }
}

template <typename T>
static int bounduvw_impl(const char cbc[18], const int n[3], const T bc[18],
                    int nh_d, int nh_u, const int isoutflow[6],
                    const T dl[3], const T *dzc, const T *dzf,
                    T *u, T *v, T *w) {
  if (!cbc || !n || !bc || !isoutflow || !dl || !dzc || !dzf || !u || !v || !w)
    return fail(FLUTAS_EINVAL, "null pointer");
  FL_TRY(check_dims(n[0], n[1], n[2]));
  if (n[0] < 2 || n[1] < 2 || n[2] < 2) return fail(FLUTAS_EINVAL, "grid sizes must be >= 2");
  if (nh_u < 1 || nh_u > kMaxHalo) return fail(FLUTAS_EINVAL, "nh_u must be in 1..8");
  if (nh_d < nh_u) return fail(FLUTAS_EINVAL, "nh_d must be >= nh_u (dzc, dzf at the halo layers)");
  for (int d = 0; d < 3; ++d)
    if (n[d] + 2 * nh_u > 65535) return fail(FLUTAS_EINVAL, "grid too large for the face launches");
  auto CB = [&](int s, int d, int c) { return cbc[s + 2 * d + 6 * c]; };
  auto BV = [&](int s, int d, int c) { return bc[s + 2 * d + 6 * c]; };
  for (int c = 0; c < 3; ++c)
    for (int d = 0; d < 3; ++d) {
      for (int s = 0; s < 2; ++s)
        if (CB(s, d, c) != 'P' && CB(s, d, c) != 'D' && CB(s, d, c) != 'N')
          return fail(FLUTAS_EINVAL, "BC must be one of P, D, N");
      if ((CB(0, d, c) == 'P') != (CB(1, d, c) == 'P'))
        return fail(FLUTAS_EINVAL, "periodic BC must be set on both sides");
    }
  hipStream_t s = current_stream();
  T *fld[3] = {u, v, w};
  const int nz = n[2];
  auto Z = [&](const T *a, int k) { return a[k + nh_d - 1]; };
  // the face stages in the reference's order, launched one by one or, when
  // every n > 2 nh, as one k_bounduvw_one launch
  std::vector<FaceArgs<T>> stages;
  auto launch = [&](FaceArgs<T> &a) -> int {
    stages.push_back(a);
    return FLUTAS_OK;
  };
  auto face = [&](int dir, int side) {
    FaceArgs<T> a{};
    for (int q = 0; q < 3; ++q) a.n[q] = n[q];
    a.nh = nh_u;
    a.dir = dir;
    a.side = side;
    return a;
  };
  // updthalo in y and z (periodic directions)
  for (int d = 1; d <= 2; ++d) {
    FaceArgs<T> a = face(d, 0);
    for (int c = 0; c < 3; ++c)
      if (CB(0, d, c) == 'P') {
        a.f[c].p = fld[c];
        a.f[c].ctype = 'P';
        a.f[c].centered = 1;
      }
    if (a.f[0].ctype || a.f[1].ctype || a.f[2].ctype) FL_TRY(launch(a));
  }
  for (int d = 0; d < 3; ++d)
    for (int side = 0; side < 2; ++side) {
      FaceArgs<T> a = face(d, side);
      for (int c = 0; c < 3; ++c) {
        if (d > 0 && CB(side, d, c) == 'P') continue;  // done by updthalo
        // a periodic set_bc copies both sides; its second call (side 1)
        // would repeat the same copies
        if (d == 0 && side == 1 && CB(side, d, c) == 'P') continue;
        FaceField<T> &F = a.f[c];
        F.p = fld[c];
        F.ctype = CB(side, d, c);
        F.centered = (c != d);
        F.value = BV(side, d, c);
        for (int q = 0; q < nh_u; ++q) {
          if (d < 2) F.dr[q] = dl[d];
          else if (c < 2) F.dr[q] = (side == 0) ? Z(dzc, -q) : Z(dzc, nz + q);
          else F.dr[q] = (side == 0) ? Z(dzf, -q) : Z(dzf, nz + q);
        }
      }
      if (a.f[0].ctype || a.f[1].ctype || a.f[2].ctype) FL_TRY(launch(a));
    }
  const bool one = n[0] > 2 * nh_u && n[1] > 2 * nh_u && n[2] > 2 * nh_u &&
                   (int)stages.size() <= kUvwStages;
  if (one && !stages.empty()) {
    UvwArgs<T> A{};
    for (int c = 0; c < 3; ++c) A.p[c] = fld[c];
    A.nst = (int)stages.size();
    A.nh = nh_u;
    for (int q = 0; q < 3; ++q) A.n[q] = n[q];
    for (int k = 0; k < A.nst; ++k) {
      const FaceArgs<T> &F = stages[k];
      UvwStage<T> &S = A.st[k];
      S.dir = F.dir;
      S.side = F.side;
      for (int c = 0; c < 3; ++c) {
        S.t[c] = F.f[c].ctype;
        S.f0[c] = F.f[c].value;
        for (int q = 0; q < kMaxHalo; ++q) S.dr[c][q] = F.f[c].dr[q];
      }
    }
    long long total = 0;
    for (int d = 0; d < 3; ++d) {
      long long m = 2 * nh_u + 1;
      for (int e = 0; e < 3; ++e)
        if (e != d) m *= (e < d) ? n[e] - 1 : n[e] + 2 * nh_u;
      A.cnt[d] = m;
      total += m;
    }
    hipLaunchKernelGGL(k_bounduvw_one<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       A);
    FL_HIP(hipGetLastError());
  } else {
    for (FaceArgs<T> &a : stages) {
      const int t1 = (a.dir == 0) ? 1 : 0, t2 = (a.dir == 2) ? 1 : 2;
      const int e1 = n[t1] + 2 * nh_u, e2 = n[t2] + 2 * nh_u;
      hipLaunchKernelGGL(k_set_bc_face<T>, dim3((e1 + 255) / 256, e2), dim3(256), 0, s, a);
      FL_HIP(hipGetLastError());
    }
  }
  // outflow faces
  bool need_dzf = false;
  for (int d = 0; d < 2; ++d) need_dzf = need_dzf || isoutflow[2 * d] || isoutflow[2 * d + 1];
  T *dzf_d = nullptr;
  const size_t nzh = (size_t)nz + 2 * nh_d;
  if (need_dzf) {
    FL_HIP(hipMalloc((void **)&dzf_d, sizeof(T) * nzh));
    hipError_t e = hipMemcpy(dzf_d, dzf, sizeof(T) * nzh, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(dzf_d);
      return fail(FLUTAS_EDEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
    }
  }
  int rc = FLUTAS_OK;
  for (int d = 0; d < 3 && rc == FLUTAS_OK; ++d)
    for (int side = 0; side < 2 && rc == FLUTAS_OK; ++side) {
      if (!isoutflow[side + 2 * d]) continue;
      OutflowArgs<T> a{};
      a.u = u; a.v = v; a.w = w;
      a.dzf = dzf_d;
      // layers written: 0..nh_u on the upper side, 0..nh_u-1 on the lower
      for (int q = 0; q <= (side == 1 ? nh_u : nh_u - 1); ++q)
        a.dzq[q] = (side == 0) ? Z(dzf, -q) : Z(dzf, nz + q);
      a.dx = dl[0]; a.dy = dl[1];
      a.dxi = T(1.0) / dl[0]; a.dyi = T(1.0) / dl[1];
      for (int q = 0; q < 3; ++q) a.n[q] = n[q];
      a.nh = nh_u; a.nh_d = nh_d; a.dir = d; a.side = side;
      const int t1 = (d == 0) ? n[1] : n[0], t2 = (d == 2) ? n[1] : n[2];
      hipLaunchKernelGGL(k_outflow<T>, dim3((t1 + 255) / 256, t2), dim3(256), 0, s, a);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) rc = fail(FLUTAS_EDEVICE, std::string("k_outflow: ") + hipGetErrorString(e));
    }
  if (dzf_d) {
    hipError_t e = hipStreamSynchronize(s);
    (void)hipFree(dzf_d);
    if (rc == FLUTAS_OK && e != hipSuccess) rc = fail(FLUTAS_EDEVICE, hipGetErrorString(e));
  }
  return rc;
}
template <typename T>
static int boundp_halo_impl(const char cbc[6], const int n[3], const T bc[6],
                       int nh_d, int nh_p, const T dl[3],
                       const T *dzc, const T *dzf, T *p) {
  (void)dzf;
  if (!cbc || !n || !bc || !dl || !dzc || !p) return fail(FLUTAS_EINVAL, "null pointer");
  FL_TRY(check_dims(n[0], n[1], n[2]));
  if (n[0] < 2 || n[1] < 2 || n[2] < 2) return fail(FLUTAS_EINVAL, "grid sizes must be >= 2");
  if (nh_p < 1 || nh_p > kMaxHalo) return fail(FLUTAS_EINVAL, "nh_p must be in 1..8");
  if (nh_d < nh_p) return fail(FLUTAS_EINVAL, "nh_d must be >= nh_p (dzc at the halo layers)");
  for (int d = 0; d < 3; ++d) {
    if (n[d] + 2 * nh_p > 65535) return fail(FLUTAS_EINVAL, "grid too large for the face launches");
    for (int s = 0; s < 2; ++s)
      if (cbc[2 * d + s] != 'P' && cbc[2 * d + s] != 'D' && cbc[2 * d + s] != 'N')
        return fail(FLUTAS_EINVAL, "BC must be one of P, D, N");
    if ((cbc[2 * d] == 'P') != (cbc[2 * d + 1] == 'P'))
      return fail(FLUTAS_EINVAL, "periodic BC must be set on both sides");
  }
  hipStream_t s = current_stream();
  const int nz = n[2];
  auto run = [&](int dir, int side, char t, T value, T dr0, bool zlo, bool zhi) -> int {
    FaceArgs<T> a{};
    for (int q = 0; q < 3; ++q) a.n[q] = n[q];
    a.nh = nh_p;
    a.dir = dir;
    a.side = side;
    FaceField<T> &F = a.f[0];
    F.p = p;
    F.ctype = t;
    F.centered = 1;
    F.value = value;
    for (int q = 0; q < nh_p; ++q)
      F.dr[q] = zlo ? dzc[-q + nh_d - 1] : (zhi ? dzc[nz + q + nh_d - 1] : dr0);
    const int t1 = (dir == 0) ? 1 : 0, t2 = (dir == 2) ? 1 : 2;
    const int e1 = n[t1] + 2 * nh_p, e2 = n[t2] + 2 * nh_p;
    hipLaunchKernelGGL(k_set_bc_face<T>, dim3((e1 + 255) / 256, e2), dim3(256), 0, s, a);
    FL_HIP(hipGetLastError());
    return FLUTAS_OK;
  };
  // updthalo in y and z (periodic), then x, y, z faces as set_bc
  if (cbc[2] == 'P') FL_TRY(run(1, 0, 'P', T(0.0), T(0.0), false, false));
  if (cbc[4] == 'P') FL_TRY(run(2, 0, 'P', T(0.0), T(0.0), false, false));
  FL_TRY(run(0, 0, cbc[0], bc[0], dl[0], false, false));
  if (cbc[1] != 'P') FL_TRY(run(0, 1, cbc[1], bc[1], dl[0], false, false));  // P: both sides done
  if (cbc[2] != 'P') FL_TRY(run(1, 0, cbc[2], bc[2], dl[1], false, false));
  if (cbc[3] != 'P') FL_TRY(run(1, 1, cbc[3], bc[3], dl[1], false, false));
  if (cbc[4] != 'P') FL_TRY(run(2, 0, cbc[4], bc[4], T(0.0), true, false));
  if (cbc[5] != 'P') FL_TRY(run(2, 1, cbc[5], bc[5], T(0.0), false, true));
  return FLUTAS_OK;
}
template <typename T>
static int updt_rhs_b_impl(int nx, int ny, int nz, const char c_or_f[3],
                      const char cbc[6], const T *rhsbx,
                      const T *rhsby, const T *rhsbz, T *p) {
  FL_TRY(check_dims(nx, ny, nz));
  if (!c_or_f || !cbc || !rhsbx || !rhsby || !rhsbz || !p) return fail(FLUTAS_EINVAL, "null pointer");
  const int q1 = (c_or_f[0] == 'f' && cbc[1] == 'D') ? 1 : 0;
  const int q2 = (c_or_f[1] == 'f' && cbc[3] == 'D') ? 1 : 0;
  const int q3 = (c_or_f[2] == 'f' && cbc[5] == 'D') ? 1 : 0;
  hipStream_t s = current_stream();
  hipLaunchKernelGGL(k_add_face<T>, dim3((ny + 255) / 256, nz), dim3(256), 0, s, nx, ny, nz, 0, q1, rhsbx, p);
  hipLaunchKernelGGL(k_add_face<T>, dim3((nx + 255) / 256, nz), dim3(256), 0, s, nx, ny, nz, 1, q2, rhsby, p);
  hipLaunchKernelGGL(k_add_face<T>, dim3((nx + 255) / 256, ny), dim3(256), 0, s, nx, ny, nz, 2, q3, rhsbz, p);
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

extern "C" {

int flutas_fillps(int nx, int ny, int nz, int nh_d, int nh_u, double dxi,
                  double dyi, double dzi, const double *dzfi, double dti,
                  double rho0, const double *u, const double *v,
                  const double *w, double *p) {
  return fillps_impl<double>(nx, ny, nz, nh_d, nh_u, dxi, dyi, dzi, dzfi, dti, rho0, u, v, w, p);
}
int flutas_fillps_f32(int nx, int ny, int nz, int nh_d, int nh_u, float dxi,
                  float dyi, float dzi, const float *dzfi, float dti,
                  float rho0, const float *u, const float *v,
                  const float *w, float *p) {
  return fillps_impl<float>(nx, ny, nz, nh_d, nh_u, dxi, dyi, dzi, dzfi, dti, rho0, u, v, w, p);
}


int flutas_pres_tw_src(int nx, int ny, int nz, double dxi, double dyi,
                       double dzi, int nh_d, int nh_u, const double *dzci,
                       double rho0i, double f_t12, double f_t12_o,
                       const double *p, const double *pold, const double *rho,
                       double *u, double *v, double *w, int const_coeff) {
  return pres_tw_src_impl<double>(nx, ny, nz, dxi, dyi, dzi, nh_d, nh_u, dzci, rho0i, f_t12, f_t12_o, p, pold, rho, u, v, w, const_coeff);
}
int flutas_pres_tw_src_f32(int nx, int ny, int nz, float dxi, float dyi,
                       float dzi, int nh_d, int nh_u, const float *dzci,
                       float rho0i, float f_t12, float f_t12_o,
                       const float *p, const float *pold, const float *rho,
                       float *u, float *v, float *w, int const_coeff) {
  return pres_tw_src_impl<float>(nx, ny, nz, dxi, dyi, dzi, nh_d, nh_u, dzci, rho0i, f_t12, f_t12_o, p, pold, rho, u, v, w, const_coeff);
}


int flutas_pres_sp_src(int nx, int ny, int nz, double f_t12, double dxi,
                       double dyi, double dzi, int nh_d, int nh_u,
                       const double *dzci, double rho0i, const double *pold,
                       double *u, double *v, double *w) {
  return pres_sp_src_impl<double>(nx, ny, nz, f_t12, dxi, dyi, dzi, nh_d, nh_u, dzci, rho0i, pold, u, v, w);
}
int flutas_pres_sp_src_f32(int nx, int ny, int nz, float f_t12, float dxi,
                       float dyi, float dzi, int nh_d, int nh_u,
                       const float *dzci, float rho0i, const float *pold,
                       float *u, float *v, float *w) {
  return pres_sp_src_impl<float>(nx, ny, nz, f_t12, dxi, dyi, dzi, nh_d, nh_u, dzci, rho0i, pold, u, v, w);
}


int flutas_correc(int nx, int ny, int nz, int nh_d, int nh_u, double dxi,
                  double dyi, double dzi, const double *dzci, double dt,
                  double rho0, const double *p, double *u, double *v,
                  double *w, const double *rho, int const_coeff) {
  return correc_impl<double>(nx, ny, nz, nh_d, nh_u, dxi, dyi, dzi, dzci, dt, rho0, p, u, v, w, rho, const_coeff);
}
int flutas_correc_f32(int nx, int ny, int nz, int nh_d, int nh_u, float dxi,
                  float dyi, float dzi, const float *dzci, float dt,
                  float rho0, const float *p, float *u, float *v,
                  float *w, const float *rho, int const_coeff) {
  return correc_impl<float>(nx, ny, nz, nh_d, nh_u, dxi, dyi, dzi, dzci, dt, rho0, p, u, v, w, rho, const_coeff);
}


int flutas_chkdiv(int nx, int ny, int nz, double dxi, double dyi, double dzi,
                  int nh_d, int nh_u, const double *dzfi, const double *u,
                  const double *v, const double *w, double *divtot,
                  double *divmax) {
  double res[2];
  FL_TRY(chkdiv_impl<double>(nx, ny, nz, dxi, dyi, dzi, nh_d, nh_u, dzfi, u, v, w, res));
  *divtot = res[0];
  *divmax = res[1];
  return FLUTAS_OK;
}

int flutas_chkdiv_f32(int nx, int ny, int nz, float dxi, float dyi, float dzi,
                      int nh_d, int nh_u, const float *dzfi, const float *u,
                      const float *v, const float *w, float *divtot,
                      float *divmax) {
  double res[2];
  if (!divtot || !divmax) return fail(FLUTAS_EINVAL, "null pointer");
  FL_TRY(chkdiv_impl<float>(nx, ny, nz, dxi, dyi, dzi, nh_d, nh_u, dzfi, u, v, w, res));
  *divtot = (float)res[0];
  *divmax = (float)res[1];
  return FLUTAS_OK;
}

int flutas_chkdiv_noalloc(int nx, int ny, int nz, double dxi, double dyi, double dzi,
                          int nh_d, int nh_u, const double *dzfi, const double *u,
                          const double *v, const double *w, double *divtot,
                          double *divmax) {
  (void)dzi;
  FL_TRY(check_dims(nx, ny, nz));
  if (!dzfi || !u || !v || !w || !divtot || !divmax) return fail(FLUTAS_EINVAL, "null pointer");
  const int nb = chkdiv_blocks((long long)nx * ny * nz);
  if (2 * nb + 2 > kNoAllocDoubles)
    return fail(FLUTAS_EINVAL, "chkdiv partials exceed the fixed scratch of " +
                                   std::to_string(kNoAllocDoubles) + " doubles");
  double *work = nullptr;
  FL_TRY(chkdiv_scratch_ptr(&work));
  hipStream_t s = current_stream();
  double res[2] = {0.0, 0.0};
  FL_TRY(chkdiv_enqueue(nx, ny, nz, nh_d, nh_u, dxi, dyi, dzfi, u, v, w, work, res, s));
  FL_HIP(hipStreamSynchronize(s));
  *divtot = res[0];
  *divmax = res[1];
  return FLUTAS_OK;
}

int flutas_boundp(const char cbc[6], const int n[3], const double bc[6],
                  const double dl[3], double dzc_0, double dzc_n, double *p) {
  return boundp_impl<double>(cbc, n, bc, dl, dzc_0, dzc_n, p);
}
int flutas_boundp_f32(const char cbc[6], const int n[3], const float bc[6],
                  const float dl[3], float dzc_0, float dzc_n, float *p) {
  return boundp_impl<float>(cbc, n, bc, dl, dzc_0, dzc_n, p);
}


// bounduvw at one rank, in the decomposition model of flutas_boundp:
// periodic y and z halos (updthalo's self-exchange) first, then the x faces
// (u face-normal, v and w centred), the y and z faces that are not periodic,
// then the outflow faces in the order x, y, z, lower side first.  One launch
// per face keeps the reference's face order, which decides the halo edges
// and corners.
int flutas_bounduvw(const char cbc[18], const int n[3], const double bc[18],
                    int nh_d, int nh_u, const int isoutflow[6],
                    const double dl[3], const double *dzc, const double *dzf,
                    double *u, double *v, double *w) {
  return bounduvw_impl<double>(cbc, n, bc, nh_d, nh_u, isoutflow, dl, dzc, dzf, u, v, w);
}
int flutas_bounduvw_f32(const char cbc[18], const int n[3], const float bc[18],
                    int nh_d, int nh_u, const int isoutflow[6],
                    const float dl[3], const float *dzc, const float *dzf,
                    float *u, float *v, float *w) {
  return bounduvw_impl<float>(cbc, n, bc, nh_d, nh_u, isoutflow, dl, dzc, dzf, u, v, w);
}


// boundp for a field with halo nh_p (src/bound.f90:146-225): the face
// launches of flutas_bounduvw on one cell-centred field.
int flutas_boundp_halo(const char cbc[6], const int n[3], const double bc[6],
                       int nh_d, int nh_p, const double dl[3],
                       const double *dzc, const double *dzf, double *p) {
  return boundp_halo_impl<double>(cbc, n, bc, nh_d, nh_p, dl, dzc, dzf, p);
}
int flutas_boundp_halo_f32(const char cbc[6], const int n[3], const float bc[6],
                       int nh_d, int nh_p, const float dl[3],
                       const float *dzc, const float *dzf, float *p) {
  return boundp_halo_impl<float>(cbc, n, bc, nh_d, nh_p, dl, dzc, dzf, p);
}


int flutas_updt_rhs_b(int nx, int ny, int nz, const char c_or_f[3],
                      const char cbc[6], const double *rhsbx,
                      const double *rhsby, const double *rhsbz, double *p) {
  return updt_rhs_b_impl<double>(nx, ny, nz, c_or_f, cbc, rhsbx, rhsby, rhsbz, p);
}
int flutas_updt_rhs_b_f32(int nx, int ny, int nz, const char c_or_f[3],
                      const char cbc[6], const float *rhsbx,
                      const float *rhsby, const float *rhsbz, float *p) {
  return updt_rhs_b_impl<float>(nx, ny, nz, c_or_f, cbc, rhsbx, rhsby, rhsbz, p);
}


int flutas_fillps_rhs(int nx, int ny, int nz, int nh_d, int nh_u, double dxi,
                      double dyi, double dzi, const double *dzfi, double dti,
                      double rho0, const double *u, const double *v,
                      const double *w, const char c_or_f[3], const char cbc[6],
                      const double *rhsbx, const double *rhsby,
                      const double *rhsbz, double *p) {
  (void)dzi;
  FL_TRY(check_dims(nx, ny, nz));
  if (nh_u < 1 || nh_d < 1) return fail(FLUTAS_EINVAL, "halo widths must be >= 1");
  if (!dzfi || !u || !v || !w || !c_or_f || !cbc || !rhsbx || !rhsby || !rhsbz || !p)
    return fail(FLUTAS_EINVAL, "null pointer");
  const int q1 = (c_or_f[0] == 'f' && cbc[1] == 'D') ? 1 : 0;
  const int q2 = (c_or_f[1] == 'f' && cbc[3] == 'D') ? 1 : 0;
  const int q3 = (c_or_f[2] == 'f' && cbc[5] == 'D') ? 1 : 0;
  const int bx = xblock(nx);
  hipLaunchKernelGGL(k_fillps_rhs, grid3(nx, ny, nz, bx), dim3(bx), 0, current_stream(), nx,
                     ny, nz, nh_d, nh_u, dxi, dyi, dzfi, dti, rho0, u, v, w, q1, q2, q3, rhsbx,
                     rhsby, rhsbz, p);
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

int flutas_correc_update(int nx, int ny, int nz, int nh_d, int nh_u, double dxi,
                         double dyi, double dzi, const double *dzci, double dt,
                         double rho0, const double *phi, const double *pold,
                         double *pnew, double *u, double *v, double *w,
                         const double *rho, int const_coeff) {
  (void)dzi;
  FL_TRY(check_dims(nx, ny, nz));
  if (nh_u < 1 || nh_d < 1) return fail(FLUTAS_EINVAL, "halo widths must be >= 1");
  if (!dzci || !phi || !pold || !pnew || !u || !v || !w || (!const_coeff && !rho))
    return fail(FLUTAS_EINVAL, "null pointer");
  if (pnew == phi) return fail(FLUTAS_EINVAL, "pnew must not be the increment array phi");
  const double rho0i = 1.0 / rho0, factori = dt * dxi, factorj = dt * dyi;
  const int bx = xblock(nx);
  hipLaunchKernelGGL(k_correc_update<false>, grid3(nx, ny, nz, bx), dim3(bx), 0,
                     current_stream(), nx, ny, nz, nh_d, nh_u, factori, factorj, dzci, dt, rho0i,
                     phi, pold, pnew, u, v, w, rho, const_coeff ? 1 : 0, HaloBc{});
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

int flutas_correc_update_bc(int nx, int ny, int nz, int nh_d, int nh_u, double dxi,
                            double dyi, double dzi, const double *dzci, double dt,
                            double rho0, const double *phi, const double *pold,
                            double *pnew, double *u, double *v, double *w,
                            const double *rho, int const_coeff, const char cbc[6],
                            const double bc[6], const double dl[3], double dzc_0,
                            double dzc_n) {
  (void)dzi;
  FL_TRY(check_dims(nx, ny, nz));
  if (nh_u < 1 || nh_d < 1) return fail(FLUTAS_EINVAL, "halo widths must be >= 1");
  if (!dzci || !phi || !pold || !pnew || !u || !v || !w || (!const_coeff && !rho) || !cbc ||
      !bc || !dl)
    return fail(FLUTAS_EINVAL, "null pointer");
  if (pnew == phi) return fail(FLUTAS_EINVAL, "pnew must not be the increment array phi");
  for (int q = 0; q < 6; ++q)
    if (cbc[q] != 'P' && cbc[q] != 'D' && cbc[q] != 'N')
      return fail(FLUTAS_EINVAL, "BC must be one of P, D, N");
  for (int d = 0; d < 3; ++d)
    if ((cbc[2 * d] == 'P') != (cbc[2 * d + 1] == 'P'))
      return fail(FLUTAS_EINVAL, "periodic BCs must be paired");
  HaloBc hb{};
  for (int d = 0; d < 3; ++d)
    for (int sd = 0; sd < 2; ++sd) {
      hb.t[d][sd] = cbc[2 * d + sd];
      hb.val[d][sd] = bc[2 * d + sd];
    }
  hb.dr[0][0] = hb.dr[0][1] = dl[0];
  hb.dr[1][0] = hb.dr[1][1] = dl[1];
  hb.dr[2][0] = dzc_0;
  hb.dr[2][1] = dzc_n;
  // boundp's face order: periodic y, periodic z, x, other y, other z
  hb.rank[0] = 2;
  hb.rank[1] = cbc[2] == 'P' ? 0 : 3;
  hb.rank[2] = cbc[4] == 'P' ? 1 : 4;
  const double rho0i = 1.0 / rho0, factori = dt * dxi, factorj = dt * dyi;
  const int bx = xblock(nx);
  hipLaunchKernelGGL(k_correc_update<true>, grid3(nx, ny, nz, bx), dim3(bx), 0,
                     current_stream(), nx, ny, nz, nh_d, nh_u, factori, factorj, dzci, dt, rho0i,
                     phi, pold, pnew, u, v, w, rho, const_coeff ? 1 : 0, hb);
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

int flutas_copy_interior(int nx, int ny, int nz, const double *src, double *dst) {
  FL_TRY(check_dims(nx, ny, nz));
  if (!src || !dst) return fail(FLUTAS_EINVAL, "null pointer");
  const int bx = xblock(nx);
  hipLaunchKernelGGL(k_interior, grid3(nx, ny, nz, bx), dim3(bx), 0, current_stream(), nx, ny, nz, src, dst, 0);
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

int flutas_add_interior(int nx, int ny, int nz, const double *pold, double *p) {
  FL_TRY(check_dims(nx, ny, nz));
  if (!pold || !p) return fail(FLUTAS_EINVAL, "null pointer");
  const int bx = xblock(nx);
  hipLaunchKernelGGL(k_interior, grid3(nx, ny, nz, bx), dim3(bx), 0, current_stream(), nx, ny, nz, pold, p, 1);
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

}  // extern "C"
