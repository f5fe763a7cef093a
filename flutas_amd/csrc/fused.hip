// Fused four-pass pressure solve for non-periodic z (the hot path).
//
// Work layout Wt[k][i][j] (j fastest): y-lines are contiguous, and an
// x-tile of P1/P4 (fixed k) touches a single 2-D plane.
//
//   P1 k_x_fwd   p (halo layout, x-lines) -> x r2r -> Wt (transposed store)
//   P2 k_yz_fwd  per x-slab i, k ascending: y r2r of each plane, then the
//                Thomas forward elimination step for all (i, j) columns
//   P3 k_zy_bwd  per x-slab i, k descending: Thomas back substitution,
//                then the inverse y r2r of each plane
//   P4 k_x_bwd   Wt (transposed load) -> inverse x r2r * normfft -> p
//
// Each pass reads and writes the field once.  The elimination factors d(k)
// are not stored: P2 writes one checkpoint d per column every GP planes and
// P3 recomputes the factors of each plane group from it with the same
// operations, so the pivots remain the reference's bit for bit
// (src/solver_cpu.f90:187-223, contraction off).
//
// Memory-level parallelism: every thread issues all global loads of a tile
// before touching LDS, and the slab kernels P2/P3 load the next plane group
// into registers while the current group is transformed and eliminated.
#include "tile_fft.hpp"

namespace flutas {

using tile::cx;

// LDS line stride: two extra elements per line so that the transposed
// accesses of P1/P4 (same element of consecutive lines) spread over banks.
template <typename T, int n>
constexpr int line_stride() {
  return n + (sizeof(T) == 8 ? 2 : 4);
}

// Lines per tile: the largest of 16, 8, 4, 2, 1 whose LDS tile fits
// kTileBytes (two tiles per CU fit the 160 KiB of gfx950 LDS).
constexpr int kTileBytes = 72 * 1024;
template <typename T, int NC>
constexpr int tile_lines() {
  constexpr int bytes = line_stride<T, 2 * NC>() * (int)sizeof(T);
  constexpr int l = kTileBytes / bytes;
  return l >= 16 ? 16 : (l >= 8 ? 8 : (l >= 4 ? 4 : (l >= 2 ? 2 : 1)));
}

// Twiddle tables staged in LDS once per workgroup: exp(-2 pi i m/N) (N),
// exp(-2 pi i k/n) (N+1) and exp(-i pi k/2n) (N+1, the entries the real
// split and the cosine/sine phases use).  Keeping them out of global memory
// leaves the vector-memory counter to the field loads, so a prefetch is not
// drained by the first twiddle read.
template <typename T, int NC>
struct Tw {
  static constexpr int ntw = 2 * tile::Plan<NC>::tw_total();
  static constexpr int nth = 2 * (NC + 1), ntq = 2 * (NC + 1);
  static constexpr int total = ntw + nth + ntq;
  const T *tw, *th, *tq;
};
template <typename T, int NC, int NT>
__device__ __forceinline__ Tw<T, NC> stage_twiddles(T *s, const T *__restrict__ tw,
                                                    const T *__restrict__ twh,
                                                    const T *__restrict__ twq) {
  using W = Tw<T, NC>;
  // fixed trip counts with clamped indices: the loads are issued back to
  // back without branches, then stored to LDS
  constexpr int Q1 = (W::ntw + NT - 1) / NT, Q2 = (W::nth + NT - 1) / NT;
  T v1[Q1], v2[Q2], v3[Q2];
#pragma unroll
  for (int q = 0; q < Q1; ++q) v1[q] = tw[min((int)threadIdx.x + q * NT, W::ntw - 1)];
#pragma unroll
  for (int q = 0; q < Q2; ++q) {
    const int t = min((int)threadIdx.x + q * NT, W::nth - 1);
    v2[q] = twh[t];
    v3[q] = twq[t];
  }
#pragma unroll
  for (int q = 0; q < Q1; ++q) {
    const int t = threadIdx.x + q * NT;
    if (t < W::ntw) s[t] = v1[q];
  }
#pragma unroll
  for (int q = 0; q < Q2; ++q) {
    const int t = threadIdx.x + q * NT;
    if (t < W::nth) {
      s[W::ntw + t] = v2[q];
      s[W::ntw + W::nth + t] = v3[q];
    }
  }
  return Tw<T, NC>{s, s + W::ntw, s + W::ntw + W::nth};
}

// ---------------------------------------------------------------- P1 -----
// One workgroup: TJ consecutive x-lines (j0..j0+TJ-1) at one k.
// in: element (i, j, k) at in[j*is1 + k*is2 + i] (i, j, k 0-based).
// Wt: element (i, j, k) at Wt[(k*nx + i)*ny + j].
template <typename T, typename Tin, int NC, int TJ, int NT>
__global__ void __launch_bounds__(NT)
    k_x_fwd(const Tin *__restrict__ in, long long is1, long long is2,
            T *__restrict__ Wt, int ny, int nz, int kind,
            const T *__restrict__ tw, const T *__restrict__ twh,
            const T *__restrict__ twq) {
  constexpr int n = 2 * NC;  // nx
  constexpr int LS = line_stride<T, n>();
  constexpr int LD = (TJ * n + NT - 1) / NT;
  __shared__ __align__(16) T sm[TJ * LS];
  const int j0 = blockIdx.x * TJ, k = blockIdx.y;
  const int nl = min(TJ, ny - j0);
  const bool trig = tile::is_trig(kind), sine = tile::is_sine(kind);
  Tin r[LD];
  // branch-free loads: lines past ny (partial tile) re-read line ny-1
#pragma unroll
  for (int q = 0; q < LD; ++q) {
    const int t = min((int)threadIdx.x + q * NT, TJ * n - 1);
    const int l = t / n, e = t - (t / n) * n;
    r[q] = in[(long long)min(j0 + l, ny - 1) * is1 + (long long)k * is2 + e];
  }
#pragma unroll
  for (int q = 0; q < LD; ++q) {
    const int t = threadIdx.x + q * NT;
    if (t >= TJ * n) continue;
    const int l = t / n, e = t - (t / n) * n;
    T x = (T)r[q];
    int pos = e;
    if (trig) {
      pos = tile::mk_pos(e, n);
      if (sine && (e & 1)) x = -x;
    }
    sm[l * LS + pos] = x;
  }
  __shared__ __align__(16) T stw[Tw<T, NC>::total];
  const Tw<T, NC> W = stage_twiddles<T, NC, NT>(stw, tw, twh, twq);
  __syncthreads();
  tile::cfft<T, -1, NC, TJ, NT, LS / 2>(reinterpret_cast<cx<T> *>(sm), W.tw);
  tile::post_fwd<T, NC, TJ, NT, LS>(sm, kind, W.th, W.tq);
  // transposed store
  T *plane = Wt + (long long)k * n * ny;
#pragma unroll
  for (int q = 0; q < LD; ++q) {
    const int t = threadIdx.x + q * NT;
    const int l = t % TJ, i = t / TJ;
    if (t < TJ * n && l < nl) plane[(long long)i * ny + j0 + l] = sm[l * LS + i];
  }
}

// ---------------------------------------------------------------- P4 -----
// Inverse of P1: out(i, j, k) = scale * r2r^-1 (x) (+ out when accumulate).
template <typename T, typename Tout, int NC, int TJ, int NT>
__global__ void __launch_bounds__(NT)
    k_x_bwd(const T *__restrict__ Wt, Tout *__restrict__ out, long long os1,
            long long os2, int ny, int nz, int kind, T scale, int accumulate,
            const T *__restrict__ tw, const T *__restrict__ twh,
            const T *__restrict__ twq) {
  constexpr int n = 2 * NC;
  constexpr int LS = line_stride<T, n>();
  constexpr int LD = (TJ * n + NT - 1) / NT;
  __shared__ __align__(16) T sm[TJ * LS];
  const int j0 = blockIdx.x * TJ, k = blockIdx.y;
  const int nl = min(TJ, ny - j0);
  const bool trig = tile::is_trig(kind), sine = tile::is_sine(kind);
  const T *plane = Wt + (long long)k * n * ny;
  T r[LD];
#pragma unroll
  for (int q = 0; q < LD; ++q) {
    const int t = min((int)threadIdx.x + q * NT, TJ * n - 1);
    const int l = t % TJ, i = t / TJ;
    r[q] = plane[(long long)i * ny + min(j0 + l, ny - 1)];
  }
#pragma unroll
  for (int q = 0; q < LD; ++q) {
    const int t = threadIdx.x + q * NT;
    if (t < TJ * n) sm[(t % TJ) * LS + t / TJ] = r[q];
  }
  __shared__ __align__(16) T stw[Tw<T, NC>::total];
  const Tw<T, NC> W = stage_twiddles<T, NC, NT>(stw, tw, twh, twq);
  __syncthreads();
  tile::pre_inv<T, NC, TJ, NT, LS>(sm, kind, W.th, W.tq);
  tile::cfft<T, +1, NC, TJ, NT, LS / 2>(reinterpret_cast<cx<T> *>(sm), W.tw);
  Tout prev[LD];
  if (accumulate) {
#pragma unroll
    for (int q = 0; q < LD; ++q) {
      const int t = min((int)threadIdx.x + q * NT, TJ * n - 1);
      const int l = t / n, e = t - (t / n) * n;
      prev[q] = out[(long long)min(j0 + l, ny - 1) * os1 + (long long)k * os2 + e];
    }
  }
#pragma unroll
  for (int q = 0; q < LD; ++q) {
    const int t = threadIdx.x + q * NT;
    const int l = t / n, e = t - (t / n) * n;
    if (t >= TJ * n || l >= nl) continue;
    int pos = e;
    T sg = T(1);
    if (trig) {
      pos = tile::mk_pos(e, n);
      if (sine && (e & 1)) sg = T(-1);
    }
    const T v = sm[l * LS + pos] * (scale * sg);
    Tout *o = out + (long long)(j0 + l) * os1 + (long long)k * os2 + e;
    if (accumulate) *o = (Tout)((double)prev[q] + (double)v);
    else *o = (Tout)v;
  }
}

// ---------------------------------------------------------------- P2 -----
// One workgroup per x-slab i; planes are processed GP at a time.
// Checkpoints: ck[(i*nck + g)*ny + j] = d of the last plane of group g
// (nck = number of groups); P3 restarts the recurrence from them.  (One
// checkpoint per group: replaying a group of divides in P3 costs more
// latency than the 1/GP extra traffic.)
template <int GP>
__host__ __device__ constexpr int checkpoints(int nz) {
  return (nz + GP - 1) / GP;
}

template <typename T, int NC, int GP, int NT>
__global__ void __launch_bounds__(NT)
    k_yz_fwd(T *__restrict__ Wt, double *__restrict__ ck, int nx, int nz,
             int kind, const double *__restrict__ lamx,
             const double *__restrict__ lamy, const double *__restrict__ a,
             const double *__restrict__ b, const double *__restrict__ c,
             int pin, const T *__restrict__ tw, const T *__restrict__ twh,
             const T *__restrict__ twq) {
#pragma clang fp contract(off)
  constexpr int n = 2 * NC;  // ny
  constexpr int LS = line_stride<T, n>();
  constexpr int CPT = (n + NT - 1) / NT;  // columns per thread
  constexpr int LD = (GP * n + NT - 1) / NT;
  __shared__ __align__(16) T sm[2 * GP * LS];
  T *const A = sm, *const B = sm + GP * LS;
  // the complex FFT ends in A after an even number of passes
  constexpr bool fft_in_b = (tile::Plan<NC>::nst % 2) == 1;
  T *const Z = fft_in_b ? B : A;  // complex FFT result
  T *const Y = fft_in_b ? A : B;  // spectral lines after the real split
  __shared__ __align__(16) T stw[Tw<T, NC>::total];
  const Tw<T, NC> W = stage_twiddles<T, NC, NT>(stw, tw, twh, twq);
  const int i = blockIdx.x;
  const int ngrp = (nz + GP - 1) / GP;
  const int nck = checkpoints<GP>(nz);
  const long long pstride = (long long)nx * n;  // plane stride
  T *line0 = Wt + (long long)i * n;             // plane 0 line of slab i
  const bool trig = tile::is_trig(kind), sine = tile::is_sine(kind);
  const double lx = lamx[i];
  double lam[CPT], dprev[CPT], pprev[CPT];
#pragma unroll
  for (int u = 0; u < CPT; ++u) {
    const int j = threadIdx.x + u * NT;
    lam[u] = (j < n) ? lx + lamy[j] : 0.0;
    dprev[u] = 0.0;
    pprev[u] = 0.0;
  }
  T r[LD];
  // branch-free prefetch: planes past nz (last group) re-read plane nz-1
  auto load = [&](int g) {
    const int k0 = g * GP;
#pragma unroll
    for (int q = 0; q < LD; ++q) {
      const int t = min((int)threadIdx.x + q * NT, GP * n - 1);
      const int l = t / n, e = t - (t / n) * n;
      r[q] = line0[(long long)min(k0 + l, nz - 1) * pstride + e];
    }
  };
  load(0);
  for (int g = 0; g < ngrp; ++g) {
    const int k0 = g * GP;
    const int np = min(GP, nz - k0);
#pragma unroll
    for (int q = 0; q < LD; ++q) {
      const int t = threadIdx.x + q * NT;
      if (t >= GP * n) continue;
      const int l = t / n, e = t - (t / n) * n;
      T x = r[q];
      int pos = e;
      if (trig) {
        pos = tile::mk_pos(e, n);
        if (sine && (e & 1)) x = -x;
      }
      A[l * LS + pos] = x;
    }
    __syncthreads();
    if (g + 1 < ngrp) load(g + 1);
    tile::cfft_pp<T, -1, NC, GP, NT, LS / 2>(reinterpret_cast<cx<T> *>(A),
                                            reinterpret_cast<cx<T> *>(B), W.tw);
    tile::post_fwd_pp<T, NC, GP, NT, LS>(Z, Y, kind, W.th, W.tq);
    // forward elimination, plane by plane (dgtsv_homebrewed order)
    for (int l = 0; l < np; ++l) {
      const int k = k0 + l;
      const double ak = a[k], bk = b[k], cc = c[k];
      T *line = line0 + (long long)k * pstride;
#pragma unroll
      for (int u = 0; u < CPT; ++u) {
        const int j = threadIdx.x + u * NT;
        if (j >= n) continue;
        const double y = (double)Y[l * LS + j];
        double pv;
        if (k == 0) {
          const double z = 1.0 / (bk + lam[u]);
          dprev[u] = cc * z;
          pv = y * z;
        } else if (k < nz - 1) {
          const double bb = bk + lam[u];
          const double z = 1.0 / (bb - ak * dprev[u]);
          dprev[u] = cc * z;
          pv = (y - ak * pprev[u]) * z;
        } else {
          const double z = (bk + lam[u]) - ak * dprev[u];
          if (z != 0.0 && !(pin && lam[u] == 0.0))
            pv = (y - ak * pprev[u]) / z;
          else
            pv = 0.0;
        }
        pprev[u] = pv;
        line[j] = (T)pv;
      }
    }
    // checkpoint after every group: ck[(i*nck + g)*ny + j]
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
      const int j = threadIdx.x + u * NT;
      if (j < n) ck[((long long)i * nck + g) * n + j] = dprev[u];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- P3 -----
template <typename T, int NC, int GP, int NT>
__global__ void __launch_bounds__(NT)
    k_zy_bwd(T *__restrict__ Wt, const double *__restrict__ ck, int nx, int nz,
             int kind, const double *__restrict__ lamx,
             const double *__restrict__ lamy, const double *__restrict__ a,
             const double *__restrict__ b, const double *__restrict__ c,
             const T *__restrict__ tw, const T *__restrict__ twh,
             const T *__restrict__ twq) {
#pragma clang fp contract(off)
  constexpr int n = 2 * NC;
  constexpr int LS = line_stride<T, n>();
  constexpr int CPT = (n + NT - 1) / NT;
  __shared__ __align__(16) T sm[2 * GP * LS];
  T *const A = sm, *const B = sm + GP * LS;
  // pre_inv A -> B, then the FFT passes start in B
  constexpr bool fft_in_a = (tile::Plan<NC>::nst % 2) == 1;
  T *const R = fft_in_a ? A : B;  // inverse FFT result
  __shared__ __align__(16) T stw[Tw<T, NC>::total];
  const Tw<T, NC> W = stage_twiddles<T, NC, NT>(stw, tw, twh, twq);
  const int i = blockIdx.x;
  const int ngrp = (nz + GP - 1) / GP;
  const int nck = checkpoints<GP>(nz);
  const long long pstride = (long long)nx * n;
  T *line0 = Wt + (long long)i * n;
  const bool trig = tile::is_trig(kind), sine = tile::is_sine(kind);
  const double lx = lamx[i];
  double lam[CPT], pnext[CPT];
#pragma unroll
  for (int u = 0; u < CPT; ++u) {
    const int j = threadIdx.x + u * NT;
    lam[u] = (j < n) ? lx + lamy[j] : 0.0;
    pnext[u] = 0.0;
  }
  // p' of a plane group in column-owner layout (thread owns columns j)
  T pv[GP][CPT];
  double dck[CPT];
  auto load = [&](int g) {
    const int k0 = g * GP;
    // branch-free: out-of-range planes/columns re-read valid entries
#pragma unroll
    for (int l = 0; l < GP; ++l)
#pragma unroll
      for (int u = 0; u < CPT; ++u) {
        const int j = min((int)threadIdx.x + u * NT, n - 1);
        pv[l][u] = line0[(long long)min(k0 + l, nz - 1) * pstride + j];
      }
    // the recurrence restarts at the end of group g - 1 (unused, but read
    // in bounds, for g = 0)
    const int c0 = max(g - 1, 0);
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
      const int j = min((int)threadIdx.x + u * NT, n - 1);
      dck[u] = ck[((long long)i * nck + c0) * n + j];
    }
  };
  load(ngrp - 1);
  for (int g = ngrp - 1; g >= 0; --g) {
    const int k0 = g * GP;
    const int np = min(GP, nz - k0);
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
      const int j = threadIdx.x + u * NT;
      if (j >= n) continue;
      // elimination factors of this group, recomputed from the checkpoint
      double d[GP];
      double dp = dck[u];
#pragma unroll
      for (int l = 0; l < GP; ++l) {
        const int k = k0 + l;
        if (l < np && k < nz - 1) {
          if (k == 0) dp = c[0] * (1.0 / (b[0] + lam[u]));
          else dp = c[k] * (1.0 / ((b[k] + lam[u]) - a[k] * dp));
        }
        d[l] = dp;
      }
      double pn = pnext[u];
#pragma unroll
      for (int l = GP - 1; l >= 0; --l) {
        if (l >= np) continue;
        const int k = k0 + l;
        const double p = (k == nz - 1) ? (double)pv[l][u] : (double)pv[l][u] - d[l] * pn;
        pn = p;
        A[l * LS + j] = (T)p;
      }
      pnext[u] = pn;
    }
    __syncthreads();
    if (g > 0) load(g - 1);
    tile::pre_inv_pp<T, NC, GP, NT, LS>(A, B, kind, W.th, W.tq);
    tile::cfft_pp<T, +1, NC, GP, NT, LS / 2>(reinterpret_cast<cx<T> *>(B),
                                            reinterpret_cast<cx<T> *>(A), W.tw);
    constexpr int LD = (GP * n + NT - 1) / NT;
#pragma unroll
    for (int q = 0; q < LD; ++q) {
      const int t = threadIdx.x + q * NT;
      const int l = t / n, e = t - (t / n) * n;
      if (t >= GP * n || l >= np) continue;
      int pos = e;
      T sg = T(1);
      if (trig) {
        pos = tile::mk_pos(e, n);
        if (sine && (e & 1)) sg = T(-1);
      }
      line0[(long long)(k0 + l) * pstride + e] = R[l * LS + pos] * sg;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- host ---

bool fused_supported(int n) {
  switch (n / 2) {
    case 16: case 24: case 32: case 48: case 64: case 96: case 128: case 192:
    case 256: case 384: case 512: case 1024:
      return (n % 2) == 0;
    default:
      return false;
  }
}

#define FL_DISPATCH(NCV, ...)                                \
  switch (NCV) {                                             \
    case 16: { constexpr int NC = 16; __VA_ARGS__; } break;   \
    case 24: { constexpr int NC = 24; __VA_ARGS__; } break;   \
    case 32: { constexpr int NC = 32; __VA_ARGS__; } break;   \
    case 48: { constexpr int NC = 48; __VA_ARGS__; } break;   \
    case 64: { constexpr int NC = 64; __VA_ARGS__; } break;   \
    case 96: { constexpr int NC = 96; __VA_ARGS__; } break;   \
    case 128: { constexpr int NC = 128; __VA_ARGS__; } break; \
    case 192: { constexpr int NC = 192; __VA_ARGS__; } break; \
    case 256: { constexpr int NC = 256; __VA_ARGS__; } break; \
    case 384: { constexpr int NC = 384; __VA_ARGS__; } break; \
    case 512: { constexpr int NC = 512; __VA_ARGS__; } break; \
    case 1024: { constexpr int NC = 1024; __VA_ARGS__; } break; \
    default: return fail(FLUTAS_EUNSUPPORTED, "no fused kernel for this length"); \
  }

// The fused kernels use the per-pass twiddle tables in place of the plain
// exp(-2 pi i m/N) table.
template <typename T>
static LinePlan<T> plan_of(const LineTables &t) {
  LinePlan<T> p;
  if constexpr (sizeof(T) == 8) { p = t.dplan(); p.tw = t.d_ptw; }
  else { p = t.fplan(); p.tw = t.f_ptw; }
  return p;
}

// Per-pass twiddles of Plan<NC>: pass s >= 1, entry (rr, j) =
// exp(-2 pi i rr j / (ns(s) r[s])), laid out [rr][j].
template <int NC>
static void pass_twiddles(std::vector<double> &out) {
  using P = tile::Plan<NC>;
  const long double pi = 3.14159265358979323846264338327950288L;
  out.assign(2 * P::tw_total(), 0.0);
  for (int s = 1; s < P::nst; ++s) {
    const int R = P::r[s], NS = P::ns(s);
    for (int rr = 0; rr < R; ++rr)
      for (int j = 0; j < NS; ++j) {
        const int m = P::tw_off(s) + rr * NS + j;
        const long double a = -2.0L * pi * (long double)(rr * j) / (long double)(NS * R);
        out[2 * m] = (double)cosl(a);
        out[2 * m + 1] = (double)sinl(a);
      }
  }
}

int build_fused_twiddles(LineTables &tab) {
  std::vector<double> h;
  FL_DISPATCH(tab.n / 2, pass_twiddles<NC>(h));
  if (tab.d_ptw) (void)hipFree(tab.d_ptw);
  if (tab.f_ptw) (void)hipFree(tab.f_ptw);
  tab.d_ptw = nullptr;
  tab.f_ptw = nullptr;
  FL_HIP(hipMalloc((void **)&tab.d_ptw, h.size() * sizeof(double)));
  FL_HIP(hipMemcpy(tab.d_ptw, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  std::vector<float> f(h.begin(), h.end());
  FL_HIP(hipMalloc((void **)&tab.f_ptw, f.size() * sizeof(float)));
  FL_HIP(hipMemcpy(tab.f_ptw, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
  return FLUTAS_OK;
}

template <typename T, typename Tin>
int launch_x_fwd(const LineTables &tx, int kind, const Tin *in, long long is1,
                 long long is2, T *Wt, int ny, int nz, hipStream_t s) {
  const LinePlan<T> pl = plan_of<T>(tx);
  constexpr int NT = 256;
  FL_DISPATCH(tx.n / 2, {
    constexpr int TJ = tile_lines<T, NC>();
    dim3 grid((ny + TJ - 1) / TJ, nz);
    hipLaunchKernelGGL((k_x_fwd<T, Tin, NC, TJ, NT>), grid, dim3(NT), 0, s, in, is1,
                       is2, Wt, ny, nz, kind, pl.tw, pl.twh, pl.twq);
  });
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

template <typename T, typename Tout>
int launch_x_bwd(const LineTables &tx, int kind, const T *Wt, Tout *out,
                 long long os1, long long os2, int ny, int nz, T scale,
                 int accumulate, hipStream_t s) {
  const LinePlan<T> pl = plan_of<T>(tx);
  constexpr int NT = 256;
  FL_DISPATCH(tx.n / 2, {
    constexpr int TJ = tile_lines<T, NC>();
    dim3 grid((ny + TJ - 1) / TJ, nz);
    hipLaunchKernelGGL((k_x_bwd<T, Tout, NC, TJ, NT>), grid, dim3(NT), 0, s, Wt, out,
                       os1, os2, ny, nz, kind, scale, accumulate, pl.tw, pl.twh, pl.twq);
  });
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

// P2/P3 plane groups: the two ping-pong tiles of a group stay within
// kTileBytes so that two slab workgroups fit one CU.
template <typename T, int NC>
constexpr int group_planes() {
  constexpr int bytes = 2 * line_stride<T, 2 * NC>() * (int)sizeof(T);
  constexpr int l = kTileBytes / bytes;
  return l >= 16 ? 16 : (l >= 8 ? 8 : (l >= 4 ? 4 : (l >= 2 ? 2 : 1)));
}

int fused_group_planes(int ny, int elem_bytes) {
  int gp = 1;
  const int nc = ny / 2;
  auto f = [&](auto tag) -> int {
    using T = decltype(tag);
    switch (nc) {
      case 16: return group_planes<T, 16>();
      case 24: return group_planes<T, 24>();
      case 32: return group_planes<T, 32>();
      case 48: return group_planes<T, 48>();
      case 64: return group_planes<T, 64>();
      case 96: return group_planes<T, 96>();
      case 128: return group_planes<T, 128>();
      case 192: return group_planes<T, 192>();
      case 256: return group_planes<T, 256>();
      case 384: return group_planes<T, 384>();
      case 512: return group_planes<T, 512>();
      case 1024: return group_planes<T, 1024>();
      default: return 1;
    }
  };
  gp = (elem_bytes == 8) ? f(double()) : f(float());
  return gp;
}

template <typename T>
int launch_yz_fwd(const LineTables &ty, int kind, T *Wt, double *ck, int nx,
                  int nz, const double *lamx, const double *lamy,
                  const double *a, const double *b, const double *c, int pin,
                  hipStream_t s) {
  const LinePlan<T> pl = plan_of<T>(ty);
  constexpr int NT = 256;
  FL_DISPATCH(ty.n / 2, {
    constexpr int GP = group_planes<T, NC>();
    hipLaunchKernelGGL((k_yz_fwd<T, NC, GP, NT>), dim3(nx), dim3(NT), 0, s, Wt, ck, nx,
                       nz, kind, lamx, lamy, a, b, c, pin, pl.tw, pl.twh, pl.twq);
  });
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

template <typename T>
int launch_zy_bwd(const LineTables &ty, int kind, T *Wt, const double *ck,
                  int nx, int nz, const double *lamx, const double *lamy,
                  const double *a, const double *b, const double *c,
                  hipStream_t s) {
  const LinePlan<T> pl = plan_of<T>(ty);
  constexpr int NT = 256;
  FL_DISPATCH(ty.n / 2, {
    constexpr int GP = group_planes<T, NC>();
    hipLaunchKernelGGL((k_zy_bwd<T, NC, GP, NT>), dim3(nx), dim3(NT), 0, s, Wt, ck, nx,
                       nz, kind, lamx, lamy, a, b, c, pl.tw, pl.twh, pl.twq);
  });
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

template int launch_x_fwd<double, double>(const LineTables &, int, const double *, long long, long long, double *, int, int, hipStream_t);
template int launch_x_fwd<float, double>(const LineTables &, int, const double *, long long, long long, float *, int, int, hipStream_t);
template int launch_x_fwd<float, float>(const LineTables &, int, const float *, long long, long long, float *, int, int, hipStream_t);
template int launch_x_bwd<double, double>(const LineTables &, int, const double *, double *, long long, long long, int, int, double, int, hipStream_t);
template int launch_x_bwd<float, double>(const LineTables &, int, const float *, double *, long long, long long, int, int, float, int, hipStream_t);
template int launch_yz_fwd<double>(const LineTables &, int, double *, double *, int, int, const double *, const double *, const double *, const double *, const double *, int, hipStream_t);
template int launch_yz_fwd<float>(const LineTables &, int, float *, double *, int, int, const double *, const double *, const double *, const double *, const double *, int, hipStream_t);
template int launch_zy_bwd<double>(const LineTables &, int, double *, const double *, int, int, const double *, const double *, const double *, const double *, const double *, hipStream_t);
template int launch_zy_bwd<float>(const LineTables &, int, float *, const double *, int, int, const double *, const double *, const double *, const double *, const double *, hipStream_t);

}  // namespace flutas
