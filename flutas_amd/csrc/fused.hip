// Single-problem kernels and host launchers of the fused four-pass
// pressure solve; the device bodies and the pipeline description are in
// fused_body.hpp.
#include "fused_body.hpp"
#include "fused_chunk.hpp"

namespace flutas {

using tile::cx;

// Tile of an x-pass workgroup (one-dimensional grid over ntile tiles per
// plane times the planes).  Workgroups are dealt to the 8 XCDs round robin,
// so the ids a and a + 8 run on one XCD at about the same time; they get
// the neighbouring tiles 2m and 2m + 1 of one plane.  An FP32 tile of 16
// lines covers 64 bytes of each transposed row of the work field, and its
// neighbour the other 64: on one XCD the 128-byte line is fetched (or
// written back) once instead of once per XCD.
struct XTile {
  int j0t, k;  // tile index within the plane, plane
};
__device__ __forceinline__ XTile x_tile_of(int ntile) {
  const int id = blockIdx.x, total = gridDim.x;
  int L = id;
  if ((ntile & 1) == 0 && id < (total & ~15)) {
    const int r = id & 15;
    L = (id & ~15) + 2 * (r & 7) + (r >> 3);
  }
  return {L % ntile, L / ntile};
}

template <typename T, typename Tin, int NC, int TJ, int NT>
__global__ void __launch_bounds__(NT, slab_waves_per_simd(NT))
    k_x_fwd(const Tin *__restrict__ in, long long is1, long long is2,
            T *__restrict__ Wt, WStride ws, int ny, int kind, const T *__restrict__ tw,
            const T *__restrict__ twh, const T *__restrict__ twq,
            double *__restrict__ absmax_part, int *nf_flag, int nf_seq) {
  __shared__ __align__(16) T sm[TJ * line_stride<T, 2 * NC>()];
  __shared__ __align__(16) T stw[x_global_tw<T, NC>() ? 2 : Tw<T, NC>::total];
  const XTile xt = x_tile_of((ny + TJ - 1) / TJ);
  double *slot = absmax_part ? absmax_part + blockIdx.x : nullptr;
  x_fwd_tile<T, Tin, NC, TJ, NT>(in, is1, is2, Wt, ws, ny, kind, tw, twh, twq, slot, sm, stw,
                                 xt.j0t * TJ, xt.k, nf_flag, nf_seq);
}

template <typename T, typename Tout, int NC, int TJ, int NT, bool ACC>
__global__ void __launch_bounds__(NT, slab_waves_per_simd(NT))
    k_x_bwd(const T *__restrict__ Wt, WStride ws, Tout *__restrict__ out, long long os1,
            long long os2, int ny, int kind, T scale, const T *__restrict__ tw,
            const T *__restrict__ twh, const T *__restrict__ twq) {
  __shared__ __align__(16) T sm[TJ * line_stride<T, 2 * NC>()];
  __shared__ __align__(16) T stw[x_global_tw<T, NC>() ? 2 : Tw<T, NC>::total];
  const XTile xt = x_tile_of((ny + TJ - 1) / TJ);
  x_bwd_tile<T, Tout, NC, TJ, NT, ACC>(Wt, ws, out, os1, os2, ny, kind, scale, tw, twh, twq, sm,
                                       stw, xt.j0t * TJ, xt.k);
}

template <typename T, int NC, int GP, int NT, bool PER>
__global__ void __launch_bounds__(NT, slab_waves_per_simd(NT))
    k_yz_fwd(T *__restrict__ Wt, WStride ws, double *__restrict__ ck, int nx, int nz, int nzs,
             int kind, const double *__restrict__ lamx, const double *__restrict__ lamy,
             const double *__restrict__ a, const double *__restrict__ b,
             const double *__restrict__ c, int pin, const T *__restrict__ tw,
             const T *__restrict__ twh, const T *__restrict__ twq) {
  __shared__ __align__(16) T sm[2 * GP * line_stride<T, 2 * NC>()];
  __shared__ __align__(16) T stw[Tw<T, NC>::total];
  yz_fwd_slab<T, NC, GP, NT, PER>(Wt, ws, ck, nx, nz, nzs, kind, lamx, lamy, a, b, c, pin, tw,
                                  twh, twq, sm, stw, blockIdx.x);
}

template <typename T, int NC, int GP, int NT, bool PER>
__global__ void __launch_bounds__(NT, slab_waves_per_simd(NT))
    k_zy_bwd(T *__restrict__ Wt, WStride ws, const double *__restrict__ ck, int nx, int nz,
             int nzs, int kind,
             const double *__restrict__ lamx, const double *__restrict__ lamy,
             const double *__restrict__ a, const double *__restrict__ b,
             const double *__restrict__ c, const T *__restrict__ tw,
             const T *__restrict__ twh, const T *__restrict__ twq) {
  __shared__ __align__(16) T sm[2 * GP * line_stride<T, 2 * NC>()];
  __shared__ __align__(16) T stw[Tw<T, NC>::total];
  zy_bwd_slab<T, NC, GP, NT, PER>(Wt, ws, ck, nx, nz, nzs, kind, lamx, lamy, a, b, c, tw, twh,
                                  twq, sm, stw, blockIdx.x);
}

template <typename T, int NC, int GP, int NT, bool PER>
__global__ void __launch_bounds__(NT, slab_waves_per_simd(NT))
    k_yz_fwd_zc(T *__restrict__ Wt, WStride ws, ZChunks zc, int nx, int nz, int msys, int kind,
                const double *__restrict__ lamx, const double *__restrict__ lamy,
                const double *__restrict__ a, const double *__restrict__ b,
                const double *__restrict__ c, int pin, const T *__restrict__ tw,
                const T *__restrict__ twh, const T *__restrict__ twq) {
  __shared__ __align__(16) T sm[2 * GP * line_stride<T, 2 * NC>()];
  __shared__ __align__(16) T stw[Tw<T, NC>::total];
  int i, ch;
  zc_slab_of(nx, zc.nchunk, &i, &ch);
  yz_fwd_chunk<T, NC, GP, NT, PER>(Wt, ws, zc, nz, msys, kind, lamx, lamy, a, b, c, pin, tw, twh,
                                   twq, sm, stw, i, ch);
}

template <typename T, int NC, int GP, int NT, bool PER>
__global__ void __launch_bounds__(NT, slab_waves_per_simd(NT))
    k_zy_bwd_zc(T *__restrict__ Wt, WStride ws, ZChunks zc, int nx, int nz, int msys, int kind,
                const double *__restrict__ lamx, const double *__restrict__ lamy,
                const double *__restrict__ a, const double *__restrict__ b,
                const double *__restrict__ c, int pin, const T *__restrict__ tw,
                const T *__restrict__ twh, const T *__restrict__ twq) {
  __shared__ __align__(16) T sm[2 * GP * line_stride<T, 2 * NC>()];
  __shared__ __align__(16) T stw[Tw<T, NC>::total];
  int i, ch;
  zc_slab_of(nx, zc.nchunk, &i, &ch);
  zy_bwd_chunk<T, NC, GP, NT, PER>(Wt, ws, zc, nz, msys, kind, lamx, lamy, a, b, c, pin, tw, twh,
                                   twq, sm, stw, i, ch);
}

// Tables of the z-chunked passes, one thread per column (i, j), with the
// operations of yz_fwd_chunk / zy_bwd_chunk (so d is bit for bit the
// passes' and the reference's): walking up, d, m (periodic: q2) before
// every plane group and the chunk sums M, SW, N; for periodic z then
// walking down group by group (factors recomputed from the stored values),
// the second solution p2 at every chunk start and the closing row's
// denominator (b(m) + lam) + c(m) p2(0) + a(m) p2(m-1).
// The pivot reciprocals are those of the chain the tables serve (zrcp<T>).
constexpr int kZcMaxGroup = 16;
template <typename T>
__global__ void __launch_bounds__(256)
    k_zchunk_tables(double *__restrict__ ckd, double *__restrict__ ckm, double *__restrict__ ckq,
                    double *__restrict__ tab, double *__restrict__ den, int nx, int ny, int nz,
                    int msys, int gp, int cpl, int nchunk, int per, int pin,
                    const double *__restrict__ lamx, const double *__restrict__ lamy,
                    const double *__restrict__ a, const double *__restrict__ b,
                    const double *__restrict__ c) {
#pragma clang fp contract(off)
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= (long long)nx * ny) return;
  const int i = (int)(col / ny), j = (int)(col - (long long)i * ny);
  const double lam = lamx[i] + lamy[j];
  const int ngrp = (nz + gp - 1) / gp;
  const int TS = zc_tab_slots(per != 0);
  double dp = 0.0, mp = 1.0, qp = 0.0, w = 1.0, sw = 0.0, ml = 0.0;
  for (int k = 0; k < nz; ++k) {
    const int ch = min(k / cpl, nchunk - 1);
    if (k == ch * cpl) {
      mp = 1.0;
      w = 1.0;
      sw = 0.0;
      ml = 0.0;
    }
    if (k % gp == 0) {
      const long long o = ((long long)i * ngrp + k / gp) * ny + j;
      ckd[o] = dp;
      ckm[o] = mp;
      if (per) ckq[o] = qp;
    }
    if (k < msys) {
      const double r2 = per ? ((k == msys - 1) ? -c[msys - 1] : ((k == 0) ? -a[0] : 0.0)) : 0.0;
      double dl = 0.0;
      if (k == 0) {
        const double z = zrcp<T>(b[0] + lam);
        dp = c[0] * z;
        mp = 0.0;
        qp = r2 * z;
        dl = dp;
      } else if (k < msys - 1) {
        const double z = zrcp<T>((b[k] + lam) - a[k] * dp);
        dp = c[k] * z;
        mp = (-a[k] * mp) * z;
        qp = (r2 - a[k] * qp) * z;
        dl = dp;
      } else {
        const double z = (b[k] + lam) - a[k] * dp;
        mp = (z != 0.0 && (per || !(pin && lam == 0.0))) ? (-a[k] * mp) / z : 0.0;
        qp = (z != 0.0) ? (r2 - a[k] * qp) / z : 0.0;
      }
      sw = sw + w * mp;
      w = -(dl * w);
      ml = mp;
    }
    const int ke = (ch == nchunk - 1) ? nz : (ch + 1) * cpl;
    if (k == ke - 1) {
      double *t = tab + ((long long)i * nchunk + ch) * TS * ny + j;
      t[0] = ml;
      t[ny] = sw;
      t[2 * ny] = w;
    }
  }
  if (!per) return;
  const int m = msys;  // nz - 1
  double p2n = 0.0, p20 = 0.0, p2m1 = 0.0;
  for (int g = ngrp - 1; g >= 0; --g) {
    const int k0 = g * gp;
    const long long o = ((long long)i * ngrp + g) * ny + j;
    double dd[kZcMaxGroup], qq[kZcMaxGroup];
    double d2 = ckd[o], q2 = ckq[o];
#pragma unroll
    for (int l = 0; l < kZcMaxGroup; ++l) {
      const int k = k0 + l;
      if (l < gp && k < m) {
        const double r2 = (k == m - 1) ? -c[m - 1] : ((k == 0) ? -a[0] : 0.0);
        if (k == 0) {
          const double z = zrcp<T>(b[0] + lam);
          d2 = c[0] * z;
          q2 = r2 * z;
        } else if (k < m - 1) {
          const double z = zrcp<T>((b[k] + lam) - a[k] * d2);
          d2 = c[k] * z;
          q2 = (r2 - a[k] * q2) * z;
        } else {
          const double z = (b[k] + lam) - a[k] * d2;
          q2 = (z != 0.0) ? (r2 - a[k] * q2) / z : 0.0;
        }
      }
      dd[l] = d2;
      qq[l] = q2;
    }
#pragma unroll
    for (int l = kZcMaxGroup - 1; l >= 0; --l) {
      const int k = k0 + l;
      if (l >= gp || k >= m) continue;
      const double p2 = (k == m - 1) ? qq[l] : qq[l] - dd[l] * p2n;
      p2n = p2;
      if (k == m - 1) p2m1 = p2;
      if (k == 0) p20 = p2;
      if (k > 0 && k % cpl == 0 && k / cpl < nchunk)
        tab[(((long long)i * nchunk + k / cpl - 1) * TS + 3) * ny + j] = p2;
    }
  }
  den[(long long)i * ny + j] = ((b[m] + lam) + c[m] * p20) + a[m] * p2m1;
}

// ---------------------------------------------------------------- host ---

bool fused_supported(int n) {
  switch (n / 2) {
    case 16: case 24: case 32: case 40: case 48: case 50: case 64: case 80: case 96:
    case 100: case 128: case 160: case 192: case 200: case 256: case 320: case 384:
    case 400: case 512: case 1024:
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
    case 40: { constexpr int NC = 40; __VA_ARGS__; } break;   \
    case 50: { constexpr int NC = 50; __VA_ARGS__; } break;   \
    case 100: { constexpr int NC = 100; __VA_ARGS__; } break; \
    case 200: { constexpr int NC = 200; __VA_ARGS__; } break; \
    case 400: { constexpr int NC = 400; __VA_ARGS__; } break; \
    case 48: { constexpr int NC = 48; __VA_ARGS__; } break;   \
    case 64: { constexpr int NC = 64; __VA_ARGS__; } break;   \
    case 80: { constexpr int NC = 80; __VA_ARGS__; } break;   \
    case 96: { constexpr int NC = 96; __VA_ARGS__; } break;   \
    case 128: { constexpr int NC = 128; __VA_ARGS__; } break; \
    case 160: { constexpr int NC = 160; __VA_ARGS__; } break; \
    case 192: { constexpr int NC = 192; __VA_ARGS__; } break; \
    case 256: { constexpr int NC = 256; __VA_ARGS__; } break; \
    case 320: { constexpr int NC = 320; __VA_ARGS__; } break; \
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

// The tables a kind's tile kernels stage as their split and phase tables:
// the DCT-IV/DST-IV pre/post twiddles for REDFT11/RODFT11.
template <typename T>
static void kind_tables(const LinePlan<T> &pl, int kind, const T **th, const T **tq) {
  if (kind == K_REDFT11 || kind == K_RODFT11) {
    *th = pl.t4;
    *tq = pl.t4 + 2 * (pl.N + 1);
  } else {
    *th = pl.twh;
    *tq = pl.twq;
  }
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
                 long long is2, T *Wt, WStride ws, int ny, int nz,
                 double *absmax_part, int *nf_flag, int nf_seq, hipStream_t s) {
  const LinePlan<T> pl = plan_of<T>(tx);
  const T *th, *tq;
  kind_tables(pl, kind, &th, &tq);
  FL_DISPATCH(tx.n / 2, {
    constexpr int NT = tile_threads<T, NC>();
    constexpr int TJ = tile_lines<T, NC>();
    dim3 grid(((ny + TJ - 1) / TJ) * nz);
    hipLaunchKernelGGL((k_x_fwd<T, Tin, NC, TJ, NT>), grid, dim3(NT), 0, s, in, is1,
                       is2, Wt, ws, ny, kind, pl.tw, th, tq, absmax_part, nf_flag,
                       nf_seq);
  });
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

template <typename T, typename Tout>
int launch_x_bwd(const LineTables &tx, int kind, const T *Wt, WStride ws,
                 Tout *out, long long os1, long long os2, int ny, int nz,
                 T scale, int accumulate, hipStream_t s) {
  const LinePlan<T> pl = plan_of<T>(tx);
  const T *th, *tq;
  kind_tables(pl, kind, &th, &tq);
  FL_DISPATCH(tx.n / 2, {
    constexpr int NT = tile_threads<T, NC>();
    constexpr int TJ = tile_lines<T, NC>();
    dim3 grid(((ny + TJ - 1) / TJ) * nz);
    if (accumulate)
      hipLaunchKernelGGL((k_x_bwd<T, Tout, NC, TJ, NT, true>), grid, dim3(NT), 0, s, Wt,
                         ws, out, os1, os2, ny, kind, scale, pl.tw, th, tq);
    else
      hipLaunchKernelGGL((k_x_bwd<T, Tout, NC, TJ, NT, false>), grid, dim3(NT), 0, s, Wt,
                         ws, out, os1, os2, ny, kind, scale, pl.tw, th, tq);
  });
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

int x_fwd_blocks(const LineTables &tx, int elem_bytes, int ny, int nz) {
  // tile_lines<T, NC>() evaluated at run time
  const int ls = tx.n + (elem_bytes == 8 ? 2 : 4);
  const int l = kTileBytes / (ls * elem_bytes);
  int tj = l >= 16 ? 16 : (l >= 8 ? 8 : (l >= 4 ? 4 : (l >= 2 ? 2 : 1)));
  if (elem_bytes == 8 && tx.n == 1024) tj = 8;
  if (elem_bytes == 8 && tx.n == 400) tj = 16;
  return ((ny + tj - 1) / tj) * nz;
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
      case 40: return group_planes<T, 40>();
      case 50: return group_planes<T, 50>();
      case 100: return group_planes<T, 100>();
      case 200: return group_planes<T, 200>();
      case 400: return group_planes<T, 400>();
      case 48: return group_planes<T, 48>();
      case 64: return group_planes<T, 64>();
      case 80: return group_planes<T, 80>();
      case 96: return group_planes<T, 96>();
      case 128: return group_planes<T, 128>();
      case 160: return group_planes<T, 160>();
      case 192: return group_planes<T, 192>();
      case 256: return group_planes<T, 256>();
      case 320: return group_planes<T, 320>();
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
int launch_yz_fwd(const LineTables &ty, int kind, T *Wt, WStride ws,
                  double *ck, int nx, int nz, int nzs, const double *lamx,
                  const double *lamy, const double *a, const double *b,
                  const double *c, int pin, int periodic, hipStream_t s) {
  const LinePlan<T> pl = plan_of<T>(ty);
  const T *th, *tq;
  kind_tables(pl, kind, &th, &tq);
  FL_DISPATCH(ty.n / 2, {
    constexpr int GP = group_planes<T, NC>();
    constexpr int NTP = slab_threads<T, NC, true>();
    constexpr int NTN = slab_threads<T, NC, false>();
    if (periodic)
      hipLaunchKernelGGL((k_yz_fwd<T, NC, GP, NTP, true>), dim3(nx), dim3(NTP), 0, s, Wt, ws,
                         ck, nx, nz, nzs, kind, lamx, lamy, a, b, c, pin, pl.tw, th,
                         tq);
    else
      hipLaunchKernelGGL((k_yz_fwd<T, NC, GP, NTN, false>), dim3(nx), dim3(NTN), 0, s, Wt, ws,
                         ck, nx, nz, nzs, kind, lamx, lamy, a, b, c, pin, pl.tw, th,
                         tq);
  });
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

template <typename T>
int launch_zy_bwd(const LineTables &ty, int kind, T *Wt, WStride ws,
                  const double *ck, int nx, int nz, int nzs, const double *lamx,
                  const double *lamy, const double *a, const double *b,
                  const double *c, int periodic, hipStream_t s) {
  const LinePlan<T> pl = plan_of<T>(ty);
  const T *th, *tq;
  kind_tables(pl, kind, &th, &tq);
  FL_DISPATCH(ty.n / 2, {
    constexpr int GP = group_planes<T, NC>();
    constexpr int NTP = slab_threads<T, NC, true>();
    constexpr int NTN = slab_threads<T, NC, false>();
    if (periodic)
      hipLaunchKernelGGL((k_zy_bwd<T, NC, GP, NTP, true>), dim3(nx), dim3(NTP), 0, s, Wt, ws,
                         ck, nx, nz, nzs, kind, lamx, lamy, a, b, c, pl.tw, th, tq);
    else
      hipLaunchKernelGGL((k_zy_bwd<T, NC, GP, NTN, false>), dim3(nx), dim3(NTN), 0, s, Wt, ws,
                         ck, nx, nz, nzs, kind, lamx, lamy, a, b, c, pl.tw, th, tq);
  });
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

// Workgroups of the chunked P2 (*r2) and P3 (*r3) slab kernels that one CU
// holds at once, from the runtime's occupancy calculator (it counts the
// kernels' registers, AGPRs included, and their LDS).
template <typename T>
static int slab_residency(int ny, int periodic, int *r2, int *r3) {
  FL_DISPATCH(ny / 2, {
    constexpr int GP = group_planes<T, NC>();
    constexpr int NTP = slab_threads<T, NC, true>();
    constexpr int NTN = slab_threads<T, NC, false>();
    if (periodic) {
      FL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
          r2, reinterpret_cast<const void *>(&k_yz_fwd_zc<T, NC, GP, NTP, true>), NTP, 0));
      FL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
          r3, reinterpret_cast<const void *>(&k_zy_bwd_zc<T, NC, GP, NTP, true>), NTP, 0));
    } else {
      FL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
          r2, reinterpret_cast<const void *>(&k_yz_fwd_zc<T, NC, GP, NTN, false>), NTN, 0));
      FL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
          r3, reinterpret_cast<const void *>(&k_zy_bwd_zc<T, NC, GP, NTN, false>), NTN, 0));
    }
  });
  return FLUTAS_OK;
}

int fused_slab_residency(int ny, int elem_bytes, int periodic, int *r2, int *r3) {
  *r2 = *r3 = 0;
  return elem_bytes == 8 ? slab_residency<double>(ny, periodic, r2, r3)
                         : slab_residency<float>(ny, periodic, r2, r3);
}

size_t zchunk_doubles(int nx, int ny, int ngrp, int nchunk, bool per) {
  const size_t col = (size_t)nx * ny;
  return col * (3 * (size_t)ngrp + (size_t)nchunk * zc_tab_slots(per) + 1 +
                (size_t)nchunk * zc_side_slots(per));
}

int build_zchunk_tables(double *mem, ZChunks *zc, int nx, int ny, int nz, int msys, int gp,
                        int cpl, int nchunk, bool per, int pin, int elem_bytes,
                        const double *lamx, const double *lamy, const double *a,
                        const double *b, const double *c, hipStream_t s) {
  if (gp < 1 || gp > kZcMaxGroup || cpl % gp != 0 || nchunk < 2 || nchunk > kMaxZChunks ||
      (nchunk - 1) * cpl > nz - 2)
    return fail(FLUTAS_EINVAL, "z chunks: bad chunk shape");
  const int ngrp = (nz + gp - 1) / gp;
  const size_t col = (size_t)nx * ny;
  double *ckd = mem, *ckm = ckd + col * ngrp, *ckq = ckm + col * ngrp;
  double *tab = ckq + col * ngrp;
  double *den = tab + col * nchunk * zc_tab_slots(per);
  double *side = den + col;
  const int threads = 256;
  const long long blocks = ((long long)col + threads - 1) / threads;
  if (elem_bytes == 8)
    hipLaunchKernelGGL(k_zchunk_tables<double>, dim3((unsigned)blocks), dim3(threads), 0, s, ckd,
                       ckm, ckq, tab, den, nx, ny, nz, msys, gp, cpl, nchunk, per ? 1 : 0, pin,
                       lamx, lamy, a, b, c);
  else
    hipLaunchKernelGGL(k_zchunk_tables<float>, dim3((unsigned)blocks), dim3(threads), 0, s, ckd,
                       ckm, ckq, tab, den, nx, ny, nz, msys, gp, cpl, nchunk, per ? 1 : 0, pin,
                       lamx, lamy, a, b, c);
  FL_HIP(hipGetLastError());
  *zc = ZChunks{nchunk, cpl, ngrp, ckd, ckm, ckq, tab, den, side};
  return FLUTAS_OK;
}

template <typename T>
int launch_yz_fwd_zc(const LineTables &ty, int kind, T *Wt, WStride ws, const ZChunks &zc,
                     int nx, int nz, int msys, const double *lamx, const double *lamy,
                     const double *a, const double *b, const double *c, int pin, int periodic,
                     hipStream_t s) {
  const LinePlan<T> pl = plan_of<T>(ty);
  const T *th, *tq;
  kind_tables(pl, kind, &th, &tq);
  const dim3 grid(nx * zc.nchunk);
  FL_DISPATCH(ty.n / 2, {
    constexpr int GP = group_planes<T, NC>();
    constexpr int NTP = slab_threads<T, NC, true>();
    constexpr int NTN = slab_threads<T, NC, false>();
    if (zc.ngrp != (nz + GP - 1) / GP) return fail(FLUTAS_EINVAL, "z chunks: plane groups");
    if (periodic)
      hipLaunchKernelGGL((k_yz_fwd_zc<T, NC, GP, NTP, true>), grid, dim3(NTP), 0, s, Wt, ws, zc,
                         nx, nz, msys, kind, lamx, lamy, a, b, c, pin, pl.tw, th, tq);
    else
      hipLaunchKernelGGL((k_yz_fwd_zc<T, NC, GP, NTN, false>), grid, dim3(NTN), 0, s, Wt, ws, zc,
                         nx, nz, msys, kind, lamx, lamy, a, b, c, pin, pl.tw, th, tq);
  });
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

template <typename T>
int launch_zy_bwd_zc(const LineTables &ty, int kind, T *Wt, WStride ws, const ZChunks &zc,
                     int nx, int nz, int msys, const double *lamx, const double *lamy,
                     const double *a, const double *b, const double *c, int pin, int periodic,
                     hipStream_t s) {
  const LinePlan<T> pl = plan_of<T>(ty);
  const T *th, *tq;
  kind_tables(pl, kind, &th, &tq);
  const dim3 grid(nx * zc.nchunk);
  FL_DISPATCH(ty.n / 2, {
    constexpr int GP = group_planes<T, NC>();
    constexpr int NTP = slab_threads<T, NC, true>();
    constexpr int NTN = slab_threads<T, NC, false>();
    if (zc.ngrp != (nz + GP - 1) / GP) return fail(FLUTAS_EINVAL, "z chunks: plane groups");
    if (periodic)
      hipLaunchKernelGGL((k_zy_bwd_zc<T, NC, GP, NTP, true>), grid, dim3(NTP), 0, s, Wt, ws, zc,
                         nx, nz, msys, kind, lamx, lamy, a, b, c, pin, pl.tw, th, tq);
    else
      hipLaunchKernelGGL((k_zy_bwd_zc<T, NC, GP, NTN, false>), grid, dim3(NTN), 0, s, Wt, ws, zc,
                         nx, nz, msys, kind, lamx, lamy, a, b, c, pin, pl.tw, th, tq);
  });
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

template int launch_x_fwd<double, double>(const LineTables &, int, const double *, long long, long long, double *, WStride, int, int, double *, int *, int, hipStream_t);
template int launch_x_fwd<float, double>(const LineTables &, int, const double *, long long, long long, float *, WStride, int, int, double *, int *, int, hipStream_t);
template int launch_x_fwd<float, float>(const LineTables &, int, const float *, long long, long long, float *, WStride, int, int, double *, int *, int, hipStream_t);
template int launch_x_bwd<double, double>(const LineTables &, int, const double *, WStride, double *, long long, long long, int, int, double, int, hipStream_t);
template int launch_x_bwd<float, float>(const LineTables &, int, const float *, WStride, float *, long long, long long, int, int, float, int, hipStream_t);
template int launch_x_bwd<float, double>(const LineTables &, int, const float *, WStride, double *, long long, long long, int, int, float, int, hipStream_t);
template int launch_yz_fwd<double>(const LineTables &, int, double *, WStride, double *, int, int, int, const double *, const double *, const double *, const double *, const double *, int, int, hipStream_t);
template int launch_yz_fwd<float>(const LineTables &, int, float *, WStride, double *, int, int, int, const double *, const double *, const double *, const double *, const double *, int, int, hipStream_t);
template int launch_zy_bwd<double>(const LineTables &, int, double *, WStride, const double *, int, int, int, const double *, const double *, const double *, const double *, const double *, int, hipStream_t);
template int launch_zy_bwd<float>(const LineTables &, int, float *, WStride, const double *, int, int, int, const double *, const double *, const double *, const double *, const double *, int, hipStream_t);

template int launch_yz_fwd_zc<double>(const LineTables &, int, double *, WStride, const ZChunks &, int, int, int, const double *, const double *, const double *, const double *, const double *, int, int, hipStream_t);
template int launch_yz_fwd_zc<float>(const LineTables &, int, float *, WStride, const ZChunks &, int, int, int, const double *, const double *, const double *, const double *, const double *, int, int, hipStream_t);
template int launch_zy_bwd_zc<double>(const LineTables &, int, double *, WStride, const ZChunks &, int, int, int, const double *, const double *, const double *, const double *, const double *, int, int, hipStream_t);
template int launch_zy_bwd_zc<float>(const LineTables &, int, float *, WStride, const ZChunks &, int, int, int, const double *, const double *, const double *, const double *, const double *, int, int, hipStream_t);

}  // namespace flutas
