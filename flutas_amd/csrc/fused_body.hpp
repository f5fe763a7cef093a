// Device bodies of the fused four-pass pressure solve (the hot path),
// shared by the single-problem kernels (fused.hip) and the batched kernels
// (batch.hip).
//
// The work field W holds (i, j, k) with j fastest (y-lines contiguous) at
// the strides of a WStride.  The FP64 single solve keeps it in blocks of a
// few x-slabs (fused_layout_fp64): within a block the planes of the slabs
// follow each other, so P2/P3 stream nearly contiguous memory while the
// transposed accesses of P1/P4 stay within few distant regions.  The FP32
// chain and the batched solves use the plane-major layout.
//
//   P1 x_fwd_tile   p (halo layout, x-lines) -> x r2r -> W (transposed
//                   store)
//   P2 yz_fwd_slab  per x-slab i, k ascending: y r2r of each plane, then
//                   the Thomas forward elimination step for all (i, j)
//                   columns
//   P3 zy_bwd_slab  per x-slab i, k descending: Thomas back substitution,
//                   then the inverse y r2r of each plane
//   P4 x_bwd_tile   W (transposed load) -> inverse x r2r * normfft -> p
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
#pragma once

#include "tile_fft.hpp"

namespace flutas {

using tile::cx;

// LDS line stride: two extra elements per line so that the transposed
// accesses of P1/P4 (same element of consecutive lines) spread over banks.
template <typename T, int n>
constexpr int line_stride() {
  return n + (sizeof(T) == 8 ? 2 : 4);
}

// Stores of the passes' outputs with the non-temporal hint: a pass writes
// its whole field before the next pass reads it, so nothing is re-read from
// the L2 soon (measured at 512^3: the FP64 solve 1.87 -> 1.84 ms, P1 0.41
// -> 0.39 ms).  The FP32 x passes keep plain stores (see x_fwd_tile).
template <typename T>
__device__ __forceinline__ void st_nt(T v, T *p) {
  __builtin_nontemporal_store(v, p);
}
template <typename T>
__device__ __forceinline__ void st_nt2(T x, T y, T *p) {
  typedef T vt2 __attribute__((ext_vector_type(2)));
  vt2 w;
  w.x = x;
  w.y = y;
  __builtin_nontemporal_store(w, reinterpret_cast<vt2 *>(p));
}

// Lines per tile: the largest of 16, 8, 4, 2, 1 whose LDS tile fits
// kTileBytes (two tiles per CU fit the 160 KiB of gfx950 LDS).  FP64
// lines of 1024 points take 8 lines (65.7 KB) and read their twiddles from
// global memory (x_global_tw) instead of staging 25.6 KB of them, so two
// tiles share a CU and each transposed row still covers a whole 64-byte
// half line (1024x1024x512 FP64: P1 2.42 -> 2.13 ms, P4 2.48 -> 2.36 ms,
// solve 9.58 -> 9.12 ms, against 16 staged-twiddle lines at one tile per
// CU; 16 lines had measured 9.18 ms against 8 staged-twiddle lines at
// 10.40 ms, also one per CU).  FP64 lines of 400 points take 16 lines
// (400^3 solve 1.37 -> 1.29 ms, P1 0.24 -> 0.20 ms); for 640- and
// 768-point lines 16-line tiles measured slower (P1 0.53 -> 0.62 ms,
// 0.58 -> 0.71 ms).
constexpr int kTileBytes = 72 * 1024;
template <typename T, int NC>
constexpr int tile_lines() {
  if (NC == 200 && sizeof(T) == 8) return 16;
  if (NC == 512 && sizeof(T) == 8) return 8;
  constexpr int bytes = line_stride<T, 2 * NC>() * (int)sizeof(T);
  constexpr int l = kTileBytes / bytes;
  return l >= 16 ? 16 : (l >= 8 ? 8 : (l >= 4 ? 4 : (l >= 2 ? 2 : 1)));
}

// The P1/P4 tiles of 1024-point FP64 lines read the twiddle tables from
// global memory (a tile has no prefetch for them to drain), which leaves
// LDS for two tiles per CU.
template <typename T, int NC>
constexpr bool x_global_tw() { return NC == 512 && sizeof(T) == 8; }

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
// The staged tables of a previous stage_twiddles call into s.
template <typename T, int NC>
__device__ __forceinline__ Tw<T, NC> twiddle_view(T *s) {
  using W = Tw<T, NC>;
  return Tw<T, NC>{s, s + W::ntw, s + W::ntw + W::nth};
}
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

// max(m, |v|) with NaN propagation (a NaN in either stays NaN)
__device__ __forceinline__ double nan_max(double m, double v) {
  const double av = fabs(v);
  return (v != v) ? v : ((av > m) ? av : m);
}

// Workgroup max of one value per thread (NaN propagating), stored by
// thread 0.  Contains a barrier; every thread must call it.
template <int NT>
__device__ __forceinline__ void block_max_store(double am, double *dst) {
  __shared__ double wmax[NT / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) am = nan_max(am, __shfl_xor(am, off));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = am;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wmax[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) m = nan_max(m, wmax[w]);
    *dst = m;
  }
}

// Minimum waves per SIMD of a workgroup size: two 512-thread workgroups
// per CU (4 waves per SIMD, at most 128 VGPRs each).  For 512-point lines
// the LDS of two tiles or slabs (with their staged twiddles) fits a CU,
// and so do two 1024-point FP64 tiles (8 lines, twiddles from global
// memory); a 1024-point FP64 slab group (2 x 8 planes of 1026 doubles,
// 131 KB, plus 25.6 KB of staged twiddles) fills one CU's LDS, so those
// run one workgroup (two waves per SIMD) per CU.
constexpr int slab_waves_per_simd(int nt) { return nt >= 512 ? 4 : 1; }

// Threads of a P1/P4 tile workgroup: 512 for lines of 512 and 1024 points
// (measured faster than 256: P4 0.49 -> 0.45 ms at 512^3; P1 3.29 ->
// 2.70 ms and P4 3.91 -> 3.35 ms at 1024x1024x512 in FP64, the FP32 P1
// 1.59 -> 1.34 ms), 256 otherwise (faster for the short lines of batched
// problems).
template <typename T, int NC>
constexpr int tile_threads() {
  if (2 * NC == 512) return 512;
  if (2 * NC == 1024) return 512;
  if (NC == 200 && sizeof(T) == 8) return 512;  // 16-line tiles (tile_lines)
  return 256;
}

// ---------------------------------------------------------------- P1 -----
// One workgroup: TJ consecutive x-lines (j0..j0+TJ-1) at one k.
// in: element (i, j, k) at in[j*is1 + k*is2 + i] (i, j, k 0-based).
// Wt: element (i, j, k) at Wt[ws.ioff(i) + k*ws.wk + j] (WStride).
// The tile body (also used by the batched kernels): sm holds TJ lines of
// line_stride<T, nx>() elements, stw the staged twiddles; absmax_slot, when
// not null, receives the tile's max |in|; nf_flag, when not null, receives
// nf_seq from every lane that loads a NaN or an infinity (a plain store of
// the same value, no ordering needed).
template <typename T, typename Tin, int NC, int TJ, int NT>
__device__ __forceinline__ void x_fwd_tile(const Tin *__restrict__ in, long long is1,
                                           long long is2, T *__restrict__ Wt, WStride ws,
                                           int ny,
                                           int kind, const T *__restrict__ tw,
                                           const T *__restrict__ twh,
                                           const T *__restrict__ twq, double *absmax_slot,
                                           T *sm, T *stw, int j0, int k,
                                           int *nf_flag = nullptr, int nf_seq = 0,
                                           bool stage_tw = true) {
  constexpr int n = 2 * NC;  // nx
  constexpr int LS = line_stride<T, n>();
  constexpr int LD = (TJ * n + NT - 1) / NT;
  const int nl = min(TJ, ny - j0);
  // d4: REDFT11 / RODFT11 (DCT-IV / DST-IV, tile_fft.hpp)
  const bool d4 = tile::is_d4(kind);
  const bool trig = tile::is_trig(kind) && !d4;
  const bool sine = tile::is_sine(kind) || kind == K_RODFT11;
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
    if (trig || d4) {
      pos = trig ? tile::mk_pos(e, n) : tile::d4_pos(e, n);
      if (sine && (e & 1)) x = -x;
    }
    sm[l * LS + pos] = x;
  }
  const Tw<T, NC> W = x_global_tw<T, NC>() ? Tw<T, NC>{tw, twh, twq}
      : (stage_tw ? stage_twiddles<T, NC, NT>(stw, tw, twh, twq) : twiddle_view<T, NC>(stw));
  if (absmax_slot) {
    // max |in| over the tile (the clamped loads only repeat tile entries),
    // NaN propagating
    double am = 0.0;
#pragma unroll
    for (int q = 0; q < LD; ++q) am = nan_max(am, (double)r[q]);
    block_max_store<NT>(am, absmax_slot);
  }
  if (nf_flag) {
    bool bad = false;
#pragma unroll
    for (int q = 0; q < LD; ++q) bad |= !(fabs((double)r[q]) <= 1.7976931348623157e308);
    if (bad) *nf_flag = nf_seq;
  }
  __syncthreads();
  if (d4) {
    tile::d4_pre<T, NC, TJ, NT, LS, false>(sm, W.th);
    __syncthreads();
  }
  tile::cfft<T, -1, NC, TJ, NT, LS / 2>(reinterpret_cast<cx<T> *>(sm), W.tw);
  if (d4) tile::d4_post<T, NC, TJ, NT, LS, false>(sm, W.tq, sine);
  else tile::post_fwd<T, NC, TJ, NT, LS>(sm, kind, W.th, W.tq);
  // transposed store
  T *plane = Wt + (long long)k * ws.wk;
#pragma unroll
  for (int q = 0; q < LD; ++q) {
    const int t = threadIdx.x + q * NT;
    const int l = t % TJ, i = t / TJ;
    if (t < TJ * n && l < nl) {
      // FP64 tiles write whole 128-byte rows; the 64-byte halves of FP32
      // rows are merged in L2 with the neighbouring tile's (same XCD)
      if constexpr (sizeof(T) == 8) st_nt(sm[l * LS + i], plane + ws.ioff(i) + j0 + l);
      else plane[ws.ioff(i) + j0 + l] = sm[l * LS + i];
    }
  }
}

// ---------------------------------------------------------------- P4 -----
// Inverse of P1: out(i, j, k) = scale * r2r^-1 (x) (+ out when ACC).
template <typename T, typename Tout, int NC, int TJ, int NT, bool ACC>
__device__ __forceinline__ void x_bwd_tile(const T *__restrict__ Wt, WStride ws,
                                           Tout *__restrict__ out,
                                           long long os1, long long os2, int ny, int kind,
                                           T scale, const T *__restrict__ tw,
                                           const T *__restrict__ twh,
                                           const T *__restrict__ twq, T *sm, T *stw, int j0,
                                           int k, bool stage_tw = true) {
  constexpr int n = 2 * NC;
  constexpr int LS = line_stride<T, n>();
  constexpr int LD = (TJ * n + NT - 1) / NT;
  const int nl = min(TJ, ny - j0);
  // d4: REDFT11 / RODFT11 (DCT-IV / DST-IV, tile_fft.hpp)
  const bool d4 = tile::is_d4(kind);
  const bool trig = tile::is_trig(kind) && !d4;
  const bool sine = tile::is_sine(kind) || kind == K_RODFT11;
  const T *plane = Wt + (long long)k * ws.wk;
  T r[LD];
#pragma unroll
  for (int q = 0; q < LD; ++q) {
    const int t = min((int)threadIdx.x + q * NT, TJ * n - 1);
    const int l = t % TJ, i = t / TJ;
    r[q] = plane[ws.ioff(i) + min(j0 + l, ny - 1)];
  }
  // the previous values of an accumulated output are loaded with the tile
  // so that their latency overlaps the transform
  // FP32 1024-point tiles load the accumulated values at the store: 64 of
  // them per lane would not fit beside the tile's registers at 512 threads
  // (refined 1024x1024x512: this pass 3.34 -> 2.63 ms with 512 threads
  // against the preloading 256-thread form)
  constexpr bool LATE = ACC && NC == 512 && sizeof(T) == 4;
  Tout prev[(ACC && !LATE) ? LD : 1];
  if constexpr (ACC && !LATE) {
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
    if (t < TJ * n) {
      const int i = t / TJ;
      T x = r[q];
      int pos = i;
      if (d4) {
        pos = tile::d4_pos(i, n);
        if (sine && (i & 1)) x = -x;
      }
      sm[(t % TJ) * LS + pos] = x;
    }
  }
  const Tw<T, NC> W = x_global_tw<T, NC>() ? Tw<T, NC>{tw, twh, twq}
      : (stage_tw ? stage_twiddles<T, NC, NT>(stw, tw, twh, twq) : twiddle_view<T, NC>(stw));
  __syncthreads();
  if (d4) {
    tile::d4_pre<T, NC, TJ, NT, LS, true>(sm, W.th);
    __syncthreads();
  } else {
    tile::pre_inv<T, NC, TJ, NT, LS>(sm, kind, W.th, W.tq);
  }
  tile::cfft<T, +1, NC, TJ, NT, LS / 2>(reinterpret_cast<cx<T> *>(sm), W.tw);
  if (d4) tile::d4_post<T, NC, TJ, NT, LS, true>(sm, W.tq, sine);
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
    if constexpr (LATE) *o = (Tout)((double)*o + (double)v);
    else if constexpr (ACC) *o = (Tout)((double)prev[q] + (double)v);
    else if constexpr (sizeof(T) == 8) st_nt((Tout)v, o);
    else *o = (Tout)v;
  }
}

// Two adjacent elements (16 bytes in FP64, 8 in FP32).
template <typename T>
struct Pair;
template <>
struct Pair<double> {
  using type = double2;
};
template <>
struct Pair<float> {
  using type = float2;
};

// The CPT adjacent columns j = CPT*tid + u of a slab line, stored/loaded in
// pairs when CPT is even (one wide access per pair), else one by one.
// Loads of columns past n re-read column n-1 (branch free).
template <typename T, int CPT, int NT>
__device__ __forceinline__ void store_cols(T *__restrict__ line, const T *v, int n) {
  const int j0 = CPT * threadIdx.x;
  if constexpr (CPT % 2 == 0) {
#pragma unroll
    for (int u = 0; u < CPT; u += 2)
      if (j0 + u < n) {
        typename Pair<T>::type w;
        w.x = v[u];
        w.y = v[u + 1];
        *reinterpret_cast<typename Pair<T>::type *>(line + j0 + u) = w;
      }
  } else {
#pragma unroll
    for (int u = 0; u < CPT; ++u)
      if (j0 + u < n) line[j0 + u] = v[u];
  }
}
template <typename T, int CPT, int NT>
__device__ __forceinline__ void store_cols_nt(T *__restrict__ line, const T *v, int n) {
  const int j0 = CPT * threadIdx.x;
  if constexpr (CPT % 2 == 0) {
#pragma unroll
    for (int u = 0; u < CPT; u += 2)
      if (j0 + u < n) st_nt2(v[u], v[u + 1], line + j0 + u);
  } else {
#pragma unroll
    for (int u = 0; u < CPT; ++u)
      if (j0 + u < n) st_nt(v[u], line + j0 + u);
  }
}
template <typename T, int CPT, int NT>
__device__ __forceinline__ void load_cols(const T *__restrict__ line, T *v, int n) {
  const int j0 = CPT * threadIdx.x;
  if constexpr (CPT % 2 == 0) {
#pragma unroll
    for (int u = 0; u < CPT; u += 2) {
      const typename Pair<T>::type w =
          *reinterpret_cast<const typename Pair<T>::type *>(line + min(j0 + u, n - 2));
      v[u] = w.x;
      v[u + 1] = w.y;
    }
  } else {
#pragma unroll
    for (int u = 0; u < CPT; ++u) v[u] = line[min(j0 + u, n - 1)];
  }
}

// Reciprocal of a z pivot (1 / (b + lambda - a d)): IEEE division in FP64
// chains, so the pivots are the reference's bit for bit; in FP32 chains
// (whose result the FP64 refinement corrects) the hardware estimate
// refined by two Newton steps, within an ulp of the quotient (FP32 P2/P3
// at 512^3: 0.330 -> 0.306 ms and 0.329 -> 0.312 ms, the refined solve
// 4.33 -> 4.22 ms with the same final residual).  The pivots of a z system
// stay below 1e300 in magnitude and away from zero, where the estimate
// needs no range scaling.
template <typename T>
__device__ __forceinline__ double zrcp(double x) {
  if constexpr (sizeof(T) == 8) {
    return 1.0 / x;
  } else {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
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

// The slab body (also used by the batched kernels): sm holds 2*GP lines
// of line_stride<T, ny>() elements, stw the staged twiddles.
template <typename T, int NC, int GP, int NT, bool PER>
__device__ __forceinline__ void yz_fwd_slab(T *__restrict__ Wt, WStride ws,
                                            double *__restrict__ ck,
                                            int nx, int nz, int nzs, int kind,
                                            const double *__restrict__ lamx,
                                            const double *__restrict__ lamy,
                                            const double *__restrict__ a,
                                            const double *__restrict__ b,
                                            const double *__restrict__ c, int pin,
                                            const T *__restrict__ tw,
                                            const T *__restrict__ twh,
                                            const T *__restrict__ twq, T *sm, T *stw,
                                            int i) {
#pragma clang fp contract(off)
  constexpr int n = 2 * NC;  // ny
  constexpr int LS = line_stride<T, n>();
  constexpr int CPT = (n + NT - 1) / NT;  // columns per thread, adjacent
  constexpr int LD = (GP * n / 2 + NT - 1) / NT;  // element pairs per thread
  using V2 = typename Pair<T>::type;
  T *const A = sm, *const B = sm + GP * LS;
  // the complex FFT ends in A after an even number of passes
  constexpr bool fft_in_b = (tile::Plan<NC>::nst % 2) == 1;
  T *const Z = fft_in_b ? B : A;  // complex FFT result
  T *const Y = fft_in_b ? A : B;  // spectral lines after the real split
  const Tw<T, NC> W = stage_twiddles<T, NC, NT>(stw, tw, twh, twq);
  const int ngrp = (nz + GP - 1) / GP;
  const int nck = checkpoints<GP>(nz);
  const long long pstride = ws.wk;       // plane stride
  T *line0 = Wt + ws.ioff(i);            // plane 0 line of slab i
  // d4: REDFT11 / RODFT11 (DCT-IV / DST-IV, tile_fft.hpp)
  const bool d4 = tile::is_d4(kind);
  const bool trig = tile::is_trig(kind) && !d4;
  const bool sine = tile::is_sine(kind) || kind == K_RODFT11;
  const double lx = lamx[i];
  double lam[CPT], dprev[CPT], pprev[CPT];
  // periodic z only: second right-hand side, sums giving p1(0) and p2(0)
  double q2prev[CPT], s1[CPT], s2[CPT], gk[CPT];
#pragma unroll
  for (int u = 0; u < CPT; ++u) {
    const int j = CPT * threadIdx.x + u;
    lam[u] = (j < n) ? lx + lamy[j] : 0.0;
    dprev[u] = 0.0;
    pprev[u] = 0.0;
    q2prev[u] = 0.0;
    s1[u] = 0.0;
    s2[u] = 0.0;
    gk[u] = 1.0;
  }
  V2 r[LD];
  // branch-free prefetch of element pairs (16-byte loads in FP64): planes
  // past nz (last group) re-read plane nz-1
  auto load = [&](int g) {
    const int k0 = g * GP;
#pragma unroll
    for (int q = 0; q < LD; ++q) {
      const int t = min((int)threadIdx.x + q * NT, GP * n / 2 - 1);
      const int l = t / (n / 2), e = 2 * (t - l * (n / 2));
      r[q] = *reinterpret_cast<const V2 *>(line0 + (long long)min(k0 + l, nz - 1) * pstride + e);
    }
  };
  // The results of a group are held in registers and stored after the
  // next group's prefetch has arrived.  Vector-memory loads and stores
  // share one completion counter, so a wait for the prefetch also waits
  // for every store issued before it; stores issued at the end of a
  // group would put their full write latency on the next group's start.
  // Branch-free: in a last group of np < GP planes, hold[l] for l >= np
  // repeats plane np-1 and goes to the same address again.
  T hold[GP][CPT];
  auto flush = [&](int g) {
    const int k0 = g * GP;
#pragma unroll
    for (int l = 0; l < GP; ++l)
      store_cols_nt<T, CPT, NT>(line0 + (long long)min(k0 + l, nz - 1) * pstride, hold[l], n);
    // checkpoint of group g: ck[(i*nck + g)*ny + j] (periodic z: the
    // second right-hand side's p' follows at offset nx*nck*ny)
    double *cg = ck + ((long long)i * nck + g) * n;
    store_cols<double, CPT, NT>(cg, dprev, n);
    if constexpr (PER) store_cols<double, CPT, NT>(cg + (long long)nx * nck * n, q2prev, n);
  };
  load(0);
  for (int g = 0; g < ngrp; ++g) {
    const int k0 = g * GP;
    const int np = min(GP, nz - k0);
    // branch-free staging: Makhoul's reordering for the cosine/sine kinds
    // (e even -> e/2, e+1 -> n-1-e/2, odd entries negated for the sine
    // kinds), natural order otherwise
#pragma unroll
    for (int q = 0; q < LD; ++q) {
      const int t = threadIdx.x + q * NT;
      if ((GP * n / 2) % NT != 0 && t >= GP * n / 2) continue;
      const int l = t / (n / 2), e = 2 * (t - l * (n / 2));
      const int p0 = trig ? (e >> 1) : e;
      const int p1 = trig ? n - 1 - (e >> 1) : (d4 ? n - 1 - e : e + 1);
      A[l * LS + p0] = r[q].x;
      A[l * LS + p1] = ((trig || d4) && sine) ? -r[q].y : r[q].y;
    }
    __syncthreads();
    if (g > 0) flush(g - 1);
    // the last group re-reads itself (the values are not used), so the
    // loop carries no branch around the prefetch
    load(min(g + 1, ngrp - 1));
    if (d4) {
      tile::d4_pre<T, NC, GP, NT, LS, false>(A, W.th);
      __syncthreads();
    }
    tile::cfft_pp<T, -1, NC, GP, NT, LS / 2>(reinterpret_cast<cx<T> *>(A),
                                            reinterpret_cast<cx<T> *>(B), W.tw);
    if (d4) tile::d4_post_pp<T, NC, GP, NT, LS, false>(Z, Y, W.tq, sine);
    else tile::post_fwd_pp<T, NC, GP, NT, LS>(Z, Y, kind, W.th, W.tq);
    // forward elimination, plane by plane (dgtsv_homebrewed order)
#pragma unroll
    for (int l = 0; l < GP; ++l) {
      if (l >= np) {
#pragma unroll
        for (int u = 0; u < CPT; ++u) hold[l][u] = hold[l - 1][u];
        continue;
      }
      const int k = k0 + l;
      const double ak = a[k], bk = b[k], cc = c[k];
#pragma unroll
      for (int u = 0; u < CPT; ++u) {
        const int j = CPT * threadIdx.x + u;
        if (CPT * NT != n && j >= n) continue;
        const double y = (double)Y[l * LS + j];
        double pv;
        if constexpr (!PER) {
          if (k >= nzs) {
            // a boundary plane outside the z system passes through
            pv = y;
          } else if (k == 0) {
            const double z = zrcp<T>(bk + lam[u]);
            dprev[u] = cc * z;
            pv = y * z;
          } else if (k < nzs - 1) {
            const double bb = bk + lam[u];
            const double z = zrcp<T>(bb - ak * dprev[u]);
            dprev[u] = cc * z;
            pv = (y - ak * pprev[u]) * z;
          } else {
            const double z = (bk + lam[u]) - ak * dprev[u];
            if (z != 0.0 && !(pin && lam[u] == 0.0))
              pv = (y - ak * pprev[u]) / z;
            else
              pv = 0.0;
          }
        } else {
          // cyclic system (src/solver_cpu.f90:147-185): the first m = nz-1
          // rows are eliminated for the field (p1) and for
          // r2 = (-a(0), 0, ..., 0, -c(m-1)) (p2); p1(0) and p2(0) are the
          // sums of g(k) p'(k), g(k) = prod_{k'<k} (-d(k')), so the last
          // plane is closed here and P3 needs no second sweep.
          const int m = nz - 1;
          if (k < m) {
            const double r2 = (k == m - 1) ? -c[m - 1] : ((k == 0) ? -a[0] : 0.0);
            double q2;
            if (k == 0) {
              const double z = zrcp<T>(bk + lam[u]);
              dprev[u] = cc * z;
              pv = y * z;
              q2 = r2 * z;
            } else if (k < m - 1) {
              const double z = zrcp<T>((bk + lam[u]) - ak * dprev[u]);
              dprev[u] = cc * z;
              pv = (y - ak * pprev[u]) * z;
              q2 = (r2 - ak * q2prev[u]) * z;
            } else {
              const double z = (bk + lam[u]) - ak * dprev[u];
              if (z != 0.0) {
                pv = (y - ak * pprev[u]) / z;
                q2 = (r2 - ak * q2prev[u]) / z;
              } else {
                pv = 0.0;
                q2 = 0.0;
              }
            }
            s1[u] = s1[u] + gk[u] * pv;
            s2[u] = s2[u] + gk[u] * q2;
            gk[u] = -(dprev[u] * gk[u]);
            q2prev[u] = q2;
          } else if (pin && lam[u] == 0.0) {
            pv = 0.0;
          } else {
            pv = (y - c[m] * s1[u] - a[m] * pprev[u]) /
                 ((bk + lam[u]) + c[m] * s2[u] + a[m] * q2prev[u]);
          }
        }
        pprev[u] = pv;
        hold[l][u] = (T)pv;
      }
    }
    __syncthreads();
  }
  flush(ngrp - 1);
}

// ---------------------------------------------------------------- P3 -----
template <typename T, int NC, int GP, int NT, bool PER>
__device__ __forceinline__ void zy_bwd_slab(T *__restrict__ Wt, WStride ws,
                                            const double *__restrict__ ck,
                                            int nx, int nz, int nzs, int kind,
                                            const double *__restrict__ lamx,
                                            const double *__restrict__ lamy,
                                            const double *__restrict__ a,
                                            const double *__restrict__ b,
                                            const double *__restrict__ c,
                                            const T *__restrict__ tw,
                                            const T *__restrict__ twh,
                                            const T *__restrict__ twq, T *sm, T *stw,
                                            int i) {
#pragma clang fp contract(off)
  constexpr int n = 2 * NC;
  constexpr int LS = line_stride<T, n>();
  constexpr int CPT = (n + NT - 1) / NT;  // columns per thread, adjacent
  using V2 = typename Pair<T>::type;
  T *const A = sm, *const B = sm + GP * LS;
  // pre_inv A -> B, then the FFT passes start in B
  constexpr bool fft_in_a = (tile::Plan<NC>::nst % 2) == 1;
  T *const R = fft_in_a ? A : B;  // inverse FFT result
  const Tw<T, NC> W = stage_twiddles<T, NC, NT>(stw, tw, twh, twq);
  const int ngrp = (nz + GP - 1) / GP;
  const int nck = checkpoints<GP>(nz);
  const long long pstride = ws.wk;
  T *line0 = Wt + ws.ioff(i);
  // d4: REDFT11 / RODFT11 (DCT-IV / DST-IV, tile_fft.hpp)
  const bool d4 = tile::is_d4(kind);
  const bool trig = tile::is_trig(kind) && !d4;
  const bool sine = tile::is_sine(kind) || kind == K_RODFT11;
  const double lx = lamx[i];
  double lam[CPT], pnext[CPT];
  // periodic z only: p2 of the next plane, the closed last plane
  double pnext2[CPT], plast[CPT];
#pragma unroll
  for (int u = 0; u < CPT; ++u) {
    const int j = CPT * threadIdx.x + u;
    lam[u] = (j < n) ? lx + lamy[j] : 0.0;
    pnext[u] = 0.0;
    pnext2[u] = 0.0;
    plast[u] = 0.0;
  }
  // p' of a plane group in column-owner layout (thread owns columns j)
  T pv[GP][CPT];
  double dck[CPT], q2ck[CPT];
  auto load = [&](int g) {
    const int k0 = g * GP;
    // the recurrence restarts at the end of group g - 1 (unused, but read
    // in bounds, for g = 0); the checkpoint is issued first, so that the
    // factors can be recomputed while the lines are still arriving (loads
    // complete in issue order)
    const int c0 = max(g - 1, 0);
    const double *cg = ck + ((long long)i * nck + c0) * n;
    load_cols<double, CPT, NT>(cg, dck, n);
    if constexpr (PER) load_cols<double, CPT, NT>(cg + (long long)nx * nck * n, q2ck, n);
    // branch-free: out-of-range planes/columns re-read valid entries
#pragma unroll
    for (int l = 0; l < GP; ++l)
      load_cols<T, CPT, NT>(line0 + (long long)min(k0 + l, nz - 1) * pstride, pv[l], n);
  };
  // Stores of group g's lines (element pairs, 16-byte stores in FP64) from
  // R.  They are issued in the next group's step, after its loads have
  // arrived: vector-memory loads and stores share one completion counter,
  // so stores issued just before a wait for loads would add their write
  // latency to it.
  constexpr int LD2 = (GP * n / 2 + NT - 1) / NT;
  auto flush = [&](int g) {
    const int k0 = g * GP;
    const int np = min(GP, nz - k0);
#pragma unroll
    for (int q = 0; q < LD2; ++q) {
      const int t = threadIdx.x + q * NT;
      const int l = t / (n / 2), e = 2 * (t - l * (n / 2));
      if (((GP * n / 2) % NT != 0 && t >= GP * n / 2) || l >= np) continue;
      V2 v;
      if (trig) {
        v.x = R[l * LS + (e >> 1)];
        v.y = sine ? -R[l * LS + n - 1 - (e >> 1)] : R[l * LS + n - 1 - (e >> 1)];
      } else {
        v = *reinterpret_cast<const V2 *>(R + l * LS + e);
      }
      st_nt2(v.x, v.y, line0 + (long long)(k0 + l) * pstride + e);
    }
  };
  load(ngrp - 1);
  for (int g = ngrp - 1; g >= 0; --g) {
    const int k0 = g * GP;
    const int np = min(GP, nz - k0);
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
      const int j = CPT * threadIdx.x + u;
      if (CPT * NT != n && j >= n) continue;
      // elimination factors of this group, recomputed from the checkpoint
      double d[GP], q2[GP];
      double dp = dck[u];
      if constexpr (!PER) {
#pragma unroll
        for (int l = 0; l < GP; ++l) {
          const int k = k0 + l;
          if (l < np && k < nzs - 1) {
            if (k == 0) dp = c[0] * zrcp<T>(b[0] + lam[u]);
            else dp = c[k] * zrcp<T>((b[k] + lam[u]) - a[k] * dp);
          }
          d[l] = dp;
        }
        double pn = pnext[u];
#pragma unroll
        for (int l = GP - 1; l >= 0; --l) {
          if (l >= np) continue;
          const int k = k0 + l;
          // the last row of the system and the planes past it keep p'
          const double p = (k >= nzs - 1) ? (double)pv[l][u] : (double)pv[l][u] - d[l] * pn;
          pn = p;
          pv[l][u] = (T)p;
        }
        pnext[u] = pn;
      } else {
        // cyclic system: p = p1 + p2 * p(last) below the last plane, with
        // p1 from the stored p' and p2 from the recomputed r2 elimination
        const int m = nz - 1;
        double qp = q2ck[u];
#pragma unroll
        for (int l = 0; l < GP; ++l) {
          const int k = k0 + l;
          if (l < np && k < m) {
            const double r2 = (k == m - 1) ? -c[m - 1] : ((k == 0) ? -a[0] : 0.0);
            if (k == 0) {
              const double z = zrcp<T>(b[0] + lam[u]);
              dp = c[0] * z;
              qp = r2 * z;
            } else if (k < m - 1) {
              const double z = zrcp<T>((b[k] + lam[u]) - a[k] * dp);
              dp = c[k] * z;
              qp = (r2 - a[k] * qp) * z;
            } else {
              const double z = (b[k] + lam[u]) - a[k] * dp;
              qp = (z != 0.0) ? (r2 - a[k] * qp) / z : 0.0;
            }
          }
          if (k == nz - 1) plast[u] = (double)pv[l][u];
          d[l] = dp;
          q2[l] = qp;
        }
        double pn = pnext[u], pn2 = pnext2[u];
#pragma unroll
        for (int l = GP - 1; l >= 0; --l) {
          if (l >= np) continue;
          const int k = k0 + l;
          double p;
          if (k == nz - 1) {
            p = (double)pv[l][u];
          } else {
            double p1, p2;
            if (k == m - 1) {
              p1 = (double)pv[l][u];
              p2 = q2[l];
            } else {
              p1 = (double)pv[l][u] - d[l] * pn;
              p2 = q2[l] - d[l] * pn2;
            }
            pn = p1;
            pn2 = p2;
            p = p1 + p2 * plast[u];
          }
          pv[l][u] = (T)p;
        }
        pnext[u] = pn;
        pnext2[u] = pn2;
      }
    }
    // the previous group's lines leave R before this group's values enter
    // A (the same buffer)
    if (g + 1 < ngrp) {
      flush(g + 1);
      __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
      const int j = CPT * threadIdx.x + u;
      if (CPT * NT != n && j >= n) continue;
      const int jp = d4 ? tile::d4_pos(j, n) : j;
      const bool neg = d4 && sine && (j & 1);
#pragma unroll
      for (int l = 0; l < GP; ++l) A[l * LS + jp] = neg ? -pv[l][u] : pv[l][u];
    }
    __syncthreads();
    if (g > 0) load(g - 1);
    if (d4) tile::d4_pre_pp<T, NC, GP, NT, LS, true>(A, B, W.th);
    else tile::pre_inv_pp<T, NC, GP, NT, LS>(A, B, kind, W.th, W.tq);
    tile::cfft_pp<T, +1, NC, GP, NT, LS / 2>(reinterpret_cast<cx<T> *>(B),
                                            reinterpret_cast<cx<T> *>(A), W.tw);
    if (d4) tile::d4_post<T, NC, GP, NT, LS, true>(R, W.tq, sine);
  }
  flush(0);
}

// P2/P3 plane groups: the two ping-pong tiles of a group stay within
// kSlabBytes so that two slab workgroups fit one CU; FP64 lines of 1024
// points take 8 planes, one workgroup per CU (see tile_lines).
constexpr int kSlabBytes = 72 * 1024;
template <typename T, int NC>
constexpr int group_planes() {
  if (NC == 512 && sizeof(T) == 8) return 8;
  constexpr int bytes = 2 * line_stride<T, 2 * NC>() * (int)sizeof(T);
  constexpr int l = kSlabBytes / bytes;
  return l >= 16 ? 16 : (l >= 8 ? 8 : (l >= 4 ? 4 : (l >= 2 ? 2 : 1)));
}

// Threads of a P2/P3 slab workgroup of a single problem.  512- and
// 1024-point y-lines: 512 threads (one or two columns each, at most 128
// registers, 4 waves per SIMD) for the non-periodic passes (at
// 1024x1024x512: P2 2.99 -> 2.43 ms, P3 3.48 -> 2.63 ms), 256 for the
// periodic-z passes (whose cyclic state does not fit 128 registers without
// spilling); 256 for other lengths (a problem with few x-slabs has few
// workgroups, and more threads shorten each one).
template <typename T, int NC, bool PER>
constexpr int slab_threads() {
  if (2 * NC != 512 && 2 * NC != 1024) return 256;
  return PER ? 256 : 512;
}

// Threads of a batched slab workgroup (lines of up to 256 points): one
// column per thread in whole waves.  A batch has workgroups enough to fill
// the GPU, so no wave should idle in the elimination.
template <int NC>
constexpr int batch_slab_threads() {
  return ((2 * NC + 63) / 64) * 64;
}

}  // namespace flutas
