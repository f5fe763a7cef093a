// z-chunked slab passes of the fused solve, for problems with few x-slabs.
//
// P2/P3 give one workgroup a whole x-slab and walk its nz planes in order.
// With few slabs (256x256xN, 128^3) that leaves most of the GPU idle: 256
// slabs are one workgroup per CU.  Here the planes of a slab are cut into
// nchunk chunks of cpl planes (a multiple of the plane group GP; the last
// chunk runs to nz), one workgroup each, and the chunks are tied together
// by a reduced system of two unknowns per chunk and column.
//
// For one column (i, j), lam = lamx(i) + lamy(j), the z system has msys
// rows (nzs, or nz - 1 for periodic z, whose last plane closes the cyclic
// system as in src/solver_cpu.f90:147-185); planes past msys pass through.
// The reference eliminates
//   p'(k) = (y(k) - a(k) p'(k-1)) z(k),  d(k) = c(k) z(k),
//   z(k) = 1 / (b(k) + lam - a(k) d(k-1))          (src/solver_cpu.f90:187-223)
// and back-substitutes x(k) = p'(k) - d(k) x(k+1).  d, z do not depend on
// the right-hand side.  A chunk [s, e) eliminates its own planes from a
// zero carry, pL(k), so that the true p'(k) = pL(k) + m(k) P with
// P = p'(s-1) and the multiplier m(k) = prod_{s<=k'<=k} (-a(k') z(k')).
// The back substitution gives the first value of the chunk as
//   X = x(s) = S + P SW + N x(e),
// with S = sum_k w(k) pL(k), SW = sum_k w(k) m(k), N = w(e) and
// w(k) = prod_{s<=k'<k} (-d(k')).  So:
//   P2 (yz_fwd_chunk)  per chunk: y r2r, pL, and E = pL of the chunk's last
//                      system row and S, written to a small side array;
//   P3 (zy_bwd_chunk)  per chunk: the carries P(c+1) = E(c) + M(c) P(c)
//                      and X(c) = S(c) + P(c) SW(c) + N(c) X(c+1) for the
//                      whole slab (a few operations per column), then the
//                      back substitution x(k) = pL(k) + m(k) P - d(k) x(k+1)
//                      from x(e) = X(c+1), and the inverse y r2r.
// d, m (and for periodic z the second solution p2 of the cyclic system) do
// not depend on the right-hand side: k_zchunk_tables computes them once per
// plan, one thread per column, with the same operations as the passes, so
// the pivots d(k) of FP64 chains stay the reference's bit for bit (FP32
// chains take the reciprocal zrcp); P3 recomputes the d, m
// (and q2) of each plane group from the tables' values at the group start.
// For periodic z the closing plane x(m), m = nz - 1, is
//   (y(m) - c(m) p1(0) - a(m) p1(m-1)) / (b(m) + lam + c(m) p2(0) + a(m) p2(m-1))
// with p1(0) = X(0) and p1(m-1) the true p' of row m-1, and every other
// plane is p1(k) + p2(k) x(m).  The results equal the sequential
// elimination's to rounding (the carries are added, not threaded through
// every plane).
#pragma once

#include "fused_body.hpp"

namespace flutas {

// Chunk rows [ks, ke) of chunk ch.
__device__ __forceinline__ int zc_begin(const ZChunks &zc, int ch) { return ch * zc.cpl; }
__device__ __forceinline__ int zc_end(const ZChunks &zc, int ch, int nz) {
  return ch == zc.nchunk - 1 ? nz : (ch + 1) * zc.cpl;
}

// Workgroup -> (slab, chunk).  Workgroups are dealt to the 8 XCDs round
// robin; with nx a multiple of 8 the chunks of one slab get ids equal
// modulo 8 and share an XCD (and its L2, where the side arrays of the
// slab's chunks are read by each of them).
__device__ __forceinline__ void zc_slab_of(int nx, int nchunk, int *i, int *ch) {
  const int b = blockIdx.x;
  if ((nx & 7) == 0) {
    *i = (b / (8 * nchunk)) * 8 + (b & 7);
    *ch = (b >> 3) % nchunk;
  } else {
    *i = b / nchunk;
    *ch = b % nchunk;
  }
}

// ---------------------------------------------------------------- P2 -----
template <typename T, int NC, int GP, int NT, bool PER>
__device__ __forceinline__ void yz_fwd_chunk(T *__restrict__ Wt, WStride ws, const ZChunks &zc,
                                             int nz, int msys, int kind,
                                             const double *__restrict__ lamx,
                                             const double *__restrict__ lamy,
                                             const double *__restrict__ a,
                                             const double *__restrict__ b,
                                             const double *__restrict__ c, int pin,
                                             const T *__restrict__ tw,
                                             const T *__restrict__ twh,
                                             const T *__restrict__ twq, T *sm, T *stw, int i,
                                             int ch) {
#pragma clang fp contract(off)
  constexpr int n = 2 * NC;  // ny
  constexpr int LS = line_stride<T, n>();
  constexpr int CPT = (n + NT - 1) / NT;
  constexpr int LD = (GP * n / 2 + NT - 1) / NT;
  using V2 = typename Pair<T>::type;
  T *const A = sm, *const B = sm + GP * LS;
  constexpr bool fft_in_b = (tile::Plan<NC>::nst % 2) == 1;
  T *const Z = fft_in_b ? B : A;
  T *const Y = fft_in_b ? A : B;
  const Tw<T, NC> W = stage_twiddles<T, NC, NT>(stw, tw, twh, twq);
  const int ks = zc_begin(zc, ch), ke = zc_end(zc, ch, nz);
  const int gb = ks / GP, ge = (ke + GP - 1) / GP;
  const long long pstride = ws.wk;
  T *line0 = Wt + ws.ioff(i);
  // d4: REDFT11 / RODFT11 (DCT-IV / DST-IV, tile_fft.hpp)
  const bool d4 = tile::is_d4(kind);
  const bool trig = tile::is_trig(kind) && !d4;
  const bool sine = tile::is_sine(kind) || kind == K_RODFT11;
  const double lx = lamx[i];
  double lam[CPT], dprev[CPT], pprev[CPT], w[CPT], S[CPT], E[CPT], ym[CPT];
  // d of the plane before the chunk (not used by the first chunk)
  load_cols<double, CPT, NT>(zc.ckd + ((long long)i * zc.ngrp + gb) * n, dprev, n);
#pragma unroll
  for (int u = 0; u < CPT; ++u) {
    const int j = CPT * threadIdx.x + u;
    lam[u] = (j < n) ? lx + lamy[j] : 0.0;
    pprev[u] = 0.0;
    w[u] = 1.0;
    S[u] = 0.0;
    E[u] = 0.0;
    ym[u] = 0.0;
  }
  V2 r[LD];
  // branch-free prefetch; planes past the chunk (last group) re-read ke-1
  auto load = [&](int g) {
    const int k0 = g * GP;
#pragma unroll
    for (int q = 0; q < LD; ++q) {
      const int t = min((int)threadIdx.x + q * NT, GP * n / 2 - 1);
      const int l = t / (n / 2), e = 2 * (t - l * (n / 2));
      r[q] = *reinterpret_cast<const V2 *>(line0 + (long long)min(k0 + l, ke - 1) * pstride + e);
    }
  };
  // a group's results leave after the next group's prefetch has arrived
  // (see yz_fwd_slab); hold[l] past the chunk repeats its last plane
  T hold[GP][CPT];
  auto flush = [&](int g) {
    const int k0 = g * GP;
#pragma unroll
    for (int l = 0; l < GP; ++l)
      store_cols_nt<T, CPT, NT>(line0 + (long long)min(k0 + l, ke - 1) * pstride, hold[l], n);
  };
  load(gb);
  for (int g = gb; g < ge; ++g) {
    const int k0 = g * GP;
    const int np = min(GP, ke - k0);
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
    if (g > gb) flush(g - 1);
    load(min(g + 1, ge - 1));
    if (d4) {
      tile::d4_pre<T, NC, GP, NT, LS, false>(A, W.th);
      __syncthreads();
    }
    tile::cfft_pp<T, -1, NC, GP, NT, LS / 2>(reinterpret_cast<cx<T> *>(A),
                                            reinterpret_cast<cx<T> *>(B), W.tw);
    if (d4) tile::d4_post_pp<T, NC, GP, NT, LS, false>(Z, Y, W.tq, sine);
    else tile::post_fwd_pp<T, NC, GP, NT, LS>(Z, Y, kind, W.th, W.tq);
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
        double pv, dk = 0.0;
        if (k >= msys) {
          // a plane outside the system (periodic z: the closing plane)
          pv = y;
          if (PER) ym[u] = y;
        } else if (k == 0) {
          const double z = zrcp<T>(bk + lam[u]);
          dprev[u] = cc * z;
          dk = dprev[u];
          pv = y * z;
        } else if (k < msys - 1) {
          const double z = zrcp<T>((bk + lam[u]) - ak * dprev[u]);
          dprev[u] = cc * z;
          dk = dprev[u];
          pv = (y - ak * pprev[u]) * z;
        } else {
          const double z = (bk + lam[u]) - ak * dprev[u];
          if (z != 0.0 && (PER || !(pin && lam[u] == 0.0)))
            pv = (y - ak * pprev[u]) / z;
          else
            pv = 0.0;
        }
        const T pt = (T)pv;
        if (k < msys) {
          // the sums see the stored (rounded) values P3 reads back
          S[u] = S[u] + w[u] * (double)pt;
          w[u] = -(dk * w[u]);
          E[u] = (double)pt;
        }
        pprev[u] = pv;
        hold[l][u] = pt;
      }
    }
    __syncthreads();
  }
  flush(ge - 1);
  double *sd = zc.side + ((long long)i * zc.nchunk + ch) * zc_side_slots(PER) * n;
  store_cols<double, CPT, NT>(sd, E, n);
  store_cols<double, CPT, NT>(sd + n, S, n);
  if (PER && ch == zc.nchunk - 1) store_cols<double, CPT, NT>(sd + 2 * n, ym, n);
}

// ---------------------------------------------------------------- P3 -----
template <typename T, int NC, int GP, int NT, bool PER>
__device__ __forceinline__ void zy_bwd_chunk(T *__restrict__ Wt, WStride ws, const ZChunks &zc,
                                             int nz, int msys, int kind,
                                             const double *__restrict__ lamx,
                                             const double *__restrict__ lamy,
                                             const double *__restrict__ a,
                                             const double *__restrict__ b,
                                             const double *__restrict__ c, int pin,
                                             const T *__restrict__ tw,
                                             const T *__restrict__ twh,
                                             const T *__restrict__ twq, T *sm, T *stw, int i,
                                             int ch) {
#pragma clang fp contract(off)
  constexpr int n = 2 * NC;
  constexpr int LS = line_stride<T, n>();
  constexpr int CPT = (n + NT - 1) / NT;
  using V2 = typename Pair<T>::type;
  T *const A = sm, *const B = sm + GP * LS;
  constexpr bool fft_in_a = (tile::Plan<NC>::nst % 2) == 1;
  T *const R = fft_in_a ? A : B;
  const Tw<T, NC> W = stage_twiddles<T, NC, NT>(stw, tw, twh, twq);
  const int ks = zc_begin(zc, ch), ke = zc_end(zc, ch, nz);
  const int gb = ks / GP, ge = (ke + GP - 1) / GP;
  const long long pstride = ws.wk;
  T *line0 = Wt + ws.ioff(i);
  // d4: REDFT11 / RODFT11 (DCT-IV / DST-IV, tile_fft.hpp)
  const bool d4 = tile::is_d4(kind);
  const bool trig = tile::is_trig(kind) && !d4;
  const bool sine = tile::is_sine(kind) || kind == K_RODFT11;
  const double lx = lamx[i];
  const int nch = zc.nchunk;
  double lam[CPT];
#pragma unroll
  for (int u = 0; u < CPT; ++u) {
    const int j = CPT * threadIdx.x + u;
    lam[u] = (j < n) ? lx + lamy[j] : 0.0;
  }
  // The reduced system of the slab's chunks: carry into this chunk (Pc),
  // x of the plane after it (pn), periodic: p2 of that plane (pn2) and the
  // closing plane's value (xm).
  double Pc[CPT], pn[CPT], pn2[CPT], xm[CPT];
  {
    constexpr int TS = zc_tab_slots(PER), SS = zc_side_slots(PER);
    const double *tb = zc.tab + (long long)i * nch * TS * n;
    const double *sd = zc.side + (long long)i * nch * SS * n;
    double Pk[kMaxZChunks][CPT], Pv[CPT], X[CPT];
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
      Pv[u] = 0.0;
      X[u] = 0.0;
      Pc[u] = 0.0;
      pn[u] = 0.0;
      pn2[u] = 0.0;
      xm[u] = 0.0;
    }
#pragma unroll
    for (int q = 0; q < kMaxZChunks; ++q) {
      if (q >= nch) break;
      double Eq[CPT], Mq[CPT];
      load_cols<double, CPT, NT>(sd + (long long)q * SS * n, Eq, n);
      load_cols<double, CPT, NT>(tb + (long long)q * TS * n, Mq, n);
#pragma unroll
      for (int u = 0; u < CPT; ++u) {
        Pk[q][u] = Pv[u];
        Pv[u] = Eq[u] + Mq[u] * Pv[u];
      }
    }
#pragma unroll
    for (int q = kMaxZChunks - 1; q >= 0; --q) {
      if (q >= nch) continue;
      double Sq[CPT], SWq[CPT], Nq[CPT];
      load_cols<double, CPT, NT>(sd + ((long long)q * SS + 1) * n, Sq, n);
      load_cols<double, CPT, NT>(tb + ((long long)q * TS + 1) * n, SWq, n);
      load_cols<double, CPT, NT>(tb + ((long long)q * TS + 2) * n, Nq, n);
#pragma unroll
      for (int u = 0; u < CPT; ++u) {
        if (q == ch) {
          pn[u] = X[u];
          Pc[u] = Pk[q][u];
        }
        X[u] = (Sq[u] + Pk[q][u] * SWq[u]) + Nq[u] * X[u];
      }
    }
    if constexpr (PER) {
      // Pv: p1 of row m-1; X: p1(0)
      const int m = nz - 1;
      double Ym[CPT], Dn[CPT];
      load_cols<double, CPT, NT>(sd + ((long long)(nch - 1) * SS + 2) * n, Ym, n);
      load_cols<double, CPT, NT>(zc.den + (long long)i * n, Dn, n);
      if (ch < nch - 1) load_cols<double, CPT, NT>(tb + ((long long)ch * TS + 3) * n, pn2, n);
#pragma unroll
      for (int u = 0; u < CPT; ++u)
        xm[u] = (pin && lam[u] == 0.0) ? 0.0 : (Ym[u] - c[m] * X[u] - a[m] * Pv[u]) / Dn[u];
    }
  }
  T pv[GP][CPT];
  double dck[CPT], mck[CPT], qck[CPT];
  auto load = [&](int g) {
    const int k0 = g * GP;
    // the group's starting values first (see zy_bwd_slab), then its lines
    const long long o = ((long long)i * zc.ngrp + g) * n;
    load_cols<double, CPT, NT>(zc.ckd + o, dck, n);
    load_cols<double, CPT, NT>(zc.ckm + o, mck, n);
    if constexpr (PER) load_cols<double, CPT, NT>(zc.ckq + o, qck, n);
#pragma unroll
    for (int l = 0; l < GP; ++l)
      load_cols<T, CPT, NT>(line0 + (long long)min(k0 + l, ke - 1) * pstride, pv[l], n);
  };
  constexpr int LD2 = (GP * n / 2 + NT - 1) / NT;
  auto flush = [&](int g) {
    const int k0 = g * GP;
    const int np = min(GP, ke - k0);
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
  load(ge - 1);
  for (int g = ge - 1; g >= gb; --g) {
    const int k0 = g * GP;
    const int np = min(GP, ke - k0);
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
      const int j = CPT * threadIdx.x + u;
      if (CPT * NT != n && j >= n) continue;
      // d, m (periodic: q2) of the group's rows from the values before it
      double d[GP], mm[GP], q2[GP];
      double dp = dck[u], mp = mck[u], qp = PER ? qck[u] : 0.0;
#pragma unroll
      for (int l = 0; l < GP; ++l) {
        const int k = k0 + l;
        double dl = 0.0, ml = 0.0;
        if (l < np && k < msys) {
          const double r2 = PER ? ((k == msys - 1) ? -c[msys - 1] : ((k == 0) ? -a[0] : 0.0)) : 0.0;
          if (k == 0) {
            const double z = zrcp<T>(b[0] + lam[u]);
            dp = c[0] * z;
            mp = 0.0;
            if (PER) qp = r2 * z;
            dl = dp;
          } else if (k < msys - 1) {
            const double z = zrcp<T>((b[k] + lam[u]) - a[k] * dp);
            dp = c[k] * z;
            mp = (-a[k] * mp) * z;
            if (PER) qp = (r2 - a[k] * qp) * z;
            dl = dp;
          } else {
            const double z = (b[k] + lam[u]) - a[k] * dp;
            mp = (z != 0.0 && (PER || !(pin && lam[u] == 0.0))) ? (-a[k] * mp) / z : 0.0;
            if (PER) qp = (z != 0.0) ? (r2 - a[k] * qp) / z : 0.0;
          }
          ml = mp;
        }
        d[l] = dl;
        mm[l] = ml;
        q2[l] = qp;
      }
      double xn = pn[u], xn2 = pn2[u];
#pragma unroll
      for (int l = GP - 1; l >= 0; --l) {
        if (l >= np) continue;
        const int k = k0 + l;
        const double pt = (double)pv[l][u];
        double x;
        if constexpr (!PER) {
          if (k >= msys) x = pt;
          else if (k == msys - 1) x = pt + mm[l] * Pc[u];
          else x = (pt + mm[l] * Pc[u]) - d[l] * xn;
          xn = x;
        } else {
          if (k >= msys) {
            x = xm[u];
          } else {
            const double p1t = pt + mm[l] * Pc[u];
            double p1, p2;
            if (k == msys - 1) {
              p1 = p1t;
              p2 = q2[l];
            } else {
              p1 = p1t - d[l] * xn;
              p2 = q2[l] - d[l] * xn2;
            }
            xn = p1;
            xn2 = p2;
            x = p1 + p2 * xm[u];
          }
        }
        pv[l][u] = (T)x;
      }
      pn[u] = xn;
      pn2[u] = xn2;
    }
    if (g + 1 < ge) {
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
    if (g > gb) load(g - 1);
    if (d4) tile::d4_pre_pp<T, NC, GP, NT, LS, true>(A, B, W.th);
    else tile::pre_inv_pp<T, NC, GP, NT, LS>(A, B, kind, W.th, W.tq);
    tile::cfft_pp<T, +1, NC, GP, NT, LS / 2>(reinterpret_cast<cx<T> *>(B),
                                            reinterpret_cast<cx<T> *>(A), W.tw);
    if (d4) tile::d4_post<T, NC, GP, NT, LS, true>(R, W.tq, sine);
  }
  flush(gb);
}

}  // namespace flutas
