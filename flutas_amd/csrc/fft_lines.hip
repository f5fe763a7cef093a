// Batched real-to-real line transforms for gfx950.
//
// Each kind FluTAS selects for cell-centred x/y directions (src/fft.f90:
// 242-264) is evaluated with FFTW's unnormalised definition:
//   R2HC / HC2R          periodic          (halfcomplex order)
//   REDFT10 / REDFT01    Neumann-Neumann   (DCT-II / DCT-III)
//   RODFT10 / RODFT01    Dirichlet-Dirichlet (DST-II / DST-III)
//   REDFT11 / RODFT11    Neumann-Dirichlet / Dirichlet-Neumann (DCT-IV /
//                        DST-IV, each its own inverse up to 2n)
// A real line of even length n is packed as N = n/2 complex values and
// transformed with a Stockham FFT in LDS; the real split (pairs k, N-k) and
// the cosine/sine pre/post phases (Makhoul's even/odd reordering,
// phases exp(-i pi k / 2n)) are fused into the same kernel, so one launch
// reads every line once from HBM and writes it once.
#include "internal.hpp"

#include <cmath>

namespace flutas {

template <typename T>
struct cpx {
  T x, y;
};
template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> cconj(cpx<T> a) { return {a.x, -a.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return {a.x * s, a.y * s}; }
// multiply by S*i (S = +1 or -1)
template <typename T, int S>
__device__ __forceinline__ cpx<T> mul_si(cpx<T> a) {
  return (S > 0) ? cpx<T>{-a.y, a.x} : cpx<T>{a.y, -a.x};
}

// Small DFTs y_p = sum_q v_q exp(S 2 pi i p q / R), in place.
template <typename T, int S>
__device__ __forceinline__ void dft2(cpx<T> *v) {
  cpx<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}
template <typename T, int S>
__device__ __forceinline__ void dft3(cpx<T> *v) {
  const T h = T(0.866025403784438646763723170752936183L);  // sqrt(3)/2
  cpx<T> t1 = cadd(v[1], v[2]);
  cpx<T> t2 = {v[0].x - T(0.5) * t1.x, v[0].y - T(0.5) * t1.y};
  cpx<T> t3 = mul_si<T, S>(cscale(csub(v[1], v[2]), h));
  v[0] = cadd(v[0], t1);
  v[1] = cadd(t2, t3);
  v[2] = csub(t2, t3);
}
template <typename T, int S>
__device__ __forceinline__ void dft4(cpx<T> *v) {
  cpx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  cpx<T> t2 = cadd(v[1], v[3]), t3 = mul_si<T, S>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[1] = cadd(t1, t3);
  v[2] = csub(t0, t2);
  v[3] = csub(t1, t3);
}
template <typename T, int S>
__device__ __forceinline__ void dft5(cpx<T> *v) {
  const T c1 = T(0.309016994374947424102293417182819059L);   // cos(2pi/5)
  const T c2 = T(-0.809016994374947424102293417182819059L);  // cos(4pi/5)
  const T s1 = T(0.951056516295153572116439333379382143L);   // sin(2pi/5)
  const T s2 = T(0.587785252292473129168705954639072769L);   // sin(4pi/5)
  cpx<T> a1 = cadd(v[1], v[4]), a2 = cadd(v[2], v[3]);
  cpx<T> b1 = csub(v[1], v[4]), b2 = csub(v[2], v[3]);
  cpx<T> v0 = v[0];
  cpx<T> r1 = {v0.x + c1 * a1.x + c2 * a2.x, v0.y + c1 * a1.y + c2 * a2.y};
  cpx<T> r2 = {v0.x + c2 * a1.x + c1 * a2.x, v0.y + c2 * a1.y + c1 * a2.y};
  cpx<T> i1 = mul_si<T, S>(cpx<T>{s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y});
  cpx<T> i2 = mul_si<T, S>(cpx<T>{s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y});
  v[0] = cadd(v0, cadd(a1, a2));
  v[1] = cadd(r1, i1);
  v[4] = csub(r1, i1);
  v[2] = cadd(r2, i2);
  v[3] = csub(r2, i2);
}
template <typename T, int S>
__device__ __forceinline__ void dft8(cpx<T> *v) {
  const T r = T(0.707106781186547524400844362104849039L);
  cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
  cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<T, S>(e);
  dft4<T, S>(o);
  // w8^k = exp(S 2 pi i k / 8)
  cpx<T> o1 = {r * (o[1].x - S * o[1].y), r * (o[1].y + S * o[1].x)};
  cpx<T> o2 = mul_si<T, S>(o[2]);
  cpx<T> o3 = {r * (-o[3].x - S * o[3].y), r * (-o[3].y + S * o[3].x)};
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

// twiddle exp(S 2 pi i m / N) from the table of exp(-2 pi i m / N)
template <typename T, int S>
__device__ __forceinline__ cpx<T> twid(const T *tw, int m) {
  cpx<T> w = {tw[2 * m], tw[2 * m + 1]};
  return (S > 0) ? cconj(w) : w;
}

// One Stockham butterfly of radix R (Govindaraju et al. 2008 ordering):
// inputs x[b + r*N/R] times w^(r*j), outputs y[(b/Ns)*Ns*R + j + r*Ns].
template <typename T, int S, int R>
__device__ __forceinline__ void bfly(const cpx<T> *x, cpx<T> *y, const T *tw,
                                     int N, int Ns, int b) {
  const int nb = N / R;
  const int j = b % Ns;
  const int tstep = j * (N / (Ns * R));
  cpx<T> v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    cpx<T> a = x[b + r * nb];
    v[r] = (r == 0) ? a : cmul(a, twid<T, S>(tw, r * tstep));
  }
  if constexpr (R == 2) dft2<T, S>(v);
  if constexpr (R == 3) dft3<T, S>(v);
  if constexpr (R == 4) dft4<T, S>(v);
  if constexpr (R == 5) dft5<T, S>(v);
  if constexpr (R == 8) dft8<T, S>(v);
  const int base = (b / Ns) * Ns * R + j;
#pragma unroll
  for (int r = 0; r < R; ++r) y[base + r * Ns] = v[r];
}

// Radix without a fixed kernel (a prime > 5): direct DFT from LDS.
template <typename T, int S>
__device__ void bfly_generic(const cpx<T> *x, cpx<T> *y, const T *tw, int N,
                             int Ns, int R, int b) {
  const int nb = N / R;
  const int j = b % Ns;
  const int tstep = j * (N / (Ns * R));
  const int base = (b / Ns) * Ns * R + j;
  for (int p = 0; p < R; ++p) {
    cpx<T> acc = {T(0), T(0)};
    for (int r = 0; r < R; ++r) {
      cpx<T> a = x[b + r * nb];
      // w^(r*j) * exp(S 2 pi i p r / R) = exp(S 2 pi i (r*tstep + p*r*nb) / N)
      int m = (int)(((long long)r * tstep + (long long)p * r * nb) % N);
      acc = cadd(acc, cmul(a, twid<T, S>(tw, m)));
    }
    y[base + p * Ns] = acc;
  }
}

template <typename T, int S>
__device__ void fft_stages(T *&src, T *&dst, const LinePlan<T> &pl, int nl,
                           int n) {
  const int N = pl.N;
  int Ns = 1;
  for (int s = 0; s < pl.nst; ++s) {
    const int R = pl.radix[s];
    const int nb = N / R;
    for (int t = threadIdx.x; t < nl * nb; t += blockDim.x) {
      const int l = t / nb, b = t - (t / nb) * nb;
      const cpx<T> *x = reinterpret_cast<const cpx<T> *>(src + l * n);
      cpx<T> *y = reinterpret_cast<cpx<T> *>(dst + l * n);
      switch (R) {
        case 2: bfly<T, S, 2>(x, y, pl.tw, N, Ns, b); break;
        case 3: bfly<T, S, 3>(x, y, pl.tw, N, Ns, b); break;
        case 4: bfly<T, S, 4>(x, y, pl.tw, N, Ns, b); break;
        case 5: bfly<T, S, 5>(x, y, pl.tw, N, Ns, b); break;
        case 8: bfly<T, S, 8>(x, y, pl.tw, N, Ns, b); break;
        default: bfly_generic<T, S>(x, y, pl.tw, N, Ns, R, b); break;
      }
    }
    Ns *= R;
    __syncthreads();
    T *tmp = src;
    src = dst;
    dst = tmp;
  }
}

// halfcomplex element k (0 <= k <= n/2) of a real array X of length n
template <typename T>
__device__ __forceinline__ cpx<T> hc_at(const T *X, int k, int n) {
  if (k == 0 || 2 * k == n) return {X[k], T(0)};
  return {X[k], X[n - k]};
}

// V'_k = exp(i pi k / 2n) (X_k - i X_{n-k}), X_n := 0 (inverse DCT-II
// spectrum of the reordered sequence); `rev` reads X'_k = X_{n-1-k}.
template <typename T>
__device__ __forceinline__ cpx<T> vprime_at(const T *X, int k, int n,
                                            const T *twq, bool rev) {
  T xk = rev ? X[n - 1 - k] : X[k];
  T xnk = (k == 0) ? T(0) : (rev ? X[k - 1] : X[n - k]);
  cpx<T> q = {twq[2 * k], -twq[2 * k + 1]};  // exp(+i pi k / 2n)
  return cmul(q, cpx<T>{xk, -xnk});
}

template <typename T, typename Tin, typename Tout>
__global__ void __launch_bounds__(256)
    k_lines(const Tin *__restrict__ in, Tout *__restrict__ out, LineGeom g,
            LinePlan<T> pl, int kind, T scale, int lpb) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n = g.n, N = pl.N;
  T *sA = reinterpret_cast<T *>(smem_raw);
  T *sB = sA + (size_t)lpb * n;
  const int l1_0 = blockIdx.x * lpb;
  const int l2 = blockIdx.y;
  const int nl = min(lpb, g.L1 - l1_0);
  // DCT-IV / DST-IV (REDFT11 / RODFT11): v_m = x_2m + i x_{n-1-2m} times
  // exp(-i pi (4m+1) / 4n), forward FFT of length N, times exp(-i pi k/n):
  // Y_2k = 2 Re, Y_{n-1-2k} = -2 Im.  The sine kind transforms (-1)^j x_j
  // and reverses the output.
  const bool quarter = (kind == K_REDFT11 || kind == K_RODFT11);
  const bool fwd = (kind == K_R2HC || kind == K_REDFT10 || kind == K_RODFT10 || quarter);
  const bool trig = (kind != K_R2HC && kind != K_HC2R && !quarter);
  const bool sine = (kind == K_RODFT10 || kind == K_RODFT01 || kind == K_RODFT11);
  const bool contig = (g.ise == 1);

  // 1. load; forward cosine/sine kinds apply Makhoul's reordering
  //    v_m = x_2m, v_{n-1-m} = x_{2m+1} (and (-1)^j for the sine kind)
  for (int t = threadIdx.x; t < nl * n; t += blockDim.x) {
    int l, e;
    if (contig) { l = t / n; e = t - l * n; }
    else { e = t / nl; l = t - e * nl; }
    T x = (T)in[(long long)(l1_0 + l) * g.is1 + (long long)l2 * g.is2 +
                (long long)e * g.ise];
    int pos = e;
    if (fwd && trig) {
      pos = (e & 1) ? (n - 1 - (e >> 1)) : (e >> 1);
      if (sine && (e & 1)) x = -x;
    } else if (quarter) {
      pos = (e & 1) ? (n - e) : e;  // x_2m -> Re v_m, x_{n-1-2m} -> Im v_m
      if (sine && (e & 1)) x = -x;
    }
    sA[l * n + pos] = x;
  }
  __syncthreads();
  T *src = sA, *dst = sB;
  if (quarter) {
    for (int t = threadIdx.x; t < nl * N; t += blockDim.x) {
      const int l = t / N, m = t - (t / N) * N;
      cpx<T> *v = reinterpret_cast<cpx<T> *>(src + l * n);
      v[m] = cmul(v[m], cpx<T>{pl.twe[2 * m], pl.twe[2 * m + 1]});
    }
    __syncthreads();
  }

  // 2. inverse kinds: Z'_k = (X_k + conj X_{N-k}) + i e^{2 pi i k/n}(X_k - conj X_{N-k})
  if (!fwd) {
    for (int t = threadIdx.x; t < nl * N; t += blockDim.x) {
      const int l = t / N, k = t - (t / N) * N;
      const T *X = src + l * n;
      cpx<T> Xk, Xm;
      if (!trig) {
        Xk = hc_at(X, k, n);
        Xm = hc_at(X, N - k, n);
      } else {
        Xk = vprime_at(X, k, n, pl.twq, sine);
        Xm = vprime_at(X, N - k, n, pl.twq, sine);
      }
      cpx<T> A = cadd(Xk, cconj(Xm)), B = csub(Xk, cconj(Xm));
      cpx<T> w = {pl.twh[2 * k], -pl.twh[2 * k + 1]};  // exp(+2 pi i k / n)
      cpx<T> Z = cadd(A, mul_si<T, 1>(cmul(w, B)));
      dst[l * n + 2 * k] = Z.x;
      dst[l * n + 2 * k + 1] = Z.y;
    }
    __syncthreads();
    T *tmp = src; src = dst; dst = tmp;
  }

  // 3. complex FFT of length N per line
  if (fwd) fft_stages<T, -1>(src, dst, pl, nl, n);
  else fft_stages<T, +1>(src, dst, pl, nl, n);

  // 4. forward kinds: real split X_k = (Z_k + conj Z_{N-k})/2
  //    - i e^{-2 pi i k/n} (Z_k - conj Z_{N-k})/2, then halfcomplex packing
  //    (R2HC) or the phase Y_k = 2 Re(e^{-i pi k/2n} X_k) (cosine/sine).
  if (quarter) {
    for (int t = threadIdx.x; t < nl * N; t += blockDim.x) {
      const int l = t / N, k = t - (t / N) * N;
      const cpx<T> *Z = reinterpret_cast<const cpx<T> *>(src + l * n);
      T *Y = dst + l * n;
      // exp(-i pi k / n) = twq[2k]
      const cpx<T> u = cmul(Z[k], cpx<T>{pl.twq[4 * k], pl.twq[4 * k + 1]});
      const int ie = 2 * k, io = n - 1 - 2 * k;
      Y[sine ? n - 1 - ie : ie] = T(2) * u.x;
      Y[sine ? n - 1 - io : io] = T(-2) * u.y;
    }
    __syncthreads();
    T *tmp = src; src = dst; dst = tmp;
  } else if (fwd) {
    for (int t = threadIdx.x; t < nl * (N + 1); t += blockDim.x) {
      const int l = t / (N + 1), k = t - (t / (N + 1)) * (N + 1);
      const cpx<T> *Z = reinterpret_cast<const cpx<T> *>(src + l * n);
      T *Y = dst + l * n;
      cpx<T> Zk = Z[(k == N) ? 0 : k];
      cpx<T> Zm = Z[(k == 0) ? 0 : N - k];
      cpx<T> A = cadd(Zk, cconj(Zm)), B = csub(Zk, cconj(Zm));
      cpx<T> wB = cmul(cpx<T>{pl.twh[2 * k], pl.twh[2 * k + 1]}, B);
      cpx<T> X = {T(0.5) * A.x + T(0.5) * wB.y, T(0.5) * A.y - T(0.5) * wB.x};
      if (!trig) {
        Y[k] = X.x;
        if (k > 0 && k < N) Y[n - k] = X.y;
      } else {
        const T c = pl.twq[2 * k], sn = -pl.twq[2 * k + 1];
        T yk = T(2) * (X.x * c + X.y * sn);
        T ynk = T(2) * (X.x * sn - X.y * c);
        int ik = k, ink = n - k;
        if (sine) { ik = n - 1 - k; ink = k - 1; }
        if (k < n) Y[ik] = yk;
        if (k > 0 && k < N) Y[ink] = ynk;
      }
    }
    __syncthreads();
    T *tmp = src; src = dst; dst = tmp;
  }

  // 5. store; inverse cosine/sine kinds undo the reordering
  //    (Y_2m = v_m, Y_{2m+1} = v_{n-1-m}, times (-1)^j for the sine kind)
  for (int t = threadIdx.x; t < nl * n; t += blockDim.x) {
    int l, e;
    if (g.ose == 1) { l = t / n; e = t - l * n; }
    else { e = t / nl; l = t - e * nl; }
    int pos = e;
    T sg = T(1);
    if (!fwd && trig) {
      pos = (e & 1) ? (n - 1 - (e >> 1)) : (e >> 1);
      if (sine && (e & 1)) sg = T(-1);
    }
    out[(long long)(l1_0 + l) * g.os1 + (long long)l2 * g.os2 +
        (long long)e * g.ose] = (Tout)(src[l * n + pos] * (scale * sg));
  }
}

// REDFT00 / RODFT00 (DCT-I / DST-I, the face-centred NN and DD kinds of
// src/fft.f90:265-289) through the real DFT of the symmetric extension of
// each line of n values, length m = 2(n-1) (even) or 2(n+1) (odd):
//   even  e_j = x_j (j < n), e_{m-j} = x_j (0 < j < n-1):
//         E_k real, y_k = E_k = x_0 + (-1)^k x_{n-1}
//                             + 2 sum_{0<j<n-1} x_j cos(pi j k/(n-1));
//   odd   o_0 = o_{n+1} = 0, o_{j+1} = x_j, o_{m-1-j} = -x_j:
//         O_k = -i 2 sum_j x_j sin(pi (j+1) k/(n+1)), y_k = -Im O_{k+1}.
// The plan pl is that of length m (N = m/2 complex points); g.n = n.
template <typename T, typename Tin, typename Tout>
__global__ void __launch_bounds__(256)
    k_lines_ext(const Tin *__restrict__ in, Tout *__restrict__ out, LineGeom g,
                LinePlan<T> pl, int kind, T scale, int lpb) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n = g.n, m = pl.n, N = pl.N;
  T *sA = reinterpret_cast<T *>(smem_raw);
  T *sB = sA + (size_t)lpb * m;
  const int l1_0 = blockIdx.x * lpb;
  const int l2 = blockIdx.y;
  const int nl = min(lpb, g.L1 - l1_0);
  const bool odd = (kind == K_RODFT00);
  if (odd) {
    for (int t = threadIdx.x; t < nl; t += blockDim.x) {
      sA[t * m] = T(0);
      sA[t * m + n + 1] = T(0);
    }
  }
  // 1. load x and place it (and its mirror image) in the extension
  for (int t = threadIdx.x; t < nl * n; t += blockDim.x) {
    int l, e;
    if (g.ise == 1) { l = t / n; e = t - l * n; }
    else { e = t / nl; l = t - e * nl; }
    const T x = (T)in[(long long)(l1_0 + l) * g.is1 + (long long)l2 * g.is2 +
                      (long long)e * g.ise];
    T *E = sA + l * m;
    if (odd) {
      E[e + 1] = x;
      E[m - 1 - e] = -x;
    } else {
      E[e] = x;
      if (e > 0 && e < n - 1) E[m - e] = x;
    }
  }
  __syncthreads();
  // 2. complex FFT of the packed extension (N points)
  T *src = sA, *dst = sB;
  fft_stages<T, -1>(src, dst, pl, nl, m);
  // 3. real split X_k = (Z_k + conj Z_{N-k})/2 - i e^{-2 pi i k/m}
  //    (Z_k - conj Z_{N-k})/2 (Z_N := Z_0) for the k the kind needs
  for (int t = threadIdx.x; t < nl * n; t += blockDim.x) {
    const int l = t / n, e = t - (t / n) * n;
    const int k = odd ? e + 1 : e;  // k <= N
    const cpx<T> *Z = reinterpret_cast<const cpx<T> *>(src + l * m);
    const cpx<T> Zk = Z[(k == N) ? 0 : k];
    const cpx<T> Zm = Z[(k == 0) ? 0 : N - k];
    const cpx<T> A = cadd(Zk, cconj(Zm)), B = csub(Zk, cconj(Zm));
    const cpx<T> wB = cmul(cpx<T>{pl.twh[2 * k], pl.twh[2 * k + 1]}, B);
    dst[l * m + e] = odd ? -(T(0.5) * A.y - T(0.5) * wB.x) : (T(0.5) * A.x + T(0.5) * wB.y);
  }
  __syncthreads();
  // 4. store
  for (int t = threadIdx.x; t < nl * n; t += blockDim.x) {
    int l, e;
    if (g.ose == 1) { l = t / n; e = t - l * n; }
    else { e = t / nl; l = t - e * nl; }
    out[(long long)(l1_0 + l) * g.os1 + (long long)l2 * g.os2 + (long long)e * g.ose] =
        (Tout)(dst[l * m + e] * scale);
  }
}

// ---- host side -------------------------------------------------------------

// Largest prime factor of N a plan accepts (direct DFT of that size per
// butterfly, O(p^2) operations).
constexpr int kMaxPrime = 1021;

static int factorize(int N, int *radix) {
  int nst = 0;
  int m = N;
  while (m % 8 == 0 && m > 8) { radix[nst++] = 8; m /= 8; }
  for (int p : {8, 4, 2, 3, 5}) {
    while (m % p == 0) {
      if (nst >= kMaxStages) return -1;
      radix[nst++] = p;
      m /= p;
    }
  }
  for (int p = 7; m > 1; p += 2) {
    while (m % p == 0) {
      if (nst >= kMaxStages || p > kMaxPrime) return -1;
      radix[nst++] = p;
      m /= p;
    }
  }
  return nst;
}

LinePlan<double> LineTables::dplan() const {
  LinePlan<double> p{};
  p.n = n; p.N = N; p.nst = nst;
  for (int i = 0; i < kMaxStages; ++i) p.radix[i] = radix[i];
  p.tw = d_tw; p.twh = d_twh; p.twq = d_twq; p.twe = d_twe; p.t4 = d_t4;
  return p;
}
LinePlan<float> LineTables::fplan() const {
  LinePlan<float> p{};
  p.n = n; p.N = N; p.nst = nst;
  for (int i = 0; i < kMaxStages; ++i) p.radix[i] = radix[i];
  p.tw = f_tw; p.twh = f_twh; p.twq = f_twq; p.twe = f_twe; p.t4 = f_t4;
  return p;
}

int LineTables::build(int n_) {
  release();
  if (n_ < 2 || (n_ & 1)) return fail(FLUTAS_EINVAL, "line length must be even");
  n = n_;
  N = n / 2;
  nst = (N == 1) ? 0 : factorize(N, radix);
  if (nst < 0) return fail(FLUTAS_EUNSUPPORTED, "n/2 has a prime factor > 1021");
  const long double pi = 3.14159265358979323846264338327950288L;
  std::vector<double> tw(2 * N), twh(2 * (N + 1)), twq(2 * (n + 1)), twe(2 * N),
      t4(4 * (N + 1));
  for (int m = 0; m < N; ++m) {
    long double a = 2.0L * pi * m / N;
    tw[2 * m] = (double)cosl(a);
    tw[2 * m + 1] = (double)-sinl(a);
  }
  for (int k = 0; k <= N; ++k) {
    long double a = 2.0L * pi * k / n;
    twh[2 * k] = (double)cosl(a);
    twh[2 * k + 1] = (double)-sinl(a);
  }
  for (int k = 0; k <= n; ++k) {
    long double a = pi * k / (2.0L * n);
    twq[2 * k] = (double)cosl(a);
    twq[2 * k + 1] = (double)-sinl(a);
  }
  for (int m = 0; m < N; ++m) {
    long double a = pi * (4.0L * m + 1.0L) / (4.0L * n);
    twe[2 * m] = (double)cosl(a);
    twe[2 * m + 1] = (double)-sinl(a);
  }
  for (int m = 0; m <= N; ++m) {
    const long double a = pi * (4.0L * m + 1.0L) / (4.0L * n), b = pi * m / (long double)n;
    t4[2 * m] = (double)cosl(a);
    t4[2 * m + 1] = (double)-sinl(a);
    t4[2 * (N + 1) + 2 * m] = (double)cosl(b);
    t4[2 * (N + 1) + 2 * m + 1] = (double)-sinl(b);
  }
  auto up = [](const std::vector<double> &h, double **dd, float **df) -> int {
    FL_HIP(hipMalloc(dd, h.size() * sizeof(double)));
    FL_HIP(hipMemcpy(*dd, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    std::vector<float> f(h.begin(), h.end());
    FL_HIP(hipMalloc(df, f.size() * sizeof(float)));
    FL_HIP(hipMemcpy(*df, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
    return FLUTAS_OK;
  };
  FL_TRY(up(tw, &d_tw, &f_tw));
  FL_TRY(up(twh, &d_twh, &f_twh));
  FL_TRY(up(twq, &d_twq, &f_twq));
  FL_TRY(up(twe, &d_twe, &f_twe));
  FL_TRY(up(t4, &d_t4, &f_t4));
  return FLUTAS_OK;
}

void LineTables::release() {
  for (void *p : {(void *)d_tw, (void *)d_twh, (void *)d_twq, (void *)f_tw,
                  (void *)f_twh, (void *)f_twq, (void *)d_ptw, (void *)f_ptw,
                  (void *)d_twe, (void *)f_twe, (void *)d_t4, (void *)f_t4})
    if (p) (void)hipFree(p);
  d_tw = d_twh = d_twq = d_twe = d_t4 = nullptr;
  f_tw = f_twh = f_twq = f_twe = f_t4 = nullptr;
  d_ptw = nullptr;
  f_ptw = nullptr;
}

template <typename T, typename Tin, typename Tout>
int launch_lines(const LineTables &tab, int kind, const Tin *in, Tout *out,
                 const LineGeom &g, T scale, hipStream_t s) {
  const int m = dft_length(kind, g.n);
  if (m != tab.n) return fail(FLUTAS_EINVAL, "line length does not match plan");
  // lines per block: as many as fit 64 KiB of LDS (two buffers), <= 16
  const size_t per_line = 2 * (size_t)m * sizeof(T);
  int lpb = (int)std::min<size_t>(16, (64 * 1024) / per_line);
  if (lpb < 1) return fail(FLUTAS_EUNSUPPORTED, "line too long for one workgroup");
  lpb = std::min(lpb, g.L1);
  dim3 grid((g.L1 + lpb - 1) / lpb, g.L2);
  const size_t shmem = per_line * lpb;
  LinePlan<T> pl;
  if constexpr (sizeof(T) == 8) pl = tab.dplan(); else pl = tab.fplan();
  // (REDFT00 of two points has m == n: dispatch on the kind)
  if (kind == K_REDFT00 || kind == K_RODFT00)
    hipLaunchKernelGGL((k_lines_ext<T, Tin, Tout>), grid, dim3(256), shmem, s, in,
                       out, g, pl, kind, scale, lpb);
  else
    hipLaunchKernelGGL((k_lines<T, Tin, Tout>), grid, dim3(256), shmem, s, in,
                       out, g, pl, kind, scale, lpb);
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

template int launch_lines<double, double, double>(const LineTables &, int, const double *, double *, const LineGeom &, double, hipStream_t);
template int launch_lines<float, double, float>(const LineTables &, int, const double *, float *, const LineGeom &, float, hipStream_t);
template int launch_lines<float, float, float>(const LineTables &, int, const float *, float *, const LineGeom &, float, hipStream_t);
template int launch_lines<float, float, double>(const LineTables &, int, const float *, double *, const LineGeom &, float, hipStream_t);

}  // namespace flutas
