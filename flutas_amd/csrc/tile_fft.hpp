// Compile-time-sized FFT building blocks for tiles of lines held in LDS.
//
// A tile is L lines of n = 2N reals (line stride LS reals) in LDS.  The whole
// workgroup (NT threads) runs a Stockham FFT of length N on every line: each
// pass reads the butterfly inputs of all its butterflies into registers,
// synchronises, and writes the outputs in place.  The real split of
// R2HC/HC2R and the cosine/sine phases of REDFT10/01, RODFT10/01 are
// separate tile passes with the same read/synchronise/write structure, so a
// line is read from HBM once and written once per kernel.
#pragma once

#include "internal.hpp"

namespace flutas {
namespace tile {

template <typename T>
struct cx {
  T x, y;
};
template <typename T>
__device__ __forceinline__ cx<T> add(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> sub(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> mul(cx<T> a, cx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cx<T> cj(cx<T> a) { return {a.x, -a.y}; }
template <typename T, int S>
__device__ __forceinline__ cx<T> msi(cx<T> a) {  // multiply by S*i
  return (S > 0) ? cx<T>{-a.y, a.x} : cx<T>{a.y, -a.x};
}

template <typename T, int S>
__device__ __forceinline__ void d2(cx<T> *v) {
  cx<T> a = v[0], b = v[1];
  v[0] = add(a, b);
  v[1] = sub(a, b);
}
template <typename T, int S>
__device__ __forceinline__ void d3(cx<T> *v) {
  const T h = T(0.866025403784438646763723170752936183L);
  cx<T> t1 = add(v[1], v[2]);
  cx<T> t2 = {v[0].x - T(0.5) * t1.x, v[0].y - T(0.5) * t1.y};
  cx<T> d = sub(v[1], v[2]);
  cx<T> t3 = msi<T, S>(cx<T>{h * d.x, h * d.y});
  v[0] = add(v[0], t1);
  v[1] = add(t2, t3);
  v[2] = sub(t2, t3);
}
template <typename T, int S>
__device__ __forceinline__ void d4(cx<T> *v) {
  cx<T> t0 = add(v[0], v[2]), t1 = sub(v[0], v[2]);
  cx<T> t2 = add(v[1], v[3]), t3 = msi<T, S>(sub(v[1], v[3]));
  v[0] = add(t0, t2);
  v[1] = add(t1, t3);
  v[2] = sub(t0, t2);
  v[3] = sub(t1, t3);
}
template <typename T, int S>
__device__ __forceinline__ void d8(cx<T> *v) {
  const T r = T(0.707106781186547524400844362104849039L);
  cx<T> e[4] = {v[0], v[2], v[4], v[6]};
  cx<T> o[4] = {v[1], v[3], v[5], v[7]};
  d4<T, S>(e);
  d4<T, S>(o);
  cx<T> o1 = {r * (o[1].x - S * o[1].y), r * (o[1].y + S * o[1].x)};
  cx<T> o2 = msi<T, S>(o[2]);
  cx<T> o3 = {r * (-o[3].x - S * o[3].y), r * (-o[3].y + S * o[3].x)};
  v[0] = add(e[0], o[0]);
  v[4] = sub(e[0], o[0]);
  v[1] = add(e[1], o1);
  v[5] = sub(e[1], o1);
  v[2] = add(e[2], o2);
  v[6] = sub(e[2], o2);
  v[3] = add(e[3], o3);
  v[7] = sub(e[3], o3);
}
// 16-point DFT as 4 x 4 with the internal twiddles w16^(n2*k1)
template <typename T, int S>
__device__ __forceinline__ void d16(cx<T> *v) {
  const T c1 = T(0.923879532511286756128183189396788933L);  // cos(pi/8)
  const T s1 = T(0.382683432365089771728459984030398867L);  // sin(pi/8)
  const T r = T(0.707106781186547524400844362104849039L);
  cx<T> A[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    cx<T> q[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
    d4<T, S>(q);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) A[n2][k1] = q[k1];
  }
  // twiddles w^m, w = exp(S 2 pi i / 16), m = n2*k1
  auto tw = [&](cx<T> a, int m) -> cx<T> {
    switch (m) {
      case 0: return a;
      case 1: return mul(a, cx<T>{c1, S * s1});
      case 2: return mul(a, cx<T>{r, S * r});
      case 3: return mul(a, cx<T>{s1, S * c1});
      case 4: return msi<T, S>(a);
      case 6: return mul(a, cx<T>{-r, S * r});
      case 9: return mul(a, cx<T>{-c1, -S * s1});
      default: return a;
    }
  };
#pragma unroll
  for (int n2 = 1; n2 < 4; ++n2)
#pragma unroll
    for (int k1 = 1; k1 < 4; ++k1) A[n2][k1] = tw(A[n2][k1], n2 * k1);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    cx<T> q[4] = {A[0][k1], A[1][k1], A[2][k1], A[3][k1]};
    d4<T, S>(q);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = q[k2];
  }
}

// 5-point DFT: with t1 = x1 + x4, t2 = x2 + x3, t3 = x1 - x4, t4 = x2 - x3
// and w = exp(S 2 pi i / 5),
//   X0 = x0 + t1 + t2,
//   X1, X4 = x0 + c1 t1 + c2 t2 +- S i (s1 t3 + s2 t4),
//   X2, X3 = x0 + c2 t1 + c1 t2 +- S i (s2 t3 - s1 t4),
// c1 = cos(2 pi/5), c2 = cos(4 pi/5), s1 = sin(2 pi/5), s2 = sin(4 pi/5).
template <typename T, int S>
__device__ __forceinline__ void d5(cx<T> *v) {
  const T c1 = T(0.309016994374947424102293417182819059L);
  const T c2 = T(-0.809016994374947424102293417182819059L);
  const T s1 = T(0.951056516295153572116439333379382143L);
  const T s2 = T(0.587785252292473129168705954639072769L);
  const cx<T> t1 = add(v[1], v[4]), t2 = add(v[2], v[3]);
  const cx<T> t3 = sub(v[1], v[4]), t4 = sub(v[2], v[3]);
  const cx<T> x0 = v[0];
  const cx<T> a1 = {x0.x + c1 * t1.x + c2 * t2.x, x0.y + c1 * t1.y + c2 * t2.y};
  const cx<T> a2 = {x0.x + c2 * t1.x + c1 * t2.x, x0.y + c2 * t1.y + c1 * t2.y};
  const cx<T> b1 = msi<T, S>(cx<T>{s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y});
  const cx<T> b2 = msi<T, S>(cx<T>{s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y});
  v[0] = add(x0, add(t1, t2));
  v[1] = add(a1, b1);
  v[4] = sub(a1, b1);
  v[2] = add(a2, b2);
  v[3] = sub(a2, b2);
}

template <typename T, int S, int R>
__device__ __forceinline__ void dft(cx<T> *v) {
  if constexpr (R == 2) d2<T, S>(v);
  else if constexpr (R == 3) d3<T, S>(v);
  else if constexpr (R == 5) d5<T, S>(v);
  else if constexpr (R == 4) d4<T, S>(v);
  else if constexpr (R == 8) d8<T, S>(v);
  else if constexpr (R == 16) d16<T, S>(v);
  else static_assert(R == 2, "radix not built");
}

// Compile-time radix plans of the complex length N = n/2 (radix 16/8/4 first,
// then 3 or 5).  Pass s reads
// with stride N/r[s] and writes with stride ns(s) = r[0]*...*r[s-1].
// Twiddles of pass s >= 1 live in their own table laid out [rr][j]
// (rr < r[s], j < ns(s)) so that lanes with consecutive j read
// consecutive entries; tw_off(s) is the table's offset in complex values.
template <int N>
struct Plan;
#define FL_PLAN(NN, A, B, C, NS)                                        \
  template <>                                                           \
  struct Plan<NN> {                                                     \
    static constexpr int nst = NS;                                      \
    static constexpr int r[3] = {A, B, C};                              \
    static constexpr int ns(int s) { return s == 0 ? 1 : (s == 1 ? A : A * B); } \
    static constexpr int tw_off(int s) { return s <= 1 ? 0 : B * A; }   \
    static constexpr int tw_total() { return NS == 1 ? 1 : (NS == 2 ? B * A : B * A + C * A * B); } \
  };
FL_PLAN(16, 16, 1, 1, 1)
FL_PLAN(24, 8, 3, 1, 2)
FL_PLAN(32, 8, 4, 1, 2)
FL_PLAN(48, 16, 3, 1, 2)
FL_PLAN(64, 16, 4, 1, 2)
FL_PLAN(96, 8, 4, 3, 3)
FL_PLAN(128, 16, 8, 1, 2)
FL_PLAN(192, 16, 4, 3, 3)
FL_PLAN(256, 8, 8, 4, 3)
FL_PLAN(384, 16, 8, 3, 3)
FL_PLAN(512, 8, 8, 8, 3)
FL_PLAN(1024, 16, 16, 4, 3)
// radix-5 lengths (real lines of 80, 160, 320 and 640 points)
FL_PLAN(40, 8, 5, 1, 2)
FL_PLAN(80, 16, 5, 1, 2)
FL_PLAN(160, 8, 4, 5, 3)
FL_PLAN(320, 8, 8, 5, 3)
// two radix-5 passes (real lines of 100, 200, 400 and 800 points)
FL_PLAN(50, 2, 5, 5, 3)
FL_PLAN(100, 4, 5, 5, 3)
FL_PLAN(200, 8, 5, 5, 3)
FL_PLAN(400, 16, 5, 5, 3)
#undef FL_PLAN

// XOR swizzle of the buffer between passes 0 and 1 (power-of-two r[0]):
// pass 0 writes element b*r0 + rr of butterfly b; swizzling the low bits by
// b spreads the r0 lanes of a write over distinct LDS banks, and pass 1's
// contiguous reads stay within one block of r0 elements.
template <int N>
__device__ __forceinline__ constexpr bool swizzled() {
  constexpr int r0 = Plan<N>::r[0];
  return Plan<N>::nst >= 2 && (r0 & (r0 - 1)) == 0;
}
template <int N>
__device__ __forceinline__ int swz(int m) {
  constexpr int r0 = Plan<N>::r[0];
  return m ^ ((m / r0) & (r0 - 1));
}
template <int N, int SI>
__device__ __forceinline__ int in_idx(int m) {
  if constexpr (SI == 1 && swizzled<N>()) return swz<N>(m);
  return m;
}
template <int N, int SI>
__device__ __forceinline__ int out_idx(int m) {
  if constexpr (SI == 0 && swizzled<N>()) return swz<N>(m);
  return m;
}

// Twiddle of pass SI (>= 1) for butterfly position j and input rr, from the
// per-pass table (see Plan).
template <typename T, int S, int N, int SI>
__device__ __forceinline__ cx<T> ptw(const T *__restrict__ tw, int rr, int j) {
  const int m = Plan<N>::tw_off(SI) + rr * Plan<N>::ns(SI) + j;
  cx<T> w = {tw[2 * m], tw[2 * m + 1]};
  if (S > 0) w.y = -w.y;
  return w;
}

// Stockham pass SI of radix R = Plan<N>::r[SI] over L lines (complex line
// stride LSC), in place: all inputs are read into registers, then written.
template <typename T, int S, int N, int SI, int L, int NT, int LSC>
__device__ __forceinline__ void pass(cx<T> *buf, const T *__restrict__ tw) {
  constexpr int R = Plan<N>::r[SI];
  constexpr int NS = Plan<N>::ns(SI);
  constexpr int NB = N / R;
  constexpr int TOT = L * NB;
  constexpr int PER = (TOT + NT - 1) / NT;
  cx<T> v[PER][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int t = tid + u * NT;
    if (TOT % NT == 0 || t < TOT) {
      const int l = t / NB, b = t - (t / NB) * NB;
      const int j = b % NS;
      const cx<T> *x = buf + l * LSC;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        cx<T> a = x[in_idx<N, SI>(b + r * NB)];
        if (SI > 0 && r > 0) a = mul(a, ptw<T, S, N, SI>(tw, r, j));
        v[u][r] = a;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int t = tid + u * NT;
    if (TOT % NT == 0 || t < TOT) {
      const int l = t / NB, b = t - (t / NB) * NB;
      const int j = b % NS;
      dft<T, S, R>(v[u]);
      cx<T> *y = buf + l * LSC;
      const int base = (b / NS) * NS * R + j;
#pragma unroll
      for (int r = 0; r < R; ++r) y[out_idx<N, SI>(base + r * NS)] = v[u][r];
    }
  }
  __syncthreads();
}

// Complex FFT of length N on L lines, sign S; tw = per-pass twiddle tables.
template <typename T, int S, int N, int L, int NT, int LSC>
__device__ __forceinline__ void cfft(cx<T> *buf, const T *__restrict__ tw) {
  using P = Plan<N>;
  pass<T, S, N, 0, L, NT, LSC>(buf, tw);
  if constexpr (P::nst > 1) pass<T, S, N, 1, L, NT, LSC>(buf, tw);
  if constexpr (P::nst > 2) pass<T, S, N, 2, L, NT, LSC>(buf, tw);
}

__device__ __forceinline__ bool is_fwd(int kind) {
  return kind == K_R2HC || kind == K_REDFT10 || kind == K_RODFT10;
}
__device__ __forceinline__ bool is_trig(int kind) {
  return kind != K_R2HC && kind != K_HC2R;
}
__device__ __forceinline__ bool is_sine(int kind) {
  return kind == K_RODFT10 || kind == K_RODFT01;
}
// Makhoul's reordering position of element e (v_m = x_2m, v_{n-1-m} = x_2m+1)
__device__ __forceinline__ int mk_pos(int e, int n) {
  return (e & 1) ? (n - 1 - (e >> 1)) : (e >> 1);
}

// ---- DCT-IV / DST-IV (REDFT11 / RODFT11: the ND and DN pairs) ----------
// Y_k = 2 sum_j x_j cos(pi (j + 1/2)(k + 1/2) / n) through a complex FFT of
// length N = n/2: v_m = x_2m + i x_{n-1-2m}, w_m = v_m exp(-i pi (4m+1)/4n),
// W = FFT_-(w), u_k = W_k exp(-i pi k/n), Y_2k = 2 Re u_k,
// Y_{n-1-2k} = -2 Im u_k.  DST-IV: the same on (-1)^j x_j, output
// reversed.  Each is its own inverse up to 2n, so the backward passes run
// the same transform; their FFT has the + sign, which gives FFT_- through
// FFT_-(w) = conj(FFT_+(conj w)) (CJ below).  The pre and post tables
// (exp(-i pi (4m+1)/4n), exp(-i pi k/n)) take the staged slots of the split
// and phase tables, which these kinds do not use.
__device__ __forceinline__ bool is_d4(int kind) { return kind == K_REDFT11 || kind == K_RODFT11; }
// position of element e in the packed line (Re v_m at 2m, Im v_m at 2m+1)
__device__ __forceinline__ int d4_pos(int e, int n) { return (e & 1) ? (n - e) : e; }

// w_m = v_m pre_m (conjugated when CJ), element-wise in place: every thread
// rewrites the values it reads, so no barrier inside.
template <typename T, int N, int L, int NT, int LS, bool CJ>
__device__ __forceinline__ void d4_pre(T *sm, const T *__restrict__ pre) {
  constexpr int TOT = L * N;
  constexpr int PER = (TOT + NT - 1) / NT;
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int t = threadIdx.x + u * NT;
    if (TOT % NT != 0 && t >= TOT) continue;
    const int l = t / N, m = t - (t / N) * N;
    cx<T> *Z = reinterpret_cast<cx<T> *>(sm + l * LS);
    const cx<T> v = mul(Z[m], cx<T>{pre[2 * m], pre[2 * m + 1]});
    Z[m] = CJ ? cj(v) : v;
  }
}
// the same from src into dst, then a barrier
template <typename T, int N, int L, int NT, int LS, bool CJ>
__device__ __forceinline__ void d4_pre_pp(const T *__restrict__ src, T *__restrict__ dst,
                                          const T *__restrict__ pre) {
  constexpr int TOT = L * N;
  constexpr int PER = (TOT + NT - 1) / NT;
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int t = threadIdx.x + u * NT;
    if (TOT % NT != 0 && t >= TOT) continue;
    const int l = t / N, m = t - (t / N) * N;
    const cx<T> *Z = reinterpret_cast<const cx<T> *>(src + l * LS);
    const cx<T> v = mul(Z[m], cx<T>{pre[2 * m], pre[2 * m + 1]});
    reinterpret_cast<cx<T> *>(dst + l * LS)[m] = CJ ? cj(v) : v;
  }
  __syncthreads();
}
// Output of slot k of the FFT (conjugated when CJ): the two real values
// and their positions (reversed for the sine kind).
template <typename T, int N, bool CJ>
__device__ __forceinline__ void d4_out(cx<T> W, int k, const T *__restrict__ post, bool sine,
                                       T &v0, int &p0, T &v1, int &p1) {
  constexpr int n = 2 * N;
  const cx<T> u = mul(CJ ? cj(W) : W, cx<T>{post[2 * k], post[2 * k + 1]});
  v0 = T(2) * u.x;
  v1 = T(-2) * u.y;
  p0 = sine ? n - 1 - 2 * k : 2 * k;
  p1 = sine ? 2 * k : n - 1 - 2 * k;
}
// in place, register staged (barrier between the reads and the writes, and
// after the writes)
template <typename T, int N, int L, int NT, int LS, bool CJ>
__device__ __forceinline__ void d4_post(T *sm, const T *__restrict__ post, bool sine) {
  constexpr int TOT = L * N;
  constexpr int PER = (TOT + NT - 1) / NT;
  T o[PER][2];
  int ps[PER][2];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int t = threadIdx.x + u * NT;
    ps[u][0] = ps[u][1] = -1;
    if (TOT % NT != 0 && t >= TOT) continue;
    const int l = t / N, k = t - (t / N) * N;
    const cx<T> W = reinterpret_cast<const cx<T> *>(sm + l * LS)[k];
    d4_out<T, N, CJ>(W, k, post, sine, o[u][0], ps[u][0], o[u][1], ps[u][1]);
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int t = threadIdx.x + u * NT;
    if (ps[u][0] < 0) continue;
    T *line = sm + (t / N) * LS;
    line[ps[u][0]] = o[u][0];
    line[ps[u][1]] = o[u][1];
  }
  __syncthreads();
}
// from src into dst, then a barrier
template <typename T, int N, int L, int NT, int LS, bool CJ>
__device__ __forceinline__ void d4_post_pp(const T *__restrict__ src, T *__restrict__ dst,
                                           const T *__restrict__ post, bool sine) {
  constexpr int TOT = L * N;
  constexpr int PER = (TOT + NT - 1) / NT;
#pragma unroll 2
  for (int u = 0; u < PER; ++u) {
    const int t = threadIdx.x + u * NT;
    if (TOT % NT != 0 && t >= TOT) continue;
    const int l = t / N, k = t - (t / N) * N;
    const cx<T> W = reinterpret_cast<const cx<T> *>(src + l * LS)[k];
    T v0, v1;
    int p0, p1;
    d4_out<T, N, CJ>(W, k, post, sine, v0, p0, v1, p1);
    T *line = dst + l * LS;
    line[p0] = v0;
    line[p1] = v1;
  }
  __syncthreads();
}

// Real split of the pair (k, N-k), 0 <= k <= N/2, of the complex FFT Z of
// a packed real line (Z_N := Z_0): with A = Z_k + conj Z_{N-k},
// B = Z_k - conj Z_{N-k} and w = exp(-2 pi i k/n),
//   X_k = A/2 - i w B/2,   X_{N-k} = conj(A/2 + i w B/2)
// (w_{N-k} = -conj w_k), so one twiddle product serves both members.
template <typename T>
__device__ __forceinline__ void split_pair(cx<T> Zk, cx<T> Zm, cx<T> w, cx<T> &Xk,
                                           cx<T> &Xm) {
  const cx<T> A = add(Zk, cj(Zm)), B = sub(Zk, cj(Zm));
  const cx<T> wB = mul(w, B);
  Xk = {T(0.5) * A.x + T(0.5) * wB.y, T(0.5) * A.y - T(0.5) * wB.x};
  Xm = {T(0.5) * A.x - T(0.5) * wB.y, -(T(0.5) * A.y + T(0.5) * wB.x)};
}
// Output of spectral value X_kk (0 <= kk <= N) for the forward kinds:
// R2HC halfcomplex (Y_kk = Re, Y_{n-kk} = Im) or the cosine/sine phase
// (Y = 2 Re(exp(-i pi kk/2n) X), natural order, reversed for the sine kind).
// p1 < 0 when the second value does not exist (kk = 0 or N).
template <typename T, int N>
__device__ __forceinline__ void fwd_out(cx<T> X, int kk, bool trig, bool sine,
                                        const T *__restrict__ twq, T &v0, int &p0, T &v1,
                                        int &p1) {
  constexpr int n = 2 * N;
  const bool two = (kk > 0 && kk < N);
  if (!trig) {
    v0 = X.x; p0 = kk;
    v1 = X.y; p1 = two ? n - kk : -1;
  } else {
    const T c = twq[2 * kk], sn = -twq[2 * kk + 1];
    v0 = T(2) * (X.x * c + X.y * sn);
    v1 = T(2) * (X.x * sn - X.y * c);
    p0 = sine ? n - 1 - kk : kk;
    p1 = two ? (sine ? kk - 1 : n - kk) : -1;
  }
}
// Inverse pre-processing of the pair (k, N-k): with A = X_k + conj X_{N-k},
// B = X_k - conj X_{N-k} and w = exp(+2 pi i k/n),
//   Z'_k = A + i w B,   Z'_{N-k} = conj(A - i w B).
template <typename T>
__device__ __forceinline__ void join_pair(cx<T> Xk, cx<T> Xm, cx<T> w, cx<T> &Zk,
                                          cx<T> &Zm) {
  const cx<T> A = add(Xk, cj(Xm)), B = sub(Xk, cj(Xm));
  const cx<T> wB = mul(w, B);
  Zk = {A.x - wB.y, A.y + wB.x};
  Zm = {A.x + wB.y, -(A.y - wB.x)};
}

// Forward real split and kind phases, in place on L lines of n = 2N reals:
// the complex FFT Z of the packed line becomes the R2HC halfcomplex output
// or the REDFT10 / RODFT10 output in natural order.
template <typename T, int N, int L, int NT, int LS>
__device__ __forceinline__ void post_fwd(T *sm, int kind, const T *__restrict__ twh,
                                         const T *__restrict__ twq) {
  constexpr int NP = N / 2 + 1;  // pairs (k, N-k), k = 0..N/2
  constexpr int TOT = L * NP;
  constexpr int PER = (TOT + NT - 1) / NT;
  const bool trig = is_trig(kind), sine = is_sine(kind);
  T o[PER][4];
  int pos[PER][4];
  const int tid = threadIdx.x;
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int t = tid + u * NT;
    pos[u][0] = pos[u][1] = pos[u][2] = pos[u][3] = -1;
    if (t < TOT) {
      const int l = t / NP, k = t - (t / NP) * NP;
      const cx<T> *Z = reinterpret_cast<const cx<T> *>(sm + l * LS);
      cx<T> Xk, Xm;
      split_pair(Z[k], Z[k == 0 ? 0 : N - k], cx<T>{twh[2 * k], twh[2 * k + 1]}, Xk, Xm);
      fwd_out<T, N>(Xk, k, trig, sine, twq, o[u][0], pos[u][0], o[u][1], pos[u][1]);
      if (2 * k != N)
        fwd_out<T, N>(Xm, N - k, trig, sine, twq, o[u][2], pos[u][2], o[u][3], pos[u][3]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int t = tid + u * NT;
    if (t < TOT) {
      T *line = sm + (t / NP) * LS;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (pos[u][q] >= 0) line[pos[u][q]] = o[u][q];
    }
  }
  __syncthreads();
}

// halfcomplex element k (0 <= k <= N) of a real line
template <typename T, int N>
__device__ __forceinline__ cx<T> hc(const T *X, int k) {
  if (k == 0 || k == N) return {X[k], T(0)};
  return {X[k], X[2 * N - k]};
}
// V'_k = exp(i pi k / 2n)(X_k - i X_{n-k}), X_n := 0; `rev` reads X_{n-1-k}
template <typename T, int N>
__device__ __forceinline__ cx<T> vprime(const T *X, int k, const T *__restrict__ twq,
                                        bool rev) {
  constexpr int n = 2 * N;
  const T xk = rev ? X[n - 1 - k] : X[k];
  const T xnk = (k == 0) ? T(0) : (rev ? X[k - 1] : X[n - k]);
  const cx<T> q = {twq[2 * k], -twq[2 * k + 1]};
  return mul(q, cx<T>{xk, -xnk});
}

// Inverse pre-processing in place: real input line X (halfcomplex for
// HC2R, natural for REDFT01/RODFT01) becomes the complex array
// Z'_k = (X_k + conj X_{N-k}) + i e^{2 pi i k/n}(X_k - conj X_{N-k}).
template <typename T, int N, int L, int NT, int LS>
__device__ __forceinline__ void pre_inv(T *sm, int kind, const T *__restrict__ twh,
                                        const T *__restrict__ twq) {
  constexpr int NP = N / 2 + 1;
  constexpr int TOT = L * NP;
  constexpr int PER = (TOT + NT - 1) / NT;
  const bool trig = is_trig(kind), sine = is_sine(kind);
  cx<T> z[PER][2];
  const int tid = threadIdx.x;
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int t = tid + u * NT;
    if (t < TOT) {
      const int l = t / NP, k = t - (t / NP) * NP;
      const T *X = sm + l * LS;
      cx<T> Xk, Xm;
      if (!trig) {
        Xk = hc<T, N>(X, k);
        Xm = hc<T, N>(X, N - k);
      } else {
        Xk = vprime<T, N>(X, k, twq, sine);
        Xm = vprime<T, N>(X, N - k, twq, sine);
      }
      join_pair(Xk, Xm, cx<T>{twh[2 * k], -twh[2 * k + 1]}, z[u][0], z[u][1]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int t = tid + u * NT;
    if (t < TOT) {
      const int l = t / NP, k = t - (t / NP) * NP;
      cx<T> *Z = reinterpret_cast<cx<T> *>(sm + l * LS);
      Z[k] = z[u][0];
      if (k != 0 && 2 * k != N) Z[N - k] = z[u][1];
    }
  }
  __syncthreads();
}

// ---- ping-pong variants: read buffer `src`, write buffer `dst`, one
// barrier at the end; no register staging (lower register pressure at the
// cost of a second LDS tile).

template <typename T, int S, int N, int SI, int L, int NT, int LSC>
__device__ __forceinline__ void pass_pp(const cx<T> *__restrict__ src, cx<T> *__restrict__ dst,
                                        const T *__restrict__ tw) {
  constexpr int R = Plan<N>::r[SI];
  constexpr int NS = Plan<N>::ns(SI);
  constexpr int NB = N / R;
  constexpr int TOT = L * NB;
  constexpr int PER = (TOT + NT - 1) / NT;
  // fully unrolled over a thread's butterflies; src and dst never overlap,
  // so the reads of every butterfly can be issued before the first write
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int t = threadIdx.x + u * NT;
    if (TOT % NT != 0 && t >= TOT) continue;
    const int l = t / NB, b = t - (t / NB) * NB;
    const int j = b % NS;
    const cx<T> *x = src + l * LSC;
    cx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      cx<T> a = x[in_idx<N, SI>(b + r * NB)];
      if (SI > 0 && r > 0) a = mul(a, ptw<T, S, N, SI>(tw, r, j));
      v[r] = a;
    }
    dft<T, S, R>(v);
    cx<T> *y = dst + l * LSC;
    const int base = (b / NS) * NS * R + j;
#pragma unroll
    for (int r = 0; r < R; ++r) y[out_idx<N, SI>(base + r * NS)] = v[r];
  }
  __syncthreads();
}

// Complex FFT with ping-pong buffers: the result ends in `b` after an odd
// number of passes (Plan<N>::nst), back in `a` after an even number.
template <typename T, int S, int N, int L, int NT, int LSC>
__device__ __forceinline__ void cfft_pp(cx<T> *a, cx<T> *b, const T *__restrict__ tw) {
  using P = Plan<N>;
  pass_pp<T, S, N, 0, L, NT, LSC>(a, b, tw);
  if constexpr (P::nst > 1) pass_pp<T, S, N, 1, L, NT, LSC>(b, a, tw);
  if constexpr (P::nst > 2) pass_pp<T, S, N, 2, L, NT, LSC>(a, b, tw);
}

// post_fwd from src (complex FFT of the packed lines) into dst.
template <typename T, int N, int L, int NT, int LS>
__device__ __forceinline__ void post_fwd_pp(const T *__restrict__ src, T *__restrict__ dst,
                                            int kind, const T *__restrict__ twh,
                                            const T *__restrict__ twq) {
  constexpr int NP = N / 2 + 1;
  constexpr int TOT = L * NP;
  constexpr int PER = (TOT + NT - 1) / NT;
  const bool trig = is_trig(kind), sine = is_sine(kind);
  // two items in flight per thread: their LDS reads overlap, at a bounded
  // register cost
#pragma unroll 2
  for (int u = 0; u < PER; ++u) {
    const int t = threadIdx.x + u * NT;
    if (TOT % NT != 0 && t >= TOT) continue;
    const int l = t / NP, k = t - (t / NP) * NP;
    const cx<T> *Z = reinterpret_cast<const cx<T> *>(src + l * LS);
    T *Y = dst + l * LS;
    cx<T> Xk, Xm;
    split_pair(Z[k], Z[k == 0 ? 0 : N - k], cx<T>{twh[2 * k], twh[2 * k + 1]}, Xk, Xm);
    T v0, v1;
    int p0, p1;
    fwd_out<T, N>(Xk, k, trig, sine, twq, v0, p0, v1, p1);
    Y[p0] = v0;
    if (p1 >= 0) Y[p1] = v1;
    if (2 * k != N) {
      fwd_out<T, N>(Xm, N - k, trig, sine, twq, v0, p0, v1, p1);
      Y[p0] = v0;
      if (p1 >= 0) Y[p1] = v1;
    }
  }
  __syncthreads();
}

// pre_inv from src (real lines) into dst (complex Z').
template <typename T, int N, int L, int NT, int LS>
__device__ __forceinline__ void pre_inv_pp(const T *__restrict__ src, T *__restrict__ dst,
                                           int kind, const T *__restrict__ twh,
                                           const T *__restrict__ twq) {
  constexpr int NP = N / 2 + 1;
  constexpr int TOT = L * NP;
  constexpr int PER = (TOT + NT - 1) / NT;
  const bool trig = is_trig(kind), sine = is_sine(kind);
#pragma unroll 2
  for (int u = 0; u < PER; ++u) {
    const int t = threadIdx.x + u * NT;
    if (TOT % NT != 0 && t >= TOT) continue;
    const int l = t / NP, k = t - (t / NP) * NP;
    const T *X = src + l * LS;
    cx<T> *Z = reinterpret_cast<cx<T> *>(dst + l * LS);
    cx<T> Xk, Xm, Zk, Zm;
    if (!trig) {
      Xk = hc<T, N>(X, k);
      Xm = hc<T, N>(X, N - k);
    } else {
      Xk = vprime<T, N>(X, k, twq, sine);
      Xm = vprime<T, N>(X, N - k, twq, sine);
    }
    join_pair(Xk, Xm, cx<T>{twh[2 * k], -twh[2 * k + 1]}, Zk, Zm);
    Z[k] = Zk;
    if (k != 0 && 2 * k != N) Z[N - k] = Zm;
  }
  __syncthreads();
}

}  // namespace tile
}  // namespace flutas
