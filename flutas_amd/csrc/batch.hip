// Batched pipeline: many independent Poisson problems of uneven sizes and
// mixed boundary conditions, each pass of the pipeline one persistent
// launch over all problems.
//
// Small problems cannot fill the GPU on their own (a slab pass has one
// workgroup per x-slab, e.g. 64 for a 64-wide grid), so every pass is a list
// of work items (problem, tile or slab) over every problem and every line
// length; a fixed grid of workgroups (as many as the CUs hold at once)
// walks the list, each running the body compiled for the item's line
// length and tuned variant.  The items of a pass are independent; the
// passes follow each other as launches on one stream, captured into one
// hipGraph per solve.
//
// Why one launch per pass and not one for the whole solve: an item of a
// later pass reads what items of the previous pass wrote on other CUs of
// other XCDs.  Inside one launch that needs a device-scope release and
// acquire around every item, and on this GPU those write back and
// invalidate the XCD's whole L2 (buffer_wbl2 / buffer_inv): a single
// persistent launch with per-problem completion counters solved the
// 64-problem set correctly in 36.9 ms, against 3.6 ms for launches per
// pass and length.  A kernel boundary orders the passes at no such cost.
//
// Stages (FP64 mode: X_FWD..X_BWD; refined mode: all ten):
//   X_FWD     f (FP64, halo 1) -> x r2r -> W (max|f| per problem, refined)
//   YZ_FWD    y r2r + forward elimination (checkpoints)
//   ZY_BWD    back substitution + inverse y r2r
//   X_BWD     inverse x r2r * normfft -> p
//   RESID     r = f - L p (FP64) -> R (FP32, dense), max|r|
//   X_FWD_R .. ZY_BWD_R   the correction solve of R
//   X_BWD_ACC p += S(r)
//   CHECK     max|f - L p| (no store)
// A correction round (RESID..CHECK) is a second graph, replayed while a
// problem's relative residual is above its tolerance; a last one-workgroup
// launch publishes the per-problem maxima and clears them.
#include "fused_body.hpp"

namespace flutas {

using tile::cx;

constexpr int kPNT = 256;          // threads of a persistent workgroup
constexpr int kResidPlanes = 16;   // planes a residual lane walks

// Tuned variants of the tile lines and the slab plane group: 0 full,
// 1 half, 2 quarter of the LDS-sized default.
template <typename T, int NC>
__host__ __device__ constexpr int tj_of(int v) {
  constexpr int f = tile_lines<T, NC>();
  return v == 0 ? f : (v == 1 ? (f > 1 ? f / 2 : 1) : (f > 3 ? f / 4 : 1));
}
template <typename T, int NC>
__host__ __device__ constexpr int gp_of(int v) {
  constexpr int f = group_planes<T, NC>();
  return v == 0 ? f : (v == 1 ? (f > 1 ? f / 2 : 1) : (f > 3 ? f / 4 : 1));
}

#define FL_BATCH_SIZES(X) X(16) X(24) X(32) X(48) X(64) X(96) X(128)

// ---- device helpers ----------------------------------------------------

// max of non-negative doubles (and NaN) through their bit patterns: for
// values >= 0 the unsigned order of the bits is the numeric order, and a
// NaN's bits exceed +inf's, so a NaN wins (and is reported)
__device__ __forceinline__ void atomic_max_abs(unsigned long long *slot, double v) {
  const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
  __hip_atomic_fetch_max(slot, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- item bodies ---------------------------------------------------------
// Each body is a function of its own (not inlined into the kernel): its
// registers are allocated for that body alone, where one inlined switch
// over every body spilled heavily.  The bodies take the dynamic LDS from
// the kernel's extern __shared__ array themselves, so their tile accesses
// stay LDS instructions.

// A pointer known to address global memory: the item bodies are called
// functions, whose pointer arguments would otherwise be generic (flat
// instructions, which also count against the LDS counter).
// (through an integer: a cast of the generic pointer itself would fold
// back to the generic pointer)
template <typename T>
__device__ __forceinline__ T *as_global(T *p) {
  return (T *)((__attribute__((address_space(1))) T *)(uintptr_t)p);
}

template <typename T>
__device__ __forceinline__ T *lds_base() {
  extern __shared__ __align__(16) double dyn_raw[];
  return reinterpret_cast<T *>(dyn_raw);
}

template <typename T>
__device__ __forceinline__ const T *tw_x(const BatchProblem &q, int which);
template <>
__device__ __forceinline__ const double *tw_x<double>(const BatchProblem &q, int which) {
  return as_global(which == 0 ? q.twx : (which == 1 ? q.twhx : q.twqx));
}
template <>
__device__ __forceinline__ const float *tw_x<float>(const BatchProblem &q, int which) {
  return as_global(which == 0 ? q.ftwx : (which == 1 ? q.ftwhx : q.ftwqx));
}
template <typename T>
__device__ __forceinline__ const T *tw_y(const BatchProblem &q, int which);
template <>
__device__ __forceinline__ const double *tw_y<double>(const BatchProblem &q, int which) {
  return as_global(which == 0 ? q.twy : (which == 1 ? q.twhy : q.twqy));
}
template <>
__device__ __forceinline__ const float *tw_y<float>(const BatchProblem &q, int which) {
  return as_global(which == 0 ? q.ftwy : (which == 1 ? q.ftwhy : q.ftwqy));
}

// x forward: item (tile u, planes v*q.kxp ..).  FROM_R: the dense FP32
// residual, else the FP64 halo right-hand side (with max|f| into *slot when
// slot is not null).  The twiddles are staged once per item.
template <typename T, int NC, int TJ, bool FROM_R, int NT>
__device__ __forceinline__ void item_x_fwd(const BatchProblem &qin, int u, int v, double *slot) {
  const BatchProblem &q = *as_global(&qin);
  T *const dyn = lds_base<T>();
  T *W = as_global(static_cast<T *>(q.W));
  const WStride ws = plane_major(q.nx, q.ny);
  T *stw = dyn + TJ * line_stride<T, 2 * NC>();
  const int k0 = v * q.kxp, k1 = min(q.nz, k0 + q.kxp);
  double am = 0.0;
  for (int k = k0; k < k1; ++k) {
    // the previous plane's transposed store has read the tile
    if (k > k0) __syncthreads();
    if constexpr (FROM_R) {
      x_fwd_tile<T, float, NC, TJ, NT>(as_global(q.R), q.nx, (long long)q.nx * q.ny, W, ws,
                                         q.ny, q.kx_f, tw_x<T>(q, 0), tw_x<T>(q, 1),
                                         tw_x<T>(q, 2), nullptr, dyn, stw, u * TJ, k, nullptr, 0,
                                         k == k0);
    } else {
      x_fwd_tile<T, double, NC, TJ, NT>(as_global(q.f), q.hx, q.hxy, W, ws, q.ny, q.kx_f,
                                          tw_x<T>(q, 0), tw_x<T>(q, 1), tw_x<T>(q, 2), slot, dyn,
                                          stw, u * TJ, k, nullptr, 0, k == k0);
      if (slot && threadIdx.x == 0) am = nan_max(am, *slot);
    }
  }
  if (slot && threadIdx.x == 0) *slot = am;
}

template <typename T, int NC, int TJ, bool ACC, int NT>
__device__ __forceinline__ void item_x_bwd(const BatchProblem &qin, int u, int v) {
  const BatchProblem &q = *as_global(&qin);
  T *const dyn = lds_base<T>();
  const T *W = as_global(static_cast<const T *>(q.W));
  T *stw = dyn + TJ * line_stride<T, 2 * NC>();
  const int k0 = v * q.kxp, k1 = min(q.nz, k0 + q.kxp);
  for (int k = k0; k < k1; ++k) {
    // the previous plane's output has left the tile
    if (k > k0) __syncthreads();
    x_bwd_tile<T, double, NC, TJ, NT, ACC>(W, plane_major(q.nx, q.ny), as_global(q.p), q.hx,
                                             q.hxy, q.ny, q.kx_b, (T)q.scale, tw_x<T>(q, 0),
                                             tw_x<T>(q, 1), tw_x<T>(q, 2), dyn, stw, u * TJ, k,
                                             k == k0);
  }
}

template <typename T, int NC, int GP, bool PER, int NT>
__device__ __forceinline__ void item_yz(const BatchProblem &qin, int i) {
  const BatchProblem &q = *as_global(&qin);
  T *const dyn = lds_base<T>();
  yz_fwd_slab<T, NC, GP, NT, PER>(as_global(static_cast<T *>(q.W)), plane_major(q.nx, q.ny),
                                    as_global(q.ck), q.nx, q.nz, q.nz, q.ky_f, as_global(q.lamx),
                                    as_global(q.lamy), as_global(q.a), as_global(q.b),
                                    as_global(q.c), q.pin,
                                    tw_y<T>(q, 0), tw_y<T>(q, 1), tw_y<T>(q, 2), dyn,
                                    dyn + 2 * GP * line_stride<T, 2 * NC>(), i);
}
template <typename T, int NC, int GP, bool PER, int NT>
__device__ __forceinline__ void item_zy(const BatchProblem &qin, int i) {
  const BatchProblem &q = *as_global(&qin);
  T *const dyn = lds_base<T>();
  zy_bwd_slab<T, NC, GP, NT, PER>(as_global(static_cast<T *>(q.W)), plane_major(q.nx, q.ny),
                                    as_global(q.ck), q.nx, q.nz, q.nz, q.ky_b, as_global(q.lamx),
                                    as_global(q.lamy), as_global(q.a), as_global(q.b),
                                    as_global(q.c),
                                    tw_y<T>(q, 0), tw_y<T>(q, 1), tw_y<T>(q, 2), dyn,
                                    dyn + 2 * GP * line_stride<T, 2 * NC>(), i);
}

// r = f - L p for rows jg*rpi .. jg*rpi + rpi - 1 and planes
// kc*kResidPlanes .., one lane per (i, row), the lane walking z with the
// next plane in flight.  L: second differences in x and y with the
// homogeneous ghost rules (P wrap, N p0 = p1, D p0 = -p1) and the z rows
// a, b, c (periodic wrap for PP).  STORE writes r (FP32, dense).  The
// lane's max|r| goes to the workgroup max in *slot.
template <bool STORE>
__device__ __forceinline__ void item_residual(const BatchProblem &qin, int jg, int kc,
                                              double *slot) {
#pragma clang fp contract(off)
  const BatchProblem &q = *as_global(&qin);
  const int nx = q.nx, ny = q.ny, nz = q.nz;
  const int t = threadIdx.x;
  const int i = t % nx, jj = t / nx;
  const int j = jg * q.rpi + jj;
  const int k0 = kc * kResidPlanes, k1 = min(nz, k0 + kResidPlanes);
  const bool zper = q.cbc[4] == 'P';
  double am = 0.0;
  if (jj < q.rpi && j < ny) {
    const long long sy = q.hx, sz = q.hxy;
    const long long c0 = (i + 1) + sy * (j + 1);
    const long long ixm = (i > 0) ? -1 : ((q.cbc[0] == 'P') ? (nx - 1) : 0);
    const long long ixp = (i < nx - 1) ? 1 : ((q.cbc[1] == 'P') ? -(nx - 1) : 0);
    const long long iym = (j > 0) ? -sy : ((q.cbc[2] == 'P') ? sy * (ny - 1) : 0);
    const long long iyp = (j < ny - 1) ? sy : ((q.cbc[3] == 'P') ? -sy * (ny - 1) : 0);
    const double *f = as_global(q.f - (1 + q.hx + q.hxy));
    const double *p = as_global(q.p - (1 + q.hx + q.hxy));
    const double *za = as_global(q.a), *zb = as_global(q.b), *zc = as_global(q.c);
    float *R = as_global(q.R);
    struct Pl {
      double f, c, xm, xp, ym, yp;
    };
    auto gather = [&](int k) {
      const long long c = c0 + sz * (k + 1);
      Pl v;
      v.f = f[c];
      v.c = p[c];
      v.xm = p[c + ixm];
      v.xp = p[c + ixp];
      v.ym = p[c + iym];
      v.yp = p[c + iyp];
      return v;
    };
    auto ghost = [&](double nb, double pc, char bc, bool inside) {
      if (inside || bc == 'P') return nb;
      return bc == 'N' ? pc : -pc;
    };
    auto knext = [&](int k) { return (k + 1 < nz) ? k + 1 : (zper ? k + 1 - nz : nz - 1); };
    double pzm = (k0 > 0) ? p[c0 + sz * k0] : (zper ? p[c0 + sz * nz] : 0.0);
    Pl cur = gather(k0);
    Pl nxt = gather(knext(k0));
    float hold[kResidPlanes];
#pragma unroll
    for (int s = 0; s < kResidPlanes; ++s) {
      const int k = k0 + s;
      if (k >= k1) continue;
      Pl nxt2 = nxt;
      if (k + 1 < k1) nxt2 = gather(knext(k + 1));
      const double pc = cur.c;
      const double lx = (ghost(cur.xp, pc, q.cbc[1], i < nx - 1) - 2.0 * pc +
                         ghost(cur.xm, pc, q.cbc[0], i > 0)) * q.dx2;
      const double ly = (ghost(cur.yp, pc, q.cbc[3], j < ny - 1) - 2.0 * pc +
                         ghost(cur.ym, pc, q.cbc[2], j > 0)) * q.dy2;
      const double pzp = (k < nz - 1 || zper) ? nxt.c : 0.0;
      double lz = zb[k] * pc;
      if (k > 0 || zper) lz += za[k] * pzm;
      if (k < nz - 1 || zper) lz += zc[k] * pzp;
      const double res = cur.f - ((lx + ly) + lz);
      hold[s] = (float)res;
      am = nan_max(am, res);
      pzm = pc;
      cur = nxt;
      nxt = nxt2;
    }
    if constexpr (STORE) {
#pragma unroll
      for (int s = 0; s < kResidPlanes; ++s)
        if (k0 + s < k1) R[i + (long long)nx * (j + (long long)ny * (k0 + s))] = hold[s];
    }
  }
  block_max_store<kPNT>(am, slot);
}

// Called (not inlined) bodies of the persistent pass kernel: each is
// register-allocated for itself.
template <typename T, int NC, int TJ, bool FROM_R>
__device__ __noinline__ void body_x_fwd(const BatchProblem &q, int u, int v, double *slot) {
  item_x_fwd<T, NC, TJ, FROM_R, kPNT>(q, u, v, slot);
}
template <typename T, int NC, int TJ, bool ACC>
__device__ __noinline__ void body_x_bwd(const BatchProblem &q, int u, int v) {
  item_x_bwd<T, NC, TJ, ACC, kPNT>(q, u, v);
}
template <typename T, int NC, int GP, bool PER>
__device__ __noinline__ void body_yz(const BatchProblem &q, int i) {
  item_yz<T, NC, GP, PER, kPNT>(q, i);
}
template <typename T, int NC, int GP, bool PER>
__device__ __noinline__ void body_zy(const BatchProblem &q, int i) {
  item_zy<T, NC, GP, PER, kPNT>(q, i);
}
template <bool STORE>
__device__ __noinline__ void body_residual(const BatchProblem &q, int jg, int kc, double *slot) {
  item_residual<STORE>(q, jg, kc, slot);
}

// ---- dispatch of one item ------------------------------------------------

template <typename T, int NC>
__device__ __forceinline__ void run_x(const BatchProblem &q, const BatchWork &w, double *slot) {
  const int v = q.tj_v;
#define FL_TJ(V, ...)                                                       \
  if (v == V) {                                                             \
    constexpr int TJ = tj_of<T, NC>(V);                                     \
    __VA_ARGS__;                                                            \
    return;                                                                 \
  }
  switch (w.stage) {
    case BS_X_FWD:
      FL_TJ(0, (body_x_fwd<T, NC, TJ, false>(q, w.u, w.v, slot)))
      FL_TJ(1, (body_x_fwd<T, NC, TJ, false>(q, w.u, w.v, slot)))
      FL_TJ(2, (body_x_fwd<T, NC, TJ, false>(q, w.u, w.v, slot)))
      return;
    case BS_X_BWD:
      FL_TJ(0, (body_x_bwd<T, NC, TJ, false>(q, w.u, w.v)))
      FL_TJ(1, (body_x_bwd<T, NC, TJ, false>(q, w.u, w.v)))
      FL_TJ(2, (body_x_bwd<T, NC, TJ, false>(q, w.u, w.v)))
      return;
    default:
      break;
  }
  if constexpr (sizeof(T) == 4) {
    switch (w.stage) {
      case BS_X_FWD_R:
        FL_TJ(0, (body_x_fwd<T, NC, TJ, true>(q, w.u, w.v, slot)))
        FL_TJ(1, (body_x_fwd<T, NC, TJ, true>(q, w.u, w.v, slot)))
        FL_TJ(2, (body_x_fwd<T, NC, TJ, true>(q, w.u, w.v, slot)))
        return;
      case BS_X_BWD_ACC:
        FL_TJ(0, (body_x_bwd<T, NC, TJ, true>(q, w.u, w.v)))
        FL_TJ(1, (body_x_bwd<T, NC, TJ, true>(q, w.u, w.v)))
        FL_TJ(2, (body_x_bwd<T, NC, TJ, true>(q, w.u, w.v)))
        return;
      default:
        break;
    }
  }
#undef FL_TJ
}

template <typename T, int NC>
__device__ __forceinline__ void run_slab(const BatchProblem &q, const BatchWork &w) {
  const int v = q.gp_v;
  const bool fwd = (w.stage == BS_YZ_FWD || w.stage == BS_YZ_FWD_R);
#define FL_GP(V)                                                            \
  if (v == V) {                                                             \
    constexpr int GP = gp_of<T, NC>(V);                                     \
    if (fwd) {                                                              \
      if (q.per) body_yz<T, NC, GP, true>(q, w.u);                     \
      else body_yz<T, NC, GP, false>(q, w.u);                          \
    } else {                                                                \
      if (q.per) body_zy<T, NC, GP, true>(q, w.u);                     \
      else body_zy<T, NC, GP, false>(q, w.u);                          \
    }                                                                       \
    return;                                                                 \
  }
  FL_GP(0)
  FL_GP(1)
  FL_GP(2)
#undef FL_GP
}

template <typename T>
__device__ __forceinline__ void run_item(const BatchProblem &q, const BatchWork &w,
                                         double *slot) {
  const bool xs = (w.stage == BS_X_FWD || w.stage == BS_X_BWD || w.stage == BS_X_FWD_R ||
                   w.stage == BS_X_BWD_ACC);
  const bool slab = (w.stage == BS_YZ_FWD || w.stage == BS_ZY_BWD ||
                     w.stage == BS_YZ_FWD_R || w.stage == BS_ZY_BWD_R);
  if (xs) {
    switch (q.nx / 2) {
#define FL_CASE(NCV) \
  case NCV:          \
    run_x<T, NCV>(q, w, slot); \
    break;
      FL_BATCH_SIZES(FL_CASE)
#undef FL_CASE
      default:
        break;
    }
  } else if (slab) {
    switch (q.ny / 2) {
#define FL_CASE(NCV) \
  case NCV:          \
    run_slab<T, NCV>(q, w); \
    break;
      FL_BATCH_SIZES(FL_CASE)
#undef FL_CASE
      default:
        break;
    }
  } else {
    // the residual stages exist in the refined (FP32) queues only
    if constexpr (sizeof(T) == 4) {
      if (w.stage == BS_RESID) body_residual<true>(q, w.u, w.v, slot);
      else if (w.stage == BS_CHECK) body_residual<false>(q, w.u, w.v, slot);
    }
  }
}

// max slot of a stage (-1: none): 0 max|f|, 1 max|r| after the solve,
// 2 max|r| after the correction
__device__ __forceinline__ int max_slot(int stage) {
  return stage == BS_X_FWD ? 0 : (stage == BS_RESID ? 1 : (stage == BS_CHECK ? 2 : -1));
}

// One pass: the persistent workgroups walk the pass's items with a grid
// stride (the list holds the largest problems' items first, so the stride
// spreads the long items over the workgroups).
template <typename T>
__global__ void __launch_bounds__(kPNT, 2)
    k_batch_pass(BatchPassArgs A) {
  __shared__ double s_max;
  for (int idx = blockIdx.x; idx < A.nitems; idx += gridDim.x) {
    // the previous item's LDS reads are complete before this item stages
    // into the same memory
    __syncthreads();
    const BatchWork w = A.items[idx];
    const BatchProblem &q = A.P[w.prob];
    // a problem already at its tolerance sits out a correction round
    if (!q.active && w.stage >= BS_RESID) continue;
    const int ms = max_slot(w.stage);
    const bool with_max = ms >= 0 && sizeof(T) == 4;
    run_item<T>(q, w, with_max ? &s_max : nullptr);
    if (with_max && threadIdx.x == 0) atomic_max_abs(A.mx + 3 * w.prob + ms, s_max);
  }
}

// ---- per-length launches --------------------------------------------------
// The other launch shape of a pass: one launch per line length present, one
// workgroup per item, the body inlined into a kernel of its own (tiles 256
// threads, slabs one column per thread in whole waves), so that short lines
// leave no lane idle and no body pays for another's registers.  The
// launches of one pass run side by side (the host forks them over
// streams).
template <int NC, int STAGE>
__host__ __device__ constexpr int seg_threads() {
  return (STAGE == BS_YZ_FWD || STAGE == BS_ZY_BWD || STAGE == BS_YZ_FWD_R ||
          STAGE == BS_ZY_BWD_R)
             ? batch_slab_threads<NC>()
             : kPNT;
}

template <typename T, int NC, int STAGE>
__global__ void __launch_bounds__((seg_threads<NC, STAGE>()))
    k_batch_seg(BatchPassArgs A) {
  constexpr int NT = seg_threads<NC, STAGE>();
  __shared__ double s_max;
  const BatchWork w = A.items[blockIdx.x];
  const BatchProblem &q = A.P[w.prob];
  if (!q.active && STAGE >= BS_RESID) return;
  if constexpr (STAGE == BS_X_FWD || STAGE == BS_X_FWD_R) {
    constexpr bool FROM_R = STAGE == BS_X_FWD_R;
    double *slot = (!FROM_R && sizeof(T) == 4) ? &s_max : nullptr;
    if (q.tj_v == 0) item_x_fwd<T, NC, tj_of<T, NC>(0), FROM_R, NT>(q, w.u, w.v, slot);
    else if (q.tj_v == 1) item_x_fwd<T, NC, tj_of<T, NC>(1), FROM_R, NT>(q, w.u, w.v, slot);
    else item_x_fwd<T, NC, tj_of<T, NC>(2), FROM_R, NT>(q, w.u, w.v, slot);
    if (slot && threadIdx.x == 0) atomic_max_abs(A.mx + 3 * w.prob, s_max);
  } else if constexpr (STAGE == BS_X_BWD || STAGE == BS_X_BWD_ACC) {
    constexpr bool ACC = STAGE == BS_X_BWD_ACC;
    if (q.tj_v == 0) item_x_bwd<T, NC, tj_of<T, NC>(0), ACC, NT>(q, w.u, w.v);
    else if (q.tj_v == 1) item_x_bwd<T, NC, tj_of<T, NC>(1), ACC, NT>(q, w.u, w.v);
    else item_x_bwd<T, NC, tj_of<T, NC>(2), ACC, NT>(q, w.u, w.v);
  } else {
    constexpr bool FWD = STAGE == BS_YZ_FWD || STAGE == BS_YZ_FWD_R;
#define FL_SLAB(V)                                                           \
  {                                                                          \
    constexpr int GP = gp_of<T, NC>(V);                                      \
    if (FWD) {                                                               \
      if (q.per) item_yz<T, NC, GP, true, NT>(q, w.u);                       \
      else item_yz<T, NC, GP, false, NT>(q, w.u);                            \
    } else {                                                                 \
      if (q.per) item_zy<T, NC, GP, true, NT>(q, w.u);                       \
      else item_zy<T, NC, GP, false, NT>(q, w.u);                            \
    }                                                                        \
  }
    if (q.gp_v == 0) FL_SLAB(0)
    else if (q.gp_v == 1) FL_SLAB(1)
    else FL_SLAB(2)
#undef FL_SLAB
  }
}

// The residual stages: one launch over the items of every problem.
template <bool STORE>
__global__ void __launch_bounds__(kPNT)
    k_batch_resid(BatchPassArgs A) {
  __shared__ double s_max;
  const BatchWork w = A.items[blockIdx.x];
  const BatchProblem &q = A.P[w.prob];
  if (!q.active) return;
  item_residual<STORE>(q, w.u, w.v, &s_max);
  if (threadIdx.x == 0) atomic_max_abs(A.mx + 3 * w.prob + (STORE ? 1 : 2), s_max);
}

// out[t] = the maxima gathered by the passes of one solve or round; the
// slots are cleared for the next.
__global__ void __launch_bounds__(256)
    k_batch_publish(unsigned long long *__restrict__ mx, double *__restrict__ out, int n) {
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    out[t] = __longlong_as_double((long long)mx[t]);
    mx[t] = 0ull;
  }
}

// ---------------------------------------------------------------- host ---

bool batch_supported(int n) {
  if (n % 2) return false;
  switch (n / 2) {
#define FL_CASE(NCV) case NCV:
    FL_BATCH_SIZES(FL_CASE)
#undef FL_CASE
    return true;
    default:
      return false;
  }
}

int batch_tile_lines(int nx, int elem_bytes, int v) {
  switch (nx / 2) {
#define FL_CASE(NCV) \
  case NCV:          \
    return elem_bytes == 8 ? tj_of<double, NCV>(v) : tj_of<float, NCV>(v);
    FL_BATCH_SIZES(FL_CASE)
#undef FL_CASE
    default:
      return 0;
  }
}

int batch_group_planes(int ny, int elem_bytes, int v) {
  switch (ny / 2) {
#define FL_CASE(NCV) \
  case NCV:          \
    return elem_bytes == 8 ? gp_of<double, NCV>(v) : gp_of<float, NCV>(v);
    FL_BATCH_SIZES(FL_CASE)
#undef FL_CASE
    default:
      return 0;
  }
}

int batch_resid_rows(int nx) { return std::max(1, kPNT / nx); }
int batch_resid_planes() { return kResidPlanes; }

// LDS of one item body: the tile (TJ lines) or the two slab buffers (2 GP
// lines), then the staged twiddles.
size_t batch_item_lds(int stage_kind, int n, int elem_bytes, int v) {
  switch (n / 2) {
#define FL_CASE(NCV)                                                                      \
  case NCV: {                                                                             \
    if (elem_bytes == 8) {                                                                \
      constexpr size_t ls = line_stride<double, 2 * NCV>(), tw = Tw<double, NCV>::total;  \
      const size_t lines = stage_kind == 0 ? tj_of<double, NCV>(v) : 2 * gp_of<double, NCV>(v); \
      return (lines * ls + tw) * sizeof(double);                                          \
    } else {                                                                              \
      constexpr size_t ls = line_stride<float, 2 * NCV>(), tw = Tw<float, NCV>::total;    \
      const size_t lines = stage_kind == 0 ? tj_of<float, NCV>(v) : 2 * gp_of<float, NCV>(v); \
      return (lines * ls + tw) * sizeof(float);                                           \
    }                                                                                     \
  }
    FL_BATCH_SIZES(FL_CASE)
#undef FL_CASE
    default:
      return 0;
  }
}

template <typename T>
static const void *pass_fn() {
  return (const void *)k_batch_pass<T>;
}

int batch_persistent_setup(int elem_bytes, size_t lds_bytes, int *blocks_per_cu) {
  const void *fn = elem_bytes == 8 ? pass_fn<double>() : pass_fn<float>();
  FL_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes));
  int nb = 0;
  FL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, kPNT, lds_bytes));
  if (nb < 1) return fail(FLUTAS_EDEVICE, "the batched pass kernel does not fit a CU");
  *blocks_per_cu = nb;
  return FLUTAS_OK;
}

int launch_batch_pass(int elem_bytes, const BatchPassArgs &A, int grid, size_t lds_bytes,
                      hipStream_t s) {
  if (A.nitems <= 0 || grid <= 0) return FLUTAS_OK;
  if (elem_bytes == 8)
    hipLaunchKernelGGL(k_batch_pass<double>, dim3(grid), dim3(kPNT), lds_bytes, s, A);
  else
    hipLaunchKernelGGL(k_batch_pass<float>, dim3(grid), dim3(kPNT), lds_bytes, s, A);
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

// Kernel of a per-length launch (T by elem_bytes, line length n, stage).
template <typename T>
static const void *seg_fn(int n, int stage) {
  switch (n / 2) {
#define FL_CASE(NCV)                                                        \
  case NCV:                                                                 \
    switch (stage) {                                                        \
      case BS_X_FWD: return (const void *)k_batch_seg<T, NCV, BS_X_FWD>;    \
      case BS_YZ_FWD: return (const void *)k_batch_seg<T, NCV, BS_YZ_FWD>;  \
      case BS_ZY_BWD: return (const void *)k_batch_seg<T, NCV, BS_ZY_BWD>;  \
      case BS_X_BWD: return (const void *)k_batch_seg<T, NCV, BS_X_BWD>;    \
      default: break;                                                       \
    }                                                                       \
    if constexpr (sizeof(T) == 4) {                                         \
      switch (stage) {                                                      \
        case BS_X_FWD_R: return (const void *)k_batch_seg<T, NCV, BS_X_FWD_R>;   \
        case BS_YZ_FWD_R: return (const void *)k_batch_seg<T, NCV, BS_YZ_FWD_R>; \
        case BS_ZY_BWD_R: return (const void *)k_batch_seg<T, NCV, BS_ZY_BWD_R>; \
        case BS_X_BWD_ACC: return (const void *)k_batch_seg<T, NCV, BS_X_BWD_ACC>; \
        default: break;                                                     \
      }                                                                     \
    }                                                                       \
    return nullptr;
    FL_BATCH_SIZES(FL_CASE)
#undef FL_CASE
    default:
      return nullptr;
  }
}

static int seg_threads_rt(int n, int stage) {
  const bool slab = stage == BS_YZ_FWD || stage == BS_ZY_BWD || stage == BS_YZ_FWD_R ||
                    stage == BS_ZY_BWD_R;
  return slab ? ((n + 63) / 64) * 64 : kPNT;
}

int launch_batch_segment(int elem_bytes, const BatchPassArgs &A, int stage, int n,
                         size_t lds_bytes, hipStream_t s) {
  if (A.nitems <= 0) return FLUTAS_OK;
  if (stage == BS_RESID || stage == BS_CHECK) {
    if (stage == BS_RESID)
      hipLaunchKernelGGL(k_batch_resid<true>, dim3(A.nitems), dim3(kPNT), 0, s, A);
    else
      hipLaunchKernelGGL(k_batch_resid<false>, dim3(A.nitems), dim3(kPNT), 0, s, A);
    FL_HIP(hipGetLastError());
    return FLUTAS_OK;
  }
  const void *fn = elem_bytes == 8 ? seg_fn<double>(n, stage) : seg_fn<float>(n, stage);
  if (!fn) return fail(FLUTAS_EINVAL, "no batched kernel for this stage and length");
  void *args[] = {(void *)&A};
  FL_HIP(hipLaunchKernel(fn, dim3(A.nitems), dim3(seg_threads_rt(n, stage)), args, lds_bytes, s));
  return FLUTAS_OK;
}

int batch_segment_setup(int elem_bytes, int stage, int n, size_t lds_bytes) {
  if (stage == BS_RESID || stage == BS_CHECK) return FLUTAS_OK;
  const void *fn = elem_bytes == 8 ? seg_fn<double>(n, stage) : seg_fn<float>(n, stage);
  if (!fn) return fail(FLUTAS_EINVAL, "no batched kernel for this stage and length");
  FL_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes));
  return FLUTAS_OK;
}

int launch_batch_publish(unsigned long long *mx, double *out, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_publish, dim3(1), dim3(256), 0, s, mx, out, n);
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

}  // namespace flutas
