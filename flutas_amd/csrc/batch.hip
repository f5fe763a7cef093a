// Batched four-pass pipeline: many independent FP64 Poisson problems of
// uneven sizes and mixed boundary conditions.
//
// Small problems cannot fill the GPU on their own (a slab pass has one
// workgroup per x-slab, e.g. 64 for a 64-wide grid), so each pass of a
// batch is one launch per line length present, whose workgroups are the
// tiles or slabs of all problems of that length.  A workgroup reads its
// (problem, tile) pair from the item table and runs the body compiled for
// that line length (one kernel instance per length and pass), with the
// tile or slab in dynamic LDS sized for that length.  Each problem also carries two tuned
// choices, the halved x tile and the halved slab plane group, which the
// batch autotuner sets per size.
#include "fused_body.hpp"

namespace flutas {

// Tile passes run 256 threads, as the single-problem kernels at these line
// lengths; slab passes one thread per column (batch_slab_threads), where
// the single-problem kernels run 256, so a batched solve may differ from
// the single solve in the last bits (different unrolling of the same
// arithmetic).
constexpr int kBatchThreads = 256;

#define FL_BATCH_SIZES(X) X(16) X(24) X(32) X(48) X(64) X(96) X(128)

template <int NC>
constexpr int tj_variant(int half) {
  constexpr int f = tile_lines<double, NC>();
  return half ? (f > 1 ? f / 2 : 1) : f;
}
template <int NC>
constexpr int gp_variant(int half) {
  constexpr int f = group_planes<double, NC>();
  return half ? (f > 1 ? f / 2 : 1) : f;
}

// The tile passes are launched once per x-line length (every item of a
// launch has nx = 2*NC), so each length gets a kernel of its own registers.
template <int NC>
__global__ void __launch_bounds__(kBatchThreads)
    k_batch_x_fwd(const BatchProblem *__restrict__ P, const BatchItem *__restrict__ items) {
  extern __shared__ __align__(16) double dyn[];
  const BatchItem w = items[blockIdx.x];
  const BatchProblem &q = P[w.prob];
  if (q.tj_half) {
    constexpr int TJ = tj_variant<NC>(1);
    x_fwd_tile<double, double, NC, TJ, kBatchThreads>(
        q.f, q.hx, q.hxy, q.W, plane_major(q.nx, q.ny), q.ny, q.kx_f, q.twx, q.twhx, q.twqx,
        nullptr, dyn,
        dyn + TJ * line_stride<double, 2 * NC>(), w.u * TJ, w.v);
  } else {
    constexpr int TJ = tj_variant<NC>(0);
    x_fwd_tile<double, double, NC, TJ, kBatchThreads>(
        q.f, q.hx, q.hxy, q.W, plane_major(q.nx, q.ny), q.ny, q.kx_f, q.twx, q.twhx, q.twqx,
        nullptr, dyn,
        dyn + TJ * line_stride<double, 2 * NC>(), w.u * TJ, w.v);
  }
}

template <int NC>
__global__ void __launch_bounds__(kBatchThreads)
    k_batch_x_bwd(const BatchProblem *__restrict__ P, const BatchItem *__restrict__ items) {
  extern __shared__ __align__(16) double dyn[];
  const BatchItem w = items[blockIdx.x];
  const BatchProblem &q = P[w.prob];
  if (q.tj_half) {
    constexpr int TJ = tj_variant<NC>(1);
    x_bwd_tile<double, double, NC, TJ, kBatchThreads, false>(
        q.W, plane_major(q.nx, q.ny), q.p, q.hx, q.hxy, q.ny, q.kx_b, q.scale, q.twx, q.twhx,
        q.twqx, dyn,
        dyn + TJ * line_stride<double, 2 * NC>(), w.u * TJ, w.v);
  } else {
    constexpr int TJ = tj_variant<NC>(0);
    x_bwd_tile<double, double, NC, TJ, kBatchThreads, false>(
        q.W, plane_major(q.nx, q.ny), q.p, q.hx, q.hxy, q.ny, q.kx_b, q.scale, q.twx, q.twhx,
        q.twqx, dyn,
        dyn + TJ * line_stride<double, 2 * NC>(), w.u * TJ, w.v);
  }
}

template <int NC, int GP, bool PER>
__device__ __forceinline__ void batch_yz(const BatchProblem &q, int i, double *dyn) {
  yz_fwd_slab<double, NC, GP, batch_slab_threads<NC>(), PER>(
      q.W, plane_major(q.nx, q.ny), q.ck, q.nx, q.nz, q.nz, q.ky_f, q.lamx, q.lamy, q.a, q.b, q.c,
      q.pin, q.twy, q.twhy,
      q.twqy, dyn, dyn + 2 * GP * line_stride<double, 2 * NC>(), i);
}
template <int NC, int GP, bool PER>
__device__ __forceinline__ void batch_zy(const BatchProblem &q, int i, double *dyn) {
  zy_bwd_slab<double, NC, GP, batch_slab_threads<NC>(), PER>(
      q.W, plane_major(q.nx, q.ny), q.ck, q.nx, q.nz, q.nz, q.ky_b,
      q.lamx, q.lamy, q.a, q.b, q.c, q.twy, q.twhy, q.twqy,
      dyn, dyn + 2 * GP * line_stride<double, 2 * NC>(), i);
}

// The slab passes are launched once per y-line length (every item of a
// launch has ny = 2*NC), with batch_slab_threads<NC>() threads.
template <int NC>
__global__ void __launch_bounds__(batch_slab_threads<NC>())
    k_batch_yz_fwd(const BatchProblem *__restrict__ P, const BatchItem *__restrict__ items) {
  extern __shared__ __align__(16) double dyn[];
  const BatchItem w = items[blockIdx.x];
  const BatchProblem &q = P[w.prob];
  if (q.gp_half) {
    if (q.per) batch_yz<NC, gp_variant<NC>(1), true>(q, w.u, dyn);
    else batch_yz<NC, gp_variant<NC>(1), false>(q, w.u, dyn);
  } else {
    if (q.per) batch_yz<NC, gp_variant<NC>(0), true>(q, w.u, dyn);
    else batch_yz<NC, gp_variant<NC>(0), false>(q, w.u, dyn);
  }
}

template <int NC>
__global__ void __launch_bounds__(batch_slab_threads<NC>())
    k_batch_zy_bwd(const BatchProblem *__restrict__ P, const BatchItem *__restrict__ items) {
  extern __shared__ __align__(16) double dyn[];
  const BatchItem w = items[blockIdx.x];
  const BatchProblem &q = P[w.prob];
  if (q.gp_half) {
    if (q.per) batch_zy<NC, gp_variant<NC>(1), true>(q, w.u, dyn);
    else batch_zy<NC, gp_variant<NC>(1), false>(q, w.u, dyn);
  } else {
    if (q.per) batch_zy<NC, gp_variant<NC>(0), true>(q, w.u, dyn);
    else batch_zy<NC, gp_variant<NC>(0), false>(q, w.u, dyn);
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

int batch_tile_lines(int nx, int half) {
  switch (nx / 2) {
#define FL_CASE(NCV) \
  case NCV:          \
    return tj_variant<NCV>(half);
    FL_BATCH_SIZES(FL_CASE)
#undef FL_CASE
    default:
      return 0;
  }
}

int batch_group_planes(int ny, int half) {
  switch (ny / 2) {
#define FL_CASE(NCV) \
  case NCV:          \
    return gp_variant<NCV>(half);
    FL_BATCH_SIZES(FL_CASE)
#undef FL_CASE
    default:
      return 0;
  }
}

size_t batch_lds_bytes(int stage, int n) {
  switch (n / 2) {
#define FL_CASE(NCV)                                                                    \
  case NCV: {                                                                           \
    constexpr size_t ls = line_stride<double, 2 * NCV>();                               \
    constexpr size_t tw = Tw<double, NCV>::total;                                       \
    if (stage == BATCH_X_FWD || stage == BATCH_X_BWD)                                   \
      return (tj_variant<NCV>(0) * ls + tw) * sizeof(double);                           \
    return (2 * gp_variant<NCV>(0) * ls + tw) * sizeof(double);                         \
  }
    FL_BATCH_SIZES(FL_CASE)
#undef FL_CASE
    default:
      return 0;
  }
}

int batch_set_lds(const size_t lds_bytes[4]) {
#define FL_CASE(NCV)                                                                     \
  FL_HIP(hipFuncSetAttribute((const void *)k_batch_x_fwd<NCV>,                           \
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes[0])); \
  FL_HIP(hipFuncSetAttribute((const void *)k_batch_x_bwd<NCV>,                           \
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes[3])); \
  FL_HIP(hipFuncSetAttribute((const void *)k_batch_yz_fwd<NCV>,                          \
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes[1])); \
  FL_HIP(hipFuncSetAttribute((const void *)k_batch_zy_bwd<NCV>,                          \
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes[2]));
  FL_BATCH_SIZES(FL_CASE)
#undef FL_CASE
  return FLUTAS_OK;
}

int launch_batch_stage(int stage, const BatchProblem *probs, const BatchItem *items,
                       int nitems, size_t lds_bytes, int n, hipStream_t s) {
  if (nitems <= 0) return FLUTAS_OK;
  if (stage < BATCH_X_FWD || stage > BATCH_X_BWD) return fail(FLUTAS_EINVAL, "bad batch stage");
  {
    switch (n / 2) {
#define FL_CASE(NCV)                                                                          \
  case NCV:                                                                                   \
    if (stage == BATCH_X_FWD)                                                                 \
      hipLaunchKernelGGL(k_batch_x_fwd<NCV>, dim3(nitems), dim3(kBatchThreads), lds_bytes, s, \
                         probs, items);                                                       \
    else if (stage == BATCH_X_BWD)                                                            \
      hipLaunchKernelGGL(k_batch_x_bwd<NCV>, dim3(nitems), dim3(kBatchThreads), lds_bytes, s, \
                         probs, items);                                                       \
    else if (stage == BATCH_YZ_FWD)                                                           \
      hipLaunchKernelGGL(k_batch_yz_fwd<NCV>, dim3(nitems),                                   \
                         dim3(batch_slab_threads<NCV>()), lds_bytes, s, probs, items); \
    else                                                                                      \
      hipLaunchKernelGGL(k_batch_zy_bwd<NCV>, dim3(nitems),                                   \
                         dim3(batch_slab_threads<NCV>()), lds_bytes, s, probs, items); \
    break;
      FL_BATCH_SIZES(FL_CASE)
#undef FL_CASE
      default:
        return fail(FLUTAS_EINVAL, "bad batch line length");
    }
  }
  FL_HIP(hipGetLastError());
  return FLUTAS_OK;
}

}  // namespace flutas
