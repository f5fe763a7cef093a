// Internal declarations shared by the HIP translation units of
// libflutas_amd.so.  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/flutas_amd.h"

namespace flutas {

// ---- error handling ----------------------------------------------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define FL_HIP(call)                                                          \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                     \
      return ::flutas::fail(FLUTAS_EDEVICE, std::string(#call) + ": " +       \
                                                hipGetErrorString(e_));       \
  } while (0)

#define FL_TRY(call)                                                          \
  do {                                                                        \
    int r_ = (call);                                                          \
    if (r_ != FLUTAS_OK) return r_;                                           \
  } while (0)

hipStream_t current_stream();
int check_device();  // EDEVICE unless a gfx950 device is current

// ---- r2r line transforms -------------------------------------------------
// Transform kinds (FFTW numbering, src/fftw.f90:38-61).
enum Kind {
  K_R2HC = 0, K_HC2R = 1,
  K_REDFT00 = 3, K_REDFT01 = 4, K_REDFT10 = 5, K_REDFT11 = 6,
  K_RODFT00 = 7, K_RODFT01 = 8, K_RODFT10 = 9, K_RODFT11 = 10
};
// Length of the real DFT that evaluates a kind on lines of n values: the
// symmetric extension 2(n-1) for REDFT00 and 2(n+1) for RODFT00, n
// otherwise.
inline int dft_length(int kind, int n) {
  return kind == K_REDFT00 ? 2 * (n - 1) : (kind == K_RODFT00 ? 2 * (n + 1) : n);
}

constexpr int kMaxStages = 16;

// Device-side plan of one real line length n: a complex Stockham FFT of
// length N = n/2 (radix list), and the twiddle tables of the real split and
// of the cosine/sine pre/post phases.
template <typename T>
struct LinePlan {
  int n, N, nst;
  int radix[kMaxStages];
  const T *tw;   // 2N values: exp(-2 pi i m / N), m < N (re, im)
  const T *twh;  // 2(N+1):  exp(-2 pi i k / n), k <= N
  const T *twq;  // 2(n+1):  exp(-i pi k / (2n)), k <= n
  const T *twe;  // 2N:      exp(-i pi (4m+1) / (4n)), m < N (DCT-IV/DST-IV)
  // 4(N+1): exp(-i pi (4m+1) / (4n)), m <= N, then exp(-i pi k / n), k <= N
  // (the DCT-IV/DST-IV tables of the tile kernels, sized as twh)
  const T *t4;
};

// Host-owned device tables for one length, in double and float.
struct LineTables {
  int n = 0, N = 0, nst = 0;
  int radix[kMaxStages] = {};
  double *d_tw = nullptr, *d_twh = nullptr, *d_twq = nullptr, *d_twe = nullptr;
  float *f_tw = nullptr, *f_twh = nullptr, *f_twq = nullptr, *f_twe = nullptr;
  double *d_t4 = nullptr;
  float *f_t4 = nullptr;
  // per-pass twiddle tables of the fused kernels (set by
  // build_fused_twiddles when the length has a compiled plan)
  double *d_ptw = nullptr;
  float *f_ptw = nullptr;
  LinePlan<double> dplan() const;
  LinePlan<float> fplan() const;
  int build(int n);
  void release();
};

// Geometry of a batch of lines: element e of line (l1,l2) lives at
// base + l1*s1 + l2*s2 + e*se (separately for input and output).
struct LineGeom {
  int n;              // line length
  int L1, L2;         // line counts
  long long is1, is2, ise;
  long long os1, os2, ose;
};

// Launch one batched line transform: in -> out (may alias only if the
// geometries are identical).  Tin/Tout storage, T arithmetic.
template <typename T, typename Tin, typename Tout>
int launch_lines(const LineTables &tab, int kind, const Tin *in, Tout *out,
                 const LineGeom &g, T scale, hipStream_t s);

// ---- z tridiagonal solves ------------------------------------------------
// Thomas solves of the line path (zsolve.hip): CK holds the elimination
// checkpoints, thomas_checkpoint_doubles values.
size_t thomas_checkpoint_doubles(int nx, int ny, int nz, bool periodic);
template <typename T>
int launch_thomas(T *W, double *CK, const double *lamx, const double *lamy,
                  const double *a, const double *b, const double *c, int nx,
                  int ny, int nz, int pin, hipStream_t s);
template <typename T>
int launch_thomas_periodic(T *W, double *CK, const double *lamx, const double *lamy,
                           const double *a, const double *b, const double *c, int nx,
                           int ny, int nz, int pin, hipStream_t s);

// ---- fused four-pass pipeline --------------------------------------------
// Layout of a work field, j contiguous: the x-slabs are grouped in blocks
// of B = 2^lb slabs (B divides nx); element (i, j, k) sits at
//   W[(i >> lb)*wib + (i & (B-1))*wii + k*wk + j],
// wib = nz*B*ny, wii = ny, wk = B*ny.  B >= nx is the plane-major layout
// (an x-tile of P1/P4 at one k touches one 2-D plane), B = 1 the slab-major
// one (the planes of an x-slab of P2/P3 form one contiguous stream).
struct WStride {
  long long wib, wii, wk;
  int lb;
  __host__ __device__ long long ioff(int i) const {
    return (long long)(i >> lb) * wib + (long long)(i & ((1 << lb) - 1)) * wii;
  }
};
// blocks of 2^lb slabs (2^lb must divide nx)
__host__ __device__ inline WStride slab_blocked(int ny, int nz, int lb) {
  return {(long long)nz * ny << lb, ny, (long long)ny << lb, lb};
}
__host__ __device__ inline WStride plane_major(int nx, int ny) {
  return {0, ny, (long long)nx * ny, 30};
}
// The layout of the FP64 single solve.  Measured on one MI355X: with
// 512-point y-lines (4 KB rows; 512^3), blocks of 4 to 128 slabs run the
// solve 3-4% faster than plane-major and blocks of 2 or 1 (slab-major) do
// not; with 256- or 1024-point y-lines (256x256x512, 1024x1024x512) blocks
// of 8 are 1-2% slower, so those stay plane-major.  The FP32 chain of the
// refined mode is 2% faster plane-major, and the batched solves measure the
// same either way (they stay plane-major).
__host__ __device__ inline WStride fused_layout_fp64(int nx, int ny, int nz) {
  if (ny != 512) return plane_major(nx, ny);
  int lb = 0;
  while (lb < 3 && ((nx >> lb) & 1) == 0) ++lb;
  return slab_blocked(ny, nz, lb);
}
bool fused_supported(int n);
int fused_group_planes(int ny, int elem_bytes);
// Workgroups per CU of the chunked P2 (*r2) and P3 (*r3) slab kernels for
// y-lines of ny points in storage of elem_bytes (occupancy calculator).
int fused_slab_residency(int ny, int elem_bytes, int periodic, int *r2, int *r3);
// Tables and side arrays of the z-chunked slab passes (fused_chunk.hpp):
// the planes of an x-slab are cut into nchunk chunks of cpl planes (a
// multiple of the plane group; the last chunk runs to nz).  Index forms:
//   ckd, ckm, ckq  [(i * ngrp + g) * ny + j]   value at the plane before
//                  plane group g: d, the chunk's carry multiplier m, the
//                  periodic q2
//   tab            [((i * nchunk + c) * zc_tab_slots + s) * ny + j]
//                  slots M, SW, N (periodic: p2 of the plane after chunk c)
//   den            [i * ny + j]  periodic: the closing row's denominator
//   side           [((i * nchunk + c) * zc_side_slots + s) * ny + j]
//                  slots E, S (periodic: y of the closing plane, last chunk)
constexpr int kMaxZChunks = 8;
struct ZChunks {
  int nchunk, cpl, ngrp;
  const double *ckd, *ckm, *ckq, *tab, *den;
  double *side;
};
__host__ __device__ constexpr int zc_tab_slots(bool per) { return per ? 4 : 3; }
__host__ __device__ constexpr int zc_side_slots(bool per) { return per ? 3 : 2; }
// Tables of a plan's chunking (one thread per column; mem holds
// 3 ngrp + nchunk tab slots + 1 + nchunk side slots values per column).
size_t zchunk_doubles(int nx, int ny, int ngrp, int nchunk, bool per);
int build_zchunk_tables(double *mem, ZChunks *zc, int nx, int ny, int nz, int msys, int gp,
                        int cpl, int nchunk, bool per, int pin, int elem_bytes,
                        const double *lamx, const double *lamy, const double *a,
                        const double *b, const double *c, hipStream_t s);
template <typename T>
int launch_yz_fwd_zc(const LineTables &ty, int kind, T *Wt, WStride ws, const ZChunks &zc,
                     int nx, int nz, int msys, const double *lamx, const double *lamy,
                     const double *a, const double *b, const double *c, int pin, int periodic,
                     hipStream_t s);
template <typename T>
int launch_zy_bwd_zc(const LineTables &ty, int kind, T *Wt, WStride ws, const ZChunks &zc,
                     int nx, int nz, int msys, const double *lamx, const double *lamy,
                     const double *a, const double *b, const double *c, int pin, int periodic,
                     hipStream_t s);
// Uploads the per-pass twiddle tables of the compiled plan of tab.n.
int build_fused_twiddles(LineTables &tab);
// absmax_part: one max|in| partial per workgroup (or null); nf_flag: set
// to nf_seq when the input holds a NaN or an infinity (or null).
template <typename T, typename Tin>
int launch_x_fwd(const LineTables &tx, int kind, const Tin *in, long long is1,
                 long long is2, T *Wt, WStride ws, int ny, int nz,
                 double *absmax_part, int *nf_flag, int nf_seq, hipStream_t s);
// workgroups of launch_x_fwd (one max|in| partial each when absmax_part)
int x_fwd_blocks(const LineTables &tx, int elem_bytes, int ny, int nz);
template <typename T, typename Tout>
int launch_x_bwd(const LineTables &tx, int kind, const T *Wt, WStride ws,
                 Tout *out, long long os1, long long os2, int ny, int nz,
                 T scale, int accumulate, hipStream_t s);
template <typename T>
int launch_yz_fwd(const LineTables &ty, int kind, T *Wt, WStride ws,
                  double *ck, int nx, int nz, int nzs, const double *lamx,
                  const double *lamy, const double *a, const double *b,
                  const double *c, int pin, int periodic, hipStream_t s);
template <typename T>
int launch_zy_bwd(const LineTables &ty, int kind, T *Wt, WStride ws,
                  const double *ck, int nx, int nz, int nzs, const double *lamx,
                  const double *lamy, const double *a, const double *b,
                  const double *c, int periodic, hipStream_t s);

// ---- batched pipeline: one persistent launch over a work queue ---------
// Stages of the batched solve (see batch.hip): the four passes of a solve,
// then, in the refined mode, the residual, the four passes of the
// correction solve and the residual check.
enum BatchStage {
  BS_X_FWD = 0, BS_YZ_FWD = 1, BS_ZY_BWD = 2, BS_X_BWD = 3,
  BS_RESID = 4, BS_X_FWD_R = 5, BS_YZ_FWD_R = 6, BS_ZY_BWD_R = 7,
  BS_X_BWD_ACC = 8, BS_CHECK = 9
};
constexpr int kBatchStages = 10;
struct BatchProblem {
  int nx, ny, nz;
  int kx_f, kx_b, ky_f, ky_b;      // r2r kinds of the four passes
  int per, pin;                    // periodic z; singular column pinned
  int tj_v, gp_v;                  // tuned variants (0 full, 1 half, 2 quarter)
  int kxp;                         // planes of one x-tile item (tuned)
  int active;                      // refined: still above its tolerance
  int rpi;                         // residual items: rows per item
  int cnt[kBatchStages];           // items of each stage (completion target)
  char cbc[8];                     // boundary conditions (residual operator)
  double scale;                    // normfft
  double dx2, dy2;                 // dli^2 (residual operator)
  long long hx, hxy;               // strides of the halo-1 fields
  const double *f;                 // right-hand side (interior origin)
  double *p;                       // solution (interior origin)
  void *W;                         // work field (FP64, or FP32 refined)
  float *R;                        // residual (refined, dense nx*ny*nz)
  double *ck;                      // z checkpoints
  const double *lamx, *lamy, *a, *b, *c;
  const double *twx, *twhx, *twqx, *twy, *twhy, *twqy;
  const float *ftwx, *ftwhx, *ftwqx, *ftwy, *ftwhy, *ftwqy;
};
struct BatchWork {
  int stage, prob, u, v;  // x stages: (tile, plane); slabs: (slab, 0);
                          // residual: (row group, plane chunk)
};
// Arguments of one pass launch: the pass's items and the per-problem
// maxima ([3 prob + m] bits of max |f|, |r| after the solve, |r| after the
// correction).
struct BatchPassArgs {
  const BatchProblem *P;
  const BatchWork *items;
  int nitems;
  unsigned long long *mx;
};
// line lengths the batched kernels are compiled for
bool batch_supported(int n);
// lines per x tile and planes per slab group of a length and variant
int batch_tile_lines(int nx, int elem_bytes, int v);
int batch_group_planes(int ny, int elem_bytes, int v);
// residual items: rows per item (nx <= 256), planes per item
int batch_resid_rows(int nx);
int batch_resid_planes();
// dynamic LDS of one item body: stage_kind 0 x tile (length n = nx),
// 1 slab (n = ny)
size_t batch_item_lds(int stage_kind, int n, int elem_bytes, int v);
// raises the dynamic-LDS limit of the pass kernel and returns the
// workgroups per CU it reaches
int batch_persistent_setup(int elem_bytes, size_t lds_bytes, int *blocks_per_cu);
int launch_batch_pass(int elem_bytes, const BatchPassArgs &A, int grid, size_t lds_bytes,
                      hipStream_t s);
// per-length launch shape: one launch of stage `stage` over items whose
// line length (nx for x stages, ny for slab stages) is n, one workgroup per
// item; the residual stages take every problem's items in one launch
int launch_batch_segment(int elem_bytes, const BatchPassArgs &A, int stage, int n,
                         size_t lds_bytes, hipStream_t s);
// raises the dynamic-LDS limit of that launch's kernel
int batch_segment_setup(int elem_bytes, int stage, int n, size_t lds_bytes);
// out[0..n) = maxima of the solve (bits in mx), mx cleared
int launch_batch_publish(unsigned long long *mx, double *out, int n, hipStream_t s);

// ---- refinement helpers ----------------------------------------------------
struct OperatorDesc {
  int nx, ny, nz;
  char cbc[6];
  double dx2, dy2;               // dli^2
  const double *a, *b, *c;       // device z coefficients
};
// r = f - L p on the interior (halo-1 arrays f, p; dense r of type Tr),
// and per-block max|r| partials (nblocks returned through *nparts).
template <typename Tr>
int launch_residual(const OperatorDesc &op, const double *f, const double *p,
                    Tr *r, double *partial, int *nparts, hipStream_t s);
// p(interior) += r (dense FP32 correction from the refined solve)
int launch_accumulate(const float *r, double *p_halo, int nx, int ny, int nz,
                      hipStream_t s);
int launch_absmax(const double *f_halo, int nx, int ny, int nz, double *partial,
                  int *nparts, hipStream_t s);
// flag = seq when the interior of the halo-1 field f holds a NaN or an
// infinity (flag untouched otherwise)
int launch_flag_nonfinite(const double *f_halo, int nx, int ny, int nz, int *flag, int seq,
                          hipStream_t s);
// partial count of launch_residual and launch_absmax
size_t residual_blocks(int nx, int ny, int nz);
// out[0] = max of n partial maxima (single workgroup)
int launch_reduce_max(const double *part, int n, double *out, hipStream_t s);

}  // namespace flutas
