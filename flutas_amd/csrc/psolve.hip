// Poisson plan (initsolver/fftini restated on the host) and the solve
// pipeline, plus the C-ABI plumbing (errors, stream, device memory).
//
// Pipeline of one FP64 solve on a one-GPU pencil (src/solver_gpu.f90:
// 107-229 replaced; the reference's transposes become strided line
// geometries, no copies):
//   x-forward  p(interior, halo layout) -> W        r2r along x
//   y-forward  W -> W                               r2r along y
//   z-solve    W -> W                               Thomas per (i,j)
//   y-backward W -> W
//   x-backward W -> p(interior) * normfft
// The refined mode runs the same chain in FP32 storage/arithmetic (the z
// elimination stays FP64) and corrects with FP64 residuals r = f - L p
// until max|r| / max|f| <= tol.
#include "internal.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <cstdlib>

namespace flutas {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

static hipStream_t g_stream = nullptr;
hipStream_t current_stream() { return g_stream; }

int check_device() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(FLUTAS_EDEVICE, std::string("no HIP device: ") + hipGetErrorString(e));
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) return fail(FLUTAS_EDEVICE, hipGetErrorString(e));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(FLUTAS_EDEVICE, std::string("libflutas_amd is built for gfx950, device is ") + prop.gcnArchName);
  return FLUTAS_OK;
}

// ---- initsolver restated (src/initsolver.f90) ------------------------------
// pi as the reference defines it: acos(-1._rp) (param.f90:99)
static double ref_pi() { return std::acos(-1.0); }

// eigenvalues, 'c' branch (src/initsolver.f90:122-186, CPU ordering); the
// face-centred PP, ND and DN pairs use the same formulas (:136-140, :180-183)
static void eigen_c(int n, char b0, char b1, std::vector<double> &lam) {
  const double pi = ref_pi();
  lam.assign(n, 0.0);
  for (int l = 1; l <= n; ++l) {
    double s = 0.0;
    if (b0 == 'P' && b1 == 'P') s = std::sin((1.0 * (l - 1)) * pi / (1.0 * n));
    else if (b0 == 'N' && b1 == 'N') s = std::sin((1.0 * (l - 1)) * pi / (2.0 * n));
    else if (b0 == 'D' && b1 == 'D') s = std::sin((1.0 * (l - 0)) * pi / (2.0 * n));
    else s = std::sin((1.0 * (2 * l - 1)) * pi / (4.0 * n));
    lam[l - 1] = -4.0 * (s * s);
  }
}

// tridmatrix (src/initsolver.f90:188-246); zci[k] == dzci(k).  Cell-centred
// ('c') unknowns couple through dzci(k-1), dzci(k); face-centred ('f') ones
// through dzfi(k), dzfi(k+1), and only an N face modifies the diagonal (a D
// face value is no unknown).
static void tridmatrix(char b0, char b1, int n, const double *zci, const double *zfi,
                       char cf, std::vector<double> &a, std::vector<double> &b,
                       std::vector<double> &c) {
  a.resize(n); b.resize(n); c.resize(n);
  for (int k = 1; k <= n; ++k) {
    if (cf == 'c') {
      a[k - 1] = zfi[k] * zci[k - 1];
      c[k - 1] = zfi[k] * zci[k];
    } else {
      a[k - 1] = zfi[k] * zci[k];
      c[k - 1] = zfi[k + 1] * zci[k];
    }
  }
  for (int k = 0; k < n; ++k) b[k] = -(a[k] + c[k]);
  auto fac = [](char t) { return t == 'P' ? 0.0 : (t == 'D' ? -1.0 : 1.0); };
  if (cf == 'c') {
    b[0] = b[0] + fac(b0) * a[0];
    b[n - 1] = b[n - 1] + fac(b1) * c[n - 1];
  } else {
    if (b0 == 'N') b[0] = b[0] + fac(b0) * a[0];
    if (b1 == 'N') b[n - 1] = b[n - 1] + fac(b1) * c[n - 1];
  }
}

// bc_rhs (src/initsolver.f90:248-298)
static void bc_rhs(const char *cbc, const double *bc, const double dlc[2],
                   const double dlf[2], char cf, long plane, std::vector<double> &rhs) {
  rhs.assign(2 * plane, 0.0);
  for (int ib = 0; ib < 2; ++ib) {
    double factor = 0.0;
    const double sgn = (ib == 0) ? 1.0 : -1.0;
    if (cbc[ib] == 'D') factor = (cf == 'c') ? -2.0 * bc[ib] : -bc[ib];
    else if (cbc[ib] == 'N') factor = (cf == 'c') ? sgn * dlc[ib] * bc[ib] : sgn * dlf[ib] * bc[ib];
    const double v = factor / dlc[ib] / dlf[ib];
    std::fill(rhs.begin() + ib * plane, rhs.begin() + (ib + 1) * plane, v);
  }
}

// find_fft (src/fft.f90:233-291) for the kinds built here; -1 for an
// unknown pair, -2 for the face-centred NN and DD pairs (REDFT00/RODFT00,
// whose eigenvalues in src/initsolver.f90:164-178 do not match them).
static int find_fft(char b0, char b1, char cf, int &kf, int &kb, double &norm0) {
  const bool P = b0 == 'P' && b1 == 'P', NN = b0 == 'N' && b1 == 'N',
             DD = b0 == 'D' && b1 == 'D', ND = b0 == 'N' && b1 == 'D',
             DN = b0 == 'D' && b1 == 'N';
  if (P) { kf = K_R2HC; kb = K_HC2R; norm0 = 1; return 0; }
  if (cf == 'c') {
    if (NN) { kf = K_REDFT10; kb = K_REDFT01; norm0 = 2; return 0; }
    if (DD) { kf = K_RODFT10; kb = K_RODFT01; norm0 = 2; return 0; }
    if (ND) { kf = K_REDFT11; kb = K_REDFT11; norm0 = 2; return 0; }
    if (DN) { kf = K_RODFT11; kb = K_RODFT11; norm0 = 2; return 0; }
    return -1;
  }
  if (ND) { kf = K_REDFT10; kb = K_REDFT01; norm0 = 2; return 0; }
  if (DN) { kf = K_RODFT01; kb = K_RODFT10; norm0 = 2; return 0; }
  return (NN || DD) ? -2 : -1;
}

// forward r2r kinds whose tile transform (post_fwd) the fused P1/P2 run
static bool fused_forward_kind(int k) {
  return k == K_R2HC || k == K_REDFT10 || k == K_RODFT10 || k == K_REDFT11 || k == K_RODFT11;
}

}  // namespace flutas

using namespace flutas;

struct flutas_psolve {
  int nx, ny, nz;
  // planes of the z system: nz, or nz - 1 for a face-centred z grid with a
  // D top face, whose last plane is the boundary (src/solver_cpu.f90:71-72)
  int nzs;
  char cf[3];  // c_or_f
  char cbc[6];
  double bc[6];
  double dli[3];
  int precision;
  double tol;
  int max_iter = 12;
  int kinds[4];
  double normfft;
  std::vector<double> lamx, lamy, a, b, c, rhsbx, rhsby, rhsbz;
  LineTables tx, ty;
  double *d_lamx = nullptr, *d_lamy = nullptr, *d_a = nullptr, *d_b = nullptr, *d_c = nullptr;
  double *d_rhsbx = nullptr, *d_rhsby = nullptr, *d_rhsbz = nullptr;
  double *W64 = nullptr;   // spectral work field (FP64 mode)
  float *W32 = nullptr;    // spectral work field (refined mode)
  double *TK = nullptr;    // z elimination checkpoints of the line path
  double *F = nullptr;     // copy of the right-hand side (refined mode)
  float *R32 = nullptr;    // residual (refined mode)
  // reduction partials: [0, nres) residual / absmax workgroups,
  // [nres, nres + nx1) max|f| of the FP32 P1, then two result slots
  double *part = nullptr;
  size_t nres = 0, nx1 = 0;
  bool fused = false;      // four-pass pipeline (fused.hip) in use
  double *CK = nullptr;    // z elimination checkpoints of the fused pipeline
  // z-chunked slab passes (fused_chunk.hpp): chunks per x-slab (1: whole
  // slabs), their tables, and the element size of the chain they serve
  int zchunks = 1;
  ZChunks zc{};
  double *zc_mem = nullptr;
  int zc_elem = 0;
  int last_iters = 0;
  double last_relres = -1.0;
  // NaN/Inf detection of the FP64 solve: P1 (or a flag pass on the line
  // pipeline) stores the solve's sequence number into a host-mapped word
  // when the right-hand side holds a non-finite value; the synchronous
  // entry points compare it after their stream synchronisation.
  int *nf_host = nullptr, *nf_dev = nullptr;
  int seq = 0;
  // stage timing (flutas_psolve_profile)
  bool profiling = false;
  std::vector<hipEvent_t> ev;
  std::vector<std::string> stage_names;
  int nmark = 0;
};

namespace {

template <typename T>
int dev_upload(T **d, const std::vector<T> &h) {
  FL_HIP(hipMalloc((void **)d, std::max<size_t>(1, h.size()) * sizeof(T)));
  if (!h.empty())
    FL_HIP(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return FLUTAS_OK;
}

void release(flutas_psolve *h) {
  for (void *p : {(void *)h->d_lamx, (void *)h->d_lamy, (void *)h->d_a,
                  (void *)h->d_b, (void *)h->d_c, (void *)h->d_rhsbx,
                  (void *)h->d_rhsby, (void *)h->d_rhsbz, (void *)h->W64,
                  (void *)h->W32, (void *)h->TK, (void *)h->F,
                  (void *)h->R32, (void *)h->part, (void *)h->CK, (void *)h->zc_mem})
    if (p) (void)hipFree(p);
  for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
  h->ev.clear();
  if (h->nf_host) (void)hipHostFree(h->nf_host);
  h->nf_host = h->nf_dev = nullptr;
  h->tx.release();
  h->ty.release();
}

// Geometry helpers: the halo-1 field (0:nx+1,0:ny+1,0:nz+1) and the dense
// (nx,ny,nz) work field.
struct Geo {
  long long hx, hxy, interior;  // halo strides and offset of (1,1,1)
};
Geo halo_geo(const flutas_psolve *h) {
  Geo g;
  g.hx = h->nx + 2;
  g.hxy = (long long)(h->nx + 2) * (h->ny + 2);
  g.interior = 1 + g.hx + g.hxy;
  return g;
}

LineGeom geom_x_from_halo(const flutas_psolve *h) {
  Geo q = halo_geo(h);
  LineGeom g;
  g.n = h->nx; g.L1 = h->ny; g.L2 = h->nz;
  g.is1 = q.hx; g.is2 = q.hxy; g.ise = 1;
  g.os1 = h->nx; g.os2 = (long long)h->nx * h->ny; g.ose = 1;
  return g;
}
LineGeom geom_x_dense(const flutas_psolve *h) {
  LineGeom g;
  g.n = h->nx; g.L1 = h->ny; g.L2 = h->nz;
  g.is1 = h->nx; g.is2 = (long long)h->nx * h->ny; g.ise = 1;
  g.os1 = g.is1; g.os2 = g.is2; g.ose = 1;
  return g;
}
LineGeom geom_x_to_halo(const flutas_psolve *h) {
  Geo q = halo_geo(h);
  LineGeom g;
  g.n = h->nx; g.L1 = h->ny; g.L2 = h->nz;
  g.is1 = h->nx; g.is2 = (long long)h->nx * h->ny; g.ise = 1;
  g.os1 = q.hx; g.os2 = q.hxy; g.ose = 1;
  return g;
}
LineGeom geom_y(const flutas_psolve *h) {
  LineGeom g;
  g.n = h->ny; g.L1 = h->nx; g.L2 = h->nz;
  g.is1 = 1; g.is2 = (long long)h->nx * h->ny; g.ise = h->nx;
  g.os1 = 1; g.os2 = g.is2; g.ose = h->nx;
  return g;
}

// Marks the end of a pipeline stage when profiling.
int mark(flutas_psolve *h, const char *name, hipStream_t s) {
  if (!h->profiling) return FLUTAS_OK;
  const size_t m = (size_t)h->nmark;
  if (h->ev.size() < m + 2) {
    hipEvent_t e;
    FL_HIP(hipEventCreate(&e));
    h->ev.push_back(e);
    if (h->ev.size() < m + 2) {
      FL_HIP(hipEventCreate(&e));
      h->ev.push_back(e);
    }
  }
  if (h->stage_names.size() < m + 1) h->stage_names.push_back(name);
  else h->stage_names[m] = name;
  FL_HIP(hipEventRecord(h->ev[m + 1], s));
  ++h->nmark;
  return FLUTAS_OK;
}

template <typename T>
int zsolve(flutas_psolve *h, T *W, hipStream_t s) {
  const bool per = (h->cbc[4] == 'P' && h->cbc[5] == 'P');
  const int pin = (h->cbc[4] == 'N' && h->cbc[5] == 'N') || per;
  if (per)
    return launch_thomas_periodic<T>(W, h->TK, h->d_lamx, h->d_lamy, h->d_a, h->d_b, h->d_c,
                                     h->nx, h->ny, h->nz, pin, s);
  // a face-centred z grid with a D top face solves nz - 1 planes; the last
  // one is left as it is (src/solver_cpu.f90:71-77)
  return launch_thomas<T>(W, h->TK, h->d_lamx, h->d_lamy, h->d_a, h->d_b, h->d_c,
                          h->nx, h->ny, h->nzs, pin, s);
}

// The slab passes P2/P3 of the fused pipeline: z-chunked when the plan
// has chunk tables for this chain's element size, whole slabs otherwise.
template <typename T>
int slab_fwd(flutas_psolve *h, T *W, WStride ws, int pin, int per, hipStream_t s) {
  if (h->zchunks > 1 && h->zc_elem == (int)sizeof(T))
    return launch_yz_fwd_zc<T>(h->ty, h->kinds[2], W, ws, h->zc, h->nx, h->nz,
                               per ? h->nz - 1 : h->nzs, h->d_lamx, h->d_lamy, h->d_a, h->d_b,
                               h->d_c, pin, per, s);
  return launch_yz_fwd<T>(h->ty, h->kinds[2], W, ws, h->CK, h->nx, h->nz, h->nzs, h->d_lamx,
                          h->d_lamy, h->d_a, h->d_b, h->d_c, pin, per, s);
}
template <typename T>
int slab_bwd(flutas_psolve *h, T *W, WStride ws, int pin, int per, hipStream_t s) {
  if (h->zchunks > 1 && h->zc_elem == (int)sizeof(T))
    return launch_zy_bwd_zc<T>(h->ty, h->kinds[3], W, ws, h->zc, h->nx, h->nz,
                               per ? h->nz - 1 : h->nzs, h->d_lamx, h->d_lamy, h->d_a, h->d_b,
                               h->d_c, pin, per, s);
  return launch_zy_bwd<T>(h->ty, h->kinds[3], W, ws, h->CK, h->nx, h->nz, h->nzs, h->d_lamx,
                          h->d_lamy, h->d_a, h->d_b, h->d_c, per, s);
}

// Chunks per x-slab for the z-chunked slab passes: n = 1 turns them off.
// The planes are cut at multiples of the plane group into n chunks (fewer
// when nz is short; the last chunk runs to nz and holds two planes or
// more), and the tables are computed for the plan's chain.
int set_zchunks(flutas_psolve *h, int n) {
  if (n < 1 || n > kMaxZChunks) return fail(FLUTAS_EINVAL, "z chunks must be 1..8");
  if (h->zc_mem) {
    FL_HIP(hipDeviceSynchronize());
    (void)hipFree(h->zc_mem);
    h->zc_mem = nullptr;
  }
  h->zchunks = 1;
  h->zc = ZChunks{};
  h->zc_elem = 0;
  if (n == 1 || !h->fused) return FLUTAS_OK;
  const int elem = (h->precision == FLUTAS_FP64) ? 8 : 4;
  const int gp = fused_group_planes(h->ny, elem);
  const int nz = h->nz;
  const int per = (h->cbc[4] == 'P');
  const int pin = (h->cbc[4] == 'N' && h->cbc[5] == 'N') || per;
  // planes per chunk: a multiple of gp near nz / n
  const int cpl = gp * std::max(1, (nz / n + gp / 2) / gp);
  int nch = std::min(n, (nz + cpl - 1) / cpl);
  while (nch > 1 && (nch - 1) * cpl > nz - 2) --nch;
  if (nch < 2) return FLUTAS_OK;
  const int ngrp = (nz + gp - 1) / gp;
  const size_t nd = zchunk_doubles(h->nx, h->ny, ngrp, nch, per != 0);
  FL_HIP(hipMalloc((void **)&h->zc_mem, nd * sizeof(double)));
  const int rc = build_zchunk_tables(h->zc_mem, &h->zc, h->nx, h->ny, nz, per ? nz - 1 : h->nzs,
                                     gp, cpl, nch, per != 0, pin, elem, h->d_lamx, h->d_lamy,
                                     h->d_a, h->d_b, h->d_c, nullptr);
  if (rc == FLUTAS_OK && hipDeviceSynchronize() != hipSuccess)
    return fail(FLUTAS_EDEVICE, "z chunk tables");
  if (rc) {
    (void)hipFree(h->zc_mem);
    h->zc_mem = nullptr;
    return rc;
  }
  h->zchunks = nch;
  h->zc_elem = elem;
  return FLUTAS_OK;
}

// Default chunking: the count that fills the GPU's workgroup slots best
// over the two slab passes.  A pass of W = nx * n workgroups, R of which
// the GPU holds at once (CUs x residency per CU from the occupancy
// calculator), runs ceil(W / R) rounds and keeps W / (ceil(W / R) R) of
// the slots busy; the score is the mean of that fill over P2 and P3, ties
// go to the smaller count, and every chunk keeps 32 planes or more.  The
// FP32 chain of the refined mode pays 0.05 of score per chunk beyond the
// first: its chunk tables and side arrays stay FP64 against half the field
// bytes, so a chunk costs it relatively more (refined 256x256x512
// periodic: 1/2/3/4/6/8 chunks 1.71, 1.30, 1.36, 1.31, 1.37, 1.34 ms; the
// unpenalised score took 6, the penalised one takes 2; refined 128^3:
// 0.309, 0.244, 0.230, 0.207, 0.249, 0.225 ms, 4 either way).
// Measured FP64 solves (ms for 1/2/3/4/6/8 chunks; * the count taken):
//   768x768x256 NN-z      2.81  2.69* 2.72  2.71  2.79  2.79
//   640x640x320 NN-z      2.85  2.55  2.56  2.49* 2.59  2.57
//   640x640x256 periodic  2.49  2.13  2.14  2.10* 2.17  2.18
//   384x768x256 DD-z      1.54  1.38  1.42  1.32* 1.38  1.38
//   512^3 NN-z            1.92* 1.90  1.92  1.98  1.96  1.95
//   1024x1024x512 NN-z    9.68* 10.0  10.1  10.1  10.3  10.5
//   256x256x512 periodic  0.78  0.55* 0.63  0.61  0.62  0.60
//   400^3 NN-z            1.26  1.23  1.15  1.19  1.16  1.21  (5 taken)
// and at 128^3: 0.115, 0.082 (2), 0.065 (4, taken), 0.072 (8).
int default_zchunks(const flutas_psolve *h) {
  if (!h->fused) return 1;
  const int elem = (h->precision == FLUTAS_FP64) ? 8 : 4;
  const int per = (h->cbc[4] == 'P');
  int r2 = 0, r3 = 0, dev = 0, ncu = 0;
  if (fused_slab_residency(h->ny, elem, per, &r2, &r3) != FLUTAS_OK || r2 < 1 || r3 < 1 ||
      hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      ncu < 1) {
    (void)hipGetLastError();
    return 1;
  }
  auto fill = [&](long long w, long long r) {
    const long long rounds = (w + r - 1) / r;
    return (double)w / (double)(rounds * r);
  };
  int best = 1;
  double best_score = -1.0;
  for (int n = 1; n <= kMaxZChunks; ++n) {
    if (n > 1 && h->nz / n < 32) break;
    const long long w = (long long)h->nx * n;
    const double score = 0.5 * (fill(w, (long long)ncu * r2) + fill(w, (long long)ncu * r3)) -
                         (elem == 4 ? 0.05 * (n - 1) : 0.0);
    if (score > best_score + 1e-9) {
      best_score = score;
      best = n;
    }
  }
  return best;
}

// One spectral solve in storage/arithmetic T: the right-hand side `in`
// (halo-1 layout at its interior origin when in_halo, else dense nx*ny*nz)
// -> out (interior origin of a halo-1 FP64 field): out = S(in), or
// out += S(in) when accumulate.  The fused FP64 passes keep W in blocks
// of x-slabs (fused_layout_fp64), the FP32 ones plane-major.
template <typename T, typename Tin, typename Tout = double>
int spectral(flutas_psolve *h, const Tin *in, bool in_halo, Tout *out,
             int accumulate, T *W, hipStream_t s, double *absmax_part = nullptr,
             int *nf_flag = nullptr) {
  const Geo q = halo_geo(h);
  const int nx = h->nx, ny = h->ny, nz = h->nz;
  const long long is1 = in_halo ? q.hx : nx;
  const long long is2 = in_halo ? q.hxy : (long long)nx * ny;
  if (h->fused) {
    // the singular column takes p(last) = 0 for NN and for periodic z
    const int per = (h->cbc[4] == 'P');
    const int pin = (h->cbc[4] == 'N' && h->cbc[5] == 'N') || per;
    const WStride ws = sizeof(T) == 8 ? fused_layout_fp64(nx, ny, nz) : plane_major(nx, ny);
    FL_TRY((launch_x_fwd<T, Tin>(h->tx, h->kinds[0], in, is1, is2, W, ws, ny, nz, absmax_part,
                                 nf_flag, h->seq, s)));
    FL_TRY(mark(h, "P1_x_fwd", s));
    FL_TRY(slab_fwd<T>(h, W, ws, pin, per, s));
    FL_TRY(mark(h, "P2_yz_fwd", s));
    FL_TRY(slab_bwd<T>(h, W, ws, pin, per, s));
    FL_TRY(mark(h, "P3_zy_bwd", s));
    FL_TRY((launch_x_bwd<T, Tout>(h->tx, h->kinds[1], W, ws, out, q.hx, q.hxy, ny, nz,
                                  (T)h->normfft, accumulate, s)));
    FL_TRY(mark(h, "P4_x_bwd", s));
    return FLUTAS_OK;
  }
  LineGeom gx = geom_x_dense(h);
  gx.is1 = is1;
  gx.is2 = is2;
  FL_TRY((launch_lines<T, Tin, T>(h->tx, h->kinds[0], in, W, gx, T(1), s)));
  FL_TRY(mark(h, "x_fwd", s));
  FL_TRY((launch_lines<T, T, T>(h->ty, h->kinds[2], W, W, geom_y(h), T(1), s)));
  FL_TRY(mark(h, "y_fwd", s));
  FL_TRY(zsolve<T>(h, W, s));
  FL_TRY(mark(h, "z_solve", s));
  FL_TRY((launch_lines<T, T, T>(h->ty, h->kinds[3], W, W, geom_y(h), T(1), s)));
  FL_TRY(mark(h, "y_bwd", s));
  if (!accumulate) {
    FL_TRY((launch_lines<T, T, Tout>(h->tx, h->kinds[1], W, out, geom_x_to_halo(h),
                                     (T)h->normfft, s)));
  } else if constexpr (sizeof(Tout) == 8) {
    // only the refined mode accumulates: through the dense FP32 residual buffer
    FL_TRY((launch_lines<T, T, T>(h->tx, h->kinds[1], W, (T *)h->R32, geom_x_dense(h),
                                  (T)h->normfft, s)));
    FL_TRY(launch_accumulate(h->R32, out - q.interior, nx, ny, nz, s));
  } else {
    return fail(FLUTAS_EINVAL, "no accumulation into an FP32 field");
  }
  FL_TRY(mark(h, "x_bwd", s));
  return FLUTAS_OK;
}

int solve_fp64(flutas_psolve *h, const double *f, double *p, hipStream_t s) {
  const Geo q = halo_geo(h);
  ++h->seq;
  if (!h->fused) FL_TRY(launch_flag_nonfinite(f, h->nx, h->ny, h->nz, h->nf_dev, h->seq, s));
  FL_TRY((spectral<double, double>(h, f + q.interior, true, p + q.interior, 0, h->W64, s,
                                   nullptr, h->fused ? h->nf_dev : nullptr)));
  h->last_iters = 0;
  h->last_relres = -1.0;
  return FLUTAS_OK;
}

// The solve of an FP32 field (a FluTAS build with _SINGLE_PRECISION, rp =
// FP32: src/types.f90:6-12), in place or out of place: the FP32 chain of
// the refined mode (FP32 transforms, FP64 z elimination) once, no
// refinement, as the reference's single-precision solver runs its whole
// solve in FP32.  The FP32 work field is allocated on first use when the
// plan is an FP64 one.
int solve_f32(flutas_psolve *h, const float *f, float *p, hipStream_t s) {
  if (!h->W32) {
    const size_t N = (size_t)h->nx * h->ny * h->nz;
    FL_HIP(hipMalloc((void **)&h->W32, N * sizeof(float)));
  }
  const Geo q = halo_geo(h);
  ++h->seq;
  FL_TRY((spectral<float, float, float>(h, f + q.interior, true, p + q.interior, 0, h->W32, s,
                                        nullptr, h->fused ? h->nf_dev : nullptr)));
  h->last_iters = 0;
  h->last_relres = -1.0;
  return FLUTAS_OK;
}

OperatorDesc op_desc(const flutas_psolve *h) {
  OperatorDesc op;
  op.nx = h->nx; op.ny = h->ny; op.nz = h->nz;
  std::memcpy(op.cbc, h->cbc, 6);
  op.dx2 = h->dli[0] * h->dli[0];
  op.dy2 = h->dli[1] * h->dli[1];
  op.a = h->d_a; op.b = h->d_b; op.c = h->d_c;
  return op;
}

// Device-side final reduction of the partial maxima (res[0] from the
// residual partials, res[1] from the P1 partials when n1 > 0); one copy
// back and one synchronisation.
int reduce_max(flutas_psolve *h, int nparts, hipStream_t s, double *out,
               int n1 = 0, double *out1 = nullptr) {
  double *res = h->part + h->nres + h->nx1;
  FL_TRY(launch_reduce_max(h->part, nparts, res, s));
  if (n1 > 0) FL_TRY(launch_reduce_max(h->part + h->nres, n1, res + 1, s));
  double host[2] = {0.0, 0.0};
  FL_HIP(hipMemcpyAsync(host, res, (n1 > 0 ? 2 : 1) * sizeof(double), hipMemcpyDeviceToHost, s));
  FL_HIP(hipStreamSynchronize(s));
  *out = host[0];
  if (n1 > 0) *out1 = host[1];
  return FLUTAS_OK;
}

int solve_refined(flutas_psolve *h, const double *fsrc, double *p, hipStream_t s) {
  const Geo q = halo_geo(h);
  const long long nhalo = (long long)(h->nx + 2) * (h->ny + 2) * (h->nz + 2);
  double *pin = p + q.interior;
  const double normf = h->normfft;
  // in place: keep a copy of the right-hand side for the residuals
  const double *F = fsrc;
  if (fsrc == p) {
    FL_HIP(hipMemcpyAsync(h->F, p, sizeof(double) * nhalo, hipMemcpyDeviceToDevice, s));
    F = h->F;
  }
  const double *fin = F + q.interior;
  int np = 0;
  double fmax = 0.0;
  // max|f| comes with the first FP32 transform on the fused path (P1 reads
  // f anyway), from its own pass otherwise
  int n1 = 0;
  if (h->fused) {
    n1 = (int)h->nx1;
  } else {
    FL_TRY(launch_absmax(F, h->nx, h->ny, h->nz, h->part, &np, s));
    FL_TRY(reduce_max(h, np, s, &fmax));
    if (!(fmax == fmax) || std::isinf(fmax)) return fail(FLUTAS_ENONFINITE, "right-hand side holds NaN or Inf");
  }
  // initial FP32 solve
  (void)normf;
  FL_TRY((spectral<float, double>(h, fin, true, pin, 0, h->W32, s,
                                  n1 > 0 ? h->part + h->nres : nullptr)));
  const OperatorDesc op = op_desc(h);
  int it = 0;
  double rel = 0.0;
  // the residual after a correction is first only measured (16 B/pt, no
  // store); it is formed again with its FP32 store when a further
  // correction needs it
  bool stored = true;
  for (;;) {
    FL_TRY(launch_residual<float>(op, F, p, it == 0 ? h->R32 : nullptr, h->part, &np, s));
    FL_TRY(mark(h, it == 0 ? "residual" : "check", s));
    stored = (it == 0);
    double rmax = 0.0;
    FL_TRY(reduce_max(h, np, s, &rmax, n1, &fmax));
    if (n1 > 0) {
      n1 = 0;
      if (!(fmax == fmax) || std::isinf(fmax)) return fail(FLUTAS_ENONFINITE, "right-hand side holds NaN or Inf");
    }
    rel = (fmax > 0.0) ? rmax / fmax : rmax;
    if (!(rel == rel)) return fail(FLUTAS_ENONFINITE, "residual is NaN");
    if (rel <= h->tol) break;
    if (it >= h->max_iter) {
      h->last_iters = it;
      h->last_relres = rel;
      return fail(FLUTAS_ENOTCONV, "refinement did not reach the tolerance");
    }
    if (!stored) {
      FL_TRY(launch_residual<float>(op, F, p, h->R32, h->part, &np, s));
      FL_TRY(mark(h, "residual", s));
    }
    // correction solve on the residual, accumulated into p
    FL_TRY((spectral<float, float>(h, h->R32, false, pin, 1, h->W32, s)));
    ++it;
  }
  h->last_iters = it;
  h->last_relres = rel;
  return FLUTAS_OK;
}

}  // namespace

extern "C" {

const char *flutas_last_error(void) { return g_err.c_str(); }
const char *flutas_version(void) { return "flutas_amd 0.1 (gfx950)"; }

int flutas_device_alloc(size_t bytes, void **ptr) {
  if (!ptr) return fail(FLUTAS_EINVAL, "null pointer");
  FL_HIP(hipMalloc(ptr, std::max<size_t>(bytes, 1)));
  return FLUTAS_OK;
}
int flutas_device_free(void *ptr) {
  if (ptr) FL_HIP(hipFree(ptr));
  return FLUTAS_OK;
}
int flutas_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  FL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return FLUTAS_OK;
}
int flutas_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  FL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return FLUTAS_OK;
}
int flutas_memset_d(void *dst, int value, size_t bytes) {
  FL_HIP(hipMemset(dst, value, bytes));
  return FLUTAS_OK;
}
int flutas_set_stream(void *hip_stream) {
  g_stream = (hipStream_t)hip_stream;
  return FLUTAS_OK;
}
int flutas_synchronize(void) {
  FL_HIP(hipStreamSynchronize(g_stream));
  return FLUTAS_OK;
}
int flutas_pointer_is_device(const void *ptr, int *is_device) {
  if (!is_device) return fail(FLUTAS_EINVAL, "null pointer");
  *is_device = 0;
  if (!ptr) return FLUTAS_OK;
  hipPointerAttribute_t at;
  hipError_t e = hipPointerGetAttributes(&at, ptr);
  if (e != hipSuccess) {
    // plain (unregistered) host memory: not an error for this query
    (void)hipGetLastError();
    return FLUTAS_OK;
  }
  *is_device = (at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged) ? 1 : 0;
  return FLUTAS_OK;
}

int flutas_psolve_create(const int ng[3], const char cbc[6], const double bc[6],
                         const char c_or_f[3], const double dli[3], int nh_d,
                         const double *dzci, const double *dzfi, int precision,
                         double tol, flutas_psolve **out) {
  if (!ng || !cbc || !bc || !c_or_f || !dli || !dzci || !dzfi || !out)
    return fail(FLUTAS_EINVAL, "null pointer");
  *out = nullptr;
  for (int d = 0; d < 3; ++d) {
    if (ng[d] < 2 || (ng[d] & 1)) return fail(FLUTAS_EINVAL, "grid sizes must be even and >= 2 (src/sanity.f90:155)");
    if (c_or_f[d] != 'c' && c_or_f[d] != 'f') return fail(FLUTAS_EINVAL, "c_or_f must be 'c' or 'f'");
  }
  for (int q = 0; q < 6; ++q)
    if (cbc[q] != 'P' && cbc[q] != 'D' && cbc[q] != 'N') return fail(FLUTAS_EINVAL, "BC must be one of P, D, N");
  for (int d = 0; d < 3; ++d)
    if ((cbc[2 * d] == 'P') != (cbc[2 * d + 1] == 'P'))
      return fail(FLUTAS_EINVAL, "periodic BCs must be paired (src/sanity.f90:180-259)");
  if (nh_d < 1) return fail(FLUTAS_EINVAL, "nh_d must be >= 1");
  if (precision != FLUTAS_FP64 && precision != FLUTAS_FP32_REFINE)
    return fail(FLUTAS_EINVAL, "precision must be 64 or 32");
  if (precision == FLUTAS_FP32_REFINE && !(tol > 0.0)) return fail(FLUTAS_EINVAL, "tol must be > 0");
  if (cbc[4] == 'P' && ng[2] < 4) return fail(FLUTAS_EINVAL, "periodic z needs nz >= 4");
  const bool face = (c_or_f[0] == 'f' || c_or_f[1] == 'f' || c_or_f[2] == 'f');
  if (face && precision == FLUTAS_FP32_REFINE)
    return fail(FLUTAS_EUNSUPPORTED, "the refined mode's residual is the cell-centred operator; "
                                     "face-centred ('f') plans are FP64 only");
  const int qz = (c_or_f[2] == 'f' && cbc[5] == 'D') ? 1 : 0;
  if (ng[2] - qz < 2) return fail(FLUTAS_EINVAL, "the z system needs at least two planes");
  int kinds[4];
  double n0x = 0.0, n0y = 0.0;
  for (int d = 0; d < 2; ++d) {
    const int r = find_fft(cbc[2 * d], cbc[2 * d + 1], c_or_f[d], kinds[2 * d], kinds[2 * d + 1],
                           d == 0 ? n0x : n0y);
    if (r == -2)
      return fail(FLUTAS_EUNSUPPORTED, "face-centred NN/DD in x or y (REDFT00/RODFT00 solves) "
                                       "is not built: src/initsolver.f90:164-178 pairs them "
                                       "with eigenvalues of other transforms");
    if (r) return fail(FLUTAS_EINVAL, "unknown boundary-condition pair in x or y");
  }
  FL_TRY(check_device());

  auto *h = new flutas_psolve();
  h->nx = ng[0]; h->ny = ng[1]; h->nz = ng[2];
  h->nzs = ng[2] - qz;
  std::memcpy(h->cf, c_or_f, 3);
  std::memcpy(h->cbc, cbc, 6);
  std::memcpy(h->bc, bc, sizeof(double) * 6);
  std::memcpy(h->dli, dli, sizeof(double) * 3);
  std::memcpy(h->kinds, kinds, sizeof(kinds));
  h->precision = precision;
  h->tol = tol;
  // normfft = 1/(norm(1)*(n+norm(2))) per direction (src/fft.f90:87,125,150)
  h->normfft = 1.0 / ((1.0 * n0x * (h->nx + 0.0)) * n0y * (h->ny + 0.0));
  eigen_c(h->nx, cbc[0], cbc[1], h->lamx);
  for (double &v : h->lamx) v = v * (dli[0] * dli[0]);
  eigen_c(h->ny, cbc[2], cbc[3], h->lamy);
  for (double &v : h->lamy) v = v * (dli[1] * dli[1]);
  const double *zci = dzci + (nh_d - 1), *zfi = dzfi + (nh_d - 1);
  tridmatrix(cbc[4], cbc[5], h->nz, zci, zfi, c_or_f[2], h->a, h->b, h->c);
  {
    const double dlx[2] = {1.0 / dli[0], 1.0 / dli[0]};
    const double dly[2] = {1.0 / dli[1], 1.0 / dli[1]};
    bc_rhs(cbc + 0, bc + 0, dlx, dlx, c_or_f[0], (long)h->ny * h->nz, h->rhsbx);
    bc_rhs(cbc + 2, bc + 2, dly, dly, c_or_f[1], (long)h->nx * h->nz, h->rhsby);
    // the z spacings of the boundary rows: dzc(0), dzc(nz) cell-centred,
    // dzc(1), dzc(nz-1) face-centred (src/initsolver.f90:109-113)
    const bool fz = (c_or_f[2] == 'f');
    const double dlc[2] = {1.0 / zci[fz ? 1 : 0], 1.0 / zci[fz ? h->nz - 1 : h->nz]};
    const double dlf[2] = {1.0 / zfi[1], 1.0 / zfi[h->nz]};
    bc_rhs(cbc + 4, bc + 4, dlc, dlf, c_or_f[2], (long)h->nx * h->ny, h->rhsbz);
  }
  int rc = FLUTAS_OK;
  const size_t N = (size_t)h->nx * h->ny * h->nz;
  const size_t nhalo = (size_t)(h->nx + 2) * (h->ny + 2) * (h->nz + 2);
  do {
    if ((rc = h->tx.build(h->nx))) break;
    if ((rc = h->ty.build(h->ny))) break;
    if (fused_supported(h->nx) && (rc = build_fused_twiddles(h->tx))) break;
    if (fused_supported(h->ny) && (rc = build_fused_twiddles(h->ty))) break;
    if ((rc = dev_upload(&h->d_lamx, h->lamx))) break;
    if ((rc = dev_upload(&h->d_lamy, h->lamy))) break;
    if ((rc = dev_upload(&h->d_a, h->a))) break;
    if ((rc = dev_upload(&h->d_b, h->b))) break;
    if ((rc = dev_upload(&h->d_c, h->c))) break;
    if ((rc = dev_upload(&h->d_rhsbx, h->rhsbx))) break;
    if ((rc = dev_upload(&h->d_rhsby, h->rhsby))) break;
    if ((rc = dev_upload(&h->d_rhsbz, h->rhsbz))) break;
    hipError_t e = hipSuccess;
    if (precision == FLUTAS_FP64) {
      e = hipMalloc((void **)&h->W64, N * sizeof(double));
    } else {
      e = hipMalloc((void **)&h->W32, N * sizeof(float));
      if (e == hipSuccess) e = hipMalloc((void **)&h->R32, N * sizeof(float));
      if (e == hipSuccess) e = hipMalloc((void **)&h->F, nhalo * sizeof(double));
    }
    // the fused pipeline covers the compiled line lengths
    const char *env = std::getenv("FLUTAS_UNFUSED");
    // (the tile transforms cover the R2HC, DCT-II/III and DST-II/III kinds
    // run forward in P1/P2 and backward in P3/P4)
    const bool tile_kinds = fused_forward_kind(h->kinds[0]) && fused_forward_kind(h->kinds[2]);
    h->fused = fused_supported(h->nx) && fused_supported(h->ny) && h->nz >= 3 && tile_kinds &&
               !(env && env[0] == '1');
    if (h->fused) {
      const int gp = fused_group_planes(h->ny, precision == FLUTAS_FP64 ? 8 : 4);
      // one checkpoint per plane group (two values per column for periodic z)
      const size_t ngrp = (h->nz + gp - 1) / gp;
      const size_t nck = std::max<size_t>(1, ngrp) * (cbc[4] == 'P' ? 2 : 1);
      if (e == hipSuccess) e = hipMalloc((void **)&h->CK, (size_t)h->nx * nck * h->ny * sizeof(double));
    } else {
      const size_t ntk = thomas_checkpoint_doubles(h->nx, h->ny, h->nz, cbc[4] == 'P');
      if (e == hipSuccess) e = hipMalloc((void **)&h->TK, ntk * sizeof(double));
    }
    if (e == hipSuccess) e = hipHostMalloc((void **)&h->nf_host, sizeof(int), hipHostMallocDefault);
    if (e == hipSuccess) {
      *h->nf_host = 0;
      e = hipHostGetDevicePointer((void **)&h->nf_dev, h->nf_host, 0);
    }
    h->nres = residual_blocks(h->nx, h->ny, h->nz);
    h->nx1 = (h->fused && precision == FLUTAS_FP32_REFINE) ? (size_t)x_fwd_blocks(h->tx, 4, h->ny, h->nz) : 0;
    if (e == hipSuccess) e = hipMalloc((void **)&h->part, (h->nres + h->nx1 + 2) * sizeof(double));
    if (e != hipSuccess) { rc = fail(FLUTAS_ENOMEM, std::string("device allocation: ") + hipGetErrorString(e)); break; }
    const char *zenv = std::getenv("FLUTAS_ZCHUNKS");
    if ((rc = set_zchunks(h, zenv ? std::atoi(zenv) : default_zchunks(h)))) break;
  } while (0);
  if (rc) {
    release(h);
    delete h;
    return rc;
  }
  *out = h;
  return FLUTAS_OK;
}

int flutas_psolve_destroy(flutas_psolve *h) {
  if (!h) return FLUTAS_OK;
  (void)hipDeviceSynchronize();
  release(h);
  delete h;
  return FLUTAS_OK;
}

int flutas_psolve_coeffs(const flutas_psolve *h, double *lambdaxy, double *a,
                         double *b, double *c, double *normfft, double *rhsbx,
                         double *rhsby, double *rhsbz) {
  if (!h) return fail(FLUTAS_EINVAL, "null plan");
  if (lambdaxy)
    for (int j = 0; j < h->ny; ++j)
      for (int i = 0; i < h->nx; ++i) lambdaxy[i + (size_t)h->nx * j] = h->lamx[i] + h->lamy[j];
  if (a) std::copy(h->a.begin(), h->a.end(), a);
  if (b) std::copy(h->b.begin(), h->b.end(), b);
  if (c) std::copy(h->c.begin(), h->c.end(), c);
  if (normfft) *normfft = h->normfft;
  if (rhsbx) std::copy(h->rhsbx.begin(), h->rhsbx.end(), rhsbx);
  if (rhsby) std::copy(h->rhsby.begin(), h->rhsby.end(), rhsby);
  if (rhsbz) std::copy(h->rhsbz.begin(), h->rhsbz.end(), rhsbz);
  return FLUTAS_OK;
}

int flutas_psolve_rhsb_device(const flutas_psolve *h, const double **rx,
                              const double **ry, const double **rz) {
  if (!h) return fail(FLUTAS_EINVAL, "null plan");
  if (rx) *rx = h->d_rhsbx;
  if (ry) *ry = h->d_rhsby;
  if (rz) *rz = h->d_rhsbz;
  return FLUTAS_OK;
}

static int solve_impl(flutas_psolve *h, const double *f, double *p) {
  if (!h || !p || !f) return fail(FLUTAS_EINVAL, "null pointer");
  hipStream_t s = current_stream();
  if (h->precision == FLUTAS_FP64) return solve_fp64(h, f, p, s);
  return solve_refined(h, f, p, s);
}

int flutas_psolve_solve_async(flutas_psolve *h, double *p) { return solve_impl(h, p, p); }

// After the stream synchronisation of a synchronous FP64 solve: a
// non-finite right-hand side was seen when the flag holds this solve's
// sequence number.
static int check_nonfinite(const flutas_psolve *h) {
  if (h->precision == FLUTAS_FP64 && *(volatile int *)h->nf_host == h->seq)
    return fail(FLUTAS_ENONFINITE, "right-hand side holds NaN or Inf");
  return FLUTAS_OK;
}

int flutas_psolve_solve_into(flutas_psolve *h, const double *f, double *p) {
  FL_TRY(solve_impl(h, f, p));
  FL_HIP(hipStreamSynchronize(current_stream()));
  return check_nonfinite(h);
}

int flutas_psolve_solve_f32(flutas_psolve *h, float *p) {
  if (!h || !p) return fail(FLUTAS_EINVAL, "null pointer");
  hipStream_t s = current_stream();
  FL_TRY(solve_f32(h, p, p, s));
  FL_HIP(hipStreamSynchronize(s));
  if (h->fused && *(volatile int *)h->nf_host == h->seq)
    return fail(FLUTAS_ENONFINITE, "right-hand side holds NaN or Inf");
  return FLUTAS_OK;
}

int flutas_psolve_solve(flutas_psolve *h, double *p) {
  FL_TRY(flutas_psolve_solve_async(h, p));
  FL_HIP(hipStreamSynchronize(current_stream()));
  return check_nonfinite(h);
}

// ---- graphs of library calls ------------------------------------------------
// No counterpart in the reference.  Between flutas_graph_begin and
// flutas_graph_end, the asynchronous library calls (FP64 solves through
// flutas_psolve_solve_async, the field kernels, boundp, bounduvw without
// outflow faces) are recorded from the library stream instead of run; the
// graph then replays the whole sequence as one launch.  Calls that
// synchronise or allocate (the synchronous solves, the refined mode,
// chkdiv) fail the capture.  The null stream cannot be captured, so a
// library stream left at the default is replaced by a private stream for
// the capture.
struct flutas_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static hipStream_t g_capture_stream = nullptr;  // private capture stream
static hipStream_t g_saved_stream = nullptr;
static bool g_capturing = false;

int flutas_graph_begin(void) {
  if (g_capturing) return fail(FLUTAS_EINVAL, "a graph capture is already open");
  hipStream_t s = g_stream;
  if (!s) {
    if (!g_capture_stream)
      FL_HIP(hipStreamCreateWithFlags(&g_capture_stream, hipStreamNonBlocking));
    // the capture starts after the work already queued on the null stream
    FL_HIP(hipStreamSynchronize(nullptr));
    s = g_capture_stream;
  }
  FL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  g_saved_stream = g_stream;
  g_stream = s;
  g_capturing = true;
  return FLUTAS_OK;
}

int flutas_graph_end(flutas_graph **out) {
  if (!out) return fail(FLUTAS_EINVAL, "null pointer");
  *out = nullptr;
  if (!g_capturing) return fail(FLUTAS_EINVAL, "no graph capture is open");
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(g_stream, &g);
  g_stream = g_saved_stream;
  g_capturing = false;
  if (e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    // the call that broke the capture left its error as the thread's last
    // HIP error; clear it, or the next launch check would report it
    (void)hipGetLastError();
    return fail(FLUTAS_EINVAL, std::string("graph capture failed (a call synchronised or "
                                           "allocated?): ") + hipGetErrorString(e));
  }
  auto *fg = new flutas_graph;
  fg->graph = g;
  const hipError_t ei = hipGraphInstantiate(&fg->exec, g, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(g);
    delete fg;
    return fail(FLUTAS_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  }
  *out = fg;
  return FLUTAS_OK;
}

int flutas_graph_launch(flutas_graph *g) {
  if (!g) return fail(FLUTAS_EINVAL, "null graph");
  if (g_capturing) return fail(FLUTAS_EINVAL, "a graph capture is open");
  FL_HIP(hipGraphLaunch(g->exec, g_stream));
  return FLUTAS_OK;
}

int flutas_graph_destroy(flutas_graph *g) {
  if (!g) return FLUTAS_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return FLUTAS_OK;
}

int flutas_psolve_profile(flutas_psolve *h, const double *f, double *p,
                          int reps, int max_stages, double *ms, int *nstages) {
  if (!h || !f || !p || !ms || !nstages || reps < 1) return fail(FLUTAS_EINVAL, "bad profile arguments");
  // The reps solves run back to back on the library stream with one event
  // after each stage (event 0 before the first); nothing synchronises until
  // the last event, so the stage times are those of the steady pipeline.
  hipStream_t s = current_stream();
  h->profiling = true;
  h->nmark = 0;
  int rc = FLUTAS_OK;
  std::vector<int> rep_end(reps, 0);
  if (h->ev.empty()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) rc = fail(FLUTAS_EDEVICE, "hipEventCreate");
    else h->ev.push_back(e);
  }
  if (rc == FLUTAS_OK && hipEventRecord(h->ev[0], s) != hipSuccess)
    rc = fail(FLUTAS_EDEVICE, "hipEventRecord");
  for (int r = 0; r < reps && rc == FLUTAS_OK; ++r) {
    rc = solve_impl(h, f, p);
    rep_end[r] = h->nmark;
  }
  h->profiling = false;
  if (rc) return rc;
  FL_HIP(hipEventSynchronize(h->ev[h->nmark]));
  const int per = rep_end[0];
  for (int r = 1; r < reps; ++r)
    if (rep_end[r] - rep_end[r - 1] != per)
      return fail(FLUTAS_EINVAL, "profile: stage count differs between solves");
  std::vector<double> acc(per, 0.0);
  for (int m = 0; m < h->nmark; ++m) {
    float t = 0.f;
    FL_HIP(hipEventElapsedTime(&t, h->ev[m], h->ev[m + 1]));
    acc[m % per] += t;
  }
  const int ns = std::min(per, max_stages);
  for (int m = 0; m < ns; ++m) ms[m] = acc[m] / reps;
  *nstages = ns;
  return FLUTAS_OK;
}

// One stage of the four-pass pipeline launched `reps` times back to back
// between two events on the library stream (after one whole solve f -> p
// has filled the work field and the checkpoints): the kernel's own average
// duration, free of the gaps that events between stages add.  P2 and P3
// run in place on the work field, so their repeats transform their own
// output again (no overflow for reps <= 16; the data of the work field are
// not used afterwards).
int flutas_psolve_time_stage(flutas_psolve *h, const double *f, double *p, int stage,
                             int reps, double *ms) {
  if (!h || !f || !p || !ms || reps < 1 || reps > 16 || stage < 0 || stage > 3)
    return fail(FLUTAS_EINVAL, "bad stage timing arguments");
  if (!h->fused || h->precision != FLUTAS_FP64)
    return fail(FLUTAS_EUNSUPPORTED, "stage timing covers the four-pass FP64 pipeline");
  hipStream_t s = current_stream();
  FL_TRY(solve_impl(h, f, p));
  const Geo q = halo_geo(h);
  const int nx = h->nx, ny = h->ny, nz = h->nz;
  const int per = (h->cbc[4] == 'P');
  const int pin = (h->cbc[4] == 'N' && h->cbc[5] == 'N') || per;
  const WStride ws = fused_layout_fp64(nx, ny, nz);
  hipEvent_t e0, e1;
  FL_HIP(hipEventCreate(&e0));
  FL_HIP(hipEventCreate(&e1));
  int rc = FLUTAS_OK;
  if (hipEventRecord(e0, s) != hipSuccess) rc = fail(FLUTAS_EDEVICE, "hipEventRecord");
  for (int r = 0; r < reps && rc == FLUTAS_OK; ++r) {
    if (stage == 0)
      rc = launch_x_fwd<double, double>(h->tx, h->kinds[0], f + q.interior, q.hx, q.hxy, h->W64,
                                        ws, ny, nz, nullptr, nullptr, 0, s);
    else if (stage == 1)
      rc = slab_fwd<double>(h, h->W64, ws, pin, per, s);
    else if (stage == 2)
      rc = slab_bwd<double>(h, h->W64, ws, pin, per, s);
    else
      rc = launch_x_bwd<double, double>(h->tx, h->kinds[1], h->W64, ws, p + q.interior, q.hx,
                                        q.hxy, ny, nz, h->normfft, 0, s);
  }
  float t = 0.f;
  if (rc == FLUTAS_OK && hipEventRecord(e1, s) != hipSuccess) rc = fail(FLUTAS_EDEVICE, "hipEventRecord");
  if (rc == FLUTAS_OK && hipEventSynchronize(e1) != hipSuccess) rc = fail(FLUTAS_EDEVICE, "hipEventSynchronize");
  if (rc == FLUTAS_OK && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = fail(FLUTAS_EDEVICE, "hipEventElapsedTime");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  *ms = (double)t / reps;
  return FLUTAS_OK;
}

const char *flutas_psolve_stage_name(const flutas_psolve *h, int stage) {
  if (!h || stage < 0 || stage >= (int)h->stage_names.size()) return "";
  return h->stage_names[stage].c_str();
}

int flutas_fft_lines(int kind, int n, int howmany, long long stride,
                     long long dist, double *data) {
  return flutas_fft_guru(kind, n, stride, howmany, dist, 1, 0, data);
}

}  // extern "C"

// fft of fftini's line geometries (src/fft.f90:181-193) on FP64 or FP32
// data (the FP32 variant serves FluTAS built with _SINGLE_PRECISION, whose
// FFTW plans are single precision).
template <typename T>
static int fft_guru_impl(int kind, int n, long long stride, int howmany1, long long dist1,
                         int howmany2, long long dist2, T *data) {
  if (!data || howmany1 < 1 || howmany2 < 1 || howmany2 > 65535 || stride < 1 || n < 2)
    return fail(FLUTAS_EINVAL, "bad fft arguments");
  if (kind != K_R2HC && kind != K_HC2R && kind != K_REDFT01 && kind != K_REDFT10 &&
      kind != K_RODFT01 && kind != K_RODFT10 && kind != K_REDFT11 && kind != K_RODFT11 &&
      kind != K_REDFT00 && kind != K_RODFT00)
    return fail(FLUTAS_EINVAL, "unknown r2r kind");
  // DCT-I / DST-I take any n >= 2 (their extension has even length), the
  // other kinds an even n
  if (kind != K_REDFT00 && kind != K_RODFT00 && (n & 1))
    return fail(FLUTAS_EINVAL, "line length must be even for this kind");
  LineTables tab;
  FL_TRY(tab.build(dft_length(kind, n)));
  LineGeom g;
  g.n = n; g.L1 = howmany1; g.L2 = howmany2;
  g.is1 = dist1; g.is2 = dist2; g.ise = stride;
  g.os1 = dist1; g.os2 = dist2; g.ose = stride;
  int rc = launch_lines<T, T, T>(tab, kind, data, data, g, T(1), current_stream());
  hipError_t e = hipStreamSynchronize(current_stream());
  tab.release();
  if (rc) return rc;
  if (e != hipSuccess) return fail(FLUTAS_EDEVICE, hipGetErrorString(e));
  return FLUTAS_OK;
}

extern "C" {

int flutas_fft_guru(int kind, int n, long long stride, int howmany1, long long dist1,
                    int howmany2, long long dist2, double *data) {
  return fft_guru_impl<double>(kind, n, stride, howmany1, dist1, howmany2, dist2, data);
}

int flutas_fft_guru_f32(int kind, int n, long long stride, int howmany1, long long dist1,
                        int howmany2, long long dist2, float *data) {
  return fft_guru_impl<float>(kind, n, stride, howmany1, dist1, howmany2, dist2, data);
}

int flutas_event_create(void **ev) {
  if (!ev) return fail(FLUTAS_EINVAL, "null pointer");
  hipEvent_t e;
  FL_HIP(hipEventCreate(&e));
  *ev = (void *)e;
  return FLUTAS_OK;
}
int flutas_event_record(void *ev) {
  FL_HIP(hipEventRecord((hipEvent_t)ev, current_stream()));
  return FLUTAS_OK;
}
int flutas_event_elapsed_ms(void *start, void *end, float *ms) {
  FL_HIP(hipEventSynchronize((hipEvent_t)end));
  FL_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
  return FLUTAS_OK;
}
int flutas_event_destroy(void *ev) {
  if (ev) FL_HIP(hipEventDestroy((hipEvent_t)ev));
  return FLUTAS_OK;
}

int flutas_psolve_set_zchunks(flutas_psolve *h, int n) {
  if (!h) return fail(FLUTAS_EINVAL, "null plan");
  return set_zchunks(h, n == 0 ? default_zchunks(h) : n);
}

int flutas_psolve_zchunks(const flutas_psolve *h, int *n) {
  if (!h || !n) return fail(FLUTAS_EINVAL, "null pointer");
  *n = h->zchunks;
  return FLUTAS_OK;
}

int flutas_psolve_tune(flutas_psolve *h, const double *f, double *p, int reps, int *chosen,
                       double *ms) {
  if (!h || !f || !p || !chosen || reps < 1 || reps > 100)
    return fail(FLUTAS_EINVAL, "bad tuning arguments");
  hipStream_t s = current_stream();
  int best = 1;
  double tbest = 1e300;
  // the first timed count should not pay for the GPU's clock ramp
  for (int r = 0; r < 3; ++r) FL_TRY(solve_impl(h, f, p));
  for (int v = 0; v < kMaxZChunks; ++v) {
    const int n = v + 1;
    if (ms) ms[v] = -1.0;
    FL_TRY(set_zchunks(h, n));
    if (h->zchunks != n) continue;
    FL_TRY(solve_impl(h, f, p));  // warm-up
    FL_HIP(hipStreamSynchronize(s));
    const auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < reps; ++r) FL_TRY(solve_impl(h, f, p));
    FL_HIP(hipStreamSynchronize(s));
    const double t = std::chrono::duration<double, std::milli>(
                         std::chrono::steady_clock::now() - t0).count() / reps;
    if (ms) ms[v] = t;
    if (t < tbest) {
      tbest = t;
      best = n;
    }
  }
  FL_TRY(set_zchunks(h, best));
  FL_TRY(solve_impl(h, f, p));
  FL_HIP(hipStreamSynchronize(s));
  *chosen = h->zchunks;
  return check_nonfinite(h);
}

int flutas_psolve_stats(const flutas_psolve *h, int *iters, double *relres) {
  if (!h) return fail(FLUTAS_EINVAL, "null plan");
  if (iters) *iters = h->last_iters;
  if (relres) *relres = h->last_relres;
  return FLUTAS_OK;
}

int flutas_psolve_residual(flutas_psolve *h, const double *f, const double *p,
                           double *relres) {
  if (!h || !f || !p || !relres) return fail(FLUTAS_EINVAL, "null pointer");
  if (h->cf[0] != 'c' || h->cf[1] != 'c' || h->cf[2] != 'c')
    return fail(FLUTAS_EUNSUPPORTED, "the residual operator is the cell-centred one");
  hipStream_t s = current_stream();
  int np = 0;
  FL_TRY(launch_absmax(f, h->nx, h->ny, h->nz, h->part, &np, s));
  double fmax = 0.0, rmax = 0.0;
  FL_TRY(reduce_max(h, np, s, &fmax));
  FL_TRY(launch_residual<double>(op_desc(h), f, p, nullptr, h->part, &np, s));
  FL_TRY(reduce_max(h, np, s, &rmax));
  *relres = (fmax > 0.0) ? rmax / fmax : rmax;
  return FLUTAS_OK;
}

// initgrid (src/initgrid.f90:17-97); arrays (1-nh_d : n+nh_d), g(k) ==
// arr[k + nh_d - 1].  one_end selects gridpoint_cluster_one_end
// (:120-136) instead of gridpoint_cluster_two_end (:102-118), as inivel
// 'hcl' / 'hcp' do (:32-39).
static int initgrid_impl(bool one_end, int n, double gr, double lz, int nh_d,
                         double *dzc, double *dzf, double *zc, double *zf) {
  if (n < 1 || nh_d < 1 || !dzc || !dzf || !zc || !zf) return fail(FLUTAS_EINVAL, "bad initgrid arguments");
  auto G = [nh_d](double *arr, int k) -> double & { return arr[k + nh_d - 1]; };
  for (int k = 1; k <= n; ++k) {
    const double z0 = (k - 0.0) / (1.0 * n);
    double z = z0;
    if (gr != 0.0)
      z = one_end ? 1.0 * (1.0 + std::tanh((z0 - 1.0) * gr) / std::tanh(gr / 1.0))
                  : 0.5 * (1.0 + std::tanh((z0 - 0.5) * gr) / std::tanh(gr / 2.0));
    G(zf, k) = z * lz;
  }
  G(zf, 0) = 0.0;
  for (int k = 1; k <= n; ++k) G(dzf, k) = G(zf, k) - G(zf, k - 1);
  G(dzf, 0) = G(dzf, 1);
  G(dzf, n + 1) = G(dzf, n);
  for (int k = 0; k <= n; ++k) G(dzc, k) = 0.5 * (G(dzf, k) + G(dzf, k + 1));
  G(dzc, n + 1) = G(dzc, n);
  G(zc, 0) = -G(dzc, 0) / 2.0;
  G(zf, 0) = 0.0;
  for (int k = 1; k <= n + 1; ++k) {
    G(zc, k) = G(zc, k - 1) + G(dzc, k - 1);
    G(zf, k) = G(zf, k - 1) + G(dzf, k);
  }
  for (int k = 1 - nh_d; k <= 0; ++k) {
    G(dzf, k) = G(dzf, -k + 1);
    G(dzc, k) = G(dzc, -k);
  }
  for (int k = n + 1; k <= n + nh_d; ++k) {
    G(dzf, k) = G(dzf, 2 * n - k - 1);
    G(dzc, k) = G(dzc, 2 * n - k);
  }
  for (int k = 0; k >= 1 - nh_d; --k) {
    G(zf, k) = G(zf, k + 1) - G(dzf, k);
    G(zc, k) = G(zf, k + 1) - G(dzc, k + 1);
  }
  for (int k = n + 1; k <= n + nh_d; ++k) {
    G(zf, k) = G(zf, k - 1) + G(dzf, k);
    G(zc, k) = G(zf, k - 1) + G(dzc, k - 1);
  }
  return FLUTAS_OK;
}

int flutas_initgrid(int n, double gr, double lz, int nh_d, double *dzc,
                    double *dzf, double *zc, double *zf) {
  return initgrid_impl(false, n, gr, lz, nh_d, dzc, dzf, zc, zf);
}

int flutas_initgrid_inivel(const char *inivel, int n, double gr, double lz,
                           int nh_d, double *dzc, double *dzf, double *zc,
                           double *zf) {
  if (!inivel) return fail(FLUTAS_EINVAL, "null inivel");
  const std::string v(inivel, strnlen(inivel, 3));
  return initgrid_impl(v == "hcl" || v == "hcp", n, gr, lz, nh_d, dzc, dzf, zc, zf);
}

}  // extern "C"

// ---- batched solves ---------------------------------------------------------
// No counterpart in the reference (one solve per call there).  Each problem
// keeps its own plan (coefficients, twiddles, work field; FP64, or FP32 for
// the refined mode); the batch adds a device descriptor table, its own
// checkpoint buffers (sized for the quarter plane group), one item list per
// pass (batch.hip) and the per-problem maxima.  Graph 0 is the whole solve
// (FP64: four pass launches; refined: solve, residual, correction solve,
// check and the publishing launch), graph 1 one more correction round of
// the refined mode; each is captured once and replayed.
struct flutas_batch {
  int nprob = 0;
  int precision = FLUTAS_FP64;
  double tol = 0.0;
  int max_rounds = 8;
  int elem = 8;  // bytes per value of the work fields
  std::vector<flutas_psolve *> plans;
  std::vector<BatchProblem> probs;
  std::vector<double *> ck;
  BatchProblem *d_probs = nullptr;
  BatchWork *d_items[kBatchStages] = {};
  int nitems[kBatchStages] = {};
  int grid[kBatchStages] = {};
  // launch shape: 0 one launch per pass and line length (segments of the
  // pass's item list, run side by side on the side streams), 1 one
  // persistent launch per pass walking the whole list
  int mode = 0;
  struct Segment {
    int off, cnt, n;
    size_t lds;
  };
  std::vector<Segment> segs[kBatchStages];
  static constexpr int kSide = 4, kEvents = 8;
  hipStream_t side[kSide] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t ev[kBatchStages][kEvents] = {};
  size_t lds = 0;
  int ncu = 0, per_cu = 0, max_per_cu = 1;  // workgroups per CU (0: occupancy)
  unsigned long long *d_mx = nullptr;       // maxima bits, 3 per problem
  double *d_out = nullptr;                  // published maxima
  double *h_out = nullptr;                  // pinned copy
  hipStream_t stream = nullptr;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  bool fields_set = false;
  std::vector<int> iters;
  std::vector<double> relres;
};

static void batch_drop_graphs(flutas_batch *b) {
  for (int g = 0; g < 2; ++g) {
    if (b->exec[g]) (void)hipGraphExecDestroy(b->exec[g]);
    if (b->graph[g]) (void)hipGraphDestroy(b->graph[g]);
    b->exec[g] = nullptr;
    b->graph[g] = nullptr;
  }
}

static void batch_release(flutas_batch *b) {
  batch_drop_graphs(b);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  for (hipStream_t t : b->side)
    if (t) (void)hipStreamDestroy(t);
  for (auto &row : b->ev)
    for (hipEvent_t e : row)
      if (e) (void)hipEventDestroy(e);
  for (BatchWork *d : b->d_items)
    if (d) (void)hipFree(d);
  for (void *d : {(void *)b->d_probs, (void *)b->d_mx, (void *)b->d_out})
    if (d) (void)hipFree(d);
  if (b->h_out) (void)hipHostFree(b->h_out);
  for (double *c : b->ck)
    if (c) (void)hipFree(c);
  for (flutas_psolve *h : b->plans) flutas_psolve_destroy(h);
  b->plans.clear();
}

static int batch_upload(flutas_batch *b) {
  FL_HIP(hipMemcpy(b->d_probs, b->probs.data(), b->probs.size() * sizeof(BatchProblem),
                   hipMemcpyHostToDevice));
  return FLUTAS_OK;
}

// stages of graph g
static void batch_stage_range(const flutas_batch *b, int g, int *s0, int *s1) {
  if (b->precision == FLUTAS_FP64) { *s0 = BS_X_FWD; *s1 = BS_X_BWD; }
  else if (g == 0) { *s0 = BS_X_FWD; *s1 = BS_CHECK; }
  else { *s0 = BS_RESID; *s1 = BS_CHECK; }
}

static bool x_stage(int s) {
  return s == BS_X_FWD || s == BS_X_BWD || s == BS_X_FWD_R || s == BS_X_BWD_ACC;
}
static bool resid_stage(int s) { return s == BS_RESID || s == BS_CHECK; }

// Item lists, one per stage: the problems with the most points first; x
// stages (tile, plane chunk) items of the problem's current tile size,
// slab stages one item per x-slab, residual stages (row group, plane
// chunk) items.  Sets the dynamic LDS and the grids, and drops the
// captured graphs.
static int batch_items(flutas_batch *b) {
  const int KP = batch_resid_planes();
  std::vector<int> order(b->nprob);
  for (int q = 0; q < b->nprob; ++q) order[q] = q;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    const BatchProblem &px = b->probs[x], &py = b->probs[y];
    return (long long)px.nx * px.ny * px.nz > (long long)py.nx * py.ny * py.nz;
  });
  size_t lds = 0;
  for (BatchProblem &p : b->probs) {
    p.rpi = batch_resid_rows(p.nx);
    lds = std::max(lds, batch_item_lds(0, p.nx, b->elem, p.tj_v));
    lds = std::max(lds, batch_item_lds(1, p.ny, b->elem, p.gp_v));
  }
  batch_drop_graphs(b);
  b->lds = lds;
  FL_TRY(batch_persistent_setup(b->elem, lds, &b->max_per_cu));
  const int per_cu = (b->per_cu > 0) ? std::min(b->per_cu, b->max_per_cu) : b->max_per_cu;
  int s0, s1;
  batch_stage_range(b, 0, &s0, &s1);
  for (int s = 0; s < kBatchStages; ++s) {
    if (b->d_items[s]) (void)hipFree(b->d_items[s]);
    b->d_items[s] = nullptr;
    b->nitems[s] = 0;
    b->grid[s] = 0;
    if (s < s0 || s > s1) continue;
    // per-length order: the problems by line length of the pass (nx for
    // the x stages, ny for the slab stages), longest first, then by size
    std::vector<int> ord = order;
    if (!resid_stage(s))
      std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
        const int nxx = x_stage(s) ? b->probs[x].nx : b->probs[x].ny;
        const int nyy = x_stage(s) ? b->probs[y].nx : b->probs[y].ny;
        return nxx > nyy;
      });
    b->segs[s].clear();
    std::vector<BatchWork> items;
    for (int q : ord) {
      const BatchProblem &p = b->probs[q];
      const int n = resid_stage(s) ? 0 : (x_stage(s) ? p.nx : p.ny);
      if (b->segs[s].empty() || b->segs[s].back().n != n) {
        const size_t l = resid_stage(s) ? 0
                         : batch_item_lds(x_stage(s) ? 0 : 1, n, b->elem,
                                          x_stage(s) ? p.tj_v : p.gp_v);
        b->segs[s].push_back({(int)items.size(), 0, n, l});
      }
      if (x_stage(s)) {
        const int tj = batch_tile_lines(p.nx, b->elem, p.tj_v);
        const int ntile = (p.ny + tj - 1) / tj, nkc = (p.nz + p.kxp - 1) / p.kxp;
        for (int kc = 0; kc < nkc; ++kc)
          for (int t = 0; t < ntile; ++t) items.push_back({s, q, t, kc});
      } else if (resid_stage(s)) {
        const int ng = (p.ny + p.rpi - 1) / p.rpi, nkc = (p.nz + KP - 1) / KP;
        for (int kc = 0; kc < nkc; ++kc)
          for (int jg = 0; jg < ng; ++jg) items.push_back({s, q, jg, kc});
      } else {
        for (int i = 0; i < p.nx; ++i) items.push_back({s, q, i, 0});
      }
      b->segs[s].back().cnt = (int)items.size() - b->segs[s].back().off;
    }
    for (const auto &sg : b->segs[s]) FL_TRY(batch_segment_setup(b->elem, s, sg.n, sg.lds));
    b->nitems[s] = (int)items.size();
    b->grid[s] = std::max(1, std::min(b->nitems[s], per_cu * b->ncu));
    FL_HIP(hipMalloc((void **)&b->d_items[s], items.size() * sizeof(BatchWork)));
    FL_HIP(hipMemcpy(b->d_items[s], items.data(), items.size() * sizeof(BatchWork),
                     hipMemcpyHostToDevice));
  }
  return batch_upload(b);
}

// The launches of graph g on stream st: one per stage, then (refined) the
// publishing launch.
static int batch_enqueue(flutas_batch *b, int g, hipStream_t st) {
  int s0, s1;
  batch_stage_range(b, g, &s0, &s1);
  for (int s = s0; s <= s1; ++s) {
    BatchPassArgs A;
    A.P = b->d_probs;
    A.items = b->d_items[s];
    A.nitems = b->nitems[s];
    A.mx = b->d_mx;
    if (b->mode == 1) {
      FL_TRY(launch_batch_pass(b->elem, A, b->grid[s], b->lds, st));
      continue;
    }
    const auto &sg = b->segs[s];
    auto launch = [&](size_t q, hipStream_t t) {
      BatchPassArgs a = A;
      a.items = A.items + sg[q].off;
      a.nitems = sg[q].cnt;
      return launch_batch_segment(b->elem, a, s, sg[q].n, sg[q].lds, t);
    };
    const size_t nseg = sg.size();
    if (nseg <= 1 || nseg > (size_t)flutas_batch::kEvents - 1) {
      for (size_t q = 0; q < nseg; ++q) FL_TRY(launch(q, st));
      continue;
    }
    // fork: segment q > 0 on side stream (q-1) % kSide after the work so
    // far; join: the next pass starts when every segment has ended
    hipEvent_t *e = b->ev[s];
    FL_HIP(hipEventRecord(e[0], st));
    for (size_t q = 1; q < nseg; ++q) {
      hipStream_t t = b->side[(q - 1) % flutas_batch::kSide];
      FL_HIP(hipStreamWaitEvent(t, e[0], 0));
      FL_TRY(launch(q, t));
      FL_HIP(hipEventRecord(e[q], t));
    }
    FL_TRY(launch(0, st));
    for (size_t q = 1; q < nseg; ++q) FL_HIP(hipStreamWaitEvent(st, e[q], 0));
  }
  if (b->precision != FLUTAS_FP64) FL_TRY(launch_batch_publish(b->d_mx, b->d_out, 3 * b->nprob, st));
  return FLUTAS_OK;
}

// Replays graph g (captured on first use), copies the published maxima to
// the host (refined mode) and synchronises.
static int batch_run(flutas_batch *b, int g) {
  if (!b->exec[g]) {
    FL_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = batch_enqueue(b, g, b->stream);
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &gr);
    if (rc) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
    }
    if (e != hipSuccess) return fail(FLUTAS_EDEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    b->graph[g] = gr;
    FL_HIP(hipGraphInstantiate(&b->exec[g], gr, nullptr, nullptr, 0));
  }
  FL_HIP(hipGraphLaunch(b->exec[g], b->stream));
  if (b->precision != FLUTAS_FP64)
    FL_HIP(hipMemcpyAsync(b->h_out, b->d_out, 3 * b->nprob * sizeof(double),
                          hipMemcpyDeviceToHost, b->stream));
  FL_HIP(hipStreamSynchronize(b->stream));
  return FLUTAS_OK;
}

// The whole batched solve: graph 0, then (refined) correction rounds for
// the problems still above their tolerance.
static int batch_solve_all(flutas_batch *b) {
  FL_TRY(batch_run(b, 0));
  const int n = b->nprob;
  b->iters.assign(n, 0);
  b->relres.assign(n, -1.0);
  if (b->precision == FLUTAS_FP64) return FLUTAS_OK;
  std::vector<double> fmax(n);
  bool more = false;
  for (int q = 0; q < n; ++q) {
    fmax[q] = b->h_out[3 * q];
    const double r = b->h_out[3 * q + 2];
    b->iters[q] = 1;
    b->relres[q] = fmax[q] > 0.0 ? r / fmax[q] : r;
    if (!(fmax[q] == fmax[q]) || std::isinf(fmax[q]))
      return fail(FLUTAS_ENONFINITE, "problem " + std::to_string(q) + ": right-hand side holds NaN or Inf");
    if (!(b->relres[q] == b->relres[q]))
      return fail(FLUTAS_ENONFINITE, "problem " + std::to_string(q) + ": residual is NaN");
    more |= b->relres[q] > b->tol;
  }
  int rc = FLUTAS_OK;
  for (int round = 1; more && round < b->max_rounds && rc == FLUTAS_OK; ++round) {
    for (int q = 0; q < n; ++q) b->probs[q].active = b->relres[q] > b->tol ? 1 : 0;
    if ((rc = batch_upload(b))) break;
    if ((rc = batch_run(b, 1))) break;
    more = false;
    for (int q = 0; q < n; ++q) {
      if (!b->probs[q].active) continue;
      const double r = b->h_out[3 * q + 2];
      b->relres[q] = fmax[q] > 0.0 ? r / fmax[q] : r;
      ++b->iters[q];
      if (!(b->relres[q] == b->relres[q])) {
        rc = fail(FLUTAS_ENONFINITE, "problem " + std::to_string(q) + ": residual is NaN");
        break;
      }
      more |= b->relres[q] > b->tol;
    }
  }
  // every problem active again for the next solve
  bool changed = false;
  for (BatchProblem &p : b->probs) {
    changed |= p.active == 0;
    p.active = 1;
  }
  if (changed) FL_TRY(batch_upload(b));
  if (rc) return rc;
  if (more) return fail(FLUTAS_ENOTCONV, "batched refinement did not reach the tolerance");
  return FLUTAS_OK;
}

extern "C" {

int flutas_batch_create_ex(int nprob, const int *ng, const char *cbc, const double *dli,
                           int nh_d, const double *const *dzci, const double *const *dzfi,
                           int precision, double tol, flutas_batch **out) {
  if (!out || nprob < 1 || !ng || !cbc || !dli || !dzci || !dzfi)
    return fail(FLUTAS_EINVAL, "bad batch arguments");
  if (precision != FLUTAS_FP64 && precision != FLUTAS_FP32_REFINE)
    return fail(FLUTAS_EINVAL, "precision must be 64 or 32");
  if (precision == FLUTAS_FP32_REFINE && !(tol > 0.0)) return fail(FLUTAS_EINVAL, "tol must be > 0");
  *out = nullptr;
  FL_TRY(check_device());
  flutas_batch *b = new flutas_batch;
  b->nprob = nprob;
  b->precision = precision;
  b->tol = tol;
  b->elem = precision == FLUTAS_FP64 ? 8 : 4;
  int rc = FLUTAS_OK;
  const double zero_bc[6] = {0, 0, 0, 0, 0, 0};
  for (int q = 0; q < nprob && rc == FLUTAS_OK; ++q) {
    const int *n = ng + 3 * q;
    if (!batch_supported(n[0]) || !batch_supported(n[1]) || n[2] < 3) {
      rc = fail(FLUTAS_EUNSUPPORTED, "batched solves need nx, ny in {32,48,64,96,128,192,256}"
                                     " and nz >= 3");
      break;
    }
    flutas_psolve *h = nullptr;
    rc = flutas_psolve_create(n, cbc + 6 * q, zero_bc, "ccc", dli + 3 * q, nh_d, dzci[q],
                              dzfi[q], precision, precision == FLUTAS_FP64 ? 0.0 : tol, &h);
    if (rc) break;
    b->plans.push_back(h);
    if (!h->fused) {
      rc = fail(FLUTAS_EUNSUPPORTED, "problem not on the fused pipeline");
      break;
    }
    BatchProblem p{};
    p.nx = h->nx; p.ny = h->ny; p.nz = h->nz;
    p.kx_f = h->kinds[0]; p.kx_b = h->kinds[1]; p.ky_f = h->kinds[2]; p.ky_b = h->kinds[3];
    p.per = (h->cbc[4] == 'P');
    p.pin = (h->cbc[4] == 'N' && h->cbc[5] == 'N') || p.per;
    p.tj_v = 0; p.gp_v = 0;
    p.kxp = 4;
    p.active = 1;
    std::memcpy(p.cbc, h->cbc, 6);
    p.scale = h->normfft;
    p.dx2 = h->dli[0] * h->dli[0];
    p.dy2 = h->dli[1] * h->dli[1];
    p.hx = h->nx + 2;
    p.hxy = (long long)(h->nx + 2) * (h->ny + 2);
    p.W = precision == FLUTAS_FP64 ? (void *)h->W64 : (void *)h->W32;
    p.R = h->R32;
    p.lamx = h->d_lamx; p.lamy = h->d_lamy; p.a = h->d_a; p.b = h->d_b; p.c = h->d_c;
    // split/phase tables, or the DCT-IV/DST-IV pre/post tables of the ND
    // and DN kinds (REDFT11/RODFT11)
    const bool d4x = (p.kx_f == K_REDFT11 || p.kx_f == K_RODFT11);
    const bool d4y = (p.ky_f == K_REDFT11 || p.ky_f == K_RODFT11);
    const size_t ox = 2 * (size_t)(h->tx.N + 1), oy = 2 * (size_t)(h->ty.N + 1);
    p.twx = h->tx.d_ptw;
    p.twhx = d4x ? h->tx.d_t4 : h->tx.d_twh;
    p.twqx = d4x ? h->tx.d_t4 + ox : h->tx.d_twq;
    p.twy = h->ty.d_ptw;
    p.twhy = d4y ? h->ty.d_t4 : h->ty.d_twh;
    p.twqy = d4y ? h->ty.d_t4 + oy : h->ty.d_twq;
    p.ftwx = h->tx.f_ptw;
    p.ftwhx = d4x ? h->tx.f_t4 : h->tx.f_twh;
    p.ftwqx = d4x ? h->tx.f_t4 + ox : h->tx.f_twq;
    p.ftwy = h->ty.f_ptw;
    p.ftwhy = d4y ? h->ty.f_t4 : h->ty.f_twh;
    p.ftwqy = d4y ? h->ty.f_t4 + oy : h->ty.f_twq;
    // checkpoints for the quarter plane group (the most groups)
    const int gp = batch_group_planes(h->ny, b->elem, 2);
    const size_t nck = (size_t)((h->nz + gp - 1) / gp) * (p.per ? 2 : 1);
    double *ck = nullptr;
    if (hipMalloc((void **)&ck, (size_t)h->nx * nck * h->ny * sizeof(double)) != hipSuccess) {
      rc = fail(FLUTAS_ENOMEM, "batch checkpoint allocation");
      break;
    }
    b->ck.push_back(ck);
    p.ck = ck;
    b->probs.push_back(p);
  }
  auto dalloc = [&](void **ptr, size_t bytes) {
    if (rc == FLUTAS_OK && hipMalloc(ptr, bytes) != hipSuccess)
      rc = fail(FLUTAS_ENOMEM, "batch table allocation");
    if (rc == FLUTAS_OK && hipMemset(*ptr, 0, bytes) != hipSuccess)
      rc = fail(FLUTAS_EDEVICE, "hipMemset");
  };
  dalloc((void **)&b->d_probs, nprob * sizeof(BatchProblem));
  dalloc((void **)&b->d_mx, 3 * (size_t)nprob * sizeof(unsigned long long));
  dalloc((void **)&b->d_out, 3 * (size_t)nprob * sizeof(double));
  if (rc == FLUTAS_OK &&
      hipHostMalloc((void **)&b->h_out, 3 * (size_t)nprob * sizeof(double),
                    hipHostMallocDefault) != hipSuccess)
    rc = fail(FLUTAS_ENOMEM, "pinned host allocation");
  if (rc == FLUTAS_OK) {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
      rc = fail(FLUTAS_EDEVICE, "hipGetDeviceProperties");
    else
      b->ncu = prop.multiProcessorCount;
  }
  if (rc == FLUTAS_OK && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    rc = fail(FLUTAS_EDEVICE, "hipStreamCreateWithFlags");
  for (int t = 0; t < flutas_batch::kSide && rc == FLUTAS_OK; ++t)
    if (hipStreamCreateWithFlags(&b->side[t], hipStreamNonBlocking) != hipSuccess)
      rc = fail(FLUTAS_EDEVICE, "hipStreamCreateWithFlags");
  for (int st = 0; st < kBatchStages && rc == FLUTAS_OK; ++st)
    for (int q = 0; q < flutas_batch::kEvents && rc == FLUTAS_OK; ++q)
      if (hipEventCreateWithFlags(&b->ev[st][q], hipEventDisableTiming) != hipSuccess)
        rc = fail(FLUTAS_EDEVICE, "hipEventCreateWithFlags");
  if (rc == FLUTAS_OK) rc = batch_items(b);
  if (rc) {
    batch_release(b);
    delete b;
    return rc;
  }
  *out = b;
  return FLUTAS_OK;
}

int flutas_batch_create(int nprob, const int *ng, const char *cbc, const double *dli,
                        int nh_d, const double *const *dzci, const double *const *dzfi,
                        flutas_batch **out) {
  return flutas_batch_create_ex(nprob, ng, cbc, dli, nh_d, dzci, dzfi, FLUTAS_FP64, 0.0, out);
}

int flutas_batch_set_fields(flutas_batch *b, const double *const *f, double *const *p) {
  if (!b || !f || !p) return fail(FLUTAS_EINVAL, "null pointer");
  for (int q = 0; q < b->nprob; ++q) {
    if (!f[q] || !p[q]) return fail(FLUTAS_EINVAL, "null field pointer");
    if (f[q] == p[q]) return fail(FLUTAS_EINVAL, "batched solves need f and p apart");
    const long long off = 1 + b->probs[q].hx + b->probs[q].hxy;
    b->probs[q].f = f[q] + off;
    b->probs[q].p = p[q] + off;
  }
  FL_TRY(batch_upload(b));
  b->fields_set = true;
  return FLUTAS_OK;
}

int flutas_batch_solve(flutas_batch *b) {
  if (!b) return fail(FLUTAS_EINVAL, "null batch");
  if (!b->fields_set) return fail(FLUTAS_EINVAL, "flutas_batch_set_fields not called");
  // the caller's pending work on the library stream comes first
  FL_HIP(hipStreamSynchronize(current_stream()));
  return batch_solve_all(b);
}

int flutas_batch_stats(const flutas_batch *b, int *iters, double *relres) {
  if (!b) return fail(FLUTAS_EINVAL, "null batch");
  for (int q = 0; q < b->nprob; ++q) {
    if (iters) iters[q] = q < (int)b->iters.size() ? b->iters[q] : 0;
    if (relres) relres[q] = q < (int)b->relres.size() ? b->relres[q] : -1.0;
  }
  return FLUTAS_OK;
}

// Autotuner: coordinate descent over the launch shape (per-length launches
// or persistent pass walks, the latter at its best workgroups per CU),
// then per x length the tile variant (full, half, quarter lines), per y
// length the plane-group variant (the z elimination's block of planes),
// the planes of an x-tile item (1, 2, 4, 8), then the launch shape again; each choice timed over `reps` replays of
// graph 0 (the whole solve in FP64; solve, residual, correction and check
// in the refined mode).
int flutas_batch_tune(flutas_batch *b, int reps) {
  if (!b || reps < 1) return fail(FLUTAS_EINVAL, "bad tune arguments");
  if (!b->fields_set) return fail(FLUTAS_EINVAL, "flutas_batch_set_fields not called");
  FL_HIP(hipStreamSynchronize(current_stream()));
  hipEvent_t e0, e1;
  FL_HIP(hipEventCreate(&e0));
  FL_HIP(hipEventCreate(&e1));
  auto timed = [&](float *ms) -> int {
    FL_TRY(batch_items(b));
    FL_TRY(batch_run(b, 0));  // capture and warm
    FL_HIP(hipEventRecord(e0, b->stream));
    for (int r = 0; r < reps; ++r) FL_HIP(hipGraphLaunch(b->exec[0], b->stream));
    FL_HIP(hipEventRecord(e1, b->stream));
    FL_HIP(hipEventSynchronize(e1));
    FL_HIP(hipEventElapsedTime(ms, e0, e1));
    return FLUTAS_OK;
  };
  int rc = FLUTAS_OK;
  auto tune_per_cu = [&]() -> int {
    FL_TRY(batch_items(b));
    const int top = b->max_per_cu;
    float best_t = 0.f;
    int best = top;
    for (int pc = 1; pc <= top; ++pc) {
      b->per_cu = pc;
      float t = 0.f;
      FL_TRY(timed(&t));
      if (pc == 1 || t < best_t) { best_t = t; best = pc; }
    }
    b->per_cu = best;
    return FLUTAS_OK;
  };
  // launch shape: per-length launches or persistent pass walks (each at
  // its own best occupancy)
  auto tune_mode = [&]() -> int {
    float t[2] = {0.f, 0.f};
    int pc1 = b->per_cu;
    for (int m = 0; m < 2; ++m) {
      b->mode = m;
      if (m == 1) {
        FL_TRY(tune_per_cu());
        pc1 = b->per_cu;
      }
      FL_TRY(timed(&t[m]));
    }
    b->mode = t[1] < t[0] ? 1 : 0;
    b->per_cu = pc1;
    return FLUTAS_OK;
  };
  rc = tune_mode();
  for (int axis = 0; axis < 2 && rc == FLUTAS_OK; ++axis) {
    std::vector<int> classes;
    for (const BatchProblem &p : b->probs) classes.push_back(axis == 0 ? p.nx : p.ny);
    std::sort(classes.begin(), classes.end());
    classes.erase(std::unique(classes.begin(), classes.end()), classes.end());
    for (int n : classes) {
      float t[3] = {0.f, 0.f, 0.f};
      for (int v = 0; v < 3 && rc == FLUTAS_OK; ++v) {
        for (BatchProblem &p : b->probs)
          if ((axis == 0 ? p.nx : p.ny) == n) (axis == 0 ? p.tj_v : p.gp_v) = v;
        rc = timed(&t[v]);
      }
      if (rc) break;
      int best = 0;
      for (int v = 1; v < 3; ++v)
        if (t[v] < t[best]) best = v;
      for (BatchProblem &p : b->probs)
        if ((axis == 0 ? p.nx : p.ny) == n) (axis == 0 ? p.tj_v : p.gp_v) = best;
    }
  }
  // planes per x-tile item (fewer: more workgroups; more: the twiddles
  // staged once for more planes)
  if (rc == FLUTAS_OK) {
    const int opts[4] = {1, 2, 4, 8};
    float t[4] = {0.f, 0.f, 0.f, 0.f};
    for (int o = 0; o < 4 && rc == FLUTAS_OK; ++o) {
      for (BatchProblem &p : b->probs) p.kxp = opts[o];
      rc = timed(&t[o]);
    }
    int best = 0;
    for (int o = 1; o < 4; ++o)
      if (t[o] < t[best]) best = o;
    for (BatchProblem &p : b->probs) p.kxp = opts[best];
  }
  if (rc == FLUTAS_OK) rc = tune_mode();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  return batch_items(b);
}

int flutas_batch_tuning(const flutas_batch *b, int *tile_variant, int *group_variant) {
  if (!b) return fail(FLUTAS_EINVAL, "null batch");
  for (int q = 0; q < b->nprob; ++q) {
    if (tile_variant) tile_variant[q] = b->probs[q].tj_v;
    if (group_variant) group_variant[q] = b->probs[q].gp_v;
  }
  return FLUTAS_OK;
}

int flutas_batch_launch_info(const flutas_batch *b, int *workgroups, int *items,
                             int *per_cu, size_t *lds_bytes) {
  if (!b) return fail(FLUTAS_EINVAL, "null batch");
  int wg = 0, it = 0;
  for (int s = 0; s < kBatchStages; ++s) {
    wg = std::max(wg, b->grid[s]);
    it += b->nitems[s];
  }
  if (workgroups) *workgroups = wg;
  if (items) *items = it;
  if (per_cu) *per_cu = b->per_cu > 0 ? std::min(b->per_cu, b->max_per_cu) : b->max_per_cu;
  if (lds_bytes) *lds_bytes = b->lds;
  return FLUTAS_OK;
}

int flutas_batch_launch_shape(const flutas_batch *b, int *shape, int *x_item_planes) {
  if (!b) return fail(FLUTAS_EINVAL, "null batch");
  if (shape) *shape = b->mode;
  if (x_item_planes) *x_item_planes = b->probs.empty() ? 0 : b->probs[0].kxp;
  return FLUTAS_OK;
}

int flutas_batch_destroy(flutas_batch *b) {
  if (!b) return FLUTAS_OK;
  (void)hipDeviceSynchronize();
  batch_release(b);
  delete b;
  return FLUTAS_OK;
}

}  // extern "C"
