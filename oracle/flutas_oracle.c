/*
 * FluTAS pressure-projection oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C FP64 restatement of the reference CPU path of FluTAS
 * (Multiphysics-Flow-Solvers/FluTAS).  It is the checker for the HIP product
 * library and is loaded only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Each function cites the reference lines it
 * follows.
 *
 * The reference transforms are FFTW 3.3.10 guru r2r plans (src/fft.f90:84-86,
 * :122-124); FFTW is not vendored, so the transforms here are written from
 * FFTW's published unnormalised r2r definitions.  They are computed two ways:
 * a long-double direct sum (the definition) and a mixed-radix FFT of the
 * symmetric extension.  Both are pinned by tests against scipy.fft/numpy
 * golden vectors.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include "flutas_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

#ifdef _OPENMP
#include <omp.h>
int or_max_threads(void) { return omp_get_max_threads(); }
#else
int or_max_threads(void) { return 1; }
#endif

/* ------------------------------------------------------------------------ */
/* Complex DFT of any length: recursive mixed-radix decimation in time.      */
/* ------------------------------------------------------------------------ */

static const long double PI_L = 3.14159265358979323846264338327950288L;

/* exp(i*pi*num/den), rounded from long double. */
static cplx expi_frac(long num, long den) {
  long double ang = PI_L * (long double)num / (long double)den;
  return (double)cosl(ang) + I * (double)sinl(ang);
}

/* Twiddle table w[m] = exp(sign*2*pi*i*m/N), m < N. */
static cplx *make_twiddles(int N, int sign) {
  cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)N);
  for (int m = 0; m < N; ++m) {
    cplx e = expi_frac(2L * m, N);
    w[m] = (sign < 0) ? conj(e) : e;
  }
  return w;
}

static int smallest_factor(int n) {
  if (n % 4 == 0) return 4;
  if (n % 2 == 0) return 2;
  for (int p = 3; (long)p * p <= n; p += 2)
    if (n % p == 0) return p;
  return n;
}

/* out[0..N) = DFT of in[0], in[s], ..., in[(N-1)s].  tw is the table of the
 * top-level length Ntop, ts = Ntop/N.  scratch holds >= 2N values. */
static void dft_rec(int N, const cplx *in, long s, cplx *out, const cplx *tw,
                    int Ntop, long ts, cplx *scratch) {
  if (N == 1) {
    out[0] = in[0];
    return;
  }
  int p = smallest_factor(N);
  int M = N / p;
  for (int r = 0; r < p; ++r)
    dft_rec(M, in + r * s, s * p, out + (long)r * M, tw, Ntop, ts * p, scratch);
  if (p == 2) {
    for (int k = 0; k < M; ++k) {
      cplx a = out[k], b = out[M + k] * tw[(long)k * ts];
      out[k] = a + b;
      out[M + k] = a - b;
    }
  } else if (p == 4) {
    cplx om = tw[(long)M * ts]; /* exp(sign*2*pi*i/4) = sign*i */
    for (int k = 0; k < M; ++k) {
      cplx a0 = out[k], a1 = out[M + k] * tw[(long)k * ts],
           a2 = out[2 * M + k] * tw[2L * k * ts], a3 = out[3 * M + k] * tw[3L * k * ts];
      cplx t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = (a1 - a3) * om;
      out[k] = t0 + t2;
      out[M + k] = t1 + t3;
      out[2 * M + k] = t0 - t2;
      out[3 * M + k] = t1 - t3;
    }
  } else {
    cplx *tt = scratch, *res = scratch + p;
    for (int k = 0; k < M; ++k) {
      for (int r = 0; r < p; ++r) tt[r] = out[(long)r * M + k] * tw[(long)r * k * ts];
      for (int q = 0; q < p; ++q) {
        cplx acc = 0;
        for (int r = 0; r < p; ++r) acc += tt[r] * tw[((long)r * q * M * ts) % Ntop];
        res[q] = acc;
      }
      for (int q = 0; q < p; ++q) out[k + (long)M * q] = res[q];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* r2r transforms, FFTW unnormalised definitions                              */
/* ------------------------------------------------------------------------ */

int or_r2r_direct(int kind, int n, const double *in, double *out) {
  const long double pi = PI_L;
  if (n < 1) return -1;
  for (int k = 0; k < n; ++k) {
    long double s = 0.0L;
    switch (kind) {
    case OR_R2HC: {
      /* r_k for k <= n/2 at out[k]; i_k for 0<k<(n+1)/2 at out[n-k] */
      int kk = (k <= n / 2) ? k : n - k;
      for (int j = 0; j < n; ++j) {
        long double ang = 2.0L * pi * (long double)(((long)j * kk) % n) / n;
        s += (k <= n / 2) ? in[j] * cosl(ang) : -in[j] * sinl(ang);
      }
    } break;
    case OR_HC2R: {
      s = in[0];
      for (int m = 1; m < (n + 1) / 2; ++m) {
        long double ang = 2.0L * pi * (long double)(((long)k * m) % n) / n;
        s += 2.0L * (in[m] * cosl(ang) - in[n - m] * sinl(ang));
      }
      if (n % 2 == 0) s += (k % 2 ? -1.0L : 1.0L) * in[n / 2];
    } break;
    case OR_REDFT00:
      if (n < 2) return -1;
      s = in[0] + ((k % 2) ? -1.0L : 1.0L) * in[n - 1];
      for (int j = 1; j < n - 1; ++j)
        s += 2.0L * in[j] * cosl(pi * (long double)(((long)j * k) % (2 * (n - 1))) / (n - 1));
      break;
    case OR_REDFT10:
      for (int j = 0; j < n; ++j)
        s += 2.0L * in[j] * cosl(pi * (long double)(((long)(2 * j + 1) * k) % (4 * n)) / (2 * n));
      break;
    case OR_REDFT01:
      s = in[0];
      for (int j = 1; j < n; ++j)
        s += 2.0L * in[j] * cosl(pi * (long double)(((long)j * (2 * k + 1)) % (4 * n)) / (2 * n));
      break;
    case OR_REDFT11:
      for (int j = 0; j < n; ++j)
        s += 2.0L * in[j] * cosl(pi * (long double)(((long)(2 * j + 1) * (2 * k + 1)) % (8 * n)) / (4 * n));
      break;
    case OR_RODFT00:
      for (int j = 0; j < n; ++j)
        s += 2.0L * in[j] * sinl(pi * (long double)(((long)(j + 1) * (k + 1)) % (2 * (n + 1))) / (n + 1));
      break;
    case OR_RODFT10:
      for (int j = 0; j < n; ++j)
        s += 2.0L * in[j] * sinl(pi * (long double)(((long)(2 * j + 1) * (k + 1)) % (4 * n)) / (2 * n));
      break;
    case OR_RODFT01:
      s = ((k % 2) ? -1.0L : 1.0L) * in[n - 1];
      for (int j = 0; j < n - 1; ++j)
        s += 2.0L * in[j] * sinl(pi * (long double)(((long)(j + 1) * (2 * k + 1)) % (4 * n)) / (2 * n));
      break;
    case OR_RODFT11:
      for (int j = 0; j < n; ++j)
        s += 2.0L * in[j] * sinl(pi * (long double)(((long)(2 * j + 1) * (2 * k + 1)) % (8 * n)) / (4 * n));
      break;
    default:
      return -1;
    }
    out[k] = (double)s;
  }
  return 0;
}

/* A plan evaluates each kind through one complex DFT of length M of the
 * (anti)symmetric extension: M = n (R2HC/HC2R), 2n (the 10/01/11 kinds),
 * 2(n-1) (REDFT00) or 2(n+1) (RODFT00).  pre/post hold the phase factors. */
typedef struct {
  int kind, n, M, sign;
  cplx *tw, *pre, *post;
} r2r_plan;

static int plan_make(r2r_plan *pl, int kind, int n) {
  memset(pl, 0, sizeof(*pl));
  pl->kind = kind;
  pl->n = n;
  switch (kind) {
  case OR_R2HC: pl->M = n; pl->sign = -1; break;
  case OR_HC2R: pl->M = n; pl->sign = +1; break;
  case OR_REDFT10: case OR_RODFT10: case OR_REDFT11: case OR_RODFT11:
    pl->M = 2 * n; pl->sign = -1; break;
  case OR_REDFT01: case OR_RODFT01: pl->M = 2 * n; pl->sign = +1; break;
  case OR_REDFT00: if (n < 2) return -1; pl->M = 2 * (n - 1); pl->sign = -1; break;
  case OR_RODFT00: pl->M = 2 * (n + 1); pl->sign = -1; break;
  default: return -1;
  }
  pl->tw = make_twiddles(pl->M, pl->sign);
  pl->pre = (cplx *)malloc(sizeof(cplx) * (size_t)(n + 2));
  pl->post = (cplx *)malloc(sizeof(cplx) * (size_t)(n + 2));
  for (int k = 0; k <= n + 1; ++k) {
    pl->pre[k] = 0;
    pl->post[k] = 0;
    if (kind == OR_REDFT10 || kind == OR_RODFT10) pl->post[k] = conj(expi_frac(k, 2L * n));
    if (kind == OR_REDFT01 || kind == OR_RODFT01) pl->pre[k] = expi_frac(k, 2L * n);
    if (kind == OR_REDFT11 || kind == OR_RODFT11) {
      pl->pre[k] = conj(expi_frac(k, 2L * n));
      pl->post[k] = conj(expi_frac(2L * k + 1, 4L * n));
    }
  }
  return 0;
}

static void plan_free(r2r_plan *pl) {
  free(pl->tw);
  free(pl->pre);
  free(pl->post);
}

/* work holds >= 4*M complex values. */
static void plan_exec(const r2r_plan *pl, const double *in, double *out, cplx *work) {
  int n = pl->n, M = pl->M;
  cplx *y = work, *Y = work + M, *scr = work + 2 * M;
  for (int m = 0; m < M; ++m) y[m] = 0;
  switch (pl->kind) {
  case OR_R2HC:
    for (int j = 0; j < n; ++j) y[j] = in[j];
    break;
  case OR_HC2R:
    y[0] = in[0];
    for (int k = 1; k < (n + 1) / 2; ++k) {
      y[k] = in[k] + I * in[n - k];
      y[n - k] = in[k] - I * in[n - k];
    }
    if (n % 2 == 0) y[n / 2] = in[n / 2];
    break;
  case OR_REDFT10: case OR_RODFT10: {
    double sg = (pl->kind == OR_REDFT10) ? 1.0 : -1.0;
    for (int j = 0; j < n; ++j) { y[j] = in[j]; y[M - 1 - j] = sg * in[j]; }
  } break;
  case OR_REDFT01:
    for (int k = 0; k < n; ++k) y[k] = ((k == 0) ? 1.0 : 2.0) * in[k] * pl->pre[k];
    break;
  case OR_RODFT01:
    for (int k = 0; k < n; ++k) y[k + 1] = ((k == n - 1) ? 1.0 : 2.0) * in[k] * pl->pre[k + 1];
    break;
  case OR_REDFT11: case OR_RODFT11:
    for (int j = 0; j < n; ++j) y[j] = in[j] * pl->pre[j];
    break;
  case OR_REDFT00:
    for (int j = 0; j < n; ++j) y[j] = in[j];
    for (int j = 1; j < n - 1; ++j) y[M - j] = in[j];
    break;
  case OR_RODFT00:
    for (int j = 0; j < n; ++j) { y[j + 1] = in[j]; y[M - (j + 1)] = -in[j]; }
    break;
  }
  dft_rec(M, y, 1, Y, pl->tw, M, 1, scr);
  switch (pl->kind) {
  case OR_R2HC:
    for (int k = 0; k <= n / 2; ++k) out[k] = creal(Y[k]);
    for (int k = 1; k < (n + 1) / 2; ++k) out[n - k] = cimag(Y[k]);
    break;
  case OR_HC2R:
    for (int j = 0; j < n; ++j) out[j] = creal(Y[j]);
    break;
  /* DFT_{2n}(y)_m = e^{i pi m/2n} C_m (even ext.) or e^{i pi m/2n}(-2i) S_m */
  case OR_REDFT10:
    for (int k = 0; k < n; ++k) out[k] = creal(pl->post[k] * Y[k]);
    break;
  case OR_RODFT10:
    for (int k = 0; k < n; ++k) out[k] = creal(I * pl->post[k + 1] * Y[k + 1]);
    break;
  case OR_REDFT01:
    for (int j = 0; j < n; ++j) out[j] = creal(Y[j]);
    break;
  case OR_RODFT01:
    for (int j = 0; j < n; ++j) out[j] = cimag(Y[j]);
    break;
  case OR_REDFT11:
    for (int k = 0; k < n; ++k) out[k] = 2.0 * creal(pl->post[k] * Y[k]);
    break;
  case OR_RODFT11:
    for (int k = 0; k < n; ++k) out[k] = -2.0 * cimag(pl->post[k] * Y[k]);
    break;
  case OR_REDFT00:
    for (int k = 0; k < n; ++k) out[k] = creal(Y[k]);
    break;
  case OR_RODFT00:
    for (int k = 0; k < n; ++k) out[k] = creal(I * Y[k + 1]);
    break;
  }
}

int or_r2r(int kind, int n, const double *in, double *out) {
  if (n < 1) return -1;
  r2r_plan pl;
  if (plan_make(&pl, kind, n)) return -1;
  cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)(4 * pl.M + 4));
  plan_exec(&pl, in, out, work);
  free(work);
  plan_free(&pl);
  return 0;
}

int or_r2r_many(int kind, int n, long is, int h1n, long h1s, int h2n, long h2s,
                double *data) {
  if (n < 1) return -1;
  r2r_plan pl;
  if (plan_make(&pl, kind, n)) return -1;
#pragma omp parallel
  {
    cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)(4 * pl.M + 4));
    double *tin = (double *)malloc(sizeof(double) * (size_t)n);
    double *tout = (double *)malloc(sizeof(double) * (size_t)n);
#pragma omp for collapse(2) schedule(static)
    for (int l2 = 0; l2 < h2n; ++l2)
      for (int l1 = 0; l1 < h1n; ++l1) {
        double *line = data + (long)l1 * h1s + (long)l2 * h2s;
        for (int e = 0; e < n; ++e) tin[e] = line[e * is];
        plan_exec(&pl, tin, tout, work);
        for (int e = 0; e < n; ++e) line[e * is] = tout[e];
      }
    free(work); free(tin); free(tout);
  }
  plan_free(&pl);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* initsolver (src/initsolver.f90)                                           */
/* ------------------------------------------------------------------------ */

/* pi as the reference defines it: acos(-1._rp)
 * (src/apps/two_phase_inc_isot/param.f90:99). */
static double ref_pi(void) { return acos(-1.0); }

/* src/initsolver.f90:122-186, CPU branch (no _OPENACC packing). */
void or_eigenvalues(int n, char bc0, char bc1, char c_or_f, double *lambda) {
  const double pi = ref_pi();
  double s;
  for (int l = 1; l <= n; ++l) lambda[l - 1] = 0.0;
  if (bc0 == 'P' && bc1 == 'P') {
    for (int l = 1; l <= n; ++l) {
      s = sin((1.0 * (l - 1)) * pi / (1.0 * n));
      lambda[l - 1] = -4.0 * (s * s);
    }
  } else if (bc0 == 'N' && bc1 == 'N') {
    if (c_or_f == 'c') {
      for (int l = 1; l <= n; ++l) {
        s = sin((1.0 * (l - 1)) * pi / (2.0 * n));
        lambda[l - 1] = -4.0 * (s * s);
      }
    } else {
      for (int l = 1; l <= n; ++l) {
        s = sin((1.0 * (l - 1)) * pi / (2.0 * (n - 1 + 1)));
        lambda[l - 1] = -4.0 * (s * s);
      }
    }
  } else if (bc0 == 'D' && bc1 == 'D') {
    if (c_or_f == 'c') {
      for (int l = 1; l <= n; ++l) {
        s = sin((1.0 * (l - 0)) * pi / (2.0 * n));
        lambda[l - 1] = -4.0 * (s * s);
      }
    } else {
      for (int l = 1; l <= n - 1; ++l) {
        s = sin((1.0 * (l - 0)) * pi / (2.0 * (n + 1 - 1)));
        lambda[l - 1] = -4.0 * (s * s);
      }
    }
  } else if ((bc0 == 'N' && bc1 == 'D') || (bc0 == 'D' && bc1 == 'N')) {
    for (int l = 1; l <= n; ++l) {
      s = sin((1.0 * (2 * l - 1)) * pi / (4.0 * n));
      lambda[l - 1] = -4.0 * (s * s);
    }
  }
}

/* src/initsolver.f90:188-246.  dzci/dzfi point at element (1-nh_d). */
void or_tridmatrix(char bc0, char bc1, int n, double dzi, int nh_d,
                   const double *dzci, const double *dzfi, char c_or_f,
                   double *a, double *b, double *c) {
  (void)dzi;
  const double *zci = dzci + (nh_d - 1); /* zci[k] == dzci(k) */
  const double *zfi = dzfi + (nh_d - 1);
  for (int k = 1; k <= n; ++k) {
    if (c_or_f == 'c') {
      a[k - 1] = zfi[k] * zci[k - 1];
      c[k - 1] = zfi[k] * zci[k];
    } else {
      a[k - 1] = zfi[k] * zci[k];
      c[k - 1] = zfi[k + 1] * zci[k];
    }
  }
  for (int k = 0; k < n; ++k) b[k] = -(a[k] + c[k]);
  double factor[2];
  char bc[2] = {bc0, bc1};
  for (int ib = 0; ib < 2; ++ib)
    factor[ib] = (bc[ib] == 'P') ? 0.0 : (bc[ib] == 'D') ? -1.0 : 1.0;
  if (c_or_f == 'c') {
    b[0] = b[0] + factor[0] * a[0];
    b[n - 1] = b[n - 1] + factor[1] * c[n - 1];
  } else {
    if (bc0 == 'N') b[0] = b[0] + factor[0] * a[0];
    if (bc1 == 'N') b[n - 1] = b[n - 1] + factor[1] * c[n - 1];
  }
}

/* src/initsolver.f90:248-298. */
void or_bc_rhs(char cbc0, char cbc1, int n1, int n2, const double bc[2],
               const double dlc[2], const double dlf[2], char c_or_f,
               double *rhs) {
  double factor[2];
  char cbc[2] = {cbc0, cbc1};
  for (int ib = 0; ib < 2; ++ib) {
    double sgn = (ib == 0) ? 1.0 : -1.0;
    if (cbc[ib] == 'P')
      factor[ib] = 0.0;
    else if (cbc[ib] == 'D')
      factor[ib] = (c_or_f == 'c') ? -2.0 * bc[ib] : -bc[ib];
    else
      factor[ib] = (c_or_f == 'c') ? sgn * dlc[ib] * bc[ib] : sgn * dlf[ib] * bc[ib];
  }
  for (int ib = 0; ib < 2; ++ib) {
    double v = factor[ib] / dlc[ib] / dlf[ib];
    for (long m = 0; m < (long)n1 * n2; ++m) rhs[ib * (long)n1 * n2 + m] = v;
  }
}

/* src/fft.f90:233-291. */
int or_find_fft(char bc0, char bc1, char c_or_f, int *kf, int *kb,
                double norm[2]) {
  char s[3] = {bc0, bc1, 0};
  if (c_or_f == 'c') {
    if (!strcmp(s, "PP")) { *kf = OR_R2HC; *kb = OR_HC2R; norm[0] = 1; norm[1] = 0; }
    else if (!strcmp(s, "NN")) { *kf = OR_REDFT10; *kb = OR_REDFT01; norm[0] = 2; norm[1] = 0; }
    else if (!strcmp(s, "DD")) { *kf = OR_RODFT10; *kb = OR_RODFT01; norm[0] = 2; norm[1] = 0; }
    else if (!strcmp(s, "ND")) { *kf = OR_REDFT11; *kb = OR_REDFT11; norm[0] = 2; norm[1] = 0; }
    else if (!strcmp(s, "DN")) { *kf = OR_RODFT11; *kb = OR_RODFT11; norm[0] = 2; norm[1] = 0; }
    else return -1;
  } else if (c_or_f == 'f') {
    if (!strcmp(s, "PP")) { *kf = OR_R2HC; *kb = OR_HC2R; norm[0] = 1; norm[1] = 0; }
    else if (!strcmp(s, "NN")) { *kf = OR_REDFT00; *kb = OR_REDFT00; norm[0] = 2; norm[1] = -1; }
    else if (!strcmp(s, "DD")) { *kf = OR_RODFT00; *kb = OR_RODFT00; norm[0] = 2; norm[1] = 1; }
    else if (!strcmp(s, "ND")) { *kf = OR_REDFT10; *kb = OR_REDFT01; norm[0] = 2; norm[1] = 0; }
    else if (!strcmp(s, "DN")) { *kf = OR_RODFT01; *kb = OR_RODFT10; norm[0] = 2; norm[1] = 0; }
    else return -1;
  } else
    return -1;
  return 0;
}

/* src/initsolver.f90:21-120 at one rank (dims = 1, ijk_start_z = 0) and
 * fftini's size/normalisation logic (src/fft.f90:71-87, :110-125, :150). */
int or_initsolver(const int ng[3], const double dli[3], int nh_d,
                  const double *dzci, const double *dzfi, const char cbc[6],
                  const double bc[6], const char c_or_f[3], double *lambdaxy,
                  double *a, double *b, double *c, int kinds[4], int sizes[2],
                  double *normfft, double *rhsbx, double *rhsby, double *rhsbz) {
  int nx = ng[0], ny = ng[1], nz = ng[2];
  double *lx = malloc(sizeof(double) * nx), *ly = malloc(sizeof(double) * ny);
  or_eigenvalues(nx, cbc[0], cbc[1], c_or_f[0], lx);
  for (int i = 0; i < nx; ++i) lx[i] = lx[i] * (dli[0] * dli[0]);
  or_eigenvalues(ny, cbc[2], cbc[3], c_or_f[1], ly);
  for (int j = 0; j < ny; ++j) ly[j] = ly[j] * (dli[1] * dli[1]);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) lambdaxy[i + (long)nx * j] = lx[i] + ly[j];
  free(lx); free(ly);
  or_tridmatrix(cbc[4], cbc[5], nz, dli[2], nh_d, dzci, dzfi, c_or_f[2], a, b, c);
  double dl[3];
  for (int d = 0; d < 3; ++d) dl[d] = 1.0 / dli[d];
  const double *zci = dzci + (nh_d - 1), *zfi = dzfi + (nh_d - 1);
  double dzc0 = 1.0 / zci[0], dzcn = 1.0 / zci[nz], dzcnm = 1.0 / zci[nz - 1],
         dzc1 = 1.0 / zci[1];
  double dzf1 = 1.0 / zfi[1], dzfn = 1.0 / zfi[nz];
  double dlx[2] = {dl[0], dl[0]}, dly[2] = {dl[1], dl[1]};
  or_bc_rhs(cbc[0], cbc[1], ny, nz, bc + 0, dlx, dlx, c_or_f[0], rhsbx);
  or_bc_rhs(cbc[2], cbc[3], nx, nz, bc + 2, dly, dly, c_or_f[1], rhsby);
  double dlf[2] = {dzf1, dzfn};
  if (c_or_f[2] == 'c') {
    double dlc[2] = {dzc0, dzcn};
    or_bc_rhs(cbc[4], cbc[5], nx, ny, bc + 4, dlc, dlf, c_or_f[2], rhsbz);
  } else {
    double dlc[2] = {dzc1, dzcnm};
    or_bc_rhs(cbc[4], cbc[5], nx, ny, bc + 4, dlc, dlf, c_or_f[2], rhsbz);
  }
  double norm[2];
  double nf = 1.0;
  int ix = (cbc[0] == 'D' && cbc[1] == 'D' && c_or_f[0] == 'f') ? 1 : 0;
  int iy = (cbc[2] == 'D' && cbc[3] == 'D' && c_or_f[1] == 'f') ? 1 : 0;
  if (or_find_fft(cbc[0], cbc[1], c_or_f[0], &kinds[0], &kinds[1], norm)) return -1;
  nf = nf * norm[0] * (nx + norm[1] - ix);
  if (or_find_fft(cbc[2], cbc[3], c_or_f[1], &kinds[2], &kinds[3], norm)) return -1;
  nf = nf * norm[0] * (ny + norm[1] - iy);
  *normfft = 1.0 / nf;
  sizes[0] = nx - ix;
  sizes[1] = ny - iy;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Tridiagonal solvers (src/solver_cpu.f90:117-223)                           */
/* ------------------------------------------------------------------------ */

/* src/solver_cpu.f90:187-223; p is strided. */
static void dgtsv_pin(int n, const double *a, const double *b, const double *c,
                      double *p, long st, int force_zero_last) {
  double *d = malloc(sizeof(double) * (size_t)n);
  double z;
  z = 1.0 / b[0];
  d[0] = c[0] * z;
  p[0] = p[0] * z;
  for (int l = 1; l < n - 1; ++l) {
    z = 1.0 / (b[l] - a[l] * d[l - 1]);
    d[l] = c[l] * z;
    p[l * st] = (p[l * st] - a[l] * p[(l - 1) * st]) * z;
  }
  z = b[n - 1] - a[n - 1] * d[n - 2];
  if (z != 0.0 && !force_zero_last)
    p[(n - 1) * st] = (p[(n - 1) * st] - a[n - 1] * p[(n - 2) * st]) / z;
  else
    p[(n - 1) * st] = 0.0;
  for (int l = n - 2; l >= 0; --l) p[l * st] = p[l * st] - d[l] * p[(l + 1) * st];
  free(d);
}

void or_dgtsv_homebrewed(int n, const double *a, const double *b,
                         const double *c, double *p, long stride) {
  dgtsv_pin(n, a, b, c, p, stride, 0);
}

/* src/solver_cpu.f90:117-145.  The singular column of an NN/NN... problem
 * has lambda == 0 and zero row sums; pin_gauge forces the p(n) = 0 branch. */
void or_gaussel(int nx, int ny, int n, const double *a, const double *b,
                const double *c, const double *lambdaxy, double *pz,
                int pin_gauge) {
  long st = (long)nx * ny;
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double *bb = malloc(sizeof(double) * (size_t)n);
      double lam = lambdaxy[i + (long)nx * j];
      for (int k = 0; k < n; ++k) bb[k] = b[k] + lam;
      /* pin_gauge != 0 is passed only for a z-matrix that is singular at
       * lambda == 0 (NN in z); that column then takes p(n) = 0. */
      int force = (pin_gauge && lam == 0.0);
      dgtsv_pin(n, a, bb, c, pz + i + (long)nx * j, st, force);
      free(bb);
    }
}

/* src/solver_cpu.f90:147-185. */
void or_gaussel_periodic(int nx, int ny, int n, const double *a,
                         const double *b, const double *c,
                         const double *lambdaxy, double *pz, int pin_gauge) {
  long st = (long)nx * ny;
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      double *bb = malloc(sizeof(double) * (size_t)n);
      double *p1 = malloc(sizeof(double) * (size_t)n);
      double *p2 = calloc((size_t)n, sizeof(double));
      double *pc = pz + i + (long)nx * j;
      double lam = lambdaxy[i + (long)nx * j];
      for (int k = 0; k < n; ++k) bb[k] = b[k] + lam;
      for (int k = 0; k < n - 1; ++k) p1[k] = pc[k * st];
      dgtsv_pin(n - 1, a, bb, c, p1, 1, 0);
      p2[0] = -a[0];
      p2[n - 2] = -c[n - 2];
      dgtsv_pin(n - 1, a, bb, c, p2, 1, 0);
      if (pin_gauge && lam == 0.0) {
        pc[(n - 1) * st] = 0.0;
      } else {
        pc[(n - 1) * st] = (pc[(n - 1) * st] - c[n - 1] * p1[0] - a[n - 1] * p1[n - 2]) /
                           (bb[n - 1] + c[n - 1] * p2[0] + a[n - 1] * p2[n - 2]);
      }
      for (int k = 0; k < n - 1; ++k) pc[k * st] = p1[k] + p2[k] * pc[(n - 1) * st];
      free(bb); free(p1); free(p2);
    }
}

/* src/solver_cpu.f90:20-115, _DECOMP_X at one rank: the transposes are
 * identity copies, so one dense buffer px holds the field throughout. */
int or_solver_cpu(const int n[3], const int kinds[4], const int sizes[2],
                  double normfft, const double *lambdaxy, const double *a,
                  const double *b, const double *c, char bcz0, char bcz1,
                  char c_or_f_z, double *p, int pin_gauge) {
  int nx = n[0], ny = n[1], nz = n[2];
  long px_ = nx + 2, py_ = ny + 2;
  double *px = malloc(sizeof(double) * (size_t)nx * ny * nz);
  if (!px) return -1;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        px[i + (long)nx * (j + (long)ny * k)] = p[(i + 1) + px_ * ((j + 1) + py_ * (k + 1))];
  int err = 0;
  err |= or_r2r_many(kinds[0], sizes[0], 1, ny, nx, nz, (long)nx * ny, px);
  err |= or_r2r_many(kinds[2], sizes[1], nx, nx, 1, nz, (long)nx * ny, px);
  int q = (c_or_f_z == 'f' && bcz1 == 'D') ? 1 : 0;
  if (bcz0 == 'P' && bcz1 == 'P')
    or_gaussel_periodic(nx, ny, nz - q, a, b, c, lambdaxy, px, pin_gauge);
  else
    or_gaussel(nx, ny, nz - q, a, b, c, lambdaxy, px,
               pin_gauge && bcz0 == 'N' && bcz1 == 'N');
  err |= or_r2r_many(kinds[3], sizes[1], nx, nx, 1, nz, (long)nx * ny, px);
  err |= or_r2r_many(kinds[1], sizes[0], 1, ny, nx, nz, (long)nx * ny, px);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        p[(i + 1) + px_ * ((j + 1) + py_ * (k + 1))] =
            px[i + (long)nx * (j + (long)ny * k)] * normfft;
  free(px);
  return err ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* Field kernels                                                              */
/* ------------------------------------------------------------------------ */

/* index of (i,j,k) in a halo-h block whose first element is (1-h,1-h,1-h) */
#define IX(i, j, k, h, nx, ny) \
  ((long)((i) - 1 + (h)) + (long)((nx) + 2 * (h)) * ((long)((j) - 1 + (h)) + (long)((ny) + 2 * (h)) * (long)((k) - 1 + (h))))
#define ZI(arr, k, nh_d) ((arr)[(k) + (nh_d) - 1])

/* src/fillps.f90:16-69 with _CONSTANT_COEFFS_POISSON (:55-57). */
void or_fillps(int nx, int ny, int nz, int nh_d, int nh_u, double dxi,
               double dyi, double dzi, const double *dzfi, double dti,
               double rho0, const double *u, const double *v, const double *w,
               double *p) {
  (void)dzi;
  double dtidxi = dti * dxi, dtidyi = dti * dyi;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        double val = ((w[IX(i, j, k, nh_u, nx, ny)] - w[IX(i, j, k - 1, nh_u, nx, ny)]) * dti * ZI(dzfi, k, nh_d) +
                      (v[IX(i, j, k, nh_u, nx, ny)] - v[IX(i, j - 1, k, nh_u, nx, ny)]) * dtidyi) +
                     (u[IX(i, j, k, nh_u, nx, ny)] - u[IX(i - 1, j, k, nh_u, nx, ny)]) * dtidxi;
        p[IX(i, j, k, 1, nx, ny)] = val * rho0;
      }
}

/* src/correc.f90:16-81. */
void or_correc(int nx, int ny, int nz, int nh_d, int nh_u, double dxi,
               double dyi, double dzi, const double *dzci, double dt,
               double rho0, const double *p, double *u, double *v, double *w,
               const double *rho, int const_coeff) {
  (void)dzi;
  double rho0i = 1.0 / rho0, factori = dt * dxi, factorj = dt * dyi;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        long c1 = IX(i, j, k, 1, nx, ny), cu = IX(i, j, k, nh_u, nx, ny);
        double pc = p[c1];
        double dpx = p[IX(i + 1, j, k, 1, nx, ny)] - pc;
        double dpy = p[IX(i, j + 1, k, 1, nx, ny)] - pc;
        double dpz = p[IX(i, j, k + 1, 1, nx, ny)] - pc;
        double fz = dt * ZI(dzci, k, nh_d);
        if (const_coeff) {
          u[cu] = u[cu] - factori * dpx * rho0i;
          v[cu] = v[cu] - factorj * dpy * rho0i;
          w[cu] = w[cu] - fz * dpz * rho0i;
        } else {
          double rc = rho[c1];
          double rhox = 0.5 * (rho[IX(i + 1, j, k, 1, nx, ny)] + rc);
          double rhoy = 0.5 * (rho[IX(i, j + 1, k, 1, nx, ny)] + rc);
          double rhoz = 0.5 * (rho[IX(i, j, k + 1, 1, nx, ny)] + rc);
          u[cu] = u[cu] - factori * dpx / rhox;
          v[cu] = v[cu] - factorj * dpy / rhoy;
          w[cu] = w[cu] - fz * dpz / rhoz;
        }
      }
}

/* src/source.f90:247-309 (pres_tw_src). */
void or_pres_tw_src(int nx, int ny, int nz, double dxi, double dyi, double dzi,
                    int nh_d, int nh_u, const double *dzci, double rho0i,
                    double f_t12, double f_t12_o, const double *p,
                    const double *pold, const double *rho, double *u,
                    double *v, double *w, int const_coeff) {
  (void)dzi;
  double f1 = 1.0 + (f_t12 / f_t12_o), f2 = (f_t12 / f_t12_o);
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        long c1 = IX(i, j, k, 1, nx, ny), cu = IX(i, j, k, nh_u, nx, ny);
        long ipx = IX(i + 1, j, k, 1, nx, ny), jpx = IX(i, j + 1, k, 1, nx, ny),
             kpx = IX(i, j, k + 1, 1, nx, ny);
        double rhoxi = 1.0 / (0.5 * (rho[ipx] + rho[c1]));
        double rhoyi = 1.0 / (0.5 * (rho[jpx] + rho[c1]));
        double rhozi = 1.0 / (0.5 * (rho[kpx] + rho[c1]));
        double dzk = ZI(dzci, k, nh_d);
        if (const_coeff) {
          double hc = f1 * p[c1] - f2 * pold[c1];
          u[cu] = u[cu] + f_t12 * ((-((p[ipx] - p[c1]) * dxi)) * rho0i -
                                   (rhoxi - rho0i) * ((f1 * p[ipx] - f2 * pold[ipx]) - hc) * dxi);
          v[cu] = v[cu] + f_t12 * ((-((p[jpx] - p[c1]) * dyi)) * rho0i -
                                   (rhoyi - rho0i) * ((f1 * p[jpx] - f2 * pold[jpx]) - hc) * dyi);
          w[cu] = w[cu] + f_t12 * ((-((p[kpx] - p[c1]) * dzk)) * rho0i -
                                   (rhozi - rho0i) * ((f1 * p[kpx] - f2 * pold[kpx]) - hc) * dzk);
        } else {
          u[cu] = u[cu] + f_t12 * (-((p[ipx] - p[c1]) * dxi)) * rhoxi;
          v[cu] = v[cu] + f_t12 * (-((p[jpx] - p[c1]) * dyi)) * rhoyi;
          w[cu] = w[cu] + f_t12 * (-((p[kpx] - p[c1]) * dzk)) * rhozi;
        }
      }
}

/* src/source.f90:311-346 (pres_sp_src). */
void or_pres_sp_src(int nx, int ny, int nz, double f_t12, double dxi,
                    double dyi, double dzi, int nh_d, int nh_u,
                    const double *dzci, double rho0i, const double *pold,
                    double *u, double *v, double *w) {
  (void)dzi;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        long c1 = IX(i, j, k, 1, nx, ny), cu = IX(i, j, k, nh_u, nx, ny);
        double pc = pold[c1];
        u[cu] = u[cu] + f_t12 * (-((pold[IX(i + 1, j, k, 1, nx, ny)] - pc) * dxi)) * rho0i;
        v[cu] = v[cu] + f_t12 * (-((pold[IX(i, j + 1, k, 1, nx, ny)] - pc) * dyi)) * rho0i;
        w[cu] = w[cu] + f_t12 * (-((pold[IX(i, j, k + 1, 1, nx, ny)] - pc) * ZI(dzci, k, nh_d))) * rho0i;
      }
}

/* src/chkdiv.f90:18-69 (serial loop order k, j, i). */
void or_chkdiv(int nx, int ny, int nz, double dxi, double dyi, double dzi,
               int nh_d, int nh_u, const double *dzfi, const double *u,
               const double *v, const double *w, double *divtot,
               double *divmax) {
  (void)dzi;
  double tot = 0.0, mx = 0.0;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        double div = ((w[IX(i, j, k, nh_u, nx, ny)] - w[IX(i, j, k - 1, nh_u, nx, ny)]) * ZI(dzfi, k, nh_d) +
                      (v[IX(i, j, k, nh_u, nx, ny)] - v[IX(i, j - 1, k, nh_u, nx, ny)]) * dyi) +
                     (u[IX(i, j, k, nh_u, nx, ny)] - u[IX(i - 1, j, k, nh_u, nx, ny)]) * dxi;
        double ad = fabs(div);
        mx = (ad > mx) ? ad : mx;
        tot = tot + div;
      }
  *divtot = tot;
  *divmax = mx;
}

/* set_bc for nh_p = 1, centred (src/bound.f90:227-646). */
static void set_bc1(int nx, int ny, int nz, char ctype, int ibound, int idir,
                    double rvalue, double dr, double *p) {
  long sx = nx + 2, sy = (long)(nx + 2) * (ny + 2);
  int n[3] = {nx, ny, nz};
  long st[3] = {1, sx, sy};
  /* extents of the two tangential directions including the halo */
  int t1 = (idir == 0) ? 1 : 0, t2 = (idir == 2) ? 1 : 2;
  double factor = rvalue, sgn = 0.0;
  if (ctype == 'D') { factor = 2.0 * factor; sgn = -1.0; }
  if (ctype == 'N') { factor = (ibound == 0) ? -dr * factor : dr * factor; sgn = 1.0; }
  int nd = n[idir];
  for (int b2 = 0; b2 <= n[t2] + 1; ++b2)
    for (int b1 = 0; b1 <= n[t1] + 1; ++b1) {
      long base = b1 * st[t1] + b2 * st[t2];
      long s = st[idir];
      if (ctype == 'P') {
        if (ibound == 0) p[base + 0 * s] = p[base + nd * s];
        else p[base + (nd + 1) * s] = p[base + 1 * s];
      } else {
        if (ibound == 0) p[base + 0 * s] = factor + sgn * p[base + 1 * s];
        else p[base + (nd + 1) * s] = factor + sgn * p[base + nd * s];
      }
    }
}

/* boundp at one rank (src/bound.f90:146-225): periodic y and z halos come
 * from updthalo (a self-exchange at one rank), then set_bc for x, y, z. */
void or_boundp(const char cbc[6], const int n[3], const double bc[6], int nh_d,
               const double dl[3], const double *dzc, const double *dzf,
               double *p) {
  (void)dzf;
  int nx = n[0], ny = n[1], nz = n[2];
  if (cbc[2] == 'P') { set_bc1(nx, ny, nz, 'P', 0, 1, 0, 0, p); set_bc1(nx, ny, nz, 'P', 1, 1, 0, 0, p); }
  if (cbc[4] == 'P') { set_bc1(nx, ny, nz, 'P', 0, 2, 0, 0, p); set_bc1(nx, ny, nz, 'P', 1, 2, 0, 0, p); }
  set_bc1(nx, ny, nz, cbc[0], 0, 0, bc[0], dl[0], p);
  set_bc1(nx, ny, nz, cbc[1], 1, 0, bc[1], dl[0], p);
  if (cbc[2] != 'P') {
    set_bc1(nx, ny, nz, cbc[2], 0, 1, bc[2], dl[1], p);
    set_bc1(nx, ny, nz, cbc[3], 1, 1, bc[3], dl[1], p);
  }
  if (cbc[4] != 'P') {
    set_bc1(nx, ny, nz, cbc[4], 0, 2, bc[4], ZI(dzc, 0, nh_d), p);
    set_bc1(nx, ny, nz, cbc[5], 1, 2, bc[5], ZI(dzc, nz, nh_d), p);
  }
}

/* set_bc (src/bound.f90:227-646) on a field with halo nh: the reference's
 * cases literally, the layer loop q outermost, so that a later layer
 * overwrites an earlier one where the reference's does (the upper face of
 * the face-normal kinds).  dr[q], q < nh: the spacings of the N factor. */
static void set_bc_nh(const int n[3], int nh, char ctype, int ibound, int idir,
                      int centered, double rvalue, const double *dr, double *p) {
  long ext[3] = {n[0] + 2 * nh, n[1] + 2 * nh, n[2] + 2 * nh};
  long st[3] = {1, ext[0], ext[0] * ext[1]};
  int d = idir, t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2, nd = n[d];
  double factor[16], sgn = 0.0;
  for (int q = 0; q < nh; ++q) factor[q] = rvalue;
  if (ctype == 'D' && centered) {
    for (int q = 0; q < nh; ++q) factor[q] = +2.0 * factor[q];
    sgn = -1.0;
  }
  if (ctype == 'N') {
    for (int q = 0; q < nh; ++q) factor[q] = (ibound == 0) ? -dr[q] * factor[q] : dr[q] * factor[q];
    sgn = +1.0;
  }
#define L(a) p[base + ((long)(a) - 1 + nh) * st[d]]
  for (int q = 0; q < nh; ++q)
    for (long b2 = 0; b2 < ext[t2]; ++b2)
      for (long b1 = 0; b1 < ext[t1]; ++b1) {
        long base = b1 * st[t1] + b2 * st[t2];
        const double f = factor[q];
        if (ctype == 'P') {
          L(0 - q) = L(nd - q);
          L(nd + 1 + q) = L(1 + q);
        } else if (centered) {
          if (ibound == 0) L(0 - q) = f + sgn * L(1 + q);
          else L(nd + 1 + q) = f + sgn * L(nd - q);
        } else if (ctype == 'D') {
          if (ibound == 0) L(0 - q) = f;
          else {
            L(nd + q) = f;
            L(nd + 1 + q) = L(nd - 1 - q);
          }
        } else {
          if (ibound == 0) L(0 - q) = 1.0 * f + L(1 + q);
          else {
            L(nd + q) = 1.0 * f + L(nd - 1 - q);
            L(nd + 1 + q) = 2.0 * f + L(nd - 1 - q);
          }
        }
      }
#undef L
}

/* outflow (src/bound.f90:649-773) at one rank: the velocity normal to the
 * face from zero divergence; side 1 writes q = 0..nh (nx + q), side 0
 * q = 0..nh-1 (0 - q). */
static void outflow_nh(const int n[3], int nh, int nh_d, int dir, int side,
                       const double dl[3], const double *dzf, double *u,
                       double *v, double *w) {
  int nx = n[0], ny = n[1], nz = n[2], qmin = nh - 1;
  long sx = nx + 2 * nh, sy = sx * (ny + 2 * nh);
#define F(a, i, j, k) (a)[((long)(i) - 1 + nh) + sx * ((long)(j) - 1 + nh) + sy * ((long)(k) - 1 + nh)]
#define DZF(k) dzf[(k) + nh_d - 1]
  double dx = dl[0], dy = dl[1], dxi = 1.0 / dx, dyi = 1.0 / dy;
  if (dir == 0 && side == 1) {
    int i = nx;
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int q = 0; q <= qmin + 1; ++q)
          F(u, i + q, j, k) = F(u, i - 1 - q, j, k) -
              dx * ((F(v, i + q, j, k) - F(v, i + q, j - 1, k)) * dyi +
                    (F(w, i + q, j, k) - F(w, i + q, j, k - 1)) * (1.0 / DZF(k)));
  } else if (dir == 1 && side == 1) {
    int j = ny;
    for (int k = 1; k <= nz; ++k)
      for (int q = 0; q <= qmin + 1; ++q)
        for (int i = 1; i <= nx; ++i)
          F(v, i, j + q, k) = F(v, i, j - 1 - q, k) -
              dy * ((F(u, i, j + q, k) - F(u, i - 1, j + q, k)) * dxi +
                    (F(w, i, j + q, k) - F(w, i, j + q, k - 1)) * (1.0 / DZF(k)));
  } else if (dir == 2 && side == 1) {
    int k = nz;
    for (int q = 0; q <= qmin + 1; ++q)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i)
          F(w, i, j, k + q) = F(w, i, j, k - 1 - q) -
              DZF(k + q) * ((F(u, i, j, k + q) - F(u, i - 1, j, k + q)) * dxi +
                            (F(v, i, j, k + q) - F(v, i, j - 1, k + q)) * dyi);
  } else if (dir == 0 && side == 0) {
    int i = 0;
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int q = 0; q <= qmin; ++q)
          F(u, i - q, j, k) = F(u, i + 1 + q, j, k) +
              dx * ((F(v, i + 1 + q, j, k) - F(v, i + 1 + q, j - 1, k)) * dyi +
                    (F(w, i + 1 + q, j, k) - F(w, i + 1 + q, j, k - 1)) * (1.0 / DZF(k)));
  } else if (dir == 1 && side == 0) {
    int j = 0;
    for (int k = 1; k <= nz; ++k)
      for (int q = 0; q <= qmin; ++q)
        for (int i = 1; i <= nx; ++i)
          F(v, i, j - q, k) = F(v, i, j + 1 + q, k) +
              dy * ((F(u, i, j + 1 + q, k) - F(u, i - 1, j + 1 + q, k)) * dxi +
                    (F(w, i, j + 1 + q, k) - F(w, i, j + 1 + q, k - 1)) * (1.0 / DZF(k)));
  } else {
    int k = 0;
    for (int q = 0; q <= qmin; ++q)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i)
          F(w, i, j, k - q) = F(w, i, j, k + 1 + q) +
              DZF(k - q) * ((F(u, i, j, k + 1 + q) - F(u, i - 1, j, k + 1 + q)) * dxi +
                            (F(v, i, j, k + 1 + q) - F(v, i, j - 1, k + 1 + q)) * dyi);
  }
#undef F
#undef DZF
}

/* bounduvw at one rank (src/bound.f90:17-144), in the same decomposition
 * model as or_boundp: periodic y and z halos of u, v, w from updthalo (a
 * self-exchange at one rank), then set_bc for the x faces (u face-normal,
 * v and w centred), the y faces and the z faces where they are not
 * periodic, then outflow in the order x, y, z, lower side first.
 * cbc/bc: the reference's (0:1,3,3) arrays, [side + 2*dir + 6*comp]. */
void or_bounduvw(const char cbc[18], const int n[3], const double bc[18],
                 int nh_d, int nh_u, const int isoutflow[6], const double dl[3],
                 const double *dzc, const double *dzf, double *u, double *v,
                 double *w) {
  double *fld[3] = {u, v, w};
  double dr[16] = {0.0};
  int nz = n[2];
#define CB(s, d, c) cbc[(s) + 2 * (d) + 6 * (c)]
#define BV(s, d, c) bc[(s) + 2 * (d) + 6 * (c)]
  for (int c = 0; c < 3; ++c) {
    if (CB(0, 1, c) == 'P') set_bc_nh(n, nh_u, 'P', 0, 1, 1, 0.0, dr, fld[c]);
    if (CB(0, 2, c) == 'P') set_bc_nh(n, nh_u, 'P', 0, 2, 1, 0.0, dr, fld[c]);
  }
  for (int s = 0; s < 2; ++s) {
    for (int q = 0; q < nh_u; ++q) dr[q] = dl[0];
    for (int c = 0; c < 3; ++c)
      set_bc_nh(n, nh_u, CB(s, 0, c), s, 0, c != 0, BV(s, 0, c), dr, fld[c]);
  }
  for (int s = 0; s < 2; ++s) {
    for (int q = 0; q < nh_u; ++q) dr[q] = dl[1];
    for (int c = 0; c < 3; ++c)
      if (CB(s, 1, c) != 'P') set_bc_nh(n, nh_u, CB(s, 1, c), s, 1, c != 1, BV(s, 1, c), dr, fld[c]);
  }
  for (int s = 0; s < 2; ++s) {
    for (int q = 0; q < nh_u; ++q) dr[q] = (s == 0) ? dzc[-q + nh_d - 1] : dzc[nz + q + nh_d - 1];
    for (int c = 0; c < 2; ++c)
      if (CB(s, 2, c) != 'P') set_bc_nh(n, nh_u, CB(s, 2, c), s, 2, 1, BV(s, 2, c), dr, fld[c]);
    for (int q = 0; q < nh_u; ++q) dr[q] = (s == 0) ? dzf[-q + nh_d - 1] : dzf[nz + q + nh_d - 1];
    if (CB(s, 2, 2) != 'P') set_bc_nh(n, nh_u, CB(s, 2, 2), s, 2, 0, BV(s, 2, 2), dr, fld[2]);
  }
  for (int d = 0; d < 3; ++d)
    for (int s = 0; s < 2; ++s)
      if (isoutflow[s + 2 * d]) outflow_nh(n, nh_u, nh_d, d, s, dl, dzf, u, v, w);
#undef CB
#undef BV
}

/* boundp at one rank for a field with halo nh_p (src/bound.f90:146-225):
 * periodic y and z halos from updthalo, then the centred set_bc of the x
 * faces (periodic included), and of the y and z faces that are not
 * periodic; z spacings dzc(-q) and dzc(nz+q). */
void or_boundp_nh(const char cbc[6], const int n[3], const double bc[6],
                  int nh_d, int nh_p, const double dl[3], const double *dzc,
                  const double *dzf, double *p) {
  (void)dzf;
  double dr[16] = {0.0};
  int nz = n[2];
  if (cbc[2] == 'P') set_bc_nh(n, nh_p, 'P', 0, 1, 1, 0.0, dr, p);
  if (cbc[4] == 'P') set_bc_nh(n, nh_p, 'P', 0, 2, 1, 0.0, dr, p);
  for (int q = 0; q < nh_p; ++q) dr[q] = dl[0];
  set_bc_nh(n, nh_p, cbc[0], 0, 0, 1, bc[0], dr, p);
  set_bc_nh(n, nh_p, cbc[1], 1, 0, 1, bc[1], dr, p);
  for (int q = 0; q < nh_p; ++q) dr[q] = dl[1];
  if (cbc[2] != 'P') set_bc_nh(n, nh_p, cbc[2], 0, 1, 1, bc[2], dr, p);
  if (cbc[3] != 'P') set_bc_nh(n, nh_p, cbc[3], 1, 1, 1, bc[3], dr, p);
  for (int q = 0; q < nh_p; ++q) dr[q] = dzc[-q + nh_d - 1];
  if (cbc[4] != 'P') set_bc_nh(n, nh_p, cbc[4], 0, 2, 1, bc[4], dr, p);
  for (int q = 0; q < nh_p; ++q) dr[q] = dzc[nz + q + nh_d - 1];
  if (cbc[5] != 'P') set_bc_nh(n, nh_p, cbc[5], 1, 2, 1, bc[5], dr, p);
}

/* src/bound.f90:829-944 at one rank. */
void or_updt_rhs_b(int nx, int ny, int nz, const char c_or_f[3],
                   const char cbc[6], const double *rhsbx, const double *rhsby,
                   const double *rhsbz, double *p) {
  int q1 = (c_or_f[0] == 'f' && cbc[1] == 'D') ? 1 : 0;
  int q2 = (c_or_f[1] == 'f' && cbc[3] == 'D') ? 1 : 0;
  int q3 = (c_or_f[2] == 'f' && cbc[5] == 'D') ? 1 : 0;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j) {
      p[IX(1, j, k, 1, nx, ny)] += rhsbx[(j - 1) + (long)ny * (k - 1)];
      p[IX(nx - q1, j, k, 1, nx, ny)] += rhsbx[(j - 1) + (long)ny * (k - 1) + (long)ny * nz];
    }
  for (int k = 1; k <= nz; ++k)
    for (int i = 1; i <= nx; ++i) {
      p[IX(i, 1, k, 1, nx, ny)] += rhsby[(i - 1) + (long)nx * (k - 1)];
      p[IX(i, ny - q2, k, 1, nx, ny)] += rhsby[(i - 1) + (long)nx * (k - 1) + (long)nx * nz];
    }
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      p[IX(i, j, 1, 1, nx, ny)] += rhsbz[(i - 1) + (long)nx * (j - 1)];
      p[IX(i, j, nz - q3, 1, nx, ny)] += rhsbz[(i - 1) + (long)nx * (j - 1) + (long)nx * ny];
    }
}

/* src/initgrid.f90:17-97 with gridpoint_cluster_two_end (:102-118). */
/* gridpoint_cluster_two_end (src/initgrid.f90:102-118) and
 * gridpoint_cluster_one_end (:120-136). */
static double gridpoint(int one_end, double alpha, double z0) {
  if (alpha == 0.0) return z0;
  if (one_end) return 1.0 * (1.0 + tanh((z0 - 1.0) * alpha) / tanh(alpha / 1.0));
  return 0.5 * (1.0 + tanh((z0 - 0.5) * alpha) / tanh(alpha / 2.0));
}

void or_initgrid(int n, double gr, double lz, int nh_d, double *dzc,
                 double *dzf, double *zc, double *zf) {
  or_initgrid_inivel("zer", n, gr, lz, nh_d, dzc, dzf, zc, zf);
}

void or_initgrid_inivel(const char *inivel, int n, double gr, double lz,
                        int nh_d, double *dzc, double *dzf, double *zc,
                        double *zf) {
  /* src/initgrid.f90:32-39: 'hcl' and 'hcp' cluster at the lower end */
  const int one_end = !strncmp(inivel, "hcl", 3) || !strncmp(inivel, "hcp", 3);
#define G(a, k) ((a)[(k) + (nh_d) - 1])
  for (int k = 1; k <= n; ++k) {
    double z0 = (k - 0.0) / (1.0 * n);
    G(zf, k) = gridpoint(one_end, gr, z0) * lz;
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
#undef G
}

void or_apply_laplacian(const int n[3], const char cbc[6], const double dli[3],
                        const double *a, const double *b, const double *c,
                        const double *p, double *out) {
  int nx = n[0], ny = n[1], nz = n[2];
  double dx2 = dli[0] * dli[0], dy2 = dli[1] * dli[1];
#define P3(i, j, k) p[(i) + (long)nx * ((j) + (long)ny * (k))]
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        double pc = P3(i, j, k), pm, pp;
        /* x: ghost values by BC pattern */
        if (i > 0) pm = P3(i - 1, j, k);
        else pm = (cbc[0] == 'P') ? P3(nx - 1, j, k) : (cbc[0] == 'N') ? pc : -pc;
        if (i < nx - 1) pp = P3(i + 1, j, k);
        else pp = (cbc[1] == 'P') ? P3(0, j, k) : (cbc[1] == 'N') ? pc : -pc;
        double lx = (pp - 2.0 * pc + pm) * dx2;
        if (j > 0) pm = P3(i, j - 1, k);
        else pm = (cbc[2] == 'P') ? P3(i, ny - 1, k) : (cbc[2] == 'N') ? pc : -pc;
        if (j < ny - 1) pp = P3(i, j + 1, k);
        else pp = (cbc[3] == 'P') ? P3(i, 0, k) : (cbc[3] == 'N') ? pc : -pc;
        double ly = (pp - 2.0 * pc + pm) * dy2;
        /* z: rows of the tridiagonal (b holds the BC modification) */
        double lz = b[k] * pc;
        if (k > 0) lz += a[k] * P3(i, j, k - 1);
        else if (cbc[4] == 'P') lz += a[k] * P3(i, j, nz - 1);
        if (k < nz - 1) lz += c[k] * P3(i, j, k + 1);
        else if (cbc[5] == 'P') lz += c[k] * P3(i, j, 0);
        out[i + (long)nx * (j + (long)ny * k)] = lx + ly + lz;
      }
#undef P3
}
