/*
 * FluTAS pressure-projection oracle: a plain-C FP64 restatement of the
 * reference CPU path.  TEST INFRASTRUCTURE ONLY.  Only the tests, the smoke
 * check and the CPU-baseline leg of the benchmark may load this library, and
 * only as the checker; the product (libflutas_amd.so) never links it.
 *
 * Array conventions (Fortran column-major, i fastest):
 *   - a 3-D field with halo h is passed as a pointer to element (1-h,1-h,1-h)
 *     of an (nx+2h)x(ny+2h)x(nz+2h) block.  p, pold, rho use h = 1
 *     (p(0:n+1), src/fillps.f90:26); u, v, w use h = nh_u
 *     (src/fillps.f90:25).
 *   - 1-D z metrics dzci/dzfi/dzc/dzf are passed as a pointer to element
 *     (1-nh_d) (src/initsolver.f90:30).
 *   - BC characters are 'P','D','N'; c_or_f is 'c' (cell) or 'f' (face).
 */
#ifndef FLUTAS_ORACLE_H
#define FLUTAS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* FFTW r2r kinds, numbered as in src/fftw.f90:38-61. */
enum {
  OR_R2HC = 0, OR_HC2R = 1, OR_DHT = 2,
  OR_REDFT00 = 3, OR_REDFT01 = 4, OR_REDFT10 = 5, OR_REDFT11 = 6,
  OR_RODFT00 = 7, OR_RODFT01 = 8, OR_RODFT10 = 9, OR_RODFT11 = 10
};

/* OpenMP threads the batched loops use (1 without OpenMP). */
int or_max_threads(void);

/* --- transforms ---------------------------------------------------------- */
/* Unnormalised FFTW r2r definitions, evaluated as direct O(n^2) sums in long
 * double.  The definition the fast path is checked against. */
int or_r2r_direct(int kind, int n, const double *in, double *out);
/* Same transform through a mixed-radix complex FFT of the symmetric
 * extension (O(n log n) for any n; prime factors use direct DFT steps). */
int or_r2r(int kind, int n, const double *in, double *out);
/* Guru-style batched in-place transform (src/fft.f90:75-86, :113-124):
 * length n with element stride `is`, over howmany lines (h1n,h1s)x(h2n,h2s). */
int or_r2r_many(int kind, int n, long is, int h1n, long h1s, int h2n, long h2s,
                double *data);

/* --- initsolver pieces (src/initsolver.f90) ------------------------------- */
void or_eigenvalues(int n, char bc0, char bc1, char c_or_f, double *lambda);
void or_tridmatrix(char bc0, char bc1, int n, double dzi, int nh_d,
                   const double *dzci, const double *dzfi, char c_or_f,
                   double *a, double *b, double *c);
/* rhs has shape (n1,n2,0:1) (src/initsolver.f90:248-298). */
void or_bc_rhs(char cbc0, char cbc1, int n1, int n2, const double bc[2],
               const double dlc[2], const double dlf[2], char c_or_f,
               double *rhs);
/* find_fft (src/fft.f90:233-291): returns 0 or -1 for an unknown pair. */
int or_find_fft(char bc0, char bc1, char c_or_f, int *kind_fwd, int *kind_bwd,
                double norm[2]);

/* Full initsolver at one rank (src/initsolver.f90:21-120).
 * cbc[6], bc[6]: Fortran (0:1,3) order, i.e. cbc[2*dir+side].
 * Outputs: lambdaxy(nx*ny), a,b,c(nz), kinds[4] = {fwd_x,bwd_x,fwd_y,bwd_y},
 * sizes[2] = transform lengths (nx-ix, ny-iy), normfft,
 * rhsbx(ny,nz,0:1), rhsby(nx,nz,0:1), rhsbz(nx,ny,0:1). */
int or_initsolver(const int ng[3], const double dli[3], int nh_d,
                  const double *dzci, const double *dzfi, const char cbc[6],
                  const double bc[6], const char c_or_f[3], double *lambdaxy,
                  double *a, double *b, double *c, int kinds[4], int sizes[2],
                  double *normfft, double *rhsbx, double *rhsby, double *rhsbz);

/* --- tridiagonal solvers (src/solver_cpu.f90:117-223) ---------------------
 * pin_gauge = 0: the reference literally.  pin_gauge = 1: additionally, the
 * singular column (lambda == 0 with NN or PP in z) takes p(n) = 0 exactly. */
void or_dgtsv_homebrewed(int n, const double *a, const double *b,
                         const double *c, double *p, long stride);
void or_gaussel(int nx, int ny, int n, const double *a, const double *b,
                const double *c, const double *lambdaxy, double *pz,
                int pin_gauge);
void or_gaussel_periodic(int nx, int ny, int n, const double *a,
                         const double *b, const double *c,
                         const double *lambdaxy, double *pz, int pin_gauge);

/* solver_cpu at one rank, _DECOMP_X (src/solver_cpu.f90:20-115).
 * p has halo 1; only the interior is read and written. */
int or_solver_cpu(const int n[3], const int kinds[4], const int sizes[2],
                  double normfft, const double *lambdaxy, const double *a,
                  const double *b, const double *c, char bcz0, char bcz1,
                  char c_or_f_z, double *p, int pin_gauge);

/* --- field kernels -------------------------------------------------------- */
void or_fillps(int nx, int ny, int nz, int nh_d, int nh_u, double dxi,
               double dyi, double dzi, const double *dzfi, double dti,
               double rho0, const double *u, const double *v, const double *w,
               double *p);
/* const_coeff = 1 is the _CONSTANT_COEFFS_POISSON branch. */
void or_correc(int nx, int ny, int nz, int nh_d, int nh_u, double dxi,
               double dyi, double dzi, const double *dzci, double dt,
               double rho0, const double *p, double *u, double *v, double *w,
               const double *rho, int const_coeff);
void or_pres_tw_src(int nx, int ny, int nz, double dxi, double dyi, double dzi,
                    int nh_d, int nh_u, const double *dzci, double rho0i,
                    double f_t12, double f_t12_o, const double *p,
                    const double *pold, const double *rho, double *u,
                    double *v, double *w, int const_coeff);
void or_pres_sp_src(int nx, int ny, int nz, double f_t12, double dxi,
                    double dyi, double dzi, int nh_d, int nh_u,
                    const double *dzci, double rho0i, const double *pold,
                    double *u, double *v, double *w);
void or_chkdiv(int nx, int ny, int nz, double dxi, double dyi, double dzi,
               int nh_d, int nh_u, const double *dzfi, const double *u,
               const double *v, const double *w, double *divtot,
               double *divmax);
/* boundp for a field with halo nh_p (src/bound.f90:146-225); p is
 * (1-nh_p : n+nh_p)^3; dzc over (1-nh_d : nz+nh_d), nh_d >= nh_p. */
void or_boundp_nh(const char cbc[6], const int n[3], const double bc[6],
                  int nh_d, int nh_p, const double dl[3], const double *dzc,
                  const double *dzf, double *p);
/* bounduvw at one rank (src/bound.f90:17-144, set_bc :227-646, outflow
 * :649-773): u, v, w with halo nh_u (<= 16), cbc/bc the (0:1,3,3) arrays
 * in memory order [side + 2*dir + 6*comp], isoutflow (0:1,3) as ints,
 * dzc/dzf over (1-nh_d : nz+nh_d), nh_d >= nh_u. */
void or_bounduvw(const char cbc[18], const int n[3], const double bc[18],
                 int nh_d, int nh_u, const int isoutflow[6], const double dl[3],
                 const double *dzc, const double *dzf, double *u, double *v,
                 double *w);
/* boundp at one rank with nh_p = 1 (src/bound.f90:146-225). */
void or_boundp(const char cbc[6], const int n[3], const double bc[6], int nh_d,
               const double dl[3], const double *dzc, const double *dzf,
               double *p);
void or_updt_rhs_b(int nx, int ny, int nz, const char c_or_f[3],
                   const char cbc[6], const double *rhsbx, const double *rhsby,
                   const double *rhsbz, double *p);
/* initgrid (src/initgrid.f90:17-97), two-end tanh clustering; arrays are
 * (1-nh_d : n+nh_d). */
void or_initgrid(int n, double gr, double lz, int nh_d, double *dzc,
                 double *dzf, double *zc, double *zf);
/* initgrid with the reference's inivel selection (src/initgrid.f90:32-39):
 * one-end clustering for "hcl"/"hcp", two-end otherwise. */
void or_initgrid_inivel(const char *inivel, int n, double gr, double lz,
                        int nh_d, double *dzc, double *dzf, double *zc,
                        double *zf);
/* The operator the spectral solver inverts: out = L p for the interior, with
 * homogeneous ghost values from the BC pattern (periodic wrap, p0 = p1 for
 * N, p0 = -p1 for D, cell-centred) and the z coefficients a, b, c of
 * tridmatrix.  p is a dense nx*ny*nz block (no halo). */
void or_apply_laplacian(const int n[3], const char cbc[6], const double dli[3],
                        const double *a, const double *b, const double *c,
                        const double *p, double *out);

#ifdef __cplusplus
}
#endif
#endif
