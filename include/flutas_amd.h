/*
 * flutas_amd -- MI355X (gfx950) pressure-projection step for FluTAS.
 *
 * C-ABI of libflutas_amd.so.  Plain pointers and sizes only.  Arrays follow
 * the reference's Fortran conventions (column-major, i fastest) and are
 * passed as a pointer to their FIRST element including the halo:
 *   p, pold, rho : p(0:nx+1, 0:ny+1, 0:nz+1)           (halo 1)
 *   u, v, w      : u(1-nh_u:nx+nh_u, ...)               (halo nh_u)
 *   dzci, dzfi   : dzci(1-nh_d : nz+nh_d)               (halo nh_d)
 * Field pointers ("device") are HIP device or managed allocations; the
 * metric arrays of flutas_psolve_create are host arrays.  BC characters are
 * 'P', 'D', 'N' in the reference's (0:1,3) order: cbc[2*dir + side].
 *
 * Every function returns 0 (FLUTAS_OK) or a status code; the message of the
 * last failure on the calling thread is flutas_last_error().  Field kernels
 * run asynchronously on the library stream (flutas_set_stream; default the
 * null stream) unless they return host values; flutas_synchronize() waits.
 */
#ifndef FLUTAS_AMD_H
#define FLUTAS_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FLUTAS_OK = 0,
  FLUTAS_ENOMEM = 12,
  FLUTAS_EINVAL = 22,
  FLUTAS_EDEVICE = 1001,      /* HIP runtime error, or no gfx950 device   */
  FLUTAS_ENONFINITE = 1002,   /* NaN/Inf met in the right-hand side       */
  FLUTAS_ENOTCONV = 1003,     /* refinement did not reach the tolerance   */
  FLUTAS_EUNSUPPORTED = 1004  /* valid for the reference, not built here  */
};

/* Precision modes of a Poisson plan. */
enum {
  FLUTAS_FP64 = 64,        /* plain FP64 spectral solve                        */
  FLUTAS_FP32_REFINE = 32  /* FP32 transforms + FP64 residual refinement       */
};

typedef struct flutas_psolve flutas_psolve;

const char *flutas_last_error(void);
const char *flutas_version(void);

/* ---- device plumbing (no counterpart in the reference, which uses CUDA
 *      managed memory for its fields: main__two_phase_inc_isot.f90:148-153) */
int flutas_device_alloc(size_t bytes, void **ptr);
int flutas_device_free(void *ptr);
int flutas_memcpy_h2d(void *dst, const void *src, size_t bytes);
int flutas_memcpy_d2h(void *dst, const void *src, size_t bytes);
int flutas_memset_d(void *dst, int value, size_t bytes);
int flutas_set_stream(void *hip_stream);
int flutas_synchronize(void);
/* *is_device = 1 when ptr is a HIP device or managed allocation (usable by
 * the field kernels as is), 0 for plain host memory. */
int flutas_pointer_is_device(const void *ptr, int *is_device);

/* ---- Poisson solver --------------------------------------------------- */

/* initsolver + fftini (src/initsolver.f90:21-44, src/fft.f90:24-32).
 * ng: global grid (one GPU = one rank); dli = 1/(dx,dy,dz);
 * dzci, dzfi: host arrays (1-nh_d : ng[2]+nh_d).  precision: FLUTAS_FP64 or
 * FLUTAS_FP32_REFINE; tol: relative residual target of the refined mode
 * (ignored for FP64).  c_or_f 'c' (cell-centred) or 'f' (face-centred) per
 * direction, with find_fft's kinds (src/fft.f90:233-291), tridmatrix's and
 * bc_rhs's 'f' branches (src/initsolver.f90:210-236, :277-287) and, for an
 * 'f' z direction with a D top face, nz - 1 planes in the z system
 * (src/solver_cpu.f90:71-77).  Unsupported here (EUNSUPPORTED): 'f' NN and
 * DD in x or y (the REDFT00/RODFT00 solves, whose eigenvalues in
 * src/initsolver.f90:164-178 belong to other transforms) and 'f' with
 * FLUTAS_FP32_REFINE (its residual is the cell-centred operator); odd ng
 * (EINVAL, src/sanity.f90:155-158).  ND/DN pairs in x or y use REDFT11 /
 * RODFT11 ('c'; DCT-IV / DST-IV tiles on the four-pass pipeline at the
 * compiled lengths) or RODFT01 forward ('f' DN, five-pass line pipeline). */
int flutas_psolve_create(const int ng[3], const char cbc[6], const double bc[6],
                         const char c_or_f[3], const double dli[3], int nh_d,
                         const double *dzci, const double *dzfi, int precision,
                         double tol, flutas_psolve **out);

/* fftend (src/fft.f90:159-179): releases the plan and its device memory. */
int flutas_psolve_destroy(flutas_psolve *h);

/* The initsolver outputs (src/initsolver.f90:34-44), host arrays:
 * lambdaxy(nx,ny) in the halfcomplex/natural order of the CPU path,
 * a,b,c(nz), normfft, rhsbx(ny,nz,0:1), rhsby(nx,nz,0:1), rhsbz(nx,ny,0:1).
 * Any pointer may be NULL. */
int flutas_psolve_coeffs(const flutas_psolve *h, double *lambdaxy, double *a,
                         double *b, double *c, double *normfft, double *rhsbx,
                         double *rhsby, double *rhsbz);

/* Device copies of rhsbx, rhsby, rhsbz owned by the plan (for
 * flutas_updt_rhs_b).  Pointers stay valid until destroy. */
int flutas_psolve_rhsb_device(const flutas_psolve *h, const double **rhsbx,
                              const double **rhsby, const double **rhsbz);

/* solver_gpu, one-GPU branch (src/solver_gpu.f90:31-47, :107-229): solves
 * the Poisson equation in place on the interior of the device array p
 * (0:nx+1,0:ny+1,0:nz+1); halo values are neither read nor written.
 * Returns when the result is in p.  A right-hand side holding NaN or Inf
 * returns FLUTAS_ENONFINITE (p then holds non-finite values). */
int flutas_psolve_solve(flutas_psolve *h, double *p);

/* Same, enqueued on the library stream; returns immediately.  The refined
 * mode still synchronises once per refinement iteration (residual norm)
 * and reports a non-finite right-hand side; the FP64 mode does not check
 * here (the synchronous entry points do). */
int flutas_psolve_solve_async(flutas_psolve *h, double *p);

/* Out-of-place variant (no counterpart in the reference, whose solver works
 * in place): reads the right-hand side from the interior of f and writes
 * the solution to the interior of p (both halo-1 device arrays); f is left
 * unchanged.  Same traffic as the in-place solve. */
int flutas_psolve_solve_into(flutas_psolve *h, const double *f, double *p);

/* Diagnostics of the last solve: refinement iterations and final relative
 * residual max|f - L p| / max|f| (FP64 plans: iters = 0, relres = -1). */
int flutas_psolve_stats(const flutas_psolve *h, int *iters, double *relres);

/* Relative residual max|f - L p| / max|f| of a given pair (device arrays
 * with halo 1), with the operator the plan inverts.  f is not modified. */
int flutas_psolve_residual(flutas_psolve *h, const double *f, const double *p,
                           double *relres);

/* Per-stage timing of the solve pipeline (no counterpart in the reference,
 * whose profiler times the whole SOLVER region, src/profiler.f90:83-205):
 * runs `reps` out-of-place solves f -> p and returns the mean duration in
 * ms of each stage, measured with HIP events on the library stream.
 * *nstages receives the number of stages (<= max_stages);
 * flutas_psolve_stage_name names them. */
int flutas_psolve_profile(flutas_psolve *h, const double *f, double *p,
                          int reps, int max_stages, double *ms, int *nstages);
const char *flutas_psolve_stage_name(const flutas_psolve *h, int stage);
/* One stage of the four-pass FP64 pipeline (0 x forward, 1 y forward and
 * z elimination, 2 z substitution and y backward, 3 x backward) launched
 * `reps` (<= 16) times back to back after one solve f -> p: *ms = the
 * average duration of one launch, from two HIP events on the library
 * stream.  The work field is left transformed; p holds the solution. */
int flutas_psolve_time_stage(flutas_psolve *h, const double *f, double *p, int stage,
                             int reps, double *ms);

/* z-chunked slab passes (no counterpart in the reference, whose z solve
 * walks whole columns, src/solver_gpu.f90:483-529 and
 * src/solver_cpu.f90:187-223): the planes of each x-slab are eliminated in
 * n chunks (1..8) by separate workgroups and joined by a reduced system of
 * two unknowns per chunk and column, which gives plans with few x-slabs
 * more workgroups than the GPU has CUs.  n = 0 restores the plan's default
 * (the count, with chunks of 32 planes or more, that best fills the GPU's
 * workgroup slots over the two slab passes, from the CU count and the
 * kernels' occupancy); n = 1 walks whole slabs.  The count may come out lower than n for
 * a short z.  Results equal the whole-slab walk's to rounding; the pivots
 * stay the reference's bit for bit.  Changing the count frees the tables a
 * recorded graph of this plan's solve refers to: record again after it.
 * flutas_psolve_zchunks reads the count in use. */
int flutas_psolve_set_zchunks(flutas_psolve *h, int n);
int flutas_psolve_zchunks(const flutas_psolve *h, int *n);

/* Autotuner of the single solve (no counterpart in the reference): times
 * `reps` (1..100) solves f -> p at each chunk count 1..8 that the grid
 * allows, keeps the fastest and returns it in *chosen; ms (may be null)
 * receives 8 values, the mean milliseconds per solve for 1..8 chunks (-1
 * where the count was not possible).  p holds a solution afterwards. */
int flutas_psolve_tune(flutas_psolve *h, const double *f, double *p, int reps, int *chosen,
                       double *ms);

/* ---- FP32 fields: FluTAS built with _SINGLE_PRECISION (rp = FP32,
 *      src/types.f90:6-12) ------------------------------------------------
 * The same entry points for FP32 fields and FP32 scalars, computing in
 * FP32 as the reference's single-precision build does (chkdiv sums in
 * FP64).  The solve runs the FP32 chain (FP32 transforms, FP64 z
 * elimination) once on a plan of either precision, in place on the
 * interior of p(0:nx+1,0:ny+1,0:nz+1); a non-finite right-hand side
 * returns FLUTAS_ENONFINITE on the four-pass pipeline. */
int flutas_psolve_solve_f32(flutas_psolve *h, float *p);
int flutas_fft_guru_f32(int kind, int n, long long stride, int howmany1,
                        long long dist1, int howmany2, long long dist2,
                        float *data);
int flutas_fillps_f32(int nx, int ny, int nz, int nh_d, int nh_u, float dxi,
                      float dyi, float dzi, const float *dzfi, float dti,
                      float rho0, const float *u, const float *v,
                      const float *w, float *p);
int flutas_pres_tw_src_f32(int nx, int ny, int nz, float dxi, float dyi,
                           float dzi, int nh_d, int nh_u, const float *dzci,
                           float rho0i, float f_t12, float f_t12_o,
                           const float *p, const float *pold, const float *rho,
                           float *u, float *v, float *w, int const_coeff);
int flutas_pres_sp_src_f32(int nx, int ny, int nz, float f_t12, float dxi,
                           float dyi, float dzi, int nh_d, int nh_u,
                           const float *dzci, float rho0i, const float *pold,
                           float *u, float *v, float *w);
int flutas_correc_f32(int nx, int ny, int nz, int nh_d, int nh_u, float dxi,
                      float dyi, float dzi, const float *dzci, float dt,
                      float rho0, const float *p, float *u, float *v,
                      float *w, const float *rho, int const_coeff);
int flutas_chkdiv_f32(int nx, int ny, int nz, float dxi, float dyi, float dzi,
                      int nh_d, int nh_u, const float *dzfi, const float *u,
                      const float *v, const float *w, float *divtot,
                      float *divmax);
int flutas_boundp_f32(const char cbc[6], const int n[3], const float bc[6],
                      const float dl[3], float dzc_0, float dzc_n, float *p);
int flutas_boundp_halo_f32(const char cbc[6], const int n[3], const float bc[6],
                           int nh_d, int nh_p, const float dl[3],
                           const float *dzc, const float *dzf, float *p);
int flutas_bounduvw_f32(const char cbc[18], const int n[3], const float bc[18],
                        int nh_d, int nh_u, const int isoutflow[6],
                        const float dl[3], const float *dzc, const float *dzf,
                        float *u, float *v, float *w);
int flutas_updt_rhs_b_f32(int nx, int ny, int nz, const char c_or_f[3],
                          const char cbc[6], const float *rhsbx,
                          const float *rhsby, const float *rhsbz, float *p);

/* ---- graphs of library calls (no counterpart in the reference, whose
 *      steps launch kernel by kernel) --------------------------------------
 * Between flutas_graph_begin and flutas_graph_end the asynchronous library
 * calls are recorded instead of run: FP64 solves through
 * flutas_psolve_solve_async / the step's field kernels (fillps, fillps_rhs,
 * updt_rhs_b, pres_*_src, correc, correc_update, copy/add_interior), boundp
 * and bounduvw without outflow faces.  Calls that synchronise or allocate
 * (synchronous solves, the refined mode, chkdiv, bounduvw with outflow)
 * make flutas_graph_end fail.  flutas_graph_launch replays the recorded
 * sequence on the library stream (asynchronous).  A non-finite right-hand
 * side met in a replayed solve is not reported. */
typedef struct flutas_graph flutas_graph;
int flutas_graph_begin(void);
int flutas_graph_end(flutas_graph **g);
int flutas_graph_launch(flutas_graph *g);
int flutas_graph_destroy(flutas_graph *g);

/* fft (src/fft.f90:181-193): executes one r2r kind in place on `howmany`
 * device lines of length n (element stride `stride`, line distance
 * `dist`), FP64, unnormalised FFTW definitions.  kind uses FFTW's
 * numbering (src/fftw.f90:38-61); built: R2HC(0), HC2R(1), REDFT00(3),
 * REDFT01(4), REDFT10(5), REDFT11(6), RODFT00(7), RODFT01(8), RODFT10(9),
 * RODFT11(10).  n even, except for REDFT00 (n >= 2) and RODFT00 (any n),
 * which run as the real DFT of their symmetric extension (2(n-1) or
 * 2(n+1) points); a prime factor of n/2 (or of the extension's half
 * length) above 1021 returns EUNSUPPORTED. */
int flutas_fft_lines(int kind, int n, int howmany, long long stride,
                     long long dist, double *data);
/* The same transform over a two-level batch, in the geometry of the guru
 * plans of fftini (src/fft.f90:75-86 for x, :113-124 for y): howmany1
 * lines at distance dist1, repeated howmany2 (<= 65535) times at distance
 * dist2; element stride `stride`.  x of an (nx, ny, nz) array: n = nx,
 * stride 1, (ny, nx), (nz, nx*ny); y: n = ny, stride nx, (nx, 1),
 * (nz, nx*ny).  Synchronous, like flutas_fft_lines. */
int flutas_fft_guru(int kind, int n, long long stride, int howmany1,
                    long long dist1, int howmany2, long long dist2,
                    double *data);

/* ---- batched solves (no counterpart in the reference, which solves one
 *      problem per call; for many independent small solves) ------------- */
typedef struct flutas_batch flutas_batch;
/* nprob independent problems, problem q described by ng[3q..3q+2],
 * cbc[6q..6q+5], dli[3q..3q+2] and the host metric arrays dzci[q],
 * dzfi[q] (1-nh_d : nz+nh_d), as for flutas_psolve_create with zero BC
 * values and c_or_f "ccc".  Every nx and ny in {32, 48, 64, 96, 128, 192,
 * 256} and nz >= 3, else EUNSUPPORTED.  precision FLUTAS_FP64 or
 * FLUTAS_FP32_REFINE (FP32 transforms, FP64 residuals and z elimination,
 * each problem refined to a relative residual <= tol). */
int flutas_batch_create_ex(int nprob, const int *ng, const char *cbc,
                           const double *dli, int nh_d,
                           const double *const *dzci,
                           const double *const *dzfi, int precision,
                           double tol, flutas_batch **out);
/* flutas_batch_create_ex with FLUTAS_FP64. */
int flutas_batch_create(int nprob, const int *ng, const char *cbc,
                        const double *dli, int nh_d,
                        const double *const *dzci, const double *const *dzfi,
                        flutas_batch **out);
/* The halo-1 device fields of every problem: f[q] is read (interior), the
 * interior of p[q] receives the solution (f[q] != p[q]).  May be called
 * again with other fields; the captured graphs are kept. */
int flutas_batch_set_fields(flutas_batch *b, const double *const *f,
                            double *const *p);
/* Solves every problem.  Each pass of the pipeline (x transform, y
 * transform + forward elimination, back substitution + inverse y
 * transform, inverse x transform; in the refined mode also the FP64
 * residual, the correction solve and the residual check) is one persistent
 * launch whose workgroups walk a list of (problem, tile or slab) items over
 * every problem and line length; the launches of a solve are captured into
 * a hipGraph on the first call and replayed.  Problems still above tol get
 * further correction rounds (one graph replay each, at most 7).  Returns
 * when every solution is in place; FLUTAS_ENOTCONV if a problem stays above
 * tol, FLUTAS_ENONFINITE for a NaN or Inf. */
int flutas_batch_solve(flutas_batch *b);
/* Per problem: correction rounds of the last solve (0 in FP64 mode) and
 * its relative residual max|f - L p| / max|f| (-1 in FP64 mode). */
int flutas_batch_stats(const flutas_batch *b, int *iters, double *relres);
/* Autotuner (the fields must be set): coordinate descent over the launch
 * shape of the passes (0: one launch per pass and line length, one
 * workgroup per item, run side by side; 1: one persistent launch per pass
 * whose workgroups walk every item, at its best workgroups per CU), the
 * x-tile variant of each x length and the plane-group variant
 * (z-elimination block) of each y length (0 full, 1 half, 2 quarter of the
 * LDS-sized default) and the planes of an x-tile work item (1, 2, 4, 8),
 * each choice timed over `reps` replays of the whole solve.
 * flutas_batch_tuning reports the variants per problem,
 * flutas_batch_launch_shape the shape and the planes per x-tile item,
 * flutas_batch_launch_info the persistent launch (workgroups, items over
 * all passes, workgroups per CU, dynamic LDS bytes). */
int flutas_batch_tune(flutas_batch *b, int reps);
int flutas_batch_tuning(const flutas_batch *b, int *tile_variant,
                        int *group_variant);
int flutas_batch_launch_info(const flutas_batch *b, int *workgroups,
                             int *items, int *per_cu, size_t *lds_bytes);
int flutas_batch_launch_shape(const flutas_batch *b, int *shape,
                              int *x_item_planes);
int flutas_batch_destroy(flutas_batch *b);

/* ---- restart files in the reference's formats ------------------------- */
/* load (src/load.f90:21-90): io 'r' reads, 'w' writes the file `filename`
 * holding the global n[0] x n[1] x n[2] FP64 field in Fortran order (no
 * header, no halo; the reference's MPI-IO layout at one rank) from / into
 * the interior of the field fld with halo `halo` (a device field, or plain
 * host memory, then with no device work).  A file of the wrong size is
 * rejected (EINVAL), as the reference aborts. */
int flutas_load(char io, const char *filename, const int n[3], int halo, double *fld);
/* load_ns (src/load.f90:132-204), the CaNS layout: u, v, w (halo
 * halo_u), p (halo 1), then the two doubles time, istep. */
int flutas_load_ns(char io, const char *filename, const int n[3], int halo_u,
                   double *u, double *v, double *w, double *p, double *time,
                   double *istep);
/* The same with one halo width per field (u, v, w, p): 0 for the dense
 * explicit-shape arrays load_ns itself receives (src/load.f90:141-142). */
int flutas_load_ns_halos(char io, const char *filename, const int n[3],
                         const int halo[4], double *u, double *v, double *w,
                         double *p, double *time, double *istep);
/* load_scalar (src/load.f90:92-130): the text file "time dto istep",
 * written with the format (2E15.7, 1I9.8).  Host values only. */
int flutas_load_scalar(char io, const char *filename, double *time, int *istep,
                       double *dto);

/* HIP events on the library stream, for host-side timing of any sequence. */
int flutas_event_create(void **ev);
int flutas_event_record(void *ev);
int flutas_event_elapsed_ms(void *start, void *end, float *ms);
int flutas_event_destroy(void *ev);

/* ---- field kernels of the projection step ------------------------------ */

/* fillps (src/fillps.f90:16-26), _CONSTANT_COEFFS_POISSON: p interior =
 * rho0 * div(u,v,w) * dti.  dzfi is a device array. */
int flutas_fillps(int nx, int ny, int nz, int nh_d, int nh_u, double dxi,
                  double dyi, double dzi, const double *dzfi, double dti,
                  double rho0, const double *u, const double *v,
                  const double *w, double *p);

/* pres_tw_src (src/source.f90:247-259), two-phase pressure splitting;
 * const_coeff = 1 is the _CONSTANT_COEFFS_POISSON branch. */
int flutas_pres_tw_src(int nx, int ny, int nz, double dxi, double dyi,
                       double dzi, int nh_d, int nh_u, const double *dzci,
                       double rho0i, double f_t12, double f_t12_o,
                       const double *p, const double *pold, const double *rho,
                       double *u, double *v, double *w, int const_coeff);

/* pres_sp_src (src/source.f90:311-323), single-phase. */
int flutas_pres_sp_src(int nx, int ny, int nz, double f_t12, double dxi,
                       double dyi, double dzi, int nh_d, int nh_u,
                       const double *dzci, double rho0i, const double *pold,
                       double *u, double *v, double *w);

/* boundp for a field with any halo nh_p (1..8), as the reference calls it
 * for temperature and VOF fields (src/bound.f90:146-225): p is the device
 * array (1-nh_p : n+nh_p)^3; dzc, dzf: HOST arrays (1-nh_d : nz+nh_d),
 * nh_d >= nh_p; the z faces use dzc(-q) and dzc(nz+q) for layer q. */
int flutas_boundp_halo(const char cbc[6], const int n[3], const double bc[6],
                       int nh_d, int nh_p, const double dl[3],
                       const double *dzc, const double *dzf, double *p);

/* bounduvw (src/bound.f90:17-144; set_bc :227-646; outflow :649-773) at
 * one rank: the halos of the device fields u, v, w (halo nh_u, 1..8).
 * cbc, bc: the reference's cbc(0:1,3,3), bc(0:1,3,3) in memory order,
 * [side + 2*dir + 6*comp]; isoutflow: the (0:1,3) flags [side + 2*dir];
 * dl = (dx, dy, dz); dzc, dzf: HOST arrays (1-nh_d : nz+nh_d), nh_d >=
 * nh_u.  Periodic y and z halos come first (updthalo's self-exchange),
 * then the x, y, z faces in the reference's order (u face-normal in x, v in
 * y, w in z, the others cell-centred), then the outflow faces x, y, z.
 * Asynchronous unless an x or y outflow face is set (then it returns after
 * the work is done). */
int flutas_bounduvw(const char cbc[18], const int n[3], const double bc[18],
                    int nh_d, int nh_u, const int isoutflow[6],
                    const double dl[3], const double *dzc, const double *dzf,
                    double *u, double *v, double *w);

/* updt_rhs_b (src/bound.f90:829-944) at one rank; rhsb* device arrays. */
int flutas_updt_rhs_b(int nx, int ny, int nz, const char c_or_f[3],
                      const char cbc[6], const double *rhsbx,
                      const double *rhsby, const double *rhsbz, double *p);

/* boundp (src/bound.f90:146-225) at one rank, nh_p = 1.  dl = (dx,dy,dz);
 * dzc_0 = dzc(0), dzc_n = dzc(nz) (the only z metrics it reads). */
int flutas_boundp(const char cbc[6], const int n[3], const double bc[6],
                  const double dl[3], double dzc_0, double dzc_n, double *p);

/* correc (src/correc.f90:16-29); const_coeff = 1 divides by rho0, else by
 * the arithmetic face density of rho (rho may be NULL when const_coeff). */
int flutas_correc(int nx, int ny, int nz, int nh_d, int nh_u, double dxi,
                  double dyi, double dzi, const double *dzci, double dt,
                  double rho0, const double *p, double *u, double *v,
                  double *w, const double *rho, int const_coeff);

/* chkdiv (src/chkdiv.f90:18-31): returns the total and the maximum absolute
 * divergence (synchronous). */
int flutas_chkdiv(int nx, int ny, int nz, double dxi, double dyi, double dzi,
                  int nh_d, int nh_u, const double *dzfi, const double *u,
                  const double *v, const double *w, double *divtot,
                  double *divmax);

/* flutas_chkdiv without any device allocation.  No counterpart in the
 * reference: it serves callers that cannot allocate (code under HIP graph
 * capture, tight time loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results (bit for bit) as flutas_chkdiv; the block
 * partials live in a fixed per-GPU device array of 2*2048+2 doubles
 * (32 KiB) that the library holds until the process ends.  The kernel
 * strides over the grid, so every field size fits those 2048 partials;
 * a partial count beyond them is rejected with FLUTAS_EINVAL.  Calls that
 * overlap on one GPU share that array and race: serialise them. */
int flutas_chkdiv_noalloc(int nx, int ny, int nz, double dxi, double dyi,
                          double dzi, int nh_d, int nh_u, const double *dzfi,
                          const double *u, const double *v, const double *w,
                          double *divtot, double *divmax);

/* Fused prologue of the projection step: fillps then updt_rhs_b
 * (src/fillps.f90:16-26 with src/bound.f90:829-944, called back to back in
 * main__single_phase.f90:707-709) in one pass; bit-identical to the two
 * calls.  rhsb* are device arrays (flutas_psolve_rhsb_device). */
int flutas_fillps_rhs(int nx, int ny, int nz, int nh_d, int nh_u, double dxi,
                      double dyi, double dzi, const double *dzfi, double dti,
                      double rho0, const double *u, const double *v,
                      const double *w, const char c_or_f[3], const char cbc[6],
                      const double *rhsbx, const double *rhsby,
                      const double *rhsbz, double *p);

/* Fused epilogue of the projection step: correc (src/correc.f90:16-29) with
 * the increment phi, then pnew = pold + phi on the interior (the correc call
 * and the update loop of main__single_phase.f90:726, :734-740), in one pass; bit-identical to
 * flutas_correc followed by flutas_add_interior.  The new pressure goes to
 * pnew, which may be pold but not phi, so phi stays readable for the
 * neighbours of every point; a caller that swaps its pressure buffers
 * afterwards (pnew becomes the next step's pold) also saves the pold = p
 * copy.  phi's halos must be set (flutas_boundp), as for correc; pnew's
 * halos are not written (flutas_boundp on pnew, as after the reference's
 * update). */
int flutas_correc_update(int nx, int ny, int nz, int nh_d, int nh_u, double dxi,
                         double dyi, double dzi, const double *dzci, double dt,
                         double rho0, const double *phi, const double *pold,
                         double *pnew, double *u, double *v, double *w,
                         const double *rho, int const_coeff);

/* flutas_correc_update with the two boundp calls around it folded in
 * (main__single_phase.f90:727-747: boundp(phi) before correc, boundp(p)
 * after the update; src/bound.f90:146-225 with nh_p = 1): the halo values
 * of phi that correc reads are computed from phi's interior with set_bc's
 * rules (phi's halo is neither read nor written), and every halo point of
 * pnew is set as boundp sets it.  cbc, bc, dl, dzc_0, dzc_n as for
 * flutas_boundp.  Bit for bit boundp(phi), flutas_correc_update,
 * boundp(pnew), in one pass. */
int flutas_correc_update_bc(int nx, int ny, int nz, int nh_d, int nh_u, double dxi,
                            double dyi, double dzi, const double *dzci, double dt,
                            double rho0, const double *phi, const double *pold,
                            double *pnew, double *u, double *v, double *w,
                            const double *rho, int const_coeff, const char cbc[6],
                            const double bc[6], const double dl[3], double dzc_0,
                            double dzc_n);

/* pold = p on the interior (main__two_phase_inc_isot.f90:656-673). */
int flutas_copy_interior(int nx, int ny, int nz, const double *src, double *dst);

/* p = pold + p on the interior (main__two_phase_inc_isot.f90:698-714). */
int flutas_add_interior(int nx, int ny, int nz, const double *pold, double *p);

/* initgrid with two-end tanh clustering (src/initgrid.f90:17-97,
 * :102-118); host arrays (1-nh_d : n+nh_d). */
int flutas_initgrid(int n, double gr, double lz, int nh_d, double *dzc,
                    double *dzf, double *zc, double *zf);
/* initgrid with its inivel argument (src/initgrid.f90:17-39): "hcl" and
 * "hcp" select gridpoint_cluster_one_end (:120-136), every other value
 * gridpoint_cluster_two_end.  inivel: the first three characters. */
int flutas_initgrid_inivel(const char *inivel, int n, double gr, double lz,
                           int nh_d, double *dzc, double *dzf, double *zc,
                           double *zf);

#ifdef __cplusplus
}
#endif
#endif
