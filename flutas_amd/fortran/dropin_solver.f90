! Drop-in Poisson-solver modules: the reference's initsolver / fftend /
! solver_gpu / solver_cpu signatures, backed by a flutas_amd plan.
!
!   initsolver  src/initsolver.f90:21-44   creates the plan (one GPU = one
!               rank), fills lambdaxy, a, b, c, normfft and rhsb{x,y,z}
!               exactly as the reference does, and keeps the plan handle
!               in arrplan(1,1) (the other three entries are null).
!   fftend      src/fft.f90:159-179        destroys the plan.
!   solver_gpu  src/solver_gpu.f90:31-47   solves in place on the interior
!   solver_cpu  src/solver_cpu.f90:20-34   of p; the transforms, eigenvalues
!               and z coefficients come from the plan, so lambdaxy, a, b,
!               c, normfft and the BC arguments are accepted and unused.
!
! flutas_set_precision(mode, tol), called before initsolver, selects
! FLUTAS_FP64 (default) or FLUTAS_FP32_REFINE with a relative-residual
! target.
module mod_initsolver
  use, intrinsic :: iso_c_binding
  use flutas_capi
  implicit none
  private
  public :: initsolver, flutas_set_precision
  integer, parameter :: rp = c_double
  integer(c_int), save :: precision_mode = FLUTAS_FP64
  real(rp), save :: refine_tol = 1.0e-10_rp
contains
  subroutine flutas_set_precision(mode, tol)
    integer, intent(in) :: mode
    real(rp), intent(in) :: tol
    precision_mode = int(mode, c_int)
    refine_tol = tol
  end subroutine flutas_set_precision
  !
  subroutine initsolver(n,dims,dims_z,dli,nh_d,dzci,dzfi,cbc,bc,c_or_f,lambdaxy,a,b,c,arrplan,normfft, &
                        rhsbx,rhsby,rhsbz)
    integer         , intent(in ), dimension(3)       :: n,dims,dims_z
    real(rp)        , intent(in ), dimension(3)       :: dli
    integer         , intent(in )                     :: nh_d
    real(rp)        , intent(in ), dimension(1-nh_d:) :: dzci,dzfi
    character(len=1), intent(in ), dimension(0:1,3)   :: cbc
    real(rp)        , intent(in ), dimension(0:1,3)   :: bc
    character(len=1), intent(in ), dimension(3)       :: c_or_f
    real(rp)        , intent(out), dimension(n(1),n(2)) :: lambdaxy
    real(rp)        , intent(out), dimension(n(3))    :: a,b,c
    type(c_ptr)     , intent(out), dimension(2,2)     :: arrplan
    real(rp)        , intent(out)                     :: normfft
    real(rp)        , intent(out), dimension(n(2),n(3),0:1) :: rhsbx
    real(rp)        , intent(out), dimension(n(1),n(3),0:1) :: rhsby
    real(rp)        , intent(out), dimension(n(1),n(2),0:1) :: rhsbz
    !
    integer(c_int) :: ng(3)
    character(kind=c_char) :: cbc_c(6), cf_c(3)
    real(c_double) :: bc_c(6)
    type(c_ptr) :: h
    integer :: q, s, d
    if (any(dims /= 1) .or. any(dims_z /= 1)) &
      error stop 'flutas_amd initsolver: one GPU per solve (dims must be 1)'
    ng = int(n, c_int)
    do d = 1, 3
      do s = 0, 1
        q = 2*(d-1) + s + 1
        cbc_c(q) = cbc(s,d)
        bc_c(q) = bc(s,d)
      end do
      cf_c(d) = c_or_f(d)
    end do
    h = c_null_ptr
    call flutas_check(flutas_psolve_create(ng, cbc_c, bc_c, cf_c, dli, int(nh_d, c_int), dzci, dzfi, &
                                           precision_mode, refine_tol, h), 'initsolver')
    call flutas_check(flutas_psolve_coeffs(h, lambdaxy, a, b, c, normfft, rhsbx, rhsby, rhsbz), &
                      'initsolver coefficients')
    arrplan(:,:) = c_null_ptr
    arrplan(1,1) = h
  end subroutine initsolver
end module mod_initsolver
!
module mod_fft
  use, intrinsic :: iso_c_binding
  use flutas_capi
  use flutas_stage, only: stage_release
  implicit none
  private
  public :: fftend
contains
  subroutine fftend(arrplan)
    type(c_ptr), intent(inout), dimension(2,2) :: arrplan
    if (c_associated(arrplan(1,1))) call flutas_check(flutas_psolve_destroy(arrplan(1,1)), 'fftend')
    arrplan(:,:) = c_null_ptr
    call stage_release()
  end subroutine fftend
end module mod_fft
!
module flutas_psolve_call
  use, intrinsic :: iso_c_binding
  use flutas_capi
  use flutas_stage
  implicit none
  private
  public :: solve_in_place
contains
  ! Solves on the interior of the halo-1 field p (c_loc of p(0,0,0)).
  subroutine solve_in_place(h, p, nbytes)
    type(c_ptr), intent(in) :: h, p
    integer(c_size_t), intent(in) :: nbytes
    type(c_ptr) :: dp
    if (.not. c_associated(h)) error stop 'flutas_amd solver: plan not initialised (initsolver)'
    dp = stage_in(4, p, nbytes, .true.)
    call flutas_check(flutas_psolve_solve(h, dp), 'solver')
    call stage_out(p, dp, nbytes)
  end subroutine solve_in_place
end module flutas_psolve_call
!
module mod_solver_gpu
  use, intrinsic :: iso_c_binding
  use flutas_psolve_call
  implicit none
  private
  public :: solver_gpu
  integer, parameter :: rp = c_double
contains
  subroutine solver_gpu(n,dims,arrplan,normfft,lambdaxy,a,b,c,bc,c_or_f,p)
    integer         , intent(in   ), dimension(3)     :: n
    integer         , intent(in   ), dimension(3)     :: dims
    type(c_ptr)     , intent(in   ), dimension(2,2)   :: arrplan
    real(rp)        , intent(in   )                   :: normfft
    real(rp)        , intent(in   ), dimension(*)     :: lambdaxy
    real(rp)        , intent(in   ), dimension(*)     :: a,b,c
    character(len=1), intent(in   ), dimension(0:1,3) :: bc
    character(len=1), intent(in   ), dimension(3)     :: c_or_f
    real(rp)        , intent(inout), dimension(0:,0:,0:), target, contiguous :: p
    if (any(dims /= 1)) error stop 'flutas_amd solver_gpu: one GPU per solve (dims must be 1)'
    if (any(shape(p) /= n + 2)) error stop 'flutas_amd solver_gpu: p must be (0:n+1)^3'
    call solve_in_place(arrplan(1,1), c_loc(p), 8_c_size_t*size(p, kind=c_size_t))
  end subroutine solver_gpu
end module mod_solver_gpu
!
module mod_solver_cpu
  use, intrinsic :: iso_c_binding
  use flutas_psolve_call
  implicit none
  private
  public :: solver_cpu
  integer, parameter :: rp = c_double
contains
  subroutine solver_cpu(n,arrplan,normfft,lambdaxy,a,b,c,bcz,c_or_f,p)
    integer         , intent(in   ), dimension(3)   :: n
    type(c_ptr)     , intent(in   ), dimension(2,2) :: arrplan
    real(rp)        , intent(in   )                 :: normfft
    real(rp)        , intent(in   ), dimension(*)   :: lambdaxy
    real(rp)        , intent(in   ), dimension(*)   :: a,b,c
    character(len=1), intent(in   ), dimension(0:1) :: bcz
    character(len=1), intent(in   ), dimension(3)   :: c_or_f
    real(rp)        , intent(inout), dimension(0:,0:,0:), target, contiguous :: p
    if (any(shape(p) /= n + 2)) error stop 'flutas_amd solver_cpu: p must be (0:n+1)^3'
    call solve_in_place(arrplan(1,1), c_loc(p), 8_c_size_t*size(p, kind=c_size_t))
  end subroutine solver_cpu
end module mod_solver_cpu
