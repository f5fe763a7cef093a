! Drop-in Poisson-solver modules: the reference's initsolver / fftend /
! solver_gpu / solver_cpu signatures, backed by a flutas_amd plan.
!
!   initsolver  src/initsolver.f90:21-44   creates the plan (one GPU = one
!               rank), fills lambdaxy, a, b, c, normfft and rhsb{x,y,z}
!               exactly as the reference does, and keeps the plan handle
!               in arrplan(1,1) (the other three entries are null).
!   fftini      src/fft.f90:24-157         the x and y r2r plans and normfft
!   fft         src/fft.f90:181-193        (find_fft's table, src/fft.f90:233-291)
!               on their own, for callers of the transforms; executed on
!               the GPU, every kind of the table included.
!   fftend      src/fft.f90:159-179        destroys either kind of arrplan.
!   solver_gpu  src/solver_gpu.f90:31-47   solves in place on the interior
!   solver_cpu  src/solver_cpu.f90:20-34   of p; the transforms, eigenvalues
!               and z coefficients come from the plan, so lambdaxy, a, b,
!               c, normfft and the BC arguments are accepted and unused.
!
! flutas_set_precision(mode, tol), called before initsolver, selects
! FLUTAS_FP64 (default) or FLUTAS_FP32_REFINE with a relative-residual
! target.  Reals are real(rp) (flutas_kinds): in a _SINGLE_PRECISION build
! the metric arrays go to the plan as FP64 copies, its coefficients come
! back rounded to FP32, and the solves and transforms run the library's
! FP32 entry points (FP32 transforms, FP64 z elimination).
module mod_initsolver
  use, intrinsic :: iso_c_binding
  use flutas_capi
  use flutas_kinds, only: rp
  implicit none
  private
  public :: initsolver, flutas_set_precision
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
    real(c_double) :: bc_c(6), dli_c(3), nf_c
    real(c_double), allocatable :: zci_c(:), zfi_c(:), lam_c(:,:), a_c(:), b_c(:), c_c(:), &
                                   rx_c(:,:,:), ry_c(:,:,:), rz_c(:,:,:)
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
    ! the plan takes and returns FP64 values (exact copies when rp is FP64)
    dli_c = real(dli, c_double)
    allocate(zci_c(size(dzci)), zfi_c(size(dzfi)))
    zci_c = real(dzci, c_double)
    zfi_c = real(dzfi, c_double)
    allocate(lam_c(n(1),n(2)), a_c(n(3)), b_c(n(3)), c_c(n(3)), rx_c(n(2),n(3),0:1), &
             ry_c(n(1),n(3),0:1), rz_c(n(1),n(2),0:1))
    h = c_null_ptr
    call flutas_check(flutas_psolve_create(ng, cbc_c, bc_c, cf_c, dli_c, int(nh_d, c_int), zci_c, zfi_c, &
                                           precision_mode, real(refine_tol, c_double), h), 'initsolver')
    call flutas_check(flutas_psolve_coeffs(h, lam_c, a_c, b_c, c_c, nf_c, rx_c, ry_c, rz_c), &
                      'initsolver coefficients')
    lambdaxy = real(lam_c, rp)
    a = real(a_c, rp)
    b = real(b_c, rp)
    c = real(c_c, rp)
    normfft = real(nf_c, rp)
    rhsbx = real(rx_c, rp)
    rhsby = real(ry_c, rp)
    rhsbz = real(rz_c, rp)
    arrplan(:,:) = c_null_ptr
    arrplan(1,1) = h
  end subroutine initsolver
end module mod_initsolver
!
module mod_fft
  use, intrinsic :: iso_c_binding
  use flutas_capi
  use flutas_stage, only: stage_in, stage_out, stage_release
  use flutas_kinds, only: rp, rp_bytes
  implicit none
  private
  public :: fftini, fftend, fft
  ! FFTW's r2r kind numbers (src/fftw.f90:38-61)
  integer(c_int), parameter :: R2HC = 0, HC2R = 1, REDFT00 = 3, REDFT01 = 4, REDFT10 = 5, &
                               REDFT11 = 6, RODFT00 = 7, RODFT01 = 8, RODFT10 = 9, RODFT11 = 10
  ! One r2r plan of fftini: lines of length n, element stride `stride`,
  ! howmany1 lines at distance dist1 repeated howmany2 times at distance
  ! dist2, over an array of `elems` elements.
  type :: line_plan
    integer(c_int) :: kind, n, howmany1, howmany2
    integer(c_long_long) :: stride, dist1, dist2
    integer(c_size_t) :: elems
  end type line_plan
  ! the plans fftini made and fftend has not yet freed
  type(c_ptr), allocatable, save :: issued(:)
contains
  ! Transform kinds and normalisation of one direction (find_fft,
  ! src/fft.f90:233-291).
  subroutine find_fft(bc, c_or_f, kind_fwd, kind_bwd, norm)
    character(len=1), intent(in ), dimension(0:1) :: bc
    character(len=1), intent(in )                 :: c_or_f
    integer(c_int)  , intent(out)                 :: kind_fwd, kind_bwd
    real(rp)        , intent(out), dimension(2)   :: norm
    character(len=2) :: pair
    pair = bc(0)//bc(1)
    if (c_or_f == 'c') then
      select case (pair)
      case ('PP'); kind_fwd = R2HC;    kind_bwd = HC2R;    norm = [1.0_rp, 0.0_rp]
      case ('NN'); kind_fwd = REDFT10; kind_bwd = REDFT01; norm = [2.0_rp, 0.0_rp]
      case ('DD'); kind_fwd = RODFT10; kind_bwd = RODFT01; norm = [2.0_rp, 0.0_rp]
      case ('ND'); kind_fwd = REDFT11; kind_bwd = REDFT11; norm = [2.0_rp, 0.0_rp]
      case ('DN'); kind_fwd = RODFT11; kind_bwd = RODFT11; norm = [2.0_rp, 0.0_rp]
      case default; error stop 'flutas_amd fftini: unknown boundary-condition pair'
      end select
    else if (c_or_f == 'f') then
      select case (pair)
      case ('PP'); kind_fwd = R2HC;    kind_bwd = HC2R;    norm = [1.0_rp, 0.0_rp]
      case ('ND'); kind_fwd = REDFT10; kind_bwd = REDFT01; norm = [2.0_rp, 0.0_rp]
      case ('DN'); kind_fwd = RODFT01; kind_bwd = RODFT10; norm = [2.0_rp, 0.0_rp]
      case ('NN'); kind_fwd = REDFT00; kind_bwd = REDFT00; norm = [2.0_rp, -1.0_rp]
      case ('DD'); kind_fwd = RODFT00; kind_bwd = RODFT00; norm = [2.0_rp, 1.0_rp]
      case default; error stop 'flutas_amd fftini: unknown boundary-condition pair'
      end select
    else
      error stop 'flutas_amd fftini: c_or_f must be c or f'
    end if
  end subroutine find_fft
  !
  function new_plan(kind, n, stride, howmany1, dist1, howmany2, dist2, elems) result(cp)
    integer(c_int), intent(in) :: kind
    integer, intent(in) :: n, stride, howmany1, dist1, howmany2, dist2
    integer(c_size_t), intent(in) :: elems
    type(c_ptr) :: cp
    type(line_plan), pointer :: lp
    type(c_ptr), allocatable :: grown(:)
    allocate(lp)
    lp = line_plan(kind, int(n, c_int), int(howmany1, c_int), int(howmany2, c_int), &
                   int(stride, c_long_long), int(dist1, c_long_long), int(dist2, c_long_long), elems)
    cp = c_loc(lp)
    if (.not. allocated(issued)) allocate(issued(0))
    allocate(grown(size(issued) + 1))
    grown(1:size(issued)) = issued
    grown(size(grown)) = cp
    call move_alloc(grown, issued)
  end function new_plan
  !
  ! index in `issued` of a plan made by fftini, 0 if it is none
  integer function plan_index(cp)
    type(c_ptr), intent(in) :: cp
    integer :: q
    plan_index = 0
    if (.not. allocated(issued)) return
    do q = 1, size(issued)
      if (c_associated(issued(q), cp)) then
        plan_index = q
        return
      end if
    end do
  end function plan_index
  !
  ! fftini (src/fft.f90:24-157): the forward and backward r2r plans of x
  ! (arrplan(:,1), over n_x arrays) and y (arrplan(:,2), over n_y arrays) and
  ! normfft = 1 / prod(norm(1) * (n + norm(2) - i)).  A face-centred DD
  ! direction transforms n - 1 points (i = 1): its last face is the boundary.
  subroutine fftini(n_x,n_y,bcxy,c_or_f,arrplan,normfft)
    integer         , intent(in ), dimension(3)     :: n_x,n_y
    character(len=1), intent(in ), dimension(0:1,2) :: bcxy
    character(len=1), intent(in ), dimension(2)     :: c_or_f
    type(c_ptr)     , intent(out), dimension(2,2)   :: arrplan
    real(rp)        , intent(out)                   :: normfft
    integer(c_int) :: kf, kb
    real(rp) :: norm(2)
    integer(c_size_t) :: ex, ey
    integer :: ix, iy
    ex = int(n_x(1), c_size_t)*n_x(2)*n_x(3)
    ey = int(n_y(1), c_size_t)*n_y(2)*n_y(3)
    normfft = 1.0_rp
    ix = 0
    if (bcxy(0,1)//bcxy(1,1) == 'DD' .and. c_or_f(1) == 'f') ix = 1
    call find_fft(bcxy(:,1), c_or_f(1), kf, kb, norm)
    arrplan(1,1) = new_plan(kf, n_x(1) - ix, 1, n_x(2), n_x(1), n_x(3), n_x(1)*n_x(2), ex)
    arrplan(2,1) = new_plan(kb, n_x(1) - ix, 1, n_x(2), n_x(1), n_x(3), n_x(1)*n_x(2), ex)
    normfft = normfft*norm(1)*(n_x(1) + norm(2) - ix)
    iy = 0
    if (bcxy(0,2)//bcxy(1,2) == 'DD' .and. c_or_f(2) == 'f') iy = 1
    call find_fft(bcxy(:,2), c_or_f(2), kf, kb, norm)
    arrplan(1,2) = new_plan(kf, n_y(2) - iy, n_y(1), n_y(1), 1, n_y(3), n_y(1)*n_y(2), ey)
    arrplan(2,2) = new_plan(kb, n_y(2) - iy, n_y(1), n_y(1), 1, n_y(3), n_y(1)*n_y(2), ey)
    normfft = normfft*norm(1)*(n_y(2) + norm(2) - iy)
    normfft = normfft**(-1)
  end subroutine fftini
  !
  ! fft (src/fft.f90:181-193): executes a plan of fftini in place on arr,
  ! unnormalised; host arrays are staged through a device buffer.
  subroutine fft(plan,arr)
    type(c_ptr), intent(in   )                                       :: plan
    real(rp)   , intent(inout), dimension(:,:,:), target, contiguous :: arr
    type(line_plan), pointer :: lp
    type(c_ptr) :: d
    integer(c_size_t) :: bytes
    if (plan_index(plan) == 0) error stop 'flutas_amd fft: plan not made by fftini'
    call c_f_pointer(plan, lp)
    if (size(arr, kind=c_size_t) /= lp%elems) error stop 'flutas_amd fft: array size differs from the plan'
    bytes = rp_bytes*lp%elems
    d = stage_in(6, c_loc(arr), bytes, .true.)
#if defined(_SINGLE_PRECISION)
    call flutas_check(flutas_fft_guru_f32(lp%kind, lp%n, lp%stride, lp%howmany1, lp%dist1, &
                                          lp%howmany2, lp%dist2, d), 'fft')
#else
    call flutas_check(flutas_fft_guru(lp%kind, lp%n, lp%stride, lp%howmany1, lp%dist1, &
                                      lp%howmany2, lp%dist2, d), 'fft')
#endif
    call stage_out(c_loc(arr), d, bytes)
  end subroutine fft
  !
  ! fftend (src/fft.f90:159-179): frees the plans of fftini, or the solver
  ! plan initsolver keeps in arrplan(1,1).
  subroutine fftend(arrplan)
    type(c_ptr), intent(inout), dimension(2,2) :: arrplan
    type(line_plan), pointer :: lp
    type(c_ptr), allocatable :: kept(:)
    integer :: i, j, q
    if (c_associated(arrplan(1,1)) .and. plan_index(arrplan(1,1)) == 0) then
      call flutas_check(flutas_psolve_destroy(arrplan(1,1)), 'fftend')
    else
      do j = 1, 2
        do i = 1, 2
          q = plan_index(arrplan(i,j))
          if (q == 0) cycle
          call c_f_pointer(arrplan(i,j), lp)
          deallocate(lp)
          kept = [issued(1:q-1), issued(q+1:)]
          call move_alloc(kept, issued)
        end do
      end do
    end if
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
#if defined(_SINGLE_PRECISION)
    call flutas_check(flutas_psolve_solve_f32(h, dp), 'solver')
#else
    call flutas_check(flutas_psolve_solve(h, dp), 'solver')
#endif
    call stage_out(p, dp, nbytes)
  end subroutine solve_in_place
end module flutas_psolve_call
!
module mod_solver_gpu
  use, intrinsic :: iso_c_binding
  use flutas_psolve_call
  use flutas_kinds, only: rp, rp_bytes
  implicit none
  private
  public :: solver_gpu
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
    call solve_in_place(arrplan(1,1), c_loc(p), rp_bytes*size(p, kind=c_size_t))
  end subroutine solver_gpu
end module mod_solver_gpu
!
module mod_solver_cpu
  use, intrinsic :: iso_c_binding
  use flutas_psolve_call
  use flutas_kinds, only: rp, rp_bytes
  implicit none
  private
  public :: solver_cpu
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
    call solve_in_place(arrplan(1,1), c_loc(p), rp_bytes*size(p, kind=c_size_t))
  end subroutine solver_cpu
end module mod_solver_cpu
