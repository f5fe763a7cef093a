! Drop-in field modules of the projection step: the reference's fillps,
! correc, pres_tw_src, pres_sp_src, chkdiv, boundp and updt_rhs_b
! signatures, backed by the flutas_amd field kernels.  Fields are real(rp)
! (flutas_kinds): FP64, or FP32 in a _SINGLE_PRECISION build, which calls
! the library's FP32 entry points.
!
! Fields that are HIP device or managed allocations go to the kernels as
! they are; plain host arrays are staged through device buffers
! (flutas_stage).  As in the reference, _CONSTANT_COEFFS_POISSON selects the
! constant-coefficient branches of fillps (the rho0 factor), correc and
! pres_tw_src.
module flutas_field_sizes
  use, intrinsic :: iso_c_binding
  use flutas_kinds, only: rp_bytes
  implicit none
  private
  public :: nbytes
contains
  pure function nbytes(nelem) result(b)
    integer, intent(in) :: nelem
    integer(c_size_t) :: b
    b = rp_bytes * int(nelem, c_size_t)
  end function nbytes
end module flutas_field_sizes
!
! src/fillps.f90:16-26
module mod_fillps
  use, intrinsic :: iso_c_binding
  use flutas_capi
  use flutas_stage
  use flutas_field_sizes
  use flutas_kinds, only: rp
  implicit none
  private
  public :: fillps
contains
  subroutine fillps(nx,ny,nz,nh_d,nh_u,dxi,dyi,dzi,dzfi,dti,rho0,u,v,w,p)
    integer , intent(in )                                     :: nx,ny,nz
    integer , intent(in )                                     :: nh_d,nh_u
    real(rp), intent(in )                                     :: dxi,dyi,dzi
    real(rp), intent(in ), dimension(1-nh_d:), target, contiguous :: dzfi
    real(rp), intent(in )                                     :: dti,rho0
    real(rp), intent(in ), dimension(1-nh_u:,1-nh_u:,1-nh_u:), target, contiguous :: u,v,w
    real(rp), intent(out), dimension(0:,0:,0:), target, contiguous :: p
    type(c_ptr) :: du,dv,dw,dp,dz
    real(rp) :: scale
    ! the kernel multiplies by its rho0 argument; the reference scales by
    ! rho0 only under _CONSTANT_COEFFS_POISSON (src/fillps.f90:55-57), and
    ! a product with 1 is exact
#if defined(_CONSTANT_COEFFS_POISSON)
    scale = rho0
#else
    scale = 1._rp
#endif
    du = stage_in(1, c_loc(u), nbytes(size(u)), .true.)
    dv = stage_in(2, c_loc(v), nbytes(size(v)), .true.)
    dw = stage_in(3, c_loc(w), nbytes(size(w)), .true.)
    dp = stage_in(4, c_loc(p), nbytes(size(p)), .true.)
    dz = stage_in(5, c_loc(dzfi), nbytes(size(dzfi)), .true.)
#if defined(_SINGLE_PRECISION)
    call flutas_check(flutas_fillps_f32(nx,ny,nz,nh_d,nh_u,dxi,dyi,dzi,dz,dti,scale,du,dv,dw,dp), 'fillps')
#else
    call flutas_check(flutas_fillps(nx,ny,nz,nh_d,nh_u,dxi,dyi,dzi,dz,dti,scale,du,dv,dw,dp), 'fillps')
#endif
    call stage_out(c_loc(p), dp, nbytes(size(p)))
  end subroutine fillps
end module mod_fillps
!
! src/correc.f90:16-29
module mod_correc
  use, intrinsic :: iso_c_binding
  use flutas_capi
  use flutas_stage
  use flutas_field_sizes
  use flutas_kinds, only: rp
  implicit none
  private
  public :: correc
contains
  subroutine correc(nx,ny,nz,nh_d,nh_u,dxi,dyi,dzi,dzci,dt,rho0,p,u,v,w,rho)
    integer , intent(in   )                                     :: nx,ny,nz
    integer , intent(in   )                                     :: nh_d,nh_u
    real(rp), intent(in   )                                     :: dxi,dyi,dzi
    real(rp), intent(in   ), dimension(1-nh_d:), target, contiguous :: dzci
    real(rp), intent(in   )                                     :: dt,rho0
    real(rp), intent(in   ), dimension(0:,0:,0:), target, contiguous :: p
    real(rp), intent(inout), dimension(1-nh_u:,1-nh_u:,1-nh_u:), target, contiguous :: u,v,w
    real(rp), intent(in   ), dimension(0:,0:,0:), target, contiguous :: rho
    type(c_ptr) :: du,dv,dw,dp,dz,dr
    integer(c_int) :: cc
#if defined(_CONSTANT_COEFFS_POISSON)
    cc = 1
    dr = c_null_ptr
#else
    cc = 0
    dr = stage_in(5, c_loc(rho), nbytes(size(rho)), .true.)
#endif
    du = stage_in(1, c_loc(u), nbytes(size(u)), .true.)
    dv = stage_in(2, c_loc(v), nbytes(size(v)), .true.)
    dw = stage_in(3, c_loc(w), nbytes(size(w)), .true.)
    dp = stage_in(4, c_loc(p), nbytes(size(p)), .true.)
    dz = stage_in(6, c_loc(dzci), nbytes(size(dzci)), .true.)
#if defined(_SINGLE_PRECISION)
    call flutas_check(flutas_correc_f32(nx,ny,nz,nh_d,nh_u,dxi,dyi,dzi,dz,dt,rho0,dp,du,dv,dw,dr,cc), 'correc')
#else
    call flutas_check(flutas_correc(nx,ny,nz,nh_d,nh_u,dxi,dyi,dzi,dz,dt,rho0,dp,du,dv,dw,dr,cc), 'correc')
#endif
    call stage_out(c_loc(u), du, nbytes(size(u)))
    call stage_out(c_loc(v), dv, nbytes(size(v)))
    call stage_out(c_loc(w), dw, nbytes(size(w)))
  end subroutine correc
end module mod_correc
!
! src/source.f90:247-259 (pres_tw_src) and :311-323 (pres_sp_src)
module mod_source
  use, intrinsic :: iso_c_binding
  use flutas_capi
  use flutas_stage
  use flutas_field_sizes
  use flutas_kinds, only: rp
  implicit none
  private
  public :: pres_tw_src, pres_sp_src
contains
  subroutine pres_tw_src(nx,ny,nz,dxi,dyi,dzi,nh_d,nh_u,dzci,rho0i,f_t12,f_t12_o,p,pold,rho,u,v,w)
    integer , intent(in   )                                     :: nx,ny,nz
    real(rp), intent(in   )                                     :: dxi,dyi,dzi
    integer , intent(in   )                                     :: nh_d,nh_u
    real(rp), intent(in   ), dimension(1-nh_d:), target, contiguous :: dzci
    real(rp), intent(in   )                                     :: rho0i,f_t12,f_t12_o
    real(rp), intent(in   ), dimension(0:,0:,0:), target, contiguous :: p,pold
    real(rp), intent(in   ), dimension(0:,0:,0:), target, contiguous :: rho
    real(rp), intent(inout), dimension(1-nh_u:,1-nh_u:,1-nh_u:), target, contiguous :: u,v,w
    type(c_ptr) :: du,dv,dw,dp,dpo,dr,dz
    integer(c_int) :: cc
#if defined(_CONSTANT_COEFFS_POISSON)
    cc = 1
#else
    cc = 0
#endif
    du  = stage_in(1, c_loc(u), nbytes(size(u)), .true.)
    dv  = stage_in(2, c_loc(v), nbytes(size(v)), .true.)
    dw  = stage_in(3, c_loc(w), nbytes(size(w)), .true.)
    dp  = stage_in(4, c_loc(p), nbytes(size(p)), .true.)
    dpo = stage_in(5, c_loc(pold), nbytes(size(pold)), .true.)
    dr  = stage_in(6, c_loc(rho), nbytes(size(rho)), .true.)
    dz  = stage_in(7, c_loc(dzci), nbytes(size(dzci)), .true.)
#if defined(_SINGLE_PRECISION)
    call flutas_check(flutas_pres_tw_src_f32(nx,ny,nz,dxi,dyi,dzi,nh_d,nh_u,dz,rho0i,f_t12,f_t12_o, &
                                         dp,dpo,dr,du,dv,dw,cc), 'pres_tw_src')
#else
    call flutas_check(flutas_pres_tw_src(nx,ny,nz,dxi,dyi,dzi,nh_d,nh_u,dz,rho0i,f_t12,f_t12_o, &
                                         dp,dpo,dr,du,dv,dw,cc), 'pres_tw_src')
#endif
    call stage_out(c_loc(u), du, nbytes(size(u)))
    call stage_out(c_loc(v), dv, nbytes(size(v)))
    call stage_out(c_loc(w), dw, nbytes(size(w)))
  end subroutine pres_tw_src
  !
  subroutine pres_sp_src(nx,ny,nz,f_t12,dxi,dyi,dzi,nh_d,nh_u,dzci,rho0i,pold,u,v,w)
    integer , intent(in   )                                     :: nx,ny,nz
    real(rp), intent(in   )                                     :: f_t12
    real(rp), intent(in   )                                     :: dxi,dyi,dzi
    integer , intent(in   )                                     :: nh_d,nh_u
    real(rp), intent(in   ), dimension(1-nh_d:), target, contiguous :: dzci
    real(rp), intent(in   )                                     :: rho0i
    real(rp), intent(in   ), dimension(0:,0:,0:), target, contiguous :: pold
    real(rp), intent(inout), dimension(1-nh_u:,1-nh_u:,1-nh_u:), target, contiguous :: u,v,w
    type(c_ptr) :: du,dv,dw,dpo,dz
    du  = stage_in(1, c_loc(u), nbytes(size(u)), .true.)
    dv  = stage_in(2, c_loc(v), nbytes(size(v)), .true.)
    dw  = stage_in(3, c_loc(w), nbytes(size(w)), .true.)
    dpo = stage_in(5, c_loc(pold), nbytes(size(pold)), .true.)
    dz  = stage_in(7, c_loc(dzci), nbytes(size(dzci)), .true.)
#if defined(_SINGLE_PRECISION)
    call flutas_check(flutas_pres_sp_src_f32(nx,ny,nz,f_t12,dxi,dyi,dzi,nh_d,nh_u,dz,rho0i,dpo,du,dv,dw), &
                      'pres_sp_src')
#else
    call flutas_check(flutas_pres_sp_src(nx,ny,nz,f_t12,dxi,dyi,dzi,nh_d,nh_u,dz,rho0i,dpo,du,dv,dw), &
                      'pres_sp_src')
#endif
    call stage_out(c_loc(u), du, nbytes(size(u)))
    call stage_out(c_loc(v), dv, nbytes(size(v)))
    call stage_out(c_loc(w), dw, nbytes(size(w)))
  end subroutine pres_sp_src
end module mod_source
!
! src/chkdiv.f90:18-31
module mod_chkdiv
  use, intrinsic :: iso_c_binding
  use flutas_capi
  use flutas_stage
  use flutas_field_sizes
  use flutas_kinds, only: rp
  implicit none
  private
  public :: chkdiv
contains
  subroutine chkdiv(nx,ny,nz,dxi,dyi,dzi,nh_d,nh_u,dzfi,u,v,w,divtot,divmax)
    integer , intent(in )                                     :: nx,ny,nz
    real(rp), intent(in )                                     :: dxi,dyi,dzi
    integer , intent(in )                                     :: nh_d,nh_u
    real(rp), intent(in ), dimension(1-nh_d:), target, contiguous :: dzfi
    real(rp), intent(in ), dimension(1-nh_u:,1-nh_u:,1-nh_u:), target, contiguous :: u,v,w
    real(rp), intent(out)                                     :: divtot,divmax
    type(c_ptr) :: du,dv,dw,dz
    du = stage_in(1, c_loc(u), nbytes(size(u)), .true.)
    dv = stage_in(2, c_loc(v), nbytes(size(v)), .true.)
    dw = stage_in(3, c_loc(w), nbytes(size(w)), .true.)
    dz = stage_in(5, c_loc(dzfi), nbytes(size(dzfi)), .true.)
#if defined(_SINGLE_PRECISION)
    call flutas_check(flutas_chkdiv_f32(nx,ny,nz,dxi,dyi,dzi,nh_d,nh_u,dz,du,dv,dw,divtot,divmax), 'chkdiv')
#else
    call flutas_check(flutas_chkdiv(nx,ny,nz,dxi,dyi,dzi,nh_d,nh_u,dz,du,dv,dw,divtot,divmax), 'chkdiv')
#endif
    ! the report the reference prints (src/chkdiv.f90:66; one rank)
    print*, 'Total divergence = ', divtot, '| Maximum divergence = ', divmax
  end subroutine chkdiv
end module mod_chkdiv
!
! src/bound.f90:17-144 (bounduvw, velocity halo nh_u <= 8), :146-225
! (boundp, any field with halo nh_p <= 8) and :829-944 (updt_rhs_b,
! pressure halo 1), one rank.
module mod_bound
  use, intrinsic :: iso_c_binding
  use flutas_capi
  use flutas_stage
  use flutas_field_sizes
  use flutas_kinds, only: rp
  implicit none
  private
  public :: bounduvw, boundp, updt_rhs_b
contains
  ! host copy of a metric array that may live in device memory
  subroutine metric_to_host(arr, out)
    real(rp), intent(in), dimension(:), target, contiguous :: arr
    real(rp), intent(out), dimension(:), target, contiguous :: out
    integer(c_int) :: isdev
    call flutas_check(flutas_pointer_is_device(c_loc(arr), isdev), 'pointer query')
    if (isdev == 1) then
      call flutas_check(flutas_memcpy_d2h(c_loc(out), c_loc(arr), nbytes(size(arr))), 'metric copy')
    else
      out = arr
    end if
  end subroutine metric_to_host
  !
  subroutine bounduvw(cbc,n,bc,nh_d,nh_u,halo,isoutflow,dl,dzc,dzf,u,v,w)
    character(len=1), intent(in   ), dimension(0:1,3,3)                 :: cbc
    integer         , intent(in   ), dimension(3)                       :: n
    real(rp)        , intent(in   ), dimension(0:1,3,3)                 :: bc
    integer         , intent(in   )                                     :: nh_d,nh_u
    integer         , intent(in   ), dimension(3)                       :: halo
    logical         , intent(in   ), dimension(0:1,3)                   :: isoutflow
    real(rp)        , intent(in   ), dimension(3)                       :: dl
    real(rp)        , intent(in   ), dimension(1-nh_d:), target, contiguous :: dzc,dzf
    real(rp)        , intent(inout), dimension(1-nh_u:,1-nh_u:,1-nh_u:), target, contiguous :: u,v,w
    character(kind=c_char) :: cbc_c(18)
    real(rp) :: bc_c(18)
    integer(c_int) :: nn(3), out_c(6)
    real(rp), allocatable, target :: hdzc(:), hdzf(:)
    type(c_ptr) :: du, dv, dw
    integer :: q, s, d, c
    do c = 1, 3
      do d = 1, 3
        do s = 0, 1
          q = s + 2*(d-1) + 6*(c-1) + 1
          cbc_c(q) = cbc(s,d,c)
          bc_c(q) = bc(s,d,c)
        end do
      end do
    end do
    do d = 1, 3
      do s = 0, 1
        out_c(s + 2*(d-1) + 1) = merge(1_c_int, 0_c_int, isoutflow(s,d))
      end do
    end do
    nn = int(n, c_int)
    allocate(hdzc(size(dzc)), hdzf(size(dzf)))
    call metric_to_host(dzc, hdzc)
    call metric_to_host(dzf, hdzf)
    du = stage_in(1, c_loc(u), nbytes(size(u)), .true.)
    dv = stage_in(2, c_loc(v), nbytes(size(v)), .true.)
    dw = stage_in(3, c_loc(w), nbytes(size(w)), .true.)
#if defined(_SINGLE_PRECISION)
    call flutas_check(flutas_bounduvw_f32(cbc_c, nn, bc_c, int(nh_d, c_int), int(nh_u, c_int), out_c, &
                                      dl, hdzc, hdzf, du, dv, dw), 'bounduvw')
#else
    call flutas_check(flutas_bounduvw(cbc_c, nn, bc_c, int(nh_d, c_int), int(nh_u, c_int), out_c, &
                                      dl, hdzc, hdzf, du, dv, dw), 'bounduvw')
#endif
    call stage_out(c_loc(u), du, nbytes(size(u)))
    call stage_out(c_loc(v), dv, nbytes(size(v)))
    call stage_out(c_loc(w), dw, nbytes(size(w)))
  end subroutine bounduvw
  !
  subroutine boundp(cbc,n,bc,nh_d,nh_p,halo,dl,dzc,dzf,p)
    character(len=1), intent(in   ), dimension(0:1,3)                   :: cbc
    integer         , intent(in   ), dimension(3)                       :: n
    real(rp)        , intent(in   ), dimension(0:1,3)                   :: bc
    integer         , intent(in   )                                     :: nh_d,nh_p
    integer         , intent(in   ), dimension(3)                       :: halo
    real(rp)        , intent(in   ), dimension(3)                       :: dl
    real(rp)        , intent(in   ), dimension(1-nh_d:), target, contiguous :: dzc,dzf
    real(rp)        , intent(inout), dimension(1-nh_p:,1-nh_p:,1-nh_p:), target, contiguous :: p
    character(kind=c_char) :: cbc_c(6)
    real(rp) :: bc_c(6)
    integer(c_int) :: nn(3)
    real(rp), allocatable, target :: hdzc(:), hdzf(:)
    type(c_ptr) :: dp
    integer :: q, s, d
    do d = 1, 3
      do s = 0, 1
        q = 2*(d-1) + s + 1
        cbc_c(q) = cbc(s,d)
        bc_c(q) = bc(s,d)
      end do
    end do
    nn = int(n, c_int)
    allocate(hdzc(size(dzc)), hdzf(size(dzf)))
    call metric_to_host(dzc, hdzc)
    call metric_to_host(dzf, hdzf)
    dp = stage_in(4, c_loc(p), nbytes(size(p)), .true.)
#if defined(_SINGLE_PRECISION)
    call flutas_check(flutas_boundp_halo_f32(cbc_c, nn, bc_c, int(nh_d, c_int), int(nh_p, c_int), dl, &
                                         hdzc, hdzf, dp), 'boundp')
#else
    call flutas_check(flutas_boundp_halo(cbc_c, nn, bc_c, int(nh_d, c_int), int(nh_p, c_int), dl, &
                                         hdzc, hdzf, dp), 'boundp')
#endif
    call stage_out(c_loc(p), dp, nbytes(size(p)))
  end subroutine boundp
  !
  subroutine updt_rhs_b(nx,ny,nz,c_or_f,cbc,nh_p,rhsbx,rhsby,rhsbz,p)
    integer         , intent(in   )                                     :: nx,ny,nz
    character       , intent(in   ), dimension(3)                       :: c_or_f
    character(len=1), intent(in   ), dimension(0:1,3)                   :: cbc
    integer         , intent(in   )                                     :: nh_p
    real(rp)        , intent(in   ), dimension(:,:,0:), target, contiguous :: rhsbx,rhsby,rhsbz
    real(rp)        , intent(inout), dimension(1-nh_p:,1-nh_p:,1-nh_p:), target, contiguous :: p
    character(kind=c_char) :: cbc_c(6), cf_c(3)
    type(c_ptr) :: dp, dbx, dby, dbz
    integer :: q, s, d
    if (nh_p /= 1) error stop 'flutas_amd updt_rhs_b: pressure halo nh_p = 1 only'
    do d = 1, 3
      do s = 0, 1
        q = 2*(d-1) + s + 1
        cbc_c(q) = cbc(s,d)
      end do
      cf_c(d) = c_or_f(d)
    end do
    dbx = stage_in(5, c_loc(rhsbx), nbytes(size(rhsbx)), .true.)
    dby = stage_in(6, c_loc(rhsby), nbytes(size(rhsby)), .true.)
    dbz = stage_in(7, c_loc(rhsbz), nbytes(size(rhsbz)), .true.)
    dp  = stage_in(4, c_loc(p), nbytes(size(p)), .true.)
#if defined(_SINGLE_PRECISION)
    call flutas_check(flutas_updt_rhs_b_f32(nx,ny,nz,cf_c,cbc_c,dbx,dby,dbz,dp), 'updt_rhs_b')
#else
    call flutas_check(flutas_updt_rhs_b(nx,ny,nz,cf_c,cbc_c,dbx,dby,dbz,dp), 'updt_rhs_b')
#endif
    call stage_out(c_loc(p), dp, nbytes(size(p)))
  end subroutine updt_rhs_b
end module mod_bound
