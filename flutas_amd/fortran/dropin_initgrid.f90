! Drop-in mod_initgrid: the reference's initgrid (src/initgrid.f90:17-97)
! with its signature, backed by flutas_initgrid_inivel (host arrays, no
! device work).  inivel 'hcl' / 'hcp' cluster the grid at the lower wall,
! every other value at both walls (src/initgrid.f90:32-39).
module mod_initgrid
  use, intrinsic :: iso_c_binding
  use flutas_capi
  use flutas_kinds, only: rp
  implicit none
  private
  public :: initgrid
contains
  subroutine initgrid(inivel,n,gr,lz,nh_d,dzc,dzf,zc,zf)
    character(len=3), intent(in )                           :: inivel
    integer         , intent(in )                           :: n
    real(rp)        , intent(in )                           :: gr,lz
    integer         , intent(in )                           :: nh_d
    real(rp)        , intent(out), dimension(1-nh_d:n+nh_d) :: dzc,dzf,zc,zf
    ! computed in FP64, returned in the kind of rp (FluTAS's single-precision
    ! build computes them in FP32: the values differ in the last bits)
    real(c_double), dimension(1-nh_d:n+nh_d) :: hzc,hzf,hc,hf
    call flutas_check(flutas_initgrid_inivel(inivel//c_null_char, int(n, c_int), &
                                             real(gr, c_double), real(lz, c_double), &
                                             int(nh_d, c_int), hzc, hzf, hc, hf), 'initgrid')
    dzc = real(hzc, rp)
    dzf = real(hzf, rp)
    zc = real(hc, rp)
    zf = real(hf, rp)
  end subroutine initgrid
end module mod_initgrid
