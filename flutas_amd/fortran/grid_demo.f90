! The drop-in mod_initgrid as the reference's apps call it
! (main__single_phase.f90:375): grid_demo INIVEL N GR LZ NH_D prints
! dzc, dzf, zc, zf over (1-nh_d : n+nh_d), one array per line.
program grid_demo
  use, intrinsic :: iso_c_binding
  use mod_initgrid, only: initgrid
  implicit none
  integer, parameter :: rp = c_double
  character(len=32) :: arg
  character(len=3) :: inivel
  integer :: n, nh_d
  real(rp) :: gr, lz
  real(rp), allocatable :: dzc(:), dzf(:), zc(:), zf(:)
  call get_command_argument(1, arg); inivel = arg(1:3)
  call get_command_argument(2, arg); read(arg, *) n
  call get_command_argument(3, arg); read(arg, *) gr
  call get_command_argument(4, arg); read(arg, *) lz
  call get_command_argument(5, arg); read(arg, *) nh_d
  allocate(dzc(1-nh_d:n+nh_d), dzf(1-nh_d:n+nh_d), zc(1-nh_d:n+nh_d), zf(1-nh_d:n+nh_d))
  call initgrid(inivel, n, gr, lz, nh_d, dzc, dzf, zc, zf)
  print '(*(es25.17))', dzc
  print '(*(es25.17))', dzf
  print '(*(es25.17))', zc
  print '(*(es25.17))', zf
end program grid_demo
