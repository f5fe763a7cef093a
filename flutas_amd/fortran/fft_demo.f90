! The drop-in mod_fft on host arrays, as solver_cpu drives it
! (src/solver_cpu.f90:59-89): fftini, the forward x and y transforms, the
! backward y and x transforms, normfft, fftend.
!
!   fft_demo IN OUT     IN holds int32 nx, ny, nz, the six characters
!                       bcx0 bcx1 bcy0 bcy1 c_or_f(1) c_or_f(2), then the
!                       nx*ny*nz FP64 array (Fortran order).  OUT receives
!                       normfft, the array after fft(arrplan(1,1)), after
!                       fft(arrplan(1,2)) and after the two backward
!                       transforms scaled by normfft.
!   fft_demo -n BCS     prints normfft of fftini for nx = 12, ny = 10 and
!                       the six characters BCS, without device work.
program fft_demo
  use, intrinsic :: iso_c_binding
  use mod_fft, only: fftini, fftend, fft
  implicit none
  integer, parameter :: rp = c_double
  character(len=512) :: fin, fout
  character(len=6) :: bcs
  character(len=1) :: bcxy(0:1,2), c_or_f(2)
  integer(c_int32_t) :: hdr(3)
  integer :: n(3), iu
  type(c_ptr) :: arrplan(2,2)
  real(rp) :: normfft
  real(rp), allocatable :: a(:,:,:), b(:,:,:)
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  if (trim(fin) == '-n') then
    bcs = fout(1:6)
    call split(bcs)
    n = [12, 10, 4]
    call fftini(n, n, bcxy, c_or_f, arrplan, normfft)
    print '(es24.16)', normfft
    call fftend(arrplan)
  else
    call run()
  end if
contains
  subroutine run()
    open(newunit=iu, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
    read(iu) hdr, bcs
    n = int(hdr)
    allocate(a(n(1), n(2), n(3)), b(n(1), n(2), n(3)))
    read(iu) a
    close(iu)
    call split(bcs)
    call fftini(n, n, bcxy, c_or_f, arrplan, normfft)
    open(newunit=iu, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
    write(iu) normfft
    b = a
    call fft(arrplan(1,1), b)
    write(iu) b
    call fft(arrplan(1,2), b)
    write(iu) b
    call fft(arrplan(2,2), b)
    call fft(arrplan(2,1), b)
    b = b*normfft
    write(iu) b
    close(iu)
    call fftend(arrplan)
    print '(a)', 'fft_demo ok'
  end subroutine run
  !
  subroutine split(s)
    character(len=6), intent(in) :: s
    bcxy(0,1) = s(1:1)
    bcxy(1,1) = s(2:2)
    bcxy(0,2) = s(3:3)
    bcxy(1,2) = s(4:4)
    c_or_f(1) = s(5:5)
    c_or_f(2) = s(6:6)
  end subroutine split
end program fft_demo
