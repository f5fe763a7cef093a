! Round trip of the drop-in mod_load on host arrays, in the call forms of
! the reference's apps (main__single_phase.f90:468-494): writes a field
! file, a CaNS-layout file and scalar.out into the directory given as the
! first argument, reads them back into other arrays, and reads the last
! checkpoint of restart_checkpoints.out.  Prints "restart_demo ok" and the
! checkpoint number, or stops with a message.
program restart_demo
  use, intrinsic :: iso_c_binding
  use mod_load, only: load, load_ns, load_scalar, cmpt_it_chkpt
  implicit none
  integer, parameter :: rp = c_double
  integer, parameter :: n(3) = [6, 5, 4]
  real(rp), allocatable :: u(:,:,:), v(:,:,:), w(:,:,:), p(:,:,:)
  real(rp), allocatable :: u2(:,:,:), v2(:,:,:), w2(:,:,:), p2(:,:,:)
  real(rp) :: time, istep, time2, istep2, t, dto, t2, dto2
  integer :: i, j, k, it, it2, iu
  character(len=512) :: dir
  call get_command_argument(1, dir)
  allocate(u(0:n(1)+1, 0:n(2)+1, 0:n(3)+1), v(0:n(1)+1, 0:n(2)+1, 0:n(3)+1), &
           w(0:n(1)+1, 0:n(2)+1, 0:n(3)+1), p(0:n(1)+1, 0:n(2)+1, 0:n(3)+1))
  allocate(u2(0:n(1)+1, 0:n(2)+1, 0:n(3)+1), v2(0:n(1)+1, 0:n(2)+1, 0:n(3)+1), &
           w2(0:n(1)+1, 0:n(2)+1, 0:n(3)+1), p2(0:n(1)+1, 0:n(2)+1, 0:n(3)+1))
  do k = 0, n(3)+1
    do j = 0, n(2)+1
      do i = 0, n(1)+1
        u(i,j,k) = i + 10*j + 100*k + 0.25_rp
        v(i,j,k) = -u(i,j,k)
        w(i,j,k) = 0.5_rp*u(i,j,k)
        p(i,j,k) = 1.0_rp/(1.0_rp + u(i,j,k))
      end do
    end do
  end do
  u2 = 0; v2 = 0; w2 = 0; p2 = 0
  ! one field, written from and read into the interior of a halo array
  call load('w', trim(dir)//'fldu.bin', n, u(1:n(1),1:n(2),1:n(3)))
  call load('r', trim(dir)//'fldu.bin', n, u2(1:n(1),1:n(2),1:n(3)))
  if (any(u2(1:n(1),1:n(2),1:n(3)) /= u(1:n(1),1:n(2),1:n(3)))) stop 'load round trip'
  if (any(u2(0,:,:) /= 0) .or. any(u2(:,:,n(3)+1) /= 0)) stop 'load wrote a halo'
  ! the CaNS layout
  time = 1.5_rp
  istep = 42.0_rp
  call load_ns('w', trim(dir)//'fld.bin', n, u(1:n(1),1:n(2),1:n(3)), v(1:n(1),1:n(2),1:n(3)), &
               w(1:n(1),1:n(2),1:n(3)), p(1:n(1),1:n(2),1:n(3)), time, istep)
  time2 = 0; istep2 = 0
  call load_ns('r', trim(dir)//'fld.bin', n, u2(1:n(1),1:n(2),1:n(3)), v2(1:n(1),1:n(2),1:n(3)), &
               w2(1:n(1),1:n(2),1:n(3)), p2(1:n(1),1:n(2),1:n(3)), time2, istep2)
  if (time2 /= time .or. istep2 /= istep) stop 'load_ns info'
  if (any(v2(1:n(1),1:n(2),1:n(3)) /= v(1:n(1),1:n(2),1:n(3))) .or. &
      any(w2(1:n(1),1:n(2),1:n(3)) /= w(1:n(1),1:n(2),1:n(3))) .or. &
      any(p2(1:n(1),1:n(2),1:n(3)) /= p(1:n(1),1:n(2),1:n(3)))) stop 'load_ns round trip'
  ! scalar.out
  t = 0.125_rp
  dto = 1.0e-3_rp
  it = 1234
  call load_scalar('w', trim(dir)//'scalar.out', t, it, dto)
  t2 = 0; dto2 = 0; it2 = 0
  call load_scalar('r', trim(dir)//'scalar.out', t2, it2, dto2)
  if (abs(t2 - t) > 1.0e-7_rp*abs(t) .or. abs(dto2 - dto) > 1.0e-7_rp*abs(dto) .or. it2 /= it) &
    stop 'load_scalar round trip'
  ! the checkpoint list: three columns per line, the last line wins
  open(newunit=iu, file=trim(dir)//'restart_checkpoints.out', status='replace', action='write')
  write(iu, *) 0.1_rp, 10.0_rp, 1.0_rp
  write(iu, *) 0.2_rp, 20.0_rp, 2.0_rp
  write(iu, *) 0.3_rp, 30.0_rp, 3.0_rp
  close(iu)
  call cmpt_it_chkpt(trim(dir), it)
  print '(a, i0)', 'restart_demo ok ', it
end program restart_demo
