! One pressure-projection step through the drop-in modules, in the order of
! the reference's single-phase time loop (main__single_phase.f90:705-727):
!   fillps -> updt_rhs_b -> solver_gpu -> boundp -> correc -> bounduvw
! plus chkdiv of the predicted velocity, with the grid of initgrid and the
! plan of initsolver.  nh_d = max(nh_u, 1) as the reference sets it
! (main__single_phase.f90:261); the velocity BCs follow the pressure
! pattern: periodic stays periodic, a pressure N wall is a no-slip (D, 0)
! wall, a pressure D face a zero-gradient (N) face, no outflow faces.
!
!   projection_demo INPUT OUTPUT MODE [SOLVER]
!
! INPUT (stream binary): int32 nx,ny,nz,nh_u; float64 lx,ly,lz,alpha,dt,rho0;
! 6 characters cbc (the (0:1,3) order); then u, v, w as
! (1-nh_u:n+nh_u)^3 arrays in Fortran order.
! OUTPUT (stream binary): f (right-hand side after updt_rhs_b), p (after
! the solve and boundp), u, v, w (after correc and bounduvw), divtot and
! divmax of the predicted velocity.
! MODE: "host" passes plain Fortran arrays (the modules stage them through
! device buffers); "device" passes HIP device allocations.  SOLVER: "gpu"
! (default) calls solver_gpu, "cpu" the drop-in solver_cpu
! (src/solver_cpu.f90:20-31), as the reference's CPU builds do.
! Built three times: with _CONSTANT_COEFFS_POISSON (projection_demo),
! without (projection_demo_vc: fillps without rho0, correc dividing by the
! face density of rho = rho0), and with _CONSTANT_COEFFS_POISSON and
! _SINGLE_PRECISION (projection_demo_sp: rp = FP32; the FP64 input is
! rounded to FP32 and OUTPUT holds FP32 values).
program projection_demo
  use, intrinsic :: iso_c_binding
  use flutas_capi
  use mod_initsolver, only: initsolver
  use mod_fft,        only: fftend
  use mod_solver_gpu, only: solver_gpu
  use mod_solver_cpu, only: solver_cpu
  use mod_fillps,     only: fillps
  use mod_correc,     only: correc
  use mod_chkdiv,     only: chkdiv
  use mod_bound,      only: bounduvw, boundp, updt_rhs_b
  use mod_initgrid,   only: initgrid
  use flutas_kinds,   only: rp, rp_bytes
  implicit none
  character(len=512) :: fin, fout, mode, solver
  integer(c_int32_t) :: hdr(4)
  real(rp) :: par(6)
  real(c_double) :: par8(6)
  real(c_double), allocatable :: f8(:,:,:)
  character(len=1) :: cbc(0:1,3), cbc6(6), cbcvel(0:1,3,3)
  real(rp) :: bc(0:1,3), bcvel(0:1,3,3)
  logical :: no_outflow(0:1,3)
  integer :: n(3), nh_u, nh_d, i, u_in, u_out
  real(rp), allocatable, target :: u(:,:,:), v(:,:,:), w(:,:,:), p(:,:,:), rho(:,:,:), f(:,:,:)
  real(rp), pointer, contiguous :: ud(:,:,:), vd(:,:,:), wd(:,:,:), pd(:,:,:)
  type(c_ptr) :: cu, cv, cw, cp
  integer(c_size_t) :: bu, bp
  !
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, mode)
  solver = 'gpu'
  if (command_argument_count() >= 4) call get_command_argument(4, solver)
  open(newunit=u_in, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u_in) hdr, par8, cbc6
  par = real(par8, rp)
  n = hdr(1:3)
  nh_u = hdr(4)
  nh_d = max(nh_u, 1)
  do i = 1, 6
    cbc(mod(i-1,2), (i-1)/2 + 1) = cbc6(i)
  end do
  bc = 0._rp
  do i = 1, 3
    where (cbc == 'P') cbcvel(:,:,i) = 'P'
    where (cbc == 'N') cbcvel(:,:,i) = 'D'
    where (cbc == 'D') cbcvel(:,:,i) = 'N'
  end do
  bcvel = 0._rp
  no_outflow = .false.
  allocate(u(1-nh_u:n(1)+nh_u,1-nh_u:n(2)+nh_u,1-nh_u:n(3)+nh_u))
  allocate(v, w, mold=u)
  allocate(p(0:n(1)+1,0:n(2)+1,0:n(3)+1))
  allocate(rho, f, mold=p)
  allocate(f8(1-nh_u:n(1)+nh_u,1-nh_u:n(2)+nh_u,1-nh_u:n(3)+nh_u))
  read(u_in) f8
  u = real(f8, rp)
  read(u_in) f8
  v = real(f8, rp)
  read(u_in) f8
  w = real(f8, rp)
  deallocate(f8)
  close(u_in)
  p = 0._rp
  rho = par(6)
  bu = rp_bytes*size(u, kind=c_size_t)
  bp = rp_bytes*size(p, kind=c_size_t)
  !
  if (trim(mode) == 'device') then
    call flutas_check(flutas_device_alloc(bu, cu), 'alloc u')
    call flutas_check(flutas_device_alloc(bu, cv), 'alloc v')
    call flutas_check(flutas_device_alloc(bu, cw), 'alloc w')
    call flutas_check(flutas_device_alloc(bp, cp), 'alloc p')
    call flutas_check(flutas_memcpy_h2d(cu, c_loc(u), bu), 'copy u')
    call flutas_check(flutas_memcpy_h2d(cv, c_loc(v), bu), 'copy v')
    call flutas_check(flutas_memcpy_h2d(cw, c_loc(w), bu), 'copy w')
    call flutas_check(flutas_memcpy_h2d(cp, c_loc(p), bp), 'copy p')
    call c_f_pointer(cu, ud, shape(u))
    call c_f_pointer(cv, vd, shape(u))
    call c_f_pointer(cw, wd, shape(u))
    call c_f_pointer(cp, pd, shape(p))
    call step(ud, vd, wd, pd)
    call flutas_check(flutas_memcpy_d2h(c_loc(u), cu, bu), 'copy back u')
    call flutas_check(flutas_memcpy_d2h(c_loc(v), cv, bu), 'copy back v')
    call flutas_check(flutas_memcpy_d2h(c_loc(w), cw, bu), 'copy back w')
    call flutas_check(flutas_memcpy_d2h(c_loc(p), cp, bp), 'copy back p')
    call flutas_check(flutas_device_free(cu), 'free')
    call flutas_check(flutas_device_free(cv), 'free')
    call flutas_check(flutas_device_free(cw), 'free')
    call flutas_check(flutas_device_free(cp), 'free')
  else
    call step(u, v, w, p)
  end if
  !
  open(newunit=u_out, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u_out) f, p, u, v, w
  close(u_out)
  print '(a)', 'projection_demo: done'
contains
  subroutine step(uu, vv, ww, pp)
    real(rp), intent(inout), dimension(1-nh_u:,1-nh_u:,1-nh_u:), target, contiguous :: uu, vv, ww
    real(rp), intent(inout), dimension(0:,0:,0:), target, contiguous :: pp
    real(rp), allocatable :: dzc(:), dzf(:), zc(:), zf(:), dzci(:), dzfi(:)
    real(rp), allocatable :: lambdaxy(:,:), a(:), b(:), c(:), rhsbx(:,:,:), rhsby(:,:,:), rhsbz(:,:,:)
    type(c_ptr) :: arrplan(2,2)
    real(rp) :: normfft, dli(3), dl(3), dt, dti, rho0, divtot, divmax
    integer :: u_div
    allocate(dzc(1-nh_d:n(3)+nh_d), dzf(1-nh_d:n(3)+nh_d), zc(1-nh_d:n(3)+nh_d), zf(1-nh_d:n(3)+nh_d))
    allocate(dzci, dzfi, mold=dzc)
    call initgrid('zer', n(3), par(4), par(3), nh_d, dzc, dzf, zc, zf)
    dzci(:) = 1._rp/dzc(:)
    dzfi(:) = 1._rp/dzf(:)
    dli = real(n, rp)/par(1:3)
    dl = 1._rp/dli
    dt = par(5)
    dti = 1._rp/dt
    rho0 = par(6)
    allocate(lambdaxy(n(1),n(2)), a(n(3)), b(n(3)), c(n(3)))
    allocate(rhsbx(n(2),n(3),0:1), rhsby(n(1),n(3),0:1), rhsbz(n(1),n(2),0:1))
    call initsolver(n, [1,1,1], [1,1,1], dli, nh_d, dzci, dzfi, cbc, bc, ['c','c','c'], lambdaxy, &
                    a, b, c, arrplan, normfft, rhsbx, rhsby, rhsbz)
    call fillps(n(1), n(2), n(3), nh_d, nh_u, dli(1), dli(2), dli(3), dzfi, dti, rho0, uu, vv, ww, pp)
    call chkdiv(n(1), n(2), n(3), dli(1), dli(2), dli(3), nh_d, nh_u, dzfi, uu, vv, ww, divtot, divmax)
    call updt_rhs_b(n(1), n(2), n(3), ['c','c','c'], cbc, 1, rhsbx, rhsby, rhsbz, pp)
    call snapshot(pp, f)
    if (trim(solver) == 'cpu') then
      call solver_cpu(n, arrplan, normfft, lambdaxy, a, b, c, cbc(:,3), ['c','c','c'], pp)
    else
      call solver_gpu(n, [1,1,1], arrplan, normfft, lambdaxy, a, b, c, cbc, ['c','c','c'], pp)
    end if
    call boundp(cbc, n, bc, nh_d, 1, [1,1,1], dl, dzc, dzf, pp)
    call correc(n(1), n(2), n(3), nh_d, nh_u, dli(1), dli(2), dli(3), dzci, dt, rho0, pp, uu, vv, ww, rho)
    call bounduvw(cbcvel, n, bcvel, nh_d, nh_u, [1,1,1], no_outflow, dl, dzc, dzf, uu, vv, ww)
    call fftend(arrplan)
    open(newunit=u_div, file=trim(fout)//'.div', status='replace')
    write(u_div, '(2es25.16)') divtot, divmax
    close(u_div)
  end subroutine step
  !
  ! host copy of a field that may live in device memory
  subroutine snapshot(src, dst)
    real(rp), intent(in), dimension(0:,0:,0:), target, contiguous :: src
    real(rp), intent(out), dimension(0:,0:,0:), target, contiguous :: dst
    integer(c_int) :: isdev
    call flutas_check(flutas_pointer_is_device(c_loc(src), isdev), 'pointer query')
    if (isdev == 1) then
      call flutas_check(flutas_memcpy_d2h(c_loc(dst), c_loc(src), rp_bytes*size(src, kind=c_size_t)), &
                        'snapshot')
    else
      dst = src
    end if
  end subroutine snapshot
end program projection_demo
