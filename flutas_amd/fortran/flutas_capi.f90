! ISO_C_BINDING interfaces of libflutas_amd.so (include/flutas_amd.h).
! Field arguments are passed as type(c_ptr) to their first element
! (including the halo); BC characters as the six characters of the
! reference's cbc(0:1,3) array in memory order.
module flutas_capi
  use, intrinsic :: iso_c_binding
  implicit none
  private :: c_int, c_double, c_float, c_char, c_ptr, c_size_t, c_long_long
  !
  integer(c_int), parameter :: FLUTAS_OK = 0, FLUTAS_FP64 = 64, FLUTAS_FP32_REFINE = 32
  !
  interface
    function flutas_last_error() bind(C, name="flutas_last_error") result(msg)
      import :: c_ptr
      type(c_ptr) :: msg
    end function
    function flutas_device_alloc(bytes, ptr) bind(C, name="flutas_device_alloc") result(rc)
      import :: c_int, c_size_t, c_ptr
      integer(c_size_t), value :: bytes
      type(c_ptr) :: ptr
      integer(c_int) :: rc
    end function
    function flutas_device_free(ptr) bind(C, name="flutas_device_free") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ptr
      integer(c_int) :: rc
    end function
    function flutas_memcpy_h2d(dst, src, bytes) bind(C, name="flutas_memcpy_h2d") result(rc)
      import :: c_int, c_size_t, c_ptr
      type(c_ptr), value :: dst, src
      integer(c_size_t), value :: bytes
      integer(c_int) :: rc
    end function
    function flutas_memcpy_d2h(dst, src, bytes) bind(C, name="flutas_memcpy_d2h") result(rc)
      import :: c_int, c_size_t, c_ptr
      type(c_ptr), value :: dst, src
      integer(c_size_t), value :: bytes
      integer(c_int) :: rc
    end function
    function flutas_synchronize() bind(C, name="flutas_synchronize") result(rc)
      import :: c_int
      integer(c_int) :: rc
    end function
    function flutas_pointer_is_device(ptr, is_device) bind(C, name="flutas_pointer_is_device") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ptr
      integer(c_int) :: is_device
      integer(c_int) :: rc
    end function
    !
    function flutas_psolve_create(ng, cbc, bc, c_or_f, dli, nh_d, dzci, dzfi, precision, tol, h) &
        bind(C, name="flutas_psolve_create") result(rc)
      import :: c_int, c_double, c_char, c_ptr
      integer(c_int), dimension(3) :: ng
      character(kind=c_char), dimension(*) :: cbc, c_or_f
      real(c_double), dimension(*) :: bc, dli, dzci, dzfi
      integer(c_int), value :: nh_d, precision
      real(c_double), value :: tol
      type(c_ptr) :: h
      integer(c_int) :: rc
    end function
    function flutas_psolve_destroy(h) bind(C, name="flutas_psolve_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int) :: rc
    end function
    function flutas_psolve_coeffs(h, lambdaxy, a, b, c, normfft, rhsbx, rhsby, rhsbz) &
        bind(C, name="flutas_psolve_coeffs") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double), dimension(*) :: lambdaxy, a, b, c, rhsbx, rhsby, rhsbz
      real(c_double) :: normfft
      integer(c_int) :: rc
    end function
    function flutas_psolve_solve(h, p) bind(C, name="flutas_psolve_solve") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: h, p
      integer(c_int) :: rc
    end function
    function flutas_psolve_stats(h, iters, relres) bind(C, name="flutas_psolve_stats") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int) :: iters
      real(c_double) :: relres
      integer(c_int) :: rc
    end function
    !
    function flutas_fillps(nx, ny, nz, nh_d, nh_u, dxi, dyi, dzi, dzfi, dti, rho0, u, v, w, p) &
        bind(C, name="flutas_fillps") result(rc)
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nx, ny, nz, nh_d, nh_u
      real(c_double), value :: dxi, dyi, dzi, dti, rho0
      type(c_ptr), value :: dzfi, u, v, w, p
      integer(c_int) :: rc
    end function
    function flutas_pres_tw_src(nx, ny, nz, dxi, dyi, dzi, nh_d, nh_u, dzci, rho0i, f_t12, f_t12_o, &
        p, pold, rho, u, v, w, const_coeff) bind(C, name="flutas_pres_tw_src") result(rc)
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nx, ny, nz, nh_d, nh_u, const_coeff
      real(c_double), value :: dxi, dyi, dzi, rho0i, f_t12, f_t12_o
      type(c_ptr), value :: dzci, p, pold, rho, u, v, w
      integer(c_int) :: rc
    end function
    function flutas_pres_sp_src(nx, ny, nz, f_t12, dxi, dyi, dzi, nh_d, nh_u, dzci, rho0i, pold, u, v, w) &
        bind(C, name="flutas_pres_sp_src") result(rc)
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nx, ny, nz, nh_d, nh_u
      real(c_double), value :: f_t12, dxi, dyi, dzi, rho0i
      type(c_ptr), value :: dzci, pold, u, v, w
      integer(c_int) :: rc
    end function
    function flutas_updt_rhs_b(nx, ny, nz, c_or_f, cbc, rhsbx, rhsby, rhsbz, p) &
        bind(C, name="flutas_updt_rhs_b") result(rc)
      import :: c_int, c_char, c_ptr
      integer(c_int), value :: nx, ny, nz
      character(kind=c_char), dimension(*) :: c_or_f, cbc
      type(c_ptr), value :: rhsbx, rhsby, rhsbz, p
      integer(c_int) :: rc
    end function
    function flutas_boundp(cbc, n, bc, dl, dzc_0, dzc_n, p) bind(C, name="flutas_boundp") result(rc)
      import :: c_int, c_double, c_char, c_ptr
      character(kind=c_char), dimension(*) :: cbc
      integer(c_int), dimension(3) :: n
      real(c_double), dimension(*) :: bc, dl
      real(c_double), value :: dzc_0, dzc_n
      type(c_ptr), value :: p
      integer(c_int) :: rc
    end function
    function flutas_boundp_halo(cbc, n, bc, nh_d, nh_p, dl, dzc, dzf, p) &
        bind(C, name="flutas_boundp_halo") result(rc)
      import :: c_int, c_double, c_char, c_ptr
      character(kind=c_char), dimension(*) :: cbc
      integer(c_int), dimension(3) :: n
      real(c_double), dimension(*) :: bc, dl, dzc, dzf
      integer(c_int), value :: nh_d, nh_p
      type(c_ptr), value :: p
      integer(c_int) :: rc
    end function
    function flutas_bounduvw(cbc, n, bc, nh_d, nh_u, isoutflow, dl, dzc, dzf, u, v, w) &
        bind(C, name="flutas_bounduvw") result(rc)
      import :: c_int, c_double, c_char, c_ptr
      character(kind=c_char), dimension(*) :: cbc
      integer(c_int), dimension(3) :: n
      real(c_double), dimension(*) :: bc, dl, dzc, dzf
      integer(c_int), value :: nh_d, nh_u
      integer(c_int), dimension(6) :: isoutflow
      type(c_ptr), value :: u, v, w
      integer(c_int) :: rc
    end function
    function flutas_correc(nx, ny, nz, nh_d, nh_u, dxi, dyi, dzi, dzci, dt, rho0, p, u, v, w, rho, &
        const_coeff) bind(C, name="flutas_correc") result(rc)
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nx, ny, nz, nh_d, nh_u, const_coeff
      real(c_double), value :: dxi, dyi, dzi, dt, rho0
      type(c_ptr), value :: dzci, p, u, v, w, rho
      integer(c_int) :: rc
    end function
    ! fused prologue (fillps + updt_rhs_b) and epilogue (correc + pnew = pold + phi)
    function flutas_fillps_rhs(nx, ny, nz, nh_d, nh_u, dxi, dyi, dzi, dzfi, dti, rho0, u, v, w, &
        c_or_f, cbc, rhsbx, rhsby, rhsbz, p) bind(C, name="flutas_fillps_rhs") result(rc)
      import :: c_int, c_double, c_char, c_ptr
      integer(c_int), value :: nx, ny, nz, nh_d, nh_u
      real(c_double), value :: dxi, dyi, dzi, dti, rho0
      type(c_ptr), value :: dzfi, u, v, w
      character(kind=c_char), dimension(*) :: c_or_f, cbc
      type(c_ptr), value :: rhsbx, rhsby, rhsbz, p
      integer(c_int) :: rc
    end function
    function flutas_correc_update(nx, ny, nz, nh_d, nh_u, dxi, dyi, dzi, dzci, dt, rho0, phi, &
        pold, pnew, u, v, w, rho, const_coeff) bind(C, name="flutas_correc_update") result(rc)
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nx, ny, nz, nh_d, nh_u, const_coeff
      real(c_double), value :: dxi, dyi, dzi, dt, rho0
      type(c_ptr), value :: dzci, phi, pold, pnew, u, v, w, rho
      integer(c_int) :: rc
    end function
    function flutas_chkdiv(nx, ny, nz, dxi, dyi, dzi, nh_d, nh_u, dzfi, u, v, w, divtot, divmax) &
        bind(C, name="flutas_chkdiv") result(rc)
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nx, ny, nz, nh_d, nh_u
      real(c_double), value :: dxi, dyi, dzi
      type(c_ptr), value :: dzfi, u, v, w
      real(c_double) :: divtot, divmax
      integer(c_int) :: rc
    end function
    ! restart files (src/load.f90)
    function flutas_load(io, filename, n, halo, fld) bind(C, name="flutas_load") result(rc)
      import :: c_int, c_char, c_ptr
      character(kind=c_char), value :: io
      character(kind=c_char), dimension(*) :: filename
      integer(c_int), dimension(3) :: n
      integer(c_int), value :: halo
      type(c_ptr), value :: fld
      integer(c_int) :: rc
    end function
    function flutas_load_ns_halos(io, filename, n, halo, u, v, w, p, time, istep) &
        bind(C, name="flutas_load_ns_halos") result(rc)
      import :: c_int, c_char, c_ptr, c_double
      character(kind=c_char), value :: io
      character(kind=c_char), dimension(*) :: filename
      integer(c_int), dimension(3) :: n
      integer(c_int), dimension(4) :: halo
      type(c_ptr), value :: u, v, w, p
      real(c_double) :: time, istep
      integer(c_int) :: rc
    end function
    function flutas_load_scalar(io, filename, time, istep, dto) bind(C, name="flutas_load_scalar") &
        result(rc)
      import :: c_int, c_char, c_double
      character(kind=c_char), value :: io
      character(kind=c_char), dimension(*) :: filename
      real(c_double) :: time, dto
      integer(c_int) :: istep
      integer(c_int) :: rc
    end function
    ! r2r lines in the geometry of fftini's guru plans (src/fft.f90:75-124)
    function flutas_fft_guru(kind, n, stride, howmany1, dist1, howmany2, dist2, dat) &
        bind(C, name="flutas_fft_guru") result(rc)
      import :: c_int, c_long_long, c_ptr
      integer(c_int), value :: kind, n, howmany1, howmany2
      integer(c_long_long), value :: stride, dist1, dist2
      type(c_ptr), value :: dat
      integer(c_int) :: rc
    end function
    function flutas_initgrid_inivel(inivel, n, gr, lz, nh_d, dzc, dzf, zc, zf) &
        bind(C, name="flutas_initgrid_inivel") result(rc)
      import :: c_int, c_double, c_char
      character(kind=c_char), dimension(*) :: inivel
      integer(c_int), value :: n, nh_d
      real(c_double), value :: gr, lz
      real(c_double), dimension(*) :: dzc, dzf, zc, zf
      integer(c_int) :: rc
    end function
    function flutas_initgrid(n, gr, lz, nh_d, dzc, dzf, zc, zf) bind(C, name="flutas_initgrid") result(rc)
      import :: c_int, c_double
      integer(c_int), value :: n, nh_d
      real(c_double), value :: gr, lz
      real(c_double), dimension(*) :: dzc, dzf, zc, zf
      integer(c_int) :: rc
    end function
    ! FP32 fields and scalars (FluTAS built with _SINGLE_PRECISION)
    function flutas_psolve_solve_f32(h, p) bind(C, name="flutas_psolve_solve_f32") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: h, p
      integer(c_int) :: rc
    end function
    function flutas_fft_guru_f32(kind, n, stride, howmany1, dist1, howmany2, dist2, data) &
        bind(C, name="flutas_fft_guru_f32") result(rc)
      import :: c_int, c_long_long, c_ptr
      integer(c_int), value :: kind, n, howmany1, howmany2
      integer(c_long_long), value :: stride, dist1, dist2
      type(c_ptr), value :: data
      integer(c_int) :: rc
    end function
    function flutas_fillps_f32(nx, ny, nz, nh_d, nh_u, dxi, dyi, dzi, dzfi, dti, rho0, u, v, w, p) &
        bind(C, name="flutas_fillps_f32") result(rc)
      import :: c_int, c_float, c_ptr
      integer(c_int), value :: nx, ny, nz, nh_d, nh_u
      real(c_float), value :: dxi, dyi, dzi, dti, rho0
      type(c_ptr), value :: dzfi, u, v, w, p
      integer(c_int) :: rc
    end function
    function flutas_pres_tw_src_f32(nx, ny, nz, dxi, dyi, dzi, nh_d, nh_u, dzci, rho0i, f_t12, f_t12_o, &
        p, pold, rho, u, v, w, const_coeff) bind(C, name="flutas_pres_tw_src_f32") result(rc)
      import :: c_int, c_float, c_ptr
      integer(c_int), value :: nx, ny, nz, nh_d, nh_u, const_coeff
      real(c_float), value :: dxi, dyi, dzi, rho0i, f_t12, f_t12_o
      type(c_ptr), value :: dzci, p, pold, rho, u, v, w
      integer(c_int) :: rc
    end function
    function flutas_pres_sp_src_f32(nx, ny, nz, f_t12, dxi, dyi, dzi, nh_d, nh_u, dzci, rho0i, pold, &
        u, v, w) bind(C, name="flutas_pres_sp_src_f32") result(rc)
      import :: c_int, c_float, c_ptr
      integer(c_int), value :: nx, ny, nz, nh_d, nh_u
      real(c_float), value :: f_t12, dxi, dyi, dzi, rho0i
      type(c_ptr), value :: dzci, pold, u, v, w
      integer(c_int) :: rc
    end function
    function flutas_correc_f32(nx, ny, nz, nh_d, nh_u, dxi, dyi, dzi, dzci, dt, rho0, p, u, v, w, &
        rho, const_coeff) bind(C, name="flutas_correc_f32") result(rc)
      import :: c_int, c_float, c_ptr
      integer(c_int), value :: nx, ny, nz, nh_d, nh_u, const_coeff
      real(c_float), value :: dxi, dyi, dzi, dt, rho0
      type(c_ptr), value :: dzci, p, u, v, w, rho
      integer(c_int) :: rc
    end function
    function flutas_chkdiv_f32(nx, ny, nz, dxi, dyi, dzi, nh_d, nh_u, dzfi, u, v, w, divtot, divmax) &
        bind(C, name="flutas_chkdiv_f32") result(rc)
      import :: c_int, c_float, c_ptr
      integer(c_int), value :: nx, ny, nz, nh_d, nh_u
      real(c_float), value :: dxi, dyi, dzi
      type(c_ptr), value :: dzfi, u, v, w
      real(c_float) :: divtot, divmax
      integer(c_int) :: rc
    end function
    function flutas_boundp_halo_f32(cbc, n, bc, nh_d, nh_p, dl, dzc, dzf, p) &
        bind(C, name="flutas_boundp_halo_f32") result(rc)
      import :: c_int, c_float, c_char, c_ptr
      character(kind=c_char), dimension(6) :: cbc
      integer(c_int), dimension(3) :: n
      real(c_float), dimension(6) :: bc
      integer(c_int), value :: nh_d, nh_p
      real(c_float), dimension(3) :: dl
      real(c_float), dimension(*) :: dzc, dzf
      type(c_ptr), value :: p
      integer(c_int) :: rc
    end function
    function flutas_bounduvw_f32(cbc, n, bc, nh_d, nh_u, isoutflow, dl, dzc, dzf, u, v, w) &
        bind(C, name="flutas_bounduvw_f32") result(rc)
      import :: c_int, c_float, c_char, c_ptr
      character(kind=c_char), dimension(18) :: cbc
      integer(c_int), dimension(3) :: n
      real(c_float), dimension(18) :: bc
      integer(c_int), value :: nh_d, nh_u
      integer(c_int), dimension(6) :: isoutflow
      real(c_float), dimension(3) :: dl
      real(c_float), dimension(*) :: dzc, dzf
      type(c_ptr), value :: u, v, w
      integer(c_int) :: rc
    end function
    function flutas_updt_rhs_b_f32(nx, ny, nz, c_or_f, cbc, rhsbx, rhsby, rhsbz, p) &
        bind(C, name="flutas_updt_rhs_b_f32") result(rc)
      import :: c_int, c_char, c_ptr
      integer(c_int), value :: nx, ny, nz
      character(kind=c_char), dimension(3) :: c_or_f
      character(kind=c_char), dimension(6) :: cbc
      type(c_ptr), value :: rhsbx, rhsby, rhsbz, p
      integer(c_int) :: rc
    end function
  end interface
  !
contains
  !
  ! Stops the program with the library's message when rc is not FLUTAS_OK.
  subroutine flutas_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (rc == FLUTAS_OK) return
    text = ''
    call c_f_pointer(flutas_last_error(), msg, [512])
    do i = 1, 512
      if (msg(i) == c_null_char) exit
      text(i:i) = msg(i)
    end do
    write(*, '(a,a,a,i0,a,a)') 'flutas_amd: ', where, ' failed (status ', rc, '): ', trim(text)
    error stop 1
  end subroutine flutas_check
end module flutas_capi
