! Drop-in restart-file module: the reference's mod_load routines
! (src/load.f90) with their signatures, backed by flutas_load*.
!
!   load          src/load.f90:21-90    one FP64 field, raw global array in
!                                       Fortran order (no header)
!   load_scalar   src/load.f90:92-130   time, dto, istep as text
!                                       ('(2E15.7, 1I9.8)' on writing)
!   load_ns       src/load.f90:132-204  u, v, w, p, then time and istep
!   cmpt_it_chkpt src/load.f90:217-268  the last checkpoint number listed in
!                                       restart_dir//'restart_checkpoints.out'
!
! The field arguments are the explicit-shape n(1) x n(2) x n(3) arrays the
! reference receives (callers pass u(1:n(1),1:n(2),1:n(3))); in host memory
! they are read and written with no device work, in device or managed
! memory through the device path.  A missing or wrongly sized file stops
! with the library's message, as the reference aborts.
module mod_load
  use, intrinsic :: iso_c_binding
  use flutas_capi
  implicit none
  private
  public :: load, load_ns, load_scalar, cmpt_it_chkpt
  integer, parameter :: rp = c_double
contains
  subroutine load(io, filename, n, fld)
    character(len=1), intent(in) :: io
    character(len=*), intent(in) :: filename
    integer, intent(in), dimension(3) :: n
    real(rp), intent(inout), dimension(n(1), n(2), n(3)), target :: fld
    integer(c_int) :: nc(3)
    nc = int(n, c_int)
    call flutas_check(flutas_load(io, trim(filename)//c_null_char, nc, 0_c_int, c_loc(fld)), &
                      'load '//trim(filename))
  end subroutine load

  subroutine load_ns(io, filename, n, u, v, w, p, time, istep)
    character(len=1), intent(in) :: io
    character(len=*), intent(in) :: filename
    integer, intent(in), dimension(3) :: n
    real(rp), intent(inout), dimension(n(1), n(2), n(3)), target :: u, v, w, p
    real(rp), intent(inout) :: time, istep
    integer(c_int) :: nc(3), halo(4)
    real(c_double) :: t, s
    nc = int(n, c_int)
    halo = 0_c_int
    t = time
    s = istep
    call flutas_check(flutas_load_ns_halos(io, trim(filename)//c_null_char, nc, halo, c_loc(u), &
                                           c_loc(v), c_loc(w), c_loc(p), t, s), &
                      'load_ns '//trim(filename))
    time = t
    istep = s
  end subroutine load_ns

  subroutine load_scalar(io, filename, time, istep, dto)
    character(len=1), intent(in) :: io
    character(len=*), intent(in) :: filename
    integer, intent(inout) :: istep
    real(rp), intent(inout) :: time, dto
    real(c_double) :: t, d
    integer(c_int) :: s
    t = time
    d = dto
    s = int(istep, c_int)
    call flutas_check(flutas_load_scalar(io, trim(filename)//c_null_char, t, s, d), &
                      'load_scalar '//trim(filename))
    time = t
    dto = d
    istep = int(s)
  end subroutine load_scalar

  ! The checkpoint to restart from: the third column of the last line of
  ! restart_dir//'restart_checkpoints.out' (src/load.f90:217-268, read here
  ! without the shell line count the reference runs first).
  subroutine cmpt_it_chkpt(restart_dir, it_chkpt)
    character(len=*), intent(in) :: restart_dir
    integer, intent(out) :: it_chkpt
    logical :: is_file
    integer :: u, ios
    real(rp) :: c1, c2, c3, last
    logical :: any_row
    inquire(file=trim(restart_dir)//'restart_checkpoints.out', exist=is_file)
    if (.not. is_file) then
      print *, 'The file <<restart_checkpoints.out>> is absent. Please, provide it'
      print *, 'In this case, the restarting files are probably missing. Check dns.in'
      error stop 1
    end if
    open(newunit=u, file=trim(restart_dir)//'restart_checkpoints.out', status='old', &
         action='read')
    any_row = .false.
    last = 0.0_rp
    do
      read(u, *, iostat=ios) c1, c2, c3
      if (ios /= 0) exit
      last = c3
      any_row = .true.
    end do
    close(u)
    if (.not. any_row) then
      print *, 'The file <<restart_checkpoints.out>> holds no checkpoint'
      error stop 1
    end if
    it_chkpt = int(last)
  end subroutine cmpt_it_chkpt
end module mod_load
