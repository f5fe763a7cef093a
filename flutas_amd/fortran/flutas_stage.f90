! Device staging for the drop-in modules.
!
! The field kernels need device-accessible memory.  A field that already is
! a HIP device or managed allocation is passed through unchanged; a plain
! host array is copied into a device buffer kept per slot (grown on demand,
! reused across calls) and, for outputs, copied back after the kernel.  Each
! call site uses its own slots for arrays that are live at the same time.
module flutas_stage
  use, intrinsic :: iso_c_binding
  use flutas_capi
  implicit none
  private
  public :: stage_in, stage_out, stage_release
  integer, parameter :: NSLOT = 8
  type(c_ptr), save :: dbuf(NSLOT) = c_null_ptr
  integer(c_size_t), save :: dsize(NSLOT) = 0
contains
  !
  ! Device address holding the `bytes` bytes at `host`: host itself when it
  ! is device memory, else slot `slot`'s buffer (filled from host if `copy`).
  function stage_in(slot, host, bytes, copy) result(d)
    integer, intent(in) :: slot
    type(c_ptr), intent(in) :: host
    integer(c_size_t), intent(in) :: bytes
    logical, intent(in) :: copy
    type(c_ptr) :: d
    integer(c_int) :: isdev
    if (slot < 1 .or. slot > NSLOT) error stop 'flutas_stage: bad slot'
    call flutas_check(flutas_pointer_is_device(host, isdev), 'pointer query')
    if (isdev == 1) then
      d = host
      return
    end if
    if (dsize(slot) < bytes) then
      if (c_associated(dbuf(slot))) call flutas_check(flutas_device_free(dbuf(slot)), 'device free')
      dbuf(slot) = c_null_ptr
      call flutas_check(flutas_device_alloc(bytes, dbuf(slot)), 'device alloc')
      dsize(slot) = bytes
    end if
    d = dbuf(slot)
    ! the blocking copy runs on the null stream; a kernel still reading this
    ! slot from an earlier call may be queued on the library stream (which
    ! flutas_set_stream may have made non-blocking), so wait for it first
    if (copy) then
      call flutas_check(flutas_synchronize(), 'synchronize')
      call flutas_check(flutas_memcpy_h2d(d, host, bytes), 'host to device copy')
    end if
  end function stage_in
  !
  ! Copies a staged output back to the host array (no-op when d is host).
  subroutine stage_out(host, d, bytes)
    type(c_ptr), intent(in) :: host, d
    integer(c_size_t), intent(in) :: bytes
    if (c_associated(host, d)) return
    call flutas_check(flutas_synchronize(), 'synchronize')
    call flutas_check(flutas_memcpy_d2h(host, d, bytes), 'device to host copy')
  end subroutine stage_out
  !
  ! Frees every staging buffer.
  subroutine stage_release()
    integer :: s
    do s = 1, NSLOT
      if (c_associated(dbuf(s))) call flutas_check(flutas_device_free(dbuf(s)), 'device free')
      dbuf(s) = c_null_ptr
      dsize(s) = 0
    end do
  end subroutine stage_release
end module flutas_stage
