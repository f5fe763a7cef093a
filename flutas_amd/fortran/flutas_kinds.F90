! Real kind of the drop-in modules: FluTAS's rp (src/types.f90:6-12), FP32
! when the build defines _SINGLE_PRECISION, FP64 otherwise.  rp_bytes is
! the size of one value.
module flutas_kinds
  use, intrinsic :: iso_c_binding, only: c_float, c_double, c_size_t
  implicit none
  private
  public :: rp, rp_bytes
#if defined(_SINGLE_PRECISION)
  integer, parameter :: rp = c_float
#else
  integer, parameter :: rp = c_double
#endif
  integer(c_size_t), parameter :: rp_bytes = storage_size(1._rp)/8
end module flutas_kinds
