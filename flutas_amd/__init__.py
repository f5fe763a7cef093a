"""flutas_amd: MI355X (gfx950) pressure-projection step of FluTAS.

The compute lives in libflutas_amd.so (hand-written HIP kernels behind a
C-ABI, include/flutas_amd.h); `flutas_amd.api` is the Python host mirror of
the reference's Fortran interface.  Importing `flutas_amd.api` loads the
library and fails loudly when it has not been built.
"""
__all__ = ["api", "build", "workloads"]
