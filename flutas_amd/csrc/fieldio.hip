// Restart-file I/O in the reference's formats (src/load.f90), for fields
// that live on the GPU with their halos.
//
//   flutas_load        load(io, filename, n, fld)      src/load.f90:21-90
//   flutas_load_scalar load_scalar(io, filename, ...)  src/load.f90:92-130
//   flutas_load_ns     load_ns(io, filename, n, ...)   src/load.f90:132-204
//
// A field file is the global array, FP64, Fortran order, no halo and no
// header (what MPI-IO writes at one rank); its size is checked against the
// grid on reading, as the reference does.  The interior of a device field
// with halo h is moved through a dense device buffer and a copy kernel; a
// field in plain host memory (the explicit-shape arrays the reference's
// load receives) is read or written directly, with no device work.
#include "internal.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>

namespace flutas {
namespace {

// dense (n0,n1,n2) <-> interior of a halo-h field (n0+2h, n1+2h, n2+2h)
__global__ void __launch_bounds__(256)
    k_dense_halo(double *__restrict__ dense, double *__restrict__ halo, int n0, int n1,
                 int h, int to_halo) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= n0) return;
  const long long d = i + (long long)n0 * (j + (long long)n1 * k);
  const long long e = (i + h) + (long long)(n0 + 2 * h) * ((j + h) + (long long)(n1 + 2 * h) * (k + h));
  if (to_halo) halo[e] = dense[d];
  else dense[d] = halo[e];
}

int move_interior(double *dense_dev, double *halo_dev, const int n[3], int h, int to_halo) {
  dim3 grid((n[0] + 255) / 256, n[1], n[2]);
  hipLaunchKernelGGL(k_dense_halo, grid, dim3(256), 0, current_stream(), dense_dev, halo_dev,
                     n[0], n[1], h, to_halo);
  FL_HIP(hipGetLastError());
  FL_HIP(hipStreamSynchronize(current_stream()));
  return FLUTAS_OK;
}

int check_args(char io, const char *filename, const int n[3], int h) {
  if (io != 'r' && io != 'w') return fail(FLUTAS_EINVAL, "io must be 'r' or 'w'");
  if (!filename || !n) return fail(FLUTAS_EINVAL, "null pointer");
  if (n[0] < 1 || n[1] < 1 || n[2] < 1 || h < 0 || n[1] > 65535 || n[2] > 65535)
    return fail(FLUTAS_EINVAL, "bad grid or halo");
  return FLUTAS_OK;
}

// dense (n0,n1,n2) <-> interior of a halo-h host field
void host_move(double *dense, double *fld, const int n[3], int h, int to_field) {
  const long long s0 = n[0] + 2 * h, s1 = (long long)s0 * (n[1] + 2 * h);
  for (int k = 0; k < n[2]; ++k)
    for (int j = 0; j < n[1]; ++j) {
      double *d = dense + (long long)n[0] * (j + (long long)n[1] * k);
      double *f = fld + h + s0 * (j + h) + s1 * (k + h);
      if (to_field) std::memcpy(f, d, sizeof(double) * n[0]);
      else std::memcpy(d, f, sizeof(double) * n[0]);
    }
}

// Reads (io 'r') or writes (io 'w') `count` fields of n points each, then
// `nextra` doubles, at the given device or host fields.
int field_file(char io, const char *filename, const int n[3], const int *halo,
               double *const *fld, int count, double *extra, int nextra) {
  const size_t npts = (size_t)n[0] * n[1] * n[2];
  const size_t want = (npts * count + nextra) * sizeof(double);
  std::vector<double> host(npts);
  std::vector<int> on_dev(count, 0);
  bool any_dev = false;
  for (int f = 0; f < count; ++f) {
    FL_TRY(flutas_pointer_is_device(fld[f], &on_dev[f]));
    any_dev = any_dev || on_dev[f];
  }
  double *dense = nullptr;
  if (any_dev) FL_HIP(hipMalloc((void **)&dense, npts * sizeof(double)));
  int rc = FLUTAS_OK;
  FILE *fh = std::fopen(filename, io == 'r' ? "rb" : "wb");
  if (!fh) {
    if (dense) (void)hipFree(dense);
    return fail(FLUTAS_EINVAL, std::string("cannot open ") + filename);
  }
  if (io == 'r') {
    std::fseek(fh, 0, SEEK_END);
    const long size = std::ftell(fh);
    std::fseek(fh, 0, SEEK_SET);
    if (size < 0 || (size_t)size != want)
      rc = fail(FLUTAS_EINVAL, std::string("checkpoint file with incorrect size: ") + filename +
                                   " | expected size: " + std::to_string(want) +
                                   " | actual size: " + std::to_string(size));
  }
  for (int f = 0; f < count && rc == FLUTAS_OK; ++f) {
    if (io == 'r') {
      if (std::fread(host.data(), sizeof(double), npts, fh) != npts) {
        rc = fail(FLUTAS_EINVAL, "short read");
        break;
      }
      if (!on_dev[f]) {
        host_move(host.data(), fld[f], n, halo[f], 1);
        continue;
      }
      if (hipMemcpy(dense, host.data(), npts * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail(FLUTAS_EDEVICE, "hipMemcpy");
        break;
      }
      rc = move_interior(dense, fld[f], n, halo[f], 1);
    } else {
      if (!on_dev[f]) {
        host_move(host.data(), fld[f], n, halo[f], 0);
      } else {
        rc = move_interior(dense, fld[f], n, halo[f], 0);
        if (rc) break;
        if (hipMemcpy(host.data(), dense, npts * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
          rc = fail(FLUTAS_EDEVICE, "hipMemcpy");
          break;
        }
      }
      if (std::fwrite(host.data(), sizeof(double), npts, fh) != npts) rc = fail(FLUTAS_EINVAL, "short write");
    }
  }
  if (rc == FLUTAS_OK && nextra > 0) {
    if (io == 'r') {
      if (std::fread(extra, sizeof(double), nextra, fh) != (size_t)nextra) rc = fail(FLUTAS_EINVAL, "short read");
    } else if (std::fwrite(extra, sizeof(double), nextra, fh) != (size_t)nextra) {
      rc = fail(FLUTAS_EINVAL, "short write");
    }
  }
  std::fclose(fh);
  if (dense) (void)hipFree(dense);
  return rc;
}

}  // namespace
}  // namespace flutas

using namespace flutas;

extern "C" {

int flutas_load(char io, const char *filename, const int n[3], int halo, double *fld) {
  FL_TRY(check_args(io, filename, n, halo));
  if (!fld) return fail(FLUTAS_EINVAL, "null field");
  return field_file(io, filename, n, &halo, &fld, 1, nullptr, 0);
}

int flutas_load_ns(char io, const char *filename, const int n[3], int halo_u, double *u,
                   double *v, double *w, double *p, double *time, double *istep) {
  const int halos[4] = {halo_u, halo_u, halo_u, 1};
  return flutas_load_ns_halos(io, filename, n, halos, u, v, w, p, time, istep);
}

int flutas_load_ns_halos(char io, const char *filename, const int n[3], const int halo[4],
                         double *u, double *v, double *w, double *p, double *time,
                         double *istep) {
  if (!halo) return fail(FLUTAS_EINVAL, "null pointer");
  for (int q = 0; q < 4; ++q) FL_TRY(check_args(io, filename, n, halo[q]));
  if (!u || !v || !w || !p || !time || !istep) return fail(FLUTAS_EINVAL, "null pointer");
  double *flds[4] = {u, v, w, p};
  const int *halos = halo;
  double info[2] = {*time, *istep};
  FL_TRY(field_file(io, filename, n, halos, flds, 4, info, 2));
  if (io == 'r') {
    *time = info[0];
    *istep = info[1];
  }
  return FLUTAS_OK;
}

int flutas_load_scalar(char io, const char *filename, double *time, int *istep, double *dto) {
  if ((io != 'r' && io != 'w') || !filename || !time || !istep || !dto)
    return fail(FLUTAS_EINVAL, "bad load_scalar arguments");
  if (io == 'r') {
    FILE *fh = std::fopen(filename, "r");
    if (!fh) return fail(FLUTAS_EINVAL, std::string("the restarting file does not exist: ") + filename);
    // list-directed read of "time, dto, istep" (src/load.f90:120)
    const int got = std::fscanf(fh, "%lf%*[ ,\t\n]%lf%*[ ,\t\n]%d", time, dto, istep);
    std::fclose(fh);
    if (got != 3) return fail(FLUTAS_EINVAL, std::string("cannot parse ") + filename);
    return FLUTAS_OK;
  }
  FILE *fh = std::fopen(filename, "w");
  if (!fh) return fail(FLUTAS_EINVAL, std::string("cannot open ") + filename);
  // Fortran '(2E15.7, 1I9.8)': two E15.7 fields, then I9.8.  Ew.d prints
  // 0.dddddddE+xx, and drops the E for exponents of three digits
  // (0.ddddddd+xxx).  The seven digits are those of one correctly rounded
  // %.6E conversion (d.dddddd x 10^e), moved behind the point in the text
  // (exponent e + 1), so no second rounding can change them.
  char buf[2][32];
  for (int q = 0; q < 2; ++q) {
    const double x = (q == 0) ? *time : *dto;
    char digits[8] = "0000000";
    int ex = 0;
    if (x != 0.0) {
      char t[40];
      std::snprintf(t, sizeof t, "%.6E", std::fabs(x));  // "d.ddddddE+xx"
      digits[0] = t[0];
      std::memcpy(digits + 1, t + 2, 6);
      std::sscanf(std::strchr(t, 'E') + 1, "%d", &ex);
      ex += 1;
    }
    const int ax = ex < 0 ? -ex : ex;
    if (ax <= 99)
      std::snprintf(buf[q], sizeof buf[q], "%s0.%sE%c%02d", x < 0 ? "-" : "", digits,
                    ex < 0 ? '-' : '+', ax);
    else
      std::snprintf(buf[q], sizeof buf[q], "%s0.%s%c%03d", x < 0 ? "-" : "", digits,
                    ex < 0 ? '-' : '+', ax);
  }
  std::fprintf(fh, "%15s%15s%9.8d\n", buf[0], buf[1], *istep);
  std::fclose(fh);
  return FLUTAS_OK;
}

}  // extern "C"
