// Host-side pieces of the dense batched LU (mhm_dense.hip) shared with the
// other translation units of libmhm_hip.so.  Not part of the C ABI.
#ifndef MHM_DENSE_INTERNAL_H
#define MHM_DENSE_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>

namespace mhm_dense_detail {

// One workspace cut into the arrays of a factorisation (lu, rs, cs, perm) and
// of a solve (y, r, colstate), each array starting on a 256-byte boundary.
struct Carve {
  double *lu = nullptr, *rs = nullptr, *cs = nullptr, *y = nullptr, *r = nullptr;
  int *perm = nullptr, *colstate = nullptr;
};

int dfail(int code, const std::string &msg);
size_t factor_bytes(size_t batch, size_t n);
size_t solve_bytes(size_t batch, size_t n, size_t nrhs);
void carve_factor(unsigned char *base, size_t batch, size_t n, Carve &c);
void carve_solve(unsigned char *base, size_t batch, size_t n, size_t nrhs, Carve &c);

// Argument checks of the one-shot factor-and-solve entry points (MHM_EINVAL
// with the message set on failure).
int check_oneshot(int device, int batch, int n, int nrhs, double rtol, double atol, const void *A, const void *rhs,
                  const void *x, const void *status);

// Enqueue the factorisation / the solve with its corrections on st.
int launch_factor(int device, int batch, int n, const double *A, const Carve &c, int *status, hipStream_t st);
int launch_solve(int batch, int n, int nrhs, const double *A, const double *rhs, double *x, double rtol, double atol,
                 const Carve &c, int *status, hipStream_t st);

// hipSetDevice(dev) for a scope, restoring the caller's device.
struct DevScope {
  int prev = -1;
  bool ok = true;
  explicit DevScope(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) ok = hipSetDevice(dev) == hipSuccess;
  }
  ~DevScope() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

}  // namespace mhm_dense_detail

#endif
