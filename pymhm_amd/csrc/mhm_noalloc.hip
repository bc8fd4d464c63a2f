// mhm_dense_factor_solve without device allocation: the workspace of the
// dense batched LU (mhm_dense.hip) comes from a fixed per-device array instead
// of hipMalloc, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The array `scratch` exists once on every GPU the library runs on and stays
// there until the process ends.  In host code its name is only a stand-in: the
// device address comes from hipGetSymbolAddress, taken once per device with
// that device current, and only that address reaches kernels.  Calls on one
// device share the array and keep the last call's contents; every array the
// kernels read before writing (status, column states) is cleared by a fill
// kernel at the start of each call.  Calls that overlap on one GPU race.

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/mhm_hip.h"
#include "mhm_dense_internal.h"

using namespace mhm_dense_detail;

constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
unsigned char *g_scratch[kMaxDevices] = {};

// Device address of `scratch` on `device` (the caller made it current).
int scratch_address(int device, unsigned char **out) {
  if (device < 0 || device >= kMaxDevices) return dfail(MHM_EINVAL, "device index outside [0, 64)");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return dfail(MHM_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = (unsigned char *)p;
  }
  *out = g_scratch[device];
  return MHM_OK;
}

}  // namespace

extern "C" {

int mhm_dense_factor_solve_noalloc(int32_t device, int32_t batch, int32_t n, int32_t nrhs, const double *d_matrices,
                                   const double *d_rhs, double *d_x, double rtol, double atol, int32_t *d_status,
                                   void *stream) {
  if (int rc = check_oneshot(device, batch, n, nrhs, rtol, atol, d_matrices, d_rhs, d_x, d_status)) return rc;
  const size_t fb = factor_bytes(batch, n), sb = solve_bytes(batch, n, nrhs);
  if (fb + sb > kScratchBytes)
    return dfail(MHM_EINVAL, "workspace of " + std::to_string(fb + sb) + " bytes exceeds the " +
                                 std::to_string(kScratchBytes) + "-byte no-allocation scratch");
  DevScope scope(device);
  if (!scope.ok) return dfail(MHM_EDEVICE, "hipSetDevice failed");
  unsigned char *work = nullptr;
  if (int rc = scratch_address(device, &work)) return rc;
  hipStream_t st = (hipStream_t)stream;
  Carve c;
  carve_factor(work, batch, n, c);
  carve_solve(work + fb, batch, n, nrhs, c);
  if (int rc = launch_factor(device, batch, n, d_matrices, c, d_status, st)) return rc;
  return launch_solve(batch, n, nrhs, d_matrices, d_rhs, d_x, rtol, atol, c, d_status, st);
}

int mhm_dense_scratch_bytes(int32_t batch, int32_t n, int32_t nrhs, int64_t *bytes, int64_t *limit) {
  if (batch <= 0 || n <= 0 || nrhs <= 0) return dfail(MHM_EINVAL, "batch, n and nrhs must be positive");
  if (bytes) *bytes = (int64_t)(factor_bytes(batch, n) + solve_bytes(batch, n, nrhs));
  if (limit) *limit = (int64_t)kScratchBytes;
  return MHM_OK;
}

}  // extern "C"
