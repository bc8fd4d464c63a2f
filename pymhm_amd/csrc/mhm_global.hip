// Global hybrid system on the device: restarted GMRES, right-preconditioned
// by weighted additive Schwarz over the contribution cells.
//
// The assembled global matrix (trace multipliers, then coarse amplitudes) is
// the sum of every cell's dense (w x w) condensed block over its global
// indices.  The preconditioner inverts, per cell, the assembled matrix
// restricted to the cell's indices (a small saddle block: the cell's trace
// rows with their neighbours' contributions, and its coarse row), and applies
//   z = W sum_c R_c^T inv(R_c S R_c^T) R_c r,   W = diag(1 / cells per row),
// which is one matrix-vector product per cell plus a fixed-order gather.
// The Krylov solve is GMRES with classical Gram-Schmidt applied twice, the
// basis kept whole in HBM (restart = maxiter by default), a binary64
// Hessenberg least-squares problem on the host, and a double-double
// residual of the original equations at every restart; a solve ends when
// ||b - S x|| <= rtol ||b||, the acceptance of linalg/linear.py:132-152.
//
// Every sum runs in a fixed order (per-row sequential sums, per-block tree
// reductions, blocks summed in order), so a solve's bits do not depend on
// timing.  Replaces the global stage of solve_hybrid_system
// (src/pymhm/core/system.py:355-528) and SaddleBlockSolver.solve
// (src/pymhm/linalg/block.py:176-200) for the assembled CSC system.

#include "mhm_hip.h"
#include "mhm_dense_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using mhm_dense_detail::DevScope;
using mhm_dense_detail::dfail;

#define GLB_MAXW 64      // widest cell (indices per contribution)
#define DOT_THREADS 256
#define DOT_ROWS 2048    // rows per block of a dot product

namespace {

__device__ __forceinline__ void dd_fma_acc(double &hi, double &lo, double a, double b) {
  // (hi, lo) += a * b in double-double
  const double p = a * b;
  const double pe = __builtin_fma(a, b, -p);
  const double s = hi + p;
  const double bv = s - hi;
  const double se = (hi - (s - bv)) + (p - bv);
  hi = s;
  lo += se + pe;
}

__global__ void __launch_bounds__(256) k_glb_perm(long nnz, const int *__restrict__ perm,
                                                  const double *__restrict__ csc, double *__restrict__ vals) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nnz) vals[k] = csc[perm[k]];
}

// Gauss-Jordan with partial pivoting on [S_c | I] in LDS (w rows, 2w
// columns; thread t owns column t).  Writes inv(S_c) transposed, so the
// apply kernel reads it with consecutive threads on consecutive entries.
__global__ void __launch_bounds__(2 * GLB_MAXW)
    k_glb_invert(const int *__restrict__ cellptr, const long *__restrict__ blkoff, const int *__restrict__ blkslot,
                 const double *__restrict__ vals, double *__restrict__ binv, int *__restrict__ bstatus) {
  extern __shared__ double aug[];  // w x 2w
  __shared__ double fcol[GLB_MAXW];
  __shared__ int spiv;
  __shared__ double smax;
  const int c = blockIdx.x;
  const int t = threadIdx.x;
  const int w = cellptr[c + 1] - cellptr[c];
  const int w2 = 2 * w;
  const long off = blkoff[c];
  double amax = 0.0;
  for (int e = t; e < w * w; e += blockDim.x) {
    const int i = e / w, j = e - i * w;
    const int s = blkslot[off + e];
    const double v = s >= 0 ? vals[s] : 0.0;
    aug[i * w2 + j] = v;
    aug[i * w2 + w + j] = i == j ? 1.0 : 0.0;
    amax = fmax(amax, fabs(v));
  }
  // block maximum of |S_c| (non-finite entries make it NaN/inf)
  for (int o = 32; o > 0; o >>= 1) amax = fmax(amax, __shfl_xor(amax, o));
  if (t == 0) smax = 0.0;
  __syncthreads();
  if ((t & 63) == 0) atomicMax((unsigned long long *)&smax, __double_as_longlong(amax));  // non-negative doubles order as integers
  __syncthreads();
  const double tiny = w * 2.220446049250313e-16 * smax;
  bool bad = !(smax < INFINITY) || smax == 0.0;
  for (int k = 0; k < w && !bad; ++k) {
    if (t < 64) {
      double v = (t >= k && t < w) ? fabs(aug[t * w2 + k]) : -1.0;
      int r = t;
      for (int o = 32; o > 0; o >>= 1) {
        const double v2 = __shfl_xor(v, o);
        const int r2 = __shfl_xor(r, o);
        if (v2 > v || (v2 == v && r2 < r)) {
          v = v2;
          r = r2;
        }
      }
      if (t == 0) spiv = (v > tiny && v < INFINITY) ? r : -1;
    }
    __syncthreads();
    const int p = spiv;
    if (p < 0) {
      bad = true;
      break;
    }
    if (p != k && t < w2) {
      const double a = aug[k * w2 + t];
      aug[k * w2 + t] = aug[p * w2 + t];
      aug[p * w2 + t] = a;
    }
    __syncthreads();
    if (t < w) fcol[t] = aug[t * w2 + k];
    __syncthreads();
    const double piv = fcol[k];
    if (t < w2) {
      const double rk = aug[k * w2 + t] / piv;
      aug[k * w2 + t] = rk;
      for (int i = 0; i < w; ++i)
        if (i != k) aug[i * w2 + t] -= fcol[i] * rk;
    }
    __syncthreads();
  }
  if (t == 0) bstatus[c] = bad ? 1 : 0;
  for (int e = t; e < w * w; e += blockDim.x) {
    const int j = e / w, i = e - j * w;  // binv[j * w + i] = inv[i][j]
    binv[off + e] = bad ? (i == j ? 1.0 : 0.0) : aug[i * w2 + w + j];
  }
}

// tloc[cell][i] = sum_j inv(S_c)[i][j] r[idx_c[j]]
__global__ void __launch_bounds__(GLB_MAXW)
    k_glb_apply(const int *__restrict__ cellptr, const int *__restrict__ cellidx, const long *__restrict__ blkoff,
                const double *__restrict__ binv, const double *__restrict__ r, double *__restrict__ tloc) {
  __shared__ double rl[GLB_MAXW];
  const int c = blockIdx.x;
  const int t = threadIdx.x;
  const int o = cellptr[c], w = cellptr[c + 1] - o;
  if (t < w) rl[t] = r[cellidx[o + t]];
  __syncthreads();
  if (t >= w) return;
  const double *bi = binv + blkoff[c];
  double s = 0.0;
  for (int j = 0; j < w; ++j) s += bi[j * w + t] * rl[j];
  tloc[o + t] = s;
}

// z[g] = wgt[g] * sum of the cells' local values at row g, in cell order
__global__ void __launch_bounds__(256) k_glb_gather(int n, const int *__restrict__ gptr, const int *__restrict__ gpos,
                                                    const double *__restrict__ wgt, const double *__restrict__ tloc,
                                                    double *__restrict__ z) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  double s = 0.0;
  for (int p = gptr[g]; p < gptr[g + 1]; ++p) s += tloc[gpos[p]];
  z[g] = wgt[g] * s;
}

// y = S x (binary64, row order)
__global__ void __launch_bounds__(256) k_glb_spmv(int n, const int *__restrict__ rowptr, const int *__restrict__ col,
                                                  const double *__restrict__ vals, const double *__restrict__ x,
                                                  double *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) s += vals[k] * x[col[k]];
  y[i] = s;
}

// r = b - S x, each row in double-double, rounded once
__global__ void __launch_bounds__(256) k_glb_residual(int n, const int *__restrict__ rowptr,
                                                      const int *__restrict__ col, const double *__restrict__ vals,
                                                      const double *__restrict__ x, const double *__restrict__ b,
                                                      double *__restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double hi = b[i], lo = 0.0;
  for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) dd_fma_acc(hi, lo, -vals[k], x[col[k]]);
  r[i] = hi + lo;
}

// partial[j][blk] = sum over the block's rows of V[j][i] * w[i]
__global__ void __launch_bounds__(DOT_THREADS) k_glb_dot_partial(int n, int nblk, const double *__restrict__ V, long ld,
                                                                 const double *__restrict__ w,
                                                                 double *__restrict__ partial) {
  __shared__ double red[DOT_THREADS];
  const int j = blockIdx.y;
  const int t = threadIdx.x;
  const long base = (long)blockIdx.x * DOT_ROWS;
  const double *v = V + (long)j * ld;
  double s = 0.0;
  for (int q = t; q < DOT_ROWS; q += DOT_THREADS) {
    const long i = base + q;
    if (i < n) s += v[i] * w[i];
  }
  red[t] = s;
  __syncthreads();
  for (int h = DOT_THREADS / 2; h > 0; h >>= 1) {
    if (t < h) red[t] += red[t + h];
    __syncthreads();
  }
  if (t == 0) partial[(long)j * nblk + blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_glb_dot_final(int ncols, int nblk, const double *__restrict__ partial,
                                                       double *__restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncols) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += partial[(long)j * nblk + b];
  out[j] = s;
}

// w -= sum_{j < k} h[j] V[j]
__global__ void __launch_bounds__(256) k_glb_orth(int n, int k, const double *__restrict__ V, long ld,
                                                  const double *__restrict__ h, double *__restrict__ w) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = w[i];
  for (int j = 0; j < k; ++j) s -= h[j] * V[(long)j * ld + i];
  w[i] = s;
}

__global__ void __launch_bounds__(256) k_glb_scale(int n, double a, const double *__restrict__ x,
                                                   double *__restrict__ y) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i];
}

// x += sum_{j < k} y[j] Z[j]
__global__ void __launch_bounds__(256) k_glb_combine(int n, int k, const double *__restrict__ Z, long ld,
                                                     const double *__restrict__ y, double *__restrict__ x) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < k; ++j) s += y[j] * Z[(long)j * ld + i];
  x[i] += s;
}

__global__ void __launch_bounds__(256) k_glb_fill(long n, double v, double *__restrict__ x) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

inline unsigned nb256(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

struct mhm_global {
  int device = 0;
  int n = 0, ncells = 0, nblk = 0;
  long nnz = 0, nloc = 0, nblkent = 0;
  std::vector<void *> owned;
  int *rowptr = nullptr, *col = nullptr, *perm = nullptr, *cellptr = nullptr, *cellidx = nullptr;
  int *blkslot = nullptr, *bstatus = nullptr, *gptr = nullptr, *gpos = nullptr;
  long *blkoff = nullptr;
  double *vals = nullptr, *binv = nullptr, *wgt = nullptr, *tloc = nullptr;
  bool ready = false;
  // Krylov workspace (grown on demand)
  int kcap = 0;
  double *V = nullptr, *Z = nullptr, *w = nullptr, *r = nullptr, *partial = nullptr, *dh = nullptr;
  double *hpin = nullptr;  // pinned host: 2 (kcap + 2) + 1 values
  std::vector<int> hstatus;
};

static int glb_alloc(mhm_global *g, void **p, size_t bytes) {
  if (bytes == 0) bytes = 8;
  if (hipMalloc(p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return dfail(MHM_ENOMEM, "global solver: device allocation of " + std::to_string(bytes) + " bytes failed");
  }
  g->owned.push_back(*p);
  return MHM_OK;
}

template <typename T>
static int glb_upload(mhm_global *g, const std::vector<T> &h, T **d) {
  if (int rc = glb_alloc(g, (void **)d, h.size() * sizeof(T))) return rc;
  if (!h.empty() && hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    return dfail(MHM_EDEVICE, "global solver: host-to-device copy failed");
  return MHM_OK;
}

static void glb_release(mhm_global *g) {
  for (void *p : g->owned) (void)hipFree(p);
  g->owned.clear();
  if (g->hpin) (void)hipHostFree(g->hpin);
  g->hpin = nullptr;
}

extern "C" int mhm_global_create(int32_t device, int32_t size, int64_t nnz, const int32_t *csc_indptr,
                                 const int32_t *csc_indices, int32_t ncells, const int64_t *cell_ptr,
                                 const int64_t *cell_idx, mhm_global **out) {
  if (!out) return dfail(MHM_EINVAL, "out is NULL");
  *out = nullptr;
  if (size <= 0 || nnz <= 0 || nnz >= INT32_MAX || ncells <= 0 || !csc_indptr || !csc_indices || !cell_ptr ||
      !cell_idx)
    return dfail(MHM_EINVAL, "global solver: size, nnz (< 2^31) and ncells must be positive, arrays non-NULL");
  const int n = size;
  if (csc_indptr[0] != 0 || csc_indptr[n] != nnz)
    return dfail(MHM_EINVAL, "global solver: csc_indptr must run from 0 to nnz");
  for (int j = 0; j < n; ++j) {
    if (csc_indptr[j + 1] < csc_indptr[j]) return dfail(MHM_EINVAL, "global solver: csc_indptr must not decrease");
    for (int k = csc_indptr[j]; k < csc_indptr[j + 1]; ++k) {
      const int i = csc_indices[k];
      if (i < 0 || i >= n || (k > csc_indptr[j] && i <= csc_indices[k - 1]))
        return dfail(MHM_EINVAL, "global solver: row indices must be in range and increase within a column");
    }
  }
  if (cell_ptr[0] != 0) return dfail(MHM_EINVAL, "global solver: cell_ptr must start at 0");
  for (int c = 0; c < ncells; ++c) {
    const int64_t w = cell_ptr[c + 1] - cell_ptr[c];
    if (w < 1 || w > GLB_MAXW)
      return dfail(MHM_EINVAL, "global solver: every cell needs 1.." + std::to_string(GLB_MAXW) + " indices");
  }
  const long nloc = (long)cell_ptr[ncells];
  if (nloc >= INT32_MAX) return dfail(MHM_EINVAL, "global solver: too many cell indices");
  // CSR of the same pattern, rows' columns ascending, and the CSC slot of each entry
  std::vector<int> rowptr(n + 1, 0), col(nnz), perm(nnz);
  for (long k = 0; k < nnz; ++k) ++rowptr[csc_indices[k] + 1];
  for (int i = 0; i < n; ++i) rowptr[i + 1] += rowptr[i];
  {
    std::vector<int> next(rowptr.begin(), rowptr.end() - 1);
    for (int j = 0; j < n; ++j)
      for (int k = csc_indptr[j]; k < csc_indptr[j + 1]; ++k) {
        const int p = next[csc_indices[k]]++;
        col[p] = j;
        perm[p] = k;
      }
  }
  for (int i = 0; i < n; ++i)
    if (rowptr[i + 1] == rowptr[i]) return dfail(MHM_EINVAL, "global solver: an empty matrix row");
  // per-cell blocks: the CSR slot of every (i, j), and the gather map
  std::vector<int> cellptr(ncells + 1), cellidx(nloc), cnt(n, 0);
  std::vector<long> blkoff(ncells + 1, 0);
  for (int c = 0; c < ncells; ++c) {
    cellptr[c] = (int)cell_ptr[c];
    const long w = cell_ptr[c + 1] - cell_ptr[c];
    blkoff[c + 1] = blkoff[c] + w * w;
    for (long a = cell_ptr[c]; a < cell_ptr[c + 1]; ++a) {
      const int64_t g = cell_idx[a];
      if (g < 0 || g >= n) return dfail(MHM_EINVAL, "global solver: cell index out of range");
      for (long b = cell_ptr[c]; b < a; ++b)
        if (cell_idx[b] == g) return dfail(MHM_EINVAL, "global solver: repeated index within a cell");
      cellidx[a] = (int)g;
      ++cnt[g];
    }
  }
  cellptr[ncells] = (int)nloc;
  std::vector<int> blkslot(blkoff[ncells]);
  for (int c = 0; c < ncells; ++c) {
    const int o = cellptr[c], w = cellptr[c + 1] - o;
    for (int i = 0; i < w; ++i) {
      const int row = cellidx[o + i];
      const int *b = col.data() + rowptr[row], *e = col.data() + rowptr[row + 1];
      for (int j = 0; j < w; ++j) {
        const int *f = std::lower_bound(b, e, cellidx[o + j]);
        blkslot[blkoff[c] + (long)i * w + j] = (f != e && *f == cellidx[o + j]) ? (int)(f - col.data()) : -1;
      }
    }
  }
  std::vector<int> gptr(n + 1, 0), gpos(nloc);
  std::vector<double> wgt(n);
  for (int g = 0; g < n; ++g) {
    if (cnt[g] == 0) return dfail(MHM_EINVAL, "global solver: a row belongs to no cell");
    gptr[g + 1] = gptr[g] + cnt[g];
    wgt[g] = 1.0 / cnt[g];
  }
  {
    std::vector<int> next(gptr.begin(), gptr.end() - 1);
    for (long a = 0; a < nloc; ++a) gpos[next[cellidx[a]]++] = (int)a;  // cells in order
  }
  DevScope scope(device);
  if (!scope.ok) return dfail(MHM_EDEVICE, "global solver: hipSetDevice failed");
  mhm_global *g = new mhm_global();
  g->device = device;
  g->n = n;
  g->nnz = nnz;
  g->ncells = ncells;
  g->nloc = nloc;
  g->nblkent = blkoff[ncells];
  g->nblk = (n + DOT_ROWS - 1) / DOT_ROWS;
  g->hstatus.resize(ncells);
  int rc;
  if ((rc = glb_upload(g, rowptr, &g->rowptr)) || (rc = glb_upload(g, col, &g->col)) ||
      (rc = glb_upload(g, perm, &g->perm)) || (rc = glb_upload(g, cellptr, &g->cellptr)) ||
      (rc = glb_upload(g, cellidx, &g->cellidx)) || (rc = glb_upload(g, blkoff, &g->blkoff)) ||
      (rc = glb_upload(g, blkslot, &g->blkslot)) || (rc = glb_upload(g, gptr, &g->gptr)) ||
      (rc = glb_upload(g, gpos, &g->gpos)) || (rc = glb_upload(g, wgt, &g->wgt)) ||
      (rc = glb_alloc(g, (void **)&g->vals, nnz * sizeof(double))) ||
      (rc = glb_alloc(g, (void **)&g->binv, g->nblkent * sizeof(double))) ||
      (rc = glb_alloc(g, (void **)&g->tloc, nloc * sizeof(double))) ||
      (rc = glb_alloc(g, (void **)&g->bstatus, ncells * sizeof(int)))) {
    glb_release(g);
    delete g;
    return rc;
  }
  *out = g;
  return MHM_OK;
}

extern "C" int mhm_global_set_values(mhm_global *g, const double *d_csc_values, int32_t *bad_cell, void *stream) {
  if (!g || !d_csc_values) return dfail(MHM_EINVAL, "global solver/values is NULL");
  DevScope scope(g->device);
  hipStream_t st = (hipStream_t)stream;
  g->ready = false;
  hipLaunchKernelGGL(k_glb_perm, dim3(nb256(g->nnz)), dim3(256), 0, st, g->nnz, (const int *)g->perm, d_csc_values,
                     g->vals);
  const size_t shm = (size_t)GLB_MAXW * 2 * GLB_MAXW * sizeof(double);
  if (hipFuncSetAttribute((const void *)k_glb_invert, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm) !=
      hipSuccess)
    return dfail(MHM_EDEVICE, "global solver: LDS attribute rejected");
  hipLaunchKernelGGL(k_glb_invert, dim3(g->ncells), dim3(2 * GLB_MAXW), shm, st, (const int *)g->cellptr,
                     (const long *)g->blkoff, (const int *)g->blkslot, (const double *)g->vals, g->binv, g->bstatus);
  if (hipGetLastError() != hipSuccess) return dfail(MHM_EDEVICE, "global solver: preconditioner launch failed");
  if (hipMemcpyAsync(g->hstatus.data(), g->bstatus, g->ncells * sizeof(int), hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return dfail(MHM_EDEVICE, "global solver: status copy failed");
  for (int c = 0; c < g->ncells; ++c)
    if (g->hstatus[c]) {
      if (bad_cell) *bad_cell = c;
      return dfail(MHM_ESINGULAR, "global solver: the assembled block of cell " + std::to_string(c) +
                                      " is singular or not finite");
    }
  if (bad_cell) *bad_cell = -1;
  g->ready = true;
  return MHM_OK;
}

static int glb_workspace(mhm_global *g, int k) {
  if (k <= g->kcap) return MHM_OK;
  // the old workspace is released (the stream was synchronised by the caller)
  std::vector<void *> keep;
  for (void *p : g->owned)
    if (p != g->V && p != g->Z && p != g->w && p != g->r && p != g->partial && p != g->dh) keep.push_back(p);
    else (void)hipFree(p);
  g->owned = keep;
  if (g->hpin) (void)hipHostFree(g->hpin);
  g->hpin = nullptr;
  g->V = g->Z = g->w = g->r = g->partial = g->dh = nullptr;
  g->kcap = 0;
  const size_t n = g->n;
  int rc;
  if ((rc = glb_alloc(g, (void **)&g->V, (size_t)(k + 1) * n * sizeof(double))) ||
      (rc = glb_alloc(g, (void **)&g->Z, (size_t)k * n * sizeof(double))) ||
      (rc = glb_alloc(g, (void **)&g->w, n * sizeof(double))) ||
      (rc = glb_alloc(g, (void **)&g->r, n * sizeof(double))) ||
      (rc = glb_alloc(g, (void **)&g->partial, (size_t)(k + 2) * g->nblk * sizeof(double))) ||
      (rc = glb_alloc(g, (void **)&g->dh, (size_t)3 * (k + 2) * sizeof(double))))
    return rc;
  if (hipHostMalloc((void **)&g->hpin, (size_t)3 * (k + 2) * sizeof(double)) != hipSuccess)
    return dfail(MHM_ENOMEM, "global solver: pinned host allocation failed");
  g->kcap = k;
  return MHM_OK;
}

// dots[0..k) = V[0..k)^T x into device array out
static void glb_dots(mhm_global *g, const double *V, int k, const double *x, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_glb_dot_partial, dim3(g->nblk, k), dim3(DOT_THREADS), 0, st, g->n, g->nblk, V, (long)g->n, x,
                     g->partial);
  hipLaunchKernelGGL(k_glb_dot_final, dim3((k + 255) / 256), dim3(256), 0, st, k, g->nblk,
                     (const double *)g->partial, out);
}

static void glb_precond(mhm_global *g, const double *r, double *z, hipStream_t st) {
  hipLaunchKernelGGL(k_glb_apply, dim3(g->ncells), dim3(GLB_MAXW), 0, st, (const int *)g->cellptr,
                     (const int *)g->cellidx, (const long *)g->blkoff, (const double *)g->binv, r, g->tloc);
  hipLaunchKernelGGL(k_glb_gather, dim3(nb256(g->n)), dim3(256), 0, st, g->n, (const int *)g->gptr,
                     (const int *)g->gpos, (const double *)g->wgt, (const double *)g->tloc, z);
}

static int glb_fetch(double *dst, const double *src, size_t count, hipStream_t st) {
  if (hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return dfail(MHM_EDEVICE, "global solver: device-to-host copy failed");
  return MHM_OK;
}

extern "C" int mhm_global_solve(mhm_global *g, const double *d_rhs, double *d_x, double rtol, int32_t maxiter,
                                int32_t restart, int32_t *iterations, double *residual, void *stream) {
  if (!g || !d_rhs || !d_x) return dfail(MHM_EINVAL, "global solver/rhs/x is NULL");
  if (!g->ready) return dfail(MHM_EINVAL, "global solver: values not set (mhm_global_set_values)");
  if (!(rtol > 0.0) || !(rtol < 1.0) || maxiter < 1 || restart < 1)
    return dfail(MHM_EINVAL, "global solver: 0 < rtol < 1, maxiter >= 1 and restart >= 1 required");
  DevScope scope(g->device);
  hipStream_t st = (hipStream_t)stream;
  const int R = std::min(restart, maxiter);
  if (hipStreamSynchronize(st) != hipSuccess) return dfail(MHM_EDEVICE, "global solver: stream failed");
  if (int rc = glb_workspace(g, R)) return rc;
  const int n = g->n;
  const unsigned nb = nb256(n);
  double *hp = g->hpin;
  double *dh = g->dh, *dh2 = g->dh + (R + 2), *dy = g->dh + 2 * (R + 2);
  // ||b||
  glb_dots(g, d_rhs, 1, d_rhs, dh, st);
  if (int rc = glb_fetch(hp, dh, 1, st)) return rc;
  const double bnorm = std::sqrt(hp[0]);
  if (!(bnorm < INFINITY)) return dfail(MHM_EINVAL, "global solver: right-hand side not finite");
  hipLaunchKernelGGL(k_glb_fill, dim3(nb), dim3(256), 0, st, (long)n, 0.0, d_x);
  if (iterations) *iterations = 0;
  if (residual) *residual = 0.0;
  if (bnorm == 0.0) return hipStreamSynchronize(st) == hipSuccess ? MHM_OK : dfail(MHM_EDEVICE, "stream failed");
  const double target = rtol * bnorm;
  std::vector<double> H((size_t)(R + 1) * R), cs(R), sn(R), gv(R + 1), y(R);
  int total = 0;
  double rel = INFINITY, last = INFINITY;
  for (;;) {
    // r = b - S x (double-double rows), beta = ||r||
    hipLaunchKernelGGL(k_glb_residual, dim3(nb), dim3(256), 0, st, n, (const int *)g->rowptr, (const int *)g->col,
                       (const double *)g->vals, (const double *)d_x, d_rhs, g->r);
    glb_dots(g, g->r, 1, g->r, dh, st);
    if (int rc = glb_fetch(hp, dh, 1, st)) return rc;
    const double beta = std::sqrt(hp[0]);
    rel = beta / bnorm;
    if (!(beta < INFINITY)) return dfail(MHM_ERESIDUAL, "global solver: residual not finite");
    if (beta <= target) break;
    if (total >= maxiter || !(beta < 0.5 * last)) {
      // out of steps, or a whole cycle no longer halves the residual (the
      // attainable accuracy of the preconditioned iteration)
      if (iterations) *iterations = total;
      if (residual) *residual = rel;
      char buf[160];
      snprintf(buf, sizeof buf, "global solver: GMRES stopped at relative residual %.3e > rtol %.3e after %d steps",
               rel, rtol, total);
      return dfail(MHM_ERESIDUAL, buf);
    }
    last = beta;
    hipLaunchKernelGGL(k_glb_scale, dim3(nb), dim3(256), 0, st, n, 1.0 / beta, (const double *)g->r, g->V);
    std::fill(gv.begin(), gv.end(), 0.0);
    gv[0] = beta;
    int k = 0;
    // a cycle aims one decade below the checked criterion, or nine decades
    // below its own starting residual, whichever is larger: later cycles
    // restart from the double-double residual, refining the solution
    const double inner = std::max(0.1 * target, 1e-9 * beta);
    for (; k < R && total < maxiter; ++k, ++total) {
      double *zk = g->Z + (size_t)k * n;
      glb_precond(g, g->V + (size_t)k * n, zk, st);
      hipLaunchKernelGGL(k_glb_spmv, dim3(nb), dim3(256), 0, st, n, (const int *)g->rowptr, (const int *)g->col,
                         (const double *)g->vals, (const double *)zk, g->w);
      // classical Gram-Schmidt, twice
      glb_dots(g, g->V, k + 1, g->w, dh, st);
      hipLaunchKernelGGL(k_glb_orth, dim3(nb), dim3(256), 0, st, n, k + 1, (const double *)g->V, (long)n,
                         (const double *)dh, g->w);
      glb_dots(g, g->V, k + 1, g->w, dh2, st);
      hipLaunchKernelGGL(k_glb_orth, dim3(nb), dim3(256), 0, st, n, k + 1, (const double *)g->V, (long)n,
                         (const double *)dh2, g->w);
      glb_dots(g, g->w, 1, g->w, dh2 + (R + 1), st);
      if (int rc = glb_fetch(hp, dh, (size_t)2 * (R + 2), st)) return rc;
      double *col = &H[(size_t)k * (R + 1)];
      for (int j = 0; j <= k; ++j) col[j] = hp[j] + hp[(R + 2) + j];
      const double hk1 = std::sqrt(hp[(R + 2) + (R + 1)]);
      col[k + 1] = hk1;
      if (!(hk1 < INFINITY)) return dfail(MHM_ERESIDUAL, "global solver: Arnoldi vector not finite");
      for (int j = 0; j < k; ++j) {
        const double a = cs[j] * col[j] + sn[j] * col[j + 1];
        col[j + 1] = -sn[j] * col[j] + cs[j] * col[j + 1];
        col[j] = a;
      }
      const double den = std::hypot(col[k], col[k + 1]);
      if (den == 0.0) return dfail(MHM_ESINGULAR, "global solver: breakdown with a singular Hessenberg matrix");
      cs[k] = col[k] / den;
      sn[k] = col[k + 1] / den;
      col[k] = den;
      col[k + 1] = 0.0;
      gv[k + 1] = -sn[k] * gv[k];
      gv[k] = cs[k] * gv[k];
      const bool lucky = hk1 <= 1e-300;
      if (!lucky)
        hipLaunchKernelGGL(k_glb_scale, dim3(nb), dim3(256), 0, st, n, 1.0 / hk1, (const double *)g->w,
                           g->V + (size_t)(k + 1) * n);
      if (std::fabs(gv[k + 1]) <= inner || lucky) {
        ++k;
        ++total;
        break;
      }
    }
    // y = H^-1 g (upper triangular), x += Z y
    for (int i = k - 1; i >= 0; --i) {
      double s = gv[i];
      for (int j = i + 1; j < k; ++j) s -= H[(size_t)j * (R + 1) + i] * y[j];
      y[i] = s / H[(size_t)i * (R + 1) + i];
    }
    std::memcpy(hp, y.data(), k * sizeof(double));
    if (hipMemcpyAsync(dy, hp, k * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess)
      return dfail(MHM_EDEVICE, "global solver: host-to-device copy failed");
    hipLaunchKernelGGL(k_glb_combine, dim3(nb), dim3(256), 0, st, n, k, (const double *)g->Z, (long)n,
                       (const double *)dy, d_x);
    if (hipStreamSynchronize(st) != hipSuccess) return dfail(MHM_EDEVICE, "global solver: stream failed");
  }
  if (iterations) *iterations = total;
  if (residual) *residual = rel;
  return hipStreamSynchronize(st) == hipSuccess ? MHM_OK : dfail(MHM_EDEVICE, "global solver: stream failed");
}

extern "C" int mhm_global_destroy(mhm_global *g) {
  if (!g) return MHM_OK;
  {
    DevScope scope(g->device);
    (void)hipDeviceSynchronize();
    glb_release(g);
  }
  delete g;
  return MHM_OK;
}
