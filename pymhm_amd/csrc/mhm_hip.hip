// MI355X (gfx950) kernels and C ABI of the MHM local-problem stage.
//
// Per macroelement the stage assembles the local stiffness A (scalar
// diffusion, P1/P2 simplices) and load f from coefficients tabulated at
// quadrature points, factors the augmented operator
//     S = [[A, c], [c^T, 0]],  c = C min(1, max|A| / max|C|),
// solves S X = [f, B, A Z] (source, one lift per multiplier, retained action),
// refines in binary64, condenses into the (m+k)^2 block of the global hybrid
// matrix and reconstructs fields u = s - L lambda + E c.
//
// Factorisation.  A is symmetric positive semidefinite with kernel Z (the
// constants), so the library factors the SPD band matrix
//     Ahat = A + sigma e_p e_p^T,  sigma = A_pp,
// after symmetric Jacobi scaling D Ahat D (unit diagonal), with a band
// Cholesky whose (b+1)x(b+1) active window stays resident in LDS while the
// band streams through it.  The saddle solution follows from three Ahat
// solves (rhs, e_p, c) and a 2x2 bordered system for (x_p, mu):
//     x = y + sigma x_p u - mu w,   y = Ahat^-1 g, u = Ahat^-1 e_p, w = Ahat^-1 c
//     [1 - sigma u_p,  w_p ] [x_p]   [ y_p         ]
//     [-sigma c^T u,  c^T w] [mu ] = [ c^T y - beta]
// which is exact for any nonsingular S (no assumption A Z == 0).  Binary64
// residuals of the true S, accumulated in double-double, drive iterative
// refinement; the factor and triangular solves run in binary64 (MHM_FP64) or
// binary32 (MHM_MIXED).
//
// Every reduction has a fixed order, so a macro's results do not depend on
// which other macros share the launch.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mhm_hip.h"

// ----------------------------------------------------------------- errors
static thread_local std::string g_error;

static int fail(int code, const std::string &msg) {
  g_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess)                                                             \
      return fail(_e == hipErrorOutOfMemory ? MHM_ENOMEM : MHM_EDEVICE,               \
                  std::string(#expr) + ": " + hipGetErrorString(_e));                 \
  } while (0)

#define LAUNCH_CHECK(name)                                                            \
  do {                                                                                \
    hipError_t _e = hipGetLastError();                                                \
    if (_e != hipSuccess)                                                             \
      return fail(MHM_EDEVICE, std::string("launch ") + name + ": " + hipGetErrorString(_e)); \
  } while (0)

// ----------------------------------------------------------------- device helpers
// The error-free transformations below are exact only with every operation
// rounded on its own: contraction into fused multiply-adds is off in them (the
// one fma of dd_fma is explicit).
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
#pragma clang fp contract(off)
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Accumulate the exact product a*b into the double-double (hi, lo).
__device__ __forceinline__ void dd_fma(double &hi, double &lo, double a, double b) {
#pragma clang fp contract(off)
  double p = a * b;
  double pe = fma(a, b, -p);
  double s, e;
  two_sum(hi, p, s, e);
  hi = s;
  lo += e + pe;
}

__device__ __forceinline__ void dd_add(double &hi, double &lo, double v) {
#pragma clang fp contract(off)
  double s, e;
  two_sum(hi, v, s, e);
  hi = s;
  lo += e;
}

// A zero in global memory, never written: loads that select between a real
// address and this one stay global loads (a zero in LDS or constant memory
// would turn the selected pointer into a flat address), and loads outside a
// band or a matrix can read it instead of branching.
__device__ double kZeroF64[1];
__device__ float kZeroF32[1];
template <typename S>
__device__ __forceinline__ const S *zero_of();
template <>
__device__ __forceinline__ const double *zero_of<double>() { return kZeroF64; }
template <>
__device__ __forceinline__ const float *zero_of<float>() { return kZeroF32; }

// A pointer known to address global memory (one read back from LDS would
// otherwise be accessed as a flat address).
template <typename P>
__device__ __forceinline__ const __attribute__((address_space(1))) P *global_ptr(const P *p) {
  return (const __attribute__((address_space(1))) P *)p;
}

// Compiler-only barrier: memory accesses are not moved across it.
#define MEM_FENCE() asm volatile("" ::: "memory")

// Cross-lane hand-off through LDS inside one wavefront: the wavefront-scope
// release/acquire fences keep the compiler from reusing or reordering LDS
// values across it (lanes read what other lanes wrote), the wave barrier keeps
// the scheduler from moving instructions over it.  LDS operations of one
// wavefront execute in order, so no block barrier is needed.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// e / d for 0 <= e < 2^22, 1 <= d < 2^16: a binary32 reciprocal estimate is
// within one of the quotient, one correction each way makes it exact.
struct FastDiv {
  int d;
  float r;
  __device__ explicit FastDiv(int dv) : d(dv), r(1.0f / (float)dv) {}
  __device__ __forceinline__ int operator()(int e) const {
    int q = (int)((float)e * r);
    q += (q + 1) * d <= e;
    q -= q * d > e;
    return q;
  }
};

// Value of v in lane l (wave-uniform l) as a scalar: v_readlane, no LDS.
__device__ __forceinline__ double lane_bcast(double v, int l) {
  const long long x = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)x, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(x >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float lane_bcast(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// 1/sqrt(d) for a positive normal d: the hardware estimate refined by Newton
// steps y <- y + y (1 - d y^2) / 2 (two in binary64, one in binary32), within
// a few ulp; the factorisations take root = d / sqrt(d) = d * y.
__device__ __forceinline__ double rsqrt_nr(double d) {
  double y = __builtin_amdgcn_rsq(d);
#pragma unroll
  for (int i = 0; i < 2; ++i) y = fma(0.5 * y, fma(-d, y * y, 1.0), y);
  return y;
}
__device__ __forceinline__ float rsqrt_nr(float d) {
  const float y = __builtin_amdgcn_rsqf(d);
  return fmaf(0.5f * y, fmaf(-d, y * y, 1.0f), y);
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Fixed-order block reductions (result broadcast to every thread).
template <int NT>
__device__ double block_sum(double v, double *sh) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < NT / 64; ++i) t += sh[i];
  return t;
}

// Double-double sum over the block in a fixed order (shuffle tree per
// wavefront, then the wavefront partials in order); the result is returned to
// every thread.  sh holds 2 * NT / 64 doubles.
template <int NT>
__device__ void block_dd_sum(double &hi, double &lo, double *sh) {
  for (int off = 32; off > 0; off >>= 1) {
    const double oh = __shfl_down(hi, off), ol = __shfl_down(lo, off);
    double s, e;
    two_sum(hi, oh, s, e);
    hi = s;
    lo = lo + ol + e;
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    sh[2 * (threadIdx.x >> 6)] = hi;
    sh[2 * (threadIdx.x >> 6) + 1] = lo;
  }
  __syncthreads();
  double th = 0.0, tl = 0.0;
#pragma unroll
  for (int i = 0; i < NT / 64; ++i) {
    double s, e;
    two_sum(th, sh[2 * i], s, e);
    th = s;
    tl += sh[2 * i + 1] + e;
  }
  hi = th;
  lo = tl;
}

template <int NT>
__device__ double block_max(double v, double *sh) {
  v = wave_max(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = sh[0];
#pragma unroll
  for (int i = 1; i < NT / 64; ++i) t = fmax(t, sh[i]);
  return t;
}

// Workgroup barrier that orders LDS only.  __syncthreads() also releases and
// acquires global memory, which waits for every outstanding vector-memory
// operation - including prefetch loads meant to stay in flight across rows.
// Used inside the band kernels' row loops, where no thread reads global data
// written by another thread of the block; full __syncthreads() remains
// wherever it does.
__device__ __forceinline__ void lds_barrier() {
  __asm__ volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// ----------------------------------------------------------------- event timing
struct TimerState {
  bool enabled = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> spans[MHM_TIMER_COUNT];
  std::vector<hipEvent_t> pool;
};
static TimerState g_timer;

static hipEvent_t timer_event() {
  if (!g_timer.pool.empty()) {
    hipEvent_t e = g_timer.pool.back();
    g_timer.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// RAII span: records start/end events around the launches in its scope
// (none for a negative id).
struct TimerSpan {
  int id;
  hipStream_t st;
  hipEvent_t a = nullptr;
  TimerSpan(int kernel, hipStream_t stream) : id(kernel), st(stream) {
    if (id >= 0 && g_timer.enabled && (a = timer_event())) (void)hipEventRecord(a, st);
  }
  ~TimerSpan() {
    if (!a) return;
    hipEvent_t b = timer_event();
    if (!b) return;
    (void)hipEventRecord(b, st);
    g_timer.spans[id].push_back({a, b});
  }
};

// ----------------------------------------------------------------- kernel arguments
struct LayoutK {
  int n, k, m, nnz, b, W, pin, ne, nb, dim, nq, ncr;
  const int *perm, *rowptr, *cols, *diagslot, *crows;
  const double *coupling, *kernel, *constraint;
  const int *slot_ptr, *slot_code, *load_ptr, *load_code;
  const double *glam, *measure, *qw, *dtab, *vtab;
  const int *bslot;  // bslot[j*W + r] = CSR slot of A(j + r, j), or -1
  const int *srow;   // row of every CSR slot
};

// Macroface skeleton of a layout (mhm_layout_set_skeleton): the boundary
// edges of the local mesh split into pieces that each lie inside one face
// segment, and the nonzero (band row, side, segment) entries of the coupling
// with the (piece, local basis) pairs that sum into each.
struct SkelK {
  int nsides, nseg, degree, edeg, ng, npieces, nentries;
  const double *breaks;  // [nseg + 1] in the face parameter (symmetric)
  const double *pf;      // [npieces][4]: piece f0, f1 and its edge's e0, e1 (side fractions)
  const double *gx, *gw; // [ng] Gauss-Legendre nodes and weights on [-1, 1]
  const int *eside, *eseg, *erow, *eptr, *contrib;  // entries and their (piece * 3 + basis) lists
};

// Per-macro scalars kept in scal[mac * NSCAL + ...].
enum { S_AMAX = 0, S_SIGMA, S_CS, S_A11, S_A12, S_A21, S_A22, S_DET, NSCAL };

struct BatchK {
  int M, R, N, ldy;
  double *aval, *load, *az, *dscale, *scal, *uw;
  int *status, *iters;  // status: factor-time bits (pivot, non-finite operator), kept until refactored
  int *cstatus;         // per-call bits of the last solve (residual, non-finite data), cleared per call
  int *astatus;         // bits of the last assembly (coefficient outside (0, inf) or non-finite source)
  const int *colmap;                          // per-macro signed column map or null
  const double *colscale, *bcoup, *cons, *kern;  // per-macro overrides (band rows) or null
  const double *glam, *meas;                     // per-macro element geometry or null
};

__device__ __forceinline__ double coupling_at(const LayoutK &L, const BatchK &B, int mac, int a,
                                              int i) {
  if (B.bcoup) return B.bcoup[((size_t)mac * L.m + a) * L.n + i];
  const int src = B.colmap ? B.colmap[(size_t)mac * L.m + a] : a;
  const double v = L.coupling[(size_t)src * L.n + i];
  return B.colscale ? v * B.colscale[(size_t)mac * L.m + a] : v;
}

__device__ __forceinline__ double constraint_at(const LayoutK &L, const BatchK &B, int mac,
                                                int i) {
  return B.cons ? B.cons[(size_t)mac * L.n + i] : L.constraint[i];
}

__device__ __forceinline__ double kernel_at(const LayoutK &L, const BatchK &B, int mac, int i) {
  if (B.kern) return B.kern[(size_t)mac * L.n + i];
  return L.kernel ? L.kernel[i] : 1.0;
}

// ----------------------------------------------------------------- assembly
// Compensated accumulation: total += value, corr += the exact rounding error of
// that addition.  Knuth's TwoSum recovers the error without the magnitude test
// of the Neumaier step (fem/scalar/_integration.py:32-40).  Both steps deliver
// the exact error - a representable binary64, +0 when the sum is exact - so
// for finite operands they leave identical totals and corrections.
__device__ __forceinline__ void comp_add(double &total, double &corr, double value) {
#pragma clang fp contract(off)
  const double s = total + value;
  const double bv = s - total;
  corr += (total - (s - bv)) + (value - bv);
  total = s;
}

// Waves per block of the assembly kernel.
#define ASM_WAVES 4

// Element stiffness, one lane per (element, basis row i), 64 / NB elements per
// wavefront; every lane owns row i of its element and the load entry i.
//
// The arithmetic repeats the reference Gram kernel operation for operation -
// physical gradients as the sequential sum sum_l dphi/dlambda_l grad(lambda_l),
// K grad(phi_j) (for K = k I the compensated sum over b of K[a][b] g[j][b]
// reduces to the single rounded product k g[j][a]), the Cartesian contraction
// compensated, ((c + corr) w_q) |T|, compensated quadrature sum
// (fem/scalar/_integration.py:43-88) - with floating contraction off, so the
// element matrices match a binary64 CPU evaluation bit for bit.  At every
// quadrature point a lane forms only its own row's gradient and hands
// k grad(phi_i) to the other lanes of its element through LDS (double-buffered
// by point parity, one wavefront fence per point).  The first step of the
// Cartesian sum starts from (0, 0), where the correction is exactly +0 for a
// finite product.  The load is sum_q ((w_q phi_i) f_q) |T| in quadrature order.
// Element matrices are stored code-major, ework[mac][i*NB+j][e].  A block keeps
// the reference tables in LDS and walks macros blockIdx.y, + gridDim.y, ...
template <int NB, int D>
__global__ void __launch_bounds__(64 * ASM_WAVES)
    k_assemble(LayoutK L, int M, const double *__restrict__ kq, const double *__restrict__ fq,
               double fconst, double *__restrict__ ework, double *__restrict__ fwork,
               const double *__restrict__ bglam, const double *__restrict__ bmeas, int *__restrict__ astatus) {
#pragma clang fp contract(off)
  constexpr int EPW = 64 / NB;
  constexpr int DP = D == 3 ? 4 : D;  // padded row of k grad(phi) in LDS
  constexpr int KG = EPW * NB * DP;   // one buffer of a wavefront
  extern __shared__ __attribute__((aligned(16))) double sh[];
  const int nq = L.nq;
  const int tab = (nq * NB * (D + 1) + nq * NB + nq + 1) & ~1;
  double *sd = sh;
  double *sv = sd + nq * NB * (D + 1);
  double *sw = sv + nq * NB;
  for (int t = threadIdx.x; t < nq * NB * (D + 1); t += blockDim.x) sd[t] = L.dtab[t];
  for (int t = threadIdx.x; t < nq * NB; t += blockDim.x) sv[t] = L.vtab[t];
  for (int t = threadIdx.x; t < nq; t += blockDim.x) sw[t] = L.qw[t];
  __syncthreads();
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int e0 = (blockIdx.x * ASM_WAVES + wv) * EPW;
  if (e0 >= L.ne) return;  // whole wavefront past the last element
  const int el = min(lane / NB, EPW - 1);
  const int i = lane - el * NB < NB ? lane - el * NB : NB - 1;
  const bool own = lane < EPW * NB && e0 + el < L.ne;
  const int e = min(e0 + el, L.ne - 1);
  double *kgb = sh + tab + wv * 2 * KG;
  // element geometry: the layout's, or the macro's own (uneven macros)
  double gl[D + 1][D];
  double meas = 0.0;
  auto geometry = [&](const double *g, const double *ms) {
#pragma unroll
    for (int l = 0; l <= D; ++l)
#pragma unroll
      for (int a = 0; a < D; ++a) gl[l][a] = g[((size_t)e * (D + 1) + l) * D + a];
    meas = ms[e];
  };
  if (!bglam) geometry(L.glam, L.measure);
  for (int mac = blockIdx.y; mac < M; mac += gridDim.y) {
    if (bglam) geometry(bglam + (size_t)mac * L.ne * (D + 1) * D, bmeas + (size_t)mac * L.ne);
    double tot[NB], cor[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) tot[j] = cor[j] = 0.0;
    double F = 0.0;
    const double *kp = kq + ((size_t)mac * L.ne + e) * nq;
    const double *fp = fq ? fq + ((size_t)mac * L.ne + e) * nq : nullptr;
    double knext = kp[0];
    double fnext = fp ? fp[0] : fconst;
    bool valid = true;  // diffusion finite and positive, source finite (materials/evaluation.py:16-29)
    for (int q = 0; q < nq; ++q) {
      const double k = knext, f = fnext;
      valid &= (k > 0.0) & (fabs(k) <= DBL_MAX) & (fabs(f) <= DBL_MAX);
      if (q + 1 < nq) {
        knext = kp[q + 1];
        if (fp) fnext = fp[q + 1];
      }
      const double *dq = sd + (q * NB + i) * (D + 1);
      double gi[D];
#pragma unroll
      for (int a = 0; a < D; ++a) {
        double acc = dq[0] * gl[0][a];
#pragma unroll
        for (int l = 1; l <= D; ++l) acc = acc + dq[l] * gl[l][a];
        gi[a] = acc;
      }
      double *buf = kgb + (q & 1) * KG;
#pragma unroll
      for (int a = 0; a < D; ++a) buf[(el * NB + i) * DP + a] = k * gi[a];
      wave_lds_sync();
      const double w = sw[q];
      const double *row = buf + el * NB * DP;
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        double ct = 0.0 + gi[0] * row[j * DP], cc = 0.0;
#pragma unroll
        for (int a = 1; a < D; ++a) comp_add(ct, cc, gi[a] * row[j * DP + a]);
        comp_add(tot[j], cor[j], ((ct + cc) * w) * meas);
      }
      F = F + ((w * sv[q * NB + i]) * f) * meas;
    }
    if (own) {
      const size_t base = (size_t)mac * NB * NB * L.ne;
#pragma unroll
      for (int j = 0; j < NB; ++j) ework[base + (size_t)(i * NB + j) * L.ne + e] = tot[j] + cor[j];
      fwork[(size_t)mac * NB * L.ne + (size_t)i * L.ne + e] = F;
      if (!valid && i == 0) atomicOr(&astatus[mac], MHM_STATUS_COEFFICIENT);
    }
  }
}

// XCD-aware block order for a grid of R parts (x) per macro (y): the chip
// hands workgroups to its 8 XCDs round-robin by linear id, and each XCD has
// its own L2, so in launch order a macro's parts would spread over all eight.
// Linear id L maps to (macro, part) with macro % 8 == L % 8 - every part of a
// macro on one XCD - for the first M - M % 8 macros; the others keep the
// launch order.  A bijection either way: correctness does not depend on the
// dispatch order.
__device__ __forceinline__ void xcd_block(int &mac, int &part) {
  const int R = gridDim.x, M = gridDim.y;
  const int L = blockIdx.x + R * blockIdx.y;
  const int Mp = M & ~7;
  if (L < Mp * R) {
    const int k = L >> 3, g = k / R;
    mac = (L & 7) + 8 * g;
    part = k - g * R;
  } else {
    const int r = L - Mp * R, g = r / R;
    mac = Mp + g;
    part = r - g * R;
  }
}

// Ordered gather of the element contributions into the CSR values of
// GATHER_ROWS band rows: one thread per CSR slot sums the slot's contributions
// in the reference's duplicate-summation order (first value, then +=, fixed
// element order -> deterministic, the reference's bits), GATHER_SPT slots per
// thread at a time with their k-th contributions loaded together; the block
// then forms the load rows and A Z (double-double, slot order) one thread per
// row from the slot values staged in LDS.
#define GATHER_ROWS 64
#define GATHER_SPT 4  // CSR slots per thread whose contributions are loaded together (8: more VGPRs, slower in 2D)
__global__ void __launch_bounds__(256)
    k_gather(LayoutK L, BatchK B, const double *__restrict__ ework,
             const double *__restrict__ fwork) {
  extern __shared__ double gsv[];
  int mac, part;
  xcd_block(mac, part);  // a macro's row blocks share one XCD's L2 for its element matrices
  const int i0 = part * GATHER_ROWS, i1 = min(L.n, i0 + GATHER_ROWS);
  const int s0 = L.rowptr[i0], s1 = L.rowptr[i1];
  const double *ew = ework + (size_t)mac * L.nb * L.nb * L.ne;
  double *av = B.aval + (size_t)mac * L.nnz;
  // a slot's sum stays first value, then += in order
  for (int sb = s0 + threadIdx.x; sb < s1; sb += GATHER_SPT * blockDim.x) {
    int t[GATHER_SPT], te[GATHER_SPT], left = 0;
    double v[GATHER_SPT];
#pragma unroll
    for (int u = 0; u < GATHER_SPT; ++u) {
      const int s = sb + u * (int)blockDim.x;
      t[u] = s < s1 ? L.slot_ptr[s] : 0;
      te[u] = s < s1 ? L.slot_ptr[s + 1] : 0;
      left = max(left, te[u] - t[u]);
      v[u] = 0.0;
    }
    for (int k = 0; k < left; ++k) {
      int code[GATHER_SPT];
      double x[GATHER_SPT];
#pragma unroll
      for (int u = 0; u < GATHER_SPT; ++u) code[u] = t[u] + k < te[u] ? L.slot_code[t[u] + k] : -1;
#pragma unroll
      for (int u = 0; u < GATHER_SPT; ++u) x[u] = code[u] >= 0 ? ew[code[u]] : 0.0;
#pragma unroll
      for (int u = 0; u < GATHER_SPT; ++u)
        if (code[u] >= 0) v[u] = k == 0 ? x[u] : v[u] + x[u];
    }
#pragma unroll
    for (int u = 0; u < GATHER_SPT; ++u) {
      const int s = sb + u * (int)blockDim.x;
      if (s < s1) {
        av[s] = v[u];
        gsv[s - s0] = v[u];
      }
    }
  }
  __syncthreads();
  const int i = i0 + threadIdx.x;
  if (i >= i1) return;
  double hi = 0.0, lo = 0.0;
  for (int s = L.rowptr[i]; s < L.rowptr[i + 1]; ++s) dd_fma(hi, lo, gsv[s - s0], kernel_at(L, B, mac, L.cols[s]));
  B.az[(size_t)mac * L.n + i] = hi + lo;
  const double *fw = fwork + (size_t)mac * L.nb * L.ne;
  double f = 0.0;
  for (int t = L.load_ptr[i]; t < L.load_ptr[i + 1]; ++t) f += fw[L.load_code[t]];
  B.load[(size_t)mac * L.n + i] = f;
}

// New source fields against resident factors, block per (macro, field c):
// element loads sum_q ((w_q phi_i) f_q) |T| in quadrature order (the same
// operations as k_assemble, so equal loads give equal bits) staged in LDS, then
// the ordered load gather of k_gather into column c of the solve's right-hand
// side rhs[mac][N][ncols] (band rows; the constraint row is 0), scaled by a
// power of two per column (unscale[mac][c] undoes it).
template <int NB, int D>
__global__ void __launch_bounds__(256)
    k_source_load(LayoutK L, int M, int ncols, const double *__restrict__ fq, double fconst,
                  double *__restrict__ rhs, int N, double *__restrict__ unscale,
                  const double *__restrict__ bmeas) {
#pragma clang fp contract(off)
  extern __shared__ double fe[];  // NB x ne, code i*ne + e
  const int mac = blockIdx.x, c = blockIdx.y;
  const int nq = L.nq;
  for (int t = threadIdx.x; t < NB * L.ne; t += blockDim.x) {
    const int i = t / L.ne, e = t - i * L.ne;
    const double *fp = fq ? fq + (((size_t)c * M + mac) * L.ne + e) * nq : nullptr;
    const double meas = bmeas ? bmeas[(size_t)mac * L.ne + e] : L.measure[e];
    double F = 0.0;
    for (int q = 0; q < nq; ++q) F = F + ((L.qw[q] * L.vtab[q * NB + i]) * (fp ? fp[q] : fconst)) * meas;
    fe[t] = F;
  }
  __syncthreads();
  double *out = rhs + (size_t)mac * N * ncols + c;
  double amax = 0.0;
  for (int row = threadIdx.x; row < N; row += blockDim.x) {
    double f = 0.0;
    if (row < L.n)
      for (int t = L.load_ptr[row]; t < L.load_ptr[row + 1]; ++t) f += fe[L.load_code[t]];
    out[(size_t)row * ncols] = f;
    amax = fmax(amax, fabs(f));
  }
  // scale the column by 2^-e, max|f| = m 2^e, m in [0.5, 1): exact, and it keeps
  // a source that is tiny on this macro inside the binary32 range of the mixed
  // solve; the results are scaled back by 2^e
  __shared__ double red[4];
  amax = block_max<256>(amax, red);
  int e = 0;
  if (amax > 0.0 && isfinite(amax)) (void)frexp(amax, &e);
  const double scale = ldexp(1.0, -e);
  for (int row = threadIdx.x; row < N; row += blockDim.x) out[(size_t)row * ncols] *= scale;
  if (threadIdx.x == 0) unscale[(size_t)mac * ncols + c] = ldexp(1.0, e);
}

// Condensed loads of new sources (core/reconstruction.py:91-127), block per
// (macro, field c), from s = X[:, c] and f = rhs[:, c]: B_t^T s over the
// coupled rows in row order (as k_condense), and with a retained mode
// (A Z)^T s - Z^T f by fixed-order block reductions; optionally the source
// responses in original rows.  Each field's arithmetic is independent of the
// others, so a request's result does not depend on its batch.
__global__ void __launch_bounds__(256)
    k_source_condense(LayoutK L, BatchK B, int ncols, const double *__restrict__ X,
                      const double *__restrict__ rhs, const double *__restrict__ unscale,
                      double *__restrict__ loads, double *__restrict__ source) {
  __shared__ double red[2][256];
  const int mac = blockIdx.x, c = blockIdx.y, tid = threadIdx.x;
  const double us = unscale[(size_t)mac * ncols + c];  // a power of two
  const int n = L.n, m = L.m, N = B.N, w = L.m + L.k;
  const double *x = X + (size_t)mac * N * ncols + c;
  const double *f = rhs + (size_t)mac * N * ncols + c;
  double *ld = loads + ((size_t)c * B.M + mac) * w;
  const int nrows = L.crows ? L.ncr : n;
  for (int a = tid; a < m; a += 256) {
    double acc = 0.0;
    for (int q = 0; q < nrows; ++q) {
      const int i = L.crows ? L.crows[q] : q;
      acc += coupling_at(L, B, mac, a, i) * x[(size_t)i * ncols];
    }
    ld[a] = acc * us;
  }
  if (source)
    for (int i = tid; i < n; i += 256) source[((size_t)c * B.M + mac) * n + L.perm[i]] = x[(size_t)i * ncols] * us;
  if (!L.k) return;
  double as = 0.0, zf = 0.0;
  for (int i = tid; i < n; i += 256) {
    as += B.az[(size_t)mac * n + i] * x[(size_t)i * ncols];
    zf += kernel_at(L, B, mac, i) * f[(size_t)i * ncols];
  }
  red[0][tid] = as;
  red[1][tid] = zf;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o) {
      red[0][tid] += red[0][tid + o];
      red[1][tid] += red[1][tid + o];
    }
    __syncthreads();
  }
  if (tid == 0) ld[m] = (red[0][0] - red[1][0]) * us;
}

// Assembled-operator path: A Z from supplied CSR values; loads permuted to band rows.
__global__ void __launch_bounds__(128)
    k_rowsum(LayoutK L, BatchK B, const double *__restrict__ loads_orig) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int mac = blockIdx.y;
  if (i >= L.n) return;
  double hi = 0.0, lo = 0.0;
  for (int s = L.rowptr[i]; s < L.rowptr[i + 1]; ++s)
    dd_fma(hi, lo, B.aval[(size_t)mac * L.nnz + s], kernel_at(L, B, mac, L.cols[s]));
  B.az[(size_t)mac * L.n + i] = hi + lo;
  B.load[(size_t)mac * L.n + i] = loads_orig[(size_t)mac * L.n + L.perm[i]];
}

// ----------------------------------------------------------------- skeleton and boundary forms
// Per-macro coupling B[i, a] = sign_s |side s| int_0^1 phi_i mu_a over the fine
// boundary edges of macro side s (trace_coupling, fem/scalar/triangle.py:160-205,
// and face_integration, fem/scalar/operators.py:200-244): thread per (coupling
// entry, macro).  Each piece of a boundary edge lies inside one face segment; a
// Gauss-Legendre rule of ng points integrates the nodal function phi (P1 or P2
// on the edge, in the side's counterclockwise fraction f) against the Legendre
// modes of the segment in the face parameter t - t = f for the face's first
// owner, t = 1 - f for the second (sign -1), which mirrors the segment and flips
// the odd modes.  Column a = s fs + segment(t) (p + 1) + j.
#define SKEL_MAXMODES 8
__global__ void __launch_bounds__(256)
    k_trace_coupling(LayoutK L, SkelK K, const double *__restrict__ side_len, const double *__restrict__ side_sign,
                     double *__restrict__ bcoup) {
  const int en = blockIdx.x * blockDim.x + threadIdx.x;
  const int mac = blockIdx.y;
  if (en >= K.nentries) return;
  const int s = K.eside[en], row = K.erow[en];
  const double len = side_len[(size_t)mac * K.nsides + s];
  const double sg = side_sign[(size_t)mac * K.nsides + s];
  const bool fwd = sg > 0.0;
  const int seg = fwd ? K.eseg[en] : K.nseg - 1 - K.eseg[en];
  const double lo = K.breaks[seg], hi = K.breaks[seg + 1];
  double acc[SKEL_MAXMODES];
#pragma unroll
  for (int j = 0; j < SKEL_MAXMODES; ++j) acc[j] = 0.0;
  for (int t = K.eptr[en]; t < K.eptr[en + 1]; ++t) {
    const int code = K.contrib[t];
    const int piece = code / 3, bas = code - 3 * piece;
    const double *pp = K.pf + 4 * (size_t)piece;
    const double f0 = pp[0], f1 = pp[1], e0 = pp[2], e1 = pp[3];
    const double half = 0.5 * (f1 - f0);
    for (int q = 0; q < K.ng; ++q) {
      const double f = f0 + (K.gx[q] + 1.0) * half;
      const double u = (f - e0) / (e1 - e0);  // coordinate along the edge, start -> end
      double phi;
      if (K.edeg == 1)
        phi = bas == 0 ? 1.0 - u : u;
      else
        phi = bas == 0 ? (1.0 - u) * (1.0 - 2.0 * u) : (bas == 1 ? u * (2.0 * u - 1.0) : 4.0 * u * (1.0 - u));
      const double tq = fwd ? f : 1.0 - f;
      const double x = 2.0 * (tq - lo) / (hi - lo) - 1.0;
      const double wq = K.gw[q] * half * phi;
      double lm1 = 1.0, lj = x;  // Bonnet's recurrence for the conventional Legendre modes
      acc[0] += wq;
#pragma unroll
      for (int j = 1; j < SKEL_MAXMODES; ++j) {
        if (j > K.degree) break;
        acc[j] += wq * lj;
        const double nxt = ((2 * j + 1) * x * lj - j * lm1) / (j + 1);
        lm1 = lj;
        lj = nxt;
      }
    }
  }
  const int fs = K.nseg * (K.degree + 1);
#pragma unroll
  for (int j = 0; j < SKEL_MAXMODES; ++j)
    if (j <= K.degree)
      bcoup[((size_t)mac * L.m + s * fs + seg * (K.degree + 1) + j) * L.n + row] = acc[j] * len * sg;
}

// C = M 1 / (1^T M 1) per macro from its own element measures (the constraint
// of the primal Darcy provider, _legacy/models/darcy/primal.py:70-72, with M 1
// the integral of each basis function): row i sums meas[e] vint[b] over its
// element contributions in the fixed load order, then every row is divided by
// the block's fixed-order total.
__global__ void __launch_bounds__(256)
    k_mass_constraints(LayoutK L, BatchK B, const double *__restrict__ vint, double *__restrict__ cons) {
  __shared__ double red[4];
  const int mac = blockIdx.x;
  const double *ms = B.meas ? B.meas + (size_t)mac * L.ne : L.measure;
  double *c = cons + (size_t)mac * L.n;
  double sum = 0.0;
  for (int i = threadIdx.x; i < L.n; i += 256) {
    double v = 0.0;
    for (int t = L.load_ptr[i]; t < L.load_ptr[i + 1]; ++t) {
      const int code = L.load_code[t];
      const int bas = code / L.ne, e = code - bas * L.ne;
      v += ms[e] * vint[bas];
    }
    c[i] = v;
    sum += v;
  }
  sum = block_sum<256>(sum, red);
  for (int i = threadIdx.x; i < L.n; i += 256) c[i] = c[i] / sum;
}

// Weak Dirichlet trace load of a scalar skeleton (boundary_data,
// fem/scalar/operators.py:247-321, without Neumann faces): thread per (boundary
// face f, mode a), load = |f| sum_q (mu_a(t_q) w_q) g_q with the compensated
// accumulation of boundary_moments (fem/scalar/_integration.py:92-108).  The
// global rows take rhs -= load and scale += |load|, as the boundary load enters
// the hybrid system (core/contributions.py:185-195).
__global__ void __launch_bounds__(256)
    k_dirichlet_load(int nfaces, int nq, int fs, const double *__restrict__ basis, const double *__restrict__ w,
                     const double *__restrict__ g, const double *__restrict__ len, const int64_t *__restrict__ dof0,
                     double *__restrict__ load, double *__restrict__ rhs, double *__restrict__ scale) {
#pragma clang fp contract(off)
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nfaces * fs) return;
  const int f = (int)(t / fs), a = (int)(t - (long)f * fs);
  double total = 0.0, corr = 0.0;
  for (int q = 0; q < nq; ++q) comp_add(total, corr, (basis[q * fs + a] * w[q]) * g[(size_t)f * nq + q]);
  const double v = (total + corr) * len[f];
  const int64_t row = dof0[f] + a;
  if (load) load[row] = v;
  if (rhs) rhs[row] -= v;
  if (scale) scale[row] += fabs(v);
}

// The effective per-macro coupling and constraint in ORIGINAL rows (inspection).
__global__ void __launch_bounds__(256)
    k_coupling_out(LayoutK L, BatchK B, double *__restrict__ coup, double *__restrict__ cons) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int mac = blockIdx.y;
  if (t >= (long)(L.m + 1) * L.n) return;
  const int a = (int)(t / L.n), i = (int)(t - (long)a * L.n);
  if (a < L.m) {
    if (coup) coup[((size_t)mac * L.m + a) * L.n + L.perm[i]] = coupling_at(L, B, mac, a, i);
  } else if (cons && L.k) {
    cons[(size_t)mac * L.n + L.perm[i]] = constraint_at(L, B, mac, i);
  }
}

// A Z again after new retained modes (same double-double slot-order sums as
// k_gather and k_rowsum, so the bits do not depend on the call order).
__global__ void __launch_bounds__(128) k_retained_action(LayoutK L, BatchK B) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int mac = blockIdx.y;
  if (i >= L.n) return;
  double hi = 0.0, lo = 0.0;
  for (int s = L.rowptr[i]; s < L.rowptr[i + 1]; ++s)
    dd_fma(hi, lo, B.aval[(size_t)mac * L.nnz + s], kernel_at(L, B, mac, L.cols[s]));
  B.az[(size_t)mac * L.n + i] = hi + lo;
}

// ----------------------------------------------------------------- factor preparation
// Block per macro: max|A| and max|C| (constraint scaling of
// core/condensation.py:160-166), Jacobi scaling of Ahat, band fill.
// Band columns per LDS tile of k_prepare's band fill.
#define PREP_COLS 16

template <typename T>
__global__ void __launch_bounds__(256) k_prepare(LayoutK L, BatchK B, T *__restrict__ band, T *__restrict__ sval) {
  __shared__ double red[4];
  const int mac = blockIdx.x;
  const double *a = B.aval + (size_t)mac * L.nnz;
  double amax = 0.0;
  bool finite = true;
  // max|A| and finiteness are taken in the pass over the slots below
  double cmax = 0.0;
  if (L.k)
    for (int i = threadIdx.x; i < L.n; i += 256) cmax = fmax(cmax, fabs(constraint_at(L, B, mac, i)));
  cmax = block_max<256>(cmax, red);
  const double sigma = (L.k && L.pin >= 0 && L.diagslot[L.pin] >= 0) ? a[L.diagslot[L.pin]] : 0.0;
  int bad = 0;
  for (int i = threadIdx.x; i < L.n; i += 256) {
    double d = L.diagslot[i] >= 0 ? a[L.diagslot[i]] : 0.0;
    if (i == L.pin) d += sigma;
    if (!(d > 0.0) || !isfinite(d)) {
      bad |= MHM_STATUS_PIVOT;
      d = 1.0;
    }
    B.dscale[(size_t)mac * L.n + i] = 1.0 / sqrt(d);
  }
  __syncthreads();  // dscale of every row written
  const double *ds = B.dscale + (size_t)mac * L.n;
  if (sval) {
    // the scaled, pinned operator in CSR slot order, ds[i] (A(i, j) + sigma
    // at the pin) ds[j]: the LDS-window factor gathers its band from it
    T *sv = sval + (size_t)mac * L.nnz;
    const int pslot = L.pin >= 0 ? L.diagslot[L.pin] : -1;
    for (int e = threadIdx.x; e < L.nnz; e += 256) {
      double v = a[e];
      finite &= isfinite(v);
      amax = fmax(amax, fabs(v));
      if (e == pslot) v += sigma;
      sv[e] = T(ds[L.srow[e]] * v * ds[L.cols[e]]);
    }
  } else {
    // The band of the symmetric operator, PREP_COLS band columns at a time
    // through an LDS tile: band[j*W + r] = ds[j+r] A(j, j+r) ds[j], the upper
    // part of CSR row j (the tile is zeroed, the rows' slots, read coalesced in
    // slot order, scattered into it, and it is written out contiguous).  The
    // element matrices are symmetric only to rounding (K_e(i, j) and K_e(j, i)
    // are formed with the roles of i and j exchanged), so A and its transpose
    // differ in the last bit of a few entries; the factor approximates A
    // either way, and the refinement's residuals use A itself.
    extern __shared__ __attribute__((aligned(16))) unsigned char prep_raw[];
    T *tile = reinterpret_cast<T *>(prep_raw);  // PREP_COLS x W
    T *bd = band + (size_t)mac * L.n * L.W;
    const int W = L.W;
    const int pslot = L.pin >= 0 ? L.diagslot[L.pin] : -1;
    for (int j0 = 0; j0 < L.n; j0 += PREP_COLS) {
      const int nc = min(PREP_COLS, L.n - j0);
      for (int e = threadIdx.x; e < nc * W; e += 256) tile[e] = T(0);
      __syncthreads();
      const int s0 = L.rowptr[j0], s1 = L.rowptr[j0 + nc];
      for (int e = s0 + threadIdx.x; e < s1; e += 256) {
        const int i = L.srow[e], c = L.cols[e];
        double v = a[e];
        finite &= isfinite(v);
        amax = fmax(amax, fabs(v));
        if (c >= i) {
          if (e == pslot) v += sigma;
          tile[(i - j0) * W + (c - i)] = T(ds[c] * v * ds[i]);
        }
      }
      __syncthreads();
      for (int e = threadIdx.x; e < nc * W; e += 256) bd[(size_t)j0 * W + e] = tile[e];
      __syncthreads();  // the tile is rewritten by the next columns
    }
  }
  amax = block_max<256>(amax, red);
  if (!finite) bad |= MHM_STATUS_NONFINITE;
  if (bad) atomicOr(&B.status[mac], bad);
  if (threadIdx.x == 0) {
    double *sc = B.scal + (size_t)mac * NSCAL;
    sc[S_AMAX] = amax;
    sc[S_SIGMA] = sigma;
    const double scale = amax > 0.0 ? amax : 1.0;
    sc[S_CS] = (L.k && cmax > 0.0) ? fmin(1.0, scale / cmax) : 1.0;
  }
}

// ----------------------------------------------------------------- band Cholesky
// Block per macro.  The lower band of column j is band[j*W + r] = L(j+r, j).
// Window slot (col % W) holds the b+1 band entries of one active column.  All
// global traffic happens at chunk boundaries (FACTOR_CHUNK columns): the
// columns entering during chunk q+1 are loaded into registers at the start of
// chunk q and staged in a double-buffered LDS area at its end; a column's
// stage slot is reused for its factor column, and the chunk's factor columns
// are written out (coalesced) at the start of the next chunk.  Inside a chunk
// each column costs two LDS-only barriers.  pairs[t] = (cc << 16) | r
// enumerates the rank-1 update triangle.
#define FACTOR_CHUNK 8

template <typename T, int NT>
__global__ void __launch_bounds__(NT)
    k_factor(int n, int b, T *__restrict__ band, T *__restrict__ rband, int *__restrict__ status, double tiny) {
  constexpr int CH = FACTOR_CHUNK;
  constexpr int MAXR = 8;  // per <= NT * MAXR because W <= NT
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int W = b + 1;
  const int npairs = b * (b + 1) / 2;
  const int per = CH * W;
  const int nreg = (per + NT - 1) / NT;
  T *win = reinterpret_cast<T *>(smem_raw);
  T *lc = win + (size_t)W * W;
  T *stage = lc + W;  // 2 * per
  unsigned *pairs = reinterpret_cast<unsigned *>(stage + 2 * per);
  __shared__ int bad;
  const int mac = blockIdx.x;
  const int tid = threadIdx.x;
  T *Lg = band + (size_t)mac * n * W;
  // row-major copy for the backward solve: rband[j*W + r] = L(j, j-r), written
  // as each column finishes; entries with j - r < 0 are zero
  T *Rg = rband ? rband + (size_t)mac * n * W : nullptr;
  if (Rg)
    for (int e = tid; e < min(b, n) * W; e += NT) {
      const int j = e / W, r = e - j * W;
      if (r > j) Rg[e] = T(0);
    }
  for (int t = tid; t < npairs; t += NT) {
    int cc = 1, rem = t;  // t -> (cc = 1..b, r = 0..b-cc), column-offset major
    while (rem >= b - cc + 1) {
      rem -= b - cc + 1;
      ++cc;
    }
    pairs[t] = ((unsigned)cc << 16) | (unsigned)rem;
  }
  for (int t = tid; t < W * W; t += NT) {
    const int c = t / W, r = t - c * W;
    win[t] = c < n ? Lg[(size_t)c * W + r] : T(0);
  }
  for (int e = tid; e < per; e += NT) {  // stage chunk 0: columns W .. W+CH-1
    const int col = W + e / W;
    stage[e] = col < n ? Lg[(size_t)col * W + (e % W)] : T(0);
  }
  if (tid == 0) bad = 0;
  T pf[MAXR];
  __syncthreads();
  int s = 0;  // slot of column j
  const int nchunks = (n + CH - 1) / CH;
  for (int q = 0; q < nchunks; ++q) {
    T *sb = stage + (q & 1) * per;
    if (q > 0) {
      // factor columns of chunk q-1 (held in the other stage buffer) -> HBM
      const T *src = stage + ((q - 1) & 1) * per;
      const size_t base = (size_t)(q - 1) * CH * W;
      for (int e = tid; e < per; e += NT) Lg[base + e] = src[e];
    }
    // columns entering during chunk q+1 (j + W for j in chunk q+1) -> registers
#pragma unroll
    for (int v = 0; v < MAXR; ++v) {
      const int e = tid + v * NT;
      const int col = (q + 1) * CH + W + e / W;
      pf[v] = (v < nreg && e < per && col < n) ? Lg[(size_t)col * W + (e % W)] : T(0);
    }
    const int jend = min(n, (q + 1) * CH);
    for (int j = q * CH; j < jend; ++j) {
      const int jj = j - q * CH;
      if (tid < W) {
        T d = win[s * W];
        if (!(d > T(tiny)) || !isfinite((double)d)) {
          if (tid == 0) bad = MHM_STATUS_PIVOT;
          d = T(1);
        }
        const T root = sqrt(d);
        lc[tid] = tid == 0 ? root : win[s * W + tid] / root;
      }
      lds_barrier();
      if (tid < W) {
        const T in = sb[jj * W + tid];
        win[s * W + tid] = in;       // column j+W enters the freed slot
        sb[jj * W + tid] = lc[tid];  // the stage slot now holds factor column j
        if (Rg && j + tid < n) Rg[(size_t)(j + tid) * W + tid] = lc[tid];
      }
      for (int t = tid; t < npairs; t += NT) {
        const unsigned p = pairs[t];
        const int cc = (int)(p >> 16), r = (int)(p & 0xffffu);
        int slot = s + cc;
        if (slot >= W) slot -= W;
        win[slot * W + r] -= lc[cc + r] * lc[cc];
      }
      lds_barrier();
      s = s + 1 == W ? 0 : s + 1;
    }
    if (q + 1 < nchunks) {
      T *dst = stage + ((q + 1) & 1) * per;
#pragma unroll
      for (int v = 0; v < MAXR; ++v) {
        const int e = tid + v * NT;
        if (v < nreg && e < per) dst[e] = pf[v];
      }
      lds_barrier();
    }
  }
  {  // factor columns of the last chunk
    const int q = nchunks - 1;
    const T *src = stage + (q & 1) * per;
    const size_t base = (size_t)q * CH * W;
    const int count = (n - q * CH) * W;
    for (int e = tid; e < count; e += NT) Lg[base + e] = src[e];
  }
  if (tid == 0 && bad) atomicOr(&status[mac], bad);
}

// ----------------------------------------------------------------- blocked band Cholesky
// Block per macro, panels of FPANEL columns, factored in place in an LDS ring
// of S2 = S + NB column slots, S = max(NB + b, 2 NB) (column c at slot c mod
// S2, entries r = 0..b of column c at win[slot * ldw + r], rows c + r).  At
// panel step k (columns j0 .. j0 + NB - 1) the ring holds panel k, the S - NB
// columns after it (so the next panel is always inside) and, in the NB slots
// of panel k - 1 (written out in step k - 1), the columns entering after
// panel k.
// Panel factorisation (one wavefront, in registers): lane q holds panel rows q
// and q + 64 (PR = NB + b <= 127 rows); left-looking, column t takes the
// updates of the finished columns k < t with their multipliers L(j0 + t,
// j0 + k) read back from the ring (LDS broadcast reads, in order within the
// wavefront), then is pivoted and scaled and written back to its slot; the
// pivot comes from its lane by v_readlane, 1/sqrt by the hardware estimate and
// Newton steps.  The step is branch-free: loads at clamped addresses with
// selects, a failed pivot flagged and replaced by 1, the columns of a short
// last panel padded with identity columns (which leave the others unchanged).
// One-panel look-ahead, one block barrier per panel.  While panel k is the
// source of the rank-NB update (C = C - L_I L_J^T over 16 x 16 tiles of the
// lower triangle, four 16x16x4 MFMA steps per tile, the C fragment read from
// and written back to the ring, the operands read from panel k's slots and
// masked to the band; ldw makes the C fragment accesses conflict-free in
// binary64):
//  - one wavefront (k mod 4, so the serial steps of the blocks sharing a CU
//    spread over its SIMDs) applies the tiles of the first tile column - the
//    columns of panel k + 1 - and then factors panel k + 1 in its slots, so
//    the serial panel step of k + 1 overlaps the rest of k's update (nobody
//    else reads or writes panel k + 1's columns during step k);
//  - the other three gather the NB columns entering after panel k (issued
//    first, consumed last), apply the other tile columns, write panel k to
//    HBM (the column band, contiguous; the row band, rband[R*W + r] =
//    L(R, R-r), runs of NB per row; with dblk set its 16 x 16 diagonal block,
//    packed (tri16), to dblk[(mac * nblocks + j0 / 16) * DBLK]) and store the
//    entering columns in panel k - 1's slots.
// The tiles of panel k + 1's columns are updated by every earlier panel before
// it is factored: panel k's by the serial wavefront in this step, earlier
// panels' by the other wavefronts in earlier steps.  Barriers order LDS only: no thread reads
// back another's HBM stores, and each wavefront waits for its own gather
// loads where it uses them.
#define FPANEL 16
// A 16 x 16 lower-triangular block, packed by rows: entry (rho, sigma), sigma
// <= rho, at rho (rho + 1) / 2 + sigma.
#define DBLK 136
__host__ __device__ constexpr int tri16(int rho, int sigma) { return rho * (rho + 1) / 2 + sigma; }

template <typename T>
struct Mfma16;
template <>
struct Mfma16<double> {  // v_mfma_f64_16x16x4f64: D[4r + lane/16][lane%16] in register r
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 step(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int r, int g) { return 4 * r + g; }
};
template <>
struct Mfma16<float> {  // v_mfma_f32_16x16x4f32: D[4(lane/16) + r][lane%16] in register r
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 step(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int r, int g) { return 4 * g + r; }
};
// (A operand: lane l holds A[l%16][l/16]; B operand: lane l holds B[l/16][l%16].)

// Ring column stride: in binary64 the C fragments of a 16 x 16 tile hit 64
// distinct banks when ldw - 1 = 2 mod 4; in binary32 an odd ldw - 1 = 3 mod 4
// leaves at most two-way conflicts.
static int panel_ldw(int b, size_t tsize) {
  int w = b + 1;
  while ((w & 3) != 3) ++w;
  if (tsize == 4 && (w & 1) == 0) ++w;
  return w;
}

// The ring: S + NB column slots, then a zero element and a write sink.
static size_t factor_panel_lds(int b, size_t tsize) {
  const size_t S = std::max((size_t)FPANEL + b, (size_t)2 * FPANEL);
  return ((S + FPANEL) * panel_ldw(b, tsize) + 2) * tsize;
}

static int lds_limit(int device);
static bool factor_panel_ok(int b, size_t tsize, int device) {
  return FPANEL + b <= 127 && factor_panel_lds(b, tsize) <= (size_t)lds_limit(device);
}

// Blocks per CU the kernel's registers allow: in binary64 three while W <= 72
// (52.5 KiB of ring at b = 66 lets three share a CU's LDS; wider rings fit
// two), four in binary32.
template <typename T, int NPF>
struct FactorPanelWaves {
  static constexpr int value = sizeof(T) == 8 ? (NPF <= 6 ? 3 : 2) : 4;
};

// NPF: entering elements per thread of the three gathering wavefronts, NB W
// <= 192 NPF (6: W <= 72; 10: W <= 120).  NARROW: b < NB, where a panel
// column's band ends inside the panel (the in-panel multipliers past it are
// zero and read as such).
template <typename T, int NPF, bool NARROW>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FactorPanelWaves<T, NPF>::value)))
    k_factor_panel(int n, int b, int ldw, T *__restrict__ band, T *__restrict__ rband,
                   int *__restrict__ status, double tiny, const T *__restrict__ sval, int nnz,
                   const int *__restrict__ bslot, double *__restrict__ dblk) {
  constexpr int NB = FPANEL;
  typedef typename Mfma16<T>::v4 V4;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int W = b + 1;
  const int S = max(NB + b, 2 * NB);  // window columns: the next panel is always inside
  const int S2 = S + NB;              // ring slots: the window and the entering columns
  const int PR = NB + b;              // panel rows
  T *win = reinterpret_cast<T *>(smem_raw);  // S2 x ldw, then the zero element and the sink
  const int zoff = S2 * ldw;                 // win[zoff] = 0: out-of-band operands read it
  const int sink = zoff + 1;                 // masked-off stores go here, so stores need no branch
  __shared__ int bad;
  const int mac = blockIdx.x, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  T *Lg = band + (size_t)mac * n * W;
  T *Rg = rband ? rband + (size_t)mac * n * W : nullptr;
  // the scaled, pinned operator (CSR slot order, written by k_prepare) at band
  // position (j, r) is sv[bslot[j*W + r]]; slot -1 (outside the pattern or
  // past row n - 1) reads as 0
  const T *sv = sval + (size_t)mac * nnz;
  auto slot_of = [&](int j, int r) -> int { return (j < n && j + r < n) ? bslot[(size_t)j * W + r] : -1; };
  // ring slot of the column x columns after the one in slot s (s, x < S2)
  auto ring = [&](int s, int x) -> int {
    const int v = s + x;
    return v >= S2 ? v - S2 : v;
  };
  const FastDiv divW(W);
  if (tid == 0) {
    bad = 0;
    win[zoff] = T(0);
  }
  // columns 0 .. S-1 in slots 0 .. S-1, eight elements per thread at a time:
  // all slot indices, then all values, then the stores (a loop of dependent
  // pairs would wait for two round trips per element)
  for (int e0 = 0; e0 < S * W; e0 += 8 * 256) {
    int qv[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      const int e = e0 + tid + 256 * v;
      const int c = divW(e), r = e - c * W;
      qv[v] = e < S * W ? slot_of(c, r) : -1;
    }
    T x[8];
#pragma unroll
    for (int v = 0; v < 8; ++v) x[v] = global_ptr(sv)[qv[v] >= 0 ? qv[v] : 0];
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      const int e = e0 + tid + 256 * v;
      const int c = divW(e), r = e - c * W;
      if (e < S * W) win[c * ldw + r] = qv[v] >= 0 ? x[v] : T(0);
    }
  }
  if (Rg)  // entries L(j, j-r) with j - r < 0
    for (int e = tid; e < min(b, n) * W; e += 256) {
      const int j = divW(e), r = e - j * W;
      if (r > j) Rg[e] = T(0);
    }
  const int nt = (b + 15) / 16;  // tile rows / columns of the trailing block
  const int ntile = nt * (nt + 1) / 2;

  // Factor the panel of columns jp .. jp + NB - 1 (one wavefront, slot sp of
  // column jp) in place, left-looking: column t takes the updates of the
  // finished columns k < t - the multipliers L(jp + t, jp + k), k < t - 1,
  // read back from the ring one step ahead (LDS broadcast reads off the
  // critical path), L(jp + t, jp + t - 1) from its lane by v_readlane - and is
  // then pivoted, scaled and stored back; the pivot also comes by v_readlane.
  // A wavefront's LDS accesses execute in order, so an entry stored in one
  // step is read back two steps later without a fence.
  auto factor_panel = [&](int jp, int sp) {
    const int np = min(NB, n - jp);
    T a0[NB], a1[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      const T *col = win + ring(sp, t) * ldw;
      // a laundered lane index: the masks are formed here, not hoisted out of
      // the panel loop as SGPR pairs
      int qa = lane;
      asm volatile("" : "+v"(qa));
      const int qb = qa + 64;
      const T v0 = col[min(max(qa - t, 0), b)];
      const T v1 = col[min(qb - t, b)];
      const bool pad = t >= np;  // identity column of a short last panel
      a0[t] = pad ? (qa == t ? T(1) : T(0)) : (qa >= t && qa - t <= b ? v0 : T(0));
      a1[t] = !pad && qb < PR && qb - t <= b ? v1 : T(0);
    }
    bool failed = false;
    // multiplier L(jp + t + 1, jp + k), k < t: an entry stored in an earlier
    // step (rows past the band of column k are zero)
    auto mult = [&](int t, int k) -> T {
      if (NARROW) {
        const T v = win[ring(sp, k) * ldw + min(t + 1 - k, b)];
        return t + 1 - k <= b ? v : T(0);
      }
      return win[ring(sp, k) * ldw + t + 1 - k];
    };
    // The next column's multipliers are read into m once the step's sums have
    // used it (reading them a step ahead into a second array, PRE, measured no
    // faster and costs the registers of a block per CU).
    constexpr bool PRE = false;
    T m[NB];  // m[k] = L(jp + t, jp + k), k < t - 1, for the current column t
    T mn[PRE ? NB : 1];
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      if constexpr (PRE) {
#pragma unroll
        for (int k = 0; k < t && t + 1 < NB; ++k) mn[k] = mult(t, k);
      }
      // two partial sums halve the dependent chain of the updates
      T s0a = T(0), s0b = T(0), s1a = T(0), s1b = T(0);
#pragma unroll
      for (int k = 0; k + 1 < t; ++k) {
        if (k & 1) {
          s0b += a0[k] * m[k];
          s1b += a1[k] * m[k];
        } else {
          s0a += a0[k] * m[k];
          s1a += a1[k] * m[k];
        }
      }
      a0[t] -= s0a + s0b;
      a1[t] -= s1a + s1b;
      // the next column's multipliers k < t, read into m now that the sums
      // have used it, so the reads overlap this step's pivot chain
      if constexpr (!PRE) {
#pragma unroll
        for (int k = 0; k < t && t + 1 < NB; ++k) m[k] = mult(t, k);
      }
      if (t > 0) {
        const T l = lane_bcast(a0[t - 1], t);  // L(jp + t, jp + t - 1), held by lane t
        a0[t] -= a0[t - 1] * l;
        a1[t] -= a1[t - 1] * l;
      }
      T d = lane_bcast(a0[t], t);  // pivot, held by lane t
      const bool ok = d > T(tiny) && isfinite((double)d);
      failed |= !ok;
      d = ok ? d : T(1);
      const T inv = rsqrt_nr(d);
      const T root = d * inv;
      int qt = lane;  // this step's masks are formed here, not hoisted
      asm volatile("" : "+v"(qt));
      a0[t] = qt > t ? a0[t] * inv : (qt == t ? root : T(0));
      a1[t] = a1[t] * inv;  // q1 >= 64 > t
      // back to the slot: entries r = row - t = 0..b of column t (other lanes
      // store to the sink: no branch splits the unrolled panel, so the next
      // column's sums can be scheduled into this column's pivot chain)
      const int colt = ring(sp, t) * ldw;
      win[qt >= t && qt - t <= b ? colt + qt - t : sink] = a0[t];
      win[qt + 64 - t <= b ? colt + qt + 64 - t : sink] = a1[t];
      if constexpr (PRE) {
#pragma unroll
        for (int k = 0; k < t && t + 1 < NB; ++k) m[k] = mn[k];
      }
    }
    if (failed && lane == 0) bad = MHM_STATUS_PIVOT;
  };

  // Up to 4 tiles (I[c], J[c]) of the rank-NB update by the panel at j0 (slot
  // s0), as independent MFMA chains: all C fragments and operands are read
  // first, then the steps of the tiles interleave.  Tile (I, J) covers
  // columns qc0 = NB + 16 J, rows qr0 = NB + 16 I of the window.  Operand
  // entry L(j0 + q, j0 + kk), kk = 4k + g16, is at ob[k] + q - kk (ob: the
  // slot base of the lane's panel column kk); past the band (q - kk > b; q >=
  // NB > kk, so never above it) the lane reads the zero element instead.
  auto update_tiles = [&](const int (&I)[4], const int (&J)[4], int count, int s0, const int (&ob)[4], int j16,
                          int g16) {
    V4 acc[4];
    T *colp[4];
    T av[4][4], bv[4][4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (c >= count) break;
      const int qr0 = NB + 16 * I[c], qc0 = NB + 16 * J[c];
      const int qc = qc0 + j16;  // this lane's column
      colp[c] = win + ring(s0, qc) * ldw;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int qr = qr0 + Mfma16<T>::row(r, g16);
        const bool in = qr >= qc && qr < PR && qc < PR;
        acc[c][r] = colp[c][in ? qr - qc : 0];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int kk = 4 * k + g16;
        const int ra = qr0 + j16 - kk, rb = qc0 + j16 - kk;
        av[c][k] = -win[ra <= b ? ob[k] + ra : zoff];  // A[i][kk] = -L(j0 + qr0 + i, j0 + kk)
        bv[c][k] = win[rb <= b ? ob[k] + rb : zoff];   // B[kk][j] = L(j0 + qc0 + j, j0 + kk)
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (c < count) acc[c] = Mfma16<T>::step(av[c][k], bv[c][k], acc[c]);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (c >= count) break;
      const int qr0 = NB + 16 * I[c], qc = NB + 16 * J[c] + j16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int qr = qr0 + Mfma16<T>::row(r, g16);
        const bool in = qr >= qc && qr < PR && qc < PR;
        win[in ? (int)(colp[c] - win) + qr - qc : sink] = acc[c][r];
      }
    }
  };

  // The serial wavefront's tiles (I, 0), I < nt - the columns of the next
  // panel - in one pass: the B operand (this lane's column of the next panel)
  // is read once, the NTC tile chains interleave, and tiles past nt (rows past
  // PR) read in bounds and store to the sink, so no branch splits the pass.
  constexpr int NTC = NARROW ? 1 : (NPF <= 6 ? 5 : 7);  // nt = ceil(b / 16) <= NTC
  auto update_col0 = [&](int s0, const int (&ob)[4], int j16, int g16) {
    const int qc = NB + j16;  // this lane's column
    const int cofs = ring(s0, qc) * ldw;
    V4 acc[NTC];
    T av[NTC][4], bv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int kk = 4 * k + g16, rb = qc - kk;
      bv[k] = win[rb <= b ? ob[k] + rb : zoff];  // B[kk][j] = L(j0 + NB + j, j0 + kk)
    }
#pragma unroll
    for (int c = 0; c < NTC; ++c) {
      const int qr0 = NB + 16 * c;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int qr = qr0 + Mfma16<T>::row(r, g16);
        const bool in = qr >= qc && qr < PR && qc < PR;
        acc[c][r] = win[in ? cofs + qr - qc : cofs];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int kk = 4 * k + g16, ra = qr0 + j16 - kk;
        av[c][k] = -win[ra <= b ? ob[k] + ra : zoff];  // A[i][kk] = -L(j0 + qr0 + i, j0 + kk)
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int c = 0; c < NTC; ++c) acc[c] = Mfma16<T>::step(av[c][k], bv[k], acc[c]);
#pragma unroll
    for (int c = 0; c < NTC; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int qr = NB + 16 * c + Mfma16<T>::row(r, g16);
        const bool in = qr >= qc && qr < PR && qc < PR;
        win[in ? cofs + qr - qc : sink] = acc[c][r];
      }
  };

  lds_barrier();  // the window holds columns 0 .. S-1
  if (wave == 0) factor_panel(0, 0);
  lds_barrier();
  T pf[NPF];
  int s0 = 0;  // slot of column j0
  for (int j0 = 0; j0 < n; j0 += NB, s0 = ring(s0, NB)) {
    const int np = min(NB, n - j0);
    // the lane's tile coordinates, laundered so that the per-lane offsets
    // derived from them are formed in each step instead of held (and
    // spilled) across the panel loop
    int ln = lane;
    asm volatile("" : "+v"(ln));
    const int j16 = ln & 15, g16 = ln >> 4;
    // slot bases of this lane's operand columns 4k + g16 of panel k
    int ob[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) ob[k] = ring(s0, 4 * k + g16) * ldw;
    // the serial role moves to the next wavefront every panel, so the serial
    // steps of the blocks sharing a CU spread over its four SIMDs
    const int pw = (j0 / NB) & 3;
    if (wave != pw) {
      // gather the columns j0 + S .. j0 + S + NB - 1, which enter after this
      // step: all slot indices, then all values, unconditional loads at
      // clamped positions masked afterwards (a branch per load waits for each)
      // entering element e = t W + r is band position (j0 + S + t, r), at
      // bslot[(j0 + S) W + e]; past row n - 1 the map holds -1
      const int u = ((wave - pw - 1) & 3) * 64 + lane;
      int qv[NPF];
      const int gbase = (j0 + S) * W;
#pragma unroll
      for (int v = 0; v < NPF; ++v) {
        const int e = u + 192 * v;
        const bool in = e < NB * W && gbase + e < n * W;
        const int raw = global_ptr(bslot)[in ? (size_t)gbase + e : 0];
        qv[v] = in ? raw : -1;
      }
#pragma unroll
      for (int v = 0; v < NPF; ++v) {
        const T val = global_ptr(sv)[qv[v] >= 0 ? qv[v] : 0];
        pf[v] = qv[v] >= 0 ? val : T(0);
      }
    }
    if (wave == pw) {
      // the next panel's columns first, then the next panel itself
      update_col0(s0, ob, j16, g16);
      if (j0 + NB < n) factor_panel(j0 + NB, ring(s0, NB));
    } else {
      const int ow = (wave - pw - 1) & 3;  // 0 .. 2 among the other wavefronts
      const int u = ow * 64 + lane;        // 0 .. 191
      // tile columns J >= 1, round-robin over the other wavefronts, four at a time
      for (int k0 = ow; k0 < ntile - nt; k0 += 12) {
        int I[4] = {0, 0, 0, 0}, J[4] = {1, 1, 1, 1}, count = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int k = k0 + 3 * c;
          if (k < ntile - nt) {
            int jj = 1, rem = k;
            while (rem >= nt - jj) {
              rem -= nt - jj;
              ++jj;
            }
            I[c] = jj + rem;
            J[c] = jj;
            count = c + 1;
          }
        }
        update_tiles(I, J, count, s0, ob, j16, g16);
      }
      // panel k to HBM: column band (contiguous, as the slots hold it) and row band
      for (int e = u; e < np * W; e += 192) {
        const int t = divW(e), r = e - t * W;
        Lg[(size_t)j0 * W + e] = win[ring(s0, t) * ldw + r];
      }
      if (Rg)
        for (int e = u; e < PR * NB; e += 192) {
          const int q = e / NB, t = e - q * NB;
          const int R = j0 + q, r = q - t;
          if (t < np && r >= 0 && r <= b && R < n) Rg[(size_t)R * W + r] = win[ring(s0, t) * ldw + r];
        }
      if (dblk)  // the diagonal block, packed (identity-padded past row n - 1)
        for (int e = u; e < NB * NB; e += 192) {
          const int rho = e >> 4, sig = e & 15;
          if (sig <= rho)
            dblk[((size_t)mac * ((n + NB - 1) / NB) + j0 / NB) * DBLK + tri16(rho, sig)] =
                rho - sig <= b ? (double)win[ring(s0, sig) * ldw + rho - sig] : 0.0;
        }
      // the entering columns take panel k - 1's slots (written out in step k - 1)
#pragma unroll
      for (int v = 0; v < NPF; ++v) {
        const int e = u + 192 * v;
        if (e < NB * W) {
          const int t = divW(e), r = e - t * W;
          win[ring(s0, S + t) * ldw + r] = pf[v];
        }
      }
    }
    lds_barrier();
  }
  if (tid == 0 && bad) atomicOr(&status[mac], bad);
}

// ----------------------------------------------------------------- wide-band Cholesky
// Bands too wide for the LDS window of k_factor_panel (NB + b > 127): the same
// blocked right-looking algorithm with the band array itself as the window,
// in pairs of panels of FPANEL columns (A at j0, B at j0 + NB):
//  (a) wavefront 0 reads panel A (PR = NB + b <= 256 rows, four per lane) from
//      the band, factors it in registers (pivots and multipliers by
//      v_readlane, as in k_factor_panel) and leaves it in LDS, transposed
//      (LpT[t * ldq + q] = L(j0 + q, j0 + t), rows relative to j0); then it
//      reads panel B, subtracts panel A's update from B's own columns (the
//      rank-16 update restricted to them, operands from LDS) and factors B
//      into LpT columns NB .. 2 NB - 1 (rows NB .. PR2 - 1, PR2 = 2 NB + b);
//  (b) both panels go back to the column band (and to the row band if one is
//      kept, and in binary64 their 16 x 16 diagonal blocks, packed (tri16)
//      and identity-padded past row n - 1, to dblk for the matrix-core solve);
//  (c) the block applies the rank-2NB update of the pair to the trailing
//      triangle in place in the band, 4x4 register tiles; consecutive threads
//      take consecutive row tiles of a tile column, so their band reads and
//      writes are contiguous and their LDS reads of LpT are conflict-free.
// The trailing triangle (about b^2 / 2 entries read and written per update)
// is touched once per pair instead of once per panel.  LpT entries outside
// the band (rows PR .. PR2 - 1 of A's columns, rows 0 .. NB - 1 of B's) are
// zeroed once and never written.  Block barriers order the in-place updates
// between pairs.
#define WIDE_ROWS 4  // panel rows per lane of wavefront 0

static int wide_ldq(int b) { return ((2 * FPANEL + b + 3) & ~3) + 4; }

static size_t factor_wide_lds(int b, size_t tsize) {
  return ((size_t)2 * FPANEL * wide_ldq(b) + 64) * tsize;
}

static bool factor_wide_ok(int b, size_t tsize, int limit) {
  return FPANEL + b <= 64 * WIDE_ROWS && factor_wide_lds(b, tsize) <= (size_t)limit;
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_factor_wide(int n, int b, int ldq, T *__restrict__ band, T *__restrict__ rband, int *__restrict__ status,
                  double tiny, double *__restrict__ dblk) {
  constexpr int NB = FPANEL, NR = WIDE_ROWS;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int W = b + 1, PR = NB + b, PR2 = 2 * NB + b;
  T *LpT = reinterpret_cast<T *>(smem_raw);  // 2 NB x ldq
  __shared__ int bad;
  const int mac = blockIdx.x, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  T *Lg = band + (size_t)mac * n * W;
  T *Rg = rband ? rband + (size_t)mac * n * W : nullptr;
  if (tid == 0) bad = 0;
  for (int e = tid; e < 2 * NB * ldq; e += 256) LpT[e] = T(0);
  if (Rg)  // entries L(j, j-r) with j - r < 0
    for (int e = tid; e < min(b, n) * W; e += 256) {
      const int j = e / W, r = e - j * W;
      if (r > j) Rg[e] = T(0);
    }
  const int side = (b + 3) / 4;  // tile rows/columns of the trailing triangle
  const int ntiles = side * (side + 1) / 2;
  __syncthreads();
  // Wavefront 0: factor the panel of columns jc .. jc + np - 1 (rows jc + q,
  // q = lane + 64 v < PR) into LpT columns t0 + t, rows q0 + q.  With upd,
  // first subtract the update of the panel already in LpT columns 0 .. NB - 1
  // (rows relative to jc - NB).
  auto factor_rows = [&](int jc, int np, int t0, int q0, bool upd) {
    T a[NR][NB];
#pragma unroll
    for (int v = 0; v < NR; ++v)
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        const int q = lane + 64 * v;
        a[v][t] = (t < np && q >= t && q - t <= b && jc + q < n) ? Lg[(size_t)(jc + t) * W + (q - t)] : T(0);
      }
    if (upd) {
#pragma unroll 1
      for (int k = 0; k < NB; ++k) {
        const T *Ak = LpT + k * ldq + NB;  // L(jc + q, jc - NB + k) at Ak[q]
        T lk[NR];
#pragma unroll
        for (int v = 0; v < NR; ++v) lk[v] = (lane + 64 * v < PR) ? Ak[lane + 64 * v] : T(0);
#pragma unroll
        for (int t = 0; t < NB; ++t) {
          const T m = Ak[t];  // L(jc + t, jc - NB + k), the same for every lane
#pragma unroll
          for (int v = 0; v < NR; ++v) a[v][t] -= lk[v] * m;
        }
      }
    }
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      if (t < np) {
        T d = lane_bcast(a[0][t], t);  // pivot, held by lane t
        if (!(d > T(tiny)) || !isfinite((double)d)) {
          if (lane == 0) bad = MHM_STATUS_PIVOT;
          d = T(1);
        }
        const T inv = rsqrt_nr(d);
        const T root = d * inv;
        a[0][t] = lane > t ? a[0][t] * inv : (lane == t ? root : T(0));
#pragma unroll
        for (int v = 1; v < NR; ++v) a[v][t] = a[v][t] * inv;  // rows >= 64 > t
#pragma unroll
        for (int v = 0; v < NR; ++v)
          if (lane + 64 * v < PR) LpT[(t0 + t) * ldq + q0 + lane + 64 * v] = a[v][t];
#pragma unroll
        for (int u = t + 1; u < NB; ++u) {
          const T l = lane_bcast(a[0][t], u);  // L(jc + u, jc + t), held by lane u
#pragma unroll
          for (int v = 0; v < NR; ++v) a[v][u] -= a[v][t] * l;
        }
      } else {
#pragma unroll
        for (int v = 0; v < NR; ++v)
          if (lane + 64 * v < PR) LpT[(t0 + t) * ldq + q0 + lane + 64 * v] = T(0);
      }
    }
  };
  for (int j0 = 0; j0 < n; j0 += 2 * NB) {
    const int npA = min(NB, n - j0);
    const int npB = max(0, min(NB, n - j0 - NB));
    const int np2 = npA + npB;
    // (a) both panels, wavefront 0 (its LDS accesses execute in order, so B
    // reads A's columns without a barrier)
    if (wave == 0) {
#pragma unroll 1
      for (int p = 0; p < 2; ++p) factor_rows(j0 + p * NB, p ? npB : npA, p * NB, p * NB, p == 1);
    }
    __syncthreads();
    // (b) the panels to the column band, the row band and the diagonal blocks
    for (int e = tid; e < np2 * W; e += 256) {
      const int t = e / W, r = e - t * W;
      Lg[(size_t)j0 * W + e] = LpT[t * ldq + t + r];
    }
    if (Rg)
      for (int e = tid; e < PR2 * 2 * NB; e += 256) {
        const int q = e / (2 * NB), t = e - q * (2 * NB);
        const int R = j0 + q, r = q - t;
        if (t < np2 && r >= 0 && r <= b && R < n) Rg[(size_t)R * W + r] = LpT[t * ldq + q];
      }
    if (dblk)
      for (int e = tid; e < 2 * NB * NB; e += 256) {
        const int p = e >> 8, rho = (e >> 4) & 15, sig = e & 15;
        const int np = p ? npB : npA;
        if (sig <= rho && (p == 0 || npB > 0))
          dblk[((size_t)mac * ((n + NB - 1) / NB) + j0 / NB + p) * DBLK + tri16(rho, sig)] =
              rho < np ? (double)LpT[(p * NB + sig) * ldq + p * NB + rho] : (rho == sig ? 1.0 : 0.0);
      }
    // (c) trailing update: columns qc in [2 NB, PR2), rows qr in [qc, PR2)
    // with j0 + qr < n; tiles numbered down the tile columns of the lower
    // triangle
    for (int id = tid; id < ntiles; id += 256) {
      const float s2 = 2.0f * side + 1.0f;
      int tc = (int)((s2 - sqrtf(s2 * s2 - 8.0f * id)) * 0.5f);
      tc = max(0, min(tc, side - 1));
      while (tc > 0 && tc * side - tc * (tc - 1) / 2 > id) --tc;
      while (tc + 1 < side && (tc + 1) * side - (tc + 1) * tc / 2 <= id) ++tc;
      const int tr = tc + (id - (tc * side - tc * (tc - 1) / 2));
      const int qr0 = 2 * NB + 4 * tr, qc0 = 2 * NB + 4 * tc;
      if (j0 + qr0 >= n) continue;  // every row of the tile is past the matrix
      T acc[4][4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const T *col = Lg + (size_t)(j0 + qc0 + j) * W - (qc0 + j);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int qr = qr0 + i, qc = qc0 + j;
          acc[i][j] = (qr < PR2 && qr >= qc && j0 + qr < n) ? col[qr] : T(0);
        }
      }
#pragma unroll 4
      for (int t = 0; t < 2 * NB; ++t) {
        T ar[4], bc[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) ar[i] = LpT[t * ldq + qr0 + i];
#pragma unroll
        for (int j = 0; j < 4; ++j) bc[j] = LpT[t * ldq + qc0 + j];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] -= ar[i] * bc[j];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        T *col = Lg + (size_t)(j0 + qc0 + j) * W - (qc0 + j);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int qr = qr0 + i, qc = qc0 + j;
          if (qr < PR2 && qr >= qc && j0 + qr < n) col[qr] = acc[i][j];
        }
      }
    }
    __syncthreads();
  }
  if (tid == 0 && bad) atomicOr(&status[mac], bad);
}

// ----------------------------------------------------------------- band solves
// Block per macro: forward L z = y (axpy form) then backward L^T x = z (dot
// form, x_j = (z_j - L(j+1..j+b, j)^T x) / L_jj) for ncols columns of the
// row-major y[mac][row][ldy].  The (b+1)-row window lives in registers:
// thread t owns column t % ncols and window slots g, g + G, ... (g = t /
// ncols, G = NT / ncols); a slot holds one row (row % (b+1)).  One LDS-only
// barrier per row.  All global traffic happens at chunk boundaries
// (SOLVE_CHUNK rows): band columns and incoming rows are loaded one chunk
// ahead into registers and staged in double-buffered LDS; finished rows are
// collected in an LDS buffer and written out (coalesced) during the next
// chunk.  Skips macros whose every column has converged.
#define SOLVE_CHUNK 16

template <typename T, int NT, int MAXR>
__device__ __forceinline__ void stage_issue(T (&regs)[MAXR], int nreg, const T *Lg, const T *Y, int W,
                                            int ldy, int ncols, int first_col, int first_row, int dir, int n) {
  // stage element e < CH*W: band column first_col + dir*(e / W), entry e % W;
  // element CH*W + e': row first_row + dir*(e' / ncols), column e' % ncols
  const int ncolvals = SOLVE_CHUNK * W;
  const int total = ncolvals + SOLVE_CHUNK * ncols;
#pragma unroll
  for (int q = 0; q < MAXR; ++q) {
    const int e = threadIdx.x + q * NT;
    T v = T(0);
    if (q < nreg) {
      if (e < ncolvals) {
        const int jj = e / W, r = e - jj * W;
        const int col = first_col + dir * jj;
        if (col >= 0 && col < n) v = Lg[(size_t)col * W + r];
      } else if (e < total) {
        const int e2 = e - ncolvals;
        const int jj = e2 / ncols, c = e2 - jj * ncols;
        const int row = first_row + dir * jj;
        if (row >= 0 && row < n) v = Y[(size_t)row * ldy + c];
      }
    }
    regs[q] = v;
  }
}

template <typename T, int NT, int MAXR>
__device__ __forceinline__ void stage_store(const T (&regs)[MAXR], int nreg, T *stage, int W, int ncols) {
  const int total = SOLVE_CHUNK * (W + ncols);
#pragma unroll
  for (int q = 0; q < MAXR; ++q) {
    const int e = threadIdx.x + q * NT;
    if (q < nreg && e < total) stage[e] = regs[q];
  }
}

// Rows first_row + dir*jj (jj < count) of the chunk output buffer -> Y.
template <typename T, int NT>
__device__ __forceinline__ void rows_out(const T *buf, T *Y, int ldy, int ncols, int first_row, int dir,
                                         int count) {
  for (int e = threadIdx.x; e < count * ncols; e += NT) {
    const int jj = e / ncols, c = e - jj * ncols;
    Y[(size_t)(first_row + dir * jj) * ldy + c] = buf[e];
  }
}

template <typename T, int NT, int MAXS>
__global__ void __launch_bounds__(NT)
    k_solve(int n, int b, int ldy, int ncols, const T *__restrict__ band, T *__restrict__ y,
            const int *__restrict__ colstate, int *__restrict__ iters) {
  constexpr int CH = SOLVE_CHUNK;
  constexpr int MAXR = 16;  // staged values per thread: CH*(W+ncols) <= NT*MAXR
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int mac = blockIdx.x;
  const int tid = threadIdx.x;
  __shared__ int active;
  if (colstate) {
    if (tid == 0) active = 0;
    __syncthreads();
    if (tid < ncols && colstate[(size_t)mac * ncols + tid] == 0) active = 1;
    __syncthreads();
    if (!active) return;
  }
  if (iters && tid == 0) iters[mac] += 1;
  const int W = b + 1;
  const int G = NT / ncols;
  const int g = tid / ncols, c = tid - g * ncols;
  const bool on = g < G;
  const int cnt = on ? (W - g + G - 1) / G : 0;  // owned slots: g + G*u, u < cnt
  const int per = CH * (W + ncols);               // staged values per chunk
  const int nreg = (per + NT - 1) / NT;
  T *stage = reinterpret_cast<T *>(smem_raw);     // 2 * per: [CH][W] columns, then [CH][ncols] rows
  T *outb = stage + 2 * per;                      // 2 * CH * ncols finished rows
  T *part = outb + 2 * CH * ncols;                // 2 * G * ncols
  const T *Lg = band + (size_t)mac * n * W;
  T *Y = y + (size_t)mac * n * ldy;
  T reg[MAXS];
  T pf[MAXR];
  const int nchunks = (n + CH - 1) / CH;
  // ---------------- forward: chunk q covers rows q*CH .. q*CH+CH-1
#pragma unroll
  for (int u = 0; u < MAXS; ++u) {
    const int row = g + G * u;
    reg[u] = (u < cnt && row < n) ? Y[(size_t)row * ldy + c] : T(0);
  }
  stage_issue<T, NT, MAXR>(pf, nreg, Lg, Y, W, ldy, ncols, 0, W, 1, n);
  stage_store<T, NT, MAXR>(pf, nreg, stage, W, ncols);
  __syncthreads();
  int s = 0;
  for (int q = 0; q < nchunks; ++q) {
    const T *sb = stage + (q & 1) * per;
    T *ob = outb + (q & 1) * CH * ncols;
    if (q > 0) rows_out<T, NT>(outb + ((q - 1) & 1) * CH * ncols, Y, ldy, ncols, (q - 1) * CH, 1, CH);
    stage_issue<T, NT, MAXR>(pf, nreg, Lg, Y, W, ldy, ncols, (q + 1) * CH, (q + 1) * CH + W, 1, n);
    const int jend = min(n, (q + 1) * CH);
    for (int j = q * CH; j < jend; ++j) {
      const int jj = j - q * CH;
      const T *col = sb + jj * W;
      // phase 1: owner of (slot s, column c) finishes row j; row j+W enters the slot
      if (on && (s % G) == g) {
        const int u = s / G;
        const T in = sb[CH * W + jj * ncols + c];
        T z = T(0);
#pragma unroll
        for (int v = 0; v < MAXS; ++v)
          if (v == u) z = reg[v];
        z = z / col[0];
        ob[jj * ncols + c] = z;
#pragma unroll
        for (int v = 0; v < MAXS; ++v)
          if (v == u) reg[v] = in;
      }
      lds_barrier();
      // phase 2: every other slot holds rows j+1..j+b: subtract L(row, j) z_j
      if (on) {
        const T p = ob[jj * ncols + c];
        T coef[MAXS];
#pragma unroll
        for (int u = 0; u < MAXS; ++u) {
          int r = g + G * u - s;
          r = r < 0 ? r + W : r;
          r = u < cnt ? r : 0;
          coef[u] = r != 0 ? col[r] : T(0);
        }
#pragma unroll
        for (int u = 0; u < MAXS; ++u) reg[u] -= coef[u] * p;
      }
      s = s + 1 == W ? 0 : s + 1;
    }
    stage_store<T, NT, MAXR>(pf, nreg, stage + ((q + 1) & 1) * per, W, ncols);
    lds_barrier();
  }
  {
    const int q = nchunks - 1;
    rows_out<T, NT>(outb + (q & 1) * CH * ncols, Y, ldy, ncols, q * CH, 1, n - q * CH);
  }
  __syncthreads();
  // ---------------- backward: step k handles row j = n-1-k; registers hold finished x rows
#pragma unroll
  for (int u = 0; u < MAXS; ++u) reg[u] = T(0);
  stage_issue<T, NT, MAXR>(pf, nreg, Lg, Y, W, ldy, ncols, n - 1, n - 1, -1, n);
  stage_store<T, NT, MAXR>(pf, nreg, stage, W, ncols);
  __syncthreads();
  int sj = (n - 1) % W;
  for (int q = 0; q < nchunks; ++q) {
    const T *sb = stage + (q & 1) * per;
    T *ob = outb + (q & 1) * CH * ncols;
    if (q > 0)
      rows_out<T, NT>(outb + ((q - 1) & 1) * CH * ncols, Y, ldy, ncols, n - 1 - (q - 1) * CH, -1, CH);
    stage_issue<T, NT, MAXR>(pf, nreg, Lg, Y, W, ldy, ncols, n - 1 - (q + 1) * CH, n - 1 - (q + 1) * CH, -1, n);
    const int kend = min(n, (q + 1) * CH);
    for (int k = q * CH; k < kend; ++k) {
      const int j = n - 1 - k;
      const int kk = k - q * CH;
      const T *col = sb + kk * W;
      // phase 1: partial dot over owned finished rows j+1..j+b
      if (on) {
        T coef[MAXS];
#pragma unroll
        for (int u = 0; u < MAXS; ++u) {
          int r = g + G * u - sj;
          r = r < 0 ? r + W : r;
          r = u < cnt ? r : 0;
          coef[u] = r != 0 ? col[r] : T(0);
        }
        T acc = T(0);
#pragma unroll
        for (int u = 0; u < MAXS; ++u) acc += coef[u] * reg[u];
        part[((k & 1) * G + g) * ncols + c] = acc;
      }
      lds_barrier();
      // phase 2: owner of (slot sj, column c) completes x_j
      if (on && (sj % G) == g) {
        T sum = T(0);
        for (int l = 0; l < G; ++l) sum += part[((k & 1) * G + l) * ncols + c];
        const T x = (sb[CH * W + kk * ncols + c] - sum) / col[0];
        const int u = sj / G;
#pragma unroll
        for (int v = 0; v < MAXS; ++v)
          if (v == u) reg[v] = x;
        ob[kk * ncols + c] = x;
      }
      sj = sj == 0 ? W - 1 : sj - 1;
    }
    stage_store<T, NT, MAXR>(pf, nreg, stage + ((q + 1) & 1) * per, W, ncols);
    lds_barrier();
  }
  {
    const int q = nchunks - 1;
    rows_out<T, NT>(outb + (q & 1) * CH * ncols, Y, ldy, ncols, n - 1 - q * CH, -1, n - q * CH);
  }
}

// Band solves, wavefront per macro (SOLVE_WAVES macros per block): lane c owns
// right-hand-side column c (ncols <= 64).  Both sweeps are right-looking:
// forward L z = y visits rows 0..n-1 and uses column j of L (band), backward
// L^T x = z visits rows n-1..0 and uses row j of L (rband); a step finishes
// the pending row at the head of the window (v = w_0 / L_jj) and updates the
// next b rows (w_r -= L_r v).  The window of pending rows is a shift register
// of MAXW >= b+1 values per lane in VGPRs (compile-time width, so every
// register index is static): a step is one division and MAXW fused
// multiply-adds against the band values, broadcast from LDS, with no
// cross-lane traffic and no barrier.  Band values are read from LDS in small
// groups, one group ahead of the arithmetic; MEM_FENCE keeps the
// compiler from hoisting more reads and pin() from sinking a step's updates
// past the next step (either would hold all staged band values in registers).
// The band rows of a chunk of steps are loaded one chunk ahead (coalesced,
// into registers) and staged in a double-buffered per-wave LDS area whose
// entries r >= b+1 stay zero; incoming and finished rows move between HBM and
// registers once per chunk.
#define SOLVE_WAVES 4

// Band values per LDS read group (two groups are in flight).
template <typename T, int MAXW>
constexpr int wave_group() {
  return sizeof(T) * MAXW > 256 ? 4 : 8;
}

// Steps per chunk: the chunk's incoming, next and finished rows (3 * CH values
// per lane) share the register file with the MAXW-value window.
template <typename T, int MAXW>
constexpr int wave_chunk() {
  return sizeof(T) * MAXW > 256 ? 4 : 8;
}


template <typename T, int N>
__device__ __forceinline__ void pin(T (&a)[N]) {
#pragma unroll
  for (int k = 0; k < N; ++k) asm volatile("" : "+v"(a[k]));
}

// A wave-uniform pointer held in SGPRs.  Passing the row base through
// readfirstlane hides its affine form from loop strength reduction, so an access
// base[lane] becomes one saddr load/store instead of a 64-bit VGPR pointer per
// unrolled access.
template <typename P>
__device__ __forceinline__ __attribute__((address_space(1))) P *uniform_ptr(P *p) {
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (__attribute__((address_space(1))) P *)(((unsigned long long)hi << 32) | lo);
}

// One sweep over the n rows of Y (ldy columns per row); step k works on row
// k (DIR > 0) or n-1-k (DIR < 0) with band values Lg[row*W + r].
template <typename T, int MAXW, int DIR>
__device__ __forceinline__ void band_sweep(int n_in, int W, int ldy, bool on, unsigned cl, int lane, const T *Lg,
                                           T *Y, T *stage) {
  // a laundered copy of n: the two sweeps' row-range tests must not be shared
  // (a shared test result would stay live, spilled, through the first sweep)
  int n = n_in;
  asm volatile("" : "+s"(n));
  constexpr int CH = wave_chunk<T, MAXW>();
  constexpr int G = wave_group<T, MAXW>();
  constexpr int NG = (MAXW + G - 1) / G;       // band-value groups per step
  constexpr int NPF = (CH * MAXW + 63) / 64;  // staged band values per lane and chunk
  const int nchunks = (n + CH - 1) / CH;
  auto row_of = [&](int k) { return DIR > 0 ? k : n - 1 - k; };
  T pf[NPF];
  // stage position of the lane's element e = lane + 64 v of a chunk (e = jj*W + r
  // for the band row of step jj), or -1 past the chunk
  int spos[NPF];
#pragma unroll
  for (int v = 0; v < NPF; ++v) {
    const int e = lane + 64 * v;
    const int jj = e / W;
    spos[v] = e < CH * W ? (DIR > 0 ? jj : CH - 1 - jj) * MAXW + (e - jj * W) : -1;
  }
  // band rows of steps q*CH .. q*CH+CH-1 -> registers; they are contiguous in
  // HBM from row lo, and rows outside [0, n) are the elements outside [elo, ehi)
  auto issue = [&](int q) {
    const int lo = DIR > 0 ? q * CH : n - (q + 1) * CH;
    const int elo = max(0, -lo) * W;
    const int ehi = min(CH, n - lo) * W;
#pragma unroll
    for (int v = 0; v < NPF; ++v) {
      const int e = lane + 64 * v;
      auto base = uniform_ptr(Lg + (ptrdiff_t)lo * W + 64 * v);
      pf[v] = (e >= elo && e < ehi) ? base[lane] : T(0);
    }
  };
  // registers -> stage slots (entries r < W)
  auto commit = [&](int buf) {
    T *sb = stage + buf * CH * MAXW;
#pragma unroll
    for (int v = 0; v < NPF; ++v)
      if (spos[v] >= 0) sb[spos[v]] = pf[v];
  };
  auto fetch = [&](T(&dst)[G], const T *col, int g) {
#pragma unroll
    for (int k = 0; k < G; ++k) dst[k] = (g * G + k < MAXW) ? col[g * G + k] : T(0);
  };
  auto load_row = [&](int k) {  // row of step k, zero past the end
    return k < n ? uniform_ptr(Y + (size_t)row_of(k) * ldy)[cl] : T(0);
  };
  T win[MAXW];
#pragma unroll
  for (int r = 0; r < MAXW; ++r) {
    if (r % 8 == 0) MEM_FENCE();  // at most 8 row pointers live at once
    win[r] = load_row(r);
  }
  T rin[CH], rnext[CH], done[CH];
#pragma unroll
  for (int i = 0; i < CH; ++i) rin[i] = load_row(MAXW + i);
  T cb0[G], cb1[G];
  issue(0);
  commit(0);
  wave_lds_sync();
  for (int q = 0; q < nchunks; ++q) {
    const int k0 = q * CH;
    issue(q + 1);
#pragma unroll
    for (int i = 0; i < CH; ++i) rnext[i] = load_row(k0 + CH + MAXW + i);
    const T *sb = stage + (q & 1) * CH * MAXW;
    fetch(cb0, sb, 0);
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      T v = T(0);
#pragma unroll
      for (int g = 0; g < NG; ++g) {
        T(&cur)[G] = ((i * NG + g) & 1) ? cb1 : cb0;
        T(&nxt)[G] = ((i * NG + g) & 1) ? cb0 : cb1;
        MEM_FENCE();
        if (g + 1 < NG)
          fetch(nxt, sb + i * MAXW, g + 1);
        else if (i + 1 < CH)
          fetch(nxt, sb + (i + 1) * MAXW, 0);
        MEM_FENCE();
        if (g == 0) v = win[0] / cur[0];
#pragma unroll
        for (int k = 0; k < G; ++k) {
          const int e = g * G + k;  // w_{e-1} <- w_e - L_e v
          if (e >= 1 && e < MAXW) win[e - 1] = win[e] - cur[k] * v;
        }
      }
      win[MAXW - 1] = rin[i];
      done[i] = v;
      pin(win);
    }
    // finished rows -> HBM after the chunk, so the step loop stays one basic block
#pragma unroll
    for (int i = 0; i < CH; ++i)
      if (k0 + i < n) {
        auto dst = uniform_ptr(Y + (size_t)row_of(k0 + i) * ldy);
        if (on) dst[cl] = done[i];
      }
    commit((q + 1) & 1);
    wave_lds_sync();  // other lanes read the staged band values
#pragma unroll
    for (int i = 0; i < CH; ++i) rin[i] = rnext[i];
  }
}

template <typename T, int MAXW>
__global__ void __launch_bounds__(64 * SOLVE_WAVES) __attribute__((amdgpu_waves_per_eu(2)))
    k_solve_wave(int n, int b, int ldy, int ncols, int macros, const T *__restrict__ band,
                 const T *__restrict__ rband, T *__restrict__ y, const int *__restrict__ colstate,
                 int *__restrict__ iters) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  constexpr int CH = wave_chunk<T, MAXW>();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int mac = __builtin_amdgcn_readfirstlane(blockIdx.x * SOLVE_WAVES + wave);
  if (mac >= macros) return;
  const int W = b + 1;
  const bool on = lane < ncols;
  if (colstate) {
    const bool pending = on && colstate[(size_t)mac * ncols + lane] == 0;
    if (!__any(pending)) return;
  }
  if (iters && lane == 0) iters[mac] += 1;
  T *stage = reinterpret_cast<T *>(smem_raw) + (size_t)wave * 2 * CH * MAXW;  // [2][CH][MAXW]
  for (int e = lane; e < 2 * CH * MAXW; e += 64) stage[e] = T(0);
  // lanes past ncols read column 0 (in bounds) and store nothing
  const unsigned cl = on ? (unsigned)lane : 0u;
  T *Y = y + (size_t)mac * n * ldy;
  band_sweep<T, MAXW, 1>(n, W, ldy, on, cl, lane, band + (size_t)mac * n * W, Y, stage);
  // a lane reads back only the rows it wrote itself
  band_sweep<T, MAXW, -1>(n, W, ldy, on, cl, lane, rband + (size_t)mac * n * W, Y, stage);
}

// ----------------------------------------------------------------- blocked solve on the matrix cores
// L z = y, then L^T x = z in binary64, one wavefront per macro, in blocks of 16
// rows; dinv holds the inverse of every 16 x 16 diagonal block of L (the
// factor writes the blocks, k_diag_inverse inverts them in place).  With the column band
// (band[j*W + r] = L(j+r, j)) for the forward and the row band
// (rband[i*W + r] = L(i, i-r)) for the backward sweep, a block I is
//   forward:  z_I = Dinv_I (y_I - L(I, J) z_J),        J = the 4 nk rows above I
//   backward: x_I = Dinv_I^T (z_I - L(K, I)^T x_K),   K = the 4 nk rows below I
// with nk = ceil(b / 4).  The coupling product is nk v_mfma_f64_16x16x4f64
// steps per 16-column tile (A = L, 16 rows x 4 window rows; B = 4 window rows
// x 16 columns), the diagonal block four more (B = y_I - product).  In binary64 the MFMA result
// layout (row 4r + lane/16 of the tile in register r) is the B operand layout
// of the k-step of rows 4r .. 4r+3, so a solved block enters the register
// window of the next block as it is: window slot t holds one k-step for every
// column tile, the newest four slots the block just solved.  Every lane loads
// distinct band values (16 consecutive entries per 16-lane group); the band
// operands of the next block are loaded during this block's products, the
// right-hand side and the diagonal inverse at the start of the block and used
// after them.  Each element of y is read and written by the same lane in both
// sweeps.  The storage type S of the bands and of y is that of the factor:
// binary64, or binary32 for the mixed path, whose values are widened to
// binary64 where the products use them (the prefetched operands stay binary32
// in registers, so converting does not wait for their loads); the arithmetic
// and the diagonal-block inverses are binary64 either way.
#define SOLVE_MFMA_WAVES 4


// Inverses of the 16 x 16 diagonal blocks of L in place: dinv[(mac * nblk +
// q) * DBLK] holds the block (rows and columns 16 q .. 16 q + 15, packed lower
// triangle, identity past row n - 1, written by the LDS-window factor) and
// receives its inverse, packed the same way.  A workgroup takes 16 blocks: thread (block, rho) stages row rho
// of its block and 1 / L(rho, rho) in LDS, then lane sigma of the block's 16
// forms column sigma of the inverse by column-oriented forward substitution.
__global__ void __launch_bounds__(256) k_diag_inverse(long nblocks, double *__restrict__ dinv) {
  __shared__ double sL[16][16][17];  // [block][column tau][L(tau + r, tau), r < 16; 1 / L(tau, tau)]
  const int lb = threadIdx.x >> 4, sigma = threadIdx.x & 15;
  const long gb = (long)blockIdx.x * 16 + lb;
  const bool on = gb < nblocks;
  double *blk = dinv + (size_t)(on ? gb : 0) * DBLK;
  {
    const int rho = sigma;  // this thread stages row rho
    double row[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) row[t] = on && t <= rho ? blk[tri16(rho, t)] : (t == rho ? 1.0 : 0.0);
#pragma unroll
    for (int t = 0; t < 16; ++t)
      if (t <= rho) sL[lb][t][rho - t] = row[t];
    sL[lb][rho][16] = 1.0 / row[rho];
  }
  __syncthreads();
  if (!on) return;
  double acc[16];
#pragma unroll
  for (int rho = 0; rho < 16; ++rho) acc[rho] = rho == sigma ? 1.0 : 0.0;
#pragma unroll
  for (int tau = 0; tau < 16; ++tau) {
    if ((tau & 3) == 0) MEM_FENCE();  // four columns of the block in registers at a time
    const double *c = sL[lb][tau];
    const double z = acc[tau] * c[16];
    if (tau >= sigma) blk[tri16(tau, sigma)] = z;  // z = 0 above the diagonal
#pragma unroll
    for (int rho = tau + 1; rho < 16; ++rho) acc[rho] -= c[rho - tau] * z;
  }
}

template <typename S, int MAXK, int NCT, int DIR>
__device__ __forceinline__ void mfma_sweep(int n, int b, int ldy, int ncols, const S *__restrict__ Lb,
                                           const double *__restrict__ Di, S *__restrict__ Y, int lane,
                                           int cbase = 0) {
  typedef Mfma16<double>::v4 V4;
  const int W = b + 1, nblk = (n + 15) / 16;
  const int a16 = lane & 15, g = lane >> 4;
  // operand A of window slot t for block q: row a16 of the block against
  // window row g of slot t, zero outside the band and the matrix (loaded
  // values are not touched before the product, so the loads stay in flight)
  auto loadA = [&](int q, int t) -> S {
    const int i0 = 16 * q;
    if (DIR > 0) {
      const int j = i0 - 4 * (MAXK - t) + g, rr = a16 + 4 * (MAXK - t) - g;
      return *((j >= 0 && rr <= b && i0 + a16 < n) ? Lb + (size_t)j * W + rr : zero_of<S>());
    }
    const int i = i0 + 16 + 4 * (MAXK - 1 - t) + g, rr = 16 + 4 * (MAXK - 1 - t) + g - a16;
    return *((i < n && rr <= b) ? Lb + (size_t)i * W + rr : zero_of<S>());
  };
  S A[MAXK];
  double win[MAXK][NCT];
#pragma unroll
  for (int t = 0; t < MAXK; ++t) {
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) win[t][ct] = 0.0;
  }
  // (slots below MAXK - ceil(b / 4) lie outside the band and read zeros)
  {
    const int q = DIR > 0 ? 0 : nblk - 1;
#pragma unroll
    for (int t = 0; t < MAXK; ++t) A[t] = loadA(q, t);
  }
  for (int s = 0; s < nblk; ++s) {
    const int q = DIR > 0 ? s : nblk - 1 - s;
    const bool more = s + 1 < nblk;
    const int qn = more ? (DIR > 0 ? q + 1 : q - 1) : q;  // the last block reloads its own operands
    const int i0 = 16 * q;
    // right-hand side rows and the diagonal inverse of this block
    S rhs[NCT][4];
    double dv[4];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) {
      const int col = cbase + 16 * ct + a16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = i0 + 4 * r + g;
        rhs[ct][r] = *((row < n && col < ncols) ? Y + (size_t)row * ldy + col : zero_of<S>());
      }
    }
    const double *dq = Di + (size_t)q * DBLK;
#pragma unroll
    for (int k = 0; k < 4; ++k) {  // Dinv[a16][4k+g] forward, Dinv[4k+g][a16] backward (zero above the diagonal)
      const int kk = 4 * k + g;
      dv[k] = DIR > 0 ? *(kk <= a16 ? dq + tri16(a16, kk) : kZeroF64) : *(a16 <= kk ? dq + tri16(kk, a16) : kZeroF64);
    }
    // these loads issue before the next block's operands (the memory counter
    // drains in order), so waiting for them leaves those in flight
    MEM_FENCE();
    // coupling product; each slot's operand is reloaded for the next block
    // once its products are issued
    V4 acc[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) acc[ct] = V4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < MAXK; ++t) {
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) acc[ct] = Mfma16<double>::step((double)A[t], win[t][ct], acc[ct]);
      A[t] = loadA(qn, t);
    }
    // diagonal block
    V4 x[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) {
      x[ct] = V4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k = 0; k < 4; ++k) x[ct] = Mfma16<double>::step(dv[k], (double)rhs[ct][k] - acc[ct][k], x[ct]);
    }
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) {
      const int col = cbase + 16 * ct + a16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = i0 + 4 * r + g;
        if (row < n && col < ncols) Y[(size_t)row * ldy + col] = (S)x[ct][r];
      }
    }
    // window: drop the four oldest slots, the block enters as the newest
#pragma unroll
    for (int t = 0; t + 4 < MAXK; ++t) {
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) win[t][ct] = win[t + 4][ct];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) win[DIR > 0 ? MAXK - 4 + r : MAXK - 1 - r][ct] = x[ct][r];
    }
  }
}

template <typename S, int MAXK, int NCT>
__global__ void __launch_bounds__(64 * SOLVE_MFMA_WAVES) __attribute__((amdgpu_waves_per_eu(2)))
    k_solve_mfma(int n, int b, int ldy, int ncols, int macros, const S *__restrict__ band,
                 const S *__restrict__ rband, const double *__restrict__ dinv, S *__restrict__ y,
                 const int *__restrict__ colstate, int *__restrict__ iters) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int mac = __builtin_amdgcn_readfirstlane(blockIdx.x * SOLVE_MFMA_WAVES + wave);
  if (mac >= macros) return;
  if (colstate) {
    const bool pending = lane < ncols && colstate[(size_t)mac * ncols + lane] == 0;
    if (!__any(pending)) return;
  }
  if (iters && lane == 0) iters[mac] += 1;
  const size_t nW = (size_t)n * (b + 1);
  const double *Di = dinv + (size_t)mac * ((n + 15) / 16) * DBLK;
  S *Y = y + (size_t)mac * n * ldy;
  mfma_sweep<S, MAXK, NCT, 1>(n, b, ldy, ncols, band + mac * nW, Di, Y, lane);
  mfma_sweep<S, MAXK, NCT, -1>(n, b, ldy, ncols, rband + mac * nW, Di, Y, lane);
}

// The same solve with one wavefront per 16-column tile of a macro (a block per
// macro): the tiles share only the band, so a macro's latency is that of one
// tile's products (each wavefront reads the band itself: for refinement passes
// that leave few macros pending).
template <typename S, int MAXK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
    k_solve_mfma_tiles(int n, int b, int ldy, int ncols, const S *__restrict__ band,
                       const S *__restrict__ rband, const double *__restrict__ dinv, S *__restrict__ y,
                       const int *__restrict__ colstate, int *__restrict__ iters) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int mac = blockIdx.x;
  if (colstate) {
    const bool pending = lane < ncols && colstate[(size_t)mac * ncols + lane] == 0;
    if (!__any(pending)) return;
  }
  if (iters && threadIdx.x == 0) iters[mac] += 1;
  const size_t nW = (size_t)n * (b + 1);
  const double *Di = dinv + (size_t)mac * ((n + 15) / 16) * DBLK;
  S *Y = y + (size_t)mac * n * ldy;
  mfma_sweep<S, MAXK, 1, 1>(n, b, ldy, ncols, band + mac * nW, Di, Y, lane, 16 * wave);
  mfma_sweep<S, MAXK, 1, -1>(n, b, ldy, ncols, rband + mac * nW, Di, Y, lane, 16 * wave);
}

// The tile solve for bands past b = 111 (36 or 48 k-steps): its operand and
// window registers alone take 144 or 192 of the 256 VGPRs that two waves per
// SIMD leave, so it runs one wave per SIMD with the whole register file.
template <typename S, int MAXK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1)))
    k_solve_mfma_tiles_wide(int n, int b, int ldy, int ncols, const S *__restrict__ band,
                            const S *__restrict__ rband, const double *__restrict__ dinv,
                            S *__restrict__ y, const int *__restrict__ colstate, int *__restrict__ iters) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int mac = blockIdx.x;
  if (colstate) {
    const bool pending = lane < ncols && colstate[(size_t)mac * ncols + lane] == 0;
    if (!__any(pending)) return;
  }
  if (iters && threadIdx.x == 0) iters[mac] += 1;
  const size_t nW = (size_t)n * (b + 1);
  const double *Di = dinv + (size_t)mac * ((n + 15) / 16) * DBLK;
  S *Y = y + (size_t)mac * n * ldy;
  mfma_sweep<S, MAXK, 1, 1>(n, b, ldy, ncols, band + mac * nW, Di, Y, lane, 16 * wave);
  mfma_sweep<S, MAXK, 1, -1>(n, b, ldy, ncols, rband + mac * nW, Di, Y, lane, 16 * wave);
}

// Register window of the one-tile-per-wavefront matrix-core solve: the
// smallest instance of at least ceil(b / 4) k-steps (48 doubles of operands
// and 48 of window per lane at most), 0 past b = 191.
static int solve_tile_window(int b) {
  for (int k : {8, 12, 17, 24, 28, 36, 48})
    if (b >= 1 && (b + 3) / 4 <= k) return k;
  return 0;
}

// Register window of the matrix-core solve: MAXK k-steps x NCT column tiles
// (at most 52 doubles per lane).
static void solve_mfma_shape(int b, int ncols, int *maxk, int *nct) {
  static const int kMaxK[] = {8, 12, 17, 24, 28};
  *maxk = *nct = 0;
  if (ncols < 1 || ncols > 64 || b < 1) return;
  const int t = (ncols + 15) / 16, nk = (b + 3) / 4;
  for (int m : kMaxK)
    if (nk <= m && m * t <= 52) {
      *maxk = m;
      *nct = t;
      return;
    }
}

// D e_p and D c into columns c0, c0+1 of the workspace.
template <typename T>
__global__ void __launch_bounds__(256) k_load_uw(LayoutK L, BatchK B, T *__restrict__ y, int c0) {
  const int mac = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.n) return;
  const double d = B.dscale[(size_t)mac * L.n + i];
  const double cs = B.scal[(size_t)mac * NSCAL + S_CS];
  T *row = y + ((size_t)mac * L.n + i) * B.ldy + c0;
  row[0] = T(i == L.pin ? d : 0.0);
  row[1] = T(d * cs * constraint_at(L, B, mac, i));
}

// u = D y_c0, w = D y_c0+1 and the 2x2 bordered matrix per macro.
template <typename T>
__global__ void __launch_bounds__(256) k_saddle_init(LayoutK L, BatchK B, const T *__restrict__ y, int c0) {
  __shared__ double red[4];
  const int mac = blockIdx.x;
  double *u = B.uw + (size_t)mac * 2 * L.n;
  double *w = u + L.n;
  double *sc = B.scal + (size_t)mac * NSCAL;
  const double cs = sc[S_CS], sigma = sc[S_SIGMA];
  double cu = 0.0, cw = 0.0;
  for (int i = threadIdx.x; i < L.n; i += 256) {
    const double d = B.dscale[(size_t)mac * L.n + i];
    const T *row = y + ((size_t)mac * L.n + i) * B.ldy + c0;
    const double ui = d * (double)row[0], wi = d * (double)row[1];
    u[i] = ui;
    w[i] = wi;
    const double ci = cs * constraint_at(L, B, mac, i);
    cu += ci * ui;
    cw += ci * wi;
  }
  cu = block_sum<256>(cu, red);
  cw = block_sum<256>(cw, red);
  __syncthreads();
  if (threadIdx.x == 0) {
    const double up = u[L.pin], wp = w[L.pin];
    const double a11 = 1.0 - sigma * up, a12 = wp, a21 = -sigma * cu, a22 = cw;
    sc[S_A11] = a11;
    sc[S_A12] = a12;
    sc[S_A21] = a21;
    sc[S_A22] = a22;
    sc[S_DET] = a11 * a22 - a12 * a21;
    if (!(fabs(sc[S_DET]) > 0.0) || !isfinite(sc[S_DET])) atomicOr(&B.status[mac], MHM_STATUS_PIVOT);
  }
}

// ----------------------------------------------------------------- refinement
// Solutions are row-major per macro: X[mac][row][col], N rows (row n = the
// constraint multiplier when k == 1).  Both refinement kernels run one block
// per macro on a (row lane x column) grid: thread t owns column t % ncols and
// rows lane, lane + lanes, ...; per-column sums are reduced over lanes in a
// fixed order.  ncols <= 64.

// Right-hand side entry: the standard columns [f, B, A Z] (rhs == NULL) or a
// caller's row-major band-order rhs.
__device__ __forceinline__ double rhs_at(const LayoutK &L, const BatchK &B, const double *rhs, int mac,
                                         int ncols, int i, int c) {
  if (rhs) return rhs[((size_t)mac * B.N + i) * ncols + c];
  if (i >= L.n) return 0.0;
  if (c == 0) return B.load[(size_t)mac * L.n + i];
  if (c <= L.m) return coupling_at(L, B, mac, c - 1, i);
  return B.az[(size_t)mac * L.n + i];
}

// r = b - S x, per-column acceptance ||r|| <= rtol ||b|| (linalg/linear.py:
// 132-152; the max|b| column scaling of the reference cancels from both
// sides), scaled correction rhs D r into y, constraint-row residual into beta,
// for x from a refinement pass (it >= 1; the first pass is k_rhs_init).  Like
// the reference, in two stages:
//  FAST: binary64 products and sums for every pending column, with the
//        rounding bound E = gam_row ||(|b| + |S||x|)|| (gam_con for the
//        constraint row).  A column that meets the criterion is accepted;
//        one whose binary64 residual is at least 16 E (so within 1/16 of the
//        exact one) keeps it as its correction (rdone = 1);
//  DD:   double-double evaluation for the pending columns left (rdone = 0,
//        near the rounding floor), and for the final check (check_only: set
//        MHM_STATUS_RESIDUAL for columns that miss the criterion).
// ||b||^2 comes from bnorm (kept by k_rhs_init).
// Block per macro, RES_WAVES wavefronts, lane c = column c (ncols <= 64).
// Rows go in chunks of RES_ROWS; wavefront w takes rows i0 + w, i0 + w +
// RES_WAVES, ... of the chunk, one at a time.  The chunk's CSR slots (values,
// column indices, row offsets) are staged in LDS with the right-hand side; a
// row's slots are the same for the whole wavefront, so their values and
// column indices are broadcast LDS reads (four slots at a time) and the x row
// of a slot is one contiguous LDS read; each (row, column) entry is
// accumulated in slot order.
// x rows stream from HBM once into an LDS ring of RES_ROWS + 2b rows (a
// chunk's CSR columns lie in rows [i0 - b, i0 + RES_ROWS + b)); the chunk's
// right-hand side and outputs go through an LDS stage with row-contiguous
// global access.  Column c's right-hand side is colptr[c][i * colstride[c]] *
// colscale[c] (explicit rhs, f, A Z, or the coupling through the per-macro
// column map and scale).  The per-column sums of every wavefront are combined
// in wavefront order at the end.  Without a ring (no LDS room) x is read from
// HBM directly.
#define RES_ROWS 48
#define RES_WAVES 8
#define RES_CPF 4  // CSR slots per thread of a chunk: at most 64 * RES_WAVES * RES_CPF per chunk

struct ResShared {
  const double *colptr[64];
  int colstride[64];
  int store[64];  // the column's residual becomes its next correction's right-hand side
  double colscale[64], mu[64];
};

// LDS of the residual past the x ring: the chunk's right-hand side rows, row
// scales and constraint entries, its CSR values, column indices and row
// offsets (at most slots per chunk).
static size_t res_stage(int ncols, int slots) {
  return ((size_t)RES_ROWS * (ncols + 2) + slots) * sizeof(double) + ((size_t)slots + RES_ROWS + 1) * sizeof(int);
}

template <typename T, bool RING, bool FAST>
__global__ void __launch_bounds__(64 * RES_WAVES) __attribute__((amdgpu_waves_per_eu(4)))
    k_residual(LayoutK L, BatchK B, int ncols, const double *__restrict__ rhs, const double *__restrict__ X,
               double rtol, double rtol_floor, int it, T *__restrict__ y, double *__restrict__ beta,
               int *__restrict__ colstate, int check_only,
               const int *__restrict__ rowptr, const int *__restrict__ cols, const double *__restrict__ aval,
               const double *__restrict__ dscale, const double *__restrict__ coupling_rm, int ring,
               int slots, double *__restrict__ bnorm, int *__restrict__ rdone, double gam_row, double gam_con) {
  constexpr int NT = 64 * RES_WAVES;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  __shared__ ResShared sh;
  const int ldr = ncols;
  double *xr = reinterpret_cast<double *>(smem_raw);  // ring x ldr
  const int mac = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int n = L.n, N = B.N, b = L.b;
  const bool on = lane < ncols;
  const int cl = on ? lane : 0;  // lanes past ncols read column 0 and store nothing
  int *state = colstate + (size_t)mac * ncols;
  // converged columns were accepted by the residual of their current X; the
  // double-double stage takes the pending columns the binary64 stage left
  int *done = rdone + (size_t)mac * ncols;
  const bool mine = tid < ncols && state[tid] == 0 && (FAST || check_only || done[tid] == 0);
  const double *x = X + (size_t)mac * N * ncols;
  if (tid < 64) {
    const int c = tid;
    const double *ptr = kZeroF64;
    int stride = 0;
    double scale = 1.0;
    if (c < ncols) {
      stride = 1;
      if (rhs) {
        ptr = rhs + (size_t)mac * N * ncols + c;
        stride = ncols;
      } else if (c == 0) {
        ptr = B.load + (size_t)mac * n;
      } else if (c > L.m) {
        ptr = B.az + (size_t)mac * n;
      } else if (B.bcoup) {
        ptr = B.bcoup + ((size_t)mac * L.m + (c - 1)) * n;
      } else {
        ptr = coupling_rm + (B.colmap ? B.colmap[(size_t)mac * L.m + c - 1] : c - 1);
        stride = L.m;
        if (B.colscale) scale = B.colscale[(size_t)mac * L.m + c - 1];
      }
    }
    sh.colptr[c] = ptr;
    sh.colstride[c] = stride;
    // the double-double stage rewrites only the columns it takes over: a
    // column whose binary64 residual was kept must not see a more accurate one
    // because another column of its macro needed it (columns stay independent)
    sh.store[c] = c < ncols && (FAST || (state[c] == 0 && done[c] == 0));
    sh.colscale[c] = scale;
    sh.mu[c] = (c < ncols && L.k) ? x[(size_t)n * ncols + c] : 0.0;
  }
  if (!__syncthreads_or(mine)) return;
  constexpr bool use_ring = RING;  // an instance per case: LDS and global reads stay apart
  const double *a = aval + (size_t)mac * L.nnz;
  const double *ds = dscale + (size_t)mac * n;
  const double cs = B.scal[(size_t)mac * NSCAL + S_CS];
  T *Y = y + (size_t)mac * n * B.ldy;
  double *stg = xr + (size_t)ring * ldr;
  double *cds = stg + (size_t)RES_ROWS * ldr;     // the chunk's row scales
  double *ccn = cds + RES_ROWS;                   // ... scaled constraint entries
  double *ca = ccn + RES_ROWS;                    // ... CSR values (negated)
  int *cj = reinterpret_cast<int *>(ca + slots);  // ... byte offsets of their x rows (ring or X)
  int *crp = cj + slots;                          // ... row offsets from the chunk's first slot
  const FastDiv divc(ncols);
  auto store_chunk = [&](int i0) {
    for (int e = tid; e < RES_ROWS * ncols; e += NT) {
      const int q = divc(e), c = e - q * ncols;
      const int i = i0 + q;
      if (i < n && sh.store[c]) Y[(size_t)i * B.ldy + c] = T(stg[q * ldr + c]);
    }
  };
  // ring rows entering before chunk i0: [i0 + b, i0 + RES_ROWS + b) (all of
  // [0, RES_ROWS + b) before the first); the next chunk's rows are loaded into
  // registers during the current chunk and committed after it
  constexpr int RPF = RES_ROWS * 64 / NT;
  double xpf[RPF];
  auto ring_issue = [&](int lo, int hi) {
#pragma unroll
    for (int v = 0; v < RPF; ++v) {
      const int e = tid + NT * v;
      const int q = divc(e), c = e - q * ncols;
      xpf[v] = (lo + q < hi) ? x[(size_t)(lo + q) * ncols + c] : 0.0;
    }
  };
  auto ring_commit = [&](int lo, int hi) {
#pragma unroll
    for (int v = 0; v < RPF; ++v) {
      const int e = tid + NT * v;
      const int q = divc(e), c = e - q * ncols;
      if (lo + q < hi) xr[((lo + q) % ring) * ldr + c] = xpf[v];
    }
  };
  // byte offset of x row j for the chunk at i0: in the ring, or in X
  auto xoff = [&](int i0, int j) -> int {
    if constexpr (RING) {
      const int rb = (max(0, i0 - b) / ring) * ring;  // ring slot of row j: j - rb (- ring)
      int slot = j - rb;
      slot = slot >= ring ? slot - ring : slot;
      return slot * ldr * (int)sizeof(double);
    } else {
      return j * ncols * (int)sizeof(double);
    }
  };
  // a chunk's right-hand side and CSR slots, loaded into registers one chunk
  // ahead and committed to the stage at the chunk boundary (the loaded values
  // are not touched before the commit, so the loads stay in flight meanwhile).
  // A chunk's first slot is the previous chunk's end, loaded with it, so no
  // load waits for another: the slots are read at clamped positions from it.
  constexpr int SPF = RES_ROWS * 64 / NT;  // right-hand side values per thread
  double spf[SPF], apf[RES_CPF], dpf = 0.0, cpf = 0.0;
  int jpf[RES_CPF], rpf = 0, cend = 0, cbase = 0, cnext = 0;  // rowptr[0] = 0
  auto chunk_issue = [&](int i0) {
#pragma unroll
    for (int v = 0; v < SPF; ++v) {
      const int e = tid + NT * v;
      const int q = divc(e), c = e - q * ncols;
      const int i = i0 + q;
      const bool in = e < RES_ROWS * ncols && i < n;
      spf[v] = *(in ? global_ptr(sh.colptr[c]) + (size_t)i * sh.colstride[c] : global_ptr(kZeroF64));
    }
    cbase = cnext;
    cend = rowptr[min(n, i0 + RES_ROWS)];
#pragma unroll
    for (int v = 0; v < RES_CPF; ++v) {
      const int e = min(cbase + tid + NT * v, L.nnz - 1);  // past the chunk: slots not committed
      apf[v] = a[e];
      jpf[v] = cols[e];
    }
    rpf = rowptr[min(n, i0 + min(tid, RES_ROWS))];
    const int ir = min(i0 + min(tid, RES_ROWS - 1), n - 1);
    dpf = ds[ir];
    cpf = L.k ? (B.cons ? B.cons[(size_t)mac * n + ir] : L.constraint[ir]) : 0.0;
  };
  auto chunk_commit = [&](int i0) {
    const int ccnt = cend - cbase;
    cnext = cend;
#pragma unroll
    for (int v = 0; v < SPF; ++v) {
      const int e = tid + NT * v;
      const int q = divc(e), c = e - q * ncols;
      if (e < RES_ROWS * ncols) stg[q * ldr + c] = spf[v] * sh.colscale[c];
    }
#pragma unroll
    for (int v = 0; v < RES_CPF; ++v) {
      const int e = tid + NT * v;
      if (e < ccnt) {
        ca[e] = -apf[v];
        cj[e] = xoff(i0, jpf[v]);
      }
    }
    if (tid <= RES_ROWS) crp[tid] = rpf - cbase;
    if (tid < RES_ROWS) {
      cds[tid] = dpf;
      ccn[tid] = cs * cpf;
    }
  };
  if (use_ring) {
    for (int e = tid; e < min(n, RES_ROWS + b) * ncols; e += NT) {
      const int q = divc(e), c = e - q * ncols;
      xr[q * ldr + c] = x[(size_t)q * ncols + c];  // rows < ring: slot = row
    }
    ring_issue(RES_ROWS + b, min(n, 2 * RES_ROWS + b));
  }
  chunk_issue(0);
  chunk_commit(0);
  double rr = 0.0, chi = 0.0, clo = 0.0;  // FAST: clo = sum |c_i x_i|
  double tt = 0.0;                         // FAST: sum of (|b| + |S||x|)^2
  bool finite = true;
  // the lane's column in the ring or in X; a slot adds its row's byte offset
  const char *xcol = RING ? reinterpret_cast<const char *>(xr + cl) : reinterpret_cast<const char *>(x + cl);
  auto xat = [&](int off) -> double { return *reinterpret_cast<const double *>(xcol + off); };
  __syncthreads();
  for (int i0 = 0; i0 < n; i0 += RES_ROWS) {
    const bool more = i0 + RES_ROWS < n;
    if (more) chunk_issue(i0 + RES_ROWS);
    // only the lanes of real columns run the rows (LDS returns data to active
    // lanes only)
    for (int q = wave; on && q < RES_ROWS; q += RES_WAVES) {
      const int i = i0 + q;
      if (i >= n) break;
      double hi = stg[q * ldr + cl], lo = FAST ? fabs(hi) : 0.0;  // FAST: lo = |b| + |S||x|
      auto term = [&](double av, double xv) {
        if constexpr (FAST) {
          hi = fma(av, xv, hi);
          lo = fma(fabs(av), fabs(xv), lo);
        } else {
          dd_fma(hi, lo, av, xv);
        }
      };
      const int s1 = __builtin_amdgcn_readfirstlane(crp[q + 1]);
      int sb = __builtin_amdgcn_readfirstlane(crp[q]);
      for (; sb + 4 <= s1; sb += 4) {
        double av[4], xv[4];
        int ov[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          av[u] = ca[sb + u];
          ov[u] = cj[sb + u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) xv[u] = xat(ov[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u) term(av[u], xv[u]);
      }
      for (; sb < s1; ++sb) term(ca[sb], xat(cj[sb]));
      if (L.k) {
        const double ci = ccn[q];
        const double xi = xat(xoff(i0, i));
        term(-ci, sh.mu[cl]);
        if constexpr (FAST) {
          chi = fma(ci, xi, chi);
          clo = fma(fabs(ci), fabs(xi), clo);
        } else {
          dd_fma(chi, clo, ci, xi);
        }
      }
      const double r = FAST ? hi : hi + lo;
      if (FAST && on) tt += lo * lo;
      if (on) {
        finite &= isfinite(r);
        rr += r * r;
        if (!check_only) stg[q * ldr + cl] = cds[q] * r;
      }
    }
    __syncthreads();  // the chunk's stage and ring reads are done
    // Every load in flight is a prefetch consumed here: wait for them once,
    // before this chunk's stores, so that no later wait covers the stores or
    // the next prefetch.  store_chunk reads and chunk_commit rewrites the same
    // stage elements per thread, in that order.
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0), expcnt and lgkmcnt untouched
    const int lo = i0 + RES_ROWS + b, hi = min(n, i0 + 2 * RES_ROWS + b);
    // rows of the next chunk into the slots of rows [i0 - b, i0 - b + RES_ROWS)
    if (use_ring && more) ring_commit(lo, hi);
    if (!check_only) store_chunk(i0);
    if (more) chunk_commit(i0 + RES_ROWS);
    if (use_ring && more) ring_issue(hi, min(n, i0 + 3 * RES_ROWS + b));
    __syncthreads();  // the ring and the stage hold the next chunk
  }
  // per-wavefront column sums over the dead ring and stage: tot[(w*5 + k)*64
  // + c] = ||r||^2, constraint row (hi, lo), nonfinite flag, FAST bound sum
  double *tot = xr;
  __syncthreads();  // the last chunk's stage is written out
  tot[(wave * 5 + 0) * 64 + lane] = rr;
  tot[(wave * 5 + 1) * 64 + lane] = chi;
  tot[(wave * 5 + 2) * 64 + lane] = clo;
  tot[(wave * 5 + 3) * 64 + lane] = finite ? 0.0 : 1.0;
  tot[(wave * 5 + 4) * 64 + lane] = tt;
  __syncthreads();
  if (!mine) return;
  const int c = tid;
  double R2 = 0.0, BAD = 0.0, h = 0.0, l = 0.0, T2 = 0.0;
  for (int w = 0; w < RES_WAVES; ++w) {
    R2 += tot[(w * 5 + 0) * 64 + c];
    if (FAST) {
      h += tot[(w * 5 + 1) * 64 + c];
      l += tot[(w * 5 + 2) * 64 + c];
      T2 += tot[(w * 5 + 4) * 64 + c];
    } else {
      double ss, ee;
      two_sum(h, tot[(w * 5 + 1) * 64 + c], ss, ee);
      h = ss;
      l = l + tot[(w * 5 + 2) * 64 + c] + ee;
    }
    BAD += tot[(w * 5 + 3) * 64 + c];
  }
  double B2 = bnorm[(size_t)mac * ncols + c];
  double rn = 0.0, E2 = 0.0;
  if (L.k) {
    const double bn = rhs ? rhs[((size_t)mac * N + n) * ncols + c] : 0.0;
    rn = FAST ? bn - h : bn - (h + l);
    R2 += rn * rn;
    B2 += bn * bn;
    if (FAST) E2 = gam_con * (fabs(bn) + l) * (gam_con * (fabs(bn) + l));
  }
  if (FAST) E2 += gam_row * gam_row * T2;
  // a column whose binary64 residual is below 16 times its rounding bound is
  // at the binary64 floor, where another binary64 pass cannot lower it: it is
  // accepted at rtol_floor (the caller's rtol); others at rtol (which a
  // refinement pass may set tighter).  The double-double stage sees only
  // columns at the floor, and the final check, and uses rtol_floor.
  const bool at_floor = !FAST || R2 < 256.0 * E2;
  const bool resid_ok = sqrt(R2) <= (at_floor ? rtol_floor : rtol) * sqrt(B2) && BAD == 0.0;
  const bool ok = resid_ok;
  if (FAST) {
    // accepted, or a correction from the binary64 residual when it is at
    // least 16 times its rounding bound; the rest go to the double-double stage
    if (ok) {
      state[c] = 1;
    } else {
      beta[(size_t)mac * ncols + c] = rn;
      done[c] = (BAD == 0.0 && R2 >= 256.0 * E2) ? 1 : 0;
    }
    return;
  }
  if (BAD != 0.0) atomicOr(&B.cstatus[mac], MHM_STATUS_NONFINITE);
  if (check_only) {
    if (!resid_ok) atomicOr(&B.cstatus[mac], MHM_STATUS_RESIDUAL);
  } else {
    state[c] = ok ? 1 : 0;
    beta[(size_t)mac * ncols + c] = rn;
  }
}

// Per-column sums over the rows of the refinement kernels go in fixed chunks
// of SADDLE_ROWS rows; their blocks have SADDLE_THREADS threads (a macro still
// refining alone spreads its chunks over SADDLE_THREADS / ncols row groups).
#define SADDLE_ROWS 32
#define SADDLE_THREADS 1024

// First pass of a refined solve (x = 0, so r = b): y = D b, ||b||^2 per column
// (kept for the later passes), the constraint entry of b (the bordered
// correction's right-hand side) and the column state, without the residual's
// CSR walk and x ring.  Block per macro; thread (row, column) over the written
// width (the columns, then u's and w's when uw_col >= 0), so a wavefront's
// stores cover whole rows of y.  Rows go in groups of gchunks chunks of
// SADDLE_ROWS rows, each thread holding up to RHS_PASSES rows of a group in
// flight (the next group's loads are issued before this group's sums).  The
// per-column sums go in fixed chunks of SADDLE_ROWS rows in row order, then in
// chunk order, so a column's bits do not depend on the other columns of the
// call.  The macro's row scales are staged in LDS once.
#define RHS_PASSES 7
#define RHS_TILE 6400  // doubles of a row group's values kept in LDS for the sums

template <typename T>
__global__ void __launch_bounds__(SADDLE_THREADS) __attribute__((amdgpu_waves_per_eu(8)))
    k_rhs_init(LayoutK L, BatchK B, int ncols, const double *__restrict__ rhs, T *__restrict__ y,
               double *__restrict__ beta, int *__restrict__ colstate, const double *__restrict__ coupling_rm,
               double *__restrict__ bnorm, int uw_col, int gchunks) {
  extern __shared__ double part[];  // nchunk x ncols chunk sums, the group's values, the row scales
  __shared__ int bad;
  const int mac = blockIdx.x, tid = threadIdx.x;
  const int n = L.n, m = L.m, N = B.N;
  const int width = uw_col >= 0 ? ncols + 2 : ncols;
  const int rp = SADDLE_THREADS / width;
  const int r = tid / width, c = tid - r * width;
  const bool act = r < rp;
  const int nchunk = (n + SADDLE_ROWS - 1) / SADDLE_ROWS;
  const int grows = gchunks * SADDLE_ROWS;
  const int ngroup = (n + grows - 1) / grows;
  double *tile = part + (size_t)nchunk * ncols;
  double *sds = tile + (size_t)grows * ncols;
  if (tid == 0) bad = 0;
  const double *ds = B.dscale + (size_t)mac * n;
  T *Y = y + (size_t)mac * n * B.ldy;
  // column c reads src[i * stride] * scale; u's column is e_pin, w's is the
  // constraint times cs (its product taken as (d cs) c)
  const double *src = kZeroF64;
  int stride = 0;
  double scale = 1.0;
  const bool pinc = c == ncols, conc = c == ncols + 1;
  if (c < ncols) {
    stride = 1;
    if (rhs) {
      src = rhs + (size_t)mac * N * ncols + c;
      stride = ncols;
    } else if (c == 0) {
      src = B.load + (size_t)mac * n;
    } else if (c > m) {
      src = B.az + (size_t)mac * n;
    } else if (B.bcoup) {
      src = B.bcoup + ((size_t)mac * m + (c - 1)) * n;
    } else {
      src = coupling_rm + (B.colmap ? B.colmap[(size_t)mac * m + c - 1] : c - 1);
      stride = m;
      if (B.colscale) scale = B.colscale[(size_t)mac * m + c - 1];
    }
  } else if (conc) {
    src = B.cons ? B.cons + (size_t)mac * n : L.constraint;
    stride = 1;
    scale = B.scal[(size_t)mac * NSCAL + S_CS];
  }
  const int ycol = c < ncols ? c : uw_col + (c - ncols);
  double v[RHS_PASSES];
  // unconditional loads of a clamped row (a select of the address becomes a
  // branch with a wait per pass); values of rows outside the group are unused
  auto issue = [&](int g) {
#pragma unroll
    for (int p = 0; p < RHS_PASSES; ++p) {
      const int i = min(g * grows + r + p * rp, n - 1);
      v[p] = global_ptr(src)[(size_t)i * stride];
    }
  };
  bool finite = true;
  issue(0);
  for (int t = tid; t < n; t += SADDLE_THREADS) sds[t] = global_ptr(ds)[t];
  __syncthreads();
  for (int g = 0; g < ngroup; ++g) {
    // every load of the group has landed: without this the branchy consume
    // loop waits for vmcnt(0) per pass, i.e. for the previous pass's stores
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0), expcnt and lgkmcnt untouched
#pragma unroll
    for (int p = 0; p < RHS_PASSES; ++p) {
      const int ir = r + p * rp, i = g * grows + ir;
      if (act && ir < grows && i < n) {
        const double d = sds[i];
        if (c < ncols) {
          const double x = v[p] * scale;
          finite &= isfinite(x);
          tile[ir * ncols + c] = x;
          Y[(size_t)i * B.ldy + ycol] = T(d * x);
        } else {
          Y[(size_t)i * B.ldy + ycol] = T(conc ? d * scale * v[p] : (pinc && i == L.pin ? d : 0.0));
        }
      }
    }
    __syncthreads();
    if (g + 1 < ngroup) issue(g + 1);
    if (tid < gchunks * ncols) {
      const int k = tid / ncols, cc = tid - k * ncols;
      const int i0 = g * grows + k * SADDLE_ROWS;
      if (i0 < n) {
        const int rows = min(SADDLE_ROWS, n - i0);
        const double *t = tile + k * SADDLE_ROWS * ncols + cc;
        double s2 = 0.0;
        int q = 0;
        for (; q + 4 <= rows; q += 4) {  // four reads in flight, summed in row order
          const double x0 = t[q * ncols], x1 = t[(q + 1) * ncols], x2 = t[(q + 2) * ncols],
                       x3 = t[(q + 3) * ncols];
          s2 += x0 * x0;
          s2 += x1 * x1;
          s2 += x2 * x2;
          s2 += x3 * x3;
        }
        for (; q < rows; ++q) s2 += t[q * ncols] * t[q * ncols];
        part[(size_t)(i0 / SADDLE_ROWS) * ncols + cc] = s2;
      }
    }
    __syncthreads();
  }
  if (!finite) bad = 1;
  __syncthreads();
  if (tid < ncols) {
    double s2 = 0.0;
    for (int k = 0; k < nchunk; ++k) s2 += part[k * ncols + tid];
    bnorm[(size_t)mac * ncols + tid] = s2;
    colstate[(size_t)mac * ncols + tid] = 0;
    beta[(size_t)mac * ncols + tid] = rhs ? rhs[((size_t)mac * N + n) * ncols + tid] : 0.0;
  }
  if (tid == 0 && bad) atomicOr(&B.cstatus[mac], MHM_STATUS_NONFINITE);
}

// Chunks per k_rhs_init row group: the most (up to 8) whose values fit
// RHS_TILE and whose rows take at most RHS_PASSES passes of the block.
static int rhs_group_chunks(int ncols, int width) {
  const int rp = SADDLE_THREADS / width;
  int g = 8;
  while (g > 1 && (g * SADDLE_ROWS * ncols > RHS_TILE || (g * SADDLE_ROWS + rp - 1) / rp > RHS_PASSES)) --g;
  return g;
}

// x = D y (+ bordered correction when k == 1); the first pass (REFINE false)
// sets X, refinement passes add the correction.
// Per-column sums over the rows go in fixed chunks of SADDLE_ROWS rows, each
// summed in row order, then the chunk sums in chunk order: the result does not
// depend on how many columns share the call (a column's bits are the same
// alone and in a batch).

#define SAP_ROWS 3

template <typename T, bool REFINE>
__global__ void __launch_bounds__(SADDLE_THREADS) __attribute__((amdgpu_waves_per_eu(8)))
    k_saddle_apply(LayoutK L, BatchK B, int ncols, const T *__restrict__ y, const double *__restrict__ beta,
                   double *__restrict__ X, const int *__restrict__ colstate) {
  extern __shared__ double part[];  // nchunk x ncols
  __shared__ double sxp[64], smu[64];
  const int mac = blockIdx.x;
  const int tid = threadIdx.x;
  const int lanes = (int)blockDim.x / ncols;  // SADDLE_THREADS, or 256 for passes with few macros left
  const int lane = tid / ncols, c = tid - lane * ncols;
  const bool on = lane < lanes;
  const int n = L.n, N = B.N;
  const int nchunk = (n + SADDLE_ROWS - 1) / SADDLE_ROWS;
  if (REFINE && !__syncthreads_or(tid < ncols && colstate[(size_t)mac * ncols + tid] == 0))
    return;  // every column of the macro has converged (uniform over the block)
  const bool skip = !on || (REFINE && colstate[(size_t)mac * ncols + c]);
  double *x = X + (size_t)mac * N * ncols;
  const double *ds = B.dscale + (size_t)mac * n;
  const T *Y = y + (size_t)mac * n * B.ldy;
  const double *sc = B.scal + (size_t)mac * NSCAL;
  const double *u = B.uw + (size_t)mac * 2 * n;
  const double *w = u + n;
  const double sigma = sc[S_SIGMA];
  if (L.k) {
    const double cs = sc[S_CS];
    if (!skip)
      for (int k = lane; k < nchunk; k += lanes) {
        double cy = 0.0;
        const int i1 = min(n, (k + 1) * SADDLE_ROWS);
#pragma unroll 4
        for (int i = k * SADDLE_ROWS; i < i1; ++i)
          cy += cs * constraint_at(L, B, mac, i) * (ds[i] * (double)Y[(size_t)i * B.ldy + c]);
        part[k * ncols + c] = cy;
      }
    __syncthreads();
    if (tid < ncols) {
      double s = 0.0;
      for (int k = 0; k < nchunk; ++k) s += part[k * ncols + tid];
      const double yp = ds[L.pin] * (double)Y[(size_t)L.pin * B.ldy + tid];
      const double r1 = yp, r2 = s - beta[(size_t)mac * ncols + tid];
      const double det = sc[S_DET];
      sxp[tid] = (r1 * sc[S_A22] - sc[S_A12] * r2) / det;
      smu[tid] = (sc[S_A11] * r2 - sc[S_A21] * r1) / det;
    }
    __syncthreads();
  }
  if (!skip) {
    const double xp = L.k ? sxp[c] : 0.0, mu = L.k ? smu[c] : 0.0;
    for (int k = lane; k < nchunk; k += lanes) {
      const int i1 = min(n, (k + 1) * SADDLE_ROWS);
      // SAP_ROWS rows in flight: unconditional loads of clamped rows, one
      // wait, then the stores (a per-row loop waits for the previous row's
      // stores before its loads land)
      for (int ib = k * SADDLE_ROWS; ib < i1; ib += SAP_ROWS) {
        double yv[SAP_ROWS], dv[SAP_ROWS], uv[SAP_ROWS], wv[SAP_ROWS], xv[SAP_ROWS];
#pragma unroll
        for (int q = 0; q < SAP_ROWS; ++q) {
          // 32-bit offsets from the macro's bases (at most n * ldy elements)
          const unsigned i = (unsigned)min(ib + q, n - 1);
          yv[q] = (double)global_ptr(Y)[i * (unsigned)B.ldy + c];
          dv[q] = global_ptr(ds)[i];
          if (L.k) {
            uv[q] = global_ptr(u)[i];
            wv[q] = global_ptr(w)[i];
          }
          if (REFINE) xv[q] = global_ptr(x)[i * (unsigned)ncols + c];
        }
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
#pragma unroll
        for (int q = 0; q < SAP_ROWS; ++q) {
          const int i = ib + q;
          if (i < i1) {
            double v = dv[q] * yv[q];
            if (L.k) v += sigma * xp * uv[q] - mu * wv[q];
            x[(size_t)i * ncols + c] = REFINE ? xv[q] + v : v;
          }
        }
      }
    }
    if (L.k && lane == 0) {
      double *xm = x + (size_t)n * ncols + c;
      if (REFINE)
        *xm += mu;
      else
        *xm = mu;
    }
  }
}

// LDS of the per-column chunk sums (nchunk x ncols)
static size_t saddle_lds(int n, int ncols) {
  return (size_t)((n + SADDLE_ROWS - 1) / SADDLE_ROWS) * ncols * sizeof(double);
}

__global__ void k_fill_int(int *p, int v, size_t count) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) p[t] = v;
}

__global__ void k_fill_f64(double *p, double v, size_t count) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) p[t] = v;
}

// done[mac] = 1 when every column of the macro has converged (its X is final).
__global__ void k_mark_done(const int *__restrict__ colstate, int ncols, int macros, int *__restrict__ done) {
  const int mac = blockIdx.x * blockDim.x + threadIdx.x;
  if (mac >= macros) return;
  int all = 1;
  for (int c = 0; c < ncols; ++c) all &= colstate[(size_t)mac * ncols + c] == 1 ? 1 : 0;
  done[mac] = all;
}

// ----------------------------------------------------------------- condensation
// Block per macro: the condensed block of core/contributions.py:40-47 in the
// general (corrected retained) form and the reduced load of
// core/reconstruction.py:122-127, from s = X[:,0], L = X[:,1..m],
// E = Z - X[:,m+1] (X row-major, R columns).
// Phase 1 (coupled rows, in chunks of COND_ROWS staged in LDS): the products
// B^T [L | s | E | Z], one output per (a, col), accumulated in LDS.
// Phase 2 (all rows): (A Z)^T [L | E | s], thread (group, col) with coalesced
// row reads and four rows in flight, then a reduction over the groups; Z^T f
// by a block reduction.
#define COND_ROWS 32

__global__ void __launch_bounds__(256)
    k_condense(LayoutK L, BatchK B, const double *__restrict__ X, double *__restrict__ blocks,
               double *__restrict__ loads, const int *__restrict__ done, int mode) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  __shared__ double fin[2];
  const int mac = blockIdx.x;
  // mode 1: only macros marked done, mode 2: only the others, 0: all
  if (mode != 0 && (done[mac] != 0) != (mode == 1)) return;
  const int tid = threadIdx.x;
  const int n = L.n, m = L.m, k = L.k, N = B.N, R = B.R;
  const int w = m + k;
  const int nb = m + 1 + 2 * k;  // phase-1 columns: L (m), s, then E and Z when k
  const int nout = m * nb;
  const double *Xm = X + (size_t)mac * N * R;
  const double *az = B.az + (size_t)mac * n;
  const double *f = B.load + (size_t)mac * n;
  double *blk = blocks + (size_t)mac * w * w;
  double *ld = loads + (size_t)mac * w;
  double *acc = reinterpret_cast<double *>(smem_raw);  // nout
  double *bsh = acc + nout;                              // COND_ROWS x m
  double *xsh = bsh + COND_ROWS * m;                     // COND_ROWS x nb
  double *red = xsh + COND_ROWS * nb;                    // 256
  for (int o = tid; o < nout; o += 256) acc[o] = 0.0;
  const int nrows = L.crows ? L.ncr : n;
  for (int q0 = 0; q0 < nrows; q0 += COND_ROWS) {
    const int cnt = min(COND_ROWS, nrows - q0);
    __syncthreads();
    for (int e = tid; e < cnt * m; e += 256) {
      const int q = e / m, a = e - q * m;
      const int i = L.crows ? L.crows[q0 + q] : q0 + q;
      bsh[e] = coupling_at(L, B, mac, a, i);
    }
    for (int e = tid; e < cnt * nb; e += 256) {
      const int q = e / nb, c = e - q * nb;
      const int i = L.crows ? L.crows[q0 + q] : q0 + q;
      double v;
      if (c < m)
        v = Xm[(size_t)i * R + 1 + c];
      else if (c == m)
        v = Xm[(size_t)i * R];
      else if (c == m + 1)
        v = kernel_at(L, B, mac, i) - Xm[(size_t)i * R + 1 + m];
      else
        v = kernel_at(L, B, mac, i);
      xsh[e] = v;
    }
    __syncthreads();
    for (int o = tid; o < nout; o += 256) {
      const int a = o / nb, c = o - a * nb;
      double s = acc[o];
      for (int q = 0; q < cnt; ++q) s += bsh[q * m + a] * xsh[q * nb + c];
      acc[o] = s;
    }
  }
  __syncthreads();
  double zb = 0.0;
  for (int o = tid; o < nout; o += 256) {
    const int a = o / nb, c = o - a * nb;
    if (c < m) blk[a * w + c] = acc[o];
    else if (c == m) ld[a] = acc[o];
    else if (c == m + 1) blk[a * w + m] = -acc[o];
  }
  if (!k) return;
  // phase 2: column c < m -> az.L_c, c == m -> az.E, c == m+1 -> az.s (X
  // column 1 + c, 1 + m, 0), four rows in flight per thread; c == m+2 -> Z.f,
  // a block reduction of its own
  const int nc = m + 3;
  const int G = 256 / (m + 2);
  const int g = tid / (m + 2), c = tid - g * (m + 2);
  const int xc = c < m ? 1 + c : (c == m ? 1 + m : 0);
  double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
  if (g < G) {
    const double *xcol = Xm + xc;
    auto term = [&](int i) {
      const double x = xcol[(size_t)i * R];
      return az[i] * (c == m ? kernel_at(L, B, mac, i) - x : x);
    };
    int i = g;
    for (; i + 3 * G < n; i += 4 * G) {
      p0 += term(i);
      p1 += term(i + G);
      p2 += term(i + 2 * G);
      p3 += term(i + 3 * G);
    }
    for (; i < n; i += G) p0 += term(i);
  }
  double zf = 0.0;
  for (int i = tid; i < n; i += 256) zf += kernel_at(L, B, mac, i) * f[i];
  red[tid] = zf;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o) red[tid] += red[tid + o];
    __syncthreads();
  }
  zf = red[0];
  __syncthreads();
  red[tid] = (p0 + p1) + (p2 + p3);
  __syncthreads();
  if (tid < nc) {
    double t = 0.0;
    if (tid < m + 2)
      for (int l = 0; l < G; ++l) t += red[l * (m + 2) + tid];
    else
      t = zf;
    if (tid < m) {
      zb = acc[tid * nb + m + 2];  // Z^T B column a = tid
      blk[m * w + tid] = -(zb - t);
    } else if (tid == m) {
      blk[m * w + m] = -t;
    } else {
      fin[tid - m - 1] = t;  // az.s, then Z.f (read below by thread m+1)
    }
  }
  __syncthreads();
  if (tid == m + 1) ld[m] = fin[0] - fin[1];
}

static size_t condense_lds(int m, int k) {
  const size_t nb = (size_t)m + 1 + 2 * k;
  return ((size_t)m * nb + COND_ROWS * ((size_t)m + nb) + 256) * sizeof(double);
}

// Responses in original rows: source, lifts (n x m row-major), retained basis.
__global__ void __launch_bounds__(256)
    k_responses(LayoutK L, BatchK B, const double *__restrict__ X, double *__restrict__ src,
                double *__restrict__ lifts, double *__restrict__ ret) {
  const int mac = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.n) return;
  const int N = B.N, R = B.R, n = L.n, m = L.m;
  const double *xr = X + ((size_t)mac * N + i) * R;
  const int o = L.perm[i];
  if (src) src[(size_t)mac * n + o] = xr[0];
  if (lifts)
    for (int a = 0; a < m; ++a) lifts[((size_t)mac * n + o) * m + a] = xr[1 + a];
  if (ret && L.k) ret[(size_t)mac * n + o] = kernel_at(L, B, mac, i) - xr[1 + m];
}

// u = s - L lambda + E c, original rows; s from src[case][macro][n] (original
// rows) when src is not NULL, else the standard source response.
__global__ void __launch_bounds__(256)
    k_reconstruct(LayoutK L, BatchK B, const double *__restrict__ X, const double *__restrict__ tr,
                  const double *__restrict__ co, double *__restrict__ fields, int ncases,
                  const double *__restrict__ src) {
  const int mac = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.n) return;
  const int N = B.N, R = B.R, n = L.n, m = L.m, k = L.k;
  const double *xr = X + ((size_t)mac * N + i) * R;
  const int o = L.perm[i];
  const double e = k ? kernel_at(L, B, mac, i) - xr[1 + m] : 0.0;
  for (int c = 0; c < ncases; ++c) {
    const double *lam = tr + ((size_t)mac * ncases + c) * m;
    double t = 0.0;
    for (int a = 0; a < m; ++a) t += xr[1 + a] * lam[a];
    double u = (src ? src[((size_t)c * B.M + mac) * n + o] : xr[0]) - t;
    if (k) u += e * co[((size_t)mac * ncases + c) * k];
    fields[((size_t)mac * ncases + c) * n + o] = u;
  }
}

// Caller columns [mac][col][N] (original rows) <-> internal row-major
// [mac][row][col] (band rows); row n (constraint) keeps its position.
__global__ void __launch_bounds__(256)
    k_permute_cols(const int *__restrict__ perm, int n, int N, int ncols, const double *__restrict__ in,
                   double *__restrict__ out, int to_band) {
  const int mac = blockIdx.y;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * ncols) return;
  const int i = t / ncols, c = t - i * ncols;  // internal (row, col)
  const int o = i < n ? perm[i] : i;
  const size_t ext = ((size_t)mac * ncols + c) * N + o;
  const size_t inr = ((size_t)mac * N + i) * ncols + c;
  if (to_band)
    out[inr] = in[ext];
  else
    out[ext] = in[inr];
}

// Copy a per-macro [macro][rows..] double array from original to band rows.
__global__ void __launch_bounds__(256)
    k_permute_rows(const int *__restrict__ perm, int n, int inner, const double *__restrict__ in,
                   double *__restrict__ out) {
  const int mac = blockIdx.y;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * inner) return;
  const int a = t / n, i = t - a * n;
  out[((size_t)mac * inner + a) * n + i] = in[((size_t)mac * inner + a) * n + perm[i]];
}

__global__ void __launch_bounds__(256)
    k_scatter(const int64_t *__restrict__ idx, int w, int count, int64_t *__restrict__ rows,
              int64_t *__restrict__ cols) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t per = (size_t)w * w;
  if (t >= per * count) return;
  const size_t c = t / per;
  const int rem = (int)(t - c * per);
  const int i = rem / w, j = rem - i * w;
  rows[t] = idx[c * w + i];
  cols[t] = idx[c * w + j];
}

__global__ void __launch_bounds__(256)
    k_ordered_gather(const int64_t *__restrict__ ptr, const int64_t *__restrict__ code,
                     const double *__restrict__ src, int64_t count, double *__restrict__ out,
                     double *__restrict__ aout) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= count) return;
  double s = 0.0, a = 0.0;
  for (int64_t t = ptr[g]; t < ptr[g + 1]; ++t) {
    const double v = src[code[t]];
    s += v;
    a += fabs(v);
  }
  out[g] = s;
  if (aout) aout[g] = a;
}

// ================================================================= host side
struct mhm_layout {
  int device = 0;
  LayoutK k{};
  int gather_slots = 0;  // most CSR slots in GATHER_ROWS consecutive rows
  int res_slots = 0;     // most CSR slots in a residual chunk (RES_ROWS rows from a multiple of RES_ROWS)
  int maxrow = 0;        // most CSR slots in one row
  const double *coupling_rm = nullptr;  // template coupling, row-major n x m (band rows)
  std::vector<void *> owned;
  double cmax_host = 0.0;
  bool has_skel = false;
  SkelK skel{};
  const double *vint = nullptr;  // [nb] integral of each reference basis function / |T|
};

struct mhm_batch {
  const mhm_layout *layout = nullptr;
  int M = 0, precision = MHM_FP64, maxcols = 0, R = 0, N = 0, ldy = 0;
  size_t tsize = 8;
  BatchK bk{};
  void *band = nullptr, *rband = nullptr, *y = nullptr;  // rband: row-major band (wavefront solve)
  void *sval = nullptr;  // scaled operator in CSR slot order (LDS-window factor only)
  double *dinv = nullptr;  // inverses of the 16 x 16 diagonal blocks (binary64 matrix-core solve)
  double *xstd = nullptr, *beta = nullptr, *bnorm = nullptr;
  double *rhsu = nullptr, *xu = nullptr, *srcscale = nullptr;
  double *glam = nullptr, *meas = nullptr;  // per-macro element geometry (uneven macros)
  int *colstate = nullptr;
  int *rdone = nullptr;  // per column: the binary64 residual stage supplied the correction
  int *done = nullptr;   // per macro: converged after the first refinement pass (early condensation)
  hipStream_t side = nullptr;                    // early condensation, beside the caller's stream
  hipEvent_t ev_mark = nullptr, ev_side = nullptr;
  hipEvent_t ev_prep = nullptr, ev_rhs = nullptr;  // the standard right-hand side built beside the factor
  bool use_side = true;  // early condensation (and the standard right-hand side) on the side stream
  bool rhs_ready = false;  // y, beta, bnorm and the per-call state hold pass 0 of the standard columns
  int solve_first = MHM_SOLVE_AUTO, solve_refine = MHM_SOLVE_AUTO;  // solve kernel family per pass kind
  double *ework = nullptr, *fwork = nullptr;
  int *colmap = nullptr;
  double *colscale = nullptr, *bcoup = nullptr, *cons = nullptr, *kern = nullptr;
  bool factored = false, condensed = false;
  bool operator_set = false;  // aval/load/az hold an assembled or supplied operator
  bool uw_pending = false;  // u, w and the bordered 2x2 not yet computed for the current factors
  std::vector<void *> owned;
};

template <typename P>
static int dev_alloc(std::vector<void *> &owned, P **ptr, size_t count) {
  void *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(&p, count * sizeof(P));
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? MHM_ENOMEM : MHM_EDEVICE,
                std::string("hipMalloc: ") + hipGetErrorString(e));
  owned.push_back(p);
  *ptr = static_cast<P *>(p);
  return MHM_OK;
}

template <typename P>
static int upload(std::vector<void *> &owned, const P *host, size_t count, const P **out) {
  if (!host) {
    *out = nullptr;
    return MHM_OK;
  }
  P *d = nullptr;
  int rc = dev_alloc(owned, &d, count);
  if (rc) return rc;
  HIP_TRY(hipMemcpy(d, host, count * sizeof(P), hipMemcpyHostToDevice));
  *out = d;
  return MHM_OK;
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur == dev) return;  // already current: nothing to restore
    prev = cur;
    (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

static size_t factor_lds(int b, size_t tsize) {
  const size_t W = (size_t)b + 1;
  return (W * W + W + 2 * FACTOR_CHUNK * W) * tsize + (size_t)(b * (b + 1) / 2) * 4 + 16;
}

static int lds_limit(int device) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess ||
      v <= 0)
    v = 65536;
  return v;
}

// A kernel launched with more than 64 KiB of dynamic LDS needs its limit
// raised first; that host call costs about as much as a launch, so it is made
// once per (device, kernel) and again only when a launch asks for more.
static hipError_t set_max_lds(const void *fn, size_t bytes) {
  static std::mutex mu;
  static std::map<std::pair<int, const void *>, size_t> granted;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(mu);
  size_t &g = granted[{dev, fn}];
  if (bytes <= g) return hipSuccess;
  e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess) g = bytes;
  return e;
}

static int block_threads(int b) { return b >= 12 ? 256 : 64; }

static size_t solve_lds(int b, int ncols, size_t tsize) {
  const size_t W = (size_t)b + 1;
  return (2 * SOLVE_CHUNK * (W + ncols) + 2 * SOLVE_CHUNK * (size_t)ncols + 2 * (size_t)block_threads(b)) *
             tsize + 16;
}

static bool solve_stage_fits(int b, int ncols) {
  return (size_t)SOLVE_CHUNK * (b + 1 + ncols) <= (size_t)block_threads(b) * 16;
}

// Register slots per thread of the solve window.
static int solve_slots(int b, int ncols) {
  const int G = block_threads(b) / ncols;
  return (b + 1 + G - 1) / G;
}

// Register window widths of the wavefront solve (the smallest >= b+1 is used).
static const int kWaveWidths[] = {8, 16, 24, 32, 48, 64, 68, 80, 96};

static int wave_width(int b, int ncols) {
  if (ncols > 64) return 0;
  for (int w : kWaveWidths)
    if (b + 1 <= w) return w;
  return 0;
}

// Blocked solve (wide bands): rows per block and columns per wavefront.
#define SBLK 16
static const int kBlkCols[] = {2, 4, 7, 8, 12, 16};  // columns per wavefront

static int blk_cpt(int ncols) {
  for (int c : kBlkCols)
    if (ncols <= 4 * c) return c;
  return 0;
}

static size_t solve_blk_lds(int b, int ncols, size_t tsize) {
  const size_t W = (size_t)b + 1, ldr = 4 * (size_t)blk_cpt(ncols) + 1, RB = (size_t)SBLK + b;
  return (SBLK * W + RB * ldr + SBLK * ldr) * tsize;
}

static bool solve_blk_ok(int b, int ncols, size_t tsize, int limit) {
  return b >= 1 && b <= 255 && blk_cpt(ncols) > 0 && solve_blk_lds(b, ncols, tsize) <= (size_t)limit;
}

// Either the wavefront kernel or the block kernel can run a solve of ncols columns.
static bool solve_supported(int b, int ncols, size_t tsize, int limit) {
  if (wave_width(b, ncols) > 0 || solve_blk_ok(b, ncols, tsize, limit)) return true;
  const int cap = block_threads(b) == 64 ? 16 : 64;
  return solve_slots(b, ncols) <= cap && solve_stage_fits(b, ncols) &&
         solve_lds(b, ncols, tsize) <= (size_t)limit;
}

// ----------------------------------------------------------------- blocked band solves
// Block per macro for bandwidths past the wavefront solve's register window
// (b + 1 > 96).  Rows go in blocks of SBLK; the band columns of a block (NB x W,
// one contiguous run of the column band) are loaded a block ahead into
// registers and staged in LDS, and the working rows of y live in an LDS ring of
// NB + b rows (row r in slot r mod (NB + b), odd row stride).  Wavefront w
// owns columns [w CPT, w CPT + CPT).
// Forward L z = y, block k = rows [k, k + NB):
//   (A) stage the band block; wavefronts 1-3 read the rows entering the
//       window, [k + b, k + b + NB), from y into the ring;
//   (B) wavefront 0 solves the NB x NB diagonal block (lane = column) and
//       writes z to the ring and to y;
//   (C) lane l updates rows k + NB + l + 64 v of the window [k + NB, k + NB + b)
//       with the block's z rows (right-looking).
// Backward L^T x = z, blocks from the last:
//   (A) stage the band block; wavefront 0 reads the block's z from y;
//   (C) lane (u, p) = (l mod 16, l / 16) sums L(r, k + u) x_r over rows
//       r = k + NB + p, + 4, ... of the window for its wavefront's columns;
//       the four partial sums are added across lanes and left in LDS;
//   (B) wavefront 0 finishes the block's x upwards (lane = column) and writes
//       x to the ring and to y.
// Three block barriers per NB rows instead of one per row.
template <typename T, int CPT>
__global__ void __launch_bounds__(256)
    k_solve_blk(int n, int b, int ldy, int ncols, const T *__restrict__ band, T *__restrict__ y,
                const int *__restrict__ colstate, int *__restrict__ iters) {
  constexpr int NB = SBLK;
  constexpr int PFR = 16;  // NB * W <= 256 * PFR for b <= 255
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int mac = blockIdx.x, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  __shared__ int active;
  if (colstate) {
    if (tid == 0) active = 0;
    __syncthreads();
    if (tid < ncols && colstate[(size_t)mac * ncols + tid] == 0) active = 1;
    __syncthreads();
    if (!active) return;
  }
  if (iters && tid == 0) iters[mac] += 1;
  const int W = b + 1, RB = NB + b, ldr = 4 * CPT + 1;
  T *bb = reinterpret_cast<T *>(smem_raw);  // NB x W: bb[t*W + d] = L(k + t + d, k + t)
  T *ring = bb + NB * W;                    // RB x ldr
  T *sred = ring + (size_t)RB * ldr;        // NB x ldr
  const T *Lg = band + (size_t)mac * n * W;
  T *Y = y + (size_t)mac * n * ldy;
  const int c0 = wave * CPT;
  const int nblk = (n + NB - 1) / NB;
  const int rpt = (b + 63) / 64;
  T pf[PFR];
  auto band_issue = [&](int k) {  // band columns [k, k + NB) of [0, n)
    const int cnt = (k >= 0 && k < n) ? min(NB, n - k) * W : 0;
#pragma unroll
    for (int v = 0; v < PFR; ++v) {
      const int e = tid + 256 * v;
      pf[v] = e < cnt ? Lg[(size_t)k * W + e] : T(0);
    }
  };
  auto band_store = [&]() {
#pragma unroll
    for (int v = 0; v < PFR; ++v) {
      const int e = tid + 256 * v;
      if (e < NB * W) bb[e] = pf[v];
    }
  };
  // ---------------- forward
  band_issue(0);
  for (int e = tid; e < min(n, RB) * ncols; e += 256) {
    const int r = e / ncols, c = e - r * ncols;
    ring[r * ldr + c] = Y[(size_t)r * ldy + c];
  }
  for (int kb = 0; kb < nblk; ++kb) {
    const int k = kb * NB, np = min(NB, n - k);
    const int sk = k % RB;
    __syncthreads();
    band_store();
    band_issue(k + NB);
    if (kb > 0 && wave > 0) {
      const int r0 = k + b, r1 = min(n, k + b + NB);
      for (int e = tid - 64; e < (r1 - r0) * ncols; e += 192) {
        const int r = r0 + e / ncols, c = e % ncols;
        ring[(r % RB) * ldr + c] = Y[(size_t)r * ldy + c];
      }
    }
    __syncthreads();
    if (wave == 0 && lane < ncols) {
      T v[NB];
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        const int st = sk + t >= RB ? sk + t - RB : sk + t;
        v[t] = t < np ? ring[st * ldr + lane] : T(0);
      }
#pragma unroll
      for (int t = 0; t < NB; ++t) {
        if (t < np) {
          const T z = v[t] / bb[t * W];
          v[t] = z;
#pragma unroll
          for (int u = t + 1; u < NB; ++u)
            if (u < np && u - t <= b) v[u] -= bb[t * W + (u - t)] * z;
        }
      }
#pragma unroll
      for (int t = 0; t < NB; ++t)
        if (t < np) {
          const int st = sk + t >= RB ? sk + t - RB : sk + t;
          ring[st * ldr + lane] = v[t];
          Y[(size_t)(k + t) * ldy + lane] = v[t];
        }
    }
    __syncthreads();
    const int rend = min(n, k + NB + b);
#pragma unroll
    for (int vv = 0; vv < 4; ++vv) {
      const int r = k + NB + lane + 64 * vv;
      if (vv < rpt && r < rend) {
        T *rp = ring + (r % RB) * ldr + c0;
        T acc[CPT];
#pragma unroll
        for (int c = 0; c < CPT; ++c) acc[c] = rp[c];
#pragma unroll 4
        for (int t = 0; t < NB; ++t) {
          if (t < np) {
            const int d = r - k - t;
            const T l = d <= b ? bb[t * W + d] : T(0);
            const int st = sk + t >= RB ? sk + t - RB : sk + t;
            const T *zp = ring + st * ldr + c0;
#pragma unroll
            for (int c = 0; c < CPT; ++c) acc[c] -= l * zp[c];
          }
        }
#pragma unroll
        for (int c = 0; c < CPT; ++c) rp[c] = acc[c];
      }
    }
  }
  // ---------------- backward
  band_issue((nblk - 1) * NB);
  for (int kb = nblk - 1; kb >= 0; --kb) {
    const int k = kb * NB, np = min(NB, n - k);
    const int sk = k % RB;
    __syncthreads();
    band_store();
    band_issue(k - NB);
    T zv[NB];
    if (wave == 0 && lane < ncols) {
#pragma unroll
      for (int t = 0; t < NB; ++t) zv[t] = t < np ? Y[(size_t)(k + t) * ldy + lane] : T(0);
    }
    __syncthreads();
    {
      const int u = lane & 15, p = lane >> 4;
      T acc[CPT];
#pragma unroll
      for (int c = 0; c < CPT; ++c) acc[c] = T(0);
      const int rend = min(n, k + NB + b);
      int sr = (k + NB + p) % RB;
      for (int r = k + NB + p; r < rend; r += 4) {
        const int d = r - k - u;
        const T l = d <= b ? bb[u * W + d] : T(0);
        const T *xp = ring + sr * ldr + c0;
#pragma unroll
        for (int c = 0; c < CPT; ++c) acc[c] += l * xp[c];
        sr += 4;
        if (sr >= RB) sr -= RB;
      }
#pragma unroll
      for (int c = 0; c < CPT; ++c) {
        acc[c] += __shfl_xor(acc[c], 16, 64);
        acc[c] += __shfl_xor(acc[c], 32, 64);
      }
      if (p == 0 && u < np)
#pragma unroll
        for (int c = 0; c < CPT; ++c) sred[u * ldr + c0 + c] = acc[c];
    }
    __syncthreads();
    if (wave == 0 && lane < ncols) {
      T xv[NB];
#pragma unroll
      for (int u = NB - 1; u >= 0; --u) {
        xv[u] = T(0);
        if (u < np) {
          T acc = zv[u] - sred[u * ldr + lane];
#pragma unroll
          for (int v = u + 1; v < NB; ++v)
            if (v < np && v - u <= b) acc -= bb[u * W + (v - u)] * xv[v];
          xv[u] = acc / bb[u * W];
        }
      }
#pragma unroll
      for (int u = 0; u < NB; ++u)
        if (u < np) {
          const int su = sk + u >= RB ? sk + u - RB : sk + u;
          ring[su * ldr + lane] = xv[u];
          Y[(size_t)(k + u) * ldy + lane] = xv[u];
        }
    }
  }
}

template <typename T>
static int launch_solve_blk(const mhm_batch *b, int ncols, const int *colstate, int *iters, hipStream_t st) {
  const LayoutK &L = b->layout->k;
  const size_t shm = solve_blk_lds(L.b, ncols, sizeof(T));
#define MHM_BLK(c)                                                                                         \
  case c:                                                                                                  \
    if (shm > 65536)                                                                                       \
      HIP_TRY(set_max_lds((const void *)k_solve_blk<T, c>, shm));                                          \
    hipLaunchKernelGGL((k_solve_blk<T, c>), dim3(b->M), dim3(256), shm, st, L.n, L.b, b->ldy, ncols,         \
                       (const T *)b->band, (T *)b->y, colstate, iters);                                    \
    break;
  switch (blk_cpt(ncols)) {
    MHM_BLK(2) MHM_BLK(4) MHM_BLK(7) MHM_BLK(8) MHM_BLK(12) MHM_BLK(16)
    default:
      return fail(MHM_EINVAL, "too many columns for the blocked solve");
  }
#undef MHM_BLK
  LAUNCH_CHECK("k_solve_blk");
  return MHM_OK;
}

template <typename T, int MAXW>
static void launch_wave(const mhm_batch *b, int ncols, const int *colstate, int *iters, hipStream_t st) {
  const LayoutK &L = b->layout->k;
  const size_t shm = (size_t)SOLVE_WAVES * 2 * wave_chunk<T, MAXW>() * MAXW * sizeof(T);
  const dim3 grid((unsigned)((b->M + SOLVE_WAVES - 1) / SOLVE_WAVES));
  hipLaunchKernelGGL((k_solve_wave<T, MAXW>), grid, dim3(64 * SOLVE_WAVES), shm, st, L.n, L.b, b->ldy, ncols,
                     b->M, (const T *)b->band, (const T *)b->rband, (T *)b->y, colstate, iters);
}

// Band solve kernel families (mhm_batch_set_solve_variants names them):
//   MHM_SOLVE_MFMA        binary64 matrix cores, one wavefront per macro
//   MHM_SOLVE_MFMA_TILES  binary64 matrix cores, one wavefront per 16-column tile
//   MHM_SOLVE_WAVE        one wavefront per macro, lane = column (b + 1 <= 96)
//   MHM_SOLVE_BLOCKED     four wavefronts per macro over column groups (b <= 255)
// Each launcher fails with MHM_EINVAL when its family cannot serve the batch.
template <typename T>
static int launch_solve_mfma(const mhm_batch *b, int ncols, const int *colstate, int *iters, hipStream_t st) {
  {
    const LayoutK &L = b->layout->k;
    int maxk, nct;
    solve_mfma_shape(L.b, ncols, &maxk, &nct);
    if (b->dinv && b->rband && maxk) {
      const dim3 grid((unsigned)((b->M + SOLVE_MFMA_WAVES - 1) / SOLVE_MFMA_WAVES));
#define MHM_MFMA(K, C)                                                                                      \
  if (maxk == K && nct == C) {                                                                              \
    hipLaunchKernelGGL((k_solve_mfma<T, K, C>), grid, dim3(64 * SOLVE_MFMA_WAVES), 0, st, L.n, L.b, b->ldy,  \
                       ncols, b->M, (const T *)b->band, (const T *)b->rband, (const double *)b->dinv,      \
                       (T *)b->y, colstate, iters);                                                         \
    LAUNCH_CHECK("k_solve_mfma");                                                                           \
    return MHM_OK;                                                                                          \
  }
      MHM_MFMA(8, 1) MHM_MFMA(8, 2) MHM_MFMA(8, 3) MHM_MFMA(8, 4)
      MHM_MFMA(12, 1) MHM_MFMA(12, 2) MHM_MFMA(12, 3) MHM_MFMA(12, 4)
      MHM_MFMA(17, 1) MHM_MFMA(17, 2) MHM_MFMA(17, 3)
      MHM_MFMA(24, 1) MHM_MFMA(24, 2) MHM_MFMA(28, 1)
#undef MHM_MFMA
    }
  }
  return fail(MHM_EINVAL, "no matrix-core solve for this band and column count");
}

template <typename T>
static int launch_solve_tiles(const mhm_batch *b, int ncols, const int *colstate, int *iters, hipStream_t st) {
  {
    const LayoutK &L = b->layout->k;
    const int kt = (b->dinv && b->rband && ncols >= 1 && ncols <= 64) ? solve_tile_window(L.b) : 0;
    const int nct = (ncols + 15) / 16;
#define MHM_TILES(K)                                                                                         \
  if (kt == K) {                                                                                             \
    hipLaunchKernelGGL((k_solve_mfma_tiles<T, K>), dim3(b->M), dim3(64 * nct), 0, st, L.n, L.b, b->ldy, ncols, \
                       (const T *)b->band, (const T *)b->rband, (const double *)b->dinv, (T *)b->y, colstate, \
                       iters);                                                                               \
    LAUNCH_CHECK("k_solve_mfma_tiles");                                                                      \
    return MHM_OK;                                                                                           \
  }
    MHM_TILES(8) MHM_TILES(12) MHM_TILES(17) MHM_TILES(24) MHM_TILES(28)
#undef MHM_TILES
#define MHM_TILES_WIDE(K)                                                                                     \
  if (kt == K) {                                                                                              \
    hipLaunchKernelGGL((k_solve_mfma_tiles_wide<T, K>), dim3(b->M), dim3(64 * nct), 0, st, L.n, L.b, b->ldy,   \
                       ncols, (const T *)b->band, (const T *)b->rband, (const double *)b->dinv, (T *)b->y,  \
                       colstate, iters);                                                                    \
    LAUNCH_CHECK("k_solve_mfma_tiles_wide");                                                                  \
    return MHM_OK;                                                                                            \
  }
    MHM_TILES_WIDE(36) MHM_TILES_WIDE(48)
#undef MHM_TILES_WIDE
  }
  return fail(MHM_EINVAL, "no matrix-core tile solve for this band and column count");
}

template <typename T>
static int launch_solve_wave(const mhm_batch *b, int ncols, const int *colstate, int *iters, hipStream_t st) {
  const LayoutK &L = b->layout->k;
  switch (b->rband ? wave_width(L.b, ncols) : 0) {
#define MHM_WAVE(w)                                   \
  case w:                                             \
    launch_wave<T, w>(b, ncols, colstate, iters, st); \
    LAUNCH_CHECK("k_solve_wave");                     \
    return MHM_OK;
    MHM_WAVE(8) MHM_WAVE(16) MHM_WAVE(24) MHM_WAVE(32) MHM_WAVE(48) MHM_WAVE(64) MHM_WAVE(68) MHM_WAVE(80)
    MHM_WAVE(96)
#undef MHM_WAVE
    default:
      return fail(MHM_EINVAL, "the band is too wide for the wavefront solve");
  }
}

template <typename T>
static int launch_solve_generic(const mhm_batch *b, int ncols, const int *colstate, int *iters, hipStream_t st) {
  const LayoutK &L = b->layout->k;
  const size_t shm = solve_lds(L.b, ncols, sizeof(T));
  const int slots = solve_slots(L.b, ncols);
  const dim3 grid(b->M);
#define MHM_SOLVE(NT, MS)                                                                              \
  hipLaunchKernelGGL((k_solve<T, NT, MS>), grid, dim3(NT), shm, st, L.n, L.b, b->ldy, ncols,        \
                     (const T *)b->band, (T *)b->y, colstate, iters)
  if (!solve_stage_fits(L.b, ncols) || shm > (size_t)lds_limit(b->layout->device))
    return fail(MHM_EINVAL, "solve stage exceeds the register or LDS budget");
  if (block_threads(L.b) == 64) {
    if (slots <= 4)
      MHM_SOLVE(64, 4);
    else if (slots <= 8)
      MHM_SOLVE(64, 8);
    else if (slots <= 16)
      MHM_SOLVE(64, 16);
    else
      return fail(MHM_EINVAL, "solve window exceeds the register budget");
  } else if (slots <= 4) {
    MHM_SOLVE(256, 4);
  } else if (slots <= 8) {
    MHM_SOLVE(256, 8);
  } else if (slots <= 12) {
    MHM_SOLVE(256, 12);
  } else if (slots <= 16) {
    MHM_SOLVE(256, 16);
  } else if (slots <= 24) {
    MHM_SOLVE(256, 24);
  } else if (slots <= 32) {
    MHM_SOLVE(256, 32);
  } else if (slots <= 64) {
    MHM_SOLVE(256, 64);
  } else {
    return fail(MHM_EINVAL, "solve window exceeds the register budget");
  }
#undef MHM_SOLVE
  LAUNCH_CHECK("k_solve");
  return MHM_OK;
}

// The family a call uses: the batch's override for its pass kind (first
// pass: no column state; refinement: a column state), else the first that
// fits of the matrix-core solve (binary64; one wavefront per tile on
// refinement passes, which leave few macros pending, one per macro on the
// first pass, which reads the band once), the wavefront solve, the blocked
// solve and the generic one.
template <typename T>
static int launch_solve(const mhm_batch *b, int ncols, const int *colstate, int *iters, hipStream_t st) {
  const LayoutK &L = b->layout->k;
  switch (colstate ? b->solve_refine : b->solve_first) {
    case MHM_SOLVE_MFMA:
      return launch_solve_mfma<T>(b, ncols, colstate, iters, st);
    case MHM_SOLVE_MFMA_TILES:
      return launch_solve_tiles<T>(b, ncols, colstate, iters, st);
    case MHM_SOLVE_WAVE:
      return launch_solve_wave<T>(b, ncols, colstate, iters, st);
    case MHM_SOLVE_BLOCKED:
      if (!solve_blk_ok(L.b, ncols, sizeof(T), lds_limit(b->layout->device)))
        return fail(MHM_EINVAL, "the blocked solve does not fit this band and column count");
      return launch_solve_blk<T>(b, ncols, colstate, iters, st);
    default:
      break;
  }
  {
    int maxk, nct;
    solve_mfma_shape(L.b, ncols, &maxk, &nct);
    if (b->dinv && b->rband && maxk)
      return colstate ? launch_solve_tiles<T>(b, ncols, colstate, iters, st)
                      : launch_solve_mfma<T>(b, ncols, colstate, iters, st);
    // bands too wide for a whole macro in one wavefront's registers: one
    // wavefront per 16-column tile on every pass
    if (b->dinv && b->rband && solve_tile_window(L.b))
      return launch_solve_tiles<T>(b, ncols, colstate, iters, st);
  }
  if (b->rband && wave_width(L.b, ncols)) return launch_solve_wave<T>(b, ncols, colstate, iters, st);
  if (solve_blk_ok(L.b, ncols, sizeof(T), lds_limit(b->layout->device)))
    return launch_solve_blk<T>(b, ncols, colstate, iters, st);
  return launch_solve_generic<T>(b, ncols, colstate, iters, st);
}


// Exported entry points: C linkage comes from their declarations in mhm_hip.h.

// Sets the calling thread's last error for the other translation units of the
// library (the dense batched LU); returns code.
int mhm_internal_fail(int code, const char *msg) { return fail(code, msg); }

int mhm_last_error(char *buf, size_t len) {
  if (buf && len) {
    std::strncpy(buf, g_error.c_str(), len - 1);
    buf[len - 1] = 0;
  }
  return MHM_OK;
}

int mhm_device_count(int32_t *count) {
  if (!count) return fail(MHM_EINVAL, "count is NULL");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  *count = e == hipSuccess ? c : 0;
  return MHM_OK;
}

int mhm_version(int32_t *major, int32_t *minor) {
  if (major) *major = 1;
  if (minor) *minor = 0;
  return MHM_OK;
}

int mhm_layout_create(int32_t device, const mhm_layout_desc *d, mhm_layout **out) {
  if (!d || !out) return fail(MHM_EINVAL, "desc/out is NULL");
  *out = nullptr;
  if (d->n <= 0 || d->m < 0 || d->k < 0 || d->k > 1 || d->nnz <= 0 || d->bandwidth < 0 ||
      d->bandwidth >= d->n + 1 || !d->perm || !d->rowptr || !d->cols)
    return fail(MHM_EINVAL, "layout sizes or pattern arrays are invalid");
  if (d->k && (d->pin < 0 || d->pin >= d->n))
    return fail(MHM_EINVAL, "k == 1 requires a pin row");
  if (d->bandwidth >= 65535) return fail(MHM_EINVAL, "bandwidth exceeds 65534");
  // host validation of the pattern: sorted rows, in range, symmetric bandwidth
  std::vector<int> diag(d->n, -1);
  for (int i = 0; i < d->n; ++i) {
    if (d->rowptr[i] > d->rowptr[i + 1]) return fail(MHM_EINVAL, "rowptr not monotone");
    for (int s = d->rowptr[i]; s < d->rowptr[i + 1]; ++s) {
      const int j = d->cols[s];
      if (j < 0 || j >= d->n || (s > d->rowptr[i] && d->cols[s - 1] >= j))
        return fail(MHM_EINVAL, "pattern columns must be sorted and in range");
      if (std::abs(i - j) > d->bandwidth) return fail(MHM_EINVAL, "entry outside bandwidth");
      if (j == i) diag[i] = s;
    }
  }
  if (d->rowptr[0] != 0 || d->rowptr[d->n] != d->nnz) return fail(MHM_EINVAL, "rowptr bounds");
  std::vector<int> seen(d->n, 0);
  for (int i = 0; i < d->n; ++i) {
    if (d->perm[i] < 0 || d->perm[i] >= d->n || seen[d->perm[i]]++)
      return fail(MHM_EINVAL, "perm must be a permutation");
  }
  if (d->ne > 0) {
    if (!((d->dim == 2 && (d->nb == 3 || d->nb == 6)) || (d->dim == 3 && (d->nb == 4 || d->nb == 10))))
      return fail(MHM_EINVAL, "assembly supports P1/P2 triangles and tetrahedra");
    if (!d->slot_ptr || !d->slot_code || !d->load_ptr || !d->load_code || !d->grad_lambda ||
        !d->measure || !d->qweights || !d->dtable || !d->vtable || d->nq <= 0)
      return fail(MHM_EINVAL, "assembly arrays missing");
    const int64_t ncode = (int64_t)d->ne * d->nb * d->nb;
    if (d->slot_ptr[0] != 0 || d->slot_ptr[d->nnz] != ncode || d->load_ptr[0] != 0 ||
        d->load_ptr[d->n] != (int64_t)d->ne * d->nb)
      return fail(MHM_EINVAL, "assembly contribution lists inconsistent");
    for (int64_t t = 0; t < ncode; ++t)
      if (d->slot_code[t] < 0 || d->slot_code[t] >= ncode) return fail(MHM_EINVAL, "slot_code range");
    for (int64_t t = 0; t < (int64_t)d->ne * d->nb; ++t)
      if (d->load_code[t] < 0 || d->load_code[t] >= (int64_t)d->ne * d->nb)
        return fail(MHM_EINVAL, "load_code range");
  }
  if (d->coupling_rows)
    for (int q = 0; q < d->ncoupling_rows; ++q)
      if (d->coupling_rows[q] < 0 || d->coupling_rows[q] >= d->n)
        return fail(MHM_EINVAL, "coupling_rows range");
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(MHM_EINVAL, "device outside visible devices");
  DeviceGuard guard(device);
  auto *L = new mhm_layout();
  L->device = device;
  LayoutK &k = L->k;
  k.n = d->n;
  k.k = d->k;
  k.m = d->m;
  k.nnz = d->nnz;
  k.b = d->bandwidth;
  k.W = d->bandwidth + 1;
  k.pin = d->k ? d->pin : -1;
  k.ne = d->ne;
  k.nb = d->nb;
  k.dim = d->dim;
  k.nq = d->nq;
  k.ncr = d->coupling_rows ? d->ncoupling_rows : 0;
  int rc = 0;
  auto cleanup = [&](int code) {
    for (void *p : L->owned) (void)hipFree(p);
    delete L;
    return code;
  };
  const size_t n = d->n;
  if ((rc = upload(L->owned, d->perm, n, &k.perm)) || (rc = upload(L->owned, d->rowptr, n + 1, &k.rowptr)) ||
      (rc = upload(L->owned, d->cols, (size_t)d->nnz, &k.cols)) ||
      (rc = upload(L->owned, diag.data(), n, &k.diagslot)) ||
      (rc = upload(L->owned, d->coupling_rows, (size_t)k.ncr, &k.crows)) ||
      (rc = upload(L->owned, d->coupling, (size_t)d->m * n, &k.coupling)) ||
      (rc = upload(L->owned, d->kernel, n, &k.kernel)) ||
      (rc = upload(L->owned, d->constraint, n, &k.constraint)))
    return cleanup(rc);
  if (d->coupling && d->m > 0) {
    std::vector<double> rm((size_t)d->m * n);
    for (size_t a = 0; a < (size_t)d->m; ++a)
      for (size_t i = 0; i < n; ++i) rm[i * d->m + a] = d->coupling[a * n + i];
    if ((rc = upload(L->owned, rm.data(), rm.size(), &L->coupling_rm))) return cleanup(rc);
  }
  {
    std::vector<int> bslot(n * (size_t)k.W, -1);
    for (int i = 0; i < d->n; ++i)
      for (int s = d->rowptr[i]; s < d->rowptr[i + 1]; ++s) {
        const int j = d->cols[s];
        if (j <= i) bslot[(size_t)j * k.W + (i - j)] = s;
      }
    if ((rc = upload(L->owned, bslot.data(), bslot.size(), &k.bslot))) return cleanup(rc);
    std::vector<int> srow((size_t)d->nnz);
    for (int i = 0; i < d->n; ++i)
      for (int s = d->rowptr[i]; s < d->rowptr[i + 1]; ++s) srow[s] = i;
    if ((rc = upload(L->owned, srow.data(), srow.size(), &k.srow))) return cleanup(rc);
    for (int i0 = 0; i0 < d->n; i0 += GATHER_ROWS)
      L->gather_slots = std::max(L->gather_slots, d->rowptr[std::min(d->n, i0 + GATHER_ROWS)] - d->rowptr[i0]);
    for (int i0 = 0; i0 < d->n; i0 += RES_ROWS)
      L->res_slots = std::max(L->res_slots, d->rowptr[std::min(d->n, i0 + RES_ROWS)] - d->rowptr[i0]);
    for (int i = 0; i < d->n; ++i) L->maxrow = std::max(L->maxrow, d->rowptr[i + 1] - d->rowptr[i]);
  }
  if (d->ne > 0) {
    const size_t ne = d->ne, nb = d->nb, nq = d->nq, dim = d->dim;
    if ((rc = upload(L->owned, d->slot_ptr, (size_t)d->nnz + 1, &k.slot_ptr)) ||
        (rc = upload(L->owned, d->slot_code, ne * nb * nb, &k.slot_code)) ||
        (rc = upload(L->owned, d->load_ptr, n + 1, &k.load_ptr)) ||
        (rc = upload(L->owned, d->load_code, ne * nb, &k.load_code)) ||
        (rc = upload(L->owned, d->grad_lambda, ne * (dim + 1) * dim, &k.glam)) ||
        (rc = upload(L->owned, d->measure, ne, &k.measure)) ||
        (rc = upload(L->owned, d->qweights, nq, &k.qw)) ||
        (rc = upload(L->owned, d->dtable, nq * nb * (dim + 1), &k.dtab)) ||
        (rc = upload(L->owned, d->vtable, nq * nb, &k.vtab)))
      return cleanup(rc);
    // the integral of each reference basis function over a unit-measure
    // element, in quadrature order (M 1 = measure * vint per element)
    std::vector<double> vint(nb, 0.0);
    for (size_t b = 0; b < nb; ++b)
      for (size_t q = 0; q < nq; ++q) vint[b] += d->qweights[q] * d->vtable[q * nb + b];
    if ((rc = upload(L->owned, vint.data(), nb, &L->vint))) return cleanup(rc);
  }
  *out = L;
  return MHM_OK;
}

int mhm_layout_set_skeleton(mhm_layout *L, const mhm_skeleton_desc *d) {
  if (!L || !d) return fail(MHM_EINVAL, "layout/desc is NULL");
  const LayoutK &k = L->k;
  if (d->nsides <= 0 || d->nseg <= 0 || d->degree < 0 || d->degree >= SKEL_MAXMODES)
    return fail(MHM_EINVAL, "skeleton: nsides, nseg >= 1 and 0 <= degree < " + std::to_string(SKEL_MAXMODES));
  if (d->nsides * d->nseg * (d->degree + 1) != k.m)
    return fail(MHM_EINVAL, "skeleton: nsides * nseg * (degree + 1) must equal the layout's m");
  if (d->edeg != 1 && d->edeg != 2) return fail(MHM_EINVAL, "skeleton: edge degree must be 1 or 2");
  if (d->ng <= 0 || d->ng > 16 || !d->gx || !d->gw) return fail(MHM_EINVAL, "skeleton: 1..16 Gauss points");
  if (!d->breaks || d->breaks[0] != 0.0 || d->breaks[d->nseg] != 1.0)
    return fail(MHM_EINVAL, "skeleton: breaks must run from 0 to 1");
  for (int i = 0; i < d->nseg; ++i) {
    if (!(d->breaks[i + 1] > d->breaks[i])) return fail(MHM_EINVAL, "skeleton: breaks must increase");
    if (std::fabs(d->breaks[i] - (1.0 - d->breaks[d->nseg - i])) > 1e-15)
      return fail(MHM_EINVAL, "skeleton: breaks must be symmetric (the second owner mirrors them)");
  }
  if (d->npieces < 0 || d->nentries < 0 || (d->nentries && (!d->entry_ptr || !d->contrib)))
    return fail(MHM_EINVAL, "skeleton: piece or entry tables missing");
  for (int e = 0; e < d->nentries; ++e) {
    if (d->entry_row[e] < 0 || d->entry_row[e] >= k.n || d->entry_side[e] < 0 || d->entry_side[e] >= d->nsides ||
        d->entry_seg[e] < 0 || d->entry_seg[e] >= d->nseg || d->entry_ptr[e + 1] < d->entry_ptr[e])
      return fail(MHM_EINVAL, "skeleton: entry out of range");
    for (int t = d->entry_ptr[e]; t < d->entry_ptr[e + 1]; ++t)
      if (d->contrib[t] < 0 || d->contrib[t] / 3 >= d->npieces || d->contrib[t] % 3 > d->edeg)
        return fail(MHM_EINVAL, "skeleton: contribution out of range");
  }
  DeviceGuard guard(L->device);
  SkelK &K = L->skel;
  K.nsides = d->nsides;
  K.nseg = d->nseg;
  K.degree = d->degree;
  K.edeg = d->edeg;
  K.ng = d->ng;
  K.npieces = d->npieces;
  K.nentries = d->nentries;
  const size_t ncontrib = d->nentries ? (size_t)d->entry_ptr[d->nentries] : 0;
  int rc;
  if ((rc = upload(L->owned, d->breaks, (size_t)d->nseg + 1, &K.breaks)) ||
      (rc = upload(L->owned, d->piece_f, (size_t)d->npieces * 4, &K.pf)) ||
      (rc = upload(L->owned, d->gx, (size_t)d->ng, &K.gx)) || (rc = upload(L->owned, d->gw, (size_t)d->ng, &K.gw)) ||
      (rc = upload(L->owned, d->entry_side, (size_t)d->nentries, &K.eside)) ||
      (rc = upload(L->owned, d->entry_seg, (size_t)d->nentries, &K.eseg)) ||
      (rc = upload(L->owned, d->entry_row, (size_t)d->nentries, &K.erow)) ||
      (rc = upload(L->owned, d->entry_ptr, (size_t)d->nentries + 1, &K.eptr)) ||
      (rc = upload(L->owned, d->contrib, ncontrib, &K.contrib)))
    return rc;
  L->has_skel = true;
  return MHM_OK;
}

int mhm_layout_destroy(mhm_layout *L) {
  if (!L) return MHM_OK;
  DeviceGuard guard(L->device);
  (void)hipDeviceSynchronize();
  for (void *p : L->owned) (void)hipFree(p);
  delete L;
  return MHM_OK;
}

int mhm_layout_info(const mhm_layout *L, int32_t *n, int32_t *bw, int32_t *nnz) {
  if (!L) return fail(MHM_EINVAL, "layout is NULL");
  if (n) *n = L->k.n;
  if (bw) *bw = L->k.b;
  if (nnz) *nnz = L->k.nnz;
  return MHM_OK;
}

int mhm_batch_create(const mhm_layout *L, int32_t macros, int32_t precision, int32_t max_columns,
                     mhm_batch **out) {
  if (!L || !out) return fail(MHM_EINVAL, "layout/out is NULL");
  *out = nullptr;
  if (macros <= 0 || macros > 65535) return fail(MHM_EINVAL, "macros must be in [1, 65535]");
  if (precision != MHM_FP64 && precision != MHM_MIXED) return fail(MHM_EINVAL, "unknown precision");
  if (max_columns < 0) return fail(MHM_EINVAL, "max_columns must be nonnegative");
  const LayoutK &lk = L->k;
  const size_t tsize = precision == MHM_FP64 ? 8 : 4;
  const int dev = L->device;
  const int limit = lds_limit(dev);
  const int nt = block_threads(lk.b);
  if (lk.b + 1 > nt || std::max(std::max(1 + lk.m + lk.k, 2), max_columns) > std::min(nt, 64))
    return fail(MHM_EINVAL, "bandwidth or column count exceeds the workgroup size (at most 64 columns)");
  const int widest = std::max(std::max(1 + lk.m + lk.k, 2), max_columns);
  if (!solve_supported(lk.b, widest, tsize, limit) || !solve_supported(lk.b, 2, tsize, limit))
    return fail(MHM_EINVAL, "solve window exceeds the register or LDS budget");
  if (saddle_lds(lk.n, widest) + (RHS_TILE + lk.n) * sizeof(double) > (size_t)limit)
    return fail(MHM_EINVAL, "too many local rows for the refinement's per-column sums");
  if (factor_lds(lk.b, tsize) > (size_t)limit &&
      !(FPANEL + lk.b <= 127 && factor_panel_lds(lk.b, tsize) <= (size_t)limit) &&
      !factor_wide_ok(lk.b, tsize, limit))
    return fail(MHM_EINVAL, "bandwidth " + std::to_string(lk.b) +
                                " too large for the LDS-resident factor window at this precision");
  DeviceGuard guard(dev);
  auto *b = new mhm_batch();
  b->layout = L;
  b->M = macros;
  b->precision = precision;
  b->tsize = tsize;
  b->R = 1 + lk.m + lk.k;
  b->N = lk.n + lk.k;
  b->maxcols = max_columns;
  b->ldy = widest;
  // room for u, w as two extra columns of the first solve
  if (lk.k && widest + 2 <= 64 && solve_supported(lk.b, widest + 2, tsize, limit)) b->ldy = widest + 2;
  const size_t M = macros, n = lk.n, N = b->N, R = b->R, ncm = std::max(b->R, max_columns);
  BatchK &k = b->bk;
  k.M = macros;
  k.R = b->R;
  k.N = b->N;
  k.ldy = b->ldy;
  int rc = 0;
  auto cleanup = [&](int code) {
    for (void *p : b->owned) (void)hipFree(p);
    delete b;
    return code;
  };
  if ((rc = dev_alloc(b->owned, &k.aval, M * lk.nnz)) || (rc = dev_alloc(b->owned, &k.load, M * n)) ||
      (rc = dev_alloc(b->owned, &k.az, M * n)) || (rc = dev_alloc(b->owned, &k.dscale, M * n)) ||
      (rc = dev_alloc(b->owned, &k.scal, M * NSCAL)) || (rc = dev_alloc(b->owned, &k.uw, M * 2 * n)) ||
      (rc = dev_alloc(b->owned, &k.status, M)) || (rc = dev_alloc(b->owned, &k.cstatus, M)) ||
      (rc = dev_alloc(b->owned, &k.astatus, M)) ||
      (rc = dev_alloc(b->owned, &k.iters, M)) ||
      (rc = dev_alloc(b->owned, &b->xstd, M * R * N)) ||
      (rc = dev_alloc(b->owned, &b->beta, M * ncm)) ||
      (rc = dev_alloc(b->owned, &b->bnorm, M * ncm)) ||
      (rc = dev_alloc(b->owned, &b->colstate, M * ncm)) || (rc = dev_alloc(b->owned, &b->rdone, M * ncm)) ||
      (rc = dev_alloc(b->owned, &b->done, M)))
    return cleanup(rc);
  if (max_columns > 0 &&
      ((rc = dev_alloc(b->owned, &b->rhsu, M * max_columns * N)) ||
       (rc = dev_alloc(b->owned, &b->xu, M * max_columns * N)) ||
       (rc = dev_alloc(b->owned, &b->srcscale, M * max_columns))))
    return cleanup(rc);
  {
    unsigned char *p = nullptr;
    if ((rc = dev_alloc(b->owned, &p, M * n * lk.W * tsize))) return cleanup(rc);
    b->band = p;
    // the matrix-core solve needs the row band and the diagonal blocks (as
    // binary64, whatever the factor's type), which the LDS-window and the
    // wide-band factors write
    const bool panel = factor_panel_ok(lk.b, tsize, dev);
    const bool mfma = solve_tile_window(lk.b) > 0 && (panel || factor_wide_ok(lk.b, tsize, limit));
    if (wave_width(lk.b, widest) > 0 || wave_width(lk.b, 2) > 0 || mfma) {
      if ((rc = dev_alloc(b->owned, &p, M * n * lk.W * tsize))) return cleanup(rc);
      b->rband = p;
    }
    if ((rc = dev_alloc(b->owned, &p, M * n * b->ldy * tsize))) return cleanup(rc);
    b->y = p;
    if (panel) {
      if ((rc = dev_alloc(b->owned, &p, M * lk.nnz * tsize))) return cleanup(rc);
      b->sval = p;
    }
    if (mfma && (rc = dev_alloc(b->owned, &b->dinv, M * (size_t)((n + 15) / 16) * DBLK))) return cleanup(rc);
  }
  HIP_TRY(hipMemset(k.status, 0, M * sizeof(int)));
  HIP_TRY(hipMemset(k.cstatus, 0, M * sizeof(int)));
  HIP_TRY(hipMemset(k.astatus, 0, M * sizeof(int)));
  HIP_TRY(hipMemset(b->rdone, 0, M * ncm * sizeof(int)));
  if (hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_mark, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_side, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_prep, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_rhs, hipEventDisableTiming) != hipSuccess) {
    for (hipEvent_t e : {b->ev_mark, b->ev_side, b->ev_prep})
      if (e) (void)hipEventDestroy(e);
    if (b->side) (void)hipStreamDestroy(b->side);
    return cleanup(fail(MHM_EDEVICE, "stream/event creation failed"));
  }
  HIP_TRY(hipMemset(k.iters, 0, M * sizeof(int)));
  HIP_TRY(hipMemset(k.scal, 0, M * NSCAL * sizeof(double)));
  *out = b;
  return MHM_OK;
}

int mhm_batch_destroy(mhm_batch *b) {
  if (!b) return MHM_OK;
  DeviceGuard guard(b->layout->device);
  (void)hipDeviceSynchronize();
  for (hipEvent_t e : {b->ev_side, b->ev_mark, b->ev_prep, b->ev_rhs})
    if (e) (void)hipEventDestroy(e);
  if (b->side) (void)hipStreamDestroy(b->side);
  for (void *p : b->owned) (void)hipFree(p);
  delete b;
  return MHM_OK;
}

static dim3 grid2(size_t work, int threads, int macros) {
  return dim3((unsigned)((work + threads - 1) / threads), (unsigned)macros);
}

template <int NB, int D>
static int launch_assemble(const mhm_batch *b, const double *kq, const double *fq, double fc,
                           hipStream_t st) {
  const LayoutK &L = b->layout->k;
  constexpr int EPW = 64 / NB, DP = D == 3 ? 4 : D;
  const size_t tab = ((size_t)L.nq * NB * (D + 1) + (size_t)L.nq * NB + L.nq + 1) & ~(size_t)1;
  const size_t shm = (tab + (size_t)ASM_WAVES * 2 * EPW * NB * DP) * sizeof(double);
  if (shm > 64 * 1024) {
    if (shm > 160 * 1024) return fail(MHM_EINVAL, "assembly tables exceed the LDS of a workgroup");
    HIP_TRY(set_max_lds((const void *)k_assemble<NB, D>, shm));
  }
  const int tasks = (L.ne + EPW - 1) / EPW;
  const dim3 grid((unsigned)((tasks + ASM_WAVES - 1) / ASM_WAVES), (unsigned)std::min(b->M, 2048));
  hipLaunchKernelGGL(k_fill_int, dim3((b->M + 255) / 256), dim3(256), 0, st, b->bk.astatus, 0, (size_t)b->M);
  hipLaunchKernelGGL((k_assemble<NB, D>), grid, dim3(64 * ASM_WAVES), shm, st, L, b->M, kq, fq, fc, b->ework,
                     b->fwork, b->bk.glam, b->bk.meas, b->bk.astatus);
  LAUNCH_CHECK("k_assemble");
  return MHM_OK;
}

int mhm_batch_assemble(mhm_batch *b, const double *kq, const double *fq, double f_const, void *stream) {
  if (!b || !kq) return fail(MHM_EINVAL, "batch/kq is NULL");
  const LayoutK &L = b->layout->k;
  if (L.ne <= 0) return fail(MHM_EINVAL, "layout has no assembly description");
  DeviceGuard guard(b->layout->device);
  hipStream_t st = (hipStream_t)stream;
  if (!b->ework) {
    int rc;
    if ((rc = dev_alloc(b->owned, &b->ework, (size_t)b->M * L.nb * L.nb * L.ne)) ||
        (rc = dev_alloc(b->owned, &b->fwork, (size_t)b->M * L.nb * L.ne)))
      return rc;
  }
  TimerSpan span(MHM_TIMER_ASSEMBLE, st);
  int rc = MHM_OK;
  if (L.dim == 2 && L.nb == 3)
    rc = launch_assemble<3, 2>(b, kq, fq, f_const, st);
  else if (L.dim == 2 && L.nb == 6)
    rc = launch_assemble<6, 2>(b, kq, fq, f_const, st);
  else if (L.dim == 3 && L.nb == 4)
    rc = launch_assemble<4, 3>(b, kq, fq, f_const, st);
  else
    rc = launch_assemble<10, 3>(b, kq, fq, f_const, st);
  if (rc) return rc;
  const size_t gshm = (size_t)b->layout->gather_slots * sizeof(double);
  if (gshm > (size_t)lds_limit(b->layout->device))
    return fail(MHM_EINVAL, "pattern too dense for the gather stage (" + std::to_string(GATHER_ROWS) + " rows)");
  if (gshm > 65536)
    HIP_TRY(set_max_lds((const void *)k_gather, gshm));
  hipLaunchKernelGGL(k_gather, grid2(L.n, GATHER_ROWS, b->M), dim3(256), gshm, st, L, b->bk, b->ework, b->fwork);
  LAUNCH_CHECK("k_gather");
  b->factored = b->condensed = false;
  b->operator_set = true;
  return MHM_OK;
}

int mhm_batch_set_operator(mhm_batch *b, const double *values, const double *loads, void *stream) {
  if (!b || !values || !loads) return fail(MHM_EINVAL, "batch/values/loads is NULL");
  const LayoutK &L = b->layout->k;
  DeviceGuard guard(b->layout->device);
  hipStream_t st = (hipStream_t)stream;
  HIP_TRY(hipMemcpyAsync(b->bk.aval, values, (size_t)b->M * L.nnz * sizeof(double),
                         hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_fill_int, dim3((b->M + 255) / 256), dim3(256), 0, st, b->bk.astatus, 0, (size_t)b->M);
  hipLaunchKernelGGL(k_rowsum, grid2(L.n, 128, b->M), dim3(128), 0, st, L, b->bk, loads);
  LAUNCH_CHECK("k_rowsum");
  b->factored = b->condensed = false;
  b->operator_set = true;
  return MHM_OK;
}

// Band rows -> original rows of a per-macro [macro][n] vector.
__global__ void __launch_bounds__(256)
    k_unpermute_vec(const int *__restrict__ perm, int n, const double *__restrict__ in,
                    double *__restrict__ out) {
  const int mac = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[(size_t)mac * n + perm[i]] = in[(size_t)mac * n + i];
}

int mhm_batch_operator(mhm_batch *b, double *values, double *loads, double *az, void *stream) {
  if (!b) return fail(MHM_EINVAL, "batch is NULL");
  const LayoutK &L = b->layout->k;
  DeviceGuard guard(b->layout->device);
  hipStream_t st = (hipStream_t)stream;
  if (values)
    HIP_TRY(hipMemcpyAsync(values, b->bk.aval, (size_t)b->M * L.nnz * sizeof(double), hipMemcpyDeviceToDevice, st));
  if (loads) {
    hipLaunchKernelGGL(k_unpermute_vec, grid2(L.n, 256, b->M), dim3(256), 0, st, L.perm, L.n,
                       (const double *)b->bk.load, loads);
    LAUNCH_CHECK("k_unpermute_vec");
  }
  if (az) {
    hipLaunchKernelGGL(k_unpermute_vec, grid2(L.n, 256, b->M), dim3(256), 0, st, L.perm, L.n,
                       (const double *)b->bk.az, az);
    LAUNCH_CHECK("k_unpermute_vec");
  }
  return MHM_OK;
}

static int permuted_copy(mhm_batch *b, const double *src, double **dst, int inner, hipStream_t st) {
  const LayoutK &L = b->layout->k;
  if (!*dst) {
    int rc = dev_alloc(b->owned, dst, (size_t)b->M * inner * L.n);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_permute_rows, grid2((size_t)L.n * inner, 256, b->M), dim3(256), 0, st, L.perm,
                     L.n, inner, src, *dst);
  LAUNCH_CHECK("k_permute_rows");
  return MHM_OK;
}

int mhm_batch_set_coupling(mhm_batch *b, const int32_t *colmap, const double *colscale,
                           const double *couplings, const double *constraints, const double *kernels,
                           void *stream) {
  if (!b) return fail(MHM_EINVAL, "batch is NULL");
  const LayoutK &L = b->layout->k;
  DeviceGuard guard(b->layout->device);
  hipStream_t st = (hipStream_t)stream;
  int rc;
  const size_t mm = (size_t)b->M * std::max(L.m, 1);
  if (colmap) {
    if (!L.coupling) return fail(MHM_EINVAL, "a column map needs a layout coupling");
    if (!b->colmap && (rc = dev_alloc(b->owned, &b->colmap, mm))) return rc;
    HIP_TRY(hipMemcpyAsync(b->colmap, colmap, (size_t)b->M * L.m * sizeof(int), hipMemcpyDeviceToDevice, st));
    b->bk.colmap = b->colmap;
  }
  if (colscale) {
    if (!b->colscale && (rc = dev_alloc(b->owned, &b->colscale, mm))) return rc;
    HIP_TRY(hipMemcpyAsync(b->colscale, colscale, (size_t)b->M * L.m * sizeof(double), hipMemcpyDeviceToDevice, st));
    b->bk.colscale = b->colscale;
  }
  if (couplings) {
    if ((rc = permuted_copy(b, couplings, &b->bcoup, L.m, st))) return rc;
    b->bk.bcoup = b->bcoup;
  }
  if (constraints) {
    if ((rc = permuted_copy(b, constraints, &b->cons, 1, st))) return rc;
    b->bk.cons = b->cons;
  }
  if (kernels) {
    if ((rc = permuted_copy(b, kernels, &b->kern, 1, st))) return rc;
    b->bk.kern = b->kern;
    if (b->operator_set) {  // A Z of the operator already in place follows the new modes
      hipLaunchKernelGGL(k_retained_action, grid2(L.n, 128, b->M), dim3(128), 0, st, L, b->bk);
      LAUNCH_CHECK("k_retained_action");
    }
  }
  // the constraint scaling, u, w and the responses depend on these data:
  // factor and condense again before using the batch
  if (colmap || colscale || couplings || constraints || kernels) b->factored = b->condensed = false;
  if (!L.coupling && !b->bk.bcoup && L.m > 0)
    return fail(MHM_EINVAL, "no coupling: the layout has none and no per-macro couplings were set");
  if (L.k && !L.constraint && !b->bk.cons) return fail(MHM_EINVAL, "no constraint vectors");
  return MHM_OK;
}

int mhm_batch_set_geometry(mhm_batch *b, const double *grad_lambda, const double *measure, void *stream) {
  if (!b || !grad_lambda || !measure) return fail(MHM_EINVAL, "batch/grad_lambda/measure is NULL");
  const LayoutK &L = b->layout->k;
  if (L.ne <= 0) return fail(MHM_EINVAL, "layout has no assembly description");
  DeviceGuard guard(b->layout->device);
  hipStream_t st = (hipStream_t)stream;
  const size_t ng = (size_t)b->M * L.ne * (L.dim + 1) * L.dim, nm = (size_t)b->M * L.ne;
  int rc;
  if (!b->glam && ((rc = dev_alloc(b->owned, &b->glam, ng)) || (rc = dev_alloc(b->owned, &b->meas, nm))))
    return rc;
  HIP_TRY(hipMemcpyAsync(b->glam, grad_lambda, ng * sizeof(double), hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(b->meas, measure, nm * sizeof(double), hipMemcpyDeviceToDevice, st));
  b->bk.glam = b->glam;
  b->bk.meas = b->meas;
  b->factored = b->condensed = false;
  return MHM_OK;
}

int mhm_batch_skeleton_coupling(mhm_batch *b, const double *side_len, const double *side_sign, void *stream) {
  if (!b || !side_len || !side_sign) return fail(MHM_EINVAL, "batch/side_len/side_sign is NULL");
  const mhm_layout *lay = b->layout;
  if (!lay->has_skel) return fail(MHM_EINVAL, "the layout has no skeleton description");
  const LayoutK &L = lay->k;
  DeviceGuard guard(lay->device);
  hipStream_t st = (hipStream_t)stream;
  if (!b->bcoup) {  // entries outside the skeleton pattern stay zero
    int rc = dev_alloc(b->owned, &b->bcoup, (size_t)b->M * L.m * L.n);
    if (rc) return rc;
    const size_t cnt = (size_t)b->M * L.m * L.n;
    hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st, b->bcoup, 0.0, cnt);
    LAUNCH_CHECK("k_fill_f64");
  }
  if (lay->skel.nentries > 0) {
    hipLaunchKernelGGL(k_trace_coupling, grid2(lay->skel.nentries, 256, b->M), dim3(256), 0, st, L, lay->skel,
                       side_len, side_sign, b->bcoup);
    LAUNCH_CHECK("k_trace_coupling");
  }
  b->bk.bcoup = b->bcoup;
  b->factored = b->condensed = false;
  return MHM_OK;
}

int mhm_batch_mass_constraints(mhm_batch *b, void *stream) {
  if (!b) return fail(MHM_EINVAL, "batch is NULL");
  const mhm_layout *lay = b->layout;
  const LayoutK &L = lay->k;
  if (L.ne <= 0 || !lay->vint) return fail(MHM_EINVAL, "layout has no assembly description");
  if (!L.k) return fail(MHM_EINVAL, "constraints need a layout with a retained mode");
  DeviceGuard guard(lay->device);
  hipStream_t st = (hipStream_t)stream;
  if (!b->cons) {
    int rc = dev_alloc(b->owned, &b->cons, (size_t)b->M * L.n);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_mass_constraints, dim3(b->M), dim3(256), 0, st, L, b->bk, lay->vint, b->cons);
  LAUNCH_CHECK("k_mass_constraints");
  b->bk.cons = b->cons;
  b->factored = b->condensed = false;
  return MHM_OK;
}

int mhm_batch_coupling(mhm_batch *b, double *coupling, double *constraints, void *stream) {
  if (!b) return fail(MHM_EINVAL, "batch is NULL");
  const LayoutK &L = b->layout->k;
  if (L.m > 0 && !L.coupling && !b->bk.bcoup) return fail(MHM_EINVAL, "no coupling set");
  if (constraints && L.k && !L.constraint && !b->bk.cons) return fail(MHM_EINVAL, "no constraint set");
  DeviceGuard guard(b->layout->device);
  hipLaunchKernelGGL(k_coupling_out, grid2((size_t)(L.m + 1) * L.n, 256, b->M), dim3(256), 0, (hipStream_t)stream,
                     L, b->bk, coupling, constraints);
  LAUNCH_CHECK("k_coupling_out");
  return MHM_OK;
}

int mhm_dirichlet_load(int32_t nfaces, int32_t nq, int32_t fs, const double *basis, const double *weights,
                       const double *g, const double *length, const int64_t *dof0, double *load, double *rhs,
                       double *scale, void *stream) {
  if (nfaces < 0 || nq <= 0 || fs <= 0) return fail(MHM_EINVAL, "nfaces >= 0, nq >= 1 and fs >= 1 required");
  if (nfaces == 0) return MHM_OK;
  if (!basis || !weights || !g || !length || !dof0) return fail(MHM_EINVAL, "dirichlet load: NULL input");
  const size_t work = (size_t)nfaces * fs;
  hipLaunchKernelGGL(k_dirichlet_load, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     nfaces, nq, fs, basis, weights, g, length, dof0, load, rhs, scale);
  LAUNCH_CHECK("k_dirichlet_load");
  return MHM_OK;
}

// Pass 0 of a solve: clear its per-call state (solve status bits, column
// states, solve counts) and build the scaled right-hand side y with its norms.
// Timed with the residual passes when it runs on the caller's stream (on the
// side stream its span would include the wait for room beside the factor).
template <typename T>
static int launch_first_pass(mhm_batch *b, const double *rhs, int ncols, bool uw_inline, hipStream_t st,
                             bool timed) {
  const LayoutK &L = b->layout->k;
  // the status of a solve is its own: bits of an earlier call do not carry over
  hipLaunchKernelGGL(k_fill_int, dim3((b->M + 255) / 256), dim3(256), 0, st, b->bk.cstatus, 0, (size_t)b->M);
  hipLaunchKernelGGL(k_fill_int, dim3((unsigned)(((size_t)b->M * ncols + 255) / 256)), dim3(256), 0, st,
                     b->colstate, 0, (size_t)b->M * ncols);
  hipLaunchKernelGGL(k_fill_int, dim3((b->M + 255) / 256), dim3(256), 0, st, b->bk.iters, 0, (size_t)b->M);
  TimerSpan span(timed ? MHM_TIMER_RESIDUAL : -1, st);
  const int gch = rhs_group_chunks(ncols, uw_inline ? ncols + 2 : ncols);
  const size_t shm = saddle_lds(L.n, ncols) + ((size_t)gch * SADDLE_ROWS * ncols + L.n) * sizeof(double);
  if (shm > 65536)
    HIP_TRY(set_max_lds((const void *)k_rhs_init<T>, shm));
  hipLaunchKernelGGL((k_rhs_init<T>), dim3(b->M), dim3(SADDLE_THREADS), shm, st, L, b->bk, ncols, rhs, (T *)b->y,
                     b->beta, b->colstate, b->layout->coupling_rm, b->bnorm, uw_inline ? ncols : -1, gch);
  LAUNCH_CHECK("k_rhs_init");
  return MHM_OK;
}

template <typename T>
static int run_factor(mhm_batch *b, hipStream_t st) {
  const mhm_layout *lay = b->layout;
  const LayoutK &L = lay->k;
  hipLaunchKernelGGL(k_fill_int, dim3((b->M + 255) / 256), dim3(256), 0, st, b->bk.status, 0, (size_t)b->M);
  hipLaunchKernelGGL(k_fill_int, dim3((b->M + 255) / 256), dim3(256), 0, st, b->bk.cstatus, 0, (size_t)b->M);
  const size_t pshm = factor_panel_lds(L.b, sizeof(T));
  const bool panel = factor_panel_ok(L.b, sizeof(T), lay->device);
  hipLaunchKernelGGL((k_prepare<T>), dim3(b->M), dim3(256), panel ? 0 : (size_t)PREP_COLS * L.W * sizeof(T), st, L,
                     b->bk, (T *)b->band, panel ? (T *)b->sval : (T *)nullptr);
  LAUNCH_CHECK("k_prepare");
  // The standard columns' pass 0 (the condensation that follows a factorization)
  // needs the prepared scales but not the factors: beside the panel factor,
  // whose blocks leave most of each CU idle, it runs on the side stream and
  // run_solve waits for it (beside the wide-band factor it slows the factor by
  // more than it saves).  Not when u and w need a solve of their own first,
  // which writes y.
  b->rhs_ready = false;
  const bool pre = panel && b->use_side && b->R > 0 && (L.k == 0 || b->R + 2 <= b->ldy);
  if (pre) HIP_TRY(hipEventRecord(b->ev_prep, st));
  const size_t shm = factor_lds(L.b, sizeof(T));
  const double tiny = (double)L.n * (sizeof(T) == 8 ? 2.220446049250313e-16 : 1.1920929e-07);
  {
  TimerSpan span(MHM_TIMER_FACTOR, st);
  if (panel) {
    auto kern = L.b < FPANEL ? k_factor_panel<T, 6, true>
                : FPANEL * L.W <= 192 * 6 ? k_factor_panel<T, 6, false> : k_factor_panel<T, 10, false>;
    if (pshm > 65536)
      HIP_TRY(set_max_lds((const void *)kern, pshm));
    hipLaunchKernelGGL(kern, dim3(b->M), dim3(256), pshm, st, L.n, L.b, panel_ldw(L.b, sizeof(T)),
                       (T *)b->band, (T *)b->rband, b->bk.status, tiny,
                       (const T *)b->sval, L.nnz, L.bslot, (double *)b->dinv);
  } else if (factor_wide_ok(L.b, sizeof(T), lds_limit(lay->device))) {
    const size_t wshm = factor_wide_lds(L.b, sizeof(T));
    if (wshm > 65536)
      HIP_TRY(set_max_lds((const void *)k_factor_wide<T>, wshm));
    hipLaunchKernelGGL((k_factor_wide<T>), dim3(b->M), dim3(256), wshm, st, L.n, L.b, wide_ldq(L.b),
                       (T *)b->band, (T *)b->rband, b->bk.status, tiny, (double *)b->dinv);
  } else if (L.b >= 12) {
    if (shm > 65536)
      HIP_TRY(set_max_lds((const void *)k_factor<T, 256>, shm));
    hipLaunchKernelGGL((k_factor<T, 256>), dim3(b->M), dim3(256), shm, st, L.n, L.b, (T *)b->band,
                       (T *)b->rband, b->bk.status, tiny);
  } else {
    hipLaunchKernelGGL((k_factor<T, 64>), dim3(b->M), dim3(64), shm, st, L.n, L.b, (T *)b->band,
                       (T *)b->rband, b->bk.status, tiny);
  }
  LAUNCH_CHECK("k_factor");
  }
  if (b->dinv) {
    const size_t blocks = (size_t)b->M * ((L.n + 15) / 16);
    hipLaunchKernelGGL(k_diag_inverse, dim3((unsigned)((blocks + 15) / 16)), dim3(256), 0, st, (long)blocks,
                       b->dinv);
    LAUNCH_CHECK("k_diag_inverse");
  }
  // u = Ahat^-1 e_p and w = Ahat^-1 c ride along with the first solve
  b->uw_pending = L.k > 0;
  if (pre) {
    HIP_TRY(hipStreamWaitEvent(b->side, b->ev_prep, 0));
    if (int rc = launch_first_pass<T>(b, nullptr, b->R, L.k > 0, b->side, false)) return rc;
    HIP_TRY(hipEventRecord(b->ev_rhs, b->side));
    b->rhs_ready = true;
  }
  return MHM_OK;
}

int mhm_batch_factor(mhm_batch *b, void *stream) {
  if (!b) return fail(MHM_EINVAL, "batch is NULL");
  const LayoutK &L = b->layout->k;
  if (L.m > 0 && !L.coupling && !b->bk.bcoup)
    return fail(MHM_EINVAL, "no coupling: set per-macro couplings before factoring");
  if (L.k && !L.constraint && !b->bk.cons) return fail(MHM_EINVAL, "no constraint vectors");
  if (!b->operator_set) return fail(MHM_EINVAL, "no operator: assemble or set the operator before factoring");
  DeviceGuard guard(b->layout->device);
  hipStream_t st = (hipStream_t)stream;
  int rc = b->precision == MHM_FP64 ? run_factor<double>(b, st) : run_factor<float>(b, st);
  if (rc) return rc;
  b->factored = true;
  b->condensed = false;
  return MHM_OK;
}

// Residual launch: column width G per wavefront and the x ring when it fits in LDS.
template <typename T>
static int launch_residual(mhm_batch *b, int ncols, const double *rhs, const double *X, double rtol, double rtol_floor,
                           int it, int check_only, hipStream_t st) {
  const LayoutK &L = b->layout->k;
  if (it < 1 || ncols < 1 || ncols > 64) return fail(MHM_EINVAL, "residual pass needs it >= 1 and 1..64 columns");
  int ring = RES_ROWS + 2 * L.b;
  const int slots = b->layout->res_slots;
  const size_t stage = res_stage(ncols, slots);
  const size_t totals = (size_t)RES_WAVES * 5 * 64 * sizeof(double);  // the final per-wavefront sums
  size_t shm = std::max((size_t)ring * ncols * sizeof(double) + stage, totals);
  if (shm + sizeof(ResShared) > (size_t)lds_limit(b->layout->device)) {
    ring = 0;
    shm = std::max(stage, totals);
  }
  if (shm + sizeof(ResShared) > (size_t)lds_limit(b->layout->device) || slots > 64 * RES_WAVES * RES_CPF)
    return fail(MHM_EINVAL, "residual chunk exceeds the LDS or register budget");
  // rounding bounds of the binary64 stage: a row's sum has at most maxrow + 2
  // terms, the constraint row n + 1
  const double u = 0x1p-53;
  const double kr = b->layout->maxrow + 2.0, kc = L.n + 1.0;
  const double gam_row = kr * u / (1.0 - kr * u), gam_con = kc * u / (1.0 - kc * u);
  for (int stage = check_only ? 1 : 0; stage < 2; ++stage) {
    auto kern = stage == 0 ? (ring > 0 ? k_residual<T, true, true> : k_residual<T, false, true>)
                           : (ring > 0 ? k_residual<T, true, false> : k_residual<T, false, false>);
    if (shm > 65536)
      HIP_TRY(set_max_lds((const void *)kern, shm));
    hipLaunchKernelGGL(kern, dim3(b->M), dim3(64 * RES_WAVES), shm, st, L, b->bk, ncols, rhs, X, rtol, rtol_floor, it,
                       (T *)b->y, b->beta, b->colstate, check_only, L.rowptr, L.cols,
                       (const double *)b->bk.aval, (const double *)b->bk.dscale, b->layout->coupling_rm, ring, slots,
                       b->bnorm, b->rdone, gam_row, gam_con);
    LAUNCH_CHECK("k_residual");
  }
  return MHM_OK;
}

// Solve ncols columns into the row-major X[mac][row][col] (band rows) with
// binary64 refinement; rhs == NULL selects the standard columns [f, B, A Z].
template <typename T>
static int run_solve(mhm_batch *b, const double *rhs, double *X, int ncols, int max_refine,
                     double rtol, hipStream_t st, const std::function<int()> *after_settled = nullptr) {
  const LayoutK &L = b->layout->k;
  // pass 0 of the standard columns built beside the factor (run_factor), or
  // built here at it == 0; any solve consumes or overwrites y, so a
  // prefetched pass serves at most one call, and a call with other columns
  // still waits for it before writing the same workspace
  const bool pre = b->rhs_ready && !rhs && ncols == b->R;
  if (b->rhs_ready) HIP_TRY(hipStreamWaitEvent(st, b->ev_rhs, 0));
  b->rhs_ready = false;
  // u, w of fresh factors: two extra columns of the first solve when the
  // workspace has room, else a solve of their own before the refinement
  const bool uw_inline = b->uw_pending && ncols + 2 <= b->ldy;
  if (b->uw_pending && !uw_inline) {
    hipLaunchKernelGGL((k_load_uw<T>), grid2(L.n, 256, b->M), dim3(256), 0, st, L, b->bk, (T *)b->y, 0);
    LAUNCH_CHECK("k_load_uw");
    if (int rc = launch_solve<T>(b, 2, nullptr, nullptr, st)) return rc;
    hipLaunchKernelGGL((k_saddle_init<T>), dim3(b->M), dim3(256), 0, st, L, b->bk, (const T *)b->y, 0);
    LAUNCH_CHECK("k_saddle_init");
    b->uw_pending = false;
  }
  // the pass whose residual finds almost every macro converged: the first
  // with binary64 factors; with binary32 ones the third (a correction
  // contracts the error by about 1e-6 on these operators, so the solution
  // and two corrections meet the 1e-10 criterion)
  const int settled = sizeof(T) == 8 ? 1 : 3;
  // With binary32 factors a refinement pass accepts a column at a tenth of
  // rtol: a column that meets rtol after the second solve can be as far from
  // the exact solution as its residual (1e-10), and the responses are meant to
  // be within 1e-10 of the binary64 solution; the tenth keeps them an order
  // below.  A macro's slowest column needs the third solve anyway, so this
  // adds almost no pass.  Columns at the binary64 rounding floor, and the
  // final check, use rtol itself.
  const double rtol_pass = sizeof(T) == 8 ? rtol : 0.1 * rtol;
  for (int it = 0; it <= max_refine; ++it) {
    if (it == 0) {
      if (!pre)
        if (int rc = launch_first_pass<T>(b, rhs, ncols, uw_inline, st, true)) return rc;
    } else {
      TimerSpan span(MHM_TIMER_RESIDUAL, st);
      if (int rc = launch_residual<T>(b, ncols, rhs, X, rtol_pass, rtol, it, 0, st)) return rc;
    }
    if (it == settled && after_settled)
      if (int rc = (*after_settled)()) return rc;
    if (it == 0 && uw_inline) {
      {
        // every column is pending at it == 0, so no column state is needed
        TimerSpan span(MHM_TIMER_SOLVE, st);
        if (int rc = launch_solve<T>(b, ncols + 2, nullptr, b->bk.iters, st)) return rc;
      }
      hipLaunchKernelGGL((k_saddle_init<T>), dim3(b->M), dim3(256), 0, st, L, b->bk, (const T *)b->y, ncols);
      LAUNCH_CHECK("k_saddle_init");
      b->uw_pending = false;
    } else {
      TimerSpan span(MHM_TIMER_SOLVE, st);
      if (int rc = launch_solve<T>(b, ncols, b->colstate, b->bk.iters, st)) return rc;
    }
    const void *sap = it > 0 ? (const void *)k_saddle_apply<T, true> : (const void *)k_saddle_apply<T, false>;
    if (saddle_lds(L.n, ncols) > 65536)
      HIP_TRY(set_max_lds(sap, saddle_lds(L.n, ncols)));
    // binary64 refinement passes leave few macros: small blocks dispatch faster
    const int sthreads = (it > 0 && sizeof(T) == 8) ? 256 : SADDLE_THREADS;
    if (it > 0)
      hipLaunchKernelGGL((k_saddle_apply<T, true>), dim3(b->M), dim3(sthreads), saddle_lds(L.n, ncols), st, L, b->bk,
                         ncols, (const T *)b->y, (const double *)b->beta, X, (const int *)b->colstate);
    else
      hipLaunchKernelGGL((k_saddle_apply<T, false>), dim3(b->M), dim3(sthreads), saddle_lds(L.n, ncols), st, L, b->bk,
                         ncols, (const T *)b->y, (const double *)b->beta, X, (const int *)b->colstate);
    LAUNCH_CHECK("k_saddle_apply");
  }
  if (int rc = launch_residual<T>(b, ncols, rhs, X, rtol, rtol, 1, 1, st)) return rc;
  return MHM_OK;
}

int mhm_batch_condense(mhm_batch *b, int32_t max_refine, double rtol, double *blocks, double *loads,
                       void *stream) {
  if (!b || !blocks || !loads) return fail(MHM_EINVAL, "batch/blocks/loads is NULL");
  if (!b->factored) return fail(MHM_EINVAL, "batch is not factored");
  if (max_refine < 0 || max_refine > 64 || !(rtol > 0.0)) return fail(MHM_EINVAL, "bad refinement arguments");
  const LayoutK &L = b->layout->k;
  DeviceGuard guard(b->layout->device);
  hipStream_t st = (hipStream_t)stream;
  const size_t cshm = condense_lds(L.m, L.k);
  if (cshm > 65536)
    HIP_TRY(set_max_lds((const void *)k_condense, cshm));
  // Once almost every macro has converged (run_solve's settled pass) those
  // are condensed on the side stream while the few others finish refining on
  // the caller's stream, then the caller's stream condenses the rest.
  bool early = false;
  const std::function<int()> early_condense = [&]() -> int {
    hipLaunchKernelGGL(k_mark_done, dim3((b->M + 255) / 256), dim3(256), 0, st, (const int *)b->colstate, b->R,
                       b->M, b->done);
    LAUNCH_CHECK("k_mark_done");
    HIP_TRY(hipEventRecord(b->ev_mark, st));
    HIP_TRY(hipStreamWaitEvent(b->side, b->ev_mark, 0));
    hipLaunchKernelGGL(k_condense, dim3(b->M), dim3(256), cshm, b->side, L, b->bk, (const double *)b->xstd, blocks,
                       loads, (const int *)b->done, 1);
    LAUNCH_CHECK("k_condense");
    HIP_TRY(hipEventRecord(b->ev_side, b->side));
    early = true;
    return MHM_OK;
  };
  const std::function<int()> *hook = b->use_side ? &early_condense : nullptr;
  int rc = b->precision == MHM_FP64 ? run_solve<double>(b, nullptr, b->xstd, b->R, max_refine, rtol, st, hook)
                                    : run_solve<float>(b, nullptr, b->xstd, b->R, max_refine, rtol, st, hook);
  if (rc) {
    if (early) (void)hipStreamWaitEvent(st, b->ev_side, 0);
    return rc;
  }
  TimerSpan span(MHM_TIMER_CONDENSE, st);
  if (early) HIP_TRY(hipStreamWaitEvent(st, b->ev_side, 0));
  hipLaunchKernelGGL(k_condense, dim3(b->M), dim3(256), cshm, st, L, b->bk, (const double *)b->xstd, blocks,
                     loads, (const int *)b->done, early ? 2 : 0);
  LAUNCH_CHECK("k_condense");
  b->condensed = true;
  return MHM_OK;
}

int mhm_batch_set_solve_variants(mhm_batch *b, int32_t first, int32_t refine) {
  if (!b) return fail(MHM_EINVAL, "batch is NULL");
  for (int v : {first, refine})
    if (v < MHM_SOLVE_AUTO || v > MHM_SOLVE_BLOCKED) return fail(MHM_EINVAL, "unknown solve kernel family");
  b->solve_first = first;
  b->solve_refine = refine;
  return MHM_OK;
}

int mhm_batch_set_side_stream(mhm_batch *b, int32_t enable) {
  if (!b) return fail(MHM_EINVAL, "batch is NULL");
  b->use_side = enable != 0;
  return MHM_OK;
}

int mhm_batch_responses(mhm_batch *b, double *src, double *lifts, double *ret, void *stream) {
  if (!b) return fail(MHM_EINVAL, "batch is NULL");
  if (!b->condensed) return fail(MHM_EINVAL, "no condensed responses");
  const LayoutK &L = b->layout->k;
  DeviceGuard guard(b->layout->device);
  hipLaunchKernelGGL(k_responses, grid2(L.n, 256, b->M), dim3(256), 0, (hipStream_t)stream, L, b->bk,
                     (const double *)b->xstd, src, lifts, ret);
  LAUNCH_CHECK("k_responses");
  return MHM_OK;
}

int mhm_batch_solve(mhm_batch *b, const double *rhs, double *x, int32_t ncols, int32_t max_refine,
                    double rtol, void *stream) {
  if (!b || !rhs || !x) return fail(MHM_EINVAL, "batch/rhs/x is NULL");
  if (!b->factored) return fail(MHM_EINVAL, "batch is not factored");
  if (ncols <= 0 || ncols > b->maxcols) return fail(MHM_EINVAL, "ncols outside [1, max_columns]");
  if (max_refine < 0 || max_refine > 64 || !(rtol > 0.0)) return fail(MHM_EINVAL, "bad refinement arguments");
  const LayoutK &L = b->layout->k;
  DeviceGuard guard(b->layout->device);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_permute_cols, grid2((size_t)b->N * ncols, 256, b->M), dim3(256), 0, st, L.perm, L.n,
                     b->N, ncols, rhs, b->rhsu, 1);
  LAUNCH_CHECK("k_permute_cols");
  int rc = b->precision == MHM_FP64 ? run_solve<double>(b, b->rhsu, b->xu, ncols, max_refine, rtol, st)
                                    : run_solve<float>(b, b->rhsu, b->xu, ncols, max_refine, rtol, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_permute_cols, grid2((size_t)b->N * ncols, 256, b->M), dim3(256), 0, st, L.perm, L.n,
                     b->N, ncols, (const double *)b->xu, x, 0);
  LAUNCH_CHECK("k_permute_cols");
  return MHM_OK;
}

template <int NB, int D>
static void launch_source_load(const mhm_batch *b, const double *fq, double fc, int ncols, size_t shm,
                               hipStream_t st) {
  hipLaunchKernelGGL((k_source_load<NB, D>), dim3(b->M, ncols), dim3(256), shm, st, b->layout->k, b->M, ncols, fq,
                     fc, b->rhsu, b->N, b->srcscale, b->bk.meas);
}

int mhm_batch_sources(mhm_batch *b, const double *fq, double f_const, int32_t ncols, int32_t max_refine,
                      double rtol, double *loads, double *source, void *stream) {
  if (!b || !loads) return fail(MHM_EINVAL, "batch/loads is NULL");
  if (!b->factored) return fail(MHM_EINVAL, "batch is not factored");
  if (ncols <= 0 || ncols > b->maxcols) return fail(MHM_EINVAL, "ncols outside [1, max_columns]");
  if (max_refine < 0 || max_refine > 64 || !(rtol > 0.0)) return fail(MHM_EINVAL, "bad refinement arguments");
  const LayoutK &L = b->layout->k;
  if (L.ne <= 0) return fail(MHM_EINVAL, "layout has no assembly description");
  const size_t shm = (size_t)L.nb * L.ne * sizeof(double);
  if (shm > (size_t)lds_limit(b->layout->device)) return fail(MHM_EINVAL, "too many element loads per macro");
  DeviceGuard guard(b->layout->device);
  hipStream_t st = (hipStream_t)stream;
  if (L.dim == 2 && L.nb == 3) {
    if (shm > 65536)
      HIP_TRY(set_max_lds((const void *)k_source_load<3, 2>, shm));
    launch_source_load<3, 2>(b, fq, f_const, ncols, shm, st);
  } else if (L.dim == 2 && L.nb == 6) {
    if (shm > 65536)
      HIP_TRY(set_max_lds((const void *)k_source_load<6, 2>, shm));
    launch_source_load<6, 2>(b, fq, f_const, ncols, shm, st);
  } else if (L.dim == 3 && L.nb == 4) {
    if (shm > 65536)
      HIP_TRY(set_max_lds((const void *)k_source_load<4, 3>, shm));
    launch_source_load<4, 3>(b, fq, f_const, ncols, shm, st);
  } else {
    if (shm > 65536)
      HIP_TRY(set_max_lds((const void *)k_source_load<10, 3>, shm));
    launch_source_load<10, 3>(b, fq, f_const, ncols, shm, st);
  }
  LAUNCH_CHECK("k_source_load");
  int rc = b->precision == MHM_FP64 ? run_solve<double>(b, b->rhsu, b->xu, ncols, max_refine, rtol, st)
                                    : run_solve<float>(b, b->rhsu, b->xu, ncols, max_refine, rtol, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_source_condense, dim3(b->M, ncols), dim3(256), 0, st, L, b->bk, ncols,
                     (const double *)b->xu, (const double *)b->rhsu, (const double *)b->srcscale, loads, source);
  LAUNCH_CHECK("k_source_condense");
  return MHM_OK;
}

int mhm_batch_reconstruct(mhm_batch *b, const double *tr, const double *co, double *fields, int32_t ncases,
                          void *stream) {
  if (!b || !tr || !fields || ncases <= 0) return fail(MHM_EINVAL, "bad reconstruct arguments");
  if (!b->condensed) return fail(MHM_EINVAL, "no condensed responses");
  const LayoutK &L = b->layout->k;
  if (L.k && !co) return fail(MHM_EINVAL, "coarse amplitudes required");
  DeviceGuard guard(b->layout->device);
  hipLaunchKernelGGL(k_reconstruct, grid2(L.n, 256, b->M), dim3(256), 0, (hipStream_t)stream, L, b->bk,
                     (const double *)b->xstd, tr, co, fields, ncases, (const double *)nullptr);
  LAUNCH_CHECK("k_reconstruct");
  return MHM_OK;
}

int mhm_batch_reconstruct_sources(mhm_batch *b, const double *src, const double *tr, const double *co,
                                  double *fields, int32_t ncases, void *stream) {
  if (!b || !src || !tr || !fields || ncases <= 0) return fail(MHM_EINVAL, "bad reconstruct arguments");
  if (!b->condensed) return fail(MHM_EINVAL, "no condensed responses");
  const LayoutK &L = b->layout->k;
  if (L.k && !co) return fail(MHM_EINVAL, "coarse amplitudes required");
  DeviceGuard guard(b->layout->device);
  hipLaunchKernelGGL(k_reconstruct, grid2(L.n, 256, b->M), dim3(256), 0, (hipStream_t)stream, L, b->bk,
                     (const double *)b->xstd, tr, co, fields, ncases, src);
  LAUNCH_CHECK("k_reconstruct");
  return MHM_OK;
}

int mhm_batch_status(mhm_batch *b, int32_t *host_status, int32_t *any, void *stream) {
  if (!b) return fail(MHM_EINVAL, "batch is NULL");
  DeviceGuard guard(b->layout->device);
  std::vector<int> tmp(3 * (size_t)b->M);
  const int *arrays[3] = {b->bk.status, b->bk.cstatus, b->bk.astatus};
  for (int a = 0; a < 3; ++a)
    HIP_TRY(hipMemcpyAsync(tmp.data() + (size_t)a * b->M, arrays[a], b->M * sizeof(int), hipMemcpyDeviceToHost,
                           (hipStream_t)stream));
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  int acc = 0;
  for (int i = 0; i < b->M; ++i) {
    const int v = tmp[i] | tmp[(size_t)b->M + i] | tmp[2 * (size_t)b->M + i];
    acc |= v;
    if (host_status) host_status[i] = v;
  }
  if (any) *any = acc;
  return MHM_OK;
}

int mhm_batch_iterations(mhm_batch *b, int32_t *host_counts, void *stream) {
  if (!b || !host_counts) return fail(MHM_EINVAL, "batch/host_counts is NULL");
  DeviceGuard guard(b->layout->device);
  HIP_TRY(hipMemcpyAsync(host_counts, b->bk.iters, b->M * sizeof(int), hipMemcpyDeviceToHost,
                         (hipStream_t)stream));
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  return MHM_OK;
}

int mhm_scatter_coordinates(const int64_t *idx, int32_t width, int32_t count, int64_t *rows, int64_t *cols,
                            void *stream) {
  if (!idx || !rows || !cols || width <= 0 || count < 0) return fail(MHM_EINVAL, "bad scatter arguments");
  if (count == 0) return MHM_OK;
  const size_t total = (size_t)width * width * count;
  hipLaunchKernelGGL(k_scatter, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream, idx,
                     width, count, rows, cols);
  LAUNCH_CHECK("k_scatter");
  return MHM_OK;
}

int mhm_ordered_gather(const int64_t *ptr, const int64_t *code, const double *src, int64_t count, double *out,
                       double *aout, void *stream) {
  if (!ptr || !code || !src || !out || count < 0) return fail(MHM_EINVAL, "bad gather arguments");
  if (count == 0) return MHM_OK;
  hipLaunchKernelGGL(k_ordered_gather, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     ptr, code, src, count, out, aout);
  LAUNCH_CHECK("k_ordered_gather");
  return MHM_OK;
}


int mhm_timing_enable(int32_t enable) {
  g_timer.enabled = enable != 0;
  return MHM_OK;
}

int mhm_timing_reset(void) {
  for (auto &v : g_timer.spans) {
    for (auto &p : v) {
      g_timer.pool.push_back(p.first);
      g_timer.pool.push_back(p.second);
    }
    v.clear();
  }
  return MHM_OK;
}

int mhm_timing_read(int32_t kernel, double *total_ms, int64_t *launches) {
  if (kernel < 0 || kernel >= MHM_TIMER_COUNT || !total_ms || !launches)
    return fail(MHM_EINVAL, "bad timing arguments");
  double total = 0.0;
  for (auto &p : g_timer.spans[kernel]) {
    HIP_TRY(hipEventSynchronize(p.second));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, p.first, p.second));
    total += ms;
  }
  *total_ms = total;
  *launches = (int64_t)g_timer.spans[kernel].size();
  return MHM_OK;
}
