// Element-wise and reduction kernels: SIMP power law, axpby, deterministic dot products,
// Dirichlet split for the linear-dependency-aware solve, and the fused CG vector updates.
//
// Reductions use a fixed grid of kReduceBlocks blocks with grid-stride loops, per-block
// partials in a caller workspace, and a single-block second stage summing the partials in a
// fixed order, so every dot product is bit-reproducible run to run.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "launch.h"

namespace pmt {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int hip_status(hipError_t err, const char* where) {
  if (err == hipSuccess) return PMT_OK;
  set_error(std::string(where) + ": " + hipGetErrorString(err));
  return PMT_ERR_HIP;
}

int check_launch(const char* where) { return hip_status(hipGetLastError(), where); }

// x^p: repeated products for a small whole exponent (ip >= 0; SIMP's p = 3 and its derivative's 2, within
// an ulp or two of pow and a fraction of its cost), else pow
__device__ inline double powlaw_pow(double x, double p, int ip) {
  if (ip < 0) return pow(x, p);
  double r = 1.0;
  for (int k = 0; k < ip; ++k) r *= x;
  return r;
}

// ip for powlaw_pow: p itself when it is a whole number in [0, 8], else -1
static int powlaw_int_exponent(double p) { return (p >= 0.0 && p <= 8.0 && p == std::floor(p)) ? (int)p : -1; }

__global__ void powlaw_fwd_kernel(long long n, double a, double b, double p, int ip, const double* __restrict__ x,
                                  double* __restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a + b * powlaw_pow(x[i], p, ip);
}

// Heaviside projection (ex_compliance_robust.py:139, a MathExpression of the reference):
//   y = (tanh(beta*eta) + tanh(beta*(x - eta))) / (tanh(beta*eta) + tanh(beta*(1 - eta)))
// dx = dy * beta * (1 - tanh^2(beta*(x - eta))) / den; optional dbeta_i = dy_i * dy_i/dbeta.
__global__ void heaviside_fwd_kernel(long long n, double beta, double eta, const double* __restrict__ x,
                                     double* __restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double te = tanh(beta * eta);
  const double den = te + tanh(beta * (1.0 - eta));
  y[i] = (te + tanh(beta * (x[i] - eta))) / den;
}

__global__ void heaviside_bwd_kernel(long long n, double beta, double eta, const double* __restrict__ x,
                                     const double* __restrict__ dy, double* __restrict__ dx, int accumulate,
                                     double* __restrict__ dbeta) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double te = tanh(beta * eta), t1 = tanh(beta * (1.0 - eta)), tx = tanh(beta * (x[i] - eta));
  const double den = te + t1, num = te + tx;
  const double g = dy[i] * beta * (1.0 - tx * tx) / den;
  dx[i] = accumulate ? dx[i] + g : g;
  if (dbeta) {
    const double dnum = eta * (1.0 - te * te) + (x[i] - eta) * (1.0 - tx * tx);
    const double dden = eta * (1.0 - te * te) + (1.0 - eta) * (1.0 - t1 * t1);
    dbeta[i] = dy[i] * (dnum / den - num * dden / (den * den));
  }
}

// OC update for one bisection value lmid (pymoto/common/optimizers.py:420-433):
//   xnew = clip(x * sqrt(-min(dg, 0) / lmid), max(xmin, x - move), min(xmax, x + move))
// and per-block partial sums of xnew (for the volume test).
__global__ void oc_update_kernel(long long n, const double* __restrict__ x, const double* __restrict__ dg,
                                 double xmin, double xmax, double move, double lmid, double* __restrict__ xnew,
                                 double* __restrict__ partial);

__global__ void powlaw_bwd_kernel(long long n, double c, double q, int iq, const double* __restrict__ x,
                                  const double* __restrict__ dy, double* __restrict__ dx, int accumulate) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double v = dy[i] * (c * powlaw_pow(x[i], q, iq));
    dx[i] = accumulate ? dx[i] + v : v;
  }
}

__global__ void axpby_kernel(long long n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (x ? a * x[i] : 0.0) + (b == 0.0 ? 0.0 : b * y[i]);
}

__device__ inline double block_sum(double v) {
  __shared__ double sh[kReduceThreads / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < kReduceThreads / 64; ++w) r += sh[w];
  }
  __syncthreads();
  return r;  // valid in thread 0
}

// stage 1: NOUT dot products sharing the grid-stride loop; partial[o*kReduceBlocks + block].  B0: type of
// the first product's second operand (an FP32 vector is widened exactly: the sums are those of its
// FP64 copy)
template <int NOUT, typename B0 = double>
__global__ void __launch_bounds__(kReduceThreads)
dot_stage1(long long n, const double* __restrict__ a0, const B0* __restrict__ b0, const double* __restrict__ a1,
           const double* __restrict__ b1, double* __restrict__ partial) {
  double s0 = 0.0, s1 = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    s0 = fma(a0[i], (double)b0[i], s0);
    if (NOUT > 1) s1 = fma(a1[i], b1[i], s1);
  }
  double r0 = block_sum(s0);
  if (threadIdx.x == 0) partial[blockIdx.x] = r0;
  if (NOUT > 1) {
    double r1 = block_sum(s1);
    if (threadIdx.x == 0) partial[kReduceBlocks + blockIdx.x] = r1;
  }
}

// stage 2: out[o] = sum of partials (fixed order), one block
template <int NOUT>
__global__ void __launch_bounds__(kReduceThreads) dot_stage2(const double* __restrict__ partial, double* __restrict__ out) {
  for (int o = 0; o < NOUT; ++o) {
    double s = 0.0;
    for (int b = threadIdx.x; b < kReduceBlocks; b += blockDim.x) s += partial[o * kReduceBlocks + b];
    double r = block_sum(s);
    if (threadIdx.x == 0) out[o] = r;
  }
}

// out[0] = sum of nparts partials (fixed order, one block of 1024 threads, each thread's strided loads
// issued eight at a time); host (optional): also hands the sum to the host through pinned memory (the
// protocol of publish_kernel: value, system-scope fence, sequence number)
constexpr int kStage2Threads = 1024;
__global__ void __launch_bounds__(kStage2Threads) parts_stage2(const double* __restrict__ partial, long long nparts,
                                                               double* __restrict__ out, double seq, double* host) {
  double s = 0.0;
  long long b = threadIdx.x;
  for (; b + 7LL * kStage2Threads < nparts; b += 8LL * kStage2Threads) {
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = partial[b + (long long)k * kStage2Threads];
#pragma unroll
    for (int k = 0; k < 8; ++k) s += v[k];
  }
  for (; b < nparts; b += kStage2Threads) s += partial[b];
  __shared__ double sh[kStage2Threads / 64];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int w = 0; w < kStage2Threads / 64; ++w) r += sh[w];
    out[0] = r;
    if (host) {
      volatile double* h = host;
      h[0] = r;
      __threadfence_system();
      h[1] = seq;
      __threadfence_system();
    }
  }
}

__global__ void oc_update_kernel(long long n, const double* __restrict__ x, const double* __restrict__ dg,
                                 double xmin, double xmax, double move, double lmid, double* __restrict__ xnew,
                                 double* __restrict__ partial) {
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double xi = x[i];
    const double lb = fmax(xmin, xi - move), ub = fmin(xmax, xi + move);
    const double v = fmin(fmax(xi * sqrt(-fmin(dg[i], 0.0) / lmid), lb), ub);
    xnew[i] = v;
    s += v;
  }
  const double t = block_sum(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

int launch_dot(long long n, const double* x, const double* y, double* out, double* work, hipStream_t s) {
  dot_stage1<1><<<kReduceBlocks, kReduceThreads, 0, s>>>(n, x, y, nullptr, nullptr, work);
  PMT_TRY(check_launch("dot_stage1"));
  dot_stage2<1><<<1, kReduceThreads, 0, s>>>(work, out);
  return check_launch("dot_stage2");
}

int launch_dot(long long n, const double* x, const float* y, double* out, double* work, hipStream_t s) {
  dot_stage1<1, float><<<kReduceBlocks, kReduceThreads, 0, s>>>(n, x, y, nullptr, nullptr, work);
  PMT_TRY(check_launch("dot_stage1<float>"));
  dot_stage2<1><<<1, kReduceThreads, 0, s>>>(work, out);
  return check_launch("dot_stage2");
}

int launch_dot2(long long n, const double* a0, const double* b0, const double* a1, const double* b1, double* out,
                double* work, hipStream_t s) {
  dot_stage1<2><<<kReduceBlocks, kReduceThreads, 0, s>>>(n, a0, b0, a1, b1, work);
  PMT_TRY(check_launch("dot_stage1<2>"));
  dot_stage2<2><<<1, kReduceThreads, 0, s>>>(work, out);
  return check_launch("dot_stage2<2>");
}

// CG: x += alpha p (WITHX; else the caller adds it later, in cg_p_update), r -= alpha q, partial ||r||^2;
// alpha = sc[1]/sc[0] read on device; r32 (optional): the new r rounded to FP32, the input of a
// reduced-precision preconditioner; u1 (optional, with r32): omega * D^-1 * r32, the preconditioner's
// first sweep from zero (same FP32 expression as jacobi0_kernel).  One pass over the vectors: a block
// owns kXrPer * 256 consecutive entries (strided by 256 per step, so every step is coalesced and the
// steps' loads are independent) and writes its sum of r^2 to partial[block]; the fixed-order second
// stage is parts_stage2.  At 12.8M dofs: 71 us (4 per thread) for the 461 MB of the mixed-precision
// update, where the fixed 1024-block grid-stride loop of the dots took 84 us.
constexpr int kXrPer = 4;

long long cg_xr_parts(long long n) { return (n + 256LL * kXrPer - 1) / (256LL * kXrPer); }

template <bool WITHX>
__global__ void __launch_bounds__(kReduceThreads)
cg_xr_update(long long n, const double* __restrict__ sc, const double* __restrict__ p, const double* __restrict__ q,
             double* __restrict__ x, double* __restrict__ r, double* __restrict__ partial, float* __restrict__ r32,
             const float* __restrict__ dinv32, float omega32, float* __restrict__ u1) {
  const double alpha = sc[1] / sc[0];
  double s = 0.0;
  const long long base = (long long)blockIdx.x * (256 * kXrPer) + threadIdx.x;
  auto step = [&](long long i) {
    if (WITHX) x[i] = fma(alpha, p[i], x[i]);
    const double rv = fma(-alpha, q[i], r[i]);
    r[i] = rv;
    if (r32) {
      const float rf = (float)rv;
      r32[i] = rf;
      if (u1) u1[i] = omega32 * dinv32[i] * rf;
    }
    s = fma(rv, rv, s);
  };
  if (base + 256LL * (kXrPer - 1) < n) {  // every block but the last: no bounds test between the steps
#pragma unroll
    for (int k = 0; k < kXrPer; ++k) step(base + 256LL * k);
  } else {
    for (int k = 0; k < kXrPer; ++k)
      if (base + 256LL * k < n) step(base + 256LL * k);
  }
  double t = block_sum(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// x.y in one flat pass (the cg_xr_update blocking: kXrPer * 256 entries per block, partial[block]); its
// second stage is parts_stage2.  For the dots of the CG loop, whose bits need not match pmt_dot's.
template <typename Y>
__global__ void __launch_bounds__(kReduceThreads)
dot_flat(long long n, const double* __restrict__ x, const Y* __restrict__ y, double* __restrict__ partial) {
  double s = 0.0;
  const long long base = (long long)blockIdx.x * (256 * kXrPer) + threadIdx.x;
  if (base + 256LL * (kXrPer - 1) < n) {
#pragma unroll
    for (int k = 0; k < kXrPer; ++k) s = fma(x[base + 256LL * k], (double)y[base + 256LL * k], s);
  } else {
    for (int k = 0; k < kXrPer; ++k)
      if (base + 256LL * k < n) s = fma(x[base + 256LL * k], (double)y[base + 256LL * k], s);
  }
  const double t = block_sum(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// CG restart branch: x += alpha p only
__global__ void cg_x_update(long long n, const double* __restrict__ sc, const double* __restrict__ p,
                            double* __restrict__ x) {
  const double alpha = sc[1] / sc[0];
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = fma(alpha, p[i], x[i]);
}

// CG: p = z + beta p with beta = -sc[3]/sc[0]   (iterative.py:394-395); z FP64 or the FP32 output of a
// reduced-precision preconditioner (widened exactly).  x (optional): first the x update the iteration's
// cg_xr_update<false> left out, x += alpha p with alpha = sc[1]/sc[0] and the old p - the same
// expression, so the iterates are bitwise those of the undeferred update, for one read of p fewer
template <typename Z, bool WITHX>
__global__ void cg_p_update(long long n, const double* __restrict__ sc, const Z* __restrict__ z,
                            double* __restrict__ p, double* __restrict__ x) {
  const double beta = -sc[3] / sc[0];
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double pv = p[i];
    if (WITHX) x[i] = fma(sc[1] / sc[0], pv, x[i]);
    p[i] = fma(beta, pv, (double)z[i]);
  }
}

// p = z / sqrt(sc[k])  (orth(z, normalize=True) for one column, iterative.py:286-290)
__global__ void scale_rsqrt(long long n, const double* __restrict__ sc, int k, const double* __restrict__ z,
                            double* __restrict__ p) {
  const double inv = 1.0 / sqrt(sc[k]);
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] * inv;
}

__global__ void bc_split_kernel(long long nnodes, int D, const uint8_t* __restrict__ bcm, double bcdiag,
                                const double* __restrict__ rhs, double* __restrict__ rl, double* __restrict__ sol) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nnodes * D) return;
  const long long n = q / D;
  const int d = (int)(q - n * D);
  const bool fixed = bcm && ((bcm[n] >> d) & 1u);
  const double v = rhs[q];
  rl[q] = fixed ? 0.0 : v;
  sol[q] = fixed ? v / bcdiag : 0.0;
}

__global__ void axpy_free_kernel(long long nnodes, int D, const uint8_t* __restrict__ bcm, double a,
                                 const double* __restrict__ x, double* __restrict__ y) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nnodes * D) return;
  const long long n = q / D;
  const int d = (int)(q - n * D);
  const bool fixed = bcm && ((bcm[n] >> d) & 1u);
  if (!fixed) y[q] = fma(a, x[q], y[q]);
}

__global__ void fill_kernel(long long n, double v, double* __restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = v;
}

int launch_axpby(long long n, double a, const double* x, double b, double* y, hipStream_t s) {
  if (n <= 0) return PMT_OK;
  axpby_kernel<<<blocks_for(n, 256), 256, 0, s>>>(n, a, x, b, y);
  return check_launch("axpby_kernel");
}

int launch_cg_xr_update(long long n, const double* sc, const double* p, const double* q, double* x, double* r,
                        double* work, double* out_rr, hipStream_t s, float* r32, const float* dinv32, float omega32,
                        float* u1, double* host, double seq) {
  if (n <= 0) return PMT_OK;
  const long long parts = cg_xr_parts(n);
  if (x)
    cg_xr_update<true><<<(unsigned)parts, kReduceThreads, 0, s>>>(n, sc, p, q, x, r, work, r32, dinv32, omega32, u1);
  else
    cg_xr_update<false><<<(unsigned)parts, kReduceThreads, 0, s>>>(n, sc, p, q, x, r, work, r32, dinv32, omega32, u1);
  PMT_TRY(check_launch("cg_xr_update"));
  parts_stage2<<<1, kStage2Threads, 0, s>>>(work, parts, out_rr, seq, host);
  return check_launch("parts_stage2");
}

// v -> host-visible memory (pinned): host[0] = value, then (after a system-scope fence) host[1] = seq, so a
// host thread that polls host[1] for the sequence number it expects reads a complete value in host[0]
// without a stream synchronisation, whatever the value is (NaN included)
__global__ void publish_kernel(const double* __restrict__ v, double seq, double* host) {
  if (threadIdx.x == 0) {
    volatile double* h = host;
    h[0] = v[0];
    __threadfence_system();
    h[1] = seq;
    __threadfence_system();
  }
}

int launch_publish(const double* v, double seq, double* host, hipStream_t s) {
  publish_kernel<<<1, 64, 0, s>>>(v, seq, host);
  return check_launch("publish_kernel");
}

template <typename Y>
int launch_dot_flat(long long n, const double* x, const Y* y, double* out, double* work, hipStream_t s) {
  if (n <= 0) return PMT_OK;
  const long long parts = cg_xr_parts(n);
  dot_flat<Y><<<(unsigned)parts, kReduceThreads, 0, s>>>(n, x, y, work);
  PMT_TRY(check_launch("dot_flat"));
  parts_stage2<<<1, kStage2Threads, 0, s>>>(work, parts, out, 0.0, nullptr);
  return check_launch("parts_stage2");
}
template int launch_dot_flat<double>(long long, const double*, const double*, double*, double*, hipStream_t);
template int launch_dot_flat<float>(long long, const double*, const float*, double*, double*, hipStream_t);

int launch_cg_x_update(long long n, const double* sc, const double* p, double* x, hipStream_t s) {
  cg_x_update<<<blocks_for(n, 256), 256, 0, s>>>(n, sc, p, x);
  return check_launch("cg_x_update");
}

int launch_cg_p_update(long long n, const double* sc, const double* z, double* p, hipStream_t s, double* x) {
  if (x)
    cg_p_update<double, true><<<blocks_for(n, 256), 256, 0, s>>>(n, sc, z, p, x);
  else
    cg_p_update<double, false><<<blocks_for(n, 256), 256, 0, s>>>(n, sc, z, p, x);
  return check_launch("cg_p_update");
}

int launch_cg_p_update(long long n, const double* sc, const float* z, double* p, hipStream_t s, double* x) {
  if (x)
    cg_p_update<float, true><<<blocks_for(n, 256), 256, 0, s>>>(n, sc, z, p, x);
  else
    cg_p_update<float, false><<<blocks_for(n, 256), 256, 0, s>>>(n, sc, z, p, x);
  return check_launch("cg_p_update<float>");
}

int launch_scale_rsqrt(long long n, const double* sc, int k, const double* z, double* p, hipStream_t s) {
  scale_rsqrt<<<blocks_for(n, 256), 256, 0, s>>>(n, sc, k, z, p);
  return check_launch("scale_rsqrt");
}

__global__ void to_double_kernel(long long n, const float* __restrict__ x, double* __restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (double)x[i];
}

int launch_to_double(long long n, const float* x, double* y, hipStream_t s) {
  if (n <= 0) return PMT_OK;
  to_double_kernel<<<blocks_for(n, 256), 256, 0, s>>>(n, x, y);
  return check_launch("to_double_kernel");
}

__global__ void jacobi_apply_kernel(long long n, double w, const double* __restrict__ diag,
                                    const double* __restrict__ r, double* __restrict__ z) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = w * r[i] / diag[i];
}

// ---------------------------------------------------------------------------------------------
// Dot products of k column pairs without a caller workspace: the per-block partial sums live in a
// file-scope device array (one copy per GPU, kept until the process ends).  Same grid, loop and
// summation order per column as dot_stage1<1> / dot_stage2<1>, so each column's result is bit-identical
// to pmt_dot's.  Every partial is written before it is read, so nothing needs clearing between calls.
constexpr int kNoAllocMaxCols = 8;
static __device__ __align__(256) double scratch[kNoAllocMaxCols * kReduceBlocks];  // 64 KiB
constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices];  // device address of `scratch`, per device
static std::mutex g_scratch_mu;

// Device address of this device's `scratch`, looked up once per device with that device current.
// In host code the array's name is only a host-side stand-in; kernels get this pointer.
static int scratch_for_current_device(double** out) {
  int dev = 0;
  PMT_TRY(hip_status(hipGetDevice(&dev), "hipGetDevice"));
  if (dev < 0 || dev >= kMaxDevices) return arg_error("pmt_dot_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    PMT_TRY(hip_status(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return PMT_OK;
}

__global__ void __launch_bounds__(kReduceThreads)
dot_cols_stage1(long long n, const double* __restrict__ x, const double* __restrict__ y,
                double* __restrict__ partial) {
  const long long off = (long long)blockIdx.y * n;
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    s = fma(x[off + i], y[off + i], s);
  const double r = block_sum(s);
  if (threadIdx.x == 0) partial[blockIdx.y * kReduceBlocks + blockIdx.x] = r;
}

__global__ void __launch_bounds__(kReduceThreads)
dot_cols_stage2(const double* __restrict__ partial, double* __restrict__ out) {
  const double* p = partial + (long long)blockIdx.x * kReduceBlocks;
  double s = 0.0;
  for (int b = threadIdx.x; b < kReduceBlocks; b += blockDim.x) s += p[b];
  const double r = block_sum(s);
  if (threadIdx.x == 0) out[blockIdx.x] = r;
}

// ---------------------------------------------------------------------------------------------
// Deflation of the preconditioned residual by a coarse space W (A-orthogonal projection):
// z <- z - W Einv (AW)^T z with Einv = (W^T A W)^-1, all in FP64.  Stage 1: per-block partial sums of the
// m dots AW_j . z (fixed grid, deterministic); stage 2: one block sums them in a fixed order and forms
// c = Einv d; then every entry of z loses sum_j W_j c_j.
__global__ void __launch_bounds__(kReduceThreads)
cs_dots_stage1(long long n, const double* __restrict__ AW, const double* __restrict__ z, double* __restrict__ partial) {
  const double* w = AW + (long long)blockIdx.y * n;
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    s = fma(w[i], z[i], s);
  const double t = block_sum(s);
  if (threadIdx.x == 0) partial[blockIdx.y * kReduceBlocks + blockIdx.x] = t;
}

__global__ void __launch_bounds__(kReduceThreads)
cs_coef_kernel(int m, const double* __restrict__ partial, SmallMat Einv, double* __restrict__ c) {
  __shared__ double d[4];
  for (int j = 0; j < m; ++j) {
    double s = 0.0;
    for (int b = threadIdx.x; b < kReduceBlocks; b += blockDim.x) s += partial[j * kReduceBlocks + b];
    const double t = block_sum(s);
    if (threadIdx.x == 0) d[j] = t;
  }
  __syncthreads();
  if (threadIdx.x < m) {
    double v = 0.0;
    for (int k = 0; k < m; ++k) v = fma(Einv.e[threadIdx.x * m + k], d[k], v);
    c[threadIdx.x] = v;
  }
}

__global__ void cs_sub_kernel(long long n, int m, const double* __restrict__ W, const double* __restrict__ c,
                              double* __restrict__ z) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = 0.0;
  for (int j = 0; j < m; ++j) v = fma(W[(long long)j * n + i], c[j], v);
  z[i] -= v;
}

int launch_coarse_correct(long long n, int m, const double* W, const double* AW, const SmallMat& Einv, double* buf,
                          double* z, hipStream_t s) {
  if (m <= 0) return PMT_OK;
  double* c = buf;
  double* partial = buf + 4;
  cs_dots_stage1<<<dim3(kReduceBlocks, m), kReduceThreads, 0, s>>>(n, AW, z, partial);
  PMT_TRY(check_launch("cs_dots_stage1"));
  cs_coef_kernel<<<1, kReduceThreads, 0, s>>>(m, partial, Einv, c);
  PMT_TRY(check_launch("cs_coef_kernel"));
  cs_sub_kernel<<<blocks_for(n, 256), 256, 0, s>>>(n, m, W, c, z);
  return check_launch("cs_sub_kernel");
}

int launch_fill(long long n, double v, double* y, hipStream_t s) {
  if (n <= 0) return PMT_OK;
  fill_kernel<<<blocks_for(n, 256), 256, 0, s>>>(n, v, y);
  return check_launch("fill_kernel");
}

}  // namespace pmt

using namespace pmt;

extern "C" {

const char* pmt_last_error(void) { return g_last_error.c_str(); }

int pmt_abi_version(void) { return 1; }

int pmt_reduce_work_size(void) { return 2 * kReduceBlocks; }

int pmt_powlaw_fwd(long long n, double a, double b, double p, const double* x, double* y, void* stream) {
  if (n < 0 || (n > 0 && (!x || !y))) return arg_error("pmt_powlaw_fwd: bad argument");
  if (n == 0) return PMT_OK;
  powlaw_fwd_kernel<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(n, a, b, p, powlaw_int_exponent(p), x, y);
  return check_launch("powlaw_fwd_kernel");
}

int pmt_powlaw_bwd(long long n, double c, double q, const double* x, const double* dy, double* dx, int accumulate,
                   void* stream) {
  if (n < 0 || (n > 0 && (!x || !dy || !dx))) return arg_error("pmt_powlaw_bwd: bad argument");
  if (n == 0) return PMT_OK;
  powlaw_bwd_kernel<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(n, c, q, powlaw_int_exponent(q), x, dy, dx,
                                                                         accumulate);
  return check_launch("powlaw_bwd_kernel");
}

int pmt_heaviside_fwd(long long n, double beta, double eta, const double* x, double* y, void* stream) {
  if (n < 0 || (n > 0 && (!x || !y))) return arg_error("pmt_heaviside_fwd: bad argument");
  if (n == 0) return PMT_OK;
  heaviside_fwd_kernel<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(n, beta, eta, x, y);
  return check_launch("heaviside_fwd_kernel");
}

int pmt_heaviside_bwd(long long n, double beta, double eta, const double* x, const double* dy, double* dx,
                      int accumulate, double* dbeta, void* stream) {
  if (n < 0 || (n > 0 && (!x || !dy || !dx))) return arg_error("pmt_heaviside_bwd: bad argument");
  if (n == 0) return PMT_OK;
  heaviside_bwd_kernel<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(n, beta, eta, x, dy, dx, accumulate,
                                                                           dbeta);
  return check_launch("heaviside_bwd_kernel");
}

int pmt_oc_update(long long n, const double* x, const double* dg, double xmin, double xmax, double move,
                  double lmid, double* xnew, double* sum_out, double* work, void* stream) {
  if (n <= 0 || !x || !dg || !xnew || !sum_out || !work || !(lmid > 0.0))
    return arg_error("pmt_oc_update: bad argument");
  hipStream_t s = (hipStream_t)stream;
  oc_update_kernel<<<kReduceBlocks, kReduceThreads, 0, s>>>(n, x, dg, xmin, xmax, move, lmid, xnew, work);
  PMT_TRY(check_launch("oc_update_kernel"));
  dot_stage2<1><<<1, kReduceThreads, 0, s>>>(work, sum_out);
  return check_launch("dot_stage2");
}

int pmt_axpby(long long n, double a, const double* x, double b, double* y, void* stream) {
  if (n < 0 || (n > 0 && !y)) return arg_error("pmt_axpby: bad argument");
  return launch_axpby(n, a, x, b, y, (hipStream_t)stream);
}

int pmt_dot(long long n, const double* x, const double* y, double* out, double* work, void* stream) {
  if (n < 0 || !x || !y || !out || !work) return arg_error("pmt_dot: bad argument");
  return launch_dot(n, x, y, out, work, (hipStream_t)stream);
}

int pmt_jacobi_apply(long long n, double w, const double* diag, const double* r, double* z, void* stream) {
  if (n < 0 || (n > 0 && (!diag || !r || !z))) return arg_error("pmt_jacobi_apply: bad argument");
  if (n == 0) return PMT_OK;
  jacobi_apply_kernel<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(n, w, diag, r, z);
  return check_launch("jacobi_apply_kernel");
}

int pmt_dot_noalloc(int k, long long n, const double* x, const double* y, double* out, void* stream) {
  if (k < 1 || n < 0 || !x || !y || !out) return arg_error("pmt_dot_noalloc: bad argument");
  if (k > kNoAllocMaxCols)
    return arg_error(("pmt_dot_noalloc: k = " + std::to_string(k) + " column pairs exceed the maximum of " +
                      std::to_string(kNoAllocMaxCols) + " (the static scratch holds 8 x 1024 partial sums)").c_str());
  double* partial = nullptr;
  PMT_TRY(scratch_for_current_device(&partial));
  hipStream_t s = (hipStream_t)stream;
  dot_cols_stage1<<<dim3(kReduceBlocks, k), kReduceThreads, 0, s>>>(n, x, y, partial);
  PMT_TRY(check_launch("dot_cols_stage1"));
  dot_cols_stage2<<<k, kReduceThreads, 0, s>>>(partial, out);
  return check_launch("dot_cols_stage2");
}

int pmt_bc_split_ndof(int nelx, int nely, int nelz, int ndof, const uint8_t* bcmask, double bcdiag,
                      const double* rhs, double* rhs_loc, double* sol, void* stream) {
  if (!grid_ok(nelx, nely, nelz) || ndof < 1 || ndof > 8 || !rhs || !rhs_loc || !sol)
    return arg_error("pmt_bc_split: bad argument");
  GridDims g = make_grid(nelx, nely, nelz);
  bc_split_kernel<<<blocks_for(g.nnodes * ndof, 256), 256, 0, (hipStream_t)stream>>>(g.nnodes, ndof, bcmask, bcdiag,
                                                                                     rhs, rhs_loc, sol);
  return check_launch("bc_split_kernel");
}

int pmt_bc_split(int nelx, int nely, int nelz, const uint8_t* bcmask, double bcdiag, const double* rhs,
                 double* rhs_loc, double* sol, void* stream) {
  return pmt_bc_split_ndof(nelx, nely, nelz, nelz > 0 ? 3 : 2, bcmask, bcdiag, rhs, rhs_loc, sol, stream);
}

int pmt_axpy_free_ndof(int nelx, int nely, int nelz, int ndof, const uint8_t* bcmask, double a, const double* x,
                       double* y, void* stream) {
  if (!grid_ok(nelx, nely, nelz) || ndof < 1 || ndof > 8 || !x || !y) return arg_error("pmt_axpy_free: bad argument");
  GridDims g = make_grid(nelx, nely, nelz);
  axpy_free_kernel<<<blocks_for(g.nnodes * ndof, 256), 256, 0, (hipStream_t)stream>>>(g.nnodes, ndof, bcmask, a, x,
                                                                                      y);
  return check_launch("axpy_free_kernel");
}

int pmt_axpy_free(int nelx, int nely, int nelz, const uint8_t* bcmask, double a, const double* x, double* y,
                  void* stream) {
  return pmt_axpy_free_ndof(nelx, nely, nelz, nelz > 0 ? 3 : 2, bcmask, a, x, y, stream);
}

}  // extern "C"
