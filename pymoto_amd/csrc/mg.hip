// Multigrid-preconditioned CG runtime (pymoto/solvers/iterative.py:124-403).
//
// A pmt_mg handle owns the grid hierarchy: per level the inverse diagonal, work vectors and,
// below the fine grid, the Galerkin element matrices and the node stencil; the coarsest level
// holds an explicit dense inverse.  update() rebuilds everything from the SIMP densities of the
// fine grid; vcycle() and pcg_solve() only launch kernels on the caller's stream and never
// allocate.  The CG loop follows CG.solve step by step (single right-hand side): alpha/beta are
// read by the kernels from a small device array, and the only host synchronisation per
// iteration is the 8-byte read of ||r||^2 for the stopping test.
//
// Precision: the V-cycle runs in FP64 (the reference's arithmetic) or, after
// pmt_mg_set_precision(mg, 32), in FP32 - operator, smoother, transfers and stencils in FP32, built
// from the FP64 Galerkin hierarchy - while the coarsest solve, the CG recurrences, the outer
// operator, dots and the stopping test stay FP64.  CG reads the FP32 preconditioner output through
// one conversion, so the solution still reaches an FP64 residual tolerance.  With
// pmt_mg_set_precision(mg, 16) the damped-Jacobi sweeps of the fine grid run on the matrix cores
// with BF16 operands and FP32 accumulation (hex8 only); the fine residual that is restricted to the
// coarse grid stays FP32, so the coarse correction sees no BF16 rounding.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <limits>
#include <new>
#include <type_traits>
#include <vector>

#include "mg_impl.h"

namespace pmt {

template <typename T>
static int dalloc(pmt_mg* mg, T** p, long long n) {
  void* ptr = nullptr;
  const long long nb = (n > 0 ? n : 1) * (long long)sizeof(T);
  PMT_HIP(hipMalloc(&ptr, (size_t)nb));
  mg->allocs.push_back(ptr);
  mg->bytes += nb;
  *p = (T*)ptr;
  return PMT_OK;
}

static void release(pmt_mg* mg) {
  if (mg->s2) (void)hipStreamSynchronize(mg->s2);  // an inverse that no solve has joined yet
  block_cg_release(mg);
  for (void* p : mg->allocs) (void)hipFree(p);
  mg->allocs.clear();
  if (mg->host_sc) (void)hipHostFree(mg->host_sc);
  mg->host_sc = nullptr;
  if (mg->host_rr) (void)hipHostFree(mg->host_rr);
  mg->host_rr = nullptr;
  if (mg->host_counts) (void)hipHostFree(mg->host_counts);
  mg->host_counts = nullptr;
  if (mg->ev_fork) (void)hipEventDestroy(mg->ev_fork);
  if (mg->ev_join) (void)hipEventDestroy(mg->ev_join);
  if (mg->s2) (void)hipStreamDestroy(mg->s2);
  mg->ev_fork = mg->ev_join = nullptr;
  mg->s2 = nullptr;
}

// frees one allocation of the handle (bytes: its size, for the accounting)
static int mg_free(pmt_mg* mg, void* p, long long bytes) {
  for (size_t q = 0; q < mg->allocs.size(); ++q)
    if (mg->allocs[q] == p) {
      mg->allocs.erase(mg->allocs.begin() + q);
      mg->bytes -= bytes;
      PMT_HIP(hipFree(p));
      return PMT_OK;
    }
  return arg_error("mg_free: not an allocation of this handle");
}

template <typename T>
static const T* fine_x(pmt_mg* mg);
template <>
const double* fine_x<double>(pmt_mg* mg) {
  return mg->xe;
}
template <>
const float* fine_x<float>(pmt_mg* mg) {
  return mg->xe32;
}

// Coarsest grid: z = A_c^-1 b with the explicit FP64 inverse.  The reduced-precision V-cycles
// convert in and out: the coarsest solve stays FP64 (an FP32 copy of the inverse of this badly
// conditioned operator loses the smooth components the coarse correction exists for).
template <typename T>
static int coarse_solve(pmt_mg* mg, const T* b, T* out, hipStream_t s);
template <>
int coarse_solve<double>(pmt_mg* mg, const double* b, double* out, hipStream_t s) {
  PMT_TRY(mg_join(mg, s));
  return launch_gemv<double>(mg->nc, mg->Ainv, b, out, s);
}
template <>
int coarse_solve<float>(pmt_mg* mg, const float* b, float* out, hipStream_t s) {
  PMT_TRY(mg_join(mg, s));
  return launch_gemv_f32io(mg->nc, mg->Ainv, b, out, s);  // FP64 inverse and sums, FP32 in and out
}

// level operator in `mode`; Jacobi / Chebyshev sweeps use `omega` (and `up`, `beta`)
template <typename T>
static int level_apply(pmt_mg* mg, int l, int mode, const T* u, const T* b, T* y, hipStream_t s, double omega,
                       const T* up = nullptr, double beta = 0.0) {
  Level& L = mg->L[l];
  LevelBuf<T>& B = L.buf<T>();
  if constexpr (std::is_same<T, float>::value) {
    if (l == 0 && mg->prec == 16 && (mode == MODE_JACOBI || mode == MODE_CHEB))
      return launch_fine_mfma(mode, L.g, mg->K0, fine_x<T>(mg), mg->bcm, mg->bcdiag, u, b, B.dinv, omega, y, s, up,
                              beta);
  }
  if (l == 0)
    return launch_fine<T>(mode, L.g, mg->K0, mg->k0pat, fine_x<T>(mg), mg->bcm, mg->bcdiag, u, b, B.dinv, omega, y,
                          s, up, beta);
  if (l == mg->nlev - 1) {  // coarsest: only its full FP64 stencil exists (it is solved directly)
    if constexpr (std::is_same<T, double>::value)
      if (mode != MODE_CHEB) return launch_stencil<double>(mode, L.g, L.d.S, u, b, B.dinv, omega, y, s);
    return arg_error("level_apply: the coarsest level is solved directly");
  }
  if (l == 1 && mg->l1free)
    return launch_coarse1<T>(mode, L.g, mg->g0pat, mg->G0ext, B.Xh, L.tnode, mg->host_counts[0],
                             L.t_cap, L.TS, u, b, B.dinv, omega, y, s, up, beta);
  return launch_stencil_half<T>(mode, L.g, B.H, u, b, B.dinv, omega, y, s, up, beta);
}

template <typename T>
static int level_apply(pmt_mg* mg, int l, int mode, const T* u, const T* b, T* y, hipStream_t s) {
  return level_apply<T>(mg, l, mode, u, b, y, s, mg->L[l].omega);
}

// `steps` smoothing sweeps on level l, from zero or from cur.  Damped Jacobi: u += omega D^-1 (b - A u)
// (iterative.py:21-47, 232-255).  Chebyshev (no reference counterpart): the same sweep count as a
// degree-`steps` Chebyshev polynomial in D^-1 A on [lo, hi], by the three-term recurrence
// u_{k+1} = u_k + alpha_k D^-1 (b - A u_k) + beta_k (u_k - u_{k-1}), u_1 = u_0 + D^-1 (b - A u_0) / theta,
// theta = (hi+lo)/2, delta = (hi-lo)/2, sigma = theta/delta, rho_0 = 1/sigma,
// rho_k = 1/(2 sigma - rho_{k-1}), alpha_k = 2 rho_k/delta, beta_k = rho_k rho_{k-1}.
// The ping-pong buffers cur/oth swap after every sweep; a Chebyshev sweep overwrites u_{k-1} in place.
template <typename T>
static int smooth(pmt_mg* mg, int l, const T* b, T*& cur, T*& oth, bool from_zero, hipStream_t s,
                  bool first_done = false) {
  Level& L = mg->L[l];
  LevelBuf<T>& B = L.buf<T>();
  if (mg->smoother == 0) {
    int it0 = 0;
    if (from_zero) {
      // first_done: cur already holds omega D^-1 b (written by the CG update)
      if (!first_done) PMT_TRY(launch_jacobi0<T>(L.ndof, B.dinv, b, L.omega, cur, s));
      it0 = 1;
    }
    for (int it = it0; it < L.steps; ++it) {
      PMT_TRY(level_apply<T>(mg, l, MODE_JACOBI, cur, b, oth, s, L.omega));
      T* tmp = cur;
      cur = oth;
      oth = tmp;
    }
    return PMT_OK;
  }
  const double theta = 0.5 * (L.cheb_hi + L.cheb_lo), delta = 0.5 * (L.cheb_hi - L.cheb_lo);
  const double sigma = theta / delta;
  double rho_prev = 1.0 / sigma;
  if (from_zero) {
    // first_done: cur already holds D^-1 b / theta (written by the CG update)
    if (!first_done) PMT_TRY(launch_jacobi0<T>(L.ndof, B.dinv, b, 1.0 / theta, cur, s));
  } else {
    PMT_TRY(level_apply<T>(mg, l, MODE_JACOBI, cur, b, oth, s, 1.0 / theta));
    T* tmp = cur;
    cur = oth;
    oth = tmp;
  }
  for (int it = 1; it < L.steps; ++it) {
    const double rho = 1.0 / (2.0 * sigma - rho_prev);
    const T* up = (from_zero && it == 1) ? nullptr : oth;  // u_{k-1} (zero before the first sweep)
    PMT_TRY(level_apply<T>(mg, l, MODE_CHEB, cur, b, oth, s, 2.0 * rho / delta, up, rho * rho_prev));
    rho_prev = rho;
    T* tmp = cur;
    cur = oth;
    oth = tmp;
  }
  return PMT_OK;
}

// the factor of a level's first smoothing sweep from zero, u = f D^-1 b: omega (damped Jacobi) or 1 / theta
// (Chebyshev) - the value smooth() passes to jacobi0 (kernels that write this sweep on the side use it)
static double first_sweep_factor(const pmt_mg* mg, int l) {
  const Level& L = mg->L[l];
  return mg->smoother == 0 ? L.omega : 1.0 / (0.5 * (L.cheb_hi + L.cheb_lo));
}

// out = V-cycle(b) on level l (GeometricMultigrid.solve with x0 = None, iterative.py:222-256)
template <typename T>
static int vcycle(pmt_mg* mg, int l, const T* b, T* out, hipStream_t s, bool first_done = false) {
  if (l == mg->nlev - 1) return coarse_solve<T>(mg, b, out, s);
  Level& L = mg->L[l];
  Level& C = mg->L[l + 1];
  LevelBuf<T>& B = L.buf<T>();
  LevelBuf<T>& CB = C.buf<T>();
  // 2*steps-1 out-of-place sweeps follow the first one: start in the buffer that ends in `out`
  T* cur = B.t;
  T* oth = out;
  PMT_TRY(smooth<T>(mg, l, b, cur, oth, true, s, first_done));
  PMT_TRY(level_apply<T>(mg, l, MODE_RESIDUAL, cur, b, B.r, s));
  // a smoothed coarse level starts from zero: the restriction also writes its first sweep into the
  // buffer that sweep would write (CB.t), with the factor the smoother would use
  const bool fuse = l + 1 < mg->nlev - 1;
  const double w0 = first_sweep_factor(mg, l + 1);
  PMT_TRY(launch_restrict<T>(L.g, C.g, B.r, CB.b, s, fuse ? CB.dinv : nullptr, w0, fuse ? CB.t : nullptr));
  PMT_TRY(vcycle<T>(mg, l + 1, CB.b, CB.u, s, fuse));
  PMT_TRY(launch_prolong<T>(L.g, C.g, CB.u, cur, s));
  PMT_TRY(smooth<T>(mg, l, b, cur, oth, false, s));
  if (cur != out) return arg_error("vcycle: internal buffer parity error");
  return PMT_OK;
}

// reduced-precision V-cycles: z32 = M r, from r rounded to FP32 (already in the fine level's FP32 right-hand
// side when r32_ready: the CG update wrote it)
static int precondition_f32(pmt_mg* mg, const double* r, bool r32_ready, hipStream_t s, bool first_done = false) {
  Level& F = mg->L[0];
  if (!r32_ready) {
    // the rounding also writes the first fine sweep from zero (the CG update does the same in the loop)
    PMT_TRY(launch_to_float_first(F.ndof, r, F.f.b, F.f.dinv, (float)first_sweep_factor(mg, 0), F.f.t, s));
    first_done = true;
  }
  return vcycle<float>(mg, 0, F.f.b, mg->z32, s, first_done);
}

// the preconditioner as CG sees it: FP64 in, FP64 out
static int precondition(pmt_mg* mg, const double* r, double* z, hipStream_t s) {
  if (mg->prec == 64) return vcycle<double>(mg, 0, r, z, s);
  PMT_TRY(precondition_f32(mg, r, false, s));
  return launch_to_double(mg->L[0].ndof, mg->z32, z, s);
}

static int read_scalar(pmt_mg* mg, int k, double* v, hipStream_t s) {
  PMT_HIP(hipMemcpyAsync(mg->host_sc + k, mg->sc + k, sizeof(double), hipMemcpyDeviceToHost, s));
  PMT_HIP(hipStreamSynchronize(s));
  *v = mg->host_sc[k];
  return PMT_OK;
}

// The CG stopping test without a stream synchronisation (whose wake-up left the GPU idle for ~60 us per
// iteration): a one-thread kernel (publish_scalar) stores sc[k] and then a fresh sequence number into
// pinned memory, and the host polls for that sequence number (wait_published).  Work enqueued between
// the two keeps the GPU busy.  After 1 s without it the host falls back to the synchronising read, so a
// value that never arrives cannot hang the solve.
static std::atomic<long long> g_publish_seq{0};

// the sequence number of the next publish (kept for wait_published); unique in the process: pinned memory
// freed by another handle (with its last sequence number still in it) and handed out again here can
// never look like this publish
static double next_publish_seq(pmt_mg* mg) {
  mg->pub_seq = (double)g_publish_seq.fetch_add(1, std::memory_order_relaxed) + 1.0;
  return mg->pub_seq;
}

static int publish_scalar(pmt_mg* mg, int k, hipStream_t s) {
  const double seq = next_publish_seq(mg);
  return launch_publish(mg->sc + k, seq, mg->host_rr, s);
}

static int wait_published(pmt_mg* mg, int k, double* v, hipStream_t s) {
  volatile double* h = mg->host_rr;
  const double want = mg->pub_seq;
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    if (h[1] == want) {
      *v = h[0];
      return PMT_OK;
    }
    if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(1)) return read_scalar(mg, k, v, s);
  }
}

static int alloc_fp32(pmt_mg* mg) {
  int rc = PMT_OK;
  for (int l = 0; l < mg->nlev && rc == PMT_OK; ++l) {
    Level& L = mg->L[l];
    LevelBuf<float>& B = L.f;
    if (l == mg->nlev - 1) {
      if (l > 0) {
        rc = dalloc(mg, &B.u, L.ndof);
        if (rc == PMT_OK) rc = dalloc(mg, &B.b, L.ndof);
      }
      continue;
    }
    rc = dalloc(mg, &B.dinv, L.ndof);
    if (rc == PMT_OK) rc = dalloc(mg, &B.t, L.ndof);
    if (rc == PMT_OK) rc = dalloc(mg, &B.r, L.ndof);
    if (rc == PMT_OK) rc = dalloc(mg, &B.b, L.ndof);
    if (l > 0) {
      if (l == 1 && mg->l1free) {
        if (rc == PMT_OK) rc = dalloc(mg, &B.Xh, 8 * L.g.nel);
      } else {
        if (rc == PMT_OK) rc = dalloc(mg, &B.H, L.g.nnodes * half_stencil_coefficients(mg->dim));
      }
      if (rc == PMT_OK) rc = dalloc(mg, &B.u, L.ndof);
    }
  }
  if (rc == PMT_OK) rc = dalloc(mg, &mg->xe32, mg->L[0].g.nel);
  if (rc == PMT_OK) rc = dalloc(mg, &mg->z32, mg->L[0].ndof);
  return rc;
}

__global__ void hash_fill_kernel(long long n, double* __restrict__ v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    unsigned h = (unsigned)(i * 2654435761ull) ^ 0x9e3779b9u;
    h ^= h >> 15;
    h *= 0x2c1b3c6du;
    h ^= h >> 12;
    v[i] = (double)(h & 0xffffffu) / 16777216.0 - 0.5;
  }
}

// y = a * x (mode 0) or y = x / a (mode 1), elementwise
__global__ void ew_kernel(long long n, int mode, const double* __restrict__ a, const double* __restrict__ x,
                          double* __restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = mode ? x[i] / a[i] : a[i] * x[i];
}

// lambda_max(D^-1 A_l) by power iteration with the Rayleigh quotient (v.Av)/(v.Dv); levels >= 1
static int estimate_lmax(pmt_mg* mg, int l, int iters, hipStream_t s, double* out) {
  Level& L = mg->L[l];
  LevelBuf<double>& B = L.d;
  const long long n = L.ndof;
  double* v = B.t;
  double* w = B.r;
  double* z = B.b;
  hash_fill_kernel<<<blocks_for(n, 256), 256, 0, s>>>(n, v);
  PMT_TRY(check_launch("hash_fill_kernel"));
  double lam = 0.0;
  for (int it = 0; it < iters; ++it) {
    PMT_TRY(level_apply<double>(mg, l, MODE_APPLY, v, nullptr, w, s));
    ew_kernel<<<blocks_for(n, 256), 256, 0, s>>>(n, 1, B.dinv, v, z);  // z = D v
    PMT_TRY(check_launch("ew_kernel"));
    PMT_TRY(launch_dot(n, v, w, mg->sc + 6, mg->work, s));
    PMT_TRY(launch_dot(n, v, z, mg->sc + 7, mg->work, s));
    ew_kernel<<<blocks_for(n, 256), 256, 0, s>>>(n, 0, B.dinv, w, z);  // z = D^-1 A v
    PMT_TRY(check_launch("ew_kernel"));
    PMT_TRY(launch_dot(n, z, z, mg->sc + 8, mg->work, s));
    PMT_HIP(hipMemcpyAsync(mg->host_sc + 6, mg->sc + 6, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    PMT_HIP(hipStreamSynchronize(s));
    lam = mg->host_sc[6] / mg->host_sc[7];
    const double nz = std::sqrt(mg->host_sc[8]);
    if (!(nz > 0.0)) break;
    PMT_TRY(launch_axpby(n, 1.0 / nz, z, 0.0, v, s));
  }
  *out = lam;
  return PMT_OK;
}

// lambda_max(diag(K0)^-1 K0) on the host (power iteration on the symmetric scaling)
static double element_lmax(int ne, const double* K0) {
  std::vector<double> d(ne), v(ne, 1.0), w(ne);
  for (int i = 0; i < ne; ++i) d[i] = 1.0 / std::sqrt(K0[i * ne + i]);
  for (int i = 0; i < ne; ++i) v[i] = 1.0 + 0.01 * i;
  double lam = 0.0;
  for (int it = 0; it < 2000; ++it) {
    double nn = 0.0, vw = 0.0, vv = 0.0;
    for (int i = 0; i < ne; ++i) {
      double acc = 0.0;
      for (int j = 0; j < ne; ++j) acc += d[i] * K0[i * ne + j] * d[j] * v[j];
      w[i] = acc;
    }
    for (int i = 0; i < ne; ++i) {
      vw += v[i] * w[i];
      vv += v[i] * v[i];
      nn += w[i] * w[i];
    }
    lam = vw / vv;
    nn = std::sqrt(nn);
    for (int i = 0; i < ne; ++i) v[i] = w[i] / nn;
  }
  return lam;
}

int mg_level_apply(pmt_mg* mg, int l, int mode, const double* u, const double* b, double* y, hipStream_t s) {
  return level_apply<double>(mg, l, mode, u, b, y, s);
}

int mg_precondition(pmt_mg* mg, const double* r, double* z, hipStream_t s) { return precondition(mg, r, z, s); }

// Every stream that reads the coarsest inverse after an update waits for its inversion once: the handle
// lists the streams that have waited since the update (a stream ordered after a joined stream is not
// thereby joined itself).  Once the inversion has finished nobody needs to wait and the list is dropped.
// A stream under graph capture must already be listed (block CG joins its stream before capturing).
int mg_join(pmt_mg* mg, hipStream_t s) {
  if (!mg->pending_join) return PMT_OK;
  if (s == nullptr) {
    // the legacy default stream: wait for the inversion on the host instead of through an event, so the
    // ordering does not depend on how the null stream treats a wait on a non-blocking stream's event
    PMT_HIP(hipStreamSynchronize(mg->s2));
    mg->pending_join = false;
    mg->joined.clear();
    return PMT_OK;
  }
  for (hipStream_t j : mg->joined)
    if (j == s) return PMT_OK;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  PMT_HIP(hipStreamIsCapturing(s, &cap));
  if (cap == hipStreamCaptureStatusNone && hipEventQuery(mg->ev_join) == hipSuccess) {
    mg->pending_join = false;
    mg->joined.clear();
    return PMT_OK;
  }
  PMT_HIP(hipStreamWaitEvent(s, mg->ev_join, 0));
  mg->joined.push_back(s);
  return PMT_OK;
}

void mg_forget_stream(pmt_mg* mg, hipStream_t s) {
  for (size_t q = 0; q < mg->joined.size(); ++q)
    if (mg->joined[q] == s) {
      mg->joined.erase(mg->joined.begin() + q);
      return;
    }
}

int mg_alloc(pmt_mg* mg, void** p, long long bytes) {
  char* q = nullptr;
  PMT_TRY(dalloc(mg, &q, bytes));
  *p = q;
  return PMT_OK;
}

// Coarse stencils of a hex8 hierarchy without element matrices (moments.hip): flags of every level
// (they depend on the mask only), one read of the flagged counts (corrections are grown if needed),
// corrections, moments, then each level's stencil.
static int hex_coarse_stencils(pmt_mg* mg, hipStream_t s) {
  const int nl = mg->nlev;
  PMT_TRY(mg_join(mg, s));  // the previous inversion still reads the coarsest stencil rewritten below
  PMT_TRY(launch_level1_flags(mg->L[0].g, mg->bcm, mg->L[1].cflag, mg->L[1].list, mg->L[1].map, s));
  for (int l = 2; l < nl; ++l)
    PMT_TRY(launch_flags_up(mg->L[l - 1].g, mg->L[l - 1].cflag, mg->L[l].cflag, mg->L[l].list, mg->L[l].map, s));
  Level& L1 = mg->L[1];
  if (mg->l1free) {
    PMT_TRY(launch_coarse1_touched(L1.g, L1.list, L1.map, L1.tnode, L1.tnode + L1.g.nnodes, s));
    PMT_HIP(hipMemcpyAsync(mg->host_counts, L1.tnode + L1.g.nnodes, sizeof(int), hipMemcpyDeviceToHost, s));
  }
  for (int l = 1; l < nl; ++l)
    PMT_HIP(hipMemcpyAsync(mg->host_counts + l, mg->L[l].list + mg->L[l].g.nel, sizeof(int), hipMemcpyDeviceToHost,
                           s));
  PMT_HIP(hipStreamSynchronize(s));
  for (int l = 1; l < nl; ++l) {
    Level& L = mg->L[l];
    const long long need = mg->host_counts[l];
    if (need > L.dk_cap) {
      PMT_TRY(mg_free(mg, L.dK, L.dk_cap * 576 * (long long)sizeof(double)));
      L.dK = nullptr;
      L.dk_cap = 0;
      PMT_TRY(mg_alloc(mg, (void**)&L.dK, need * 576 * (long long)sizeof(double)));
      L.dk_cap = need;
    }
  }
  if (mg->l1free && mg->host_counts[0] > L1.t_cap) {
    if (L1.TS) PMT_TRY(mg_free(mg, L1.TS, L1.t_cap * 243 * (long long)sizeof(double)));
    L1.TS = nullptr;
    L1.t_cap = 0;
    PMT_TRY(mg_alloc(mg, (void**)&L1.TS, (long long)mg->host_counts[0] * 243 * (long long)sizeof(double)));
    L1.t_cap = mg->host_counts[0];
  }
  PMT_TRY(launch_level1_dk(mg->L[0].g, mg->K0, mg->xe, mg->bcm, mg->bcdiag, mg->L[1].cflag, mg->L[1].list,
                           mg->host_counts[1], mg->L[1].dK, s));
  if (mg->l1free)
    PMT_TRY(launch_coarse1_tstencil(L1.g, L1.tnode, mg->host_counts[0], L1.t_cap, L1.map, L1.dK, L1.TS, s));
  for (int l = 2; l < nl; ++l)
    PMT_TRY(launch_dk_up(mg->L[l - 1].g, mg->L[l].cflag, mg->L[l].list, mg->host_counts[l], mg->L[l - 1].map,
                         mg->L[l - 1].dK, mg->L[l].dK, s));
  PMT_TRY(launch_moments1(mg->L[0].g, mg->xe, mg->L[1].cflag, mg->L[1].M, s));
  for (int l = 2; l < nl; ++l) PMT_TRY(launch_moments_up(mg->L[l - 1].g, mg->L[l - 1].M, mg->L[l].M, s));
  // smoothing levels get the half stencil and D^-1 (FP64, and FP32 for reduced-precision cycles); the
  // coarsest level the full stencil for its dense inverse
  auto target = [&](int l) {
    Level& L = mg->L[l];
    StencilOut o;
    if (l == nl - 1) {
      o.S = L.d.S;
    } else {
      o.H = L.d.H;  // null for a level without a stencil: D^-1 only
      o.dinv = L.d.dinv;
      if (mg->prec != 64) {
        o.H32 = L.f.H;
        o.dinv32 = L.f.dinv;
      }
    }
    return o;
  };
  // the coarsest level first; its dense inverse (a chain of small, latency-bound launches) then runs on
  // the handle's second stream while this one builds the other levels and, after update, starts the
  // solve (coarse_solve joins it)
  const int lc = nl - 1;
  Level& Lc = mg->L[lc];
  if (lc == 1)
    PMT_TRY(launch_stencil_from_fine(mg->L[0].g, mg->G, mg->xe, L1.cflag, L1.map, L1.dK, target(1), s));
  else
    PMT_TRY(launch_stencil_from_moments(Lc.g, Lc.C, Lc.M, Lc.map, Lc.dK, target(lc), s));
  PMT_HIP(hipEventRecord(mg->ev_fork, s));
  // this stream's remaining work is enqueued first (a handful of launches), so it runs while the host
  // enqueues the ~45 launches of the inversion
  if (mg->l1free)
    PMT_TRY(launch_coarse1_xhat(mg->L[0].g, L1.g, mg->xe, L1.cflag, L1.d.Xh, mg->prec != 64 ? L1.f.Xh : nullptr, s));
  if (lc > 1)
    PMT_TRY(launch_stencil_from_fine(mg->L[0].g, mg->G, mg->xe, L1.cflag, L1.map, L1.dK, target(1), s));
  for (int l = 2; l < lc; ++l) {
    Level& L = mg->L[l];
    PMT_TRY(launch_stencil_from_moments(L.g, L.C, L.M, L.map, L.dK, target(l), s));
  }
  PMT_HIP(hipStreamWaitEvent(mg->s2, mg->ev_fork, 0));
  PMT_TRY(launch_dense_from_stencil(Lc.g, Lc.d.S, mg->Ainv, mg->s2));
  PMT_TRY(gj_invert(mg->nc, mg->Ainv, mg->gjP, mg->gjC, mg->s2));
  PMT_HIP(hipEventRecord(mg->ev_join, mg->s2));
  // joined by the first coarsest solve on each stream: the V-cycle descent before it hides the inversion
  mg->pending_join = true;
  mg->joined.clear();
  return PMT_OK;
}

}  // namespace pmt

using namespace pmt;

extern "C" {

int pmt_mg_create(int nelx, int nely, int nelz, int nlevels, const double* omega, const int* steps,
                  const double* K0_host, pmt_mg** out) {
  if (!out || !K0_host || !grid_ok(nelx, nely, nelz) || nlevels < 2 || nlevels > kMaxLevels)
    return arg_error("pmt_mg_create: bad argument (nlevels must be 2..12)");
  *out = nullptr;
  pmt_mg* mg = new (std::nothrow) pmt_mg();
  if (!mg) return arg_error("pmt_mg_create: out of host memory");
  mg->dim = nelz > 0 ? 3 : 2;
  mg->nlev = nlevels;
  mg->NE = (1 << mg->dim) * mg->dim;
  std::vector<double> K0ext(PMT_K0_EXT_LEN);
  mg->k0pat = pmt_k0_extend(mg->dim, K0_host, K0ext.data());
  mg->l1free = mg->dim == 3 && nlevels >= 3 && (mg->k0pat & PMT_K0_REFLECTION) && get_tuning(PMT_TUNE_COARSE1) == 1;
  mg->lmax_elem = element_lmax(mg->NE, K0_host);
  int ex = nelx, ey = nely, ez = nelz;
  for (int l = 0; l < nlevels; ++l) {
    Level& L = mg->L[l];
    L.g = make_grid(ex, ey, ez);
    L.ndof = L.g.nnodes * mg->dim;
    if (l < nlevels - 1) {
      if (ex % 2 || ey % 2 || ez % 2 || ex < 2 || ey < 2) {
        delete mg;
        return arg_error("pmt_mg_create: every level above the coarsest needs even element counts");
      }
      L.omega = omega ? omega[l] : 0.5;
      L.steps = steps ? steps[l] : 1;
      if (L.steps < 1 || !(L.omega > 0.0 && L.omega <= 1.0)) {
        delete mg;
        return arg_error("pmt_mg_create: smoothing steps must be >= 1 and 0 < omega <= 1");
      }
      ex /= 2;
      ey /= 2;
      ez /= 2;
    }
  }
  mg->nc = mg->L[nlevels - 1].ndof;
  if (mg->nc > 20000) {
    delete mg;
    return arg_error("pmt_mg_create: coarsest grid too large for the dense direct solve (> 20000 dofs); add levels");
  }
  int rc = PMT_OK;
#define ALLOC(ptr, n) \
  if (rc == PMT_OK) rc = dalloc(mg, &(ptr), (n));
  const int NE = mg->NE;
  const int NS = mg->dim == 3 ? 243 : 36;
  ALLOC(mg->K0, (long long)PMT_K0_EXT_LEN);  // K0 followed by its reflection-basis blocks
  ALLOC(mg->G, (long long)(1 << mg->dim) * NE * NE);
  if (mg->l1free) ALLOC(mg->G0ext, 864);
  const bool hex = mg->dim == 3;
  for (int l = 1; l < nlevels && hex; ++l) {
    Level& L = mg->L[l];
    if (rc == PMT_OK) rc = dalloc(mg, &L.cflag, L.g.nel);
    if (rc == PMT_OK) rc = dalloc(mg, &L.list, L.g.nel + 1);
    if (rc == PMT_OK) rc = dalloc(mg, &L.map, L.g.nel);
    ALLOC(L.M, 27 * L.g.nel);
    if (l >= 2) ALLOC(L.C, 64 * 27 * 9);
    // corrections for the elements along the largest face of the level's grid; grown at update if needed
    const long long face = std::max({(long long)L.g.ex * L.g.ey, (long long)L.g.ey * L.g.ez, (long long)L.g.ex * L.g.ez});
    L.dk_cap = std::max(1LL, std::min(L.g.nel, face));
    ALLOC(L.dK, L.dk_cap * NE * NE);
    if (l == 1 && mg->l1free) {
      if (rc == PMT_OK) rc = dalloc(mg, &L.tnode, L.g.nnodes + 1);
      // correction stencils for the nodes along the largest face; grown at update if needed
      L.t_cap = std::max({(long long)L.g.nx * L.g.ny, (long long)L.g.ny * L.g.nz, (long long)L.g.nx * L.g.nz});
      ALLOC(L.TS, L.t_cap * 243);
    }
  }
  if (rc == PMT_OK)
    rc = hip_status(hipHostMalloc((void**)&mg->host_counts, kMaxLevels * sizeof(int), hipHostMallocDefault),
                    "hipHostMalloc");
  if (rc == PMT_OK && hex) rc = hip_status(hipStreamCreateWithFlags(&mg->s2, hipStreamNonBlocking), "hipStreamCreate");
  if (rc == PMT_OK && hex) rc = hip_status(hipEventCreateWithFlags(&mg->ev_fork, hipEventDisableTiming), "hipEventCreate");
  if (rc == PMT_OK && hex) rc = hip_status(hipEventCreateWithFlags(&mg->ev_join, hipEventDisableTiming), "hipEventCreate");
  for (int l = 0; l < nlevels; ++l) {
    Level& L = mg->L[l];
    if (l == nlevels - 1) {
      if (l > 0) {
        if (!hex) ALLOC(L.Ke, L.g.nel * NE * NE);
        ALLOC(L.d.S, L.g.nnodes * NS);
        ALLOC(L.d.u, L.ndof);
        ALLOC(L.d.b, L.ndof);
      }
      continue;
    }
    ALLOC(L.d.dinv, L.ndof);
    ALLOC(L.d.t, L.ndof);
    ALLOC(L.d.r, L.ndof);
    if (l > 0) {
      if (!hex) ALLOC(L.Ke, L.g.nel * NE * NE);
      if (!hex) ALLOC(L.d.S, L.g.nnodes * NS);  // hex8: half stencil only above the coarsest level
      if (l == 1 && mg->l1free) {
        ALLOC(L.d.Xh, 8 * L.g.nel);
      } else {
        ALLOC(L.d.H, L.g.nnodes * half_stencil_coefficients(mg->dim));
      }
      ALLOC(L.d.u, L.ndof);
      ALLOC(L.d.b, L.ndof);
    }
  }
  ALLOC(mg->Ainv, mg->nc * mg->nc);
  ALLOC(mg->gjP, 64 * 64);
  ALLOC(mg->gjC, mg->nc * 64);
  const long long n0 = mg->L[0].ndof;
  ALLOC(mg->p, n0);
  ALLOC(mg->q, n0);
  ALLOC(mg->z, n0);
  ALLOC(mg->r, n0);
  ALLOC(mg->sc, 16);
  ALLOC(mg->work, 2 * kReduceBlocks);
  ALLOC(mg->rwork, cg_xr_parts(n0));
  ALLOC(mg->cs_buf, 4 + 4 * kReduceBlocks);
  if (hex && (mg->k0pat & PMT_K0_REFLECTION)) {
    mg->dotws_parts = fine_apply_dots_parts(mg->L[0].g);
    ALLOC(mg->dotws, 2 * mg->dotws_parts);
  }
#undef ALLOC
  if (rc == PMT_OK)
    rc = hip_status(hipHostMalloc((void**)&mg->host_rr, 2 * sizeof(double), hipHostMallocCoherent), "hipHostMalloc");
  if (rc == PMT_OK) {  // pinned memory is not cleared: no sequence number may be left in it
    mg->host_rr[0] = 0.0;
    mg->host_rr[1] = -1.0;
  }
  if (rc == PMT_OK) rc = hip_status(hipHostMalloc((void**)&mg->host_sc, 16 * sizeof(double), hipHostMallocDefault),
                                    "hipHostMalloc");
  if (rc == PMT_OK)
    rc = hip_status(hipMemcpy(mg->K0, K0ext.data(), K0ext.size() * sizeof(double), hipMemcpyHostToDevice),
                    "K0 upload");
  if (rc == PMT_OK) {
    std::vector<double> G((size_t)(1 << mg->dim) * NE * NE);
    galerkin_child_matrices(mg->dim, K0_host, G.data());
    rc = hip_status(hipMemcpy(mg->G, G.data(), G.size() * sizeof(double), hipMemcpyHostToDevice), "G upload");
    if (rc == PMT_OK && mg->l1free) {
      std::vector<double> ext(864);
      coarse1_table(G.data(), ext.data());  // from G_0, the child at the element's corner 0
      mg->g0pat = coarse1_table_pattern(ext.data(), mg->k0pat);
      rc = hip_status(hipMemcpy(mg->G0ext, ext.data(), ext.size() * sizeof(double), hipMemcpyHostToDevice),
                      "level-1 table upload");
    }
  }
  for (int l = 2; l < nlevels && hex && rc == PMT_OK; ++l) {
    std::vector<double> C(64 * 27 * 9);
    moment_tables(l, K0_host, C.data());
    rc = hip_status(hipMemcpy(mg->L[l].C, C.data(), C.size() * sizeof(double), hipMemcpyHostToDevice),
                    "moment table upload");
  }
  if (rc != PMT_OK) {
    release(mg);
    delete mg;
    return rc;
  }
  *out = mg;
  return PMT_OK;
}

int pmt_mg_destroy(pmt_mg* mg) {
  if (!mg) return PMT_OK;
  (void)hipDeviceSynchronize();
  release(mg);
  delete mg;
  return PMT_OK;
}

int pmt_mg_set_precision(pmt_mg* mg, int bits) {
  if (!mg || (bits != 64 && bits != 32 && bits != 16))
    return arg_error("pmt_mg_set_precision: bits must be 64, 32 or 16");
  if (bits == 16 && mg->dim != 3)
    return arg_error("pmt_mg_set_precision: the BF16 matrix-core operator is implemented for hex8 grids only");
  if (bits != 64 && !mg->xe32) PMT_TRY(alloc_fp32(mg));
  mg->prec = bits;
  mg->ready = false;
  block_cg_invalidate(mg);
  return PMT_OK;
}

int pmt_mg_set_smoother(pmt_mg* mg, int type, double lower, double upper) {
  if (!mg || (type != 0 && type != 1) || !(lower > 0.0 && lower < 1.0) || !(upper >= 1.0))
    return arg_error("pmt_mg_set_smoother: type 0 (damped Jacobi) or 1 (Chebyshev), 0 < lower < 1 <= upper");
  mg->smoother = type;
  mg->cheb_lower = lower;
  mg->cheb_upper = upper;
  mg->ready = false;
  block_cg_invalidate(mg);
  return PMT_OK;
}

int pmt_mg_smoother_interval(pmt_mg* mg, int level, double* lo, double* hi) {
  if (!mg || level < 0 || level >= mg->nlev - 1 || !lo || !hi) return arg_error("pmt_mg_smoother_interval: bad argument");
  *lo = mg->L[level].cheb_lo;
  *hi = mg->L[level].cheb_hi;
  return PMT_OK;
}

int pmt_mg_update(pmt_mg* mg, const double* xe, const uint8_t* bcmask, double bcdiag, void* stream) {
  if (!mg || !xe || !bcmask) return arg_error("pmt_mg_update: bad argument (the Dirichlet mask is mandatory)");
  hipStream_t s = (hipStream_t)stream;
  mg->xe = xe;
  mg->bcm = bcmask;
  mg->bcdiag = bcdiag;
  mg->ready = false;
  block_cg_invalidate(mg);
  Level& F = mg->L[0];
  PMT_TRY(launch_fine_diag(F.g, mg->K0, xe, bcmask, bcdiag, F.d.dinv, 1, s, mg->prec != 64 ? F.f.dinv : nullptr));
  if (mg->dim == 3) {
    PMT_TRY(hex_coarse_stencils(mg, s));
  } else {
    for (int l = 1; l < mg->nlev; ++l) {
      Level& L = mg->L[l];
      Level& P = mg->L[l - 1];
      if (l == 1)
        PMT_TRY(launch_galerkin_fine(P.g, mg->K0, xe, bcmask, bcdiag, L.Ke, s));
      else
        PMT_TRY(launch_galerkin_coarse(P.g, P.Ke, L.Ke, s));
      PMT_TRY(launch_stencil_from_elements(L.g, L.Ke, L.d.S, s));
    }
  }
  for (int l = 1; l < mg->nlev && mg->dim != 3; ++l) {  // hex8: written by the stencil kernels
    Level& L = mg->L[l];
    if (l < mg->nlev - 1) {
      PMT_TRY(launch_stencil_dinv<double>(L.g, L.d.S, L.d.dinv, s));
      PMT_TRY(launch_stencil_compress<double>(L.g, L.d.S, L.d.H, s));
    }
  }
  if (mg->dim != 3) {  // hex8: inverted on the second stream inside hex_coarse_stencils
    Level& Cl = mg->L[mg->nlev - 1];
    PMT_TRY(launch_dense_from_stencil(Cl.g, Cl.d.S, mg->Ainv, s));
    PMT_TRY(gj_invert(mg->nc, mg->Ainv, mg->gjP, mg->gjC, s));
  }
  if (mg->smoother == 1) {
    for (int l = 0; l < mg->nlev - 1; ++l) {
      double lam = mg->lmax_elem;  // the fine grid: lambda_max(D^-1 K) <= lambda_max(D0^-1 K0)
      if (l > 0) PMT_TRY(estimate_lmax(mg, l, 8, s, &lam));
      Level& L = mg->L[l];
      L.cheb_hi = (l > 0 ? mg->cheb_upper : 1.0) * lam;
      L.cheb_lo = mg->cheb_lower * L.cheb_hi;
    }
  }
  if (mg->prec != 64) {
    PMT_TRY(launch_to_float(F.g.nel, xe, mg->xe32, s));  // (F.f.dinv: written with F.d.dinv above)
    for (int l = 1; l < mg->nlev - 1 && mg->dim != 3; ++l) {
      Level& L = mg->L[l];
      PMT_TRY(launch_stencil_compress<float>(L.g, L.d.S, L.f.H, s));
      PMT_TRY(launch_stencil_dinv<float>(L.g, L.d.S, L.f.dinv, s));
    }
  }
  mg->ready = true;
  return PMT_OK;
}

int pmt_mg_vcycle(pmt_mg* mg, const double* r, double* z, void* stream) {
  if (!mg || !r || !z) return arg_error("pmt_mg_vcycle: bad argument");
  if (!mg->ready) {
    set_error("pmt_mg_vcycle: call pmt_mg_update first");
    return PMT_ERR_STATE;
  }
  return precondition(mg, r, z, (hipStream_t)stream);
}

int pmt_mg_level_apply(pmt_mg* mg, int level, const double* u, double* y, void* stream) {
  if (!mg || !u || !y || level < 0 || level >= mg->nlev) return arg_error("pmt_mg_level_apply: bad argument");
  if (!mg->ready) {
    set_error("pmt_mg_level_apply: call pmt_mg_update first");
    return PMT_ERR_STATE;
  }
  return level_apply<double>(mg, level, MODE_APPLY, u, nullptr, y, (hipStream_t)stream);
}

long long pmt_mg_level_ndof(pmt_mg* mg, int level) {
  if (!mg || level < 0 || level >= mg->nlev) return PMT_ERR_ARG;
  return mg->L[level].ndof;
}

long long pmt_mg_device_bytes(pmt_mg* mg) { return mg ? mg->bytes : PMT_ERR_ARG; }

int pmt_mg_set_coarse_space(pmt_mg* mg, int m, const double* W, const double* AW, const double* Einv_host) {
  if (!mg || m < 0 || m > 4 || (m > 0 && (!W || !AW || !Einv_host)))
    return arg_error("pmt_mg_set_coarse_space: bad argument (0 <= m <= 4)");
  mg->cs_m = m;
  mg->cs_W = m > 0 ? W : nullptr;
  mg->cs_AW = m > 0 ? AW : nullptr;
  for (int q = 0; q < 16; ++q) mg->cs_Einv.e[q] = q < m * m ? Einv_host[q] : 0.0;
  return PMT_OK;
}

int pmt_pcg_solve(pmt_mg* mg, const double* b, double* x, int x0_is_zero, double tol, int maxit, int restart,
                  int* iters, double* relres, double* history, void* stream) {
  if (!mg || !b || !x || maxit < 0 || restart < 1) return arg_error("pmt_pcg_solve: bad argument");
  if (!mg->ready) {
    set_error("pmt_pcg_solve: call pmt_mg_update first");
    return PMT_ERR_STATE;
  }
  hipStream_t s = (hipStream_t)stream;
  Level& F = mg->L[0];
  const long long n = F.ndof;
  double* sc = mg->sc;
  // sc: 0 p.q  1 p.r  2 r.r  3 q.z  4 b.b  5 z.z
  if (x0_is_zero) {
    PMT_TRY(launch_fill(n, 0.0, x, s));
    PMT_TRY(launch_axpby(n, 1.0, b, 0.0, mg->r, s));
  } else {
    PMT_TRY(level_apply<double>(mg, 0, MODE_RESIDUAL, x, b, mg->r, s));
  }
  PMT_TRY(launch_dot2(n, b, b, mg->r, mg->r, sc + 4, mg->work, s));  // sc[4] = b.b, sc[5] = r.r
  PMT_HIP(hipMemcpyAsync(mg->host_sc + 4, sc + 4, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  PMT_HIP(hipStreamSynchronize(s));
  const double bn = std::sqrt(mg->host_sc[4]);
  if (iters) *iters = 0;
  if (!std::isfinite(bn)) {
    if (relres) *relres = bn;
    set_error("pmt_pcg_solve: the right-hand side has non-finite entries");
    return PMT_ERR_NUMERIC;
  }
  if (bn == 0.0) {  // zero right-hand side: x = 0 is the solution
    PMT_TRY(launch_fill(n, 0.0, x, s));
    if (history) history[0] = 0.0;
    if (relres) *relres = 0.0;
    return PMT_OK;
  }
  double tval = std::sqrt(mg->host_sc[5]) / bn;
  if (history) history[0] = tval;
  if (relres) *relres = tval;
  if (!std::isfinite(tval)) {
    set_error("pmt_pcg_solve: the initial residual is not finite (non-finite start vector or operator)");
    return PMT_ERR_NUMERIC;
  }
  if (tval <= tol) return PMT_OK;
  PMT_TRY(precondition(mg, mg->r, mg->z, s));
  PMT_TRY(launch_coarse_correct(n, mg->cs_m, mg->cs_W, mg->cs_AW, mg->cs_Einv, mg->cs_buf, mg->z, s));
  PMT_TRY(launch_dot(n, mg->z, mg->z, sc + 5, mg->work, s));
  PMT_TRY(launch_scale_rsqrt(n, sc, 5, mg->z, mg->p, s));
  int i = 0;
  double tprev = 0.0;  // the relative residual one iteration before tval (0: none yet)
  bool x_pending = false;  // x lacks this iteration's alpha p (added by the p update or after the loop)
  for (i = 0; i < maxit; ++i) {
    // q = A p, sc[0] = p.q, sc[1] = p.r: fused into the reflection-basis operator when it is the fine
    // operator, else K.u and one two-dot pass
    const int frc = mg->dotws ? launch_fine_apply_dots(F.g, mg->K0, mg->k0pat, mg->xe, mg->bcm, mg->bcdiag, mg->p,
                                                       mg->r, mg->q, mg->dotws, mg->dotws_parts, sc, s)
                              : PMT_ERR_STATE;
    if (frc == PMT_ERR_STATE) {
      PMT_TRY(level_apply<double>(mg, 0, MODE_APPLY, mg->p, nullptr, mg->q, s));
      PMT_TRY(launch_dot2(n, mg->p, mg->q, mg->p, mg->r, sc, mg->work, s));
    } else {
      PMT_TRY(frc);
    }
    // reduced-precision V-cycles: the update also writes r in FP32 (the V-cycle's input), and z stays
    // FP32 (q.z and p = z + beta p widen it exactly: the same sums as through an FP64 copy)
    const bool mixed = mg->prec != 64;
    // the update also writes the V-cycle's first fine sweep from zero into the fine level's work vector:
    // omega D^-1 r32 (damped Jacobi) or D^-1 r32 / theta (Chebyshev), the same FP32 expression as the
    // sweep kernel
    const bool first = mixed && (mg->smoother == 0 || mg->smoother == 1);
    const double first_factor = first_sweep_factor(mg, 0);
    bool r32_ready = false;
    bool published = false;  // r.r already on its way to the host (with the r update's reduction)
    if (i % restart == 0) {
      PMT_TRY(launch_cg_x_update(n, sc, mg->p, x, s));
      PMT_TRY(level_apply<double>(mg, 0, MODE_RESIDUAL, x, b, mg->r, s));
      PMT_TRY(launch_dot(n, mg->r, mg->r, sc + 2, mg->work, s));
    } else {
      // x += alpha p is deferred to the p update, which reads p anyway (x_pending until it runs)
      const double seq = next_publish_seq(mg);
      PMT_TRY(launch_cg_xr_update(n, sc, mg->p, mg->q, nullptr, mg->r, mg->rwork, sc + 2, s, mixed ? F.f.b : nullptr,
                                  first ? F.f.dinv : nullptr, (float)first_factor, first ? F.f.t : nullptr,
                                  mg->host_rr, seq));
      r32_ready = mixed;
      x_pending = true;
      published = true;
    }
    // z = M r, sc[3] = q.z, p = z + beta p: reads r and q, writes only z, sc[3] and p (no result)
    auto next_direction = [&]() -> int {
      if (mixed) {
        PMT_TRY(precondition_f32(mg, mg->r, r32_ready, s, r32_ready && first));
        if (mg->cs_m == 0) {
          PMT_TRY(launch_dot_flat(n, mg->q, mg->z32, sc + 3, mg->rwork, s));
          PMT_TRY(launch_cg_p_update(n, sc, mg->z32, mg->p, s, x_pending ? x : nullptr));
          x_pending = false;
          return PMT_OK;
        }
        // with a coarse space the projection runs in FP64: rounded to FP32 it would leave components
        // along W of relative size 1e-7 in every direction, and CG would stall at that level
        PMT_TRY(launch_to_double(n, mg->z32, mg->z, s));
      } else {
        PMT_TRY(precondition(mg, mg->r, mg->z, s));
      }
      PMT_TRY(launch_coarse_correct(n, mg->cs_m, mg->cs_W, mg->cs_AW, mg->cs_Einv, mg->cs_buf, mg->z, s));
      PMT_TRY(launch_dot_flat(n, mg->q, mg->z, sc + 3, mg->rwork, s));
      PMT_TRY(launch_cg_p_update(n, sc, mg->z, mg->p, s, x_pending ? x : nullptr));
      x_pending = false;
      return PMT_OK;
    };
    if (!published) PMT_TRY(publish_scalar(mg, 2, s));
    // when the residual's rate of decrease says this iteration will not stop, the next direction is
    // enqueued before the host waits for r.r, so the GPU stays busy meanwhile; if it does stop after all,
    // that work is discarded (the iterates and the iteration count are the same either way)
    const double rate = tprev > 0.0 ? tval / tprev : 0.3;
    const bool ahead = tval * rate > 2.0 * tol;
    if (ahead) PMT_TRY(next_direction());
    double rr = 0.0;
    PMT_TRY(wait_published(mg, 2, &rr, s));
    tprev = tval;
    tval = std::sqrt(rr) / bn;
    if (history) history[i + 1] = tval;
    if (!std::isfinite(tval)) {
      PMT_HIP(hipStreamSynchronize(s));  // nothing enqueued ahead is left running on the caller's buffers
      if (iters) *iters = i + 1;
      if (relres) *relres = tval;
      set_error("pmt_pcg_solve: the residual became non-finite at iteration " + std::to_string(i + 1) +
                " (non-finite operator, preconditioner or breakdown)");
      return PMT_ERR_NUMERIC;
    }
    if (tval <= tol) break;
    if (!ahead) PMT_TRY(next_direction());
  }
  // stopped before the p update of the last iteration ran: its x update
  if (x_pending) PMT_TRY(launch_cg_x_update(n, sc, mg->p, x, s));
  if (iters) *iters = i < maxit ? i + 1 : maxit;
  if (relres) *relres = tval;
  return PMT_OK;
}

}  // extern "C"
