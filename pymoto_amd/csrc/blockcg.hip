// Block preconditioned CG for k right-hand sides that share one operator: CG.solve with an
// (ndof, k) right-hand side (pymoto/solvers/iterative.py:259-403, Ji & Li's breakdown-free block
// CG).  Per iteration, with P, Q, R, Z, W the k-column blocks:
//   Q = A P;  alpha = (P^T Q)^-1 P^T R;  X += P alpha;
//   R = B - A X every `restart` iterations (including the first), else R -= Q alpha;
//   stop when max_j ||r_j|| / ||b_j|| <= tol;
//   Z = M R (one V-cycle per column);  beta = -(P^T Q)^-1 Q^T Z;  P = orth(Z + P beta).
// orth() is the reference's modified Gram-Schmidt (iterative.py:259-292) on the vectors: column i of
// the result is v_i minus its projections on the kept columns before it, dropped when its squared norm
// falls below 1e-15 of ||v_i||^2 (normalised to unit length in the initial orth, as the reference
// does).  Dropped columns are kept as zero vectors with their rows and columns masked out of the
// k x k algebra, which is the same arithmetic as compacting them.
//
// Everything runs on the device: the k x k products come from one deterministic two-stage reduction
// per block pair, the small dense algebra from a one-thread kernel, the block updates from one fused
// kernel per block, the Gram-Schmidt coefficients from device dot products.  So an iteration is two fixed launch sequences (before and after
// the convergence test) that are captured once as HIP graphs and replayed; the host reads 8 bytes
// per iteration for the stopping test, as the single-right-hand-side runtime does.
#include <cmath>
#include <new>

#include "mg_impl.h"

namespace pmt {

constexpr int kMaxRhs = 8;

struct VecSet {
  const double* v[kMaxRhs];
};
struct OutSet {
  double* v[kMaxRhs];
};

// small device state, in doubles
struct Small {
  enum {
    G0 = 0,                          // Gram results, 4 slots of kMaxRhs^2
    PQINV = 4 * kMaxRhs * kMaxRhs,   // (P^T Q)^-1 on the active set
    ALPHA = PQINV + kMaxRhs * kMaxRhs,
    BETA = ALPHA + kMaxRhs * kMaxRhs,
    COEF = BETA + kMaxRhs * kMaxRhs,  // orth coefficients, column i of p over the input columns
    ACTIVE = COEF + kMaxRhs * kMaxRhs,
    BNORM = ACTIVE + kMaxRhs,
    TVAL = BNORM + kMaxRhs,
    DONE = TVAL + kMaxRhs,  // [DONE] = 1 when converged, [DONE+1] = max tval
    MGS_B0 = DONE + 8,      // ||v_i||^2 before orthogonalisation
    MGS_D = MGS_B0 + kMaxRhs,  // p_j . p_j of the kept columns (1 when normalised)
    MGS_T = MGS_D + kMaxRhs,   // the current dot product
    SIZE = MGS_T + 8
  };
};

struct BlockCG {
  long long n = 0;
  int kcap = 0;
  double* mem = nullptr;  // X, B, P, Q, R, Z, W blocks of kcap columns + a zero column
  double* X = nullptr;
  double* B = nullptr;
  double* P = nullptr;
  double* Q = nullptr;
  double* R = nullptr;
  double* Z = nullptr;
  double* W = nullptr;
  double* zero = nullptr;
  double* sm = nullptr;       // Small::SIZE doubles
  double* partial = nullptr;  // kReduceBlocks * kMaxRhs^2
  double* host = nullptr;     // pinned mirror of the convergence test
  // captured iterations: [0] operator + alpha + restart residual, [1] same with recurrence, [2] beta + orth
  hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
  int graph_k = 0, graph_epoch = -1;
  // captured work runs on a stream of the handle (the caller's may be the legacy default stream,
  // which cannot be captured), ordered after and before the caller's stream with two events
  hipStream_t cs = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
};

// ------------------------------------------------------------------------------------------------
// kernels
// ------------------------------------------------------------------------------------------------

// partial Gram block G[a][b] = sum_i A_a[i] * B_b[i] over this block's rows, a, b < K
template <int K>
__global__ void __launch_bounds__(256) gram_stage1(long long n, VecSet A, VecSet Bv, double* __restrict__ partial) {
  double acc[K][K];
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int b = 0; b < K; ++b) acc[a][b] = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double va[K], vb[K];
#pragma unroll
    for (int a = 0; a < K; ++a) va[a] = A.v[a][i];
#pragma unroll
    for (int b = 0; b < K; ++b) vb[b] = Bv.v[b][i];
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
      for (int b = 0; b < K; ++b) acc[a][b] = fma(va[a], vb[b], acc[a][b]);
  }
  __shared__ double red[4][K * K];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int b = 0; b < K; ++b) {
      double v = acc[a][b];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if (lane == 0) red[wave][a * K + b] = v;
    }
  __syncthreads();
  for (int t = threadIdx.x; t < K * K; t += blockDim.x)
    partial[(long long)blockIdx.x * K * K + t] = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
}

// G[t] = sum over blocks of partial[block][t] in a fixed order (one block per entry)
__global__ void __launch_bounds__(256) gram_stage2(const double* __restrict__ partial, int nblocks, int m,
                                                   double* __restrict__ G) {
  const int t = blockIdx.x;
  double s = 0.0;
  for (int bi = threadIdx.x; bi < nblocks; bi += blockDim.x) s += partial[(long long)bi * m + t];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) G[t] = red[0];
}

// out_j = base_scale * base_j + gamma * sum_i in_i * C[i][j]   (j < kout, i < K; C row stride kMaxRhs)
template <int K>
__global__ void __launch_bounds__(256) combine_kernel(long long n, int kout, VecSet in, const double* __restrict__ C,
                                                      double gamma, VecSet base, double base_scale, OutSet out) {
  __shared__ double c[K * kMaxRhs];
  for (int t = threadIdx.x; t < K * kMaxRhs; t += blockDim.x) c[t] = gamma * C[(t / kMaxRhs) * kMaxRhs + t % kMaxRhs];
  __syncthreads();
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v[K];
#pragma unroll
  for (int a = 0; a < K; ++a) v[a] = in.v[a][i];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    if (j >= kout) break;
    double s = base_scale != 0.0 ? base_scale * base.v[j][i] : 0.0;
#pragma unroll
    for (int a = 0; a < K; ++a) s = fma(v[a], c[a * kMaxRhs + j], s);
    out.v[j][i] = s;
  }
}

// Gauss-Jordan inverse of the active principal submatrix of M (k x k, stride kMaxRhs) into Minv;
// inactive rows/columns of Minv are zero.  Single thread.
__device__ void small_inverse(int k, const double* M, const double* active, double* Minv) {
  double a[kMaxRhs][2 * kMaxRhs];
  int idx[kMaxRhs];
  int m = 0;
  for (int i = 0; i < k; ++i)
    if (active[i] != 0.0) idx[m++] = i;
  for (int r = 0; r < m; ++r)
    for (int c = 0; c < m; ++c) {
      a[r][c] = M[idx[r] * kMaxRhs + idx[c]];
      a[r][m + c] = r == c ? 1.0 : 0.0;
    }
  for (int c = 0; c < m; ++c) {
    int piv = c;
    for (int r = c + 1; r < m; ++r)
      if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
    if (piv != c)
      for (int t = 0; t < 2 * m; ++t) {
        const double tmp = a[c][t];
        a[c][t] = a[piv][t];
        a[piv][t] = tmp;
      }
    const double inv = 1.0 / a[c][c];
    for (int t = 0; t < 2 * m; ++t) a[c][t] *= inv;
    for (int r = 0; r < m; ++r) {
      if (r == c) continue;
      const double f = a[r][c];
      if (f != 0.0)
        for (int t = 0; t < 2 * m; ++t) a[r][t] -= f * a[c][t];
    }
  }
  for (int i = 0; i < kMaxRhs * kMaxRhs; ++i) Minv[i] = 0.0;
  for (int r = 0; r < m; ++r)
    for (int c = 0; c < m; ++c) Minv[idx[r] * kMaxRhs + idx[c]] = a[r][m + c];
}

// Modified Gram-Schmidt on the vectors, as orth() does (iterative.py:259-292): column i starts as
// v_i and loses its projection on every kept column j < i in turn; it is dropped (zeroed, ACTIVE = 0)
// when its squared norm falls below 1e-15 of ||v_i||^2, and scaled to unit length when normalising.
// (Orthogonalising through the Gram matrix V^T V instead squares the condition number of V, which the
// block recurrence does not tolerate near convergence.)
__global__ void mgs_axpy_kernel(long long n, const double* __restrict__ sm, int j, int normalize,
                                const double* __restrict__ pj, double* __restrict__ vi) {
  const double coef =
      sm[Small::ACTIVE + j] != 0.0 ? sm[Small::MGS_T] / (normalize ? 1.0 : sm[Small::MGS_D + j]) : 0.0;
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) vi[r] = fma(-coef, pj[r], vi[r]);
}

__global__ void mgs_finish_kernel(long long n, double* __restrict__ sm, int i, int normalize,
                                  double* __restrict__ vi) {
  const double b0 = sm[Small::MGS_B0 + i], b1 = sm[Small::MGS_T];
  const bool keep = b0 > 0.0 && b1 / b0 >= 1e-15;
  const double scale = keep ? (normalize ? 1.0 / sqrt(b1) : 1.0) : 0.0;
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) vi[r] *= scale;
  if (r == 0) {
    sm[Small::ACTIVE + i] = keep ? 1.0 : 0.0;
    sm[Small::MGS_D + i] = normalize ? 1.0 : b1;
  }
}

// alpha = (P^T Q)^-1 P^T R  (G slot 0 = P^T Q, slot 1 = P^T R)
__global__ void bcg_alpha_kernel(int k, double* sm) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double* act = sm + Small::ACTIVE;
  double* Minv = sm + Small::PQINV;
  small_inverse(k, sm + Small::G0, act, Minv);
  const double* PR = sm + Small::G0 + kMaxRhs * kMaxRhs;
  double* al = sm + Small::ALPHA;
  for (int i = 0; i < kMaxRhs * kMaxRhs; ++i) al[i] = 0.0;
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) {
      double s = 0.0;
      for (int m = 0; m < k; ++m) s += Minv[i * kMaxRhs + m] * PR[m * kMaxRhs + j];
      al[i * kMaxRhs + j] = s;
    }
}

// beta = -(P^T Q)^-1 Q^T Z  (G slot 2 = Q^T Z)
__global__ void bcg_beta_kernel(int k, double* sm) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double* Minv = sm + Small::PQINV;
  const double* QZ = sm + Small::G0 + 2 * kMaxRhs * kMaxRhs;
  double* be = sm + Small::BETA;
  for (int i = 0; i < kMaxRhs * kMaxRhs; ++i) be[i] = 0.0;
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) {
      double s = 0.0;
      for (int m = 0; m < k; ++m) s += Minv[i * kMaxRhs + m] * QZ[m * kMaxRhs + j];
      be[i * kMaxRhs + j] = -s;
    }
}

// per-column relative residuals from G slot 3 = R^T R; first call (init != 0) also sets ||b_j||
// from G slot 2 = B^T B.  DONE = 1 when max_j tval_j <= tol.
__global__ void bcg_check_kernel(int k, int init, double tol, double* sm) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double* RR = sm + Small::G0 + 3 * kMaxRhs * kMaxRhs;
  double* bn = sm + Small::BNORM;
  if (init) {
    const double* BB = sm + Small::G0 + 2 * kMaxRhs * kMaxRhs;
    for (int j = 0; j < k; ++j) bn[j] = sqrt(BB[j * kMaxRhs + j]);
  }
  // a non-finite column (NaN/inf in b, x0 or the iterates) makes the maximum NaN, which the host turns into
  // an error; a zero column counts as converged
  double mx = 0.0;
  for (int j = 0; j < k; ++j) {
    const double t = bn[j] == 0.0 ? 0.0 : sqrt(RR[j * kMaxRhs + j]) / bn[j];
    sm[Small::TVAL + j] = t;
    mx = (t != t || mx != mx) ? t + mx : fmax(mx, t);
  }
  sm[Small::DONE] = mx <= tol ? 1.0 : 0.0;
  sm[Small::DONE + 1] = mx;
}

// ------------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------------

static VecSet vset(const double* base, long long n, int k, const double* zero) {
  VecSet s;
  for (int j = 0; j < kMaxRhs; ++j) s.v[j] = j < k ? base + (long long)j * n : zero;
  return s;
}
static OutSet oset(double* base, long long n, int k) {
  OutSet s;
  for (int j = 0; j < kMaxRhs; ++j) s.v[j] = j < k ? base + (long long)j * n : nullptr;
  return s;
}

static int kround(int k) { return k <= 1 ? 1 : k <= 2 ? 2 : k <= 4 ? 4 : 8; }

// G slot <- A^T B over the first k columns of each block
static int gram(BlockCG* w, int k, const double* A, const double* Bv, int slot, hipStream_t s) {
  const int K = kround(k);
  const VecSet a = vset(A, w->n, k, w->zero), b = vset(Bv, w->n, k, w->zero);
  switch (K) {
    case 1: gram_stage1<1><<<kReduceBlocks, 256, 0, s>>>(w->n, a, b, w->partial); break;
    case 2: gram_stage1<2><<<kReduceBlocks, 256, 0, s>>>(w->n, a, b, w->partial); break;
    case 4: gram_stage1<4><<<kReduceBlocks, 256, 0, s>>>(w->n, a, b, w->partial); break;
    default: gram_stage1<8><<<kReduceBlocks, 256, 0, s>>>(w->n, a, b, w->partial); break;
  }
  PMT_TRY(check_launch("gram_stage1"));
  // stage 2 writes a K x K block; scatter into the kMaxRhs-strided slot with a tiny kernel pass
  double* G = w->sm + Small::G0 + slot * kMaxRhs * kMaxRhs;
  if (K == kMaxRhs) {
    gram_stage2<<<K * K, 256, 0, s>>>(w->partial, kReduceBlocks, K * K, G);
    return check_launch("gram_stage2");
  }
  // K < kMaxRhs: reduce row by row into the strided layout
  for (int a = 0; a < K; ++a) {
    gram_stage2<<<K, 256, 0, s>>>(w->partial + a * K, kReduceBlocks, K * K, G + a * kMaxRhs);
    PMT_TRY(check_launch("gram_stage2"));
  }
  return PMT_OK;
}

// out_j = base_scale*base_j + gamma * sum_i in_i C[i][j], j < k
static int combine(BlockCG* w, int k, const double* in, const double* C, double gamma, const double* base,
                   double base_scale, double* out, hipStream_t s) {
  const int K = kround(k);
  const VecSet vi = vset(in, w->n, k, w->zero);
  const VecSet vb = vset(base ? base : w->zero, base ? w->n : 0, base ? k : 0, w->zero);
  const OutSet vo = oset(out, w->n, k);
  const unsigned blocks = blocks_for(w->n, 256);
  switch (K) {
    case 1: combine_kernel<1><<<blocks, 256, 0, s>>>(w->n, k, vi, C, gamma, vb, base_scale, vo); break;
    case 2: combine_kernel<2><<<blocks, 256, 0, s>>>(w->n, k, vi, C, gamma, vb, base_scale, vo); break;
    case 4: combine_kernel<4><<<blocks, 256, 0, s>>>(w->n, k, vi, C, gamma, vb, base_scale, vo); break;
    default: combine_kernel<8><<<blocks, 256, 0, s>>>(w->n, k, vi, C, gamma, vb, base_scale, vo); break;
  }
  return check_launch("combine_kernel");
}

static int single(int which, int k, int arg, double tol, double* sm, hipStream_t s) {
  switch (which) {
    case 1: bcg_alpha_kernel<<<1, 64, 0, s>>>(k, sm); break;
    case 2: bcg_beta_kernel<<<1, 64, 0, s>>>(k, sm); break;
    default: bcg_check_kernel<<<1, 64, 0, s>>>(k, arg, tol, sm); break;
  }
  return check_launch("block CG small algebra");
}

static double* col(double* base, long long n, int j) { return base + (long long)j * n; }

// P = orth(V) (iterative.py:259-292), modified Gram-Schmidt column by column
static int orth_into_P(BlockCG* w, int k, bool normalize, const double* V, hipStream_t s) {
  const long long n = w->n;
  const unsigned blocks = blocks_for(n, 256);
  const int nz = normalize ? 1 : 0;
  for (int i = 0; i < k; ++i) {
    double* pi = w->P + (long long)i * n;
    const double* vi = V + (long long)i * n;
    PMT_TRY(launch_axpby(n, 1.0, vi, 0.0, pi, s));
    PMT_TRY(launch_dot(n, vi, vi, w->sm + Small::MGS_B0 + i, w->partial, s));
    for (int j = 0; j < i; ++j) {
      const double* pj = w->P + (long long)j * n;
      PMT_TRY(launch_dot(n, pi, pj, w->sm + Small::MGS_T, w->partial, s));
      mgs_axpy_kernel<<<blocks, 256, 0, s>>>(n, w->sm, j, nz, pj, pi);
      PMT_TRY(check_launch("mgs_axpy_kernel"));
    }
    PMT_TRY(launch_dot(n, pi, pi, w->sm + Small::MGS_T, w->partial, s));
    mgs_finish_kernel<<<blocks, 256, 0, s>>>(n, w->sm, i, nz, pi);
    PMT_TRY(check_launch("mgs_finish_kernel"));
  }
  return PMT_OK;
}

// iteration part 1: Q = A P, alpha, X += P alpha, R update, convergence test
static int iter_part1(pmt_mg* mg, BlockCG* w, int k, bool restart_now, double tol, hipStream_t s) {
  const long long n = w->n;
  for (int j = 0; j < k; ++j) PMT_TRY(mg_level_apply(mg, 0, MODE_APPLY, col(w->P, n, j), nullptr, col(w->Q, n, j), s));
  PMT_TRY(gram(w, k, w->P, w->Q, 0, s));
  PMT_TRY(gram(w, k, w->P, w->R, 1, s));
  PMT_TRY(single(1, k, 0, tol, w->sm, s));
  PMT_TRY(combine(w, k, w->P, w->sm + Small::ALPHA, 1.0, w->X, 1.0, w->X, s));
  if (restart_now) {
    for (int j = 0; j < k; ++j)
      PMT_TRY(mg_level_apply(mg, 0, MODE_RESIDUAL, col(w->X, n, j), col(w->B, n, j), col(w->R, n, j), s));
  } else {
    PMT_TRY(combine(w, k, w->Q, w->sm + Small::ALPHA, -1.0, w->R, 1.0, w->R, s));
  }
  PMT_TRY(gram(w, k, w->R, w->R, 3, s));
  return single(3, k, 0, tol, w->sm, s);
}

// iteration part 2: Z = M R, beta, P = orth(Z + P beta)
static int iter_part2(pmt_mg* mg, BlockCG* w, int k, double tol, hipStream_t s) {
  const long long n = w->n;
  for (int j = 0; j < k; ++j) PMT_TRY(mg_precondition(mg, col(w->R, n, j), col(w->Z, n, j), s));
  PMT_TRY(gram(w, k, w->Q, w->Z, 2, s));
  PMT_TRY(single(2, k, 0, tol, w->sm, s));
  PMT_TRY(combine(w, k, w->P, w->sm + Small::BETA, 1.0, w->Z, 1.0, w->W, s));
  return orth_into_P(w, k, false, w->W, s);
}

static void drop_graphs(BlockCG* w) {
  for (auto& e : w->exec) {
    if (e) (void)hipGraphExecDestroy(e);
    e = nullptr;
  }
  w->graph_k = 0;
  w->graph_epoch = -1;
}

static void free_workspace(BlockCG* w) {
  drop_graphs(w);
  if (w->mem) (void)hipFree(w->mem);
  if (w->host) (void)hipHostFree(w->host);
  if (w->ev_in) (void)hipEventDestroy(w->ev_in);
  if (w->ev_out) (void)hipEventDestroy(w->ev_out);
  if (w->cs) (void)hipStreamDestroy(w->cs);
  delete w;
}

static int ensure_workspace(pmt_mg* mg, int k, hipStream_t s) {
  BlockCG* w = mg->bcg;
  const long long n = mg->L[0].ndof;
  if (w && w->kcap >= k && w->n == n) return PMT_OK;
  const int kc = kround(k);
  if (w) {
    PMT_HIP(hipStreamSynchronize(s));
    mg_forget_stream(mg, w->cs);
    free_workspace(w);
    mg->bcg = nullptr;
  }
  w = new (std::nothrow) BlockCG();
  if (!w) return arg_error("block CG: out of host memory");
  w->n = n;
  w->kcap = kc;
  const long long doubles = 7LL * kc * n + n + Small::SIZE + (long long)kReduceBlocks * kMaxRhs * kMaxRhs;
  hipError_t e = hipMalloc((void**)&w->mem, (size_t)doubles * sizeof(double));
  if (e != hipSuccess) {
    delete w;
    return hip_status(e, "block CG workspace");
  }
  double* p = w->mem;
  w->X = p;
  p += (long long)kc * n;
  w->B = p;
  p += (long long)kc * n;
  w->P = p;
  p += (long long)kc * n;
  w->Q = p;
  p += (long long)kc * n;
  w->R = p;
  p += (long long)kc * n;
  w->Z = p;
  p += (long long)kc * n;
  w->W = p;
  p += (long long)kc * n;
  w->zero = p;
  p += n;
  w->sm = p;
  p += Small::SIZE;
  w->partial = p;
  e = hipHostMalloc((void**)&w->host, 16 * sizeof(double), hipHostMallocDefault);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&w->cs, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&w->ev_in, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&w->ev_out, hipEventDisableTiming);
  if (e != hipSuccess) {
    free_workspace(w);
    return hip_status(e, "block CG stream/events");
  }
  mg->bcg = w;
  // zero column and zeroed blocks (a fill kernel, not a memset: the buffers are captured later)
  return launch_fill(7LL * kc * n + n + Small::SIZE, 0.0, w->mem, s);
}

template <typename F>
static int capture(hipGraphExec_t* exec, hipStream_t s, F&& body) {
  hipGraph_t graph = nullptr;
  PMT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int rc = body();
  const hipError_t e = hipStreamEndCapture(s, &graph);
  if (rc != PMT_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  PMT_HIP(e);
  const hipError_t ei = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  return hip_status(ei, "hipGraphInstantiate");
}

static int read_done(BlockCG* w, hipStream_t s, double* mx) {
  PMT_HIP(hipMemcpyAsync(w->host, w->sm + Small::DONE, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  PMT_HIP(hipStreamSynchronize(s));
  *mx = w->host[1];
  return w->host[0] != 0.0 ? 1 : 0;
}

void block_cg_release(pmt_mg* mg) {
  if (!mg->bcg) return;
  mg_forget_stream(mg, mg->bcg->cs);
  free_workspace(mg->bcg);
  mg->bcg = nullptr;
}

void block_cg_invalidate(pmt_mg* mg) {
  if (mg->bcg) drop_graphs(mg->bcg);
}

}  // namespace pmt

using namespace pmt;

extern "C" {

int pmt_block_pcg_solve(pmt_mg* mg, int k, const double* B, double* X, int x0_is_zero, double tol, int maxit,
                        int restart, int use_graph, int* iters, double* relres, double* history, void* stream) {
  if (!mg || !B || !X || k < 1 || k > kMaxRhs || maxit < 0 || restart < 1)
    return arg_error("pmt_block_pcg_solve: bad argument (1 <= k <= 8 right-hand sides)");
  if (!mg->ready) {
    set_error("pmt_block_pcg_solve: call pmt_mg_update first");
    return PMT_ERR_STATE;
  }
  hipStream_t caller = (hipStream_t)stream;
  PMT_TRY(ensure_workspace(mg, k, caller));
  BlockCG* w = mg->bcg;
  // all work on the handle's stream, after what the caller queued before this call
  hipStream_t s = w->cs;
  PMT_HIP(hipEventRecord(w->ev_in, caller));
  PMT_HIP(hipStreamWaitEvent(s, w->ev_in, 0));
  PMT_TRY(mg_join(mg, s));  // the coarsest inverse of the last update, before anything is captured on s
  const long long n = w->n;
  const size_t bytes = (size_t)k * n * sizeof(double);
  PMT_HIP(hipMemcpyAsync(w->B, B, bytes, hipMemcpyDeviceToDevice, s));
  if (x0_is_zero) {
    PMT_TRY(launch_fill((long long)k * n, 0.0, w->X, s));
    PMT_HIP(hipMemcpyAsync(w->R, B, bytes, hipMemcpyDeviceToDevice, s));
  } else {
    PMT_HIP(hipMemcpyAsync(w->X, X, bytes, hipMemcpyDeviceToDevice, s));
    for (int j = 0; j < k; ++j)
      PMT_TRY(mg_level_apply(mg, 0, MODE_RESIDUAL, col(w->X, n, j), col(w->B, n, j), col(w->R, n, j), s));
  }
  PMT_TRY(gram(w, k, w->B, w->B, 2, s));
  PMT_TRY(gram(w, k, w->R, w->R, 3, s));
  PMT_TRY(single(3, k, 1, tol, w->sm, s));
  double mx = 0.0;
  int done = read_done(w, s, &mx);
  if (done < 0) return PMT_ERR_HIP;
  if (history) history[0] = mx;
  auto non_finite = [&](int at) -> int {
    PMT_HIP(hipEventRecord(w->ev_out, s));
    PMT_HIP(hipStreamWaitEvent(caller, w->ev_out, 0));
    if (iters) *iters = at;
    if (relres)
      for (int j = 0; j < k; ++j) relres[j] = mx;
    set_error("pmt_block_pcg_solve: non-finite residual at iteration " + std::to_string(at) +
              " (non-finite right-hand side, start vector or operator)");
    return PMT_ERR_NUMERIC;
  };
  if (!std::isfinite(mx)) return non_finite(0);
  int it = 0;
  if (!done) {
    // P = orth(Z, normalize=True), Z = M R
    for (int j = 0; j < k; ++j) PMT_TRY(mg_precondition(mg, col(w->R, n, j), col(w->Z, n, j), s));
    PMT_TRY(orth_into_P(w, k, true, w->Z, s));
    const bool graphs = use_graph != 0;
    if (graphs && (w->graph_k != k || w->graph_epoch != tuning_epoch())) {
      drop_graphs(w);
      PMT_TRY(capture(&w->exec[0], s, [&] { return iter_part1(mg, w, k, true, tol, s); }));
      PMT_TRY(capture(&w->exec[1], s, [&] { return iter_part1(mg, w, k, false, tol, s); }));
      PMT_TRY(capture(&w->exec[2], s, [&] { return iter_part2(mg, w, k, tol, s); }));
      w->graph_k = k;
      w->graph_epoch = tuning_epoch();
    }
    for (it = 0; it < maxit; ++it) {
      const bool rs = it % restart == 0;
      if (graphs)
        PMT_HIP(hipGraphLaunch(w->exec[rs ? 0 : 1], s));
      else
        PMT_TRY(iter_part1(mg, w, k, rs, tol, s));
      done = read_done(w, s, &mx);
      if (history) history[it + 1] = mx;
      if (!std::isfinite(mx)) return non_finite(it + 1);
      if (done) break;
      if (graphs)
        PMT_HIP(hipGraphLaunch(w->exec[2], s));
      else
        PMT_TRY(iter_part2(mg, w, k, tol, s));
    }
    if (it < maxit) ++it;
  }
  PMT_HIP(hipMemcpyAsync(X, w->X, bytes, hipMemcpyDeviceToDevice, s));
  PMT_HIP(hipMemcpyAsync(w->host + 2, w->sm + Small::TVAL, (size_t)k * sizeof(double), hipMemcpyDeviceToHost, s));
  PMT_HIP(hipEventRecord(w->ev_out, s));
  PMT_HIP(hipStreamWaitEvent(caller, w->ev_out, 0));
  PMT_HIP(hipStreamSynchronize(s));
  if (relres)
    for (int j = 0; j < k; ++j) relres[j] = w->host[2 + j];
  if (iters) *iters = it;
  return PMT_OK;
}

}  // extern "C"
