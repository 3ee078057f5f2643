// Internal state of a pmt_mg handle (multigrid hierarchy + CG work vectors), shared by the
// single- and multi-right-hand-side CG runtimes.
#pragma once
#include <vector>

#include "launch.h"

namespace pmt {

constexpr int kMaxLevels = 12;
struct BlockCG;

template <typename T>
struct LevelBuf {
  T* dinv = nullptr;  // inverse diagonal (levels above the coarsest)
  T* S = nullptr;     // coarse stencil, full (levels >= 1, FP64 only: setup, diagonal, coarsest)
  T* H = nullptr;     // coarse stencil, symmetric half (levels >= 1 above the coarsest: the operator)
  T* u = nullptr;     // V-cycle correction (levels >= 1)
  T* t = nullptr;     // ping-pong buffer
  T* b = nullptr;     // restricted right-hand side (levels >= 1; level 0 in FP32: converted input)
  T* r = nullptr;     // residual
  T* Xh = nullptr;    // hex8 level 1 without a stencil: transformed child densities [layer][8][element]
};

struct Level {
  GridDims g;
  long long ndof = 0;
  double omega = 0.5;
  int steps = 1;
  double cheb_lo = 0.0, cheb_hi = 0.0;  // Chebyshev interval of D^-1 A (set at update)
  double* Ke = nullptr;  // Galerkin element matrices (Q4 levels >= 1, [E][64]; hex8 stencils come from moments)
  // hex8 levels >= 1: children with Dirichlet descendants per element, list (+ count) and map of the
  // flagged elements, their corrections dK (dk_cap x 576), density moments (27 x nel) and, levels >= 2,
  // the moment tables C (64 x 27 x 9)
  uint8_t* cflag = nullptr;
  int* list = nullptr;
  int* map = nullptr;
  double* dK = nullptr;
  // level 1 without a stencil: the nodes of elements with Dirichlet children (nnodes + 1 ints, count
  // last) and their correction stencils TS[243][t_cap]
  int* tnode = nullptr;
  double* TS = nullptr;
  long long t_cap = 0;
  long long dk_cap = 0;
  double* M = nullptr;
  double* C = nullptr;
  LevelBuf<double> d;
  LevelBuf<float> f;
  template <typename T>
  LevelBuf<T>& buf();
};
template <>
inline LevelBuf<double>& Level::buf<double>() {
  return d;
}
template <>
inline LevelBuf<float>& Level::buf<float>() {
  return f;
}

}  // namespace pmt

struct pmt_mg {
  int dim = 3;
  int nlev = 0;
  int prec = 64;
  int smoother = 0;            // 0 damped Jacobi, 1 Chebyshev
  double cheb_lower = 0.25;    // lo = cheb_lower * hi
  double cheb_upper = 1.1;     // hi = cheb_upper * lambda_max estimate
  double lmax_elem = 0.0;      // lambda_max(diag(K0)^-1 K0): bounds lambda_max(D^-1 K) of the fine grid
  pmt::Level L[pmt::kMaxLevels];
  int NE = 24;
  double* K0 = nullptr;
  double* G = nullptr;  // P_c^T K0 P_c per child position (level-1 stencil of hex8 grids)
  bool l1free = false;     // hex8 level 1 applied element by element in the reflection basis (no stencil)
  double* G0ext = nullptr;  // its 24 x 24 table (coarse1_table), FP64 then FP32
  int g0pat = 0;            // zero pattern of the table (coarse1_table_pattern)
  int* host_counts = nullptr;  // pinned: [0] touched nodes of level 1 (without a stencil), [l] flagged elements
                               // of level l (read once per update)
  int k0pat = 0;
  const double* xe = nullptr;
  float* xe32 = nullptr;
  const uint8_t* bcm = nullptr;
  double bcdiag = 0.0;
  bool ready = false;
  long long nc = 0;
  double* Ainv = nullptr;
  // hex8 update: the coarsest inverse (a chain of small launches) runs on a second stream of the
  // handle, forked after the coarsest stencil and joined before update returns
  hipStream_t s2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // the inverse is waited for by every stream that runs a coarsest solve (or mg_join) after an update: the
  // streams that have waited since the last update are listed, so each waits once and none is skipped
  bool pending_join = false;
  std::vector<hipStream_t> joined;
  double* gjP = nullptr;
  double* gjC = nullptr;
  float* z32 = nullptr;  // FP32 V-cycle output on the fine grid
  double* p = nullptr;
  double* q = nullptr;
  double* z = nullptr;
  double* r = nullptr;
  double* sc = nullptr;    // device scalars
  double* work = nullptr;  // reduction partials
  double* rwork = nullptr;  // per-block partials of the CG r update (cg_xr_parts(ndof) entries)
  double* dotws = nullptr;  // partial sums of the fused fine K.u + dot products (2 x dotws_parts)
  long long dotws_parts = 0;
  double* host_sc = nullptr;
  double* host_rr = nullptr;  // pinned [value, sequence number], written by publish_kernel: the CG stopping
                              // test polls the sequence number
  double pub_seq = 0.0;       // sequence number of the last publish
  long long bytes = 0;
  std::vector<void*> allocs;
  // multi-right-hand-side CG (blockcg.hip): workspace for kcap columns and captured iterations
  pmt::BlockCG* bcg = nullptr;
  // deflation of the single-right-hand-side PCG's preconditioned residual by a coarse space,
  // z -= W Einv (AW)^T z (pmt_mg_set_coarse_space): cs_m caller-owned columns W and AW, Einv kept on the
  // host (kernel argument)
  int cs_m = 0;
  const double* cs_W = nullptr;
  const double* cs_AW = nullptr;
  pmt::SmallMat cs_Einv{};
  double* cs_buf = nullptr;  // device: [c (4)][dot partials (4 x kReduceBlocks)]
};

namespace pmt {

// frees the block-CG workspace and graphs of a handle (blockcg.hip)
void block_cg_release(pmt_mg* mg);
// drops captured block-CG iterations (after an update or a change of precision or launch shapes)
void block_cg_invalidate(pmt_mg* mg);

// y = A_l u (mode apply), b - A_l u (residual) or a damped-Jacobi sweep on level l, FP64
int mg_level_apply(pmt_mg* mg, int l, int mode, const double* u, const double* b, double* y, hipStream_t s);
// z = M r: one V-cycle in the handle's precision, FP64 in and out
int mg_precondition(pmt_mg* mg, const double* r, double* z, hipStream_t s);
// device allocation owned (and freed) by the handle
int mg_alloc(pmt_mg* mg, void** p, long long bytes);
// make stream s wait for the coarsest inverse of the last update (no-op once joined); callers that
// capture graphs call it before capturing
int mg_join(pmt_mg* mg, hipStream_t s);
// drop a stream that is about to be destroyed from the list of joined streams
void mg_forget_stream(pmt_mg* mg, hipStream_t s);

}  // namespace pmt
