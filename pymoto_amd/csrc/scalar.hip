// Scalar (one unknown per node) finite-element operators for AssemblePoisson (thermal conduction,
// electric permittivity; pymoto/modules/assembly.py:523-560) and a geometric-multigrid PCG for them.
//
// Operator: K(x) = sum_e x_e P0 with P0 the (2^D x 2^D) element matrix, Dirichlet rows and columns
// zeroed and bcdiag on their diagonal (AssembleGeneral, assembly.py:209-230,245-296), applied matrix
// free and node-centric: thread n gathers the 2^D elements around node n and the 3^D neighbour values
// (no scatter, no atomics, fixed summation order).
//
// Multigrid: every level stores its operator as a node stencil S[c][node] (3^D coefficients, c =
// (dz+1)*9 + (dy+1)*3 + dx+1, coefficient-major so a wave reads consecutive nodes of one coefficient);
// level 0 from the elements, level l+1 as the Galerkin product R^T S_l R with the trilinear (bilinear)
// transfer of GeometricMultigrid.setup_interpolation (pymoto/solvers/iterative.py:178-220, weights
// 1, 1/2, 1/4, 1/8), which keeps the reference's Dirichlet diagonal inside the coarse operators
// (iterative.py:173).  V-cycle: damped-Jacobi pre-smoothing whose first sweep starts from zero, residual,
// restriction, inner level, prolongation, post-smoothing (iterative.py:222-256); the coarsest grid is
// solved with an explicit FP64 inverse (blocked Gauss-Jordan of coarse.hip).  The CG loop follows
// CG.solve for one right-hand side (iterative.py:340-403).
#include <cmath>
#include <vector>

#include "launch.h"

namespace pmt {

int gj_invert(long long n, double* A, double* wsP, double* wsC, hipStream_t s);
template <typename T>
int launch_gemv(long long n, const T* A, const T* x, T* y, hipStream_t s);
int launch_dot(long long n, const double* x, const double* y, double* out, double* work, hipStream_t s);
int launch_axpby(long long n, double a, const double* x, double b, double* y, hipStream_t s);

namespace {

constexpr int kScalarMaxLevels = 12;

template <int DIM>
struct SCfg {
  static constexpr int NN = 1 << DIM;           // nodes per element
  static constexpr int NS = DIM == 3 ? 27 : 9;  // stencil coefficients per node
};

__device__ inline void node_ijk(const GridDims& g, long long n, int& i, int& j, int& k) {
  i = (int)(n % g.nx);
  const long long t = n / g.nx;
  j = (int)(t % g.ny);
  k = (int)(t / g.ny);
}

// ---------------------------------------------------------------------------------------------
// fine operator from the elements
// MODE 0: y = K u ; MODE 1: y = b - K u
template <int DIM, int MODE>
__global__ void __launch_bounds__(256)
scalar_apply_kernel(GridDims g, const double* __restrict__ P0, const double* __restrict__ xe,
                    const uint8_t* __restrict__ bcm, double bcdiag, const double* __restrict__ u,
                    const double* __restrict__ b, double* __restrict__ y) {
  constexpr int NN = SCfg<DIM>::NN;
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nnodes) return;
  int i, j, k;
  node_ijk(g, n, i, j, k);
  double acc = 0.0;
  if (bcm && (bcm[n] & 1u)) {
    acc = bcdiag * u[n];
  } else {
    // elements (ei, ej, ek) = (i - 1 + ox, ...), local index of node n inside it a = (1-ox) + 2(1-oy) + 4(1-oz)
#pragma unroll
    for (int oz = 0; oz < (DIM == 3 ? 2 : 1); ++oz)
#pragma unroll
      for (int oy = 0; oy < 2; ++oy)
#pragma unroll
        for (int ox = 0; ox < 2; ++ox) {
          const int ei = i - 1 + ox, ej = j - 1 + oy, ek = DIM == 3 ? k - 1 + oz : 0;
          if (ei < 0 || ei >= g.ex || ej < 0 || ej >= g.ey || ek < 0 || ek >= g.ez) continue;
          const int a = (1 - ox) + 2 * (1 - oy) + (DIM == 3 ? 4 * (1 - oz) : 0);
          const double xv = xe[((long long)ek * g.ey + ej) * g.ex + ei];
          double s = 0.0;
#pragma unroll
          for (int c = 0; c < NN; ++c) {
            const long long m = ((long long)(ek + ((c >> 2) & 1)) * g.ny + (ej + ((c >> 1) & 1))) * g.nx + (ei + (c & 1));
            const double um = (bcm && (bcm[m] & 1u)) ? 0.0 : u[m];
            s = fma(P0[a * NN + c], um, s);
          }
          acc = fma(xv, s, acc);
        }
  }
  y[n] = MODE == 0 ? acc : b[n] - acc;
}

// diag(K)
template <int DIM>
__global__ void scalar_diag_kernel(GridDims g, const double* __restrict__ P0, const double* __restrict__ xe,
                                   const uint8_t* __restrict__ bcm, double bcdiag, double* __restrict__ d) {
  constexpr int NN = SCfg<DIM>::NN;
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nnodes) return;
  int i, j, k;
  node_ijk(g, n, i, j, k);
  if (bcm && (bcm[n] & 1u)) {
    d[n] = bcdiag;
    return;
  }
  double acc = 0.0;
  for (int oz = 0; oz < (DIM == 3 ? 2 : 1); ++oz)
    for (int oy = 0; oy < 2; ++oy)
      for (int ox = 0; ox < 2; ++ox) {
        const int ei = i - 1 + ox, ej = j - 1 + oy, ek = DIM == 3 ? k - 1 + oz : 0;
        if (ei < 0 || ei >= g.ex || ej < 0 || ej >= g.ey || ek < 0 || ek >= g.ez) continue;
        const int a = (1 - ox) + 2 * (1 - oy) + (DIM == 3 ? 4 * (1 - oz) : 0);
        acc = fma(xe[((long long)ek * g.ey + ej) * g.ex + ei], P0[a * NN + a], acc);
      }
  d[n] = acc;
}

// dx_e (+)= alpha * u_e^T P0 v_e, Dirichlet entries of u and v zeroed (assembly.py:298-315)
template <int DIM>
__global__ void __launch_bounds__(256)
scalar_sens_kernel(GridDims g, const double* __restrict__ P0, const uint8_t* __restrict__ bcm,
                   const double* __restrict__ u, const double* __restrict__ v, double alpha,
                   double* __restrict__ dx, int accumulate) {
  constexpr int NN = SCfg<DIM>::NN;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.nel) return;
  const int ei = (int)(e % g.ex);
  const long long t = e / g.ex;
  const int ej = (int)(t % g.ey);
  const int ek = (int)(t / g.ey);
  double ue[NN], ve[NN];
#pragma unroll
  for (int a = 0; a < NN; ++a) {
    const long long m = ((long long)(ek + ((a >> 2) & 1)) * g.ny + (ej + ((a >> 1) & 1))) * g.nx + (ei + (a & 1));
    const bool fixed = bcm && (bcm[m] & 1u);
    ue[a] = fixed ? 0.0 : u[m];
    ve[a] = fixed ? 0.0 : v[m];
  }
  double s = 0.0;
#pragma unroll
  for (int a = 0; a < NN; ++a) {
    double r = 0.0;
#pragma unroll
    for (int c = 0; c < NN; ++c) r = fma(P0[a * NN + c], ve[c], r);
    s = fma(ue[a], r, s);
  }
  dx[e] = accumulate ? fma(alpha, s, dx[e]) : alpha * s;
}

// ---------------------------------------------------------------------------------------------
// stencils

// level-0 stencil from the elements: S[c][n] = sum over elements e around n of x_e P0[a(n,e)][b] for the
// neighbour m = corner b of e at offset c; Dirichlet rows: bcdiag on the centre only; Dirichlet columns 0
template <int DIM>
__global__ void scalar_stencil_fine_kernel(GridDims g, const double* __restrict__ P0, const double* __restrict__ xe,
                                           const uint8_t* __restrict__ bcm, double bcdiag, double* __restrict__ S) {
  constexpr int NN = SCfg<DIM>::NN, NS = SCfg<DIM>::NS;
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nnodes) return;
  int i, j, k;
  node_ijk(g, n, i, j, k);
  double st[NS];
#pragma unroll
  for (int c = 0; c < NS; ++c) st[c] = 0.0;
  if (bcm && (bcm[n] & 1u)) {
    st[NS / 2] = bcdiag;
  } else {
#pragma unroll
    for (int oz = 0; oz < (DIM == 3 ? 2 : 1); ++oz)
#pragma unroll
      for (int oy = 0; oy < 2; ++oy)
#pragma unroll
        for (int ox = 0; ox < 2; ++ox) {
          const int ei = i - 1 + ox, ej = j - 1 + oy, ek = DIM == 3 ? k - 1 + oz : 0;
          if (ei < 0 || ei >= g.ex || ej < 0 || ej >= g.ey || ek < 0 || ek >= g.ez) continue;
          const int a = (1 - ox) + 2 * (1 - oy) + (DIM == 3 ? 4 * (1 - oz) : 0);
          const double xv = xe[((long long)ek * g.ey + ej) * g.ex + ei];
#pragma unroll
          for (int b = 0; b < NN; ++b) {
            const int bx = b & 1, by = (b >> 1) & 1, bz = DIM == 3 ? (b >> 2) & 1 : 0;
            const long long m = ((long long)(ek + bz) * g.ny + (ej + by)) * g.nx + (ei + bx);
            if (bcm && (bcm[m] & 1u)) continue;
            // offset of m from n per axis: (e + corner) - node = corner - (1 - o)
            const int dx = bx - (1 - ox), dy = by - (1 - oy), dz = DIM == 3 ? bz - (1 - oz) : 0;
            const int c = DIM == 3 ? (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1) : (dy + 1) * 3 + (dx + 1);
            st[c] = fma(xv, P0[a * NN + b], st[c]);
          }
        }
  }
#pragma unroll
  for (int c = 0; c < NS; ++c) S[(long long)c * g.nnodes + n] = st[c];
}

// Galerkin coarse stencil Sc = R^T Sf R, one thread per coarse node C.  With F = 2C the fine image of C,
// R[i][C] = prod_d w(i_d - F_d) (w(0) = 1, w(+-1) = 1/2) and Sc[c'][C] = sum_{i, j} R[i][C] Sf[j - i][i]
// R[j][C + c'].  Per axis the pairs (i_d - F_d, j_d - 2(C + c')_d) with |j - i| <= 1 are fixed: 7 pairs
// for c'_d = 0, 3 for c'_d = +-1; a fine stencil entry pointing outside the grid is 0 by construction,
// and so is a coarse entry pointing outside the coarse grid.
struct AxisPairs {
  // for c' = -1, 0, +1: count and (di, dj, w) of each pair
  int cnt[3];
  int di[3][7], dj[3][7];
  double w[3][7];
};

__host__ AxisPairs make_axis_pairs() {
  AxisPairs p{};
  for (int cp = -1; cp <= 1; ++cp) {
    int n = 0;
    for (int di = -1; di <= 1; ++di)
      for (int dj = -1; dj <= 1; ++dj) {
        const int off = 2 * cp + dj - di;  // (j - i) along the axis
        if (off < -1 || off > 1) continue;
        p.di[cp + 1][n] = di;
        p.dj[cp + 1][n] = dj;
        p.w[cp + 1][n] = (di ? 0.5 : 1.0) * (dj ? 0.5 : 1.0);
        ++n;
      }
    p.cnt[cp + 1] = n;
  }
  return p;
}

template <int DIM>
__global__ void scalar_galerkin_kernel(GridDims gf, GridDims gc, AxisPairs ap, const double* __restrict__ Sf,
                                       double* __restrict__ Sc) {
  constexpr int NS = SCfg<DIM>::NS;
  const long long N = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (N >= gc.nnodes) return;
  int I, J, K;
  node_ijk(gc, N, I, J, K);
  for (int c = 0; c < NS; ++c) {
    const int cx = c % 3 - 1, cy = (c / 3) % 3 - 1, cz = DIM == 3 ? c / 9 - 1 : 0;
    double acc = 0.0;
    const bool inside = I + cx >= 0 && I + cx < gc.nx && J + cy >= 0 && J + cy < gc.ny && K + cz >= 0 &&
                        K + cz < gc.nz;
    if (inside) {
      const int nzp = DIM == 3 ? ap.cnt[cz + 1] : 1;
      for (int pz = 0; pz < nzp; ++pz) {
        const int diz = DIM == 3 ? ap.di[cz + 1][pz] : 0, djz = DIM == 3 ? ap.dj[cz + 1][pz] : 0;
        const double wz = DIM == 3 ? ap.w[cz + 1][pz] : 1.0;
        const int fk = 2 * K + diz;
        if (fk < 0 || fk >= gf.nz) continue;
        const int oz = 2 * cz + djz - diz;
        for (int py = 0; py < ap.cnt[cy + 1]; ++py) {
          const int diy = ap.di[cy + 1][py], djy = ap.dj[cy + 1][py];
          const double wy = ap.w[cy + 1][py] * wz;
          const int fj = 2 * J + diy;
          if (fj < 0 || fj >= gf.ny) continue;
          const int oy = 2 * cy + djy - diy;
          for (int px = 0; px < ap.cnt[cx + 1]; ++px) {
            const int dix = ap.di[cx + 1][px], djx = ap.dj[cx + 1][px];
            const int fi = 2 * I + dix;
            if (fi < 0 || fi >= gf.nx) continue;
            const int ox = 2 * cx + djx - dix;
            const int cf = DIM == 3 ? (oz + 1) * 9 + (oy + 1) * 3 + (ox + 1) : (oy + 1) * 3 + (ox + 1);
            const long long fnode = ((long long)fk * gf.ny + fj) * gf.nx + fi;
            acc = fma(ap.w[cx + 1][px] * wy, Sf[(long long)cf * gf.nnodes + fnode], acc);
          }
        }
      }
    }
    Sc[(long long)c * gc.nnodes + N] = acc;
  }
}

// stencil operator on a level. MODE 0: y = S u; 1: y = b - S u; 2: damped Jacobi y = u + w D^-1 (b - S u)
template <int DIM, int MODE>
__global__ void __launch_bounds__(256)
scalar_stencil_kernel(GridDims g, const double* __restrict__ S, const double* __restrict__ u,
                      const double* __restrict__ b, double omega, double* __restrict__ y) {
  constexpr int NS = SCfg<DIM>::NS;
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nnodes) return;
  int i, j, k;
  node_ijk(g, n, i, j, k);
  double acc = 0.0;
#pragma unroll
  for (int c = 0; c < NS; ++c) {
    const int dx = c % 3 - 1, dy = (c / 3) % 3 - 1, dz = DIM == 3 ? c / 9 - 1 : 0;
    const int ii = i + dx, jj = j + dy, kk = k + dz;
    if (ii < 0 || ii >= g.nx || jj < 0 || jj >= g.ny || kk < 0 || kk >= g.nz) continue;
    acc = fma(S[(long long)c * g.nnodes + n], u[((long long)kk * g.ny + jj) * g.nx + ii], acc);
  }
  if (MODE == 0) {
    y[n] = acc;
  } else if (MODE == 1) {
    y[n] = b[n] - acc;
  } else {
    y[n] = u[n] + omega * (b[n] - acc) / S[(long long)(NS / 2) * g.nnodes + n];
  }
}

// first sweep from zero: u = w b / diag
template <int DIM>
__global__ void scalar_jacobi0_kernel(GridDims g, const double* __restrict__ S, const double* __restrict__ b,
                                      double omega, double* __restrict__ u) {
  constexpr int NS = SCfg<DIM>::NS;
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < g.nnodes) u[n] = omega * b[n] / S[(long long)(NS / 2) * g.nnodes + n];
}

template <int DIM>
__global__ void scalar_restrict_kernel(GridDims gf, GridDims gc, const double* __restrict__ rf,
                                       double* __restrict__ rc) {
  const long long N = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (N >= gc.nnodes) return;
  int I, J, K;
  node_ijk(gc, N, I, J, K);
  double acc = 0.0;
  for (int dz = (DIM == 3 ? -1 : 0); dz <= (DIM == 3 ? 1 : 0); ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int fi = 2 * I + dx, fj = 2 * J + dy, fk = 2 * K + dz;
        if (fi < 0 || fi >= gf.nx || fj < 0 || fj >= gf.ny || fk < 0 || fk >= gf.nz) continue;
        const double w = (dx ? 0.5 : 1.0) * (dy ? 0.5 : 1.0) * (dz ? 0.5 : 1.0);
        acc = fma(w, rf[((long long)fk * gf.ny + fj) * gf.nx + fi], acc);
      }
  rc[N] = acc;
}

template <int DIM>
__global__ void scalar_prolong_kernel(GridDims gf, GridDims gc, const double* __restrict__ uc,
                                      double* __restrict__ uf) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= gf.nnodes) return;
  int i, j, k;
  node_ijk(gf, n, i, j, k);
  double acc = 0.0;
  for (int cz = 0; cz <= (DIM == 3 ? (k & 1) : 0); ++cz)
    for (int cy = 0; cy <= (j & 1); ++cy)
      for (int cx = 0; cx <= (i & 1); ++cx) {
        const double w = ((i & 1) ? 0.5 : 1.0) * ((j & 1) ? 0.5 : 1.0) * ((DIM == 3 && (k & 1)) ? 0.5 : 1.0);
        const long long N = ((long long)((k >> 1) + cz) * gc.ny + ((j >> 1) + cy)) * gc.nx + ((i >> 1) + cx);
        acc = fma(w, uc[N], acc);
      }
  uf[n] += acc;
}

// dense coarsest matrix from its stencil
template <int DIM>
__global__ void scalar_dense_kernel(GridDims g, const double* __restrict__ S, double* __restrict__ A) {
  constexpr int NS = SCfg<DIM>::NS;
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nnodes) return;
  int i, j, k;
  node_ijk(g, n, i, j, k);
  for (int c = 0; c < NS; ++c) {
    const int dx = c % 3 - 1, dy = (c / 3) % 3 - 1, dz = DIM == 3 ? c / 9 - 1 : 0;
    const int ii = i + dx, jj = j + dy, kk = k + dz;
    if (ii < 0 || ii >= g.nx || jj < 0 || jj >= g.ny || kk < 0 || kk >= g.nz) continue;
    A[n * g.nnodes + ((long long)kk * g.ny + jj) * g.nx + ii] = S[(long long)c * g.nnodes + n];
  }
}

// x += a p ; r -= a q (a = sc[0] / sc[1], device scalars) and rr = r.r partials
__global__ void scalar_fill_kernel(long long n, double v, double* __restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = v;
}

template <int DIM>
int stencil_launch(int mode, const GridDims& g, const double* S, const double* u, const double* b, double omega,
                   double* y, hipStream_t s) {
  const unsigned blocks = blocks_for(g.nnodes, 256);
  if (mode == 0)
    scalar_stencil_kernel<DIM, 0><<<blocks, 256, 0, s>>>(g, S, u, b, omega, y);
  else if (mode == 1)
    scalar_stencil_kernel<DIM, 1><<<blocks, 256, 0, s>>>(g, S, u, b, omega, y);
  else
    scalar_stencil_kernel<DIM, 2><<<blocks, 256, 0, s>>>(g, S, u, b, omega, y);
  return check_launch("scalar_stencil_kernel");
}

int launch_scalar_fine(int mode, const GridDims& g, const double* P0, const double* xe, const uint8_t* bcm,
                       double bcdiag, const double* u, const double* b, double* y, hipStream_t s) {
  const unsigned blocks = blocks_for(g.nnodes, 256);
  if (g.dim == 3) {
    if (mode == 0)
      scalar_apply_kernel<3, 0><<<blocks, 256, 0, s>>>(g, P0, xe, bcm, bcdiag, u, b, y);
    else
      scalar_apply_kernel<3, 1><<<blocks, 256, 0, s>>>(g, P0, xe, bcm, bcdiag, u, b, y);
  } else {
    if (mode == 0)
      scalar_apply_kernel<2, 0><<<blocks, 256, 0, s>>>(g, P0, xe, bcm, bcdiag, u, b, y);
    else
      scalar_apply_kernel<2, 1><<<blocks, 256, 0, s>>>(g, P0, xe, bcm, bcdiag, u, b, y);
  }
  return check_launch("scalar_apply_kernel");
}

}  // namespace
}  // namespace pmt

using namespace pmt;

// ---------------------------------------------------------------------------------------------
// scalar multigrid handle
struct pmt_smg {
  int dim = 3;
  int nlev = 0;
  GridDims g[kScalarMaxLevels];
  double omega[kScalarMaxLevels] = {};
  int steps[kScalarMaxLevels] = {};
  double* S[kScalarMaxLevels] = {};  // stencils (levels 0..nlev-1)
  double* u[kScalarMaxLevels] = {};  // corrections (levels >= 1)
  double* b[kScalarMaxLevels] = {};  // right-hand sides (levels >= 1)
  double* t[kScalarMaxLevels] = {};  // work
  double* P0 = nullptr;              // element matrix (device)
  double* Ainv = nullptr;            // coarsest inverse
  double* gjP = nullptr;
  double* gjC = nullptr;
  const double* xe = nullptr;
  const uint8_t* bcm = nullptr;
  double bcdiag = 0.0;
  bool ready = false;
  // CG vectors and scalars
  double *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *sc = nullptr, *work = nullptr;
  long long bytes = 0;
  std::vector<void*> allocs;
};

namespace {

int smg_alloc(pmt_smg* h, double** p, long long count) {
  void* q = nullptr;
  PMT_HIP(hipMalloc(&q, (size_t)(count > 0 ? count : 1) * sizeof(double)));
  h->allocs.push_back(q);
  h->bytes += count * (long long)sizeof(double);
  *p = static_cast<double*>(q);
  return PMT_OK;
}

int smg_stencil(pmt_smg* h, int l, int mode, const double* u, const double* b, double omega, double* y,
                hipStream_t s) {
  return h->dim == 3 ? stencil_launch<3>(mode, h->g[l], h->S[l], u, b, omega, y, s)
                     : stencil_launch<2>(mode, h->g[l], h->S[l], u, b, omega, y, s);
}

// z = V-cycle(r) on level l (output written to z; rhs r; both length nnodes of level l)
int smg_vcycle(pmt_smg* h, int l, const double* r, double* z, hipStream_t s) {
  const GridDims& g = h->g[l];
  if (l == h->nlev - 1) return launch_gemv<double>(g.nnodes, h->Ainv, r, z, s);
  const unsigned blocks = blocks_for(g.nnodes, 256);
  // pre-smoothing: first sweep from zero (iterative.py:232-233), then steps-1 sweeps
  if (h->dim == 3)
    scalar_jacobi0_kernel<3><<<blocks, 256, 0, s>>>(g, h->S[l], r, h->omega[l], z);
  else
    scalar_jacobi0_kernel<2><<<blocks, 256, 0, s>>>(g, h->S[l], r, h->omega[l], z);
  PMT_TRY(check_launch("scalar_jacobi0_kernel"));
  double* t = h->t[l];
  for (int k = 1; k < h->steps[l]; ++k) {
    PMT_TRY(smg_stencil(h, l, 2, z, r, h->omega[l], t, s));
    PMT_HIP(hipMemcpyAsync(z, t, g.nnodes * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  // residual, restriction, inner level, prolongation
  PMT_TRY(smg_stencil(h, l, 1, z, r, 0.0, t, s));
  const GridDims& gc = h->g[l + 1];
  const unsigned cblocks = blocks_for(gc.nnodes, 256);
  if (h->dim == 3)
    scalar_restrict_kernel<3><<<cblocks, 256, 0, s>>>(g, gc, t, h->b[l + 1]);
  else
    scalar_restrict_kernel<2><<<cblocks, 256, 0, s>>>(g, gc, t, h->b[l + 1]);
  PMT_TRY(check_launch("scalar_restrict_kernel"));
  PMT_TRY(smg_vcycle(h, l + 1, h->b[l + 1], h->u[l + 1], s));
  if (h->dim == 3)
    scalar_prolong_kernel<3><<<blocks, 256, 0, s>>>(g, gc, h->u[l + 1], z);
  else
    scalar_prolong_kernel<2><<<blocks, 256, 0, s>>>(g, gc, h->u[l + 1], z);
  PMT_TRY(check_launch("scalar_prolong_kernel"));
  // post-smoothing
  for (int k = 0; k < h->steps[l]; ++k) {
    PMT_TRY(smg_stencil(h, l, 2, z, r, h->omega[l], t, s));
    PMT_HIP(hipMemcpyAsync(z, t, g.nnodes * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  return PMT_OK;
}

int read_scalars(pmt_smg* h, double* out, int count, hipStream_t s) {
  PMT_HIP(hipMemcpyAsync(out, h->sc, count * sizeof(double), hipMemcpyDeviceToHost, s));
  PMT_HIP(hipStreamSynchronize(s));
  return PMT_OK;
}

}  // namespace

extern "C" {

int pmt_scalar_apply(int nelx, int nely, int nelz, const double* P0, const double* xe, const uint8_t* bcmask,
                     double bcdiag, const double* u, double* y, void* stream) {
  if (!grid_ok(nelx, nely, nelz) || !P0 || !xe || !u || !y) return arg_error("pmt_scalar_apply: bad argument");
  return launch_scalar_fine(0, make_grid(nelx, nely, nelz), P0, xe, bcmask, bcdiag, u, nullptr, y,
                            (hipStream_t)stream);
}

int pmt_scalar_residual(int nelx, int nely, int nelz, const double* P0, const double* xe, const uint8_t* bcmask,
                        double bcdiag, const double* b, const double* u, double* r, void* stream) {
  if (!grid_ok(nelx, nely, nelz) || !P0 || !xe || !b || !u || !r) return arg_error("pmt_scalar_residual: bad argument");
  return launch_scalar_fine(1, make_grid(nelx, nely, nelz), P0, xe, bcmask, bcdiag, u, b, r, (hipStream_t)stream);
}

int pmt_scalar_diagonal(int nelx, int nely, int nelz, const double* P0, const double* xe, const uint8_t* bcmask,
                        double bcdiag, double* diag, void* stream) {
  if (!grid_ok(nelx, nely, nelz) || !P0 || !xe || !diag) return arg_error("pmt_scalar_diagonal: bad argument");
  GridDims g = make_grid(nelx, nely, nelz);
  hipStream_t s = (hipStream_t)stream;
  if (g.dim == 3)
    scalar_diag_kernel<3><<<blocks_for(g.nnodes, 256), 256, 0, s>>>(g, P0, xe, bcmask, bcdiag, diag);
  else
    scalar_diag_kernel<2><<<blocks_for(g.nnodes, 256), 256, 0, s>>>(g, P0, xe, bcmask, bcdiag, diag);
  return check_launch("scalar_diag_kernel");
}

int pmt_scalar_sens_contract(int nelx, int nely, int nelz, const double* P0, const uint8_t* bcmask, const double* u,
                             const double* v, double alpha, double* dx, int accumulate, void* stream) {
  if (!grid_ok(nelx, nely, nelz) || !P0 || !u || !v || !dx) return arg_error("pmt_scalar_sens_contract: bad argument");
  GridDims g = make_grid(nelx, nely, nelz);
  hipStream_t s = (hipStream_t)stream;
  if (g.dim == 3)
    scalar_sens_kernel<3><<<blocks_for(g.nel, 256), 256, 0, s>>>(g, P0, bcmask, u, v, alpha, dx, accumulate);
  else
    scalar_sens_kernel<2><<<blocks_for(g.nel, 256), 256, 0, s>>>(g, P0, bcmask, u, v, alpha, dx, accumulate);
  return check_launch("scalar_sens_kernel");
}

int pmt_smg_create(int nelx, int nely, int nelz, int nlevels, const double* omega, const int* steps,
                   const double* P0_host, pmt_smg** out) {
  if (!out || !P0_host || !grid_ok(nelx, nely, nelz) || nlevels < 2 || nlevels > kScalarMaxLevels || !omega ||
      !steps)
    return arg_error("pmt_smg_create: bad argument");
  *out = nullptr;
  pmt_smg* h = new pmt_smg();
  h->dim = nelz > 0 ? 3 : 2;
  h->nlev = nlevels;
  int ex = nelx, ey = nely, ez = nelz;
  for (int l = 0; l < nlevels; ++l) {
    if (l > 0) {
      if (ex % 2 || ey % 2 || (h->dim == 3 && ez % 2)) {
        delete h;
        return arg_error("pmt_smg_create: every level above the coarsest needs even element counts");
      }
      ex /= 2;
      ey /= 2;
      ez = h->dim == 3 ? ez / 2 : 0;
    }
    h->g[l] = make_grid(ex, ey, ez);
    if (l < nlevels - 1) {
      h->omega[l] = omega[l];
      h->steps[l] = steps[l] < 1 ? 1 : steps[l];
    }
  }
  const int NN = 1 << h->dim, NS = h->dim == 3 ? 27 : 9;
  const long long nc = h->g[nlevels - 1].nnodes;
  int rc = PMT_OK;
  auto A = [&](double** p, long long n) {
    if (rc == PMT_OK) rc = smg_alloc(h, p, n);
  };
  A(&h->P0, NN * NN);
  for (int l = 0; l < nlevels; ++l) {
    A(&h->S[l], NS * h->g[l].nnodes);
    A(&h->t[l], h->g[l].nnodes);
    if (l > 0) {
      A(&h->u[l], h->g[l].nnodes);
      A(&h->b[l], h->g[l].nnodes);
    }
  }
  A(&h->Ainv, nc * nc);
  A(&h->gjP, 64 * nc + 64 * 64);
  A(&h->gjC, 64 * nc);
  const long long n0 = h->g[0].nnodes;
  A(&h->r, n0);
  A(&h->z, n0);
  A(&h->p, n0);
  A(&h->q, n0);
  A(&h->sc, 8);
  A(&h->work, 2 * kReduceBlocks);
  if (rc == PMT_OK) rc = hip_status(hipMemcpy(h->P0, P0_host, NN * NN * sizeof(double), hipMemcpyHostToDevice),
                                    "pmt_smg_create: copy P0");
  if (rc != PMT_OK) {
    pmt_smg_destroy(h);
    return rc;
  }
  *out = h;
  return PMT_OK;
}

int pmt_smg_destroy(pmt_smg* h) {
  if (!h) return PMT_OK;
  (void)hipDeviceSynchronize();
  for (void* p : h->allocs) (void)hipFree(p);
  delete h;
  return PMT_OK;
}

long long pmt_smg_level_nodes(pmt_smg* h, int level) {
  if (!h || level < 0 || level >= h->nlev) return arg_error("pmt_smg_level_nodes: bad argument");
  return h->g[level].nnodes;
}

int pmt_smg_update(pmt_smg* h, const double* xe, const uint8_t* bcmask, double bcdiag, void* stream) {
  if (!h || !xe) return arg_error("pmt_smg_update: bad argument");
  hipStream_t s = (hipStream_t)stream;
  h->xe = xe;
  h->bcm = bcmask;
  h->bcdiag = bcdiag;
  const GridDims& g0 = h->g[0];
  if (h->dim == 3)
    scalar_stencil_fine_kernel<3><<<blocks_for(g0.nnodes, 128), 128, 0, s>>>(g0, h->P0, xe, bcmask, bcdiag, h->S[0]);
  else
    scalar_stencil_fine_kernel<2><<<blocks_for(g0.nnodes, 128), 128, 0, s>>>(g0, h->P0, xe, bcmask, bcdiag, h->S[0]);
  PMT_TRY(check_launch("scalar_stencil_fine_kernel"));
  const AxisPairs ap = make_axis_pairs();
  for (int l = 0; l + 1 < h->nlev; ++l) {
    const GridDims& gf = h->g[l];
    const GridDims& gc = h->g[l + 1];
    if (h->dim == 3)
      scalar_galerkin_kernel<3><<<blocks_for(gc.nnodes, 64), 64, 0, s>>>(gf, gc, ap, h->S[l], h->S[l + 1]);
    else
      scalar_galerkin_kernel<2><<<blocks_for(gc.nnodes, 64), 64, 0, s>>>(gf, gc, ap, h->S[l], h->S[l + 1]);
    PMT_TRY(check_launch("scalar_galerkin_kernel"));
  }
  const int L = h->nlev - 1;
  const long long nc = h->g[L].nnodes;
  scalar_fill_kernel<<<blocks_for(nc * nc, 256), 256, 0, s>>>(nc * nc, 0.0, h->Ainv);
  PMT_TRY(check_launch("scalar_fill_kernel"));
  if (h->dim == 3)
    scalar_dense_kernel<3><<<blocks_for(nc, 128), 128, 0, s>>>(h->g[L], h->S[L], h->Ainv);
  else
    scalar_dense_kernel<2><<<blocks_for(nc, 128), 128, 0, s>>>(h->g[L], h->S[L], h->Ainv);
  PMT_TRY(check_launch("scalar_dense_kernel"));
  PMT_TRY(gj_invert(nc, h->Ainv, h->gjP, h->gjC, s));
  h->ready = true;
  return PMT_OK;
}

int pmt_smg_vcycle(pmt_smg* h, const double* r, double* z, void* stream) {
  if (!h || !r || !z) return arg_error("pmt_smg_vcycle: bad argument");
  if (!h->ready) {
    set_error("pmt_smg_vcycle: call pmt_smg_update first");
    return PMT_ERR_STATE;
  }
  return smg_vcycle(h, 0, r, z, (hipStream_t)stream);
}

int pmt_smg_level_apply(pmt_smg* h, int level, const double* u, double* y, void* stream) {
  if (!h || level < 0 || level >= h->nlev || !u || !y) return arg_error("pmt_smg_level_apply: bad argument");
  if (!h->ready) {
    set_error("pmt_smg_level_apply: call pmt_smg_update first");
    return PMT_ERR_STATE;
  }
  return smg_stencil(h, level, 0, u, nullptr, 0.0, y, (hipStream_t)stream);
}

int pmt_smg_pcg_solve(pmt_smg* h, const double* b, double* x, int x0_is_zero, double tol, int maxit, int restart,
                      int* iters, double* relres, double* history, void* stream) {
  if (!h || !b || !x || maxit < 0 || restart < 1) return arg_error("pmt_smg_pcg_solve: bad argument");
  if (!h->ready) {
    set_error("pmt_smg_pcg_solve: call pmt_smg_update first");
    return PMT_ERR_STATE;
  }
  hipStream_t s = (hipStream_t)stream;
  const GridDims& g = h->g[0];
  const long long n = g.nnodes;
  double hs[4];
  if (iters) *iters = 0;
  if (relres) *relres = 0.0;
  if (x0_is_zero) {
    scalar_fill_kernel<<<blocks_for(n, 256), 256, 0, s>>>(n, 0.0, x);
    PMT_TRY(check_launch("scalar_fill_kernel"));
  }
  // r = b - A x ; bn = |b| ; |r|
  PMT_TRY(launch_scalar_fine(1, g, h->P0, h->xe, h->bcm, h->bcdiag, x, b, h->r, s));
  PMT_TRY(launch_dot(n, b, b, h->sc, h->work, s));
  PMT_TRY(launch_dot(n, h->r, h->r, h->sc + 1, h->work, s));
  PMT_TRY(read_scalars(h, hs, 2, s));
  const double bn = std::sqrt(hs[0]);
  if (!std::isfinite(hs[0]) || !std::isfinite(hs[1])) {
    set_error("pmt_smg_pcg_solve: non-finite right-hand side or start vector");
    return PMT_ERR_NUMERIC;
  }
  if (bn == 0.0) {  // b = 0: x = 0 (CG.solve returns x0 only when it already meets tol)
    scalar_fill_kernel<<<blocks_for(n, 256), 256, 0, s>>>(n, 0.0, x);
    return check_launch("scalar_fill_kernel");
  }
  double tval = std::sqrt(hs[1]) / bn;
  if (history) history[0] = tval;
  if (tval <= tol) {
    if (relres) *relres = tval;
    return PMT_OK;
  }
  // z = M r ; p = orth(z) = z / |z| (CG.solve works on (n, 1) blocks and orth normalises the first
  // direction, iterative.py:259-292, 360-362)
  PMT_TRY(smg_vcycle(h, 0, h->r, h->z, s));
  PMT_TRY(launch_dot(n, h->z, h->z, h->sc, h->work, s));
  PMT_TRY(read_scalars(h, hs, 1, s));
  PMT_TRY(launch_axpby(n, 1.0 / std::sqrt(hs[0]), h->z, 0.0, h->p, s));
  int it = 0;
  for (it = 0; it < maxit; ++it) {
    PMT_TRY(launch_scalar_fine(0, g, h->P0, h->xe, h->bcm, h->bcdiag, h->p, nullptr, h->q, s));
    PMT_TRY(launch_dot(n, h->p, h->q, h->sc, h->work, s));
    PMT_TRY(launch_dot(n, h->p, h->r, h->sc + 1, h->work, s));
    PMT_TRY(read_scalars(h, hs, 2, s));
    const double pq = hs[0], alpha = hs[1] / pq;
    PMT_TRY(launch_axpby(n, alpha, h->p, 1.0, x, s));
    if (it % restart == 0) {
      PMT_TRY(launch_scalar_fine(1, g, h->P0, h->xe, h->bcm, h->bcdiag, x, b, h->r, s));
    } else {
      PMT_TRY(launch_axpby(n, -alpha, h->q, 1.0, h->r, s));
    }
    PMT_TRY(launch_dot(n, h->r, h->r, h->sc, h->work, s));
    PMT_TRY(read_scalars(h, hs, 1, s));
    if (!std::isfinite(hs[0])) {
      set_error(("pmt_smg_pcg_solve: non-finite residual at iteration " + std::to_string(it + 1)).c_str());
      if (iters) *iters = it + 1;
      return PMT_ERR_NUMERIC;
    }
    tval = std::sqrt(hs[0]) / bn;
    if (history) history[it + 1] = tval;
    if (tval <= tol) break;
    PMT_TRY(smg_vcycle(h, 0, h->r, h->z, s));
    // beta = -(q.z) / (p.q) ; p = z + beta p
    PMT_TRY(launch_dot(n, h->q, h->z, h->sc, h->work, s));
    PMT_TRY(read_scalars(h, hs, 1, s));
    const double beta = -hs[0] / pq;
    PMT_TRY(launch_axpby(n, 1.0, h->z, beta, h->p, s));
  }
  if (iters) *iters = it < maxit ? it + 1 : maxit;
  if (relres) *relres = tval;
  return PMT_OK;
}

}  // extern "C"
