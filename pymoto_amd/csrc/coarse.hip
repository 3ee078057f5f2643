// Coarse-grid operators of the geometric multigrid (pymoto/solvers/iterative.py:168-176).
//
// Galerkin R^T A R element by element: every fine element c lies in one coarse element E and
// the trilinear prolongation restricted to c only reads E's nodes, u_c = P_c u_E with
// P_c = p_c (x) I_D, p_c[a][A] = prod_dims w(f, A_bit), f = child_bit + a_bit in {0,1,2},
// w(0,0) = w(2,1) = 1, w(1,*) = 1/2, else 0.  Hence R^T A R = assemble_E sum_c P_c^T A_c P_c.
// On the fine level A_c = x_c M_c K0 M_c + D_c with M_c the Dirichlet mask and D_c the bcdiag
// diagonal of Dirichlet dofs shared equally by the mult(node) elements around each node, so
// that the element sum reproduces A = P K P + bcdiag I_bc of the reference assembly
// (assembly.py:209-230, 267-272).  On coarser levels A_c is the stored element matrix of the
// child.  The coarse element matrices are then summed per node into a 3^D block stencil.
//
// Coarsest level: dense matrix from the stencil, explicit inverse by blocked Gauss-Jordan
// (no pivoting; the Galerkin operator is SPD), applied as a dense GEMV in every V-cycle -
// the direct solve of auto_determine_solver (iterative.py:174-176).
#include "launch.h"

namespace pmt {

// One block per coarse element; thread t = (A, B) node pair owns the D x D block (A, B) of K_E.
// FINE: A_c built from x_c, K0, the Dirichlet mask and bcdiag/mult; else read from Kf.
// Does coarse element (CI, CJ, CK) contain a Dirichlet node among its 3^D fine nodes?
template <int DIM>
__device__ inline bool touches_bc(const GridDims& gf, const uint8_t* __restrict__ bcm, int CI, int CJ, int CK) {
  for (int q = 0; q < (DIM == 3 ? 27 : 9); ++q) {
    const int i = 2 * CI + q % 3, j = 2 * CJ + (q / 3) % 3, k = DIM == 3 ? 2 * CK + q / 9 : 0;
    if (bcm[((long long)k * gf.ny + j) * gf.nx + i]) return true;
  }
  return false;
}

// Fine -> first coarse level (hex8).  A child without Dirichlet nodes contributes x_c G_c with
// G_c = P_c^T K0 P_c fixed (P_c^T (x_c K0) P_c); the children with Dirichlet nodes contribute the masked
// P_c^T (x_c M K0 M + D_c) P_c, collected per flagged coarse element as a correction dK in a compact
// list (the order of the list is irrelevant: each entry only touches its own element's correction).
//  1. galerkin_bc_flags_kernel: per coarse element the bit set of children with a Dirichlet node, the
//     list of flagged elements and the element -> list position map (-1: not flagged);
//  2. galerkin_fine_bc_kernel: dK of the listed elements (one wave per element, lane = (A, B)).
__global__ void galerkin_bc_flags_kernel(GridDims gf, GridDims gc, const uint8_t* __restrict__ bcm,
                                         uint8_t* __restrict__ cflag, int* __restrict__ list, int* __restrict__ count,
                                         int* __restrict__ map) {
  const long long E = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (E >= gc.nel) return;
  const int CI = (int)(E % gc.ex);
  const long long tt = E / gc.ex;
  const int CJ = (int)(tt % gc.ey), CK = (int)(tt / gc.ey);
  unsigned m27 = 0;  // bit q: fine node (q%3, q/3%3, q/9) of the element is Dirichlet
  for (int q = 0; q < 27; ++q) {
    const int i = 2 * CI + q % 3, j = 2 * CJ + (q / 3) % 3, k = 2 * CK + q / 9;
    if (bcm[((long long)k * gf.ny + j) * gf.nx + i]) m27 |= 1u << q;
  }
  unsigned flags = 0;
  for (int c = 0; c < 8; ++c)
    for (int a = 0; a < 8; ++a) {
      const int q = ((c >> 2) + (a >> 2)) * 9 + (((c >> 1) & 1) + ((a >> 1) & 1)) * 3 + (c & 1) + (a & 1);
      if ((m27 >> q) & 1u) flags |= 1u << c;
    }
  cflag[E] = (uint8_t)flags;
  int idx = -1;
  if (flags) {
    idx = atomicAdd(count, 1);
    list[idx] = (int)E;
  }
  map[E] = idx;
}

__global__ void zero_int_kernel(int* p) { *p = 0; }

void zero_int(int* p, hipStream_t s) { zero_int_kernel<<<1, 1, 0, s>>>(p); }

// dK of a listed element: block (A, B) of the sum over its flagged children of P_c^T (x_c M K0 M + D_c) P_c.
// One 64-lane block per listed element, lane = (A, B) 3x3 block; per child in two passes through LDS:
// T[a][B] = sum_b w(c, b, B) K_c[a][b] (lane a = A), then dK[A][B] += sum_a w(c, a, A) T[a][B] - no
// divergent loops (the first form, with data-dependent skips over zero weights, took 200 us per update at
// 256x128x128).
__global__ void __launch_bounds__(64)
galerkin_fine_bc_kernel(GridDims gf, GridDims gc, const double* __restrict__ K0, const double* __restrict__ xe,
                        const uint8_t* __restrict__ bcm, double bcdiag, const uint8_t* __restrict__ cflag,
                        const int* __restrict__ list, double* __restrict__ dK) {
  constexpr int NE = 24;
  __shared__ double sK0[NE * NE];
  __shared__ double sT[8][8][9];  // [a][B][d * 3 + d2]
  const int lane = threadIdx.x, A = lane >> 3, B = lane & 7;
  for (int q = lane; q < NE * NE; q += 64) sK0[q] = K0[q];
  const long long w = blockIdx.x;
  const long long E = list[w];
  const int CI = (int)(E % gc.ex);
  const long long tt = E / gc.ex;
  const int CJ = (int)(tt % gc.ey), CK = (int)(tt / gc.ey);
  const unsigned flags = cflag[E];
  double acc[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) acc[q] = 0.0;
  __syncthreads();
  for (int c = 0; c < 8; ++c) {
    if (!((flags >> c) & 1u)) continue;  // uniform over the block
    const int fi = 2 * CI + (c & 1), fj = 2 * CJ + ((c >> 1) & 1), fk = 2 * CK + ((c >> 2) & 1);
    const double xv = xe[((long long)fk * gf.ey + fj) * gf.ex + fi];
    unsigned mk[8];
#pragma unroll
    for (int b = 0; b < 8; ++b)
      mk[b] = bcm[((long long)(fk + ((b >> 2) & 1)) * gf.ny + fj + ((b >> 1) & 1)) * gf.nx + fi + (b & 1)];
    // row corner a = A of the child: its Dirichlet mask and the number of elements sharing its node
    const int ai = fi + (A & 1), aj = fj + ((A >> 1) & 1), ak = fk + ((A >> 2) & 1);
    const unsigned ma = mk[A];
    const double dsplit =
        bcdiag / (double)(((ai > 0) + (ai < gf.nx - 1)) * ((aj > 0) + (aj < gf.ny - 1)) * ((ak > 0) + (ak < gf.nz - 1)));
    double t[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) t[q] = 0.0;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const double wb = pweight<3>(c, b, B);
      const unsigned mb = mk[b];
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int d2 = 0; d2 < 3; ++d2) {
          const bool fa = (ma >> d) & 1u, fb = (mb >> d2) & 1u;
          double v = (fa || fb) ? 0.0 : xv * sK0[(A * 3 + d) * NE + b * 3 + d2];
          if (A == b && d == d2 && fa) v += dsplit;
          t[d * 3 + d2] = fma(wb, v, t[d * 3 + d2]);
        }
    }
    __syncthreads();  // the previous child's reads of sT are done
#pragma unroll
    for (int q = 0; q < 9; ++q) sT[A][B][q] = t[q];
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const double wa = pweight<3>(c, a, A);
#pragma unroll
      for (int q = 0; q < 9; ++q) acc[q] = fma(wa, sT[a][B][q], acc[q]);
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int d2 = 0; d2 < 3; ++d2) dK[w * NE * NE + (A * 3 + d) * NE + B * 3 + d2] = acc[d * 3 + d2];
}

template <int DIM, bool FINE>
__global__ void __launch_bounds__(64)
galerkin_kernel(GridDims gf, GridDims gc, const double* __restrict__ K0, const double* __restrict__ xe,
                const uint8_t* __restrict__ bcm, double bcdiag, const double* __restrict__ Kf,
                double* __restrict__ Kc, int bc_only) {
  constexpr int NN = 1 << DIM;
  constexpr int NE = NN * DIM;
  __shared__ double sA[NE * NE];
  __shared__ double sT[NE * NE];
  __shared__ double sK0[FINE ? NE * NE : 1];
  const long long E = blockIdx.x;
  const int tid = threadIdx.x;  // NN*NN threads
  const int CI = (int)(E % gc.ex);
  const long long tt = E / gc.ex;
  const int CJ = (int)(tt % gc.ey);
  const int CK = (int)(tt / gc.ey);
  if (FINE && bc_only && !(bcm && touches_bc<DIM>(gf, bcm, CI, CJ, CK))) return;  // whole block leaves
  if (FINE) {
    for (int q = tid; q < NE * NE; q += NN * NN) sK0[q] = K0[q];
  }
  const int tA = tid / NN, tB = tid % NN;
  double acc[DIM][DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d)
#pragma unroll
    for (int d2 = 0; d2 < DIM; ++d2) acc[d][d2] = 0.0;

  for (int c = 0; c < NN; ++c) {
    const int fi = 2 * CI + (c & 1), fj = 2 * CJ + ((c >> 1) & 1), fk = DIM == 3 ? 2 * CK + ((c >> 2) & 1) : 0;
    const long long fe = ((long long)fk * gf.ey + fj) * gf.ex + fi;
    __syncthreads();  // previous child's sA/sT fully consumed (and sK0 loaded)
    if (FINE) {
      const double xc = xe[fe];
      for (int q = tid; q < NE * NE; q += NN * NN) {
        const int r = q / NE, s = q % NE;
        const int a = r / DIM, d = r % DIM, b = s / DIM, d2 = s % DIM;
        const int ai = fi + (a & 1), aj = fj + ((a >> 1) & 1), ak = DIM == 3 ? fk + ((a >> 2) & 1) : 0;
        const int bi = fi + (b & 1), bj = fj + ((b >> 1) & 1), bk = DIM == 3 ? fk + ((b >> 2) & 1) : 0;
        const long long na = ((long long)ak * gf.ny + aj) * gf.nx + ai;
        const long long nb = ((long long)bk * gf.ny + bj) * gf.nx + bi;
        const bool fa = bcm && ((bcm[na] >> d) & 1u);
        const bool fb = bcm && ((bcm[nb] >> d2) & 1u);
        double v = (fa || fb) ? 0.0 : xc * sK0[q];
        if (r == s && fa) {
          const int mult = ((ai > 0) + (ai < gf.nx - 1)) * ((aj > 0) + (aj < gf.ny - 1)) *
                           (DIM == 3 ? ((ak > 0) + (ak < gf.nz - 1)) : 1);
          v += bcdiag / (double)mult;
        }
        sA[q] = v;
      }
    } else {
      for (int q = tid; q < NE * NE; q += NN * NN) sA[q] = Kf[fe * NE * NE + q];
    }
    __syncthreads();
    // T = A_c P_c : thread (a = tA, B = tB) computes rows a*D+d, cols B*D+d2
    {
      const int a = tA, B = tB;
      double t[DIM][DIM];
#pragma unroll
      for (int d = 0; d < DIM; ++d)
#pragma unroll
        for (int d2 = 0; d2 < DIM; ++d2) t[d][d2] = 0.0;
      for (int b = 0; b < NN; ++b) {
        const double w = pweight<DIM>(c, b, B);
        if (w == 0.0) continue;
#pragma unroll
        for (int d = 0; d < DIM; ++d)
#pragma unroll
          for (int d2 = 0; d2 < DIM; ++d2) t[d][d2] = fma(w, sA[(a * DIM + d) * NE + b * DIM + d2], t[d][d2]);
      }
#pragma unroll
      for (int d = 0; d < DIM; ++d)
#pragma unroll
        for (int d2 = 0; d2 < DIM; ++d2) sT[(a * DIM + d) * NE + B * DIM + d2] = t[d][d2];
    }
    __syncthreads();
    // K_E(A, B) += sum_a p_c[a][A] T(a, B)
    for (int a = 0; a < NN; ++a) {
      const double w = pweight<DIM>(c, a, tA);
      if (w == 0.0) continue;
#pragma unroll
      for (int d = 0; d < DIM; ++d)
#pragma unroll
        for (int d2 = 0; d2 < DIM; ++d2) acc[d][d2] = fma(w, sT[(a * DIM + d) * NE + tB * DIM + d2], acc[d][d2]);
    }
  }
#pragma unroll
  for (int d = 0; d < DIM; ++d)
#pragma unroll
    for (int d2 = 0; d2 < DIM; ++d2) Kc[E * NE * NE + (tA * DIM + d) * NE + tB * DIM + d2] = acc[d][d2];
}

// Level-1 stencil straight from the fine densities (hex8): S[o](n) = sum over the coarse elements E
// holding n and n+o of K_E(loc n, loc n+o), K_E = sum over unflagged children of x_c G_c plus the
// correction dK of flagged elements.  Thread = coarse node; the 64 child densities of its 8 elements
// are held in registers, the offset/element/child loops are unrolled so every G entry is a
// compile-time offset (scalar loads).  KIND 0: full stencil (out.S); 1: half stencil and D^-1; 2: D^-1
// only (the level's operator is applied without a stencil).
template <int KIND>
__global__ void __launch_bounds__(256)
stencil_from_fine_kernel(GridDims gf, GridDims gc, const double* __restrict__ G, const double* __restrict__ xe,
                         const uint8_t* __restrict__ cflag, const int* __restrict__ map, const double* __restrict__ dK,
                         StencilOut out) {
  constexpr int NE = 24;
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= gc.nnodes) return;
  const int i = (int)(n % gc.nx);
  const long long t = n / gc.nx;
  const int j = (int)(t % gc.ny);
  const int k = (int)(t / gc.ny);
  const long long nn = gc.nnodes;
  double xv[8][8];
  unsigned fl[8];
  long long Eix[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int ei = i - 1 + (e & 1), ej = j - 1 + ((e >> 1) & 1), ek = k - 1 + ((e >> 2) & 1);
    const bool ok = ei >= 0 && ei < gc.ex && ej >= 0 && ej < gc.ey && ek >= 0 && ek < gc.ez;
    const long long E = ok ? ((long long)ek * gc.ey + ej) * gc.ex + ei : 0;
    Eix[e] = E;
    fl[e] = ok ? cflag[E] : 0u;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      double v = 0.0;
      if (ok) v = xe[((long long)(2 * ek + ((c >> 2) & 1)) * gf.ey + (2 * ej + ((c >> 1) & 1))) * gf.ex + 2 * ei + (c & 1)];
      xv[e][c] = ((fl[e] >> c) & 1u) ? 0.0 : v;  // flagged children come from dK below
    }
  }
  int o = 0;
#pragma unroll
  for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx, ++o) {
        if (KIND == 1 && o < 13) continue;  // backward blocks: the neighbours' forward blocks transposed
        if (KIND == 2 && o != 13) continue;
        double blk[3][3];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int d2 = 0; d2 < 3; ++d2) blk[d][d2] = 0.0;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int ex = e & 1, ey = (e >> 1) & 1, ez = (e >> 2) & 1;
          const int bx = dx + 1 - ex, by = dy + 1 - ey, bz = dz + 1 - ez;
          if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;  // compile time
          const int A = (1 - ex) + 2 * (1 - ey) + 4 * (1 - ez), B = bx + 2 * by + 4 * bz;
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            const double* g = G + c * NE * NE + (A * 3) * NE + B * 3;
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
              for (int d2 = 0; d2 < 3; ++d2) blk[d][d2] = fma(xv[e][c], g[d * NE + d2], blk[d][d2]);
          }
          if (fl[e]) {
            const double* K = dK + (long long)map[Eix[e]] * NE * NE + (A * 3) * NE + B * 3;
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
              for (int d2 = 0; d2 < 3; ++d2) blk[d][d2] += K[d * NE + d2];
          }
        }
        if (KIND == 2) {
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double v = 1.0 / blk[d][d];
            out.dinv[n * 3 + d] = v;
            if (out.dinv32) out.dinv32[n * 3 + d] = (float)v;
          }
        } else {
          emit_block<KIND == 1>(out, o, blk, n, nn);
        }
      }
}

// Per-node block stencil S[(o*D+d)*D+d2][n] = sum over elements E holding n and n+o of K_E(loc n, loc n+o)
template <int DIM>
__global__ void stencil_from_elements_kernel(GridDims g, const double* __restrict__ Ke, double* __restrict__ S) {
  constexpr int NN = 1 << DIM;
  constexpr int NE = NN * DIM;
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nnodes) return;
  const int i = (int)(n % g.nx);
  const long long t = n / g.nx;
  const int j = (int)(t % g.ny);
  const int k = (int)(t / g.ny);
  const long long nn = g.nnodes;
  constexpr int ZLO = DIM == 3 ? -1 : 0, ZHI = DIM == 3 ? 1 : 0;
  int o = 0;
  for (int dz = ZLO; dz <= ZHI; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx, ++o) {
        double blk[DIM][DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d)
#pragma unroll
          for (int d2 = 0; d2 < DIM; ++d2) blk[d][d2] = 0.0;
        for (int e = 0; e < NN; ++e) {
          const int ex = e & 1, ey = (e >> 1) & 1, ez = DIM == 3 ? (e >> 2) & 1 : 1;
          const int bx = dx + 1 - ex, by = dy + 1 - ey, bz = DIM == 3 ? dz + 1 - ez : 0;
          if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
          const int ei = i - 1 + ex, ej = j - 1 + ey, ek = DIM == 3 ? k - 1 + ez : 0;
          if (ei < 0 || ei >= g.ex || ej < 0 || ej >= g.ey || ek < 0 || ek >= g.ez) continue;
          const long long E = ((long long)ek * g.ey + ej) * g.ex + ei;
          const int locn = (1 - ex) + 2 * (1 - ey) + (DIM == 3 ? 4 * (1 - ez) : 0);
          const int locm = bx + 2 * by + 4 * bz;
          const double* KE = Ke + E * NE * NE;
#pragma unroll
          for (int d = 0; d < DIM; ++d)
#pragma unroll
            for (int d2 = 0; d2 < DIM; ++d2) blk[d][d2] += KE[(locn * DIM + d) * NE + locm * DIM + d2];
        }
#pragma unroll
        for (int d = 0; d < DIM; ++d)
#pragma unroll
          for (int d2 = 0; d2 < DIM; ++d2) S[((long long)(o * DIM + d) * DIM + d2) * nn + n] = blk[d][d2];
      }
}

// dense row-major A (n x n, n = nnodes*D) from the stencil; A must be zero-filled first
template <int DIM>
__global__ void dense_from_stencil_kernel(GridDims g, const double* __restrict__ S, double* __restrict__ A) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nnodes) return;
  const int i = (int)(n % g.nx);
  const long long t = n / g.nx;
  const int j = (int)(t % g.ny);
  const int k = (int)(t / g.ny);
  const long long N = g.nnodes * DIM;
  constexpr int ZLO = DIM == 3 ? -1 : 0, ZHI = DIM == 3 ? 1 : 0;
  int o = 0;
  for (int dz = ZLO; dz <= ZHI; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx, ++o) {
        const int ii = i + dx, jj = j + dy, kk = k + dz;
        if (ii < 0 || ii >= g.nx || jj < 0 || jj >= g.ny || kk < 0 || kk >= g.nz) continue;
        const long long m = ((long long)kk * g.ny + jj) * g.nx + ii;
        for (int d = 0; d < DIM; ++d)
          for (int d2 = 0; d2 < DIM; ++d2)
            A[(n * DIM + d) * N + m * DIM + d2] = S[((long long)(o * DIM + d) * DIM + d2) * g.nnodes + n];
      }
}

// ---------------- blocked Gauss-Jordan inversion (in place), block size GJB ----------------
constexpr int GJB = 64;

// invert the kk x kk diagonal block at (k0, k0) in LDS, write it to P (GJB x GJB, row-major).
// 16 x 16 threads, each owning a 4 x 4 set of entries (rows tr + 16 i, columns tc + 16 j); the block
// is padded with the identity beyond kk, so every pivot step is the same 16 multiply-adds per thread
// between two barriers (in-place Gauss-Jordan, no pivoting: the Galerkin operator is SPD).
__global__ void __launch_bounds__(1024) gj_diag_kernel(long long n, long long k0, int kk, const double* __restrict__ A,
                                                        double* __restrict__ P) {
  // 32 x 32 threads, each owning a 2 x 2 sub-block; two LDS copies of the block alternate between
  // pivot steps, so one barrier separates a step's reads from the next step's writes
  __shared__ double sM[2][GJB][GJB + 1];
  const int tr = threadIdx.x / 32, tc = threadIdx.x % 32;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int r = tr + 32 * i, c = tc + 32 * j;
      sM[0][r][c] = (r < kk && c < kk) ? A[(k0 + r) * n + k0 + c] : (r == c ? 1.0 : 0.0);
    }
  __syncthreads();
  for (int p = 0; p < kk; ++p) {
    double(*const src)[GJB + 1] = sM[p & 1];
    double(*const dst)[GJB + 1] = sM[(p + 1) & 1];
    const double inv = 1.0 / src[p][p];
    double mrp[2], mpc[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) mrp[i] = src[tr + 32 * i][p];
#pragma unroll
    for (int j = 0; j < 2; ++j) mpc[j] = src[p][tc + 32 * j];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int r = tr + 32 * i, c = tc + 32 * j;
        const double m = src[r][c];
        double v;
        if (r == p)
          v = c == p ? inv : m * inv;
        else
          v = c == p ? -mrp[i] * inv : m - mrp[i] * (mpc[j] * inv);
        dst[r][c] = v;
      }
    __syncthreads();
  }
  for (int q = threadIdx.x; q < GJB * GJB; q += blockDim.x) {  // zero outside kk x kk: consumers read full tiles
    const int r = q / GJB, c = q % GJB;
    P[q] = (r < kk && c < kk) ? sM[kk & 1][r][c] : 0.0;
  }
}

// C[i][c] = A[i][k0+c] (old column panel), all rows i
__global__ void gj_colcopy_kernel(long long n, long long k0, int kk, const double* __restrict__ A,
                                  double* __restrict__ C) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n * kk) return;
  const long long i = q / kk;
  const int c = (int)(q % kk);
  C[i * GJB + c] = A[i * n + k0 + c];
}

// row panel A[K][j] = P A[K][j] for j outside K.  Block = 16 columns, staged in LDS with P (the panel
// is updated in place); thread = one column and 4 of the 64 rows (r, r + 16, r + 32, r + 48).
constexpr int GJ_RP_COLS = 16;
__global__ void __launch_bounds__(256) gj_rowpanel_kernel(long long n, long long k0, int kk, double* __restrict__ A,
                                                           const double* __restrict__ P) {
  __shared__ double sP[GJB][GJB + 1];
  __shared__ double sA[GJB][GJ_RP_COLS + 1];
  const int tid = threadIdx.x;
  const long long j0 = (long long)blockIdx.x * GJ_RP_COLS;
  for (int q = tid; q < GJB * GJB; q += blockDim.x) sP[q / GJB][q % GJB] = P[q];
  for (int q = tid; q < GJB * GJ_RP_COLS; q += blockDim.x) {
    const int r = q / GJ_RP_COLS, c = q % GJ_RP_COLS;
    const long long j = j0 + c;
    sA[r][c] = (r < kk && j < n) ? A[(k0 + r) * n + j] : 0.0;
  }
  __syncthreads();
  const int jj = tid % GJ_RP_COLS, rg = tid / GJ_RP_COLS;
  const long long j = j0 + jj;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int c = 0; c < kk; ++c) {
    const double a = sA[c][jj];
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] = fma(sP[rg + 16 * i][c], a, acc[i]);
  }
  if (j >= n || (j >= k0 && j < k0 + kk)) return;  // the pivot columns are set by gj_update_kernel
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = rg + 16 * i;
    if (r < kk) A[(k0 + r) * n + j] = acc[i];
  }
}

// rank-kk update of all tiles outside the pivot row block:
//   j outside K: A[i][j] -= C[i][:] A[K][j] ; j inside K: A[i][K] = -C[i][:] P ; pivot block = P
__global__ void __launch_bounds__(256) gj_update_kernel(long long n, long long k0, int kk, double* __restrict__ A,
                                                         const double* __restrict__ C, const double* __restrict__ P) {
  __shared__ double sC[64][GJB + 1];
  __shared__ double sR[GJB][64 + 1];
  const long long i0 = (long long)blockIdx.y * 64, j0 = (long long)blockIdx.x * 64;
  const int tid = threadIdx.x;
  const bool rowK = (i0 == k0);  // tiles are aligned to GJB, so the pivot row block is one tile row
  const bool colK = (j0 == k0);
  if (rowK && !colK) return;  // already holds the new row panel
  if (rowK && colK) {
    for (int q = tid; q < kk * kk; q += blockDim.x) {
      const int r = q / kk, c = q % kk;
      A[(k0 + r) * n + k0 + c] = P[r * GJB + c];
    }
    return;
  }
  // stage C tile (64 x kk) and R tile (kk x 64): R = new row panel, or P for the pivot columns
  for (int q = tid; q < 64 * GJB; q += blockDim.x) {
    const int r = q / GJB, c = q % GJB;
    const long long i = i0 + r;
    sC[r][c] = (i < n && c < kk) ? C[i * GJB + c] : 0.0;
  }
  for (int q = tid; q < GJB * 64; q += blockDim.x) {
    const int r = q / 64, c = q % 64;
    const long long j = j0 + c;
    double v = 0.0;
    if (r < kk && j < n) v = colK ? (c < kk ? P[r * GJB + c] : 0.0) : A[(k0 + r) * n + j];
    sR[r][c] = v;
  }
  __syncthreads();
  const int tr = (tid / 16) * 4, tc = (tid % 16) * 4;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  for (int c = 0; c < kk; ++c) {
    double cv[4], rv[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) cv[a] = sC[tr + a][c];
#pragma unroll
    for (int b = 0; b < 4; ++b) rv[b] = sR[c][tc + b];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = fma(cv[a], rv[b], acc[a][b]);
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const long long i = i0 + tr + a;
    if (i >= n) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const long long j = j0 + tc + b;
      if (j >= n) continue;
      if (colK) {
        if (j < k0 + kk) A[i * n + j] = -acc[a][b];
      } else {
        A[i * n + j] -= acc[a][b];
      }
    }
  }
}

// y = A x, one 64-lane wave per row (T = double for the FP64 V-cycle, float for FP32)
template <typename T>
__global__ void __launch_bounds__(256) gemv_kernel(long long n, const T* __restrict__ A, const T* __restrict__ x,
                                                    T* __restrict__ y) {
  const long long row = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const T* a = A + row * n;
  T s = 0;
  for (long long c = lane; c < n; c += 64) s = fma(a[c], x[c], s);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) y[row] = s;
}

// y = (float)(A (double)x): the FP64 coarsest inverse applied to an FP32 right-hand side, FP64 sums (the
// widening, gemv_kernel<double> and the rounding of separate launches, in one)
__global__ void __launch_bounds__(256) gemv_f32io_kernel(long long n, const double* __restrict__ A,
                                                         const float* __restrict__ x, float* __restrict__ y) {
  const long long row = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const double* a = A + row * n;
  double s = 0;
  for (long long c = lane; c < n; c += 64) s = fma(a[c], (double)x[c], s);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) y[row] = (float)s;
}

// y = (float) x
__global__ void to_float_kernel(long long n, const double* __restrict__ x, float* __restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}

void galerkin_child_matrices(int dim, const double* K0, double* G) {
  const int NN = 1 << dim, NE = NN * dim;
  for (int c = 0; c < NN; ++c) {
    double* g = G + (long long)c * NE * NE;
    for (int q = 0; q < NE * NE; ++q) g[q] = 0.0;
    // G_c[A d][B d2] = sum_{a, b} p_c[a][A] K0[a d][b d2] p_c[b][B]
    for (int A = 0; A < NN; ++A)
      for (int B = 0; B < NN; ++B)
        for (int a = 0; a < NN; ++a) {
          const double wa = dim == 3 ? pweight<3>(c, a, A) : pweight<2>(c, a, A);
          if (wa == 0.0) continue;
          for (int b = 0; b < NN; ++b) {
            const double wb = dim == 3 ? pweight<3>(c, b, B) : pweight<2>(c, b, B);
            if (wb == 0.0) continue;
            for (int d = 0; d < dim; ++d)
              for (int d2 = 0; d2 < dim; ++d2)
                g[(A * dim + d) * NE + B * dim + d2] += wa * wb * K0[(a * dim + d) * NE + b * dim + d2];
          }
        }
  }
}

int launch_galerkin_fine(const GridDims& gf, const double* K0, const double* xe, const uint8_t* bcm, double bcdiag,
                         double* Kc, hipStream_t s) {
  GridDims gc = make_grid(gf.ex / 2, gf.ey / 2, gf.dim == 3 ? gf.ez / 2 : 0);
  if (gf.dim == 3)
    galerkin_kernel<3, true><<<(unsigned)gc.nel, 64, 0, s>>>(gf, gc, K0, xe, bcm, bcdiag, nullptr, Kc, 0);
  else
    galerkin_kernel<2, true><<<(unsigned)gc.nel, 16, 0, s>>>(gf, gc, K0, xe, bcm, bcdiag, nullptr, Kc, 0);
  return check_launch("galerkin_kernel<fine>");
}

int launch_level1_flags(const GridDims& gf, const uint8_t* bcm, uint8_t* cflag, int* list, int* map, hipStream_t s) {
  GridDims gc = make_grid(gf.ex / 2, gf.ey / 2, gf.ez / 2);
  int* count = list + gc.nel;
  zero_int(count, s);
  galerkin_bc_flags_kernel<<<blocks_for(gc.nel, 256), 256, 0, s>>>(gf, gc, bcm, cflag, list, count, map);
  return check_launch("galerkin_bc_flags_kernel");
}

int launch_level1_dk(const GridDims& gf, const double* K0, const double* xe, const uint8_t* bcm, double bcdiag,
                     const uint8_t* cflag, const int* list, int nlist, double* dK, hipStream_t s) {
  if (nlist <= 0) return PMT_OK;
  GridDims gc = make_grid(gf.ex / 2, gf.ey / 2, gf.ez / 2);
  galerkin_fine_bc_kernel<<<(unsigned)nlist, 64, 0, s>>>(gf, gc, K0, xe, bcm, bcdiag, cflag, list, dK);
  return check_launch("galerkin_fine_bc_kernel");
}

int launch_galerkin_coarse(const GridDims& gf, const double* Kf, double* Kc, hipStream_t s) {
  GridDims gc = make_grid(gf.ex / 2, gf.ey / 2, gf.dim == 3 ? gf.ez / 2 : 0);
  if (gf.dim == 3)
    galerkin_kernel<3, false><<<(unsigned)gc.nel, 64, 0, s>>>(gf, gc, nullptr, nullptr, nullptr, 0.0, Kf, Kc, 0);
  else
    galerkin_kernel<2, false><<<(unsigned)gc.nel, 16, 0, s>>>(gf, gc, nullptr, nullptr, nullptr, 0.0, Kf, Kc, 0);
  return check_launch("galerkin_kernel<coarse>");
}

int launch_stencil_from_fine(const GridDims& gf, const double* G, const double* xe, const uint8_t* cflag,
                             const int* map, const double* dK, const StencilOut& out, hipStream_t s) {
  GridDims gc = make_grid(gf.ex / 2, gf.ey / 2, gf.ez / 2);
  if (out.S)
    stencil_from_fine_kernel<0><<<blocks_for(gc.nnodes, 256), 256, 0, s>>>(gf, gc, G, xe, cflag, map, dK, out);
  else if (out.H)
    stencil_from_fine_kernel<1><<<blocks_for(gc.nnodes, 256), 256, 0, s>>>(gf, gc, G, xe, cflag, map, dK, out);
  else
    stencil_from_fine_kernel<2><<<blocks_for(gc.nnodes, 256), 256, 0, s>>>(gf, gc, G, xe, cflag, map, dK, out);
  return check_launch("stencil_from_fine_kernel");
}

int launch_stencil_from_elements(const GridDims& g, const double* Ke, double* S, hipStream_t s) {
  dim3 grid(blocks_for(g.nnodes, 128));
  if (g.dim == 3)
    stencil_from_elements_kernel<3><<<grid, 128, 0, s>>>(g, Ke, S);
  else
    stencil_from_elements_kernel<2><<<grid, 128, 0, s>>>(g, Ke, S);
  return check_launch("stencil_from_elements_kernel");
}

int launch_dense_from_stencil(const GridDims& g, const double* S, double* A, hipStream_t s) {
  const long long N = g.nnodes * g.dim;
  PMT_TRY(launch_fill(N * N, 0.0, A, s));
  dim3 grid(blocks_for(g.nnodes, 128));
  if (g.dim == 3)
    dense_from_stencil_kernel<3><<<grid, 128, 0, s>>>(g, S, A);
  else
    dense_from_stencil_kernel<2><<<grid, 128, 0, s>>>(g, S, A);
  return check_launch("dense_from_stencil_kernel");
}

int gj_invert(long long n, double* A, double* wsP, double* wsC, hipStream_t s) {
  const long long nb = (n + GJB - 1) / GJB;
  for (long long kb = 0; kb < nb; ++kb) {
    const long long k0 = kb * GJB;
    const int kk = (int)((n - k0) < GJB ? (n - k0) : GJB);
    gj_diag_kernel<<<1, 1024, 0, s>>>(n, k0, kk, A, wsP);
    PMT_TRY(check_launch("gj_diag_kernel"));
    gj_colcopy_kernel<<<blocks_for(n * kk, 256), 256, 0, s>>>(n, k0, kk, A, wsC);
    PMT_TRY(check_launch("gj_colcopy_kernel"));
    gj_rowpanel_kernel<<<(unsigned)((n + GJ_RP_COLS - 1) / GJ_RP_COLS), 256, 0, s>>>(n, k0, kk, A, wsP);
    PMT_TRY(check_launch("gj_rowpanel_kernel"));
    dim3 grid((unsigned)nb, (unsigned)nb);
    gj_update_kernel<<<grid, 256, 0, s>>>(n, k0, kk, A, wsC, wsP);
    PMT_TRY(check_launch("gj_update_kernel"));
  }
  return PMT_OK;
}

template <typename T>
int launch_gemv(long long n, const T* A, const T* x, T* y, hipStream_t s) {
  gemv_kernel<T><<<blocks_for(n, 4), 256, 0, s>>>(n, A, x, y);
  return check_launch("gemv_kernel");
}
template int launch_gemv<double>(long long, const double*, const double*, double*, hipStream_t);
template int launch_gemv<float>(long long, const float*, const float*, float*, hipStream_t);

int launch_gemv_f32io(long long n, const double* A, const float* x, float* y, hipStream_t s) {
  gemv_f32io_kernel<<<blocks_for(n, 4), 256, 0, s>>>(n, A, x, y);
  return check_launch("gemv_f32io_kernel");
}

// y = (float) x and u1 = w * dinv * y: an FP32 V-cycle's right-hand side and its first fine sweep from zero
// (the expression of jacobi0_kernel)
__global__ void to_float_first_kernel(long long n, const double* __restrict__ x, float* __restrict__ y,
                                      const float* __restrict__ dinv, float w, float* __restrict__ u1) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const float v = (float)x[i];
    y[i] = v;
    u1[i] = w * dinv[i] * v;
  }
}

int launch_to_float_first(long long n, const double* x, float* y, const float* dinv, float w, float* u1,
                          hipStream_t s) {
  if (n <= 0) return PMT_OK;
  to_float_first_kernel<<<blocks_for(n, 256), 256, 0, s>>>(n, x, y, dinv, w, u1);
  return check_launch("to_float_first_kernel");
}

int launch_to_float(long long n, const double* x, float* y, hipStream_t s) {
  if (n <= 0) return PMT_OK;
  to_float_kernel<<<blocks_for(n, 256), 256, 0, s>>>(n, x, y);
  return check_launch("to_float_kernel");
}

}  // namespace pmt
