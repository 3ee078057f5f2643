// Node-centric operator kernels for the structured Q4/hex8 grid.
//
// Fine level (matrix-free): every thread owns one node and gathers the rows of
//   K(x) = sum_e x_e K0, with Dirichlet rows/cols zeroed and bcdiag on their diagonal,
// over the 2^D elements around it.  Neighbour values are visited once (3^D of them) and the
// partial products are kept per adjacent element, so a node costs exactly
// 2^D * D * (2^D*D) multiply-adds (576 in 3D) - the flop count of an element-by-element
// product - without any scatter, atomics or colouring.  The same kernel is fused with the
// residual (r = b - K u) and the damped-Jacobi sweep (u' = u + w D^-1 (b - K u)) of
// pymoto/solvers/iterative.py:21-47, 238-255.
//
// Coarse levels: a 3^D-point block stencil (D x D blocks) stored coefficient-major
// S[c][node], c = (offset*D + d)*D + d2, which is the Galerkin operator R^T A R of
// iterative.py:173 assembled per node.
#include "k0_classes.h"
#include "launch.h"

namespace pmt {

template <int DIM>
struct Cfg {
  static constexpr int NN = 1 << DIM;     // nodes per element
  static constexpr int NE = NN * DIM;     // dofs per element
  static constexpr int NS = DIM == 3 ? 27 : 9;  // stencil points
  static constexpr int NC = NS * DIM * DIM;     // stencil coefficients per node
};

// Sign/class pattern of an isotropic box element matrix (k0_classes.h): K0[i] = sgn[i]*val[cls[i]].
template <int DIM>
struct K0Pattern;
template <>
struct K0Pattern<3> {
  static constexpr int N = K0NCLS3;
  __device__ static constexpr int cls(int i) { return K0CLS3[i]; }
  __device__ static constexpr int sgn(int i) { return K0SGN3[i]; }
  __device__ static constexpr int first(int c) { return K0FIRST3[c]; }
};
template <>
struct K0Pattern<2> {
  static constexpr int N = K0NCLS2;
  __device__ static constexpr int cls(int i) { return K0CLS2[i]; }
  __device__ static constexpr int sgn(int i) { return K0SGN2[i]; }
  __device__ static constexpr int first(int c) { return K0FIRST2[c]; }
};

inline unsigned xcd_blocks(long long n, int threads) {
  unsigned b = blocks_for(n, threads);
  return (b + 7u) & ~7u;
}

// Node tile of a block (TX*TY*TZ = 256 threads) and its one-node halo.
template <int DIM>
struct Tile;
template <>
struct Tile<3> {
  static constexpr int X = 16, Y = 4, Z = 4;
};
template <>
struct Tile<2> {
  static constexpr int X = 16, Y = 16, Z = 1;
};

// One block = one node tile.  The block stages the Dirichlet-masked displacements of the tile plus
// a one-node halo (structure of arrays, zero outside the grid) and the densities of the elements
// touching the tile (zero outside the grid) in LDS; every thread then gathers its node's rows from
// LDS with no bounds checks, so the 2^D*D*2^D*D multiply-adds run branch-free.
template <int DIM, int MODE, bool CLS, typename T>
__global__ void __launch_bounds__(256)
fine_node_kernel(GridDims g, int tiles_x, int tiles_y, int ntiles, const double* __restrict__ K0,
                 const T* __restrict__ xe, const uint8_t* __restrict__ bcm, T bcdiag,
                 const T* __restrict__ u, const T* __restrict__ b, const T* __restrict__ dinv,
                 T omega, T* y, const T* up, T beta) {
  constexpr int NN = Cfg<DIM>::NN;
  constexpr int NE = Cfg<DIM>::NE;
  constexpr int TX = Tile<DIM>::X, TY = Tile<DIM>::Y, TZ = Tile<DIM>::Z;
  constexpr int HX = TX + 2, HY = TY + 2, HZ = DIM == 3 ? TZ + 2 : 1;
  constexpr int HN = HX * HY * HZ;
  constexpr int EX = TX + 1, EY = TY + 1, EZ = DIM == 3 ? TZ + 1 : 1;
  constexpr int EN = EX * EY * EZ;
  __shared__ T su[DIM][HN];
  __shared__ T sx[EN];

  const int tile = (int)xcd_block();
  if (tile >= ntiles) return;  // whole block leaves together (grid padded for the XCD remap)
  const int tx0 = (tile % tiles_x) * TX;
  const int trest = tile / tiles_x;
  const int ty0 = (trest % tiles_y) * TY;
  const int tz0 = DIM == 3 ? (trest / tiles_y) * TZ : 0;

  // with the box pattern the whole element matrix is K0Pattern::N scalars (uniform, SGPR-resident)
  T kv[CLS ? K0Pattern<DIM>::N : 1];
  if (CLS) {
#pragma unroll
    for (int c = 0; c < K0Pattern<DIM>::N; ++c)
      kv[c] = (T)(K0Pattern<DIM>::sgn(K0Pattern<DIM>::first(c)) * K0[K0Pattern<DIM>::first(c)]);
  }

  for (int h = threadIdx.x; h < HN; h += 256) {
    const int hx = h % HX, hy = (h / HX) % HY, hz = h / (HX * HY);
    const int gi = tx0 - 1 + hx, gj = ty0 - 1 + hy, gk = DIM == 3 ? tz0 - 1 + hz : 0;
    T v[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) v[d] = 0.0;
    if (gi >= 0 && gi < g.nx && gj >= 0 && gj < g.ny && gk >= 0 && gk < g.nz) {
      const long long m = ((long long)gk * g.ny + gj) * g.nx + gi;
      const unsigned mk = bcm[m];  // the mask is mandatory: no branch around the compute below
#pragma unroll
      for (int d = 0; d < DIM; ++d) v[d] = ((mk >> d) & 1u) ? 0.0 : u[m * DIM + d];
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d) su[d][h] = v[d];
  }
  for (int q = threadIdx.x; q < EN; q += 256) {
    const int ex = q % EX, ey = (q / EX) % EY, ez = q / (EX * EY);
    const int gi = tx0 - 1 + ex, gj = ty0 - 1 + ey, gk = DIM == 3 ? tz0 - 1 + ez : 0;
    T v = 0;
    if (gi >= 0 && gi < g.ex && gj >= 0 && gj < g.ey && gk >= 0 && gk < g.ez)
      v = xe[((long long)gk * g.ey + gj) * g.ex + gi];
    sx[q] = v;
  }
  __syncthreads();

  const int lx = threadIdx.x % TX, ly = (threadIdx.x / TX) % TY, lz = threadIdx.x / (TX * TY);
  const int i = tx0 + lx, j = ty0 + ly, k = tz0 + lz;
  if (i >= g.nx || j >= g.ny || k >= g.nz) return;

  T acc[NN][DIM];
#pragma unroll
  for (int e = 0; e < NN; ++e)
#pragma unroll
    for (int d = 0; d < DIM; ++d) acc[e][d] = 0.0;

  constexpr int ZLO = DIM == 3 ? -1 : 0, ZHI = DIM == 3 ? 1 : 0;
#pragma unroll
  for (int dz = ZLO; dz <= ZHI; ++dz) {
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        const int h = ((DIM == 3 ? lz + 1 + dz : 0) * HY + (ly + 1 + dy)) * HX + (lx + 1 + dx);
        T um[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) um[d] = su[d][h];
        // elements (i-1+ex, j-1+ey, k-1+ez) containing both this node and the neighbour
#pragma unroll
        for (int e = 0; e < NN; ++e) {
          const int ex = e & 1, ey = (e >> 1) & 1, ez = DIM == 3 ? (e >> 2) & 1 : 1;
          const int bx = dx + 1 - ex, by = dy + 1 - ey, bz = DIM == 3 ? dz + 1 - ez : 0;
          if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
          const int locn = (1 - ex) + 2 * (1 - ey) + (DIM == 3 ? 4 * (1 - ez) : 0);
          const int locm = bx + 2 * by + 4 * bz;
#pragma unroll
          for (int d = 0; d < DIM; ++d)
#pragma unroll
            for (int d2 = 0; d2 < DIM; ++d2) {
              const int idx = (locn * DIM + d) * NE + locm * DIM + d2;
              T kval;
              if (CLS)
                kval = K0Pattern<DIM>::sgn(idx) > 0 ? kv[K0Pattern<DIM>::cls(idx)] : -kv[K0Pattern<DIM>::cls(idx)];
              else
                kval = (T)K0[idx];
              acc[e][d] = fma(kval, um[d2], acc[e][d]);
            }
        }
      }
    }
  }

  T out[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) out[d] = 0.0;
#pragma unroll
  for (int e = 0; e < NN; ++e) {
    const int ex = e & 1, ey = (e >> 1) & 1, ez = DIM == 3 ? (e >> 2) & 1 : 0;
    const T xv = sx[((DIM == 3 ? lz + ez : 0) * EY + (ly + ey)) * EX + (lx + ex)];
#pragma unroll
    for (int d = 0; d < DIM; ++d) out[d] = fma(xv, acc[e][d], out[d]);
  }

  const long long n = ((long long)k * g.ny + j) * g.nx + i;
  const unsigned mk = bcm[n];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const long long q = n * DIM + d;
    const T ku = ((mk >> d) & 1u) ? bcdiag * u[q] : out[d];
    if (MODE == MODE_APPLY) {
      y[q] = ku;
    } else if (MODE == MODE_RESIDUAL) {
      y[q] = b[q] - ku;
    } else if (MODE == MODE_JACOBI) {
      y[q] = u[q] + omega * dinv[q] * (b[q] - ku);
    } else {
      const T uq = u[q], pq = up ? up[q] : (T)0;
      y[q] = uq + omega * dinv[q] * (b[q] - ku) + beta * (uq - pq);
    }
  }
}

// 3D fine level, streamed along z: one block owns an SX x SY column of nodes and walks a chunk of
// node planes.  LDS holds a ring of four Dirichlet-masked displacement planes (the three a node
// plane reads plus the one being filled) and three density planes, each with its one-node
// (one-element) halo.  While the block multiplies plane k, every thread already has the global
// loads of displacement plane k+2 and density plane k+1 in flight in registers; they are written
// to the free ring slots after the multiply, so one barrier per plane separates producer and
// consumers and global latency hides behind the 576 multiply-adds per node.  The halo is loaded
// once per plane instead of once per 3D tile.
// Column shapes (SX x SY nodes per plane, 256 threads) selectable at run time (autotuner).
template <int SX_, int SY_>
struct StreamTile {
  static constexpr int SX = SX_, SY = SY_;                      // nodes per plane of a block
  static constexpr int HX = SX + 2, HY = SY + 2, HP = HX * HY;  // displacement plane with halo
  static constexpr int EX = SX + 1, EY = SY + 1, EP = EX * EY;  // density plane with halo
  static constexpr int UL = (HP + 255) / 256, XL = (EP + 255) / 256;  // slots per thread
  static_assert(SX * SY == 256, "one thread per node of a plane");
};

template <class Stream, int MODE, bool CLS, typename T>
__global__ void __launch_bounds__(256)
fine_stream_kernel(GridDims g, int tiles_x, int ntiles_xy, int nzc, int nwork, const double* __restrict__ K0,
                   const T* __restrict__ xe, const uint8_t* __restrict__ bcm, T bcdiag, const T* __restrict__ u,
                   const T* __restrict__ b, const T* __restrict__ dinv, T omega, T* y, const T* up, T beta) {
  constexpr int HX = Stream::HX, HP = Stream::HP, EX = Stream::EX, EP = Stream::EP;
  constexpr int UL = Stream::UL, XL = Stream::XL;
  __shared__ T su[4][3][HP];
  __shared__ T sx[3][EP];

  const int work = (int)xcd_block();
  if (work >= nwork) return;  // whole block leaves together
  const int txy = work % ntiles_xy;
  const int i0 = (txy % tiles_x) * Stream::SX, j0 = (txy / tiles_x) * Stream::SY;
  const int kbeg = (work / ntiles_xy) * nzc;
  const int kend = min(kbeg + nzc, g.nz);
  const int tid = threadIdx.x;

  T kv[CLS ? K0Pattern<3>::N : 1];
  if (CLS) {
#pragma unroll
    for (int c = 0; c < K0Pattern<3>::N; ++c)
      kv[c] = (T)(K0Pattern<3>::sgn(K0Pattern<3>::first(c)) * K0[K0Pattern<3>::first(c)]);
  }

  // per-thread halo slots of a displacement plane and of a density plane (plane-independent)
  long long uoff[UL];
  bool uin[UL];
#pragma unroll
  for (int l = 0; l < UL; ++l) {
    const int h = tid + 256 * l;
    const int gi = i0 - 1 + h % HX, gj = j0 - 1 + h / HX;
    uin[l] = h < HP && gi >= 0 && gi < g.nx && gj >= 0 && gj < g.ny;
    uoff[l] = (long long)gj * g.nx + gi;
  }
  long long xoff[XL];
  bool xin[XL];
#pragma unroll
  for (int l = 0; l < XL; ++l) {
    const int q = tid + 256 * l;
    const int gi = i0 - 1 + q % EX, gj = j0 - 1 + q / EX;
    xin[l] = q < EP && gi >= 0 && gi < g.ex && gj >= 0 && gj < g.ey;
    xoff[l] = (long long)gj * g.ex + gi;
  }
  const long long uplane = (long long)g.nx * g.ny, xplane = (long long)g.ex * g.ey;

  T ru[UL][3];
  unsigned rm[UL];
  T rx[XL];
  // issue the loads of displacement plane p and density plane pe into registers (no waits)
  auto fetch = [&](int p, int pe) {
#pragma unroll
    for (int l = 0; l < UL; ++l) {
      rm[l] = 7u;
      ru[l][0] = ru[l][1] = ru[l][2] = (T)0;
      if (uin[l] && p >= 0 && p < g.nz) {
        const long long m = (long long)p * uplane + uoff[l];
        rm[l] = bcm[m];
        ru[l][0] = u[3 * m];
        ru[l][1] = u[3 * m + 1];
        ru[l][2] = u[3 * m + 2];
      }
    }
#pragma unroll
    for (int l = 0; l < XL; ++l) {
      rx[l] = (T)0;
      if (xin[l] && pe >= 0 && pe < g.ez) rx[l] = xe[(long long)pe * xplane + xoff[l]];
    }
  };
  auto store_u = [&](int p) {
    T(*dst)[HP] = su[(p + 1) & 3];
#pragma unroll
    for (int l = 0; l < UL; ++l) {
      const int h = tid + 256 * l;
      if (h < HP) {
#pragma unroll
        for (int d = 0; d < 3; ++d) dst[d][h] = ((rm[l] >> d) & 1u) ? (T)0 : ru[l][d];
      }
    }
  };
  auto store_x = [&](int pe) {
    T* dst = sx[(pe + 1) % 3];
#pragma unroll
    for (int l = 0; l < XL; ++l) {
      const int q = tid + 256 * l;
      if (q < EP) dst[q] = rx[l];
    }
  };

  // prologue: displacement planes kbeg-1, kbeg, kbeg+1 and density planes kbeg-1, kbeg
  fetch(kbeg - 1, kbeg - 1);
  store_u(kbeg - 1);
  store_x(kbeg - 1);
  fetch(kbeg, kbeg);
  store_u(kbeg);
  store_x(kbeg);
  fetch(kbeg + 1, -1);
  store_u(kbeg + 1);
  __syncthreads();

  const int lx = tid % Stream::SX, ly = tid / Stream::SX;
  const int i = i0 + lx, j = j0 + ly;
  const bool own = i < g.nx && j < g.ny;

  for (int k = kbeg; k < kend; ++k) {
    fetch(k + 2, k + 1);  // in flight during the multiply
    // own-node operands of the update (also in flight)
    const long long n = ((long long)k * g.ny + j) * g.nx + i;
    unsigned mk = 0;
    T uo[3] = {0, 0, 0}, bo[3] = {0, 0, 0}, dv[3] = {0, 0, 0}, po[3] = {0, 0, 0};
    if (own) {
      mk = bcm[n];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        uo[d] = u[3 * n + d];
        if (MODE != MODE_APPLY) bo[d] = b[3 * n + d];
        if (MODE == MODE_JACOBI || MODE == MODE_CHEB) dv[d] = dinv[3 * n + d];
        if (MODE == MODE_CHEB && up) po[d] = up[3 * n + d];
      }
    }

    T acc[8][3];
#pragma unroll
    for (int e = 0; e < 8; ++e)
#pragma unroll
      for (int d = 0; d < 3; ++d) acc[e][d] = 0.0;
#pragma unroll
    for (int dz = -1; dz <= 1; ++dz) {
      T(*pl)[HP] = su[(k + dz + 1) & 3];
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy) {
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
          const int h = (ly + 1 + dy) * HX + (lx + 1 + dx);
          T um[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) um[d] = pl[d][h];
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const int ex = e & 1, ey = (e >> 1) & 1, ez = (e >> 2) & 1;
            const int bx = dx + 1 - ex, by = dy + 1 - ey, bz = dz + 1 - ez;
            if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
            const int locn = (1 - ex) + 2 * (1 - ey) + 4 * (1 - ez);
            const int locm = bx + 2 * by + 4 * bz;
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
              for (int d2 = 0; d2 < 3; ++d2) {
                const int idx = (locn * 3 + d) * 24 + locm * 3 + d2;
                T kval;
                if (CLS)
                  kval = K0Pattern<3>::sgn(idx) > 0 ? kv[K0Pattern<3>::cls(idx)] : -kv[K0Pattern<3>::cls(idx)];
                else
                  kval = (T)K0[idx];
                acc[e][d] = fma(kval, um[d2], acc[e][d]);
              }
          }
        }
      }
    }
    T out[3] = {0, 0, 0};
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int ex = e & 1, ey = (e >> 1) & 1, ez = (e >> 2) & 1;
      const T xv = sx[(k - 1 + ez + 1) % 3][(ly + ey) * EX + (lx + ex)];
#pragma unroll
      for (int d = 0; d < 3; ++d) out[d] = fma(xv, acc[e][d], out[d]);
    }
    if (own) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const T ku = ((mk >> d) & 1u) ? bcdiag * uo[d] : out[d];
        if (MODE == MODE_APPLY)
          y[3 * n + d] = ku;
        else if (MODE == MODE_RESIDUAL)
          y[3 * n + d] = bo[d] - ku;
        else if (MODE == MODE_JACOBI)
          y[3 * n + d] = uo[d] + omega * dv[d] * (bo[d] - ku);
        else
          y[3 * n + d] = uo[d] + omega * dv[d] * (bo[d] - ku) + beta * (uo[d] - po[d]);
      }
    }
    store_u(k + 2);
    store_x(k + 1);
    __syncthreads();
  }
}

template <int DIM>
__global__ void fine_diag_kernel(GridDims g, const double* __restrict__ K0, const double* __restrict__ xe,
                                 const uint8_t* __restrict__ bcm, double bcdiag, double* __restrict__ diag,
                                 int invert, float* __restrict__ diag32) {
  constexpr int NN = Cfg<DIM>::NN;
  constexpr int NE = Cfg<DIM>::NE;
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nnodes) return;
  const int i = (int)(n % g.nx);
  const long long t = n / g.nx;
  const int j = (int)(t % g.ny);
  const int k = (int)(t / g.ny);
  double dg[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) dg[d] = 0.0;
#pragma unroll
  for (int e = 0; e < NN; ++e) {
    const int ex = e & 1, ey = (e >> 1) & 1, ez = DIM == 3 ? (e >> 2) & 1 : 1;
    const int ei = i - 1 + ex, ej = j - 1 + ey, ek = DIM == 3 ? k - 1 + ez : 0;
    if (ei < 0 || ei >= g.ex || ej < 0 || ej >= g.ey || ek < 0 || ek >= g.ez) continue;
    const double xv = xe[((long long)ek * g.ey + ej) * g.ex + ei];
    const int loc = (1 - ex) + 2 * (1 - ey) + (DIM == 3 ? 4 * (1 - ez) : 0);
#pragma unroll
    for (int d = 0; d < DIM; ++d) dg[d] = fma(xv, K0[(loc * DIM + d) * NE + loc * DIM + d], dg[d]);
  }
  const unsigned mk = bcm ? bcm[n] : 0u;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const double v = ((mk >> d) & 1u) ? bcdiag : dg[d];
    const double w = invert ? 1.0 / v : v;
    diag[n * DIM + d] = w;
    if (diag32) diag32[n * DIM + d] = (float)w;  // the FP32 copy a reduced-precision V-cycle uses
  }
}

template <int DIM, int MODE, typename T>
__global__ void __launch_bounds__(256)
stencil_node_kernel(GridDims g, const T* __restrict__ S, const T* __restrict__ u,
                    const T* __restrict__ b, const T* __restrict__ dinv, T omega,
                    T* __restrict__ y) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nnodes) return;
  const int i = (int)(n % g.nx);
  const long long t = n / g.nx;
  const int j = (int)(t % g.ny);
  const int k = (int)(t / g.ny);
  const long long nn = g.nnodes;
  T out[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) out[d] = 0.0;
  constexpr int ZLO = DIM == 3 ? -1 : 0, ZHI = DIM == 3 ? 1 : 0;
  int o = 0;
#pragma unroll
  for (int dz = ZLO; dz <= ZHI; ++dz) {
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx, ++o) {
        const int ii = i + dx, jj = j + dy, kk = k + dz;
        if (ii < 0 || ii >= g.nx || jj < 0 || jj >= g.ny || kk < 0 || kk >= g.nz) continue;
        const long long m = n + dx + dy * (long long)g.nx + dz * (long long)g.nx * g.ny;
        T um[DIM];
#pragma unroll
        for (int d2 = 0; d2 < DIM; ++d2) um[d2] = u[m * DIM + d2];
#pragma unroll
        for (int d = 0; d < DIM; ++d)
#pragma unroll
          for (int d2 = 0; d2 < DIM; ++d2)
            out[d] = fma(S[((long long)(o * DIM + d) * DIM + d2) * nn + n], um[d2], out[d]);
      }
    }
  }
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const long long q = n * DIM + d;
    if (MODE == MODE_APPLY) {
      y[q] = out[d];
    } else if (MODE == MODE_RESIDUAL) {
      y[q] = b[q] - out[d];
    } else {
      y[q] = u[q] + omega * dinv[q] * (b[q] - out[d]);
    }
  }
}

// Half stencil: the Galerkin operators are symmetric, A(n, n+o) = A(n+o, n)^T, so only the blocks of
// the centre and the forward offsets o > centre are stored per node, H[c][node] with
// c = ((o - centre)*D + d)*D + d2 (126 coefficients per node in 3D instead of 243, 20 in 2D).  The
// block towards a backward neighbour m = n - off(o) is the transpose of m's forward block o; for a
// wave of 64 consecutive nodes those reads are 64 consecutive entries of the same coefficient row
// (coalesced) that the neighbour rows of the same sweep have just brought into L2 / the last-level
// cache, so HBM carries each stored coefficient about once.
template <int DIM>
struct Half {
  static constexpr int C = (Cfg<DIM>::NS - 1) / 2;   // centre offset index
  static constexpr int NB = Cfg<DIM>::NS - C;        // stored blocks
  static constexpr int NC = NB * DIM * DIM;          // stored coefficients per node
};

template <int DIM, typename T>
__global__ void stencil_compress_kernel(GridDims g, const double* __restrict__ S, T* __restrict__ H) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nnodes) return;
  constexpr int C = Half<DIM>::C;
#pragma unroll
  for (int c = 0; c < Half<DIM>::NC; ++c)
    H[(long long)c * g.nnodes + n] = (T)S[((long long)C * DIM * DIM + c) * g.nnodes + n];
}

// Loads go through buffer descriptors (one per coefficient row, one per vector): a neighbour outside
// the grid uses offset kOOB and reads 0, so every lane issues the same loads with no branch around
// them (the bounds-checked version waited on each neighbour).  The neighbours are processed in rows
// of three (fixed dz, dy); the loads of the next row are issued before the current row is multiplied.
template <int DIM, int MODE, typename T>
__global__ void __launch_bounds__(256)
stencil_half_kernel(GridDims g, const T* __restrict__ H, const T* __restrict__ u, const T* __restrict__ b,
                    const T* __restrict__ dinv, T omega, T* y, const T* up, T beta) {
  const long long n = (long long)xcd_block() * blockDim.x + threadIdx.x;
  if (n >= g.nnodes) return;
  const int i = (int)(n % g.nx);
  const long long t = n / g.nx;
  const int j = (int)(t % g.ny);
  const int k = (int)(t / g.ny);
  const long long nn = g.nnodes;
  constexpr int TS = (int)sizeof(T);
  constexpr int C = Half<DIM>::C;
  constexpr int NROW = DIM == 3 ? 9 : 3;  // rows of three neighbours
  constexpr int NV = DIM + DIM * DIM;     // values per neighbour: its dofs and one block
  const __amdgpu_buffer_rsrc_t ru = plane_rsrc(u, 0, nn * DIM, true);

  // loads of row r (dz = r / 3 - 1, dy = r % 3 - 1; DIM == 2: dy = r - 1) into v[3][NV]
  auto load_row = [&](int r, T(&v)[3][NV]) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int dx = q - 1, dy = (DIM == 3 ? r % 3 : r) - 1, dz = DIM == 3 ? r / 3 - 1 : 0;
      const int o = r * 3 + q;
      const int ii = i + dx, jj = j + dy, kk = k + dz;
      const bool in = ii >= 0 && ii < g.nx && jj >= 0 && jj < g.ny && kk >= 0 && kk < g.nz;
      const long long m = n + dx + dy * (long long)g.nx + dz * (long long)g.nx * g.ny;
      const int om = in ? (int)m * TS : kOOB;          // byte offset of node m in a coefficient row
      const int ou = in ? (int)(m * DIM) * TS : kOOB;  // byte offset of node m's dofs
#pragma unroll
      for (int d2 = 0; d2 < DIM; ++d2) v[q][d2] = bload(ru, ou + d2 * TS, T());
#pragma unroll
      for (int d = 0; d < DIM; ++d)
#pragma unroll
        for (int d2 = 0; d2 < DIM; ++d2) {
          if (o >= C) {  // stored at n: block (n, n+o), entry (d, d2)
            const __amdgpu_buffer_rsrc_t rh = plane_rsrc(H, (long long)(((o - C) * DIM + d) * DIM + d2) * nn, nn, true);
            v[q][DIM + d * DIM + d2] = bload(rh, (int)n * TS, T());
          } else {  // backward: (n, m) = (m, m + o')^T with o' the mirrored offset, stored at m
            const __amdgpu_buffer_rsrc_t rh =
                plane_rsrc(H, (long long)((((2 * C - o) - C) * DIM + d2) * DIM + d) * nn, nn, true);
            v[q][DIM + d * DIM + d2] = bload(rh, om, T());
          }
        }
    }
  };
  T out[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) out[d] = 0.0;
  auto mul_row = [&](const T(&v)[3][NV]) {
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int d = 0; d < DIM; ++d)
#pragma unroll
        for (int d2 = 0; d2 < DIM; ++d2) out[d] = fma(v[q][DIM + d * DIM + d2], v[q][d2], out[d]);
  };
  T va[3][NV], vb[3][NV];
  load_row(0, va);
#pragma unroll
  for (int r = 0; r < NROW; r += 2) {
    if (r + 1 < NROW) load_row(r + 1, vb);
    mul_row(va);
    if (r + 1 < NROW) {
      if (r + 2 < NROW) load_row(r + 2, va);
      mul_row(vb);
    }
  }
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const long long q = n * DIM + d;
    if (MODE == MODE_APPLY) {
      y[q] = out[d];
    } else if (MODE == MODE_RESIDUAL) {
      y[q] = b[q] - out[d];
    } else if (MODE == MODE_JACOBI) {
      y[q] = u[q] + omega * dinv[q] * (b[q] - out[d]);
    } else {
      const T uq = u[q], pq = up ? up[q] : (T)0;
      y[q] = uq + omega * dinv[q] * (b[q] - out[d]) + beta * (uq - pq);
    }
  }
}

template <int DIM, typename T>
__global__ void stencil_diag_kernel(GridDims g, const double* __restrict__ S, T* __restrict__ dinv) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.nnodes) return;
  constexpr int center = Cfg<DIM>::NS / 2;
#pragma unroll
  for (int d = 0; d < DIM; ++d) dinv[n * DIM + d] = (T)(1.0 / S[((long long)(center * DIM + d) * DIM + d) * g.nnodes + n]);
}

// first damped-Jacobi sweep from zero: u = w D^-1 b (iterative.py:232-233)
template <typename T>
__global__ void jacobi0_kernel(long long n, const T* __restrict__ dinv, const T* __restrict__ b,
                               T omega, T* __restrict__ u) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) u[q] = omega * dinv[q] * b[q];
}

// rc = R^T rf : coarse node gathers the 3^D fine nodes around its fine image (weights 1, 1/2, 1/4, 1/8).
// u0 (optional): also the coarse level's first smoothing sweep from zero, u0 = w0 D^-1 rc (the
// expression of jacobi0_kernel, so the bits are the same as a separate launch)
template <int DIM, typename T>
__global__ void restrict_kernel(GridDims gf, GridDims gc, const T* __restrict__ rf, T* __restrict__ rc,
                                const T* __restrict__ dinv, T w0, T* __restrict__ u0) {
  const long long N = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (N >= gc.nnodes) return;
  const int I = (int)(N % gc.nx);
  const long long t = N / gc.nx;
  const int J = (int)(t % gc.ny);
  const int K = (int)(t / gc.ny);
  T acc[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) acc[d] = 0;
  constexpr int ZLO = DIM == 3 ? -1 : 0, ZHI = DIM == 3 ? 1 : 0;
#pragma unroll
  for (int dz = ZLO; dz <= ZHI; ++dz)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        const int fi = 2 * I + dx, fj = 2 * J + dy, fk = 2 * K + dz;
        if (fi < 0 || fi >= gf.nx || fj < 0 || fj >= gf.ny || fk < 0 || fk >= gf.nz) continue;
        const T w = (T)((dx ? 0.5 : 1.0) * (dy ? 0.5 : 1.0) * (dz ? 0.5 : 1.0));
        const long long m = ((long long)fk * gf.ny + fj) * gf.nx + fi;
#pragma unroll
        for (int d = 0; d < DIM; ++d) acc[d] = fma(w, rf[m * DIM + d], acc[d]);
      }
#pragma unroll
  for (int d = 0; d < DIM; ++d) rc[N * DIM + d] = acc[d];
  if (u0) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) u0[N * DIM + d] = w0 * dinv[N * DIM + d] * acc[d];
  }
}

// uf += R uc : fine node interpolates from the 1, 2, 4 or 8 coarse nodes around it.  Branch free: every
// lane reads 2^D corners, the second one along an axis of even fine index repeating the first with
// weight 0 (the weights are powers of two, so the sum equals w * (sum of the nonzero corners) exactly).
template <int DIM, typename T>
__global__ void prolong_kernel(GridDims gf, GridDims gc, const T* __restrict__ uc, T* __restrict__ uf) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= gf.nnodes) return;
  const int i = (int)(n % gf.nx);
  const long long t = n / gf.nx;
  const int j = (int)(t % gf.ny);
  const int k = (int)(t / gf.ny);
  const int oi = i & 1, oj = j & 1, ok = DIM == 3 ? (k & 1) : 0;
  const T wx[2] = {(T)(oi ? 0.5 : 1.0), (T)(oi ? 0.5 : 0.0)};
  const T wy[2] = {(T)(oj ? 0.5 : 1.0), (T)(oj ? 0.5 : 0.0)};
  const T wz[2] = {(T)(ok ? 0.5 : 1.0), (T)(ok ? 0.5 : 0.0)};
  T acc[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) acc[d] = 0;
#pragma unroll
  for (int c = 0; c < (DIM == 3 ? 2 : 1); ++c)
#pragma unroll
    for (int bb = 0; bb < 2; ++bb)
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const T w = wx[a] * wy[bb] * wz[c];
        const long long N =
            ((long long)((k >> 1) + c * ok) * gc.ny + ((j >> 1) + bb * oj)) * gc.nx + ((i >> 1) + a * oi);
#pragma unroll
        for (int d = 0; d < DIM; ++d) acc[d] = fma(w, uc[N * DIM + d], acc[d]);
      }
#pragma unroll
  for (int d = 0; d < DIM; ++d) uf[n * DIM + d] += acc[d];
}

// dx_e (+)= alpha * u_e^T K0 v_e with Dirichlet entries zeroed (element-centric, no scatter)
template <int DIM>
__global__ void __launch_bounds__(256)
sens_kernel(GridDims g, const double* __restrict__ K0, const uint8_t* __restrict__ bcm,
            const double* __restrict__ u, const double* __restrict__ v, double alpha,
            double* __restrict__ dx, int accumulate) {
  constexpr int NN = Cfg<DIM>::NN;
  constexpr int NE = Cfg<DIM>::NE;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.nel) return;
  const int ei = (int)(e % g.ex);
  const long long t = e / g.ex;
  const int ej = (int)(t % g.ey);
  const int ek = (int)(t / g.ey);
  double ue[NE], ve[NE];
#pragma unroll
  for (int a = 0; a < NN; ++a) {
    const long long m = ((long long)(ek + ((a >> 2) & 1)) * g.ny + (ej + ((a >> 1) & 1))) * g.nx + (ei + (a & 1));
    const unsigned mk = bcm ? bcm[m] : 0u;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const bool fixed = (mk >> d) & 1u;
      ue[a * DIM + d] = fixed ? 0.0 : u[m * DIM + d];
      ve[a * DIM + d] = fixed ? 0.0 : v[m * DIM + d];
    }
  }
  double s = 0.0;
#pragma unroll
  for (int r = 0; r < NE; ++r) {
    double kv = 0.0;
#pragma unroll
    for (int c = 0; c < NE; ++c) kv = fma(K0[r * NE + c], ve[c], kv);
    s = fma(ue[r], kv, s);
  }
  s *= alpha;
  dx[e] = accumulate ? dx[e] + s : s;
}

// ---------------------------------------------------------------------------------------------
// host launchers (shared with the multigrid runtime)
// ---------------------------------------------------------------------------------------------
template <int DIM, bool CLS, typename T>
static void launch_fine_t(int mode, const GridDims& g, const double* K0, const T* xe, const uint8_t* bcm, T bcdiag,
                          const T* u, const T* b, const T* dinv, T omega, T* y, hipStream_t s, const T* up, T beta) {
  const int tx = (g.nx + Tile<DIM>::X - 1) / Tile<DIM>::X;
  const int ty = (g.ny + Tile<DIM>::Y - 1) / Tile<DIM>::Y;
  const int tz = DIM == 3 ? (g.nz + Tile<DIM>::Z - 1) / Tile<DIM>::Z : 1;
  const int nt = tx * ty * tz;
  dim3 grid((unsigned)((nt + 7) & ~7));
#define PMT_FINE_T(M) \
  fine_node_kernel<DIM, M, CLS, T><<<grid, 256, 0, s>>>(g, tx, ty, nt, K0, xe, bcm, bcdiag, u, b, dinv, omega, y, up, beta)
  if (mode == MODE_APPLY)
    PMT_FINE_T(MODE_APPLY);
  else if (mode == MODE_RESIDUAL)
    PMT_FINE_T(MODE_RESIDUAL);
  else if (mode == MODE_JACOBI)
    PMT_FINE_T(MODE_JACOBI);
  else
    PMT_FINE_T(MODE_CHEB);
#undef PMT_FINE_T
}

extern int g_filter_tiled;  // filter.hip

// Launch-shape tuning of the fine operator (set by pmt_set_tuning; read at every launch).
static int g_fine_kernel = 2;  // 0 = 3D tiles, 1 = z-streamed columns, 2 = reflection basis (if available)
static int g_refl_rows = 16;   // block shape of the reflection-basis kernel (node rows per block; 0 = by grid)
static int g_coarse1 = 1;      // hex8 level 1 without a stencil (reflection basis), new hierarchies
static int g_stream_nzc = 0;   // node planes per streamed block (0 = chosen from the CU count)
static int g_stream_tile = 0;  // streamed column shape: 0 = 32x8, 1 = 64x4, 2 = 16x16 nodes

static int g_tuning_epoch = 0;  // bumped by every change: captured launches must be re-captured

int tuning_epoch() { return g_tuning_epoch; }

int set_tuning(int param, int value) {
  if (param == PMT_TUNE_FINE_KERNEL && value >= 0 && value <= 2) {
    g_fine_kernel = value;
    ++g_tuning_epoch;
    return 0;
  }
  if (param == PMT_TUNE_STREAM_PLANES && value >= 0) {
    g_stream_nzc = value;
    ++g_tuning_epoch;
    return 0;
  }
  if (param == PMT_TUNE_STREAM_TILE && value >= 0 && value <= 2) {
    g_stream_tile = value;
    ++g_tuning_epoch;
    return 0;
  }
  if (param == PMT_TUNE_REFL_ROWS && (value == 0 || value == 12 || value == 16 || value == 24 || value == 32)) {
    g_refl_rows = value;
    ++g_tuning_epoch;
    return 0;
  }
  if (param == PMT_TUNE_COARSE1 && (value == 0 || value == 1)) {
    g_coarse1 = value;
    ++g_tuning_epoch;
    return 0;
  }
  if (param == PMT_TUNE_FILTER && (value == 0 || value == 1)) {
    g_filter_tiled = value;
    ++g_tuning_epoch;
    return 0;
  }
  return -1;
}

int tuned_stream_planes() { return g_stream_nzc; }

int get_tuning(int param) {
  if (param == PMT_TUNE_FINE_KERNEL) return g_fine_kernel;
  if (param == PMT_TUNE_STREAM_PLANES) return g_stream_nzc;
  if (param == PMT_TUNE_STREAM_TILE) return g_stream_tile;
  if (param == PMT_TUNE_REFL_ROWS) return g_refl_rows;
  if (param == PMT_TUNE_COARSE1) return g_coarse1;
  if (param == PMT_TUNE_FILTER) return g_filter_tiled;
  return -1;
}

int device_cu_count() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cus[dev] == 0) {
    int c = 0;
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) c = 256;
    cus[dev] = c;
  }
  return cus[dev];
}

// Planes per streamed block: minimise (rounds of resident blocks) x (planes + 2 prologue planes),
// assuming four resident 256-thread blocks per CU.
static int stream_planes(int ntiles_xy, int nz) {
  if (g_stream_nzc > 0) return g_stream_nzc < nz ? g_stream_nzc : nz;
  const long long slots = 4LL * device_cu_count();
  int best = nz;
  long long best_cost = -1;
  for (int c = 2; c <= nz; ++c) {
    const long long blocks = (long long)ntiles_xy * ((nz + c - 1) / c);
    const long long cost = ((blocks + slots - 1) / slots) * (c + 2);
    if (best_cost < 0 || cost < best_cost || (cost == best_cost && c < best)) {
      best_cost = cost;
      best = c;
    }
  }
  return best;
}

template <class Stream, bool CLS, typename T>
static void launch_stream_shape(int mode, const GridDims& g, const double* K0, const T* xe, const uint8_t* bcm,
                                T bcdiag, const T* u, const T* b, const T* dinv, T omega, T* y, hipStream_t s,
                                const T* up, T beta) {
  const int tx = (g.nx + Stream::SX - 1) / Stream::SX;
  const int ty = (g.ny + Stream::SY - 1) / Stream::SY;
  const int nxy = tx * ty;
  const int nzc = stream_planes(nxy, g.nz);
  const int nwork = nxy * ((g.nz + nzc - 1) / nzc);
  dim3 grid((unsigned)((nwork + 7) & ~7));
#define PMT_STREAM(M)                                                                                             \
  fine_stream_kernel<Stream, M, CLS, T><<<grid, 256, 0, s>>>(g, tx, nxy, nzc, nwork, K0, xe, bcm, bcdiag, u, b, \
                                                             dinv, omega, y, up, beta)
  if (mode == MODE_APPLY)
    PMT_STREAM(MODE_APPLY);
  else if (mode == MODE_RESIDUAL)
    PMT_STREAM(MODE_RESIDUAL);
  else if (mode == MODE_JACOBI)
    PMT_STREAM(MODE_JACOBI);
  else
    PMT_STREAM(MODE_CHEB);
#undef PMT_STREAM
}

template <bool CLS, typename T>
static void launch_stream_t(int mode, const GridDims& g, const double* K0, const T* xe, const uint8_t* bcm, T bcdiag,
                            const T* u, const T* b, const T* dinv, T omega, T* y, hipStream_t s, const T* up, T beta) {
  if (g_stream_tile == 1)
    launch_stream_shape<StreamTile<64, 4>, CLS, T>(mode, g, K0, xe, bcm, bcdiag, u, b, dinv, omega, y, s, up, beta);
  else if (g_stream_tile == 2)
    launch_stream_shape<StreamTile<16, 16>, CLS, T>(mode, g, K0, xe, bcm, bcdiag, u, b, dinv, omega, y, s, up, beta);
  else
    launch_stream_shape<StreamTile<32, 8>, CLS, T>(mode, g, K0, xe, bcm, bcdiag, u, b, dinv, omega, y, s, up, beta);
}

template <typename T>
int launch_fine(int mode, const GridDims& g, const double* K0, int k0_pattern, const T* xe, const uint8_t* bcm,
                double bcdiag, const T* u, const T* b, const T* dinv, double omega, T* y, hipStream_t s, const T* up,
                double beta) {
  const T bd = (T)bcdiag, om = (T)omega, be = (T)beta;
  const bool cls = (k0_pattern & PMT_K0_CLASS_PATTERN) != 0;
  if (g.dim == 3 && g_fine_kernel == 2 && (k0_pattern & PMT_K0_REFLECTION)) {
    return launch_fine_refl<T>(mode, g, g_stream_nzc, g_refl_rows, refl_pattern(k0_pattern), K0, xe, bcm, bcdiag, u,
                               b, dinv, omega, y, s, up, beta);
  }
  if (g.dim == 3 && g_fine_kernel >= 1) {
    if (cls)
      launch_stream_t<true, T>(mode, g, K0, xe, bcm, bd, u, b, dinv, om, y, s, up, be);
    else
      launch_stream_t<false, T>(mode, g, K0, xe, bcm, bd, u, b, dinv, om, y, s, up, be);
  } else if (g.dim == 3) {
    if (cls)
      launch_fine_t<3, true, T>(mode, g, K0, xe, bcm, bd, u, b, dinv, om, y, s, up, be);
    else
      launch_fine_t<3, false, T>(mode, g, K0, xe, bcm, bd, u, b, dinv, om, y, s, up, be);
  } else {
    if (cls)
      launch_fine_t<2, true, T>(mode, g, K0, xe, bcm, bd, u, b, dinv, om, y, s, up, be);
    else
      launch_fine_t<2, false, T>(mode, g, K0, xe, bcm, bd, u, b, dinv, om, y, s, up, be);
  }
  return check_launch("fine_node_kernel");
}

// Does a host element matrix follow the isotropic box pattern (so the SGPR kernel applies)?
int k0_matches_pattern(int dim, const double* K0) {
  const int ne = (1 << dim) * dim;
  const int8_t* cls = dim == 3 ? K0CLS3 : K0CLS2;
  const int8_t* sgn = dim == 3 ? K0SGN3 : K0SGN2;
  const int16_t* first = dim == 3 ? K0FIRST3 : K0FIRST2;
  double scale = 0.0;
  for (int i = 0; i < ne * ne; ++i) scale = fmax(scale, fabs(K0[i]));
  if (!(scale > 0.0)) return 0;
  for (int i = 0; i < ne * ne; ++i) {
    const double v = sgn[first[cls[i]]] * K0[first[cls[i]]] * sgn[i];
    if (fabs(v - K0[i]) > 1e-14 * scale) return 0;
  }
  return 1;
}

int launch_fine_diag(const GridDims& g, const double* K0, const double* xe, const uint8_t* bcm, double bcdiag,
                     double* diag, int invert, hipStream_t s, float* diag32) {
  dim3 grid(blocks_for(g.nnodes, 256));
  if (g.dim == 3)
    fine_diag_kernel<3><<<grid, 256, 0, s>>>(g, K0, xe, bcm, bcdiag, diag, invert, diag32);
  else
    fine_diag_kernel<2><<<grid, 256, 0, s>>>(g, K0, xe, bcm, bcdiag, diag, invert, diag32);
  return check_launch("fine_diag_kernel");
}

int half_stencil_coefficients(int dim) { return dim == 3 ? Half<3>::NC : Half<2>::NC; }

template <typename T>
int launch_stencil_compress(const GridDims& g, const double* S, T* H, hipStream_t s) {
  dim3 grid(blocks_for(g.nnodes, 256));
  if (g.dim == 3)
    stencil_compress_kernel<3, T><<<grid, 256, 0, s>>>(g, S, H);
  else
    stencil_compress_kernel<2, T><<<grid, 256, 0, s>>>(g, S, H);
  return check_launch("stencil_compress_kernel");
}

template <typename T>
int launch_stencil_half(int mode, const GridDims& g, const T* H, const T* u, const T* b, const T* dinv, double omega,
                        T* y, hipStream_t s, const T* up, double beta) {
  dim3 grid(xcd_blocks(g.nnodes, 256));
  const T om = (T)omega, be = (T)beta;
#define PMT_HALF(D, M) stencil_half_kernel<D, M, T><<<grid, 256, 0, s>>>(g, H, u, b, dinv, om, y, up, be)
  if (g.dim == 3) {
    if (mode == MODE_APPLY)
      PMT_HALF(3, MODE_APPLY);
    else if (mode == MODE_RESIDUAL)
      PMT_HALF(3, MODE_RESIDUAL);
    else if (mode == MODE_JACOBI)
      PMT_HALF(3, MODE_JACOBI);
    else
      PMT_HALF(3, MODE_CHEB);
  } else {
    if (mode == MODE_APPLY)
      PMT_HALF(2, MODE_APPLY);
    else if (mode == MODE_RESIDUAL)
      PMT_HALF(2, MODE_RESIDUAL);
    else if (mode == MODE_JACOBI)
      PMT_HALF(2, MODE_JACOBI);
    else
      PMT_HALF(2, MODE_CHEB);
  }
#undef PMT_HALF
  return check_launch("stencil_half_kernel");
}

template <typename T>
int launch_stencil(int mode, const GridDims& g, const T* S, const T* u, const T* b, const T* dinv, double omega, T* y,
                   hipStream_t s) {
  dim3 grid(blocks_for(g.nnodes, 256));
  const T om = (T)omega;
  if (g.dim == 3) {
    if (mode == MODE_APPLY)
      stencil_node_kernel<3, MODE_APPLY, T><<<grid, 256, 0, s>>>(g, S, u, b, dinv, om, y);
    else if (mode == MODE_RESIDUAL)
      stencil_node_kernel<3, MODE_RESIDUAL, T><<<grid, 256, 0, s>>>(g, S, u, b, dinv, om, y);
    else
      stencil_node_kernel<3, MODE_JACOBI, T><<<grid, 256, 0, s>>>(g, S, u, b, dinv, om, y);
  } else {
    if (mode == MODE_APPLY)
      stencil_node_kernel<2, MODE_APPLY, T><<<grid, 256, 0, s>>>(g, S, u, b, dinv, om, y);
    else if (mode == MODE_RESIDUAL)
      stencil_node_kernel<2, MODE_RESIDUAL, T><<<grid, 256, 0, s>>>(g, S, u, b, dinv, om, y);
    else
      stencil_node_kernel<2, MODE_JACOBI, T><<<grid, 256, 0, s>>>(g, S, u, b, dinv, om, y);
  }
  return check_launch("stencil_node_kernel");
}

template <typename T>
int launch_stencil_dinv(const GridDims& g, const double* S, T* dinv, hipStream_t s) {
  dim3 grid(blocks_for(g.nnodes, 256));
  if (g.dim == 3)
    stencil_diag_kernel<3, T><<<grid, 256, 0, s>>>(g, S, dinv);
  else
    stencil_diag_kernel<2, T><<<grid, 256, 0, s>>>(g, S, dinv);
  return check_launch("stencil_diag_kernel");
}

template <typename T>
int launch_jacobi0(long long n, const T* dinv, const T* b, double omega, T* u, hipStream_t s) {
  jacobi0_kernel<T><<<blocks_for(n, 256), 256, 0, s>>>(n, dinv, b, (T)omega, u);
  return check_launch("jacobi0_kernel");
}

template <typename T>
int launch_restrict(const GridDims& gf, const GridDims& gc, const T* rf, T* rc, hipStream_t s, const T* dinv,
                    double w0, T* u0) {
  dim3 grid(blocks_for(gc.nnodes, 256));
  if (gf.dim == 3)
    restrict_kernel<3, T><<<grid, 256, 0, s>>>(gf, gc, rf, rc, dinv, (T)w0, u0);
  else
    restrict_kernel<2, T><<<grid, 256, 0, s>>>(gf, gc, rf, rc, dinv, (T)w0, u0);
  return check_launch("restrict_kernel");
}

template <typename T>
int launch_prolong(const GridDims& gf, const GridDims& gc, const T* uc, T* uf, hipStream_t s) {
  dim3 grid(blocks_for(gf.nnodes, 256));
  if (gf.dim == 3)
    prolong_kernel<3, T><<<grid, 256, 0, s>>>(gf, gc, uc, uf);
  else
    prolong_kernel<2, T><<<grid, 256, 0, s>>>(gf, gc, uc, uf);
  return check_launch("prolong_kernel");
}

#define PMT_INSTANTIATE_LEVEL(T)                                                                                   \
  template int launch_fine<T>(int, const GridDims&, const double*, int, const T*, const uint8_t*, double, const T*, \
                              const T*, const T*, double, T*, hipStream_t, const T*, double);                      \
  template int launch_stencil<T>(int, const GridDims&, const T*, const T*, const T*, const T*, double, T*,         \
                                 hipStream_t);                                                                     \
  template int launch_stencil_dinv<T>(const GridDims&, const double*, T*, hipStream_t);                            \
  template int launch_stencil_compress<T>(const GridDims&, const double*, T*, hipStream_t);                        \
  template int launch_stencil_half<T>(int, const GridDims&, const T*, const T*, const T*, const T*, double, T*,    \
                                      hipStream_t, const T*, double);                                              \
  template int launch_jacobi0<T>(long long, const T*, const T*, double, T*, hipStream_t);                          \
  template int launch_restrict<T>(const GridDims&, const GridDims&, const T*, T*, hipStream_t, const T*, double, T*); \
  template int launch_prolong<T>(const GridDims&, const GridDims&, const T*, T*, hipStream_t);
PMT_INSTANTIATE_LEVEL(double)
PMT_INSTANTIATE_LEVEL(float)
#undef PMT_INSTANTIATE_LEVEL

static bool even_grid(int nelx, int nely, int nelz) {
  return nelx % 2 == 0 && nely % 2 == 0 && nelz % 2 == 0 && nelx >= 2 && nely >= 2;
}

}  // namespace pmt

using namespace pmt;

extern "C" {

int pmt_k0_pattern(int dim, const double* K0_host) {
  if ((dim != 2 && dim != 3) || !K0_host) return arg_error("pmt_k0_pattern: bad argument");
  return k0_matches_pattern(dim, K0_host);
}

int pmt_stiffness_apply_f32(int nelx, int nely, int nelz, const double* K0, int k0_pattern, const float* xe,
                            const uint8_t* bcmask, double bcdiag, const float* u, float* y, void* stream) {
  if (!grid_ok(nelx, nely, nelz) || !K0 || !xe || !bcmask || !u || !y)
    return arg_error("pmt_stiffness_apply_f32: bad argument (the Dirichlet mask is mandatory)");
  return launch_fine<float>(MODE_APPLY, make_grid(nelx, nely, nelz), K0, k0_pattern, xe, bcmask, bcdiag, u, nullptr,
                            nullptr, 0.0, y, (hipStream_t)stream);
}

int pmt_stiffness_apply_bf16(int nelx, int nely, int nelz, const double* K0, const float* xe, const uint8_t* bcmask,
                             double bcdiag, const float* u, float* y, void* stream) {
  if (!grid_ok(nelx, nely, nelz) || nelz <= 0 || !K0 || !xe || !bcmask || !u || !y)
    return arg_error("pmt_stiffness_apply_bf16: bad argument (hex8 grids only; the Dirichlet mask is mandatory)");
  return launch_fine_mfma(MODE_APPLY, make_grid(nelx, nely, nelz), K0, xe, bcmask, bcdiag, u, nullptr, nullptr, 0.0, y,
                          (hipStream_t)stream);
}

int pmt_set_tuning(int param, int value) {
  if (set_tuning(param, value) != 0) return arg_error("pmt_set_tuning: unknown parameter or value out of range");
  return PMT_OK;
}

int pmt_get_tuning(int param) {
  const int v = get_tuning(param);
  return v < 0 ? arg_error("pmt_get_tuning: unknown parameter") : v;
}

int pmt_stiffness_apply(int nelx, int nely, int nelz, const double* K0, int k0_pattern, const double* xe,
                        const uint8_t* bcmask, double bcdiag, const double* u, double* y, void* stream) {
  if (!grid_ok(nelx, nely, nelz) || !K0 || !xe || !bcmask || !u || !y)
    return arg_error("pmt_stiffness_apply: bad argument (the Dirichlet mask is mandatory)");
  return launch_fine<double>(MODE_APPLY, make_grid(nelx, nely, nelz), K0, k0_pattern, xe, bcmask, bcdiag, u, nullptr, nullptr,
                     0.0, y, (hipStream_t)stream);
}

int pmt_stiffness_residual(int nelx, int nely, int nelz, const double* K0, int k0_pattern, const double* xe,
                           const uint8_t* bcmask, double bcdiag, const double* b, const double* u, double* r,
                           void* stream) {
  if (!grid_ok(nelx, nely, nelz) || !K0 || !xe || !bcmask || !u || !b || !r)
    return arg_error("pmt_stiffness_residual: bad argument (the Dirichlet mask is mandatory)");
  return launch_fine<double>(MODE_RESIDUAL, make_grid(nelx, nely, nelz), K0, k0_pattern, xe, bcmask, bcdiag, u, b, nullptr,
                     0.0, r, (hipStream_t)stream);
}

int pmt_stiffness_diagonal(int nelx, int nely, int nelz, const double* K0, const double* xe, const uint8_t* bcmask,
                           double bcdiag, double* diag, void* stream) {
  if (!grid_ok(nelx, nely, nelz) || !K0 || !xe || !diag) return arg_error("pmt_stiffness_diagonal: bad argument");
  return launch_fine_diag(make_grid(nelx, nely, nelz), K0, xe, bcmask, bcdiag, diag, 0, (hipStream_t)stream);
}

int pmt_sens_contract(int nelx, int nely, int nelz, const double* K0, const uint8_t* bcmask, const double* u,
                      const double* v, double alpha, double* dx, int accumulate, void* stream) {
  if (!grid_ok(nelx, nely, nelz) || !K0 || !u || !v || !dx) return arg_error("pmt_sens_contract: bad argument");
  GridDims g = make_grid(nelx, nely, nelz);
  dim3 grid(blocks_for(g.nel, 256));
  if (g.dim == 3)
    sens_kernel<3><<<grid, 256, 0, (hipStream_t)stream>>>(g, K0, bcmask, u, v, alpha, dx, accumulate);
  else
    sens_kernel<2><<<grid, 256, 0, (hipStream_t)stream>>>(g, K0, bcmask, u, v, alpha, dx, accumulate);
  return check_launch("sens_kernel");
}

int pmt_sens_contract_k0(int nelx, int nely, int nelz, const double* K0, int k0_pattern, const uint8_t* bcmask,
                         const double* u, const double* v, double alpha, double* dx, int accumulate, void* stream) {
  if (!grid_ok(nelx, nely, nelz) || !K0 || !u || !v || !dx) return arg_error("pmt_sens_contract: bad argument");
  GridDims g = make_grid(nelx, nely, nelz);
  if (g.dim != 3 || !(k0_pattern & PMT_K0_REFLECTION))
    return pmt_sens_contract(nelx, nely, nelz, K0, bcmask, u, v, alpha, dx, accumulate, stream);
  const double* tab = K0 + kK0ReflOffset64;  // the FP64 reflection-basis blocks behind K0 (pmt_k0_extend layout)
  return launch_sens_stream(g, tab, bcmask, u, v, alpha, dx, accumulate, (hipStream_t)stream);
}

int pmt_restrict(int nelx, int nely, int nelz, const double* rf, double* rc, void* stream) {
  if (!grid_ok(nelx, nely, nelz) || !even_grid(nelx, nely, nelz) || !rf || !rc)
    return arg_error("pmt_restrict: grid must have even element counts");
  return launch_restrict(make_grid(nelx, nely, nelz), make_grid(nelx / 2, nely / 2, nelz / 2), rf, rc,
                         (hipStream_t)stream);
}

int pmt_prolong_add(int nelx, int nely, int nelz, const double* uc, double* uf, void* stream) {
  if (!grid_ok(nelx, nely, nelz) || !even_grid(nelx, nely, nelz) || !uc || !uf)
    return arg_error("pmt_prolong_add: grid must have even element counts");
  return launch_prolong(make_grid(nelx, nely, nelz), make_grid(nelx / 2, nely / 2, nelz / 2), uc, uf,
                        (hipStream_t)stream);
}

}  // extern "C"
