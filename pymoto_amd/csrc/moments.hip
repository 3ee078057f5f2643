// Coarse-grid Galerkin operators of hex8 hierarchies from density moments (no level-1 element matrices).
//
// Composite trilinear prolongation from a level-l element (h = 2^l fine elements per side) to one of
// its fine elements f is again trilinear, so R^T A R = assemble_E sum_{f in E} P_f^T A_f P_f with
// P_f[a][A] = prod_d w(f_d + a_d, A_d), w(p, 0) = 1 - p/h, w(p, 1) = p/h.  For a fine element without
// Dirichlet nodes A_f = x_f K0.  In the centred coordinate s_d = (f_d + 1/2)/h - 1/2 each factor
// w(f+a, A) w(f+b, B) is a quadratic in s_d, hence
//   sum_f x_f P_f^T K0 P_f = sum_{m = (i,j,k) in {0,1,2}^3} M_m(E) C^(l)_m,
//   M_m(E) = sum_{f in E} x_f s_x^i s_y^j s_z^k,
// with 27 fixed 24x24 tables C^(l)_m per level (moment_tables).  Moments of a level-(l+1) element
// follow from its children's by the binomial shift s' = s/2 + t, t = +-1/4 per dimension.
// Fine elements touching Dirichlet nodes are left out of the moments; their masked products
// P_f^T (x_f M K0 M + D_f) P_f are kept per flagged element as a correction dK_E (level 1 directly,
// level l+1 as the Galerkin product of the children's corrections), in a compact list.
// Level-1 stencils are summed from the densities directly (coarse.hip); levels >= 2 materialise their
// (small) element matrices in structure-of-arrays layout and sum those.
#include "launch.h"

namespace pmt {

constexpr int kNM = 27;  // moments s_x^i s_y^j s_z^k, m = i + 3 j + 9 k

void moment_tables(int level, const double* K0, double* C) {
  const double h = (double)(1 << level);
  for (int q = 0; q < 64 * kNM * 9; ++q) C[q] = 0.0;
  auto coefs = [h](int a, int A, int b, int B, double* c) {
    const double sA = A ? 1.0 : -1.0, sB = B ? 1.0 : -1.0;
    const double aa = (2 * a - 1) / (2.0 * h), ab = (2 * b - 1) / (2.0 * h);
    c[0] = 0.25 + (sA * aa + sB * ab) / 2.0 + sA * sB * aa * ab;
    c[1] = (sA + sB) / 2.0 + sA * sB * (aa + ab);
    c[2] = sA * sB;
  };
  for (int A = 0; A < 8; ++A)
    for (int B = 0; B < 8; ++B)
      for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b) {
          double cx[3], cy[3], cz[3];
          coefs(a & 1, A & 1, b & 1, B & 1, cx);
          coefs((a >> 1) & 1, (A >> 1) & 1, (b >> 1) & 1, (B >> 1) & 1, cy);
          coefs((a >> 2) & 1, (A >> 2) & 1, (b >> 2) & 1, (B >> 2) & 1, cz);
          for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 3; ++j)
              for (int i = 0; i < 3; ++i) {
                const double w = cx[i] * cy[j] * cz[k];
                double* c = C + ((A * 8 + B) * kNM + i + 3 * j + 9 * k) * 9;
                for (int d = 0; d < 3; ++d)
                  for (int d2 = 0; d2 < 3; ++d2) c[d * 3 + d2] += w * K0[(a * 3 + d) * 24 + b * 3 + d2];
              }
        }
}

__device__ inline long long child_index(const GridDims& gf, long long E, const GridDims& gc, int c) {
  const int CI = (int)(E % gc.ex);
  const long long tt = E / gc.ex;
  const int CJ = (int)(tt % gc.ey), CK = (int)(tt / gc.ey);
  return ((long long)(2 * CK + ((c >> 2) & 1)) * gf.ey + (2 * CJ + ((c >> 1) & 1))) * gf.ex + 2 * CI + (c & 1);
}

// level-1 moments from the fine densities (flagged children left out); M[m * nel + E]
__global__ void moments1_kernel(GridDims gf, GridDims gc, const double* __restrict__ xe,
                                const uint8_t* __restrict__ cflag, double* __restrict__ M) {
  const long long E = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (E >= gc.nel) return;
  const unsigned fl = cflag[E];
  double m[kNM];
#pragma unroll
  for (int q = 0; q < kNM; ++q) m[q] = 0.0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const double x = ((fl >> c) & 1u) ? 0.0 : xe[child_index(gf, E, gc, c)];
    const double tx = (c & 1) ? 0.25 : -0.25, ty = ((c >> 1) & 1) ? 0.25 : -0.25, tz = ((c >> 2) & 1) ? 0.25 : -0.25;
    const double px[3] = {1.0, tx, tx * tx}, py[3] = {1.0, ty, ty * ty}, pz[3] = {1.0, tz, tz * tz};
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) m[i + 3 * j + 9 * k] = fma(x, px[i] * py[j] * pz[k], m[i + 3 * j + 9 * k]);
  }
#pragma unroll
  for (int q = 0; q < kNM; ++q) M[q * gc.nel + E] = m[q];
}

// level-l moments from the level-(l-1) moments of the 8 children: s' = s/2 + t per dimension
__global__ void moments_up_kernel(GridDims gf, GridDims gc, const double* __restrict__ Mf, double* __restrict__ M) {
  const long long E = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (E >= gc.nel) return;
  double m[kNM];
#pragma unroll
  for (int q = 0; q < kNM; ++q) m[q] = 0.0;
  for (int c = 0; c < 8; ++c) {
    const long long ch = child_index(gf, E, gc, c);
    double v[kNM];
#pragma unroll
    for (int q = 0; q < kNM; ++q) v[q] = Mf[q * gf.nel + ch];
    // (s/2 + t)^i = sum_{i' <= i} binom(i, i') 2^-i' t^(i-i') s^i', one dimension at a time
#pragma unroll
    for (int dim = 0; dim < 3; ++dim) {
      const double t = ((c >> dim) & 1) ? 0.25 : -0.25;
      const int st = dim == 0 ? 1 : (dim == 1 ? 3 : 9);
#pragma unroll
      for (int q = 0; q < kNM; ++q) {
        const int i = (q / st) % 3;
        if (i != 0) continue;  // handle each line (q, q + st, q + 2 st) once
        const double v0 = v[q], v1 = v[q + st], v2 = v[q + 2 * st];
        v[q] = v0;
        v[q + st] = t * v0 + 0.5 * v1;
        v[q + 2 * st] = t * t * v0 + t * v1 + 0.25 * v2;
      }
    }
#pragma unroll
    for (int q = 0; q < kNM; ++q) m[q] += v[q];
  }
#pragma unroll
  for (int q = 0; q < kNM; ++q) M[q * gc.nel + E] = m[q];
}

// flags of level l from level l-1: bit c set when child c has Dirichlet descendants; list and map
__global__ void flags_up_kernel(GridDims gf, GridDims gc, const uint8_t* __restrict__ cflag_f,
                                uint8_t* __restrict__ cflag, int* __restrict__ list, int* __restrict__ count,
                                int* __restrict__ map) {
  const long long E = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (E >= gc.nel) return;
  unsigned bits = 0;
  for (int c = 0; c < 8; ++c)
    if (cflag_f[child_index(gf, E, gc, c)]) bits |= 1u << c;
  cflag[E] = (uint8_t)bits;
  int idx = -1;
  if (bits) {
    idx = atomicAdd(count, 1);
    list[idx] = (int)E;
  }
  map[E] = idx;
}

// dK of a listed level-l element: sum over flagged children c of P_c^T dK_c P_c.  One 64-lane block per
// listed element, lane = (A, B) 3x3 block; per child in two passes through LDS: T[a][B] = sum_b
// w(c, b, B) dK_c[a][b] (lane a = A), then dK[A][B] += sum_a w(c, a, A) T[a][B] - 144 multiply-adds per
// lane and child, no divergent loops.
__global__ void __launch_bounds__(64)
dk_up_kernel(GridDims gf, GridDims gc, const uint8_t* __restrict__ cflag, const int* __restrict__ list,
             const int* __restrict__ map_f, const double* __restrict__ dKf, double* __restrict__ dK) {
  __shared__ double sT[8][8][9];  // [a][B][d * 3 + d2]
  const long long w = blockIdx.x;
  const int lane = threadIdx.x, A = lane >> 3, B = lane & 7;
  const long long E = list[w];
  const unsigned bits = cflag[E];
  double acc[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) acc[q] = 0.0;
  for (int c = 0; c < 8; ++c) {
    if (!((bits >> c) & 1u)) continue;  // uniform over the block
    const double* K = dKf + (long long)map_f[child_index(gf, E, gc, c)] * 576;
    double t[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) t[q] = 0.0;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const double wb = pweight3(c, b, B);
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int d2 = 0; d2 < 3; ++d2) t[d * 3 + d2] = fma(wb, K[(A * 3 + d) * 24 + b * 3 + d2], t[d * 3 + d2]);
    }
    __syncthreads();  // the previous child's reads of sT are done
#pragma unroll
    for (int q = 0; q < 9; ++q) sT[A][B][q] = t[q];
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const double wa = pweight3(c, a, A);
#pragma unroll
      for (int q = 0; q < 9; ++q) acc[q] = fma(wa, sT[a][B][q], acc[q]);
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int d2 = 0; d2 < 3; ++d2) dK[w * 576 + (A * 3 + d) * 24 + B * 3 + d2] = acc[d * 3 + d2];
}


int launch_moments1(const GridDims& gf, const double* xe, const uint8_t* cflag, double* M, hipStream_t s) {
  GridDims gc = make_grid(gf.ex / 2, gf.ey / 2, gf.ez / 2);
  moments1_kernel<<<blocks_for(gc.nel, 256), 256, 0, s>>>(gf, gc, xe, cflag, M);
  return check_launch("moments1_kernel");
}

int launch_moments_up(const GridDims& gf, const double* Mf, double* M, hipStream_t s) {
  GridDims gc = make_grid(gf.ex / 2, gf.ey / 2, gf.ez / 2);
  moments_up_kernel<<<blocks_for(gc.nel, 256), 256, 0, s>>>(gf, gc, Mf, M);
  return check_launch("moments_up_kernel");
}

int launch_flags_up(const GridDims& gf, const uint8_t* cflag_f, uint8_t* cflag, int* list, int* map, hipStream_t s) {
  GridDims gc = make_grid(gf.ex / 2, gf.ey / 2, gf.ez / 2);
  int* count = list + gc.nel;
  zero_int(count, s);
  flags_up_kernel<<<blocks_for(gc.nel, 256), 256, 0, s>>>(gf, gc, cflag_f, cflag, list, count, map);
  return check_launch("flags_up_kernel");
}

int launch_dk_up(const GridDims& gf, const uint8_t* cflag, const int* list, int nlist, const int* map_f,
                 const double* dKf, double* dK, hipStream_t s) {
  if (nlist <= 0) return PMT_OK;
  GridDims gc = make_grid(gf.ex / 2, gf.ey / 2, gf.ez / 2);
  dk_up_kernel<<<(unsigned)nlist, 64, 0, s>>>(gf, gc, cflag, list, map_f, dKf, dK);
  return check_launch("dk_up_kernel");
}

// Node stencils straight from the density moments, the element matrices never stored: block (n, n+o)
// sums, over the elements e holding n and n+o in index order, K_e[A][B] = sum_m M_m(e) C[A][B][m] (m
// ascending) plus dK of a flagged element, summed over e in index order - the order in which element
// matrices formed first and assembled afterwards would be summed (bitwise the same stencils), without
// the 576 x nel element matrices' round trip through memory (302 MB at 64x32x32).  A block covers one
// offset (blockIdx.y) for 256 nodes, so the <= 8 table blocks (A, B) it needs, 27 x 9 entries each, are
// staged in LDS and read as broadcasts; the moments are per-lane loads that neighbouring nodes share
// through the cache.
template <bool HALF>
__global__ void __launch_bounds__(256)
stencil_moments_kernel(GridDims g, const double* __restrict__ C, const double* __restrict__ M,
                       const int* __restrict__ map, const double* __restrict__ dK, StencilOut out) {
  // a thread per (node, block offset), the offset from blockIdx.y: uniform, so the table reads below are
  // scalar loads
  const long long nn = g.nnodes;
  const int o = (HALF ? 13 : 0) + (int)blockIdx.y;
  // the offset's (A, B) table blocks, one per element position e that holds both nodes, staged in LDS
  __shared__ double sC[8][kNM * 9];
  {
    const int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
    for (int e = 0; e < 8; ++e) {
      const int ex = e & 1, ey = (e >> 1) & 1, ez = (e >> 2) & 1;
      const int bx = dx + 1 - ex, by = dy + 1 - ey, bz = dz + 1 - ez;
      if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
      const int A = (1 - ex) + 2 * (1 - ey) + 4 * (1 - ez), B = bx + 2 * by + 4 * bz;
      for (int q = threadIdx.x; q < kNM * 9; q += blockDim.x) sC[e][q] = C[(A * 8 + B) * kNM * 9 + q];
    }
  }
  __syncthreads();
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  const int i = (int)(n % g.nx);
  const long long t = n / g.nx;
  const int j = (int)(t % g.ny);
  const int k = (int)(t / g.ny);
  {
    const int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
    double blk[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int d2 = 0; d2 < 3; ++d2) blk[d][d2] = 0.0;
    for (int e = 0; e < 8; ++e) {
      const int ex = e & 1, ey = (e >> 1) & 1, ez = (e >> 2) & 1;
      const int bx = dx + 1 - ex, by = dy + 1 - ey, bz = dz + 1 - ez;
      if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;  // uniform
      const int ei = i - 1 + ex, ej = j - 1 + ey, ek = k - 1 + ez;
      if (ei < 0 || ei >= g.ex || ej < 0 || ej >= g.ey || ek < 0 || ek >= g.ez) continue;
      const long long E = ((long long)ek * g.ey + ej) * g.ex + ei;
      const int A = (1 - ex) + 2 * (1 - ey) + 4 * (1 - ez), B = bx + 2 * by + 4 * bz;
      const double* Cab = sC[e];
      double m[kNM];  // the element's moments, loaded together
#pragma unroll
      for (int mm = 0; mm < kNM; ++mm) m[mm] = M[mm * g.nel + E];
      double ke[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) ke[q] = 0.0;
#pragma unroll
      for (int mm = 0; mm < kNM; ++mm)
#pragma unroll
        for (int q = 0; q < 9; ++q) ke[q] = fma(m[mm], Cab[mm * 9 + q], ke[q]);
      const int w = map[E];
      if (w >= 0) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int d2 = 0; d2 < 3; ++d2) ke[d * 3 + d2] += dK[(long long)w * 576 + (A * 3 + d) * 24 + B * 3 + d2];
      }
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int d2 = 0; d2 < 3; ++d2) blk[d][d2] += ke[d * 3 + d2];
    }
    emit_block<HALF>(out, o, blk, n, nn);
  }
}

int launch_stencil_from_moments(const GridDims& g, const double* C, const double* M, const int* map,
                                const double* dK, const StencilOut& out, hipStream_t s) {
  const unsigned bx = (unsigned)blocks_for(g.nnodes, 256);
  if (out.S)
    stencil_moments_kernel<false><<<dim3(bx, 27), 256, 0, s>>>(g, C, M, map, dK, out);
  else
    stencil_moments_kernel<true><<<dim3(bx, 14), 256, 0, s>>>(g, C, M, map, dK, out);
  return check_launch("stencil_moments_kernel");
}

}  // namespace pmt
