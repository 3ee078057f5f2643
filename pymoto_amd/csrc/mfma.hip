// Fine-grid hex8 operator on the matrix cores: element-by-element K(x) u as a dense contraction
// Y = K0 * [u_e1 u_e2 ...] with BF16 operands and FP32 accumulation (v_mfma_f32_16x16x32_bf16),
// for the damped-Jacobi sweeps of the reduced-precision V-cycle.  Same operator as the node-centric FP64 kernel of
// operator.hip (K = sum_e x_e M_e K0 M_e with Dirichlet rows/columns replaced by bcdiag on the
// diagonal, pymoto/modules/assembly.py:245-296), same fused modes (apply, residual b - K u and
// the damped-Jacobi sweep of pymoto/solvers/iterative.py:21-47).
//
// Layout of the contraction.  Element dofs are ordered k = 4*a + d (local node a = x + 2y + 4z,
// component d = x, y, z and a zero pad at d = 3), so one element's displacement is 8 nodes x
// 4 BF16 = 64 bytes and K0 becomes a 32 x 32 matrix (24 x 24 real).  A block owns a 32 x 8
// column of nodes and walks it along z.  Per element plane it:
//   1. contracts 16 elements at a time: B operand (32 dofs x 16 elements) gathered from the LDS
//      displacement planes with two 8-byte reads per lane, A operand (K0, two 16-row tiles)
//      held in registers for the whole kernel as a BF16 hi + lo pair; 4 MFMAs per 16 elements;
//   2. scales each element's 32 results by x_e and writes them to LDS, [node][element] float4s;
//   3. every thread sums the 8 element contributions of its node: the four elements of this
//      plane below the node finish the node's value, the four above start the next plane's
//      (kept in registers), so each element is contracted once per sweep.
// The next displacement and density planes are loaded into registers while the current plane is
// contracted (one plane of look-ahead), as in the FP64 streamed kernel.
#include "launch.h"

namespace pmt {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));

struct MTile {
  static constexpr int SX = 32, SY = 8;                         // nodes per plane of a block
  static constexpr int HX = SX + 2, HY = SY + 2, HP = HX * HY;  // displacement plane with halo
  static constexpr int EX = SX + 1, EY = SY + 1, EP = EX * EY;  // element plane with halo (297)
  static constexpr int NG = (EP + 15) / 16;                     // groups of 16 elements (19)
  static constexpr int ES = NG * 16;                            // element slots (304)
  static constexpr int UL = (HP + 255) / 256, XL = (EP + 255) / 256;
};

// float -> bfloat16, round to nearest even (finite inputs)
__device__ inline unsigned bf16_bits(float f) {
  const unsigned u = __float_as_uint(f);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

template <int MODE>
__global__ void __launch_bounds__(256)
fine_mfma_kernel(GridDims g, int tiles_x, int ntiles_xy, int nzc, int nwork, const double* __restrict__ K0,
                 const float* __restrict__ xe, const uint8_t* __restrict__ bcm, float bcdiag,
                 const float* __restrict__ u, const float* __restrict__ b, const float* __restrict__ dinv,
                 float omega, float* y, const float* up, float beta) {
  constexpr int HX = MTile::HX, HP = MTile::HP, EX = MTile::EX, EP = MTile::EP, ES = MTile::ES;
  constexpr int UL = MTile::UL, XL = MTile::XL;
  __shared__ uint2 su[3][HP];    // BF16 (x, y, z, 0) per node, ring of 3 planes
  __shared__ float sx[2][ES];    // densities, ring of 2 planes
  __shared__ f4 sy[8 * ES];      // x_e * K0 u_e per local node and element slot ([a][slot])

  const int work = (int)xcd_block();
  if (work >= nwork) return;  // whole block leaves together
  const int txy = work % ntiles_xy;
  const int i0 = (txy % tiles_x) * MTile::SX, j0 = (txy / tiles_x) * MTile::SY;
  const int kbeg = (work / ntiles_xy) * nzc;
  const int kend = min(kbeg + nzc, g.nz);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  // A operands: rows m = 16*t + lane%16 = 4a + d, columns k = 8*(lane/16) + s = 4b + d2.  K0 is
  // split as K0 = hi + lo with both parts in BF16 (16 significant bits together), so the rounding of
  // the element matrix does not break its rigid-body null space; u is rounded once to BF16.
  bf16x8 A[2][2];  // [tile][hi, lo]
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int m = 16 * t + (lane & 15), a = m >> 2, d = m & 3;
    unsigned wh[4], wl[4];
#pragma unroll
    for (int s2 = 0; s2 < 4; ++s2) {
      unsigned h[2], l2[2];
#pragma unroll
      for (int o = 0; o < 2; ++o) {
        const int kk = 8 * (lane >> 4) + 2 * s2 + o, bb = kk >> 2, d2 = kk & 3;
        const float v = (d < 3 && d2 < 3) ? (float)K0[(3 * a + d) * 24 + 3 * bb + d2] : 0.0f;
        h[o] = bf16_bits(v);
        l2[o] = bf16_bits(v - __uint_as_float(h[o] << 16));
      }
      wh[s2] = h[0] | (h[1] << 16);
      wl[s2] = l2[0] | (l2[1] << 16);
    }
    u4v rh = {wh[0], wh[1], wh[2], wh[3]}, rl = {wl[0], wl[1], wl[2], wl[3]};
    A[t][0] = __builtin_bit_cast(bf16x8, rh);
    A[t][1] = __builtin_bit_cast(bf16x8, rl);
  }

  // per-thread halo slots (plane-independent)
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

  float ru[UL][3];
  unsigned rm[UL];
  float rx[XL];
  auto fetch = [&](int p, int pe) {
#pragma unroll
    for (int l = 0; l < UL; ++l) {
      rm[l] = 7u;
      ru[l][0] = ru[l][1] = ru[l][2] = 0.0f;
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
      rx[l] = 0.0f;
      if (xin[l] && pe >= 0 && pe < g.ez) rx[l] = xe[(long long)pe * xplane + xoff[l]];
    }
  };
  auto store_u = [&](int p) {
    uint2* dst = su[(p + 1) % 3];
#pragma unroll
    for (int l = 0; l < UL; ++l) {
      const int h = tid + 256 * l;
      if (h < HP) {
        const unsigned bx = (rm[l] & 1u) ? 0u : bf16_bits(ru[l][0]);
        const unsigned by = (rm[l] & 2u) ? 0u : bf16_bits(ru[l][1]);
        const unsigned bz = (rm[l] & 4u) ? 0u : bf16_bits(ru[l][2]);
        dst[h] = make_uint2(bx | (by << 16), bz);
      }
    }
  };
  auto store_x = [&](int pe) {
    float* dst = sx[(pe + 1) & 1];
#pragma unroll
    for (int l = 0; l < XL; ++l) {
      const int q = tid + 256 * l;
      if (q < EP) dst[q] = rx[l];
    }
    // element slots beyond the plane stay zero
    for (int q = EP + tid; q < ES; q += 256) dst[q] = 0.0f;
  };

  // prologue: displacement planes kbeg-1, kbeg and density plane kbeg-1 (element plane kbeg-1
  // starts the first node plane of the chunk)
  fetch(kbeg - 1, kbeg - 1);
  store_u(kbeg - 1);
  store_x(kbeg - 1);
  fetch(kbeg, -1);
  store_u(kbeg);
  __syncthreads();

  const int lx = tid % MTile::SX, ly = tid / MTile::SX;
  const int i = i0 + lx, j = j0 + ly;
  const bool own = i < g.nx && j < g.ny;
  float carry[3] = {0.0f, 0.0f, 0.0f};  // contributions of the element plane below the node plane

  for (int k = kbeg - 1; k < kend; ++k) {
    fetch(k + 2, k + 1);  // in flight during the contraction
    const bool out_plane = k >= kbeg;
    const long long n = ((long long)k * g.ny + j) * g.nx + i;
    unsigned mk = 0;
    float uo[3] = {0, 0, 0}, bo[3] = {0, 0, 0}, dv[3] = {0, 0, 0}, po[3] = {0, 0, 0};
    if (own && out_plane) {
      mk = bcm[n];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        uo[d] = u[3 * n + d];
        if (MODE != MODE_APPLY) bo[d] = b[3 * n + d];
        if (MODE == MODE_JACOBI || MODE == MODE_CHEB) dv[d] = dinv[3 * n + d];
        if (MODE == MODE_CHEB && up) po[d] = up[3 * n + d];
      }
    }

    // 1-2: contract element plane k, 16 elements per MFMA pair
    {
      const uint2* lo = su[(k + 1) % 3];      // node plane k
      const uint2* hi = su[(k + 2) % 3];      // node plane k+1
      const float* xs = sx[(k + 1) & 1];
      const int kq = lane >> 4;
      const uint2* pl = (kq >> 1) ? hi : lo;  // nodes 2kq, 2kq+1: z = kq>>1, y = kq&1, x = 0/1
#pragma unroll 1
      for (int grp = wave; grp < MTile::NG; grp += 4) {
        const int s = 16 * grp + (lane & 15);
        u4v raw = {0u, 0u, 0u, 0u};
        if (s < EP) {
          const int ex = s % EX, ey = s / EX;
          const int hn = (ey + (kq & 1)) * HX + ex;
          const uint2 n0 = pl[hn], n1 = pl[hn + 1];
          raw = u4v{n0.x, n0.y, n1.x, n1.y};
        }
        const bf16x8 B = __builtin_bit_cast(bf16x8, raw);
        f4 c0 = {0.0f, 0.0f, 0.0f, 0.0f}, c1 = {0.0f, 0.0f, 0.0f, 0.0f};
        c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0][1], B, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[1][1], B, c1, 0, 0, 0);
        c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0][0], B, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[1][0], B, c1, 0, 0, 0);
        // D[row = 4*(lane/16) + v][col = lane%16]: this lane holds node a = lane/16 (tile 0) and
        // a = 4 + lane/16 (tile 1) of element slot s, components v = 0..3
        const float xv = xs[s];
        sy[kq * ES + s] = c0 * xv;
        sy[(4 + kq) * ES + s] = c1 * xv;
      }
    }
    __syncthreads();

    // 3: node sums
    float acc[3] = {carry[0], carry[1], carry[2]};
    float nxt[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int ex = lx + 1 - (a & 1), ey = ly + 1 - ((a >> 1) & 1);
      const f4 v = sy[a * ES + ey * EX + ex];
      if (a < 4) {
        acc[0] += v[0];
        acc[1] += v[1];
        acc[2] += v[2];
      } else {
        nxt[0] += v[0];
        nxt[1] += v[1];
        nxt[2] += v[2];
      }
    }
    carry[0] = nxt[0];
    carry[1] = nxt[1];
    carry[2] = nxt[2];
    if (own && out_plane) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const float ku = ((mk >> d) & 1u) ? bcdiag * uo[d] : acc[d];
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

// Planes per block: minimise (rounds of resident blocks) x (planes + 1 prologue element plane),
// three resident 256-thread blocks per CU (LDS-bound).
static int mfma_planes(int ntiles_xy, int nz, int cus) {
  const long long slots = 3LL * cus;
  int best = nz;
  long long best_cost = -1;
  for (int c = 2; c <= nz; ++c) {
    const long long blocks = (long long)ntiles_xy * ((nz + c - 1) / c);
    const long long cost = ((blocks + slots - 1) / slots) * (c + 1);
    if (best_cost < 0 || cost < best_cost || (cost == best_cost && c < best)) {
      best_cost = cost;
      best = c;
    }
  }
  return best;
}

int launch_fine_mfma(int mode, const GridDims& g, const double* K0, const float* xe, const uint8_t* bcm,
                     double bcdiag, const float* u, const float* b, const float* dinv, double omega, float* y,
                     hipStream_t s, const float* up, double beta) {
  if (g.dim != 3) {
    set_error("the BF16 matrix-core operator is implemented for hex8 grids only");
    return PMT_ERR_ARG;
  }
  const int tx = (g.nx + MTile::SX - 1) / MTile::SX;
  const int ty = (g.ny + MTile::SY - 1) / MTile::SY;
  const int nxy = tx * ty;
  const int planes = tuned_stream_planes();
  const int nzc = planes > 0 ? (planes < g.nz ? planes : g.nz) : mfma_planes(nxy, g.nz, device_cu_count());
  const int nwork = nxy * ((g.nz + nzc - 1) / nzc);
  dim3 grid((unsigned)((nwork + 7) & ~7));
  const float bd = (float)bcdiag, om = (float)omega, be = (float)beta;
#define PMT_MFMA(M) \
  fine_mfma_kernel<M><<<grid, 256, 0, s>>>(g, tx, nxy, nzc, nwork, K0, xe, bcm, bd, u, b, dinv, om, y, up, be)
  if (mode == MODE_APPLY)
    PMT_MFMA(MODE_APPLY);
  else if (mode == MODE_RESIDUAL)
    PMT_MFMA(MODE_RESIDUAL);
  else if (mode == MODE_JACOBI)
    PMT_MFMA(MODE_JACOBI);
  else
    PMT_MFMA(MODE_CHEB);
#undef PMT_MFMA
  return check_launch("fine_mfma_kernel");
}

}  // namespace pmt
