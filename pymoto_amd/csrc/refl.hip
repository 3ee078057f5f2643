// Fine-level K(x) u for hex8 element matrices that are symmetric under the three mid-plane
// reflections of the element (every box element of an orthotropic, axis-aligned material - the
// isotropic unit cube of AssembleStiffness included).
//
// Such a K0 commutes with the reflections R_x, R_y, R_z (node bit flipped, the matching displacement
// component negated), so in the basis of their joint eigenvectors it is block diagonal: the 24 dofs
// split into 8 characters chi of 3 dofs each, dof (d, k) of the corner-sign transform
//   U_d[k] = sum_a (-1)^(k.a) u_d[a]   (a, k: 3-bit corner / frequency indices)
// belonging to chi = k ^ (1 << d).  Then
//   K0 u = WHT( blockdiag(M_chi) WHT(u) ),   M_chi[d][e] = 1/64 sum_{a,b} (-1)^(k_d.a + k_e.b) K0[a d][b e],
// with WHT the unnormalised 8-point Walsh-Hadamard transform over the corner bits, per component.
// An element costs 2*72 add/sub + 72 multiply-adds + 24 multiplies (by x_e) instead of 576
// multiply-adds, which moves the FP64 operator from the FMA roof to the HBM roof.
//
// Kernel layout: a block owns an RX x SY column of nodes (RX consecutive lanes per node row) and walks
// z.  Thread (tx, ty) owns node (i, j) and the element (i, j) whose corner 0 it is.  Per element
// layer k it forms the 8 Dirichlet-masked corner vectors (own node from registers, the +x neighbour
// by a DPP shift, the +y row through LDS), multiplies, and sums the corner results into the nodes one
// axis at a time: z first (the upper half is carried to the next layer in registers), then y (LDS),
// then x (DPP), so every value crosses HBM once per application and only 3 values per node cross
// lanes by DPP.  Edge rows/columns of the column are halo: (RX-2) x (SY-2) nodes are written per block.
#include <type_traits>

#include "launch.h"

namespace pmt {

constexpr int pow2_floor(int n) {
  int p = 1;
  while (2 * p <= n) p *= 2;
  return p;
}

// DPP lane shifts by one node along x.  RX = 16: row shifts within the 16-lane rows of the wave; RX = 32
// or 64: shifts of the whole wave.  Lanes shifted in from outside the row (wave) read 0; with RX = 32 the
// lane at a row boundary receives the other row's edge node, which only ever lands in halo lanes (the
// last column's element and the first column's node sum are never used).
// from_lower: lane n receives lane n-1; from_upper: lane n receives lane n+1.
template <int CTRL>
__device__ inline float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, true));
}
template <int CTRL>
__device__ inline double dpp_move(double v) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(bits & 0xffffffffLL), CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(bits >> 32), CTRL, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int RX, typename T>
__device__ inline T from_lower(T v) {
  static_assert(RX == 16 || RX == 32 || RX == 64, "node rows of 16, 32 or 64 lanes");
  return dpp_move<RX == 16 ? 0x111 : 0x138>(v);  // row_shr:1 / wave_shr:1
}
template <int RX, typename T>
__device__ inline T from_upper(T v) {
  static_assert(RX == 16 || RX == 32 || RX == 64, "node rows of 16, 32 or 64 lanes");
  return dpp_move<RX == 16 ? 0x101 : 0x130>(v);  // row_shl:1 / wave_shl:1
}

// Entry (d, e) of block chi vanishes for every K0 with the box sign/class pattern: characters 1..6
// split into a single dof (iso) and a coupled pair.
__host__ __device__ constexpr int refl_iso(int chi) {
  return chi == 1 || chi == 6 ? 0 : (chi == 2 || chi == 5 ? 1 : 2);
}
__host__ __device__ constexpr bool refl_zero(int chi, int d, int e) {
  return chi != 0 && chi != 7 && d != e && (d == refl_iso(chi) || e == refl_iso(chi));
}
// The rigid translation along d (frequency 0 of component d, character 2^d) is in K0's null space, so
// its block entry vanishes for a stiffness matrix (pmt_k0_extend checks it: PMT_K0_RIGID).
__host__ __device__ constexpr bool refl_rigid(int chi, int d, int e) { return chi == (1 << d) && e == d; }

// Fine-kernel switches (both keep the results): skip the rigid-translation entries along with the box
// pattern's zeros; skip the Dirichlet selects in waves whose nodes of the plane have none.
constexpr bool kReflSkipRigid = true;
constexpr bool kReflMaskBallot = true;
constexpr bool kReflCarry64 = true;  // FP64 also carries the transformed lower plane in registers

// Isotropic cube element (PMT_K0_CUBE): the blocks hold 8 numbers in a fixed pattern - chi 0 and 7:
// a on the diagonal, b off it; chi 1, 2, 4: the pair of non-translation dofs coupled as [[c, d], [d, c]];
// chi 3, 5, 6: the pair all e, the single dof f.  The kernel reads them once (as a - b, b, ...) and
// folds the element factor into them per element.
template <typename T>
struct CubeBlocks {
  T amb0, b0, amb7, b7, c, d, e, f;
};
template <typename T>
__device__ inline CubeBlocks<T> cube_blocks(const T* __restrict__ tab) {
  return {tab[0] - tab[1], tab[1], tab[63] - tab[64], tab[64], tab[13], tab[14], tab[27], tab[35]};
}

// v[0..3] / v[4..7]: the lower / upper plane's masked corner displacements after the x and y stages.
// Returns x * (K0 u_e) in v before the inverse transform (frequency index k = kx + 2 ky + 4 kz): the
// caller interleaves its stages with the node sums.  PAT skips the entries that the box pattern makes
// zero, and with them the rigid translations' entries (its callers take PAT only with PMT_K0_RIGID).
template <bool PAT, typename T>
__device__ inline void refl_element(const T* __restrict__ tab, T x, T (&v)[8][3]) {
  constexpr auto zero = [](int chi, int d, int e) {
    return PAT && (refl_zero(chi, d, e) || (kReflSkipRigid && refl_rigid(chi, d, e)));
  };
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T p = v[a][d], q = v[a + 4][d];
      v[a][d] = p + q;
      v[a + 4][d] = p - q;
    }
  T w[8][3];
#pragma unroll
  for (int chi = 0; chi < 8; ++chi) {
    const T in0 = x * v[chi ^ 1][0], in1 = x * v[chi ^ 2][1], in2 = x * v[chi ^ 4][2];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T* m = tab + chi * 9 + d * 3;
      T acc = zero(chi, d, 0) ? (T)0 : m[0] * in0;
      if (!zero(chi, d, 1)) acc = zero(chi, d, 0) ? m[1] * in1 : fma(m[1], in1, acc);
      if (!zero(chi, d, 2)) acc = (zero(chi, d, 0) && zero(chi, d, 1)) ? m[2] * in2 : fma(m[2], in2, acc);
      w[chi ^ (1 << d)][d] = acc;
    }
  }
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) v[a][d] = w[a][d];
}

// refl_element for the isotropic cube (CubeBlocks): 8 multiplies fold x into the numbers, then 33
// operations instead of 72; the rigid translations' outputs (frequency 0) are zero.
template <typename T>
__device__ inline void refl_element_cube(const CubeBlocks<T>& cb, T x, T (&v)[8][3]) {
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T p = v[a][d], q = v[a + 4][d];
      v[a][d] = p + q;
      v[a + 4][d] = p - q;
    }
  const T A0 = cb.amb0 * x, B0 = cb.b0 * x, A7 = cb.amb7 * x, B7 = cb.b7 * x;
  const T C = cb.c * x, D = cb.d * x, E = cb.e * x, F = cb.f * x;
  T w[8][3];
  // chi 0: inputs (v[1][0], v[2][1], v[4][2]); chi 7: (v[6][0], v[5][1], v[3][2])
  {
    const T bs = B0 * (v[1][0] + v[2][1] + v[4][2]);
    w[1][0] = fma(A0, v[1][0], bs);
    w[2][1] = fma(A0, v[2][1], bs);
    w[4][2] = fma(A0, v[4][2], bs);
  }
  {
    const T bs = B7 * (v[6][0] + v[5][1] + v[3][2]);
    w[6][0] = fma(A7, v[6][0], bs);
    w[5][1] = fma(A7, v[5][1], bs);
    w[3][2] = fma(A7, v[3][2], bs);
  }
  // chi 1, 2, 4: the translation dof (frequency 0) drops out; the pair is [[c, d], [d, c]]
  w[0][0] = w[0][1] = w[0][2] = (T)0;
  w[3][1] = fma(C, v[3][1], D * v[5][2]);  // chi 1: dofs 1, 2
  w[5][2] = fma(C, v[5][2], D * v[3][1]);
  w[3][0] = fma(C, v[3][0], D * v[6][2]);  // chi 2: dofs 0, 2
  w[6][2] = fma(C, v[6][2], D * v[3][0]);
  w[5][0] = fma(C, v[5][0], D * v[6][1]);  // chi 4: dofs 0, 1
  w[6][1] = fma(C, v[6][1], D * v[5][0]);
  // chi 3, 5, 6: the pair is e (x + y) in both dofs, the single dof f
  w[2][0] = w[1][1] = E * (v[2][0] + v[1][1]);  // chi 3: pair 0, 1; single 2
  w[7][2] = F * v[7][2];
  w[4][0] = w[1][2] = E * (v[4][0] + v[1][2]);  // chi 5: pair 0, 2; single 1
  w[7][1] = F * v[7][1];
  w[4][1] = w[2][2] = E * (v[4][1] + v[2][2]);  // chi 6: pair 1, 2; single 0
  w[7][0] = F * v[7][0];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) v[a][d] = w[a][d];
}

// First coarse level of a hex8 hierarchy, element by element in the same basis.  With K0 reflection
// symmetric, the fixed child matrices G_c = P_c^T K0 P_c are G_0 mirrored: G_c = S_c G_0 S_c.  In the
// transformed basis S_c is diagonal with the characters of class chi, so
//   K^_E[a][b] = G^_0[a][b] X^[chi(a) ^ chi(b)],   X^[m] = sum_c (-1)^(m.c) x_c   (children c of E),
// a fixed 24 x 24 table (G0tab, the 1/64 of both transforms folded in) modulated by the 8 transformed
// child densities.  Entry (a, b) (a = k*3 + d) can only be nonzero if some k' within k and k'' within
// the k of b (bitwise) share a class (G_0 = P_0^T K0 P_0 with P_0 bit-triangular in this basis); with
// the box class pattern also the pattern's zero entries of K0's blocks drop out.
__host__ __device__ constexpr bool g0_nonzero(int a, int b, bool pat) {
  const int K = a / 3, d = a % 3, K2 = b / 3, d2 = b % 3;
  for (int k = 0; k < 8; ++k) {
    if (k & ~K) continue;
    for (int k2 = 0; k2 < 8; ++k2) {
      if (k2 & ~K2) continue;
      const int c1 = k ^ (1 << d), c2 = k2 ^ (1 << d2);
      if (c1 == c2 && !(pat && refl_zero(c1, d, d2))) return true;
    }
  }
  return false;
}

// Rows of the entries of the table that can be nonzero for box elements (279 of the 360 above): the
// rows and columns of the zero frequency (rigid translations, which K0 annihilates) and further
// entries vanish.  Measured on the isotropic unit cube (Poisson ratios 0.2 to 0.45) and a 1 x 1 x 2
// box; coarse1_table_pattern checks a given table against it on the host, so the pattern is only used
// where it holds.
constexpr unsigned kG0NumericRows[24] = {
    0x000000, 0x000000, 0x000000, 0xfac688, 0x650650, 0xa69820, 0x650650, 0xfac688, 0xd92900, 0xffc6d8, 0xffc6d8, 0xffb920,
    0xa69820, 0xd92900, 0xfac688, 0xfedea8, 0xfd2f50, 0xfedea8, 0xe79e70, 0xfbef88, 0xfbef88, 0xffdef8, 0xffefd8, 0xffffa8};

// nz[p]: p = 0 structural, 1 with the box class pattern, 2 numerical (kG0NumericRows)
struct G0Pattern {
  bool nz[3][576];
  constexpr G0Pattern() : nz() {
    for (int p = 0; p < 2; ++p)
      for (int q = 0; q < 576; ++q) nz[p][q] = g0_nonzero(q / 24, q % 24, p == 1);
    for (int q = 0; q < 576; ++q) nz[2][q] = nz[1][q] && ((kG0NumericRows[q / 24] >> (q % 24)) & 1u);
  }
};
constexpr G0Pattern kG0Pattern{};

// v: the element's corners after the plane stages (as for refl_element); X: its X^.  Returns K_E u_e
// before the inverse transform (as refl_element), without the Dirichlet corrections.
template <int PAT, typename T>
__device__ inline void coarse1_element(const T* tab, const T (&X)[8], T (&v)[8][3]) {
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T p = v[a][d], q = v[a + 4][d];
      v[a][d] = p + q;
      v[a + 4][d] = p - q;
    }
  T w[8][3];
  int zoff = 0;  // 0, laundered through the previous class's result (see below)
#pragma unroll
  for (int ca = 0; ca < 8; ++ca) {
    // One output class at a time.  Its table entries are read through an offset the compiler cannot
    // see is 0 and that exists only once the previous class's results do, so the reads of all ~400
    // entries are not hoisted to the start (they would not fit the register file).  FP32: the offset
    // is scalar and the entries come through the scalar cache (SGPR operands); FP64: LDS broadcasts.
    const T* tb = tab + zoff;
    T acc[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int al = (ca ^ (1 << d)) * 3 + d;
      acc[d] = (T)0;
#pragma unroll
      for (int cb = 0; cb < 8; ++cb) {
        T t = (T)0;
        bool any = false;
#pragma unroll
        for (int d2 = 0; d2 < 3; ++d2) {
          const int be = (cb ^ (1 << d2)) * 3 + d2;  // the dof of class cb with component d2
          if (kG0Pattern.nz[PAT][al * 24 + be]) {
            t = any ? fma(tb[al * 24 + be], v[be / 3][d2], t) : tb[al * 24 + be] * v[be / 3][d2];
            any = true;
          }
        }
        if (any) acc[d] = fma(X[ca ^ cb], t, acc[d]);
      }
    }
    // the class's three results are complete before the next class's offset exists
    if constexpr (sizeof(T) == 4)
      asm volatile("s_mov_b32 %0, 0" : "=s"(zoff) : "v"(acc[0]), "v"(acc[1]), "v"(acc[2]));
    else
      asm volatile("v_mov_b32 %0, 0" : "=v"(zoff), "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]));
#pragma unroll
    for (int d = 0; d < 3; ++d) w[ca ^ (1 << d)][d] = acc[d];
  }
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) v[a][d] = w[a][d];
}

// Every global access of the kernel goes through a raw buffer descriptor of one node plane (or element
// layer), whose hardware range check returns 0 for loads and drops stores at an out-of-range offset
// (buffer helpers in common.h).  Halo lanes and planes outside the grid therefore issue the same
// instructions as the others (offset kOOB, or a zero-length plane), no branch surrounds a memory
// instruction, and the wait counters stay exact: a prefetch is waited for only where it is used,
// never together with the stores issued after it.

// OP 0: the fine grid (element factor x_e, Dirichlet-masked nodes, bcdiag rows).  OP 1: the first
// coarse level of a hex8 hierarchy (element factors X^[8], layout [layer][m][element]; no mask: the
// elements with Dirichlet children get their corrections from coarse1_correct_kernel afterwards).
// Waves per SIMD the register allocation must leave room for: the fine FP64 residual and the CG product
// fit 168 VGPRs (3 waves) without spilling (2 waves otherwise); every other instance keeps the
// compiler's choice (a bound on the sweeps makes them spill, which costs more than the wave gains).
template <int OP, int MODE, typename T>
constexpr int refl_min_waves() {
  return (OP == 0 && sizeof(T) == 8 && (MODE == MODE_RESIDUAL || MODE == MODE_APPLY_DOTS)) ? 3 : 1;
}

template <int OP, int RX, int SY, int MODE, int PAT, typename T>
__global__ void __launch_bounds__(RX* SY) __attribute__((amdgpu_waves_per_eu(refl_min_waves<OP, MODE, T>())))
fine_refl_kernel(GridDims g, int tiles_x, int ntiles_xy, int nzc, int nwork, const T* __restrict__ tab, int zero_stride,
                 const T* __restrict__ xe, const uint8_t* __restrict__ bcm, T bcdiag, const T* __restrict__ u,
                 const T* __restrict__ b, const T* __restrict__ dinv, T omega, T* y, const T* up, T beta,
                 double* __restrict__ partial) {
  constexpr int NT = RX * SY;
  // MODE_APPLY_DOTS: y = A u, and the block's sums of u.y and u.r over its output nodes to
  // partial[work] and partial[nwork + work] (b holds r)
  constexpr bool DOTS = MODE == MODE_APPLY_DOTS;
  constexpr int XN = OP == 0 ? 1 : 8;  // element factors per element
  constexpr int TS = (int)sizeof(T);
  __shared__ T su[2][6][NT];  // x-stage sums / differences of the block's node planes, both parities
  // fine-grid sweeps take D^-1 from the densities instead of a stored vector: with K0 reflection
  // symmetric every corner has the same diagonal k_d, so D(n, d) = k_d * (sum of the 8 element factors
  // around n), bcdiag on Dirichlet dofs
  constexpr bool DIAG = OP == 0 && (MODE == MODE_JACOBI || MODE == MODE_CHEB);
  __shared__ T so[DIAG ? 7 : 6][NT];  // y-stage differences (corners ay = 1) of the layer's results, both kz;
                                      // DIAG: then the x-summed element factors
  constexpr bool LDS_TAB = OP == 1 && sizeof(T) == 8;
  // carry the transformed lower plane to the next layer in registers (else rebuild it from LDS): the
  // products and residuals carry it; the sweeps (which hold the density sums as well) rebuild it, which
  // gives them one more wave per SIMD (FP64 3, FP32 5) and measured 5% less time
  constexpr bool CARRY = OP == 0 && !DIAG && (kReflCarry64 || sizeof(T) == 4);
  // element-factor ring depth: 3 (loaded with the node planes, three layers ahead) on the fine grid; 1
  // on the first coarse level, whose 8 factors per element would otherwise hold 48 (FP64) or 24 (FP32)
  // registers (with the LDS rebuild of the lower plane: 2 instead of 1 wave per SIMD in FP64, 5
  // instead of 4 in FP32); loaded one layer ahead, before the node prefetch, so that waiting for them
  // does not wait for the later node planes
  constexpr int XR = OP == 0 ? 3 : 1;
  __shared__ T sG0[LDS_TAB ? 576 : 1];  // OP 1, FP64: the 24 x 24 table (read as broadcasts)

  const int work = (int)xcd_block();
  if (work >= nwork) return;  // whole block leaves together (grid padded for the XCD remap)
  const int txy = work % ntiles_xy;
  const int tid = threadIdx.x, tx = tid % RX, ty = tid / RX;
  // PAT 2 (isotropic cube): the 8 block numbers, read once and kept in scalar registers
  const CubeBlocks<T> cube = (OP == 0 && PAT == 2) ? cube_blocks<T>(tab) : CubeBlocks<T>{};
  if (OP == 1)
    for (int q = tid; q < 576 && LDS_TAB; q += NT) sG0[q] = tab[q];  // visible after the first barrier below
  const int i = (txy % tiles_x) * (RX - 2) - 1 + tx;
  const int j = (txy / tiles_x) * (SY - 2) - 1 + ty;
  const int kbeg = (work / ntiles_xy) * nzc;
  const int kend = min(kbeg + nzc, g.nz);
  const bool nin = i >= 0 && i < g.nx && j >= 0 && j < g.ny;
  const bool ein = i >= 0 && i < g.ex && j >= 0 && j < g.ey;
  const bool outp = nin && tx >= 1 && tx <= RX - 2 && ty >= 1 && ty <= SY - 2;
  const int slot_up = min(ty + 1, SY - 1) * RX + tx;  // node (i, j+1); the last row reads itself (halo)
  const int slot_dn = max(ty - 1, 0) * RX + tx;       // node (i, j-1); the first row reads itself (halo)
  const long long uplane = (long long)g.nx * g.ny, xplane = (long long)g.ex * g.ey;
  // byte offsets within a plane (kOOB for lanes outside the grid / without an output node)
  const int o_node = nin ? (int)((long long)j * g.nx + i) : kOOB;
  const int o_vec = nin ? 3 * TS * o_node : kOOB;
  const int o_out = outp ? o_vec : kOOB;
  const int o_el = ein ? TS * (int)((long long)j * g.ex + i) : kOOB;

  auto load_node = [&](int p, T(&ur)[3], unsigned& mk) {
    const bool pin = p >= 0 && p < g.nz;
    const __amdgpu_buffer_rsrc_t ru = plane_rsrc(u, 3 * (long long)p * uplane, 3 * uplane, pin);
    mk = 0u;
    if (OP == 0) {
      const __amdgpu_buffer_rsrc_t rm = plane_rsrc(bcm, (long long)p * uplane, uplane, pin);
      mk = __builtin_amdgcn_raw_buffer_load_b8(rm, o_node, 0, 0);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) ur[d] = bload(ru, o_vec + d * TS, T());
  };
  // element factors of layer l
  auto load_x = [&](int l, T(&xv)[XN]) {
    const bool lin = l >= 0 && l < g.ez;
#pragma unroll
    for (int m = 0; m < XN; ++m) {
      const __amdgpu_buffer_rsrc_t rx = plane_rsrc(xe, ((long long)l * XN + m) * xplane, xplane, lin);
      xv[m] = bload(rx, o_el, T());
    }
  };
  auto load_out = [&](int p, T(&bo)[3], T(&dv)[3], T(&po)[3]) {
    const bool pin = p >= kbeg && p < kend;
    const long long first = 3 * (long long)p * uplane, cnt = 3 * uplane;
#pragma unroll
    for (int d = 0; d < 3; ++d) bo[d] = dv[d] = po[d] = (T)0;
    if (MODE != MODE_APPLY) {  // (DOTS: r)
      const __amdgpu_buffer_rsrc_t rb = plane_rsrc(b, first, cnt, pin);
#pragma unroll
      for (int d = 0; d < 3; ++d) bo[d] = bload(rb, o_out + d * TS, T());
    }
    if ((MODE == MODE_JACOBI || MODE == MODE_CHEB) && !DIAG) {
      const __amdgpu_buffer_rsrc_t rd = plane_rsrc(dinv, first, cnt, pin);
#pragma unroll
      for (int d = 0; d < 3; ++d) dv[d] = bload(rd, o_out + d * TS, T());
    }
    if (MODE == MODE_CHEB) {
      const __amdgpu_buffer_rsrc_t rp = plane_rsrc(up, first, cnt, pin && up != nullptr);
#pragma unroll
      for (int d = 0; d < 3; ++d) po[d] = bload(rp, o_out + d * TS, T());
    }
  };
  auto masked = [](const T(&ur)[3], unsigned mk, T(&c)[3]) {
    if (kReflMaskBallot && __ballot(mk != 0u) == 0ull) {  // no constrained dof among the wave's nodes
#pragma unroll
      for (int d = 0; d < 3; ++d) c[d] = ur[d];
      return;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) c[d] = ((mk >> d) & 1u) ? (T)0 : ur[d];
  };

  // Software pipeline over a ring of three slots (plane k in slot (k - kbeg + 1) % 3): the node values
  // of plane k+3 and the densities of layer k+3 are loaded at the end of layer k, after the stores of
  // plane k, into the slot plane k frees; the node values are used two layers later, the densities
  // three.  The loop is unrolled by three so every slot is a fixed register set: no copies, and no
  // wait falls on a memory operation issued within the last layer.  The update operands (b, D^-1,
  // previous iterate) of plane k are loaded at the start of layer k and used at its end.
  T raw[3][3], xs[XR][XN];
  unsigned mk[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    load_node(kbeg - 1 + r, raw[r], mk[r]);
    if (XR == 3 || r == 0) load_x(kbeg - 1 + r, xs[r % XR]);
  }
  // corners of a node plane after the x and y transform stages, from the own masked values c: the x
  // stage with the +x neighbour (DPP), its sum and difference to LDS plane buffer pl, and after the
  // barrier inside the y stage with the +y row's
  auto plane_corners = [&](const T(&c)[3], T(*const pl)[NT], T(&P)[4][3]) {
    T sx[3], dx[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T cx = from_upper<RX>(c[d]);
      sx[d] = c[d] + cx;
      dx[d] = c[d] - cx;
      pl[d][tid] = sx[d];
      pl[3 + d][tid] = dx[d];
    }
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T sy = pl[d][slot_up], dy = pl[3 + d][slot_up];
      P[0][d] = sx[d] + sy;
      P[2][d] = sx[d] - sy;
      P[1][d] = dx[d] + dy;
      P[3][d] = dx[d] - dy;
    }
  };
  T Wlo[4][3];  // plane k after the x and y stages, carried from the previous layer's upper plane
  {
    T c[3];
    masked(raw[0], mk[0], c);
    plane_corners(c, su[(kbeg - 1) & 1], Wlo);
  }
  T Z[4][3];  // upper (az = 1) corners of element layer k-1, xy frequency kx + 2 ky: plane k's share
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) Z[a][d] = (T)0;
  double spq = 0.0, spr = 0.0;  // DOTS: this thread's output nodes
  // DIAG: 1 / k_d (k_d = sum over the classes of the blocks' (d, d) entries: K0 = H M H with H^2 = 8 I
  // per axis and M holding the 1/64) and 1 / bcdiag, once per block; XC: the x sum of layer k-1
  T ikd[3] = {(T)1, (T)1, (T)1}, ibc = (T)1, XC = (T)0;
  if constexpr (DIAG) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      T kd = (T)0;
#pragma unroll
      for (int chi = 0; chi < 8; ++chi) kd += tab[chi * 9 + d * 4];
      ikd[d] = (T)1 / kd;
    }
    ibc = (T)1 / bcdiag;
  }

  auto layer = [&](int k, auto slot) {
    constexpr int SA = decltype(slot)::value, SB = (SA + 1) % 3, SX = SA % XR;
    T bo[3], dv[3], po[3];
    load_out(k, bo, dv, po);
    T c1[3];
    masked(raw[SB], mk[SB], c1);
    T Whi[4][3];
    plane_corners(c1, su[(k + 1) & 1], Whi);
    if constexpr (!CARRY) {
      // rebuild the lower plane from the x-stage values the previous layer left in LDS (its own and the
      // +y row's): 24 fewer live registers than carrying it
      T(*const pl)[NT] = su[k & 1];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const T sx = pl[d][tid], dx = pl[3 + d][tid];
        const T sy = pl[d][slot_up], dy = pl[3 + d][slot_up];
        Wlo[0][d] = sx + sy;
        Wlo[2][d] = sx - sy;
        Wlo[1][d] = dx + dy;
        Wlo[3][d] = dx - dy;
      }
    }
    T v[8][3];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        v[a][d] = Wlo[a][d];
        v[4 + a][d] = Whi[a][d];
        if constexpr (CARRY) Wlo[a][d] = Whi[a][d];
      }
    // the table entries are re-read through the scalar cache each layer (zero_stride is 0; the
    // compiler cannot hoist the loads, which would exceed the scalar register file)
    if constexpr (OP == 0 && PAT == 2) {
      refl_element_cube<T>(cube, xs[SX][0], v);
    } else if constexpr (OP == 0) {
      refl_element<PAT != 0>(tab + zero_stride * k, xs[SX][0], v);
    } else {
      coarse1_element<PAT>((LDS_TAB ? sG0 : tab) + zero_stride * k, xs[SX], v);  // (re-read per layer)
    }

    // Inverse transform interleaved with the node sums, one axis at a time (all stages are linear).
    // z: corner az = 0 completes plane k with the carried upper half of layer k-1, az = 1 is carried to
    // plane k+1; y: the node takes its own corner ay = 0 plus the row below's ay = 1 (LDS); x: its own
    // ax = 0 plus the left neighbour's ax = 1 (DPP).  21 add/sub per component, as for the transform and
    // the sums separately minus the shared stages; 3 values per node cross lanes.
    T zs[4][3];  // plane k, xy frequency kx + 2 ky
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const T p = v[a][d] + v[a + 4][d], q = v[a][d] - v[a + 4][d];
        zs[a][d] = Z[a][d] + p;
        Z[a][d] = q;
      }
    T ys[2][3];  // [kx]
#pragma unroll
    for (int kx = 0; kx < 2; ++kx)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        ys[kx][d] = zs[kx][d] + zs[kx + 2][d];
        so[3 * kx + d][tid] = zs[kx][d] - zs[kx + 2][d];
      }
    T xsum = (T)0;  // DIAG: element factors of layer k around the node, x then y
    if constexpr (DIAG) {
      xsum = xs[SX][0] + from_lower<RX>(xs[SX][0]);
      so[6][tid] = xsum;
    }
    __syncthreads();
#pragma unroll
    for (int kx = 0; kx < 2; ++kx)
#pragma unroll
      for (int d = 0; d < 3; ++d) ys[kx][d] += so[3 * kx + d][slot_dn];
    T au[3];  // (K u) of the own node, plane k
#pragma unroll
    for (int d = 0; d < 3; ++d) au[d] = (ys[0][d] + ys[1][d]) + from_lower<RX>(ys[0][d] - ys[1][d]);
    if constexpr (DIAG) {
      xsum += so[6][slot_dn];
      const T inv = (T)1 / (XC + xsum);  // plane k: layers k-1 and k
      XC = xsum;
#pragma unroll
      for (int d = 0; d < 3; ++d) dv[d] = inv * ikd[d];
    }
    // Dirichlet rows: bcdiag u (and 1 / bcdiag as D^-1); skipped in waves without constrained dofs
    const bool anyfix = !kReflMaskBallot || __ballot(mk[SA] != 0u) != 0ull;
    if (anyfix) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const bool fx = (mk[SA] >> d) & 1u;
        au[d] = fx ? bcdiag * raw[SA][d] : au[d];
        if constexpr (DIAG) dv[d] = fx ? ibc : dv[d];
      }
    }
    {
      // plane k is complete (stores of the prologue layer and of halo lanes are dropped)
      const __amdgpu_buffer_rsrc_t ry = plane_rsrc(y, 3 * (long long)k * uplane, 3 * uplane, k >= kbeg);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const T uk = raw[SA][d];
        const T ku = au[d];
        T out;
        if (MODE == MODE_APPLY || DOTS)
          out = ku;
        else if (MODE == MODE_RESIDUAL)
          out = bo[d] - ku;
        else if (MODE == MODE_JACOBI)
          out = uk + omega * dv[d] * (bo[d] - ku);
        else
          out = uk + omega * dv[d] * (bo[d] - ku) + beta * (uk - po[d]);
        bstore(ry, o_out + d * TS, out);
        if constexpr (DOTS) {
          const bool own = outp && k >= kbeg;
          spq = fma((double)uk, own ? (double)out : 0.0, spq);
          spr = fma((double)uk, own ? (double)bo[d] : 0.0, spr);
        }
      }
    }

    __builtin_amdgcn_sched_barrier(0);  // keep the prefetch behind the stores
    if (XR == 1) load_x(k + 1, xs[0]);
    load_node(k + 3, raw[SA], mk[SA]);
    if (XR == 3) load_x(k + 3, xs[SX]);
  };

  for (int k = kbeg - 1;; k += 3) {
    layer(k, std::integral_constant<int, 0>{});
    if (k + 1 >= kend) break;
    layer(k + 1, std::integral_constant<int, 1>{});
    if (k + 2 >= kend) break;
    layer(k + 2, std::integral_constant<int, 2>{});
    if (k + 3 >= kend) break;
  }
  if constexpr (DOTS) {  // fixed-order tree over the block (LDS rows 0 and 1 of so are free now)
    // NT need not be a power of two (384-thread shapes): the threads above the largest power of two
    // P2 <= NT are folded onto the first ones before the tree
    constexpr int P2 = pow2_floor(NT);
    __syncthreads();
    so[0][tid] = spq;
    so[1][tid] = spr;
    __syncthreads();
    if constexpr (P2 < NT) {
      if (tid < NT - P2) {
        so[0][tid] += so[0][tid + P2];
        so[1][tid] += so[1][tid + P2];
      }
      __syncthreads();
    }
#pragma unroll
    for (int st = P2 / 2; st > 0; st >>= 1) {
      if (tid < st) {
        so[0][tid] += so[0][tid + st];
        so[1][tid] += so[1][tid + st];
      }
      __syncthreads();
    }
    if (tid == 0) {
      partial[work] = so[0][0];
      partial[nwork + work] = so[1][0];
    }
  }
}

// The isotropic cube's block pattern (CubeBlocks): every class of entries equal to 1e-12 of the largest
// entry and the other entries below 1e-13 of it.  On success the classes are set to their means and the
// others to 0, so every kernel sees the same numbers.
static bool k0_cube_form(double* tab, double scale) {
  static const int kClass[8][12] = {
      {0, 4, 8, -1},                                           // chi 0 diagonal
      {1, 2, 3, 5, 6, 7, -1},                                  // chi 0 off-diagonal
      {63, 67, 71, -1},                                        // chi 7 diagonal
      {64, 65, 66, 68, 69, 70, -1},                            // chi 7 off-diagonal
      {13, 17, 18, 26, 36, 40, -1},                            // c: chi 1, 2, 4 pair diagonal
      {14, 16, 20, 24, 37, 39, -1},                            // d: chi 1, 2, 4 pair coupling
      {27, 28, 30, 31, 45, 47, 51, 53, 58, 59, 61, 62},        // e: chi 3, 5, 6 pair
      {35, 49, 54, -1}};                                       // f: chi 3, 5, 6 single dof
  bool member[72] = {false};
  double mean[8];
  for (int c = 0; c < 8; ++c) {
    double sum = 0.0;
    int n = 0;
    for (int q = 0; q < 12 && kClass[c][q] >= 0; ++q, ++n) sum += tab[kClass[c][q]];
    mean[c] = sum / n;
    for (int q = 0; q < 12 && kClass[c][q] >= 0; ++q) {
      if (fabs(tab[kClass[c][q]] - mean[c]) > 1e-12 * scale) return false;
      member[kClass[c][q]] = true;
    }
  }
  for (int q = 0; q < 72; ++q)
    if (!member[q] && fabs(tab[q]) > 1e-13 * scale) return false;
  for (int q = 0; q < 72; ++q) tab[q] = 0.0;
  for (int c = 0; c < 8; ++c)
    for (int q = 0; q < 12 && kClass[c][q] >= 0; ++q) tab[kClass[c][q]] = mean[c];
  return true;
}

// Block tables of an element matrix in the reflection basis; false if K0 does not commute with the
// three mid-plane reflections (to 1e-13 of its largest entry).
bool k0_reflection_table(const double* K0, double* tab) {
  double scale = 0.0;
  for (int q = 0; q < 576; ++q) scale = fmax(scale, fabs(K0[q]));
  if (!(scale > 0.0)) return false;
  auto sgn = [](int k, int a) { return (__builtin_popcount(k & a) & 1) ? -1.0 : 1.0; };
  for (int d = 0; d < 3; ++d)
    for (int k = 0; k < 8; ++k)
      for (int e = 0; e < 3; ++e)
        for (int l = 0; l < 8; ++l) {
          double m = 0.0;
          for (int a = 0; a < 8; ++a)
            for (int bb = 0; bb < 8; ++bb) m += sgn(k, a) * sgn(l, bb) * K0[(a * 3 + d) * 24 + bb * 3 + e];
          m /= 64.0;
          const int chi = k ^ (1 << d), chi2 = l ^ (1 << e);
          if (chi != chi2) {
            if (fabs(m) > 1e-13 * scale) return false;
          } else {
            tab[chi * 9 + d * 3 + e] = m;
          }
        }
  return true;
}

// Resident blocks per CU of one kernel instance (cached per device).
template <int OP, int RX, int SY, int MODE, int PAT, typename T>
static int refl_resident_blocks() {
  static int cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 2;
  if (cache[dev] == 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fine_refl_kernel<OP, RX, SY, MODE, PAT, T>, RX * SY, 0) !=
            hipSuccess ||
        nb < 1)
      nb = 2;
    cache[dev] = nb;
  }
  return cache[dev];
}

// Planes per block.  Every CU runs `resident` blocks at once; the launch takes
// ceil(blocks / (CUs * resident)) rounds of (planes + 1) layers (one prologue layer per block).
// Minimise that, among the chunkings that still give every CU at least two resident blocks (latency
// hiding) whenever the grid allows it; ties go to the fewer blocks.
static int refl_planes(int nxy, int nz, int resident) {
  const long long cus = device_cu_count();
  const long long slots = cus * (resident > 0 ? resident : 1);
  const long long want = cus * (resident < 2 ? resident : 2);
  int best = 0;
  long long best_cost = -1;
  for (int pass = 0; pass < 2 && best == 0; ++pass)
    for (int c = nz; c >= 2; --c) {
      const long long blocks = (long long)nxy * ((nz + c - 1) / c);
      if (pass == 0 && blocks < want) continue;
      const long long cost = ((blocks + slots - 1) / slots) * (c + 1);
      if (best_cost < 0 || cost < best_cost) {
        best_cost = cost;
        best = c;
      }
    }
  return best > 0 ? best : nz;
}

template <int OP, int RX, int SY, int PAT, typename T>
static void launch_refl_shape(int mode, const GridDims& g, int planes, const T* tab, const T* xe, const uint8_t* bcm,
                              T bcdiag, const T* u, const T* b, const T* dinv, T omega, T* y, hipStream_t s,
                              const T* up, T beta, double* partial = nullptr, int* nwork_out = nullptr) {
  const int tx = (g.nx + RX - 3) / (RX - 2);
  const int ty = (g.ny + SY - 3) / (SY - 2);
  const int nxy = tx * ty;
#define PMT_REFL(M)                                                                                             \
  do {                                                                                                          \
    const int nzc = planes > 0 ? (planes < g.nz ? planes : g.nz)                                               \
                               : refl_planes(nxy, g.nz, refl_resident_blocks<OP, RX, SY, M, PAT, T>());         \
    const int nwork = nxy * ((g.nz + nzc - 1) / nzc);                                                          \
    if (nwork_out) *nwork_out = nwork;                                                                          \
    fine_refl_kernel<OP, RX, SY, M, PAT, T><<<(unsigned)((nwork + 7) & ~7), RX * SY, 0, s>>>(                  \
        g, tx, nxy, nzc, nwork, tab, 0, xe, bcm, bcdiag, u, b, dinv, omega, y, up, beta, partial);               \
  } while (0)
  if (mode == MODE_APPLY_DOTS) {
    if constexpr (OP == 0 && sizeof(T) == 8) PMT_REFL(MODE_APPLY_DOTS);
  } else if (mode == MODE_APPLY)
    PMT_REFL(MODE_APPLY);
  else if (mode == MODE_RESIDUAL)
    PMT_REFL(MODE_RESIDUAL);
  else if (mode == MODE_JACOBI)
    PMT_REFL(MODE_JACOBI);
  else
    PMT_REFL(MODE_CHEB);
#undef PMT_REFL
}

// Block shapes: `rows` node rows per block of refl_cols(rows) nodes each (12 rows of 32; 16, 24 or 32
// rows of 16).  Every shape does the same arithmetic per node: the results are bitwise the same.
static int refl_cols(int rows) { return rows == 12 ? 32 : 16; }
static const int kReflShapes[4] = {16, 24, 12, 32};

static long long refl_tiles(const GridDims& g, int rows) {
  const int rx = refl_cols(rows);
  return (long long)((g.nx + rx - 3) / (rx - 2)) * ((g.ny + rows - 3) / (rows - 2));
}

// Shape 0 (automatic): the fewest threads per node plane, halo columns and rows of every block and the
// padding of the last tiles included; ties go to the earlier shape of kReflShapes.
int refl_auto_rows(const GridDims& g) {
  int best = 16;
  long long best_threads = -1;
  for (int rows : kReflShapes) {
    const long long threads = refl_tiles(g, rows) * refl_cols(rows) * rows;
    if (best_threads < 0 || threads < best_threads) {
      best_threads = threads;
      best = rows;
    }
  }
  return best;
}

template <int OP, int PAT, typename T>
static void launch_refl_rows(int rows, int mode, const GridDims& g, int planes, const T* tab, const T* xe,
                             const uint8_t* bcm, T bcdiag, const T* u, const T* b, const T* dinv, T omega, T* y,
                             hipStream_t s, const T* up, T beta, double* partial = nullptr, int* nwork_out = nullptr) {
  switch (rows > 0 ? rows : refl_auto_rows(g)) {
    case 12:
      launch_refl_shape<OP, 32, 12, PAT, T>(mode, g, planes, tab, xe, bcm, bcdiag, u, b, dinv, omega, y, s, up, beta,
                                            partial, nwork_out);
      break;
    case 24:
      launch_refl_shape<OP, 16, 24, PAT, T>(mode, g, planes, tab, xe, bcm, bcdiag, u, b, dinv, omega, y, s, up, beta,
                                            partial, nwork_out);
      break;
    case 32:
      launch_refl_shape<OP, 16, 32, PAT, T>(mode, g, planes, tab, xe, bcm, bcdiag, u, b, dinv, omega, y, s, up, beta,
                                            partial, nwork_out);
      break;
    default:
      launch_refl_shape<OP, 16, 16, PAT, T>(mode, g, planes, tab, xe, bcm, bcdiag, u, b, dinv, omega, y, s, up, beta,
                                            partial, nwork_out);
  }
}

template <typename T>
int launch_fine_refl(int mode, const GridDims& g, int planes, int rows, int pattern, const double* K0ext, const T* xe,
                     const uint8_t* bcm, double bcdiag, const T* u, const T* b, const T* dinv, double omega, T* y,
                     hipStream_t s, const T* up, double beta) {
  const T* tab = sizeof(T) == 8 ? (const T*)(K0ext + kK0ReflOffset64) : (const T*)(K0ext + kK0ReflOffset32);
  const T bd = (T)bcdiag, om = (T)omega, be = (T)beta;
  if (pattern == 2)
    launch_refl_rows<0, 2, T>(rows, mode, g, planes, tab, xe, bcm, bd, u, b, dinv, om, y, s, up, be);
  else if (pattern == 1)
    launch_refl_rows<0, 1, T>(rows, mode, g, planes, tab, xe, bcm, bd, u, b, dinv, om, y, s, up, be);
  else
    launch_refl_rows<0, 0, T>(rows, mode, g, planes, tab, xe, bcm, bd, u, b, dinv, om, y, s, up, be);
  return check_launch("fine_refl_kernel");
}

long long fine_apply_dots_parts(const GridDims& g) {
  long long most = 0;  // blocks of the shape with the most tiles, one plane each: an upper bound
  for (int rows : kReflShapes) most = refl_tiles(g, rows) > most ? refl_tiles(g, rows) : most;
  return most * (long long)g.nz;
}

// out[o] = sum over i of partial[o * n + i], o = 0, 1, in a fixed order (one block)
__global__ void __launch_bounds__(256) sum_parts_kernel(const double* __restrict__ partial, int n,
                                                          double* __restrict__ out) {
  __shared__ double sh[256];
  for (int o = 0; o < 2; ++o) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc += partial[(long long)o * n + i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[o] = sh[0];
    __syncthreads();
  }
}

int launch_fine_apply_dots(const GridDims& g, const double* K0ext, int k0_pattern, const double* xe,
                           const uint8_t* bcm, double bcdiag, const double* p, const double* r, double* q,
                           double* ws, long long ws_parts, double* out, hipStream_t s) {
  if (g.dim != 3 || get_tuning(PMT_TUNE_FINE_KERNEL) != 2 || !(k0_pattern & PMT_K0_REFLECTION) ||
      ws_parts < fine_apply_dots_parts(g))
    return PMT_ERR_STATE;
  const double* tab = K0ext + kK0ReflOffset64;
  int nwork = 0;
  const int lvl = refl_pattern(k0_pattern);
  const int rows = get_tuning(PMT_TUNE_REFL_ROWS);
  if (lvl == 2)
    launch_refl_rows<0, 2, double>(rows, MODE_APPLY_DOTS, g, tuned_stream_planes(), tab, xe, bcm, bcdiag, p, r,
                                   nullptr, 0.0, q, s, nullptr, 0.0, ws, &nwork);
  else if (lvl == 1)
    launch_refl_rows<0, 1, double>(rows, MODE_APPLY_DOTS, g, tuned_stream_planes(), tab, xe, bcm, bcdiag, p, r,
                                   nullptr, 0.0, q, s, nullptr, 0.0, ws, &nwork);
  else
    launch_refl_rows<0, 0, double>(rows, MODE_APPLY_DOTS, g, tuned_stream_planes(), tab, xe, bcm, bcdiag, p, r,
                                   nullptr, 0.0, q, s, nullptr, 0.0, ws, &nwork);
  PMT_TRY(check_launch("fine_refl_kernel<apply + dots>"));
  sum_parts_kernel<<<1, 256, 0, s>>>(ws, nwork, out);
  return check_launch("sum_parts_kernel");
}

// Sensitivity contraction dx_e = alpha * u_e^T K0 v_e (u, v Dirichlet-masked) over z-streamed node
// columns: the same transforms as the operator (x stage with the +x neighbour by DPP, y stage through
// LDS, both shared by the two element layers of a node plane, z stage per element) and
// sum_chi U^_chi^T M_chi V^_chi with U^_chi = (U^_0[chi ^ 1], U^_1[chi ^ 2], U^_2[chi ^ 4]) (K0 = H blockdiag(M) H,
// H the 8-point transform over the corner bits per component, M holding the 1/64): 72 add/sub per
// transformed vector and 96 multiply-adds per element instead of 600.  A block of 16 x 16 threads owns the
// 15 x 15 elements whose corner 0 it holds (the last row and column only supply corners) and walks a
// chunk of element layers; one barrier per layer (the exchange buffer alternates by parity).
template <bool SAME>
__global__ void __launch_bounds__(256)
sens_stream_kernel(GridDims g, int tiles_x, int ntiles_xy, int nzc, int nwork, const double* __restrict__ tab,
                   int zero_stride, const uint8_t* __restrict__ bcm, const double* __restrict__ u,
                   const double* __restrict__ v, double alpha, double* dx, int accumulate) {
  constexpr int RX = 16, SY = 16, NT = RX * SY, NV = SAME ? 1 : 2;
  __shared__ double sx_[2][NV][6][NT];  // x-stage sums / differences of the node plane, both parities
  const int work = (int)xcd_block();
  if (work >= nwork) return;
  const int txy = work % ntiles_xy;
  const int tid = threadIdx.x, tx = tid % RX, ty = tid / RX;
  const int i = (txy % tiles_x) * (RX - 1) + tx;
  const int j = (txy / tiles_x) * (SY - 1) + ty;
  const int kbeg = (work / ntiles_xy) * nzc;
  const int kend = min(kbeg + nzc, g.ez);
  const bool nin = i < g.nx && j < g.ny;
  const bool eout = i < g.ex && j < g.ey && tx < RX - 1 && ty < SY - 1;
  const int slot_up = min(ty + 1, SY - 1) * RX + tx;
  const long long uplane = (long long)g.nx * g.ny, xplane = (long long)g.ex * g.ey;
  const int o_node = nin ? (int)((long long)j * g.nx + i) : kOOB;
  const int o_vec = nin ? 24 * (int)((long long)j * g.nx + i) : kOOB;
  const int o_el = eout ? 8 * (int)((long long)j * g.ex + i) : kOOB;
  const double* vecs[2] = {u, v};

  // masked node values of plane p (zero outside the grid)
  auto load_node = [&](int p, double (&c)[NV][3]) {
    const bool pin = p >= 0 && p < g.nz;
    const __amdgpu_buffer_rsrc_t rm = plane_rsrc(bcm, (long long)p * uplane, uplane, pin && bcm != nullptr);
    const unsigned mk = __builtin_amdgcn_raw_buffer_load_b8(rm, o_node, 0, 0);
#pragma unroll
    for (int w = 0; w < NV; ++w) {
      const __amdgpu_buffer_rsrc_t rv = plane_rsrc(vecs[w], 3 * (long long)p * uplane, 3 * uplane, pin);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double val = bload(rv, o_vec + 8 * d, 0.0);
        c[w][d] = ((mk >> d) & 1u) ? 0.0 : val;
      }
    }
  };
  // x stage (DPP) to the LDS buffer of parity par, barrier, y stage with the +y row: P[kx + 2 ky][d]
  auto plane_corners = [&](const double (&c)[NV][3], int par, double (&P)[NV][4][3]) {
    double sx[NV][3], dx_[NV][3];
#pragma unroll
    for (int w = 0; w < NV; ++w)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double cx = from_upper<RX>(c[w][d]);
        sx[w][d] = c[w][d] + cx;
        dx_[w][d] = c[w][d] - cx;
        sx_[par][w][d][tid] = sx[w][d];
        sx_[par][w][3 + d][tid] = dx_[w][d];
      }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < NV; ++w)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double sy = sx_[par][w][d][slot_up], dy = sx_[par][w][3 + d][slot_up];
        P[w][0][d] = sx[w][d] + sy;
        P[w][2][d] = sx[w][d] - sy;
        P[w][1][d] = dx_[w][d] + dy;
        P[w][3][d] = dx_[w][d] - dy;
      }
  };
  double raw[2][NV][3];  // node planes k+1 (in use) and k+2 (in flight)
  double Wlo[NV][4][3];
  {
    double c0[NV][3];
    load_node(kbeg, c0);
    load_node(kbeg + 1, raw[0]);
    plane_corners(c0, kbeg & 1, Wlo);
  }
  for (int k = kbeg; k < kend; ++k) {
    double Whi[NV][4][3];
    const int cur = (k - kbeg) & 1;
    plane_corners(raw[cur], (k + 1) & 1, Whi);
    load_node(k + 2, raw[cur ^ 1]);  // in flight during this layer's products
    double U[NV][8][3];
#pragma unroll
    for (int w = 0; w < NV; ++w)
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          U[w][a][d] = Wlo[w][a][d] + Whi[w][a][d];
          U[w][a + 4][d] = Wlo[w][a][d] - Whi[w][a][d];
          Wlo[w][a][d] = Whi[w][a][d];
        }
    const double* M0 = tab + zero_stride * k;  // re-read per layer (zero_stride is 0)
    double sum = 0.0;
#pragma unroll
    for (int chi = 0; chi < 8; ++chi) {
      const double u0 = U[0][chi ^ 1][0], u1 = U[0][chi ^ 2][1], u2 = U[0][chi ^ 4][2];
      const double v0 = U[NV - 1][chi ^ 1][0], v1 = U[NV - 1][chi ^ 2][1], v2 = U[NV - 1][chi ^ 4][2];
      const double* M = M0 + chi * 9;
      const double w0 = fma(M[0], v0, fma(M[1], v1, M[2] * v2));
      const double w1 = fma(M[3], v0, fma(M[4], v1, M[5] * v2));
      const double w2 = fma(M[6], v0, fma(M[7], v1, M[8] * v2));
      sum = fma(u0, w0, fma(u1, w1, fma(u2, w2, sum)));
    }
    const __amdgpu_buffer_rsrc_t rx = plane_rsrc(dx, (long long)k * xplane, xplane, true);
    const double prev = accumulate ? bload(rx, o_el, 0.0) : 0.0;
    bstore(rx, o_el, accumulate ? fma(alpha, sum, prev) : alpha * sum);
  }
}

int launch_sens_stream(const GridDims& g, const double* tab, const uint8_t* bcm, const double* u, const double* v,
                       double alpha, double* dx, int accumulate, hipStream_t s) {
  if (g.dim != 3 || g.nel <= 0) return arg_error("launch_sens_stream: 3D grids only");
  const int tiles_x = (g.ex + 14) / 15, tiles_y = (g.ey + 14) / 15;
  const int nxy = tiles_x * tiles_y;
  // element layers per block: about four blocks per CU, at least 8 layers (a block re-reads one plane)
  const long long want = 4LL * device_cu_count();
  int nzc = (int)(((long long)g.ez * nxy + want - 1) / want);
  nzc = nzc < 8 ? 8 : nzc;
  nzc = nzc > g.ez ? g.ez : nzc;
  const int nwork = nxy * ((g.ez + nzc - 1) / nzc);
  const unsigned blocks = (unsigned)((nwork + 7) & ~7);
  if (u == v)
    sens_stream_kernel<true><<<blocks, 256, 0, s>>>(g, tiles_x, nxy, nzc, nwork, tab, 0, bcm, u, v, alpha, dx,
                                                    accumulate);
  else
    sens_stream_kernel<false><<<blocks, 256, 0, s>>>(g, tiles_x, nxy, nzc, nwork, tab, 0, bcm, u, v, alpha, dx,
                                                     accumulate);
  return check_launch("sens_stream_kernel");
}

// Corrections of the first coarse level's elements with Dirichlet children (those along constrained
// faces).  At update, the nodes of the flagged elements ("touched" nodes) get a 27-block correction
// stencil T(n, n+o) = sum over the flagged elements E holding n and n+o of dK_E(loc n, loc n+o); every
// application then adds y(n, d) += s(n, d) * (T u)(n, d) in one launch, with s = 1 (apply), -1
// (residual) or -omega D^-1(n, d) (Jacobi and Chebyshev sweeps, which are affine in A u with that
// factor).  The sums run over the elements in index order: no atomics in the values, results
// independent of the order of the lists.
//  1. coarse1_touched_kernel: thread = (listed element, corner); it lists the corner's node if no
//     flagged element holding the node has a smaller index (the list's order is irrelevant);
//  2. coarse1_tstencil_kernel: thread = (touched node, coefficient), T in [243][cap] layout;
//  3. coarse1_tapply_kernel: four lanes per (touched node, component), 27 multiply-adds each.
__global__ void __launch_bounds__(256)
coarse1_touched_kernel(GridDims g, const int* __restrict__ list, const int* __restrict__ count,
                       const int* __restrict__ map, int* __restrict__ tnode, int* __restrict__ tcount) {
  const long long nq = 8LL * (*count);
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += (long long)gridDim.x * blockDim.x) {
    const long long E = list[q >> 3];
    const int a = (int)(q & 7);
    const int CI = (int)(E % g.ex);
    const long long tt = E / g.ex;
    const int CJ = (int)(tt % g.ey), CK = (int)(tt / g.ey);
    const int I = CI + (a & 1), J = CJ + ((a >> 1) & 1), K = CK + (a >> 2);
    long long owner = -1;
    for (int e = 0; e < 8 && owner < 0; ++e) {  // e ascending = element index ascending
      const int ei = I - 1 + (e & 1), ej = J - 1 + ((e >> 1) & 1), ek = K - 1 + (e >> 2);
      if (ei < 0 || ei >= g.ex || ej < 0 || ej >= g.ey || ek < 0 || ek >= g.ez) continue;
      const long long Ee = ((long long)ek * g.ey + ej) * g.ex + ei;
      if (map[Ee] >= 0) owner = Ee;
    }
    if (owner == E) tnode[atomicAdd(tcount, 1)] = (int)(((long long)K * g.ny + J) * g.nx + I);
  }
}

__global__ void __launch_bounds__(256)
coarse1_tstencil_kernel(GridDims g, const int* __restrict__ tnode, int nt, long long cap, const int* __restrict__ map,
                        const double* __restrict__ dK, double* __restrict__ T) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 243LL * nt) return;
  const long long t = q / 243;
  const int c = (int)(q % 243), o = c / 9, d = (c / 3) % 3, d2 = c % 3;
  const long long n = tnode[t];
  const int I = (int)(n % g.nx);
  const long long tt = n / g.nx;
  const int J = (int)(tt % g.ny), K = (int)(tt / g.ny);
  const int ox = o % 3 - 1, oy = (o / 3) % 3 - 1, oz = o / 9 - 1;
  double v = 0.0;
  for (int e = 0; e < 8; ++e) {  // elements holding n, index order
    const int ex = e & 1, ey = (e >> 1) & 1, ez = e >> 2;
    const int ei = I - 1 + ex, ej = J - 1 + ey, ek = K - 1 + ez;
    if (ei < 0 || ei >= g.ex || ej < 0 || ej >= g.ey || ek < 0 || ek >= g.ez) continue;
    const int bx = 1 - ex + ox, by = 1 - ey + oy, bz = 1 - ez + oz;  // n + o's corner in the element
    if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
    const int m = map[((long long)ek * g.ey + ej) * g.ex + ei];
    if (m < 0) continue;
    const int A = (1 - ex) + 2 * (1 - ey) + 4 * (1 - ez), B = bx + 2 * by + 4 * bz;
    v += dK[(long long)m * 576 + (A * 3 + d) * 24 + B * 3 + d2];
  }
  T[(long long)c * cap + t] = v;
}

// Four lanes per (touched node, component): lane g < 3 sums the 9 neighbours of the z offset g - 1 (27
// multiply-adds), the fourth idles, and the group (aligned in the wave) adds its three sums in order -
// 4x the threads of one lane per output (latency: 81 dependent loads and multiply-adds per lane then).
template <int MODE, typename T>
__global__ void __launch_bounds__(256)
coarse1_tapply_kernel(GridDims g, const int* __restrict__ tnode, int nt, long long cap, const double* __restrict__ TS,
                      const T* __restrict__ u, const T* __restrict__ dinv, double omega, T* __restrict__ y) {
  const long long q4 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int grp = (int)(q4 & 3);
  const long long q = q4 >> 2;
  const bool live = q < 3LL * nt;
  const long long t = live ? q / 3 : 0;
  const int d = (int)(q % 3);
  const long long n = tnode[t];
  const int I = (int)(n % g.nx);
  const long long tt = n / g.nx;
  const int J = (int)(tt % g.ny), K = (int)(tt / g.ny);
  double acc = 0.0;
  if (live && grp < 3) {
#pragma unroll
    for (int oo = 0; oo < 9; ++oo) {
      const int o = grp * 9 + oo;
      // neighbours outside the grid have zero coefficients: read a clamped (valid) node instead
      const int ni = min(max(I + o % 3 - 1, 0), g.nx - 1), nj = min(max(J + (o / 3) % 3 - 1, 0), g.ny - 1),
                nk = min(max(K + o / 9 - 1, 0), g.nz - 1);
      const long long m = ((long long)nk * g.ny + nj) * g.nx + ni;
#pragma unroll
      for (int d2 = 0; d2 < 3; ++d2)
        acc = fma(TS[(long long)((o * 3 + d) * 3 + d2) * cap + t], (double)u[m * 3 + d2], acc);
    }
  }
  const double a1 = __shfl_down(acc, 1, 64), a2 = __shfl_down(acc, 2, 64);  // within the aligned group of 4
  if (!live || grp != 0) return;
  double f = (acc + a1) + a2;
  if (MODE == MODE_RESIDUAL)
    f = -f;
  else if (MODE == MODE_JACOBI || MODE == MODE_CHEB)
    f = -omega * (double)dinv[n * 3 + d] * f;
  y[n * 3 + d] = (T)((double)y[n * 3 + d] + f);
}

int launch_coarse1_touched(const GridDims& g, const int* list, const int* map, int* tnode, int* tcount,
                           hipStream_t s) {
  zero_int(tcount, s);
  coarse1_touched_kernel<<<1024, 256, 0, s>>>(g, list, list + g.nel, map, tnode, tcount);
  return check_launch("coarse1_touched_kernel");
}

int launch_coarse1_tstencil(const GridDims& g, const int* tnode, int nt, long long cap, const int* map,
                            const double* dK, double* TS, hipStream_t s) {
  if (nt <= 0) return PMT_OK;
  coarse1_tstencil_kernel<<<(unsigned)blocks_for(243LL * nt, 256), 256, 0, s>>>(g, tnode, nt, cap, map, dK, TS);
  return check_launch("coarse1_tstencil_kernel");
}

template <typename T>
int launch_coarse1(int mode, const GridDims& g, int pattern, const double* G0ext, const T* Xh, const int* tnode, int nt,
                   long long cap, const double* TS, const T* u, const T* b, const T* dinv, double omega, T* y,
                   hipStream_t s, const T* up, double beta) {
  const T* tab = sizeof(T) == 8 ? (const T*)G0ext : (const T*)(G0ext + 576);
  const T om = (T)omega, be = (T)beta;
  if (pattern == 2)
    launch_refl_shape<1, 16, 16, 2, T>(mode, g, 0, tab, Xh, nullptr, (T)0, u, b, dinv, om, y, s, up, be);
  else if (pattern == 1)
    launch_refl_shape<1, 16, 16, 1, T>(mode, g, 0, tab, Xh, nullptr, (T)0, u, b, dinv, om, y, s, up, be);
  else
    launch_refl_shape<1, 16, 16, 0, T>(mode, g, 0, tab, Xh, nullptr, (T)0, u, b, dinv, om, y, s, up, be);
  PMT_TRY(check_launch("fine_refl_kernel<coarse level 1>"));
  if (nt <= 0) return PMT_OK;
  const unsigned blocks = (unsigned)blocks_for(12LL * nt, 256);  // 4 lanes per (node, component)
  if (mode == MODE_APPLY)
    coarse1_tapply_kernel<MODE_APPLY, T><<<blocks, 256, 0, s>>>(g, tnode, nt, cap, TS, u, dinv, omega, y);
  else if (mode == MODE_RESIDUAL)
    coarse1_tapply_kernel<MODE_RESIDUAL, T><<<blocks, 256, 0, s>>>(g, tnode, nt, cap, TS, u, dinv, omega, y);
  else
    coarse1_tapply_kernel<MODE_JACOBI, T><<<blocks, 256, 0, s>>>(g, tnode, nt, cap, TS, u, dinv, omega, y);
  return check_launch("coarse1_tapply_kernel");
}

// X^[m] = sum_c (-1)^(m.c) x_c over the unflagged children c of every first-coarse-level element (the
// flagged ones are in dK), layout [layer][m][element of the layer], FP64 and (X32 set) FP32.
__global__ void coarse1_xhat_kernel(GridDims gf, GridDims gc, const double* __restrict__ xe,
                                    const uint8_t* __restrict__ cflag, double* __restrict__ X64, float* __restrict__ X32) {
  const long long E = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (E >= gc.nel) return;
  const int CI = (int)(E % gc.ex);
  const long long tt = E / gc.ex;
  const int CJ = (int)(tt % gc.ey), CK = (int)(tt / gc.ey);
  const unsigned fl = cflag[E];
  double x[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const long long f = ((long long)(2 * CK + (c >> 2)) * gf.ey + 2 * CJ + ((c >> 1) & 1)) * gf.ex + 2 * CI + (c & 1);
    x[c] = ((fl >> c) & 1u) ? 0.0 : xe[f];
  }
#pragma unroll
  for (int bit = 1; bit < 8; bit <<= 1)
#pragma unroll
    for (int c = 0; c < 8; ++c)
      if (!(c & bit)) {
        const double p = x[c], q = x[c | bit];
        x[c] = p + q;
        x[c | bit] = p - q;
      }
  const long long xplane = (long long)gc.ex * gc.ey, e2 = (long long)CJ * gc.ex + CI;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const long long o = ((long long)CK * 8 + m) * xplane + e2;
    X64[o] = x[m];
    if (X32) X32[o] = (float)x[m];
  }
}

int launch_coarse1_xhat(const GridDims& gf, const GridDims& gc, const double* xe, const uint8_t* cflag, double* X64,
                        float* X32, hipStream_t s) {
  coarse1_xhat_kernel<<<blocks_for(gc.nel, 256), 256, 0, s>>>(gf, gc, xe, cflag, X64, X32);
  return check_launch("coarse1_xhat_kernel");
}

template int launch_fine_refl<double>(int, const GridDims&, int, int, int, const double*, const double*, const uint8_t*,
                                      double, const double*, const double*, const double*, double, double*,
                                      hipStream_t, const double*, double);
template int launch_fine_refl<float>(int, const GridDims&, int, int, int, const double*, const float*, const uint8_t*,
                                     double, const float*, const float*, const float*, double, float*, hipStream_t,
                                     const float*, double);
template int launch_coarse1<double>(int, const GridDims&, int, const double*, const double*, const int*, int,
                                    long long, const double*, const double*, const double*, const double*, double,
                                    double*, hipStream_t, const double*, double);
template int launch_coarse1<float>(int, const GridDims&, int, const double*, const float*, const int*, int, long long,
                                   const double*, const float*, const float*, const float*, double, float*, hipStream_t,
                                   const float*, double);

// Zero pattern the level-1 kernel may skip for a table (coarse1_table layout): 2 when K0 has the box class
// pattern and every entry outside the numerical pattern is below 1e-14 of the largest, 1 with the class
// pattern only, else 0.
int coarse1_table_pattern(const double* ext, int k0_pattern) {
  if (!(k0_pattern & PMT_K0_CLASS_PATTERN)) return 0;
  double scale = 0.0;
  for (int q = 0; q < 576; ++q) scale = fmax(scale, fabs(ext[q]));
  for (int q = 0; q < 576; ++q)
    if (!kG0Pattern.nz[2][q] && fabs(ext[q]) > 1e-14 * scale) return 1;
  return 2;
}

// G0tab: (1/64) H G_0 H of the child-0 matrix G_0 = P_0^T K0 P_0 (G0: host, 24 x 24), FP64 then FP32
void coarse1_table(const double* G0, double* ext) {
  auto sgn = [](int k, int a) { return (__builtin_popcount(k & a) & 1) ? -1.0 : 1.0; };
  float* t32 = (float*)(ext + 576);
  for (int al = 0; al < 24; ++al)
    for (int be = 0; be < 24; ++be) {
      const int k = al / 3, d = al % 3, l = be / 3, e = be % 3;
      double acc = 0.0;
      for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b) acc += sgn(k, a) * sgn(l, b) * G0[(a * 3 + d) * 24 + b * 3 + e];
      ext[al * 24 + be] = acc / 64.0;
      t32[al * 24 + be] = (float)(acc / 64.0);
    }
}

}  // namespace pmt

using namespace pmt;

extern "C" {

int pmt_k0_extend(int dim, const double* K0_host, double* ext_host) {
  if ((dim != 2 && dim != 3) || !K0_host || !ext_host) return arg_error("pmt_k0_extend: bad argument");
  const int ne = (1 << dim) * dim;
  for (int q = 0; q < PMT_K0_EXT_LEN; ++q) ext_host[q] = 0.0;
  for (int q = 0; q < ne * ne; ++q) ext_host[q] = K0_host[q];
  int bits = k0_matches_pattern(dim, K0_host) ? PMT_K0_CLASS_PATTERN : 0;
  if (dim == 3) {
    double tab[72];
    if (k0_reflection_table(K0_host, tab)) {
      float* t32 = (float*)(ext_host + kK0ReflOffset32);
      for (int q = 0; q < 72; ++q) {
        ext_host[kK0ReflOffset64 + q] = tab[q];
        t32[q] = (float)tab[q];
      }
      bits |= PMT_K0_REFLECTION;
      double scale = 0.0;
      for (int q = 0; q < 72; ++q) scale = fmax(scale, fabs(tab[q]));
      bool rigid = true;  // the translations' entries M_(2^d)[d][d] vanish (K0 annihilates translations)
      for (int d = 0; d < 3; ++d) rigid = rigid && fabs(tab[(1 << d) * 9 + d * 4]) <= 1e-13 * scale;
      if (rigid) bits |= PMT_K0_RIGID;
      if (rigid && k0_cube_form(tab, scale)) {
        bits |= PMT_K0_CUBE;
        float* t32 = (float*)(ext_host + kK0ReflOffset32);
        for (int q = 0; q < 72; ++q) {  // the symmetrised table for every kernel that reads it
          ext_host[kK0ReflOffset64 + q] = tab[q];
          t32[q] = (float)tab[q];
        }
      }
    }
  }
  return bits;
}

}  // extern "C"
