// Kernels of the learned warm start: a 3D convolutional encoder-decoder on the element grid with
// BF16 activations and weights and FP32 accumulation (forward, backward to the input, backward to
// the weights), plus the pooling, upsampling, masking and optimiser kernels around them.
//
// Layout: activations channels-last, X[voxel][c] with voxel = (z*ny + y)*nx + x (the element order
// of pymoto/common/domain.py:200-251) and c contiguous, BF16; weights W[tap][cout][cin] BF16 with
// tap = (dz+1)*9 + (dy+1)*3 + (dx+1); zero padding outside the grid.
//
// Convolution as an implicit GEMM on the matrix cores (v_mfma_f32_16x16x32_bf16): a wave owns 16
// consecutive voxels (the voxel count is a multiple of 16) and all output channels; the reduction
// runs over k = tap*cin + ci in chunks of 32, each lane loading 8 consecutive channels of one
// neighbour voxel (16 bytes) for the A operand and 8 consecutive input channels of one output
// channel for the B operand.  The backward pass to the input is the same kernel on the flipped,
// transposed weights.  The weight gradient is a per-tap GEMM over voxels (M = cout, N = cin,
// K = voxels), accumulated per block and reduced over blocks in a fixed order (deterministic).
#include "launch.h"

namespace pmt {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));

__device__ inline unsigned short f2bf(float f) {  // round to nearest even (finite inputs)
  const unsigned u = __float_as_uint(f);
  return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline float bf2f(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }

struct Vox {
  int nx, ny, nz;
  long long n;
};

__device__ inline bool tap_neighbour(const Vox& g, int x, int y, int z, int tap, long long* m) {
  const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
  const int xx = x + dx, yy = y + dy, zz = z + dz;
  if (xx < 0 || xx >= g.nx || yy < 0 || yy >= g.ny || zz < 0 || zz >= g.nz) return false;
  *m = ((long long)zz * g.ny + yy) * g.nx + xx;
  return true;
}

// Y[v][co] = act(bias[co] + sum_{tap, ci} W[tap][co][ci] * X[v + off(tap)][ci])
template <int CIN, int COUT, bool RELU>
__global__ void __launch_bounds__(256) conv3d_fwd_kernel(Vox g, const unsigned short* __restrict__ X,
                                                         const unsigned short* __restrict__ W,
                                                         const float* __restrict__ bias,
                                                         unsigned short* __restrict__ Y) {
  constexpr int KCH = (27 * CIN + 31) / 32;
  constexpr int NT = COUT / 16;
  const int lane = threadIdx.x & 63;
  const long long seg = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const long long v0 = seg * 16;
  if (v0 >= g.n) return;
  const long long va = v0 + (lane & 15);  // A-operand row (voxel) of this lane
  const int xa = (int)(va % g.nx);
  const long long t = va / g.nx;
  const int y = (int)(t % g.ny), z = (int)(t / g.ny);
  const int kq = lane >> 4;
  f4 acc[NT];
#pragma unroll
  for (int ti = 0; ti < NT; ++ti) acc[ti] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int q = 0; q < KCH; ++q) {
    const int k0 = 32 * q + 8 * kq;
    const int tap = k0 / CIN, ci0 = k0 % CIN;
    u4v ra = {0u, 0u, 0u, 0u};
    long long m;
    if (tap < 27 && tap_neighbour(g, xa, y, z, tap, &m)) ra = *(const u4v*)(X + m * CIN + ci0);
    const bf16x8 A = __builtin_bit_cast(bf16x8, ra);
#pragma unroll
    for (int ti = 0; ti < NT; ++ti) {
      u4v rb = {0u, 0u, 0u, 0u};
      if (tap < 27) rb = *(const u4v*)(W + ((long long)tap * COUT + 16 * ti + (lane & 15)) * CIN + ci0);
      acc[ti] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, __builtin_bit_cast(bf16x8, rb), acc[ti], 0, 0, 0);
    }
  }
  // D[row = 4*(lane/16) + r][col = lane%16]: voxel v0 + row, channel 16*ti + col
#pragma unroll
  for (int ti = 0; ti < NT; ++ti) {
    const int co = 16 * ti + (lane & 15);
    const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float v = acc[ti][r] + bv;
      if (RELU) v = v > 0.0f ? v : 0.0f;
      Y[(v0 + 4 * kq + r) * COUT + co] = f2bf(v);
    }
  }
}

// Wt[tap][ci][co] = W[26 - tap][co][ci]: the backward pass to the input is a forward convolution of
// the output gradient with these weights
__global__ void conv3d_flip_kernel(int cin, int cout, const unsigned short* __restrict__ W,
                                   unsigned short* __restrict__ Wt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = 27 * cin * cout;
  if (i >= n) return;
  const int tap = i / (cin * cout), rem = i % (cin * cout), ci = rem / cout, co = rem % cout;
  Wt[i] = W[((26 - tap) * cout + co) * cin + ci];
}

// partial dW[tap][co][ci] of one block = sum over its voxels of G[v][co] * X[v + off(tap)][ci];
// grid (nblocks, 27).  Each wave walks groups of 32 voxels; the block's 4 waves are summed in LDS.
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) conv3d_wgrad_kernel(Vox g, const unsigned short* __restrict__ X,
                                                           const unsigned short* __restrict__ G,
                                                           float* __restrict__ partial) {
  constexpr int NI = COUT / 16, NJ = (CIN + 15) / 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, kq = lane >> 4;
  const int tap = blockIdx.y;
  f4 acc[NI][NJ];
#pragma unroll
  for (int a = 0; a < NI; ++a)
#pragma unroll
    for (int b = 0; b < NJ; ++b) acc[a][b] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  const long long ngroups = (g.n + 31) / 32;
  for (long long grp = (long long)blockIdx.x * 4 + wave; grp < ngroups; grp += (long long)gridDim.x * 4) {
    // this lane's 8 voxels: 32*grp + 8*kq + s
    long long vv[8], mm[8];
    bool in[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      vv[s] = 32 * grp + 8 * kq + s;
      in[s] = false;
      mm[s] = 0;
      if (vv[s] < g.n) {
        const int x = (int)(vv[s] % g.nx);
        const long long t = vv[s] / g.nx;
        in[s] = tap_neighbour(g, x, (int)(t % g.ny), (int)(t / g.ny), tap, &mm[s]);
      }
    }
    bf16x8 A[NI], B[NJ];
#pragma unroll
    for (int a = 0; a < NI; ++a) {
      unsigned w[4];
#pragma unroll
      for (int s2 = 0; s2 < 4; ++s2) {
        const int co = 16 * a + (lane & 15);
        const unsigned lo = vv[2 * s2] < g.n ? G[vv[2 * s2] * COUT + co] : 0u;
        const unsigned hi = vv[2 * s2 + 1] < g.n ? G[vv[2 * s2 + 1] * COUT + co] : 0u;
        w[s2] = lo | (hi << 16);
      }
      A[a] = __builtin_bit_cast(bf16x8, u4v{w[0], w[1], w[2], w[3]});
    }
#pragma unroll
    for (int b = 0; b < NJ; ++b) {
      const int ci = 16 * b + (lane & 15);
      unsigned w[4];
#pragma unroll
      for (int s2 = 0; s2 < 4; ++s2) {
        const unsigned lo = (ci < CIN && in[2 * s2]) ? X[mm[2 * s2] * CIN + ci] : 0u;
        const unsigned hi = (ci < CIN && in[2 * s2 + 1]) ? X[mm[2 * s2 + 1] * CIN + ci] : 0u;
        w[s2] = lo | (hi << 16);
      }
      B[b] = __builtin_bit_cast(bf16x8, u4v{w[0], w[1], w[2], w[3]});
    }
#pragma unroll
    for (int a = 0; a < NI; ++a)
#pragma unroll
      for (int b = 0; b < NJ; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[a], B[b], acc[a][b], 0, 0, 0);
  }
  // D[row = 4*(lane/16) + r][col = lane%16] -> co = 16a + row, ci = 16b + col
  __shared__ float red[4][COUT * CIN];
#pragma unroll
  for (int a = 0; a < NI; ++a)
#pragma unroll
    for (int b = 0; b < NJ; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int co = 16 * a + 4 * kq + r, ci = 16 * b + (lane & 15);
        if (ci < CIN) red[wave][co * CIN + ci] = acc[a][b][r];
      }
  __syncthreads();
  float* out = partial + ((long long)blockIdx.x * 27 + tap) * COUT * CIN;
  for (int i = threadIdx.x; i < COUT * CIN; i += 256) out[i] = ((red[0][i] + red[1][i]) + red[2][i]) + red[3][i];
}

// dW[i] = sum_b partial[b][i] (fixed order); dbias[co] = sum_v G[v][co] in the same pass (grid 1)
__global__ void conv3d_wgrad_reduce(int nblocks, int m, const float* __restrict__ partial, float* __restrict__ dW) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  float s = 0.0f;
  for (int b = 0; b < nblocks; ++b) s += partial[(long long)b * m + i];
  dW[i] = s;
}

// dbias[co] = sum_v G[v][co]: one block per channel, fixed-order tree
__global__ void __launch_bounds__(256) channel_sum_kernel(long long n, int c, const unsigned short* __restrict__ G,
                                                          float* __restrict__ out) {
  const int co = blockIdx.x;
  float s = 0.0f;
  for (long long v = threadIdx.x; v < n; v += blockDim.x) s += bf2f(G[v * c + co]);
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[co] = red[0];
}

// G *= (Y > 0)  (ReLU backward from the layer output)
__global__ void relu_mask_kernel(long long n, const unsigned short* __restrict__ Y, unsigned short* __restrict__ G) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(bf2f(Y[i]) > 0.0f)) G[i] = 0;
}

// 2x2x2 average pooling (fine grid nx, ny, nz even) and its adjoint
__global__ void pool2_kernel(Vox f, int c, const unsigned short* __restrict__ X, unsigned short* __restrict__ Y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int cx = f.nx / 2, cy = f.ny / 2, cz = f.nz / 2;
  const long long n = (long long)cx * cy * cz * c;
  if (i >= n) return;
  const int ch = (int)(i % c);
  const long long v = i / c;
  const int x = (int)(v % cx), y = (int)((v / cx) % cy), z = (int)(v / ((long long)cx * cy));
  float s = 0.0f;
  for (int d = 0; d < 8; ++d) {
    const long long m = ((long long)(2 * z + (d >> 2)) * f.ny + (2 * y + ((d >> 1) & 1))) * f.nx + 2 * x + (d & 1);
    s += bf2f(X[m * c + ch]);
  }
  Y[i] = f2bf(0.125f * s);
}

__global__ void pool2_bwd_kernel(Vox f, int c, const unsigned short* __restrict__ dY, unsigned short* __restrict__ dX) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.n * c) return;
  const int ch = (int)(i % c);
  const long long v = i / c;
  const int x = (int)(v % f.nx), y = (int)((v / f.nx) % f.ny), z = (int)(v / ((long long)f.nx * f.ny));
  const long long p = ((long long)(z / 2) * (f.ny / 2) + y / 2) * (f.nx / 2) + x / 2;
  dX[i] = f2bf(0.125f * bf2f(dY[p * c + ch]));
}

// Y[fine] = X[parent] + S[fine] (nearest upsampling plus skip connection); the adjoint to X sums the
// 8 children (the adjoint to S is the identity)
__global__ void up2_add_kernel(Vox f, int c, const unsigned short* __restrict__ X, const unsigned short* __restrict__ S,
                               unsigned short* __restrict__ Y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.n * c) return;
  const int ch = (int)(i % c);
  const long long v = i / c;
  const int x = (int)(v % f.nx), y = (int)((v / f.nx) % f.ny), z = (int)(v / ((long long)f.nx * f.ny));
  const long long p = ((long long)(z / 2) * (f.ny / 2) + y / 2) * (f.nx / 2) + x / 2;
  Y[i] = f2bf(bf2f(X[p * c + ch]) + bf2f(S[i]));
}

__global__ void up2_bwd_kernel(Vox f, int c, const unsigned short* __restrict__ dY, unsigned short* __restrict__ dX) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int cx = f.nx / 2, cy = f.ny / 2, cz = f.nz / 2;
  if (i >= (long long)cx * cy * cz * c) return;
  const int ch = (int)(i % c);
  const long long v = i / c;
  const int x = (int)(v % cx), y = (int)((v / cx) % cy), z = (int)(v / ((long long)cx * cy));
  float s = 0.0f;
  for (int d = 0; d < 8; ++d) {
    const long long m = ((long long)(2 * z + (d >> 2)) * f.ny + (2 * y + ((d >> 1) & 1))) * f.nx + 2 * x + (d & 1);
    s += bf2f(dY[m * c + ch]);
  }
  dX[i] = f2bf(s);
}

// Nodal displacement from the first 3 channels of the element output: the mean over the elements
// around each node; and its adjoint (hex8, dof = node*3 + d)
__global__ void elem_to_node_kernel(Vox e, int c, const unsigned short* __restrict__ Xe, float* __restrict__ u) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = e.nx + 1, ny = e.ny + 1, nz = e.nz + 1;
  if (q >= (long long)nx * ny * nz * 3) return;
  const int d = (int)(q % 3);
  const long long n = q / 3;
  const int i = (int)(n % nx), j = (int)((n / nx) % ny), k = (int)(n / ((long long)nx * ny));
  float s = 0.0f;
  int cnt = 0;
  for (int a = 0; a < 8; ++a) {
    const int ei = i - (a & 1), ej = j - ((a >> 1) & 1), ek = k - ((a >> 2) & 1);
    if (ei < 0 || ei >= e.nx || ej < 0 || ej >= e.ny || ek < 0 || ek >= e.nz) continue;
    s += bf2f(Xe[(((long long)ek * e.ny + ej) * e.nx + ei) * c + d]);
    ++cnt;
  }
  u[q] = s / (float)cnt;
}

__global__ void elem_to_node_bwd_kernel(Vox e, int c, const float* __restrict__ du, unsigned short* __restrict__ dXe) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= e.n * c) return;
  const int ch = (int)(q % c);
  const long long v = q / c;
  if (ch >= 3) {
    dXe[q] = 0;
    return;
  }
  const int nx = e.nx + 1, ny = e.ny + 1, nz = e.nz + 1;
  const int ei = (int)(v % e.nx), ej = (int)((v / e.nx) % e.ny), ek = (int)(v / ((long long)e.nx * e.ny));
  float s = 0.0f;
  for (int a = 0; a < 8; ++a) {
    const int i = ei + (a & 1), j = ej + ((a >> 1) & 1), k = ek + ((a >> 2) & 1);
    const int cnt = (i > 0 && i < nx - 1 ? 2 : 1) * (j > 0 && j < ny - 1 ? 2 : 1) * (k > 0 && k < nz - 1 ? 2 : 1);
    s += du[(((long long)k * ny + j) * nx + i) * 3 + ch] / (float)cnt;
  }
  dXe[q] = f2bf(s);
}

// Adam step on FP32 master weights, with the BF16 copy refreshed (t is the 1-based step)
__global__ void adam_kernel(long long n, float lr, float b1, float b2, float eps, int t, const float* __restrict__ grad,
                            float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                            unsigned short* __restrict__ wb) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gr = grad[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gr;
  const float vi = b2 * v[i] + (1.0f - b2) * gr * gr;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.0f - powf(b1, (float)t)), vh = vi / (1.0f - powf(b2, (float)t));
  const float wi = w[i] - lr * mh / (sqrtf(vh) + eps);
  w[i] = wi;
  if (wb) wb[i] = f2bf(wi);
}

__global__ void add_bf16_kernel(long long n, const unsigned short* __restrict__ a, const unsigned short* __restrict__ b,
                                unsigned short* __restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(bf2f(a[i]) + bf2f(b[i]));
}

__global__ void f32_to_bf16_kernel(long long n, const float* __restrict__ x, unsigned short* __restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

__global__ void bf16_to_f32_kernel(long long n, const unsigned short* __restrict__ x, float* __restrict__ y) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = bf2f(x[i]);
}

// ---------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------
static Vox vox(int nx, int ny, int nz) {
  Vox g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.n = (long long)nx * ny * nz;
  return g;
}

static bool conv_shape_ok(int nx, int ny, int nz, int cin, int cout) {
  const bool c_ok = (cin == 8 || cin == 16 || cin == 32 || cin == 64) && (cout == 16 || cout == 32 || cout == 64);
  return c_ok && nx > 0 && ny > 0 && nz > 0 && ((long long)nx * ny * nz) % 16 == 0;
}

#define PMT_CONV_CASES(F) \
  F(8, 16) F(16, 16) F(16, 32) F(16, 64) F(32, 16) F(32, 32) F(32, 64) F(64, 16) F(64, 32) F(64, 64)

static int conv_fwd(Vox g, int cin, int cout, const unsigned short* X, const unsigned short* W, const float* bias,
                    int relu, unsigned short* Y, hipStream_t s) {
  const long long segs = g.n / 16;
  dim3 grid((unsigned)((segs + 3) / 4));
#define F(CI, CO)                                                                   \
  if (cin == CI && cout == CO) {                                                    \
    if (relu)                                                                       \
      conv3d_fwd_kernel<CI, CO, true><<<grid, 256, 0, s>>>(g, X, W, bias, Y);       \
    else                                                                            \
      conv3d_fwd_kernel<CI, CO, false><<<grid, 256, 0, s>>>(g, X, W, bias, Y);      \
    return check_launch("conv3d_fwd_kernel");                                       \
  }
  PMT_CONV_CASES(F)
#undef F
  return arg_error("conv3d: unsupported channel counts");
}

}  // namespace pmt

using namespace pmt;

extern "C" {

int pmt_conv3d_fwd(int nx, int ny, int nz, int cin, int cout, const void* X, const void* W, const float* bias,
                   int relu, void* Y, void* stream) {
  if (!conv_shape_ok(nx, ny, nz, cin, cout) || !X || !W || !Y)
    return arg_error("pmt_conv3d_fwd: bad argument (cin in 8/16/32/64, cout in 16/32/64, voxels % 16 == 0)");
  return conv_fwd(vox(nx, ny, nz), cin, cout, (const unsigned short*)X, (const unsigned short*)W, bias, relu,
                  (unsigned short*)Y, (hipStream_t)stream);
}

int pmt_conv3d_bwd_data(int nx, int ny, int nz, int cin, int cout, const void* dY, const void* W, void* Wt, void* dX,
                        void* stream) {
  if (!conv_shape_ok(nx, ny, nz, cout, cin) || cin < 16 || !dY || !W || !Wt || !dX)
    return arg_error("pmt_conv3d_bwd_data: bad argument (cin in 16/32/64, cout in 16/32/64, voxels % 16 == 0)");
  hipStream_t s = (hipStream_t)stream;
  const int n = 27 * cin * cout;
  conv3d_flip_kernel<<<blocks_for(n, 256), 256, 0, s>>>(cin, cout, (const unsigned short*)W, (unsigned short*)Wt);
  PMT_TRY(check_launch("conv3d_flip_kernel"));
  return conv_fwd(vox(nx, ny, nz), cout, cin, (const unsigned short*)dY, (const unsigned short*)Wt, nullptr, 0,
                  (unsigned short*)dX, s);
}

int pmt_conv3d_wgrad_work_size(int cin, int cout) { return 64 * 27 * cin * cout; }

int pmt_conv3d_bwd_weight(int nx, int ny, int nz, int cin, int cout, const void* X, const void* dY, float* dW,
                          float* dbias, float* work, void* stream) {
  if (!conv_shape_ok(nx, ny, nz, cin, cout) || !X || !dY || !dW || !work)
    return arg_error("pmt_conv3d_bwd_weight: bad argument");
  hipStream_t s = (hipStream_t)stream;
  const Vox g = vox(nx, ny, nz);
  const int nb = 64;
  dim3 grid(nb, 27);
  bool done = false;
#define F(CI, CO)                                                                                             \
  if (!done && cin == CI && cout == CO) {                                                                     \
    conv3d_wgrad_kernel<CI, CO><<<grid, 256, 0, s>>>(g, (const unsigned short*)X, (const unsigned short*)dY, \
                                                     work);                                                   \
    done = true;                                                                                              \
  }
  PMT_CONV_CASES(F)
#undef F
  if (!done) return arg_error("pmt_conv3d_bwd_weight: unsupported channel counts");
  PMT_TRY(check_launch("conv3d_wgrad_kernel"));
  const int m = 27 * cin * cout;
  conv3d_wgrad_reduce<<<blocks_for(m, 256), 256, 0, s>>>(nb, m, work, dW);
  PMT_TRY(check_launch("conv3d_wgrad_reduce"));
  if (dbias) {
    channel_sum_kernel<<<cout, 256, 0, s>>>(g.n, cout, (const unsigned short*)dY, dbias);
    PMT_TRY(check_launch("channel_sum_kernel"));
  }
  return PMT_OK;
}

int pmt_relu_mask_bf16(long long n, const void* Y, void* G, void* stream) {
  if (n < 0 || (n > 0 && (!Y || !G))) return arg_error("pmt_relu_mask_bf16: bad argument");
  if (n == 0) return PMT_OK;
  relu_mask_kernel<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(n, (const unsigned short*)Y,
                                                                        (unsigned short*)G);
  return check_launch("relu_mask_kernel");
}

int pmt_pool2_bf16(int nx, int ny, int nz, int c, const void* X, void* Y, void* stream) {
  if (nx % 2 || ny % 2 || nz % 2 || c <= 0 || !X || !Y) return arg_error("pmt_pool2_bf16: bad argument");
  const long long n = (long long)(nx / 2) * (ny / 2) * (nz / 2) * c;
  pool2_kernel<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(vox(nx, ny, nz), c, (const unsigned short*)X,
                                                                    (unsigned short*)Y);
  return check_launch("pool2_kernel");
}

int pmt_pool2_bwd_bf16(int nx, int ny, int nz, int c, const void* dY, void* dX, void* stream) {
  if (nx % 2 || ny % 2 || nz % 2 || c <= 0 || !dY || !dX) return arg_error("pmt_pool2_bwd_bf16: bad argument");
  const Vox f = vox(nx, ny, nz);
  pool2_bwd_kernel<<<blocks_for(f.n * c, 256), 256, 0, (hipStream_t)stream>>>(f, c, (const unsigned short*)dY,
                                                                              (unsigned short*)dX);
  return check_launch("pool2_bwd_kernel");
}

int pmt_up2_add_bf16(int nx, int ny, int nz, int c, const void* X, const void* S, void* Y, void* stream) {
  if (nx % 2 || ny % 2 || nz % 2 || c <= 0 || !X || !S || !Y) return arg_error("pmt_up2_add_bf16: bad argument");
  const Vox f = vox(nx, ny, nz);
  up2_add_kernel<<<blocks_for(f.n * c, 256), 256, 0, (hipStream_t)stream>>>(f, c, (const unsigned short*)X,
                                                                            (const unsigned short*)S,
                                                                            (unsigned short*)Y);
  return check_launch("up2_add_kernel");
}

int pmt_up2_bwd_bf16(int nx, int ny, int nz, int c, const void* dY, void* dX, void* stream) {
  if (nx % 2 || ny % 2 || nz % 2 || c <= 0 || !dY || !dX) return arg_error("pmt_up2_bwd_bf16: bad argument");
  const long long n = (long long)(nx / 2) * (ny / 2) * (nz / 2) * c;
  up2_bwd_kernel<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(vox(nx, ny, nz), c, (const unsigned short*)dY,
                                                                      (unsigned short*)dX);
  return check_launch("up2_bwd_kernel");
}

int pmt_elem_to_node(int nelx, int nely, int nelz, int c, const void* Xe, float* u, void* stream) {
  if (nelx <= 0 || nely <= 0 || nelz <= 0 || c < 3 || !Xe || !u) return arg_error("pmt_elem_to_node: bad argument");
  const long long n = (long long)(nelx + 1) * (nely + 1) * (nelz + 1) * 3;
  elem_to_node_kernel<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(vox(nelx, nely, nelz), c,
                                                                          (const unsigned short*)Xe, u);
  return check_launch("elem_to_node_kernel");
}

int pmt_elem_to_node_bwd(int nelx, int nely, int nelz, int c, const float* du, void* dXe, void* stream) {
  if (nelx <= 0 || nely <= 0 || nelz <= 0 || c < 3 || !du || !dXe)
    return arg_error("pmt_elem_to_node_bwd: bad argument");
  const Vox e = vox(nelx, nely, nelz);
  elem_to_node_bwd_kernel<<<blocks_for(e.n * c, 256), 256, 0, (hipStream_t)stream>>>(e, c, du, (unsigned short*)dXe);
  return check_launch("elem_to_node_bwd_kernel");
}

int pmt_adam_step(long long n, float lr, float b1, float b2, float eps, int t, const float* grad, float* w, float* m,
                  float* v, void* wb, void* stream) {
  if (n < 0 || t < 1 || (n > 0 && (!grad || !w || !m || !v))) return arg_error("pmt_adam_step: bad argument");
  if (n == 0) return PMT_OK;
  adam_kernel<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(n, lr, b1, b2, eps, t, grad, w, m, v,
                                                                  (unsigned short*)wb);
  return check_launch("adam_kernel");
}

int pmt_add_bf16(long long n, const void* a, const void* b, void* y, void* stream) {
  if (n < 0 || (n > 0 && (!a || !b || !y))) return arg_error("pmt_add_bf16: bad argument");
  if (n == 0) return PMT_OK;
  add_bf16_kernel<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(n, (const unsigned short*)a,
                                                                       (const unsigned short*)b, (unsigned short*)y);
  return check_launch("add_bf16_kernel");
}

int pmt_f32_to_bf16(long long n, const float* x, void* y, void* stream) {
  if (n < 0 || (n > 0 && (!x || !y))) return arg_error("pmt_f32_to_bf16: bad argument");
  if (n == 0) return PMT_OK;
  f32_to_bf16_kernel<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(n, x, (unsigned short*)y);
  return check_launch("f32_to_bf16_kernel");
}

int pmt_bf16_to_f32(long long n, const void* x, float* y, void* stream) {
  if (n < 0 || (n > 0 && (!x || !y))) return arg_error("pmt_bf16_to_f32: bad argument");
  if (n == 0) return PMT_OK;
  bf16_to_f32_kernel<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(n, (const unsigned short*)x, y);
  return check_launch("bf16_to_f32_kernel");
}

}  // extern "C"
