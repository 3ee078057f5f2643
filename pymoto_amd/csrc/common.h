// Shared helpers for the gfx950 kernels of the FE hot path.
//
// Grid conventions (pymoto/common/domain.py:200-251):
//   element e = (k*nely + j)*nelx + i, node n = (k*(nely+1) + j)*(nelx+1) + i, dof = n*D + d,
//   local node a of an element sits at offset (a&1, (a>>1)&1, (a>>2)&1).
// In 2D (nelz == 0) the z extent is one node layer and D = 2.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "pymoto_amd.h"

namespace pmt {

void set_error(const std::string& msg);
int hip_status(hipError_t err, const char* where);
int check_launch(const char* where);

#define PMT_TRY(expr)                          \
  do {                                         \
    int _rc = (expr);                          \
    if (_rc != PMT_OK) return _rc;             \
  } while (0)

#define PMT_HIP(expr) PMT_TRY(::pmt::hip_status((expr), #expr))

inline int arg_error(const char* msg) {
  set_error(msg);
  return PMT_ERR_ARG;
}

struct GridDims {
  int dim;          // 2 or 3
  int ex, ey, ez;   // elements per direction (ez = 1 in 2D)
  int nx, ny, nz;   // nodes per direction (nz = 1 in 2D)
  long long nel, nnodes;
};

inline GridDims make_grid(int nelx, int nely, int nelz) {
  GridDims g;
  g.dim = nelz > 0 ? 3 : 2;
  g.ex = nelx;
  g.ey = nely;
  g.ez = nelz > 0 ? nelz : 1;
  g.nx = nelx + 1;
  g.ny = nely + 1;
  g.nz = nelz > 0 ? nelz + 1 : 1;
  g.nel = (long long)g.ex * g.ey * g.ez;
  g.nnodes = (long long)g.nx * g.ny * g.nz;
  return g;
}

inline int grid_ok(int nelx, int nely, int nelz) { return nelx > 0 && nely > 0 && nelz >= 0; }

inline unsigned blocks_for(long long n, int threads) {
  long long b = (n + threads - 1) / threads;
  return (unsigned)(b < 1 ? 1 : b);
}

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that each XCD walks one
// contiguous range of tiles and neighbouring tiles share its L2.  The grid is padded to a
// multiple of 8 blocks.
__device__ inline long long xcd_block() {
  const unsigned per = gridDim.x >> 3;
  return (long long)(blockIdx.x & 7u) * per + (blockIdx.x >> 3);
}

// Trilinear (bilinear) prolongation weight of coarse corner A at local fine node a of child c of a
// coarse element: per dimension f = child bit + node bit in {0,1,2}, w(0,0) = w(2,1) = 1, w(1,*) = 1/2.
__host__ __device__ inline double interp_w(int f, int Abit) {
  return f == 1 ? 0.5 : ((f == 0) == (Abit == 0) ? 1.0 : 0.0);
}
template <int DIM>
__host__ __device__ inline double pweight(int c, int a, int A) {
  double w = 1.0;
#pragma unroll
  for (int b = 0; b < DIM; ++b) w *= interp_w(((c >> b) & 1) + ((a >> b) & 1), (A >> b) & 1);
  return w;
}
__host__ __device__ inline double pweight3(int c, int a, int A) { return pweight<3>(c, a, A); }

// Output of one 3 x 3 block (offset o = (dz+1)*9 + (dy+1)*3 + dx+1) of a hex8 node stencil.  Full form:
// S[(o*3+d)*3+d2][n].  Half form (levels that smooth): the centre and forward blocks o >= 13 as
// H[((o-13)*3+d)*3+d2][n] (FP64 and, if H32 is given, FP32) and D^-1 from the centre block.
struct StencilOut {
  double* S = nullptr;
  double* H = nullptr;
  double* dinv = nullptr;
  float* H32 = nullptr;
  float* dinv32 = nullptr;
};
template <bool HALF>
__device__ inline void emit_block(const StencilOut& out, int o, const double (&blk)[3][3], long long n, long long nn) {
  if (!HALF) {
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int d2 = 0; d2 < 3; ++d2) out.S[((long long)(o * 3 + d) * 3 + d2) * nn + n] = blk[d][d2];
    return;
  }
  if (o < 13) return;
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int d2 = 0; d2 < 3; ++d2) {
      const long long c = (long long)((o - 13) * 3 + d) * 3 + d2;
      out.H[c * nn + n] = blk[d][d2];
      if (out.H32) out.H32[c * nn + n] = (float)blk[d][d2];
    }
  if (o == 13) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double v = 1.0 / blk[d][d];
      out.dinv[n * 3 + d] = v;
      if (out.dinv32) out.dinv32[n * 3 + d] = (float)v;
    }
  }
}

// Raw buffer descriptors (gfx9 family): the hardware range check returns 0 for a load and drops a
// store whose byte offset is outside [0, records), so a lane with nothing to access uses offset kOOB
// and every lane issues the same memory instructions - no branch around them, exact wait counters.
// A descriptor covers at most 2 GiB: callers build one per node plane or per coefficient row.
constexpr int kOOB = 0x7fffffff;
constexpr int kRsrcWord3 = 0x00020000;  // raw 32-bit buffer

template <typename T>
__device__ inline __amdgpu_buffer_rsrc_t plane_rsrc(const T* base, long long first, long long count, bool valid) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)(valid ? base + first : base), 0,
                                           valid ? (int)(count * (long long)sizeof(T)) : 0, kRsrcWord3);
}

typedef unsigned int pmt_u32x2 __attribute__((ext_vector_type(2)));

__device__ inline double bload(__amdgpu_buffer_rsrc_t r, int off, double) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ inline float bload(__amdgpu_buffer_rsrc_t r, int off, float) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ inline void bstore(__amdgpu_buffer_rsrc_t r, int off, double v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(pmt_u32x2, v), r, off, 0, 0);
}
__device__ inline void bstore(__amdgpu_buffer_rsrc_t r, int off, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, off, 0, 0);
}

constexpr int kReduceBlocks = 1024;  // fixed grid for deterministic two-stage reductions
constexpr int kReduceThreads = 256;

}  // namespace pmt
