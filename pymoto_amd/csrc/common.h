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

constexpr int kReduceBlocks = 1024;  // fixed grid for deterministic two-stage reductions
constexpr int kReduceThreads = 256;

}  // namespace pmt
