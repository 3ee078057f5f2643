// Assembled CSR stiffness matrix in pyMOTO's exact format (pymoto/modules/assembly.py:100-296).
//
// Pattern (closed form, one thread per row): the columns of row (node n, dof d) are the dofs of
// the clipped 3^D node neighbourhood of n, in ascending node order (k, j, i lexicographic) and
// ascending dof - exactly the sorted, de-duplicated union of the element dofs the reference builds
// with argsort/unique.  Dirichlet rows/cols stay in the pattern.  indptr is a closed-form prefix
// sum over the (k, j, i) node order, so no scan is needed.
//
// Values (one thread per row): each slot sums x_e * K0[ij] over the elements holding both nodes,
// in ascending element number, every product rounded before the add (no FMA), starting from +0.0:
// bit-identical to `np.add.at(data, datamap, K0.ravel()[None,:] * x[:,None])` whose unbuffered
// accumulation visits contributions element-major (assembly.py:257, 267-268).  Dirichlet slots
// are 0 except the diagonal, which holds bcdiag (assembly.py:209-230, 270-272).
#include "common.h"

// Products must be rounded before the sums (np.add.at adds rounded products): no contraction
// into fused multiply-adds anywhere in this file.
#pragma clang fp contract(off)

namespace pmt {

__device__ __host__ inline long long line_cnt(int i, int n) {  // neighbours of coordinate i on a line of n nodes
  if (n == 1) return 1;
  return 3 - (i == 0) - (i == n - 1);
}

__device__ __host__ inline long long line_prefix(int i, int n) {  // sum_{i' < i} line_cnt(i', n)
  if (i <= 0) return 0;
  if (n == 1) return i;
  return 3LL * i - 1 - (i > n - 1 ? 1 : 0);
}

template <typename IT>
__global__ void csr_pattern_kernel(GridDims g, int D, IT* __restrict__ indptr, IT* __restrict__ indices) {
  const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nrows = g.nnodes * D;
  if (row > nrows) return;
  const long long Sx = line_prefix(g.nx, g.nx), Sy = line_prefix(g.ny, g.ny);
  if (row == nrows) {
    const long long Sz = line_prefix(g.nz, g.nz);
    indptr[nrows] = (IT)((long long)D * D * Sx * Sy * Sz);
    return;
  }
  const long long n = row / D;
  const int d = (int)(row - n * D);
  const int i = (int)(n % g.nx);
  const long long t = n / g.nx;
  const int j = (int)(t % g.ny);
  const int k = (int)(t / g.ny);
  const long long cx = line_cnt(i, g.nx), cy = line_cnt(j, g.ny), cz = line_cnt(k, g.nz);
  const long long before_nodes_cnt =
      line_prefix(k, g.nz) * Sy * Sx + cz * line_prefix(j, g.ny) * Sx + cz * cy * line_prefix(i, g.nx);
  const long long start = (long long)D * D * before_nodes_cnt + (long long)d * D * cx * cy * cz;
  indptr[row] = (IT)start;
  long long pos = start;
  const int k0 = k > 0 ? k - 1 : 0, k1 = k < g.nz - 1 ? k + 1 : g.nz - 1;
  const int j0 = j > 0 ? j - 1 : 0, j1 = j < g.ny - 1 ? j + 1 : g.ny - 1;
  const int i0 = i > 0 ? i - 1 : 0, i1 = i < g.nx - 1 ? i + 1 : g.nx - 1;
  for (int kk = k0; kk <= k1; ++kk)
    for (int jj = j0; jj <= j1; ++jj)
      for (int ii = i0; ii <= i1; ++ii) {
        const long long m = ((long long)kk * g.ny + jj) * g.nx + ii;
        for (int d2 = 0; d2 < D; ++d2) indices[pos++] = (IT)(m * D + d2);
      }
}

// D: unknowns per node (DIM for elasticity, 1 for scalar problems)
template <int DIM, int D, typename IT>
__global__ void csr_values_kernel(GridDims g, const double* __restrict__ K0, const double* __restrict__ xe,
                                  const uint8_t* __restrict__ bcm, double bcdiag, const IT* __restrict__ indptr,
                                  double* __restrict__ data) {
  constexpr int NN = 1 << DIM;
  constexpr int NE = NN * D;
  const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= g.nnodes * D) return;
  const long long n = row / D;
  const int d = (int)(row - n * D);
  const int i = (int)(n % g.nx);
  const long long t = n / g.nx;
  const int j = (int)(t % g.ny);
  const int k = (int)(t / g.ny);
  const bool row_fixed = bcm && ((bcm[n] >> d) & 1u);
  long long pos = (long long)indptr[row];
  const int k0 = k > 0 ? k - 1 : 0, k1 = k < g.nz - 1 ? k + 1 : g.nz - 1;
  const int j0 = j > 0 ? j - 1 : 0, j1 = j < g.ny - 1 ? j + 1 : g.ny - 1;
  const int i0 = i > 0 ? i - 1 : 0, i1 = i < g.nx - 1 ? i + 1 : g.nx - 1;
  for (int kk = k0; kk <= k1; ++kk)
    for (int jj = j0; jj <= j1; ++jj)
      for (int ii = i0; ii <= i1; ++ii) {
        const long long m = ((long long)kk * g.ny + jj) * g.nx + ii;
        const unsigned mk = bcm ? bcm[m] : 0u;
        for (int d2 = 0; d2 < D; ++d2) {
          double v = 0.0;
          if (row_fixed || ((mk >> d2) & 1u)) {
            v = (m == n && d2 == d && row_fixed) ? bcdiag : 0.0;
          } else {
            // elements around node n in ascending element number: (ez, ey, ex) lexicographic
            for (int ez = 0; ez < (DIM == 3 ? 2 : 1); ++ez)
              for (int ey = 0; ey < 2; ++ey)
                for (int ex = 0; ex < 2; ++ex) {
                  const int ei = i - 1 + ex, ej = j - 1 + ey, ek = DIM == 3 ? k - 1 + ez : 0;
                  if (ei < 0 || ei >= g.ex || ej < 0 || ej >= g.ey || ek < 0 || ek >= g.ez) continue;
                  const int bx = ii - ei, by = jj - ej, bz = DIM == 3 ? kk - ek : 0;
                  if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
                  const int locn = (i - ei) + 2 * (j - ej) + (DIM == 3 ? 4 * (k - ek) : 0);
                  const int locm = bx + 2 * by + 4 * bz;
                  const double xv = xe[((long long)ek * g.ey + ej) * g.ex + ei];
                  const double prod = K0[(locn * D + d) * NE + locm * D + d2] * xv;
                  v = v + prod;
                }
          }
          data[pos++] = v;
        }
      }
}

}  // namespace pmt

using namespace pmt;

extern "C" {

long long pmt_csr_nnz(int nelx, int nely, int nelz, int ndof) {
  if (!grid_ok(nelx, nely, nelz) || ndof <= 0) return PMT_ERR_ARG;
  GridDims g = make_grid(nelx, nely, nelz);
  return (long long)ndof * ndof * line_prefix(g.nx, g.nx) * line_prefix(g.ny, g.ny) * line_prefix(g.nz, g.nz);
}

int pmt_csr_pattern(int nelx, int nely, int nelz, int ndof, int index_bytes, void* indptr, void* indices,
                    void* stream) {
  if (!grid_ok(nelx, nely, nelz) || ndof <= 0 || !indptr || !indices || (index_bytes != 4 && index_bytes != 8))
    return arg_error("pmt_csr_pattern: bad argument");
  GridDims g = make_grid(nelx, nely, nelz);
  const long long nnz = pmt_csr_nnz(nelx, nely, nelz, ndof);
  if (index_bytes == 4 && nnz > 2147483647LL) return arg_error("pmt_csr_pattern: nnz does not fit int32 indices");
  const long long rows = g.nnodes * ndof + 1;
  if (index_bytes == 4)
    csr_pattern_kernel<int32_t><<<blocks_for(rows, 256), 256, 0, (hipStream_t)stream>>>(
        g, ndof, (int32_t*)indptr, (int32_t*)indices);
  else
    csr_pattern_kernel<int64_t><<<blocks_for(rows, 256), 256, 0, (hipStream_t)stream>>>(
        g, ndof, (int64_t*)indptr, (int64_t*)indices);
  return check_launch("csr_pattern_kernel");
}

int pmt_csr_values_ndof(int nelx, int nely, int nelz, int ndof, const double* K0, const double* xe,
                        const uint8_t* bcmask, double bcdiag, int index_bytes, const void* indptr, double* data,
                        void* stream) {
  if (!grid_ok(nelx, nely, nelz) || !K0 || !xe || !indptr || !data || (index_bytes != 4 && index_bytes != 8))
    return arg_error("pmt_csr_values: bad argument");
  GridDims g = make_grid(nelx, nely, nelz);
  if (ndof != 1 && ndof != g.dim) return arg_error("pmt_csr_values: ndof must be 1 or the grid dimension");
  const long long rows = g.nnodes * ndof;
  hipStream_t s = (hipStream_t)stream;
  const unsigned blocks = blocks_for(rows, 256);
#define PMT_CSRV(DIM, D, IT) \
  csr_values_kernel<DIM, D, IT><<<blocks, 256, 0, s>>>(g, K0, xe, bcmask, bcdiag, (const IT*)indptr, data)
  if (g.dim == 3 && ndof == 3) {
    if (index_bytes == 4) PMT_CSRV(3, 3, int32_t); else PMT_CSRV(3, 3, int64_t);
  } else if (g.dim == 3) {
    if (index_bytes == 4) PMT_CSRV(3, 1, int32_t); else PMT_CSRV(3, 1, int64_t);
  } else if (ndof == 2) {
    if (index_bytes == 4) PMT_CSRV(2, 2, int32_t); else PMT_CSRV(2, 2, int64_t);
  } else {
    if (index_bytes == 4) PMT_CSRV(2, 1, int32_t); else PMT_CSRV(2, 1, int64_t);
  }
#undef PMT_CSRV
  return check_launch("csr_values_kernel");
}

int pmt_csr_values(int nelx, int nely, int nelz, const double* K0, const double* xe, const uint8_t* bcmask,
                   double bcdiag, int index_bytes, const void* indptr, double* data, void* stream) {
  return pmt_csr_values_ndof(nelx, nely, nelz, nelz > 0 ? 3 : 2, K0, xe, bcmask, bcdiag, index_bytes, indptr, data,
                             stream);
}

}  // extern "C"
