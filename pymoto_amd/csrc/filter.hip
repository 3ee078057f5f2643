// FilterConv forward and adjoint (pymoto/modules/filter.py:8-220).
//
// The reference pads an element-index array axis by axis with numpy padding modes and then
// convolves.  Padding an index array along one axis applies the same 1-D index map to every
// line, so the padded field is xpad[a,b,c] = x[mx[a], my[b], mz[c]] with one small map per axis
// (built on the host at setup); constant-padded cells carry map = -1 and a value, with the later
// axis winning at corners as the reference's override order does.  Forward: one thread per element
// gathers the kx*ky*kz window ('valid' convolution).  Adjoint: one thread per element visits the
// padded cells that map onto it (per-axis inverse lists; interior elements have exactly one) and
// evaluates the 'full' correlation there - a gather, so no atomics and a deterministic sum.
#include <algorithm>

#include "launch.h"

namespace pmt {

__global__ void __launch_bounds__(256)
filter_fwd_kernel(int nx, int ny, int nz, int kx, int ky, int kz, const double* __restrict__ w,
                  const int32_t* __restrict__ map, const double* __restrict__ cval, const double* __restrict__ x,
                  double* __restrict__ y) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nel = (long long)nx * ny * nz;
  if (e >= nel) return;
  const int i = (int)(e % nx);
  const long long t = e / nx;
  const int j = (int)(t % ny);
  const int k = (int)(t / ny);
  const int Lx = nx + kx - 1, Ly = ny + ky - 1;
  const int32_t* mx = map;
  const int32_t* my = map + Lx;
  const int32_t* mz = map + Lx + Ly;
  const double* cx = cval;
  const double* cy = cval + Lx;
  const double* cz = cval + Lx + Ly;
  double acc = 0.0;
  for (int r = 0; r < kz; ++r) {
    const int c = k + kz - 1 - r;
    const int mc = mz[c];
    for (int q = 0; q < ky; ++q) {
      const int b = j + ky - 1 - q;
      const int mb = my[b];
      for (int p = 0; p < kx; ++p) {
        const int a = i + kx - 1 - p;
        const int ma = mx[a];
        double v;
        if (mc < 0)
          v = cz[c];
        else if (mb < 0)
          v = cy[b];
        else if (ma < 0)
          v = cx[a];
        else
          v = x[((long long)mc * ny + mb) * nx + ma];
        acc = fma(w[((long long)p * ky + q) * kz + r], v, acc);
      }
    }
  }
  y[e] = acc;
}

__global__ void __launch_bounds__(256)
filter_bwd_kernel(int nx, int ny, int nz, int kx, int ky, int kz, const double* __restrict__ w,
                  const int32_t* __restrict__ inv_ptr, const int32_t* __restrict__ inv_idx,
                  const double* __restrict__ dy, double* __restrict__ dx) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nel = (long long)nx * ny * nz;
  if (e >= nel) return;
  const int i = (int)(e % nx);
  const long long t = e / nx;
  const int j = (int)(t % ny);
  const int k = (int)(t / ny);
  const int32_t* px = inv_ptr;
  const int32_t* py = inv_ptr + (nx + 1);
  const int32_t* pz = inv_ptr + (nx + 1) + (ny + 1);
  double acc = 0.0;
  for (int ic = pz[k]; ic < pz[k + 1]; ++ic) {
    const int c = inv_idx[ic];
    for (int ib = py[j]; ib < py[j + 1]; ++ib) {
      const int b = inv_idx[ib];
      for (int ia = px[i]; ia < px[i + 1]; ++ia) {
        const int a = inv_idx[ia];
        // correlate 'full' at padded cell (a, b, c)
        for (int r = 0; r < kz; ++r) {
          const int cc = c - (kz - 1) + r;
          if (cc < 0 || cc >= nz) continue;
          for (int q = 0; q < ky; ++q) {
            const int bb = b - (ky - 1) + q;
            if (bb < 0 || bb >= ny) continue;
            for (int p = 0; p < kx; ++p) {
              const int aa = a - (kx - 1) + p;
              if (aa < 0 || aa >= nx) continue;
              acc = fma(w[((long long)p * ky + q) * kz + r], dy[((long long)cc * ny + bb) * nx + aa], acc);
            }
          }
        }
      }
    }
  }
  dx[e] = acc;
}

// The same two kernels with the kernel extents fixed at compile time (KXY x KXY x KZ, the shapes that
// FilterConv builds from a radius): the per-axis maps are read once per element and the KXY^2 KZ taps
// are unrolled, so every load of a thread is independent and in flight at once (the general kernels
// chain them through their loops).  Same summation order as the general kernels.
template <int KXY, int KZ>
__global__ void __launch_bounds__(256)
filter_fwd_fixed_kernel(int nx, int ny, int nz, const double* __restrict__ w, const int32_t* __restrict__ map,
                        const double* __restrict__ cval, const double* __restrict__ x, double* __restrict__ y) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nel = (long long)nx * ny * nz;
  if (e >= nel) return;
  const int i = (int)(e % nx);
  const long long t = e / nx;
  const int j = (int)(t % ny);
  const int k = (int)(t / ny);
  const int Lx = nx + KXY - 1, Ly = ny + KXY - 1;
  int ma[KXY], mb[KXY], mc[KZ];
#pragma unroll
  for (int p = 0; p < KXY; ++p) ma[p] = map[i + KXY - 1 - p];
#pragma unroll
  for (int q = 0; q < KXY; ++q) mb[q] = map[Lx + j + KXY - 1 - q];
#pragma unroll
  for (int r = 0; r < KZ; ++r) mc[r] = map[Lx + Ly + k + KZ - 1 - r];
  double acc = 0.0;
#pragma unroll
  for (int r = 0; r < KZ; ++r)
#pragma unroll
    for (int q = 0; q < KXY; ++q)
#pragma unroll
      for (int p = 0; p < KXY; ++p) {
        double v;
        if (mc[r] < 0)
          v = cval[Lx + Ly + k + KZ - 1 - r];
        else if (mb[q] < 0)
          v = cval[Lx + j + KXY - 1 - q];
        else if (ma[p] < 0)
          v = cval[i + KXY - 1 - p];
        else
          v = x[((long long)mc[r] * ny + mb[q]) * nx + ma[p]];
        acc = fma(w[(p * KXY + q) * KZ + r], v, acc);
      }
  y[e] = acc;
}

// The element's padded cells are exactly its own position shifted by the half widths (the element is at
// least (K-1)/2 from every mirrored, replicated or wrapped face, or the padding is constant).
template <int KXY, int KZ>
__device__ inline bool filter_single_cell(const int32_t* __restrict__ inv_ptr, const int32_t* __restrict__ inv_idx,
                                          int nx, int ny, int i, int j, int k) {
  const int32_t* px = inv_ptr;
  const int32_t* py = inv_ptr + (nx + 1);
  const int32_t* pz = inv_ptr + (nx + 1) + (ny + 1);
  return px[i + 1] - px[i] == 1 && inv_idx[px[i]] == i + (KXY - 1) / 2 && py[j + 1] - py[j] == 1 &&
         inv_idx[py[j]] == j + (KXY - 1) / 2 && pz[k + 1] - pz[k] == 1 && inv_idx[pz[k]] == k + (KZ - 1) / 2;
}

// dx_e of element (i, j, k) by its inverse lists: the 'full' correlation of dy at every padded cell
// that maps onto it, in list order (the taps in the order r, q, p)
template <int KXY, int KZ>
__device__ inline double filter_bwd_gather(int nx, int ny, int nz, const double* __restrict__ w,
                                           const int32_t* __restrict__ inv_ptr, const int32_t* __restrict__ inv_idx,
                                           const double* __restrict__ dy, int i, int j, int k) {
  const int32_t* px = inv_ptr;
  const int32_t* py = inv_ptr + (nx + 1);
  const int32_t* pz = inv_ptr + (nx + 1) + (ny + 1);
  double acc = 0.0;
  for (int ic = pz[k]; ic < pz[k + 1]; ++ic) {
    const int c = inv_idx[ic];
    for (int ib = py[j]; ib < py[j + 1]; ++ib) {
      const int b = inv_idx[ib];
      for (int ia = px[i]; ia < px[i + 1]; ++ia) {
        const int a = inv_idx[ia];
#pragma unroll
        for (int r = 0; r < KZ; ++r) {
          const int cc = c - (KZ - 1) + r;
#pragma unroll
          for (int q = 0; q < KXY; ++q) {
            const int bb = b - (KXY - 1) + q;
#pragma unroll
            for (int p = 0; p < KXY; ++p) {
              const int aa = a - (KXY - 1) + p;
              const bool in = cc >= 0 && cc < nz && bb >= 0 && bb < ny && aa >= 0 && aa < nx;
              const long long m = in ? ((long long)cc * ny + bb) * nx + aa : 0;
              if (in) acc = fma(w[(p * KXY + q) * KZ + r], dy[m], acc);
            }
          }
        }
      }
    }
  }
  return acc;
}

template <int KXY, int KZ>
__global__ void __launch_bounds__(256)
filter_bwd_fixed_kernel(int nx, int ny, int nz, const double* __restrict__ w, const int32_t* __restrict__ inv_ptr,
                        const int32_t* __restrict__ inv_idx, const double* __restrict__ dy, double* __restrict__ dx) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nel = (long long)nx * ny * nz;
  if (e >= nel) return;
  const int i = (int)(e % nx);
  const long long t = e / nx;
  const int j = (int)(t % ny);
  const int k = (int)(t / ny);
  dx[e] = filter_bwd_gather<KXY, KZ>(nx, ny, nz, w, inv_ptr, inv_idx, dy, i, j, k);
}

// Shell of the tiled adjoint: the elements within the half widths of a face, where every padding mode puts
// the elements with more than one padded cell (mirrored, replicated or wrapped cells map there), in three
// disjoint slab sets - x slabs (all j, k), y slabs (interior i), z slabs (interior i, j).  Those with a
// single cell are left to the tiled kernel.
template <int KXY, int KZ>
__global__ void __launch_bounds__(256)
filter_bwd_shell_kernel(int nx, int ny, int nz, const double* __restrict__ w, const int32_t* __restrict__ inv_ptr,
                        const int32_t* __restrict__ inv_idx, const double* __restrict__ dy, double* __restrict__ dx) {
  constexpr int HX = (KXY - 1) / 2, HZ = (KZ - 1) / 2;
  const int wx = min(nx, 2 * HX), wy = min(ny, 2 * HX), wz = min(nz, 2 * HZ);
  const long long cx = (long long)wx * ny * nz, cy = (long long)(nx - wx) * wy * nz;
  const long long cz = (long long)(nx - wx) * (ny - wy) * wz;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  // slab index a of width wa on an axis of extent n: the first half-width, then the last
  auto side = [](int a, int wa, int n, int h) { return (a < h || wa == n) ? a : n - wa + a; };
  int i, j, k;
  if (t < cx) {
    i = side((int)(t % wx), wx, nx, HX);
    t /= wx;
    j = (int)(t % ny);
    k = (int)(t / ny);
  } else if ((t -= cx) < cy) {
    i = wx / 2 + (int)(t % (nx - wx));
    t /= (nx - wx);
    j = side((int)(t % wy), wy, ny, HX);
    k = (int)(t / wy);
  } else if ((t -= cy) < cz) {
    i = wx / 2 + (int)(t % (nx - wx));
    t /= (nx - wx);
    j = wy / 2 + (int)(t % (ny - wy));
    k = side((int)(t / (ny - wy)), wz, nz, HZ);
  } else {
    return;
  }
  if (filter_single_cell<KXY, KZ>(inv_ptr, inv_idx, nx, ny, i, j, k)) return;
  dx[((long long)k * ny + j) * nx + i] = filter_bwd_gather<KXY, KZ>(nx, ny, nz, w, inv_ptr, inv_idx, dy, i, j, k);
}

// LDS-tiled forms of the fixed-extent kernels (default).  A block of 64 x 4 threads owns 64 x 16 output
// columns (a thread the column of 4 outputs above its row) and walks a chunk of nzc element planes in z;
// the input planes the tile needs pass through a ring of KZ LDS planes of (64 + KXY - 1) x (16 + KXY - 1)
// values, each fetched once from HBM while the previous plane is reduced.  A thread reads the
// (4 + KXY - 1) x KXY values of a plane that its 4 outputs share once and carries the 4 tap chains
// together (a quarter of the LDS reads of one output per thread), and every chain sums its taps in the
// order of the gather kernels above (r, then q, then p), so both forms give the same bits.
constexpr int kFiltTX = 64, kFiltTY = 4, kFiltOY = 4;
constexpr int kFiltRY = kFiltTY * kFiltOY;  // output rows of a tile
int g_filter_tiled = 1;                     // PMT_TUNE_FILTER: 1 tiled (default), 0 per-element gathers

// Forward: the padded value at (a, b, c) is cz[c] if mz[c] < 0, else cy[b] if my[b] < 0, else cx[a] if
// mx[a] < 0, else x[mz[c], my[b], mx[a]]; everything but mz[c] is the same for every plane, so each
// thread resolves its cells' in-plane sources once and a plane costs one load per cell.
template <int KXY, int KZ>
__global__ void __launch_bounds__(kFiltTX * kFiltTY)
filter_fwd_tiled_kernel(int nx, int ny, int nz, int tiles_x, int tiles_y, int nzc, int zero, const double* __restrict__ w,
                        const int32_t* __restrict__ map, const double* __restrict__ cval, const double* __restrict__ x,
                        double* __restrict__ y) {
  constexpr int LX = kFiltTX + KXY - 1, LY = kFiltRY + KXY - 1, LP = LX * LY, NT = kFiltTX * kFiltTY;
  constexpr int PER = (LP + NT - 1) / NT;  // plane values loaded per thread
  constexpr int WY = kFiltOY + KXY - 1;    // rows of a thread's window
  __shared__ double ring[KZ][LP];
  const int tid = threadIdx.x, tx = tid % kFiltTX, ty = tid / kFiltTX;
  const int tile = blockIdx.x % (tiles_x * tiles_y), chunk = blockIdx.x / (tiles_x * tiles_y);
  const int i0 = (tile % tiles_x) * kFiltTX, j0 = (tile / tiles_x) * kFiltRY;
  const int kbeg = chunk * nzc, kend = min(kbeg + nzc, nz);
  const int Lx = nx + KXY - 1, Ly = ny + KXY - 1, Lz = nz + KZ - 1;
  const int nxy = nx * ny;
  const int32_t* mx = map;
  const int32_t* my = map + Lx;
  const int32_t* mz = map + Lx + Ly;
  const double* cz = cval + Lx + Ly;
  // in-plane source of each cell this thread loads: element offset (>= 0), or -1 with its constant in
  // cst (constant row or column), or -2 outside the padded grid / the tile (value 0)
  int src[PER];
  double cst[PER];
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int q = tid + t * NT;
    const int a = i0 + q % LX, b = j0 + q / LX;
    src[t] = -2;
    cst[t] = 0.0;
    if (q < LP && a < Lx && b < Ly) {
      const int mb = my[b];
      if (mb < 0) {
        src[t] = -1;
        cst[t] = cval[Lx + b];
      } else {
        const int ma = mx[a];
        if (ma < 0) {
          src[t] = -1;
          cst[t] = cval[a];
        } else {
          src[t] = mb * nx + ma;
        }
      }
    }
  }
  auto fetch = [&](int c, double (&v)[PER]) {
    const int mc = c < Lz ? mz[c] : 0;
    const double czc = (c < Lz && mc < 0) ? cz[c] : 0.0;
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      double val = src[t] == -1 ? cst[t] : 0.0;
      if (src[t] >= 0 && mc >= 0) val = x[mc * nxy + src[t]];  // a constant plane (mc < 0) reads nothing
      v[t] = (mc < 0 && src[t] != -2) ? czc : val;
    }
  };
  auto store = [&](int c, const double (&v)[PER]) {
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int q = tid + t * NT;
      if (q < LP) ring[c % KZ][q] = v[t];
    }
  };
  double v[PER];
  for (int c = kbeg; c < kbeg + KZ - 1; ++c) {
    fetch(c, v);
    store(c, v);
  }
  fetch(kbeg + KZ - 1, v);
  const int i = i0 + tx, jb = j0 + ty * kFiltOY;
  for (int k = kbeg; k < kend; ++k) {
    store(k + KZ - 1, v);  // padded plane k + KZ - 1 completes the window of output plane k
    __syncthreads();
    if (k + 1 < kend) fetch(k + KZ, v);  // next plane's loads in flight during the reduction
    // the weights are re-read per plane through the scalar cache (zero is 0; hoisted out of the loop they
    // would occupy KXY^2 KZ register pairs per thread)
    const double* wk = w + zero * k;
    double acc[kFiltOY];
#pragma unroll
    for (int o = 0; o < kFiltOY; ++o) acc[o] = 0.0;
#pragma unroll
    for (int r = 0; r < KZ; ++r) {
      const double* pl = ring[(k + KZ - 1 - r) % KZ] + ty * kFiltOY * LX + tx;
      double win[WY][KXY];
#pragma unroll
      for (int yy = 0; yy < WY; ++yy)
#pragma unroll
        for (int p = 0; p < KXY; ++p) win[yy][p] = pl[yy * LX + p];
#pragma unroll
      for (int q = 0; q < KXY; ++q)
#pragma unroll
        for (int p = 0; p < KXY; ++p) {
          const double wt = wk[(p * KXY + q) * KZ + r];
#pragma unroll
          for (int o = 0; o < kFiltOY; ++o) acc[o] = fma(wt, win[o + KXY - 1 - q][KXY - 1 - p], acc[o]);
        }
      if constexpr (KXY > 3) __builtin_amdgcn_sched_barrier(0);  // one plane's window live at a time
    }
#pragma unroll
    for (int o = 0; o < kFiltOY; ++o)
      if (i < nx && jb + o < ny) y[(k * ny + jb + o) * nx + i] = acc[o];
    __syncthreads();  // the slot of plane k is overwritten next
  }
}

// Adjoint: dx_e = sum over the padded cells mapping to e of the 'full' correlation of dy there.  The
// tiled kernel reduces, from an LDS ring of dy planes (zero outside the grid), the elements whose only
// padded cell is their own position (all but a shell of (K-1)/2 along mirrored, replicated or wrapped
// faces; every element with constant padding); filter_bwd_shell_kernel then does the shell's elements
// by their inverse lists, so no wave of the tiled kernel diverges into the list walk.
template <int KXY, int KZ>
__global__ void __launch_bounds__(kFiltTX * kFiltTY)
filter_bwd_tiled_kernel(int nx, int ny, int nz, int tiles_x, int tiles_y, int nzc, int zero, const double* __restrict__ w,
                        const int32_t* __restrict__ inv_ptr, const int32_t* __restrict__ inv_idx,
                        const double* __restrict__ dy, double* __restrict__ dx) {
  constexpr int HX = (KXY - 1) / 2, HZ = (KZ - 1) / 2;
  constexpr int LX = kFiltTX + KXY - 1, LY = kFiltRY + KXY - 1, LP = LX * LY, NT = kFiltTX * kFiltTY;
  constexpr int PER = (LP + NT - 1) / NT;
  constexpr int WY = kFiltOY + KXY - 1;
  __shared__ double ring[KZ][LP];
  const int tid = threadIdx.x, tx = tid % kFiltTX, ty = tid / kFiltTX;
  const int tile = blockIdx.x % (tiles_x * tiles_y), chunk = blockIdx.x / (tiles_x * tiles_y);
  const int i0 = (tile % tiles_x) * kFiltTX, j0 = (tile / tiles_x) * kFiltRY;
  const int kbeg = chunk * nzc, kend = min(kbeg + nzc, nz);
  const int32_t* px = inv_ptr;
  const int32_t* py = inv_ptr + (nx + 1);
  const int32_t* pz = inv_ptr + (nx + 1) + (ny + 1);
  // element (in-plane) offset of each cell this thread loads; -1 outside the grid (value 0)
  int src[PER];
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int q = tid + t * NT;
    const int ii = i0 - HX + q % LX, jj = j0 - HX + q / LX;
    src[t] = (q < LP && ii >= 0 && ii < nx && jj >= 0 && jj < ny) ? jj * nx + ii : -1;
  }
  const int nxy = nx * ny;
  auto fetch = [&](int kk, double (&v)[PER]) {
    const bool zin = kk >= 0 && kk < nz;
#pragma unroll
    for (int t = 0; t < PER; ++t) v[t] = (zin && src[t] >= 0) ? dy[kk * nxy + src[t]] : 0.0;
  };
  auto store = [&](int kk, const double (&v)[PER]) {
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int q = tid + t * NT;
      if (q < LP) ring[(kk + KZ) % KZ][q] = v[t];
    }
  };
  const int i = i0 + tx, jb = j0 + ty * kFiltOY;
  // x and y parts of the single-cell test, once per thread
  const bool xin = i < nx && px[i + 1] - px[i] == 1 && inv_idx[px[i]] == i + HX;
  bool yin[kFiltOY];
#pragma unroll
  for (int o = 0; o < kFiltOY; ++o) {
    const int j = jb + o;
    yin[o] = xin && j < ny && py[j + 1] - py[j] == 1 && inv_idx[py[j]] == j + HX;
  }
  double v[PER];
  for (int kk = kbeg - HZ; kk < kbeg + HZ; ++kk) {
    fetch(kk, v);
    store(kk, v);
  }
  fetch(kbeg + HZ, v);
  for (int k = kbeg; k < kend; ++k) {
    store(k + HZ, v);  // dy plane k + HZ completes the window of element plane k
    __syncthreads();
    if (k + 1 < kend) fetch(k + HZ + 1, v);
    const bool zone = pz[k + 1] - pz[k] == 1 && inv_idx[pz[k]] == k + HZ;  // the same for the whole block
    const double* wk = w + zero * k;  // re-read per plane (see the forward kernel)
    if (zone) {
      // element (i, j, k) at cell (i + HX, j + HX, k + HZ): dy at (i - HX + p, j - HX + q, k - HZ + r)
      double acc[kFiltOY];
#pragma unroll
      for (int o = 0; o < kFiltOY; ++o) acc[o] = 0.0;
#pragma unroll
      for (int r = 0; r < KZ; ++r) {
        const double* pl = ring[(k - HZ + r + KZ) % KZ] + ty * kFiltOY * LX + tx;
        double win[WY][KXY];
#pragma unroll
        for (int yy = 0; yy < WY; ++yy)
#pragma unroll
          for (int p = 0; p < KXY; ++p) win[yy][p] = pl[yy * LX + p];
#pragma unroll
        for (int q = 0; q < KXY; ++q)
#pragma unroll
          for (int p = 0; p < KXY; ++p) {
            const double wt = wk[(p * KXY + q) * KZ + r];
#pragma unroll
            for (int o = 0; o < kFiltOY; ++o) acc[o] = fma(wt, win[o + q][p], acc[o]);
          }
        if constexpr (KXY > 3) __builtin_amdgcn_sched_barrier(0);  // one plane's window live at a time
      }
#pragma unroll
      for (int o = 0; o < kFiltOY; ++o)
        if (yin[o]) dx[(k * ny + jb + o) * nx + i] = acc[o];
    }
    __syncthreads();
  }
}

template <int KXY, int KZ>
void launch_filter_tiled(bool fwd, int nx, int ny, int nz, const double* w, const int32_t* m0, const int32_t* m1,
                         const double* cv, const double* in, double* out, hipStream_t s) {
  const int tiles_x = (nx + kFiltTX - 1) / kFiltTX, tiles_y = (ny + kFiltRY - 1) / kFiltRY;
  // planes per block: about three blocks per CU (a block re-reads KZ - 1 halo planes), 4 to 16 planes
  const long long want = 3LL * device_cu_count();
  const int nzc = (int)std::max(4LL, std::min(16LL, (long long)nz * tiles_x * tiles_y / std::max(1LL, want)));
  const unsigned blocks = (unsigned)(tiles_x * tiles_y * ((nz + nzc - 1) / nzc));
  if (fwd) {
    filter_fwd_tiled_kernel<KXY, KZ><<<blocks, kFiltTX * kFiltTY, 0, s>>>(nx, ny, nz, tiles_x, tiles_y, nzc, 0, w, m0,
                                                                        cv, in, out);
  } else {
    filter_bwd_tiled_kernel<KXY, KZ><<<blocks, kFiltTX * kFiltTY, 0, s>>>(nx, ny, nz, tiles_x, tiles_y, nzc, 0, w, m0,
                                                                        m1, in, out);
    constexpr int HX = (KXY - 1) / 2, HZ = (KZ - 1) / 2;
    const long long wx = std::min(nx, 2 * HX), wy = std::min(ny, 2 * HX), wz = std::min(nz, 2 * HZ);
    const long long shell = wx * ny * nz + (nx - wx) * wy * nz + (nx - wx) * (ny - wy) * wz;
    if (shell > 0)
      filter_bwd_shell_kernel<KXY, KZ><<<blocks_for(shell, 256), 256, 0, s>>>(nx, ny, nz, w, m0, m1, in, out);
  }
}

// the tiled kernels index with 32-bit element numbers
inline bool filter_tiled_ok(long long nel) { return g_filter_tiled && nel < (1LL << 31); }

// DensityFilter (pymoto/modules/filter.py:223-378): y_i = so_i * sum_j H_ij (si_j * x_j) over the
// clipped window |j - i|_inf <= int(radius) with H_ij = max(0, radius - |i - j|) in element units.
// Forward: si = 1, so = 1/s; adjoint: si = 1/s, so = 1 (H symmetric); s = H 1 with si = so = 1.
__global__ void density_filter_kernel(int nx, int ny, int nz, double radius, int dxy, int dz,
                                      const double* __restrict__ x, const double* __restrict__ si,
                                      const double* __restrict__ so, double* __restrict__ y) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)nx * ny * nz) return;
  const int i = (int)(e % nx), j = (int)((e / nx) % ny), k = (int)(e / ((long long)nx * ny));
  double acc = 0.0;
  for (int kk = max(k - dz, 0); kk <= min(k + dz, nz - 1); ++kk)
    for (int jj = max(j - dxy, 0); jj <= min(j + dxy, ny - 1); ++jj)
      for (int ii = max(i - dxy, 0); ii <= min(i + dxy, nx - 1); ++ii) {
        const double di = ii - i, dj = jj - j, dk = kk - k;
        const double w = radius - sqrt(di * di + dj * dj + dk * dk);
        if (w <= 0.0) continue;
        const long long m = ((long long)kk * ny + jj) * nx + ii;
        acc = fma(w, si ? si[m] * x[m] : x[m], acc);
      }
  y[e] = so ? acc * so[e] : acc;
}

}  // namespace pmt

using namespace pmt;

extern "C" {

int pmt_filter_fwd(int nelx, int nely, int nelz, int kx, int ky, int kz, const double* w, const int32_t* map,
                   const double* cval, const double* x, double* y, void* stream) {
  if (!grid_ok(nelx, nely, nelz) || kx < 1 || ky < 1 || kz < 1 || !(kx & 1) || !(ky & 1) || !(kz & 1) || !w ||
      !map || !cval || !x || !y)
    return arg_error("pmt_filter_fwd: bad argument");
  const int nz = nelz > 0 ? nelz : 1;
  if (nelz == 0 && kz != 1) return arg_error("pmt_filter_fwd: 2D grid needs kz == 1");
  const long long nel = (long long)nelx * nely * nz;
  const unsigned nb = blocks_for(nel, 256);
  hipStream_t s = (hipStream_t)stream;
  if (filter_tiled_ok(nel) && kx == ky && ((kx == 3 && (kz == 3 || kz == 1)) || (kx == 5 && (kz == 5 || kz == 1)))) {
    if (kx == 3)
      (kz == 3 ? launch_filter_tiled<3, 3> : launch_filter_tiled<3, 1>)(true, nelx, nely, nz, w, map, nullptr, cval, x,
                                                                       y, s);
    else
      (kz == 5 ? launch_filter_tiled<5, 5> : launch_filter_tiled<5, 1>)(true, nelx, nely, nz, w, map, nullptr, cval, x,
                                                                       y, s);
    return check_launch("filter_fwd_tiled_kernel");
  }
  if (kx == 3 && ky == 3 && kz == 3)
    filter_fwd_fixed_kernel<3, 3><<<nb, 256, 0, s>>>(nelx, nely, nz, w, map, cval, x, y);
  else if (kx == 3 && ky == 3 && kz == 1)
    filter_fwd_fixed_kernel<3, 1><<<nb, 256, 0, s>>>(nelx, nely, nz, w, map, cval, x, y);
  else if (kx == 5 && ky == 5 && kz == 5)
    filter_fwd_fixed_kernel<5, 5><<<nb, 256, 0, s>>>(nelx, nely, nz, w, map, cval, x, y);
  else if (kx == 5 && ky == 5 && kz == 1)
    filter_fwd_fixed_kernel<5, 1><<<nb, 256, 0, s>>>(nelx, nely, nz, w, map, cval, x, y);
  else
    filter_fwd_kernel<<<nb, 256, 0, s>>>(nelx, nely, nz, kx, ky, kz, w, map, cval, x, y);
  return check_launch("filter_fwd_kernel");
}

int pmt_density_filter(int nelx, int nely, int nelz, double radius, const double* x, const double* scale_in,
                       const double* scale_out, double* y, void* stream) {
  if (!grid_ok(nelx, nely, nelz) || !(radius > 0.0) || radius > 15.0 || !x || !y)
    return arg_error("pmt_density_filter: bad argument (0 < radius <= 15)");
  const int nz = nelz > 0 ? nelz : 1;
  const long long nel = (long long)nelx * nely * nz;
  density_filter_kernel<<<blocks_for(nel, 256), 256, 0, (hipStream_t)stream>>>(nelx, nely, nz, radius, (int)radius,
                                                                              nelz > 0 ? (int)radius : 0, x,
                                                                              scale_in, scale_out, y);
  return check_launch("density_filter_kernel");
}

int pmt_filter_bwd(int nelx, int nely, int nelz, int kx, int ky, int kz, const double* w, const int32_t* inv_ptr,
                   const int32_t* inv_idx, const double* dy, double* dx, void* stream) {
  if (!grid_ok(nelx, nely, nelz) || kx < 1 || ky < 1 || kz < 1 || !(kx & 1) || !(ky & 1) || !(kz & 1) || !w ||
      !inv_ptr || !inv_idx || !dy || !dx)
    return arg_error("pmt_filter_bwd: bad argument");
  const int nz = nelz > 0 ? nelz : 1;
  if (nelz == 0 && kz != 1) return arg_error("pmt_filter_bwd: 2D grid needs kz == 1");
  const long long nel = (long long)nelx * nely * nz;
  const unsigned nb = blocks_for(nel, 256);
  hipStream_t s = (hipStream_t)stream;
  if (filter_tiled_ok(nel) && kx == ky && ((kx == 3 && (kz == 3 || kz == 1)) || (kx == 5 && (kz == 5 || kz == 1)))) {
    if (kx == 3)
      (kz == 3 ? launch_filter_tiled<3, 3> : launch_filter_tiled<3, 1>)(false, nelx, nely, nz, w, inv_ptr, inv_idx,
                                                                       nullptr, dy, dx, s);
    else
      (kz == 5 ? launch_filter_tiled<5, 5> : launch_filter_tiled<5, 1>)(false, nelx, nely, nz, w, inv_ptr, inv_idx,
                                                                       nullptr, dy, dx, s);
    return check_launch("filter_bwd_tiled_kernel");
  }
  if (kx == 3 && ky == 3 && kz == 3)
    filter_bwd_fixed_kernel<3, 3><<<nb, 256, 0, s>>>(nelx, nely, nz, w, inv_ptr, inv_idx, dy, dx);
  else if (kx == 3 && ky == 3 && kz == 1)
    filter_bwd_fixed_kernel<3, 1><<<nb, 256, 0, s>>>(nelx, nely, nz, w, inv_ptr, inv_idx, dy, dx);
  else if (kx == 5 && ky == 5 && kz == 5)
    filter_bwd_fixed_kernel<5, 5><<<nb, 256, 0, s>>>(nelx, nely, nz, w, inv_ptr, inv_idx, dy, dx);
  else if (kx == 5 && ky == 5 && kz == 1)
    filter_bwd_fixed_kernel<5, 1><<<nb, 256, 0, s>>>(nelx, nely, nz, w, inv_ptr, inv_idx, dy, dx);
  else
    filter_bwd_kernel<<<nb, 256, 0, s>>>(nelx, nely, nz, kx, ky, kz, w, inv_ptr, inv_idx, dy, dx);
  return check_launch("filter_bwd_kernel");
}

}  // extern "C"
