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
#include "common.h"

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
  dx[e] = acc;
}

// LDS-tiled forms of the fixed-extent kernels (default).  A block owns TX x TY output columns and walks
// a chunk of NZC element planes in z; the padded input planes it needs pass through a ring of KZ LDS
// planes of (TX + KXY - 1) x (TY + KXY - 1) values, each loaded once from HBM (through the per-axis maps,
// 32-bit index math) while the previous plane is reduced, so an element is read about once per chunk
// instead of KXY^2 KZ times through the caches.  Taps are summed in the order of the gather kernels
// above (r, then q, then p), so both forms give the same bits.
constexpr int kFiltTX = 64, kFiltTY = 4, kFiltNZC = 16;
int g_filter_tiled = 1;  // PMT_TUNE_FILTER: 1 tiled (default), 0 per-element gathers

template <int KXY, int KZ>
__global__ void __launch_bounds__(kFiltTX * kFiltTY)
filter_fwd_tiled_kernel(int nx, int ny, int nz, int tiles_x, int tiles_y, const double* __restrict__ w,
                        const int32_t* __restrict__ map, const double* __restrict__ cval, const double* __restrict__ x,
                        double* __restrict__ y) {
  constexpr int LX = kFiltTX + KXY - 1, LY = kFiltTY + KXY - 1, LP = LX * LY, NT = kFiltTX * kFiltTY;
  constexpr int PER = (LP + NT - 1) / NT;  // plane values loaded per thread
  __shared__ double ring[KZ][LP];
  const int tid = threadIdx.x, tx = tid % kFiltTX, ty = tid / kFiltTX;
  const int tile = blockIdx.x % (tiles_x * tiles_y), chunk = blockIdx.x / (tiles_x * tiles_y);
  const int i0 = (tile % tiles_x) * kFiltTX, j0 = (tile / tiles_x) * kFiltTY;
  const int kbeg = chunk * kFiltNZC, kend = min(kbeg + kFiltNZC, nz);
  const int Lx = nx + KXY - 1, Ly = ny + KXY - 1, Lz = nz + KZ - 1;
  const int32_t* mx = map;
  const int32_t* my = map + Lx;
  const int32_t* mz = map + Lx + Ly;
  const double* cx = cval;
  const double* cy = cval + Lx;
  const double* cz = cval + Lx + Ly;
  // padded plane c (padded z index) of the tile: value at local (la, lb) = padded (i0 + la, j0 + lb)
  auto fetch = [&](int c, double (&v)[PER]) {
    const bool zin = c < Lz;
    const int mc = zin ? mz[c] : 0;
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int q = tid + t * NT;
      const int la = q % LX, lb = q / LX;
      const int a = i0 + la, b = j0 + lb;
      double val = 0.0;
      if (q < LP && zin && a < Lx && b < Ly) {
        if (mc < 0) {
          val = cz[c];
        } else {
          const int mb = my[b];
          if (mb < 0) {
            val = cy[b];
          } else {
            const int ma = mx[a];
            val = ma < 0 ? cx[a] : x[(mc * ny + mb) * nx + ma];
          }
        }
      }
      v[t] = val;
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
  const int i = i0 + tx, j = j0 + ty;
  const bool out = i < nx && j < ny;
  for (int k = kbeg; k < kend; ++k) {
    store(k + KZ - 1, v);  // padded plane k + KZ - 1 completes the window of output plane k
    __syncthreads();
    if (k + 1 < kend) fetch(k + KZ, v);  // next plane's loads in flight during the reduction
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < KZ; ++r) {
      const double* pl = ring[(k + KZ - 1 - r) % KZ];
#pragma unroll
      for (int q = 0; q < KXY; ++q)
#pragma unroll
        for (int p = 0; p < KXY; ++p)
          acc = fma(w[(p * KXY + q) * KZ + r], pl[(ty + KXY - 1 - q) * LX + tx + KXY - 1 - p], acc);
    }
    if (out) y[((long long)k * ny + j) * nx + i] = acc;
    __syncthreads();  // the slot of plane k is overwritten next
  }
}

// Adjoint: dx_e = sum over the padded cells mapping to e of the 'full' correlation of dy there.  An
// element whose three per-axis inverse lists hold only its own padded position (every element at least
// (K-1)/2 from a mirrored, edge-replicated or wrapped face) is reduced from an LDS ring of dy planes
// (zero outside the grid); the others take the per-element path of filter_bwd_fixed_kernel.
template <int KXY, int KZ>
__global__ void __launch_bounds__(kFiltTX * kFiltTY)
filter_bwd_tiled_kernel(int nx, int ny, int nz, int tiles_x, int tiles_y, const double* __restrict__ w,
                        const int32_t* __restrict__ inv_ptr, const int32_t* __restrict__ inv_idx,
                        const double* __restrict__ dy, double* __restrict__ dx) {
  constexpr int HX = (KXY - 1) / 2, HZ = (KZ - 1) / 2;
  constexpr int LX = kFiltTX + KXY - 1, LY = kFiltTY + KXY - 1, LP = LX * LY, NT = kFiltTX * kFiltTY;
  constexpr int PER = (LP + NT - 1) / NT;
  __shared__ double ring[KZ][LP];
  const int tid = threadIdx.x, tx = tid % kFiltTX, ty = tid / kFiltTX;
  const int tile = blockIdx.x % (tiles_x * tiles_y), chunk = blockIdx.x / (tiles_x * tiles_y);
  const int i0 = (tile % tiles_x) * kFiltTX, j0 = (tile / tiles_x) * kFiltTY;
  const int kbeg = chunk * kFiltNZC, kend = min(kbeg + kFiltNZC, nz);
  const int32_t* px = inv_ptr;
  const int32_t* py = inv_ptr + (nx + 1);
  const int32_t* pz = inv_ptr + (nx + 1) + (ny + 1);
  // dy plane kk (element z index, zero outside) at local (la, lb) = element (i0 - HX + la, j0 - HX + lb)
  auto fetch = [&](int kk, double (&v)[PER]) {
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int q = tid + t * NT;
      const int ii = i0 - HX + q % LX, jj = j0 - HX + q / LX;
      const bool in = q < LP && kk >= 0 && kk < nz && ii >= 0 && ii < nx && jj >= 0 && jj < ny;
      v[t] = in ? dy[(kk * ny + jj) * nx + ii] : 0.0;
    }
  };
  auto store = [&](int kk, const double (&v)[PER]) {
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int q = tid + t * NT;
      if (q < LP) ring[(kk + KZ) % KZ][q] = v[t];
    }
  };
  const int i = i0 + tx, j = j0 + ty;
  const bool out = i < nx && j < ny;
  // the only padded cell of an element away from the faces is its own position shifted by the half width
  const bool xy1 = out && px[i + 1] - px[i] == 1 && inv_idx[px[i]] == i + HX && py[j + 1] - py[j] == 1 &&
                   inv_idx[py[j]] == j + HX;
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
    const bool one = xy1 && pz[k + 1] - pz[k] == 1 && inv_idx[pz[k]] == k + HZ;
    double acc = 0.0;
    if (one) {
      // cell (a, b, c) = (i + HX, j + HX, k + HZ): dy at (a - (KXY-1) + p, ...) = (i - HX + p, ...)
#pragma unroll
      for (int r = 0; r < KZ; ++r) {
        const double* pl = ring[(k - HZ + r + KZ) % KZ];
#pragma unroll
        for (int q = 0; q < KXY; ++q)
#pragma unroll
          for (int p = 0; p < KXY; ++p) acc = fma(w[(p * KXY + q) * KZ + r], pl[(ty + q) * LX + tx + p], acc);
      }
    } else if (out) {
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
                  if (in) acc = fma(w[(p * KXY + q) * KZ + r], dy[(cc * ny + bb) * nx + aa], acc);
                }
              }
            }
          }
        }
      }
    }
    if (out) dx[((long long)k * ny + j) * nx + i] = acc;
    __syncthreads();
  }
}

template <int KXY, int KZ>
void launch_filter_tiled(bool fwd, int nx, int ny, int nz, const double* w, const int32_t* m0, const int32_t* m1,
                         const double* cv, const double* in, double* out, hipStream_t s) {
  const int tiles_x = (nx + kFiltTX - 1) / kFiltTX, tiles_y = (ny + kFiltTY - 1) / kFiltTY;
  const unsigned blocks = (unsigned)(tiles_x * tiles_y * ((nz + kFiltNZC - 1) / kFiltNZC));
  if (fwd)
    filter_fwd_tiled_kernel<KXY, KZ><<<blocks, kFiltTX * kFiltTY, 0, s>>>(nx, ny, nz, tiles_x, tiles_y, w, m0, cv, in,
                                                                        out);
  else
    filter_bwd_tiled_kernel<KXY, KZ><<<blocks, kFiltTX * kFiltTY, 0, s>>>(nx, ny, nz, tiles_x, tiles_y, w, m0, m1, in,
                                                                        out);
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
