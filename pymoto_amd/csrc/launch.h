// Host-side launchers shared between translation units (each kernel is launched from the file
// that defines it).
#pragma once
#include "common.h"

namespace pmt {

// a small dense matrix passed to kernels by value (up to 4 x 4, row-major)
struct SmallMat {
  double e[16];
};

// vector.hip: z -= W Einv (AW)^T z for m <= 4 columns W and AW = A W (column j at W + j*n); buf holds
// 4 + 4*kReduceBlocks doubles (the coefficients and the dot partials)
int launch_coarse_correct(long long n, int m, const double* W, const double* AW, const SmallMat& Einv, double* buf,
                          double* z, hipStream_t s);

// MODE_CHEB: y = u + omega D^-1 (b - A u) + beta (u - up), the second-order (Chebyshev) sweep;
// up may be NULL (= 0) and may alias y (each entry reads its own up before writing it).
// MODE_APPLY_DOTS (fine reflection kernel, FP64): y = A u plus per-block partial sums of u.y and u.r.
enum { MODE_APPLY = 0, MODE_RESIDUAL = 1, MODE_JACOBI = 2, MODE_CHEB = 3, MODE_APPLY_DOTS = 4 };

// operator.hip (level kernels exist for T = double and T = float)
template <typename T>
int launch_fine(int mode, const GridDims& g, const double* K0, int k0_pattern, const T* xe, const uint8_t* bcm,
                double bcdiag, const T* u, const T* b, const T* dinv, double omega, T* y, hipStream_t s,
                const T* up = nullptr, double beta = 0.0);
int k0_matches_pattern(int dim, const double* K0_host);
// refl.hip: hex8 operator in the mid-plane reflection basis (K0ext = pmt_k0_extend layout)
constexpr int kK0ReflOffset64 = 576;  // 72 FP64 block entries after K0
constexpr int kK0ReflOffset32 = 648;  // 72 FP32 block entries (36 doubles of storage)
bool k0_reflection_table(const double* K0_host, double* tab);
// first coarse level of a hex8 hierarchy in the same basis (K0 reflection symmetric), without a stencil:
// G0ext = coarse1_table (864 doubles), Xh[layer][8][element] the transformed densities of the unflagged
// children (launch_coarse1_xhat); the elements with Dirichlet children enter through the correction
// stencils TS[243][cap] of the nt touched nodes tnode (launch_coarse1_touched, then _tstencil)
void coarse1_table(const double* G0_host, double* ext_host);
int coarse1_table_pattern(const double* ext_host, int k0_pattern);  // the launch_coarse1 pattern (0, 1, 2)
int launch_coarse1_xhat(const GridDims& gf, const GridDims& gc, const double* xe, const uint8_t* cflag, double* X64,
                        float* X32, hipStream_t s);
int launch_coarse1_touched(const GridDims& g, const int* list, const int* map, int* tnode, int* tcount,
                           hipStream_t s);
int launch_coarse1_tstencil(const GridDims& g, const int* tnode, int nt, long long cap, const int* map,
                            const double* dK, double* TS, hipStream_t s);
template <typename T>
int launch_coarse1(int mode, const GridDims& g, int pattern, const double* G0ext, const T* Xh, const int* tnode, int nt,
                   long long cap, const double* TS, const T* u, const T* b, const T* dinv, double omega, T* y,
                   hipStream_t s, const T* up = nullptr, double beta = 0.0);
template <typename T>
int launch_fine_refl(int mode, const GridDims& g, int planes /* 0 = from the occupancy */, int rows,
                     int pattern /* K0 has the box class pattern: skip its structural zeros */, const double* K0ext,
                     const T* xe,
                     const uint8_t* bcm, double bcdiag, const T* u, const T* b, const T* dinv, double omega, T* y,
                     hipStream_t s, const T* up, double beta);
// q = A p on the fine grid with out[0] = p.q and out[1] = p.r (deterministic two-stage sums; ws holds
// 2 * ws_parts doubles).  Returns PMT_ERR_STATE (nothing launched) when the reflection kernel is not the
// fine operator or the launch needs more partial sums than ws holds; the caller then runs K.u and
// the dot products separately.
// element product of the fine reflection-basis kernel: 2 = the isotropic cube's 8 numbers, 1 = the box
// class pattern's zero block entries skipped (with vanishing rigid-translation entries), 0 = dense blocks
inline int refl_pattern(int k0_pattern) {
  if (k0_pattern & PMT_K0_CUBE) return 2;
  return ((k0_pattern & PMT_K0_CLASS_PATTERN) && (k0_pattern & PMT_K0_RIGID)) ? 1 : 0;
}
int launch_fine_apply_dots(const GridDims& g, const double* K0ext, int k0_pattern, const double* xe,
                           const uint8_t* bcm, double bcdiag, const double* p, const double* r, double* q,
                           double* ws, long long ws_parts, double* out, hipStream_t s);
long long fine_apply_dots_parts(const GridDims& g);  // a sufficient ws_parts for the grid
// refl.hip: dx = alpha u_e^T K0 v_e (accumulate: dx += ...) in the reflection basis over z-streamed node
// columns; tab = the FP64 block entries behind K0; hex8 only
int launch_sens_stream(const GridDims& g, const double* tab, const uint8_t* bcm, const double* u, const double* v,
                       double alpha, double* dx, int accumulate, hipStream_t s);
int set_tuning(int param, int value);
int get_tuning(int param);
int tuned_stream_planes();  // 0 = automatic
int tuning_epoch();         // changes whenever a launch shape changes
int device_cu_count();      // of the current device
int launch_fine_diag(const GridDims& g, const double* K0, const double* xe, const uint8_t* bcm, double bcdiag,
                     double* diag, int invert, hipStream_t s,
                     float* diag32 = nullptr);
template <typename T>
int launch_stencil(int mode, const GridDims& g, const T* S, const T* u, const T* b, const T* dinv, double omega, T* y,
                   hipStream_t s);  // MODE_APPLY / RESIDUAL / JACOBI
template <typename T>
int launch_stencil_dinv(const GridDims& g, const double* S, T* dinv, hipStream_t s);
// symmetric half stencil (centre + forward blocks) built from the full one, and its operator
int half_stencil_coefficients(int dim);
template <typename T>
int launch_stencil_compress(const GridDims& g, const double* S, T* H, hipStream_t s);
template <typename T>
int launch_stencil_half(int mode, const GridDims& g, const T* H, const T* u, const T* b, const T* dinv, double omega,
                        T* y, hipStream_t s, const T* up = nullptr, double beta = 0.0);
template <typename T>
int launch_jacobi0(long long n, const T* dinv, const T* b, double omega, T* u, hipStream_t s);
template <typename T>
// u0 (optional): also u0 = w0 D^-1 rc (dinv of the coarse level), its first smoothing sweep from zero
int launch_restrict(const GridDims& gf, const GridDims& gc, const T* rf, T* rc, hipStream_t s, const T* dinv = nullptr,
                    double w0 = 0.0, T* u0 = nullptr);
template <typename T>
int launch_prolong(const GridDims& gf, const GridDims& gc, const T* uc, T* uf, hipStream_t s);

// mfma.hip (hex8 only; BF16 operands, FP32 accumulation and vectors)
int launch_fine_mfma(int mode, const GridDims& g, const double* K0, const float* xe, const uint8_t* bcm,
                     double bcdiag, const float* u, const float* b, const float* dinv, double omega, float* y,
                     hipStream_t s, const float* up = nullptr, double beta = 0.0);

// vector.hip
int launch_dot(long long n, const double* x, const double* y, double* out, double* work, hipStream_t s);
int launch_dot(long long n, const double* x, const float* y, double* out, double* work, hipStream_t s);
int launch_dot2(long long n, const double* a0, const double* b0, const double* a1, const double* b1, double* out,
                double* work, hipStream_t s);
int launch_axpby(long long n, double a, const double* x, double b, double* y, hipStream_t s);
// x = nullptr: r only (the x update deferred to launch_cg_p_update).  work: cg_xr_parts(n) partial sums
long long cg_xr_parts(long long n);
// x.y in one flat pass, work: cg_xr_parts(n) partials (the CG loop's dots; bits differ from launch_dot's)
template <typename Y>
int launch_dot_flat(long long n, const double* x, const Y* y, double* out, double* work, hipStream_t s);
int launch_cg_xr_update(long long n, const double* sc, const double* p, const double* q, double* x, double* r,
                        double* work, double* out_rr, hipStream_t s, float* r32 = nullptr,
                        const float* dinv32 = nullptr, float omega32 = 0.0f, float* u1 = nullptr,
                        double* host = nullptr, double seq = 0.0);  // host: also publish r.r (launch_publish)
int launch_cg_x_update(long long n, const double* sc, const double* p, double* x, hipStream_t s);
// x (optional): also x += (sc[1]/sc[0]) p_old first, the update a cg_xr_update with x = nullptr deferred
int launch_cg_p_update(long long n, const double* sc, const double* z, double* p, hipStream_t s, double* x = nullptr);
int launch_cg_p_update(long long n, const double* sc, const float* z, double* p, hipStream_t s, double* x = nullptr);
int launch_scale_rsqrt(long long n, const double* sc, int k, const double* z, double* p, hipStream_t s);
int launch_fill(long long n, double v, double* y, hipStream_t s);
// host[0] = *v, then host[1] = seq (pinned host memory; the host polls host[1])
int launch_publish(const double* v, double seq, double* host, hipStream_t s);
int launch_to_double(long long n, const float* x, double* y, hipStream_t s);

// coarse.hip
// G: the 2^D fixed matrices P_c^T K0 P_c (galerkin_child_matrices), for the level-1 stencil of hex8 grids
void galerkin_child_matrices(int dim, const double* K0_host, double* G_host);
// element matrices of the first coarse level by triple products (Q4 grids)
int launch_galerkin_fine(const GridDims& gf, const double* K0, const double* xe, const uint8_t* bcm, double bcdiag,
                         double* Kc, hipStream_t s);
void zero_int(int* p, hipStream_t s);
// hex8 first coarse level: children flags, list (nel + 1 ints, count last) and map of flagged elements
int launch_level1_flags(const GridDims& gf, const uint8_t* bcm, uint8_t* cflag, int* list, int* map, hipStream_t s);
// corrections dK (576 per listed element) of the flagged level-1 elements
int launch_level1_dk(const GridDims& gf, const double* K0, const double* xe, const uint8_t* bcm, double bcdiag,
                     const uint8_t* cflag, const int* list, int nlist, double* dK, hipStream_t s);
int launch_galerkin_coarse(const GridDims& gf, const double* Kf, double* Kc, hipStream_t s);
int launch_stencil_from_elements(const GridDims& g, const double* Ke, double* S, hipStream_t s);
// level-1 stencil of a hex8 grid from the fine densities, G and the corrections of flagged elements
// (out.S set: full stencil, for the coarsest level; else out.H set: half stencil and D^-1, see StencilOut;
// else D^-1 only)
int launch_stencil_from_fine(const GridDims& gf, const double* G, const double* xe, const uint8_t* cflag,
                             const int* map, const double* dK, const StencilOut& out, hipStream_t s);

// moments.hip: Galerkin operators of hex8 levels >= 2 from density moments (27 per element)
void moment_tables(int level, const double* K0_host, double* C_host);  // 64 * 27 * 9 doubles
int launch_moments1(const GridDims& gf, const double* xe, const uint8_t* cflag, double* M, hipStream_t s);
int launch_moments_up(const GridDims& gf, const double* Mf, double* M, hipStream_t s);
int launch_flags_up(const GridDims& gf, const uint8_t* cflag_f, uint8_t* cflag, int* list, int* map, hipStream_t s);
int launch_dk_up(const GridDims& gf, const uint8_t* cflag, const int* list, int nlist, const int* map_f,
                 const double* dKf, double* dK, hipStream_t s);
// element matrices (Ke: 576 x nel, structure of arrays) from the moments plus the corrections, then S
int launch_stencil_from_moments(const GridDims& g, const double* C, const double* M, const int* map,
                                const double* dK, const StencilOut& out, hipStream_t s);
int launch_dense_from_stencil(const GridDims& g, const double* S, double* A, hipStream_t s);
int gj_invert(long long n, double* A, double* wsP, double* wsC, hipStream_t s);
template <typename T>
int launch_gemv(long long n, const T* A, const T* x, T* y, hipStream_t s);
int launch_gemv_f32io(long long n, const double* A, const float* x, float* y, hipStream_t s);
int launch_to_float_first(long long n, const double* x, float* y, const float* dinv, float w, float* u1,
                          hipStream_t s);
int launch_to_float(long long n, const double* x, float* y, hipStream_t s);

}  // namespace pmt
