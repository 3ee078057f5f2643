/*
 * C-ABI of the MI355X (gfx950) FE design-iteration hot path.
 *
 * Every entry point takes plain device pointers (allocated by the caller, e.g. torch tensors),
 * sizes and a hipStream_t passed as void*; it returns PMT_OK (0) or a negative status and
 * leaves a message for pmt_last_error().  All floating point is IEEE binary64 unless the name
 * says otherwise.  Grid arguments are element counts (nelz = 0 for a 2D Q4 grid); numbering is
 * pyMOTO's: element (k*nely+j)*nelx+i, node (k*(nely+1)+j)*(nelx+1)+i, dof node*D+d
 * (pymoto/common/domain.py:200-251).  Boundary conditions are a per-node byte mask whose bit d
 * marks dof d as Dirichlet (pymoto/modules/assembly.py:86-98).
 *
 * No entry point allocates device memory except pmt_mg_create/pmt_mg_update (setup) and the first
 * pmt_block_pcg_solve call on a handle with a given number of right-hand sides (its workspace).
 * Calls on one pmt_mg handle are not thread-safe.
 */
#ifndef PYMOTO_AMD_H
#define PYMOTO_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PMT_OK 0
#define PMT_ERR_ARG (-1)
#define PMT_ERR_HIP (-2)
#define PMT_ERR_STATE (-3)
/* a non-finite right-hand side, start vector or residual (the solve stops at once) */
#define PMT_ERR_NUMERIC (-4)

/* Message of the last failing call on this thread. */
const char* pmt_last_error(void);
/* ABI revision of this header. */
int pmt_abi_version(void);

/* ---------------------------------------------------------------------------------------------
 * FilterConv forward: y = W * xpad (scipy.signal.convolve 'valid' of the padded field).
 * Replaces FilterConv.__call__ (pymoto/modules/filter.py:207-212) and get_padded_vector (183-187).
 * w: kx*ky*kz weights in [x][y][z] order.  map/cval: per-axis padded-index maps concatenated
 * x|y|z, lengths (nelx+kx-1), (nely+ky-1), (max(nelz,1)+kz-1); map<0 marks a constant-padded
 * cell whose value is cval (z overrides y overrides x, filter.py:98-160).
 */
int pmt_filter_fwd(int nelx, int nely, int nelz, int kx, int ky, int kz, const double* w,
                   const int32_t* map, const double* cval, const double* x, double* y, void* stream);
/* FilterConv adjoint: correlate 'full', drop constant cells, fold back onto elements.
 * Replaces FilterConv._sensitivity (filter.py:214-220).  inv_ptr/inv_idx: per-axis lists of the
 * padded coordinates mapping to each element coordinate, concatenated x|y|z (inv_ptr holds
 * nelx+1, nely+1, max(nelz,1)+1 entries, absolute offsets into inv_idx).  Every element coordinate
 * farther than (k-1)/2 from both ends of its axis must map from its own padded coordinate only (true
 * for the reference's padding modes): the default kernels reduce those from LDS tiles and walk the lists
 * only within that distance of a face. */
int pmt_filter_bwd(int nelx, int nely, int nelz, int kx, int ky, int kz, const double* w,
                   const int32_t* inv_ptr, const int32_t* inv_idx, const double* dy, double* dx,
                   void* stream);

/* DensityFilter (pymoto/modules/filter.py:223-378): y_i = so_i * sum_j H_ij * si_j * x_j with
 * H_ij = max(0, radius - |i - j|) over the window |j - i|_inf <= int(radius) clipped to the grid
 * (element units).  si / so may be NULL (= 1): forward si = NULL, so = 1/s; adjoint si = 1/s,
 * so = NULL; the row sums s = H 1 with both NULL and x = 1. */
int pmt_density_filter(int nelx, int nely, int nelz, double radius, const double* x, const double* scale_in,
                       const double* scale_out, double* y, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Power-law material interpolation y = a + b*x^p (SIMP, MathExpression "xmin + (1-xmin)*inp0^3",
 * pymoto/modules/generic.py:94-99) and its adjoint dx (+)= dy*c*x^q (generic.py:101-140).
 */
int pmt_powlaw_fwd(long long n, double a, double b, double p, const double* x, double* y, void* stream);
int pmt_powlaw_bwd(long long n, double c, double q, const double* x, const double* dy, double* dx,
                   int accumulate, void* stream);

/* Heaviside projection y = (tanh(beta*eta) + tanh(beta*(x-eta))) / (tanh(beta*eta) + tanh(beta*(1-eta)))
 * (the MathExpression of examples/topology_optimization/ex_compliance_robust.py:139) and its adjoint
 * dx (+)= dy*dy/dx; dbeta (may be NULL) receives dy_i * dy_i/dbeta per entry. */
int pmt_heaviside_fwd(long long n, double beta, double eta, const double* x, double* y, void* stream);
int pmt_heaviside_bwd(long long n, double beta, double eta, const double* x, const double* dy, double* dx,
                      int accumulate, double* dbeta, void* stream);

/* One bisection step of the OC update (pymoto/common/optimizers.py:404-433): xnew = clip(x *
 * sqrt(-min(dg,0)/lmid), max(xmin, x-move), min(xmax, x+move)) and *sum_out = sum(xnew) (device
 * scalar, deterministic; work: pmt_reduce_work_size() doubles). */
int pmt_oc_update(long long n, const double* x, const double* dg, double xmin, double xmax, double move,
                  double lmid, double* xnew, double* sum_out, double* work, void* stream);

/* Vector kernels.  y = a*x + b*y (x may be NULL: y = b*y). */
int pmt_axpby(long long n, double a, const double* x, double b, double* y, void* stream);
/* *out = x.y, deterministic two-stage reduction; work holds pmt_reduce_work_size() doubles.
 * Replaces numpy einsum('i,i->') of EinSum (generic.py:193-194) and the CG dots (iterative.py:377-394). */
int pmt_dot(long long n, const double* x, const double* y, double* out, double* work, void* stream);
int pmt_reduce_work_size(void);
/* pmt_dot for 1 <= k <= PMT_DOT_NOALLOC_MAX_COLS column pairs without a workspace: out[j] = x_j.y_j with
 * column j at x + j*n, y + j*n; each out[j] is bit-identical to pmt_dot(n, x_j, y_j).  Allocates nothing
 * (no hipMalloc, no caller workspace): the partial sums live in a static 64 KiB device array, one per
 * GPU, whose address is looked up on the first call on each device.  Calls that overlap in time on one
 * GPU (other streams or host threads) share that array and race; serialise them.  k above the maximum
 * returns PMT_ERR_ARG.  No counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). */
#define PMT_DOT_NOALLOC_MAX_COLS 8
int pmt_dot_noalloc(int k, long long n, const double* x, const double* y, double* out, void* stream);
/* z = w * r / diag: DampedJacobi.solve (pymoto/solvers/iterative.py:41-47) as a preconditioner of CG. */
int pmt_jacobi_apply(long long n, double w, const double* diag, const double* r, double* z, void* stream);
/* Dirichlet split used by the linear-dependency-aware solve (pymoto/solvers/solvers.py:174-176):
 * for bc dofs sol = rhs/bcdiag and rhs_loc = 0, elsewhere rhs_loc = rhs and sol = 0. */
int pmt_bc_split(int nelx, int nely, int nelz, const uint8_t* bcmask, double bcdiag, const double* rhs,
                 double* rhs_loc, double* sol, void* stream);
/* y += a*x on free (non-bc) dofs only (solvers.py:233-236). */
int pmt_axpy_free(int nelx, int nely, int nelz, const uint8_t* bcmask, double a, const double* x,
                  double* y, void* stream);
/* The two above for ndof unknowns per node (1 <= ndof <= 8; the forms above use ndof = the grid
 * dimension): the scalar systems of AssemblePoisson use ndof = 1. */
int pmt_bc_split_ndof(int nelx, int nely, int nelz, int ndof, const uint8_t* bcmask, double bcdiag,
                      const double* rhs, double* rhs_loc, double* sol, void* stream);
int pmt_axpy_free_ndof(int nelx, int nely, int nelz, int ndof, const uint8_t* bcmask, double a, const double* x,
                       double* y, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Matrix-free stiffness operator K(x) = sum_e x_e K0 with Dirichlet rows/cols zeroed and bcdiag on
 * their diagonal.  Replaces the assembled CSR of AssembleGeneral.__call__ (assembly.py:245-296)
 * and its SpMV `A @ p` (pymoto/solvers/iterative.py:374-395).  K0: (2^D*D)^2 element matrix.
 */
int pmt_stiffness_apply(int nelx, int nely, int nelz, const double* K0, int k0_pattern, const double* xe,
                        const uint8_t* bcmask, double bcdiag, const double* u, double* y, void* stream);
/* r = b - K u (LinearSolver.residual / CG restart, solvers.py:52-85, iterative.py:384) */
int pmt_stiffness_residual(int nelx, int nely, int nelz, const double* K0, int k0_pattern, const double* xe,
                           const uint8_t* bcmask, double bcdiag, const double* b, const double* u,
                           double* r, void* stream);
/* Reduced-precision forms of pmt_stiffness_apply for the mixed-precision V-cycle (xe, u, y in
 * FP32): _f32 computes in FP32; _bf16 (hex8 only) contracts each element on the matrix cores with
 * BF16 operands (u, K0) and FP32 accumulation, then scales by x_e in FP32.
 * (No reference counterpart: the reduced-precision element products.) */
int pmt_stiffness_apply_f32(int nelx, int nely, int nelz, const double* K0, int k0_pattern, const float* xe,
                            const uint8_t* bcmask, double bcdiag, const float* u, float* y, void* stream);
int pmt_stiffness_apply_bf16(int nelx, int nely, int nelz, const double* K0, const float* xe,
                             const uint8_t* bcmask, double bcdiag, const float* u, float* y, void* stream);
/* 1 if a host element matrix follows the isotropic box sign/class pattern (36 classes in 3D, 10 in
 * 2D); then k0_pattern = 1 selects the kernel that keeps K0 in scalar registers.  Host only.
 * (No reference counterpart: a property of K0 from AssembleStiffness.__init__, assembly.py:446-461.) */
int pmt_k0_pattern(int dim, const double* K0_host);
/* k0_pattern is a bit set.  PMT_K0_CLASS_PATTERN: K0 follows the sign/class pattern above.
 * PMT_K0_REFLECTION (hex8 only): the device K0 buffer is the PMT_K0_EXT_LEN-double extended form
 * written by pmt_k0_extend - K0 itself, then its 8 diagonal 3x3 blocks in the basis of the three
 * mid-plane reflections (FP64, then FP32) - and the fine operator multiplies each element in that
 * basis (2*72 add/sub + 72 multiply-adds instead of 576 multiply-adds).  Kernels that only need K0
 * read the first (2^D*D)^2 entries of the same buffer.
 * PMT_K0_RIGID (with PMT_K0_REFLECTION): the blocks' entries of the three rigid translations vanish (to
 * 1e-13 of the largest block entry; true for every stiffness matrix); with PMT_K0_CLASS_PATTERN the
 * reflection-basis kernels then skip all the pattern's zero entries.
 * PMT_K0_CUBE (with PMT_K0_RIGID): the blocks have the isotropic cube's form - 8 numbers in a fixed
 * pattern, equal within classes to 1e-12 of the largest entry; pmt_k0_extend then writes the class
 * means (and exact zeros) and the fine kernel multiplies with the 8 numbers (33 operations per element).
 * pmt_k0_extend fills ext_host[PMT_K0_EXT_LEN] from a host K0 and returns the k0_pattern bits it
 * supports (PMT_K0_REFLECTION when K0 commutes with the reflections to 1e-13 of its largest entry).
 * Host only.  (No reference counterpart: a property of K0 from AssembleStiffness.__init__,
 * assembly.py:446-461.) */
#define PMT_K0_CLASS_PATTERN 1
#define PMT_K0_REFLECTION 2
#define PMT_K0_RIGID 4
#define PMT_K0_CUBE 8
#define PMT_K0_EXT_LEN 684
int pmt_k0_extend(int dim, const double* K0_host, double* ext_host);
/* Launch-shape tuning of the fine-grid operator, process-wide, read at every launch.
 * PMT_TUNE_FINE_KERNEL: 0 = 3D node tiles, 1 = xy node columns streamed along z, 2 = streamed
 * columns multiplying in the reflection basis (default; used when k0_pattern has PMT_K0_REFLECTION,
 * else 1; 3D only).
 * PMT_TUNE_REFL_ROWS: block shape of the reflection-basis kernel, node rows per block: 16 (default),
 * 24 or 32 rows of 16 nodes, 12 rows of 32 nodes, or 0 = the shape with the fewest threads per node
 * plane.  All shapes give the same bits for the operator and the sweeps (the fused dots of PCG are
 * grouped by block).
 * PMT_TUNE_STREAM_PLANES: node planes per streamed column, 0 = chosen from the CU count (default).
 * PMT_TUNE_STREAM_TILE: streamed column shape, 0 = 32x8 (default), 1 = 64x4, 2 = 16x16 nodes.
 * PMT_TUNE_COARSE1: operator of the first coarse level of hex8 hierarchies, 1 = element by element in
 * the reflection basis from the transformed child densities, no stencil stored (default; used when
 * k0_pattern has PMT_K0_REFLECTION and the level is not the coarsest), 0 = half stencil.  Read by
 * pmt_mg_create: it applies to hierarchies created afterwards.
 * PMT_TUNE_FILTER: FilterConv kernels for the 3^3 / 5^3 (3^2 / 5^2) extents, 1 = LDS-tiled z-streamed
 * blocks (default), 0 = one thread per element gathering through the caches; both sum the taps in the
 * same order.
 * pmt_get_tuning returns the current value or a negative status.
 * (No reference counterpart: the launch shapes picked by the autotuner, pymoto_amd/tune.py.) */
#define PMT_TUNE_FINE_KERNEL 1
#define PMT_TUNE_STREAM_PLANES 2
#define PMT_TUNE_STREAM_TILE 3
#define PMT_TUNE_REFL_ROWS 4
#define PMT_TUNE_COARSE1 5
#define PMT_TUNE_FILTER 6
int pmt_set_tuning(int param, int value);
int pmt_get_tuning(int param);
/* diag(K) (DampedJacobi.update, iterative.py:38-39) */
int pmt_stiffness_diagonal(int nelx, int nely, int nelz, const double* K0, const double* xe,
                           const uint8_t* bcmask, double bcdiag, double* diag, void* stream);

/* Sensitivity dx_e (+)= alpha * u_e^T K0 v_e with Dirichlet entries of u, v zeroed.
 * Replaces AssembleGeneral._sensitivity + DyadicMatrix.contract
 * (assembly.py:298-315, pymoto/common/dyadcarrier.py:392-414). */
int pmt_sens_contract(int nelx, int nely, int nelz, const double* K0, const uint8_t* bcmask,
                      const double* u, const double* v, double alpha, double* dx, int accumulate,
                      void* stream);
/* The same contraction with the k0_pattern of pmt_stiffness_apply: with PMT_K0_REFLECTION (hex8, K0 the
 * extended buffer of pmt_k0_extend) each element is contracted in the reflection basis (two 8-point
 * transforms and 96 multiply-adds; one transform when u == v); otherwise as pmt_sens_contract. */
int pmt_sens_contract_k0(int nelx, int nely, int nelz, const double* K0, int k0_pattern, const uint8_t* bcmask,
                         const double* u, const double* v, double alpha, double* dx, int accumulate, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Assembled CSR (the reference's output format).  Pattern: per row the sorted union of the dofs
 * of every element touching the row's node, Dirichlet rows/cols kept as explicit zeros
 * (assembly.py:100-230).  index_bytes = 4 (int32, what scipy stores) or 8 (int64).
 */
long long pmt_csr_nnz(int nelx, int nely, int nelz, int ndof);
int pmt_csr_pattern(int nelx, int nely, int nelz, int ndof, int index_bytes, void* indptr, void* indices,
                    void* stream);
/* Values bit-identical to np.add.at (assembly.py:257-272): per slot, products x_e*K0_ij rounded
 * and summed in ascending element order; bc slots 0 except bcdiag on the diagonal. */
int pmt_csr_values(int nelx, int nely, int nelz, const double* K0, const double* xe,
                   const uint8_t* bcmask, double bcdiag, int index_bytes, const void* indptr,
                   double* data, void* stream);
/* pmt_csr_values for ndof unknowns per node: ndof = the grid dimension (as pmt_csr_values) or 1 (the
 * scalar matrix of AssemblePoisson, assembly.py:523-560, bit 0 of bcmask). */
int pmt_csr_values_ndof(int nelx, int nely, int nelz, int ndof, const double* K0, const double* xe,
                        const uint8_t* bcmask, double bcdiag, int index_bytes, const void* indptr, double* data,
                        void* stream);

/* ---------------------------------------------------------------------------------------------
 * Geometric multigrid transfers with the trilinear (bilinear) weights 1, 1/2, 1/4, 1/8 of
 * GeometricMultigrid.setup_interpolation (iterative.py:178-220); grids given by the FINE element
 * counts (all even).  rc = R^T rf ; uf += R uc.
 */
int pmt_restrict(int nelx, int nely, int nelz, const double* rf, double* rc, void* stream);
int pmt_prolong_add(int nelx, int nely, int nelz, const double* uc, double* uf, void* stream);

/* ---------------------------------------------------------------------------------------------
 * MG-preconditioned CG runtime.  Replaces CG(preconditioner=GeometricMultigrid(...)) with nested
 * GeometricMultigrid levels, DampedJacobi smoothers and a direct coarsest solve
 * (iterative.py:21-47, 124-256, 295-403).  nlevels counts grids (fine grid included); level l
 * (l < nlevels-1) smooths with omega[l], steps[l] sweeps; the coarsest grid is solved directly.
 * Coarse operators are Galerkin R^T A R (iterative.py:173) built element by element.
 */
typedef struct pmt_mg pmt_mg;
int pmt_mg_create(int nelx, int nely, int nelz, int nlevels, const double* omega, const int* steps,
                  const double* K0_host, pmt_mg** out);
int pmt_mg_destroy(pmt_mg* mg);
/* Arithmetic of the V-cycle: 64 (FP64, the reference's), 32 (FP32 operator, smoother, transfers and
 * stencils built from the FP64 Galerkin hierarchy; the coarsest solve and CG stay FP64) or 16 (as 32,
 * but the damped-Jacobi sweeps of the fine grid run on the matrix cores with BF16 operands and FP32
 * accumulation; hex8 grids only).  Takes effect at the next pmt_mg_update.
 * (No reference counterpart: the mixed-precision variants.) */
int pmt_mg_set_precision(pmt_mg* mg, int bits);
/* Smoother of every level: type 0 = damped Jacobi with the per-level omega (the reference's), type 1 =
 * Chebyshev polynomial in D^-1 A of degree `steps` (same number of operator sweeps) on
 * [lower*hi, hi], hi = lambda_max(diag(K0)^-1 K0) on the fine grid (an upper bound of
 * lambda_max(D^-1 K) for any positive densities) and upper * (power-iteration estimate) on the
 * Galerkin levels.  Takes effect at the next pmt_mg_update.  pmt_mg_smoother_interval returns the
 * interval in use on a level.  (No reference counterpart.) */
int pmt_mg_set_smoother(pmt_mg* mg, int type, double lower, double upper);
int pmt_mg_smoother_interval(pmt_mg* mg, int level, double* lo, double* hi);
/* Rebuild diagonals, Galerkin coarse operators and the coarsest inverse for new densities xe
 * (GeometricMultigrid.update, iterative.py:168-176).  xe and bcmask must stay alive until the
 * next update.  Hex8 grids: the coarsest dense inverse may still be in progress on the handle's own
 * stream when this returns; pmt_mg_vcycle, pmt_pcg_solve, pmt_block_pcg_solve and the next update
 * make their stream wait for it (with the legacy default stream, NULL, the first of them waits for it on
 * the host).  Calls on one handle share its work vectors: the caller orders them (one stream, or events
 * between streams); they must not overlap on the device. */
int pmt_mg_update(pmt_mg* mg, const double* xe, const uint8_t* bcmask, double bcdiag, void* stream);
/* z = M r, one V-cycle (GeometricMultigrid.solve with x0=None, iterative.py:222-256). */
int pmt_mg_vcycle(pmt_mg* mg, const double* r, double* z, void* stream);
/* y = A_l u on level l (A_0 is the fine operator) - exposes the Galerkin operators. */
int pmt_mg_level_apply(pmt_mg* mg, int level, const double* u, double* y, void* stream);
/* Number of dofs of level l, or a negative status. */
long long pmt_mg_level_ndof(pmt_mg* mg, int level);
/* Bytes of device memory held by the handle. */
long long pmt_mg_device_bytes(pmt_mg* mg);
/* PCG with the V-cycle preconditioner, CG.solve semantics (iterative.py:340-403): x holds x0 on
 * entry unless x0_is_zero; stops when ||r||/||b|| <= tol; explicit residual every `restart`
 * iterations including the first.  iters/relres receive the iteration count and final relative
 * residual; history (host, maxit+1 doubles, may be NULL) the relative residual per iteration.  On
 * PMT_OK (maxit reached included) x holds the last iterate; after PMT_ERR_NUMERIC (a non-finite
 * right-hand side, start or residual) its content is unspecified. */
int pmt_pcg_solve(pmt_mg* mg, const double* b, double* x, int x0_is_zero, double tol, int maxit,
                  int restart, int* iters, double* relres, double* history, void* stream);

/* Coarse-space deflation inside pmt_pcg_solve (no reference counterpart: the learned coarse correction
 * of pymoto_amd/learned.py): with 1 <= m <= 4 columns W and AW = A W (FP64, caller-owned device memory,
 * column j at W + j*ndof, kept alive while set) and Einv = (W^T A W)^-1 (host, m x m row-major, copied),
 * every preconditioned residual is projected, z <- z - W Einv (AW)^T z, so the search directions stay
 * A-orthogonal to span W (with the start x0 = W Einv W^T b the residuals stay orthogonal to it: deflated
 * PCG).  m = 0 removes it.  Not used by pmt_block_pcg_solve. */
int pmt_mg_set_coarse_space(pmt_mg* mg, int m, const double* W, const double* AW, const double* Einv_host);
/* Block PCG for 1 <= k <= 8 right-hand sides sharing the operator: CG.solve with an (ndof, k)
 * right-hand side (iterative.py:259-403), Gram-Schmidt orth with rank-deficient columns dropped.
 * B and X hold k columns of pmt_mg_level_ndof(mg, 0) doubles each, column j at B + j*ndof (the
 * transpose of the reference's (ndof, k) array).  Stops when max_j ||r_j||/||b_j|| <= tol; relres
 * (host, k doubles) receives the final per-column values, history (host, maxit+1 doubles, may be
 * NULL) the maximum per iteration.  use_graph = 1 replays each iteration from HIP graphs captured
 * on the first call (re-captured after pmt_mg_update, pmt_mg_set_precision or pmt_set_tuning).
 * Allocates its workspace on the first call with a larger k. */
int pmt_block_pcg_solve(pmt_mg* mg, int k, const double* B, double* X, int x0_is_zero, double tol, int maxit,
                        int restart, int use_graph, int* iters, double* relres, double* history, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Scalar problems (one unknown per node): the matrix K(x) = sum_e x_e P0 of AssemblePoisson
 * (pymoto/modules/assembly.py:523-560; AssembleGeneral with a (2^D x 2^D) element matrix,
 * assembly.py:38-296) - thermal conduction, electric permittivity.  bcmask bit 0 marks the node's
 * unknown as Dirichlet (row and column zeroed, bcdiag on the diagonal).  P0 is a device copy of the
 * element matrix (row-major, local node order of pymoto/common/domain.py:131-142).
 */
int pmt_scalar_apply(int nelx, int nely, int nelz, const double* P0, const double* xe, const uint8_t* bcmask,
                     double bcdiag, const double* u, double* y, void* stream);
int pmt_scalar_residual(int nelx, int nely, int nelz, const double* P0, const double* xe, const uint8_t* bcmask,
                        double bcdiag, const double* b, const double* u, double* r, void* stream);
int pmt_scalar_diagonal(int nelx, int nely, int nelz, const double* P0, const double* xe, const uint8_t* bcmask,
                        double bcdiag, double* diag, void* stream);
/* dx_e (+)= alpha * u_e^T P0 v_e with Dirichlet entries zeroed: AssembleGeneral._sensitivity +
 * DyadicMatrix.contract (assembly.py:298-315, pymoto/common/dyadcarrier.py:392-414) for ndof = 1. */
int pmt_scalar_sens_contract(int nelx, int nely, int nelz, const double* P0, const uint8_t* bcmask, const double* u,
                             const double* v, double alpha, double* dx, int accumulate, void* stream);
/* Multigrid-preconditioned CG for the scalar matrix: nested GeometricMultigrid levels with damped
 * Jacobi and Galerkin R^T A R stencils (27-point hex8, 9-point Q4), explicit inverse on the coarsest
 * grid, CG.solve for one right-hand side (pymoto/solvers/iterative.py:21-47,124-256,340-403).  Same
 * argument meaning as pmt_mg_create / pmt_mg_update / pmt_mg_vcycle / pmt_pcg_solve; a non-finite
 * right-hand side, start vector or residual returns PMT_ERR_NUMERIC.  Allocates at create only. */
typedef struct pmt_smg pmt_smg;
int pmt_smg_create(int nelx, int nely, int nelz, int nlevels, const double* omega, const int* steps,
                   const double* P0_host, pmt_smg** out);
int pmt_smg_destroy(pmt_smg* h);
long long pmt_smg_level_nodes(pmt_smg* h, int level);
int pmt_smg_update(pmt_smg* h, const double* xe, const uint8_t* bcmask, double bcdiag, void* stream);
int pmt_smg_vcycle(pmt_smg* h, const double* r, double* z, void* stream);
int pmt_smg_level_apply(pmt_smg* h, int level, const double* u, double* y, void* stream);
int pmt_smg_pcg_solve(pmt_smg* h, const double* b, double* x, int x0_is_zero, double tol, int maxit, int restart,
                      int* iters, double* relres, double* history, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Learned warm start (no reference counterpart): kernels of a 3D convolutional encoder-decoder on
 * the element grid.  Activations are channels-last BF16 arrays X[voxel][c] (voxel in pyMOTO's
 * element order, c contiguous), weights BF16 W[tap][cout][cin] with tap = (dz+1)*9 + (dy+1)*3 +
 * (dx+1), 3x3x3 stencil with zero padding; products on the matrix cores with FP32 accumulation.
 * BF16 buffers are passed as void* (2 bytes per value).  The voxel count must be a multiple of 16.
 */
/* Y = act(bias + conv(X, W)); cin in {8,16,32,64}, cout in {16,32,64}; bias (FP32) may be NULL;
 * relu = 1 applies max(0, .). */
int pmt_conv3d_fwd(int nx, int ny, int nz, int cin, int cout, const void* X, const void* W, const float* bias,
                   int relu, void* Y, void* stream);
/* dX = conv^T(dY): the forward kernel on the flipped, transposed weights written to Wt (27*cin*cout
 * BF16, caller-provided); cin in {16,32,64}. */
int pmt_conv3d_bwd_data(int nx, int ny, int nz, int cin, int cout, const void* dY, const void* W, void* Wt, void* dX,
                        void* stream);
/* dW[tap][cout][cin] (FP32) = sum_v dY[v][co] X[v+off(tap)][ci]; dbias (FP32, may be NULL) = sum_v dY;
 * work: pmt_conv3d_wgrad_work_size(cin, cout) floats.  Deterministic. */
int pmt_conv3d_wgrad_work_size(int cin, int cout);
int pmt_conv3d_bwd_weight(int nx, int ny, int nz, int cin, int cout, const void* X, const void* dY, float* dW,
                          float* dbias, float* work, void* stream);
/* G *= (Y > 0) elementwise (ReLU adjoint from the layer output). */
int pmt_relu_mask_bf16(long long n, const void* Y, void* G, void* stream);
/* 2x2x2 average pooling of a (nx, ny, nz, c) field and its adjoint (nx, ny, nz even). */
int pmt_pool2_bf16(int nx, int ny, int nz, int c, const void* X, void* Y, void* stream);
int pmt_pool2_bwd_bf16(int nx, int ny, int nz, int c, const void* dY, void* dX, void* stream);
/* Y = nearest-upsample(X) + S on the fine (nx, ny, nz) grid; adjoint to X sums the 8 children. */
int pmt_up2_add_bf16(int nx, int ny, int nz, int c, const void* X, const void* S, void* Y, void* stream);
int pmt_up2_bwd_bf16(int nx, int ny, int nz, int c, const void* dY, void* dX, void* stream);
/* Nodal displacement u (FP32, dof = node*3 + d) = mean of channels 0..2 over the elements around each
 * node; and its adjoint into a (nel, c) BF16 gradient (channels >= 3 zeroed). */
int pmt_elem_to_node(int nelx, int nely, int nelz, int c, const void* Xe, float* u, void* stream);
int pmt_elem_to_node_bwd(int nelx, int nely, int nelz, int c, const float* du, void* dXe, void* stream);
/* Adam on FP32 master weights (t = 1-based step), refreshing the BF16 copy wb (may be NULL). */
int pmt_adam_step(long long n, float lr, float b1, float b2, float eps, int t, const float* grad, float* w, float* m,
                  float* v, void* wb, void* stream);
int pmt_add_bf16(long long n, const void* a, const void* b, void* y, void* stream);
int pmt_f32_to_bf16(long long n, const float* x, void* y, void* stream);
int pmt_bf16_to_f32(long long n, const void* x, float* y, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PYMOTO_AMD_H */
