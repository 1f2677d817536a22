/*
 * rgf.h -- C ABI of the MI355X (gfx950) recursive Green's function solver.
 *
 * Drop-in compute layer for qttools' GF solver plugin:
 *   GFSolver.selected_inv / selected_solve  (src/qttools/greens_function_solver/solver.py:31-113)
 *   RGF                                      (src/qttools/greens_function_solver/rgf.py:10-438)
 *   linalg.inv                               (src/qttools/kernels/linalg/inv.py:24-40)
 * The Python class quatrex_amd.greens_function_solver.RGF binds these symbols
 * with ctypes and keeps the reference's Python signatures.
 *
 * Conventions
 *  - Complex values are interleaved (re, im): complex128 = 2 x double,
 *    complex64 = 2 x float -- the numpy/torch memory layout.
 *  - All matrix pointers are DEVICE pointers; host arrays are marked "host".
 *  - Block-tridiagonal operands use the packed layout: for block sizes
 *    b_0..b_{N-1}, block (i,i) (b_i x b_i, row-major) sits at element offset
 *    sum_{k<i} b_k^2 of `diag`; blocks (i,i+1) (b_i x b_{i+1}) and (i+1,i)
 *    (b_{i+1} x b_i) sit at offset sum_{k<i} b_k b_{k+1} of `upper` and
 *    `lower`.  Energies are the slowest axis (stride = total diag / offdiag
 *    size).  These are the dense blocks of DSDBSparse.blocks[i, j]
 *    (src/qttools/datastructures/dsdbsparse.py:1285-1363).
 *  - Calls are asynchronous on the context's HIP stream; results are ready
 *    after the stream is synchronised.  A singular pivot does not stop the
 *    sweep (outputs then hold inf/NaN, as with cupy); rgf_status() reports it.
 *  - No C++ exception crosses this interface; every call returns an rgf_status.
 */
#ifndef QUATREX_AMD_RGF_H
#define QUATREX_AMD_RGF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (the Python layer maps them to the reference's exceptions:
 * EINVAL -> ValueError/TypeError, ESINGULAR -> numpy.linalg.LinAlgError). */
#define RGF_OK 0
#define RGF_EINVAL 1
#define RGF_ESINGULAR 2
#define RGF_EHIP 3
#define RGF_ENOMEM 4
#define RGF_EUNSUPPORTED 5

/* Options for rgf_set_option. */
#define RGF_OPT_MAX_CHUNK 1        /* max energies per internal sweep (0 = auto)           */
#define RGF_OPT_WORKSPACE_BUDGET 2 /* bytes the context may allocate for workspace (auto)  */
#define RGF_OPT_TILE_CFG 3         /* GEMM tile config for all block sizes (-1 = autotune) */
#define RGF_OPT_UNFUSED 5          /* 1: launch every product group on its own (the plain,
                                      ungrouped op sequence; with RGF_OPT_MAX_CHUNK = 1 this is
                                      the one-energy-at-a-time baseline)                      */
#define RGF_OPT_OVERLAP 6          /* 1 (default): each block inversion runs on a second stream of
                                      default priority, overlapped with the independent
                                      lesser/greater products of its forward step              */
#define RGF_OPT_PROFILE 4          /* 1: time every sweep GEMM launch and every block inversion
                                      with HIP events on the context stream (0: off)          */
#define RGF_OPT_PERSISTENT 7       /* 1 (default): FP64 selected solves of geometries whose blocks
                                      are all <= 32 run as one persistent kernel, one workgroup
                                      per energy with every intermediate block in LDS; 0: the
                                      launch-per-product-group sweep for every geometry        */

typedef struct rgf_ctx rgf_ctx;

/* Block partition of one device: n_blocks blocks of block_sizes[i] (host array). */
typedef struct rgf_geometry {
  int32_t n_blocks;
  const int64_t* block_sizes;
} rgf_geometry;

/* Packed block-tridiagonal operand (device pointers, see Conventions). */
typedef struct rgf_bt_const {
  const void* diag;
  const void* upper;
  const void* lower;
} rgf_bt_const;

typedef struct rgf_bt {
  void* diag;
  void* upper;
  void* lower;
} rgf_bt;

/* Open-boundary blocks, the content of OBCBlocks
 * (src/qttools/greens_function_solver/solver.py:12-28): each member is NULL
 * or a host array of n_blocks device pointers, NULL entries meaning "no OBC
 * on this block"; a non-NULL entry i points to an (n_energies, b_i, b_i)
 * complex128 array.  RGF applies A_ii - B^R_i and Sigma^<>_ii + B^<>_i
 * (rgf.py:216-262). */
typedef struct rgf_obc {
  const void* const* retarded;
  const void* const* lesser;
  const void* const* greater;
} rgf_obc;

/* Mixed-precision options for rgf_selected_solve_mixed. */
typedef struct rgf_mixed_opts {
  int32_t refine;            /* >= 0: FP64 residual refinement steps per inverted block;
                                -1: the Schur complement S_j = A_jj - A_ji g_i A_ij (and its
                                factor A_ji g_i) in FP64, no refinement step */
  double fallback_tol;       /* FP64 re-solve when an energy's refinement residual exceeds this
                                (<= 0 disables the fallback) */
  int32_t* fallback_flags;   /* optional device int32[n_energies]: 1 where the FP64 fallback ran */
  double* residuals;         /* optional device double[n_energies]: max relative refinement residual */
} rgf_mixed_opts;

/* Context: one per device and stream.  stream = NULL uses the null stream.
 * Replaces the implicit cupy device/stream state of the reference. */
int rgf_create(int device, void* stream, rgf_ctx** out);
int rgf_destroy(rgf_ctx* ctx);
int rgf_set_stream(rgf_ctx* ctx, void* stream);
int rgf_set_option(rgf_ctx* ctx, int32_t option, int64_t value);

/* Copies the last error message of ctx (NUL-terminated) into buf. */
int rgf_last_error(const rgf_ctx* ctx, char* buf, size_t len);

/* Synchronises the stream; info (host int64[3]) receives (energy, block,
 * column) of the first exactly-zero pivot since the last call, or -1s.
 * Returns RGF_ESINGULAR in that case, like numpy.linalg.inv raising
 * LinAlgError (src/qttools/kernels/linalg/inv.py:6-11). */
int rgf_status(rgf_ctx* ctx, int64_t* info);

/* Workspace (bytes) the context needs to run a sweep over n_energies
 * energies at once (selected_solve when solve != 0, with current and
 * retarded-output choices as given); the admission check of the serving
 * loop.  No allocation happens. */
int rgf_workspace_size(const rgf_geometry* geom, int64_t n_energies, int32_t solve,
                       int32_t with_current, int32_t with_retarded, size_t* bytes);

/* Caller-owned workspace (e.g. a torch tensor); NULL/0 returns to
 * context-owned workspace.  The context never frees caller memory. */
int rgf_set_workspace(rgf_ctx* ctx, void* ptr, size_t bytes);

/* Batched inverse of n_mats contiguous n x n complex128 matrices: out = a^{-1}.
 * Replaces linalg.inv (src/qttools/kernels/linalg/inv.py:24-40).  a is not
 * modified.  n <= 1024. */
int rgf_inv_z(rgf_ctx* ctx, int64_t n, int64_t n_mats, const void* a, void* out);

/* System-matrix assembly on the device for n_energies energies:
 *   A_e = z_e I - H - S_B
 * z: device complex128[n_energies] (E + i eta); h: ONE energy-independent
 * block-tridiagonal H (diag, upper, lower); sigma_b: NULL or a host array of
 * n_blocks device pointers to single b_i x b_i retarded boundary blocks
 * (NULL entries: none).  Mirrors the assembly in the electron solver
 * (src/quatrex/electron/solver.py:668-709) for a fixed boundary term. */
int rgf_assemble_z(rgf_ctx* ctx, const rgf_geometry* geom, int64_t n_energies, const void* z,
                   const rgf_bt_const* h, const void* const* sigma_b, rgf_bt* a);

/* Full system-matrix assembly of the electron solver on the tridiagonal
 * blocks, replacing ElectronSolver._assemble_system_matrix
 * (src/quatrex/electron/solver.py:668-709 with :541-665), per energy e:
 *   M = z_e S - P - (Sigma^R_herm + (Sigma^> - Sigma^<) / 2 + H)
 * z: device complex128[n_energies] (E + i eta); S: the overlap (NULL: the
 * identity); P: S_ab (v_a + v_b) / 2 with an overlap, diag(v) without one
 * (potential NULL: no term); H: one packed tridiagonal Hamiltonian for all
 * energies, or one per energy (hamiltonian_per_energy = 1, energy strides of
 * the packed layout); each self-energy NULL (zero) or packed per energy.
 * Overlap and self-energies need diag, upper and lower. */
typedef struct rgf_assembly {
  const void* z;
  rgf_bt_const hamiltonian;
  int32_t hamiltonian_per_energy;
  const rgf_bt_const* overlap;
  const double* potential;
  const rgf_bt_const* sigma_retarded_herm;
  const rgf_bt_const* sigma_lesser;
  const rgf_bt_const* sigma_greater;
} rgf_assembly;
int rgf_assemble_system_z(rgf_ctx* ctx, const rgf_geometry* geom, int64_t n_energies,
                          const rgf_assembly* assembly, rgf_bt* a);

/* Selected inversion X = A^{-1} on the tridiagonal pattern, complex128.
 * Replaces RGF.selected_inv (rgf.py:25-113).  obc_retarded: NULL or a host
 * array of n_blocks device pointers (see rgf_obc).  x->upper / x->lower may be
 * NULL to skip those outputs. */
int rgf_selected_inv_z(rgf_ctx* ctx, const rgf_geometry* geom, int64_t n_energies,
                       const rgf_bt_const* a, const void* const* obc_retarded, rgf_bt* x);

/* Selected solve, complex128 throughout.  Replaces RGF.selected_solve
 * (rgf.py:115-438):
 *   G^R = (A - B^R)^{-1},  G^<> = G^R (Sigma^<> + B^<>) G^R^H
 * on the tridiagonal pattern.  Only sigma->diag and sigma->upper are read
 * (skew-Hermitian inputs, solver.py:87-91).
 *  g_retarded: NULL, or diag (and optionally upper/lower) outputs.  The
 *              reference writes diagonal G^R only; upper/lower extend it.
 *  g_lesser / g_greater: diag and upper required; lower optional (written as
 *              -G^<_{i,i+1}^H, rgf.py:357-359; pass NULL for symmetric storage).
 *              Stored diagonal blocks are (X - X^H)/2 (rgf.py:305-310).
 *  current:    NULL, or device complex128[n_energies][n_blocks+1]: the
 *              Meir-Wingreen layer currents (rgf.py:388-428); then obc->lesser
 *              and obc->greater must hold blocks 0 and n_blocks-1.
 */
int rgf_selected_solve_z(rgf_ctx* ctx, const rgf_geometry* geom, int64_t n_energies,
                         const rgf_bt_const* a, const rgf_bt_const* sigma_lesser,
                         const rgf_bt_const* sigma_greater, const rgf_obc* obc,
                         rgf_bt* g_retarded, rgf_bt* g_lesser, rgf_bt* g_greater, void* current);

/* rgf_selected_solve_z without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot call hipMalloc (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).
 * Same arguments, kernels and results as rgf_selected_solve_z; its
 * workspace and tables live in one fixed array per GPU of
 * RGF_NOALLOC_SCRATCH_BYTES, held until the process ends.  The energies run
 * in chunks that fit the array; a geometry whose single energy point does
 * not fit returns RGF_EUNSUPPORTED with the sizes in rgf_last_error.  The
 * array keeps the last call's contents, and every call on one GPU shares it:
 * calls that overlap in time on one GPU (several contexts or streams) race
 * and must be serialised by the caller.  Profiling and the second-stream
 * inversion overlap are off inside this call (both would create HIP
 * objects). */
#define RGF_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int rgf_selected_solve_noalloc_z(rgf_ctx* ctx, const rgf_geometry* geom, int64_t n_energies,
                                 const rgf_bt_const* a, const rgf_bt_const* sigma_lesser,
                                 const rgf_bt_const* sigma_greater, const rgf_obc* obc,
                                 rgf_bt* g_retarded, rgf_bt* g_lesser, rgf_bt* g_greater,
                                 void* current);

/* Several selected solves in one call, e.g. the energy points of different
 * devices (geometries) that arrived together.  Each problem has the meaning
 * of one rgf_selected_solve_z call (g_retarded.diag NULL: no retarded output;
 * upper / lower NULL: not written).  Problems whose blocks are all <= 32 share
 * ONE persistent kernel launch: one workgroup per energy point with every
 * intermediate block in LDS, energy points of all those problems pulled
 * from one device work queue, the most expensive geometries first.  The other
 * problems run one after the other by the batched sweep.  No reference
 * counterpart (uneven batching of the serving path; the reference solves one
 * device per call, rgf.py:115-438). */
typedef struct rgf_solve_problem {
  rgf_geometry geom;
  int64_t n_energies;
  rgf_bt_const a, sigma_lesser, sigma_greater;
  const rgf_obc* obc;
  rgf_bt g_retarded, g_lesser, g_greater;
  void* current;
} rgf_solve_problem;
int rgf_selected_solve_batch_z(rgf_ctx* ctx, int32_t n_problems, const rgf_solve_problem* problems);

/* Mixed-precision selected solve: same interface and outputs (complex128) as
 * rgf_selected_solve_z; the Schur-complement GEMMs and G Sigma G^H triple
 * products run in complex64 on FP32 matrix cores with FP32 accumulation,
 * the diagonal-block inversions in FP64, followed by `opts->refine` FP64
 * residual-refinement steps per inverted block (or, with refine = -1, the
 * Schur complement that feeds each inversion computed in FP64, which is
 * what those steps converge to, for fewer products); energies whose refinement
 * residual exceeds opts->fallback_tol are re-solved in FP64.  No reference
 * counterpart: the precision mode of this build's serving path. */
int rgf_selected_solve_mixed(rgf_ctx* ctx, const rgf_geometry* geom, int64_t n_energies,
                             const rgf_bt_const* a, const rgf_bt_const* sigma_lesser,
                             const rgf_bt_const* sigma_greater, const rgf_obc* obc,
                             const rgf_mixed_opts* opts, rgf_bt* g_retarded, rgf_bt* g_lesser,
                             rgf_bt* g_greater, void* current);

/* GEMM tile configuration for products of b x b blocks over n_energies
 * energies, chosen by timing every configuration on a representative set of
 * grouped products (f32 != 0: the FP32-compute products of the mixed path).
 * The winner is stored in ctx's table for block size b and returned in
 * best_cfg; ms_per_cfg (optional, RGF_NUM_TILE_CFGS floats) receives the
 * timings, -1 for configurations that do not apply.  Allocates temporary
 * device buffers.  No reference counterpart (the autotuner of this build). */
/* Configurations (index: output tile, threads per workgroup): 0 64x64 (256),
 * 1 32x32 (64), 2 64x32 (128), 3 128x64 (512); 4 64x64, 5 32x32 and
 * 6 128x64 with three real MFMAs per complex product (3M: ar br, ai bi and
 * (ar + ai)(br + bi)) instead of four.  The FP32-compute products are built
 * for 0, 1 and 3. */
#define RGF_NUM_TILE_CFGS 7
int rgf_autotune(rgf_ctx* ctx, int64_t b, int64_t n_energies, int32_t f32, int32_t* best_cfg,
                 float* ms_per_cfg);

/* Accumulated launch statistics while RGF_OPT_PROFILE is on (synchronises
 * the stream).  Flops are algorithmic: 8*m*n*k real flops per complex
 * product term, 8*n^3 per n x n inversion.  gemm_bytes are algorithmic too:
 * per problem, op(A) and op(B) of each term and each addend read once, D
 * (and D2) written once, 16 bytes per element.  No reference counterpart. */
typedef struct rgf_profile {
  int64_t gemm_launches;
  double gemm_ms;
  double gemm_flops;
  int64_t inv_calls;
  double inv_ms;
  double inv_flops;
  double gemm_bytes;
} rgf_profile;
int rgf_profile_read(rgf_ctx* ctx, rgf_profile* out, int32_t reset);

/* Tridiagonal blocks of a block-sparse stack <-> the packed layout of rgf_bt.
 * Replaces the per-block densify / sparsify of DSDBCOO._get_block /
 * _set_block (src/qttools/datastructures/dsdbcoo.py:333-412, 466-521) and
 * DSDBCSR._get_block / _set_block (dsdbcsr.py:285-345, 420-472), for the
 * blocks RGF reads and writes (a.blocks[i,i], [i,i+1], [i+1,i]).
 * data: complex128, n_energies rows of data_stride elements (nnz per energy).
 * The map (device arrays, built once per sparsity structure) has one part per
 * output part (0 diag, 1 upper, 2 lower): entry k moves data[e][src[k]] to
 * part[e][dst[k]], dst being the packed offset inside that part.  op[k] flags:
 * 1 = mirror image of a stored entry under symmetric storage (densify only,
 * value op(v)); 2 = stored diagonal element of a symmetric diagonal block
 * (densifies to (v + op(v)) / 2); op(v) = s * conj(v), or s * v with flag 8,
 * s = -1 with flag 4.  rgf_densify_z zeroes every non-NULL output part whose
 * map does not cover it (blocks absent from the pattern read as zeros);
 * rgf_sparsify_z writes the
 * stored entries only (values outside the pattern are dropped, as in the
 * reference). */
typedef struct rgf_block_map_part {
  int64_t n;
  const int64_t* src;
  const int64_t* dst;
  const uint8_t* op;
  int32_t covers_part; /* 1: the entries write every element of the part (no zero fill needed) */
} rgf_block_map_part;
typedef struct rgf_block_map {
  rgf_block_map_part part[3];
} rgf_block_map;
int rgf_densify_z(rgf_ctx* ctx, const rgf_geometry* geom, int64_t n_energies, const void* data,
                  int64_t data_stride, const rgf_block_map* map, rgf_bt* out);
int rgf_sparsify_z(rgf_ctx* ctx, const rgf_geometry* geom, int64_t n_energies,
                   const rgf_bt_const* in, const rgf_block_map* map, void* data,
                   int64_t data_stride);

/* Open boundary conditions.  All matrices are n_energies stacked b x b
 * complex128 device matrices (row-major, energy stride b*b).
 *
 * rgf_sancho_rubio_z: surface Green's function by Sancho-Rubio decimation,
 * replacing SanchoRubio.__call__ (src/qttools/boundary_conditions/obc/
 * sancho_rubio.py:33-85) for a whole energy batch.  Per iteration:
 * inverse = inv(eps); eps -= alpha inverse beta + beta inverse alpha;
 * eps_s -= alpha inverse beta; alpha = alpha inverse alpha; beta = beta
 * inverse beta, starting from eps = eps_s = a_ii, alpha = a_ji, beta = a_ij.
 * The loop stops after the first iteration whose delta = max over energies of
 * ||(|alpha| + |beta|)||_F / 2 is below convergence_tol, or after
 * max_iterations; x_ii = inv(eps_s).  *iterations (host) receives the number
 * of iterations run and *delta (host) the last delta (NaN if none ran): the
 * batch did not converge unless *delta < convergence_tol (the reference then
 * warns).  Synchronises the context stream once per iteration (the stopping
 * rule is evaluated on the host, as in the reference).
 *
 * rgf_obc_fixed_point_z: one refinement step x_ref = inv(a_ii - a_ji x a_ij)
 * with the residuals abs = ||x_ref - x||_F and rel = abs / ||x_ref||_F per
 * energy (device arrays, both NULL to skip), replacing
 * OBCSystem._fixed_point_step / _get_residuals (obc.py:102-127, 154-202).
 *
 * rgf_contact_obc_z: contact self-energies from the surface Green's function,
 * replacing ElectronSolver._compute_contact_obc (src/quatrex/electron/
 * solver.py:486-495): sigma = m_10 g_00 m_01, gamma = i (sigma - sigma^H),
 * lesser = i f gamma, greater = i (f - 1) gamma with f = occupancies[e]
 * (device, float64).  lesser and greater may be NULL.
 *
 * rgf_boundary_system_z: the boundary system handed to the OBC solver,
 * a_ii = m_00 + i eta s_00, a_ij = m_01 + i eta s_01, a_ji = m_10 + i eta s_10
 * (solver.py:468-483): with the overlap blocks s_* NULL, s_00 is the identity
 * and s_01 = s_10 = 0.  a_ij / a_ji may be NULL (not written); outputs may
 * alias their inputs. */
/* Beyn's contour-integral NEVP solver, replacing Beyn.__call__
 * (src/qttools/nevp/beyn.py:14-379, SVD projection) for a whole batch:
 * quadrature of T(z)^-1 = (sum_n z^n a_n)^-1 on the circles |z| = r_o and
 * |z| = r_i (num_quad_points nodes each, beyn.py:183-238), moments
 * P_0 = q_0 Y, P_1 = q_1 Y with a d x m_0 probe block Y of uniform [0,1)
 * real and imaginary parts drawn from `seed` (beyn.py:307-318), thin SVD
 * of P_0 with the rank cut s > s_max m_0 eps, and the eigenpairs of
 * U^H P_1 V S^-1 (beyn.py:101-145, 240-282).  m_0 <= 64. */
typedef struct rgf_beyn_opts {
  double r_o, r_i;
  int32_t m_0, num_quad_points;
  uint64_t seed;
} rgf_beyn_opts;

/* Mode selection of the spectral OBC solver (spectral.py:62-99 defaults). */
typedef struct rgf_spectral_opts {
  double min_decay, max_decay, min_propagation, residual_tolerance, eta_decay;
  int32_t num_ref_iterations, residual_normalization;
} rgf_spectral_opts;

/* a_xx: host array of n_coef (3, 5 or 7) device pointers to (n_energies, d, d)
 * coefficient blocks, lowest order first (n = -(n_coef/2) .. n_coef/2).
 * ws: device complex128[n_energies][m_0] (0 beyond the rank); vs: device
 * complex128[n_energies][d][m_0], unit columns within the rank, zero beyond;
 * n_valid: NULL or device int32[n_energies], the rank per energy. */
int rgf_beyn_z(rgf_ctx* ctx, int64_t d, int64_t n_energies, int32_t n_coef, const void* const* a_xx,
               const rgf_beyn_opts* opts, void* ws, void* vs, int32_t* n_valid);

/* Spectral surface Green's function, replacing Spectral.__call__
 * (src/qttools/boundary_conditions/obc/spectral.py:482-629) with
 * nevp = Beyn and block_sections = 1: Beyn on (a_ji, a_ii, a_ij), unit-norm
 * modes, the reflected-mode mask (spectral.py:221-327),
 * x = inv(a_ii + a_ji (vr / w) inv(vr^H vr) vr^H) (spectral.py:386-436), then
 * num_ref_iterations steps x = inv(a_ii - a_ji x a_ij).  eigenvalues: NULL or
 * device complex128[n_energies][m_0]; n_reflected: NULL or device
 * int32[n_energies], the number of reflected modes. */
int rgf_spectral_obc_z(rgf_ctx* ctx, int64_t b, int64_t n_energies, const void* a_ii,
                       const void* a_ij, const void* a_ji, const rgf_beyn_opts* beyn,
                       const rgf_spectral_opts* opts, void* x_ii, void* eigenvalues,
                       int32_t* n_reflected);

int rgf_sancho_rubio_z(rgf_ctx* ctx, int64_t b, int64_t n_energies, const void* a_ii,
                       const void* a_ij, const void* a_ji, int32_t max_iterations,
                       double convergence_tol, void* x_ii, int32_t* iterations, double* delta);
int rgf_obc_fixed_point_z(rgf_ctx* ctx, int64_t b, int64_t n_energies, const void* a_ii,
                          const void* a_ij, const void* a_ji, const void* x, void* x_ref,
                          double* abs_residuals, double* rel_residuals);
int rgf_contact_obc_z(rgf_ctx* ctx, int64_t b, int64_t n_energies, const void* m_10,
                      const void* g_00, const void* m_01, const double* occupancies, void* sigma,
                      void* lesser, void* greater);
int rgf_boundary_system_z(rgf_ctx* ctx, int64_t b, int64_t n_energies, double eta_obc,
                          const void* m_00, const void* m_01, const void* m_10, const void* s_00,
                          const void* s_01, const void* s_10, void* a_ii, void* a_ij, void* a_ji);

/* HIP graph of a complete selected-solve sweep for fixed buffers.
 * rgf_graph_capture_solve records every launch of one sweep over n_energies
 * energies (precision 0: FP64; 1: FP32-compute products with FP64
 * inversions and `refine` Newton steps, or FP64 Schur complements with refine
 * = -1, without the a-posteriori check) into
 * a graph that owns its geometry and job tables and its workspace - or, for
 * FP64, uses the context's caller-provided workspace (rgf_set_workspace) when
 * that is large enough, which must then outlive the graph; rgf_graph_launch
 * replays it on the context's stream with whatever data the buffers hold at
 * that time.  current: NULL, or device complex128[n_energies][n_blocks+1]
 * written by every replay (the Meir-Wingreen currents, as in
 * rgf_selected_solve_z; obc->lesser / obc->greater must then hold blocks 0 and
 * n_blocks-1).  The context must stay alive while the graph is launched;
 * rgf_graph_destroy waits for the last launch and may come after rgf_destroy
 * of the context.  No reference counterpart (launch-overhead removal for the
 * serving loop). */
typedef struct rgf_graph rgf_graph;
int rgf_graph_capture_solve(rgf_ctx* ctx, const rgf_geometry* geom, int64_t n_energies,
                            const rgf_bt_const* a, const rgf_bt_const* sigma_lesser,
                            const rgf_bt_const* sigma_greater, const rgf_obc* obc,
                            int32_t precision, int32_t refine, rgf_bt* g_retarded,
                            rgf_bt* g_lesser, rgf_bt* g_greater, void* current,
                            rgf_graph** out);
int rgf_graph_launch(rgf_graph* graph);
int rgf_graph_destroy(rgf_graph* graph);

/* Library version string. */
const char* rgf_version(void);

#ifdef __cplusplus
}
#endif

#endif /* QUATREX_AMD_RGF_H */
