/*
 * feaxhip.h -- C ABI of libfeaxhip.so, the MI355X (gfx950) implementation of
 * feax's HEX8 assembly + Krylov hot path.
 *
 * Conventions
 *   - Every entry point returns 0 on success, else a hipError_t value or one
 *     of the FX_E* codes below; fx_last_error() returns the message of the
 *     last failure on the calling thread.
 *   - Device entry points are stream-ordered on the hipStream_t passed as
 *     `stream` (void* here so the header needs no HIP include) and allocate
 *     no device memory: scratch comes in through `workspace`, whose size is
 *     given by the matching *_workspace_bytes() query.
 *   - The kernels-only entry points (pattern queries aside) are
 *     asynchronous.  The iterative drivers (fx_pcg_*, fx_pcg_matfree_f64,
 *     fx_pcg_gmg_f64, fx_bicgstab_f64, fx_gmres_f64, fx_pcg_batched_*)
 *     synchronise the stream to read their convergence flags every
 *     `check_every` iterations, and each call creates a private stream (of
 *     default priority) and instantiates a HIP graph of one chunk of
 *     iterations, released before it returns (fx_pcg_*: not with
 *     FX_PCG_NO_GRAPH or FX_PCG_TIME).  The batched drivers and
 *     fx_seg_project_f32 also synchronise after uploading their host-side
 *     instance tables.  The banded direct solver calls rocBLAS / rocSOLVER
 *     (one rocBLAS handle per device, created on first use and kept; rocBLAS
 *     manages its own workspace) and synchronises at the end of a
 *     factorisation; fx_band_solve_inv_f64 captures its launches once as a
 *     HIP graph per (band, inverse blocks, workspace) address set and keeps
 *     it until the process ends.  fx_pcg_f64_noalloc creates no stream,
 *     graph or event.
 *   - Index arrays are int32 (feax keeps all structure in int32,
 *     feax/problem.py:236-244, 441-451).
 *   - (d) marks device pointers, (h) host pointers.
 */
#ifndef FEAXHIP_H
#define FEAXHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FX_OK 0
#define FX_EINVAL 1001     /* bad argument / shape */
#define FX_ELIMIT 1002     /* a size limit of a kernel exceeded */
#define FX_EWORKSPACE 1003 /* workspace too small */
#define FX_ELIB 1004       /* a rocBLAS / rocSOLVER call failed */

/* Material kinds understood by the element kernels. */
#define FX_MAT_LINEAR_ELASTIC 0 /* sigma = lam tr(eps) I + 2 mu eps, E = param        */
#define FX_MAT_SIMP 1           /* as above with E = Emin + rho^p (E0 - Emin)         */
#define FX_MAT_NEO_HOOKEAN 2    /* P = mu (F - F^-T) + kappa (J-1) J F^-T, mu = param */
#define FX_MAT_CONDUCTION 3     /* vec = 1: flux k grad T, k = param                  */
#define FX_MAT_SIMP_CONDUCTION 4 /* vec = 1: k = Emin + rho^penal (E0 - Emin)          */

/* Where the per-material parameter lives. */
#define FX_PARAM_NODE 0 /* (N,) interpolated with the shape functions at quads */
#define FX_PARAM_CELL 1 /* (E,) constant over the element                       */

typedef struct fx_material {
  int32_t kind;           /* FX_MAT_*                               */
  int32_t param_loc;      /* FX_PARAM_*                             */
  double nu;              /* Poisson ratio (linear elastic / SIMP)  */
  double E0, Emin, penal; /* SIMP law                               */
  double kappa;           /* Neo-Hookean bulk modulus               */
  /* Thermal strain of the linear / SIMP laws (HEX8 only):
   * sigma = C(E) : (eps - alpha (T - t_ref) I) with T the node variable
   * `temperature` (device, num_nodes), interpolated at the Gauss points;
   * NULL: no thermal strain.  The tangent does not depend on T. */
  const double* temperature;
  double alpha, t_ref;
} fx_material;

/* Last error message of this thread (NUL-terminated, truncated to len). */
int fx_last_error(char* buf, size_t len);
/* Library identification string. */
int fx_version(char* buf, size_t len);

/* ------------------------------------------------------------------------
 * Structure (host).  Replaces Problem._build_csr_assembly_structure
 * (feax/problem.py:316-484) and the per-batch slot sorts
 * (feax/assembler.py:1164-1250) for HEX8, vec = 3, MatrixView.FULL: the
 * deduplicated CSR pattern of all (I, J) = element DOF pairs, row-major with
 * sorted columns, identical to feax's argsort + dedupe construction.
 *
 * Two calls: first with indices == NULL to get *nse, *max_degree and indptr
 * (3N+1); then with indices (nse), elem_rank (E*64: rank[e][a][b] = position
 * of node cells[e][b] in the sorted neighbour list of node cells[e][a]),
 * inc_ptr (N+1) and inc (8E, packed e*8 + a, ascending e for each node).
 * ---------------------------------------------------------------------- */
int fx_hex8_pattern(int32_t num_nodes, int32_t num_cells, const int32_t* cells /*(h) E*8*/,
                    int64_t* nse, int32_t* max_degree, int32_t* indptr /*(h) 3N+1*/,
                    int32_t* indices /*(h) nse or NULL*/, uint8_t* elem_rank /*(h) E*64 or NULL*/,
                    int32_t* inc_ptr /*(h) N+1 or NULL*/, int32_t* inc /*(h) 8E or NULL*/);

/* Node -> loaded-face incidence over a face list (F,2) of (cell, local face)
 * as built by feax/fe.py:267-318.  finc_ptr (N+1), finc (4F, packed f*8 + a
 * with a the cell-local node index, ascending f per node). */
int fx_hex8_face_incidence(int32_t num_nodes, int32_t num_cells, const int32_t* cells /*(h)*/,
                           int32_t num_faces, const int32_t* faces /*(h) F*2*/,
                           int32_t* finc_ptr /*(h) N+1*/, int32_t* finc /*(h) 4F*/);

/* Host check: rows 3m..3m+2 share one sorted column list of whole 3-blocks. */
int fx_csr_is_block3(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t* is_block3);

/* ------------------------------------------------------------------------
 * CSR kernels.  Replace CSRMatrix.__matmul__ (feax/csr.py:81-90) and
 * CSRMatrix.diagonal + the Jacobi preconditioner (feax/csr.py:115-121,
 * feax/solvers/common.py:86-94).  nse (= indptr[n]) only sizes the launch.
 * ---------------------------------------------------------------------- */
int fx_csr_spmv_f64(void* stream, int32_t n, int64_t nse, const int32_t* indptr, const int32_t* indices,
                    const double* data, const double* x, double* y);
int fx_csr_spmv_f32(void* stream, int32_t n, int64_t nse, const int32_t* indptr, const int32_t* indices,
                    const float* data, const float* x, float* y);

/* Block-3 column index of a block-3 pattern: bcol (nse/9 int32, (d)),
 * bcol[indptr[3m]/9 + k] = column node of the k-th 3x3 block of node m. */
int fx_csr_block3_index(void* stream, int32_t n, const int32_t* indptr, const int32_t* indices,
                        int32_t* bcol);
/* Same product as fx_csr_spmv_* reading one column index per 3x3 block. */
int fx_csr_spmv_b3_f64(void* stream, int32_t n, int64_t nse, const int32_t* indptr, const int32_t* bcol,
                       const double* data, const double* x, double* y);
int fx_csr_spmv_b3_f32(void* stream, int32_t n, int64_t nse, const int32_t* indptr, const int32_t* bcol,
                       const float* data, const float* x, float* y);
/* Difference-form operator (FX_ASM_DIFF3 values, FP32) on FP64 vectors:
 * y = sum_{m != n} K_nm (x_m - x_n) + S_n x_n, S_n stored in the diagonal
 * block slot.  Equals K x; FP32 rounding acts on differences only.  The
 * difference form is stored block by block: node m's values start at
 * indptr[3 m] and hold its 3x3 blocks in bcol order, nine values each,
 * row-major (the CSR layout's footprint, reordered within the node). */
int fx_csr_spmv_diff3_f32f64(void* stream, int32_t n, int64_t nse, const int32_t* indptr,
                             const int32_t* bcol, const float* data, const double* x, double* y);

/* dinv[i] = 1 / (A[i,i] + shift);  diag[i] = A[i,i]. */
int fx_csr_jacobi_f64(void* stream, int32_t n, const int32_t* indptr, const int32_t* indices,
                      const double* data, double shift, double* dinv);
int fx_csr_jacobi_f32(void* stream, int32_t n, const int32_t* indptr, const int32_t* indices,
                      const float* data, double shift, float* dinv);
int fx_csr_diagonal_f64(void* stream, int32_t n, const int32_t* indptr, const int32_t* indices,
                        const double* data, double* diag);
int fx_csr_diagonal_f32(void* stream, int32_t n, const int32_t* indptr, const int32_t* indices,
                        const float* data, float* diag);

/* ------------------------------------------------------------------------
 * Jacobi-preconditioned CG with jax.scipy.sparse.linalg.cg semantics
 * (called at feax/solvers/common.py:229-239): stop when
 * ||r||^2 <= max(tol^2 ||b||^2, atol^2) or k == maxiter; returns the last x.
 * dinv == NULL: no preconditioner.  x0 == NULL: x0 = 0.  Dot products are
 * reduced in FP64 in a fixed order (bitwise repeatable).  *iters = k,
 * *rs = ||r_k||^2 of the recursive residual.  flags select the SpMV:
 * 0 scalar CSR (indices), FX_PCG_BLOCK3 block-3 (bcol), FX_PCG_DIFF3
 * difference form (bcol, fx_pcg_f32f64 only).
 * ---------------------------------------------------------------------- */
#define FX_PCG_BLOCK3 1
#define FX_PCG_DIFF3 2
/* FX_PCG_TIME: time every SpMV launch of the executed iterations with HIP
 * events on `stream`; rs must then point to 3 doubles: rs[0] = ||r||^2,
 * rs[1] = mean SpMV kernel time in microseconds, rs[2] = launches timed.
 * Without FX_PCG_TIME the iterations run as a HIP graph of check_every
 * iterations captured once per call and replayed on `stream`;
 * FX_PCG_NO_GRAPH (16, below) launches them directly (identical results). */
#define FX_PCG_TIME 4
/* Optional recording (all three drivers): when rec_p != NULL, iteration k <
 * rec_max stores its search direction p_k as FP32 in rec_p + k*n (d) and
 * p_k . A p_k in rec_pq[k] (d), for fx_krylov_project_f32. */
size_t fx_pcg_workspace_bytes(int32_t n, int32_t vec_elem_bytes);
int fx_pcg_f64(void* stream, int32_t n, int64_t nse, int32_t flags, const int32_t* indptr,
               const int32_t* indices, const int32_t* bcol, const double* data, const double* dinv,
               const double* b, const double* x0, double tol, double atol, int32_t maxiter,
               int32_t check_every, double* x, int32_t* iters, double* rs, float* rec_p, double* rec_pq,
               int32_t rec_max, void* workspace, size_t ws_bytes);
/* fx_pcg_f64_noalloc: fx_pcg_f64 that allocates nothing, for callers that
 * cannot allocate device memory (code that must not call hipMalloc: serving
 * loops, harnesses that forbid allocation).  It has no counterpart in the
 * reference.  The PCG state, partials and the r, z, p, q vectors live in a
 * 40 MiB static device array per GPU, so n <= 1,231,168 (FX_ELIMIT
 * beyond).  No HIP graph, stream or event is created: the check_every
 * chunks are launched directly on `stream` (results equal fx_pcg_f64's);
 * FX_PCG_TIME and recording are not available.  Calls on one GPU share
 * the array: calls that overlap on one GPU (two streams, two host threads)
 * race and must be serialised by the caller.  Every value the solve reads
 * from the array is written earlier in the same call. */
int fx_pcg_f64_noalloc(void* stream, int32_t n, int64_t nse, int32_t flags, const int32_t* indptr,
                       const int32_t* indices, const int32_t* bcol, const double* data, const double* dinv,
                       const double* b, const double* x0, double tol, double atol, int32_t maxiter,
                       int32_t check_every, double* x, int32_t* iters, double* rs);
/* FP32 matrix and FP32 vectors (FP64 dot products). */
int fx_pcg_f32(void* stream, int32_t n, int64_t nse, int32_t flags, const int32_t* indptr,
               const int32_t* indices, const int32_t* bcol, const float* data, const float* dinv,
               const float* b, const float* x0, double tol, double atol, int32_t maxiter,
               int32_t check_every, float* x, int32_t* iters, double* rs, float* rec_p, double* rec_pq,
               int32_t rec_max, void* workspace, size_t ws_bytes);
/* FP32 matrix (plain or difference form), FP64 vectors. */
int fx_pcg_f32f64(void* stream, int32_t n, int64_t nse, int32_t flags, const int32_t* indptr,
                  const int32_t* indices, const int32_t* bcol, const float* data, const double* dinv,
                  const double* b, const double* x0, double tol, double atol, int32_t maxiter,
                  int32_t check_every, double* x, int32_t* iters, double* rs, float* rec_p,
                  double* rec_pq, int32_t rec_max, void* workspace, size_t ws_bytes);

/* Block-column map of a block-3 CSR for the SpMV of the PCG: the SpMV runs
 * nodes_per_block consecutive nodes per workgroup (256 or 512 threads
 * divided by the lane count of fx_csr_block3_lanes: the map's size picks
 * the workgroup); the workgroup first copies x of the sorted unique
 * column nodes of its nodes (ucol[uptr[b] .. uptr[b+1])) into LDS, then each
 * 3x3 block reads x from LDS at lidx (its column node's position in that
 * list; one entry per block, in bcol order) instead of gathering it through
 * bcol.  Same products in the same order: results equal the bcol path's
 * bitwise.  All pointers (d); max_unique = max_b (uptr[b+1] - uptr[b]). */
typedef struct fx_b3map {
  const uint16_t* lidx;
  const int32_t* uptr;
  const int32_t* ucol;
  int32_t nodes_per_block;
  int32_t max_unique;
} fx_b3map;
/* Lanes per node the block-3 SpMV kernels use for this matrix
 * (value_bytes 8: FP64 values, 4: FP32). */
int fx_csr_block3_lanes(int32_t n, int64_t nse, int32_t value_bytes);
/* fx_pcg_f64 / fx_pcg_f32f64 on a block-3 matrix (flags FX_PCG_BLOCK3 or
 * FX_PCG_DIFF3) with its block-column map in place of bcol.  A map whose
 * nodes_per_block x lanes is not 256 or 512 is refused (FX_EINVAL). */
int fx_pcg_b3map_f64(void* stream, int32_t n, int64_t nse, int32_t flags, const int32_t* indptr,
                     const fx_b3map* map, const double* data, const double* dinv, const double* b,
                     const double* x0, double tol, double atol, int32_t maxiter, int32_t check_every, double* x,
                     int32_t* iters, double* rs, float* rec_p, double* rec_pq, int32_t rec_max, void* workspace,
                     size_t ws_bytes);
int fx_pcg_b3map_f32f64(void* stream, int32_t n, int64_t nse, int32_t flags, const int32_t* indptr,
                        const fx_b3map* map, const float* data, const double* dinv, const double* b,
                        const double* x0, double tol, double atol, int32_t maxiter, int32_t check_every, double* x,
                        int32_t* iters, double* rs, float* rec_p, double* rec_pq, int32_t rec_max,
                        void* workspace, size_t ws_bytes);

/* Upper-block storage of a bitwise-symmetric FP32 difference-form matrix
 * (the mixed form's inner operator; no counterpart in feax, whose CSR is
 * feax/csr.py:81-90: an internal layout of the same operator).  Node n's
 * blocks with column node c >= n (its S block, then its upper neighbours in
 * column order) are stored contiguously from block up_ptr[n], nine FP32
 * values each (row-major); the lower block (n, c), c < n, is the transpose
 * of the stored block (c, n).  code has 32 entries per node: code[32 n + kb]
 * describes node n's block kb of the CSR pattern (bcol order) as
 * (l << 6) | (lower << 5) | r, with l the position of the column node in
 * the workgroup's sorted unique column list ucol[uptr[b] .. uptr[b+1]) and
 * r the position of the stored block inside its owner's upper list
 * (owner = c for a lower block, else n); entries past the node's degree
 * hold 0xFFFF.  usup[u] = up_ptr[ucol[u]] (same length as ucol).  The
 * kernels run nodes_per_block = 64 nodes per workgroup for FP32 values and
 * 32 for FP64 values, node degree <= 32, max_unique <= 1023 (FX_EINVAL /
 * FX_ELIMIT otherwise).  up_ptr has n/3 + 1 entries; num_upper =
 * up_ptr[n/3].  All pointers (d). */
typedef struct fx_b3sym {
  const uint16_t* code;
  const int32_t* uptr;
  const int32_t* ucol;
  const int32_t* usup;
  const int32_t* up_ptr;
  int32_t nodes_per_block;
  int32_t max_unique;
  int32_t max_degree;
  int32_t num_upper;
} fx_b3sym;
/* Upper blocks of a full block-3 matrix into upper (9 num_upper values),
 * with a bitwise symmetry check: bad[0] (d) = lower blocks differing from
 * the transpose of their mirror, bad[1] = blocks whose code disagrees with
 * the pattern (both zero: the storage is exact).  bad is zeroed on the
 * stream first.  _f32: a difference-form matrix (FX_ASM_DIFF3 values, block
 * layout); _f64: feax's CSR values. */
int fx_b3sym_compact_f32(void* stream, int32_t n, const int32_t* indptr, const int32_t* bcol, const fx_b3sym* sym,
                         const float* data, float* upper, int32_t* bad);
int fx_b3sym_compact_f64(void* stream, int32_t n, const int32_t* indptr, const int32_t* bcol, const fx_b3sym* sym,
                         const double* data, double* upper, int32_t* bad);
/* y = A x on upper-block storage: the operator of fx_csr_spmv_diff3_f32f64
 * (_f32f64) or of fx_csr_spmv_b3_f64 (_f64). */
int fx_csr_spmv_b3sym_f32f64(void* stream, int32_t n, const int32_t* indptr, const fx_b3sym* sym, const float* upper,
                             const double* x, double* y);
int fx_csr_spmv_b3sym_f64(void* stream, int32_t n, const int32_t* indptr, const fx_b3sym* sym, const double* upper,
                          const double* x, double* y);
/* fx_pcg_f32f64 (flags must hold FX_PCG_DIFF3) / fx_pcg_f64 (flags must hold
 * FX_PCG_BLOCK3) on upper-block storage. */
int fx_pcg_b3sym_f32f64(void* stream, int32_t n, int64_t nse, int32_t flags, const int32_t* indptr,
                        const fx_b3sym* sym, const float* upper, const double* dinv, const double* b,
                        const double* x0, double tol, double atol, int32_t maxiter, int32_t check_every, double* x,
                        int32_t* iters, double* rs, float* rec_p, double* rec_pq, int32_t rec_max, void* workspace,
                        size_t ws_bytes);
int fx_pcg_b3sym_f64(void* stream, int32_t n, int64_t nse, int32_t flags, const int32_t* indptr, const fx_b3sym* sym,
                     const double* upper, const double* dinv, const double* b, const double* x0, double tol,
                     double atol, int32_t maxiter, int32_t check_every, double* x, int32_t* iters, double* rs,
                     float* rec_p, double* rec_pq, int32_t rec_max, void* workspace, size_t ws_bytes);

/* Krylov recycling: Galerkin projection of r onto nvec recorded A-conjugate
 * directions, x0 = sum_k p_k (p_k . r) / (p_k . A p_k) (init-CG, Erhel and
 * Guyomarc'h 2000).  P: (d) nvec*n FP32, pq: (d) nvec, r, x0: (d) n FP64.
 * Fixed-order FP64 reductions. */
size_t fx_krylov_project_workspace_bytes(int32_t nvec);
int fx_krylov_project_f32(void* stream, int32_t n, int32_t nvec, const float* P, const double* pq,
                          const double* r, double* x0, void* workspace, size_t ws_bytes);

/* ------------------------------------------------------------------------
 * Vector helpers.
 * ---------------------------------------------------------------------- */
/* out(d) = sum_i a_i b_i in FP64, fixed order; workspace >= 8*2048 bytes. */
int fx_dot_f64(void* stream, int32_t n, const double* a, const double* b, double* out, void* workspace,
               size_t ws_bytes);
/* y64 += scale * (double) x32 */
int fx_axpy_f32_into_f64(void* stream, int32_t n, double scale, const float* x, double* y);
/* y32 = (float) (scale * x64) */
int fx_convert_f64_to_f32(void* stream, int32_t n, double scale, const double* x, float* y);
/* out = a x + b y (FP64); out may alias x or y.  u = u0 + du of
 * feax/solvers/linear.py:417. */
int fx_axpby_f64(void* stream, int32_t n, double a, const double* x, double b, const double* y,
                 double* out);
/* x0 = 0 except x0[bc_rows] = bc_vals - u0[bc_rows] (u0 may be NULL):
 * create_x0 of feax/solvers/common.py:23-53 (u0 NULL: zero_like_initial_guess,
 * feax/utils.py:464-499). */
int fx_bc_lift_f64(void* stream, int32_t n, int32_t nbc, const int32_t* bc_rows, const double* bc_vals,
                   const double* u0, double* x0);

/* ------------------------------------------------------------------------
 * HEX8 element kernels.  One 64-lane wavefront per mesh node gathers the
 * contributions of the node's elements in ascending element order (the
 * summation order of feax/assembler.py:1253-1318): no float atomics,
 * bitwise repeatable.
 *
 * Assembly and residual take an optional workspace.  With one
 * (fx_hex8_assemble_workspace_bytes / fx_hex8_residual_workspace_bytes
 * bytes) they run in two passes: one wave per element computes the element
 * matrix (or residual) once into the workspace, then one wave (thread) per
 * node gathers it in the same order.  workspace == NULL computes every
 * element once per node instead (no scratch memory).  Both ways do the
 * same arithmetic in the same order (equal up to the rounding of a fused
 * multiply-add the compiler may form in one and not the other). 
 *
 * x, y, z: (d) node coordinates (SoA, FP64).  cells: (d) E*8 int32.
 * param: (d) N or E values (mat->param_loc).  u: (d) 3N FP64 or NULL.
 * bc_mask: (d) 3N uint8 or NULL.
 * ---------------------------------------------------------------------- */

/* CSR values of the Jacobian, BC-eliminated when bc_mask != NULL (symmetric:
 * zero BC rows and columns, BC diagonal 1; else rows only;
 * feax/DCboundary.py:147-199).  Replaces _get_J_csr + apply_boundary_to_J_csr
 * (feax/assembler.py:1362-1404, 1695-1713).  flags FX_ASM_DIFF3: store the
 * difference form of fx_csr_spmv_diff3_f32f64.  dinv (d, 3N, nullable):
 * 1 / (K_bc[r, r] + jacobi_shift). */
#define FX_ASM_DIFF3 1
/* fx_hex8_assemble_f32: element matrices are formed in FP64 and rounded to
 * FP32 once at the store; FX_ASM_COMPUTE32 forms them in FP32 arithmetic. */
#define FX_ASM_COMPUTE32 2
/* FX_ASM_UPPER (fx_hex8_assemble_f32 with FX_ASM_DIFF3, a linear law,
 * symmetric elimination and the workspace): store the difference form as
 * the upper-block storage of fx_b3sym (data: 9 num_upper FP32, node n's
 * blocks with column >= n from block up_ptr[n]); bitwise the array
 * fx_b3sym_compact_f32 makes from the full difference form.
 * node_order (d, N, nullable): the order in which the per-node gather runs
 * the nodes (a space-filling-curve order of the mesh; the result does not
 * depend on it).  up_ptr (d, N + 1): with FX_ASM_UPPER only. */
#define FX_ASM_UPPER 4
int fx_hex8_assemble_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x,
                         const double* y, const double* z, const int32_t* cells, const int32_t* inc_ptr,
                         const int32_t* inc, const uint8_t* elem_rank, const int32_t* indptr,
                         const int32_t* indices, const fx_material* mat, const double* param,
                         const double* u, const uint8_t* bc_mask, int32_t symmetric, int32_t flags,
                         double jacobi_shift, double* dinv, double* data, void* workspace, size_t ws_bytes,
                         const int32_t* node_order, const int32_t* up_ptr);
int fx_hex8_assemble_f32(void* stream, int32_t num_nodes, int32_t num_cells, const double* x,
                         const double* y, const double* z, const int32_t* cells, const int32_t* inc_ptr,
                         const int32_t* inc, const uint8_t* elem_rank, const int32_t* indptr,
                         const int32_t* indices, const fx_material* mat, const double* param,
                         const double* u, const uint8_t* bc_mask, int32_t symmetric, int32_t flags,
                         double jacobi_shift, double* dinv, float* data, void* workspace, size_t ws_bytes,
                         const int32_t* node_order, const int32_t* up_ptr);
/* 576 element-matrix values per element (the linear laws use 324 of them:
 * their element matrices are exactly symmetric and only the upper 3x3
 * block triangle is stored): value_bytes 8 (f64) or 4 (f32; FP32 workspace
 * only with FX_ASM_COMPUTE32). */
size_t fx_hex8_assemble_workspace_bytes(int32_t num_cells, int32_t value_bytes, int32_t flags);

/* Global residual R(u) = volume + traction (FP64): feax get_res
 * (feax/assembler.py:1618-1667).  Surface part: the loaded faces of all
 * boundaries concatenated in location_fns order, faces (F,2) (cell, local
 * face), node->face incidence finc_ptr/finc over that list, dead traction
 * t(f, q) = fdir[f] * svar[f, q] (svar (d) F*4, fdir (d) F*3): the surface
 * map `lambda u, x, t: dir * t` of the benchmark.  num_faces = 0 for none.
 * If bc_rows != NULL, applies res[bc] = u[bc] - bc_vals
 * (feax/DCboundary.py:202-251); negate != 0 returns -R_bc, the right-hand
 * side b of feax/solvers/linear.py:400. */
int fx_hex8_residual_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x,
                         const double* y, const double* z, const int32_t* cells, const int32_t* inc_ptr,
                         const int32_t* inc, const fx_material* mat, const double* param, const double* u,
                         int32_t num_faces, const int32_t* faces, const int32_t* finc_ptr,
                         const int32_t* finc, const double* svar, const double* fdir, int32_t num_bc,
                         const int32_t* bc_rows, const double* bc_vals, int32_t negate, double* res,
                         void* workspace, size_t ws_bytes);
size_t fx_hex8_residual_workspace_bytes(int32_t num_cells);

/* ------------------------------------------------------------------ */
/* TET4 (feax/basis.py:78-83: basix P1 tetrahedron, one Gauss point; facet i
 * opposite vertex i, one point): the structure builders and the two-pass
 * kernels of the HEX8 entry points above, with cells (E, 4), elem_rank
 * E * 16 (rank of node b in node a's neighbour list at 16 e + 4 a + b),
 * node incidence entries e * 8 + a, faces (F, 2) (cell, local facet) and
 * one surface value per face (svar F, fdir F * 3).  The workspace is
 * required (fx_tet4_*_workspace_bytes: 144 values per element for the
 * matrix, 12 doubles for the residual).  Same materials, flags and
 * Dirichlet handling as fx_hex8_assemble_* / fx_hex8_residual_f64
 * (feax/assembler.py:1362-1404, 1618-1667; feax/DCboundary.py:147-251). */
int fx_tet4_pattern(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int64_t* nse,
                    int32_t* max_degree, int32_t* indptr, int32_t* indices, uint8_t* elem_rank,
                    int32_t* inc_ptr, int32_t* inc);
int fx_tet4_face_incidence(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int32_t num_faces,
                           const int32_t* faces, int32_t* finc_ptr, int32_t* finc);
int fx_tet4_assemble_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                         const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                         const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                         const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                         int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, double* data,
                         void* workspace, size_t ws_bytes);
int fx_tet4_assemble_f32(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                         const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                         const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                         const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                         int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, float* data,
                         void* workspace, size_t ws_bytes);
size_t fx_tet4_assemble_workspace_bytes(int32_t num_cells, int32_t value_bytes, int32_t flags);
int fx_tet4_residual_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                         const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                         const fx_material* mat, const double* param, const double* u, int32_t num_faces,
                         const int32_t* faces, const int32_t* finc_ptr, const int32_t* finc, const double* svar,
                         const double* fdir, int32_t num_bc, const int32_t* bc_rows, const double* bc_vals,
                         int32_t negate, double* res, void* workspace, size_t ws_bytes);
size_t fx_tet4_residual_workspace_bytes(int32_t num_cells);
/* TET4 pullbacks, as fx_hex8_param_vjp_f64 / fx_simp_compliance_sens_f64 /
 * fx_hex8_traction_vjp_f64 at the single Gauss point (node variables need
 * fx_tet4_param_vjp_workspace_bytes: one double per element); the traction
 * gradient has one value per face. */
size_t fx_tet4_param_vjp_workspace_bytes(int32_t num_cells, int32_t param_loc);
int fx_tet4_param_vjp_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const fx_material* mat, const double* param, const double* u, const double* lam,
                          const uint8_t* bc_mask, double scale, double* grad, void* workspace, size_t ws_bytes);
int fx_tet4_simp_compliance_sens_f64(void* stream, int32_t num_cells, const double* x, const double* y,
                                     const double* z, const int32_t* cells, const fx_material* mat,
                                     const double* param, const double* u, double* dc);
int fx_tet4_traction_vjp_f64(void* stream, int32_t num_faces, const double* x, const double* y, const double* z,
                             const int32_t* cells, const int32_t* faces, const double* fdir, const double* lam,
                             const uint8_t* bc_mask, double scale, double* grad_svar);
/* ------------------------------------------------------------------ */
/* HEX20 (feax/basis.py:71-77: basix serendipity degree 2 -- DOFs the corner
 * values and the edge means --, meshio node order: 8 corners then the edge
 * nodes of (0,1) (1,2) (2,3) (3,0) (4,5) (5,6) (6,7) (7,4) (0,4) (1,5)
 * (2,6) (3,7); 2x2x2 Gauss points, 2x2 per facet; faces selected by their 4
 * corners): the structure builders and two-pass kernels of the TET4 entry
 * points, with cells (E, 20), elem_rank E * 400 (rank of node b in node
 * a's neighbour list at 400 e + 20 a + b), node incidence entries
 * (e << 5) | a, face incidence entries (f << 5) | a over each facet's 8
 * nodes, and 4 surface values per face (svar F * 4, fdir F * 3).  The
 * workspace is required (3600 doubles per element for the matrix, 60 for
 * the residual); FP64 element arithmetic (FX_ASM_DIFF3 honoured,
 * FX_ASM_COMPUTE32 ignored); nodes may have up to 96 neighbours.  Materials:
 * linear elastic, SIMP, Neo-Hookean; no thermal strain. */
int fx_hex20_pattern(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int64_t* nse,
                     int32_t* max_degree, int32_t* indptr, int32_t* indices, uint8_t* elem_rank,
                     int32_t* inc_ptr, int32_t* inc);
int fx_hex20_face_incidence(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int32_t num_faces,
                            const int32_t* faces, int32_t* finc_ptr, int32_t* finc);
int fx_hex20_assemble_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                          const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                          int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, double* data,
                          void* workspace, size_t ws_bytes);
int fx_hex20_assemble_f32(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                          const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                          int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, float* data,
                          void* workspace, size_t ws_bytes);
size_t fx_hex20_assemble_workspace_bytes(int32_t num_cells, int32_t value_bytes, int32_t flags);
int fx_hex20_residual_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const fx_material* mat, const double* param, const double* u, int32_t num_faces,
                          const int32_t* faces, const int32_t* finc_ptr, const int32_t* finc, const double* svar,
                          const double* fdir, int32_t num_bc, const int32_t* bc_rows, const double* bc_vals,
                          int32_t negate, double* res, void* workspace, size_t ws_bytes);
size_t fx_hex20_residual_workspace_bytes(int32_t num_cells);
/* TET4 pullbacks, as fx_hex8_param_vjp_f64 / fx_simp_compliance_sens_f64 /
 * fx_hex8_traction_vjp_f64 at the single Gauss point (node variables need
 * fx_hex20_param_vjp_workspace_bytes: one double per element); the traction
 * gradient has one value per face. */
size_t fx_hex20_param_vjp_workspace_bytes(int32_t num_cells, int32_t param_loc);
int fx_hex20_param_vjp_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                           const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                           const fx_material* mat, const double* param, const double* u, const double* lam,
                           const uint8_t* bc_mask, double scale, double* grad, void* workspace, size_t ws_bytes);
int fx_hex20_simp_compliance_sens_f64(void* stream, int32_t num_cells, const double* x, const double* y,
                                      const double* z, const int32_t* cells, const fx_material* mat,
                                      const double* param, const double* u, double* dc);
int fx_hex20_traction_vjp_f64(void* stream, int32_t num_faces, const double* x, const double* y, const double* z,
                              const int32_t* cells, const int32_t* faces, const double* fdir, const double* lam,
                              const uint8_t* bc_mask, double scale, double* grad_svar);
/* ------------------------------------------------------------------ */
/* A monolithic two-variable HEX8 problem (feax's Problem(mesh=[m, m],
 * vec=[3, 1]), feax/problem.py:155-212, feax/assembler.py:591-647):
 * displacement u at DOF 3 n + i and temperature T at DOF 3 N + n, weak form
 * sigma = E (lam0 tr eps I + 2 mu0 eps) - E alpha (T - t_ref) / (1 - 2 nu) I
 * (variable 0) and k grad T (variable 1); E the volume variable (kind
 * FX_MAT_LINEAR_ELASTIC or FX_MAT_SIMP, node or cell), nu, alpha, t_ref from
 * fx_material and the conductivity k in fx_material.kappa (temperature
 * NULL: T is part of the state).  The pattern holds in every row of node n
 * the u columns of its neighbours, then their T columns (4 deg entries;
 * indptr has 4N + 1 entries); elem_rank and the node incidence are the
 * HEX8 ones, faces as HEX8.  Assembly: FP64 only, flags 0, the Jacobian
 * does not depend on u (u may be NULL), workspace 1024 doubles per
 * element.  Residual: per face, fdir[4 f + i] (i < 3) scales the surface
 * value on u and fdir[4 f + 3] on T (4 surface values per face, svar F * 4),
 * workspace 32 doubles per element. */
int fx_hex8ut_pattern(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int64_t* nse,
                      int32_t* max_degree, int32_t* indptr, int32_t* indices, uint8_t* elem_rank,
                      int32_t* inc_ptr, int32_t* inc);
int fx_hex8ut_face_incidence(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int32_t num_faces,
                             const int32_t* faces, int32_t* finc_ptr, int32_t* finc);
size_t fx_hex8ut_assemble_workspace_bytes(int32_t num_cells, int32_t value_bytes, int32_t flags);
int fx_hex8ut_assemble_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                           const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                           const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                           const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                           int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, double* data,
                           void* workspace, size_t ws_bytes);
size_t fx_hex8ut_residual_workspace_bytes(int32_t num_cells);
int fx_hex8ut_residual_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                           const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                           const fx_material* mat, const double* param, const double* u, int32_t num_faces,
                           const int32_t* faces, const int32_t* finc_ptr, const int32_t* finc, const double* svar,
                           const double* fdir, int32_t num_bc, const int32_t* bc_rows, const double* bc_vals,
                           int32_t negate, double* res, void* workspace, size_t ws_bytes);

/* TET4 Neo-Hookean strain energy, as fx_hex8_energy_f64 at the single
 * Gauss point (workspace fx_tet4_energy_workspace_bytes). */
size_t fx_tet4_energy_workspace_bytes(int32_t num_cells);
int fx_tet4_energy_f64(void* stream, int32_t num_cells, const double* x, const double* y, const double* z,
                       const int32_t* cells, const fx_material* mat, const double* param, const double* u,
                       double* energy, double* dpsi_dmu, void* workspace, size_t ws_bytes);

/* ------------------------------------------------------------------ */
/* QUAD4, dim = 2, vec = 2 (feax/basis.py:90-95: basix P1 quadrilateral
 * through re-order [0, 1, 3, 2], 2x2 Gauss points; edges bottom (0,1),
 * left (0,3), right (1,2), top (3,2), 2 points each): the structure
 * builders (rows 2 n + i; incidence e * 8 + a; elem_rank 16 per element)
 * and the two-pass kernels with the signatures of the HEX8 / TET4 entry
 * points (z is ignored and may be NULL; svar holds 2 values per edge,
 * fdir 2 per edge).  The laws are the plane-strain restrictions of the
 * 3-D ones (du_z = 0).  FP64 element arithmetic only (no FX_ASM_DIFF3 /
 * FX_ASM_COMPUTE32); the workspace is required (64 values per element for
 * the matrix, 8 doubles for the residual, 4 per element for node-variable
 * pullbacks). */
int fx_quad4_pattern(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int64_t* nse,
                     int32_t* max_degree, int32_t* indptr, int32_t* indices, uint8_t* elem_rank,
                     int32_t* inc_ptr, int32_t* inc);
int fx_quad4_face_incidence(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int32_t num_faces,
                            const int32_t* faces, int32_t* finc_ptr, int32_t* finc);
int fx_quad4_assemble_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                          const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                          int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, double* data,
                          void* workspace, size_t ws_bytes);
int fx_quad4_assemble_f32(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                          const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                          int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, float* data,
                          void* workspace, size_t ws_bytes);
size_t fx_quad4_assemble_workspace_bytes(int32_t num_cells, int32_t value_bytes, int32_t flags);
int fx_quad4_residual_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const fx_material* mat, const double* param, const double* u, int32_t num_faces,
                          const int32_t* faces, const int32_t* finc_ptr, const int32_t* finc, const double* svar,
                          const double* fdir, int32_t num_bc, const int32_t* bc_rows, const double* bc_vals,
                          int32_t negate, double* res, void* workspace, size_t ws_bytes);
size_t fx_quad4_residual_workspace_bytes(int32_t num_cells);
size_t fx_quad4_param_vjp_workspace_bytes(int32_t num_cells, int32_t param_loc);
int fx_quad4_param_vjp_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                           const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                           const fx_material* mat, const double* param, const double* u, const double* lam,
                           const uint8_t* bc_mask, double scale, double* grad, void* workspace, size_t ws_bytes);
int fx_quad4_simp_compliance_sens_f64(void* stream, int32_t num_cells, const double* x, const double* y,
                                      const double* z, const int32_t* cells, const fx_material* mat,
                                      const double* param, const double* u, double* dc);
int fx_quad4_traction_vjp_f64(void* stream, int32_t num_faces, const double* x, const double* y, const double* z,
                              const int32_t* cells, const int32_t* faces, const double* fdir, const double* lam,
                              const uint8_t* bc_mask, double scale, double* grad_svar);

/* Matrix-free operator out = J_bc(u) v (FP64): feax's Krylov operator
 * create_matfree_res_J_parametric (feax/assembler.py:1747-1778),
 * J_bc v = where(bc, v, JVP_R(u)[v * !bc]) under symmetric elimination: rows
 * with bc_mask set return v, the others the linearised volume residual of v
 * with the BC entries of v read as 0.  u is the linearisation point (needed
 * for the Neo-Hookean law, ignored and may be NULL for linear laws).  out
 * must not alias v. */
int fx_hex8_jvp_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                    const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                    const fx_material* mat, const double* param, const double* u, const double* v,
                    const uint8_t* bc_mask, double* out);

/* Jacobi-PCG (jax cg semantics, as fx_pcg_f64) on the matrix-free operator
 * J_bc(u) of fx_hex8_jvp_f64 instead of an assembled CSR: feax's default
 * Krylov path (feax/solvers/linear.py:394-398 with the operator of
 * feax/assembler.py:1747-1778).  dinv may be NULL (plain CG, which is what
 * feax runs there: its Jacobi flag is ignored for a callable operator,
 * feax/solvers/common.py:221-227).  flags: 0 or FX_PCG_TIME (rs as in
 * fx_pcg_f64).  Workspace: fx_pcg_workspace_bytes(3 * num_nodes, 8). */
int fx_pcg_matfree_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                       const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                       const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                       int32_t flags, const double* dinv, const double* b, const double* x0, double tol, double atol,
                       int32_t maxiter, int32_t check_every, double* xout, int32_t* iters, double* rs,
                       void* workspace, size_t ws_bytes);

/* ------------------------------------------------------------------------
 * Batched instances.  feax batches solves with jax.vmap over the parameters
 * of one mesh; its batched while_loop iterates until every lane has
 * converged (tests/test_vmap.py:63-106, feax/solvers/common.py:229-239 per
 * lane).  Here the instances (different meshes allowed) are concatenated
 * into one block-diagonal system: instance s owns nodes
 * seg_node_ptr[s] .. seg_node_ptr[s+1] (HOST array, nseg+1 entries, first 0),
 * and the CSR (block-3 form, bcol from fx_csr_block3_index) is the
 * concatenation of the instances' CSR.  Every CG scalar, the stopping test
 * (jax cg semantics, tol/atol/maxiter shared) and the convergence flag are
 * per instance; iters and rs are HOST arrays of nseg entries.  An
 * iteration is three launches (SpMV; alpha and the x, r, z updates; beta
 * and the direction), each block summing its own instance's partials of
 * the previous launch.  A chunk of check_every iterations (rounded up to
 * even) is captured once as a HIP graph and replayed (FX_PCG_NO_GRAPH:
 * launch kernels directly; results are identical).  The SpMV runs L lanes
 * per node in workgroups of 256 or 512 threads, each over consecutive nodes
 * of one instance, reading x through bcol or through a block-column map
 * that the call builds on the device for that partition (x of the
 * workgroup's unique column nodes staged in LDS).  Default: FP64 values 32
 * lanes / 256 threads / bcol, FP32 values 16 / 256 / map.
 * FX_PCG_AUTOTUNE: time lanes 8/16/32 x 256/512 threads x map off/on on
 * this instance set (cached per shape); tuned (nullable, HOST, 3 entries)
 * receives lanes, threads per workgroup and map (0/1) of the run.
 * FX_PCG_SEG_FORCE: run the configuration given in tuned[0..2] instead
 * (lanes 8/16/32, threads 256/512, map 0/1; a map that cannot be used --
 * a workgroup's column nodes spanning more than 65536 node numbers, or
 * its LDS copy above 64 KiB -- falls back to bcol, and tuned[2] says so).
 * fx_pcg_batched_f32f64: FP32 matrix (FX_PCG_DIFF3: difference form), FP64
 * vectors.  Workspace: fx_pcg_batched_workspace_bytes(nseg, seg_node_ptr,
 * nse, 8) (map storage included: 6 bytes per 3x3 block).
 * ------------------------------------------------------------------------ */
#define FX_PCG_AUTOTUNE 8
#define FX_PCG_NO_GRAPH 16
#define FX_PCG_SEG_FORCE 64
size_t fx_pcg_batched_workspace_bytes(int32_t nseg, const int32_t* seg_node_ptr, int64_t nse,
                                      int32_t vec_elem_bytes);
int fx_pcg_batched_f64(void* stream, int32_t nseg, const int32_t* seg_node_ptr, int64_t nse, int32_t flags,
                       const int32_t* indptr, const int32_t* bcol, const double* data, const double* dinv,
                       const double* b, const double* x0, double tol, double atol, int32_t maxiter,
                       int32_t check_every, double* x, int32_t* iters, double* rs, int32_t* tuned,
                       void* workspace, size_t ws_bytes);
int fx_pcg_batched_f32f64(void* stream, int32_t nseg, const int32_t* seg_node_ptr, int64_t nse, int32_t flags,
                          const int32_t* indptr, const int32_t* bcol, const float* data, const double* dinv,
                          const double* b, const double* x0, double tol, double atol, int32_t maxiter,
                          int32_t check_every, double* x, int32_t* iters, double* rs, int32_t* tuned,
                          void* workspace, size_t ws_bytes);
/* fx_pcg_batched_record_f32f64: fx_pcg_batched_f32f64 that also records
 * each instance's search directions at its own iteration count k <
 * rec_max: rec_p[k n + i] = p_k[i] (FP32, n = 3 seg_node_ptr[nseg]) for the
 * instance's rows, rec_pq[k nseg + s] = p_k . A p_k (device arrays of
 * rec_max x n and rec_max x nseg).  Instance s holds min(iters[s], rec_max)
 * directions afterwards.  No reference counterpart (feax has no mixed
 * precision): Krylov recycling for the later inner solves of the batched
 * defect correction. */
int fx_pcg_batched_record_f32f64(void* stream, int32_t nseg, const int32_t* seg_node_ptr, int64_t nse,
                                 int32_t flags, const int32_t* indptr, const int32_t* bcol, const float* data,
                                 const double* dinv, const double* b, const double* x0, double tol, double atol,
                                 int32_t maxiter, int32_t check_every, double* x, int32_t* iters, double* rs,
                                 int32_t* tuned, float* rec_p, double* rec_pq, int32_t rec_max,
                                 void* workspace, size_t ws_bytes);
/* x0 = per-instance Galerkin projection of r onto the recorded directions:
 * x0[i] = sum_{k < counts[s]} (p_k . r)_s / rec_pq[k nseg + s] p_k[i] for
 * the rows i of instance s (counts: HOST array, each in [0, kmax]; P, pq as
 * recorded, with at least kmax directions).  Workspace:
 * fx_seg_project_workspace_bytes(nseg, seg_node_ptr, kmax).  Synchronises
 * the stream before returning. */
size_t fx_seg_project_workspace_bytes(int32_t nseg, const int32_t* seg_node_ptr, int32_t kmax);
int fx_seg_project_f32(void* stream, int32_t nseg, const int32_t* seg_node_ptr, int32_t kmax, const int32_t* counts,
                       const float* P, const double* pq, const double* r, double* x0, void* workspace,
                       size_t ws_bytes);
/* out[s] = sum over the rows row_ptr[s] .. row_ptr[s+1] of a_i b_i
 * (row_ptr: DEVICE array, nseg+1 entries; out: device, nseg). */
int fx_seg_dot_f64(void* stream, int32_t nseg, const int32_t* row_ptr, const double* a, const double* b,
                   double* out);

/* ------------------------------------------------------------------------
 * Adjoint.  feax's backward pass of the linear solve (jax.custom_vjp,
 * feax/solvers/linear.py:426-481) is lambda = J_bc^-T v followed by
 * grads = -(dR_bc/dparams)^T lambda, the jax.vjp of the residual
 * (feax/assembler.py:1833-1858).  These are that pullback, hand-written.
 * ------------------------------------------------------------------------ */

/* grad = scale * (dR/dp)^T lambda for the volume variable p = param of mat
 * (cell (E,) or node (N,)).  lambda entries on rows with bc_mask set are
 * read as 0 (rows of R_bc there do not depend on p, feax/DCboundary.py:
 * 202-251).  Node variables need inc_ptr/inc and a workspace of
 * fx_hex8_param_vjp_workspace_bytes(E, FX_PARAM_NODE) bytes; cell variables
 * need neither.  With scale = -1 and lambda = J_bc^-1 v this is the
 * parameter gradient of feax's f_bwd. */
size_t fx_hex8_param_vjp_workspace_bytes(int32_t num_cells, int32_t param_loc);
int fx_hex8_param_vjp_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x,
                          const double* y, const double* z, const int32_t* cells, const int32_t* inc_ptr,
                          const int32_t* inc, const fx_material* mat, const double* param, const double* u,
                          const double* lam, const uint8_t* bc_mask, double scale, double* grad,
                          void* workspace, size_t ws_bytes);

/* Fused compliance sensitivity for a linear law with a cell variable:
 * dc[e] = -u_e^T (dK_e/dp_e) u_e, i.e. -p rho^(p-1) (E0 - Emin) u_e^T K0_e u_e
 * for SIMP (feax/solvers/cmg.py:312-323).  This is the self-adjoint case
 * (lambda = u) of the gradient of C = f.u (feax/gene/responses.py:10-105)
 * when the Dirichlet values are zero and the load does not depend on p. */
/* d(lambda . R)/dT for the temperature node variable of a thermo-elastic
 * material (mat->temperature set; linear / SIMP law): grad[b] = scale *
 * sum over the elements and Gauss points at node b of
 * -N_b E (3 lam0 + 2 mu0) alpha div(lambda) JxW (lambda read as 0 on
 * bc_mask rows).  Workspace: fx_hex8_param_vjp_workspace_bytes(E,
 * FX_PARAM_NODE).  Replaces the temperature cotangent jax.vjp forms through
 * the thermo-elastic tensor map (examples/basic/thermo_mech_coupling.py:
 * 95-103, feax/solvers/linear.py:466-483). */
int fx_hex8_thermal_vjp_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                            const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                            const fx_material* mat, const double* param, const double* lam, const uint8_t* bc_mask,
                            double scale, double* grad, void* workspace, size_t ws_bytes);
int fx_simp_compliance_sens_f64(void* stream, int32_t num_cells, const double* x, const double* y,
                                const double* z, const int32_t* cells, const fx_material* mat,
                                const double* param, const double* u, double* dc);

/* grad_svar[f*4+q] = scale * d(lambda . R)/d svar[f, q] for the dead traction
 * t = fdir[f] * svar[f, q] of fx_hex8_residual_f64 (feax/assembler.py:
 * 494-535): nanson_q * sum_a N_a(q) (fdir[f] . lambda_a), lambda read as 0
 * on rows with bc_mask set. */
int fx_hex8_traction_vjp_f64(void* stream, int32_t num_faces, const double* x, const double* y,
                             const double* z, const int32_t* cells, const int32_t* faces,
                             const double* fdir, const double* lam, const uint8_t* bc_mask, double scale,
                             double* grad_svar);

/* Neo-Hookean strain energy Pi(u) = sum_e sum_q psi JxW with
 * psi = mu/2 (I1 - 3) - mu ln J + kappa/2 (J - 1)^2 (feax/assembler.py:
 * 1866-1961), written to *energy (device), and, if dpsi_dmu != NULL, its
 * partial derivative per element dPi/dmu_e = sum_q (I1/2 - 3/2 - ln J) JxW
 * (cell variable mu).  Fixed-order reduction; workspace of
 * fx_hex8_energy_workspace_bytes(E) bytes. */
size_t fx_hex8_energy_workspace_bytes(int32_t num_cells);
int fx_hex8_energy_f64(void* stream, int32_t num_cells, const double* x, const double* y, const double* z,
                       const int32_t* cells, const fx_material* mat, const double* param, const double* u,
                       double* energy, double* dpsi_dmu, void* workspace, size_t ws_bytes);

/* ------------------------------------------------------------------ */
/* BiCGSTAB with the iteration and stopping rule of
 * jax.scipy.sparse.linalg.bicgstab (the bicgstab branch of feax's
 * create_iterative_solve_fn, feax/solvers/common.py:241-252): for the
 * non-symmetric operator of symmetric_elimination=False or a caller's
 * solver="bicgstab".  FP64; flags: FX_PCG_BLOCK3 (bcol, n % 3 == 0) or 0
 * (scalar CSR, indices), FX_PCG_NO_GRAPH.  dinv == NULL: no
 * preconditioner.  Stops when ||r||^2 <= max(tol^2 ||b||^2, atol^2), at
 * maxiter, or on breakdown (alpha = 0 or omega = 0); *iters = k, *rs =
 * ||r_k||^2.  workspace >= fx_bicgstab_workspace_bytes(n). */
size_t fx_bicgstab_workspace_bytes(int32_t n);
int fx_bicgstab_f64(void* stream, int32_t n, int64_t nse, int32_t flags, const int32_t* indptr,
                    const int32_t* indices, const int32_t* bcol, const double* data, const double* dinv,
                    const double* b, const double* x0, double tol, double atol, int32_t maxiter,
                    int32_t check_every, double* x, int32_t* iters, double* rs, void* workspace, size_t ws_bytes);

/* ------------------------------------------------------------------ */
/* Restarted GMRES with the iteration and stopping rule of
 * jax.scipy.sparse.linalg.gmres, solve_method='batched' (the gmres branch
 * of feax's create_iterative_solve_fn, feax/solvers/common.py:254-265,
 * restart = options.restart or 20): left-preconditioned (M = Jacobi dinv,
 * or the identity when dinv == NULL), restart clipped to n, maxiter counts
 * restart cycles, stops when ||M (b - A x)|| <= ||M b|| min(1, max(tol
 * ||b||, atol) / ||b||).  FP64; flags as fx_bicgstab_f64.  1 <= restart
 * <= 64.  *cycles = restart cycles run, *rn = ||M (b - A x)||.  workspace
 * >= fx_gmres_workspace_bytes(n, restart, flags) (the (restart+1) x n
 * Krylov basis included); 0 for an invalid restart. */
size_t fx_gmres_workspace_bytes(int32_t n, int32_t restart, int32_t flags);
int fx_gmres_f64(void* stream, int32_t n, int64_t nse, int32_t flags, const int32_t* indptr, const int32_t* indices,
                 const int32_t* bcol, const double* data, const double* dinv, const double* b, const double* x0,
                 double tol, double atol, int32_t restart, int32_t maxiter, int32_t check_every, double* x,
                 int32_t* cycles, double* rn, void* workspace, size_t ws_bytes);

/* ------------------------------------------------------------------ */
/* Direct solve of the SPD J_bc (symmetric elimination): banded Cholesky on
 * the device, the stand-in for feax's GPU direct solver (DirectSolverOptions
 * "auto" / "cudss", feax/solvers/common.py:116-210; the direct branch of
 * create_linear_solver, feax/solvers/linear.py:308-417: u = u0 +
 * J_bc^-1 (-R_bc(u0)), the factor reused by the adjoint since J_bc is
 * symmetric, linear.py:351-378).
 *
 * Band storage (LAPACK lower band): column j at ab + j * ldab, entry (i, j)
 * (0 <= i - j <= bw) at offset i - j, in the caller's ordering: row i' of
 * the band is row perm[i'] of the CSR (inv_perm[perm[i']] = i'; NULL =
 * identity).  ab holds n * ldab doubles, ldab >= fx_band_ldab(bw).
 * workspace >= fx_band_workspace_bytes(n, bw) for each call.
 *
 * fx_band_from_csr_f64: zero ab and scatter the lower triangle of the
 *   (symmetric) CSR; FX_EINVAL if an entry lies past bw.
 * fx_band_cholesky_f64: in-place L (A = L L^T), blocked right-looking
 *   (a one-workgroup Cholesky of each 128-column diagonal block, rocBLAS
 *   trsm / syrk on the band below); *info = 0, or the 1-based column of the
 *   first non-positive pivot (the matrix is not positive definite).  Uses
 *   one rocBLAS handle per device, set to `stream` at each call.
 * fx_band_ldlt_f64: in-place L D L^T without pivoting (unit L below the
 *   diagonal, D on it) for a symmetric indefinite matrix such as a
 *   Neo-Hookean tangent away from equilibrium; *info = 0, or the 1-based
 *   column of the first zero pivot.
 * fx_band_solve_f64: x = P^T L^-T L^-1 P b, or with FX_BAND_LDLT
 *   x = P^T L^-T D^-1 L^-1 P b, where (P b)[i'] = b[perm[i']]; x may alias
 *   b. */
#define FX_BAND_LDLT 1
int64_t fx_band_ldab(int32_t bw);
size_t fx_band_workspace_bytes(int32_t n, int32_t bw);
int fx_band_from_csr_f64(void* stream, int32_t n, const int32_t* indptr, const int32_t* indices, const double* data,
                         const int32_t* inv_perm, int32_t bw, int64_t ldab, double* ab, void* workspace,
                         size_t ws_bytes);
int fx_band_cholesky_f64(void* stream, int32_t n, int32_t bw, int64_t ldab, double* ab, int32_t* info,
                         void* workspace, size_t ws_bytes);
int fx_band_ldlt_f64(void* stream, int32_t n, int32_t bw, int64_t ldab, double* ab, int32_t* info, void* workspace,
                     size_t ws_bytes);
int fx_band_solve_f64(void* stream, int32_t n, int32_t bw, int64_t ldab, int32_t flags, const double* ab,
                      const int32_t* perm, const double* b, double* x, void* workspace, size_t ws_bytes);
/* Triangular solves through the inverses of the diagonal blocks (the
 * default solve of the Python layer): nbs = fx_band_inv_block(bw) rows per
 * block (512, or the largest power of two below ldab), inv holds
 * ceil(n / nbs) nbs x nbs FP64 blocks, column-major (fx_band_inv_bytes),
 * formed after a factorisation by fx_band_invert_diag_f64 (rocBLAS trtri,
 * strided batched; flags FX_BAND_LDLT: unit L).  fx_band_solve_inv_f64: the
 * x of fx_band_solve_f64, two launches per block and direction (the
 * block's coupling to the band as partial sums over 64 columns / rows, then
 * the partials subtracted and the inverse block applied), workspace >=
 * fx_band_solve_inv_workspace_bytes(n, bw). */
int32_t fx_band_inv_block(int32_t bw);
size_t fx_band_inv_bytes(int32_t n, int32_t bw);
size_t fx_band_solve_inv_workspace_bytes(int32_t n, int32_t bw);
int fx_band_invert_diag_f64(void* stream, int32_t n, int32_t bw, int64_t ldab, int32_t flags, const double* ab,
                            double* inv);
int fx_band_solve_inv_f64(void* stream, int32_t n, int32_t bw, int64_t ldab, int32_t flags, const double* ab,
                          const double* inv, const int32_t* perm, const double* b, double* x, void* workspace,
                          size_t ws_bytes);

/* ------------------------------------------------------------------ */
/* Geometric multigrid on a structured HEX8 box (StructuredGrid numbering,
 * feax/spgrid.py:28-41).  The analogue of feax's NarrowBandCMG
 * (feax/solvers/cmg.py:95-300: rediscretised coarse operators on cell
 * averages, trilinear transfers, V-cycle preconditioned CG); here the
 * smoother is a Chebyshev polynomial in D^-1 A and the coarsest level is a
 * dense inverse (or a long Chebyshev polynomial).
 *
 * Level l has nx x ny x nz cells; level l+1 halves or keeps each
 * direction (nx_{l+1} = ceil(nx_l / 2) or nx_l, one direction at least
 * halved; an odd count's last coarse cell reaches past the fine grid).  Every level holds its BC-eliminated block-3 CSR (indptr, bcol
 * of fx_csr_block3_index), dinv = 1 / diag, the constrained-DOF mask,
 * lmax (largest eigenvalue of D^-1 A, fx_gmg_eig_bounds_f64; lmin, used
 * by a Chebyshev coarsest solve only, the smallest or 0) and
 * four vectors of 3 * nodes entries.  Values and vectors are double, or
 * float with FX_GMG_F32.  Level 0's x and b are used only with FX_GMG_F32
 * (the V-cycle's own copy of the CG residual and correction); a coarsest
 * level with dense_inv uses only dense_inv, x, b and bc_mask. */
#define FX_GMG_F32 32
typedef struct fx_gmg_level {
  int32_t nx, ny, nz;
  const int32_t* indptr;
  const int32_t* bcol;
  const void* data;
  const void* dinv;
  const uint8_t* bc_mask;   /* NULL: no constrained DOFs */
  const void* dense_inv;    /* coarsest level: (3N)^2 row-major inverse, or NULL */
  double lmax;
  double lmin;              /* coarsest Chebyshev solve: lower bound (0: lmax / cheb_ratio) */
  void* x;                  /* correction */
  void* b;                  /* right-hand side */
  void* r;                  /* scratch */
  void* d;                  /* scratch */
} fx_gmg_level;

/* Bounds of the spectrum of D^-1 A (A: block-3 CSR of num_nodes nodes,
 * SPD; D^-1 = dinv): `steps` Lanczos steps carried out as Jacobi-PCG on
 * A x = v from a fixed v; *lmin, *lmax = extreme eigenvalues of the
 * Lanczos tridiagonal (inside the spectrum, converging to its ends).
 * Synchronises the stream. */
size_t fx_gmg_eig_bounds_workspace_bytes(int32_t num_nodes, int32_t steps);
int fx_gmg_eig_bounds_f64(void* stream, int32_t num_nodes, const int32_t* indptr, const int32_t* bcol,
                          const double* data, const double* dinv, int32_t steps, double* lmin, double* lmax,
                          void* workspace, size_t ws_bytes);

/* z = V(r): one symmetric V-cycle over nlev levels (Chebyshev of degree
 * smooth_degree on [lmax / cheb_ratio, 1.1 lmax] before and after the
 * coarse correction; coarsest: dense_inv, else Chebyshev of coarse_degree).
 * r, z: double (3N of level 0). */
int fx_gmg_vcycle(void* stream, int32_t flags, int32_t nlev, const fx_gmg_level* levels, int32_t smooth_degree,
                  int32_t coarse_degree, double cheb_ratio, const double* r, double* z);

/* CG preconditioned by the V-cycle, with the stopping rule and iteration of
 * jax.scipy.sparse.linalg.cg (the solve of feax/solvers/common.py:229-239;
 * MGPCG of feax/solvers/cmg.py:276-300).  The CG operator is the FP64 CSR
 * (levels[0].indptr, levels[0].bcol, data); x0 may be NULL (zeros).
 * Iterations run as a HIP graph of check_every iterations unless
 * FX_PCG_NO_GRAPH.  workspace >= fx_pcg_gmg_workspace_bytes(n). */
size_t fx_pcg_gmg_workspace_bytes(int32_t n);
int fx_pcg_gmg_f64(void* stream, int32_t flags, int32_t nlev, const fx_gmg_level* levels, int32_t smooth_degree,
                   int32_t coarse_degree, double cheb_ratio, const double* data, const double* b,
                   const double* x0, double tol, double atol, int32_t maxiter, int32_t check_every, double* x,
                   int32_t* iters, double* rs, void* workspace, size_t ws_bytes);

/* Coarse-grid copy of a fine-grid field for rediscretisation: loc
 * FX_PARAM_NODE -> value at the same node (ncomp values per node: a node
 * parameter, or ncomp = 3 for a displacement); FX_PARAM_CELL -> mean over
 * the child cells (the density averaging of feax/solvers/cmg.py:375-386,
 * 426-431).  Grids as for fx_gmg_level. */
int fx_gmg_coarsen_f64(void* stream, int32_t loc, int32_t ncomp, int32_t nxf, int32_t nyf, int32_t nzf, int32_t nxc,
                       int32_t nyc, int32_t nzc, const double* fine, double* coarse);

/* Dense Galerkin operator Ac = P'^T A P' of the next coarser grid
 * (P' = the trilinear prolongation with the constrained DOFs of both grids
 * masked; constrained coarse DOFs get identity rows and columns): the exact
 * coarse operator for the dense coarsest level, also across a ceil-halved
 * direction.  A: the fine level's BC-eliminated CSR (scalar indices).
 * Ac: (3 Nc)^2 row-major, at most 20000 coarse DOFs.  workspace >=
 * fx_gmg_galerkin_workspace_bytes (the dense A P', 3 Nf x 3 Nc). */
size_t fx_gmg_galerkin_workspace_bytes(int32_t nxf, int32_t nyf, int32_t nzf, int32_t nxc, int32_t nyc,
                                       int32_t nzc);
int fx_gmg_galerkin_dense_f64(void* stream, int32_t nxf, int32_t nyf, int32_t nzf, int32_t nxc, int32_t nyc,
                              int32_t nzc, const int32_t* indptr, const int32_t* indices, const double* data,
                              const uint8_t* mask_f, const uint8_t* mask_c, double* Ac, void* workspace,
                              size_t ws_bytes);

/* dense[i * n + j] = A[i, j] of a CSR matrix (double), zeros elsewhere. */
int fx_csr_to_dense_f64(void* stream, int32_t n, const int32_t* indptr, const int32_t* indices,
                        const double* data, double* dense);

/* ------------------------------------------------------------------------
 * Scalar HEX8 (vec = 1): steady conduction div(k grad T) = 0 with a surface
 * flux -- feax's scalar Problem (vec=1, tensor map ``k * grad_T``,
 * examples/basic/thermo_mech_coupling.py:64-69, tests/test_params_bridge.py:
 * 131-150) on the kernels of feax/assembler.py:406-448, 494-535, 808-818.
 * Same arguments as the fx_hex8_* entry points; DOF a = node a; the material
 * is FX_MAT_CONDUCTION or FX_MAT_SIMP_CONDUCTION; fdir is one number per
 * face (flux = fdir[f] * svar[4 f + q]); u is ignored by the assembly (the
 * law is linear).  The pattern and the face incidence are those of
 * feax/problem.py:316-484 with vec = 1.
 * ------------------------------------------------------------------------ */
int fx_hex8s_pattern(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int64_t* nse,
                     int32_t* max_degree, int32_t* indptr, int32_t* indices, uint8_t* elem_rank,
                     int32_t* inc_ptr, int32_t* inc);
int fx_hex8s_face_incidence(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int32_t num_faces,
                            const int32_t* faces, int32_t* finc_ptr, int32_t* finc);
size_t fx_hex8s_assemble_workspace_bytes(int32_t num_cells, int32_t value_bytes, int32_t flags);
int fx_hex8s_assemble_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                          const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                          int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, double* data,
                          void* workspace, size_t ws_bytes);
int fx_hex8s_assemble_f32(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                          const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                          int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, float* data,
                          void* workspace, size_t ws_bytes);
size_t fx_hex8s_residual_workspace_bytes(int32_t num_cells);
int fx_hex8s_residual_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const fx_material* mat, const double* param, const double* u, int32_t num_faces,
                          const int32_t* faces, const int32_t* finc_ptr, const int32_t* finc, const double* svar,
                          const double* fdir, int32_t num_bc, const int32_t* bc_rows, const double* bc_vals,
                          int32_t negate, double* res, void* workspace, size_t ws_bytes);
size_t fx_hex8s_param_vjp_workspace_bytes(int32_t num_cells, int32_t param_loc);
int fx_hex8s_param_vjp_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                           const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                           const fx_material* mat, const double* param, const double* u, const double* lam,
                           const uint8_t* bc_mask, double scale, double* grad, void* workspace, size_t ws_bytes);
int fx_hex8s_simp_compliance_sens_f64(void* stream, int32_t num_cells, const double* x, const double* y,
                                      const double* z, const int32_t* cells, const fx_material* mat,
                                      const double* param, const double* u, double* dc);
int fx_hex8s_traction_vjp_f64(void* stream, int32_t num_faces, const double* x, const double* y, const double* z,
                              const int32_t* cells, const int32_t* faces, const double* fdir, const double* lam,
                              const uint8_t* bc_mask, double scale, double* grad_svar);

/* Scalar QUAD4 (vec = 1, dim = 2): conduction in the plane, the fx_hex8s_*
 * entry points on QUAD4 geometry (2x2 Gauss points, 2 per edge; flux
 * = fdir[f] * svar[2 f + q]; z is ignored). */
int fx_quad4s_pattern(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int64_t* nse,
                     int32_t* max_degree, int32_t* indptr, int32_t* indices, uint8_t* elem_rank,
                     int32_t* inc_ptr, int32_t* inc);
int fx_quad4s_face_incidence(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int32_t num_faces,
                            const int32_t* faces, int32_t* finc_ptr, int32_t* finc);
size_t fx_quad4s_assemble_workspace_bytes(int32_t num_cells, int32_t value_bytes, int32_t flags);
int fx_quad4s_assemble_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                          const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                          int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, double* data,
                          void* workspace, size_t ws_bytes);
int fx_quad4s_assemble_f32(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                          const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                          int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, float* data,
                          void* workspace, size_t ws_bytes);
size_t fx_quad4s_residual_workspace_bytes(int32_t num_cells);
int fx_quad4s_residual_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const fx_material* mat, const double* param, const double* u, int32_t num_faces,
                          const int32_t* faces, const int32_t* finc_ptr, const int32_t* finc, const double* svar,
                          const double* fdir, int32_t num_bc, const int32_t* bc_rows, const double* bc_vals,
                          int32_t negate, double* res, void* workspace, size_t ws_bytes);
size_t fx_quad4s_param_vjp_workspace_bytes(int32_t num_cells, int32_t param_loc);
int fx_quad4s_param_vjp_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                           const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                           const fx_material* mat, const double* param, const double* u, const double* lam,
                           const uint8_t* bc_mask, double scale, double* grad, void* workspace, size_t ws_bytes);
int fx_quad4s_simp_compliance_sens_f64(void* stream, int32_t num_cells, const double* x, const double* y,
                                      const double* z, const int32_t* cells, const fx_material* mat,
                                      const double* param, const double* u, double* dc);
int fx_quad4s_traction_vjp_f64(void* stream, int32_t num_faces, const double* x, const double* y, const double* z,
                              const int32_t* cells, const int32_t* faces, const double* fdir, const double* lam,
                              const uint8_t* bc_mask, double scale, double* grad_svar);

/* Scalar TET4 (vec = 1): the fx_hex8s_* entry points on TET4 geometry
 * (one Gauss point; flux = fdir[f] * svar[f] at the facet centroid). */
int fx_tet4s_pattern(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int64_t* nse,
                     int32_t* max_degree, int32_t* indptr, int32_t* indices, uint8_t* elem_rank,
                     int32_t* inc_ptr, int32_t* inc);
int fx_tet4s_face_incidence(int32_t num_nodes, int32_t num_cells, const int32_t* cells, int32_t num_faces,
                            const int32_t* faces, int32_t* finc_ptr, int32_t* finc);
size_t fx_tet4s_assemble_workspace_bytes(int32_t num_cells, int32_t value_bytes, int32_t flags);
int fx_tet4s_assemble_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                          const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                          int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, double* data,
                          void* workspace, size_t ws_bytes);
int fx_tet4s_assemble_f32(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const uint8_t* elem_rank, const int32_t* indptr, const int32_t* indices,
                          const fx_material* mat, const double* param, const double* u, const uint8_t* bc_mask,
                          int32_t symmetric, int32_t flags, double jacobi_shift, double* dinv, float* data,
                          void* workspace, size_t ws_bytes);
size_t fx_tet4s_residual_workspace_bytes(int32_t num_cells);
int fx_tet4s_residual_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                          const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                          const fx_material* mat, const double* param, const double* u, int32_t num_faces,
                          const int32_t* faces, const int32_t* finc_ptr, const int32_t* finc, const double* svar,
                          const double* fdir, int32_t num_bc, const int32_t* bc_rows, const double* bc_vals,
                          int32_t negate, double* res, void* workspace, size_t ws_bytes);
size_t fx_tet4s_param_vjp_workspace_bytes(int32_t num_cells, int32_t param_loc);
int fx_tet4s_param_vjp_f64(void* stream, int32_t num_nodes, int32_t num_cells, const double* x, const double* y,
                           const double* z, const int32_t* cells, const int32_t* inc_ptr, const int32_t* inc,
                           const fx_material* mat, const double* param, const double* u, const double* lam,
                           const uint8_t* bc_mask, double scale, double* grad, void* workspace, size_t ws_bytes);
int fx_tet4s_simp_compliance_sens_f64(void* stream, int32_t num_cells, const double* x, const double* y,
                                      const double* z, const int32_t* cells, const fx_material* mat,
                                      const double* param, const double* u, double* dc);
int fx_tet4s_traction_vjp_f64(void* stream, int32_t num_faces, const double* x, const double* y, const double* z,
                              const int32_t* cells, const int32_t* faces, const double* fdir, const double* lam,
                              const uint8_t* bc_mask, double scale, double* grad_svar);

#ifdef __cplusplus
}
#endif
#endif /* FEAXHIP_H */
