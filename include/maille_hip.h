/* maille_hip — MI355X (gfx950) implementation of maille-core's 3D linear-elastic solve path:
 * Tet4/Tet10 element stiffness, CSR assembly in the reverse-Cuthill-McKee lower-triangle
 * layout, IC(0) factorisation with the diagonal-shift ladder, and preconditioned conjugate
 * gradients.
 *
 * The C ABI is what a Rust FFI crate binds (extern "C", plain pointers and sizes). Each entry
 * point cites the reference item it replaces as crates/maille-core/<file>:<line> of
 * timothee-faget/maille v0.14.0.
 *
 * Conventions
 *  - Host buffers are caller-allocated; entry points with the _device suffix take device
 *    pointers (hipMalloc'd, on the handle's device) instead.
 *  - Handles own their device memory (as Pcg owns its Csr by move, iterative.rs:194-197).
 *    A handle may be used from any one thread at a time; calls on one handle are serialised by
 *    the caller, as in the reference (single-threaded, Send types).
 *  - No exception or abort crosses the ABI; every entry point returns an ml_status.
 *  - Indices are int64_t at the boundary (Rust usize); the device stores int32 and the entry
 *    points reject sizes that do not fit.
 */
#ifndef MAILLE_HIP_H
#define MAILLE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The first three mirror MailleError (error.rs:224-229). */
typedef enum ml_status {
  ML_OK = 0,
  ML_SINGULAR = 1,            /* MailleError::Singular (error.rs:224) */
  ML_NOT_CONVERGED = 2,       /* MailleError::IterativeNotConverged (error.rs:226-229) */
  ML_INVALID_ARGUMENT = 3,    /* null pointer, size out of range, bad element family */
  ML_OUT_OF_PATTERN = 4,      /* Csr::add outside the profile (iterative.rs:126 panics) */
  ML_HIP_ERROR = 5,           /* a HIP runtime call failed; see ml_last_error() */
  ML_INVALID_MESH = 6,        /* Mesh::new refusals (mesh.rs:257-308), see ml_solve_info */
  ML_DEVICE_TIMEOUT = 7,      /* a dataflow kernel's bounded wait expired (never expected) */
  ML_INVALID_FILE = 8,        /* an input file refused by ml_msh_read (IoError, error.rs:9-65) */
  ML_NEWTON_NOT_CONVERGED = 9,/* MailleError::NotConverged (error.rs:158-163): Newton-Raphson
                                 failed even on the smallest cut of an increment */
  ML_PLASTIC_COLLAPSE = 10,   /* MailleError::PlasticCollapse (error.rs:164-168): the tangent
                                 stopped being positive definite while points yield */
  ML_DIRECT_REQUIRED = 11     /* ml_model_create with ML_AUTO below ML_ITERATIVE_FROM free DOFs:
                                 SolverKind::Auto takes the direct skyline there (solve.rs:
                                 185-194), the caller's own arm; no handle was created */
} ml_status;

/* Arithmetic of the solve. ML_FP64 is the drop-in (the reference computes in f64).
 * The mixed modes factor IC(0) in FP32 (the shift ladder in FP32 arithmetic) and store the
 * factor in FP32; they apply it with FP64 vectors and sums (each coefficient widened exactly),
 * which is the reference's FP64 triangular solve of the FP32-rounded factor:
 * ML_MIXED: FP32 SpMV inside one PCG whose iterate x, residual r, alpha/beta dot products and
 * refinement stay FP64: the recursive residual is replaced by the FP64 true residual b - A x
 * each time it has dropped tenfold (reliable updates), until a replacement stops reducing it.
 * ML_MIXED_IC0: FP64 SpMV and PCG, exact triangular solves of the FP32 factor.
 * ML_MIXED_JACOBI: FP64 SpMV and PCG; the FP32 factor applied by k Jacobi sweeps per
 * triangular system instead of exact solves (bandwidth-bound instead of latency-bound; no
 * counterpart in the reference): y = D^-1 r, then k times y <- D^-1 (r - N y) with N the
 * strict lower part of L, and the transposed iteration on L^T z = y, so that the
 * preconditioner stays symmetric. k is 4 until ml_model_autotune / ml_pcg_autotune times
 * whole solves and keeps the fastest k (0, the exact solves, included), or the caller sets it
 * (ml_pcg_set_sweeps). */
typedef enum ml_precision {
  ML_FP64 = 0, ML_MIXED = 1, ML_MIXED_IC0 = 2, ML_MIXED_JACOBI = 3
} ml_precision;

/* SolverKind (solve.rs:171-195). The library is the iterative arm: ML_ITERATIVE always runs the
 * device IC(0)-PCG, ML_AUTO runs it from ML_ITERATIVE_FROM free DOFs on and is answered with
 * ML_DIRECT_REQUIRED below (where Auto is Direct), ML_DIRECT is refused. */
typedef enum ml_solver_kind { ML_AUTO = 0, ML_DIRECT = 1, ML_ITERATIVE = 2 } ml_solver_kind;
#define ML_ITERATIVE_FROM 100000 /* solve.rs:185 */
#define ML_TOLERANCE 1e-12       /* iterative.rs:9 */
#define ML_MAX_ITERATIONS 10000  /* iterative.rs:11 */

typedef struct ml_solve_info {
  int64_t iterations;        /* PCG iterations (both phases of an escalated ML_MIXED solve) */
  double residual;           /* recursive relative residual |r|/|b| at exit (iterative.rs:246) */
  double true_residual;      /* |b - A x| / |b| computed in FP64 after the solve */
  int32_t shift_index;       /* index into the shift ladder {0,1e-3,1e-2,0.1,1,10} used */
  int32_t refinements;       /* FP64 true-residual replacements (ML_MIXED), else 0 */
  int64_t bad_element;       /* ML_INVALID_MESH: the first offending element (the node for kinds
                                4 and 5, as NonFiniteCoordinate and UnknownNode carry it), else -1 */
  int32_t mesh_error;        /* ML_INVALID_MESH: 1 degenerate, 2 inverted, 3 curved Tet10 edge,
                                4 non-finite coordinate, 5 unknown node, 6 bad material,
                                7 empty mesh (mesh.rs:258-300, material.rs:57-60,145-153) */
  int32_t escalated;         /* ML_MIXED: 1 when a residual replacement stopped halving the FP64
                                true residual and the solve continued from its iterate with the
                                FP64 operator (FP32 preconditioner kept), else 0 */
  int32_t device_retries;    /* dataflow launches (IC(0), triangular solves) that stalled and were
                                repeated on a halved grid during this call, normally 0 */
  int32_t reserved;
  double seconds[6];         /* device wall time: assemble, factor, solve, reactions, stresses,
                                host pattern build */
} ml_solve_info;

/* Last error text of the calling thread (HIP error string or argument diagnosis). */
const char* ml_last_error(void);
/* Number of visible HIP devices (0 without a GPU); never fails. */
int ml_device_count(void);

/* ------------------------------------------------------------------ element kernels */

/* solid_d (element.rs:193-206): row-major 6x6 isotropic D, Voigt (xx,yy,zz,yz,xz,xy). Host. */
ml_status ml_solid_d(double young, double nu, double* d);

/* solid_stiffness (element.rs:223-243) for a batch, as element_stiffness dispatches it
 * (assembly.rs:64-79, 204-248): Tet4 with TETRA_1, Tet10 with TETRA_4.
 * coords: n_elements x nen x 3, young/nu: n_elements, ke: n_elements x (3nen)^2 row-major
 * (ke[e][j][i] as the reference's Vec<Vec<f64>>). Host buffers; computed on the device. */
ml_status ml_stiffness_batch(int nen, int64_t n_elements, const double* coords,
                             const double* young, const double* nu, double* ke);
/* No counterpart in the reference: ml_stiffness_batch for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc). Same
 * arguments and results, bit for bit; it calls no hipMalloc, hipFree or stream creation and
 * stages the batch in a file-scope device array of ML_STIFFNESS_SCRATCH_BYTES, one per GPU,
 * held until the process ends, on the null stream. A batch needs (24 nen + 16 + 4 nen +
 * 2 (3 nen)^2 x 8) bytes per element (plus 256-byte alignment of six regions): at most 27,593
 * Tet4 or 4,566 Tet10 elements per call, larger batches are refused with ML_INVALID_ARGUMENT.
 * Calls on one GPU share the array: calls that overlap on one GPU (from several threads)
 * race. */
#define ML_STIFFNESS_SCRATCH_BYTES ((size_t)64 << 20)
ml_status ml_stiffness_batch_noalloc(int nen, int64_t n_elements, const double* coords,
                                     const double* young, const double* nu, double* ke);

/* Diagnostics (no counterpart in the reference; host only): out[i] = num[i] / den[i] by the
 * sequence the FP64 triangular solves divide with (the reciprocal of the diagonal and two
 * exact-remainder corrections: the division's result bit for bit for den > 0 and
 * 2^-969 <= |num| (or num = +-0) with |num / den| < 2^1020, within one ulp outside that),
 * so that it can be compared with the division on the host. */
ml_status ml_quotient64(int64_t n, const double* num, const double* den, double* out);

/* ------------------------------------------------------------------ pattern */

typedef struct ml_pattern ml_pattern;

/* Graph::from_elements (iterative.rs:38-76) + reverse_cuthill_mckee (solve.rs:200-245) +
 * Csr::zeros (iterative.rs:91-114) on the host: n vertices, element vertex lists given as
 * CSR (elem_offsets[n_elements+1], elem_items). */
ml_status ml_pattern_build(int64_t n, int64_t n_elements, const int64_t* elem_offsets,
                           const int64_t* elem_items, ml_pattern** out);
int64_t ml_pattern_n(const ml_pattern* p);
int64_t ml_pattern_nnz(const ml_pattern* p);
/* order[n] (order[new] = old, solve.rs:200), offsets[n+1], cols[nnz] (lower triangle,
 * ascending columns, diagonal last — iterative.rs:79-86). Any pointer may be null. */
ml_status ml_pattern_get(const ml_pattern* p, int64_t* order, int64_t* offsets, int64_t* cols);
/* Graph neighbour lists (iterative.rs:16-22): offsets[n+1], targets[ml_pattern_ntargets]. */
int64_t ml_pattern_ntargets(const ml_pattern* p);
ml_status ml_pattern_graph(const ml_pattern* p, int64_t* offsets, int64_t* targets);
void ml_pattern_free(ml_pattern* p);

/* Diagnostics (no counterpart in the reference; host only, no device needed): builds every
 * index array the device kernels of ml_pcg_create would read for the lower CSR pattern
 * (offsets[n+1], cols[]; ascending columns, diagonal last) and checks each bound they rely
 * on: index ranges, block capacities, each row's entries covered once and in order, and a
 * topological processing order of the triangular-solve blocks; likewise the two Jacobi-sweep
 * layouts (ML_MIXED_JACOBI) and, where the pattern fits them, the fused ones. ML_INVALID_ARGUMENT with the
 * failed check in ml_last_error otherwise. stats[8] (optional): forward and backward block
 * counts, their largest external and internal entry counts (forward, then backward), the
 * DAG depth and the longest row. With lvalues (the factor, lower-CSR order), r and z
 * non-null, also applies z = (L L^T)^-1 r on the host through the triangular-solve layouts,
 * in the device's block order and summation order. */
ml_status ml_layout_verify(int64_t n, const int64_t* offsets, const int64_t* cols,
                           const double* lvalues, const double* r, double* z, int64_t* stats);
/* Diagnostics (no counterpart in the reference; host only): builds and checks the fused
 * Jacobi-sweep layouts of the pattern in blocks of `rows` rows (256 or 512;
 * ML_INVALID_ARGUMENT, with the reason, when the pattern does not fit them) and, with
 * lvalues, r and z non-null, applies `sweeps` sweeps of each triangular system on the host
 * through those layouts, staging every value the way the fused kernel does, with the factor
 * rounded to FP32 (z = P^T P r). stats[5] (optional): blocks, the longest row, the most
 * external columns of a lower and of an upper block, and the code of the fused variant a
 * handle on this pattern starts with before autotuning (0: one launch per sweep). */
ml_status ml_layout_sweeps(int64_t n, const int64_t* offsets, const int64_t* cols, int rows,
                           const double* lvalues, const double* r, int sweeps, double* z,
                           int64_t* stats);

/* ------------------------------------------------------------------ Pcg on a given Csr */

typedef struct ml_pcg ml_pcg;

/* Pcg::new (iterative.rs:201-207): uploads a lower-triangle CSR in the reference layout
 * (n rows, offsets[n+1], cols[nnz], values[nnz]) and factors IC(0) on the device, trying the
 * shifts {0, 1e-3, 1e-2, 0.1, 1, 10} in turn. ML_SINGULAR if every shift fails. */
ml_status ml_pcg_create(int64_t n, const int64_t* offsets, const int64_t* cols,
                        const double* values, ml_precision precision, ml_pcg** out);
/* Pcg::solve (iterative.rs:210-248) for nrhs right-hand sides stored one after the other
 * (b[k*n .. k*n+n)), x likewise. tol <= 0 and maxit <= 0 select the reference's 1e-12 and
 * 10000. A zero right-hand side gives exactly x = 0. info may be null; with nrhs > 1 it
 * holds the largest iteration count and residual. With nrhs > 1 the right-hand sides are
 * solved together, up to 256 per set of launches: one PCG per right-hand side (its own
 * alpha, beta and stop, as sequential Pcg::solve calls would run them), every SpMV and
 * triangular solve covering all of them, with an FP64 operator. */
ml_status ml_pcg_solve(ml_pcg* h, const double* b, double* x, int64_t nrhs, double tol,
                       int64_t maxit, ml_solve_info* info);
ml_status ml_pcg_solve_device(ml_pcg* h, const double* b_device, double* x_device,
                              double tol, int64_t maxit, ml_solve_info* info);
/* y = A x (Csr::multiply, iterative.rs:131-142), host buffers. */
ml_status ml_pcg_multiply(ml_pcg* h, const double* x, double* y);
/* The same for nrhs vectors stored one after the other (x[k n + i], y likewise), as the
 * handle's batched solves multiply (every vector in one launch); each y equals ml_pcg_multiply
 * of its x bit for bit. n * nrhs below 2^31. */
ml_status ml_pcg_multiply_many(ml_pcg* h, int64_t nrhs, const double* x, double* y);
/* z = (L L^T)^-1 r (Csr::precondition, iterative.rs:173-188), host buffers. */
ml_status ml_pcg_precondition(ml_pcg* h, const double* r, double* z);
/* The same for nrhs vectors stored one after the other (r[k n + i], z likewise), as the handle's
 * batched solves (ml_pcg_solve with nrhs > 1) apply it: the exact solves over every vector, or
 * the handle's Jacobi sweeps over every vector in one launch per sweep; each z equals
 * ml_pcg_precondition of its r bit for bit. n * nrhs below 2^31. */
ml_status ml_pcg_precondition_many(ml_pcg* h, int64_t nrhs, const double* r, double* z);
/* The IC(0) factor values aligned with the CSR (Pcg::l, iterative.rs:196) and the shift. */
ml_status ml_pcg_factor(ml_pcg* h, double* l, int32_t* shift_index);
void ml_pcg_destroy(ml_pcg* h);
/* Diagnostics (no counterpart in the reference): one preconditioner application with a
 * per-block timeline of both triangular solves. nblocks[2] receives the block counts; with
 * non-null traces (capacity int64 each, >= 10 per block + n) each block b records in
 * trace[10 b ..] its start, the end of its staging, its end (device wall-clock ticks),
 * (workgroup << 8 | rows still waiting on an external value after staging), the ticks the
 * solving wave spent waiting, the first polling wave's rounds and its ticks in loads, the
 * solving wave's waiting iterations, the loads the first polling wave issued and the block's
 * distinct external columns; trace[10 nblocks + s] is the publish time of the row of slot s (the
 * block layout's s-th row, in logical row order of that solve). */
ml_status ml_pcg_trace(ml_pcg* h, const double* r, int64_t* nblocks, int64_t* trace_fw,
                       int64_t* trace_bw, int64_t capacity);
/* Diagnostics (no counterpart in the reference): the bound of every dataflow wait of this
 * handle, in device wall-clock ticks (0 restores the default, 2 s). With a tiny bound every
 * wait gives up at once and every launch still ends normally: the handle repeats the work on
 * halved grids up to three times, then the call fails with ML_DEVICE_TIMEOUT. */
ml_status ml_pcg_set_wait_limit(ml_pcg* h, int64_t ticks);
/* No counterpart in the reference: the Jacobi sweeps per triangular system with which a mixed
 * handle applies its FP32 factor (0..64; 0: the exact triangular solves). An FP64 handle
 * takes 0 only (ML_INVALID_ARGUMENT otherwise). A count set here is kept by the autotuner.
 * Batched solves apply the same sweeps: a batch of models (ml_batch) by the fused kernel over
 * the block-diagonal operator, several right-hand sides (nrhs > 1) by one launch per sweep
 * over all of them (ml_pcg_precondition_many). ml_pcg_sweeps returns the current count. */
ml_status ml_pcg_set_sweeps(ml_pcg* h, int sweeps);
int ml_pcg_sweeps(const ml_pcg* h);
/* No counterpart in the reference: the weight w of the sweeps (0 < w <= 1; 1, the default, is
 * undamped). A damped sweep is y <- (1 - w) y + w D^-1 (r - N y), in both systems, so the
 * preconditioner stays symmetric positive definite; it converges in fewer PCG iterations where
 * the factor's off-diagonal part is large against its diagonal (Tet10). The autotuner picks the
 * weight with the count (1 or 0.8) unless the caller set either; ML_INVALID_ARGUMENT on an FP64
 * handle or a weight outside (0, 1]. */
ml_status ml_pcg_set_sweep_weight(ml_pcg* h, double w);
double ml_pcg_sweep_weight(const ml_pcg* h);
/* No counterpart in the reference: how the sweeps are launched. fused 1 (the default): all
 * sweeps of a triangular system in one launch of a persistent kernel that keeps each row's
 * coefficients in registers, so the factor is read once per system instead of once per sweep,
 * wherever the pattern fits it (rows of at most 252 strict entries; blocks of 256 rows that
 * read at most 1792 distinct columns outside themselves), in the variant ml_pcg_autotune
 * timed fastest (before that, one row per lane with every row in registers); a variant code
 * 100 r + c: the fused kernel with r rows per lane (1, or 2 in blocks of 512 rows for rows of
 * at most 32 strict entries) and c coefficients per row held in registers (a longer row
 * re-reads the rest each sweep): 116 (registers bounded for 8 waves per SIMD), 124, 132, 148,
 * 180, 216 (bounded for 5 waves), 224, 232; 0: one launch per sweep. Every form gives bit-identical results.
 * ml_pcg_sweeps_fused returns the code of the fused variant applying the handle's sweeps, 0
 * when they are launched one by one (fused 0, no sweeps, or a pattern the fused kernel does
 * not fit). */
ml_status ml_pcg_set_sweeps_fused(ml_pcg* h, int fused);
int ml_pcg_sweeps_fused(const ml_pcg* h);
/* dot (solve.rs:100-110) as PCG's scalars are computed: on 1, every PCG dot product (|b|, r.z,
 * p.Ap, |r|) runs in the reference's order (eight sequential lane sums over the chunks of
 * eight, their fixed combination, the tail in order; one workgroup, about 1 ms per dot on the
 * benchmark cube) instead of fixed-tree sums of block partials. With ML_FP64, whose SpMV, IC(0),
 * triangular solves and vector updates already follow the reference's operations, the whole
 * Pcg::solve (iterative.rs:210-248) is then the reference's bit for bit: the same iterates,
 * iteration count and solution. 0 (the default) restores the tree sums. Refused for ML_MIXED
 * (its FP32 operator and residual replacements have no reference counterpart). */
ml_status ml_pcg_set_reference_dots(ml_pcg* h, int on);

/* ------------------------------------------------------------------ the static model */

typedef struct ml_model ml_model;

/* Model::new + FreeSystem::new up to the factorisation (model.rs:251, 1442-1499) for a mesh
 * of one element family: nodes[n_nodes*3], conn[n_elements*nen] (nen 4 = Tet4, 10 = Tet10),
 * young/nu per element (material.rs:16-34), fixed[3*n_nodes] (0/1) with imposed values
 * (model.rs:1133-1139). Validates the mesh (mesh.rs:257-348), builds free numbering, Graph,
 * RCM and the CSR pattern on the host, uploads and assembles K_ff on the device.
 * solver: ML_ITERATIVE runs the device IC(0)-PCG; ML_AUTO does so from ML_ITERATIVE_FROM free
 * DOFs on and returns ML_DIRECT_REQUIRED below (solve.rs:185-194: Auto is Direct there; the
 * check follows the host free numbering, before any device work); ML_DIRECT is refused
 * (ML_INVALID_ARGUMENT: the skyline Cholesky is not part of this library). A refused mesh
 * returns ML_INVALID_MESH with info->mesh_error and info->bad_element set. */
ml_status ml_model_create(int64_t n_nodes, const double* nodes, int64_t n_elements, int nen,
                          const int64_t* conn, const double* young, const double* nu,
                          const uint8_t* fixed, const double* imposed, ml_solver_kind solver,
                          ml_precision precision, ml_model** out, ml_solve_info* info);
/* ml_pcg_set_sweeps / ml_pcg_sweeps for the model's operator. */
ml_status ml_model_set_sweeps(ml_model* m, int sweeps);
int ml_model_sweeps(const ml_model* m);
/* ml_pcg_set_sweep_weight / ml_pcg_sweep_weight for the model's operator. */
ml_status ml_model_set_sweep_weight(ml_model* m, double w);
double ml_model_sweep_weight(const ml_model* m);
/* ml_pcg_set_reference_dots for the model's operator: with ML_FP64 and zero imposed
 * displacements, Model::solve's displacements are then the reference's bit for bit. */
ml_status ml_model_set_reference_dots(ml_model* m, int on);
/* ml_pcg_set_sweeps_fused / ml_pcg_sweeps_fused for the model's operator. */
ml_status ml_model_set_sweeps_fused(ml_model* m, int fused);
int ml_model_sweeps_fused(const ml_model* m);
/* Bytes per column index the kernels read: out[0] the SpMV's (2: 16-bit offsets from the row,
 * when the matrix band fits; 4: int32; 0: the grouped streams, see ml_model_spmv_index),
 * out[1] the sweeps' (2 or 4, always 2 in the fused kernel, whose 16-bit indices name LDS
 * slots; 0 before any sweep layout is built). */
ml_status ml_model_index_bytes(const ml_model* m, int32_t* out);
/* No counterpart in the reference: the SpMV's index form in use, out[0] = 1 when the rows of a
 * node share one 16-bit column stream (grouped, chosen by ml_model_autotune when faster), 2 for
 * 16-bit offsets per entry, 4 for int32 columns; out[1] = the bytes of column indices and
 * per-row fields one SpMV reads in that form. */
ml_status ml_model_spmv_index(const ml_model* m, int64_t* out);
/* Diagnostics (no counterpart in the reference): out[8] = workgroups per CU that the runtime's
 * occupancy figure reports and that the dataflow grids use, for the forward triangular solve,
 * the backward one, IC(0) and the fused sweeps (zeros until fused sweep layouts are built):
 * one fewer than the largest count every workgroup of which was resident at once in three
 * co-residency probes of the kernel itself. Zeros for a model without free DOFs. */
ml_status ml_model_residency(const ml_model* m, int32_t* out);
/* Recompute the element stiffnesses and K_ff values on the device (after ml_model_set_young). */
ml_status ml_model_assemble(ml_model* m);
/* Replace the per-element Young's moduli (n_elements values, host) without touching the
 * pattern; call ml_model_assemble and ml_model_factor afterwards. */
ml_status ml_model_set_young(ml_model* m, const double* young);
/* kff.factorize (model.rs:1496) -> Pcg::new: device IC(0) with the shift ladder. */
ml_status ml_model_factor(ml_model* m, ml_solve_info* info);
/* FreeSystem::solve (model.rs:1502-1516) + reactions (model.rs:586-593): f[3*n_nodes] nodal
 * loads (model.rs:1090-1121, pressures included by the caller or ml_model_add_pressures);
 * u[3*n_nodes] out; reactions[3*n_nodes] out or null. Host buffers. */
ml_status ml_model_solve(ml_model* m, const double* f, double* u, double* reactions,
                         ml_solve_info* info);
/* Same with device buffers (f, u, reactions of 3*n_nodes doubles; reactions may be null). */
ml_status ml_model_solve_device(ml_model* m, const double* f_device, double* u_device,
                                double* reactions_device, ml_solve_info* info);
/* element_stresses for every element (model.rs:596-598, element.rs:270-302): out[n_elements
 * x ngp x 9] = x, y, z, sxx, syy, szz, tyz, txz, txy per Gauss point (1 for Tet4, 4 Tet10). */
ml_status ml_model_stresses(ml_model* m, const double* u, double* out);
/* Same with device buffers (u of 3*n_nodes doubles, out of n_elements x ngp x 9). */
ml_status ml_model_stresses_device(ml_model* m, const double* u_device, double* out_device);
/* nodal_stresses (results.rs:138-166, smooth :184-221): the Gauss-point stresses of u
 * extrapolated to each element's nodes and averaged over the elements sharing each node, every
 * sum in the reference's element and Gauss-point order; out[n_nodes x 9] = x, y, z of the node,
 * sxx, syy, szz, tyz, txz, txy (zeros at a node no element uses). */
ml_status ml_model_nodal_stresses(ml_model* m, const double* u, double* out);
/* face_pressure + FacePressure::add_load (assembly.rs:638-674) for dead pressures
 * (element, face, value), added to f[3*n_nodes] in the given order. Host. */
ml_status ml_model_add_pressures(ml_model* m, int64_t count, const int64_t* elements,
                                 const int64_t* faces, const double* values, double* f);
/* edge_load + FacePressure::add_load (assembly.rs:680-710, 638-644; model.rs:1123-1127) for dead
 * line loads (element, edge 0-5 as mesh.rs:204-211, q[3] per unit length), added to f[3*n_nodes]
 * in the given order after the pressures, as Model::assemble does. Host. */
ml_status ml_model_add_edge_loads(ml_model* m, int64_t count, const int64_t* elements,
                                  const int64_t* edges, const double* q, double* f);
/* Sizes: free DOFs, stored lower-triangle entries, DAG depth of the IC(0) factor. */
int64_t ml_model_nfree(const ml_model* m);
int64_t ml_model_nnz(const ml_model* m);
int64_t ml_model_levels(const ml_model* m);
/* K_ff in the reference layout (after assembly): order[nfree] (RCM, over free indices),
 * offsets[nfree+1], cols[nnz], values[nnz]; any pointer may be null. */
ml_status ml_model_csr(ml_model* m, int64_t* order, int64_t* offsets, int64_t* cols,
                       double* values);
/* Adjoint compliance sensitivity (no counterpart in the reference; the backward() of the path):
 * for C = f.u with u the last solution of this model and homogeneous essential conditions,
 * dC/dE_e = -u_e^T (K_e/E_e) u_e, since compliance is self-adjoint (adjoint state = u) and K_e
 * is linear in E_e. out[n_elements] (host). A SIMP caller chains dE_e/drho_e itself.
 * ML_INVALID_ARGUMENT before a solve or with nonzero imposed displacements. */
ml_status ml_model_compliance_gradient(ml_model* m, double* out);
void ml_model_destroy(ml_model* m);

/* ------------------------------------------------------------------ quasi-static plasticity */

typedef struct ml_quasi_static_info {
  int64_t newton_iterations;  /* linear solves of every equilibrium attempt, cuts included */
  int64_t cuts;               /* increment halvings (Event::Cut, model.rs:840-851) */
  int64_t pcg_iterations;     /* PCG iterations of all those solves */
  int64_t failed_increment;   /* on ML_NEWTON_NOT_CONVERGED / ML_PLASTIC_COLLAPSE / a solver
                                 error: the Step increment (1-based) that failed, else -1 */
  double failed_time;         /* its end time (the `time` of the error) */
  int32_t device_retries;     /* stalled dataflow launches repeated on smaller grids */
  int32_t reserved;
  double seconds[4];          /* element responses + residuals, assembly + IC(0), PCG, total */
} ml_quasi_static_info;

/* Model::solve_quasi_static (model.rs:715-797) for small strains on this model's tetrahedra:
 * von Mises plasticity with linear isotropic hardening (plasticity.rs:34-112; per element
 * yield_stress > 0, infinite for an elastic element, and hardening >= 0, material.rs:117-130),
 * loads f[3*n_nodes] and the model's BC values scaled by the amplitude lambda(t) (npoints
 * points amp_t/amp_v, the first (0, 0), times increasing: Amplitude::new, model.rs:191-213) over
 * `increments` equal increments of time_period (Step::quasi_static, model.rs:136-150). Each
 * increment is an equilibrium solved by Newton-Raphson on the device (model.rs:923-1022: at
 * least one iteration, converged at relative residual <= 1e-8, at most 25), cut in halves up
 * to 5 times when Newton fails (model.rs:800-857). Outputs for increments 0..n (increment 0 is
 * the mesh at rest), host arrays, any may be null:
 *   u[(n+1) x 3*n_nodes], reactions[(n+1) x 3*n_nodes] (F_int - lambda F on the fixed DOFs),
 *   stresses[(n+1) x n_elements x ngp x 9] (x, y, z, sxx, syy, szz, tyz, txz, txy per Gauss
 *   point), p[(n+1) x n_elements x ngp] (equivalent plastic strain),
 *   residuals[n x 26] with residual_counts[n] (Increment::residuals of the increment's last
 *   equilibrium). The norms behind the residuals are fixed-order tree sums, so they agree with
 *   the reference's sequential sums to rounding.
 * The model's matrix holds tangents afterwards; a later ml_model_solve reassembles K. */
ml_status ml_model_solve_quasi_static(ml_model* m, const double* yield_stress,
                                      const double* hardening, const double* f,
                                      double time_period, int64_t increments, int64_t npoints,
                                      const double* amp_t, const double* amp_v, double* u,
                                      double* reactions, double* stresses, double* p,
                                      double* residuals, int64_t* residual_counts,
                                      ml_quasi_static_info* info);

/* Model::solve_quasi_static with Kinematics::Large (model.rs:732-1022, with_kinematics) on this
 * model's tetrahedra: total-Lagrangian elements (element_large_response, finite_strain.rs:119-214)
 * with the Saint-Venant-Kirchhoff law, or the compressible Neo-Hookean law where neo_hookean[e]
 * is 1 (Hyperelasticity::NeoHookean, material.rs:36-42; null: all Saint-Venant-Kirchhoff), Lame
 * constants from the model's E and nu; follower pressures (element, face 0-3, value) whose forces
 * and symmetrised stiffness follow the deformed faces (follower_pressure, assembly.rs:732-862;
 * follower_force, model.rs:868-903); loads f (dead: nodal loads and edge loads, no pressures) and
 * BC values scaled by the amplitude, Newton-Raphson and increment cutting as
 * ml_model_solve_quasi_static. Outputs as there: u, reactions, stresses (Cauchy stresses at the
 * current Gauss-point positions), residuals / residual_counts; any may be null. The follower
 * terms are computed on the host from the iterate (one copy of u per Newton iteration). */
ml_status ml_model_solve_large_strain(ml_model* m, const uint8_t* neo_hookean, const double* f,
                                      int64_t npressures, const int64_t* pres_elements,
                                      const int64_t* pres_faces, const double* pres_values,
                                      double time_period, int64_t increments, int64_t npoints,
                                      const double* amp_t, const double* amp_v, double* u,
                                      double* reactions, double* stresses, double* residuals,
                                      int64_t* residual_counts, ml_quasi_static_info* info);

/* ------------------------------------------------------------------ transient heat conduction */

typedef struct ml_thermal ml_thermal;

/* The thermal model of Model::solve_step (model.rs:620-706) on tetrahedra: one DOF (temperature)
 * per node (model.rs:1074-1078). Per element: conductivity k > 0 and rho*c > 0 (density times
 * specific heat, material.rs:21-30); convection faces (element, face 0-3 as
 * Element::face_nodes, film h, sink temperature) whose h int N N^T is added to the element's K
 * (thermal_matrix, model.rs:1320-1326, assembly.rs:598-624); fixed[n_nodes] (0/1) with the BC
 * temperatures imposed[n_nodes]. Validates the mesh, builds the free numbering, RCM pattern and
 * assembly plan on the host and the conduction and consistent capacity matrices (TETRA_4 /
 * TETRA_14 for the capacity, assembly.rs:413-431) on the device. A model whose every
 * temperature is fixed has no free system and is refused (ML_INVALID_ARGUMENT), unlike the
 * mechanical ml_model_create, which follows FreeSystem::new in accepting zero unknowns. */
ml_status ml_thermal_create(int64_t n_nodes, const double* nodes, int64_t n_elements, int nen,
                            const int64_t* conn, const double* conductivity, const double* rho_c,
                            int64_t nconv, const int64_t* conv_elements, const int64_t* conv_faces,
                            const double* film, const double* sink, const uint8_t* fixed,
                            const double* imposed, ml_precision precision, ml_thermal** out,
                            ml_solve_info* info);
/* Model::assemble's thermal F (model.rs:1089-1127) after the nodal heat loads already in
 * f[n_nodes]: the heat sources (element, q per unit volume; element_heat_source,
 * assembly.rs:465-484) in order, then the convection loads h T_inf int N, faces grouped per
 * element in element order. Pass the same convection faces as to ml_thermal_create. Host. */
ml_status ml_thermal_add_loads(ml_thermal* h, int64_t nsources, const int64_t* elements,
                               const double* values, int64_t nconv, const int64_t* conv_elements,
                               const int64_t* conv_faces, const double* film, const double* sink,
                               double* f);
/* solve_step_with (model.rs:628-706): `increments` equal steps of time_period with the theta
 * method, theta in [0.5, 1] (Step::with_theta, model.rs:152-159): (C/dt + theta K) T_{n+1} =
 * (C/dt - (1 - theta) K) T_n + F, the left matrix assembled and factored (IC(0), shift ladder) on
 * the device once per (dt, theta), each step a device PCG solve. initial[n_nodes]: the starting
 * field (the BC temperatures replace it at the fixed nodes). Outputs for increments 0..n, host
 * arrays: temperatures[(n+1) x n_nodes]; reactions[(n+1) x n_nodes] (C dT/dt + K(theta T_{n+1} +
 * (1 - theta) T_n) - F on the fixed nodes) or null; fluxes[(n+1) x n_elements x ngp x 6] (x, y,
 * z, qx, qy, qz with q = -k grad T at the Gauss points, element.rs:478-499) or null.
 * info->iterations sums the PCG iterations of the steps. */
ml_status ml_thermal_solve_step(ml_thermal* h, double time_period, int64_t increments,
                                double theta, const double* f, const double* initial,
                                double* temperatures, double* reactions, double* fluxes,
                                ml_solve_info* info);
/* Model::solve for the thermal model (Analysis::Thermal, model.rs:548-610): K T = F with K the
 * conduction and convection matrices, the BC temperatures at the fixed nodes, on the device
 * (the handle's left matrix becomes K and is factored; a later ml_thermal_solve_step refactors).
 * temperatures[n_nodes]; reactions[n_nodes] (K T - F on the fixed nodes) and
 * fluxes[n_elements x ngp x 6] may be null. */
ml_status ml_thermal_solve_steady(ml_thermal* h, const double* f, double* temperatures,
                                  double* reactions, double* fluxes, ml_solve_info* info);
void ml_thermal_destroy(ml_thermal* h);

/* ------------------------------------------------------------------ batched solves
 * (no single counterpart in the reference: the equivalent of sequential Model::solve /
 * Pcg::solve calls, model.rs:548 and iterative.rs:210, run together) */

/* nrhs load cases of one model: f[k*ndof ..] nodal loads of case k, u likewise; infos[nrhs]
 * (optional) per case. The cases are solved together as ml_pcg_solve does with nrhs > 1;
 * reactions are not computed. A model without free DOFs is refused (ML_INVALID_ARGUMENT). */
ml_status ml_model_solve_many(ml_model* m, int64_t nrhs, const double* f, double* u,
                              ml_solve_info* infos);

typedef struct ml_batch ml_batch;
/* Several models of any sizes (already created; they must outlive the batch and keep their
 * stiffness while it is used) solved in one set of launches: their K_ff are concatenated
 * block-diagonally into one operator, each model climbs its own shift ladder in shared IC(0)
 * launches, and one PCG per model runs with its own scalars. precision: ML_FP64,
 * ML_MIXED_IC0 or ML_MIXED_JACOBI (the FP32 factor applied by Jacobi sweeps over the whole
 * block-diagonal operator, which couple no two models). */
ml_status ml_batch_create(int64_t count, ml_model* const* models, ml_precision precision,
                          ml_batch** out);
/* f[k], u[k]: host arrays of 3*n_nodes values of model k; infos[count] optional. The
 * displacements also become each model's last solution (ml_model_compliance_gradient). */
ml_status ml_batch_solve(ml_batch* b, const double* const* f, double* const* u,
                         ml_solve_info* infos);
void ml_batch_destroy(ml_batch* b);

/* Autotuner (no counterpart in the reference): times the triangular solves of this handle's
 * factor on candidate grids (waves per CU) and, for nrhs > 1, the number of right-hand sides
 * per batched pass, and keeps the fastest for later solves. Results do not change (the
 * solves are bit-exact under any schedule). chosen[3] (optional): waves per CU, right-hand
 * sides per pass, reported occupancy per CU. Factorises first if needed. */
ml_status ml_pcg_autotune(ml_pcg* h, int64_t nrhs, int32_t* chosen);
ml_status ml_model_autotune(ml_model* m, int64_t nrhs, int32_t* chosen);
/* No counterpart in the reference: lanes per row of the batched SpMV and Jacobi sweeps over
 * several right-hand sides (8, 16, 32 or 64: a wave takes 64 / lanes rows), as the autotuner
 * timed fastest for nrhs > 1 at its right-hand sides per pass, or as set here; 0 (before
 * either) when the width follows the vectors of a pass (up to a wave). Every width gives
 * bit-identical results. */
ml_status ml_pcg_set_batch_lanes(ml_pcg* h, int lanes);
int ml_pcg_batch_lanes(const ml_pcg* h);
ml_status ml_model_set_batch_lanes(ml_model* m, int lanes);
int ml_model_batch_lanes(const ml_model* m);

/* Kernel-duration accounting with HIP events on the model's stream (no counterpart in the
 * reference; for benchmarks). Returns the accumulated milliseconds and launch counts per kind
 * (0 SpMV FP64, 1 SpMV FP32, 2 preconditioner FP64 (forward + backward solve), 3
 * preconditioner FP32, 4 IC(0), 5 element stiffness or elasto-plastic response, 6 gather
 * assembly, 7 PCG vector updates,
 * 8 Gauss-point stresses, 9 nodal stress averaging) into ms[10] and counts[10] (either may be
 * null), then resets and enables (1) or disables (0) accounting; enable < 0 only reads. */
ml_status ml_model_profile(ml_model* m, int enable, double* ms, int64_t* counts);

/* ------------------------------------------------------------------ Gmsh .msh input */
/* The mesh half of maille-io's input (crates/maille-io/src/msh): a Gmsh 4.1 ASCII file read into
 * nodes, cells in Maille's numbering (Tet10 mid-nodes 8 and 9 swapped, numbering.rs:11-15) with
 * positive orientation (numbering.rs:20-46), and the physical groups as named sets. Host only;
 * no device is needed. */
typedef struct ml_msh ml_msh;

/* Refusals of ml_msh_read, one per reader variant of IoError (error.rs:9-65). */
typedef enum ml_msh_error {
  ML_MSH_SYNTAX = 1,                   /* IoError::Syntax: malformed line */
  ML_MSH_MISSING_SECTION = 2,          /* IoError::MissingSection */
  ML_MSH_UNSUPPORTED_VERSION = 3,      /* IoError::UnsupportedVersion: not 4.1 */
  ML_MSH_BINARY = 4,                   /* IoError::BinaryFormat */
  ML_MSH_UNSUPPORTED_ELEMENT = 5,      /* IoError::UnsupportedElement: a Gmsh type not read */
  ML_MSH_NO_CELLS = 6,                 /* IoError::NoCells: no triangle and no tetrahedron */
  ML_MSH_MIXED_CELL_KINDS = 7,         /* IoError::MixedCellKinds: two element orders */
  ML_MSH_NON_PLANAR = 8,               /* IoError::NonPlanar: triangles off z = 0 */
  ML_MSH_DUPLICATE_SET_NAME = 9,       /* IoError::DuplicateSetName */
  ML_MSH_DUPLICATE_NODE_TAG = 10,      /* IoError::DuplicateNodeTag */
  ML_MSH_UNKNOWN_NODE = 11,            /* IoError::UnknownNode */
  ML_MSH_BOUNDARY_FACE_NOT_FOUND = 12, /* IoError::BoundaryFaceNotFound */
  ML_MSH_BOUNDARY_EDGE_NOT_FOUND = 13  /* IoError::BoundaryEdgeNotFound */
} ml_msh_error;

/* msh::read (reader.rs:23-39) of text[length] (no terminating NUL needed). On a refused file
 * returns ML_INVALID_FILE with *error_kind (ml_msh_error) and *error_line (1-based, 0 when the
 * refusal has no line) set and the message in ml_last_error(); error_kind and error_line may be
 * null. */
ml_status ml_msh_read(const char* text, int64_t length, ml_msh** out, int32_t* error_kind,
                      int64_t* error_line);
/* sizes[5]: nodes, nodes per cell (3 Tri3, 6 Tri6, 4 Tet4, 10 Tet10), cells, sets, dimension. */
ml_status ml_msh_sizes(const ml_msh* m, int64_t* sizes);
/* MshFile::nodes and cells (mod.rs:111-120): xyz[3*nodes], cells[cells*nodes_per_cell]; either
 * pointer may be null. */
ml_status ml_msh_geometry(const ml_msh* m, double* xyz, int64_t* cells);
/* Set k of MshFile::sets (mod.rs:97-109), sets in byte order of their names: counts[5] = nodes,
 * faces, edges, elements, name length in bytes; the name (UTF-8, as written between the quotes)
 * is copied with a terminating NUL into name[capacity] when name is not null, and
 * ML_INVALID_ARGUMENT is returned when it does not fit. */
ml_status ml_msh_set_info(const ml_msh* m, int64_t k, char* name, int64_t capacity,
                          int64_t* counts);
/* The sorted lists of set k: nodes[], faces[2*count] and edges[2*count] as (element, local
 * index) pairs (faces as Element::face_nodes, edges as Element::edge_nodes), elements[]; any
 * pointer may be null. */
ml_status ml_msh_set_items(const ml_msh* m, int64_t k, int64_t* nodes, int64_t* faces,
                           int64_t* edges, int64_t* elements);
void ml_msh_destroy(ml_msh* m);

#ifdef __cplusplus
}
#endif
#endif /* MAILLE_HIP_H */
