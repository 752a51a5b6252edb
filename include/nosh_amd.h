/* nosh_amd.h -- C ABI of the MI355X (gfx950) NLS Newton-Krylov library.
 *
 * Replaces the operator layer of nosh (the mesh edge cache, the
 * explicit_values magnetic field, the parameter matrices keo / DkeoDP, the
 * jacobian_operator, keo_regularized and the model_evaluator::nls) with
 * HIP-resident objects behind opaque handles, plus the in-repo caller
 * (Newton-MINRES) that replaces NOX/Belos.
 *
 * Conventions
 *  - Every function returns NOSH_OK (0) or an error code; nosh_last_error()
 *    gives the message of the last failure on the calling thread. No C++
 *    exception crosses this boundary.
 *  - Vectors are real R^{2n}, interleaved complex: x[2k] = Re psi_k,
 *    x[2k+1] = Im psi_k, in the vertex order of the caller's mesh ("reference
 *    order"; src/mesh.cpp:595-606). The bandwidth-reducing vertex permutation
 *    (reverse Cuthill-McKee) is internal.
 *  - Host-pointer variants copy through PCIe and synchronise. *_device
 *    variants take device pointers in reference order and a hipStream_t
 *    (NULL = the library stream) and are asynchronous.
 *  - Parameters are passed as (count, names, values), the C form of the
 *    reference's std::map<std::string,double>; a missing required name returns
 *    NOSH_ERR_MISSING_PARAM (the reference throws std::out_of_range from
 *    params.at(), e.g. src/jacobian_operator.cpp:156).
 *  - Operator applies support only mode = NOSH_NO_TRANS, alpha = 1, beta = 0,
 *    as the reference (src/jacobian_operator.cpp:48-59,
 *    src/keo_regularized.cpp:98-100); anything else returns
 *    NOSH_ERR_UNSUPPORTED. Y is overwritten. Multi-column X is applied column
 *    by column (src/jacobian_operator.cpp:80).
 *  - Threading: apply calls on one handle may run concurrently on distinct
 *    streams once its rebuild/set_parameters has completed; rebuild,
 *    set_parameters and destroy must be externally synchronised.
 */
#ifndef NOSH_AMD_H
#define NOSH_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NOSH_OK 0
#define NOSH_ERR_INVALID 1
#define NOSH_ERR_UNSUPPORTED 2
#define NOSH_ERR_MISSING_PARAM 3
#define NOSH_ERR_HIP 4
#define NOSH_ERR_NOT_BUILT 5
#define NOSH_ERR_NOT_CONVERGED 6

#define NOSH_NO_TRANS 0
#define NOSH_TRANS 1
#define NOSH_CONJ_TRANS 2

typedef struct nosh_mesh_s* nosh_mesh_t;
typedef struct nosh_vfield_s* nosh_vfield_t;
typedef struct nosh_sfield_s* nosh_sfield_t;
typedef struct nosh_keo_s* nosh_keo_t;
typedef struct nosh_jac_s* nosh_jac_t;
typedef struct nosh_prec_s* nosh_prec_t;
typedef struct nosh_nls_s* nosh_nls_t;

/* Message of the last error on this thread ("" if none). */
const char* nosh_last_error(void);

/* ---------------------------------------------------------------- mesh ---
 * Builds the edge/covolume cache on the device: edges numbered by first
 * appearance over the cells in order, local edges in MOAB canonical order,
 * endpoints (lower id, higher id) (nosh::read + mesh::build_entity_relations_,
 * src/mesh_reader.cpp:117-125, src/mesh.cpp:629-691); edge lengths and
 * covolumes (mesh_tri/mesh_tetra::compute_edge_data_,
 * src/mesh_tri.cpp:44-151, src/mesh_tetra.cpp:41-160); control volumes
 * (src/mesh_tri.cpp:153-210, src/mesh_tetra.cpp:162-271); and the CSR edge
 * graph, RCM-reordered when reorder != 0 (src/mesh.cpp:787-881).
 * coords: num_vertices x 3; cells: num_cells x vertices_per_cell (3 or 4). */
int nosh_mesh_create(const double* coords, int64_t num_vertices, const int32_t* cells,
                     int64_t num_cells, int vertices_per_cell, int reorder, nosh_mesh_t* out);
int nosh_mesh_destroy(nosh_mesh_t mesh);
int nosh_mesh_sizes(nosh_mesh_t mesh, int64_t* num_vertices, int64_t* num_edges,
                    int64_t* num_cells);
/* mesh::my_edges() (src/mesh.hpp:103-107): num_edges x 2 (lower, higher). */
int nosh_mesh_edges(nosh_mesh_t mesh, int32_t* edges);
/* relations_.cell_edges (src/mesh.hpp:259 area): num_cells x (3|6), ascending. */
int nosh_mesh_cell_edges(nosh_mesh_t mesh, int32_t* cell_edges);
/* mesh::get_edge_data() (src/mesh.hpp:41-44,192-194). */
int nosh_mesh_edge_data(nosh_mesh_t mesh, double* length, double* covolume);
/* mesh::control_volumes() (src/mesh.hpp:188-190). */
int nosh_mesh_control_volumes(nosh_mesh_t mesh, double* control_volumes);
/* Internal vertex order: perm[i] = reference vertex stored at internal slot i. */
int nosh_mesh_permutation(nosh_mesh_t mesh, int32_t* perm);
/* Internal sparse layout: padded sliced-ELLPACK slots, slices of 64 rows,
 * widest row. */
int nosh_mesh_layout(nosh_mesh_t mesh, int64_t* nnz_pad, int64_t* nslices, int* max_width);
/* The split layout the operator kernels read: padded slots of the upper part
 * (column above the row: column + the value) and of the lower part (column +
 * 4 B position of the mirrored upper slot, whose conjugate value it uses);
 * the column is a 16- or 32-bit index (nosh_mesh_layout_column_bytes). */
int nosh_mesh_layout_split(nosh_mesh_t mesh, int64_t* nnz_upper, int64_t* nnz_lower);
/* Bytes per column index the split layout's kernels read: 2 when every
 * column fits a 16-bit offset from its row (the reordered graph's bandwidth
 * is below 32768), else 4. No counterpart in the reference (its Tpetra
 * matrices hold 32- or 64-bit global indices); a measurement helper. */
int nosh_mesh_layout_column_bytes(nosh_mesh_t mesh, int* bytes);
/* Device vectors between reference order and the internal order (stream
 * may be NULL). Callers that keep state on the device in internal order use
 * the *_apply_internal entry points and skip the permutation per apply. */
int nosh_mesh_to_internal(nosh_mesh_t mesh, const double* x_ref, double* x_int, void* stream);
int nosh_mesh_to_reference(nosh_mesh_t mesh, const double* x_int, double* x_ref, void* stream);

/* -------------------------------------------------------------- fields ---
 * vector_field::explicit_values(mesh, "A", mu)
 * (src/vector_field_explicit_values.cpp:13-50); A: num_vertices x 3. The
 * only scalar parameter is "mu" (:56-69). */
int nosh_vfield_explicit_values_create(nosh_mesh_t mesh, const double* A, double mu,
                                       nosh_vfield_t* out);
/* get_edge_projection(k) = mu * 0.5 (A_lo + A_hi).(x_lo - x_hi) for all edges
 * at the field's current mu (src/vector_field_explicit_values.cpp:71-76). */
int nosh_vfield_edge_projections(nosh_vfield_t field, double* out);
int nosh_vfield_destroy(nosh_vfield_t field);
/* scalar_field::constant(mesh, c, param_name, param_init)
 * (src/scalar_field_constant.cpp:14-75): value c (+ params[param_name] if that
 * name is given and present). param_name may be NULL or "". */
int nosh_sfield_constant_create(nosh_mesh_t mesh, double c, const char* param_name,
                                double param_init, nosh_sfield_t* out);
/* scalar_field::explicit_values (src/scalar_field_explicit_values.cpp:33-57). */
int nosh_sfield_explicit_values_create(nosh_mesh_t mesh, const double* values,
                                       nosh_sfield_t* out);
int nosh_sfield_destroy(nosh_sfield_t field);

/* ------------------------------------------------------------ keo / dK ---
 * parameter_matrix::keo(mesh, thickness, mvp)
 * (src/parameter_matrix_keo.cpp:25-42). */
int nosh_keo_create(nosh_mesh_t mesh, nosh_sfield_t thickness, nosh_vfield_t mvp,
                    nosh_keo_t* out);
/* parameter_matrix::DkeoDP(mesh, thickness, mvp, param_name)
 * (src/parameter_matrix_dkeo_dp.cpp:60-155): dK/d(param_name). */
int nosh_dkeo_dp_create(nosh_mesh_t mesh, nosh_sfield_t thickness, nosh_vfield_t mvp,
                        const char* param_name, nosh_keo_t* out);
/* parameter_object::set_parameters (src/parameter_object.cpp:6-47): needs
 * "mu"; re-assembles (keo::refill_, src/parameter_matrix_keo.cpp:74-184) only
 * when mu changed. */
int nosh_keo_set_parameters(nosh_keo_t keo, int nparams, const char* const* names,
                            const double* values);
/* Tpetra::CrsMatrix::apply on the KEO (called at
 * src/model_evaluator_nls.cpp:537, src/jacobian_operator.cpp:65). x, y:
 * ncols columns of 2n values with leading dimension ld (>= 2n). */
int nosh_keo_apply(nosh_keo_t keo, const double* x, double* y, int ncols, int64_t ld, int mode,
                   double alpha, double beta);
int nosh_keo_apply_device(nosh_keo_t keo, const double* x, double* y, void* stream);
/* Device pointers in internal order. */
int nosh_keo_apply_internal(nosh_keo_t keo, const double* x, double* y, void* stream);
int nosh_keo_destroy(nosh_keo_t keo);

/* --------------------------------------------------- jacobian_operator ---
 * jacobian_operator(mesh, scalar_potential, thickness, keo)
 * (src/jacobian_operator.hpp:29-34). The keo handle is shared, as in the
 * reference (src/model_evaluator_nls.cpp:257-266). */
int nosh_jac_create(nosh_mesh_t mesh, nosh_sfield_t scalar_potential, nosh_sfield_t thickness,
                    nosh_keo_t keo, nosh_jac_t* out);
/* jacobian_operator::rebuild(params, psi) (src/jacobian_operator.cpp:110-200):
 * needs "g" and "mu". psi: 2n values (host). */
int nosh_jac_rebuild(nosh_jac_t jac, int nparams, const char* const* names,
                     const double* values, const double* psi);
/* jacobian_operator::apply (src/jacobian_operator.cpp:38-108). */
int nosh_jac_apply(nosh_jac_t jac, const double* x, double* y, int ncols, int64_t ld, int mode,
                   double alpha, double beta);
int nosh_jac_apply_device(nosh_jac_t jac, const double* x, double* y, void* stream);
/* Device pointers in internal order; the _fp32 form applies the FP32 copy of
 * the operator (values and blocks rounded to float) to FP32 vectors. */
int nosh_jac_apply_internal(nosh_jac_t jac, const double* x, double* y, void* stream);
int nosh_jac_apply_internal_fp32(nosh_jac_t jac, const float* x, float* y, void* stream);
/* Measurement helper, no reference counterpart: the average duration of the
 * internal-order J.phi kernel (FP64, or FP32 with fp32 != 0; x and y device
 * vectors of that type) over `reps` launches on `stream`, each timed by start
 * and end events that the runtime stamps on the kernel dispatch itself, so
 * launch gaps do not count. cold != 0: before each launch, a read of a 1 GiB
 * device buffer (allocated by the call) replaces the contents of the L2s and
 * the 256 MB Infinity Cache with clean lines, so the kernel streams its
 * operands from HBM. Two untimed launches come first. */
int nosh_jac_bench_apply(nosh_jac_t jac, const void* x, void* y, int fp32, int cold, int reps,
                         void* stream, double* seconds_per_launch);
int nosh_jac_destroy(nosh_jac_t jac);

/* ----------------------------------------------------- keo_regularized ---
 * keo_regularized(mesh, thickness, mvp) (src/keo_regularized.cpp:39-66).
 * Its inverse, which the reference approximates with one MueLu AMG V-cycle
 * (:106-108, :290-336), is a Chebyshev iteration with 2x2 block-Jacobi
 * scaling here: `iters` iterations (iters-1 products with P) on the
 * interval [upper/ratio, upper] of D^{-1}P; upper <= 0 selects the bound
 * 1 + max_k sum_{j!=k}|K_kj| / lambda_min(D_k) (<= 2 for nonnegative edge
 * weights). Defaults: iters 6, ratio 300, upper 0. */
int nosh_prec_create(nosh_mesh_t mesh, nosh_sfield_t thickness, nosh_vfield_t mvp,
                     nosh_prec_t* out);
int nosh_prec_set_chebyshev(nosh_prec_t prec, int iters, double ratio, double upper);
/* keo_regularized::rebuild(params, psi) (src/keo_regularized.cpp:181-264):
 * P = K + 2x2 blocks from g*cv*t*psi (added only when g > 0). */
int nosh_prec_rebuild(nosh_prec_t prec, int nparams, const char* const* names,
                      const double* values, const double* psi);
/* keo_regularized::apply: y = M^{-1} x (the approximate inverse). */
int nosh_prec_apply(nosh_prec_t prec, const double* x, double* y, int ncols, int64_t ld,
                    int mode, double alpha, double beta);
int nosh_prec_apply_device(nosh_prec_t prec, const double* x, double* y, void* stream);
/* y = P x, the assembled regularized KEO itself. */
int nosh_prec_apply_operator(nosh_prec_t prec, const double* x, double* y);
/* Spectral upper bound in use after the last rebuild. */
int nosh_prec_upper_bound(nosh_prec_t prec, double* upper);
int nosh_prec_destroy(nosh_prec_t prec);

/* ------------------------------------------------- model_evaluator::nls ---
 * nls(mesh, mvp, scalar_potential, g, thickness, initial_x, deriv_parameter)
 * (src/model_evaluator_nls.hpp:62-70, src/model_evaluator_nls.cpp:56-138).
 * initial_x may be NULL (psi = 1). Parameter names are the sorted union
 * {"g"} + potential params + {"mu"} (src/model_evaluator_nls.cpp:102-111). */
int nosh_nls_create(nosh_mesh_t mesh, nosh_vfield_t mvp, nosh_sfield_t scalar_potential,
                    double g, nosh_sfield_t thickness, const double* initial_x,
                    const char* deriv_parameter, nosh_nls_t* out);
/* get_p_names(0) / nominal p (src/model_evaluator_nls.cpp:196-205). */
int nosh_nls_num_params(nosh_nls_t nls, int* count);
int nosh_nls_param_name(nosh_nls_t nls, int index, char* buf, int buflen);
int nosh_nls_param_nominal(nosh_nls_t nls, int index, double* value);
/* evalModel with out_args.f (compute_f_, src/model_evaluator_nls.cpp:527-628). */
int nosh_nls_eval_f(nosh_nls_t nls, const double* x, int nparams, const char* const* names,
                    const double* values, double* f);
int nosh_nls_eval_f_device(nosh_nls_t nls, const double* x, int nparams,
                           const char* const* names, const double* values, double* f,
                           void* stream);
/* The same with x and f device vectors in internal order (no permutation),
 * for device-resident callers that keep state in internal order. */
int nosh_nls_eval_f_internal(nosh_nls_t nls, const double* x, int nparams,
                             const char* const* names, const double* values, double* f,
                             void* stream);
/* Column `col` of DfDp, in get_p_names order (computeDFDP_,
 * src/model_evaluator_nls.cpp:630-695, including its use of the dK/dp of the
 * model's deriv_parameter for every column). */
int nosh_nls_eval_dfdp(nosh_nls_t nls, const double* x, int nparams, const char* const* names,
                       const double* values, int col, double* dfdp);
/* create_W_op (src/model_evaluator_nls.cpp:253-267) and create_W_prec
 * (:301-314); evalModel with W_op / W_prec rebuilds them (:491-522). */
int nosh_nls_create_w_op(nosh_nls_t nls, nosh_jac_t* out);
int nosh_nls_create_w_prec(nosh_nls_t nls, nosh_prec_t* out);
int nosh_nls_destroy(nosh_nls_t nls);

/* --------------------------------------------------------- solvers -------
 * The in-repo replacement of the Belos linear solver and the NOX full-step
 * Newton (examples/conf.xml:77-190). The reference's example selects Pseudo
 * Block GMRES (examples/conf.xml:104; the MINRES line :105 is commented
 * out, its parameters :117-123 remain); J is real symmetric and
 * keo_regularized SPD, so MINRES is the short-recurrence Krylov method for
 * the same systems, with the same tolerance 1e-10 and limit 1000. */
typedef struct {
  double tol;           /* stop when ||F(psi)||_2 <= tol (default 1e-10) */
  int max_iters;        /* Newton steps (default 30) */
  double minres_rtol;   /* MINRES relative tolerance in the M^{-1} norm (1e-10) */
  int minres_max_iters; /* (default 1000) */
  int cheb_iters;       /* Chebyshev iterations of the preconditioner, 0 = none (6) */
  double cheb_ratio;    /* interval ratio (300) */
  int precision;        /* 0 = FP64 throughout; 1 = FP32 inner MINRES / FP64 outer */
  int forcing;          /* 0 = constant minres_rtol; 1 = Eisenstat-Walker choice 2 */
  double forcing_min;   /* lower bound of the forcing term (mixed: 1e-6) */
  double forcing_max;   /* upper bound of the forcing term (0.1) */
} nosh_newton_options;

typedef struct {
  int newton_iters;
  int minres_iters_total;
  int converged;
  double final_norm_f;
  double seconds;
  double norm_f_history[64];
  int minres_history[64];
} nosh_newton_stats;

/* Fills *opts with the defaults listed above. */
int nosh_newton_default_options(nosh_newton_options* opts);
/* Newton-MINRES on F(psi) = 0 at the given parameters, preconditioned by
 * keo_regularized. psi: 2n values (host), overwritten by the solution. */
int nosh_newton_solve(nosh_nls_t nls, double* psi, int nparams, const char* const* names,
                      const double* values, const nosh_newton_options* opts,
                      nosh_newton_stats* stats);
int nosh_newton_solve_device(nosh_nls_t nls, double* psi, int nparams,
                             const char* const* names, const double* values,
                             const nosh_newton_options* opts, nosh_newton_stats* stats,
                             void* stream);
/* MINRES on J x = b (x starts at 0) with preconditioner prec (NULL = none).
 * Both operators must have been rebuilt. */
int nosh_minres_solve(nosh_jac_t jac, nosh_prec_t prec, const double* b, double* x,
                      double rtol, int max_iters, int* iters, double* relres);

/* ---------------------------------------------------------- batches ------
 * No counterpart in the reference (which solves one instance per process):
 * many Newton-MINRES solves on meshes of uneven size advanced in lockstep,
 * one kernel launch per phase for all of them. Group k is the model
 * models[k] (mesh, A, V, thickness) with counts[k] (1..64) instances; mu and
 * g hold one value per instance, group-major. The inner MINRES is FP32
 * (vectors, blocks, on-the-fly KEO values) with FP64 scalars, F and update. */
typedef struct nosh_batch_s* nosh_batch_t;

typedef struct {
  int newton_iters;
  int minres_iters;
  int converged;
  double final_norm_f;
} nosh_batch_instance_stats;

int nosh_batch_create(int ngroups, const nosh_nls_t* models, const int* counts, const double* mu,
                      const double* g, nosh_batch_t* out);
/* block_threads: threads per block (multiple of 64 and of the lane count);
 * fold: row tiles (block_threads / lanes rows each) per launched block, 1..64
 * (larger folds give fewer per-block partial sums to reduce);
 * mode 0: each MINRES iteration is a captured hipGraph replayed graph_iters
 * at a time, mode 1: one persistent cooperative kernel per MINRES solve
 * (only for grids of at most 64 launched blocks; larger ones get
 * NOSH_ERR_UNSUPPORTED and the batch stays in mode 0), mode 2: as mode 0
 * with every thread of the gathering kernels taking 2 lanes of a row (every
 * group needs at least 2 lanes, else NOSH_ERR_UNSUPPORTED and mode 0); the
 * modes give the same results bit for bit;
 * poll: iterations between host checks of the done flags (mode 0). */
int nosh_batch_set_tuning(nosh_batch_t batch, int block_threads, int fold, int mode,
                          int graph_iters, int poll);
/* Times every (block_threads, fold, mode) candidate on one Newton step from
 * psi and keeps the fastest; with use_cache the choice is read from /
 * appended to $NOSH_AUTOTUNE_CACHE (default ~/.cache/nosh_amd_autotune.tsv). */
int nosh_batch_autotune(nosh_batch_t batch, const double* psi, const nosh_newton_options* opts,
                        int use_cache, int* best_threads, int* best_fold, int* best_mode);
/* Measurement (no counterpart in the reference): average duration of one
 * batched Chebyshev step (the kernel the batched MINRES spends most time in),
 * every instance active, rebuilt at the batch's current states, timed with
 * HIP events on the library stream over reps launches (step 2, or step 1 at
 * Chebyshev degree 2); bytes_per_launch (nullable) receives its algorithmic
 * bytes: 12 per padded slot (packed column, alpha, a~, shared by the lanes)
 * + 44 per row and lane (z_in, r, z_{j-2} read and overwritten by z_j, the
 * P block; FP32; 36 at step 1). Resets the MINRES state. */
int nosh_batch_time_cheb_step(nosh_batch_t batch, int reps, double* us_per_launch,
                              double* bytes_per_launch);
/* psi: all instances concatenated (each 2 n_k values, reference order);
 * overwritten by the solutions. stats: one record per instance (may be NULL).
 * A batch of 16 or more instances compacts: once at most half of them are
 * still iterating, those continue in a batch of their own (fewer lanes
 * per row, hipGraph mode (2 lanes per thread if the batch uses them and the
 * sub-batch has the lanes), the same block size and fold), which compacts in
 * turn down to 4 instances; their Eisenstat-Walker forcing restarts there. */
int nosh_batch_solve(nosh_batch_t batch, double* psi, const nosh_newton_options* opts,
                     nosh_batch_instance_stats* stats);
int nosh_batch_solve_device(nosh_batch_t batch, double* psi, const nosh_newton_options* opts,
                            nosh_batch_instance_stats* stats, void* stream);
int nosh_batch_destroy(nosh_batch_t batch);

/* ------------------------------------------------------ MLP surrogate ----
 * No counterpart in the reference: the per-vertex network that seeds Newton
 * during continuation. x (n_in) -> [Linear(width) + GELU] x depth ->
 * Linear(n_out); GELU in its tanh form. BF16 weights and activations,
 * FP32 accumulation (MFMA), FP32 master weights, Adam. Parameters: all
 * weights (layer by layer, each [out][in] row-major) then all biases. */
typedef struct nosh_mlp_s* nosh_mlp_t;

int nosh_mlp_create(int n_in, int width, int depth, int n_out, uint64_t seed, nosh_mlp_t* out);
int nosh_mlp_num_params(nosh_mlp_t mlp, int64_t* nparams, int64_t* nweights);
int nosh_mlp_get_params(nosh_mlp_t mlp, float* params);
int nosh_mlp_set_params(nosh_mlp_t mlp, const float* params);
/* X: M x n_in, Y: M x n_out (host). */
int nosh_mlp_forward(nosh_mlp_t mlp, const float* X, int64_t M, float* Y);
int nosh_mlp_forward_device(nosh_mlp_t mlp, const float* X, int64_t M, float* Y, void* stream);
/* Inference only: the same outputs as nosh_mlp_forward from one fused kernel
 * per 128 rows (activations in LDS, hidden layers on MFMA) when width is 64,
 * 128 or 256 and n_in <= 32; otherwise the layer-by-layer path. */
int nosh_mlp_infer(nosh_mlp_t mlp, const float* X, int64_t M, float* Y);
int nosh_mlp_infer_device(nosh_mlp_t mlp, const float* X, int64_t M, float* Y, void* stream);
/* Mean-squared-error loss over the M x n_out outputs against T and its
 * gradient (nparams values, parameter layout). grad may be NULL. */
int nosh_mlp_gradients(nosh_mlp_t mlp, const float* X, const float* T, int64_t M, float* grad,
                       double* loss);
/* One Adam step (lr; beta 0.9 / 0.999, eps 1e-8) on device minibatch X, T. */
int nosh_mlp_train_step_device(nosh_mlp_t mlp, const float* X, const float* T, int64_t M,
                               float lr, double* loss, void* stream);
int nosh_mlp_destroy(nosh_mlp_t mlp);

/* ------------------------------------------------ energy functionals -----
 * The functionals nls::gibbs_energy and nls::inner_product compute but do
 * not return (src/model_evaluator_nls.cpp:697-770; they return 0.0): the
 * intended values -sum_k c_k |psi_k|^4 / sum_k c_k and
 * sum_k c_k Re(conj(phi_k) psi_k) / sum_k c_k. */
int nosh_nls_gibbs_energy(nosh_nls_t nls, const double* psi, double* energy);
int nosh_nls_inner_product(nosh_nls_t nls, const double* phi, const double* psi, double* value);
/* The Gibbs energy without any allocation. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). psi: 2n
 * doubles of device memory in reference vertex order; energy: one double of
 * device memory, written on `stream` (the caller synchronizes). The partial
 * sums live in a fixed device array of the library (256 KiB per GPU, kept
 * until the process ends), so meshes of more than 4,194,304 vertices are
 * refused (NOSH_ERR_INVALID), and calls that overlap on one GPU race. Same
 * value, bit for bit, as nosh_nls_gibbs_energy. */
int nosh_nls_gibbs_energy_noalloc(nosh_nls_t nls, const double* psi, double* energy,
                                  void* stream);
/* ------------------------------------------------ measurement helper -----
 * No reference counterpart. Average seconds per launch of the library's own
 * streaming kernel `kind` (0 read, 1 write, 2 copy) over buffers of `bytes`
 * bytes (16-byte vectors, 256-thread workgroups, blocks_per_cu per CU, 0 =
 * 8; nt != 0: non-temporal accesses), timed with events over `reps`
 * launches: the attainable HBM bandwidth the operator kernels are measured
 * against (bytes, or 2 bytes for copy, per launch). */
int nosh_bench_stream(int kind, int64_t bytes, int reps, int blocks_per_cu, int nt,
                      double* seconds_per_launch);


/* ------------------------------------------------------ continuation -----
 * Natural-parameter continuation, the in-repo stand-in for the LOCA stepper
 * the reference drives through mikado (examples/conf.xml:35-75): the
 * parameter goes from start to stop in `steps` equal steps; each step
 * predicts (previous solution, tangent J t = -dF/dp, or secant) and corrects
 * with Newton-MINRES; a failed corrector halves the step, up to
 * max_halvings times (LOCA's adaptive step with a minimum size). */
typedef struct {
  const char* parameter; /* "mu" or "g" (default "mu") */
  double start, stop;
  int steps;
  int predictor;    /* 0 previous solution, 1 tangent (default), 2 secant */
  int max_halvings; /* default 4 */
  nosh_newton_options newton;
} nosh_continuation_options;

typedef struct {
  double param;       /* parameter value of this accepted grid point */
  int newton_iters;   /* Newton steps of the corrector (sum over sub-steps) */
  int minres_iters;   /* MINRES iterations incl. the tangent solves */
  int halvings;       /* sub-steps needed */
  int converged;
  double norm_f;      /* ||F(psi)||_2 at the accepted point */
  double gibbs;       /* Gibbs energy (see nosh_nls_gibbs_energy) */
  double seconds;     /* wall time of this step (predictor + corrector) */
  int predictor_used; /* 0 previous, 1 tangent, 2 secant, 3 surrogate */
  int fallback;       /* 1 if the surrogate guess was rejected for the tangent */
  double pred_error;  /* ||psi_surrogate_guess - psi_solution||_2 where the surrogate was
                         evaluated, else equal to tan_error */
  double tan_error;   /* ||psi_tangent - psi_solution||_2 (surrogate runs) */
} nosh_continuation_step;

int nosh_continuation_default_options(nosh_continuation_options* opts);
/* psi: 2n host values, the initial state in, the last accepted solution out.
 * record: steps + 1 entries (entry 0 = the solve at `start`). The other
 * parameters (e.g. "g") are fixed at the given values. Returns
 * NOSH_ERR_NOT_CONVERGED if a step failed after max_halvings (record and
 * steps_done describe what was accepted). */
int nosh_continuation_run(nosh_nls_t nls, double* psi, int nparams, const char* const* names,
                          const double* values, const nosh_continuation_options* opts,
                          nosh_continuation_step* record, int* steps_done);

/* Surrogate-seeded continuation (no counterpart in the reference): the
 * tangent run above, plus a per-vertex MLP that predicts the second-order
 * correction (psi_{k+1} - psi_k - h t_k) / h^2 from (x, y[, z], Re/Im psi,
 * h Re/Im t, mu). It trains on the run's own first train_after accepted
 * steps (train_steps Adam steps on random vertex minibatches), then
 * evaluates the guess psi_k + h t_k + h^2 MLP(...) at every step and starts
 * Newton from it when its ||F|| is at most half the tangent guess's; the
 * guess is rejected for the plain tangent predictor when Newton needs more
 * than newton_budget steps or the converged Gibbs energy leaves the linear
 * extrapolation of the last two steps by more than
 * max(energy_tol, 10 |E_k - E_{k-1}|). */
typedef struct {
  int width, depth;     /* 256, 4 */
  int train_after;      /* 40 */
  int train_steps;      /* 300 */
  int batch;            /* vertices per minibatch, 16384 */
  float lr;             /* 1e-3 */
  int newton_budget;    /* 4 */
  double energy_tol;    /* 1e-8 */
  uint64_t seed;
  int retrain_steps;    /* 0; > 0: after the first training, every accepted
                           step replaces the oldest of the train_after stored
                           pairs and runs this many more Adam steps */
  int smooth_sweeps;    /* batched driver: damped Jacobi sweeps on the residual
                           of every surrogate guess before it is compared with
                           the predictor guess (default 0) */
  double accept_ratio;  /* a surrogate guess is used when its ||F|| is at most
                           this fraction of the predictor guess's (0.5) */
} nosh_surrogate_options;

typedef struct {
  int trained;
  double train_seconds;
  double final_loss;
  int steps_seeded;     /* steps that used the surrogate guess */
  int fallbacks;
} nosh_surrogate_stats;

int nosh_surrogate_default_options(nosh_surrogate_options* opts);
int nosh_continuation_run_surrogate(nosh_nls_t nls, double* psi, int nparams,
                                    const char* const* names, const double* values,
                                    const nosh_continuation_options* opts,
                                    const nosh_surrogate_options* sopts,
                                    nosh_continuation_step* record, int* steps_done,
                                    nosh_surrogate_stats* stats);

/* ---- Batched continuation: many branches stepped in mu together ----
 * Replaces, for a set of independent (mesh, g, mu_0) branches, the LOCA
 * natural-parameter stepper the reference is driven by
 * (executables/nosh-cont/nosh-cont.cpp:226-290,344; step control
 * examples/conf.xml:68-74). Instance i of `batch` runs mu from its mu_0 to
 * mu_0 + span in `steps` equal steps; every step solves all branches with
 * the lockstep batched Newton (FP32 inner MINRES, FP64 outer) from the
 * predictor guess -- predictor 1: the tangent psi_k + h dpsi/dmu, with
 * J(psi_k) dpsi/dmu = -dF/dmu solved for all branches in one lockstep FP32
 * MINRES (LOCA's "Tangent", examples/conf.xml:38-41); 2: the secant through
 * the last two states; 3: the tangent on the first step and after a step
 * that a branch did not pass from its first guess (a transition, which the
 * secant would extrapolate across), the secant otherwise -- or, once
 * trained, from the MLP surrogate guess
 * (sopts != NULL). Branches that miss are retried in lockstep: from the
 * predictor guess and from psi_k (at most 20 Newton steps each), then by
 * pseudo-transient continuation ((J + sigma c t) steps) from both; then, one
 * branch at a time, by the single-instance FP64 Newton-MINRES (inner solves
 * to minres_rtol) from the predictor guess, from psi_k, and by the FP64
 * tangent continuation over the step (up to 3 step halvings), and last by 2,
 * 4, 8 batched sub-steps of the branches that still fail. psi
 * (host, the instances' reference-order vectors concatenated) holds the
 * start states and receives the last accepted states. record has
 * steps + 1 entries (entry 0: the solve at mu_0). Statuses as for
 * nosh_continuation_run. */
typedef struct {
  double param;         /* mu - mu_0 after the step */
  int newton_iters;     /* lockstep Newton steps (all attempts) */
  int minres_iters;     /* lockstep MINRES iterations (all attempts) */
  int converged;        /* branches converged */
  int seeded;           /* branches accepted from the surrogate guess */
  int fallbacks;        /* branches re-solved after the surrogate guess */
  double max_norm_f;    /* largest final ||F|| over the branches */
  double seconds;
  double pred_error;    /* mean over branches of ||surrogate guess - psi|| where the
                           surrogate was evaluated, else equal to sec_error */
  double sec_error;     /* mean over branches of ||predictor guess - psi|| (tangent or
                           secant, as chosen) */
  int halvings;         /* step halvings needed by the branches no guess reached (FP64
                           fallback or batched sub-steps) */
  int retries;          /* branch retries by pseudo-transient continuation */
  int serial_fallbacks; /* branches handed to the single-instance FP64 fallback */
} nosh_batch_continuation_step;

int nosh_batch_continuation_run(nosh_batch_t batch, double* psi, double span, int steps,
                                int predictor, const nosh_newton_options* opts,
                                const nosh_surrogate_options* sopts,
                                nosh_batch_continuation_step* record, int* steps_done,
                                nosh_surrogate_stats* stats);

/* ------------------------------------------------ stability eigenvalues --
 * Replaces the eigensolve of executables/nosh-eig/nosh-eig.cpp:212-333
 * (Anasazi LOBPCG on the Hermitian Jacobian W = J(psi), preconditioned by
 * W_prec when "prec" is set, block size 2, 10 eigenvalues) and LOCA's
 * "Compute Eigenvalues" of examples/conf.xml:50-66: LOBPCG on the GPU with
 * the keo_regularized Chebyshev preconditioner, the Rayleigh-Ritz problems
 * solved on the host. J is real-symmetric on R^{2n}; at a solution it has
 * the phase mode i psi in its null space (J(i psi) = i F(psi)), which
 * deflate_gauge keeps out of the search space, so the reported eigenvalues
 * are those that decide stability (a negative one: unstable). */
typedef struct {
  int nev;            /* eigenpairs wanted (1..16), default 4 */
  int block_size;     /* LOBPCG block, >= nev (default nev + 2, at most 16) */
  int max_iters;      /* default 500 */
  double tol;         /* converged when ||J x - theta x|| <= tol * scale, scale the largest
                         |Ritz value| seen (an estimate of ||J||); default 1e-8 */
  int which;          /* 0 smallest algebraic (default), 1 largest algebraic */
  int deflate_gauge;  /* 1 (default): search orthogonal to i psi */
  int use_prec;       /* 1 (default): Chebyshev keo_regularized preconditioner */
  int cheb_iters;     /* default 6 */
  double cheb_ratio;  /* default 300 */
  uint64_t seed;      /* SplitMix64 seed of the start block */
} nosh_eigen_options;

typedef struct {
  int iters;          /* LOBPCG iterations */
  int converged;      /* eigenpairs converged (of nev) */
  int applies;        /* J applies */
  double max_residual;/* largest ||J x - theta x|| / scale over the nev pairs */
  double scale;       /* the ||J|| estimate used by the test */
  double seconds;
} nosh_eigen_stats;

int nosh_eigen_default_options(nosh_eigen_options* opts);
/* psi: 2n host values (reference order). evals: nev values, ascending for
 * which = 0, descending for which = 1. evecs (nullable): nev x 2n host
 * values, reference order, unit 2-norm. Returns NOSH_ERR_NOT_CONVERGED
 * when fewer than nev pairs met tol in max_iters (evals then hold the
 * current Ritz values). */
int nosh_nls_jacobian_eigenpairs(nosh_nls_t nls, const double* psi, int nparams,
                                 const char* const* names, const double* values,
                                 const nosh_eigen_options* opts, double* evals, double* evecs,
                                 nosh_eigen_stats* stats);

/* ------------------------------------------ arc-length continuation -----
 * Pseudo-arclength continuation, replacing LOCA's "Arc Length" stepper with
 * the tangent predictor and the adaptive step the reference is driven by
 * (examples/conf.xml:35-75; nosh-cont.cpp:226-290). Unknowns (psi, p) with
 * the arclength metric theta^2 <a, b> + alpha beta; theta^2 is fixed at the
 * first step so that psi and p contribute equally to the first tangent.
 * Tangent: J z = -dF/dp, (psi', p') = (z, 1) / norm, oriented along the
 * previous tangent (first: sign(p') = direction). Corrector: Newton on
 * F(psi, p) = 0, theta^2 <psi', psi - psi_0> + p' (p - p_0) - ds = 0 by
 * bordering (two MINRES solves per Newton step: J a = -F to the Newton
 * forcing term, J b = -dF/dp to border_rtol, since near a fold an inexact b
 * lets the corrector leave the branch). Step size
 * (LOCA's adaptive rule): after a step that took k Newton steps,
 * ds <- min(max_step, ds (1 + a ((N - k) / (N - 1))^2)), N = newton.max_iters,
 * a = aggressiveness; a failed corrector halves ds (below min_step: stop).
 * A step whose predictor leaves [min_value, max_value] is replaced by a
 * natural-parameter step onto the bound, which ends the run. */
typedef struct {
  const char* parameter;  /* "mu" (default) or "g" */
  double start;           /* the solve at p = start comes first */
  double min_value, max_value;
  int max_steps;
  int direction;          /* +1 (default): p increases first, -1: decreases */
  double initial_step, min_step, max_step, aggressiveness;
  int predictor;          /* 1 tangent (default), 2 secant */
  double border_rtol;     /* MINRES rtol of the tangent and dF/dp solves, default 1e-10 */
  int eigenvalues;        /* > 0: that many (<= 8) stability eigenvalues per accepted
                             step, deflated, smallest first */
  nosh_newton_options newton;
} nosh_arclength_options;

typedef struct {
  double param;           /* parameter of the accepted point */
  double step;            /* arclength step ds that was accepted */
  double param_dot;       /* parameter component p' of the unit tangent at this point (the
                             secant through it with predictor 2; for a step onto a bound,
                             the tangent its guess used) */
  int newton_iters;       /* corrector steps (all tries) */
  int minres_iters;       /* MINRES iterations (tangent + bordering, all tries) */
  int failures;           /* step halvings before acceptance */
  int converged;
  double norm_f;
  double gibbs;
  double seconds;
  int num_eigs;           /* eigenvalues stored below */
  int num_negative;       /* eigenvalues < 0 among them (unstable directions) */
  double eigs[8];
} nosh_arclength_step;

int nosh_arclength_default_options(nosh_arclength_options* opts);
/* psi: 2n host values, the start state in, the last accepted state out.
 * record: max_steps + 1 entries (entry 0 = the solve at start). states
 * (nullable): (max_steps + 1) x 2n host values receiving every accepted
 * state (reference order), the per-step solution output that
 * continuation_data_saver writes. Returns NOSH_ERR_NOT_CONVERGED when the
 * step size fell below min_step. */
int nosh_arclength_run(nosh_nls_t nls, double* psi, int nparams, const char* const* names,
                       const double* values, const nosh_arclength_options* opts,
                       nosh_arclength_step* record, double* states, int* steps_done);

#ifdef __cplusplus
}
#endif
#endif /* NOSH_AMD_H */
