/* bartels_hip.h -- C ABI of the MI355X (gfx950) linear-tetmesh hyperelastic hot path.
 *
 * Plain pointers and sizes only.  All host arrays use the reference's Eigen default
 * (column-major) layouts, so an Eigen caller passes .data() without copies:
 *   V       nV x 3  double, column-major          (V(v,c) at v + c*nV)
 *   E       nT x 4  int32,  column-major          (E(e,k) at e + k*nT)
 *   q       3*nV    double, interleaved xyz       (q[3v+c])
 *   dphidX  nT x 12 double, column-major; row e is the column-major flattening of the 4x3 matrix
 *           of shape-function gradients: dphidX(e, m + 4j) = d phi_m / d X_j
 *   volume  nT      double
 *   params  nT x 2  double, column-major; per-model meaning:
 *             BH_NEOHOOKEAN, BH_NEOHOOKEAN_LEGACY : [C, D]       = [mu/2, lambda/2]
 *             BH_STVK, BH_COROTATIONAL            : [lambda, mu]
 *             BH_ARAP                             : [k, unused]
 *           nT x 6 for BH_NEOHOOKEAN_MUSCLE: [C, D, a, d0, d1, d2] (the MATLAB gateway's params_nh,
 *           params_a and directions side by side)
 * Sparse outputs use the reference's pattern: full 3x3 blocks for every pair of vertices that
 * share a tet, explicit zeros kept, sorted int32 indices.  The pattern is symmetric, so the CSR
 * arrays returned here equal the CSC arrays of Eigen::SparseMatrix<double> (outerIndexPtr,
 * innerIndexPtr); values follow CSR order (row-major within the pattern).
 *
 * Every function returns a bh_status.  Nothing calls exit().  Device (_device) variants take
 * device pointers and a hipStream_t and are asynchronous on that stream; host variants copy
 * in/out and synchronise.
 */
#ifndef BARTELS_HIP_H
#define BARTELS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* bh_stream_t; /* == hipStream_t */

typedef enum {
  BH_OK = 0,
  BH_BAD_ARG = 1,
  BH_NONFINITE = 2,
  BH_SOLVE_FAILED = 3,
  BH_MAXIT = 4,
  BH_DEVICE_ERROR = 5,
  BH_UNSUPPORTED = 6
} bh_status;

typedef enum {
  BH_NEOHOOKEAN = 0,        /* psi = C(tr F^T F - 3 - 2 ln J) + D (J-1)^2   src/psi_neohookean_F.cpp:31 */
  BH_STVK = 1,              /* psi = mu E:E + lambda/2 (tr E)^2            src/psi_stvk_F.cpp:5-13 */
  BH_COROTATIONAL = 2,      /* psi = mu sum (s-1)^2 + lambda/2 (sum s - 3)^2  src/psi_corotational_S.cpp:5-9 */
  BH_ARAP = 3,              /* psi = k sum (s-1)^2                         src/psi_arap_S.cpp:5-11 */
  BH_NEOHOOKEAN_LEGACY = 4, /* psi = C(J^{-2/3} tr F^T F - 3) + D (J-1)^2  src/psi_neohookean_F.cpp:8-14 (commented form) */
  BH_NEOHOOKEAN_MUSCLE = 5, /* Neo-Hookean plus a muscle fibre, psi + a/2 |F d|^2 per tet: the generic_elastic
                               functor of matlab/linear_tetmesh_neohookean_muscle_{q,dq,dq2}.cpp (:53-77 for dq2)
                               with src/psi_muscle_fiber_F.cpp:5-10, src/dpsi_muscle_fiber_dF.cpp,
                               src/d2psi_muscle_fiber_dF2.cpp; params nT x 6 [C, D, a, d0, d1, d2]; not
                               available as a per-tet model id */
  BH_MODEL_PER_TET = 255    /* per-tet model ids passed separately (generic_elastic) */
} bh_model;

/* The element-matrix callback every in-repo caller uses: simple_psd_fix(K, tol) on the
 * volume-scaled 12x12 (src/simple_psd_fix.cpp:5-21; MATLAB passes tol = 1e-3).
 * NULL or enabled == 0 means the reference's default_callback (no-op). */
typedef struct {
  int enabled;
  double tol;
} bh_psd;

/* Options that are not part of the reference's argument lists. */
typedef struct {
  /* Stretch models only (corotational, ARAP).  1 (default): reproduce the reference's row-major
   * write of the 9x9 (src/d2psi_stretch_dF2.cpp:36-45), i.e. the Hessian evaluated at F^T.
   * 0: the true Hessian at F. */
  int stretch_reference_layout;
} bh_element_opts;

typedef struct bh_mesh bh_mesh;

/* ---- setup ------------------------------------------------------------------------------ */

/* Replaces sim::linear_tetmesh_dphi_dX (src/linear_tetmesh_dphi_dX.cpp:5-18,
 * src/linear_tet_dphi_dX.cpp:5-18) and the signed volume igl::volume
 * (tests/test_assemble.cpp:27): vol = -(a-d).((b-d)x(c-d))/6.  Host arrays. */
int bh_linear_tetmesh_dphi_dX(int nV, const double* V, int nT, const int* E, double* dphidX, double* volume);

/* Same, computed on the device (device pointers; V column-major nV x 3). */
int bh_linear_tetmesh_dphi_dX_device(bh_mesh* mesh, const double* d_V, double* d_dphidX, double* d_volume,
                                     bh_stream_t stream);

/* Builds the per-mesh device state once: the reference pattern (CSR == CSC), the vertex -> incident
 * tet lists in ascending element order and the block ranks that map each element block onto its
 * CSR slots.  Copies E, dphidX and volume (host arrays, reference layouts) to the device.
 * Replaces the per-call triplet list of sim::assemble (src/assemble.cpp:36-97). */
int bh_mesh_create(int nV, int nT, const int* E, const double* dphidX, const double* volume, bh_mesh** mesh);
/* The same mesh state built on the device from device arrays (no host copy of E): the incidence lists by
 * a stable radix sort of the (vertex, element) pairs, the pattern by a sort-unique of the 16 nT vertex
 * pairs, the ranks by binary search -- the same arrays as bh_mesh_create, bit for bit.  Replaces the
 * reference's per-call triplet pattern (src/assemble.cpp:57-94) and, when d_dphidX and d_volume are
 * NULL, sim::linear_tetmesh_dphi_dX (src/linear_tetmesh_dphi_dX.cpp:5-18) with the signed volume from
 * d_V (nV x 3, column-major).  Synchronous. */
int bh_mesh_create_device(int nV, int nT, const int32_t* d_E, const double* d_V, const double* d_dphidX,
                          const double* d_volume, bh_mesh** mesh);
int bh_mesh_destroy(bh_mesh* mesh);
/* Frees the buffers the mesh makes on first use and keeps between calls -- the PSD records and Q
 * factors, the gradient's corner forces, the Newton and coarse-space workspaces, the FP32 geometry, the
 * host-API value staging -- after a device synchronisation.  The next call that needs one makes it
 * again.  For callers that hold many meshes between solves.  No reference counterpart: the reference
 * keeps no state between calls. */
int bh_mesh_trim(bh_mesh* mesh);
int bh_mesh_nnz(const bh_mesh* mesh, int64_t* nnz);
/* rowptr: 3*nV + 1 entries, colind: nnz entries (host arrays). */
int bh_mesh_pattern(const bh_mesh* mesh, int32_t* rowptr, int32_t* colind);
/* device copies of the block pattern (vertex level): bptr nV+1, bcol nblocks; returned pointers
 * are owned by the mesh. */
int bh_mesh_block_pattern_device(const bh_mesh* mesh, const int32_t** d_bptr, const int32_t** d_bcol,
                                 int64_t* nblocks);
int bh_mesh_set_element_opts(bh_mesh* mesh, const bh_element_opts* opts);
/* Rest positions V (host, nV x 3 column-major; NULL clears).  Optional: the FP32 precision form
 * then forms G = F - I from displacement edges in FP64 and evaluates the Neo-Hookean / StVK energy
 * and stress from G, keeping their relative accuracy at small strain (without V, F is rounded to
 * FP32 first and the stress carries an absolute error of about 1e-7 of the modulus).  No
 * counterpart in the reference (its V argument is passed on every call and unused by the
 * element math). */
int bh_mesh_set_rest_positions(bh_mesh* mesh, const double* V);
/* Assembly-kernel shape: 1, 2, 4 or 8 lanes per vertex (with 2, each lane owns two of the four corner
 * blocks of every incident tet; with 4, one; with 8, the vertex's two quads take two incidences per
 * step), or 0, the default: 8 in the element-math gathers, 4 in the PSD-record gathers.  The callback
 * path's gather takes 4 for 8.  Same per-slot fold order. */
int bh_mesh_set_gather_lanes(bh_mesh* mesh, int lanes);

/* Kernel-schedule knobs of a mesh (no reference counterpart: the reference has one CPU schedule).
 * Every setting keeps each slot's ascending-element fold, so results agree to the last bits; a
 * fixed setting is bitwise reproducible. */
typedef struct bh_tuning {
  int gather_lanes; /* 0 keep, else 1, 2, 4 or 8 lanes per vertex in the assembly gathers */
  int xcd_groups;   /* -1 keep, 0 off, 1 on: deal vertex groups to the 8 XCDs in contiguous runs */
  int coloring;     /* 0 keep, 1 DSatur (default), 2 greedy in vertex order: the vertex colouring that
                       bh_precond_create uses for the colour sweeps */
  int sweep_lanes;  /* 0 keep, 4 (default) or 8: lanes per vertex in the symmetric Gauss-Seidel sweeps of
                       preconditioners created afterwards */
  int gather_order; /* 0 keep, 1 every gather, 2 none, 3 the PSD-record gathers only (default): visit the
                       vertex groups in the Morton order of their rest centroids (made when rest positions
                       are set), dealt to the XCDs in contiguous runs, instead of vertex order */
} bh_tuning;
int bh_mesh_set_tuning(bh_mesh* mesh, const bh_tuning* tuning);

/* ---- file formats (host only; no device call) ------------------------------------------------
 * Two-call convention: with NULL output arrays only the sizes are returned; called again with arrays
 * of those sizes they are filled.  V nV x 3 double and E nT x 4 int32, column-major (Eigen's layout). */

/* Replaces sim::read_tetgen (src/read_tetgen.cpp:105-124, loadTet :32-99): TetGen .node / .ele.  Lines
 * shorter than 3 characters and '#' comments are skipped; each record's leading index is dropped and the
 * vertex indices are returned as written in the file (the reference does not renumber them; TetGen
 * files written with -z are 0-based).  Only the first 4 nodes of each element are kept. */
int bh_read_tetgen(const char* node_file, const char* ele_file, int* nV, int* nT, double* V, int32_t* E);
/* Replaces igl::readMESH as the reference's tests use it (tests/test_assemble.cpp:25,
 * tests/test_constitutive_model.cpp:26): MEDIT .mesh, "Vertices" and "Tetrahedra" sections (other
 * sections skipped), indices returned 0-based. */
int bh_read_medit(const char* path, int* nV, int* nT, double* V, int32_t* E);
/* Replaces Eigen::saveMarket for the reference's CSC matrices (the format of the files in data/matrices):
 * "%%MatrixMarket matrix coordinate  real general", "rows cols nnz", then one "row col value" line per
 * stored entry (explicit zeros kept), 1-based, column by column with rows ascending, values "%.64e". */
int bh_save_market(const char* path, int64_t rows, int64_t cols, const int32_t* colptr, const int32_t* rowind,
                   const double* vals);
/* Replaces Eigen::loadMarket (tests/test_assemble.cpp:33): coordinate real general or symmetric (the
 * mirror entries added); duplicates summed in file order, rows sorted within each column (the semantics
 * of setFromTriplets).  colptr: cols + 1 entries, rowind / vals: nnz entries. */
int bh_load_market(const char* path, int64_t* rows, int64_t* cols, int64_t* nnz, int32_t* colptr, int32_t* rowind,
                   double* vals);

/* ---- energy / gradient / Hessian (host arrays) -------------------------------------------- */

/* Replaces sim::linear_tetmesh_{neohookean,stvk,corotational,arap}_q
 * (include/linear_tetmesh_neohookean_q.h:13-18, src/linear_tetmesh_neohookean_q.cpp:8-24 and
 * siblings).  V is implied by the mesh.  nT == 0 gives 0. */
int bh_linear_tetmesh_q(bh_mesh* mesh, int model, const double* q, const double* params, double* energy);
/* Replaces sim::linear_tetmesh_*_dq (src/linear_tetmesh_neohookean_dq.cpp:5-22, ...).  g: 3*nV. */
int bh_linear_tetmesh_dq(bh_mesh* mesh, int model, const double* q, const double* params, double* g);
/* Replaces sim::linear_tetmesh_*_dq2 (include/linear_tetmesh_neohookean_dq2.h:14-21,
 * src/linear_tetmesh_neohookean_dq2.cpp:8-28, ...).  vals: nnz values in the pattern of
 * bh_mesh_pattern; psd NULL = default_callback. */
int bh_linear_tetmesh_dq2(bh_mesh* mesh, int model, const double* q, const double* params, const bh_psd* psd,
                          double* vals);

/* An arbitrary element-matrix callback, the reference's ElementMatrixCallback on the Hessian
 * (include/EigenTypes.h:133 default_callback; called in src/linear_tetmesh_neohookean_dq2.cpp:17-23 and
 * its siblings): a host function called once per tet, in ascending element order, with the tet's
 * 12x12 element Hessian vol B^T d2psi B (column-major, row / column 3 a + c for corner a, component c,
 * the order sim::assemble maps to 3 E(e, a) + c, src/assemble.cpp:57-70), which it may change in place;
 * the changed matrices are assembled.  element is the tet's index. */
typedef void (*bh_element_callback)(double* element_matrix, int element, void* user);
/* Replaces sim::linear_tetmesh_*_dq2 with a callback other than simple_psd_fix (bh_psd covers that one
 * on the device).  The element matrices are made on the device in chunks of up to 262,144 tets, copied
 * to pinned host memory, handed to cb, copied back and folded by the vertex gather, every CSR slot in
 * ascending element order across chunks: with a callback that changes nothing the values equal
 * bh_linear_tetmesh_dq2 without PSD to rounding (the same closed-form element blocks).  model_ids NULL: model for every tet (0..5);
 * otherwise per-tet ids 0..4.  Synchronous; the callback runs on the calling thread. */
int bh_linear_tetmesh_dq2_callback(bh_mesh* mesh, int model, const uint8_t* model_ids, const double* q,
                                   const double* params, bh_element_callback cb, void* user, double* vals);
/* The same with device inputs and output; returns after the last chunk is assembled. */
int bh_linear_tetmesh_dq2_callback_device(bh_mesh* mesh, int model, const uint8_t* d_model_ids, const double* d_q,
                                          const double* d_params, bh_element_callback cb, void* user,
                                          double* d_vals, bh_stream_t stream);

/* Replaces sim::linear_tetmesh_generic_elastic_{q,dq,dq2} (src/linear_tetmesh_generic_elastic_q.cpp:5-30,
 * _dq.cpp:5-36, _dq2.cpp:8-38) for the built-in models: the user functor becomes a per-tet model id
 * (uint8, bh_model values 0..4), params stay nT x 2 per tet. */
int bh_linear_tetmesh_generic_elastic_q(bh_mesh* mesh, const uint8_t* model_ids, const double* q,
                                        const double* params, double* energy);
int bh_linear_tetmesh_generic_elastic_dq(bh_mesh* mesh, const uint8_t* model_ids, const double* q,
                                         const double* params, double* g);
int bh_linear_tetmesh_generic_elastic_dq2(bh_mesh* mesh, const uint8_t* model_ids, const double* q,
                                          const double* params, const bh_psd* psd, double* vals);

/* ---- device variants (device pointers, asynchronous on stream) ----------------------------- */
/* model_ids may be NULL (then model is used for every tet; model must not be BH_MODEL_PER_TET). */
int bh_linear_tetmesh_q_device(bh_mesh* mesh, int model, const uint8_t* d_model_ids, const double* d_q,
                               const double* d_params, double* d_energy, bh_stream_t stream);
int bh_linear_tetmesh_dq_device(bh_mesh* mesh, int model, const uint8_t* d_model_ids, const double* d_q,
                                const double* d_params, double* d_g, bh_stream_t stream);
int bh_linear_tetmesh_dq2_device(bh_mesh* mesh, int model, const uint8_t* d_model_ids, const double* d_q,
                                 const double* d_params, const bh_psd* psd, double* d_vals, bh_stream_t stream);

/* Allocation-free variant of bh_linear_tetmesh_q_device.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving or
 * line-search loops, harnesses that forbid hipMalloc).  It calls no hipMalloc / hipFree: the
 * per-workgroup partial sums of the energy live in a fixed device array inside the library
 * (BH_Q_NOALLOC_MAX_TETS / 1024 doubles, 128 KiB, held by each GPU the library runs on until the
 * process ends); the mesh (from bh_mesh_create, on the current device) and the inputs and output are
 * the caller's.  Same kernels and summation order as bh_linear_tetmesh_q_device, so the result is
 * bit-identical to it.  A mesh with more than BH_Q_NOALLOC_MAX_TETS tets is rejected with
 * BH_UNSUPPORTED (and a message on stderr).  Every call on one GPU shares that array: calls that
 * overlap on one GPU (concurrent streams, threads) race and must be serialised by the caller. */
#define BH_Q_NOALLOC_MAX_TETS (16 * 1024 * 1024)
int bh_linear_tetmesh_q_device_noalloc(bh_mesh* mesh, int model, const uint8_t* d_model_ids, const double* d_q,
                                       const double* d_params, double* d_energy, bh_stream_t stream);

/* ---- precision form (device pointers) --------------------------------------------------------------
 * FP32 element path: FP32 copies of dphidX / volume (made on first use), FP32 params (nT x 2,
 * column-major), F formed from FP64 edge differences then FP32; J = det F, 1/J, log J and the
 * energy invariants in FP64; energy and gradient accumulated and returned in FP64; Hessian values
 * FP32.  Same pattern, same models, same PSD semantics as the FP64 entry points. */
int bh_linear_tetmesh_q_f32_device(bh_mesh* mesh, int model, const uint8_t* d_model_ids, const double* d_q,
                                   const float* d_params, double* d_energy, bh_stream_t stream);
int bh_linear_tetmesh_dq_f32_device(bh_mesh* mesh, int model, const uint8_t* d_model_ids, const double* d_q,
                                    const float* d_params, double* d_g, bh_stream_t stream);
int bh_linear_tetmesh_dq2_f32_device(bh_mesh* mesh, int model, const uint8_t* d_model_ids, const double* d_q,
                                     const float* d_params, const bh_psd* psd, float* d_vals, bh_stream_t stream);

/* ---- objective terms (config objectives) ---------------------------------------------------- */

/* Replaces sim::linear_tetmesh_mass_matrix (src/linear_tetmesh_mass_matrix.cpp:7-24,
 * src/linear_tet_mass_matrix.cpp:8-44): m/10 diagonal, m/20 off-diagonal identity blocks,
 * same pattern as the Hessian.  density: nT. */
int bh_linear_tetmesh_mass_matrix_device(bh_mesh* mesh, const double* d_density, double* d_vals,
                                         bh_stream_t stream);
/* Replaces sim::linear_tetmesh_const_ext_force (src/linear_tetmesh_const_ext_force.cpp:5-18):
 * f = sum_e traction * vol_e / 4 on each vertex of e.  traction: 3 doubles (host). */
int bh_linear_tetmesh_const_ext_force_device(bh_mesh* mesh, const double* traction, double* d_f,
                                             bh_stream_t stream);

/* ---- Newton ------------------------------------------------------------------------------------ */

typedef enum {
  BH_PRECOND_JACOBI = 0,       /* 3x3 block-Jacobi (block diagonal of H) */
  BH_PRECOND_POINT_JACOBI = 1,
  BH_PRECOND_BIC0 = 2,         /* block IC(0) in colour order, refactored from H every Newton iteration */
  BH_PRECOND_FACTOR = 3,       /* a block lower-triangular factor supplied through precond_handle */
  BH_PRECOND_LEARNED = 4,      /* factor written by the mesh-graph network opts.mgn from each Hessian */
  BH_PRECOND_SGS = 5,          /* block symmetric Gauss-Seidel in colour order, M = (D + L) D^-1 (D + L)^T
                                  with L the strictly lower part of H, applied with Eisenstat's trick (one read
                                  of H per CG iteration, no separate product); needs precond_handle */
  BH_PRECOND_LEARNED_SGS = 6,  /* M = (Dt + L) Dt^-1 (Dt + L)^T with the block diagonal Dt written by the
                                  mesh-graph network opts.mgn (Dt = D at zero decoders) */
  BH_PRECOND_JACOBI_2L = 7     /* two-level: block Jacobi plus an additive coarse correction Z E^-1 Z^T,
                                  Z the 30 quadratic modes (each displacement component times 1,
                                  d, d^2 and the products of d, d = q about the centroid) of each
                                  aggregate of consecutive vertices, E = Z^T H Z; no precond_handle
                                  needed (coarse_aggregates sets the aggregate count) */
} bh_precond_kind;

/* Block lower-triangular preconditioner M = L L^T on the Hessian's own 3x3-block pattern, in a
 * greedy vertex-colour order (no two neighbours share a colour, so each triangular sweep is one
 * parallel kernel per colour).  No counterpart in the reference (its direction is a direct solve). */
typedef struct bh_precond bh_precond;
int bh_precond_create(bh_mesh* mesh, bh_precond** precond);
int bh_precond_destroy(bh_precond* precond);
int bh_precond_info(const bh_precond* precond, int* ncolors, int64_t* nlower);
/* block IC(0) of the masked H (fixed DOFs -> identity); shift > 0 adds shift*|H_ii| to the diagonal.
 * failures (host, optional): number of 3x3 pivots that were not positive. */
int bh_precond_factor_bic0_device(bh_precond* precond, const double* d_vals, const uint8_t* d_fixed_dof, double shift,
                                  int* failures, bh_stream_t stream);
/* install a factor: d_lower 9 floats per lower block (row-major, in the order of the lower lists),
 * d_diag_inv 6 floats per vertex (inverse of L_vv, packed lower triangular, row-major). */
int bh_precond_set_factor_device(bh_precond* precond, const float* d_lower, const float* d_diag_inv, bh_stream_t stream);
/* device views of the lower lists: lptr (nV+1), lcol and lslot (mesh block index of (v, n)) per lower
 * block, dslot (mesh block index of (v, v)), vertices grouped by colour */
int bh_precond_lists_device(const bh_precond* precond, const int32_t** lptr, const int32_t** lcol,
                            const int32_t** lslot, const int32_t** dslot, const int32_t** colverts);
/* z = (L L^T)^{-1} r */
int bh_precond_apply_device(bh_precond* precond, const double* d_r, double* d_z, bh_stream_t stream);

/* ---- learned preconditioner (no counterpart in the reference) -----------------------------------
 * A mesh-graph network on the vertices and lower pairs of a bh_precond: node / pair features from
 * the masked, diagonally normalised Hessian, 3 message-passing layers of width 32, decoders that
 * write the factor L (L L^T approximates H; at initialisation the decoders are zero and L L^T is
 * the block diagonal of H).  Parameters are FP32 (bh_mgn_num_params floats, layout fixed by the
 * library).  The loss of a system H x = b is |L L^T x - b|^2 / |b|^2 in the diagonally scaled norm;
 * bh_mgn_loss_device accumulates its gradient until bh_mgn_adam_step, bh_mgn_train_step_device
 * trains on probe systems of a Hessian the plain solver produced. */
typedef struct bh_mgn bh_mgn;
int bh_mgn_create(uint64_t seed, bh_mgn** mgn);
int bh_mgn_destroy(bh_mgn* mgn);
/* The preconditioner form the loss and training steps use: 0 the general block lower-triangular factor
 * (BH_PRECOND_LEARNED), 1 the Eisenstat form M = (Dt + L) Dt^-1 (Dt + L)^T whose off-diagonal blocks
 * are tied to H (BH_PRECOND_LEARNED_SGS; only the node decoder acts), 2 the tied factor
 * (BH_PRECOND_LEARNED): L_vv from the node decoder as in form 0, L_vn = H_vn L_nn^-T + s_v O_k with O_k
 * the pair decoder's output, so that at zero decoders L L^T is block symmetric Gauss-Seidel and the
 * coarse correction applies.  probe_is_solution = 1: the vector
 * bh_mgn_train_step_device receives is a solution x of the Newton system (probe 0 is x itself) instead of
 * its right-hand side.  No counterpart in the reference (the learned preconditioner is this build's
 * addition).  energy_loss = 1 (form 1 only): the loss is the mean over probes x of
 * (log(x^T M x) - log(x^T H x))^2 instead of |M x - H x|^2 / |H x|^2.  energy_loss = 2 (forms 0 and 2):
 * the IC(0) teacher, |Lhat - T|_F^2 / |T|_F^2 with T the block IC(0) factor of the same Hessian in the
 * diagonally scaled space; bh_mgn_train_step_device then factors each Hessian it is given (no probes).
 * energy_loss = 3 (forms 0 and 2): Kaporin's condition number of L L^T against H (see
 * bh_mgn_train_step_device).  A training step whose loss is not
 * finite returns BH_NONFINITE and leaves the parameters. */
int bh_mgn_set_form(bh_mgn* mgn, int form, int probe_is_solution, int energy_loss);
/* The MLPs of inference (bh_mgn_factor_device and the Newton directions BH_PRECOND_LEARNED /
 * BH_PRECOND_LEARNED_SGS): precision 1 (the default) BF16 inputs, weights and hidden activations on the
 * matrix cores with FP32 accumulation and an FP32 residual stream; 0 the FP32 forward that training uses.
 * Training and the loss always run FP32.  No counterpart in the reference. */
int bh_mgn_set_inference(bh_mgn* mgn, int precision);
int bh_mgn_num_params(const bh_mgn* mgn, int64_t* nparams);
/* host copies of the parameters and of the accumulated gradient (either may be NULL) */
int bh_mgn_get_params(const bh_mgn* mgn, float* params, float* grad);
/* sets the parameters and clears the gradient and the Adam state */
int bh_mgn_set_params(bh_mgn* mgn, const float* params);
/* loss of one system (d_vals: FP32 Hessian values in the mesh pattern, d_b / d_x: 3 nV, fixed DOFs
 * zero); backward != 0 also accumulates the gradient.  loss (host, optional) synchronises. */
int bh_mgn_loss_device(bh_mgn* mgn, bh_precond* precond, const float* d_vals, const uint8_t* d_fixed_dof,
                       const double* d_b, const double* d_x, int backward, double* loss, bh_stream_t stream);
/* One training step on a Hessian: nprobe (<= 16) probe systems -- the normalised right-hand side
 * d_rhs first when given, then standard normal vectors drawn from seed -- give a Hutchinson estimate
 * of |L L^T - H|_F^2 / |H|_F^2 in the diagonally scaled norm; backward and one Adam step (lr,
 * beta 0.9 / 0.999, eps 1e-8).  With the Kaporin loss (bh_mgn_set_form, energy_loss = 3; factor forms)
 * the probes instead give the Hutchinson estimate of tr((L L^T)^-1 H) in
 * log(tr / n) + (2 / n) log det L (n = 3 nV), Kaporin's condition number up to log det H / n.
 * lr = 0: the gradient is accumulated and no Adam step taken (bh_mgn_get_params returns it).
 * loss (host, optional) synchronises. */
int bh_mgn_train_step_device(bh_mgn* mgn, bh_precond* precond, const float* d_vals, const uint8_t* d_fixed_dof,
                             const double* d_rhs, int nprobe, uint64_t seed, double lr, double* loss,
                             bh_stream_t stream);
/* one Adam update from the accumulated gradient (which it then clears) */
int bh_mgn_adam_step(bh_mgn* mgn, double lr, double beta1, double beta2, double eps, bh_stream_t stream);
/* inference: writes the factor of the given Hessian into the preconditioner */
int bh_mgn_factor_device(bh_mgn* mgn, bh_precond* precond, const float* d_vals, const uint8_t* d_fixed_dof,
                         bh_stream_t stream);
/* device views of the last forward pass: scales s (nV), chol of the normalised diagonal blocks
 * (6 per vertex, packed lower), decoder outputs d (6 per vertex) and O (9 per lower pair) */
int bh_mgn_debug_buffers(bh_mgn* mgn, const float** s, const float** C, const float** d, const float** O);
/* device views of the last forward pass's latents: layer -1 the node (8 per vertex) and pair (10 per
 * lower pair) features; layer 0..3 the node latents h (32 per vertex) and pair latents e (32 per pair)
 * after that many message-passing layers, and (layers 0..2) the messages m the next node update reads */
int bh_mgn_debug_latents(bh_mgn* mgn, int layer, const float** h, const float** e, const float** m);

typedef struct {
  int model;               /* bh_model, or BH_MODEL_PER_TET with model_ids */
  const uint8_t* d_model_ids;
  bh_psd psd;              /* PSD fix applied to every element Hessian */
  double gradient_tol;     /* reference default 1e-3 (absolute, on |g|_2) */
  int relative_tol;        /* 0: |g| < tol (reference); 1: |g| < tol * |g_0| */
  int max_iterations;      /* reference default 100; the loop runs up to max+1 times */
  int max_iterations_ls;   /* reference default 100 */
  double alpha;            /* 1.0 */
  double c;                /* 1e-8 */
  double p;                /* 0.5 */
  int armijo_with_alpha;   /* 0: reference test f(x+a d) > f(x) + c g^T d; 1: textbook c a g^T d */
  int reject_nonfinite;    /* 1: a non-finite trial energy counts as "not decreased" (the reference accepts it) */
  double pcg_tol;          /* relative residual of the inner PCG solve */
  int pcg_max_iterations;
  int precond;             /* bh_precond_kind */
  bh_precond* precond_handle; /* BH_PRECOND_BIC0 / BH_PRECOND_FACTOR */
  double precond_shift;    /* BH_PRECOND_BIC0 diagonal shift */
  /* 0: FP64 element path.
   * 1: the precision form for the solve -- FP32 element Hessian and FP32 Hessian blocks for every
   *    product and sweep (FP64 diagonal blocks rebuilt from exact row sums), element energy and
   *    gradient in FP64, so the equilibrium Newton converges to is the FP64 one.
   * 2: the full precision form -- element energy, gradient and Hessian all in FP32 (J = det F and
   *    log J in FP64; G = F - I from FP64 displacement edges when rest positions are set).  Its
   *    equilibrium carries the FP32 stress error, largest under big rotations.
   * In both, the energy sum of the line search, the gradient accumulation, every CG vector and dot
   * product, the residual check and the convergence test stay FP64; params stay FP64 at the
   * interface (converted once per solve). */
  int element_precision;
  /* 1: when a factor or Gauss-Seidel preconditioned PCG (BIC0, FACTOR, LEARNED, SGS, LEARNED_SGS) hits
   * pcg_max_iterations, or breaks down (r.u or p.Ap not positive or not finite: an indefinite factor or
   * diagonal), or block IC(0) still has non-positive pivots after its last diagonal shift, solve again
   * with block-Jacobi PCG from zero (the fallback), counted in bh_newton_report.pcg_fallbacks.
   * 0: the capped or broken-down direction is used as it is (IC(0) with failed pivots returns
   * BH_SOLVE_FAILED) */
  int jacobi_fallback;
  bh_mgn* mgn;             /* BH_PRECOND_LEARNED (precond_handle supplies the colour lists) */
  /* optional, called after every direction solve with device pointers: the Hessian values (FP32 in
   * the precision form, else FP64; the other pointer is NULL), the masked right-hand side b = -g
   * and the PCG solution x (3 nV each); how the plain solver's systems reach a trainer */
  void (*on_system)(void* user, int iteration, const float* d_vals32, const double* d_vals64, const double* d_b,
                    const double* d_x);
  void* user;
  /* Backward-Euler inertia (matlab/examples/test_tet_materials.m:97-112 in position form): when
   * d_mass_vals != NULL the objective gains 1/2 inertia (q - q_pred)^T M (q - q_pred), with M the
   * consistent mass matrix (values in the mesh pattern, bh_linear_tetmesh_mass_matrix_device),
   * q_pred = q_t + dt v_t and inertia = 1 / dt^2.  bh_newton_static_device then performs one step. */
  const double* d_mass_vals;
  const double* d_q_pred;
  double inertia;
  double relaxation;       /* BH_PRECOND_SGS: SSOR relaxation w in (0, 2), Dt = D / w (0 means 1) */
  int coarse_aggregates;   /* BH_PRECOND_JACOBI_2L and coarse_correction: number of aggregates (0: one
                              per 4096 vertices, 4 to 8; larger requests are capped at 16) */
  int coarse_correction;   /* 1: add the two-level coarse term Z E^-1 Z^T of BH_PRECOND_JACOBI_2L to the
                              factor preconditioners (BH_PRECOND_BIC0, _FACTOR, _LEARNED); 0 (default):
                              the factor alone */
} bh_newton_opts;

typedef struct {
  int iterations;          /* Newton iterations run */
  int converged;           /* 1 = SolverExitStatus::CONVERGED */
  int pcg_iterations_total;
  int pcg_failures;        /* inner solves that hit their cap */
  int pcg_fallbacks;       /* factor-preconditioned solves redone with block Jacobi */
  int ls_iterations_total;
  double energy;           /* final objective */
  double gradient_norm;    /* final |g| */
  double gradient_norm0;
} bh_newton_report;

/* Fills opts with the reference defaults (include/optimization_gradient_based.h:51-58). */
void bh_newton_default_opts(bh_newton_opts* opts);

/* Static equilibrium on the device: minimise Psi(q) - f_ext^T q over the DOFs with
 * fixed_dof[i] == 0 (fixed DOFs keep their input values).  Replaces
 * sim::newtons_method_bisection (include/optimization_gradient_based.h:33-110) with
 * search_direction_newton_sparse (src/search_direction_newton_sparse.cpp:5-26; the direct solve
 * becomes a preconditioned CG to opts->pcg_tol) and linesearch_backtracking_bisection
 * (src/linesearch_backtracking_bisection.cpp:5-31).  All pointers are device pointers.
 * d_f_ext may be NULL.  trace (host, optional, 4*(max_iterations+1) doubles):
 * per iteration [objective, |g|, alpha, pcg iterations]. */
int bh_newton_static_device(bh_mesh* mesh, const bh_newton_opts* opts, const double* d_params,
                            const uint8_t* d_fixed_dof, const double* d_f_ext, double* d_q, bh_newton_report* report,
                            double* trace, bh_stream_t stream);

/* ---- instrumentation ---------------------------------------------------------------------------- */

/* Kernel classes timed with HIP events (recorded on the launch stream) while profiling is on. */
typedef enum {
  BH_KT_ENERGY = 0,
  BH_KT_GRADIENT = 1,
  BH_KT_HESSIAN_GATHER = 2, /* the CSR assembly kernel (element or record blocks + deterministic gather) */
  BH_KT_PSD = 3,            /* simple_psd_fix passes: each tet's reduced matrix, eigen-clamped into its record */
  BH_KT_COUNT = 4
} bh_kernel_class;

/* on != 0 starts recording (and clears previous records); on == 0 stops. */
int bh_mesh_set_profiling(bh_mesh* mesh, int on);
/* Synchronises the recorded events and returns, per kernel class, the summed duration in ms and the
 * number of launches since profiling was switched on.  ms, launches: BH_KT_COUNT entries. */
int bh_mesh_kernel_times(bh_mesh* mesh, double* ms, int64_t* launches);

/* number of tets whose element Hessian needed eigenvalue clamping in the last PSD call: FP64, tets
 * with an eigenvalue below tol; FP32, tets whose M - tol I failed the Cholesky test (synchronises
 * the device) */
int bh_mesh_psd_count(bh_mesh* mesh, int64_t* count);

/* Device sparse helpers on the mesh pattern (values in bh_mesh_pattern order). */
int bh_spmv_device(bh_mesh* mesh, const double* d_vals, const double* d_x, double* d_y, bh_stream_t stream);

/* ---- triangle siblings: 2-D plane-strain triangles (tri2d) and the StVK shell triangle (tri3d) ----
 * Element arrays in the reference's layouts (column-major): E nT x 3, dphidX nT x 6 (tri2d) or nT x 9
 * (tri3d) with Phi(m, j) = d phi_m / d X_j at column m + 3j, volume nT (the element area times the
 * thickness the caller folds in), params nT x 2 with the tet conventions ([C, D] Neo-Hookean, [lambda,
 * mu] StVK and corotational, [k, .] ARAP), q dim nV interleaved.  The Hessian values are in
 * bh_trimesh_pattern order: full dim x dim blocks per vertex pair sharing an element, explicit zeros
 * kept, sorted indices (the reference's CSC, by symmetry).  Models: BH_NEOHOOKEAN, BH_STVK,
 * BH_COROTATIONAL, BH_ARAP.  The elements evaluate the 3-D energies at F3 = [[F2, 0], [0, 1]]. */
typedef struct bh_trimesh bh_trimesh;
/* Replaces sim::linear_tri2dmesh_dphi_dX (src/linear_tri2dmesh_dphi_dX.cpp:5-16, element
 * src/linear_tri2d_dphi_dX.cpp:5-17).  V nV x 2.  BH_NONFINITE for a degenerate triangle. */
int bh_linear_tri2dmesh_dphi_dX(int nV, const double* V, int nT, const int* E, double* dphidX);
/* Replaces sim::linear_tri3dmesh_dphi_dX (src/linear_tri3dmesh_dphi_dX.cpp:5-19, element
 * src/linear_tri3d_dphi_dX.cpp:5-19): rows 1..2 of Phi = (T^T T)^-1 T^T, T = [x1 - x0, x2 - x0]. */
int bh_linear_tri3dmesh_dphi_dX(int nV, const double* V, int nT, const int* E, double* dphidX);
/* The pattern and incidence lists of a triangle mesh with dim (2 or 3) DOFs per vertex, built once
 * (the reference rebuilds its triplet list per call, src/assemble.cpp:36-97). */
int bh_trimesh_create(int dim, int nV, int nT, const int* E, bh_trimesh** mesh);
int bh_trimesh_destroy(bh_trimesh* mesh);
int bh_trimesh_nnz(const bh_trimesh* mesh, int64_t* nnz);
/* rowptr dim nV + 1 entries, colind nnz entries (host arrays). */
int bh_trimesh_pattern(const bh_trimesh* mesh, int32_t* rowptr, int32_t* colind);
/* Replace sim::linear_tri2dmesh_{neohookean,stvk,corotational,arap}_q / _dq / _dq2
 * (src/linear_tri2dmesh_neohookean_q.cpp, _dq.cpp, _dq2.cpp:5-24 and siblings; elements
 * src/linear_tri2d_*.cpp) on host arrays; dq2 with the simple_psd_fix callback as bh_psd. */
int bh_linear_tri2dmesh_q(bh_trimesh* mesh, int model, const double* q, const double* dphidX, const double* volume,
                          const double* params, double* energy);
int bh_linear_tri2dmesh_dq(bh_trimesh* mesh, int model, const double* q, const double* dphidX, const double* volume,
                           const double* params, double* g);
int bh_linear_tri2dmesh_dq2(bh_trimesh* mesh, int model, const double* q, const double* dphidX, const double* volume,
                            const double* params, const bh_psd* psd, double* vals);
/* The same on device arrays (d_energy: one double on the device); they synchronise the stream. */
int bh_linear_tri2dmesh_q_device(bh_trimesh* mesh, int model, const double* d_q, const double* d_dphidX,
                                 const double* d_volume, const double* d_params, double* d_energy, bh_stream_t stream);
int bh_linear_tri2dmesh_dq_device(bh_trimesh* mesh, int model, const double* d_q, const double* d_dphidX,
                                  const double* d_volume, const double* d_params, double* d_g, bh_stream_t stream);
int bh_linear_tri2dmesh_dq2_device(bh_trimesh* mesh, int model, const double* d_q, const double* d_dphidX,
                                   const double* d_volume, const double* d_params, const bh_psd* psd, double* d_vals,
                                   bh_stream_t stream);
/* Replaces sim::linear_tri2dmesh_mass_matrix (src/linear_tri2dmesh_mass_matrix.cpp, element
 * src/linear_tri2d_mass_matrix.cpp:5-29): blocks rho vol / 6 I2 on the diagonal, rho vol / 12 I2 off it. */
int bh_linear_tri2dmesh_mass_matrix(bh_trimesh* mesh, const double* density, const double* volume, double* vals);
/* Replaces sim::linear_tri3dmesh_stvk_dq2 (src/linear_tri3dmesh_stvk_dq2.cpp:5-30, element
 * src/linear_tri3d_stvk_dq2.cpp:6-40) on a dim-3 mesh: per element the rest normal N and current
 * normal n (nT x 3) and the normal's derivative dndq (nT x 27, the 3 x 9 matrix column-major), as the
 * reference takes them; F3 = B q_e + vec(n N^T), the Hessian B^T d2psi B vol (the normal's second
 * derivative ignored, as in the reference). */
int bh_linear_tri3dmesh_stvk_dq2(bh_trimesh* mesh, const double* q, const double* dphidX, const double* N,
                                 const double* n, const double* dndq, const double* volume, const double* params,
                                 const bh_psd* psd, double* vals);
/* The same on device arrays; synchronises the stream.  The triangle calls keep their element results
 * in a workspace of the mesh handle: calls one after another on a handle are safe, overlapping ones
 * are not. */
int bh_linear_tri3dmesh_stvk_dq2_device(bh_trimesh* mesh, const double* d_q, const double* d_dphidX,
                                        const double* d_N, const double* d_n, const double* d_dndq,
                                        const double* d_volume, const double* d_params, const bh_psd* psd,
                                        double* d_vals, bh_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* BARTELS_HIP_H */
