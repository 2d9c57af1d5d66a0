/*
 * hgdg — C ABI of the MI355X explicit DG operator.
 *
 * The library replaces the right-hand-side path of hpGEM's explicit
 * Runge-Kutta solvers: the per-element volume and per-face quadrature
 * integrals, the block-diagonal inverse-mass apply and the RK stage
 * combination/update, for linear upwind advection and the first-order acoustic
 * wave system. Everything is plain C: opaque handles, host pointers and sizes,
 * int status codes (0 = success), and a thread-local error string.
 *
 * Threading: one context per host thread. Calls enqueue work on the context's
 * HIP stream; hgdg_state_get and hgdg_sync block.
 *
 * State layout seen by callers ("reference layout"): element-major, and inside
 * an element field-major, iVB = iV * nB + iB, with the basis functions in
 * hpGEM's own order (kernel/Base/ElementData.h:209-225;
 * kernel/FE/BasisFunctions2DH1ConformingSquare.cpp:164-184,
 * BasisFunctions3DH1ConformingCube.cpp:434-528). On the device the state is
 * kept in the library's own layout; set/get convert.
 */
#ifndef HGDG_H
#define HGDG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    HGDG_OK = 0,
    HGDG_ERR_ARG = 1,         /* invalid argument / inconsistent sizes        */
    HGDG_ERR_HIP = 2,         /* HIP runtime error (message has the detail)   */
    HGDG_ERR_UNSUPPORTED = 3, /* valid input outside what the kernels handle  */
    HGDG_ERR_STATE = 4,       /* call out of order (no mesh, no operator ...) */
    HGDG_ERR_MEMORY = 5       /* device allocation failed / over budget       */
};

/* ---- element shapes (hpGEM reference geometries) ----------------------- */
enum {
    HGDG_SHAPE_SQUARE = 2,      /* kernel/Geometry/ReferenceSquare.cpp      */
    HGDG_SHAPE_CUBE = 4,        /* kernel/Geometry/ReferenceCube.cpp        */
    HGDG_SHAPE_TETRAHEDRON = 5, /* kernel/Geometry/ReferenceTetrahedron.cpp */
    /* shapes a mesh file may hold but the device operators do not take */
    HGDG_SHAPE_LINE = 1,
    HGDG_SHAPE_TRIANGLE = 3,
    HGDG_SHAPE_PRISM = 6,
    HGDG_SHAPE_PYRAMID = 7
};

/* ---- operators ----------------------------------------------------------- */
enum {
    HGDG_OP_ADVECTION_UPWIND = 1, /* u_t + a.grad u = 0, upwind flux
                                     (tests/self/Basic/080AdvectionLinear_SelfTest.cpp:108-170) */
    HGDG_OP_ACOUSTIC_CENTRAL = 2, /* first-order wave system, central flux, reflecting walls
                                     (applications/ExampleMultipleVariableProblem/AcousticWaveLinear_Impl.h:102-283) */
    HGDG_OP_MATRIX = 3            /* any linear operator given by its matrices (hgdg_operator_set_matrices) */
};

/* ---- arithmetic of the bulk contractions ------------------------------- */
enum {
    HGDG_PREC_FP64 = 0,       /* everything FP64 (the parity path)                  */
    HGDG_PREC_MIXED_FP32 = 1  /* FP32 contractions, FP32 stage derivatives,
                                 compensated FP32 face sums; FP64 geometry, state
                                 and RK update                                      */
};

/*
 * Mesh as hpGEM's MeshManipulator sees it after reading a .hpgem file:
 * per element its vertices' physical coordinates (periodic copies already
 * resolved) and its topological node ids (periodic images share an id), in
 * the reference geometry's local node order. Faces are found from shared node
 * ids, as kernel/Base/MeshManipulator_Impl.h:1860-1950 does.
 */
typedef struct hgdg_mesh_desc {
    int32_t shape;            /* HGDG_SHAPE_*                                  */
    int32_t dim;              /* 2 or 3 (must match the shape)                 */
    int64_t n_elements;
    const double* coords;     /* [n_elements][n_vertices][dim]                 */
    const int64_t* node_ids;  /* [n_elements][n_vertices]                      */
    const int32_t* order;     /* [n_elements] polynomial order p, or NULL      */
    int32_t uniform_order;    /* used when order == NULL                       */
} hgdg_mesh_desc;

typedef struct hgdg_op_desc {
    int32_t kind;             /* HGDG_OP_*                                     */
    int32_t precision;        /* HGDG_PREC_*                                   */
    double velocity[3];       /* advection velocity a                          */
    const double* c_inv;      /* acoustic: 1/c^2-type weight per element (getCInv), [n_elements] */
} hgdg_op_desc;

typedef struct hgdg_ctx hgdg_ctx;

/* Thread-local message of the last failing call on this thread. */
const char* hgdg_last_error(void);

/* Context on a HIP device; owns a stream, device buffers and graphs.
 * Replaces the ownership role of HpgemAPIBase (kernel/Base/HpgemAPIBase_Impl.h:48-90). */
int hgdg_ctx_create(hgdg_ctx** out, int device);
int hgdg_ctx_destroy(hgdg_ctx* ctx);

/* Upload the mesh; builds faces, geometry and reference tables.
 * Replaces HpgemAPIBase::addMesh + MeshManipulator::readMesh face setup
 * (kernel/Base/HpgemAPIBase_Impl.h:75-90, MeshManipulator_Impl.h:1505-1950). */
int hgdg_mesh_upload(hgdg_ctx* ctx, const hgdg_mesh_desc* mesh);

/* Sizes: elements, faces (internal + boundary), internal faces, DOF per field
 * set (sum over elements of nB * n_fields). */
int hgdg_mesh_info(hgdg_ctx* ctx, int64_t* n_elements, int64_t* n_faces, int64_t* n_internal_faces,
                   int64_t* n_dof);

/* Choose the operator (integrands) and the arithmetic.
 * Replaces HpgemAPILinear::createMassMatrices/createStiffnessMatrices
 * (kernel/Base/HpgemAPILinear_Impl.h:131-139,250-288). */
int hgdg_operator_set(hgdg_ctx* ctx, const hgdg_op_desc* op);

/* Launch shape of the tensor-product advection kernel: each workgroup takes
 * `tiles` consecutive tiles of elements (1, 2, 4 or 8), loading the next
 * tile's stage input while it computes the current one. 0 (the default)
 * lets hgdg_operator_set time every choice on the mesh and keep the fastest.
 * Results are bitwise the same for every choice. No counterpart in the
 * reference (a launch parameter of this implementation); applies to the next
 * hgdg_operator_set, or at once if an operator is set. */
int hgdg_set_advection_tiles(hgdg_ctx* ctx, int tiles);

/* The tile count in use and the timed average launch (ms) of each candidate
 * 1, 2, 4, 8 (0 where not timed). ms may be NULL. */
int hgdg_get_advection_tiles(hgdg_ctx* ctx, int* tiles, double* ms);

/* Whether hgdg_rk_steps may fuse (1, the default) or always replays the
 * captured one-step graph (0). Fusing means: all steps in one
 * single-workgroup launch when the whole state fits in LDS (tensor advection
 * meshes of a few thousand degrees of freedom); otherwise pairs of steps
 * whose RK update runs in each step's last-stage kernel (every operator).
 * Results are bitwise the same. No counterpart in the reference (a launch
 * choice of this implementation). */
int hgdg_set_rk_fused(hgdg_ctx* ctx, int enable);

/* The setting above and how many single-workgroup RK launches the context
 * has made. launches may be NULL. */
int hgdg_get_rk_fused(hgdg_ctx* ctx, int* enable, int64_t* launches);

/* A linear operator given by the matrices HpgemAPILinear builds from the
 * application's own integrands (createMassMatrices / createStiffnessMatrices,
 * kernel/Base/HpgemAPILinear_Impl.h:131-139,250-288, from
 * computeIntegrandMassMatrix and computeIntegrandStiffnessMatrixAtElement /
 * AtFace): the device then runs multiplyStiffnessMatrices (:360-434),
 * solveMassMatrixEquations (:143-156) and everything built on them
 * (hgdg_time_derivative, hgdg_right_hand_side, hgdg_solve_mass, hgdg_rk_steps,
 * hgdg_rk_adaptive_step, hgdg_step_error, hgdg_state_diff). Vectors are in
 * hpGEM's layout (elements in order, iVB = iV nB + iB inside each), with no
 * permutation. All matrices column-major, rows = test functions, columns =
 * trial functions. Projection and error norms against a function need the
 * geometry and stay with the caller (hpGEM's own integrators).
 * Replaces the context's mesh and operator; reallocate vectors afterwards. */
typedef struct hgdg_matrix_desc {
    int64_t n_elements;
    int32_t n_fields;             /* unknowns per basis function (informative)             */
    const int32_t* block_size;    /* [E] unknowns of element e, nB(e) * n_fields, <= 1024   */
    const double* mass;           /* M_e, element after element, n_e x n_e each             */
    const double* stiffness;      /* S_e (the element part), same layout                    */
    int64_t n_faces;
    const int64_t* face_elements; /* [F][2] left and right element (right = -1: boundary)   */
    const double* face_blocks;    /* per face, the FaceMatrix blocks in order LL (nL x nL),
                                     LR (nL x nR), RL (nR x nL), RR (nR x nR); boundary faces
                                     LL only (FaceMatrix::getElementMatrix, FaceMatrix.h:61-142) */
} hgdg_matrix_desc;

int hgdg_operator_set_matrices(hgdg_ctx* ctx, const hgdg_matrix_desc* desc, int precision);

/* Quadrilateral and hexahedral meshes take one of two kernel families:
 * sum-factorised kernels for affine elements (parallelograms,
 * parallelepipeds) whose neighbours share each face with the opposite local
 * face in the same orientation (every structured mesh), and a Gauss-point
 * family for everything else hpGEM accepts: bilinear/trilinear elements
 * (MappingToPhysHypercubeLinear.cpp:222-304) and neighbours meeting through
 * any local face in any of the square's orientations (FaceGeometry.cpp:148-155,
 * ReferenceSquare.cpp:131-180), with hpGEM's pairing of face points. `force`
 * = 1 makes meshes uploaded afterwards take the Gauss-point family even when
 * the sum-factorised one applies (both give hpGEM's operator to rounding);
 * 0 (the default) chooses by the mesh. No counterpart in the reference (a
 * kernel choice of this implementation). */
int hgdg_set_tensor_general(hgdg_ctx* ctx, int force);

/* The setting above, and whether the current mesh runs on the Gauss-point
 * family (1) or the sum-factorised one (0; also for tetrahedra). */
int hgdg_get_tensor_general(hgdg_ctx* ctx, int* force, int* in_use);

/* Threads per element of the tetrahedral RHS kernel, per order 1..4
 * (threads[p-1]): 0 (the default) lets hgdg_operator_set time both candidates
 * of that order on the mesh and keep the faster; otherwise one of the two
 * candidates, 8 or 16 (order 1), 12 or 32 (order 2), 32 or 64 (order 3),
 * 35 or 64 (order 4). Results are bitwise the same for every choice. No
 * counterpart in the reference (a launch parameter of this implementation);
 * applies to the next hgdg_operator_set, or at once if an operator is set. */
int hgdg_set_simplex_threads(hgdg_ctx* ctx, const int32_t* threads);

/* The threads per element in use for orders 1..4 and the timed average launch
 * (ms) of both candidates of each order, ms[2(p-1) + k] (0 where not timed).
 * ms may be NULL. */
int hgdg_get_simplex_threads(hgdg_ctx* ctx, int32_t* threads, double* ms);

/* Allocate n time-integration vectors on the device (all zero).
 * Replaces setNumberOfTimeIntegrationVectorsGlobally (HpgemAPIBase_Impl.h:141-156). */
int hgdg_state_alloc(hgdg_ctx* ctx, int n_vectors);

/* Copy a vector from / to the host in reference layout (n_dof doubles).
 * Replaces Element::setTimeIntegrationVector / getTimeIntegrationVector. */
int hgdg_state_set(hgdg_ctx* ctx, int vec, const double* host);
int hgdg_state_get(hgdg_ctx* ctx, int vec, double* host);

/* Storage of one vector: HGDG_PREC_FP64 (default) or HGDG_PREC_MIXED_FP32
 * (FP32 storage, for stage derivatives of the mixed path). Contents are reset
 * to zero. */
int hgdg_state_set_storage(hgdg_ctx* ctx, int vec, int precision);

/* out = M^{-1} r( sum_i coeffs[i] * vec[in_ids[i]] ): stage combination,
 * volume + face integrals and inverse-mass apply in one pass.
 * Replaces HpgemAPISimplified::computeTimeDerivative(ids, coeffs, out, t)
 * (kernel/Base/HpgemAPISimplified_Impl.h:777-786), i.e.
 * getLinearCombinationOfVectors + computeRightHandSide + solveMassMatrixEquations. */
int hgdg_time_derivative(hgdg_ctx* ctx, const int* in_ids, const double* coeffs, int n_in, int out_id,
                         double time);

/* out = r( sum_i coeffs[i] * vec[in_ids[i]] ): the stage combination and the
 * volume + face integrals, before the inverse mass (r = M k, the block-diagonal
 * mass of the operator, acoustic: c_inv on the s fields; the fused kernel's k
 * multiplied back by M). out is stored in FP64. The operators are autonomous:
 * time is taken for the signature's sake; a time-dependent source is added to
 * r by the caller (hgdg_scale_and_add of a source vector), as
 * HpgemAPILinear::addSourceTerm does.
 * Replaces HpgemAPILinear::computeRightHandSide(ids, coeffs, out, t)
 * (kernel/Base/HpgemAPILinear_Impl.h:485-498 -> multiplyStiffnessMatrices,
 * :360-434). */
int hgdg_right_hand_side(hgdg_ctx* ctx, const int* in_ids, const double* coeffs, int n_in, int out_id,
                         double time);

/* vec = M^{-1} vec in place, M the operator's block-diagonal mass matrix
 * (acoustic: c_inv on the s fields), with the reference element's inverse
 * mass formed in long double; vec stored in FP64.
 * Replaces HpgemAPILinear::solveMassMatrixEquations(id)
 * (kernel/Base/HpgemAPILinear_Impl.h:143-156). */
int hgdg_solve_mass(hgdg_ctx* ctx, int vec);

/* vec[to] += scale * vec[add]  (HpgemAPISimplified::scaleAndAddVector, :707-717). */
int hgdg_scale_and_add(hgdg_ctx* ctx, int to, int add, double scale);

/* n_steps steps of an explicit Runge-Kutta tableau with `stages` stages on
 * vector `solution`, stage derivatives in vectors aux_ids[0..stages-1].
 * a is stages x stages row-major (strictly lower), b has `stages` entries.
 * Replaces HpgemAPISimplified::computeOneTimeStep (:788-820), looped.
 * The step is captured once into a HIP graph and replayed. */
int hgdg_rk_steps(hgdg_ctx* ctx, int solution, const int* aux_ids, int stages, const double* a,
                  const double* b, double dt, int n_steps);

/* Error estimate of the last step of an embedded tableau and the size of
 * the solution: error = || sum_j dt err_coef[j] k_j ||_2 over all
 * coefficients, norm = sum over elements of the Euclidean norm of the
 * element's coefficients. Replaces HpgemAPISimplified::
 * computeErrorAndNormOfUpdate (kernel/Base/HpgemAPISimplified_Impl.h:718-745). */
int hgdg_step_error(hgdg_ctx* ctx, int solution, const int* aux_ids, int stages, const double* err_coef, double dt,
                    double* error, double* norm);

/* hgdg_step_error with the same inputs and results, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). No counterpart in the reference. It calls no hipMalloc,
 * hipFree or other allocating routine: the per-block partial sums live in a
 * fixed device array of 1 MiB that the library holds on every GPU it runs on
 * until the process ends, reduced on the device, and 16 bytes are read back.
 * Limit: meshes of at most 262144 elements (HGDG_ERR_UNSUPPORTED beyond).
 * All calls on one GPU share that array: calls that overlap on one GPU race.
 * The sums are reduced in a different order than hgdg_step_error's, so the
 * two agree to rounding (~1e-16 relative), not bit for bit. */
int hgdg_step_error_noalloc(hgdg_ctx* ctx, int solution, const int* aux_ids, int stages, const double* err_coef,
                            double dt, double* error, double* norm);

/* One adaptive step with an embedded tableau (err_coef = b - b_low, order =
 * the tableau's order): try dt = min(estimate, dt_max), rewind and retry with
 * dt *= 0.8 (tol * norm / error)^(1/order) while error > max_rel_error * norm,
 * then set the next estimate to 0.9 dt (tol * norm / error)^(1/order) and
 * advance *time by the accepted dt. The first call starts from dt_initial;
 * the estimate is kept per context. Replaces HpgemAPISimplified::
 * computeOneTimeStep(time, maximumRelativeError, dtMax) (:822-861).
 * dt_used, error and norm may be NULL. */
int hgdg_rk_adaptive_step(hgdg_ctx* ctx, int solution, const int* aux_ids, int stages, const double* a,
                          const double* b, const double* err_coef, int order, double max_rel_error, double dt_max,
                          double dt_initial, double* time, double* dt_used, double* error, double* norm);

/* vec = L2 projection of f(x) = sum_m amp[m] sin(2 pi k[m] . x), k given as
 * [n_modes][3] (unused components ignored), with hpGEM's element rule of
 * order 2p+1 (squares, cubes and tetrahedra), on the device. Replaces HpgemAPISimplified::setInitialSolution
 * (kernel/Base/HpgemAPISimplified_Impl.h:227-291,751-758) for this family of
 * initial conditions. */
int hgdg_project_modes(hgdg_ctx* ctx, int vec, int n_modes, const double* amp, const double* k);

/* The same projection into one field of a multi-field operator (field = -1:
 * every field); the other fields of the vector are left as they are. */
int hgdg_project_field_modes(hgdg_ctx* ctx, int vec, int field, int n_modes, const double* amp, const double* k);

/* Error of vec against the exact solution f(x - shift), f(x) = sum_m amp[m]
 * sin(2 pi k[m] . x) (for advection, shift = a t), in one field or summed over
 * all fields (field = -1): l2 = sqrt(sum_e int_e |f - u_h|^2) with the rule of
 * order 2p+2, max_abs = the largest |f - u_h| at the element rule's points.
 * Replaces HpgemAPISimplified::computeTotalError / computeMaxError
 * (kernel/Base/HpgemAPISimplified_Impl.h:293-530), with hpGEM's rules, on the
 * device (squares, cubes and tetrahedra). */
int hgdg_error_modes(hgdg_ctx* ctx, int vec, int field, int n_modes, const double* amp, const double* k,
                     const double* shift, double* l2, double* max_abs);

/* vec (one field, or every field with field = -1) = L2 projection of the
 * Gaussian pulse f(x) = amplitude exp(-|x - center|^2 / width^2), center
 * given as 3 numbers (the third ignored in 2D). Same role as
 * hgdg_project_field_modes (HpgemAPISimplified::setInitialSolution with a
 * pulse as the initial function). */
int hgdg_project_gaussian(hgdg_ctx* ctx, int vec, int field, const double* center, double width, double amplitude);

/* L2 norm of (vec[x] - vec[y]) under the element mass matrices, and the max
 * abs coefficient difference (FP64, on the device). */
int hgdg_state_diff(hgdg_ctx* ctx, int x, int y, double* l2, double* max_abs);

/* vec[dst] = vec[src] (same storage), on the device.
 * Replaces HpgemAPIBase::copyTimeIntegrationToTimeLevelData-style vector copies
 * (kernel/Base/HpgemAPIBase_Impl.h:158-176). */
int hgdg_copy(hgdg_ctx* ctx, int dst, int src);

/* Block until all queued work on the context's stream is done. */
int hgdg_sync(hgdg_ctx* ctx);

/* The context's hipStream_t, for callers timing with HIP events. */
int hgdg_get_stream(hgdg_ctx* ctx, void** stream);

/* ---- serving: many independent solves with continuous batching ----------
 * A solve: periodic unit square/cube with cells[d] cells per direction, one
 * operator and order, sine-mode initial data (acoustic: into the pressure-like
 * field v), n_steps classic RK4 steps of size dt. Solves are admitted in
 * arrival order when their HBM footprint fits the budget and a slot is free;
 * resident solves advance together, one batched launch per operator family
 * and stage. No reference counterpart: hpGEM runs one problem per process
 * (HpgemAPISimplified::solve, kernel/Base/HpgemAPISimplified_Impl.h:936-1074). */
typedef struct hgdg_solve_desc {
    int32_t dim;            /* 2 or 3                                          */
    int32_t kind;           /* HGDG_OP_*                                       */
    int32_t order;          /* 1..5                                            */
    int32_t cells[3];
    double velocity[3];     /* advection                                       */
    double wave_speed;      /* acoustic: c (c_inv = 1/c^2 in every element)    */
    int32_t n_modes;
    const double* amp;      /* [n_modes]                                       */
    const double* k;        /* [n_modes][3]                                    */
    double dt;
    int32_t n_steps;
    int32_t keep_state;     /* 1: keep the final state for hgdg_server_result */
} hgdg_solve_desc;

typedef struct hgdg_server_stats {
    int64_t submitted, completed, failed, active;
    double latency_p50_s, latency_p99_s, latency_mean_s; /* submit -> result ready */
    double dof_rhs_per_s;   /* sum of DOF x RHS evaluations / wall time        */
    double wall_s;          /* first submit -> last completion                 */
    int64_t peak_bytes;     /* peak admitted footprint                         */
    int64_t ticks;          /* worker ticks (RK steps of the resident set)     */
    int64_t restages;       /* re-stagings (resident set changed: new graph)   */
    double restage_s;       /* worker wall time spent re-staging               */
    double tick_s;          /* worker wall time in ticks, re-staging included  */
} hgdg_server_stats;

typedef struct hgdg_server hgdg_server;

int hgdg_server_create(hgdg_server** out, int device, int precision, int64_t hbm_budget_bytes, int max_active);
/* Non-blocking; the ticket identifies the solve. */
int hgdg_server_submit(hgdg_server* s, const hgdg_solve_desc* d, int64_t* ticket);
/* Wait up to timeout_s for one ticket (or every submitted one with ticket = -1). */
int hgdg_server_wait(hgdg_server* s, int64_t ticket, double timeout_s, int* done);
/* Result of a finished solve: DOF count, final state in reference layout
 * (if kept; state may be NULL), Euclidean norm of the final coefficients and
 * the submit-to-done latency. */
int hgdg_server_result(hgdg_server* s, int64_t ticket, int64_t* n_dof, double* state, double* norm,
                       double* latency_s);
int hgdg_server_get_stats(hgdg_server* s, hgdg_server_stats* out);
/* Kernel time per family: with profiling on, each RK step is issued as its
 * separate launches between HIP events instead of one graph launch (the same
 * launches: results do not change; host overhead grows), and the events'
 * elapsed times are summed per (operator kind, dim, order) family and for the
 * RK update pass. No counterpart in the reference (a measurement of this
 * implementation's batched launches). */
int hgdg_server_set_profiling(hgdg_server* s, int enable);
/* Up to `capacity` families: kind, dim, order, summed kernel ms and launch
 * count each (any array may be NULL); n_families = how many there are;
 * update_ms = summed time of separate RK update passes (may be NULL): 0, the
 * updates run inside each family's last-stage launch and count in its time. */
int hgdg_server_family_times(hgdg_server* s, int capacity, int32_t* kind, int32_t* dim, int32_t* order,
                             double* kernel_ms, int64_t* launches, int* n_families, double* update_ms);
int hgdg_server_destroy(hgdg_server* s);

/* ---- host-only utilities (no device needed) ------------------------------ */

/* Number of DG basis functions of hpGEM's H1 set for shape and order. */
int hgdg_basis_size(int shape, int p);

/* Values of the library's basis (reference order) at one reference point. */
int hgdg_basis_eval(int shape, int p, const double* point, double* values);

/* Structured mesh of [lo,hi] with n[d] cells per direction, periodic flags,
 * optional simplex split (5 tets per cube); element and node order of
 * hpGEM's Preprocessor (applications/Preprocessor/lib/hpgem.cpp:100-235).
 * Two-phase: call with coords == NULL to get n_elements and n_vertices. */
int hgdg_structured_mesh(int dim, const int32_t* n, const double* lo, const double* hi,
                         const int32_t* periodic, int simplices, int64_t* n_elements,
                         int32_t* n_vertices, int32_t* shape, double* coords, int64_t* node_ids);

/* ---- interior-penalty Maxwell assembly on tetrahedra ---------------------
 * DG-Max's curl-curl stiffness and permittivity mass
 * (applications/DG-Max/Algorithms/DGMaxDiscretization.cpp:193-260 element
 * matrices, :344-410 face matrices: -1/2 ({curl u}.[v]_T + [u]_T.{curl v})
 * + stab/diam [u]_T.[v]_T inside, the same with factor -1 on the boundary)
 * with hpGEM's Ainsworth-Coyle H(curl) set. Replaces
 * DGMaxDiscretization::computeElementIntegrands/computeFaceIntegrals with
 * the assembly into a global matrix (kernel/Utilities/GlobalMatrix.cpp:235-366). */
typedef struct hgdg_maxwell_desc {
    int32_t order;            /* 1 or 2                                          */
    const double* epsilon;    /* [n_elements] relative permittivity, mesh order  */
    double stab;              /* penalty: stab / (face diameter)                 */
} hgdg_maxwell_desc;

/* Needs a tetrahedral mesh on the context (the element orders of the mesh
 * description are not used). First call with row_ptr == NULL for n_blocks and
 * block_size (n = 12 or 30); then row_ptr [n_elements + 1], col_idx
 * [n_blocks] (block rows and columns are mesh elements, columns ascending),
 * K [n_blocks][n][n] and M [n_elements][n][n], row = test function of the row
 * element, column = trial function of the column element; kernel_ms (may be
 * NULL) receives the assembly kernel's duration. */
int hgdg_maxwell_assemble(hgdg_ctx* ctx, const hgdg_maxwell_desc* d, int64_t* n_blocks, int32_t* block_size,
                          int64_t* row_ptr, int64_t* col_idx, double* K, double* M, double* kernel_ms);

/* hpGEM's Ainsworth-Coyle H(curl) DG set on the reference tetrahedron
 * (createDGBasisFunctionSet3DAinsworthCoyle, kernel/FE/BasisFunctions3D
 * AinsworthCoyle.cpp:417-505), orders 1 and 2: size, and values and curls
 * ([n][3] each) at one reference point. */
int hgdg_hcurl_basis_size(int p);
int hgdg_hcurl_basis_eval(int p, const double* point, double* values, double* curls);

/* Sizes of a .hpgem mesh file (format version 1 or 2, one partition). */
typedef struct hgdg_mesh_file_info {
    int32_t version, dim, shape, n_vertices; /* shape: HGDG_SHAPE_*, one type per file */
    int64_t n_nodes, n_elements;
    int64_t n_faces, n_boundary_faces;       /* codim-1 entities listed in the file     */
    int64_t n_edges;                         /* codim-2 entities (3D)                   */
    int32_t n_zones;
    int32_t zone_name_bytes;                 /* '\n'-separated names plus the final NUL */
} hgdg_mesh_file_info;

/* Read a .hpgem file. Replaces MeshManipulator::readMesh and
 * MeshFileInformation::readInformation (kernel/Base/MeshManipulator_Impl.h:1505-1800,
 * kernel/Base/MeshFileInformation.cpp:48-103). Fills info; every array that is
 * not NULL receives, in file order:
 *   coords        [n_elements][n_vertices][dim]  coordinate copy each element uses
 *   node_ids      [n_elements][n_vertices]       node (periodic copies share one)
 *   zones         [n_elements]                   zone index
 *   face_elements [n_faces][2]                   element on each side, -1 for none
 *   face_local    [n_faces][2]                   local face index, -1 for none
 *   zone_names    zone_name_bytes chars
 * coords and node_ids are what hgdg_mesh_desc takes. */
int hgdg_mesh_read(const char* path, hgdg_mesh_file_info* info, double* coords, int64_t* node_ids,
                   int32_t* zones, int64_t* face_elements, int32_t* face_local, char* zone_names);

/* Write a mesh as a .hpgem file (format version 2, one partition), as the
 * Preprocessor's outputMesh does (applications/Preprocessor/lib/output_impl.h:
 * 91-208): nodes with their distinct coordinate copies in ascending order,
 * elements, then the codim-1 and codim-2 entities numbered by first
 * appearance with the Preprocessor's local entity order
 * (applications/Preprocessor/lib/mesh/ElementShapes.cpp:74-169). Arrays as
 * hgdg_mesh_read returns them: coords [n_elements][vertices][dim], node_ids
 * [n_elements][vertices] numbering the nodes 0 .. n-1; zones may be NULL (all
 * in zone 0), zone_names '\n'-separated or NULL ("Main"). */
int hgdg_mesh_write(const char* path, int shape, int64_t n_elements, const double* coords, const int64_t* node_ids,
                    const int32_t* zones, const char* zone_names);

/* Write a state as VTK XML (order 1), as hpGEM's VTKSpecificTimeWriter does
 * (kernel/Output/VTKSpecificTimeWriter_Impl.h, VTKStandardElements.cpp):
 * <base_name>.pvtu naming one piece <base_name>.0.vtu, one cell per element
 * (VTK quad / hexahedron / tetra), the points being each element's own
 * vertices and each field's value there, arrays base64 "binary" with a uint32
 * byte count. coords as hgdg_mesh_desc takes them; orders per element or NULL
 * (uniform_order); state in hpGEM's layout (as hgdg_state_get returns it),
 * state_len doubles;
 * field_names '\n'-separated, one per field, or NULL ("field0", ...). Host
 * only: needs no device. */
int hgdg_vtk_write(const char* base_name, int shape, int64_t n_elements, const double* coords, const int32_t* orders,
                   int uniform_order, int n_fields, const double* state, int64_t state_len, const char* field_names);

#ifdef __cplusplus
}
#endif

#endif /* HGDG_H */
