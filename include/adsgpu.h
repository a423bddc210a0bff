/*
 * adsgpu -- C ABI of the MI355X-native alternating-direction (ADS) time step.
 *
 * One handle = one simulation_2d / simulation_3d of the reference (IGA-ADS 0.1.0) on one
 * GPU: the B-spline dimensions, their 1D band matrices and FP64 factors, and the device
 * state vectors u, u_prev and buffer. Each entry point replaces the reference call cited
 * next to it; "file:line" refers to the reference tree.
 *
 * Conventions (all entry points):
 *  - return value: ADSG_OK (0) on success; > 0 is a LAPACK-style zero-pivot index from a
 *    factorisation; < 0 is one of the ADSG_E* codes. Nothing throws across this ABI.
 *    adsg_last_error(h) gives the message of the handle's last failure,
 *    adsg_last_error(NULL) the message of the last failed adsg_create on this thread.
 *  - host arrays are borrowed for the duration of the call only; layouts are the
 *    reference's: vectors first-index-fastest (lin/tensor/base.hpp:14,
 *    util/multi_array/ordering/reverse.hpp:28-31), u(i0,i1,i2) at i0 + n0*(i1 + n1*i2);
 *    band matrices in LAPACK band storage (lin/band_matrix.hpp:66-73).
 *  - a handle owns one HIP stream; calls on a handle are asynchronous with respect to the
 *    host except the ones that return host data (adsg_get_vector) and adsg_synchronize.
 *    A handle is not thread-safe; different handles may be used concurrently.
 */
#ifndef ADSGPU_H
#define ADSGPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define ADSG_OK 0
#define ADSG_EINVAL (-1)   /* bad argument or unsupported configuration */
#define ADSG_EHIP (-2)     /* HIP runtime error (message in adsg_last_error) */
#define ADSG_ENOMEM (-3)   /* device allocation failed */
#define ADSG_ESTATE (-4)   /* call out of order (e.g. solve before prepare_matrices) */
#define ADSG_ENONFINITE (-5) /* a state check found NaN/Inf */

/* arithmetic of the step */
#define ADSG_FP64 0 /* FP64 throughout */
#define ADSG_FP32 1 /* FP32 contractions and substitution; factors computed in FP64 */

/* vectors held by a handle */
#define ADSG_U 0      /* the state / right-hand side being solved (simulation's u) */
#define ADSG_U_PREV 1 /* previous state (simulation's u_prev) */
#define ADSG_BUFFER 2 /* scratch of the same shape (simulation_2d.cpp:19 buffer) */
#define ADSG_KFIELD 3 /* coefficients of a diffusion field k(x) on the same basis */
#define ADSG_SOURCE 4 /* coefficients of a source f(x) on the same basis */
#define ADSG_GRAD_K 5 /* accumulated dL/dk of adsg_adjoint_step (FP64 in both precisions) */
#define ADSG_GRAD_F 6 /* accumulated dL/df of the adjoint steps for forms with a source (FP64;
                         accumulated only when the caller has set it, e.g. to zeros) */

typedef struct adsg_sim adsg_sim;

/* ads::dim_config (simulation/config.hpp:11-29). quad_order <= 0 means p + 1; accepted:
 * p 1..3 per direction, quad_order 2..64 (quad/gauss_data.cpp:8-38), repeated_nodes 0..p
 * (bspline.cpp:26-43: every interior knot repeated repeated_nodes + 1 times). A handle with
 * quad_order != p + 1, repeated nodes or a degree per direction assembles its rhs with the
 * general element loop (one lane per element, the reference's per-point loop); on such a
 * handle adsg_norm, adsg_error, adsg_norm_noalloc, adsg_project(_device) and the k-field
 * gradient of the explicit adjoint return ADSG_EINVAL, and so do adsg_sample and FP32
 * Dirichlet steps when the degree differs between directions. */
typedef struct adsg_dim_config {
    int p;
    int elements;
    double a;
    double b;
    int quad_order;
    int repeated_nodes;
} adsg_dim_config;

/* ads::config_2d / config_3d (simulation/config.hpp:46-56) plus the dimension count and the
 * arithmetic (ADSG_FP64 / ADSG_FP32). */
typedef struct adsg_config {
    int dim; /* 2 or 3 */
    adsg_dim_config x, y, z;
    int step_count;
    double dt;
    int derivatives;
    int precision;
} adsg_config;

/* The integrand of compute_rhs at a Gauss point, for test function v and state u:
 *   mass*u*v - dt*k*(grad[0]*u_x*v_x + grad[1]*u_y*v_y + grad[2]*u_z*v_z)
 *            + dt*reaction*u*(1-u)*v + dt*f*v
 * with k = 1 unless use_kfield (k from ADSG_KFIELD) and f = 0 unless use_source.
 * Explicit heat (validation.hpp:110-115, heat_3d.hpp:60-63): mass 1, grad (1,1,1).
 * Directional split terms (implicit.hpp:132-182): dt -> dt/2 and one grad entry. */
typedef struct adsg_rhs_form {
    double mass;
    double dt;
    double grad[3];
    double reaction;
    int use_kfield;
    int use_source;
} adsg_rhs_form;

/* simulation_2d(config) / simulation_3d(config) (simulation_2d.cpp:8-19,
 * simulation_3d.cpp:8-22): builds each dimension (dimension.cpp:8-18: knots, tabulated
 * basis, Gram matrix) and allocates u, u_prev and buffer on the current HIP device. */
int adsg_create(const adsg_config* cfg, adsg_sim** out);
void adsg_destroy(adsg_sim* h);
const char* adsg_last_error(const adsg_sim* h);

int adsg_dofs(const adsg_sim* h, int dir);   /* dimension::dofs (dimension.hpp:35) */
long adsg_size(const adsg_sim* h);           /* number of coefficients */
int adsg_precision(const adsg_sim* h);

/* dimension::fix_dof (dimension.cpp:23-29) on the direction's Gram matrix (slot 0). */
int adsg_fix_dof(adsg_sim* h, int dir, int k);

/* A further 1D matrix of direction dir in slot >= 1: mass*M + stiff*K built from the
 * direction's tables like implicit.hpp:46-64 (matrix(): va*vb + h*da*db). */
int adsg_set_matrix(adsg_sim* h, int dir, int slot, double mass, double stiff);

/* A user band matrix for direction dir, slot >= 0: lin::band_matrix(kl, ku, n) full buffer
 * (band_matrix.hpp:31-32, column size 2kl+ku+1), n = adsg_dofs(h, dir). */
int adsg_set_band_matrix(adsg_sim* h, int dir, int slot, int kl, int ku, const double* ab);

/* The generalized ADS (solver.hpp:43-195, tests solver_test.cpp:257-534): direction dir gets
 * one band matrix per line in slot `slot` (>= 1). ab holds the lines' LAPACK full buffers
 * ((2kl+ku+1) x n each) one after another, lines numbered by the remaining indices in the
 * reference's cyclic order (dir+1 fastest: x lines by (y, z), y lines by (z, x), z lines by
 * (x, y)). ads_solve with that slot solves this direction first and the others after it in
 * cyclic order, as the reference does with a callable in that position. */
int adsg_set_line_matrices(adsg_sim* h, int dir, int slot, int kl, int ku, const double* ab);

/* prepare_matrices (simulation_2d.hpp:56-59, simulation_3d.hpp:58-62): factor every slot
 * in FP64 (lin::factorize -> dgbtrf, band_solve.hpp:16-18) and upload the factors.
 * Returns the first nonzero LAPACK info (zero pivot) if any. */
int adsg_prepare_matrices(adsg_sim* h);

/* copy a vector in/out (FP64 host data whatever the precision). */
int adsg_set_vector(adsg_sim* h, int which, const double* host);
int adsg_get_vector(adsg_sim* h, int which, double* host);
/* device pointer of a vector (double* or float* by precision), for zero-copy interop.
 * Steps with a source use its load vector (integral of f v), computed once per source:
 * taking ADSG_SOURCE's pointer marks it stale, so write the source through the pointer
 * before the next step, or take the pointer again after writing. */
void* adsg_device_ptr(adsg_sim* h, int which);
/* the same copies from/to FP64 device memory of the caller (e.g. a framework's tensors);
 * they complete before returning. */
int adsg_set_vector_device(adsg_sim* h, int which, const double* dev);
int adsg_get_vector_device(adsg_sim* h, int which, double* dev);

/* before_step's swap(u, u_prev) (validation.hpp:73-76). */
int adsg_swap(adsg_sim* h);

/* compute_rhs: u := element-loop assembly over u_prev (validation.hpp:95-119). */
int adsg_compute_rhs(adsg_sim* h, const adsg_rhs_form* form);

/* validation.hpp:31-41: zero the boundary coefficients of u (all faces in 3D). */
int adsg_zero_boundary(adsg_sim* h);

/* ads_solve(u, buffer, dims...) (solver.hpp:222-226): u := (A_x (x) A_y [(x) A_z])^-1 u
 * with A_d the factored matrix in slots[d] (NULL = slot 0 everywhere = solve(),
 * simulation_2d.hpp:41). */
int adsg_ads_solve(adsg_sim* h, const int* slots);

/* normL2 / normH1 of basic_simulation_2d.hpp:291-316 (basic_simulation_3d.hpp likewise) on
 * the device: sqrt(sum_e sum_q w J phi(u)) of u (which = ADSG_U or ADSG_U_PREV) with
 * phi = u^2 (ADSG_NORM_L2), |grad u|^2 (ADSG_NORM_H10, the H10 struct of :46-48) or
 * u^2 + |grad u|^2 (ADSG_NORM_H1). Summed in FP64 in a fixed order (deterministic); FP32
 * handles evaluate with their FP32 basis tables. */
#define ADSG_NORM_L2 0
#define ADSG_NORM_H10 1
#define ADSG_NORM_H1 2
int adsg_norm(adsg_sim* h, int which, int kind, double* out);
/* error(u, Ux, Uy, norm, fun) and norm(Ux, Uy, norm, fun) (basic_simulation_2d.hpp:266-362,
 * basic_simulation_3d.hpp:281-400): out = sqrt(sum_e sum_q w J phi(u_h(x_q) - s(x_q))) with
 * u_h the spline of vector `which` (ADSG_U or ADSG_U_PREV; which = -1 takes u_h = 0, the
 * norm of s) and phi of kind ADSG_NORM_L2 / _H10 / _H1 as in adsg_norm. s is given at the
 * Gauss points (layout of adsg_project) as planes of Q = prod_d elements_d*(p+1) values:
 * [value] for ADSG_NORM_L2, [value | d/dx | d/dy (| d/dz)] otherwise; host memory. The
 * reference's relative errors (error_relative_L2/H1, :338-356) are the ratio of the two. */
int adsg_error(adsg_sim* h, int which, int kind, const double* sq, double* out);
/* adsg_norm (sq_dev = NULL) and adsg_error with the function values already in device memory
 * (sq_dev, same layout), allocating nothing: the block partial sums live in a static device
 * array of 1025 doubles per GPU (up to 1024 partials and the total; the kernel strides over
 * the elements, so every mesh fits and the same partials as adsg_norm / adsg_error are
 * formed, giving the same bits). No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). All calls on one GPU share that array: calls that overlap on one
 * GPU (handles on different streams used concurrently) race. Returns the host result after
 * synchronising the handle's stream. */
int adsg_norm_noalloc(adsg_sim* h, int which, int kind, const double* sq_dev, double* out);

/* projection(v, f) (simulation_2d.hpp:43-46 / simulation_3d.hpp:45-48 -> projector,
 * projection.hpp:11-153): adds the load vector of the L2 projection of f,
 *   v_i += sum_e sum_q f(x_q) B_i(x_q) w_q J_e,
 * into vector `which`, which must hold FP64 coefficients (ADSG_U, ADSG_U_PREV, ADSG_GRAD_K, ADSG_GRAD_F,
 * and in FP64 handles every vector; it is allocated and zeroed if absent). f is given at
 * the Gauss points: fq[ix + NX*(iy + NY*iz)] with ix = e_x*(p+1) + q_x, NX = elements_x*(p+1)
 * (likewise y and z; adsg_quad_points returns the abscissae). As in the reference's
 * before() (validation.hpp:65-71), adsg_ads_solve then turns the load into coefficients
 * (FP64 handles). Accumulated in FP64; FP32 handles use their FP32 basis tables.
 * adsg_project reads fq from host memory, adsg_project_device from device memory. */
int adsg_project(adsg_sim* h, int which, const double* fq);
int adsg_project_device(adsg_sim* h, int which, const double* fq);
/* bspline::eval / eval_ders (bspline/eval.hpp:73-234), as output_manager samples a solution
 * (output_manager.hpp:28-146): the spline of vector `which` (an FP64 vector) on the tensor
 * grid px[0..counts[0]) x py [x pz] (host abscissae; points outside [a, b] use the end
 * elements, as find_span does), into out (host): prod(counts) values, x fastest, and with
 * ders = 1 the d/dx, d/dy [, d/dz] planes after them. */
int adsg_sample(adsg_sim* h, int which, const int* counts, const double* px, const double* py,
                const double* pz, int ders, double* out);
/* the Gauss points of direction dir (basis_data x[e][q], basis_data.cpp:63-114):
 * elements*(p+1) abscissae, element-major. */
int adsg_quad_points(const adsg_sim* h, int dir, double* x);

/* The transpose of ads_solve: u := (A_x (x) A_y [(x) A_z])^-T u, direction by direction
 * with dgbtrs 'T' semantics on the same FP64 factors (band_solve.hpp:20-31 with
 * trans = 'T'; the reference calls dgbtrs only with 'N'). Needed by adjoint steps. FP64 handles; every
 * slot used must have unpivoted factors and lines of 2..16 segments (true for M and
 * M + h K of degree-p B-splines with 32 <= n <= 512). */
int adsg_ads_solve_transposed(adsg_sim* h, const int* slots);

/* nsteps of simulation_base::run's loop body (simulation_base.cpp:14-18) as in
 * validation.hpp:73-81: swap; compute_rhs(form); [zero_boundary]; ads_solve(slot 0).
 * Replayed from a HIP graph captured once per (form, flags). */
#define ADSG_STEP_DIRICHLET 1
/* ADSG_STEP_CHECK_FINITE (adsg_step, adsg_step_delta): after the steps, count the NaN / Inf
 * coefficients of u and return ADSG_ENONFINITE if there are any (the message gives the
 * count). A NaN or Inf reaches every coefficient its support touches within a step and no
 * step removes it, so one check after nsteps steps finds it; the reference ignores such
 * states (and LAPACK's info, band_solve.hpp:16-31). The check reads u once (about 3 us for
 * 2.2 M coefficients) and synchronises the handle's stream. */
#define ADSG_STEP_CHECK_FINITE 4
int adsg_step(adsg_sim* h, int nsteps, const adsg_rhs_form* form, int flags);

/* nsteps of the linearly implicit step in increment (delta) form:
 *   swap; buffer := r(u_prev) (the integrand of form, mass term included as given);
 *   [zero the boundary entries of buffer]; buffer := (A_x (x) A_y [(x) A_z])^-1 buffer with
 *   A_d the matrix in slots[d]; u := u_prev + buffer [boundary coefficients of u := 0].
 * With form {mass = 0, dt, source} and slots holding M_d + theta*dt*K_d this is the
 * approximately factored theta scheme, the all-direction counterpart of the directional
 * splitting of implicit.hpp:102-111 and elasticity/implicit.hpp:464-511. In FP32 the
 * increment is assembled and solved in FP32 on the FP64 factors; ADSG_STEP_REFINE adds one
 * refinement of every line solve against its FP64 residual. Replayed from captured graphs. */
#define ADSG_STEP_REFINE 2
int adsg_step_delta(adsg_sim* h, int nsteps, const adsg_rhs_form* form, const int* slots, int flags);

/* One step of the adjoint of adsg_step (explicit step, FP64 handles) for a form linear in u
 * (reaction = 0): with u holding lam(n+1) and u_prev holding the forward state u(n) of that
 * step, for u(n+1) = S r(u(n)), S = M^-1 Z (slot-0 solve, Z the boundary zeroing of
 * ADSG_STEP_DIRICHLET) and r(u) = R(k) u + c:
 *   w = Z M^-T lam(n+1);  lam(n) = R(k) w  (R symmetric);
 *   if use_kfield: grad_k += d/dk [w^T R(k) u(n)] = -dt sum_e sum_q w_q J_e B_b sum_d grad_d
 *   (d_d u)(d_d w)  (added into ADSG_GRAD_K, which the caller sets, e.g. to zeros, first);
 *   if use_source and ADSG_GRAD_F is set: grad_f += dt M w (c = dt M f, M the Gram matrix).
 * buffer is overwritten, and u_prev too when grad_f is accumulated. No counterpart in the
 * reference (it has no adjoint code). */
int adsg_adjoint_step(adsg_sim* h, const adsg_rhs_form* form, int flags);

/* Checkpoint-and-recompute for the explicit-step adjoint (FP64 handles). The forward call
 * runs nsteps steps of adsg_step(form, flags) from u, keeping u every `every` steps in
 * device memory held by the handle. The adjoint call takes u = dL/du(T), runs
 * adsg_adjoint_step backwards over the same steps -- each segment of `every` steps is
 * recomputed from its checkpoint -- and returns u = dL/du(0), adding dL/dk into
 * ADSG_GRAD_K when the form uses the k field (and dL/df into ADSG_GRAD_F when the form uses
 * a source and that vector is set). u_prev and buffer are overwritten. The segments after
 * the first are replayed from HIP graphs the handle captures and keeps (freed with the
 * handle's other step graphs). No counterpart in the reference (it keeps no checkpoints
 * and has no adjoint). */
int adsg_forward_checkpointed(adsg_sim* h, int nsteps, int every, const adsg_rhs_form* form, int flags);
int adsg_adjoint_checkpointed(adsg_sim* h, const adsg_rhs_form* form, int flags);

/* nsteps of the adjoint of adsg_step_delta with the same form, slots and flags, applied to
 * u in place (no swap; u_prev and buffer are overwritten): for the step
 * u' = Z (u + A^-1 Z r(u)) with r(u) = R u + c linear (reaction = 0) and Z the boundary
 * zeroing of ADSG_STEP_DIRICHLET, u := Z u + R Z A^-T Z u. R = mass*M - dt*K is symmetric,
 * so R applies as the form's integrand without its source. Starting from dL/du(T) this
 * yields dL/du(0) after the same number of steps; for a form with a source and ADSG_GRAD_F
 * set, dL/df += dt M Z A^-T Z u per step (c = dt M f). No counterpart in the reference (it
 * has no adjoint code); FP64 handles. */
int adsg_adjoint_step_delta(adsg_sim* h, int nsteps, const adsg_rhs_form* form, const int* slots,
                            int flags);

int adsg_synchronize(adsg_sim* h);

/* Checkpoint and resume (no counterpart in the reference, whose text dumps of coefficients,
 * implicit.hpp:121-127, are not read back into a run): adsg_save_state writes u and u_prev
 * (FP64) with a 128-byte header -- "ADSGSTAT", version, the configuration (dim, p, elements,
 * dofs, quad_order, repeated_nodes per direction), precision, the caller's step and time and
 * an FNV-1a checksum of the payload -- to a temporary file renamed over path when complete.
 * adsg_load_state checks the header against the handle's configuration and the checksum
 * before it changes the state, returns step and time, and fails with ADSG_EINVAL (message in
 * adsg_last_error) on a foreign, truncated or corrupted file. Both synchronise the handle. */
int adsg_save_state(adsg_sim* h, const char* path, long step, double t);
int adsg_load_state(adsg_sim* h, const char* path, long* step, double* t);

/* The same NaN / Inf count for any FP64 vector of the handle (ADSG_U .. ADSG_GRAD_F; u and
 * u_prev in both precisions): ADSG_OK, or ADSG_ENONFINITE with the count in the message.
 * No counterpart in the reference. Synchronises the handle's stream. */
int adsg_check_finite(adsg_sim* h, int which);

/* Device time per phase of nsteps eagerly launched steps -- adsg_step's sequence when
 * slots is NULL, adsg_step_delta's with these slots otherwise -- measured with HIP events on
 * the handle's stream: the counterpart of the reference's StatTimer around compute_rhs
 * (examples/scalability/test3d.hpp:67-101). Averages per step in milliseconds: ms[0] the
 * element-loop kernels (the 2^dim colour launches, or the one launch over all tiles),
 * ms[1] boundary handling, ms[2], ms[3], ms[4] line sweeps in x, y, z, ms[5] the whole step,
 * ms[6] the start of the rhs (a source's load vector scaled in; the zero fill, where one is
 * needed, is part of ms[0]). With the one-launch tiling the spill fix-up, and the FP64
 * boundary zeroing with it, run at the start of the x sweep and count in ms[2]; without
 * the fix-up the FP64 zeroing is done by the x sweep as it loads its lines. */
int adsg_time_phases(adsg_sim* h, int nsteps, const adsg_rhs_form* form, const int* slots, int flags,
                     double* ms);

/* Kernel tuning (no counterpart in the reference, whose CPU loops have no launch shape):
 * ADSG_TUNE_RHS_VARIANT selects the 3D element-kernel variant (0: element coefficients held
 * in registers, 1: re-read from LDS, 2: as 1 with the element's contribution added to LDS
 * per x Gauss point, 3: two tiles per workgroup, the second tile's data copied to LDS while
 * the first is integrated, integrated like 1 (like 2 at p = 3), 4: each wave integrates two
 * element layers and carries their shared DOF planes in registers (p <= 2; 2 at p = 3));
 * they compute the same sums (variants 2 and 4 add them in other orders; in FP64 variant 3
 * equals 1, or 2 at p = 3, bit for bit). A new handle starts with the variant measured
 * fastest for its degree and precision (p = 3: 2; FP64 p <= 2: 4 from 2^21 elements, else
 * 1; FP32: 1).
 * ADSG_TUNE_SWEEP_VARIANT selects the line sweep (0: lines cut into segments solved by
 * separate waves and joined through FP64-computed influence coefficients, where the factors
 * are unpivoted and the lines are 2..16 segments of at most 32 rows long (up to 512 DOFs),
 * x lines staged through LDS where that fits; 1: one lane per whole line; 2: as 0 with the x
 * lines read lane by lane, faster on some small meshes).
 * adsg_autotune times every variant on the handle's current state with the given form
 * (explicit steps, or delta steps with slots) and keeps the fastest of each; the state
 * vectors are restored afterwards. */
#define ADSG_TUNE_RHS_VARIANT 0
#define ADSG_TUNE_SWEEP_VARIANT 1
/* ADSG_TUNE_RHS_KRONECKER = 1 (default 0; 3D; the autotuner leaves it): forms without a k
 * field and without a reaction term (mass u v - dt sum_d g_d (d_d u)(d_d v), plus a source)
 * are assembled as (mass (x)_d M_d - dt sum_d g_d K_d (x) M_others) u from the 1D Gram and
 * stiffness matrices (form_matrix.cpp:8-42). The p+1-point Gauss rule integrates these
 * products exactly per element, so this is the element loop's sum in another order (equal
 * to rounding) at a few dozen operations per coefficient instead of thousands per element. */
#define ADSG_TUNE_RHS_KRONECKER 3
/* ADSG_TUNE_RHS_COLUMNS (0..3): the shape of element-kernel variant 5 (one wave per column
 * of elements along z, the x-y contractions on the matrix cores): bit 0 two columns per
 * wave, bit 1 three chunks of 16 DOF planes per workgroup; the autotuner times all four. */
#define ADSG_TUNE_RHS_COLUMNS 4
/* ADSG_TUNE_RHS_TILING (2D every form; 3D variants 0, 1, 2 and 4 with forms without a k
 * field or source in the kernel): 0 = 2^dim parity-colour launches, whose tiles share no
 * DOF; 1 = one launch over all tiles, each tile adding the entries of its DOF block that it
 * owns and storing the others in a spill array (allocated with the handle, about one to
 * three values per element), which a second kernel adds in a fixed order. Both are
 * deterministic; the sums are grouped differently (equal to rounding). Default 1 in 3D and
 * for 2D meshes of at most 2048 tiles of 16 x 16 elements, 0 for larger 2D meshes (and in
 * batches, whose problems therefore equal the same problems run alone with 0). The
 * autotuner times both. */
#define ADSG_TUNE_RHS_TILING 5
/* ADSG_TUNE_CACHE (default 1): adsg_autotune records its choices per configuration (device,
 * dimension, precision, degree / elements / Gauss points / repeated nodes per direction, the
 * kind of form, delta or explicit steps, the step flags) for the process, and a handle of a
 * configuration already tuned takes them without timing; 0 on a handle times afresh (and
 * records nothing). */
#define ADSG_TUNE_CACHE 6
int adsg_set_tuning(adsg_sim* h, int key, int value);
int adsg_get_tuning(const adsg_sim* h, int key);
int adsg_autotune(adsg_sim* h, const adsg_rhs_form* form, const int* slots, int flags);

/* ---- batches: many independent 2D problems advanced together --------------------------
 * No counterpart in the reference, which runs one simulation per process; each problem is
 * the explicit step of adsg_step (swap; compute_rhs(form); ads_solve with the Gram matrices,
 * validation.hpp:73-81) with natural boundaries, its own mesh size and degree p = 1..3 (the
 * same in x and y), and a form without k field or source (e.g. the reaction-diffusion form
 * {mass 1, dt, grad eps, reaction 1}). The state of all problems lives in shared arenas;
 * one step is a fixed sequence of grouped launches over all problems (element kernels per
 * degree and colour, x and y sweeps per degree, one update), and nsteps steps are replayed
 * from HIP graphs of at most 50 steps, each captured once per step count. FP32 batches use
 * the increment scheme of ADSG_FP32. The set-up runs on the device: adsg_batch_create
 * tabulates every problem's basis (basis_data.cpp:63-114) in one launch, adsg_batch_prepare
 * assembles the Gram matrices (form_matrix.cpp:8-42), factors them in FP64 (dgbtrf without
 * interchanges, which these matrices never need) and packs the segmented sweep rows in two
 * more; the results equal the host set-up of adsg_create / adsg_prepare_matrices bit for bit.
 * Configurations: dim 2, one precision, p 1..3 the same in x and y, elements >= 1, a < b,
 * quad_order 0 or p + 1, repeated_nodes 0; others are refused with ADSG_EINVAL. */
typedef struct adsg_batch adsg_batch;
int adsg_batch_create(const adsg_config* cfgs, int count, adsg_batch** out);
void adsg_batch_destroy(adsg_batch* b);
const char* adsg_batch_last_error(const adsg_batch* b); /* NULL: last create error */
int adsg_batch_count(const adsg_batch* b);
long adsg_batch_size(const adsg_batch* b, int problem);
int adsg_batch_dofs(const adsg_batch* b, int problem, int dir);
/* assemble and factor every problem's Gram matrices (FP64, on the device) and build the
 * launch lists; > 0: a zero pivot (its row), ADSG_EINVAL: a factor that needs interchanges */
int adsg_batch_prepare(adsg_batch* b);
/* one form per problem (default: explicit heat, dt = 0) */
int adsg_batch_set_forms(adsg_batch* b, const adsg_rhs_form* forms);
/* ADSG_U or ADSG_U_PREV of one problem, FP64 host data */
int adsg_batch_set_vector(adsg_batch* b, int problem, int which, const double* host);
int adsg_batch_get_vector(adsg_batch* b, int problem, int which, double* host);
int adsg_batch_step(adsg_batch* b, int nsteps);
int adsg_batch_synchronize(adsg_batch* b);
/* ADSG_TUNE_BATCH_TILE: element tile of the batched kernels, 0 = 16 x 16, 1 = 8 x 8.
 * ADSG_TUNE_SWEEP_VARIANT: 0 = segmented line sweeps (9 waves per 64 lines; used when every
 * problem's lines split into segments), 1 = one lane per line.
 * adsg_batch_autotune times one step with each tile shape and each line grouping (the two
 * sweep variants), keeps the fastest pair and restores the state; ms_out[0..1] (may be
 * NULL) receives each tile shape's time with its better sweep variant. */
#define ADSG_TUNE_BATCH_TILE 2
int adsg_batch_set_tuning(adsg_batch* b, int key, int value);
int adsg_batch_get_tuning(const adsg_batch* b, int key);
int adsg_batch_autotune(adsg_batch* b, double* ms_out);

/* lin::factorize + lin::solve_with_factorized (band_solve.hpp:16-31) on the host in
 * FP64, for callers that hold their own band matrices: factor ab (n x n, kl, ku, full
 * buffer as above) in place, pivots 1-based like dgbtrf. Returns LAPACK info. */
int adsg_band_factorize(int n, int kl, int ku, double* ab, int* ipiv);

/* Measured FMA ceilings of the current HIP device in TFLOP/s, from kernels of independent FMA
 * chains that fill it: tflops[0] vector FP64, [1] vector FP32 (plain FMAs, no packed
 * instructions), [2] MFMA FP64 16x16x4, [3] MFMA FP32 16x16x4. No counterpart in the
 * reference; the benchmark reports them beside the element kernels' rates. */
int adsg_device_rates(double* tflops);

#ifdef __cplusplus
}
#endif

#endif /* ADSGPU_H */
