/*
 * subfv_amd C-ABI: the explicit residual path of subfv (subface fluxes and
 * nodal solvers, node->cell gather, limited gradients, stage update) and its
 * BCSR / block-Jacobi / GMRES routines on one MI355X (gfx950).
 *
 * Plain pointers and sizes only.  Every function returns 0 on success or a
 * SUBFV_E_* code; subfv_error() gives the message.  Arrays crossing the
 * boundary use the reference's Fortran layouts: sol(5,C) is C blocks of 5
 * doubles, grad(3,5,C) is C blocks of 15 (d fastest), diag(5,5,C) is C
 * column-major 5x5 blocks, vp(3,V) is V blocks of 3.  Vertex, element and
 * sub-face ids are those of the reference's build_mesh (1-based in exported
 * integer arrays).  A context is not thread-safe; use one per host thread.
 *
 * Reference interfaces replaced (paths relative to the subfv source tree):
 *   subfv_mesh_build            read_mesh_msh_2 (src/mesh/mesh_reading_module.F90:989-1324)
 *                               + build_mesh (src/mesh/mesh_connectivity_module.F90:12-59)
 *                               + compute_geometry_mesh (src/mesh/mesh_geometry_module.F90:16-57)
 *   subfv_set_config            read_input_parameters / init_flags
 *                               (src/ns/ns_global_data_module.F90:170-330), namelist subset
 *   subfv_state_put/get         the host <-> device copies of sol(5, n_elems)
 *   subfv_compute_euler_flux    compute_euler_flux (src/ns/ns_euler_module.F90:151-339)
 *   subfv_compute_cell_grad     compute_cell_grad_from_nodal_grad (src/ns/ns_euler_recon_module.F90:8-59)
 *   subfv_flux_limiting         flux_limiting (src/ns/ns_euler_recon_module.F90:560-601)
 *   subfv_add_timedisc_and_solve add_timedisc_and_solve (src/ns/ns_euler_module.F90:521-609)
 *   subfv_compute_rhs_ns        compute_rhs_ns (src/ns/ns_main.F90:130-179)
 *   subfv_set_diffusion         activate_diffusion and the viscous namelist entries
 *   subfv_set_sutherland        use_sutherland, mu0, T0 (compute_mu)
 *                               (src/ns/ns_global_data_module.F90:8-12,177-213)
 *   subfv_run                   the time loop of src/ns/ns_main.F90:83-107
 *   subfv_bcsr_*                bcsr_matrix_type, bcsr_matmul, bcsr_diag, bcsr_gmres_diag, bcsr_gmres,
 *                               bcsr_gmres_lusgs, bcsr_gmres_parallel_lusgs, bcsr_lusgs,
 *                               bcsr_parallel_lusgs, bcsr_pcgm, add_block_to_bcsr
 *                               (src/linear_algebra/sparse_bcsr_linear_module.F90:6-29,56-71,194-290,
 *                               447-548,885-1037,1110-1169,1268-1335,1414-1502,1592-1691)
 */
#ifndef SUBFV_H
#define SUBFV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SUBFV_OK = 0,
  SUBFV_E_INVALID = 1,      /* bad argument or call order */
  SUBFV_E_DEVICE = 2,       /* HIP runtime failure */
  SUBFV_E_BAD_OMEGA = 3,    /* add_timedisc_and_solve "BAD OMEGA" stop (ns_euler_module.F90:595-598) */
  SUBFV_E_NEG_STAR = 4,     /* multi_point "Negative specific volume" stop (ns_euler_rs_module.F90:257-266) */
  SUBFV_E_MESH = 5,         /* check_mesh failure (mesh_geometry_module.F90:403-408) */
  SUBFV_E_BAD_BC = 6,       /* unknown boundary type (ns_euler_recon_module.F90:404-406) */
  SUBFV_E_NOT_CONVERGED = 7 /* GMRES returned conv = .false. (soft) */
};

/* scheme ids of ns_global_data_module.F90:63-72 */
enum {
  SUBFV_SCHEME_MULTI_POINT = 1,
  SUBFV_SCHEME_MULTI_POINT_ISO = 2,
  SUBFV_SCHEME_THREE_WAVE = 3,
  SUBFV_SCHEME_TWO_WAVE = 4,
  SUBFV_SCHEME_MODIFIED_THREE_WAVE = 5,
  SUBFV_SCHEME_MULTI_POINT_PRESSURE = 6,
  SUBFV_SCHEME_MULTI_POINT_PRESSURE_PH = 8
};

/* Euler BC ids of ns_global_data_module.F90:50-58 (potential-flow BCs excluded) */
enum {
  SUBFV_BC_WALL = 1,
  SUBFV_BC_ADHERENCE_WALL = 2,
  SUBFV_BC_FREESTREAM = 3,
  SUBFV_BC_OUTFLOWSUPERSONIC = 4,
  SUBFV_BC_INFLOW_POND = 5,
  SUBFV_BC_INOUT_DOUBLE_MACH = 6,
  SUBFV_BC_DOUBLE_MACH_BOTTOM = 7
};

#define SUBFV_MNBC 10

/* The &INPUT_PARAM entries the path reads (ns_global_data_module.F90:177-213). */
typedef struct subfv_config {
  int32_t scheme_id;
  int32_t bc_style;
  int32_t boundary_2d;
  int32_t exclude_bound_vert;
  int32_t local_time_step;
  int32_t second_order;
  int32_t method;
  int32_t n_bc;
  int32_t bc_euler_id[SUBFV_MNBC]; /* bc_euler_id(1..n_bc) */
  double bc_val[SUBFV_MNBC][5];    /* bc_val(:, 1..n_bc), primitive */
  double cfl;
  double tmax;
} subfv_config;

typedef struct subfv_ctx subfv_ctx;
typedef struct subfv_host_mesh subfv_host_mesh;

int subfv_create(subfv_ctx **ctx, int device);
void subfv_destroy(subfv_ctx *ctx);
const char *subfv_error(const subfv_ctx *ctx);

/* Mesh in msh 2.2 form: coord(3, n_vert) in file order, gmsh element kinds
 * (4 tet, 5 hex, 6 prism, 7 pyramid) with 1-based vertex ids, optional tagged
 * boundary faces (tag = BC index, 0 = default wall).  Builds the subface
 * connectivity on the device (the reference's numbering), and computes the
 * geometry there. */
int subfv_mesh_build(subfv_ctx *ctx, int32_t n_vert, const double *coord, int32_t n_elems,
                     const int32_t *kind, const int32_t *elem_vert, int32_t n_bfaces,
                     const int32_t *bf_nvert, const int32_t *bf_vert, const int32_t *bf_tag,
                     int32_t boundary_2d);
/* The same from a connectivity built beforehand on the host
 * (subfv_host_mesh_build): upload, device geometry and BC data only. */
int subfv_mesh_upload(subfv_ctx *ctx, const subfv_host_mesh *hm, int32_t boundary_2d);
/* "n_vert", "n_elems", "n_faces", "n_sub_faces", "n_sub_elems", "n_max_sub_faces", "n_groups",
 * "diff_block_doubles" (size of the packed viscous blocks, 0 before the first viscous residual),
 * "device_bytes_used" (device memory the mesh and its work arrays hold in the context) */
int64_t subfv_mesh_count(const subfv_ctx *ctx, const char *what);
/* Connectivity in the reference numbering (same names as the mesh_type
 * components, e.g. "vert_sub_face", "sub_face_left_elem"); returns the number
 * of values or -1. */
int64_t subfv_mesh_get_i32(subfv_ctx *ctx, const char *name, int32_t *out, int64_t cap);
/* Geometry computed on the device: "elem_coord", "elem_volume",
 * "sub_face_norm", "sub_face_area", "face_coord", ... */
int64_t subfv_mesh_get_f64(subfv_ctx *ctx, const char *name, double *out, int64_t cap);

/* The connectivity of subfv_mesh_build (build_mesh, mesh_connectivity_module.F90:12-59)
 * built on the context's device, the same numbering, into a mesh handle that
 * subfv_mesh_upload takes (served requests: build on arrival, then upload). */
int subfv_host_mesh_build_on_device(subfv_host_mesh **out, subfv_ctx *ctx, int32_t n_vert, const double *coord,
                                    int32_t n_elems, const int32_t *kind, const int32_t *elem_vert, int32_t n_bfaces,
                                    const int32_t *bf_nvert, const int32_t *bf_vert, const int32_t *bf_tag);
/* Host-only connectivity build (no device needed). */
int subfv_host_mesh_build(subfv_host_mesh **out, int32_t n_vert, const double *coord, int32_t n_elems,
                          const int32_t *kind, const int32_t *elem_vert, int32_t n_bfaces,
                          const int32_t *bf_nvert, const int32_t *bf_vert, const int32_t *bf_tag);
/* periodic_mesh = .true. of build_mesh (src/mesh/mesh_connectivity_module.F90:44-54):
 * stitch_periodic_faces (:976-1073) pairs the faces on x, y, z = -0.5 / 0.5
 * whose other centre coordinates agree within 1e-12, and build_periodic_stencil
 * (:1075-1328) appends each paired vertex's partner sub-faces and sub-elements to
 * its lists.  Applies to a built host mesh before subfv_mesh_upload; contexts
 * that uploaded it before keep the non-periodic one.  The viscous terms and the
 * Lagrange / acoustic sweeps return SUBFV_E_INVALID on a periodic mesh. */
int subfv_host_mesh_make_periodic(subfv_host_mesh *m);
int64_t subfv_host_mesh_get_i32(subfv_host_mesh *m, const char *name, int32_t *out, int64_t cap);
void subfv_host_mesh_free(subfv_host_mesh *m);

int subfv_set_config(subfv_ctx *ctx, const subfv_config *cfg);
/* No counterpart in the reference: the arithmetic of the node sweep (the
 * per-subface flux and nodal-solver kernel).  SUBFV_ARITH_EXACT (the default)
 * evaluates the reference's expressions as written, bit-identical to it;
 * SUBFV_ARITH_FAST compiles it with products contracted into fused
 * multiply-adds and divisions evaluated as products with reciprocals (one
 * reciprocal where a divisor repeats, as across the nodal-solver
 * iterations): results within 1e-12 relative of the reference's after 200
 * steps (tested), not bit-identical.  Every other kernel stays exact. */
#define SUBFV_ARITH_EXACT 0
#define SUBFV_ARITH_FAST 1
int subfv_set_arith(subfv_ctx *ctx, int32_t mode);
/* physical time t used by time-dependent boundaries and the tmax clamp */
int subfv_set_time(subfv_ctx *ctx, double t);

int subfv_state_put(subfv_ctx *ctx, const double *sol);
/* init_sol with init_1drp (src/ns/ns_euler_module.F90:36-43) on the device:
 * the state of each cell is primit_to_conserv(w_left) where its centroid's x
 * is below x1drp, primit_to_conserv(w_right) elsewhere ((rho, u, v, w, p)). */
int subfv_init_1drp(subfv_ctx *ctx, double x1drp, const double *w_left, const double *w_right);
int subfv_state_get(subfv_ctx *ctx, double *sol);

/* Device arrays in reference layout, by name (scalars: "dt_solve", the dt of
 * the last subfv_add_timedisc_and_solve solve before its positivity shrink;
 * "local_time_step", 1 or 0 as configured): "sol", "rhs", "rhs_o2",
 * "sum_lambda", "sum_lambda_o2", "cell_grad", "nodal_grad", "vp", and with
 * the viscous terms "diag" (5, 5, n_elems) and "rhs_diff" (4, n_elems: the
 * viscous rhs of rho u, rho v, rho w, E). */
int64_t subfv_get_f64(subfv_ctx *ctx, const char *name, double *out, int64_t cap);
int64_t subfv_put_f64(subfv_ctx *ctx, const char *name, const double *in, int64_t n);

/* compute_euler_flux on the device state.  second_order = 0 writes "rhs"
 * and "sum_lambda"; second_order = 1 reconstructs with "cell_grad" and writes
 * "rhs_o2" and "sum_lambda_o2".  want_vp also writes the nodal velocity. */
int subfv_compute_euler_flux(subfv_ctx *ctx, int32_t second_order, int32_t want_vp);
/* compute_cell_grad_from_nodal_grad: writes "nodal_grad" and "cell_grad" */
int subfv_compute_cell_grad(subfv_ctx *ctx, int32_t method);
/* flux_limiting(sol, rhs, rhs_o2, sum_lambda, cfl): updates "rhs" */
int subfv_flux_limiting(subfv_ctx *ctx, double cfl);
/* add_timedisc_and_solve(sol, rhs, diag, sum_lambda, ..., a1, a2, cfl, t, dt)
 * on the device state; *dt receives the value the reference leaves in dt,
 * *omega the positivity factor. */
int subfv_add_timedisc_and_solve(subfv_ctx *ctx, double a1, double a2, double cfl, double t, double *dt,
                                 double *omega);
/* compute_rhs_ns: Euler fluxes, second-order reconstruction and flux
 * limiting, then the viscous block when subfv_set_diffusion activated it */
int subfv_compute_rhs_ns(subfv_ctx *ctx);

/* Viscous terms of compute_rhs_ns (activate_diffusion, ns_main.F90:170-175):
 * build_momentum_diff_mat_into_global_mat_diag (ns_vectorial_diffusion_module.F90:200-320)
 * and build_diff_mat_into_global_mat_diag (ns_diffusion_module.F90:112-184) with a
 * viscosity mu_p (compute_mu without Sutherland, :322-347; subfv_set_sutherland
 * switches to the per-cell law), Cv_p and the
 * Prandtl number (a compile-time 0.71 in the reference).  Per BC index
 * i = 1..n_bc (arrays 0-based): bc_v_id / bc_t_id 1 = neumann, 2 = dirichlet
 * (bc_type_V / bc_type_T), bc_val_v(3, n_bc), bc_val_t(n_bc).  Applies to
 * subfv_compute_rhs_ns (rhs and "diag") and to the stages of subfv_run.
 * Vertices may have at most 64 sub-faces and 64 sub-elements (a mesh whose
 * vertices all have at most 16 takes kernels with the setup's matrices in LDS;
 * others, e.g. tetrahedral meshes, keep them in a global workspace). */
int subfv_set_diffusion(subfv_ctx *ctx, int32_t activate, double mu_p, double cv_p, double prandtl,
                        int32_t n_bc, const int32_t *bc_v_id, const double *bc_val_v, const int32_t *bc_t_id,
                        const double *bc_val_t);

/* use_sutherland of compute_mu (ns_vectorial_diffusion_module.F90:322-359,
 * namelist entries use_sutherland, mu0, T0 of ns_global_data_module.F90:13-20):
 * with activate != 0 the viscous terms take each cell's viscosity
 * mu0 (T/T0)^1.5 (T0 + 110.4) / (T + 110.4), T = temp_u of the cell state,
 * and mu0corr = the minimum over the cells; mu_p of subfv_set_diffusion is
 * then unused.  The nodal systems are re-formed before every residual.
 * Stays set across subfv_set_diffusion calls. */
int subfv_set_sutherland(subfv_ctx *ctx, int32_t activate, double mu0, double T0);
/* compute_mu (ns_vectorial_diffusion_module.F90:322-347): mu_arr(n_elems) and
 * mu0corr for the state sol(5, n_elems): Sutherland's law per cell on the
 * device when subfv_set_sutherland activated it (sol then becomes the
 * context's state), else mu_p of subfv_set_diffusion (mu0corr = mu_p). */
int subfv_compute_mu(subfv_ctx *ctx, const double *sol, double *mu_arr, double *mu0corr);
/* build_momentum_diff_mat_into_global_mat_diag (ns_vectorial_diffusion_module.F90:200-320,
 * parts = 1) and build_diff_mat_into_global_mat_diag (ns_diffusion_module.F90:112-184,
 * parts = 2; 3 = both) for the state sol(5, n_elems), the cell viscosities
 * mu_arr(n_elems) and mu0corr: their contributions are added to the caller's
 * rhs(5, n_elems) (rows 2-5) and diag(5, 5, n_elems), as the reference's
 * builders add them.  The nodal systems are formed on the device for the given
 * viscosities; the other viscous parameters come from subfv_set_diffusion;
 * sol becomes the context's state. */
int subfv_add_diffusion(subfv_ctx *ctx, const double *sol, const double *mu_arr, double mu0corr, int32_t parts,
                        double *rhs, double *diag);

/* The time loop on device-resident state without host synchronisation.
 * time_scheme 0: the reference loop (forward Euler, or its two-stage scheme
 * when second_order is set); 1: SSP-RK3 with one dt per step.  Runs n_steps
 * steps or until t >= tmax; *t and *dt receive the final values. */
int subfv_run(subfv_ctx *ctx, int32_t n_steps, int32_t time_scheme, double *t, double *dt);

/* Kernel timing (no counterpart in the reference): with profiling on, HIP
 * events bracket each phase ("node_flux", "node_flux_o2", "grad", "gather",
 * "diffusion" and, inside it, "diff_apply", "update") on the context's
 * stream; read returns the summed milliseconds
 * and the number of launches since profiling was (re)enabled. */
int subfv_profile_enable(subfv_ctx *ctx, int32_t on);
int subfv_profile_read(subfv_ctx *ctx, const char *phase, double *ms_total, int64_t *count);

/* ------------------------------------------------ lagrange / acoustic sweeps */
/* compute_rhs_lagrange (src/lagrange/lagrange_module.F90:14-147) on the
 * context's mesh: sol[C][5] = (rho, u, v, w, E), mass[C]; writes rhs[C][5]
 * and the nodal velocities vp[V][3].  b2d as boundary_2d.  Each cell's rhs
 * receives its terms in the reference's vertex / sub-face order. */
int subfv_compute_rhs_lagrange(subfv_ctx *ctx, const double *sol, const double *mass, int32_t b2d, double *rhs,
                               double *vp);
/* n_steps of the acoustic loop (src/acoustic/acoustic_main.F90:171-290:
 * nodal pressure, multi_point_pressure fluxes in the face frame of rotate,
 * dt = min vol / (face area * a), sol += cfl dt / vol * flux, t += dt) on
 * sol[C][4] = (p, u, v, w) in place.  Boundary tag t (1..n_bc) has
 * bc_kind[t-1]: 0 wall, 1 freestream (state bc_val[t-1][4]), 2 outflow;
 * untagged faces are walls.  *t advances; *dt_last gets the last step. */
int subfv_acoustic_run(subfv_ctx *ctx, double *sol, int32_t n_steps, double cfl, int32_t n_bc, const int32_t *bc_kind,
                       const double *bc_val, double *t, double *dt_last);

/* ------------------------------------------------------------------ BCSR */
typedef struct subfv_bcsr subfv_bcsr;
/* bcsr_matrix_type: m block rows, b x b blocks, 1-based row(m+1), col(nnzb),
 * v(b,b,nnzb) column-major blocks.  Copies to the device. */
int subfv_bcsr_create(subfv_ctx *ctx, subfv_bcsr **A, int32_t m, int32_t b, int32_t nnzb, const int32_t *row,
                      const int32_t *col, const double *v);
void subfv_bcsr_destroy(subfv_bcsr *A);
/* y = A x (bcsr_matmul), host vectors of length b*m */
int subfv_bcsr_matmul(subfv_bcsr *A, const double *x, double *y);
/* x = D^-1 rhs, block-Jacobi via lu_solve_noswp (bcsr_diag) */
int subfv_bcsr_diag(subfv_bcsr *A, double *x, const double *rhs);
/* bcsr_gmres_diag(A, x, b, tol, m, conv, ...): restarted right-preconditioned
 * GMRES(restart) from x = 0 with absolute tolerance tol on ||b - Ax||_2.
 * precision 0: FP64 throughout; 1: FP32 SpMV, block-Jacobi and Krylov basis
 * inside each cycle with FP64 orthogonalisation sums, residuals and restarts
 * (iterative refinement).  stats[0] = restarts, stats[1] = final ||b - Ax||_2,
 * stats[2] = SpMV count. Returns SUBFV_E_NOT_CONVERGED (soft) if not conv. */
int subfv_bcsr_gmres_diag(subfv_bcsr *A, double *x, const double *rhs, double tol, int32_t restart,
                          int32_t precision, int32_t *conv, double *stats);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): subfv_bcsr_gmres_diag with its Krylov basis V and
 * Hessenberg matrix H in a file-scope device array of 4 Mi doubles (32 MiB per
 * GPU, held until the process ends) instead of memory the library allocates.
 * Requires n'*(restart + 1) + (restart + 1)*restart <= 4194304 (4 Mi), with n'
 * = m*b rounded up to even, else SUBFV_E_INVALID with a message; results
 * equal subfv_bcsr_gmres_diag's.
 * Calls that overlap on one GPU share the array and race. */
int subfv_bcsr_gmres_diag_noalloc(subfv_bcsr *A, double *x, const double *rhs, double tol, int32_t restart,
                                  int32_t precision, int32_t *conv, double *stats);
/* The other GMRES loops on the same interface: precond 0 = bcsr_gmres
 * (:194-290, no preconditioner, 20 restarts), 1 = bcsr_gmres_diag (as above),
 * 2 = bcsr_gmres_lusgs (:447-548, bcsr_lusgs inside every Arnoldi step and on
 * the combination, 20 restarts), 3 = bcsr_gmres_parallel_lusgs (:885-1037,
 * with the colour sets of subfv_bcsr_set_colors).  precision 1 (FP32 inner)
 * is accepted for precond 0 and 1 only. */
int subfv_bcsr_gmres(subfv_bcsr *A, double *x, const double *rhs, double tol, int32_t restart, int32_t precond,
                     int32_t precision, int32_t *conv, double *stats);
/* bcsr_lusgs (:1268-1335) as the reference writes it: (L + I) y = rhs, then
 * (L + D) x = y with D^-1 by lu_solve_noswp, L the blocks j = row(i), ...
 * while col(j) < i.  The device runs the rows of each dependency level of L
 * together (the level schedule is built by subfv_bcsr_create). */
int subfv_bcsr_lusgs(subfv_bcsr *A, double *x, const double *rhs);
/* colour sets of bcsr_parallel_lusgs (color_set_type): color[i] in 1..n_sets
 * per block row, or NULL for a greedy colouring of the symmetric closure of the
 * block pattern (rows in order, smallest free colour).  *n_sets receives the
 * number of sets. */
int subfv_bcsr_set_colors(subfv_bcsr *A, const int32_t *color, int32_t *n_sets);
/* bcsr_parallel_lusgs (:1438-1502): forward over the colour sets with the
 * blocks of lower colour, backward with the blocks of higher colour and
 * D^-1 by lu_solve_noswp. */
int subfv_bcsr_parallel_lusgs(subfv_bcsr *A, double *x, const double *rhs);
/* bcsr_pcgm (:1110-1169): preconditioned CG from the given x (in/out) with
 * lu_solve(3, ...) on the first nine values of each diagonal block, as the
 * reference hard-codes.  stats[0] = iterations, stats[1] = sqrt(M^-1 r . r).
 * SUBFV_E_NOT_CONVERGED where the reference stops with "did not converge",
 * SUBFV_E_INVALID where lu_solve stops on a singular block. */
int subfv_bcsr_pcgm(subfv_bcsr *A, double *x, const double *rhs, double tol, double *stats);
/* New values v(b,b,nnzb) for the same pattern (the reference refills
 * bcsr_matrix_type%v in place between solves): device copy, FP32 copy and
 * block-Jacobi factors refreshed; no allocation. */
int subfv_bcsr_set_values(subfv_bcsr *A, const double *v);
/* add_block_to_bcsr (:56-71): v(:,:,k) += M (b x b column-major) for the first
 * stored block (i, j), 1-based; nothing if the block is not stored. */
int subfv_bcsr_add_block(subfv_bcsr *A, int32_t i, int32_t j, const double *M);

/* ------------------------------------------------------- served batches */
/* No counterpart in the reference (it runs one solve per process): a batch
 * advances many independent contexts together, one multi-instance launch per
 * kernel of a step instead of one launch per context.  Each context keeps
 * its own mesh, state, dt and omega, so its results are bit-identical to
 * subfv_run(ctx, n_steps, 0, ...) alone.  Members must share the device and
 * the scheme and run the first-order inviscid forward-Euler loop.
 * subfv_batch_run enqueues n_steps steps on the batch's stream and returns.
 * use_graph 0: six multi-instance launches per step; 1: whole 3-step cycles
 * replay one captured hipGraph (re-captured when the membership changes);
 * 2: one persistent cooperative launch for all n_steps (a resident grid
 * with device-side work queues and grid barriers between the phases).
 * Call subfv_batch_wait before using a member context again.
 * subfv_batch_set_grid: workgroups per CU of the persistent launch (0, the
 * default: as many as fit). */
typedef struct subfv_batch subfv_batch;
int subfv_batch_create(subfv_batch **b, int device);
void subfv_batch_destroy(subfv_batch *b);
const char *subfv_batch_error(const subfv_batch *b);
int subfv_batch_set_grid(subfv_batch *b, int32_t wgs_per_cu);
int subfv_batch_run(subfv_batch *b, subfv_ctx **ctxs, int32_t n, int32_t n_steps, int32_t use_graph);
int subfv_batch_wait(subfv_batch *b);

#ifdef __cplusplus
}
#endif
#endif
