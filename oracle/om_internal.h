/*
 * TEST INFRASTRUCTURE ONLY -- the oracle.
 *
 * CPU restatement of subfv's explicit-residual path (mesh build, geometry,
 * subface fluxes + nodal solvers, gradients, limiters, stage update) and of
 * its BCSR / block-Jacobi / GMRES routines.  Only tests/, the smoke check and
 * the cpu_baseline leg of bench.py may load it, and only as the checker.
 * The product (subfv_amd) never links or calls anything in this directory.
 *
 * All ids are 1-based and every array is allocated with one spare slot so
 * that the loops read like the Fortran they restate.  Boundary encoding of a
 * face's right neighbour follows the reference: >0 cell id, 0 default wall,
 * <0 minus the boundary-condition index.
 */
#ifndef OM_INTERNAL_H
#define OM_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

#define OM_MNBC 10

/* scheme ids: src/ns/ns_global_data_module.F90:63-72 */
enum {
  OM_SCHEME_MULTI_POINT = 1,
  OM_SCHEME_MULTI_POINT_ISO = 2,
  OM_SCHEME_THREE_WAVE = 3,
  OM_SCHEME_TWO_WAVE = 4,
  OM_SCHEME_MODIFIED_THREE_WAVE = 5,
  OM_SCHEME_MULTI_POINT_PRESSURE = 6,
  OM_SCHEME_MULTI_POINT_PRESSURE_PH = 8
};

/* Euler BC ids: src/ns/ns_global_data_module.F90:50-58 */
enum {
  OM_BC_WALL = 1,
  OM_BC_ADHERENCE_WALL = 2,
  OM_BC_FREESTREAM = 3,
  OM_BC_OUTFLOWSUPERSONIC = 4,
  OM_BC_INFLOW_POND = 5,
  OM_BC_INOUT_DOUBLE_MACH = 6,
  OM_BC_DOUBLE_MACH_BOTTOM = 7,
  OM_BC_POTENTIAL_FLOW_2D = 8,
  OM_BC_POTENTIAL_FLOW_3D = 9
};

/* Error codes shared with the product's C-ABI numbering. */
enum {
  OM_OK = 0,
  OM_ERR_BAD_OMEGA = 3,
  OM_ERR_NEG_STAR_DENSITY = 4,
  OM_ERR_MESH_GEOMETRY = 5,
  OM_ERR_BAD_INPUT = 6
};

typedef struct {
  int scheme_id;
  int bc_style;
  int boundary_2d;
  int exclude_bound_vert;
  int local_time_step;
  int second_order;
  int method;
  int n_bc;
  int bc_euler_id[OM_MNBC + 1];      /* 1-based */
  double bc_val[OM_MNBC + 1][5];     /* primitive (rho,u,v,w,p) per BC */
  double cfl;
  double tmax;
  double t;                          /* physical time seen by time-dependent BCs */
  /* viscous terms (compute_rhs_ns with activate_diffusion, ns_main.F90:170-175) */
  int activate_diffusion;
  double mu_p, Cv_p, prandtl;
  int bc_V_id[OM_MNBC + 1];          /* 1 neumann, 2 dirichlet (ns_global_data_module.F90:40-41) */
  int bc_T_id[OM_MNBC + 1];
  double bc_val_V[OM_MNBC + 1][3];
  double bc_val_T[OM_MNBC + 1];
  /* compute_mu with use_sutherland (ns_vectorial_diffusion_module.F90:322-359):
   * mu0 at T0 (ns_global_data_module.F90:17-20) */
  int use_sutherland;
  double mu0, T0;
} om_config;

typedef struct {
  int nv, ne, nf, nsf, nse, nbf, nmf;
  int nmax_sfn;
  /* vertices */
  double *vcoord;   /* 3*(nv+1) */
  double *vvol;
  int *vbound;
  int *v_nen, **v_elem;
  int *v_nfn, **v_face;
  int *v_nsfn, **v_sf;
  int *v_nsen, **v_se;
  /* elements */
  int *e_kind, *e_nv, *e_nf, *e_nsf, *e_nse, *e_tag, *e_nnbv;
  int **e_vert, **e_face, **e_sf, **e_se, **e_nbv;
  double *ecoord;   /* 3*(ne+1) */
  double *evol;
  /* faces */
  int *f_nv, *f_left, *f_leftf, *f_right, *f_rightf, *f_nsf;
  int **f_vert, **f_sf;
  double *fcoord, *fnorm, *farea;
  /* sub-faces */
  int *sf_vert, *sf_face, *sf_loc, *sf_le, *sf_re, *sf_lse, *sf_rse;
  double *sf_area, *sf_norm;
  /* sub-elements */
  int *se_loc, *se_vert, *se_elem, *se_nsf;
  int **se_sf;
  double *se_vol;
  /* reading-time data (minimal faces, boundary faces) */
  int *mf_nv, *mf_left, *mf_leftf, *mf_right, *mf_rightf;
  int **mf_vert;
  int *bf_nv, *bf_tag;
  int **bf_vert;
  /* renumbering maps kept for inspection */
  int *vert_nto;    /* new vertex id -> file vertex id */
  int *elem_nto;    /* new element id -> file element id */
  /* arena */
  void *arena;
} om_mesh;

void *om_arena_alloc(om_mesh *m, size_t bytes);

/* viscous block of compute_rhs_ns (om_diffusion.c) */
int om_add_diffusion(const om_mesh *m, const om_config *cfg, const double *sol, double *rhs, double *diag);
int om_add_diffusion_mu(const om_mesh *m, const om_config *cfg, const double *sol, const double *mu_cells,
                        double mu0corr, int parts, double *rhs, double *diag);

#endif
