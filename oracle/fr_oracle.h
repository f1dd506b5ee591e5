/*
 * fr_oracle.h -- CPU restatement of the HiFiLES flux-reconstruction residual and
 * RK update.  TEST INFRASTRUCTURE ONLY: this library is the checker that the GPU
 * path is compared with.  Only tests/, __graft_entry__.smoke() and the bench's CPU
 * baseline leg may load it; the product never links it.
 *
 * The data flow follows the reference CPU build (-D_CPU -D_NO_BLAS):
 *   CalcResidual            src/solver.cpp:65-225
 *   eles::*                 src/eles.cpp:1256-2453, 3294-3368, 3620-4052, 4557-4847
 *   eles_hexas / eles_quads src/eles_hexas.cpp, src/eles_quads.cpp
 *   int_inters              src/int_inters.cpp:225-470
 *   inters (fluxes, LUT)    src/inters.cpp:311-678
 *   flux.cpp                src/flux.cpp:36-431
 *   setup_params            src/input.cpp:345-515
 * Arrays use the reference's column-major layouts, first index fastest:
 *   disu_upts(Nu,N,nf), disu_fpts(Nfp,N,nf), grad_disu_upts(Nu,N,nf,d),
 *   grad_disu_fpts(Nfp,N,nf,d), div_tconf_upts(Nu,N,nf), detjac_upts(Nu,N).
 *
 * Parity status of each function is recorded in the header comment of
 * fr_oracle.c.
 */
#ifndef FR_ORACLE_H
#define FR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Raw input-deck values (src/input.cpp:89-331, same defaults). */
typedef struct {
  int equation;               /* 0 Euler/NS */
  int viscous;
  int riemann_solve_type;     /* 0 Rusanov, 2 Roe */
  int vis_riemann_solve_type; /* 0 LDG */
  int ic_form;                /* 0 isentropic vortex (forces mu_inf=0.1 etc.) */
  int order;
  int adv_type;               /* 0 forward Euler, 3 RK45 */
  double dt;
  double tau, pen_fact;
  double gamma, prandtl, S_gas, T_gas, R_gas, mu_gas, fix_vis;
  double Mach_free_stream, nx_free_stream, ny_free_stream, nz_free_stream;
  double Re_free_stream, L_free_stream, T_free_stream, T_wall;
  double Mach_wall, nx_wall, ny_wall, nz_wall;
  int upts_type;              /* 0 Gauss (only supported point set) */
  int vcjh_scheme;            /* 0 user eta, 1 DG, 2 SD, 3 Hu, 4 c+ */
  double eta;                 /* used when vcjh_scheme == 0 */
  int LES, SGS_model;         /* LES 1 with SGS_model 0 Smagorinsky, 1 WALE (no wall model);
                                 10: supplied eddy viscosity (or_set_sgs_nu), GPU-build form */
  double filter_ratio;
  double const_src;
  double dx_cyclic, dy_cyclic, dz_cyclic;
} or_deck;

/* Non-dimensional parameters actually used by the residual. */
typedef struct {
  double gamma, prandtl, prandtl_t, mu_inf, rt_inf, c_sth, R_ref, fix_vis;
  double tau, pen_fact, T_wall, const_src, dt;
  double v_wall[3];
} or_phys;

typedef struct or_solver or_solver;

void or_deck_defaults(or_deck* d);
/* src/input.cpp:345-515 */
void or_setup_params(const or_deck* d, or_phys* out);

/* ele_type uses the reference's numbering: 1 quad, 4 hex.
 * xv: n_verts x n_dims, row per vertex.
 * c2v: n_eles x n_verts_per_ele, in the reference's INTERNAL shape-node order
 *      (i.e. after the Gambit permutation of src/geometry.cpp:1800-1855).
 * Periodic faces are matched by translation over dx/dy/dz_cyclic of the deck. */
or_solver* or_create(const or_deck* deck, int ele_type, int n_verts, const double* xv,
                     int n_eles, const int* c2v);
/* As or_create, with boundary faces (element, local face, bctype 11 isothermal
 * no-slip wall) taken out of the face matching. */
or_solver* or_create_bc(const or_deck* deck, int ele_type, int n_verts, const double* xv, int n_eles,
                        const int* c2v, int n_bfaces, const int* bf_ele, const int* bf_face,
                        const int* bf_type);
void or_destroy(or_solver* s);

int or_n_upts(const or_solver* s);
int or_n_fpts(const or_solver* s);
int or_n_fields(const or_solver* s);
int or_n_dims(const or_solver* s);
int or_n_eles(const or_solver* s);
int or_n_inters(const or_solver* s);
void or_get_phys(const or_solver* s, or_phys* out);

/* what: 0 disu_upts(level0) 1 disu_upts(level1) 2 div_tconf_upts 3 disu_fpts
 *       4 grad_disu_upts 5 grad_disu_fpts 6 detjac_upts 7 JGinv_upts(d,d,Nu,N)
 *       8 tdA_fpts 9 norm_fpts(Nfp,N,d) 10 pos_upts(Nu,N,d) 11 norm_tconf_fpts
 *       12 opp_0 (Nfp,Nu) 13 opp_3 (Nu,Nfp) 14 opp_4 (Nu,Nu,d) 15 loc_upts(d,Nu)
 *       16 tloc_fpts(d,Nfp) 17 pos_fpts(Nfp,N,d)
 * Returns the number of doubles copied (out may be NULL to query). */
long or_get(const or_solver* s, int what, double* out);
void or_set_state(or_solver* s, const double* disu_upts);
void or_zero_rk_register(or_solver* s);
/* Taylor-Green initial state of src/eles.cpp:555-568 (ic_form 7, 3-D). */
void or_set_ics_tgv(or_solver* s, double rho);

/* Interface table: for each interior/periodic face point, the element flux-point
 * indices (fpt + Nfp*ele) of the left and right sides.  n = n_inters*pts/face. */
long or_get_face_pairs(const or_solver* s, int* left, int* right);

/* One CalcResidual (src/solver.cpp:65-225) on disu_upts level 0. */
void or_residual(or_solver* s);
/* AdvanceSolution (src/eles.cpp:1003-1198): adv_type 0 (forward Euler) or 3 (RK45). */
void or_advance(or_solver* s, int rk_stage, int adv_type);
/* n_steps full time steps with the deck's adv_type. */
void or_run_steps(or_solver* s, int n_steps);
/* CalcNormResidual (src/output.cpp:1452-1522), norm_type 0 inf, 1 L1, 2 L2. */
void or_norm_residual(const or_solver* s, int norm_type, double* out5);
/* SGS_model 10: kinematic eddy viscosity at every solution point (Nu,N). */
void or_set_sgs_nu(or_solver* s, const double* nu_t);
/* Replaces the computed metrics (layouts of or_get 6..9); the residual reads them as given. */
void or_set_metrics(or_solver* s, const double* detjac_upts, const double* JGinv_upts, const double* tdA_fpts,
                    const double* norm_fpts);
/* Number of OpenMP threads used by the dense loops (1 = serial). */
void or_set_threads(int n);
/* CalcIntegralQuantities (src/eles.cpp:6503-6643) for the state and the corrected
 * gradient of the last or_residual, with the cubature rule loc (d, ncub), weights
 * (ncub); kinds 0 kineticenergy, 1 vorticity, 2 pressuredilatation,
 * 3 straincolonproduct, 4 devstraincolonproduct; out[nq]. */
void or_integral_quantities(const or_solver* s, int ncub, const double* loc, const double* wts, int nq,
                            const int* kinds, double* out);
/* Common normal flux at one point (static mesh): riemann 0 Rusanov, 2 Roe;
 * ul, ur, fn: nd + 2 conserved values; n: unit normal (nd). */
void or_common_flux(int nd, int riemann, double gamma, const double* ul, const double* ur, const double* n,
                    double* fn);

#ifdef __cplusplus
}
#endif
#endif
