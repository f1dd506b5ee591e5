/*
 * hifr_solver.h -- C API of the host solver that sits above the device ABI
 * (hifr.h).  It restates the reference's driver: read the input deck and mesh
 * (src/input.cpp, src/geometry.cpp GeoPreprocess), set up the elements and
 * interfaces, set the initial condition or read a restart (InitSolution,
 * src/solver.cpp:352-444), then CalcResidual / AdvanceSolution / CalcNormResidual.
 * The reference has no library API (only the HiFiLES executable); this API is what
 * the bundled `hifr_run` executable and the Python tests call.
 *
 * All functions return 0 on success; on failure a message is copied into err
 * (errlen bytes) when err is non-NULL.
 */
#ifndef HIFR_SOLVER_H
#define HIFR_SOLVER_H

#include "hifr.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hifr_solver hifr_solver;

/* Deck file + the Gambit mesh it names. */
int hifr_solver_from_deck(const char* deck_path, hifr_solver** out, char* err, int errlen);

/* Deck given as text; mesh given as arrays (one linear element type; c2v in the
 * internal shape-node order). */
int hifr_solver_from_arrays(const char* deck_text, int ele_type, int n_verts, const double* xv, int n_eles,
                            const int* c2v, hifr_solver** out, char* err, int errlen);

/* As hifr_solver_from_arrays, with n_bfaces physical boundary faces given as
 * (element, local face in the reference's numbering, bctype) triples (Gambit
 * boundary groups, src/geometry.cpp:1122-1290); bctype 11 (isothermal no-slip
 * wall) is built. */
int hifr_solver_from_arrays_bc(const char* deck_text, int ele_type, int n_verts, const double* xv, int n_eles,
                               const int* c2v, int n_bfaces, const int* bfaces, hifr_solver** out, char* err,
                               int errlen);

/* Deck given as text; structured box mesh of n[0] x n[1] (x n[2]) elements on
 * [o, o+l], vertex coordinates rounded to `digits` significant digits (0: exact). */
int hifr_solver_box(const char* deck_text, int ele_type, const int* n, const double* l, const double* o,
                    int digits, hifr_solver** out, char* err, int errlen);

void hifr_solver_destroy(hifr_solver* s);

/* n_upts, n_fpts per element, n_fields, n_dims, n_eles, n_inters */
void hifr_solver_sizes(const hifr_solver* s, int* out6);

/* Host arrays (reference layouts); a host-only solver (deck key hifr_host_only 1)
 * has no device context and returns the state it holds for 0, and -1 for 1-3, 14:
 *  0 disu_upts (Nu,N,nf)  1 div_tconf_upts  2 disu_fpts (Nfp,N,nf)  3 grad_disu_fpts (Nfp,N,nf,d)
 *  4 detjac_upts (Nu,N)   5 pos_upts (Nu,N,d)  6 opp_0  7 opp_3  8 opp_4
 *  9 JGinv_upts (d,d,Nu,N) 10 tdA_fpts (Nfp,N) 11 norm_fpts (Nfp,N,d) 12 xv (n_verts,d) 13 c2v (as doubles)
 *  14 rk register (Nu,N,nf)
 *  15 loc_volume_cubpts (d,ncub) 16 weight_volume_cubpts (ncub) 17 opp_volume_cubpts (ncub,Nu)
 *  18 detjac at the cubature points (ncub,N)   (15-18: decks with integral_quantities)
 * Returns the number of doubles (out may be NULL to query). */
long hifr_solver_get(hifr_solver* s, int what, double* out);
int hifr_solver_set_state(hifr_solver* s, const double* disu_upts);
/* Flux-point pairs of the interior/periodic interfaces (fpt + Nfp*ele). */
long hifr_solver_interfaces(const hifr_solver* s, int* left, int* right);

int hifr_solver_residual(hifr_solver* s);
int hifr_solver_advance(hifr_solver* s, int stage);
int hifr_solver_run(hifr_solver* s, int n_steps);
int hifr_solver_norms(hifr_solver* s, int norm_type, double* out);
/* Writes the state in the reference's restart format (src/output.cpp:1057-1107;
 * the reference names the file Rest_<iter>_p0000.dat), 15 significant digits. */
int hifr_solver_write_restart(hifr_solver* s, const char* path);
/* Reads a restart file in the reference's format (src/solver.cpp:390-444,
 * src/eles.cpp:623-716) into the state; its solution points must be the solver's.
 * Sets the solution time from the file. */
int hifr_solver_read_restart(hifr_solver* s, const char* path, char* err, int errlen);
double hifr_solver_time(const hifr_solver* s);
/* The deck's integral_quantities (src/input.cpp:133-143) with its volume_cub_order
 * rule: their number, and their values for the current state
 * (CalcIntegralQuantities, src/output.cpp:1241-1270, src/eles.cpp:6503-6643). */
int hifr_solver_n_integrals(const hifr_solver* s);
int hifr_solver_integrals(hifr_solver* s, double* out);
/* Flux points on physical boundaries (walls). */
long hifr_solver_n_boundary_points(const hifr_solver* s);
hifr_ctx* hifr_solver_ctx(hifr_solver* s);
double hifr_solver_dt(const hifr_solver* s);
int hifr_solver_adv_type(const hifr_solver* s);

#ifdef __cplusplus
}
#endif
#endif
