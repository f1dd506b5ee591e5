/*
 * hifr.h -- C ABI of the MI355X (gfx950) flux-reconstruction residual and RK update.
 *
 * This is the drop-in boundary for the GPU side of HiFiLES's hot path.  In the
 * reference the element and interface classes call C++-linkage kernel wrappers
 * and legacy cuBLAS directly (include/cuda_kernels.h:29-118, src/eles.cpp:1313,
 * 1470, 1566, 1644, 1648, 1738, 1922, 2025).  Here the same classes hand their
 * operators, metrics, face tables and state to one context and call one entry
 * point per residual / RK stage; all device memory is owned by the context.
 *
 * Plain C types only: host arrays are double / int32 in the reference's
 * column-major layouts (first index fastest), as filled by eles_* and int_inters.
 * Every call returns a hifr_status; nothing calls exit().  One context drives one
 * element type on the current HIP device; calls on one context must not overlap.
 */
#ifndef HIFR_H
#define HIFR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  HIFR_OK = 0,
  HIFR_ERR_ARG = 1,          /* bad argument / size / order of calls */
  HIFR_ERR_UNSUPPORTED = 2,  /* element type, order or option not built */
  HIFR_ERR_DEVICE = 3,       /* HIP runtime error (no device, launch failure, OOM) */
  HIFR_ERR_NOT_TENSOR = 4,   /* tensor-product element operators lack the expected structure */
  HIFR_ERR_STATE = 5,        /* entry point called before its inputs were set */
  HIFR_ERR_NAN = 6           /* NaN in the residual (src/eles.cpp:1634-1638) */
} hifr_status;

/* Element types use the reference's numbering (src/eles_*.cpp setup_ele_type_specific).
 * Quads and hexes use sum-factorised operators (orders 1-6 / 1-5); tets use the
 * dense operators (orders 1-3, DG correction). */
enum { HIFR_QUAD = 1, HIFR_TET = 2, HIFR_HEX = 4 };

/* Arithmetic of the residual kernels.  HIFR_MIXED (no counterpart in the
 * reference): state, RK register, detJ, face normals / tdA and the RK update in
 * FP64; arithmetic, LDS, JGinv and the face traces in FP32, on the deviation of
 * the state from a constant reference at rest (the mean density and energy of the
 * uploaded state), so that FP32 never holds the O(p_ref) part of the fluxes. */
enum { HIFR_FP64 = 0, HIFR_MIXED = 1 };

/* What hifr_download can copy out (reference layouts). */
enum {
  HIFR_DISU_UPTS = 0,       /* disu_upts(0)     (Nu,N,nf)      */
  HIFR_RK_REGISTER = 1,     /* disu_upts(1)     (Nu,N,nf)      */
  HIFR_DIV_TCONF_UPTS = 2,  /* div_tconf_upts(0)(Nu,N,nf)     after hifr_residual */
  HIFR_DISU_FPTS = 3,       /* disu_fpts        (Nfp,N,nf)     current traces */
  HIFR_GRAD_DISU_FPTS = 4,  /* grad_disu_fpts   (Nfp,N,nf,d)   after hifr_residual (viscous; not
                                                                 refreshed by hifr_advance/hifr_run) */
  /* The metrics as the kernels read them (no counterpart in the reference), in
   * the layouts of hifr_set_metrics.  With per-element storage every point of an
   * element gets its element's value and every flux point its face's value; the
   * normal of a face is the outward normal its element uses in the interface
   * flux (on an interface, the left element's normal and its negation). */
  HIFR_DETJAC_UPTS = 5,     /* detjac_upts (Nu,N)     */
  HIFR_JGINV_UPTS = 6,      /* JGinv_upts  (d,d,Nu,N) */
  HIFR_TDA_FPTS = 7,        /* tdA_fpts    (Nfp,N)    */
  HIFR_NORM_FPTS = 8        /* norm_fpts   (Nfp,N,d)  */
};

/* Physical and numerical parameters after non-dimensionalisation
 * (run_input after input::setup_params, src/input.cpp:345-515). */
typedef struct {
  int ele_type;              /* HIFR_QUAD, HIFR_TET or HIFR_HEX */
  int order;                 /* polynomial order P */
  int n_eles;
  int viscous;               /* 0 Euler, 1 Navier-Stokes (LDG) */
  int riemann_solve_type;    /* 0 Rusanov; 2 Roe (FP64 without LES) */
  int vis_riemann_solve_type;/* 0 LDG */
  int precision;             /* HIFR_FP64 or HIFR_MIXED */
  double gamma, prandtl, mu_inf, rt_inf, c_sth, fix_vis;
  double tau, pen_fact;      /* LDG penalty and switch */
  double const_src;
  int metric_storage;        /* HIFR_METRICS_AUTO or HIFR_METRICS_PER_POINT */
  /* LES (src/eles.cpp:2321-2778, 3294-3368; src/int_inters.cpp:415-427): les 1 with
   * sgs_model 0 Smagorinsky or 1 WALE, eddy viscosity at the solution points with
   * filter width filter_ratio * vol^(1/d) / (P+1); requires viscous and constant
   * viscosity (fix_vis 1).  No wall model.  sgs_model HIFR_SGS_LEARNED (no
   * counterpart in the reference; hexes, P = 2, 3): the eddy viscosity of a
   * learned network (include/hifr_sgs.h), with the WALE model of the reference's
   * GPU build (src/cuda_kernels.cu:1466-1660) where no network is attached or its
   * inputs leave the training range. */
  int les, sgs_model;
  double filter_ratio;
  /* isothermal walls (bctype 11): gas constant R_ref and wall temperature T_wall,
   * non-dimensional (run_input after setup_params, src/input.cpp:476-512) */
  double R_ref, T_wall;
} hifr_config;

/* How the context stores the metrics it is given (no counterpart in the reference,
 * which always keeps them per point).  AUTO: when every element is affine (its
 * JGinv and detJ are the same at all solution points, and each face's normal and
 * tdA the same at all its flux points, to 1e-13 relative), one value per element
 * and per face is kept and read by the kernels; otherwise, and with PER_POINT,
 * the per-point arrays are used as given. */
enum { HIFR_METRICS_AUTO = 0, HIFR_METRICS_PER_POINT = 1 };

enum { HIFR_SGS_SMAGORINSKY = 0, HIFR_SGS_WALE = 1, HIFR_SGS_LEARNED = 10 };

typedef struct hifr_ctx hifr_ctx;

const char* hifr_status_string(hifr_status s);

/* Creates a context on the current HIP device.
 * Replaces eles::setup's device allocations (src/eles.cpp:84-405, mv_all_cpu_gpu :775). */
hifr_status hifr_create(const hifr_config* cfg, hifr_ctx** out);
hifr_status hifr_destroy(hifr_ctx* ctx);

/* Stream for all later work (a hipStream_t; NULL = default stream). */
hifr_status hifr_set_stream(hifr_ctx* ctx, void* hip_stream);

/* Element operators in the reference layouts, as built by eles::set_opp_0/3/4
 * (src/eles.cpp:3620-4052): opp_0 (Nfp,Nu), opp_3 (Nu,Nfp), opp_4 (Nu,Nu,d) with the
 * d matrices stacked.  opp_1, opp_2, opp_5, opp_6 follow from these
 * (src/eles.cpp:3700, 3771, 3956, 4018).  Replaces the opp_* cp_cpu_gpu uploads and
 * every cublasDgemm/bespoke_SPMV of src/eles.cpp.  For quads and hexes the
 * matrices must have the tensor-product structure of Lagrange bases on a 1-D
 * point set; the context keeps only their 1-D factors.  For tets it keeps the
 * dense matrices combined with the reference tet's face normals, as MFMA operand
 * tiles; HIFR_ERR_ARG if an entry is not finite. */
hifr_status hifr_set_operators(hifr_ctx* ctx, const double* opp_0, const double* opp_3,
                               const double* opp_4);

/* Static-mesh metrics of eles::set_transforms (src/eles.cpp:4557-4847):
 * detjac_upts (Nu,N), JGinv_upts (d,d,Nu,N), tdA_fpts (Nfp,N), norm_fpts (Nfp,N,d). */
hifr_status hifr_set_metrics(hifr_ctx* ctx, const double* detjac_upts, const double* JGinv_upts,
                             const double* tdA_fpts, const double* norm_fpts);

/* 1 if the context keeps one metric set per element (see HIFR_METRICS_AUTO), else 0. */
hifr_status hifr_metrics_per_element(const hifr_ctx* ctx, int* per_element);

/* Physical boundary flux points (bdy_inters::set_boundary pointer tables,
 * src/bdy_inters.cpp:114-148): the element flux-point index (fpt + Nfp*ele) and the
 * bctype of each point; bctype 11 (isothermal no-slip wall, static mesh) is built.
 * Call before hifr_set_interfaces, which then covers the remaining points; a
 * new call replaces the list (n_pts 0 clears it). */
hifr_status hifr_set_boundaries(hifr_ctx* ctx, int n_pts, const int32_t* fpt, const int32_t* bctype);

/* Interior and periodic interfaces: for each of n_pts interface points, the
 * element flux-point index (fpt + Nfp*ele) on the left and on the right.  This is
 * the resolved form of the pointer tables int_inters::set_interior builds with the
 * rot_tag lookup (src/int_inters.cpp:95-178, src/inters.cpp:187-296).  The left
 * side's normal is used for the whole point, as in the reference. */
hifr_status hifr_set_interfaces(hifr_ctx* ctx, int n_pts, const int32_t* left, const int32_t* right);

/* disu_upts(0) (Nu,N,nf); zeroes the RK register disu_upts(1)
 * (set_disu_upts_to_zero_other_levels, src/eles.cpp:407-425) and refreshes the traces. */
hifr_status hifr_upload_state(hifr_ctx* ctx, const double* disu_upts);
hifr_status hifr_download(hifr_ctx* ctx, int what, double* host);

/* One CalcResidual (src/solver.cpp:65-225): div_tconf_upts(0) on the device. */
hifr_status hifr_residual(hifr_ctx* ctx);

/* CalcResidual followed by AdvanceSolution (src/eles.cpp:1003-1198, dt_type 0):
 * adv_type 3 = low-storage RK45 stage `stage` (0..4), adv_type 0 = forward Euler. */
hifr_status hifr_advance(hifr_ctx* ctx, int stage, int adv_type, double dt);

/* n_steps full time steps (5 stages each for adv_type 3), src/HiFiLES.cpp:154-187.
 * adv_type 0 or 3 (else HIFR_ERR_UNSUPPORTED).  On meshes of at most 16,384
 * solution points every step but the last replays a HIP graph of one step,
 * captured on the first call (per trace-buffer parity, adv_type and dt; any
 * set-up call or new state drops it); the results equal the eager stages bit for
 * bit.  The last step runs eagerly and leaves HIFR_DIV_TCONF_UPTS current. */
hifr_status hifr_run(hifr_ctx* ctx, int n_steps, int adv_type, double dt);

/* CalcNormResidual (src/output.cpp:1452-1522) on the last residual:
 * norm_type 0 max, 1 L1, 2 L2 of div/J - const_src per field; out[nf]. */
hifr_status hifr_norm_residual(hifr_ctx* ctx, int norm_type, double* out);

/* No counterpart in the reference: hifr_norm_residual for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops).  It allocates
 * nothing and does not synchronise: div (n_pts points x n_fields fields, the layout
 * of HIFR_DIV_TCONF_UPTS) and detjac (one value per pts_per_detjac consecutive
 * points: 1 for per-point, Nu for per-element storage) are device arrays, out is a
 * device array of n_fields doubles written on hip_stream (NULL: the default
 * stream); a NaN in the residual propagates into out.  The per-block partial sums
 * live in a file-scope device array of 512 x 5 doubles per GPU, which limits
 * n_fields to 4 or 5 (HIFR_ERR_ARG otherwise); calls that overlap on one GPU race
 * on it. */
hifr_status hifr_norm_residual_noalloc(const double* div, const double* detjac, long n_pts, int pts_per_detjac,
                                       int n_fields, int norm_type, double const_src, void* hip_stream,
                                       double* out);

/* Integral diagnostics (eles::CalcIntegralQuantities, src/eles.cpp:6503-6643;
 * integral_quantities deck key, src/input.cpp:133-143): volume-cubature integrals
 * over the mesh of rho|v|^2/2 (kineticenergy), rho|omega|^2/2 (vorticity),
 * p div v (pressuredilatation), S:S (straincolonproduct) and S^d:S^d
 * (devstraincolonproduct). */
enum {
  HIFR_KINETIC_ENERGY = 0,
  HIFR_VORTICITY = 1,
  HIFR_PRESSURE_DILATATION = 2,
  HIFR_STRAIN_COLON = 3,
  HIFR_DEV_STRAIN_COLON = 4
};

/* The volume cubature the integrals use (eles_*::set_volume_cubpts,
 * eles::set_opp_volume_cubpts and the cubature-point detJ, src/eles_hexas.cpp:418-440,
 * src/eles.cpp:4111-4130, 5318-5352): n_cub points, the nodal basis there
 * opp_volume_cubpts (n_cub, Nu), weights (n_cub) and detjac (n_cub, N).  Quads and
 * hexes; HIFR_ERR_UNSUPPORTED for tets. */
hifr_status hifr_set_cubature(hifr_ctx* ctx, int n_cub, const double* opp_volume_cubpts, const double* weights,
                              const double* detjac_cubpts);

/* out[i] = integral of quantity kinds[i] (at most 5) for the current state.  The
 * velocity gradient is the corrected gradient (a viscous gradient pass of the
 * current state; viscous contexts only, except for the kinetic energy alone).  The
 * reference takes it from its last residual evaluation, which in its time loop is
 * the last RK stage's input state.  Synchronises the context's stream. */
hifr_status hifr_integral_quantities(hifr_ctx* ctx, int n_q, const int* kinds, double* out);

hifr_status hifr_synchronize(hifr_ctx* ctx);

/* Times n_steps of hifr_run (eager launches) with HIP events on the context's
 * stream.  total_ms: whole loop.  kernel_ms[i] / kernel_launches[i]: summed
 * duration and count of the timed launches of kernel kind i (0 trace, 1 gradient
 * trace, 2 flux/divergence/RK): the launches of the first stage of every step,
 * each bracketed by its own events.  Arrays may be NULL. */
hifr_status hifr_run_timed(hifr_ctx* ctx, int n_steps, int adv_type, double dt, double* total_ms,
                           double* kernel_ms, int* kernel_launches);

/* Measurement aid (no counterpart in the reference): allocates two buffers of
 * n_doubles (at most 2^34), runs a plain streaming copy kernel (one element per
 * thread) `reps` times on the default stream and returns the average kernel time
 * in *ms_per_copy.  Used to calibrate HBM byte counters and to state the
 * achievable copy bandwidth. */
hifr_status hifr_calibrate_copy(long n_doubles, int reps, double* ms_per_copy);

/* Sizes for callers that allocate host buffers. */
hifr_status hifr_sizes(const hifr_ctx* ctx, int* n_upts, int* n_fpts, int* n_fields, int* n_dims);

#ifdef __cplusplus
}
#endif
#endif
